!=======================================================================
! a12gpu_iface -- Fortran interface to liba12gpu.so (include/a12gpu.h),
! the MI355X implementation of the ATLAS12/SYNTHE hot path.
!
! The library takes the drivers' own arrays in their Fortran layout:
! REAL(4) <-> float, REAL(8) <-> double, default INTEGER <-> int32_t.
! Array arguments are passed by reference (assumed-size dummies), so a
! module array such as XLINES(kw, 30000) or opacity_matrix(L, nrhox) goes
! straight through.  Structs hold C pointers (C_LOC of TARGET arrays).
!
! Every function returns 0 (A12_OK) or a negative status; a12_errmsg()
! returns the context's last message as a Fortran string.
!=======================================================================
MODULE a12gpu_iface
  USE, INTRINSIC :: iso_c_binding
  IMPLICIT NONE

  INTEGER(C_INT32_T), PARAMETER :: A12_OK = 0, A12_ERR_ARG = -1, A12_ERR_HIP = -2, &
                                   A12_ERR_OOM = -3, A12_ERR_STATE = -4, &
                                   A12_ERR_NONFINITE = -5
  INTEGER(C_INT32_T), PARAMETER :: A12_PREC_MIXED = 0, A12_PREC_FP64 = 1

  ! lte_line_t, src/mod_mklinelist.f90:30-41 (32 bytes)
  TYPE, BIND(C) :: a12_lte_line
    INTEGER(C_INT32_T) :: nbuff
    REAL(C_FLOAT)      :: cgf
    INTEGER(C_INT32_T) :: nelion
    REAL(C_FLOAT)      :: elo, gamrf, gamsf, gamwf, code
  END TYPE a12_lte_line

  ! MOL2 32-byte record (SELECTLINES' TiO path)
  TYPE, BIND(C) :: a12_mol2_rec
    REAL(C_FLOAT)      :: wl, loggf, e1, e2
    INTEGER(C_INT32_T) :: code, iso, gr100, flag
  END TYPE a12_mol2_rec

  ! diatomic_record_t (src/mod_mklinelist.f90:71-78), 48 bytes: the
  ! reference's own type carries the same fields in this order
  TYPE, BIND(C) :: a12_diatomic_rec
    REAL(C_DOUBLE)     :: wlvac_nm, gflog_dex, elo_cm, gammar_log, gammaw_log
    INTEGER(C_INT32_T) :: molcode, pad
  END TYPE a12_diatomic_rec

  ! h2opokazatel.bin record (src/atlas12_modules.f90:19659-19660)
  TYPE, BIND(C) :: a12_h2o_rec
    INTEGER(C_INT32_T) :: iwl
    INTEGER(C_INT16_T) :: ielo, igflog
  END TYPE a12_h2o_rec

  ! SELECTLINES' seven lists (C_LOC of each array, or C_NULL_PTR with 0)
  TYPE, BIND(C) :: a12_sel_lists
    TYPE(C_PTR)        :: pred
    INTEGER(C_INT64_T) :: npred
    TYPE(C_PTR)        :: lowobs
    INTEGER(C_INT64_T) :: nlowobs
    TYPE(C_PTR)        :: hilines
    INTEGER(C_INT64_T) :: nhilines
    TYPE(C_PTR)        :: dia
    INTEGER(C_INT64_T) :: ndia
    TYPE(C_PTR)        :: mol2
    INTEGER(C_INT64_T) :: nmol2
    TYPE(C_PTR)        :: h2o
    INTEGER(C_INT64_T) :: nh2o
    TYPE(C_PTR)        :: h3p
    INTEGER(C_INT64_T) :: nh3p
  END TYPE a12_sel_lists

  ! State of SYNTHE's two hot loops (src/synthe.f90:476-718)
  TYPE, BIND(C) :: a12_synthe_atm
    INTEGER(C_INT32_T) :: nrhox, length
    REAL(C_DOUBLE)     :: wbegin, ratio, resolu
    INTEGER(C_INT32_T) :: nedge
    TYPE(C_PTR)        :: wledge
    INTEGER(C_INT32_T) :: nf
    TYPE(C_PTR)        :: continall, contabs, contscat
    INTEGER(C_INT32_T) :: mw6
    TYPE(C_PTR)        :: xnfpel, dopple
    TYPE(C_PTR)        :: t, xne, rho, xnfh, xnfhe1, xnfh2, hckt, hkt
    TYPE(C_PTR)        :: hkt_atlas, rhox
  END TYPE a12_synthe_atm

  ! State LINOP1 and the RUN_RT_PASS frequency loop read
  TYPE, BIND(C) :: a12_atlas_atm
    INTEGER(C_INT32_T) :: nrhox, numnu, mion, ldion, ldtab, nwave1
    TYPE(C_PTR)        :: waveset, rcoset, tabcont, iwavetab, xnfdop, dopple
    TYPE(C_PTR)        :: t, xne, hckt, hkt, rhox
    TYPE(C_PTR)        :: xnf1, xnf3, xnf841
    TYPE(C_PTR)        :: h0tab, h1tab, h2tab
    REAL(C_DOUBLE)     :: teff
  END TYPE a12_atlas_atm

  ! What TCORR(3) sections A-F read (src/atlas12_modules.f90:2528-2752);
  ! flxrad, rjmins, rdiagj C_NULL_PTR: the device sums of the last RT pass
  TYPE, BIND(C) :: a12_tcorr_in
    INTEGER(C_INT32_T) :: nrhox, ifconv, iter
    REAL(C_DOUBLE)     :: flux, teff, mixlth
    TYPE(C_PTR)        :: tauros, rhox, t, dltdlp, abross, flxcnv, flxcnv0, grdadb, ptotal
    TYPE(C_PTR)        :: rho, dlrdlt, heatcp, hscale
    TYPE(C_PTR)        :: flxrad, rjmins, rdiagj
  END TYPE a12_tcorr_in

  ! CZ_CONSTRUCT's SAVE state and the run flags CZC_RUN_ENGAGED / _JC
  ! (src/atlas12_modules.f90:3054-3069); start: all 0, last_dtmax 1D30
  TYPE, BIND(C) :: a12_czc_state
    INTEGER(C_INT32_T) :: prev_engaged, jc_sticky, nbot_bad, run_engaged, run_jc
    REAL(C_DOUBLE)     :: last_dtmax, rel_err
  END TYPE a12_czc_state

  ! What CZ_CONSTRUCT and CZC_DELREQ_LAYER read (:3012-3358)
  TYPE, BIND(C) :: a12_czc_in
    INTEGER(C_INT32_T) :: nrhox, ifconv, iter, use_constructor, use4pt, irosstab
    REAL(C_DOUBLE)     :: flux, mixlth, grav
    TYPE(C_PTR)        :: flxcnv, flxrad, dltdlp, grdadb, flxerr, t, rhox, ptotal, rho
    TYPE(C_PTR)        :: heatcp, dlrdlt, p, vturb, abross, hscale
  END TYPE a12_czc_in

  ! TCORR(3) sections (G2)-(K) (:2816-2967): what they read, and the
  ! module arrays they update in place (bhyd: BHYD(kw, 6) with kw = nrhox)
  TYPE, BIND(C) :: a12_tcorr_fin_in
    INTEGER(C_INT32_T) :: nrhox, ifconv, iter, use_constructor, use4pt, irosstab
    INTEGER(C_INT32_T) :: j1smooth, j2smooth
    REAL(C_DOUBLE)     :: flux, mixlth, grav, tau1lg, steplg, pradk0, tfloor, wtjm1, wtj, wtjp1
    TYPE(C_PTR)        :: flxcnv, flxrad, dltdlp, grdadb, flxerr, ptotal, rho, heatcp
    TYPE(C_PTR)        :: dlrdlt, hscale
  END TYPE a12_tcorr_fin_in
  TYPE, BIND(C) :: a12_tcorr_fin_io
    TYPE(C_PTR)        :: tauros, rhox, t, p, xne, abross, prad, vturb, bmin, pturb
    TYPE(C_PTR)        :: accrad, bhyd
    TYPE(C_PTR)        :: tk, hkt, hckt, tkev, tlog, pradk
  END TYPE a12_tcorr_fin_io

  ! CONVEC (:10323-10953) on the four EOS evaluations it makes: edens,
  ! rho4 are (nrhox, 4) -- EDENS and RHO after COMPUTE_ONE_POP at
  ! T*1.001, T*0.999, P*1.001, P*0.999
  TYPE, BIND(C) :: a12_convec_in
    INTEGER(C_INT32_T) :: nrhox, ifconv, ifcorr, nconv, use4pt, smooth, irosstab
    REAL(C_DOUBLE)     :: grav, mixlth, flux, overwt
    TYPE(C_PTR)        :: ptotal, rho, pradk, tauros, rhox, abross, vturb, height, flxrad
    TYPE(C_PTR)        :: edens, rho4
  END TYPE a12_convec_in
  TYPE, BIND(C) :: a12_convec_io
    TYPE(C_PTR)        :: t, tk, hkt, hckt, tkev, tlog, p
    TYPE(C_PTR)        :: dltdlp, heatcp, velsnd, dlrdlt, grdadb, hscale, vconv, flxcnv
    TYPE(C_PTR)        :: flxcnv0, flxcnv1, abconv, deltat
  END TYPE a12_convec_io

  ! One model of the model iteration (a12_model_converge)
  TYPE, BIND(C) :: a12_model_par
    REAL(C_DOUBLE)     :: teff, logg, mh, vturb
  END TYPE a12_model_par

  ! READMOL's molecule table for NMOLEC's Newton solve (1-based arrays)
  TYPE, BIND(C) :: a12_moltab
    INTEGER(C_INT32_T) :: nequa, nummol
    TYPE(C_PTR)        :: locj, kcomps, idequa
  END TYPE a12_moltab

  ! KAPP's tables (a12_kapp_tables): the Stehle Stark tables of one series
  TYPE, BIND(C) :: a12_kapp_stark
    INTEGER(C_INT32_T) :: loaded, n_upper_min, n_upper_max, n_dens
    TYPE(C_PTR)        :: dens, temp, log_dalpha, max_dens_idx, k_alpha, profiles
  END TYPE a12_kapp_stark

  ! KAPP's data tables (C_LOC of the arrays the sources hold; see a12gpu.h)
  TYPE, BIND(C) :: a12_kapp_tables
    TYPE(C_PTR)        :: k_freqn, k_xn, k_xl, k_ekarzas, gff_logtab
    INTEGER(C_INT32_T) :: cia_npair, cia_ntmax, cia_nnu
    REAL(C_DOUBLE)     :: cia_nu0, cia_dnu
    TYPE(C_PTR)        :: cia_nt, cia_temp, cia_tab, cia_use, cross_ch, cross_oh
    INTEGER(C_INT32_T) :: hi_nedge
    TYPE(C_PTR)        :: hi_edge
    INTEGER(C_INT32_T) :: felo_nlev(2), felo_xnfp_idx(2)
    TYPE(C_PTR)        :: felo_nuth, felo_g, felo_e, felo_ifancy, felo_npts, felo_x, felo_logsig, &
                          felo_seaton
    INTEGER(C_INT32_T) :: tb_nspecies, tb_nlevels, tb_npoints
    TYPE(C_PTR)        :: tb_xnfp_idx, tb_nlev, tb_lowest_thr, tb_g, tb_np, tb_ebind, tb_eexc, &
                          tb_logemin, tb_logestride, tb_ephot, tb_sigma
    TYPE(a12_kapp_stark) :: stark(4)
  END TYPE a12_kapp_tables

  ! KAPP's depth state: C_LOC of T, TKEV, ... (kw), BHYD (kw,6), XNF/XNFP (kw,mion)
  TYPE, BIND(C) :: a12_kapp_state
    INTEGER(C_INT32_T) :: nrhox, ld, mion
    TYPE(C_PTR)        :: t, tkev, tlog, hkt, hckt, xne, rho, bmin, bhyd, equilh2, xnf, xnfp
  END TYPE a12_kapp_state

  INTERFACE
    INTEGER(C_INT32_T) FUNCTION a12_ctx_create(device, ctx) BIND(C, NAME='a12_ctx_create')
      IMPORT :: C_INT32_T, C_PTR
      INTEGER(C_INT32_T), VALUE :: device
      TYPE(C_PTR), INTENT(OUT)  :: ctx
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_ctx_destroy(ctx) BIND(C, NAME='a12_ctx_destroy')
      IMPORT :: C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    TYPE(C_PTR) FUNCTION a12_last_error(ctx) BIND(C, NAME='a12_last_error')
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_set_operators(ctx, coefj, coefh) &
        BIND(C, NAME='a12_set_operators')
      IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(IN) :: coefj(*), coefh(*)
    END FUNCTION

    ! ---- SYNTHE ---------------------------------------------------------
    INTEGER(C_INT32_T) FUNCTION a12_synthe_line_opacity(ctx, atm, lines, nlines, prec, &
        opacity_out, evals_out) BIND(C, NAME='a12_synthe_line_opacity')
      IMPORT :: C_INT32_T, C_INT64_T, C_PTR, a12_synthe_atm, a12_lte_line
      TYPE(C_PTR), VALUE :: ctx
      TYPE(a12_synthe_atm), INTENT(IN) :: atm
      TYPE(a12_lte_line), INTENT(IN) :: lines(*)
      INTEGER(C_INT32_T), VALUE :: nlines, prec
      TYPE(C_PTR), VALUE :: opacity_out
      INTEGER(C_INT64_T), INTENT(OUT) :: evals_out
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_synthe_load(ctx, atm, lines, nlines) &
        BIND(C, NAME='a12_synthe_load')
      IMPORT :: C_INT32_T, C_PTR, a12_synthe_atm, a12_lte_line
      TYPE(C_PTR), VALUE :: ctx
      TYPE(a12_synthe_atm), INTENT(IN) :: atm
      TYPE(a12_lte_line), INTENT(IN) :: lines(*)
      INTEGER(C_INT32_T), VALUE :: nlines
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_synthe_opacity_run(ctx, prec) &
        BIND(C, NAME='a12_synthe_opacity_run')
      IMPORT :: C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: prec
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_synthe_opacity_fetch(ctx, opacity_out) &
        BIND(C, NAME='a12_synthe_opacity_fetch')
      IMPORT :: C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: ctx, opacity_out
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_synthe_transfer_run(ctx) &
        BIND(C, NAME='a12_synthe_transfer_run')
      IMPORT :: C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_synthe_transfer_fetch(ctx, hnu_line, hnu_cont) &
        BIND(C, NAME='a12_synthe_transfer_fetch')
      IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(OUT) :: hnu_line(*), hnu_cont(*)
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_synthe_transfer(ctx, atm, opacity, prec, hnu_line, &
        hnu_cont) BIND(C, NAME='a12_synthe_transfer')
      IMPORT :: C_INT32_T, C_PTR, C_DOUBLE, a12_synthe_atm
      TYPE(C_PTR), VALUE :: ctx
      TYPE(a12_synthe_atm), INTENT(IN) :: atm
      TYPE(C_PTR), VALUE :: opacity
      INTEGER(C_INT32_T), VALUE :: prec
      REAL(C_DOUBLE), INTENT(OUT) :: hnu_line(*), hnu_cont(*)
    END FUNCTION

    ! ---- JOSH, batched ----------------------------------------------------
    INTEGER(C_INT32_T) FUNCTION a12_josh_batch(ctx, ifscat, ifsurf, nrhox, nbatch, rhox, &
        acont, aline, sigmac, sigmal, scont, sline, bnu, hnu1, knu1, taunu, snu, jnu, hnu, &
        jmins) BIND(C, NAME='a12_josh_batch')
      IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: ifscat, ifsurf, nrhox, nbatch
      REAL(C_DOUBLE), INTENT(IN) :: rhox(*), acont(*), aline(*), sigmac(*), sigmal(*), &
                                    scont(*), sline(*), bnu(*)
      REAL(C_DOUBLE), INTENT(OUT) :: hnu1(*), knu1(*), taunu(*), snu(*), jnu(*), hnu(*), &
                                     jmins(*)
    END FUNCTION

    INTEGER(C_INT32_T) FUNCTION a12_solvit_batched(ctx, n, nsys, a, b, x) &
        BIND(C, NAME='a12_solvit_batched')
      IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: n, nsys
      REAL(C_DOUBLE), INTENT(IN) :: a(*), b(*)
      REAL(C_DOUBLE), INTENT(OUT) :: x(*)
    END FUNCTION

    ! ---- ATLAS12 ----------------------------------------------------------
    INTEGER(C_INT32_T) FUNCTION a12_atlas_select_lines(ctx, atm, wavetab, pred, npred, mol2, &
        nmol2, linedata_out, cap, nkept) BIND(C, NAME='a12_atlas_select_lines')
      IMPORT :: C_INT32_T, C_INT64_T, C_PTR, C_DOUBLE, a12_atlas_atm, a12_mol2_rec
      TYPE(C_PTR), VALUE :: ctx
      TYPE(a12_atlas_atm), INTENT(IN) :: atm
      REAL(C_DOUBLE), INTENT(IN) :: wavetab(*)
      INTEGER(C_INT32_T), INTENT(IN) :: pred(*)
      INTEGER(C_INT64_T), VALUE :: npred
      TYPE(a12_mol2_rec), INTENT(IN) :: mol2(*)
      INTEGER(C_INT64_T), VALUE :: nmol2
      INTEGER(C_INT32_T), INTENT(OUT) :: linedata_out(*)
      INTEGER(C_INT64_T), VALUE :: cap
      INTEGER(C_INT64_T), INTENT(OUT) :: nkept
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_atlas_select_all(ctx, atm, wavetab, lists, linedata_out, &
        cap, nkept) BIND(C, NAME='a12_atlas_select_all')
      IMPORT :: C_INT32_T, C_INT64_T, C_PTR, C_DOUBLE, a12_atlas_atm, a12_sel_lists
      TYPE(C_PTR), VALUE :: ctx
      TYPE(a12_atlas_atm), INTENT(IN) :: atm
      REAL(C_DOUBLE), INTENT(IN) :: wavetab(*)
      TYPE(a12_sel_lists), INTENT(IN) :: lists
      INTEGER(C_INT32_T), INTENT(OUT) :: linedata_out(*)
      INTEGER(C_INT64_T), VALUE :: cap
      INTEGER(C_INT64_T), INTENT(OUT) :: nkept
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_atlas_load(ctx, atm, linedata, nlines) &
        BIND(C, NAME='a12_atlas_load')
      IMPORT :: C_INT32_T, C_PTR, a12_atlas_atm
      TYPE(C_PTR), VALUE :: ctx
      TYPE(a12_atlas_atm), INTENT(IN) :: atm
      INTEGER(C_INT32_T), INTENT(IN) :: linedata(*)
      INTEGER(C_INT32_T), VALUE :: nlines
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_atlas_load_lines(ctx, atm, linedata, nlines) &
        BIND(C, NAME='a12_atlas_load_lines')
      IMPORT :: C_INT32_T, C_PTR, a12_atlas_atm
      TYPE(C_PTR), VALUE :: ctx
      TYPE(a12_atlas_atm), INTENT(IN) :: atm
      INTEGER(C_INT32_T), INTENT(IN) :: linedata(*)
      INTEGER(C_INT32_T), VALUE :: nlines
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_atlas_set_state(ctx, atm) BIND(C, NAME='a12_atlas_set_state')
      IMPORT :: C_INT32_T, C_PTR, a12_atlas_atm
      TYPE(C_PTR), VALUE :: ctx
      TYPE(a12_atlas_atm), INTENT(IN) :: atm
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_atlas_linop1_run(ctx, prec) &
        BIND(C, NAME='a12_atlas_linop1_run')
      IMPORT :: C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: prec
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_atlas_linop1_fetch(ctx, xlines, evals) &
        BIND(C, NAME='a12_atlas_linop1_fetch')
      IMPORT :: C_INT32_T, C_INT64_T, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(OUT) :: xlines(*)
      INTEGER(C_INT64_T), INTENT(OUT) :: evals
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_atlas_set_kapp(ctx, acont, sigmac, scont, sline, aline, &
        sigmal) BIND(C, NAME='a12_atlas_set_kapp')
      IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(IN) :: acont(*), sigmac(*), scont(*), sline(*), aline(*), sigmal(*)
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_atlas_rt_run(ctx, xlines) BIND(C, NAME='a12_atlas_rt_run')
      IMPORT :: C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: ctx, xlines   ! C_NULL_PTR: use the device XLINES of linop1
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_atlas_rt_fetch(ctx, acc, hnu1) &
        BIND(C, NAME='a12_atlas_rt_fetch')
      IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(OUT) :: acc(*), hnu1(*)
    END FUNCTION

    ! device arrays the caller owns (C_PTR from its own hipMalloc)
    INTEGER(C_INT32_T) FUNCTION a12_atlas_rt_pass_noalloc(ctx, nrhox, numnu, teff, d_waveset, &
        d_rcoset, d_depth3, d_kapp, d_xlines, d_hnu1, d_acc) &
        BIND(C, NAME='a12_atlas_rt_pass_noalloc')
      IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: nrhox, numnu
      REAL(C_DOUBLE), VALUE :: teff
      TYPE(C_PTR), VALUE :: d_waveset, d_rcoset, d_depth3, d_kapp, d_xlines, d_hnu1, d_acc
    END FUNCTION

    ! ---- model iteration ---------------------------------------------------
    INTEGER(C_INT32_T) FUNCTION a12_model_converge(ctx, nmodel, par, t_start, itmax, t_out, &
        s_out, iters, hist) BIND(C, NAME='a12_model_converge')
      IMPORT :: C_INT32_T, C_PTR, C_DOUBLE, a12_model_par
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: nmodel, itmax
      TYPE(a12_model_par), INTENT(IN) :: par(*)
      REAL(C_DOUBLE), INTENT(IN) :: t_start(*)
      REAL(C_DOUBLE), INTENT(OUT) :: t_out(*), s_out(*)
      INTEGER(C_INT32_T), INTENT(OUT) :: iters(*)
      TYPE(C_PTR), VALUE :: hist             ! C_NULL_PTR or C_LOC of (3, itmax, nmodel)
    END FUNCTION

    ! ---- KAPP ---------------------------------------------------------------
    INTEGER(C_INT32_T) FUNCTION a12_kapp_set_tables(ctx, tables) BIND(C, NAME='a12_kapp_set_tables')
      IMPORT :: C_INT32_T, C_PTR, a12_kapp_tables
      TYPE(C_PTR), VALUE :: ctx
      TYPE(a12_kapp_tables), INTENT(IN) :: tables
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_kapp_set_state(ctx, state) BIND(C, NAME='a12_kapp_set_state')
      IMPORT :: C_INT32_T, C_PTR, a12_kapp_state
      TYPE(C_PTR), VALUE :: ctx
      TYPE(a12_kapp_state), INTENT(IN) :: state
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_atlas_kapp_run(ctx) BIND(C, NAME='a12_atlas_kapp_run')
      IMPORT :: C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_atlas_kapp_fetch(ctx, kapp6) BIND(C, NAME='a12_atlas_kapp_fetch')
      IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(OUT) :: kapp6(*)        ! (nrhox, numnu, 6)
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_kapp_eval(ctx, nwave, wave_nm, out) BIND(C, NAME='a12_kapp_eval')
      IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: nwave
      REAL(C_DOUBLE), INTENT(IN) :: wave_nm(*)
      REAL(C_DOUBLE), INTENT(OUT) :: out(*)           ! (nrhox, 23, nwave)
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_kapp_constants(out12) BIND(C, NAME='a12_kapp_constants')
      IMPORT :: C_INT32_T, C_DOUBLE
      REAL(C_DOUBLE), INTENT(OUT) :: out12(12)
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_kapp_ms(ctx, ms) BIND(C, NAME='a12_kapp_ms')
      IMPORT :: C_INT32_T, C_PTR, C_FLOAT
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_FLOAT), INTENT(OUT) :: ms
    END FUNCTION

    ! ---- TCORR(3), SCALE_MODEL, NMOLEC -----------------------------------
    INTEGER(C_INT32_T) FUNCTION a12_atlas_tcorr3(ctx, tin, rdabh, oldt1, out) &
        BIND(C, NAME='a12_atlas_tcorr3')
      IMPORT :: C_INT32_T, C_PTR, C_DOUBLE, a12_tcorr_in
      TYPE(C_PTR), VALUE :: ctx
      TYPE(a12_tcorr_in), INTENT(IN) :: tin
      REAL(C_DOUBLE), INTENT(INOUT) :: rdabh(*), oldt1(*)
      REAL(C_DOUBLE), INTENT(OUT) :: out(*)
    END FUNCTION
    ! ---- ROSSTAB, TTAUP, CONVEC, TCORR(3) (G2)-(K) --------------------------
    INTEGER(C_INT32_T) FUNCTION a12_rosstab_reset(ctx) BIND(C, NAME='a12_rosstab_reset')
      IMPORT :: C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_rosstab_store(ctx, nrhox, t, p, abross) &
        BIND(C, NAME='a12_rosstab_store')
      IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: nrhox
      REAL(C_DOUBLE), INTENT(IN) :: t(*), p(*), abross(*)
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_rosstab_load(ctx, nross, zero, tab) &
        BIND(C, NAME='a12_rosstab_load')
      IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: nross
      REAL(C_DOUBLE), INTENT(IN) :: zero(4), tab(*)
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_rosstab_fetch(ctx, nross, zero, tab) &
        BIND(C, NAME='a12_rosstab_fetch')
      IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), INTENT(OUT) :: nross
      REAL(C_DOUBLE), INTENT(OUT) :: zero(4), tab(*)
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_rosstab_lookup(ctx, mode, nq, temp, pressure, kappa) &
        BIND(C, NAME='a12_rosstab_lookup')
      IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: mode, nq
      REAL(C_DOUBLE), INTENT(IN) :: temp(*), pressure(*)
      REAL(C_DOUBLE), INTENT(OUT) :: kappa(*)
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_ttaup(ctx, irosstab, n, t, tau, prad, pturb, grav, abstd, &
        ptotal, p, nfail) BIND(C, NAME='a12_ttaup')
      IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: irosstab, n
      REAL(C_DOUBLE), INTENT(IN) :: t(*), tau(*), prad(*), pturb(*)
      REAL(C_DOUBLE), VALUE :: grav
      REAL(C_DOUBLE), INTENT(OUT) :: abstd(*), ptotal(*), p(*)
      INTEGER(C_INT32_T), INTENT(OUT) :: nfail
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_cz_construct(ctx, cin, st, t1, engaged) &
        BIND(C, NAME='a12_cz_construct')
      IMPORT :: C_INT32_T, C_PTR, C_DOUBLE, a12_czc_in, a12_czc_state
      TYPE(C_PTR), VALUE :: ctx
      TYPE(a12_czc_in), INTENT(IN) :: cin
      TYPE(a12_czc_state), INTENT(INOUT) :: st
      REAL(C_DOUBLE), INTENT(INOUT) :: t1(*)
      INTEGER(C_INT32_T), INTENT(OUT) :: engaged
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_atlas_tcorr3_finish(ctx, fin, fio, st, t1, drhox, nfail) &
        BIND(C, NAME='a12_atlas_tcorr3_finish')
      IMPORT :: C_INT32_T, C_PTR, C_DOUBLE, a12_tcorr_fin_in, a12_tcorr_fin_io, a12_czc_state
      TYPE(C_PTR), VALUE :: ctx
      TYPE(a12_tcorr_fin_in), INTENT(IN) :: fin
      TYPE(a12_tcorr_fin_io), INTENT(INOUT) :: fio
      TYPE(a12_czc_state), INTENT(INOUT) :: st
      REAL(C_DOUBLE), INTENT(INOUT) :: t1(*)
      REAL(C_DOUBLE), INTENT(OUT) :: drhox(*)
      INTEGER(C_INT32_T), INTENT(OUT) :: nfail
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_convec(ctx, cin, cio, mlt_only) BIND(C, NAME='a12_convec')
      IMPORT :: C_INT32_T, C_PTR, a12_convec_in, a12_convec_io
      TYPE(C_PTR), VALUE :: ctx
      TYPE(a12_convec_in), INTENT(IN) :: cin
      TYPE(a12_convec_io), INTENT(INOUT) :: cio
      INTEGER(C_INT32_T), VALUE :: mlt_only
    END FUNCTION

    INTEGER(C_INT32_T) FUNCTION a12_scale_model(ctx, nmodel, nrhox, kw, tau1lg, steplg, sc, &
        fields, out, so, taustd) BIND(C, NAME='a12_scale_model')
      IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: nmodel, nrhox, kw
      REAL(C_DOUBLE), VALUE :: tau1lg, steplg
      REAL(C_DOUBLE), INTENT(IN) :: sc(*), fields(*)
      REAL(C_DOUBLE), INTENT(OUT) :: out(*), so(*), taustd(*)
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_molec_set_table(ctx, tab) BIND(C, NAME='a12_molec_set_table')
      IMPORT :: C_INT32_T, C_PTR, a12_moltab
      TYPE(C_PTR), VALUE :: ctx
      TYPE(a12_moltab), INTENT(IN) :: tab
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_molec_newton_batched(ctx, nsys, nchain, chain_off, xntot, &
        pratio, xab, equilj, xn, xnmol, iters, status) BIND(C, NAME='a12_molec_newton_batched')
      IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: nsys, nchain
      INTEGER(C_INT32_T), INTENT(IN) :: chain_off(*)
      REAL(C_DOUBLE), INTENT(IN) :: xntot(*), pratio(*), xab(*), equilj(*)
      REAL(C_DOUBLE), INTENT(INOUT) :: xn(*)
      REAL(C_DOUBLE), INTENT(OUT) :: xnmol(*)
      INTEGER(C_INT32_T), INTENT(OUT) :: iters(*), status(*)
    END FUNCTION

    ! ---- batched SYNTHE and gather shapes ---------------------------------
    INTEGER(C_INT32_T) FUNCTION a12_synthe_set_shape(ctx, shape) &
        BIND(C, NAME='a12_synthe_set_shape')
      IMPORT :: C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: shape
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_synthe_batch_load(ctx, nmodel, atms, nb_shift, line_lo, &
        line_hi, lines, nlines) BIND(C, NAME='a12_synthe_batch_load')
      IMPORT :: C_INT32_T, C_INT64_T, C_PTR, a12_synthe_atm, a12_lte_line
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: nmodel
      TYPE(a12_synthe_atm), INTENT(IN) :: atms(*)
      INTEGER(C_INT32_T), INTENT(IN) :: nb_shift(*)
      INTEGER(C_INT64_T), INTENT(IN) :: line_lo(*), line_hi(*)
      TYPE(a12_lte_line), INTENT(IN) :: lines(*)
      INTEGER(C_INT64_T), VALUE :: nlines
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_synthe_batch_run(ctx, prec, transfer, use_graph) &
        BIND(C, NAME='a12_synthe_batch_run')
      IMPORT :: C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: prec, transfer, use_graph
    END FUNCTION
    INTEGER(C_INT32_T) FUNCTION a12_synthe_batch_fetch(ctx, model, opacity_out, hnu_line, &
        hnu_cont, evals) BIND(C, NAME='a12_synthe_batch_fetch')
      IMPORT :: C_INT32_T, C_INT64_T, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx, opacity_out
      INTEGER(C_INT32_T), VALUE :: model
      REAL(C_DOUBLE), INTENT(OUT) :: hnu_line(*), hnu_cont(*)
      INTEGER(C_INT64_T), INTENT(OUT) :: evals(2)
    END FUNCTION
  END INTERFACE

CONTAINS

  ! The context's last error message as a Fortran string.
  FUNCTION a12_errmsg(ctx) RESULT(msg)
    TYPE(C_PTR), INTENT(IN) :: ctx
    CHARACTER(LEN=:), ALLOCATABLE :: msg
    CHARACTER(KIND=C_CHAR), POINTER :: s(:)
    TYPE(C_PTR) :: p
    INTEGER :: n
    INTERFACE
      INTEGER(C_SIZE_T) FUNCTION c_strlen(str) BIND(C, NAME='strlen')
        IMPORT :: C_SIZE_T, C_PTR
        TYPE(C_PTR), VALUE :: str
      END FUNCTION
    END INTERFACE
    p = a12_last_error(ctx)
    IF (.NOT. C_ASSOCIATED(p)) THEN
      msg = ''
      RETURN
    END IF
    n = INT(c_strlen(p))
    CALL C_F_POINTER(p, s, [n])
    ALLOCATE(CHARACTER(LEN=n) :: msg)
    msg = TRANSFER(s(1:n), msg)
  END FUNCTION a12_errmsg

  ! Fail-fast check in the drivers' style: print and stop on an error.
  SUBROUTINE a12_check(ctx, rc, what)
    TYPE(C_PTR), INTENT(IN) :: ctx
    INTEGER(C_INT32_T), INTENT(IN) :: rc
    CHARACTER(LEN=*), INTENT(IN) :: what
    IF (rc .NE. A12_OK) THEN
      WRITE(6, '(A,A,A,I3,A,A)') ' ', what, ' failed (', rc, '): ', a12_errmsg(ctx)
      CALL EXIT(1)
    END IF
  END SUBROUTINE a12_check

END MODULE a12gpu_iface
