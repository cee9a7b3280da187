"""A Fortran driver calling the library through atlas12_amd/fortran/
a12gpu_iface.f90 with real data: the RUN_RT_PASS call sequence of the
driver (a12_atlas_load, a12_atlas_set_kapp, a12_atlas_rt_run on the
reference's XLINES, a12_atlas_rt_fetch) on the golden 72-layer, 240-frequency
case, with every array passed in its Fortran layout through iso_c_binding.
The accumulators and HNU(1) it writes equal those of the same calls made
through ctypes bit for bit.  Skipped, with the reason, where no Fortran
compiler is installed."""
import os
import shutil
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = """
PROGRAM rtpass
  USE a12gpu_iface
  IMPLICIT NONE
  INTEGER, PARAMETER :: N = 72, NF = 240, NW = 362, MION = 1006, NL = 3000
  REAL(C_DOUBLE), TARGET :: waveset(NF), rcoset(NF), tabcont(N, NW), xnfdop(N, MION), &
       dopple(N, MION), t(N), xne(N), hckt(N), hkt(N), rhox(N), xnf1(N), xnf3(N), xnf841(N), &
       h0(2001), h1(2001), h2(2001), coefj(51, 51), coefh(51, 51), teff
  REAL(C_DOUBLE), TARGET :: kapp(N, NF, 6), xlines(N, NF)
  REAL(C_DOUBLE) :: acc(N, 9), hnu1(NF)
  INTEGER(C_INT32_T), TARGET :: iwavetab(NW), linedata(4, NL)
  TYPE(a12_atlas_atm) :: aa
  TYPE(C_PTR) :: ctx
  INTEGER :: u
  OPEN(NEWUNIT=u, FILE='in.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  READ(u) waveset, rcoset, tabcont, xnfdop, dopple, t, xne, hckt, hkt, rhox, xnf1, xnf3, &
          xnf841, h0, h1, h2, coefj, coefh, teff, kapp, xlines, iwavetab, linedata
  CLOSE(u)
  CALL a12_check(ctx, a12_ctx_create(0_C_INT32_T, ctx), 'a12_ctx_create')
  CALL a12_check(ctx, a12_set_operators(ctx, coefj, coefh), 'a12_set_operators')
  aa%nrhox = N;  aa%numnu = NF;  aa%mion = MION;  aa%ldion = N;  aa%ldtab = N;  aa%nwave1 = NW
  aa%waveset = C_LOC(waveset);  aa%rcoset = C_LOC(rcoset)
  aa%tabcont = C_LOC(tabcont);  aa%iwavetab = C_LOC(iwavetab)
  aa%xnfdop = C_LOC(xnfdop);  aa%dopple = C_LOC(dopple)
  aa%t = C_LOC(t);  aa%xne = C_LOC(xne);  aa%hckt = C_LOC(hckt);  aa%hkt = C_LOC(hkt)
  aa%rhox = C_LOC(rhox)
  aa%xnf1 = C_LOC(xnf1);  aa%xnf3 = C_LOC(xnf3);  aa%xnf841 = C_LOC(xnf841)
  aa%h0tab = C_LOC(h0);  aa%h1tab = C_LOC(h1);  aa%h2tab = C_LOC(h2)
  aa%teff = teff
  CALL a12_check(ctx, a12_atlas_load(ctx, aa, linedata, INT(NL, C_INT32_T)), 'a12_atlas_load')
  CALL a12_check(ctx, a12_atlas_set_kapp(ctx, kapp(:, :, 1), kapp(:, :, 2), kapp(:, :, 3), &
                 kapp(:, :, 4), kapp(:, :, 5), kapp(:, :, 6)), 'a12_atlas_set_kapp')
  CALL a12_check(ctx, a12_atlas_rt_run(ctx, C_LOC(xlines)), 'a12_atlas_rt_run')
  CALL a12_check(ctx, a12_atlas_rt_fetch(ctx, acc, hnu1), 'a12_atlas_rt_fetch')
  CALL a12_check(ctx, a12_ctx_destroy(ctx), 'a12_ctx_destroy')
  OPEN(NEWUNIT=u, FILE='out.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(u) acc, hnu1
  CLOSE(u)
END PROGRAM rtpass
"""


def test_fortran_driver_rt_pass(tmp_path, golden, coef):
    fc = shutil.which("amdflang") or shutil.which("gfortran")
    if fc is None:
        pytest.skip("no Fortran compiler (amdflang or gfortran) on this machine")
    from atlas12_amd import native
    lib = native.load_library()
    z = golden("atlas_small.npz")
    d = native.atlas_state_from(z)
    cj, ch = coef
    kap = np.stack([np.asarray(z[k], np.float64) for k in
                    ("acont", "sigmac", "scont", "sline", "aline", "sigmal")])  # (6, NF, N)
    assert kap.shape == (6, 240, 72) and d["xnfdop"].shape == (1006, 72)
    ld = np.ascontiguousarray(z["linedata"], np.int32)
    assert ld.shape == (3000, 4)
    parts = [d["waveset"], d["rcoset"], d["tabcont"], d["xnfdop"], d["dopple"]] + \
            [d[k] for k in ("t", "xne", "hckt", "hkt", "rhox", "xnf1", "xnf3", "xnf841",
                            "h0", "h1", "h2")] + [cj, ch, np.array([d["teff"]]), kap,
                                                  np.asarray(z["xlines"], np.float64)]
    with open(tmp_path / "in.bin", "wb") as fh:
        for a in parts:
            fh.write(np.ascontiguousarray(a, np.float64).tobytes())
        fh.write(np.ascontiguousarray(d["iwavetab"], np.int32).tobytes())
        fh.write(ld.tobytes())
    (tmp_path / "rtpass.f90").write_text(PROGRAM)
    iface = os.path.join(REPO, "atlas12_amd", "fortran", "a12gpu_iface.f90")
    mod = ["-module-dir", str(tmp_path)] if "flang" in fc else ["-J", str(tmp_path)]
    subprocess.run([fc, *mod, iface, "rtpass.f90", "-o", "rtpass", lib._name,
                    "-Wl,-rpath," + os.path.dirname(lib._name)], check=True, cwd=tmp_path)
    subprocess.run([str(tmp_path / "rtpass")], check=True, cwd=tmp_path, timeout=120)
    out = np.fromfile(tmp_path / "out.bin", np.float64)
    acc_f, h1_f = out[:9 * 72], out[9 * 72:]
    # the same call sequence through ctypes in this process
    gpu_ctx = native.Context(0)
    try:
        gpu_ctx.set_operators(cj, ch)
        native.atlas_load(gpu_ctx, d, ld)
        native.atlas_set_kapp(gpu_ctx, *kap)
        acc, h1 = native.atlas_rt_pass(gpu_ctx, z["xlines"])
    finally:
        gpu_ctx.close()
    ref = np.concatenate([acc[k] for k in native.ACC_NAMES] + [[acc["pradk0"]], np.zeros(71)])
    assert np.array_equal(h1_f, h1)
    assert np.array_equal(acc_f, ref)
    assert np.all(np.isfinite(h1_f)) and np.all(h1_f > 0)
