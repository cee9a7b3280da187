"""The C-ABI library loads and exports every entry point include/a12gpu.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(REPO, "include", "a12gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(a12_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("a12_ctx_create", "a12_synthe_line_opacity", "a12_synthe_transfer", "a12_josh_batch"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from atlas12_amd import native
    lib = native.load_library()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_struct_layouts_match_header(tmp_path):
    """Sizes and field offsets of the ctypes mirrors equal the C compiler's
    layout of the header's structs."""
    import subprocess
    from atlas12_amd import native
    from atlas12_amd.workloads import LTE_DTYPE
    fields = [f[0] for f in native.SyntheAtm._fields_]
    src = tmp_path / "layout.c"
    body = "".join(f'printf("%zu\\n", offsetof(a12_synthe_atm, {f}));' for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "a12gpu.h"\n'
                   'int main(void){printf("%zu\\n%zu\\n", sizeof(a12_lte_line), sizeof(a12_synthe_atm));'
                   + body + "return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert vals[0] == ctypes.sizeof(native.LteLine) == LTE_DTYPE.itemsize == 32
    assert vals[1] == ctypes.sizeof(native.SyntheAtm)
    assert vals[2:] == [getattr(native.SyntheAtm, f).offset for f in fields]


def test_missing_library_fails_loudly(tmp_path):
    from atlas12_amd import native
    import pytest
    saved = native._lib
    native._lib = None
    try:
        with pytest.raises(native.A12Error):
            native.load_library(str(tmp_path / "nope.so"))
    finally:
        native._lib = saved


def test_struct_layouts_atlas(tmp_path):
    """a12_atlas_atm and a12_mol2_rec: the ctypes mirrors and MOL2_DTYPE
    match the C layout."""
    import subprocess
    from atlas12_amd import native
    from atlas12_amd.workloads import MOL2_DTYPE
    fields = [f[0] for f in native.AtlasAtm._fields_]
    src = tmp_path / "layout2.c"
    body = "".join(f'printf("%zu\\n", offsetof(a12_atlas_atm, {f}));' for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "a12gpu.h"\n'
                   'int main(void){printf("%zu\\n%zu\\n", sizeof(a12_mol2_rec), sizeof(a12_atlas_atm));'
                   + body + "return 0;}\n")
    exe = tmp_path / "layout2"
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert vals[0] == MOL2_DTYPE.itemsize == 32
    assert vals[1] == ctypes.sizeof(native.AtlasAtm)
    assert vals[2:] == [getattr(native.AtlasAtm, f).offset for f in fields]


def test_struct_layouts_select(tmp_path):
    """a12_diatomic_rec, a12_h2o_rec and a12_sel_lists: the numpy record
    types and the ctypes mirror match the C layout."""
    import subprocess
    from atlas12_amd import native
    from atlas12_amd.records import DIATOMIC_DTYPE
    from atlas12_amd.workloads import H2O_DTYPE
    fields = [f[0] for f in native.SelLists._fields_]
    dfields = [f for f in DIATOMIC_DTYPE.names]
    src = tmp_path / "layout3.c"
    body = "".join(f'printf("%zu\\n", offsetof(a12_sel_lists, {f}));' for f in fields)
    body += "".join(f'printf("%zu\\n", offsetof(a12_diatomic_rec, {f}));' for f in dfields)
    body += "".join(f'printf("%zu\\n", offsetof(a12_h2o_rec, {f}));' for f in H2O_DTYPE.names)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "a12gpu.h"\n'
                   'int main(void){printf("%zu\\n%zu\\n%zu\\n", sizeof(a12_diatomic_rec), '
                   'sizeof(a12_h2o_rec), sizeof(a12_sel_lists));' + body + "return 0;}\n")
    exe = tmp_path / "layout3"
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert vals[:3] == [DIATOMIC_DTYPE.itemsize, H2O_DTYPE.itemsize, ctypes.sizeof(native.SelLists)]
    k = 3 + len(fields)
    assert vals[3:k] == [getattr(native.SelLists, f).offset for f in fields]
    assert vals[k:k + len(dfields)] == [DIATOMIC_DTYPE.fields[f][1] for f in dfields]
    assert vals[k + len(dfields):] == [H2O_DTYPE.fields[f][1] for f in H2O_DTYPE.names]


def test_fortran_interface_module(tmp_path):
    """atlas12_amd/fortran/a12gpu_iface.f90 compiles, its BIND(C) types have
    the C sizes, and calls through it reach the library (a null context is
    rejected with A12_ERR_ARG before any device work)."""
    import shutil
    import subprocess

    import pytest
    fc = shutil.which("amdflang") or shutil.which("gfortran")
    if fc is None:
        pytest.skip("no Fortran compiler")
    from atlas12_amd import native
    lib = native.load_library()
    prog = tmp_path / "t.f90"
    prog.write_text("""
PROGRAM t
  USE a12gpu_iface
  IMPLICIT NONE
  TYPE(a12_lte_line) :: l(1)
  TYPE(a12_synthe_atm) :: sa
  TYPE(a12_atlas_atm) :: aa
  TYPE(a12_mol2_rec) :: m
  TYPE(a12_tcorr_in) :: tc
  TYPE(a12_moltab) :: mt
  TYPE(a12_czc_state) :: cs
  TYPE(a12_czc_in) :: ci
  TYPE(a12_tcorr_fin_in) :: fi
  TYPE(a12_tcorr_fin_io) :: fo
  TYPE(a12_convec_in) :: vi
  TYPE(a12_convec_io) :: vo
  TYPE(a12_diatomic_rec) :: dr
  TYPE(a12_h2o_rec) :: hr
  TYPE(a12_sel_lists) :: sl
  INTEGER(C_INT32_T) :: rc, ld(4)
  INTEGER(C_INT64_T) :: nk
  PRINT *, C_SIZEOF(l(1)), C_SIZEOF(sa), C_SIZEOF(aa), C_SIZEOF(m), C_SIZEOF(tc), C_SIZEOF(mt)
  PRINT *, C_SIZEOF(cs), C_SIZEOF(ci), C_SIZEOF(fi), C_SIZEOF(fo), C_SIZEOF(vi), C_SIZEOF(vo)
  PRINT *, C_SIZEOF(dr), C_SIZEOF(hr), C_SIZEOF(sl)
  rc = a12_synthe_opacity_run(C_NULL_PTR, A12_PREC_MIXED)
  PRINT *, rc
  rc = a12_atlas_linop1_run(C_NULL_PTR, A12_PREC_FP64)
  PRINT *, rc
  rc = a12_convec(C_NULL_PTR, vi, vo, 0)
  PRINT *, rc
  rc = a12_atlas_select_all(C_NULL_PTR, aa, [1.0D0], sl, ld, 0_C_INT64_T, nk)
  PRINT *, rc
END PROGRAM t
""")
    iface = os.path.join(REPO, "atlas12_amd", "fortran", "a12gpu_iface.f90")
    exe = tmp_path / "t"
    mod = ["-module-dir", str(tmp_path)] if "flang" in fc else ["-J", str(tmp_path)]
    subprocess.run([fc, *mod, iface, str(prog), "-o", str(exe), lib._name,
                    "-Wl,-rpath," + os.path.dirname(lib._name)], check=True, cwd=tmp_path)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    sizes = [int(x) for x in out[:15]]
    assert sizes == [32, ctypes.sizeof(native.SyntheAtm), ctypes.sizeof(native.AtlasAtm), 32,
                     ctypes.sizeof(native.TcorrIn), ctypes.sizeof(native.MolTab),
                     ctypes.sizeof(native.CzcState), ctypes.sizeof(native.CzcIn),
                     ctypes.sizeof(native.TcorrFinIn), ctypes.sizeof(native.TcorrFinIo),
                     ctypes.sizeof(native.ConvecIn), ctypes.sizeof(native.ConvecIo),
                     48, 8, ctypes.sizeof(native.SelLists)]
    assert [int(x) for x in out[15:19]] == [-1, -1, -1, -1]


def test_atlas_wrappers_reject_wrong_shapes():
    """KAPP tables and XLINES of another state's size are refused on the
    host before the library is called (the C side copies nrhox * numnu
    doubles from each pointer and cannot see a short array)."""
    import numpy as np
    import pytest
    from atlas12_amd import native

    class Ctx:  # a context whose state is 72 layers x 240 frequencies; no GPU needed
        _atl = (72, 240)
        lib = None
        h = None

    good = np.zeros((240, 72))
    bad = np.zeros((200, 72))
    with pytest.raises(ValueError, match="sigmac"):
        native.atlas_set_kapp(Ctx(), good, bad, good, good, good, good)
    with pytest.raises(ValueError, match="xlines"):
        native.atlas_rt_pass(Ctx(), bad)
    with pytest.raises(native.A12Error, match="no ATLAS12 state"):
        native.atlas_set_kapp(type("C0", (), {"lib": None, "h": None})(), good, good, good,
                              good, good, good)
