"""KAPP's host-side pieces, without a GPU: the ctypes mirrors of the
a12_kapp_* structs, the physical constants the library uses against the
values the reference computes (bitwise), the table checks, and the KAPP
source (a12_kapp_core.h) compiled for the host against the reference's own
KAPP (src/atlas12_modules.f90:11202-11324) on every per-source array of
tests/golden/kapp_ref.npz: 3 decks x 349 wavelengths x 25 depths, bit for
bit (same expressions, same evaluation order, host libm like the
reference's)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
DECKS = ("plain", "hot", "cool")


def _layout(tmp_path, cname, fields, mirror, name):
    src = tmp_path / f"{name}.c"
    body = "".join(f'printf("%zu\\n", offsetof({cname}, {f}));' for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "a12gpu.h"\n'
                   f'int main(void){{printf("%zu\\n", sizeof({cname}));' + body + "return 0;}\n")
    exe = tmp_path / name
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert vals[0] == ctypes.sizeof(mirror)
    assert vals[1:] == [getattr(mirror, f).offset for f in fields]


def test_struct_layouts_match_header(tmp_path):
    from atlas12_amd import native
    for cname, mirror in (("a12_kapp_tables", native.KappTables), ("a12_kapp_stark", native.KappStark),
                          ("a12_kapp_state", native.KappState)):
        _layout(tmp_path, cname, [f[0] for f in mirror._fields_], mirror, cname)


def test_constants_equal_the_reference():
    """HCK, BNU_PREFAC, SAHA_PREFAC (a libm power), ... as the library forms
    them, against the values the compiled reference holds."""
    from atlas12_amd import native
    t = np.load(os.path.join(GOLDEN, "kapp_tables.npz"))
    c = native.kapp_constants()
    for k in native.KAPP_CONST:
        assert c[k] == float(t["const_" + k]), k


@pytest.fixture(scope="module")
def host_lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("kapp_host")
    so = d / "libkapp_host.so"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fPIC", "-shared",
                    "-I", os.path.join(REPO, "atlas12_amd", "csrc"), "-I", os.path.join(REPO, "include"),
                    os.path.join(REPO, "tests", "kapp_host", "kapp_host_eval.cpp"), "-o", str(so)],
                   check=True)
    L = ctypes.CDLL(str(so))
    L.kcpu_check_tables.restype = ctypes.c_char_p
    return L


@pytest.fixture(scope="module")
def tables():
    from atlas12_amd import native
    t = np.load(os.path.join(GOLDEN, "kapp_tables.npz"))
    return t, native.kapp_tables(t)


def deck_state(ref, deck, mion=1006):
    """The deck's KAPP inputs with XNF/XNFP rebuilt from the stored columns
    (the species the sources read; the rest zero)."""
    from atlas12_amd.workloads import kapp_deck_state
    return kapp_deck_state(ref, deck, mion)


@pytest.mark.parametrize("deck", DECKS)
def test_host_build_reproduces_reference_bitwise(host_lib, tables, deck):
    from atlas12_amd import native
    ref = np.load(os.path.join(GOLDEN, "kapp_ref.npz"))
    waves, depths = ref["waves"], ref["depths"]
    st = native.kapp_state(deck_state(ref, deck))
    n = st.struct.nrhox
    out = np.zeros((len(waves), 23, n))
    rc = host_lib.kcpu_eval(ctypes.byref(tables[1].struct), ctypes.byref(st.struct), len(waves),
                            waves.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    for i, k in enumerate(native.KAPP_SRC):
        r = ref[f"{deck}_{k}"]
        assert np.array_equal(out[:, i][:, depths], r), k
        if k not in ("sigmal",):
            assert np.count_nonzero(r) > 0, k


def test_table_checks(host_lib, tables):
    """a12_kapp_set_tables' consistency checks."""
    from atlas12_amd import native
    t = dict(tables[0])
    assert host_lib.kcpu_check_tables(ctypes.byref(tables[1].struct)) == b""
    bad = dict(t)
    bad["tb_np"] = np.asarray(t["tb_np"]).copy()
    bad["tb_np"][0] += 1
    assert b"point count" in host_lib.kcpu_check_tables(ctypes.byref(native.kapp_tables(bad).struct))
    bad = dict(t)
    bad["hi_edge"] = np.asarray(t["hi_edge"]).copy()
    bad["hi_edge"][0, 6] = 5000
    assert b"hotop" in host_lib.kcpu_check_tables(ctypes.byref(native.kapp_tables(bad).struct))
    bad = dict(t)
    bad["felo_npts"] = np.asarray(t["felo_npts"]).copy()
    bad["felo_npts"][0, 0] = 201
    assert b"Fe I/II" in host_lib.kcpu_check_tables(ctypes.byref(native.kapp_tables(bad).struct))
