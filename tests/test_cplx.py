"""SYNTHE's complex-profile lines without a GPU: the a12_cplx_line layout,
and the library's source (atlas12_amd/csrc/a12_cplx_core.h) compiled for
the host against the reference's own compute_line_opacity
(src/synthe_module.f90:3167-3702) -- BUFFER after the call, bit for bit, at
every point of three windows (tests/golden/cplx*_ref.npz: 365-660 nm with
Balmer, D alpha, He I, metals, PRD, autoionising and merged-continuum
records; 1450-2400 nm with Brackett and Pfund on the Peterson/Kurucz path;
118-180 nm with the Lyman series, Ly alpha's H2/H2+ satellites and m > 30
past the Stehle tables).  The same functions, in the same order, run in the
device kernels (a12_cplx.hip)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from test_kapp import _layout

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
WINDOWS = ("", "_ir", "_uv")


def test_line_layout_matches_header(tmp_path):
    from atlas12_amd import native
    fields = [f[0] for f in native.CplxLine._fields_]
    _layout(tmp_path, "a12_cplx_line", fields, native.CplxLine, "a12_cplx_line")
    assert native.CPLX_DTYPE.itemsize == ctypes.sizeof(native.CplxLine) == 64
    assert [native.CPLX_DTYPE.fields[f][1] for f in fields] == [getattr(native.CplxLine, f).offset
                                                                 for f in fields]


@pytest.fixture(scope="module")
def host_lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("cplx_host")
    so = d / "libcplx_host.so"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fPIC", "-shared",
                    "-I", os.path.join(REPO, "atlas12_amd", "csrc"), "-I", os.path.join(REPO, "include"),
                    os.path.join(REPO, "tests", "cplx_host", "cplx_host_eval.cpp"), "-o", str(so)], check=True)
    L = ctypes.CDLL(str(so))
    L.ccpu_voigt.restype = ctypes.c_float
    L.ccpu_voigt.argtypes = [ctypes.c_float, ctypes.c_float]
    L.ccpu_stark_qs.restype = ctypes.c_float
    L.ccpu_stark_qs.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int]
    L.ccpu_expint1.restype = ctypes.c_float
    L.ccpu_expint1.argtypes = [ctypes.c_float]
    return L


@pytest.fixture(scope="module")
def stark():
    from atlas12_amd import native
    return native.kapp_stark(np.load(os.path.join(GOLDEN, "kapp_tables.npz")))


def host_run(lib, stark, suf, flags=0, cont=None):
    """The host build on the fixture window's depths; returns (buffer,
    extents, depth values, convolutions built, reference)."""
    from atlas12_amd import native
    ref = np.load(os.path.join(GOLDEN, f"cplx{suf}_ref.npz"))
    st = np.load(os.path.join(GOLDEN, f"cplx{suf}_state.npz"))
    lines = native.cplx_lines(ref["lines"])
    L, js, mw6 = int(ref["length"]), ref["depths"], int(st["mw6"])
    nd = len(js)
    ratio = float(st["ratio"])
    grid_d = np.array([float(st["wbegin"]), ratio, np.log(ratio), float(ref["wlend"]), float(ref["cutoff"]), 1.0])
    grid_i = np.array([L, int(st["ixwlbeg"]), 1], np.int32)
    idx = list(st["ion_idx"])
    q = st["xnfpel_sub"]
    dx = np.stack([st[k][js] for k in ("t", "xne", "hckt", "rho", "xnfh", "xnfhe1", "xnfh2")] +
                  [q[js, idx.index(0)], q[js, idx.index(1)]], 1).copy()
    arr = {}
    for k in ("xnfdop", "dopple", "xnfpel"):
        a = np.zeros((nd, mw6))
        a[:, ref["ion_idx"]] = ref[k]
        arr[k] = a
    cont = np.ascontiguousarray(ref["continuum"] if cont is None else cont, np.float32)
    dout = np.zeros((nd, 4))
    ext = np.zeros((nd, len(lines), 4), np.int32)
    buf = np.zeros((nd, L), np.float32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    nconv = lib.ccpu_run(p(lines), len(lines), p(grid_d), p(grid_i), nd, p(dx), mw6, p(arr["xnfdop"]),
                         p(arr["dopple"]), p(arr["xnfpel"]), p(cont), ctypes.cast(stark.struct, ctypes.c_void_p),
                         p(dout), p(ext), p(buf), flags)
    return buf, ext, dout, nconv, ref


@pytest.mark.parametrize("suf", WINDOWS)
def test_host_build_reproduces_reference_bitwise(host_lib, stark, suf):
    buf, ext, dout, nconv, ref = host_run(host_lib, stark, suf)
    # the driver's per-depth values (src/synthe.f90:520-548)
    assert np.array_equal(dout[:, 0].astype(np.float32), ref["txnxn"])
    assert np.array_equal(dout[:, 1], ref["xnfph1"]) and np.array_equal(dout[:, 2], ref["xnfph2"])
    # BUFFER after compute_line_opacity, every point of every kept depth
    assert np.array_equal(buf, ref["buffer"]), np.argwhere(buf != ref["buffer"])[:5]
    assert np.count_nonzero(ref["buffer"]) > 40_000
    assert nconv > 0


def test_window_coverage(host_lib, stark):
    """The three windows reach every branch: each kind of line adds points,
    in both wings where the branch has two."""
    kinds = {}
    for suf in WINDOWS:
        buf, ext, _, _, ref = host_run(host_lib, stark, suf)
        lines = ref["lines"]
        red = (ext[:, :, 1] >= ext[:, :, 0]).sum(0)
        blue = (ext[:, :, 3] >= ext[:, :, 2]).sum(0)
        for l, ln in enumerate(lines):
            t = int(ln["itype"])
            key = ("H" if t in (-1, -2) else "metal" if t in (0, 3) or t <= -3 else "aut" if t == 1
                   else "coronal" if t == 2 else "cont")
            if key == "H":
                key += f"{ln['nblo']}" + ("iso" if ln["ncon"] == 0 or ln["nbup"] == ln["nblo"] + 1 else
                                         "beta" if ln["nbup"] == ln["nblo"] + 2 else "merge")
            r, b = kinds.get(key, (0, 0))
            kinds[key] = (r + int(red[l]), b + int(blue[l]))
    for k in ("H1iso", "H2iso", "H2beta", "H2merge", "H4merge", "H5merge", "metal", "aut"):
        assert kinds[k][0] > 0 and kinds[k][1] > 0, (k, kinds.get(k))
    # Ly beta and the merged Lyman lines lie blueward of 118 nm: red wings only
    assert kinds["H1beta"][0] > 0 and kinds["H1merge"][0] > 0
    assert kinds["cont"][0] > 0 and kinds["coronal"] == (0, 0)


def test_planted_mistakes_are_caught(host_lib, stark):
    """Two plausible simplifications each change a window's buffer: giving
    every H line its own convolution (the reference's cache serves H alpha
    with D alpha's), and closed-form wing wavelengths instead of the running
    product (points of the infrared and ultraviolet windows)."""
    for flags in (1, 2):
        changed = 0
        for suf in WINDOWS:
            good = host_run(host_lib, stark, suf)[0]
            bad = host_run(host_lib, stark, suf, flags=flags)[0]
            changed += int(np.sum(bad != good))
        assert changed > 0, flags


def test_voigt_matches_faddeeva(host_lib):
    """voigt_profile's Weideman N=32 / Humlicek form against scipy's
    Faddeeva function (1e-4: the asymptotic form's error at |z| = 15)."""
    from scipy.special import wofz
    rng = np.random.default_rng(3)
    v = np.concatenate([rng.uniform(0, 30, 400), rng.uniform(0, 3, 400)]).astype(np.float32)
    a = np.concatenate([10 ** rng.uniform(-4, 1.5, 400), 10 ** rng.uniform(-4, 0, 400)]).astype(np.float32)
    h = np.array([host_lib.ccpu_voigt(float(x), float(y)) for x, y in zip(v, a)])
    ref = wofz(v.astype(np.float64) + 1j * a.astype(np.float64)).real
    assert np.max(np.abs(h - ref) / ref) < 1e-4


def test_quasistatic_and_expint_closed_forms(host_lib):
    """stark_quasistatic_profile's Holtsmark asymptote beyond beta = 500 and
    its small-beta limit, and expint1 against scipy's E1 within the
    published accuracy of its Abramowitz & Stegun fits (5.1.53, 5.1.56)."""
    from scipy.special import exp1
    for b in (600.0, 2.0e3, 1.0e5):
        bf = np.float32(b)
        want = (np.float32(1.5) / np.sqrt(bf) + np.float32(27.0) / (bf * bf)) / (bf * bf)
        assert host_lib.ccpu_stark_qs(b, 0.3, 2, 3) == np.float32(want)
    # beta = 1, p = 0: 8 / (83 + 2.95) times 1 + PROB1's first entry
    assert abs(host_lib.ccpu_stark_qs(1.0, 0.0, 1, 2) / (8.0 / 85.95 * (1.0 - 0.980)) - 1.0) < 1e-5
    x = np.concatenate([np.linspace(0.01, 1.0, 50), np.linspace(1.0, 29.0, 60)])
    e = np.array([host_lib.ccpu_expint1(float(t)) for t in x])
    assert np.all(np.abs(e[:50] - exp1(x[:50])) < 1e-6)
    assert np.all(np.abs(e[50:] - exp1(x[50:])) / exp1(x[50:]) < 6e-5)
    assert host_lib.ccpu_expint1(31.0) == 0.0 and host_lib.ccpu_expint1(-1.0) == 0.0
