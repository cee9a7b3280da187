"""The no-allocation RT pass (a12_atlas_rt_pass_noalloc) against the
allocating entry point it varies (a12_atlas_set_kapp + a12_atlas_rt_run) on
the same inputs: accumulators and HNU(1) bit for bit, on two calls in a row,
the second over fewer frequencies (so the fixed scratch array still holds
the first call's larger contents).  The caller's device buffers come
from hipMalloc of the HIP runtime the library itself uses; the library
allocates nothing in the call."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class DevBuf:
    """A caller-owned device buffer (hipMalloc/hipMemcpy through ctypes)."""

    def __init__(self, hip, a=None, n=None, fill=None):
        self.hip = hip
        src = None if a is None else np.ascontiguousarray(a, np.float64)
        self.n = src.size if src is not None else int(n)
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), C.c_size_t(8 * self.n)) == 0
        if src is None:
            src = np.full(self.n, np.nan if fill is None else fill)
        assert hip.hipMemcpy(self.p, src.ctypes.data_as(C.c_void_p), C.c_size_t(8 * self.n),
                             1) == 0  # hipMemcpyHostToDevice

    def get(self):
        out = np.empty(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p,
                                  C.c_size_t(8 * self.n), 2) == 0  # DeviceToHost
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _case(z, nf):
    """The fixture's first nf frequencies: (state dict, kapp tables, xlines)."""
    from atlas12_amd import native
    zz = {k: z[k] for k in z.files}
    for k in ("waveset", "rcoset"):
        zz[k] = np.asarray(z[k])[:nf]
    for k in ("xlines", "acont", "sigmac", "scont", "sline", "aline", "sigmal"):
        zz[k] = np.asarray(z[k])[:nf]
    return zz, native.atlas_state_from(zz)


@pytest.mark.parametrize("nf", [240, 97])
def test_rt_pass_noalloc_matches_rt_run(gpu_ctx, golden, coef, nf):
    from atlas12_amd import native
    z = golden("atlas_small.npz")
    gpu_ctx.set_operators(*coef)
    zz, d = _case(z, nf)
    n = len(zz["t"])
    # the allocating entry point
    native.atlas_load(gpu_ctx, d, zz["linedata"])
    kap = [zz[k] for k in ("acont", "sigmac", "scont", "sline", "aline", "sigmal")]
    native.atlas_set_kapp(gpu_ctx, *kap)
    acc, h1 = native.atlas_rt_pass(gpu_ctx, zz["xlines"])
    # the no-allocation one on caller-owned device buffers
    hip = C.CDLL("libamdhip64.so.7")  # the runtime liba12gpu.so is linked against
    bufs = [DevBuf(hip, zz["waveset"]), DevBuf(hip, zz["rcoset"]),
            DevBuf(hip, np.concatenate([zz["t"], zz["hkt"], zz["rhox"]])),
            DevBuf(hip, np.stack(kap)), DevBuf(hip, zz["xlines"]), DevBuf(hip, n=nf),
            DevBuf(hip, n=9 * n)]
    try:
        assert hip.hipDeviceSynchronize() == 0
        native.atlas_rt_pass_noalloc(gpu_ctx, n, nf, float(zz["teff"]),
                                     *[b.p.value for b in bufs])
        gpu_ctx.synchronize()
        h2, a2 = bufs[5].get(), bufs[6].get()
    finally:
        for b in bufs:
            b.free()
    ref = np.concatenate([acc[k] for k in native.ACC_NAMES] + [[acc["pradk0"]], np.zeros(n - 1)])
    assert np.array_equal(h2, h1)
    assert np.array_equal(a2[:8 * n], ref[:8 * n])
    assert a2[8 * n] == acc["pradk0"] and np.all(a2[8 * n + 1:] == 0.0)


def test_rt_pass_noalloc_rejects_oversize(gpu_ctx, coef):
    """Sizes beyond the fixed scratch array: a clear error, nothing launched
    (the pointers are never dereferenced)."""
    from atlas12_amd import native
    gpu_ctx.set_operators(*coef)
    nf = 2 * (1179648 // 72) + 2
    with pytest.raises(native.A12Error, match="exceeds"):
        native.atlas_rt_pass_noalloc(gpu_ctx, 72, nf, 5777.0, 8, 8, 8, 8, 8, 8, 8)
