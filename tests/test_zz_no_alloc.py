"""gmi_evaluate_grand_mobility_noalloc against gmi_evaluate_grand_mobility.

The entry point takes device arrays the caller owns (SoA) and uses the
library's fixed device scratch array instead of allocating.  Two calls in a
row, the second on a smaller input, must reproduce the context entry point
(same kernels, same source splits): FP64 within 1e-12 relative and the
FP32-Kahan apply within 1e-6 (both are expected to agree bit for bit).  The
device arrays come from the HIP runtime libgmi.so is bound to, through
ctypes (test plumbing only).
"""
import ctypes as C

import numpy as np
import pytest

from generic_mobility_interface_class_amd import capi, geometry

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libgmi.so uses."""

    def __init__(self):
        capi.load()
        self.lib = C.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipSetDevice.argtypes = [C.c_int]
        self.bufs = []

    def check(self, rc, what):
        assert rc == 0, f"{what}: hip error {rc}"

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        p = C.c_void_p()
        self.check(self.lib.hipMalloc(C.byref(p), max(host.nbytes, 8)), "hipMalloc")
        self.bufs.append(p)
        if host.nbytes:
            self.check(self.lib.hipMemcpy(p, host.ctypes.data_as(C.c_void_p), host.nbytes, 1),
                       "H2D")
        return p

    def download(self, p, shape):
        out = np.empty(shape, dtype=np.float64)
        if out.nbytes:
            self.check(self.lib.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2),
                       "D2H")
        return out

    def sync(self):
        self.check(self.lib.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def free(self):
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


def _soa(aos, nc, ld):
    n = aos.shape[0]
    out = np.zeros((nc, ld))
    out[:, :n] = aos[:, :nc].T
    return out


def _context_apply(xs, fs, xb, fb, nseg, precision):
    ctx = capi.Context([xs.shape[0]], [xb.shape[0]], nseg, precision=precision, device=0)
    ctx.upload(xs, fs, xb, fb)
    ctx.call("gmi_evaluate_grand_mobility")
    vs, vb = ctx.velocities()
    ctx.close()
    return vs, vb


def _noalloc_apply(hip, xs, fs, xb, fb, precision, pad=7):
    ns, nb = xs.shape[0], xb.shape[0]
    lds, ldb = ns + pad, nb + pad          # any leading dimension >= n
    dxs, dfs = hip.upload(_soa(xs, 3, lds)), hip.upload(_soa(fs, 6, lds))
    dxb, dfb = hip.upload(_soa(xb, 3, ldb)), hip.upload(_soa(fb, 3, ldb))
    dvs, dvb = hip.upload(np.full((6, lds), np.nan)), hip.upload(np.full((3, ldb), np.nan))
    rc = capi.load().gmi_evaluate_grand_mobility_noalloc(
        dxs, dfs, ns, lds, dxb, dfb, nb, ldb, dvs, dvb, 1.0, 1.0, 1.0, precision, None)
    assert rc == capi.GMI_OK, rc
    hip.sync()
    vs = hip.download(dvs, (6, lds))[:, :ns].T.copy()
    vb = hip.download(dvb, (3, ldb))[:, :nb].T.copy()
    return vs, vb


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("precision,tol", [(capi.GMI_FP64, 1e-12), (capi.GMI_FP32_KAHAN, 1e-6)])
def test_noalloc_matches_context_apply_twice(gpu, precision, tol):
    hip = _Hip()
    hip.check(hip.lib.hipSetDevice(0), "hipSetDevice")
    try:
        # the larger input first (2,052 split partial slots), then a smaller
        # one: the second call must not see the first call's scratch contents
        for name, seed in (("step_9122", 11), ("apply_962", 12)):
            g = geometry.workload_geometry(name)
            xs, xb = g.positions()
            fs, fb = geometry.random_forces(g.ns, g.nblob, seed=seed)
            print(f"noalloc: {name} precision {precision}", flush=True)
            rs, rb = _context_apply(xs, fs, xb, fb, g.nseg, precision)
            vs, vb = _noalloc_apply(hip, xs, fs, xb, fb, precision)
            es, eb = _rel(vs, rs), _rel(vb, rb)
            print(f"  rel diff segs {es:.2e} blobs {eb:.2e} bitwise "
                  f"{np.array_equal(vs, rs) and np.array_equal(vb, rb)}", flush=True)
            assert es <= tol and eb <= tol, (name, es, eb)
    finally:
        hip.free()


def test_noalloc_rejects_oversize_before_device_work(gpu):
    lib = capi.load()
    n = capi.GMI_NOALLOC_MAX_TILES * 256 + 1
    fake = C.c_void_p(256)   # never dereferenced: the size check comes first
    rc = lib.gmi_evaluate_grand_mobility_noalloc(fake, fake, n, n, fake, fake, 0, 0, fake, fake,
                                                 1.0, 1.0, 1.0, capi.GMI_FP64, None)
    assert rc == capi.GMI_ERR_TOO_LARGE
    rc = lib.gmi_evaluate_grand_mobility_noalloc(fake, fake, 10, 5, fake, fake, 0, 0, fake, fake,
                                                 1.0, 1.0, 1.0, capi.GMI_FP64, None)
    assert rc == capi.GMI_ERR_ARG   # ld < n
