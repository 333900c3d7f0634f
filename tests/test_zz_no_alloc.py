"""wcfe_force_fa_noalloc (device pointers, scratch in a static device array,
no allocation) against wcfe_force_fa (host arrays, the existing entry point):
the same bits, on two calls in a row, the second on a smaller mesh; sizes
beyond the scratch array's capacity are refused.

The device buffers are the caller's: allocated here through the HIP runtime
the library itself links (libamdhip64 from /opt/rocm), never by the entry
point."""
import ctypes as C

import numpy as np
import pytest

import cells
import wholecellfe_amd as W

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.bufs = []

    def check(self, rc, what):
        if rc != 0:
            raise RuntimeError("%s failed: %d" % (what, rc))

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.check(self.L.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 8))), "hipMalloc")
        self.bufs.append(p)
        self.check(self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), H2D), "hipMemcpy H2D")
        return p

    def empty(self, nbytes):
        p = C.c_void_p()
        self.check(self.L.hipMalloc(C.byref(p), C.c_size_t(max(nbytes, 8))), "hipMalloc")
        self.bufs.append(p)
        return p

    def download(self, p, n, dtype):
        out = np.zeros(n, dtype=dtype)
        self.check(self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), D2H), "hipMemcpy D2H")
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _inputs(K, seed):
    d = cells.ul_cell(K=K, seed=seed, table=False, deform=0.02)
    N = len(d["X"]) // 2
    rng = np.random.default_rng(seed)
    x3 = np.zeros((N, 3))
    x3[:, :2] = d["x"].reshape(-1, 2)
    fau3 = np.zeros((N, 3))
    fau3[:, :2] = 0.006 * rng.standard_normal((N, 2))  # spans several MD-table segments
    cC = rng.uniform(0.1, 1.0, N)
    return d["tris"].reshape(-1, 3).astype(np.int32), x3, cC, fau3


def _noalloc(ctx, hip, tris, x3, cC, fau3, law, kint):
    N, T = len(x3), len(tris)
    dt_, dx, dc, df = hip.upload(tris), hip.upload(x3), hip.upload(cC), hip.upload(fau3)
    dF = hip.empty(8 * 3 * N)
    dbad = hip.upload(np.array([-1], dtype=np.int32))
    rc = ctx._L.wcfe_force_fa_noalloc(ctx._h, N, T, dt_, dx, dc, df, law, C.c_double(kint), dF, dbad)
    ctx._check(rc)
    ctx.synchronize()
    F = hip.download(dF, 3 * N, np.float64)
    bad = int(hip.download(dbad, 1, np.int32)[0])
    return F, bad


@pytest.mark.parametrize("law,kint", [(1, 0.0), (0, 1000.0)])
def test_noalloc_force_fa_equals_force_fa(law, kint):
    hip = Hip()
    try:
        with W.Context() as ctx:
            for K, seed in ((14, 3), (5, 4)):  # the second call on a smaller mesh
                tris, x3, cC, fau3 = _inputs(K, seed)
                print("case: law %d, %d nodes, %d triangles" % (law, len(x3), len(tris)), flush=True)
                F_ref = ctx.force_fa(tris, x3, cC, fau3, law=law, kint=kint)
                F, bad = _noalloc(ctx, hip, tris, x3, cC, fau3, law, kint)
                assert bad == 0
                assert np.array_equal(F.view(np.int64), np.asarray(F_ref).reshape(-1).view(np.int64))
                assert np.any(F != 0.0)
            # beyond the scratch array's capacity: refused before any launch
            rc = ctx._L.wcfe_force_fa_noalloc(ctx._h, 262145, 10, hip.bufs[0], hip.bufs[1], hip.bufs[2],
                                              hip.bufs[3], law, C.c_double(kint), hip.bufs[4], None)
            assert rc == W.ERR_ARG
            assert "capacity" in ctx._L.wcfe_last_error(ctx._h).decode()
    finally:
        hip.free()
