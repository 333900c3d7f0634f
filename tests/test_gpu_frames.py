"""Asynchronous frame output on the GPU.

A captured frame is the batch's committed state at the capture's place in
stream order (positions, strain, bound fraction, spring extension, load
time), copied to pinned host memory while the following Newton steps run.
Checks: a frame taken before another load step holds the earlier state bit
for bit; the frame's VTK file equals the host writer's file for the same
state (itself byte-identical to the reference's write_vtk, tests/test_vtk.py);
slots, replacement and relayout are refused as documented.
"""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

W = pytest.importorskip("wholecellfe_amd")

DT = 0.01


def cell_desc(seed):
    z = golden("bound_disk296")
    X = z["X"][:, :2].copy()
    N = len(X)
    rng = np.random.default_rng(seed)
    x = X + 0.01 * rng.standard_normal(X.shape)
    return dict(X=X.reshape(-1), x=x.reshape(-1), tris=z["tris"].reshape(-1).astype(np.int32), law=0, mu=1000.0,
                thick=1.0, ea=(0.01 * rng.standard_normal((N, 3))).reshape(-1), node_springs=1,
                cC=rng.uniform(0.2, 1.0, N), fau=(0.004 * rng.standard_normal((N, 2))).reshape(-1), rho_max=100.0,
                spring_law=W.SPRING_LINEAR, k_lin=1000.0, kinetics=W.KIN_MEANFIELD, k_bend=20.0, k_area=1.0,
                drag=0.001)


def make_ctx():
    g = W.Context(precision=W.PREC_MIXED, t0=1.0, dt=DT, newton_rtol=1e-10, newton_atol=1e-8, max_refine=16)
    for s in (1, 2):
        g.add_cell(**cell_desc(s))
    g.finalize()
    return g


def state(g, c):
    X, ea, cC, fau = g.ul_state(c)
    return X, ea, cC, fau, g.cell_info(c)["time"]


def to3(a2):
    a = np.zeros((len(a2) // 2, 3))
    a[:, :2] = a2.reshape(-1, 2)
    return a


def voigt(ea3):
    e = ea3.reshape(-1, 3)
    v = np.zeros((len(e), 6))
    v[:, 0], v[:, 1], v[:, 3] = e[:, 0], e[:, 1], e[:, 2]
    return v


def test_frame_holds_the_state_at_capture(tmp_path):
    g = make_ctx()
    try:
        g.load_step()
        s0 = [state(g, c) for c in (0, 1)]
        f0 = g.frame_capture()
        g.load_step()  # runs while frame 0 is copied
        s1 = [state(g, c) for c in (0, 1)]
        f1 = g.frame_capture()
        for c in (0, 1):
            got0, got1 = g.frame_get(f0, c), g.frame_get(f1, c)
            for a, b in zip(got0, s0[c]):
                assert np.array_equal(a, b)
            for a, b in zip(got1, s1[c]):
                assert np.array_equal(a, b)
        assert s1[0][4] > s0[0][4]
        assert not np.array_equal(s1[0][0], s0[0][0])
        # the frame's file is the host writer's file for the same state
        d = cell_desc(1)
        X, ea, cC, fau, t = s0[0]
        cA = np.linspace(50.0, 100.0, len(cC))
        for style, law, kint in [(W.VTK_CELL, 1, -1.0), (W.VTK_CELL, 0, 31000.0), (W.VTK_SUBSTRATE, 1, -1.0)]:
            p_frame, p_host = tmp_path / "frame.vtk", tmp_path / "host.vtk"
            g.frame_write_vtk(f0, 0, p_frame, style=style, cA=cA, law=law, kint=kint)
            W.write_vtk(p_host, d["tris"], to3(X), ea=voigt(ea), fau=to3(fau), cA=cA, cC=cC, style=style, law=law,
                        kint=kint, time=t)
            assert p_frame.read_bytes() == p_host.read_bytes()
        # two frames in flight: a third capture replaces frame 0
        f2 = g.frame_capture()
        with pytest.raises(W.WcfeError) as e:
            g.frame_get(f0, 0)
        assert e.value.code == 1
        for a, b in zip(g.frame_get(f2, 1), s1[1]):
            assert np.array_equal(a, b)
        with pytest.raises(W.WcfeError):
            g.frame_get(f2 + 1, 0)
        # a new layout invalidates earlier frames
        g.replace_cell(1, **cell_desc(3))
        with pytest.raises(W.WcfeError) as e:
            g.frame_get(f2, 0)
        assert e.value.code == W.ERR_STATE
        f3 = g.frame_capture()
        assert np.array_equal(g.frame_get(f3, 1)[0], cell_desc(3)["X"])
    finally:
        g.close()
