"""wholecellfe_amd -- MI355X (gfx950) per-step cell mechanics hot path.

Thin ctypes binding of libwcfe.so (the C ABI declared in include/wcfe.h).
The compute lives in the HIP library; this module only marshals numpy arrays.
It raises if the library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwcfe.so")

OK = 0
ERR_ARG, ERR_HIP, ERR_OOM, ERR_DEGENERATE, ERR_ZERO_PIVOT, ERR_NOT_CONVERGED, ERR_STATE, ERR_NO_DEVICE = range(1, 9)
ERR_NAMES = {1: "ARG", 2: "HIP", 3: "OOM", 4: "DEGENERATE", 5: "ZERO_PIVOT", 6: "NOT_CONVERGED", 7: "STATE",
             8: "NO_DEVICE"}
LAW_UL, LAW_CST = 0, 1
SPRING_LINEAR, SPRING_TABLE = 0, 1
KIN_NONE, KIN_MEANFIELD, KIN_BELL = 0, 1, 2
PREC_MIXED, PREC_FP64 = 0, 1
PIVOT_STATIC, PIVOT_PARTIAL = 0, 1
SOLVER_AUTO, SOLVER_DENSE, SOLVER_CSR = 0, 1, 2

D = C.POINTER(C.c_double)
I32 = C.POINTER(C.c_int32)

# every symbol include/wcfe.h declares
EXPORTS = [
    "wcfe_default_options", "wcfe_create", "wcfe_destroy", "wcfe_last_error", "wcfe_device_arch", "wcfe_set_table",
    "wcfe_add_cell", "wcfe_finalize", "wcfe_newton_step", "wcfe_load_step", "wcfe_synchronize",
    "wcfe_get_positions", "wcfe_set_positions", "wcfe_get_ul_state", "wcfe_get_springs", "wcfe_get_cell_info",
    "wcfe_residual_fp64", "wcfe_element_systems", "wcfe_assembled_system", "wcfe_get_colors", "wcfe_dense_solve",
    "wcfe_calculate_res", "wcfe_force_fa", "wcfe_force_fa_subs", "wcfe_update_fa", "wcfe_set_external_forces",
    "wcfe_write_vtk", "wcfe_frame_capture", "wcfe_frame_get", "wcfe_frame_write_vtk", "wcfe_profile_step",
    "wcfe_copy_probe", "wcfe_autotune", "wcfe_set_ul_state", "wcfe_replace_cell", "wcfe_checkpoint_save",
    "wcfe_checkpoint_load", "wcfe_area_tot", "wcfe_bending_force_2d", "wcfe_area_constraint_force",
    "wcfe_force_ap", "wcfe_get_schedule", "wcfe_num_cells", "wcfe_get_cell_sizes",
    "wcfe_eval_exp", "wcfe_get_cell_layout", "wcfe_force_fa_noalloc", "wcfe_remesh_cell",
]
TUNE_SOLVER, TUNE_LU, TUNE_PCG = 1, 2, 4
# reference constants of the boundary and body forces (UL_growth_explicit.cpp)
K_BEND_REF = 20.0              # pN/um, :1653
K_AREA_REF = 1.0               # pN/um^3, :1776
AREA_0_REF = 10 * 10 * 3.1416  # um^2, :1777
DRAG_REF = 0.001               # pN s/um, :569
PHASES = ["fill", "assemble", "lu_gemm", "lu_panel", "solve", "refine", "residual", "commit", "lu_strip"]


class Options(C.Structure):
    _fields_ = [
        ("device", C.c_int32), ("precision", C.c_int32), ("pivoting", C.c_int32), ("batched", C.c_int32),
        ("use_graph", C.c_int32), ("max_refine", C.c_int32), ("lin_rtol", C.c_double),
        ("newton_rtol", C.c_double), ("newton_atol", C.c_double), ("max_newton", C.c_int32), ("dt", C.c_double),
        ("t0", C.c_double), ("seed_lo", C.c_uint32), ("seed_hi", C.c_uint32), ("k_on", C.c_double),
        ("k_off0", C.c_double), ("f_bell", C.c_double), ("kappa", C.c_double), ("symmetric", C.c_int32),
        ("dense_max_dof", C.c_int32), ("cg_rtol", C.c_double), ("cg_max_iter", C.c_int32),
        ("lu_groups", C.c_int32), ("lu_depth", C.c_int32), ("band", C.c_int32), ("fused_step", C.c_int32),
        ("twist", C.c_int32), ("pcg_lanes", C.c_int32),
    ]


class CellDesc(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int32), ("n_tris", C.c_int32), ("X", D), ("x", D), ("tris", I32), ("law", C.c_int32),
        ("mu", C.c_double), ("E", C.c_double), ("nu", C.c_double), ("thick", C.c_double), ("ea", D),
        ("node_springs", C.c_int32), ("cC", D), ("fau", D), ("rho_max", C.c_double), ("n_spr", C.c_int32),
        ("spr_tri", I32), ("spr_bary", D), ("spr_anchor", D), ("spr_state", D), ("spring_law", C.c_int32),
        ("k_lin", C.c_double), ("table_id", C.c_int32), ("kinetics", C.c_int32), ("solver", C.c_int32),
        ("k_bend", C.c_double), ("k_area", C.c_double), ("area_0", C.c_double), ("drag", C.c_double),
        ("ap", C.c_int32), ("ap_rate", C.c_double), ("ap_quantum", C.c_double), ("ap_density", C.c_double),
        ("passive", C.c_int32), ("f_ext", D),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("n_cells", C.c_int32), ("n_converged", C.c_int32), ("max_refine_used", C.c_int32),
        ("n_static_pivots", C.c_int32), ("n_failed", C.c_int32), ("max_lin_relres", C.c_double),
        ("max_newton_rel", C.c_double), ("max_cg_iters", C.c_int32), ("n_cg_breakdown", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


class WcfeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("wcfe error %d (%s): %s" % (code, ERR_NAMES.get(code, "?"), msg))
        self.code = code


def lib():
    """Load libwcfe.so; raise if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libwcfe.so is not built: run `python -m wholecellfe_amd.build` "
                              "(or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.wcfe_last_error.restype = C.c_char_p
        L.wcfe_create.argtypes = [C.POINTER(Options), C.POINTER(C.c_void_p)]
        for name in EXPORTS:
            getattr(L, name)  # raises AttributeError if missing
        _lib = L
    return _lib


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def _dp(a):
    return None if a is None else a.ctypes.data_as(D)


def _ip(a):
    return None if a is None else a.ctypes.data_as(I32)


def default_options(**kw):
    o = Options()
    lib().wcfe_default_options(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


VTK_CELL, VTK_SUBSTRATE = 0, 1


def write_vtk(path, tris, x, ea=None, fau=None, cA=None, cC=None, style=VTK_CELL, law=1, kint=-1.0, time=0.0):
    """The reference's write_vtk (iomesh.cpp:844 cell frame / :730 substrate
    frame) from host arrays: x, fau 3 per node, ea 6 per node (Voigt)."""
    tris, x = _i32(tris).reshape(-1), _f64(x).reshape(-1)
    arr = [None if a is None else _f64(a).reshape(-1) for a in (ea, fau, cA, cC)]
    rc = lib().wcfe_write_vtk(str(path).encode(), style, len(x) // 3, len(tris) // 3, _ip(tris), _dp(x),
                              *[_dp(a) for a in arr], law, C.c_double(kint), C.c_double(time))
    if rc != OK:
        raise WcfeError(rc, "wcfe_write_vtk failed")


class Context:
    """A batch of cells resident on one GPU (wcfe_ctx)."""

    def __init__(self, **options):
        self._L = lib()
        self.opt = default_options(**options)
        h = C.c_void_p()
        rc = self._L.wcfe_create(C.byref(self.opt), C.byref(h))
        self._h = h
        if rc != OK:
            msg = self._L.wcfe_last_error(h).decode() if h.value else "no device"
            if h.value:
                self._L.wcfe_destroy(h)
            self._h = C.c_void_p()
            raise WcfeError(rc, msg)
        self._keep = []
        self.cells = []

    def _check(self, rc):
        if rc != OK:
            raise WcfeError(rc, self._L.wcfe_last_error(self._h).decode())

    def close(self):
        if self._h and self._h.value:
            self._L.wcfe_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def arch(self):
        buf = C.create_string_buffer(64)
        self._check(self._L.wcfe_device_arch(self._h, buf, 64))
        return buf.value.decode()

    def set_table(self, table_id, px, py, k):
        px, py, k = _f64(px), _f64(py), _f64(k)
        self._check(self._L.wcfe_set_table(self._h, table_id, len(px), _dp(px), _dp(py), _dp(k)))

    def add_cell(self, X, tris, **kw):
        d = self._desc(X, tris, **kw)
        cid = C.c_int32(-1)
        self._check(self._L.wcfe_add_cell(self._h, C.byref(d), C.byref(cid)))
        self.cells.append(dict(n_nodes=d.n_nodes, n_tris=d.n_tris, n_spr=d.n_spr))
        return cid.value

    def replace_cell(self, cell, X, tris, **kw):
        """New mesh and state for one cell after finalize (e.g. after
        remeshing); the other cells keep their state."""
        d = self._desc(X, tris, **kw)
        self._check(self._L.wcfe_replace_cell(self._h, cell, C.byref(d)))
        self.cells[cell] = dict(n_nodes=d.n_nodes, n_tris=d.n_tris, n_spr=d.n_spr)
        self._keep = []

    def remesh_cell(self, cell, tris, src0=None, src1=None, w=None, x_new=None, spr_tri=None, spr_bary=None):
        """One cell's new triangles after remeshing, its state mapped on the
        device (new node i = w_i old[src0_i] + (1 - w_i) old[src1_i]);
        returns 1 when the batch was laid out again, 0 when the topology did
        not change."""
        tris = _i32(tris).reshape(-1)
        n = self.cells[cell]["n_nodes"] if src0 is None else len(src0)
        arrs = [_i32(src0), _i32(src1), _f64(w), None if x_new is None else _f64(x_new).reshape(-1),
                None if spr_tri is None else _i32(spr_tri).reshape(-1),
                None if spr_bary is None else _f64(spr_bary).reshape(-1)]
        relaid = C.c_int32(-1)
        self._check(self._L.wcfe_remesh_cell(self._h, cell, n, len(tris) // 3, _ip(tris), _ip(arrs[0]), _ip(arrs[1]),
                                             _dp(arrs[2]), _dp(arrs[3]), _ip(arrs[4]), _dp(arrs[5]),
                                             C.byref(relaid)))
        self.cells[cell] = dict(n_nodes=n, n_tris=len(tris) // 3, n_spr=self.cells[cell]["n_spr"])
        return relaid.value

    def checkpoint_save(self, path):
        self._check(self._L.wcfe_checkpoint_save(self._h, str(path).encode()))

    def checkpoint_load(self, path):
        """Load a checkpoint into this (empty) context; the per-cell sizes
        that size the getters' buffers come from the library."""
        self._check(self._L.wcfe_checkpoint_load(self._h, str(path).encode()))
        self.cells = self._library_sizes()

    def _library_sizes(self):
        n = C.c_int32(0)
        self._check(self._L.wcfe_num_cells(self._h, C.byref(n)))
        out = []
        for c in range(n.value):
            a, b, s = C.c_int32(0), C.c_int32(0), C.c_int32(0)
            self._check(self._L.wcfe_get_cell_sizes(self._h, c, C.byref(a), C.byref(b), C.byref(s)))
            out.append(dict(n_nodes=a.value, n_tris=b.value, n_spr=s.value))
        return out

    def _desc(self, X, tris, law=LAW_CST, x=None, mu=1000.0, E=1000.0, nu=0.3, thick=0.2, ea=None,
              node_springs=0, cC=None, fau=None, rho_max=100.0, spr_tri=None, spr_bary=None, spr_anchor=None,
              spr_state=None, spring_law=SPRING_LINEAR, k_lin=1000.0, table_id=0, kinetics=KIN_NONE,
              solver=SOLVER_AUTO, k_bend=0.0, k_area=0.0, area_0=AREA_0_REF, drag=0.0, ap=0, ap_rate=10.0,
              ap_quantum=5.0, ap_density=100.0, passive=0, f_ext=None):
        X = _f64(X).reshape(-1)
        tris = _i32(tris).reshape(-1)
        arrs = dict(X=X, x=_f64(x), tris=tris, ea=_f64(ea), cC=_f64(cC), fau=_f64(fau), spr_tri=_i32(spr_tri),
                    spr_bary=_f64(spr_bary), spr_anchor=_f64(spr_anchor), spr_state=_f64(spr_state),
                    f_ext=_f64(f_ext))
        for k in ("x", "ea", "cC", "fau", "spr_bary", "spr_anchor", "spr_state", "f_ext"):
            if arrs[k] is not None:
                arrs[k] = arrs[k].reshape(-1)
        if arrs["spr_tri"] is not None:
            arrs["spr_tri"] = arrs["spr_tri"].reshape(-1)
        self._keep.append(arrs)
        d = CellDesc()
        d.n_nodes = len(X) // 2
        d.n_tris = len(tris) // 3
        d.X, d.x, d.tris = _dp(X), _dp(arrs["x"]), _ip(tris)
        d.law, d.mu, d.E, d.nu, d.thick = law, mu, E, nu, thick
        d.ea = _dp(arrs["ea"])
        d.node_springs, d.cC, d.fau, d.rho_max = node_springs, _dp(arrs["cC"]), _dp(arrs["fau"]), rho_max
        d.n_spr = 0 if arrs["spr_tri"] is None else len(arrs["spr_tri"])
        d.spr_tri, d.spr_bary, d.spr_anchor, d.spr_state = (_ip(arrs["spr_tri"]), _dp(arrs["spr_bary"]),
                                                            _dp(arrs["spr_anchor"]), _dp(arrs["spr_state"]))
        d.spring_law, d.k_lin, d.table_id, d.kinetics = spring_law, k_lin, table_id, kinetics
        d.solver = solver
        d.k_bend, d.k_area, d.area_0, d.drag = k_bend, k_area, area_0, drag
        d.ap, d.ap_rate, d.ap_quantum, d.ap_density = int(ap), ap_rate, ap_quantum, ap_density
        d.passive, d.f_ext = int(passive), _dp(arrs["f_ext"])
        return d

    def finalize(self):
        self._check(self._L.wcfe_finalize(self._h))
        self._keep = []

    def newton_step(self, stats=True, check=True):
        """One batched Newton iteration.  With stats, the step's statistics;
        check=False returns them (with "error": code) instead of raising when
        the step reports an error (e.g. a linear solve that missed lin_rtol)."""
        if not stats:
            self._check(self._L.wcfe_newton_step(self._h, None))
            return None
        st = StepStats()
        rc = self._L.wcfe_newton_step(self._h, C.byref(st))
        if check:
            self._check(rc)
        d = st.as_dict()
        d["error"] = rc
        return d

    def load_step(self):
        st = StepStats()
        it = C.c_int32(0)
        self._check(self._L.wcfe_load_step(self._h, C.byref(it), C.byref(st)))
        d = st.as_dict()
        d["iterations"] = it.value
        return d

    def schedule(self):
        """The dense-LU schedule in effect: lu_depth, lu_groups, max_band
        (widest block band of the dense cells), band_solve (1: one workgroup
        per cell along the band), fused_step (1: one launch per block step)."""
        v = [C.c_int32(0) for _ in range(5)]
        self._check(self._L.wcfe_get_schedule(self._h, *[C.byref(x) for x in v]))
        return dict(lu_depth=v[0].value, lu_groups=v[1].value, max_band=v[2].value, band_solve=v[3].value,
                    fused_step=v[4].value)

    def autotune(self, flags=TUNE_SOLVER | TUNE_LU | TUNE_PCG, steps=2):
        """Time candidate settings on this batch and keep the fastest (call
        after finalize, before any Newton step)."""
        chosen = np.zeros(6, dtype=np.int32)
        ms = C.c_double(0.0)
        self._check(self._L.wcfe_autotune(self._h, flags, steps, _ip(chosen), C.byref(ms)))
        return dict(lu_depth=int(chosen[0]), lu_groups=int(chosen[1]), dense_max_dof=int(chosen[2]),
                    fused_step=int(chosen[3]), two_chain_cells=int(chosen[4]), pcg_lanes=int(chosen[5]),
                    ms_per_step=ms.value)

    def set_ul_state(self, cell, X=None, ea=None, cC=None, fau=None, time=float("nan")):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (X, ea, cC, fau)]
        self._check(self._L.wcfe_set_ul_state(self._h, cell, *[_dp(a) for a in arrs], C.c_double(time)))

    def profile_step(self):
        """One eager Newton step with per-phase HIP-event timing."""
        n = len(PHASES)
        ms = np.zeros(n)
        launches = np.zeros(n, dtype=np.int32)
        work = np.zeros(n)
        self._check(self._L.wcfe_profile_step(self._h, _dp(ms), _ip(launches), _dp(work)))
        return {p: dict(ms=float(ms[i]), launches=int(launches[i]), work=float(work[i])) for i, p in enumerate(PHASES)}

    def copy_probe(self, nbytes, reps=10):
        """Milliseconds per plain device copy of nbytes (read + write)."""
        ms = C.c_double(0)
        self._check(self._L.wcfe_copy_probe(self._h, C.c_int64(nbytes), reps, C.byref(ms)))
        return ms.value

    def eval_exp(self, x):
        """The library's FP64 exp (the kinetics' exp), evaluated on the GPU."""
        x = _f64(x).reshape(-1)
        y = np.zeros_like(x)
        self._check(self._L.wcfe_eval_exp(self._h, C.c_int64(len(x)), _dp(x), _dp(y)))
        return y

    def synchronize(self):
        self._check(self._L.wcfe_synchronize(self._h))

    def positions(self, cell):
        n = self.cells[cell]["n_nodes"]
        x = np.zeros(2 * n)
        self._check(self._L.wcfe_get_positions(self._h, cell, _dp(x)))
        return x

    def frame_capture(self):
        """Snapshot every cell's committed state; the copy to host memory
        overlaps the following Newton steps.  Returns the frame id."""
        fr = C.c_int32(-1)
        self._check(self._L.wcfe_frame_capture(self._h, C.byref(fr)))
        return fr.value

    def frame_get(self, frame, cell):
        """(X 2N, ea 3N, cC N, fau 2N, time) of one cell in a captured frame."""
        n = self.cells[cell]["n_nodes"]
        X, ea, cC, fau, t = np.zeros(2 * n), np.zeros(3 * n), np.zeros(n), np.zeros(2 * n), C.c_double(0)
        self._check(self._L.wcfe_frame_get(self._h, frame, cell, _dp(X), _dp(ea), _dp(cC), _dp(fau), C.byref(t)))
        return X, ea, cC, fau, t.value

    def frame_write_vtk(self, frame, cell, path, style=0, cA=None, law=1, kint=-1.0):
        """One cell of a captured frame as the reference's VTK file."""
        cA = _f64(cA)
        self._check(self._L.wcfe_frame_write_vtk(self._h, frame, cell, str(path).encode(), style, _dp(cA), law,
                                                 C.c_double(kint)))

    def set_external_forces(self, cell, f_ext):
        """New constant nodal load (2N, None = zero) of a cell added with f_ext."""
        f = _f64(f_ext)
        self._check(self._L.wcfe_set_external_forces(self._h, cell, _dp(None if f is None else f.reshape(-1))))

    def set_positions(self, cell, x):
        x = _f64(x).reshape(-1)
        self._check(self._L.wcfe_set_positions(self._h, cell, _dp(x)))

    def ul_state(self, cell):
        n = self.cells[cell]["n_nodes"]
        X, ea, cC, fau = np.zeros(2 * n), np.zeros(3 * n), np.zeros(n), np.zeros(2 * n)
        self._check(self._L.wcfe_get_ul_state(self._h, cell, _dp(X), _dp(ea), _dp(cC), _dp(fau)))
        return X, ea, cC, fau

    def springs(self, cell):
        n = self.cells[cell]["n_spr"]
        st, an = np.zeros(n), np.zeros(2 * n)
        if n:
            self._check(self._L.wcfe_get_springs(self._h, cell, _dp(st), _dp(an)))
        return st, an

    def cell_info(self, cell):
        t = C.c_double(0)
        step, nit, nc, ndof = C.c_int32(0), C.c_int32(0), C.c_int32(0), C.c_int32(0)
        self._check(self._L.wcfe_get_cell_info(self._h, cell, C.byref(t), C.byref(step), C.byref(nit), C.byref(nc),
                                               C.byref(ndof)))
        return dict(time=t.value, step=step.value, newton_iter=nit.value, n_colors=nc.value, ndof=ndof.value)

    def cell_layout(self, cell):
        """solver (SOLVER_DENSE / SOLVER_CSR), nblk (64-row blocks), bw (block
        bandwidth in the band numbering), sym (symmetric factorization),
        chain_top / chain_bottom (two-chain band LU: blocks of each chain; 0:
        one chain)."""
        v = [C.c_int32(0) for _ in range(6)]
        self._check(self._L.wcfe_get_cell_layout(self._h, cell, *[C.byref(x) for x in v]))
        return dict(solver=v[0].value, nblk=v[1].value, bw=v[2].value, sym=v[3].value, chain_top=v[4].value,
                    chain_bottom=v[5].value)

    def residual(self, cell):
        G = np.zeros(2 * self.cells[cell]["n_nodes"])
        self._check(self._L.wcfe_residual_fp64(self._h, cell, _dp(G)))
        return G

    def element_systems(self, cell):
        T = self.cells[cell]["n_tris"]
        Ge, Ke = np.zeros(6 * T), np.zeros(36 * T)
        self._check(self._L.wcfe_element_systems(self._h, cell, _dp(Ge), _dp(Ke)))
        return Ge.reshape(T, 6), Ke.reshape(T, 6, 6)

    def assembled_system(self, cell):
        n = 2 * self.cells[cell]["n_nodes"]
        K, G = np.zeros((n, n)), np.zeros(n)
        self._check(self._L.wcfe_assembled_system(self._h, cell, _dp(K), _dp(G)))
        return K, G

    def colors(self, cell):
        c = np.zeros(self.cells[cell]["n_tris"], dtype=np.int32)
        self._check(self._L.wcfe_get_colors(self._h, cell, _ip(c)))
        return c

    def dense_solve(self, mats, rhss):
        n = np.array([m.shape[0] for m in mats], dtype=np.int32)
        A = np.concatenate([np.ascontiguousarray(m, dtype=np.float64).reshape(-1) for m in mats])
        b = np.concatenate([np.ascontiguousarray(r, dtype=np.float64).reshape(-1) for r in rhss])
        x = np.zeros_like(b)
        rel = np.zeros(len(mats))
        it = np.zeros(len(mats), dtype=np.int32)
        piv = np.zeros(int(n.sum()), dtype=np.int32)
        self._check(self._L.wcfe_dense_solve(self._h, len(mats), _ip(n), _dp(A), _dp(b), _dp(x), _dp(rel), _ip(it),
                                             _ip(piv)))
        out, o = [], 0
        for k in n:
            out.append(x[o:o + k])
            o += k
        return out, rel, it, piv

    # ---- reference-shaped stateless calls (3 components per node) ----
    def calculate_res(self, tris, X, x, ea, time, mat=1):
        """GPU counterpart of calculateRes (UL_growth_explicit.cpp:632 cell,
        mat 1; :931 substrate, mat 2); returns Res (3N), Res_ea (6N), MM (N)."""
        tris, X, x, ea = _i32(tris).reshape(-1), _f64(X).reshape(-1), _f64(x).reshape(-1), _f64(ea).reshape(-1)
        n = len(X) // 3
        Res, Rea, MM = np.zeros(3 * n), np.zeros(6 * n), np.zeros(n)
        self._check(self._L.wcfe_calculate_res(self._h, n, len(tris) // 3, _ip(tris), _dp(X), _dp(x), _dp(ea),
                                               mat, C.c_double(time), _dp(Res), _dp(Rea), _dp(MM)))
        return Res, Rea, MM

    def force_fa(self, tris, x, cC, fau, law, kint):
        """GPU counterpart of calculateForce_FA (UL:1198 / UL1:1196)."""
        tris, x, cC, fau = _i32(tris).reshape(-1), _f64(x).reshape(-1), _f64(cC), _f64(fau).reshape(-1)
        n = len(x) // 3
        F = np.zeros(3 * n)
        self._check(self._L.wcfe_force_fa(self._h, n, len(tris) // 3, _ip(tris), _dp(x), _dp(cC), _dp(fau), law,
                                          C.c_double(kint), _dp(F)))
        return F

    def force_fa_subs(self, tris, x_cell, cC, fau, stris, Sx, law, kint):
        """GPU counterpart of calculateForce_FA_subs (UL:1305 / UL1:1297):
        the cell's integrin forces on the substrate nodes (3 per node)."""
        tris, x_cell, cC, fau = _i32(tris).reshape(-1), _f64(x_cell).reshape(-1), _f64(cC), _f64(fau).reshape(-1)
        stris, Sx = _i32(stris).reshape(-1), _f64(Sx).reshape(-1)
        F = np.zeros(len(Sx))
        self._check(self._L.wcfe_force_fa_subs(self._h, len(x_cell) // 3, len(tris) // 3, _ip(tris), _dp(x_cell),
                                               _dp(cC), _dp(fau), len(Sx) // 3, len(stris) // 3, _ip(stris),
                                               _dp(Sx), law, C.c_double(kint), _dp(F)))
        return F

    def area_tot(self, tris, x):
        """GPU counterpart of calculateAreaTot (UL:1662): [area, cx, cy, cz]."""
        tris, x = _i32(tris).reshape(-1), _f64(x).reshape(-1)
        out = np.zeros(4)
        self._check(self._L.wcfe_area_tot(self._h, len(x) // 3, len(tris) // 3, _ip(tris), _dp(x), _dp(out)))
        return out

    def bending_force_2d(self, tris, x, k_bend=K_BEND_REF):
        """GPU counterpart of calculateBendingForce2D (UL:1571); returns (F 3N, n_open)."""
        tris, x = _i32(tris).reshape(-1), _f64(x).reshape(-1)
        n = len(x) // 3
        F, op = np.zeros(3 * n), C.c_int32(0)
        self._check(self._L.wcfe_bending_force_2d(self._h, n, len(tris) // 3, _ip(tris), _dp(x), C.c_double(k_bend),
                                                  _dp(F), C.byref(op)))
        return F, op.value

    def area_constraint_force(self, tris, x, tot_area, k_area=K_AREA_REF, area_0=AREA_0_REF):
        """GPU counterpart of calculateAreaConstraintForce (UL:1691); returns (F 3N, n_open)."""
        tris, x = _i32(tris).reshape(-1), _f64(x).reshape(-1)
        n = len(x) // 3
        F, op = np.zeros(3 * n), C.c_int32(0)
        self._check(self._L.wcfe_area_constraint_force(self._h, n, len(tris) // 3, _ip(tris), _dp(x),
                                                       C.c_double(tot_area), C.c_double(k_area),
                                                       C.c_double(area_0), _dp(F), C.byref(op)))
        return F, op.value

    def force_ap(self, tris, x, dt, u=None, draw=0, rate=10.0, quantum=5.0, density=100.0):
        """GPU counterpart of calculateForce_AP (UL:1477) with per-vertex draws
        u (None: the context's Philox stream, counter (draw, vertex, 0, 2))."""
        tris, x = _i32(tris).reshape(-1), _f64(x).reshape(-1)
        n = len(x) // 3
        u = _f64(u)
        F, op = np.zeros(3 * n), C.c_int32(0)
        self._check(self._L.wcfe_force_ap(self._h, n, len(tris) // 3, _ip(tris), _dp(x), C.c_double(dt),
                                          C.c_double(rate), C.c_double(quantum), C.c_double(density), _dp(u),
                                          C.c_uint32(draw), _dp(F), C.byref(op)))
        return F, op.value

    def update_fa(self, X, x, cA, cC, fau, dt, stris, SX, Sx, law, kint):
        """GPU counterpart of updateFA, 11-argument form (UL:2026 / UL1:2005)."""
        X, x = _f64(X).reshape(-1), _f64(x).reshape(-1)
        cA, cC, fau = _f64(cA).copy(), _f64(cC).copy(), _f64(fau).reshape(-1).copy()
        stris, SX, Sx = _i32(stris).reshape(-1), _f64(SX).reshape(-1), _f64(Sx).reshape(-1)
        un = C.c_int32(0)
        self._check(self._L.wcfe_update_fa(self._h, len(X) // 3, _dp(X), _dp(x), _dp(cA), _dp(cC), _dp(fau),
                                           C.c_double(dt), len(SX) // 3, len(stris) // 3, _ip(stris), _dp(SX),
                                           _dp(Sx), law, C.c_double(kint), C.byref(un)))
        return cA, cC, fau, un.value
