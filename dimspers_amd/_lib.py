"""ctypes binding of libdimspers_stokes.so (the C ABI in include/dimspers_stokes.h).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libdimspers_stokes.so"

STK_OK = 0
STK_E_HIP, STK_E_SHAPE, STK_E_STATE, STK_E_NOTCONV, STK_E_NAN, STK_E_ARG, STK_E_NOMESH = -1, -2, -3, -4, -5, -6, -7

ASM_FP64, ASM_FP32, ASM_CARTESIAN, ASM_NO_BC, ASM_ISOTROPIC, ASM_NO_MG = 1, 2, 4, 8, 16, 32
ASM_FP16_SMOOTHER = 64
ASM_7POINT, ASM_CLAMP = 128, 256
SMOOTH_GS, SMOOTH_CHEB = 0, 1
MAT_K, MAT_G, MAT_GT, MAT_M, MAT_S, MAT_KT, MAT_MT, MAT_R, MAT_RK = 0, 1, 2, 3, 4, 5, 6, 7, 8
MAT_LEVEL, MAT_PROLONG, MAT_LEVEL_RA, MAT_COARSE_INV = 16, 48, 80, 112
SOLVE_PLAIN, SOLVE_MIXED, SOLVE_LEARNED = 0, 1, 2
INNER_MG, INNER_JACOBI = 0, 1
STOP_SCHUR, STOP_FULL = 0, 1
NULL_PROJECT, NULL_DEFLATE = 0, 1
PPC_LUMPED, PPC_IC = 0, 1

EXPORTED = [
    "stk_default_opts", "stk_create", "stk_destroy", "stk_last_error", "stk_info", "stk_set_fields",
    "stk_assemble", "stk_export_csc", "stk_export_rhs", "stk_spmv", "stk_solve", "stk_get_solution",
    "stk_updatevelocity", "stk_time_kernel", "stk_default_train_opts", "stk_set_fabric", "stk_mlp_train",
    "stk_mlp_params", "stk_mlp_forward", "stk_mlp_gradient", "stk_local_schur", "stk_learned_setup",
    "stk_set_stress_fields", "stk_stress", "stk_deformation_gradient", "stk_stats", "stk_thermal_solve",
    "stk_fse", "stk_set_dem", "stk_anisotropic_tensor", "stk_colouring", "stk_set_solution", "stk_set_operators",
    "stk_stats_noalloc",
]
N_PARAM = 9094
KERNEL_SPMV_K32, KERNEL_SPMV_K64, KERNEL_CHEB_K32, KERNEL_CHEB_K64, KERNEL_ASSEMBLY, KERNEL_CHEB_KH16 = 0, 1, 2, 3, 4, 5
KERNEL_COPY, KERNEL_VCYCLE32, KERNEL_INNER_IT32, KERNEL_STRESS, KERNEL_DEFGRAD, KERNEL_KP32 = 6, 7, 8, 9, 10, 11
KERNEL_GS_BWD32, KERNEL_FSE, KERNEL_ANISO = 12, 13, 14
KERNEL_G32, KERNEL_GT32, KERNEL_COARSE32, KERNEL_CGUPD32, KERNEL_RESTRICT32 = 15, 16, 17, 18, 19
STRESS_FIELDS = ("exx", "ezz", "exz", "eII", "txx", "tzz", "txz", "tII")
FSE_NO_HEALING, FSE_FABRIC_DESTRUCTION, FSE_ANNEALING_DESTRUCTION = 0, 1, 2
INSTALL_ASSEMBLY, INSTALL_STRESS = 1, 2
COLOUR_ELEMENTS, COLOUR_NODES = 0, 1


class StkOpts(C.Structure):
    _fields_ = [
        ("method", C.c_int32), ("inner", C.c_int32), ("stop", C.c_int32), ("itmax", C.c_int32),
        ("itmin", C.c_int32), ("inner_itmax", C.c_int32), ("refine_max", C.c_int32), ("cheb_degree", C.c_int32),
        ("null_mode", C.c_int32), ("pc_sweeps", C.c_int32),
        ("rtol", C.c_double), ("inner_rtol", C.c_double), ("mixed_inner_rtol", C.c_double),
        ("mixed_schur_rtol", C.c_double), ("stall_ratio", C.c_double),
        ("inner_warm", C.c_int32), ("fine_smoother", C.c_int32), ("warm_start", C.c_int32),
        ("pressure_pc", C.c_int32),
    ]


class StkReport(C.Structure):
    _fields_ = [
        ("converged", C.c_int32), ("outer_iterations", C.c_int32), ("inner_iterations", C.c_int64),
        ("refinements", C.c_int32), ("fallbacks", C.c_int32), ("schur_residual", C.c_double),
        ("full_residual", C.c_double), ("solve_ms", C.c_double), ("bytes", C.c_double),
        ("launches", C.c_int64), ("launches_floor", C.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class StkTrainOpts(C.Structure):
    _fields_ = [
        ("snapshots", C.c_int32), ("per_snapshot", C.c_int32), ("epochs", C.c_int32), ("batch", C.c_int32),
        ("ratio_kind", C.c_int32), ("reserved", C.c_int32), ("lr", C.c_double), ("lr_final", C.c_double),
        ("seed", C.c_uint64),
    ]


class StkTrainReport(C.Structure):
    _fields_ = [
        ("initial_loss", C.c_double), ("final_loss", C.c_double), ("validation_loss", C.c_double),
        ("ms", C.c_double), ("samples", C.c_int64), ("steps", C.c_int32), ("reserved", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class StkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None

dp = C.POINTER(C.c_double)
i64p = C.POINTER(C.c_int64)
vp = C.c_void_p


def load(path: Path | str | None = None):
    """Load (once) and return the shared library with argument types set."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(str(p))
    lib.stk_default_opts.argtypes = [C.POINTER(StkOpts)]
    lib.stk_default_opts.restype = None
    lib.stk_create.argtypes = [C.POINTER(vp), C.c_int64, C.c_int64, C.c_int64, i64p, dp, dp, dp, dp, i64p, C.c_int64]
    lib.stk_destroy.argtypes = [vp]
    lib.stk_destroy.restype = None
    lib.stk_last_error.argtypes = [vp]
    lib.stk_last_error.restype = C.c_char_p
    lib.stk_info.argtypes = [vp, i64p]
    lib.stk_set_fields.argtypes = [vp, C.c_double, dp, dp, C.c_int64, C.POINTER(dp)]
    lib.stk_assemble.argtypes = [vp, C.c_int32]
    lib.stk_set_solution.argtypes = [vp, dp, dp]
    lib.stk_stats_noalloc.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.stk_set_operators.argtypes = [vp, C.c_int32, i64p, i64p, dp, C.c_int32, i64p, i64p, dp, i64p, i64p, dp, dp,
                                      C.c_int32]
    lib.stk_export_csc.argtypes = [vp, C.c_int32, C.c_int32, i64p, i64p, i64p, i64p, i64p, dp]
    lib.stk_export_rhs.argtypes = [vp, dp]
    lib.stk_spmv.argtypes = [vp, C.c_int32, C.c_int32, dp, dp]
    lib.stk_solve.argtypes = [vp, C.POINTER(StkOpts), C.POINTER(StkReport)]
    lib.stk_get_solution.argtypes = [vp, dp, dp]
    lib.stk_updatevelocity.argtypes = [vp, dp, dp, dp]
    lib.stk_time_kernel.argtypes = [vp, C.c_int32, C.c_int32, dp, dp]
    fp = C.POINTER(C.c_float)
    lib.stk_default_train_opts.argtypes = [C.POINTER(StkTrainOpts)]
    lib.stk_default_train_opts.restype = None
    lib.stk_set_fabric.argtypes = [vp, dp, dp]
    lib.stk_mlp_train.argtypes = [vp, C.POINTER(StkTrainOpts), C.POINTER(StkTrainReport)]
    lib.stk_mlp_params.argtypes = [vp, fp, fp, C.c_int32]
    lib.stk_mlp_forward.argtypes = [vp, C.c_int64, fp, fp]
    lib.stk_mlp_gradient.argtypes = [vp, C.c_int64, fp, fp, fp, dp]
    lib.stk_local_schur.argtypes = [vp, dp, fp]
    lib.stk_learned_setup.argtypes = [vp]
    lib.stk_set_stress_fields.argtypes = [vp, dp, C.c_int32, C.POINTER(dp)]
    lib.stk_stress.argtypes = [vp, dp, C.POINTER(dp)]
    lib.stk_deformation_gradient.argtypes = [vp, dp, C.c_double, dp]
    lib.stk_stats.argtypes = [vp, dp, dp, dp]
    lib.stk_thermal_solve.argtypes = [vp, dp, dp, C.c_double, C.c_double, C.c_double, C.c_double, i64p, dp,
                                      C.c_int64, C.c_double, C.POINTER(StkReport)]
    lib.stk_fse.argtypes = [vp, dp, C.c_int32, C.c_int32, C.c_double, dp]
    lib.stk_colouring.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    lib.stk_set_dem.argtypes = [vp, dp, dp, C.c_int64, C.c_double]
    lib.stk_anisotropic_tensor.argtypes = [vp, dp, C.c_int32, dp, dp, dp, C.c_int64, C.c_int32, C.POINTER(dp)]
    for name in EXPORTED:
        if name not in ("stk_default_opts", "stk_destroy", "stk_last_error", "stk_default_train_opts"):
            getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


def _d(a: np.ndarray):
    return a.ctypes.data_as(dp)


def _i(a: np.ndarray):
    return a.ctypes.data_as(i64p)


def _f(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def default_train_opts(**kw) -> StkTrainOpts:
    o = StkTrainOpts()
    load().stk_default_train_opts(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        setattr(o, k, v)
    return o


def default_opts(**kw) -> StkOpts:
    o = StkOpts()
    load().stk_default_opts(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        setattr(o, k, v)
    return o


class Context:
    """Owns one stk_ctx (device memory + stream on the current device)."""

    def __init__(self, nel, nnod, nvert, e2n_1based, el_theta, el_r, node_theta, node_r, fixed_1based):
        self.lib = load()
        e2n = np.ascontiguousarray(np.asfortranarray(e2n_1based, dtype=np.int64).T)   # element-major = Julia 6 x nel
        self._keep = [
            e2n,
            np.ascontiguousarray(np.asarray(el_theta, dtype=np.float64).T),
            np.ascontiguousarray(np.asarray(el_r, dtype=np.float64).T),
            np.ascontiguousarray(node_theta, dtype=np.float64),
            np.ascontiguousarray(node_r, dtype=np.float64),
            np.ascontiguousarray(fixed_1based, dtype=np.int64),
        ]
        h = vp()
        rc = self.lib.stk_create(C.byref(h), nel, nnod, nvert, _i(self._keep[0]), _d(self._keep[1]),
                                 _d(self._keep[2]), _d(self._keep[3]), _d(self._keep[4]), _i(self._keep[5]),
                                 self._keep[5].size)
        if rc != STK_OK:
            raise StkError(rc, self.lib.stk_last_error(None).decode())
        self.h = h
        info = np.zeros(7, dtype=np.int64)
        self.lib.stk_info(self.h, _i(info))
        self.nel, self.nnod, self.nvert, self.nU, self.nP, self.ncolours, self.mg_levels = (int(x) for x in info)

    @classmethod
    def from_mesh(cls, mesh, fixed=None):
        """Context of a mesh with the 0-based fixed dofs ``fixed`` (default: free slip)."""
        if fixed is None:
            fixed, _ = mesh.free_slip()
        fixed = np.asarray(fixed, dtype=np.int64)
        return cls(mesh.nel, mesh.nnod, mesh.nvert, mesh.e2n + 1, mesh.el_theta, mesh.el_r, mesh.theta, mesh.r,
                   fixed + 1)

    def close(self):
        if getattr(self, "h", None):
            self.lib.stk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, allow=()):
        if rc != STK_OK and rc not in allow:
            raise StkError(rc, self.lib.stk_last_error(self.h).decode())
        return rc

    def set_fields(self, g, rho, eta_ip, C6=None):
        eta = np.asfortranarray(eta_ip, dtype=np.float64)
        nip = eta.shape[1]
        self._fields = [np.ascontiguousarray(rho, dtype=np.float64), eta]
        arr = None
        if C6 is not None:
            comps = [np.asfortranarray(c, dtype=np.float64) for c in C6]
            self._fields += comps
            arr = (dp * 6)(*[_d(c) for c in comps])
        self._check(self.lib.stk_set_fields(self.h, float(g), _d(self._fields[0]), _d(eta), nip, arr))

    def assemble(self, flags=ASM_FP64):
        self._check(self.lib.stk_assemble(self.h, int(flags)))

    def export_csc(self, which, lower_only=False):
        """Assembled operator as a 0-based scipy CSC matrix."""
        import scipy.sparse as sp
        nr, nc, nnz = C.c_int64(), C.c_int64(), C.c_int64()
        self._check(self.lib.stk_export_csc(self.h, which, int(lower_only), C.byref(nr), C.byref(nc), C.byref(nnz),
                                            None, None, None))
        colptr = np.zeros(nc.value + 1, dtype=np.int64)
        rowval = np.zeros(nnz.value, dtype=np.int64)
        val = np.zeros(nnz.value)
        self._check(self.lib.stk_export_csc(self.h, which, int(lower_only), None, None, None, _i(colptr), _i(rowval),
                                            _d(val)))
        return sp.csc_matrix((val, rowval - 1, colptr - 1), shape=(nr.value, nc.value))

    def export_rhs(self):
        F = np.zeros(self.nU)
        self._check(self.lib.stk_export_rhs(self.h, _d(F)))
        return F

    def spmv(self, which, x, precision=64):
        x = np.ascontiguousarray(x, dtype=np.float64)
        ny = {MAT_K: self.nU, MAT_G: self.nU, MAT_GT: self.nP, MAT_M: self.nP, MAT_S: self.nP,
              MAT_R: 2 * self.nP, MAT_RK: 2 * self.nP}[which]
        y = np.zeros(ny)
        self._check(self.lib.stk_spmv(self.h, which, precision, _d(x), _d(y)))
        return y

    def solve(self, opts: StkOpts | None = None, allow_notconv=False):
        rep = StkReport()
        o = opts if opts is not None else default_opts()
        rc = self.lib.stk_solve(self.h, C.byref(o), C.byref(rep))
        self._check(rc, allow=(STK_E_NOTCONV,) if allow_notconv else ())
        return rep

    def set_solution(self, U=None, P=None):
        """The (U, P) a solve with warm_start = 1 starts from (None: zero)."""
        u = None if U is None else np.ascontiguousarray(U, dtype=np.float64)
        p = None if P is None else np.ascontiguousarray(P, dtype=np.float64)
        self._check(self.lib.stk_set_solution(self.h, None if u is None else _d(u), None if p is None else _d(p)))

    def set_operators(self, K, MM, GG, Rhs, flags=ASM_FP64, K_lower=False, M_lower=True):
        """StokesPcCG's operators from scipy sparse matrices (0-based), sent as 1-based CSC."""
        import scipy.sparse as sp

        def csc(A):
            A = sp.csc_matrix(A)
            A.sort_indices()
            return (np.ascontiguousarray(A.indptr + 1, dtype=np.int64),
                    np.ascontiguousarray(A.indices + 1, dtype=np.int64),
                    np.ascontiguousarray(A.data, dtype=np.float64))
        k, m, g = csc(K), csc(MM), csc(GG)
        f = np.ascontiguousarray(Rhs, dtype=np.float64)
        self._check(self.lib.stk_set_operators(self.h, int(K_lower), _i(k[0]), _i(k[1]), _d(k[2]), int(M_lower),
                                               _i(m[0]), _i(m[1]), _d(m[2]), _i(g[0]), _i(g[1]), _d(g[2]), _d(f),
                                               flags))

    def solution(self):
        U = np.zeros(self.nU)
        P = np.zeros(self.nP)
        self._check(self.lib.stk_get_solution(self.h, _d(U), _d(P)))
        return U, P

    def time_kernel(self, kernel, reps=20):
        """(mean ms per launch, algorithmic bytes per launch) of one hot kernel."""
        ms, by = C.c_double(), C.c_double()
        self._check(self.lib.stk_time_kernel(self.h, kernel, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    # ---- learned local-Schur preconditioner
    def set_fabric(self, phi, delta):
        self._fab = [np.ascontiguousarray(phi, dtype=np.float64), np.ascontiguousarray(delta, dtype=np.float64)]
        self._check(self.lib.stk_set_fabric(self.h, _d(self._fab[0]), _d(self._fab[1])))

    def mlp_train(self, opts: StkTrainOpts | None = None):
        rep = StkTrainReport()
        self._check(self.lib.stk_mlp_train(self.h, C.byref(opts if opts is not None else default_train_opts()),
                                           C.byref(rep)))
        return rep

    def mlp_params(self, params=None, norm=None):
        """Get (no arguments) or set the FP32 parameters and normalisation."""
        if params is None:
            p = np.zeros(N_PARAM, dtype=np.float32)
            nm = np.zeros(22, dtype=np.float32)
            self._check(self.lib.stk_mlp_params(self.h, _f(p), _f(nm), 0))
            return p, nm
        p = np.ascontiguousarray(params, dtype=np.float32)
        nm = np.ascontiguousarray(norm, dtype=np.float32)
        self._check(self.lib.stk_mlp_params(self.h, _f(p), _f(nm), 1))
        return p, nm

    def mlp_forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], 6), dtype=np.float32)
        self._check(self.lib.stk_mlp_forward(self.h, X.shape[0], _f(X), _f(Y)))
        return Y

    def mlp_gradient(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(N_PARAM, dtype=np.float32)
        loss = C.c_double()
        self._check(self.lib.stk_mlp_gradient(self.h, X.shape[0], _f(X), _f(T), _f(g), C.byref(loss)))
        return loss.value, g

    def local_schur(self):
        S = np.zeros((self.nel, 6))
        X = np.zeros((self.nel, 5), dtype=np.float32)
        self._check(self.lib.stk_local_schur(self.h, _d(S), _f(X)))
        return S, X

    def learned_setup(self):
        self._check(self.lib.stk_learned_setup(self.h))

    def updatevelocity(self):
        U = np.zeros(self.nU)
        Uc = np.zeros(self.nU)
        w = C.c_double()
        self._check(self.lib.stk_updatevelocity(self.h, _d(U), _d(Uc), C.byref(w)))
        return U, Uc, w.value

    def set_stress_fields(self, eta_ip, C6=None, nip=7):
        """eta.ip and the tensor arrays (nel x ncol, ncol >= nip) at the nip-point stress rule."""
        eta = np.asfortranarray(eta_ip, dtype=np.float64)
        if eta.ndim != 2 or eta.shape[0] != self.nel or eta.shape[1] < nip:
            raise ValueError(f"eta.ip must be nel x >= {nip}")
        self._sfields = [eta]
        arr = None
        if C6 is not None:
            comps = [np.asfortranarray(c, dtype=np.float64) for c in C6]
            for c in comps:
                if c.shape != eta.shape:
                    raise ValueError("tensor components must have the shape of eta.ip")
            self._sfields += comps
            arr = (dp * 6)(*[_d(c) for c in comps])
        self._check(self.lib.stk_set_stress_fields(self.h, _d(eta), int(nip), arr))
        self.s_nip = int(nip)

    def stress(self, Ucart=None):
        """stress!: dict of the eight nel x nip arrays (Fortran order), see STRESS_FIELDS."""
        outs = {k: np.zeros((self.nel, self.s_nip), order="F") for k in STRESS_FIELDS}
        arr = (dp * 8)(*[_d(outs[k]) for k in STRESS_FIELDS])
        u = None if Ucart is None else np.ascontiguousarray(Ucart, dtype=np.float64)
        self._check(self.lib.stk_stress(self.h, None if u is None else _d(u), arr))
        return outs

    def thermal_solve(self, T0, rho, K, Cp, dt, fixed_nodes_1based, fixed_values, dQdT=0.0, rtol=0.0):
        """One thermal diffusion step; returns (T, report)."""
        T = np.array(T0, dtype=np.float64, copy=True)
        r = np.ascontiguousarray(rho, dtype=np.float64)
        if T.shape != (self.nnod,) or r.shape != (self.nnod,):
            raise ValueError("T0 and rho must hold one value per node")
        fx = np.ascontiguousarray(fixed_nodes_1based, dtype=np.int64)
        fv = np.ascontiguousarray(fixed_values, dtype=np.float64)
        rep = StkReport()
        rc = self.lib.stk_thermal_solve(self.h, _d(T), _d(r), float(K), float(Cp), float(dQdT), float(dt), _i(fx),
                                        _d(fv), fx.size, float(rtol), C.byref(rep))
        self._check(rc)
        return T, rep

    def stats(self, T, U=None):
        """write_stats: (Urms, <T>, Nusselt top, Nusselt bottom)."""
        t = np.ascontiguousarray(T, dtype=np.float64)
        if t.shape != (self.nnod,):
            raise ValueError("T must hold one value per node")
        u = None if U is None else np.ascontiguousarray(U, dtype=np.float64)
        out = np.zeros(4)
        self._check(self.lib.stk_stats(self.h, None if u is None else _d(u), _d(t), _d(out)))
        return tuple(out)

    def deformation_gradient(self, F, dt, U=None):
        """One RK4 step of F (nel x nip x 2 x 2 as float64[nel, nip, 2, 2] with F[e, q] the 2x2 matrix);
        returns the updated copy."""
        F = np.asarray(F, dtype=np.float64)
        if F.shape != (self.nel, self.s_nip, 2, 2):
            raise ValueError("F must be nel x nip x 2 x 2")
        # Julia Matrix{SMatrix{2,2}}: points column-major, each 2x2 column-major
        buf = np.ascontiguousarray(np.transpose(F, (1, 0, 3, 2))).reshape(-1)
        u = None if U is None else np.ascontiguousarray(U, dtype=np.float64)
        self._check(self.lib.stk_deformation_gradient(self.h, None if u is None else _d(u), float(dt), _d(buf)))
        return np.transpose(buf.reshape(self.s_nip, self.nel, 2, 2), (1, 0, 3, 2)).copy()

    def fse(self, F, mode=FSE_FABRIC_DESTRUCTION, eps=1e3):
        """getFSE on F (float64[nel, nip, 2, 2], nip 6 or 7).  Returns (ellipses float64[nel, nip, 6]
        with the fields x1, x2, y1, y2, a1, a2, F after the healing rule)."""
        F = np.asarray(F, dtype=np.float64)
        if F.ndim != 4 or F.shape[0] != self.nel or F.shape[2:] != (2, 2) or F.shape[1] not in (6, 7):
            raise ValueError("F must be nel x nip x 2 x 2 with nip 6 or 7")
        nip = F.shape[1]
        buf = np.ascontiguousarray(np.transpose(F, (1, 0, 3, 2))).reshape(-1)
        out = np.zeros(6 * self.nel * nip)
        self._check(self.lib.stk_fse(self.h, _d(buf), nip, int(mode), float(eps), _d(out)))
        self.fse_nip = nip
        ell = np.transpose(out.reshape(nip, self.nel, 6), (1, 0, 2)).copy()
        return ell, np.transpose(buf.reshape(nip, self.nel, 2, 2), (1, 0, 3, 2)).copy()

    def set_dem(self, tensor, axes, w):
        """getDEM from the DEMloader arrays: tensor n x 9, axes n x 3; w the viscosity cut-off."""
        t = np.asfortranarray(tensor, dtype=np.float64)
        a = np.asfortranarray(axes, dtype=np.float64)
        if t.ndim != 2 or t.shape[1] != 9 or a.shape != (t.shape[0], 3):
            raise ValueError("tensor must be n x 9 and axes n x 3")
        self._check(self.lib.stk_set_dem(self.h, _d(t), _d(a), t.shape[0], float(w)))

    def anisotropic_tensor(self, fse=None, nip=None, shape_table=None, shape_points=None, install=0):
        """anisotropic_tensor: the six nel x 7 arrays (eta11, eta33, eta55, eta13, eta15, eta35).

        fse: float64[nel, nip, 6] or None for the device ellipses of the last ``fse`` call.
        shape_table = (a1 nodes, r1, r2) or shape_points = (r1, r2), each float64[nel, nip]."""
        e = None
        if fse is not None:
            fse = np.asarray(fse, dtype=np.float64)
            if fse.ndim != 3 or fse.shape[0] != self.nel or fse.shape[2] != 6:
                raise ValueError("fse must be nel x nip x 6")
            nip = fse.shape[1]
            e = np.ascontiguousarray(np.transpose(fse, (1, 0, 2))).reshape(-1)
        elif nip is None:
            nip = getattr(self, "fse_nip", 0)
        keep = []
        if shape_table is not None:
            nodes, r1, r2 = (np.ascontiguousarray(v, dtype=np.float64) for v in shape_table)
            keep += [nodes, r1, r2]
            args = (_d(nodes), _d(r1), _d(r2), nodes.size)
        elif shape_points is not None:
            r1, r2 = (np.asfortranarray(v, dtype=np.float64).ravel(order="F") for v in shape_points)
            keep += [r1, r2]
            args = (None, _d(r1), _d(r2), r1.size)
        else:
            raise ValueError("give shape_table or shape_points")
        outs = [np.zeros((self.nel, 7), order="F") for _ in range(6)]
        arr = (dp * 6)(*[_d(o) for o in outs])
        self._check(self.lib.stk_anisotropic_tensor(self.h, None if e is None else _d(e), int(nip), *args,
                                                    int(install), arr))
        return outs

    def colouring(self, which=COLOUR_ELEMENTS):
        """(colours, number of colours): per element (COLOUR_ELEMENTS) or per node (COLOUR_NODES)."""
        n = self.nel if which == COLOUR_ELEMENTS else self.nnod
        col = np.zeros(n, dtype=np.int32)
        nc = C.c_int32()
        self._check(self.lib.stk_colouring(self.h, int(which), col.ctypes.data_as(C.POINTER(C.c_int32)),
                                           C.byref(nc)))
        return col, nc.value
