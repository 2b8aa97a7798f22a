"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, as the
checker and the CPU baseline.  Product code never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libpbs_oracle.so")

REF_GPU, REF_CPU = 0, 1
CLAMP_NONE, CLAMP_GPU, CLAMP_CPU = 0, 1, 2
MAT_NEOHOOKEAN, MAT_STRETCH = 0, 1


class Params(C.Structure):
    _fields_ = [
        ("dt", C.c_float), ("grav", C.c_float * 3), ("alpha0", C.c_float),
        ("alpha_decay", C.c_float), ("alpha_min", C.c_float), ("hd_refresh", C.c_int),
        ("energy_check", C.c_int), ("damping", C.c_float), ("init_momentum", C.c_float),
        ("profile_n", C.c_int), ("profile_k", C.c_int * 8), ("profile_rho", C.c_float * 8),
        ("semantics", C.c_int), ("iterations", C.c_int), ("relgrad_tol", C.c_double),
        ("trace_cap", C.c_int), ("gather_order", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("passes", C.c_int), ("iterations", C.c_int), ("rollbacks", C.c_int), ("converged", C.c_int),
        ("alpha", C.c_float), ("energy", C.c_double), ("g0_norm", C.c_double), ("g_norm", C.c_double),
        ("used_learned", C.c_int),
    ]


class Trace(C.Structure):
    _fields_ = [
        ("k", C.c_int), ("rollback", C.c_int), ("hd_refresh", C.c_int), ("checked", C.c_int),
        ("alpha", C.c_float), ("omega", C.c_float), ("energy", C.c_double), ("g_norm", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_create.restype = P
        L.orc_create.argtypes = [C.c_int, P, C.c_int, P, P, C.c_float, C.c_float, C.c_float, C.c_float,
                                 C.c_int, C.c_int, C.POINTER(Params)]
        L.orc_destroy.argtypes = [P]
        L.orc_set_params.argtypes = [P, C.POINTER(Params)]
        L.orc_set_threads.argtypes = [P, C.c_int]
        L.orc_set_state.argtypes = [P, P, P, P]
        L.orc_get_state.argtypes = [P, P, P, P]
        L.orc_step.argtypes = [P, P, P, C.POINTER(Stats), P]
        L.orc_step.restype = C.c_int
        L.orc_gradient_norm.argtypes = [P, P]
        L.orc_gradient_norm.restype = C.c_double
        L.orc_default_params.argtypes = [C.POINTER(Params)]
        L.orc_lame.argtypes = [C.c_float, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_precompute.argtypes = [C.c_int, P, C.c_int, P, P, C.c_float, P, C.c_float, P, P, P, P, P]
        for sfx in ("f32", "f64"):
            getattr(L, "orc_tet_forces_" + sfx).argtypes = [C.c_int, P, P, P, P, P, C.c_float, C.c_float,
                                                             C.c_int, C.c_int, P, P]
            getattr(L, "orc_svd3_" + sfx).argtypes = [P, P, P, P]
            getattr(L, "orc_dpdf_" + sfx).argtypes = [P, C.c_float, C.c_float, C.c_int, P]
            getattr(L, "orc_hessian_diag_" + sfx).argtypes = [C.c_int, C.c_int, P, P, P, P, C.c_float,
                                                               C.c_float, C.c_int, P, P]
        L.orc_csr_pattern.argtypes = [C.c_int, C.c_int, P, P, P, P, P]
        L.orc_csr_pattern.restype = C.c_long
        L.orc_incidence.argtypes = [C.c_int, C.c_int, P, P, P]
        for sfx in ("f32", "f64"):
            getattr(L, "orc_pd_rotation_" + sfx).argtypes = [P, P]
            getattr(L, "orc_pd_project_" + sfx).argtypes = [C.c_int, C.c_int, P, P, P, P, C.c_float, P, P]
            getattr(L, "orc_pd_matrix_" + sfx).argtypes = [C.c_int, C.c_int, P, P, P, P, C.c_float, C.c_float,
                                                           C.c_long, P, P, P]
        for sfx in ("f32", "f64"):
            getattr(L, "orc_implicit_system_" + sfx).argtypes = [
                C.c_int, P, C.c_int, P, P, P, P, P, P, C.c_float, C.c_float, C.c_int, P, P, P,
                C.c_float, C.c_float, C.c_float, C.c_long, P, P, P, P, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def default_params(**kw):
    p = Params()
    lib().orc_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "grav":
            for i in range(3):
                p.grav[i] = v[i]
        elif k == "profile":
            p.profile_n = len(v)
            for i, (kk, rr) in enumerate(v):
                p.profile_k[i] = kk
                p.profile_rho[i] = rr
        else:
            setattr(p, k, v)
    return p


def lame(E, nu):
    mu, lam = C.c_float(), C.c_float()
    lib().orc_lame(E, nu, C.byref(mu), C.byref(lam))
    return mu.value, lam.value


def precompute(rest, tets, mask, density=1000.0, grav=(0.0, -9.8, 0.0), fixed_mass=1e8):
    rest = np.ascontiguousarray(rest, np.float32).reshape(-1, 3)
    tets = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
    mask = np.ascontiguousarray(mask, np.int32)
    nn, nt = len(rest), len(tets)
    dminv = np.zeros((nt, 9), np.float32)
    vol = np.zeros(nt, np.float32)
    an = np.zeros((nt, 9), np.float32)
    mass = np.zeros(nn, np.float32)
    gf = np.zeros((nn, 3), np.float32)
    g = np.asarray(grav, np.float32)
    lib().orc_precompute(nn, _p(rest), nt, _p(tets), _p(mask), density, _p(g), fixed_mass,
                         _p(dminv), _p(vol), _p(an), _p(mass), _p(gf))
    return dict(dminv=dminv, vol=vol, an=an, mass=mass, gforce=gf)


def tet_forces(x, tets, pre, mu, lam, material=MAT_NEOHOOKEAN, clamp=CLAMP_GPU, precision=32):
    dt = np.float32 if precision == 32 else np.float64
    x = np.ascontiguousarray(x, dt).reshape(-1, 3)
    tets = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
    nt = len(tets)
    f = np.zeros((nt, 12), dt)
    e = np.zeros(nt, dt)
    fn = lib().orc_tet_forces_f32 if precision == 32 else lib().orc_tet_forces_f64
    fn(nt, _p(tets), _p(x), _p(pre["dminv"]), _p(pre["an"]), _p(pre["vol"]), mu, lam, material, clamp,
       _p(f), _p(e))
    return f, e


def svd3(F, precision=64):
    dt = np.float32 if precision == 32 else np.float64
    F = np.ascontiguousarray(F, dt).reshape(9)
    U = np.zeros(9, dt)
    S = np.zeros(3, dt)
    V = np.zeros(9, dt)
    (lib().orc_svd3_f32 if precision == 32 else lib().orc_svd3_f64)(_p(F), _p(U), _p(S), _p(V))
    return U.reshape(3, 3), S, V.reshape(3, 3)


def dpdf(F, mu, lam, material=MAT_NEOHOOKEAN, precision=64):
    dt = np.float32 if precision == 32 else np.float64
    F = np.ascontiguousarray(F, dt).reshape(9)
    out = np.zeros(81, dt)
    (lib().orc_dpdf_f32 if precision == 32 else lib().orc_dpdf_f64)(_p(F), mu, lam, material, _p(out))
    return out.reshape(9, 9)


def hessian_diag(x, tets, pre, mu, lam, material=MAT_NEOHOOKEAN, precision=32):
    dt = np.float32 if precision == 32 else np.float64
    x = np.ascontiguousarray(x, dt).reshape(-1, 3)
    tets = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
    nn, nt = len(x), len(tets)
    hd = np.zeros((nn, 3), dt)
    th = np.zeros((nt, 12), dt)
    fn = lib().orc_hessian_diag_f32 if precision == 32 else lib().orc_hessian_diag_f64
    fn(nn, nt, _p(tets), _p(x), _p(pre["dminv"]), _p(pre["an"]), mu, lam, material, _p(hd), _p(th))
    return hd, th


def csr_pattern(nn, tets):
    tets = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
    nt = len(tets)
    nnz = lib().orc_csr_pattern(nn, nt, _p(tets), None, None, None, None)
    rowptr = np.zeros(3 * nn + 1, np.int32)
    col = np.zeros(nnz, np.int32)
    diag = np.zeros(3 * nn, np.int32)
    kid = np.zeros(48 * nt, np.int32)
    lib().orc_csr_pattern(nn, nt, _p(tets), _p(rowptr), _p(col), _p(diag), _p(kid))
    return rowptr, col, diag, kid


def incidence(nn, tets):
    tets = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
    off = np.zeros(nn + 1, np.int32)
    ent = np.zeros(4 * len(tets), np.int32)
    lib().orc_incidence(nn, len(tets), _p(tets), _p(off), _p(ent))
    return off, ent


def implicit_system(rest, tets, mask, pre, mu, lam, x, v, h=0.03, alpha=0.02, beta=0.02, ext=None,
                    material=MAT_NEOHOOKEAN, precision=64):
    """Backward-Euler system of the reference's GPU implicit path: (rowptr, col, K values,
    system-matrix values, right-hand side), CSR in the reference's layout."""
    dt = np.float32 if precision == 32 else np.float64
    rest = np.ascontiguousarray(rest, np.float32).reshape(-1, 3)
    tets = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
    mask = np.ascontiguousarray(mask, np.int32)
    nn, nt = len(rest), len(tets)
    rowptr, col, diag, kid = csr_pattern(nn, tets)
    nnz = len(col)
    x = np.ascontiguousarray(x, dt).reshape(-1)
    v = np.ascontiguousarray(v, dt).reshape(-1)
    e = None if ext is None else np.ascontiguousarray(ext, dt).reshape(-1)
    kval = np.zeros(nnz)
    aval = np.zeros(nnz)
    b = np.zeros(3 * nn)
    fn = lib().orc_implicit_system_f32 if precision == 32 else lib().orc_implicit_system_f64
    fn(nn, _p(rest), nt, _p(tets), _p(mask), _p(pre["dminv"]), _p(pre["an"]), _p(pre["mass"]),
       _p(np.ascontiguousarray(pre["gforce"], np.float32)), mu, lam, material, _p(x), _p(v), _p(e), h, alpha, beta,
       nnz, _p(kid), _p(diag), _p(kval), _p(aval), _p(b))
    return rowptr, col, kval, aval, b


def pd_rotation(F, precision=64):
    dt = np.float32 if precision == 32 else np.float64
    F = np.ascontiguousarray(F, dt).reshape(9)
    out = np.zeros(9, dt)
    (lib().orc_pd_rotation_f32 if precision == 32 else lib().orc_pd_rotation_f64)(_p(F), _p(out))
    return out.reshape(3, 3)


def pd_project(x, tets, pre, stiffness=1e6, s=None, precision=64):
    dt = np.float32 if precision == 32 else np.float64
    x = np.ascontiguousarray(x, dt).reshape(-1, 3)
    tets = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
    sa = None if s is None else np.ascontiguousarray(s, dt).reshape(-1)
    out = np.zeros_like(x)
    fn = lib().orc_pd_project_f32 if precision == 32 else lib().orc_pd_project_f64
    fn(len(x), len(tets), _p(tets), _p(x), _p(pre["dminv"]), _p(pre["vol"]), stiffness, _p(sa), _p(out))
    return out


def pd_matrix(nn, tets, pre, stiffness=1e6, t=0.03, precision=64):
    """(rowptr, col, values) of the projective-dynamics global matrix."""
    tets = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
    rowptr, col, diag, kid = csr_pattern(nn, tets)
    g = np.zeros(len(col))
    fn = lib().orc_pd_matrix_f32 if precision == 32 else lib().orc_pd_matrix_f64
    fn(nn, len(tets), _p(tets), _p(pre["dminv"]), _p(pre["vol"]), _p(pre["mass"]), stiffness, t, len(col), _p(kid),
       _p(diag), _p(g))
    return rowptr, col, g


class Solver:
    """Oracle descent solver: REF_GPU / REF_CPU semantics, FP32 or FP64 state."""

    def __init__(self, rest, tets, mask, E=1e5, nu=0.45, density=1000.0, fixed_mass=1e8,
                 material=MAT_NEOHOOKEAN, precision=32, params=None, threads=1):
        self.rest = np.ascontiguousarray(rest, np.float32).reshape(-1, 3)
        self.tets = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
        self.mask = np.ascontiguousarray(mask, np.int32)
        self.nn, self.nt = len(self.rest), len(self.tets)
        self.params = params if params is not None else default_params()
        self.h = lib().orc_create(self.nn, _p(self.rest), self.nt, _p(self.tets), _p(self.mask), E, nu,
                                  density, fixed_mass, material, precision, C.byref(self.params))
        lib().orc_set_threads(self.h, threads)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set_params(self, params):
        self.params = params
        lib().orc_set_params(self.h, C.byref(params))

    def set_state(self, x=None, v=None, v_last=None):
        a = [None if z is None else np.ascontiguousarray(z, np.float64).reshape(-1) for z in (x, v, v_last)]
        lib().orc_set_state(self.h, *[_p(z) for z in a])

    def state(self):
        x = np.zeros((self.nn, 3))
        v = np.zeros((self.nn, 3))
        vl = np.zeros((self.nn, 3))
        lib().orc_get_state(self.h, _p(x), _p(v), _p(vl))
        return x, v, vl

    def step(self, ext=None, x0=None, trace=False):
        st = Stats()
        tr = None
        if trace:
            tr = (Trace * max(1, self.params.trace_cap))()
        e = None if ext is None else np.ascontiguousarray(ext, np.float64).reshape(-1)
        x0a = None if x0 is None else np.ascontiguousarray(x0, np.float64).reshape(-1)
        lib().orc_step(self.h, _p(e), _p(x0a), C.byref(st), C.cast(tr, C.c_void_p) if tr is not None else None)
        if trace:
            return st, [tr[i] for i in range(min(st.passes, self.params.trace_cap))]
        return st

    def gradient_norm(self, x):
        x = np.ascontiguousarray(x, np.float64).reshape(-1)
        return lib().orc_gradient_norm(self.h, _p(x))
