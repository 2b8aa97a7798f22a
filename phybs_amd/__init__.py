"""phybs_amd -- Python binding of the MI355X descent solver's C ABI (include/phybs.h).

The solver itself is libphybs.so (HIP kernels for gfx950 + the C ABI).  This module only
marshals numpy arrays through ctypes; there is no Python or CPU fallback: if the native
library is missing, importing the binding raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libphybs.so")

OK = 0
ERR_ARG, ERR_HIP, ERR_MESH, ERR_IO, ERR_STATE = -1, -2, -3, -4, -5
SEM_REF_GPU, SEM_REF_CPU = 0, 1
FP32, FP64 = 32, 64
MAT_NEOHOOKEAN, MAT_STRETCH = 0, 1


class PhybsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"phybs error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("dt", C.c_float), ("grav", C.c_float * 3), ("youngs", C.c_float), ("nu", C.c_float),
        ("density", C.c_float), ("fixed_mass", C.c_float), ("alpha0", C.c_float),
        ("alpha_decay", C.c_float), ("alpha_min", C.c_float), ("hd_refresh", C.c_int),
        ("energy_check", C.c_int), ("damping", C.c_float), ("init_momentum", C.c_float),
        ("profile_n", C.c_int), ("profile_k", C.c_int * 8), ("profile_rho", C.c_float * 8),
        ("semantics", C.c_int), ("precision", C.c_int), ("material", C.c_int),
        ("iterations", C.c_int), ("relgrad_tol", C.c_double), ("max_iterations", C.c_int),
        ("block_size", C.c_int), ("tile_tets", C.c_int), ("graph_passes", C.c_int), ("mixed_switch", C.c_double),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("passes", C.c_int), ("iterations", C.c_int), ("rollbacks", C.c_int), ("converged", C.c_int),
        ("used_learned", C.c_int), ("alpha", C.c_float), ("energy", C.c_double), ("g0_norm", C.c_double),
        ("g_norm", C.c_double), ("gpu_ms", C.c_double), ("g_learned", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ImplicitParams(C.Structure):
    _fields_ = [
        ("damping_alpha", C.c_float), ("damping_beta", C.c_float), ("cg_tolerance", C.c_double),
        ("cg_max_iterations", C.c_int), ("warm_start", C.c_int),
    ]


class ImplicitStats(C.Structure):
    _fields_ = [
        ("cg_iterations", C.c_int), ("converged", C.c_int), ("residual0", C.c_double), ("residual", C.c_double),
        ("rhs_norm", C.c_double), ("gpu_ms", C.c_float),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PdParams(C.Structure):
    _fields_ = [
        ("stiffness", C.c_float), ("iterations", C.c_int), ("damping", C.c_float), ("cg_tolerance", C.c_double),
        ("cg_max_iterations", C.c_int),
    ]


class PdStats(C.Structure):
    _fields_ = [("cg_iterations", C.c_int), ("converged", C.c_int), ("gpu_ms", C.c_float)]


def pd_params(**kw):
    p = PdParams()
    _check(lib().pbs_pd_default_params(C.byref(p)))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def implicit_params(**kw):
    p = ImplicitParams()
    _check(lib().pbs_implicit_default_params(C.byref(p)))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


# name -> (restype, argtypes); every symbol include/phybs.h declares.
_P = C.c_void_p
SIGNATURES = {
    "pbs_default_params": (C.c_int, [_P]),
    "pbs_last_error": (C.c_char_p, []),
    "pbs_abi_version": (C.c_int, []),
    "pbs_create": (C.c_int, [C.c_int, _P, _P, _P, C.c_int, _P, _P, C.POINTER(_P)]),
    "pbs_create_batch": (C.c_int, [C.c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.POINTER(_P)]),
    "pbs_create_batch_tiled": (C.c_int, [C.c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.POINTER(_P)]),
    "pbs_num_instances": (C.c_int, [_P]),
    "pbs_instance_stats": (C.c_int, [_P, C.c_int, _P]),
    "pbs_autotune_report": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "pbs_instance_tile_tets": (C.c_int, [_P, _P]),
    "pbs_state_save": (C.c_int, [_P, C.c_char_p]),
    "pbs_state_load": (C.c_int, [_P, C.c_char_p]),
    "pbs_destroy": (None, [_P]),
    "pbs_set_params": (C.c_int, [_P, _P]),
    "pbs_get_params": (C.c_int, [_P, _P]),
    "pbs_sizes": (C.c_int, [_P, _P, _P]),
    "pbs_get_rest_data": (C.c_int, [_P, _P, _P, _P, _P]),
    "pbs_reset": (C.c_int, [_P]),
    "pbs_set_positions": (C.c_int, [_P, _P]),
    "pbs_get_positions": (C.c_int, [_P, _P]),
    "pbs_get_positions_f64": (C.c_int, [_P, _P]),
    "pbs_set_velocities": (C.c_int, [_P, _P, _P]),
    "pbs_get_velocities": (C.c_int, [_P, _P, _P]),
    "pbs_step": (C.c_int, [_P, _P, _P, _P]),
    "pbs_do_descent_opt": (C.c_int, [_P, C.c_float, C.c_int, _P, _P]),
    "pbs_compute_forces": (C.c_int, [_P, _P, _P, _P, _P]),
    "pbs_compute_hessian_diag": (C.c_int, [_P, _P, _P, _P]),
    "pbs_tile_forces": (C.c_int, [_P, _P, _P, _P, _P]),
    "pbs_kernel_times": (C.c_int, [_P, C.c_int, _P, _P, _P, _P]),
    "pbs_reset_kernel_times": (C.c_int, [_P]),
    "pbs_set_profiling": (C.c_int, [_P, C.c_int]),
    "pbs_mesh_bar": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_float, _P, _P]),
    "pbs_mesh_twist": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_float, C.c_double, C.c_double, C.c_uint64,
                                 C.c_int, _P]),
    "pbs_tetgen_count": (C.c_int, [C.c_char_p, _P, _P, _P]),
    "pbs_tetgen_load": (C.c_int, [C.c_char_p, C.c_float, _P, _P, _P]),
    "pbs_tetgen_read_nodes": (C.c_int, [C.c_char_p, C.c_float, _P, _P]),
    "pbs_tetgen_read_tets": (C.c_int, [C.c_char_p, _P, _P]),
    "pbs_tetgen_read_faces": (C.c_int, [C.c_char_p, _P, _P]),
    "pbs_tet_stiffness": (C.c_int, [_P, _P, C.c_int, _P, _P]),
    "pbs_tet_stiffness_noalloc": (C.c_int, [_P, _P, C.c_int, _P, _P]),
    "pbs_node_write": (C.c_int, [C.c_char_p, C.c_int, _P]),
    "pbs_node_read": (C.c_int, [C.c_char_p, C.c_int, _P]),
    "pbs_csr_pattern": (C.c_int64, [C.c_int, C.c_int, _P, _P, _P, _P, _P]),
    "pbs_incidence": (C.c_int, [C.c_int, C.c_int, _P, _P, _P]),
    "pbs_implicit_default_params": (C.c_int, [_P]),
    "pbs_implicit_step": (C.c_int, [_P, _P, _P, _P, _P]),
    "pbs_stiffness_matrix": (C.c_int, [_P, C.c_int64, _P]),
    "pbs_pd_default_params": (C.c_int, [_P]),
    "pbs_pd_step": (C.c_int, [_P, _P, _P, _P, _P]),
    "pbs_pd_project": (C.c_int, [_P, C.c_float, _P, _P]),
    "pbs_face_ring": (C.c_int64, [C.c_int, C.c_int, _P, _P, _P]),
    "pbs_pick_face": (C.c_int, [C.c_int, _P, C.c_int, _P, _P, _P, _P]),
    "pbs_drag_face_ring": (C.c_int, [C.c_int, _P, _P, C.c_int, _P, _P, _P, C.c_int, _P, _P]),
    "pbs_learn_set_model": (C.c_int, [_P, _P, _P]),
    "pbs_learn_get_model": (C.c_int, [_P, _P, _P]),
    "pbs_learn_init_model": (C.c_int, [_P, C.c_uint64]),
    "pbs_learn_enable": (C.c_int, [_P, C.c_int]),
    "pbs_learn_set_options": (C.c_int, [_P, C.c_int, C.c_int, _P]),
    "pbs_learn_collect": (C.c_int, [_P, C.c_int64]),
    "pbs_learn_collect_every": (C.c_int, [_P, C.c_int64, C.c_int]),
    "pbs_learn_add_samples": (C.c_int, [_P, C.c_int64, _P, _P]),
    "pbs_learn_samples": (C.c_int64, [_P]),
    "pbs_learn_get_samples": (C.c_int, [_P, C.c_int64, C.c_int64, _P, _P]),
    "pbs_learn_fit_scales": (C.c_int, [_P]),
    "pbs_learn_train": (C.c_int, [_P, C.c_int, C.c_int, C.c_float, C.c_uint64, _P]),
    "pbs_learn_forward": (C.c_int, [_P, C.c_int, _P, _P]),
    "pbs_learn_loss_grad": (C.c_int, [_P, C.c_int, _P, _P, _P, _P]),
    "pbs_learn_features": (C.c_int, [_P, _P]),
    "pbs_learn_guess": (C.c_int, [_P, _P, _P]),
}
MLP_PARAMS, MLP_SCALES = 9731, 7

_lib = None


def lib():
    """Load libphybs.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PhybsError(ERR_STATE, f"{LIB_PATH} not built: run __graft_entry__.build() or make -C phybs_amd/csrc")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc < 0:
        raise PhybsError(rc, lib().pbs_last_error().decode())
    return rc


def default_params(**kw):
    p = Params()
    _check(lib().pbs_default_params(C.byref(p)))
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
            if not hasattr(p, k):
                raise AttributeError(k)
            setattr(p, k, v)
    return p


def _f32(a, n=None):
    a = np.ascontiguousarray(a, np.float32)
    return a.reshape(-1) if n is None else a.reshape(n)


# ---- host mesh utilities ----
def face_ring(nn, faces):
    faces = np.ascontiguousarray(faces, np.int32).reshape(-1, 3)
    total = _check(lib().pbs_face_ring(nn, len(faces), _p(faces), None, None))
    off = np.zeros(len(faces) + 1, np.int64)
    ring = np.zeros(total, np.int32)
    _check(lib().pbs_face_ring(nn, len(faces), _p(faces), _p(off), _p(ring)))
    return off, ring


def pick_face(nodes, faces, origin, direction):
    nodes = _f32(nodes, (-1, 3))
    faces = np.ascontiguousarray(faces, np.int32).reshape(-1, 3)
    o = _f32(origin)
    d = _f32(direction)
    out = C.c_int(-1)
    _check(lib().pbs_pick_face(len(nodes), _p(nodes), len(faces), _p(faces), _p(o), _p(d), C.byref(out)))
    return out.value


def drag_face_ring(nodes, mass, faces, ring_offsets, ring, face, drag):
    nodes = _f32(nodes, (-1, 3))
    mass = _f32(mass)
    faces = np.ascontiguousarray(faces, np.int32).reshape(-1, 3)
    off = np.ascontiguousarray(ring_offsets, np.int64)
    ring = np.ascontiguousarray(ring, np.int32)
    ext = np.zeros_like(nodes)
    _check(lib().pbs_drag_face_ring(len(nodes), _p(nodes), _p(mass), len(faces), _p(faces), _p(off), _p(ring), face,
                                    _p(_f32(drag)), _p(ext)))
    return ext
def mesh_bar(nx, ny, nz, edge=0.05):
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    nodes = np.zeros((nn, 3), np.float32)
    tets = np.zeros((6 * nx * ny * nz, 4), np.int32)
    _check(lib().pbs_mesh_bar(nx, ny, nz, edge, _p(nodes), _p(tets)))
    return nodes, tets


def mesh_twist(nodes, nx, ny, nz, edge=0.05, twist_deg=90.0, jitter=0.01, seed=1):
    x = np.array(nodes, np.float32, copy=True, order="C")
    _check(lib().pbs_mesh_twist(nx, ny, nz, edge, twist_deg, jitter, seed, len(x), _p(x)))
    return x


def tetgen_load(prefix, scale=1.0, faces=False):
    nn, nt, nf = C.c_int(), C.c_int(), C.c_int()
    _check(lib().pbs_tetgen_count(prefix.encode(), C.byref(nn), C.byref(nt), C.byref(nf)))
    nodes = np.zeros((nn.value, 3), np.float32)
    tets = np.zeros((nt.value, 4), np.int32)
    fc = np.zeros((nf.value, 3), np.int32) if faces else None
    _check(lib().pbs_tetgen_load(prefix.encode(), scale, _p(nodes), _p(tets), _p(fc)))
    return (nodes, tets, fc) if faces else (nodes, tets)


def node_write(path, nodes):
    nodes = _f32(nodes)
    _check(lib().pbs_node_write(path.encode(), len(nodes) // 3, _p(nodes)))


def node_read(path, nn):
    out = np.zeros((nn, 3), np.float32)
    _check(lib().pbs_node_read(path.encode(), nn, _p(out)))
    return out


def csr_pattern(nn, tets):
    tets = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
    nnz = _check(lib().pbs_csr_pattern(nn, len(tets), _p(tets), None, None, None, None))
    rowptr = np.zeros(3 * nn + 1, np.int32)
    col = np.zeros(nnz, np.int32)
    diag = np.zeros(3 * nn, np.int32)
    kid = np.zeros(48 * len(tets), np.int32)
    _check(lib().pbs_csr_pattern(nn, len(tets), _p(tets), _p(rowptr), _p(col), _p(diag), _p(kid)))
    return rowptr, col, diag, kid


def incidence(nn, tets):
    tets = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
    off = np.zeros(nn + 1, np.int32)
    ent = np.zeros(4 * len(tets), np.int32)
    _check(lib().pbs_incidence(nn, len(tets), _p(tets), _p(off), _p(ent)))
    return off, ent


class Solver:
    """GPU descent solver handle (pbs_create ... pbs_destroy)."""

    def __init__(self, rest, tets, mask=None, params=None, nodes=None):
        self.rest = _f32(rest, (-1, 3))
        self.tets = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
        self.nn, self.nt = len(self.rest), len(self.tets)
        self.mask = None if mask is None else np.ascontiguousarray(mask, np.int32)
        self.params = params if params is not None else default_params()
        h = C.c_void_p()
        x0 = None if nodes is None else _f32(nodes, (-1, 3))
        _check(lib().pbs_create(self.nn, _p(x0), _p(self.rest), _p(self.mask), self.nt, _p(self.tets),
                                C.byref(self.params), C.byref(h)))
        self.h = h

    @classmethod
    def batch(cls, meshes, youngs, nu, params=None, nodes=None, tile_tets=None):
        """Batched solver over independent meshes: meshes = [(rest, tets, mask), ...] with tets
        indexing each mesh's own nodes; youngs / nu per mesh.  Arrays of the solver (positions,
        velocities, ...) are the meshes' arrays concatenated in order.  tile_tets: optional tile
        size per mesh (else params.tile_tets for all, -1 = autotuned)."""
        self = cls.__new__(cls)
        rests, tl, masks, nns, nts = [], [], [], [], []
        off = 0
        for rest, tets, mask in meshes:
            r = _f32(rest, (-1, 3))
            t = np.ascontiguousarray(tets, np.int32).reshape(-1, 4) + off
            m = np.zeros(len(r), np.int32) if mask is None else np.ascontiguousarray(mask, np.int32)
            rests.append(r)
            tl.append(t)
            masks.append(m)
            nns.append(len(r))
            nts.append(len(t))
            off += len(r)
        self.rest = np.concatenate(rests)
        self.tets = np.concatenate(tl)
        self.mask = np.concatenate(masks)
        self.nn, self.nt = len(self.rest), len(self.tets)
        self.inst_nodes = np.array(nns, np.int32)
        self.inst_tets = np.array(nts, np.int32)
        self.params = params if params is not None else default_params()
        E = np.ascontiguousarray(youngs, np.float32)
        v = np.ascontiguousarray(nu, np.float32)
        x0 = None if nodes is None else _f32(nodes, (-1, 3))
        h = C.c_void_p()
        if tile_tets is None:
            _check(lib().pbs_create_batch(len(meshes), _p(self.inst_nodes), _p(self.inst_tets), _p(x0),
                                          _p(self.rest), _p(self.mask), _p(self.tets), _p(E), _p(v),
                                          C.byref(self.params), C.byref(h)))
        else:
            tt = np.ascontiguousarray(tile_tets, np.int32)
            if tt.shape != (len(meshes),):
                raise ValueError("tile_tets: one tile size per mesh")
            _check(lib().pbs_create_batch_tiled(len(meshes), _p(self.inst_nodes), _p(self.inst_tets), _p(x0),
                                                _p(self.rest), _p(self.mask), _p(self.tets), _p(E), _p(v), _p(tt),
                                                C.byref(self.params), C.byref(h)))
        self.h = h
        return self

    def instance_stats(self, i):
        st = StepStats()
        _check(lib().pbs_instance_stats(self.h, i, C.byref(st)))
        return st

    def autotune_report(self):
        n = 16
        tiles = np.zeros(n, np.int32)
        us = np.zeros(n)
        c = C.c_int()
        _check(lib().pbs_autotune_report(self.h, n, _p(tiles), _p(us), C.byref(c)))
        return {(int(t) if t > 0 else "per_instance"): float(u) for t, u in zip(tiles[:c.value], us[:c.value])}

    def save_state(self, path):
        """Binary snapshot of (x, v, v_last) at the solver's precision."""
        _check(lib().pbs_state_save(self.h, os.fsencode(path)))

    def load_state(self, path):
        """Restore a snapshot written by save_state of a solver on the same meshes."""
        _check(lib().pbs_state_load(self.h, os.fsencode(path)))

    def instance_tile_tets(self):
        """Tile size of each instance (the autotuner may give a batch's instances different sizes)."""
        out = np.zeros(lib().pbs_num_instances(self.h), np.int32)
        _check(lib().pbs_instance_tile_tets(self.h, _p(out)))
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().pbs_destroy(self.h)
            self.h = None

    __del__ = close

    def get_params(self):
        p = Params()
        _check(lib().pbs_get_params(self.h, C.byref(p)))
        return p

    def set_params(self, params):
        _check(lib().pbs_set_params(self.h, C.byref(params)))
        self.params = params

    def reset(self):
        _check(lib().pbs_reset(self.h))

    def set_positions(self, x):
        x = _f32(x)
        _check(lib().pbs_set_positions(self.h, _p(x)))

    def positions(self, f64=False):
        out = np.zeros((self.nn, 3), np.float64 if f64 else np.float32)
        _check((lib().pbs_get_positions_f64 if f64 else lib().pbs_get_positions)(self.h, _p(out)))
        return out

    def set_velocities(self, v=None, v_last=None):
        v = None if v is None else _f32(v)
        vl = None if v_last is None else _f32(v_last)
        _check(lib().pbs_set_velocities(self.h, _p(v), _p(vl)))

    def velocities(self):
        v = np.zeros((self.nn, 3), np.float32)
        vl = np.zeros((self.nn, 3), np.float32)
        _check(lib().pbs_get_velocities(self.h, _p(v), _p(vl)))
        return v, vl

    def step(self, ext=None, want_positions=False):
        st = StepStats()
        e = None if ext is None else _f32(ext)
        out = np.zeros((self.nn, 3), np.float32) if want_positions else None
        _check(lib().pbs_step(self.h, _p(e), _p(out), C.byref(st)))
        return (st, out) if want_positions else st

    def implicit_step(self, ext=None, params=None, want_positions=False):
        """One backward-Euler step; returns ImplicitStats (and the new positions)."""
        st = ImplicitStats()
        p = params if params is not None else implicit_params()
        e = None if ext is None else _f32(ext)
        out = np.zeros((self.nn, 3), np.float32) if want_positions else None
        _check(lib().pbs_implicit_step(self.h, C.byref(p), _p(e), _p(out), C.byref(st)))
        return (st, out) if want_positions else st

    def pd_step(self, ext=None, params=None, want_positions=False):
        """One projective-dynamics step; returns PdStats (and the new positions)."""
        st = PdStats()
        p = params if params is not None else pd_params()
        e = None if ext is None else _f32(ext)
        out = np.zeros((self.nn, 3), np.float32) if want_positions else None
        _check(lib().pbs_pd_step(self.h, C.byref(p), _p(e), _p(out), C.byref(st)))
        return (st, out) if want_positions else st

    def pd_project(self, positions, stiffness=1e6):
        """Local step alone: per-node sums of w_k R_k c_i (3N doubles, caller order)."""
        pos = _f32(positions)
        out = np.zeros((self.nn, 3))
        _check(lib().pbs_pd_project(self.h, stiffness, _p(pos), _p(out)))
        return out

    def stiffness_matrix(self, tets):
        """K = -df/dx at the current positions: (rowptr, col, values) in the reference's CSR layout."""
        rowptr, col, _, _ = csr_pattern(self.nn, tets)
        val = np.zeros(len(col))
        _check(lib().pbs_stiffness_matrix(self.h, len(col), _p(val)))
        return rowptr, col, val

    def tet_stiffness(self, pos, tet_ids, noalloc=False):
        """Per-tet 12 x 12 stiffness K = -df/dx of the listed tets at pos (FP32); noalloc: through
        pbs_tet_stiffness_noalloc (no device allocation, at most 8192 tets)."""
        pos = _f32(pos)
        ids = np.ascontiguousarray(tet_ids, np.int32).reshape(-1)
        K = np.zeros((len(ids), 12, 12), np.float32)
        fn = lib().pbs_tet_stiffness_noalloc if noalloc else lib().pbs_tet_stiffness
        _check(fn(self.h, _p(pos), len(ids), _p(ids), _p(K)))
        return K

    def do_descent_opt(self, h, iterations, ext):
        e = _f32(ext)
        out = np.zeros((self.nn, 3), np.float32)
        _check(lib().pbs_do_descent_opt(self.h, h, iterations, _p(e), _p(out)))
        return out

    def compute_forces(self, pos):
        pos = _f32(pos)
        tf = np.zeros((self.nt, 12), np.float32)
        nf = np.zeros((self.nn, 3), np.float32)
        te = np.zeros(self.nt, np.float32)
        _check(lib().pbs_compute_forces(self.h, _p(pos), _p(tf), _p(nf), _p(te)))
        return tf, nf, te

    def compute_hessian_diag(self, pos):
        pos = _f32(pos)
        nh = np.zeros((self.nn, 3), np.float32)
        th = np.zeros((self.nt, 12), np.float32)
        _check(lib().pbs_compute_hessian_diag(self.h, _p(pos), _p(nh), _p(th)))
        return nh, th

    def tile_forces(self, pos, hessian=False):
        """Per-node elastic forces (and stiffness diagonal) as one descent pass forms them: the tile
        kernel's owned and halo partials summed in the node pass's order.  Returns (forces (N, 3)
        float64, hd (N, 3) float64 or None, elastic energy)."""
        pos = _f32(pos)
        nf = np.zeros((self.nn, 3), np.float64)
        nh = np.zeros((self.nn, 3), np.float64) if hessian else None
        e = C.c_double()
        _check(lib().pbs_tile_forces(self.h, _p(pos), _p(nf), _p(nh), C.byref(e)))
        return nf, nh, e.value

    def rest_data(self):
        dminv = np.zeros((self.nt, 9), np.float32)
        vol = np.zeros(self.nt, np.float32)
        mass = np.zeros(self.nn, np.float32)
        gf = np.zeros((self.nn, 3), np.float32)
        _check(lib().pbs_get_rest_data(self.h, _p(dminv), _p(vol), _p(mass), _p(gf)))
        return dict(dminv=dminv, vol=vol, mass=mass, gforce=gf)

    def set_profiling(self, period=1):
        """period: 0/False off, 1/True every pass, n: every n-th graph replay."""
        _check(lib().pbs_set_profiling(self.h, int(period)))

    def reset_kernel_times(self):
        _check(lib().pbs_reset_kernel_times(self.h))

    # ---- learned warm start ----
    def learn_set_model(self, params, scales):
        p = np.ascontiguousarray(params, np.float32)
        sc = np.ascontiguousarray(scales, np.float32)
        assert p.size == MLP_PARAMS and sc.size == MLP_SCALES
        _check(lib().pbs_learn_set_model(self.h, _p(p), _p(sc)))

    def learn_get_model(self):
        p = np.zeros(MLP_PARAMS, np.float32)
        sc = np.zeros(MLP_SCALES, np.float32)
        _check(lib().pbs_learn_get_model(self.h, _p(p), _p(sc)))
        return p, sc

    def learn_init_model(self, seed=0):
        _check(lib().pbs_learn_init_model(self.h, seed))

    def learn_enable(self, on=True):
        _check(lib().pbs_learn_enable(self.h, 1 if on else 0))

    def learn_set_options(self, sweeps=0, betas=(1.0,)):
        b = np.ascontiguousarray(betas, np.float32)
        _check(lib().pbs_learn_set_options(self.h, int(sweeps), len(b), _p(b)))

    def learn_collect(self, capacity, stride=1):
        """Collect (features, converged step) samples from the following steps; with stride > 1 each
        step keeps every stride-th free node (a different residue class each step)."""
        _check(lib().pbs_learn_collect_every(self.h, int(capacity), int(stride)))

    def learn_add_samples(self, features, targets):
        f = np.ascontiguousarray(features, np.float32).reshape(-1, 18)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        assert len(f) == len(t)
        _check(lib().pbs_learn_add_samples(self.h, len(f), _p(f), _p(t)))

    def learn_samples(self):
        n = lib().pbs_learn_samples(self.h)
        if n < 0:
            _check(int(n))
        return int(n)

    def learn_get_samples(self, start=0, n=None):
        n = self.learn_samples() - start if n is None else n
        f = np.zeros((n, 18), np.float32)
        t = np.zeros((n, 3), np.float32)
        _check(lib().pbs_learn_get_samples(self.h, start, n, _p(f), _p(t)))
        return f, t

    def learn_fit_scales(self):
        _check(lib().pbs_learn_fit_scales(self.h))

    def learn_train(self, epochs, batch=4096, lr=1e-3, seed=0):
        loss = np.zeros(max(epochs, 1), np.float32)
        _check(lib().pbs_learn_train(self.h, epochs, batch, lr, seed, _p(loss)))
        return loss[:epochs]

    def learn_forward(self, features):
        f = np.ascontiguousarray(features, np.float32).reshape(-1, 18)
        out = np.zeros((len(f), 3), np.float32)
        _check(lib().pbs_learn_forward(self.h, len(f), _p(f), _p(out)))
        return out

    def learn_loss_grad(self, features, targets):
        f = np.ascontiguousarray(features, np.float32).reshape(-1, 18)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        loss = C.c_double()
        g = np.zeros(MLP_PARAMS, np.float32)
        _check(lib().pbs_learn_loss_grad(self.h, len(f), _p(f), _p(t), C.byref(loss), _p(g)))
        return loss.value, g

    def learn_features(self):
        out = np.zeros((self.nn, 18), np.float32)
        _check(lib().pbs_learn_features(self.h, _p(out)))
        return out

    def learn_guess(self):
        a = np.zeros((self.nn, 3), np.float64)
        b = np.zeros((self.nn, 3), np.float64)
        _check(lib().pbs_learn_guess(self.h, _p(a), _p(b)))
        return a, b

    def kernel_times(self):
        n = 16
        names = (C.c_char_p * n)()
        avg = np.zeros(n)
        cnt = np.zeros(n, np.int64)
        c = C.c_int()
        _check(lib().pbs_kernel_times(self.h, n, C.cast(names, C.c_void_p), _p(avg), _p(cnt), C.byref(c)))
        return {names[i].decode(): (float(avg[i]), int(cnt[i])) for i in range(c.value)}
