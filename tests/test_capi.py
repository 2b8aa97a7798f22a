"""The C-ABI library loads without a GPU, exports every symbol include/phybs.h declares, and
rejects bad arguments before touching the device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import phybs_amd as P

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "phybs.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pbs_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    L = P.lib()
    syms = declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # the Python binding types every declared entry point
    assert set(syms) <= set(P.SIGNATURES), set(syms) - set(P.SIGNATURES)


def test_abi_version_and_defaults_match_reference_constants():
    assert P.lib().pbs_abi_version() == 1
    p = P.default_params()
    # CUDAInterface.cu:644, 914, 948, 919, 921, 942, 637, 491, 732-738; fvm.cpp:187
    assert p.dt == pytest.approx(0.03) and p.alpha0 == pytest.approx(0.1)
    assert p.alpha_decay == pytest.approx(0.7) and p.alpha_min == pytest.approx(0.01)
    assert (p.hd_refresh, p.energy_check, p.iterations) == (32, 8, 96)
    assert p.damping == pytest.approx(0.99) and p.init_momentum == pytest.approx(0.2)
    assert p.profile_n == 3 and list(p.profile_k[:3]) == [1, 7, 10]
    assert np.allclose(list(p.profile_rho[:3]), [0.96, 0.991, 0.999])
    assert tuple(p.grav) == pytest.approx((0.0, -9.8, 0.0))
    assert p.fixed_mass == pytest.approx(1e8)
    assert p.semantics == P.SEM_REF_GPU and p.precision == P.FP32


def test_create_rejects_bad_arguments_without_a_device():
    rest, tets = P.mesh_bar(1, 1, 1)
    h = C.c_void_p()
    L = P.lib()
    p = P.default_params()
    # null rest
    assert L.pbs_create(len(rest), None, None, None, len(tets), P._p(tets), C.byref(p), C.byref(h)) == P.ERR_ARG
    # tet index out of range
    bad = tets.copy()
    bad[0, 0] = len(rest)
    assert L.pbs_create(len(rest), None, P._p(rest), None, len(bad), P._p(bad), C.byref(p), C.byref(h)) == P.ERR_MESH
    assert b"out of range" in L.pbs_last_error()
    # bad enum
    p2 = P.default_params(precision=16)
    assert L.pbs_create(len(rest), None, P._p(rest), None, len(tets), P._p(tets), C.byref(p2), C.byref(h)) == P.ERR_ARG
    p3 = P.default_params(block_size=96)
    assert L.pbs_create(len(rest), None, P._p(rest), None, len(tets), P._p(tets), C.byref(p3), C.byref(h)) == P.ERR_ARG
    assert not h.value


def test_handle_calls_reject_null_handle():
    L = P.lib()
    assert L.pbs_step(None, None, None, None) == P.ERR_ARG
    assert L.pbs_reset(None) == P.ERR_ARG
    assert L.pbs_set_positions(None, None) == P.ERR_ARG
    L.pbs_destroy(None)


def test_create_batch_rejects_bad_arguments_without_a_device():
    p = P.default_params()
    rest, tets = P.mesh_bar(1, 1, 1)
    nodes = np.array([len(rest), len(rest)], np.int32)
    ntet = np.array([len(tets), len(tets)], np.int32)
    allrest = np.concatenate([rest, rest])
    alltets = np.concatenate([tets, tets + len(rest)]).astype(np.int32)
    E = np.array([1e5, 1e5], np.float32)
    bad_nu = np.array([0.3, 0.5], np.float32)
    h = C.c_void_p()
    assert P.lib().pbs_create_batch(2, P._p(nodes), P._p(ntet), None, P._p(allrest), None, P._p(alltets), P._p(E),
                                    P._p(bad_nu), C.byref(p), C.byref(h)) == P.ERR_ARG
    cross = alltets.copy()
    cross[0, 0] = len(rest)  # instance 0 tet pointing into instance 1
    nu = np.array([0.3, 0.45], np.float32)
    assert P.lib().pbs_create_batch(2, P._p(nodes), P._p(ntet), None, P._p(allrest), None, P._p(cross), P._p(E),
                                    P._p(nu), C.byref(p), C.byref(h)) == P.ERR_MESH
    assert P.lib().pbs_create_batch(0, P._p(nodes), P._p(ntet), None, P._p(allrest), None, P._p(alltets), P._p(E),
                                    P._p(nu), C.byref(p), C.byref(h)) == P.ERR_ARG
