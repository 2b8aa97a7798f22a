"""The entry point that allocates nothing (pbs_tet_stiffness_noalloc) against pbs_tet_stiffness,
through the C ABI: bitwise equal results on two calls in a row on one handle (the second with
fewer tets and other positions, so stale scratch contents would show), in FP32 and FP64, and a
clear error beyond its 8192-tet limit."""
import sys

import numpy as np
import pytest

import meshgen as M
import phybs_amd as P

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("precision", [P.FP32, P.FP64])
def test_noalloc_stiffness_equals_allocating_entry(precision):
    rest, tets = M.bar(12, 12, 50, 0.05)  # 43,200 tets
    mask = M.clamp_x0(rest)
    say("no-alloc per-tet stiffness, precision", precision)
    s = P.Solver(rest, tets, mask, P.default_params(precision=precision))
    rng = np.random.default_rng(4)
    calls = [(M.twisted_state(rest, 12, 12, 50, 0.05, 60.0, 0.01, 2), rng.choice(len(tets), 8192, replace=False)),
             (M.twisted_state(rest, 12, 12, 50, 0.05, 20.0, 0.005, 3), rng.choice(len(tets), 37, replace=False))]
    for pos, ids in calls:
        a = s.tet_stiffness(pos, ids, noalloc=True)
        b = s.tet_stiffness(pos, ids)
        say("  %d tets: max |K| %.3e" % (len(ids), np.abs(b).max()))
        assert np.isfinite(b).all() and np.abs(b).max() > 0
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    with pytest.raises(P.PhybsError) as e:
        s.tet_stiffness(rest, np.arange(8193), noalloc=True)
    assert e.value.code == P.ERR_ARG and "8192" in str(e.value)
    s.close()
