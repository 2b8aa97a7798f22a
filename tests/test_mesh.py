"""Host mesh utilities of the C ABI against independent restatements: the numpy bar
generator, brute-force adjacency, the TetGen fixtures, and the oracle's CSR / incidence."""
import os

import numpy as np
import pytest

import meshgen as M
import phybs_amd as P
from oracle import pyoracle as O


@pytest.mark.parametrize("dims", [(1, 1, 1), (4, 4, 16), (3, 5, 2)])
def test_bar_matches_numpy_generator_bit_exact(dims):
    rest, tets = P.mesh_bar(*dims, 0.05)
    r2, t2 = M.bar(*dims, 0.05)
    assert np.array_equal(tets, t2)
    assert np.array_equal(rest.view(np.uint32), r2.view(np.uint32))
    nx, ny, nz = dims
    assert len(rest) == (nx + 1) * (ny + 1) * (nz + 1) and len(tets) == 6 * nx * ny * nz


def test_bar_tets_positive_and_fill_the_box():
    rest, tets = P.mesh_bar(3, 2, 4, 0.05)
    r = rest.astype(np.float64)
    D = np.stack([r[tets[:, k]] - r[tets[:, 0]] for k in (1, 2, 3)], -1)
    det = np.linalg.det(D)
    assert (det > 0).all()
    assert np.allclose(det / 6, 0.05 ** 3 / 6)
    assert np.isclose((det / 6).sum(), 3 * 2 * 4 * 0.05 ** 3)
    # each cube's 6 tets share its main diagonal (000 -> 111): vertex 0 is the cube's low corner
    # and one of the others its high corner
    v = r[tets]
    far = np.abs(v - v[:, :1]).sum(2).max(1)
    assert np.allclose(far, 3 * 0.05)
    assert np.allclose(v.min(1), v[:, 0])


def test_config_node_and_clamp_counts():
    for dims, nn, ncl in [((4, 4, 16), 425, 85), ((12, 12, 50), 8619, 663)]:
        rest, tets = P.mesh_bar(*dims, 0.05)
        assert len(rest) == nn and int((rest[:, 0] == 0).sum()) == ncl


def test_twist_matches_numpy_bit_exact():
    rest, _ = P.mesh_bar(4, 4, 16, 0.05)
    x = P.mesh_twist(rest, 4, 4, 16, 0.05, 90.0, 0.01, 1)
    x2 = M.twisted_state(rest, 4, 4, 16, 0.05, 90.0, 0.01, 1)
    assert np.array_equal(x.view(np.uint32), x2.view(np.uint32))
    # pure twist preserves distance to the axis and z
    y = P.mesh_twist(rest, 4, 4, 16, 0.05, 90.0, 0.0, 1)
    c = np.array([0.1, 0.1])
    assert np.allclose(np.linalg.norm(y[:, :2] - c, axis=1), np.linalg.norm(rest[:, :2] - c, axis=1), atol=1e-6)
    assert np.array_equal(y[:, 2], rest[:, 2])


@pytest.mark.parametrize("name", ["model_tet/tet.1", "model_tet2/tet.2", "model_box/box.1", "model_bar/bar.1"])
def test_tetgen_loader_matches_fixture(name):
    prefix = os.path.join(M.GOLDEN, name)
    nodes, tets, faces = P.tetgen_load(prefix, 1.0, faces=True)
    n2, t2 = M.read_tetgen(prefix)
    assert np.array_equal(nodes, n2) and np.array_equal(tets, t2)
    assert faces.shape[1] == 3 and faces.min() >= 0 and faces.max() < len(nodes)
    scaled, _ = P.tetgen_load(prefix, 0.1)  # TetMesh.cpp:38 scale
    assert np.array_equal(scaled, (n2 * np.float32(0.1)).astype(np.float32))


def test_tetgen_missing_file_is_io_error():
    with pytest.raises(P.PhybsError) as e:
        P.tetgen_load("/nonexistent/mesh")
    assert e.value.code == P.ERR_IO


def test_node_file_round_trip(tmp_path):
    rest, _ = P.mesh_bar(2, 3, 4)
    x = P.mesh_twist(rest, 2, 3, 4, 0.05, 45.0, 0.02, 3)
    path = str(tmp_path / "state.node")
    P.node_write(path, x)
    assert open(path).readline().split()[:2] == [str(len(x)), "3"]
    y = P.node_read(path, len(x))
    assert np.array_equal(x, y)
    with pytest.raises(P.PhybsError):
        P.node_read(path, len(x) + 1)


def brute_pattern(nn, tets):
    nb = [set() for _ in range(nn)]
    for t in tets:
        for a in t:
            nb[a].update(int(b) for b in t)
    rows = []
    for i in range(nn):
        cols = [3 * j + e for j in sorted(nb[i]) for e in range(3)]
        rows += [cols] * 3
    return rows


@pytest.mark.parametrize("src", ["bar", "model_bar/bar.1", "model_box/box.1"])
def test_csr_pattern_bit_exact(src):
    if src == "bar":
        rest, tets = P.mesh_bar(4, 4, 16)
    else:
        rest, tets = M.read_tetgen(os.path.join(M.GOLDEN, src))
    nn = len(rest)
    rowptr, col, diag, kid = P.csr_pattern(nn, tets)
    r2, c2, d2, k2 = O.csr_pattern(nn, tets)
    assert np.array_equal(rowptr, r2) and np.array_equal(col, c2)
    assert np.array_equal(diag, d2) and np.array_equal(kid, k2)
    rows = brute_pattern(nn, tets)
    assert np.array_equal(rowptr, np.concatenate([[0], np.cumsum([len(r) for r in rows])]))
    assert np.array_equal(col, np.concatenate(rows))
    # diagonalIdx points at (i, i); kID at (3*node_fi + fj, 3*node_ni)
    r = np.repeat(np.arange(3 * nn), np.diff(rowptr))
    assert np.array_equal(r[diag], np.arange(3 * nn)) and np.array_equal(col[diag], np.arange(3 * nn))
    kid = kid.reshape(-1, 4, 3, 4)  # tet, fi, fj, ni
    rows_expect = 3 * tets[:, :, None, None] + np.arange(3)[None, None, :, None]
    cols_expect = 3 * tets[:, None, None, :]
    assert np.array_equal(r[kid], np.broadcast_to(rows_expect, kid.shape))
    assert np.array_equal(col[kid], np.broadcast_to(cols_expect, kid.shape))


def test_incidence_bit_exact():
    rest, tets = P.mesh_bar(4, 4, 16)
    nn = len(rest)
    off, ent = P.incidence(nn, tets)
    o2, e2 = O.incidence(nn, tets)
    assert np.array_equal(off, o2) and np.array_equal(ent, e2)
    lists = [[] for _ in range(nn)]
    for t, q in enumerate(tets):
        for j, n in enumerate(q):
            lists[n].append(12 * t + 3 * j)
    assert np.array_equal(ent, np.concatenate(lists))
    assert np.array_equal(np.diff(off), [len(l) for l in lists])


def test_adjacency_rejects_out_of_range():
    _, tets = P.mesh_bar(1, 1, 1)
    with pytest.raises(P.PhybsError) as e:
        P.incidence(3, tets)
    assert e.value.code == P.ERR_MESH


def faces_fixture():
    nodes, tets, faces = P.tetgen_load(os.path.join(M.GOLDEN, "model_bar/bar.1"), 1.0, faces=True)
    return nodes, tets, faces


def test_face_ring_brute_force():
    nodes, _, faces = faces_fixture()
    off, ring = P.face_ring(len(nodes), faces)
    nb = [set() for _ in range(len(nodes))]
    for f in faces:
        for a in f:
            nb[a].update(int(b) for b in f if b != a)
    for i, f in enumerate(faces):
        want = list(f) + sorted(set().union(*(nb[a] for a in f)) - set(int(a) for a in f))
        assert list(ring[off[i]:off[i + 1]]) == want


def test_pick_face_and_drag_forces():
    nodes, _, faces = faces_fixture()
    off, ring = P.face_ring(len(nodes), faces)
    # a ray straight at the centre of face 7, from outside along its inward normal
    f = faces[7]
    p0, p1, p2 = nodes[f[0]], nodes[f[1]], nodes[f[2]]
    n = np.cross(p2 - p0, p1 - p0)
    n /= np.linalg.norm(n)
    c = (p0 + p1 + p2) / 3
    assert P.pick_face(nodes, faces, c + 5 * n, -n) == 7
    assert P.pick_face(nodes, faces, c + 5 * n + 1e3, -n) == -1  # passes far from every face
    mass = np.linspace(1.0, 2.0, len(nodes)).astype(np.float32)
    drag = np.array([0.0, 0.3, 0.4], np.float32)
    ext = P.drag_face_ring(nodes, mass, faces, off, ring, 7, drag)
    r = np.linalg.norm(nodes[faces[0][1]] - nodes[faces[0][0]])
    factor = min(0.5 / r * 100, 800)
    want = np.zeros_like(nodes)
    for k, nd in enumerate(ring[off[7]:off[8]]):
        want[nd] += (1.0 if k < 3 else 0.5) * factor * mass[nd] * drag / 0.5
    np.testing.assert_allclose(ext, want, rtol=1e-5)
    assert not ext[np.setdiff1d(np.arange(len(nodes)), ring[off[7]:off[8]])].any()
