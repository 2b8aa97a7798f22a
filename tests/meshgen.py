"""Independent numpy restatement of the synthetic inputs (bar meshes and seeded initial states).

The product's C++ generator (phybs_amd/csrc/pbs_mesh.cpp) must agree with these bit for bit;
tests/test_mesh.py checks that.  Conventions:
  node id = i + (nx+1) * (j + (ny+1) * k), position = (i, j, k) * edge;
  cubes in the same (i fastest) order, 6 tets per cube (Freudenthal / Kuhn split): for each
  permutation (a, b, c) of the axes in lexicographic order the tet is
  [000, e_a, e_a+e_b, 111], with the last two vertices swapped for odd permutations so that
  det(Dm) > 0 for every tet.
"""
import itertools
import os

import numpy as np

_PERMS = list(itertools.permutations(range(3)))


def _perm_sign(p):
    s = 1
    p = list(p)
    for i in range(3):
        for j in range(i + 1, 3):
            if p[i] > p[j]:
                s = -s
    return s


def bar(nx, ny, nz, edge=0.05):
    edge = float(np.float32(edge))  # the C ABI takes the edge as a float
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    nid = lambda a, b, c: a + (nx + 1) * (b + (ny + 1) * c)
    idx = np.stack([i, j, k], -1).transpose(2, 1, 0, 3).reshape(-1, 3)  # i fastest
    rest = (idx.astype(np.float64) * edge).astype(np.float32)
    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cubes = np.stack([ci, cj, ck], -1).transpose(2, 1, 0, 3).reshape(-1, 3)
    tets = []
    for p in _PERMS:
        ea = np.eye(3, dtype=np.int64)[p[0]]
        eb = np.eye(3, dtype=np.int64)[p[1]]
        v = [np.zeros(3, np.int64), ea, ea + eb, np.ones(3, np.int64)]
        if _perm_sign(p) < 0:
            v[2], v[3] = v[3], v[2]
        tets.append(np.stack([nid(*(cubes + o).T) for o in v], -1))
    tets = np.stack(tets, 1).reshape(-1, 4).astype(np.int32)
    return rest, tets


def clamp_x0(rest):
    return (rest[:, 0] == 0).astype(np.int32)


def splitmix64(seed, n):
    """n uniform doubles in [0,1) from SplitMix64 (Steele, Lea, Flood 2014)."""
    mask = (1 << 64) - 1
    out = np.empty(n)
    x = seed & mask
    for i in range(n):
        x = (x + 0x9E3779B97F4A7C15) & mask
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z = z ^ (z >> 31)
        out[i] = (z >> 11) * (1.0 / (1 << 53))
    return out


def twisted_state(rest, nx, ny, nz, edge, twist_deg=90.0, jitter=0.01, seed=1):
    """Twist about the bar's long (z) axis through its centre, theta(z) = twist * z / L, plus a
    uniform +-jitter*edge perturbation per coordinate (node-major x,y,z order), in FP64, rounded
    to FP32."""
    edge = float(np.float32(edge))
    r = rest.astype(np.float64)
    L = nz * edge
    cx, cy = 0.5 * nx * edge, 0.5 * ny * edge
    th = np.deg2rad(twist_deg) * r[:, 2] / L
    dx, dy = r[:, 0] - cx, r[:, 1] - cy
    x = np.empty_like(r)
    x[:, 0] = cx + np.cos(th) * dx - np.sin(th) * dy
    x[:, 1] = cy + np.sin(th) * dx + np.cos(th) * dy
    x[:, 2] = r[:, 2]
    if jitter:
        u = splitmix64(seed, x.size).reshape(-1, 3)
        x += (2.0 * u - 1.0) * jitter * edge
    return x.astype(np.float32)


def read_tetgen(prefix):
    """TetGen .node/.ele reader (format of TetMesh.cpp:14-121, without the x0.1 scale)."""
    with open(prefix + ".node") as f:
        hdr = f.readline().split()
        n = int(hdr[0])
        nodes = np.array([[float(v) for v in f.readline().split()[1:4]] for _ in range(n)], np.float32)
    with open(prefix + ".ele") as f:
        hdr = f.readline().split()
        t = int(hdr[0])
        tets = np.array([[int(v) for v in f.readline().split()[1:5]] for _ in range(t)], np.int32)
    return nodes, tets


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
