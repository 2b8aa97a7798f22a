// pbs_mesh.cpp -- host-side mesh utilities of the C ABI (no GPU work): synthetic bars,
// seeded initial states, TetGen text I/O, and the integer adjacency structures the reference
// builds on the CPU (stiffness CSR pattern with its diagonal / per-tet value indices, and the
// per-node tet incidence).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "phybs.h"
#include "pbs_host.h"

using namespace pbs;

extern "C" int pbs_mesh_bar(int nx, int ny, int nz, float edge, float *nodes, int *tets) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !(edge > 0) || !nodes || !tets) return set_error(PBS_ERR_ARG, "pbs_mesh_bar: bad arguments");
  const long sx = nx + 1, sy = ny + 1;
  auto nid = [&](long i, long j, long k) { return (int)(i + sx * (j + sy * k)); };
  for (long k = 0; k <= nz; ++k)
    for (long j = 0; j <= ny; ++j)
      for (long i = 0; i <= nx; ++i) {
        long n = nid(i, j, k);
        nodes[3 * n + 0] = (float)((double)i * (double)edge);
        nodes[3 * n + 1] = (float)((double)j * (double)edge);
        nodes[3 * n + 2] = (float)((double)k * (double)edge);
      }
  // Freudenthal (Kuhn) split: one tet per axis permutation (a,b,c), path 0 -> e_a -> e_a+e_b -> 1;
  // odd permutations swap the last two vertices so det(Dm) > 0.
  static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  static const int sign[6] = {1, -1, -1, 1, 1, -1};
  long t = 0;
  for (long k = 0; k < nz; ++k)
    for (long j = 0; j < ny; ++j)
      for (long i = 0; i < nx; ++i)
        for (int p = 0; p < 6; ++p) {
          long o[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 1}};
          o[1][perms[p][0]] = 1;
          o[2][perms[p][0]] = 1;
          o[2][perms[p][1]] = 1;
          int v[4];
          for (int q = 0; q < 4; ++q) v[q] = nid(i + o[q][0], j + o[q][1], k + o[q][2]);
          if (sign[p] < 0) std::swap(v[2], v[3]);
          for (int q = 0; q < 4; ++q) tets[4 * t + q] = v[q];
          ++t;
        }
  return PBS_OK;
}

// SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number generators", 2014).
static inline uint64_t splitmix64(uint64_t &x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

extern "C" int pbs_mesh_twist(int nx, int ny, int nz, float edge, double twist_deg, double jitter, uint64_t seed,
                              int num_nodes, float *nodes) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !nodes || num_nodes <= 0) return set_error(PBS_ERR_ARG, "pbs_mesh_twist: bad arguments");
  const double L = (double)nz * edge, cx = 0.5 * nx * edge, cy = 0.5 * ny * edge;
  const double pi = 3.14159265358979323846;
  uint64_t st = seed;
  for (int n = 0; n < num_nodes; ++n) {
    double r[3] = {nodes[3 * n], nodes[3 * n + 1], nodes[3 * n + 2]};
    double th = twist_deg * pi / 180.0 * r[2] / L;
    double dx = r[0] - cx, dy = r[1] - cy;
    double x[3] = {cx + std::cos(th) * dx - std::sin(th) * dy, cy + std::sin(th) * dx + std::cos(th) * dy, r[2]};
    for (int d = 0; d < 3; ++d) {
      if (jitter != 0.0) {
        double u = (double)(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
        x[d] += (2.0 * u - 1.0) * jitter * edge;
      }
      nodes[3 * n + d] = (float)x[d];
    }
  }
  return PBS_OK;
}

// TetGen readers (format of TetMesh.cpp:14-151: first line holds the count, then one
// "index v..." line per item, 0-based).
static bool read_header(std::ifstream &f, long &count) {
  std::string line;
  while (std::getline(f, line)) {
    size_t h = line.find('#');
    if (h != std::string::npos) line = line.substr(0, h);
    std::istringstream ss(line);
    if (ss >> count) return true;
  }
  return false;
}

static bool read_rows(std::ifstream &f, long count, int ncol, std::vector<double> &out) {
  out.clear();
  out.reserve((size_t)count * ncol);
  std::string line;
  long got = 0;
  while (got < count && std::getline(f, line)) {
    size_t h = line.find('#');
    if (h != std::string::npos) line = line.substr(0, h);
    std::istringstream ss(line);
    double idx;
    if (!(ss >> idx)) continue;
    for (int c = 0; c < ncol; ++c) {
      double v;
      if (!(ss >> v)) return false;
      out.push_back(v);
    }
    ++got;
  }
  return got == count;
}

extern "C" int pbs_tetgen_count(const char *prefix, int *nn, int *nt, int *nf) {
  if (!prefix) return set_error(PBS_ERR_ARG, "pbs_tetgen_count: null prefix");
  std::string p(prefix);
  long c;
  {
    std::ifstream f(p + ".node");
    if (!f || !read_header(f, c)) return set_error(PBS_ERR_IO, "cannot read " + p + ".node");
    if (nn) *nn = (int)c;
  }
  {
    std::ifstream f(p + ".ele");
    if (!f || !read_header(f, c)) return set_error(PBS_ERR_IO, "cannot read " + p + ".ele");
    if (nt) *nt = (int)c;
  }
  if (nf) {
    std::ifstream f(p + ".face");
    *nf = (f && read_header(f, c)) ? (int)c : 0;
  }
  return PBS_OK;
}

extern "C" int pbs_tetgen_load(const char *prefix, float scale, float *nodes, int *tets, int *faces) {
  if (!prefix || !nodes || !tets) return set_error(PBS_ERR_ARG, "pbs_tetgen_load: null argument");
  std::string p(prefix);
  long nn, nt, nf;
  std::vector<double> rows;
  {
    std::ifstream f(p + ".node");
    if (!f || !read_header(f, nn) || !read_rows(f, nn, 3, rows)) return set_error(PBS_ERR_IO, "cannot parse " + p + ".node");
    for (long i = 0; i < 3 * nn; ++i) nodes[i] = (float)rows[i] * scale;
  }
  {
    std::ifstream f(p + ".ele");
    if (!f || !read_header(f, nt) || !read_rows(f, nt, 4, rows)) return set_error(PBS_ERR_IO, "cannot parse " + p + ".ele");
    for (long i = 0; i < 4 * nt; ++i) {
      tets[i] = (int)rows[i];
      if (tets[i] < 0 || tets[i] >= nn) return set_error(PBS_ERR_MESH, "tet index out of range in " + p + ".ele");
    }
  }
  if (faces) {
    std::ifstream f(p + ".face");
    if (!f || !read_header(f, nf) || !read_rows(f, nf, 3, rows)) return set_error(PBS_ERR_IO, "cannot parse " + p + ".face");
    for (long i = 0; i < 3 * nf; ++i) faces[i] = (int)rows[i];
  }
  return PBS_OK;
}

// One TetGen file at a time, as TetMesh::initNodesFromFile / initTetsFromFile / initFacesFromFile
// (TetMesh.cpp:14-56, 96-125, 127-151) read them: with a NULL output only the count is returned.
template <typename T>
static int read_one(const char *path, int ncol, int *count, T *out, double scale, const char *what) {
  if (!path || !count) return set_error(PBS_ERR_ARG, std::string(what) + ": null argument");
  std::ifstream f(path);
  long c;
  if (!f || !read_header(f, c) || c < 0 || c > 2147483647L) return set_error(PBS_ERR_IO, std::string("cannot read ") + path);
  *count = (int)c;
  if (!out) return PBS_OK;
  std::vector<double> rows;
  if (!read_rows(f, c, ncol, rows)) return set_error(PBS_ERR_IO, std::string("cannot parse ") + path);
  for (long i = 0; i < (long)ncol * c; ++i) out[i] = (T)(rows[i] * scale);
  return PBS_OK;
}
extern "C" int pbs_tetgen_read_nodes(const char *path, float scale, int *num_nodes, float *nodes) {
  int rc = read_one<float>(path, 3, num_nodes, nullptr, 1.0, "pbs_tetgen_read_nodes");
  if (rc || !nodes) return rc;
  // positions are scaled in FP32 after parsing, as pbs_tetgen_load does
  rc = read_one<float>(path, 3, num_nodes, nodes, 1.0, "pbs_tetgen_read_nodes");
  if (rc) return rc;
  for (long i = 0; i < 3L * *num_nodes; ++i) nodes[i] = nodes[i] * scale;
  return PBS_OK;
}
extern "C" int pbs_tetgen_read_tets(const char *path, int *num_tets, int *tets) {
  return read_one<int>(path, 4, num_tets, tets, 1.0, "pbs_tetgen_read_tets");
}
extern "C" int pbs_tetgen_read_faces(const char *path, int *num_faces, int *faces) {
  return read_one<int>(path, 3, num_faces, faces, 1.0, "pbs_tetgen_read_faces");
}

extern "C" int pbs_node_write(const char *path, int nn, const float *nodes) {
  if (!path || !nodes || nn < 0) return set_error(PBS_ERR_ARG, "pbs_node_write: bad arguments");
  FILE *f = std::fopen(path, "w");
  if (!f) return set_error(PBS_ERR_IO, std::string("cannot write ") + path);
  std::fprintf(f, "%d 3 0 0\n", nn);
  for (int i = 0; i < nn; ++i) std::fprintf(f, "%d %.9g %.9g %.9g\n", i, nodes[3 * i], nodes[3 * i + 1], nodes[3 * i + 2]);
  std::fclose(f);
  return PBS_OK;
}

extern "C" int pbs_node_read(const char *path, int nn, float *nodes) {
  if (!path || !nodes || nn < 0) return set_error(PBS_ERR_ARG, "pbs_node_read: bad arguments");
  std::ifstream f(path);
  long c;
  std::vector<double> rows;
  if (!f || !read_header(f, c)) return set_error(PBS_ERR_IO, std::string("cannot read ") + path);
  if (c != nn) return set_error(PBS_ERR_ARG, "pbs_node_read: node count mismatch");
  if (!read_rows(f, c, 3, rows)) return set_error(PBS_ERR_IO, std::string("cannot parse ") + path);
  for (long i = 0; i < 3 * c; ++i) nodes[i] = (float)rows[i];
  return PBS_OK;
}

extern "C" int pbs_incidence(int nn, int nt, const int *tets, int *offsets, int *entries) {
  if (nn < 0 || nt < 0 || (nt > 0 && !tets) || !offsets || (nt > 0 && !entries)) return set_error(PBS_ERR_ARG, "pbs_incidence: bad arguments");
  for (long i = 0; i < 4L * nt; ++i)
    if (tets[i] < 0 || tets[i] >= nn) return set_error(PBS_ERR_MESH, "pbs_incidence: tet index out of range");
  std::fill(offsets, offsets + nn + 1, 0);
  for (long t = 0; t < nt; ++t)
    for (int j = 0; j < 4; ++j) offsets[tets[4 * t + j] + 1]++;
  for (int i = 0; i < nn; ++i) offsets[i + 1] += offsets[i];
  std::vector<int> fill(nn, 0);
  for (long t = 0; t < nt; ++t)
    for (int j = 0; j < 4; ++j) {
      int n = tets[4 * t + j];
      entries[offsets[n] + fill[n]++] = (int)(12 * t + 3 * j);
    }
  return PBS_OK;
}

// getCSRvalueIndex (IsotropicMaterial.h:120-138): the reference's binary search, kept as is so
// that diag_idx / kid equal the reference's arrays for every valid pattern.
static int csr_value_index(int i, int j, const int *rowptr, const int *col) {
  int l = rowptr[i], h = rowptr[i + 1] - 1, k = (l + h) / 2;
  while (l < h) {
    int kc = col[k];
    if (kc == j) break;
    if (kc < j) l = k + 1;
    else h = k - 1;
    k = (l + h) / 2;
  }
  return k;
}

extern "C" int64_t pbs_csr_pattern(int nn, int nt, const int *tets, int *rowptr, int *colidx, int *diag_idx, int *kid) {
  if (nn < 0 || nt < 0 || (nt > 0 && !tets)) return set_error(PBS_ERR_ARG, "pbs_csr_pattern: bad arguments");
  for (long i = 0; i < 4L * nt; ++i)
    if (tets[i] < 0 || tets[i] >= nn) return set_error(PBS_ERR_MESH, "pbs_csr_pattern: tet index out of range");
  // sorted unique node neighbourhoods (each node is its own neighbour if it is in any tet)
  std::vector<std::vector<int>> nb(nn);
  for (long t = 0; t < nt; ++t)
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) nb[tets[4 * t + a]].push_back(tets[4 * t + b]);
  int64_t nnz = 0;
  for (int i = 0; i < nn; ++i) {
    std::sort(nb[i].begin(), nb[i].end());
    nb[i].erase(std::unique(nb[i].begin(), nb[i].end()), nb[i].end());
    nnz += 9 * (int64_t)nb[i].size();
  }
  if (nnz > 2147483647LL) return set_error(PBS_ERR_ARG, "pbs_csr_pattern: nnz exceeds int32");
  if (!rowptr || !colidx) return nnz;
  int64_t pos = 0;
  for (int i = 0; i < nn; ++i)
    for (int d = 0; d < 3; ++d) {
      rowptr[3 * i + d] = (int)pos;
      for (int j : nb[i])
        for (int e = 0; e < 3; ++e) colidx[pos++] = 3 * j + e;
    }
  rowptr[3 * nn] = (int)pos;
  if (diag_idx)
    for (int r = 0; r < 3 * nn; ++r) diag_idx[r] = csr_value_index(r, r, rowptr, colidx);
  if (kid)
    for (long t = 0; t < nt; ++t)
      for (int fi = 0; fi < 4; ++fi)
        for (int fj = 0; fj < 3; ++fj) {
          int gki = tets[4 * t + fi] * 3 + fj, ki = fi * 3 + fj;
          for (int ni = 0; ni < 4; ++ni) kid[48 * t + ki * 4 + ni] = csr_value_index(gki, tets[4 * t + ni] * 3, rowptr, colidx);
        }
  return nnz;
}

// ---- interactive external forces (TetMesh.cpp:229-276, 378-470; RenderWidget.cpp:150-195) ----

// Face ring of each boundary face: its 3 nodes in face order, then the other nodes adjacent to
// any of them along boundary-face edges, ascending (computeBoundfaceRingIndices).
extern "C" int64_t pbs_face_ring(int nn, int nf, const int *faces, int64_t *offsets, int *ring) {
  if (nn <= 0 || nf < 0 || (nf > 0 && !faces)) return set_error(PBS_ERR_ARG, "pbs_face_ring: bad arguments");
  for (long i = 0; i < 3L * nf; ++i)
    if (faces[i] < 0 || faces[i] >= nn) return set_error(PBS_ERR_MESH, "face index out of range");
  std::vector<std::vector<int>> nb(nn);
  for (int f = 0; f < nf; ++f)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        if (a != b) nb[faces[3 * f + a]].push_back(faces[3 * f + b]);
  for (auto &v : nb) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
  int64_t total = 0;
  std::vector<int> u;
  for (int f = 0; f < nf; ++f) {
    const int *q = faces + 3 * f;
    u.clear();
    for (int a = 0; a < 3; ++a) u.insert(u.end(), nb[q[a]].begin(), nb[q[a]].end());
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    u.erase(std::remove_if(u.begin(), u.end(), [&](int n) { return n == q[0] || n == q[1] || n == q[2]; }), u.end());
    if (offsets) offsets[f] = total;
    if (ring) {
      for (int a = 0; a < 3; ++a) ring[total + a] = q[a];
      for (size_t k = 0; k < u.size(); ++k) ring[total + 3 + (int64_t)k] = u[k];
    }
    total += 3 + (int64_t)u.size();
  }
  if (offsets) offsets[nf] = total;
  return total;
}

// Face picked by a ray (pickFacebyRay): among faces whose outward normal (v2 x v1, v1 = p1 - p0,
// v2 = p2 - p0) points against the direction, the one whose centre has the smallest squared
// distance to the ray, if below the squared length of face 0's first edge; -1 if none.
extern "C" int pbs_pick_face(int nn, const float *nodes, int nf, const int *faces, const float *orig,
                             const float *dir, int *face) {
  if (nn <= 0 || !nodes || nf <= 0 || !faces || !orig || !dir || !face)
    return set_error(PBS_ERR_ARG, "pbs_pick_face: bad arguments");
  for (long i = 0; i < 3L * nf; ++i)
    if (faces[i] < 0 || faces[i] >= nn) return set_error(PBS_ERR_MESH, "face index out of range");
  auto P = [&](int n, int d) { return nodes[3 * (size_t)n + d]; };
  float e[3];
  for (int d = 0; d < 3; ++d) e[d] = P(faces[1], d) - P(faces[0], d);
  const float r = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
  const float r2 = r * r;
  float best = 1e8f;
  int pick = -1;
  for (int f = 0; f < nf; ++f) {
    const int *q = faces + 3 * f;
    float v1[3], v2[3], c[3];
    for (int d = 0; d < 3; ++d) {
      v1[d] = P(q[1], d) - P(q[0], d);
      v2[d] = P(q[2], d) - P(q[0], d);
      c[d] = (P(q[0], d) + P(q[1], d) + P(q[2], d)) / 3.0f;
    }
    float nrm[3] = {v2[1] * v1[2] - v2[2] * v1[1], v2[2] * v1[0] - v2[0] * v1[2], v2[0] * v1[1] - v2[1] * v1[0]};
    const float len = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    if (len > 0)
      for (float &x : nrm) x /= len;
    if (nrm[0] * dir[0] + nrm[1] * dir[1] + nrm[2] * dir[2] >= 0) continue;
    float ov[3];
    for (int d = 0; d < 3; ++d) ov[d] = orig[d] - c[d];
    const float a = ov[0] * ov[0] + ov[1] * ov[1] + ov[2] * ov[2];
    const float b = ov[0] * dir[0] + ov[1] * dir[1] + ov[2] * dir[2];
    const float d2 = a - b * b;
    if (d2 < r2 && d2 < best) {
      pick = f;
      best = d2;
    }
  }
  *face = pick;
  return PBS_OK;
}

// External forces of dragging a picked face (dragFaceRing): factor = min(|drag| / r * 100, 800)
// with r = |p1 - p0| of face 0; ext = 0, then += factor * mass * drag/|drag| on the face's 3
// nodes and half of that on the rest of its ring.  ext: 3N AoS floats.
extern "C" int pbs_drag_face_ring(int nn, const float *nodes, const float *mass, int nf, const int *faces,
                                  const int64_t *ring_offsets, const int *ring, int face, const float *drag,
                                  float *ext) {
  if (nn <= 0 || !nodes || !mass || nf <= 0 || !faces || !ring_offsets || !ring || !drag || !ext)
    return set_error(PBS_ERR_ARG, "pbs_drag_face_ring: bad arguments");
  if (face < 0 || face >= nf) return set_error(PBS_ERR_ARG, "face out of range");
  std::fill(ext, ext + 3 * (size_t)nn, 0.0f);
  float e[3];
  for (int d = 0; d < 3; ++d) e[d] = nodes[3 * (size_t)faces[1] + d] - nodes[3 * (size_t)faces[0] + d];
  const float r = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
  const float dl = std::sqrt(drag[0] * drag[0] + drag[1] * drag[1] + drag[2] * drag[2]);
  if (!(r > 0) || !(dl > 0)) return PBS_OK;
  float factor = dl / r * 100.0f;
  if (factor > 800.0f) factor = 800.0f;
  const float dir[3] = {drag[0] / dl, drag[1] / dl, drag[2] / dl};
  for (int64_t k = ring_offsets[face]; k < ring_offsets[face + 1]; ++k) {
    const int n = ring[k];
    if (n < 0 || n >= nn) return set_error(PBS_ERR_MESH, "ring index out of range");
    const float w = (k - ring_offsets[face] < 3 ? 1.0f : 0.5f) * factor * mass[n];
    for (int d = 0; d < 3; ++d) ext[3 * (size_t)n + d] += w * dir[d];
  }
  return PBS_OK;
}
