// stl.hip — implicit-union STL export of a lattice on the GPU
// (projects/metamaterial-compression/industry-grade/src/sim.cpp:579-888,
// Simulator::exportSTL, MC/ below).
//
// The printable solid is the union of the beams' capsules and the joints'
// spheres at the initial positions.  Its signed distance is sampled on a voxel
// grid, the zero set is extracted cell by cell, and the mesh is audited for
// boundary and non-manifold edges before it is written as binary STL with a
// validation report next to it.
//
//  * Grid and field: the reference's voxel size, bounds, shape boxes and
//    single-precision distances (MC/src/sim.cpp:647-709).  One workgroup per
//    shape writes min(distance) into its box with an integer atomic min on an
//    order-preserving encoding of the float, so the field is the reference's
//    bit for bit whatever the order the shapes land in.  Floating-point
//    contraction is off in this file: every distance is the reference's
//    sequence of correctly rounded double operations.
//  * Extraction: cells in the reference's order (y outer, then z, then x), a
//    count pass, an exclusive scan and an emit pass, so the triangles of a
//    cell are contiguous and cells follow one another as in the reference's
//    loop.  Edge vertices are interpolated and snapped to a 1e-8 m grid as the
//    reference does (MC/src/sim.cpp:627-640), so every vertex of a sign-
//    changing cell edge is the reference's vertex.  The triangulation of a
//    cell is this file's own: a case table generated at first use from a
//    face-consistent rule (below), watertight by construction, in place of
//    the classic 256-case table.
//  * Audit and repair (MC/src/sim.cpp:758-833): zero-area triangles are
//    dropped, the orientation is made outward by the signed volume, boundary
//    loops (if any) are closed, and every edge must be shared by exactly two
//    triangles.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <limits>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/mmc.h"

#pragma clang fp contract(off)

namespace {

std::string g_stl_error;

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
#define STLCALL(x)                                                                                  \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess) throw HipError(std::string(#x) + ": " + hipGetErrorString(e_));           \
  } while (0)

// ---------------------------------------------------------------------------
// Case table: corner c of a cell sits at (c & 1, c >> 1 & 1, c >> 2 & 1); edge
// 4 a + m joins the m-th corner without bit a (in increasing order) to that
// corner with bit a.  For every one of the 256 inside/outside patterns the
// table lists triangles as triples of edge ids.
//
// Rule: on each of the six faces, walked counter-clockwise as seen from
// outside the cell, a sign-changing edge is an exit (inside corner first) or
// an entry; every exit is joined to the entry that precedes it, so that the
// contour keeps the inside on its left and, where a face has two inside
// corners on a diagonal, cuts each of them off.  The rule depends only on the
// face's four corners, so the two cells that share a face draw the same
// contour on it (in opposite directions).  Every crossing edge starts one
// segment and ends one, so the segments close into loops, each started at its
// lowest edge id.  A loop is triangulated without a diagonal between two
// vertices on a common face of the cell (such a diagonal could coincide with
// one of the neighbour's and give an edge of four triangles): the first split
// vertex that keeps both sides free of face diagonals, recursively; every one
// of the 256 cases has such a triangulation.  Loops of neighbouring cells then
// meet edge to edge, so the mesh is closed and every edge has two triangles.
// ---------------------------------------------------------------------------
constexpr int kMaxCellTris = 10;
struct CaseTable {
  int32_t count[256];
  int8_t tri[256][kMaxCellTris][3];
};

int edge_of(int c0, int c1) {  // c0 and c1 differ in one bit
  const int x = c0 ^ c1, a = x == 1 ? 0 : (x == 2 ? 1 : 2), lo = std::min(c0, c1);
  int m = 0;
  for (int c = 0; c < lo; c++)
    if (!(c & x)) m++;
  return 4 * a + m;
}

// faces of the cell (as a bit set) that contain edge e
int edge_faces(int e) {
  const int a = e >> 2, m = e & 3;
  int c0 = 0, seen = 0;
  for (int c = 0; c < 8; c++)
    if (!((c >> a) & 1)) {
      if (seen == m) {
        c0 = c;
        break;
      }
      seen++;
    }
  int f = 0;
  for (int b = 0; b < 3; b++)
    if (b != a) f |= 1 << (2 * b + ((c0 >> b) & 1));  // face (axis b, side bit b of c0)
  return f;
}

// Triangulates loop[i..j] (edge i-j is a side of the polygon) without a
// diagonal inside a face of the cell; false if there is none.
bool triangulate(const std::vector<int>& loop, int i, int j, std::vector<std::array<int, 3>>& out) {
  if (j - i < 2) return true;
  for (int k = i + 1; k < j; k++) {
    if (k - i > 1 && (edge_faces(loop[size_t(i)]) & edge_faces(loop[size_t(k)]))) continue;
    if (j - k > 1 && (edge_faces(loop[size_t(k)]) & edge_faces(loop[size_t(j)]))) continue;
    std::vector<std::array<int, 3>> t;
    if (!triangulate(loop, i, k, t)) continue;
    t.push_back({loop[size_t(i)], loop[size_t(k)], loop[size_t(j)]});
    if (!triangulate(loop, k, j, t)) continue;
    out.insert(out.end(), t.begin(), t.end());
    return true;
  }
  return false;
}

CaseTable make_case_table() {
  CaseTable T{};
  // faces: axis a, side s; corners counter-clockwise seen from outside
  // (e_u x e_v = outward normal)
  std::vector<std::array<int, 4>> faces;
  for (int a = 0; a < 3; a++)
    for (int s = 0; s < 2; s++) {
      int u = (a + 1) % 3, v = (a + 2) % 3;  // e_u x e_v = +e_a
      if (s == 0) std::swap(u, v);
      std::array<int, 4> q{};
      const int uv[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
      for (int k = 0; k < 4; k++) q[k] = (s << a) | (uv[k][0] << u) | (uv[k][1] << v);
      faces.push_back(q);
    }
  for (int cs = 0; cs < 256; cs++) {
    auto inside = [&](int c) { return (cs >> c) & 1; };
    int next[12];
    std::fill(next, next + 12, -1);
    for (const auto& q : faces) {
      int kind[4];
      for (int k = 0; k < 4; k++) {
        const int c0 = q[k], c1 = q[(k + 1) % 4];
        kind[k] = inside(c0) == inside(c1) ? 0 : (inside(c0) ? 1 : 2);  // 1 exit, 2 entry
      }
      for (int k = 0; k < 4; k++) {
        if (kind[k] != 1) continue;
        // the entry that precedes this exit, walking backwards
        int j = (k + 3) % 4;
        while (kind[j] != 2) j = (j + 3) % 4;
        const int from = edge_of(q[k], q[(k + 1) % 4]), to = edge_of(q[j], q[(j + 1) % 4]);
        next[from] = to;
      }
    }
    bool seen[12] = {};
    int nt = 0;
    for (int e = 0; e < 12; e++) {
      if (next[e] < 0 || seen[e]) continue;
      std::vector<int> loop;
      for (int f = e; !seen[f]; f = next[f]) {
        seen[f] = true;
        loop.push_back(f);
        if (next[f] < 0) throw std::logic_error("open contour in the case table");
      }
      std::vector<std::array<int, 3>> tl;
      if (!triangulate(loop, 0, int(loop.size()) - 1, tl)) throw std::logic_error("no face-free triangulation");
      for (const auto& t : tl) {
        if (nt >= kMaxCellTris) throw std::logic_error("case table overflow");
        for (int w = 0; w < 3; w++) T.tri[cs][nt][w] = int8_t(t[size_t(w)]);
        nt++;
      }
    }
    T.count[cs] = nt;
  }
  return T;
}

const CaseTable& case_table() {
  static const CaseTable T = make_case_table();
  return T;
}

// ---------------------------------------------------------------------------
// Device side
// ---------------------------------------------------------------------------
struct V3 {
  double x, y, z;
};
__device__ __forceinline__ V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 vadd(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 vscale(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double vlen(V3 a) { return __dsqrt_rn(vdot(a, a)); }

// order-preserving int key of a float (smaller float, smaller key)
__device__ __forceinline__ int fkey(float f) {
  const int i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float funkey(int k) { return __int_as_float(k >= 0 ? k : k ^ 0x7fffffff); }

struct Grid {
  double lox, loy, loz, voxel;
  int nx, ny, nz;
};

__device__ __forceinline__ V3 grid_point(const Grid& g, int i, int j, int k) {
  return {g.lox + double(i) * g.voxel, g.loy + double(j) * g.voxel, g.loz + double(k) * g.voxel};
}

// MC/src/sim.cpp:669-705: one workgroup per shape (beams' capsules, then the
// joints' spheres), its box of grid points, min(distance - radius) as float
__global__ __launch_bounds__(256) void k_stl_raster(Grid g, const double* __restrict__ x, const int* __restrict__ ba,
                                                    const int* __restrict__ bb, const double* __restrict__ rad,
                                                    int N, int B, double jr, int* __restrict__ field) {
  const int s = blockIdx.x;
  V3 a, b;
  double r;
  if (s < B) {
    const int ia = ba[s], ib = bb[s];
    a = {x[3 * size_t(ia)], x[3 * size_t(ia) + 1], x[3 * size_t(ia) + 2]};
    b = {x[3 * size_t(ib)], x[3 * size_t(ib) + 1], x[3 * size_t(ib) + 2]};
    r = rad[s];
  } else {
    const int i = s - B;
    a = b = {x[3 * size_t(i)], x[3 * size_t(i) + 1], x[3 * size_t(i) + 2]};
    r = jr;
  }
  const double pad = r + 1.8 * g.voxel;
  const V3 b0{fmin(a.x, b.x) - pad, fmin(a.y, b.y) - pad, fmin(a.z, b.z) - pad};
  const V3 b1{fmax(a.x, b.x) + pad, fmax(a.y, b.y) + pad, fmax(a.z, b.z) + pad};
  const int i0 = max(0, int(floor((b0.x - g.lox) / g.voxel)));
  const int j0 = max(0, int(floor((b0.y - g.loy) / g.voxel)));
  const int k0 = max(0, int(floor((b0.z - g.loz) / g.voxel)));
  const int i1 = min(g.nx - 1, int(ceil((b1.x - g.lox) / g.voxel)));
  const int j1 = min(g.ny - 1, int(ceil((b1.y - g.loy) / g.voxel)));
  const int k1 = min(g.nz - 1, int(ceil((b1.z - g.loz) / g.voxel)));
  if (i1 < i0 || j1 < j0 || k1 < k0) return;
  const int wi = i1 - i0 + 1, wk = k1 - k0 + 1;
  const long count = long(wi) * wk * (j1 - j0 + 1);
  const V3 ab = vsub(b, a);
  const double ab2 = fmax(1e-30, vdot(ab, ab));
  for (long t = threadIdx.x; t < count; t += blockDim.x) {
    const int i = i0 + int(t % wi), k = k0 + int((t / wi) % wk), j = j0 + int(t / (long(wi) * wk));
    const V3 p = grid_point(g, i, j, k);
    double d;
    if (s < B) {
      double u = vdot(vsub(p, a), ab) / ab2;
      u = fmax(0.0, fmin(1.0, u));
      d = vlen(vsub(p, vadd(a, vscale(ab, u))));
    } else {
      d = vlen(vsub(p, a));
    }
    const float sdf = float(d - r);
    atomicMin(field + (size_t(j) * g.nz + k) * g.nx + i, fkey(sdf));
  }
}

__global__ void k_stl_fill(int* __restrict__ field, size_t n, int key) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) field[i] = key;
}

struct DevTable {
  int32_t count[256];
  int8_t tri[256][kMaxCellTris][3];
};

__device__ __forceinline__ int cell_case(const Grid& g, const int* __restrict__ field, int i, int j, int k,
                                         double (&v)[8]) {
  int cs = 0;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const int ii = i + (c & 1), jj = j + ((c >> 1) & 1), kk = k + ((c >> 2) & 1);
    v[c] = double(funkey(field[(size_t(jj) * g.nz + kk) * g.nx + ii]));
    if (v[c] < 0.0) cs |= 1 << c;
  }
  return cs;
}

// cell id in the reference's loop order: y outer, z, then x inner
__device__ __forceinline__ void cell_ijk(const Grid& g, size_t id, int& i, int& j, int& k) {
  const size_t cx = size_t(g.nx - 1), cz = size_t(g.nz - 1);
  i = int(id % cx);
  k = int((id / cx) % cz);
  j = int(id / (cx * cz));
}

__global__ __launch_bounds__(256) void k_stl_count(Grid g, const int* __restrict__ field, size_t cells,
                                                   const DevTable* __restrict__ T, unsigned* __restrict__ counts) {
  const size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= cells) return;
  int i, j, k;
  cell_ijk(g, id, i, j, k);
  double v[8];
  counts[id] = unsigned(T->count[cell_case(g, field, i, j, k, v)]);
}

// MC/src/sim.cpp:627-640: interpolation along a cell edge, snapped to 1e-8 m
__device__ __forceinline__ V3 iso_vertex(V3 a, V3 b, double va, double vb) {
  const double den = va - vb;
  double t = fabs(den) > 1e-30 ? va / den : 0.5;
  t = fmax(0.0, fmin(1.0, t));
  V3 p = vadd(a, vscale(vsub(b, a), t));
  const double q = 1e-8;
  p.x = round(p.x / q) * q;
  p.y = round(p.y / q) * q;
  p.z = round(p.z / q) * q;
  return p;
}

__global__ __launch_bounds__(256) void k_stl_emit(Grid g, const int* __restrict__ field, size_t cells,
                                                  const DevTable* __restrict__ T, const unsigned* __restrict__ offs,
                                                  double* __restrict__ tris) {
  const size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= cells) return;
  int i, j, k;
  cell_ijk(g, id, i, j, k);
  double v[8];
  const int cs = cell_case(g, field, i, j, k, v);
  const int nt = T->count[cs];
  if (nt == 0) return;
  V3 p[8];
#pragma unroll
  for (int c = 0; c < 8; c++) p[c] = grid_point(g, i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1));
  double* out = tris + 9 * size_t(offs[id]);
  for (int t = 0; t < nt; t++) {
    for (int w = 0; w < 3; w++) {
      const int e = T->tri[cs][t][w], a = e >> 2, m = e & 3;
      // the m-th corner without bit a, then that corner with bit a
      int c0 = 0, seen = 0;
      for (int c = 0; c < 8; c++)
        if (!((c >> a) & 1)) {
          if (seen == m) {
            c0 = c;
            break;
          }
          seen++;
        }
      const int c1 = c0 | (1 << a);
      const V3 q = iso_vertex(p[c0], p[c1], v[c0], v[c1]);
      out[9 * t + 3 * w] = q.x;
      out[9 * t + 3 * w + 1] = q.y;
      out[9 * t + 3 * w + 2] = q.z;
    }
  }
}

// ---------------------------------------------------------------------------
// Host side: audit, repair, output
// ---------------------------------------------------------------------------
struct Vec {
  double x, y, z;
};
Vec operator-(Vec a, Vec b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
Vec cross(Vec a, Vec b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double dot(Vec a, Vec b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
struct Tri {
  Vec a, b, c;
};

// vertices on the 1e-8 m grid (MC/src/sim.cpp:611-616)
struct QV {
  int64_t x, y, z;
  bool operator==(const QV& o) const { return x == o.x && y == o.y && z == o.z; }
  bool operator<(const QV& o) const { return x != o.x ? x < o.x : (y != o.y ? y < o.y : z < o.z); }
};
QV quant(const Vec& p) {
  const double q = 1e-8;
  return {int64_t(std::llround(p.x / q)), int64_t(std::llround(p.y / q)), int64_t(std::llround(p.z / q))};
}
Vec dequant(const QV& q) { return {double(q.x) * 1e-8, double(q.y) * 1e-8, double(q.z) * 1e-8}; }
uint64_t mix(uint64_t v) {
  v += 0x9e3779b97f4a7c15ULL;
  v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
  v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
  return v ^ (v >> 31);
}
struct QVHash {
  size_t operator()(const QV& v) const { return size_t(mix(uint64_t(v.x) ^ mix(uint64_t(v.y) ^ mix(uint64_t(v.z))))); }
};
struct QEdge {
  QV a, b;  // a < b
  bool operator==(const QEdge& o) const { return a == o.a && b == o.b; }
};
struct QEdgeHash {
  size_t operator()(const QEdge& e) const { return QVHash()(e.a) * 31 + QVHash()(e.b); }
};
QEdge qedge(const QV& p, const QV& q) { return q < p ? QEdge{q, p} : QEdge{p, q}; }

double signed_volume(const std::vector<Tri>& T) {
  double v = 0.0;
  for (const Tri& t : T) v += dot(t.a, cross(t.b, t.c)) / 6.0;
  return v;
}
void orient_outward(std::vector<Tri>& T, double& vol) {
  vol = signed_volume(T);
  if (vol < 0.0) {
    for (Tri& t : T) std::swap(t.b, t.c);
    vol = -vol;
  }
}

// Closes the boundary loops of the mesh, as MC/src/sim.cpp:758-829 does: a
// loop of three vertices becomes a triangle, a longer one a fan around its
// centroid; at most five passes.  Returns the boundary edges found first and
// the loops closed.
void repair(std::vector<Tri>& T, size_t& initial_boundary, size_t& loops_closed) {
  initial_boundary = 0;
  loops_closed = 0;
  for (int pass = 0; pass < 5; pass++) {
    // directed boundary edges: an edge used once, in the direction it was used
    std::unordered_map<QEdge, std::pair<int, std::pair<QV, QV>>, QEdgeHash> uses;
    uses.reserve(T.size() * 2);
    auto reg = [&](const Vec& p, const Vec& q) {
      const QV a = quant(p), b = quant(q);
      auto& r = uses[qedge(a, b)];
      if (r.first == 0) r.second = {a, b};
      r.first++;
    };
    for (const Tri& t : T) {
      reg(t.a, t.b);
      reg(t.b, t.c);
      reg(t.c, t.a);
    }
    // boundary edges in a fixed order (sorted), each followed back to front
    std::vector<std::pair<QV, QV>> bnd;
    for (const auto& kv : uses)
      if (kv.second.first == 1) bnd.push_back(kv.second.second);
    std::sort(bnd.begin(), bnd.end(), [](const auto& u, const auto& v) {
      return u.first < v.first || (u.first == v.first && u.second < v.second);
    });
    if (pass == 0) initial_boundary = bnd.size();
    if (bnd.empty()) break;
    std::unordered_map<QV, QV, QVHash> prev;  // to -> from
    for (const auto& e : bnd) prev[e.second] = e.first;
    std::unordered_set<QV, QVHash> used;
    size_t closed = 0;
    for (const auto& e : bnd) {
      const QV start = e.second;
      if (used.count(start)) continue;
      std::vector<QV> loop;
      QV cur = start;
      for (size_t guard = 0; guard <= prev.size(); guard++) {
        if (used.count(cur) && !(cur == start)) break;
        used.insert(cur);
        loop.push_back(cur);
        auto it = prev.find(cur);
        if (it == prev.end()) break;
        cur = it->second;
        if (cur == start) break;
      }
      if (loop.size() < 3 || !(cur == start)) continue;
      closed++;
      std::vector<Vec> P;
      Vec c{0, 0, 0};
      for (const QV& q : loop) {
        P.push_back(dequant(q));
        c = {c.x + P.back().x, c.y + P.back().y, c.z + P.back().z};
      }
      c = dequant(quant({c.x / double(P.size()), c.y / double(P.size()), c.z / double(P.size())}));
      if (P.size() == 3)
        T.push_back({P[0], P[1], P[2]});
      else
        for (size_t i = 0; i < P.size(); i++) T.push_back({P[i], P[(i + 1) % P.size()], c});
    }
    loops_closed += closed;
    if (closed == 0) break;
  }
}

void write_stl(const std::string& path, const std::vector<Tri>& T) {
  std::ofstream out(path, std::ios::binary);
  if (!out) throw std::runtime_error("cannot write STL " + path);
  char header[80] = {};
  const char* title = "Metamaterial engineering v2 - watertight implicit union";
  std::memcpy(header, title, std::min<size_t>(std::strlen(title), sizeof header));
  out.write(header, sizeof header);
  const uint32_t n = uint32_t(T.size());
  out.write(reinterpret_cast<const char*>(&n), sizeof n);
  auto put = [&](double v) {
    const float f = float(v);
    out.write(reinterpret_cast<const char*>(&f), sizeof f);
  };
  for (const Tri& t : T) {
    Vec nrm = cross(t.b - t.a, t.c - t.a);
    const double l2 = dot(nrm, nrm);
    nrm = (std::isfinite(l2) && l2 > 1e-30) ? Vec{nrm.x / std::sqrt(l2), nrm.y / std::sqrt(l2), nrm.z / std::sqrt(l2)}
                                             : Vec{0, 1, 0};
    put(nrm.x), put(nrm.y), put(nrm.z);
    for (const Vec& p : {t.a, t.b, t.c}) put(p.x * 1000.0), put(p.y * 1000.0), put(p.z * 1000.0);
    const uint16_t attr = 0;
    out.write(reinterpret_cast<const char*>(&attr), sizeof attr);
  }
  if (!out) throw std::runtime_error("cannot write STL " + path);
}

template <class X>
struct DevBuf {
  X* p = nullptr;
  explicit DevBuf(size_t n) { STLCALL(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(X))); }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

void export_stl(const mmc_lattice* L, double jr, int segments, const char* path, mmc_stl_report* rep) {
  if (!L || !path) throw std::invalid_argument("NULL argument");
  const int N = L->n_nodes, B = L->n_beams;
  if (N <= 0 || B <= 0) throw std::runtime_error("cannot mesh empty lattice");
  const std::filesystem::path fp(path);
  if (fp.has_parent_path()) std::filesystem::create_directories(fp.parent_path());
  // MC/src/sim.cpp:651-667: voxel size, bounds and grid
  const double voxel = std::max(0.00030, std::min(0.00072, 0.006 / std::max(8, segments)));
  const double inf = std::numeric_limits<double>::infinity();
  double lo[3] = {inf, inf, inf};
  const double m = jr + 2.5 * voxel;
  for (int i = 0; i < N; i++)
    for (int d = 0; d < 3; d++) lo[d] = std::min(lo[d], L->x[3 * size_t(i) + d] - m);
  double hi[3] = {-inf, -inf, -inf};
  for (int i = 0; i < N; i++)
    for (int d = 0; d < 3; d++) hi[d] = std::max(hi[d], L->x[3 * size_t(i) + d] + m);
  auto cnt = [&](double extent) { return int(std::ceil(extent / voxel)) + 1; };
  Grid g{lo[0], lo[1], lo[2], voxel, cnt(hi[0] - lo[0]), cnt(hi[1] - lo[1]), cnt(hi[2] - lo[2])};
  const size_t points = size_t(g.nx) * size_t(g.ny) * size_t(g.nz);
  if (int64_t(g.nx) * g.ny * g.nz > 12000000LL) throw std::runtime_error("implicit STL grid exceeds safety limit");
  // device: lattice, field, case table
  DevBuf<double> dx{3 * size_t(N)}, drad{size_t(B)};
  DevBuf<int> dba{size_t(B)}, dbb{size_t(B)}, field{points};
  STLCALL(hipMemcpy(dx.p, L->x, 3 * size_t(N) * sizeof(double), hipMemcpyHostToDevice));
  STLCALL(hipMemcpy(drad.p, L->radius, size_t(B) * sizeof(double), hipMemcpyHostToDevice));
  STLCALL(hipMemcpy(dba.p, L->beam_a, size_t(B) * sizeof(int), hipMemcpyHostToDevice));
  STLCALL(hipMemcpy(dbb.p, L->beam_b, size_t(B) * sizeof(int), hipMemcpyHostToDevice));
  for (int k = 0; k < B; k++)
    if (L->beam_a[k] < 0 || L->beam_a[k] >= N || L->beam_b[k] < 0 || L->beam_b[k] >= N)
      throw std::invalid_argument("beam endpoint out of range");
  const float finf = std::numeric_limits<float>::infinity();
  int inf_key;
  std::memcpy(&inf_key, &finf, sizeof inf_key);  // +inf is positive: its key is its bit pattern
  k_stl_fill<<<unsigned((points + 255) / 256), 256>>>(field.p, points, inf_key);
  STLCALL(hipGetLastError());
  k_stl_raster<<<unsigned(B + N), 256>>>(g, dx.p, dba.p, dbb.p, drad.p, N, B, jr, field.p);
  STLCALL(hipGetLastError());
  const CaseTable& ct = case_table();
  DevTable ht;
  std::memcpy(ht.count, ct.count, sizeof ht.count);
  std::memcpy(ht.tri, ct.tri, sizeof ht.tri);
  DevBuf<DevTable> dt(1);
  STLCALL(hipMemcpy(dt.p, &ht, sizeof ht, hipMemcpyHostToDevice));
  // count, scan, emit over the cells in the reference's order
  const size_t cells = size_t(g.nx - 1) * size_t(g.ny - 1) * size_t(g.nz - 1);
  DevBuf<unsigned> counts(cells + 1), offs(cells + 1);
  const unsigned cblocks = unsigned((cells + 255) / 256);
  if (cells > 0) {
    k_stl_count<<<cblocks, 256>>>(g, field.p, cells, dt.p, counts.p);
    STLCALL(hipGetLastError());
  }
  STLCALL(hipMemset(counts.p + cells, 0, sizeof(unsigned)));
  size_t tmp_bytes = 0;
  STLCALL(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, counts.p, offs.p, int(cells + 1)));
  DevBuf<char> tmp(tmp_bytes);
  STLCALL(hipcub::DeviceScan::ExclusiveSum(tmp.p, tmp_bytes, counts.p, offs.p, int(cells + 1)));
  unsigned total = 0;
  STLCALL(hipMemcpy(&total, offs.p + cells, sizeof total, hipMemcpyDeviceToHost));
  std::vector<Tri> T(total);
  if (total > 0) {
    DevBuf<double> dtris(9 * size_t(total));
    k_stl_emit<<<cblocks, 256>>>(g, field.p, cells, dt.p, offs.p, dtris.p);
    STLCALL(hipGetLastError());
    static_assert(sizeof(Tri) == 9 * sizeof(double), "triangle layout");
    STLCALL(hipMemcpy(T.data(), dtris.p, 9 * size_t(total) * sizeof(double), hipMemcpyDeviceToHost));
  }
  // MC/src/sim.cpp:751-756: zero-area triangles out, outward orientation
  T.erase(std::remove_if(T.begin(), T.end(),
                         [](const Tri& t) {
                           const Vec c = cross(t.b - t.a, t.c - t.a);
                           return dot(c, c) < 1e-22;
                         }),
          T.end());
  double vol = 0.0;
  orient_outward(T, vol);
  size_t initial_boundary = 0, loops = 0;
  repair(T, initial_boundary, loops);
  orient_outward(T, vol);
  // edge audit (MC/src/sim.cpp:837-848)
  std::unordered_map<QEdge, int, QEdgeHash> edges;
  edges.reserve(T.size() * 2);
  for (const Tri& t : T) {
    const QV a = quant(t.a), b = quant(t.b), c = quant(t.c);
    ++edges[qedge(a, b)];
    ++edges[qedge(b, c)];
    ++edges[qedge(c, a)];
  }
  size_t boundary = 0, nonmanifold = 0;
  for (const auto& kv : edges) {
    if (kv.second == 1) ++boundary;
    else if (kv.second != 2) ++nonmanifold;
  }
  write_stl(path, T);
  std::filesystem::path report = fp;
  report.replace_extension(".validation.txt");
  std::ofstream val(report);
  val << "Implicit-union STL validation\n"
      << "voxel_mm=" << voxel * 1000.0 << '\n'
      << "grid=" << g.nx << 'x' << g.ny << 'x' << g.nz << '\n'
      << "triangles=" << T.size() << '\n'
      << "unique_edges=" << edges.size() << '\n'
      << "initial_ambiguous_boundary_edges=" << initial_boundary << '\n'
      << "repaired_boundary_loops=" << loops << '\n'
      << "boundary_edges=" << boundary << '\n'
      << "nonmanifold_edges=" << nonmanifold << '\n'
      << "signed_volume_mm3=" << std::abs(vol) * 1e9 << '\n'
      << "watertight=" << ((boundary == 0 && nonmanifold == 0) ? "PASS" : "FAIL") << '\n';
  if (rep) {
    rep->voxel = voxel;
    rep->signed_volume = vol;
    rep->nx = g.nx, rep->ny = g.ny, rep->nz = g.nz;
    rep->triangles = int64_t(T.size());
    rep->unique_edges = int64_t(edges.size());
    rep->initial_boundary_edges = int64_t(initial_boundary);
    rep->repaired_loops = int64_t(loops);
    rep->boundary_edges = int64_t(boundary);
    rep->nonmanifold_edges = int64_t(nonmanifold);
    rep->watertight = boundary == 0 && nonmanifold == 0;
  }
  if (boundary != 0 || nonmanifold != 0) throw std::runtime_error("implicit STL failed watertight edge audit");
}

}  // namespace

extern "C" {

int mmc_export_stl(const mmc_lattice* lattice, double joint_radius, int32_t radial_segments, const char* path,
                   mmc_stl_report* report) {
  try {
    export_stl(lattice, joint_radius, radial_segments, path, report);
    g_stl_error.clear();
    return MMC_OK;
  } catch (const HipError& e) {
    g_stl_error = e.what();
    return MMC_ERR_HIP;
  } catch (const std::invalid_argument& e) {
    g_stl_error = e.what();
    return MMC_ERR_ARG;
  } catch (const std::exception& e) {
    g_stl_error = e.what();
    return MMC_ERR_STATE;
  }
}

const char* mmc_stl_last_error(void) { return g_stl_error.c_str(); }

void mmc_stl_case_table(int32_t* counts, int8_t* tris) {
  const CaseTable& t = case_table();
  if (counts) std::memcpy(counts, t.count, sizeof t.count);
  if (tris) std::memcpy(tris, t.tri, sizeof t.tri);
}

}  // extern "C"
