// generate.hip — lattice construction on the device (see generate.hpp).
//
// The graded BCC family follows projects/metamaterial-compression/
// industry-grade/src/sim.cpp:51-132 (MC/ below) operation for operation and
// the simple cubic family this package's host generator (lattice.cpp), with no
// fused multiply-adds (built with -ffp-contract=off), so that what is made of
// sums, products, quotients and square roots (all correctly rounded) equals
// the host lattice bit for bit:
//   * the RNG stream (MC/include/core.hpp:46) is sequential, so one thread
//     per lattice walks it and stores its draws; every other step is one
//     thread per node, per physical strut or per beam;
//   * the few libm values that depend only on a layer or a segment position
//     (sin of the layer envelope, sin of the strut bow, the radius grading
//     pow(1 - y, 0.72)) come from host tables, so they are the host's;
//   * what needs libm per item is the device's own: the Box-Muller jitter of
//     the cubic family (log, cos, sin), which may differ from glibc's in the
//     last place; the second moment pi r^4 / 4 takes r^4 rounded correctly
//     (pow4), which glibc's pow misses by one unit in about one case in a
//     thousand.
// Then, per node, the beam incidence ordered by beam, the lumped masses of
// distributeMass (MC/src/sim.cpp:134) summed in the reference's beam order,
// the lattice's total mass in the reference's order, the anti-drift locks, and
// the self-contact exclusion adjacency.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

#include "generate.hpp"

namespace mmc {
namespace {

constexpr double kPi = 3.1415926535897932384626433832795;
constexpr int kT = 256;

#define GEN_HIP(call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      throw std::runtime_error(std::string("lattice generation: ") + #call + ": " +         \
                               hipGetErrorString(e_));                                     \
  } while (0)

// One lattice of a generation launch.
struct GenLat {
  int family, nx, ny, nz, segs, radius_mode;
  int n0, N, b0, B;
  int base;          // nodes placed before the struts (BCC corners + centres, cubic all)
  int corners, cells, struts;
  long c0, d0;       // first exclusion pair, first draw
  int tab;           // first table entry
  double h, jr, width, height, depth, imperf, member_imperf, trigger, sqrt_trigger, rmin, rmax, sigma, density, jm;
  uint64_t seed;
};
// table layout per lattice: env_c[ny + 1], env_m[ny], grade_cell[2 ny],
// grade_face[2], bow1[5], bow2[5]

struct P3 {
  double x, y, z;
};
__device__ __forceinline__ P3 add(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ P3 sub(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ P3 mul(P3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double dot(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ P3 cross(P3 a, P3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ P3 unit(P3 v) {
  const double l2 = dot(v, v);
  return (isfinite(l2) && l2 > 1e-30) ? mul(v, 1.0 / sqrt(l2)) : P3{0, 1, 0};
}
__device__ __forceinline__ double clamp01(double v) { return fmax(0.0, fmin(1.0, v)); }
__host__ __device__ __forceinline__ uint64_t splitmix(uint64_t v) {
  v += 0x9e3779b97f4a7c15ULL;
  v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
  v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
  return v ^ (v >> 31);
}
__device__ __forceinline__ double uniform(uint32_t u) { return (double(u) + 0.5) / 4294967296.0; }
__device__ __forceinline__ double signed_uniform(uint32_t u) { return 2 * uniform(u) - 1; }

// the lattice holding item g of a range whose lattice starts are first[0..n)
__device__ __forceinline__ int owner(const int* __restrict__ first, int n, long g) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ void put3(double* p, int stride, int i, P3 v) {
  p[i] = v.x;
  p[size_t(stride) + i] = v.y;
  p[2 * size_t(stride) + i] = v.z;
}
__device__ __forceinline__ P3 get3(const double* p, int stride, int i) {
  return {p[i], p[size_t(stride) + i], p[2 * size_t(stride) + i]};
}

// The RNG stream of each lattice (MC/include/core.hpp:46): state splitmix64(seed),
// each draw advances it by splitmix64 and takes its high 32 bits.
__global__ __launch_bounds__(kT) void k_draws(const GenLat* __restrict__ L, int n, const long* __restrict__ count,
                                               uint32_t* __restrict__ draws) {
  const int l = int(blockIdx.x) * kT + int(threadIdx.x);
  if (l >= n) return;
  uint64_t s = splitmix(L[l].seed);
  uint32_t* d = draws + L[l].d0;
  for (long k = 0; k < count[l]; k++) {
    s = splitmix(s);
    d[k] = uint32_t(s >> 32);
  }
}

// Corner and BCC-centre nodes (MC/src/sim.cpp:60-76), or every node of a cubic lattice.
__global__ __launch_bounds__(kT) void k_base_nodes(const GenLat* __restrict__ L, int n, const int* __restrict__ first,
                                                    long total, const double* __restrict__ tab,
                                                    const uint32_t* __restrict__ draws, GenOut o, int N) {
  const long g = long(blockIdx.x) * kT + threadIdx.x;
  if (g >= total) return;
  const int li = owner(first, n, g);
  const GenLat& G = L[li];
  const int c = int(g - first[li]);
  const double* T = tab + G.tab;
  const uint32_t* D = draws + G.d0;
  const int nx = G.nx, ny = G.ny, nz = G.nz;
  const P3 origin{-0.5 * G.width, G.jr, -0.5 * G.depth};
  P3 p;
  int layer;
  if (G.family == MMC_LATTICE_BCC && c >= G.corners) {
    // BCC centre with its coherent alternating offsets
    const int m = c - G.corners, i = m % nx, k = (m / nx) % nz, j = m / (nx * nz);
    p = add(origin, P3{(i + 0.5) * G.h, (j + 0.5) * G.h, (k + 0.5) * G.h});
    const double yf = (j + 0.5) / double(ny), env = T[(ny + 1) + j], sg = ((i + k + j) & 1) ? 1.0 : -1.0;
    const double sx = (j & 1) ? -1.0 : 1.0, sz = ((j / 2) & 1) ? -1.0 : 1.0;
    const double ux = signed_uniform(D[2 * long(c)]), uz = signed_uniform(D[2 * long(c) + 1]);
    p.x += G.h * (0.082 * (1 - 0.32 * yf) * sx + 0.018 * env * sg) + 1.2 * G.imperf * env * ux;
    p.z += G.h * (0.032 * (0.65 + 0.35 * (1 - yf)) * sz - 0.012 * env * sg) + 1.2 * G.imperf * env * uz;
    layer = j;
  } else {
    const int i = c % (nx + 1), k = (c / (nx + 1)) % (nz + 1), j = c / ((nx + 1) * (nz + 1));
    p = add(origin, P3{i * G.h, j * G.h, k * G.h});
    if (G.family == MMC_LATTICE_CUBIC && G.sigma > 0) {
      // Gaussian jitter, Box-Muller on the stream: four draws per node
      const long d = 4 * long(c);
      double g0, g1, g2;
      {
        const double u1 = uniform(D[d]), u2 = uniform(D[d + 1]);
        const double rr = sqrt(-2.0 * log(u1));
        g0 = rr * cos(2 * kPi * u2);
        g1 = rr * sin(2 * kPi * u2);
      }
      {
        const double u1 = uniform(D[d + 2]), u2 = uniform(D[d + 3]);
        const double rr = sqrt(-2.0 * log(u1));
        g2 = rr * cos(2 * kPi * u2);
      }
      p.x += G.sigma * g0;
      p.z += G.sigma * g1;
      if (j > 0 && j < ny) p.y += G.sigma * g2;
    } else {
      // corner imperfection pattern (MC/src/sim.cpp:65-67)
      const double env = T[j], sg = ((i + 2 * k + 3 * j) & 1) ? 1.0 : -1.0;
      p.x += G.imperf * env * (0.55 * sg + 0.45 * signed_uniform(D[2 * long(c)]));
      p.z += G.imperf * env * (0.35 * sg + 0.65 * signed_uniform(D[2 * long(c) + 1]));
    }
    layer = j;
  }
  const int gi = G.n0 + c;
  put3(o.x, N, gi, p);
  o.q[gi] = 1.0;
  o.q[size_t(N) + gi] = 0.0;
  o.q[2 * size_t(N) + gi] = 0.0;
  o.q[3 * size_t(N) + gi] = 0.0;
  o.node_layer[gi] = layer;
}

// r^4 to within a hair of correct rounding: r^2 = h2 + e2 exactly (fma), then
// (h2 + e2)^2 = h2^2 + 2 h2 e2 + e2^2 with h2^2 split exactly again; the
// dropped e2^2 and the rounding of the small terms stay far below half a unit
// of the result.  (libm pow(r, 4) is not correctly rounded on either side.)
__device__ __forceinline__ double pow4(double r) {
  const double h2 = r * r, e2 = fma(r, r, -h2);
  const double h4 = h2 * h2, e4 = fma(h2, h2, -h4);
  return h4 + (e4 + 2.0 * h2 * e2);
}

// One beam between global nodes u and v (add_beam / element of MC/src/sim.cpp:114).
__device__ void put_beam(const GenOut& o, int N, int B, int k, int u, int v, int layer, int family, double r) {
  const P3 d = sub(get3(o.x, N, v), get3(o.x, N, u));
  const P3 t = unit(d);
  o.a[k] = u;
  o.b[k] = v;
  o.beam_layer[k] = layer;
  o.beam_family[k] = family;
  o.L0[k] = sqrt(dot(d, d));
  put3(o.e0, B, k, t);
  o.radius[k] = r;
  o.area[k] = kPi * r * r;
  o.inertia[k] = 0.25 * kPi * pow4(r);
}

// Physical struts of the BCC family (addPhysical, MC/src/sim.cpp:78-118): cell
// struts centre -> 8 corners, then the two face grids; each adds its bowed
// middle nodes, its chain of beams and its exclusion pairs.
__global__ __launch_bounds__(kT) void k_struts(const GenLat* __restrict__ L, int n, const int* __restrict__ first,
                                                long total, const double* __restrict__ tab, GenOut o, int N, int B,
                                                int2* __restrict__ pairs) {
  const long g = long(blockIdx.x) * kT + threadIdx.x;
  if (g >= total) return;
  const int li = owner(first, n, g);
  const GenLat& G = L[li];
  const int s = int(g - first[li]);
  const int nx = G.nx, ny = G.ny, nz = G.nz;
  auto corner = [&](int i, int j, int k) { return (j * (nz + 1) + k) * (nx + 1) + i; };
  const double* T = tab + G.tab;
  int u, v, layer, family;
  double scale, gy;
  if (s < 8 * G.cells) {
    const int cell = s / 8, q = s % 8, i = cell % nx, k = (cell / nx) % nz, j = cell / (nx * nz);
    const int dx = q & 1, dy = (q >> 1) & 1, dz = (q >> 2) & 1;
    u = G.corners + cell;
    v = corner(i + dx, j + dy, k + dz);
    layer = j;
    family = q;
    scale = 1.0;
    gy = T[(ny + 1) + ny + 2 * j + dy];
  } else {
    int e = s - 8 * G.cells;
    const int per_face = (nz + 1) * nx + (nx + 1) * nz;
    const int fi = e / per_face, face = fi ? ny : 0;
    e -= fi * per_face;
    if (e < (nz + 1) * nx) {
      const int k = e / nx, i = e % nx;
      u = corner(i, face, k);
      v = corner(i + 1, face, k);
      family = 30;
    } else {
      e -= (nz + 1) * nx;
      const int i = e / nz, k = e % nz;
      u = corner(i, face, k);
      v = corner(i, face, k + 1);
      family = 31;
    }
    layer = min(max(face, 0), ny - 1);
    scale = 0.62;
    gy = T[(ny + 1) + ny + 2 * ny + fi];
  }
  const uint64_t key = (uint64_t(uint32_t(min(u, v))) << 32) | uint32_t(max(u, v));
  const P3 pu = get3(o.x, N, G.n0 + u), pv = get3(o.x, N, G.n0 + v);
  const P3 tangent = unit(sub(pv, pu));
  const uint64_t hs = splitmix(G.seed ^ key ^ uint64_t(family + 31) * 0x9e3779b97f4a7c15ULL);
  const P3 helper{double(int((hs >> 0) & 1023) - 511), double(int((hs >> 10) & 1023) - 511),
                  double(int((hs >> 20) & 1023) - 511)};
  P3 n1 = cross(tangent, helper);
  if (dot(n1, n1) < 1e-12) {
    const P3 up = fabs(tangent.z) < 0.8 ? P3{0, 0, 1} : P3{0, 1, 0};
    n1 = unit(cross(up, tangent));
  } else {
    n1 = unit(n1);
  }
  const P3 n2 = unit(cross(tangent, n1));
  const double lf = layer == 0 ? 2.60 : (layer == 1 ? 1.60 : 1.0);
  const double amp = G.member_imperf * lf * (0.55 + 0.90 * double((hs >> 32) & 65535) / 65535.0);
  const int segs = G.segs;
  const double* bow1 = T + (ny + 1) + ny + 2 * ny + 2;
  const double* bow2 = bow1 + 5;
  const int mid0 = G.base + s * (segs - 1);  // local index of the strut's first middle node
  for (int m = 1; m < segs; m++) {
    const double t = double(m) / segs;
    const P3 p = add(add(add(mul(pu, 1 - t), mul(pv, t)), mul(n1, bow1[m] * amp)), mul(n2, 0.22 * bow2[m] * amp));
    const int gi = G.n0 + mid0 + m - 1;
    put3(o.x, N, gi, p);
    o.q[gi] = 1.0;
    o.q[size_t(N) + gi] = 0.0;
    o.q[2 * size_t(N) + gi] = 0.0;
    o.q[3 * size_t(N) + gi] = 0.0;
    o.node_layer[gi] = layer;
  }
  // the graded radius (MC/src/sim.cpp:79-88): the y grade from the table,
  // the radial factor and the trigger layers here
  const P3 mp = mul(add(pu, pv), 0.5);
  const double radial = clamp01(sqrt(mp.x * mp.x + mp.z * mp.z) / (0.72 * fmax(G.width, G.depth)));
  double grade = gy;
  grade *= 0.94 + 0.12 * radial;
  if (layer == 0)
    grade *= G.trigger;
  else if (layer == 1)
    grade *= G.sqrt_trigger;
  else if (layer == 2)
    grade *= 0.88;
  const double tt = clamp01(grade);
  const double r = scale * (G.rmin * (1.0 - tt) + G.rmax * tt);
  int2* P = pairs + G.c0 + long(s) * (segs + 1);
  for (int e = 0; e < segs; e++) {
    const int cu = e == 0 ? u : mid0 + e - 1, cv = e == segs - 1 ? v : mid0 + e;
    put_beam(o, N, B, G.b0 + s * segs + e, G.n0 + cu, G.n0 + cv, layer, family, r);
    P[e] = int2{G.n0 + cu, G.n0 + cv};
  }
  P[segs] = int2{G.n0 + u, G.n0 + v};
}

// Beams of the cubic family: node by node (j, k, i), its +x, +y and +z edges.
__global__ __launch_bounds__(kT) void k_cubic_beams(const GenLat* __restrict__ L, int n, const int* __restrict__ first,
                                                     long total, const uint32_t* __restrict__ draws, GenOut o, int N,
                                                     int B, int2* __restrict__ pairs) {
  const long g = long(blockIdx.x) * kT + threadIdx.x;
  if (g >= total) return;
  const int li = owner(first, n, g);
  const GenLat& G = L[li];
  const int c = int(g - first[li]);
  const int nx = G.nx, ny = G.ny, nz = G.nz;
  const int i = c % (nx + 1), k = (c / (nx + 1)) % (nz + 1), j = c / ((nx + 1) * (nz + 1));
  // edges emitted before this node: full layers below, rows before it in its
  // layer, nodes before it in its row
  const long layer_full = long(nx + 1) * (nz + 1) + long(nx) * (nz + 1) + long(nx + 1) * nz;
  const long row = long(nx) + (j < ny ? nx + 1 : 0) + (nx + 1);
  long k0 = long(j) * layer_full + long(k) * row + long(i) * (1 + (j < ny) + (k < nz));
  const long node_draws = (G.sigma > 0 ? 4 : 2) * long(G.N);
  auto id = [&](int ii, int jj, int kk) { return (jj * (nz + 1) + kk) * (nx + 1) + ii; };
  const int nbr[3] = {i < nx ? id(i + 1, j, k) : -1, j < ny ? id(i, j + 1, k) : -1, k < nz ? id(i, j, k + 1) : -1};
  const P3 pc = get3(o.x, N, G.n0 + c);
  for (int f = 0; f < 3; f++) {
    if (nbr[f] < 0) continue;
    double r;
    if (G.radius_mode == 1) {
      r = G.rmin + (G.rmax - G.rmin) * uniform(draws[G.d0 + node_draws + k0]);
    } else {
      const P3 m = mul(add(pc, get3(o.x, N, G.n0 + nbr[f])), 0.5);
      const double t = clamp01((m.y - G.jr) / G.height);
      r = G.rmin * (1.0 - t) + G.rmax * t;
    }
    put_beam(o, N, B, G.b0 + int(k0), G.n0 + c, G.n0 + nbr[f], min(j, ny - 1), f, r);
    pairs[G.c0 + k0] = int2{G.n0 + c, G.n0 + nbr[f]};
    k0++;
  }
}

__global__ __launch_bounds__(kT) void k_zero_int(int* __restrict__ a, long n) {
  const long i = long(blockIdx.x) * kT + threadIdx.x;
  if (i < n) a[i] = 0;
}
// degrees (deg[N] stays 0, so that an exclusive scan of N + 1 entries gives
// the offsets and the total)
__global__ __launch_bounds__(kT) void k_beam_degree(const int* __restrict__ a, const int* __restrict__ b, int B,
                                                     int* __restrict__ deg) {
  const int k = int(blockIdx.x) * kT + threadIdx.x;
  if (k >= B) return;
  atomicAdd(deg + a[k], 1);
  atomicAdd(deg + b[k], 1);
}
__global__ __launch_bounds__(kT) void k_beam_fill(const int* __restrict__ a, const int* __restrict__ b, int B,
                                                   int* __restrict__ cur, int* __restrict__ inc) {
  const int k = int(blockIdx.x) * kT + threadIdx.x;
  if (k >= B) return;
  inc[atomicAdd(cur + a[k], 1)] = 2 * k;
  inc[atomicAdd(cur + b[k], 1)] = 2 * k + 1;
}
__global__ __launch_bounds__(kT) void k_pair_degree(const int2* __restrict__ pairs, long C, int* __restrict__ deg) {
  const long k = long(blockIdx.x) * kT + threadIdx.x;
  if (k >= C) return;
  atomicAdd(deg + pairs[k].x, 1);
  atomicAdd(deg + pairs[k].y, 1);
}
__global__ __launch_bounds__(kT) void k_pair_fill(const int2* __restrict__ pairs, long C, int* __restrict__ cur,
                                                   int* __restrict__ adj) {
  const long k = long(blockIdx.x) * kT + threadIdx.x;
  if (k >= C) return;
  adj[atomicAdd(cur + pairs[k].x, 1)] = pairs[k].y;
  adj[atomicAdd(cur + pairs[k].y, 1)] = pairs[k].x;
}
// per node: its lists sorted ascending (incidence by beam, adjacency by
// node); a repeated exclusion pair is flagged; the node's lumped mass summed
// over its beams in beam order, then its joint share (distributeMass,
// MC/src/sim.cpp:134)
__global__ __launch_bounds__(kT) void k_node_lists(int N, const int* __restrict__ inc_off, int* __restrict__ inc,
                                                    const int* __restrict__ adj_off, int* __restrict__ adj,
                                                    const double* __restrict__ area, const double* __restrict__ L0,
                                                    const GenLat* __restrict__ L, int n, const int* __restrict__ first,
                                                    double* __restrict__ mass, int* __restrict__ flags) {
  const int i = int(blockIdx.x) * kT + threadIdx.x;
  if (i >= N) return;
  auto sort = [](int* v, int m) {
    for (int p = 1; p < m; p++) {
      const int x = v[p];
      int q = p - 1;
      while (q >= 0 && v[q] > x) {
        v[q + 1] = v[q];
        q--;
      }
      v[q + 1] = x;
    }
  };
  int* in = inc + inc_off[i];
  const int di = inc_off[i + 1] - inc_off[i];
  sort(in, di);
  int* ad = adj + adj_off[i];
  const int da = adj_off[i + 1] - adj_off[i];
  sort(ad, da);
  for (int p = 1; p < da; p++)
    if (ad[p] == ad[p - 1]) atomicOr(flags, 1);
  const GenLat& G = L[owner(first, n, i)];
  double m = 0.0;
  for (int p = 0; p < di; p++) {
    const int k = in[p] >> 1;
    const double mb = G.density * area[k] * L0[k];
    m += 0.5 * mb;
  }
  m += G.jm;
  mass[i] = m;
}
// per lattice, one thread: the total mass in the reference's order (every
// beam, then every joint)
__global__ __launch_bounds__(kT) void k_total_mass(const GenLat* __restrict__ L, int n, const double* __restrict__ area,
                                                    const double* __restrict__ L0, double* __restrict__ out) {
  const int l = int(blockIdx.x) * kT + threadIdx.x;
  if (l >= n) return;
  const GenLat& G = L[l];
  double total = 0;
  for (int k = G.b0; k < G.b0 + G.B; k++) total += G.density * area[k] * L0[k];
  for (int i = 0; i < G.N; i++) total += G.jm;
  out[l] = total;
}
// per lattice, one block: the anti-drift nodes (MC/src/sim.cpp:134): among
// the layer-0 nodes the first nearest to (-w/2, -d/2) and the first nearest to
// (+w/2, -d/2) in the x-z plane; x and z of the first, z of the second locked
__global__ __launch_bounds__(kT) void k_locks(const GenLat* __restrict__ L, const double* __restrict__ x, int N,
                                               const int* __restrict__ layer, uint8_t* __restrict__ lock,
                                               int* __restrict__ which) {
  const GenLat& G = L[blockIdx.x];
  __shared__ double sq[2][kT];
  __shared__ int si[2][kT];
  double best[2] = {INFINITY, INFINITY};
  int bi[2] = {-1, -1};
  for (int i = threadIdx.x; i < G.N; i += kT) {
    const int gi = G.n0 + i;
    if (layer[gi] != 0) continue;
    const double px = x[gi], pz = x[2 * size_t(N) + gi];
    const double qa = (px + 0.5 * G.width) * (px + 0.5 * G.width) + (pz + 0.5 * G.depth) * (pz + 0.5 * G.depth);
    const double qb = (px - 0.5 * G.width) * (px - 0.5 * G.width) + (pz + 0.5 * G.depth) * (pz + 0.5 * G.depth);
    // strided visit in increasing i: strictly smaller keeps the first
    if (qa < best[0]) best[0] = qa, bi[0] = i;
    if (qb < best[1]) best[1] = qb, bi[1] = i;
  }
  for (int c = 0; c < 2; c++) sq[c][threadIdx.x] = best[c], si[c][threadIdx.x] = bi[c];
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w)
      for (int c = 0; c < 2; c++) {
        const double a = sq[c][threadIdx.x], b = sq[c][threadIdx.x + w];
        const int ia = si[c][threadIdx.x], ib = si[c][threadIdx.x + w];
        // (value, index) lexicographic minimum; an empty slot has index -1
        const bool take = ib >= 0 && (ia < 0 || b < a || (b == a && ib < ia));
        if (take) sq[c][threadIdx.x] = b, si[c][threadIdx.x] = ib;
      }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int ia = si[0][0], ib = si[1][0];
    which[2 * blockIdx.x] = ia;
    which[2 * blockIdx.x + 1] = ib;
    if (ia >= 0 && ib >= 0) {
      lock[G.n0 + ia] |= uint8_t(1u | 4u);
      lock[G.n0 + ib] |= uint8_t(4u);
    }
  }
}
__global__ __launch_bounds__(kT) void k_zero_bytes(uint8_t* __restrict__ a, long n) {
  const long i = long(blockIdx.x) * kT + threadIdx.x;
  if (i < n) a[i] = 0;
}

int grid(long n) { return int((n + kT - 1) / kT); }

// exclusive scan of deg[0..n] (n + 1 entries) in place into off
void scan(int* deg, int* off, int n, hipStream_t stream) {
  size_t bytes = 0;
  GEN_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, deg, off, n + 1, stream));
  void* tmp = nullptr;
  GEN_HIP(hipMalloc(&tmp, std::max<size_t>(bytes, 16)));
  GEN_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, deg, off, n + 1, stream));
  GEN_HIP(hipStreamSynchronize(stream));
  GEN_HIP(hipFree(tmp));
}

template <class X>
struct DevBuf {
  X* p = nullptr;
  explicit DevBuf(size_t n) { GEN_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(X))); }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

}  // namespace

GenCounts gen_counts(const mmc_config& c) {
  GenCounts k;
  const long nx = c.cells_x, ny = c.cells_y, nz = c.cells_z;
  if (nx < 1 || ny < 1 || nz < 1) throw std::runtime_error("cell counts must be positive");
  const long corners = (nx + 1) * (ny + 1) * (nz + 1);
  if (c.lattice == MMC_LATTICE_BCC) {
    const long cells = nx * ny * nz;
    const long struts = 8 * cells + 2 * ((nz + 1) * nx + (nx + 1) * nz);
    const long segs = std::clamp(c.elements_per_strut, 2, 5);
    k.nodes = int(corners + cells + struts * (segs - 1));
    k.beams = int(struts * segs);
    k.connected = struts * (segs + 1);
    k.draws = 2 * (corners + cells);
  } else if (c.lattice == MMC_LATTICE_CUBIC) {
    k.nodes = int(corners);
    k.beams = int(nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz);
    k.connected = k.beams;
    k.draws = (c.position_sigma > 0 ? 4 : 2) * corners + (c.radius_mode == 1 ? k.beams : 0);
  } else {
    throw std::runtime_error("unknown lattice family");
  }
  return k;
}

std::vector<GenResult> generate_lattices(const std::vector<mmc_config>& cfgs, const std::vector<int>& n0,
                                         const std::vector<int>& b0, int N, int B, const GenOut& o,
                                         hipStream_t stream) {
  const int n = int(cfgs.size());
  std::vector<GenLat> L(static_cast<size_t>(n));
  std::vector<double> tab;
  std::vector<long> draws(static_cast<size_t>(n));
  std::vector<int> first_base(static_cast<size_t>(n)), first_strut(static_cast<size_t>(n)), first_cubic(static_cast<size_t>(n));
  long c_total = 0, d_total = 0, base_total = 0, strut_total = 0, cubic_total = 0;
  for (int l = 0; l < n; l++) {
    const mmc_config& c = cfgs[size_t(l)];
    if (!(c.cell_size > 0) || !(c.max_radius > 0)) throw std::runtime_error("cell size and radius must be positive");
    const GenCounts k = gen_counts(c);
    GenLat& G = L[size_t(l)];
    G.family = c.lattice;
    G.nx = c.cells_x, G.ny = c.cells_y, G.nz = c.cells_z;
    G.segs = std::clamp(c.elements_per_strut, 2, 5);
    G.radius_mode = c.radius_mode;
    G.n0 = n0[size_t(l)], G.N = k.nodes, G.b0 = b0[size_t(l)], G.B = k.beams;
    G.corners = (G.nx + 1) * (G.ny + 1) * (G.nz + 1);
    G.cells = G.family == MMC_LATTICE_BCC ? G.nx * G.ny * G.nz : 0;
    G.base = G.corners + G.cells;
    G.struts = G.family == MMC_LATTICE_BCC ? (k.beams / G.segs) : 0;
    G.c0 = c_total, G.d0 = d_total;
    G.h = c.cell_size;
    G.jr = c.joint_scale * c.max_radius;
    G.width = c.cells_x * c.cell_size;
    G.height = c.cells_y * c.cell_size;
    G.depth = c.cells_z * c.cell_size;
    G.imperf = c.imperfection;
    G.member_imperf = c.member_imperfection;
    G.trigger = c.trigger_scale;
    G.sqrt_trigger = std::sqrt(c.trigger_scale);
    G.rmin = c.min_radius, G.rmax = c.max_radius;
    G.sigma = G.family == MMC_LATTICE_CUBIC ? c.position_sigma : 0.0;
    G.density = c.material.density;
    G.jm = c.material.density * (4.0 / 3.0) * kPi * std::pow(G.jr, 3) * 0.28;
    G.seed = c.seed;
    // host tables: layer envelopes, y grades of the strut radius, strut bows
    G.tab = int(tab.size());
    const int ny = G.ny;
    for (int j = 0; j <= ny; j++) tab.push_back(std::sin(kPi * (double(j) / ny)));
    for (int j = 0; j < ny; j++) tab.push_back(std::sin(kPi * ((j + 0.5) / double(ny))));
    auto grade_y = [&](double ya, double yb) {
      const double my = (ya + yb) * 0.5;
      const double y = std::max(0.0, std::min(1.0, (my - G.jr) / G.height));
      return 0.58 + 0.42 * std::pow(1 - y, 0.72);
    };
    for (int j = 0; j < ny; j++)
      for (int dy = 0; dy < 2; dy++) tab.push_back(grade_y(G.jr + (j + 0.5) * G.h, G.jr + (j + dy) * G.h));
    for (int face : {0, ny}) tab.push_back(grade_y(G.jr + face * G.h, G.jr + face * G.h));
    for (int m = 0; m < 5; m++) tab.push_back(m ? std::sin(kPi * (double(m) / G.segs)) : 0.0);
    for (int m = 0; m < 5; m++) tab.push_back(m ? std::sin(2 * kPi * (double(m) / G.segs)) : 0.0);
    draws[size_t(l)] = k.draws;
    c_total += k.connected;
    d_total += k.draws;
    first_base[size_t(l)] = int(base_total);
    base_total += G.base;
    first_strut[size_t(l)] = int(strut_total);
    strut_total += G.struts;
    first_cubic[size_t(l)] = int(cubic_total);
    cubic_total += G.family == MMC_LATTICE_CUBIC ? G.N : 0;
    if (n0[size_t(l)] + G.N > N || b0[size_t(l)] + G.B > B) throw std::runtime_error("lattice ranges exceed the totals");
  }
  if (base_total > 0x7fffffffL || strut_total > 0x7fffffffL || d_total > (1L << 40))
    throw std::runtime_error("lattices too large to generate");
  DevBuf<GenLat> dL{static_cast<size_t>(n)};
  DevBuf<double> dtab{tab.size()};
  DevBuf<long> dcount{static_cast<size_t>(n)};
  DevBuf<int> dfb{static_cast<size_t>(n)}, dfs{static_cast<size_t>(n)}, dfc{static_cast<size_t>(n)}, dn0{static_cast<size_t>(n)};
  DevBuf<uint32_t> ddraws{static_cast<size_t>(d_total)};
  DevBuf<int2> dpairs{static_cast<size_t>(c_total)};
  DevBuf<int> deg{size_t(N) + 1}, cur{size_t(N) + 1}, flags{1};
  DevBuf<double> dmass{static_cast<size_t>(n)};
  DevBuf<int> dwhich{2 * size_t(n)};
  GEN_HIP(hipMemcpy(dL.p, L.data(), L.size() * sizeof(GenLat), hipMemcpyHostToDevice));
  GEN_HIP(hipMemcpy(dtab.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
  GEN_HIP(hipMemcpy(dcount.p, draws.data(), draws.size() * sizeof(long), hipMemcpyHostToDevice));
  GEN_HIP(hipMemcpy(dfb.p, first_base.data(), size_t(n) * sizeof(int), hipMemcpyHostToDevice));
  GEN_HIP(hipMemcpy(dfs.p, first_strut.data(), size_t(n) * sizeof(int), hipMemcpyHostToDevice));
  GEN_HIP(hipMemcpy(dfc.p, first_cubic.data(), size_t(n) * sizeof(int), hipMemcpyHostToDevice));
  GEN_HIP(hipMemcpy(dn0.p, n0.data(), size_t(n) * sizeof(int), hipMemcpyHostToDevice));
  // lattices without struts (cubic) or cubic beams (BCC) have zero-length
  // ranges that share the next lattice's start; owner() takes the last
  // lattice whose start is at or below an item, which is the one holding it
  k_draws<<<grid(n), kT, 0, stream>>>(dL.p, n, dcount.p, ddraws.p);
  k_base_nodes<<<grid(base_total), kT, 0, stream>>>(dL.p, n, dfb.p, base_total, dtab.p, ddraws.p, o, N);
  if (strut_total) k_struts<<<grid(strut_total), kT, 0, stream>>>(dL.p, n, dfs.p, strut_total, dtab.p, o, N, B, dpairs.p);
  if (cubic_total)
    k_cubic_beams<<<grid(cubic_total), kT, 0, stream>>>(dL.p, n, dfc.p, cubic_total, ddraws.p, o, N, B, dpairs.p);
  GEN_HIP(hipGetLastError());
  // beam incidence, ordered by beam
  k_zero_int<<<grid(long(N) + 1), kT, 0, stream>>>(deg.p, long(N) + 1);
  k_beam_degree<<<grid(B), kT, 0, stream>>>(o.a, o.b, B, deg.p);
  scan(deg.p, o.inc_off, N, stream);
  GEN_HIP(hipMemcpyAsync(cur.p, o.inc_off, size_t(N) * sizeof(int), hipMemcpyDeviceToDevice, stream));
  k_beam_fill<<<grid(B), kT, 0, stream>>>(o.a, o.b, B, cur.p, o.inc);
  // exclusion adjacency
  k_zero_int<<<grid(long(N) + 1), kT, 0, stream>>>(deg.p, long(N) + 1);
  k_pair_degree<<<grid(c_total), kT, 0, stream>>>(dpairs.p, c_total, deg.p);
  scan(deg.p, o.adj_off, N, stream);
  GEN_HIP(hipMemcpyAsync(cur.p, o.adj_off, size_t(N) * sizeof(int), hipMemcpyDeviceToDevice, stream));
  k_pair_fill<<<grid(c_total), kT, 0, stream>>>(dpairs.p, c_total, cur.p, o.adj);
  // sorted lists, masses, total masses, locks
  k_zero_int<<<1, kT, 0, stream>>>(flags.p, 1);
  k_node_lists<<<grid(N), kT, 0, stream>>>(N, o.inc_off, o.inc, o.adj_off, o.adj, o.area, o.L0, dL.p, n, dn0.p,
                                           o.mass, flags.p);
  k_total_mass<<<grid(n), kT, 0, stream>>>(dL.p, n, o.area, o.L0, dmass.p);
  k_zero_bytes<<<grid(N), kT, 0, stream>>>(o.lock, N);
  k_locks<<<n, kT, 0, stream>>>(dL.p, o.x, N, o.node_layer, o.lock, dwhich.p);
  GEN_HIP(hipGetLastError());
  std::vector<double> tm(static_cast<size_t>(n));
  std::vector<int> which(2 * size_t(n));
  int flag = 0;
  GEN_HIP(hipMemcpyAsync(tm.data(), dmass.p, tm.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
  GEN_HIP(hipMemcpyAsync(which.data(), dwhich.p, which.size() * sizeof(int), hipMemcpyDeviceToHost, stream));
  GEN_HIP(hipMemcpyAsync(&flag, flags.p, sizeof(int), hipMemcpyDeviceToHost, stream));
  GEN_HIP(hipStreamSynchronize(stream));
  if (flag) throw std::runtime_error("lattice generation produced a repeated exclusion pair");
  std::vector<GenResult> res(static_cast<size_t>(n));
  for (int l = 0; l < n; l++) {
    res[size_t(l)].total_mass = tm[size_t(l)];
    res[size_t(l)].lock_a = which[2 * size_t(l)];
    res[size_t(l)].lock_b = which[2 * size_t(l) + 1];
    if (res[size_t(l)].lock_a < 0 || res[size_t(l)].lock_b < 0)
      throw std::runtime_error("failed to create anti-drift constraints");
  }
  return res;
}

}  // namespace mmc
