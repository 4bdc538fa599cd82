// mmc.hip — gfx950 kernels and the C-ABI context of the beam-lattice
// equilibrium solver (the hot path of Simulator::objective/minimize/commit,
// projects/metamaterial-compression/industry-grade/src/sim.cpp:136-212).
//
// A context holds one or many independent lattices in one concatenated SoA;
// every launch covers all of them through block tables (one lattice per
// block).  One objective evaluation:
//   k_state    node DOFs -> pose (x, R), gravity, platen contact, centroid sums
//   k_beam     per-beam axial/shear/bending/torsion energy, gradient record and
//              trial plastic state
//   k_contact  per-shape narrowphase over the Verlet partner list; each shape
//              accumulates its own contact force in a fixed order
//   k_gather   per-node deterministic assembly over the CSR incidence list,
//              gradient pack, lock mask and convergence norms
//   k_final    one block per lattice: fixed-order sum of its block partials
// The Verlet list is rebuilt (k_count / scan / k_fill / k_pairs on a uniform
// spatial hash) only when some node has moved half the skin.  The L-BFGS loop
// runs on the device as a phase machine of four predicated kernels per trial
// (k_dstate_g, k_elem_g, k_gather_g, k_decide_g), replayed from a hipGraph;
// the host-driven loop (minimize_host) is its cross-check.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/mmc.h"
#include "capture.hpp"
#include "generate.hpp"
#include "kernels.hpp"

using namespace mmc;

namespace {

__device__ __forceinline__ unsigned cell_hash(int i, int j, int k, int inst, unsigned mask) {
  return ((unsigned)i * 73856093u ^ (unsigned)j * 19349663u ^ (unsigned)k * 83492791u ^ (unsigned)inst * 2654435761u) &
         mask;
}
__device__ __forceinline__ float round_down(double v) {
  float f = (float)v;
  return ((double)f > v) ? nextafterf(f, -INFINITY) : f;
}
__device__ __forceinline__ float round_up(double v) {
  float f = (float)v;
  return ((double)f < v) ? nextafterf(f, INFINITY) : f;
}
__device__ __forceinline__ int cell_of(float v, double inv) { return (int)floor((double)v * inv); }

// Per-item energies for the per-term Armijo test: the item's energy at the
// reference point (the current iterate) and at this evaluation.  Either may be
// null (reference-style total-energy test).
struct TermBufs {
  const double* ref;
  double* out;
};

template <class T>
__device__ __forceinline__ v3<double> node_x(const double* x, int N, int i) {
  return v3<double>{x[i], x[N + i], x[2 * N + i]};
}
template <class T>
__device__ __forceinline__ quat<T> node_R(const T* R, int N, int i) {
  return quat<T>{R[i], R[N + i], R[2 * N + i], R[3 * N + i]};
}

// ---------------------------------------------------------------------------
// k_state: z (+ step p) -> x, R; gravity, platens, centroid.  MC/src/sim.cpp:136,142-143,168-171
// partials [9][P]: E_gravity, E_platen, reaction, max_pen, platen_contacts,
// sum_x, sum_z, max displacement from xref (the positions of the last contact
// list build), sum of per-node energy changes against nt.ref
// ---------------------------------------------------------------------------
// The per-node inputs of state_node (base pose, mass, list reference
// position), loaded ahead of the DOF values.
struct NodeBase {
  v3<double> p;
  quat<double> q;
  double mass;
  v3<double> ref;
};
template <class T>
__device__ __forceinline__ NodeBase node_base(Lat<T> L, const double* __restrict__ xref, int i) {
  const int N = L.N;
  NodeBase nb;
  nb.p = v3<double>{L.bpos[i], L.bpos[N + i], L.bpos[2 * N + i]};
  nb.q = quat<double>{L.bquat[i], L.bquat[N + i], L.bquat[2 * N + i], L.bquat[3 * N + i]};
  nb.mass = L.mass[i];
  nb.ref = xref ? v3<double>{xref[i], xref[N + i], xref[2 * N + i]} : v3<double>{0, 0, 0};
  return nb;
}
// Node i at DOF values q: pose written to x, R; its gravity and platen terms,
// centroid sums and displacement from xref in v[0..8] (the rows above).
template <class T>
__device__ __forceinline__ void state_node(Lat<T> L, const Scalars& S, const double (&q)[6], const NodeBase& nb,
                                           double* __restrict__ x, double* __restrict__ R, bool has_ref,
                                           TermBufs nt, int i, double (&v)[9]) {
  const int N = L.N;
  v3<double> pos{nb.p.x + q[0] * S.ts, nb.p.y + q[1] * S.ts, nb.p.z + q[2] * S.ts};
  x[i] = pos.x;
  x[N + i] = pos.y;
  x[2 * N + i] = pos.z;
  quat<double> r = qnormalize(qmul(qexp(v3<double>{q[3], q[4], q[5]}), nb.q));
  R[i] = r.w;
  R[N + i] = r.x;
  R[2 * N + i] = r.y;
  R[3 * N + i] = r.z;
  const double m = nb.mass;
  v[0] = m * S.gravity * pos.y;
  double pb = S.jr - pos.y;
  if (pb > 0) {
    v[1] += 0.5 * S.kc * pb * pb;
    v[3] = fmax(v[3], pb);
    v[4] += 1;
  }
  double pt = pos.y + S.jr - S.top;
  if (pt > 0) {
    v[1] += 0.5 * S.kc * pt * pt;
    v[2] += S.kc * pt;
    v[3] = fmax(v[3], pt);
    v[4] += 1;
  }
  v[5] = pos.x;
  v[6] = pos.z;
  if (has_ref) {
    v3<double> d{pos.x - nb.ref.x, pos.y - nb.ref.y, pos.z - nb.ref.z};
    double dn = norm(d);
    v[7] = isfinite(dn) ? dn : INFINITY;
  }
  const double en = v[0] + v[1];
  if (nt.out) nt.out[i] = en;
  if (nt.ref) v[8] = en - nt.ref[i];
}

template <class T>
__device__ __forceinline__ void state_body(Lat<T> L, Scalars S, const double* __restrict__ z,
                                           const double* __restrict__ p, double step, double* __restrict__ zc,
                                           double* __restrict__ x, double* __restrict__ R,
                                           const double* __restrict__ xref, TermBufs nt,
                                           double* __restrict__ part, int i, bool on) {
  const int N = L.N, P = gridDim.x;
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (on) {
    const unsigned lk = L.lock[i];
    double q[6];
#pragma unroll
    for (int d = 0; d < 6; d++) {
      double zz = z[d * N + i];
      if (p) zz = zz + step * p[d * N + i];
      if ((lk >> d) & 1u) zz = 0.0;
      q[d] = zz;
      if (zc) zc[d * N + i] = zz;
    }
    state_node<T>(L, S, q, node_base(L, xref, i), x, R, xref != nullptr, nt, i, v);
  }
  block_reduce<9>(v, (1u << 3) | (1u << 7));
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 9; k++) part[k * P + blockIdx.x] = v[k];
}
template <class T>
__global__ __launch_bounds__(kBlock) void k_state(Lat<T> L, const Inst* __restrict__ I, const Blk* __restrict__ tn,
                                                  const double* __restrict__ tops, const double* __restrict__ z,
                                                  const double* __restrict__ p, double step, double* __restrict__ zc,
                                                  double* __restrict__ x, double* __restrict__ R,
                                                  const double* __restrict__ xref, TermBufs nt,
                                                  double* __restrict__ part) {
  const Blk bk = tn[blockIdx.x];
  Scalars S = I[bk.inst].S;
  S.top = tops[bk.inst];
  state_body<T>(L, S, z, p, step, zc, x, R, xref, nt, part, bk.first + threadIdx.x, int(threadIdx.x) < bk.count);
}


// ---------------------------------------------------------------------------
// Contact.  Shapes 0..N-1 are joint spheres, N..N+B-1 strut capsules
// (MC/src/sim.cpp:148-151).  Candidate partners come from a Verlet list: every
// non-excluded pair closer than target + skin at the positions x_ref of the
// last build.  While no node has moved by skin/2 or more since that build,
// every pair closer than its contact target is in the list (a segment point
// moves by at most the larger of its endpoint displacements), so evaluating
// the list gives exactly the contact set of the full broadphase.  The list is
// rebuilt with a uniform-grid spatial hash: each shape's float AABB, padded by
// r + margin + skin/2, is counted and scattered into hashed cell buckets; a pair
// is enumerated only in the cell that holds the low corner of the two boxes'
// overlap (each overlapping pair once per partner); each shape's partner list
// is sorted, so the partner order, and therefore every contact sum, does not
// depend on the order of the atomic bucket fill.
// ---------------------------------------------------------------------------
// Shape s of instance in: its spheres (one per node) come first, then its
// capsules (one per beam).
template <class T>
__device__ __forceinline__ void shape_geom(const Lat<T>& L, const Inst& in, const double* x, int s, v3<double>& a,
                                           v3<double>& b, double& r, int& n0, int& n1) {
  const int N = L.N, ls = s - in.s0;
  if (ls < in.N) {
    n0 = in.n0 + ls;
    a = b = node_x<T>(x, N, n0);
    r = in.S.jr;
    n1 = -1;
  } else {
    const int k = in.b0 + ls - in.N;
    n0 = L.ba[k];
    n1 = L.bb[k];
    a = node_x<T>(x, N, n0);
    b = node_x<T>(x, N, n1);
    r = L.rad[k];
  }
}

constexpr int kMaxCellsPerShape = 512;

template <class T>
__global__ __launch_bounds__(kBlock) void k_count(Lat<T> L, const Inst* __restrict__ I, const Blk* __restrict__ tc,
                                                  const double* __restrict__ x, float* __restrict__ box,
                                                  unsigned* __restrict__ cnt, int* __restrict__ iflags) {
  const Blk bk = tc[blockIdx.x];
  if (int(threadIdx.x) >= bk.count) return;
  const int s = bk.first + threadIdx.x;
  const Inst& in = I[bk.inst];
  const Scalars& S = in.S;
  v3<double> a, b;
  double r;
  int n0, n1;
  shape_geom(L, in, x, s, a, b, r, n0, n1);
  double pad = r + S.margin + 0.5 * S.skin;
  float lo[3] = {round_down(fmin(a.x, b.x) - pad), round_down(fmin(a.y, b.y) - pad), round_down(fmin(a.z, b.z) - pad)};
  float hi[3] = {round_up(fmax(a.x, b.x) + pad), round_up(fmax(a.y, b.y) + pad), round_up(fmax(a.z, b.z) + pad)};
  // (the low corner's w carries the shape's instance, for the pair walk)
  reinterpret_cast<float4*>(box)[2 * s] = make_float4(lo[0], lo[1], lo[2], __int_as_float(bk.inst));
  reinterpret_cast<float4*>(box)[2 * s + 1] = make_float4(hi[0], hi[1], hi[2], 0.f);
  int c0[3], c1[3];
  bool ok = true;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    ok = ok && isfinite(lo[d]) && isfinite(hi[d]);
    c0[d] = cell_of(lo[d], S.inv_cell);
    c1[d] = cell_of(hi[d], S.inv_cell);
  }
  long ncell = ok ? (long)(c1[0] - c0[0] + 1) * (c1[1] - c0[1] + 1) * (c1[2] - c0[2] + 1) : 0;
  if (!ok || ncell > kMaxCellsPerShape) {
    // Left out of the grid.  A non-finite pose already makes the energy
    // non-finite; an absurdly large box is reported (flag 1) and the
    // evaluation is marked non-finite so that the line search rejects it.
    if (ok) atomicOr(iflags + bk.inst, 1);
    reinterpret_cast<float4*>(box)[2 * s] = make_float4(INFINITY, INFINITY, INFINITY, 0.f);
    return;
  }
  for (int j = c0[1]; j <= c1[1]; j++)
    for (int k = c0[2]; k <= c1[2]; k++)
      for (int i = c0[0]; i <= c1[0]; i++) atomicAdd(&cnt[cell_hash(i, j, k, bk.inst, S.table_mask)], 1u);
}

// Boxes are packed as two float4 (lo, hi) per shape: one cache line per lookup
// (lo.w: the shape's instance).
__device__ __forceinline__ bool load_cells(const float* box, int s, double inv, float lo[3], float hi[3], int c0[3],
                                           int c1[3]) {
  const float4 l = reinterpret_cast<const float4*>(box)[2 * s], h = reinterpret_cast<const float4*>(box)[2 * s + 1];
  lo[0] = l.x;
  lo[1] = l.y;
  lo[2] = l.z;
  hi[0] = h.x;
  hi[1] = h.y;
  hi[2] = h.z;
  if (!(lo[0] < INFINITY)) return false;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    c0[d] = cell_of(lo[d], inv);
    c1[d] = cell_of(hi[d], inv);
  }
  return true;
}

__global__ __launch_bounds__(kBlock) void k_fill(const Inst* __restrict__ I, const Blk* __restrict__ tc,
                                                 const float* __restrict__ box, unsigned* __restrict__ cnt,
                                                 const unsigned* __restrict__ start, int* __restrict__ entries,
                                                 int* __restrict__ flags) {
  const Blk bk = tc[blockIdx.x];
  if (int(threadIdx.x) >= bk.count) return;
  const int s = bk.first + threadIdx.x;
  const Scalars& S = I[bk.inst].S;
  float lo[3], hi[3];
  int c0[3], c1[3];
  if (!load_cells(box, s, S.inv_cell, lo, hi, c0, c1)) return;
  for (int j = c0[1]; j <= c1[1]; j++)
    for (int k = c0[2]; k <= c1[2]; k++)
      for (int i = c0[0]; i <= c1[0]; i++) {
        unsigned h = cell_hash(i, j, k, bk.inst, S.table_mask);
        unsigned slot = atomicSub(&cnt[h], 1u) - 1u;  // also returns cnt[] to zero
        unsigned idx = start[h] + slot;
        if (idx < (unsigned)S.entry_cap)
          entries[idx] = s;
        else
          atomicOr(flags, 4);
      }
}

template <class T>
__device__ __forceinline__ bool adjacent(const Lat<T>& L, int u, int v) {
  for (int e = L.adj_off[u]; e < L.adj_off[u + 1]; e++)
    if (L.adj[e] == v) return true;
  return false;
}
// localShapes, MC/src/sim.cpp:24-39
template <class T>
__device__ __forceinline__ bool excluded(const Lat<T>& L, int a0, int a1, int b0, int b1) {
  int ea[2] = {a0, a1}, eb[2] = {b0, b1};
  // shared nodes first (no memory traffic), then the adjacency lists
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      if (ea[i] >= 0 && ea[i] == eb[j]) return true;
  for (int i = 0; i < 2; i++) {
    if (ea[i] < 0) continue;
    for (int j = 0; j < 2; j++) {
      if (eb[j] < 0) continue;
      if (adjacent(L, ea[i], eb[j])) return true;
    }
  }
  return false;
}

// closestShapes, MC/src/sim.cpp:17-22 (A is the lower shape index)
__device__ __forceinline__ closest shape_closest(bool capA, v3<double> a0, v3<double> a1, bool capB, v3<double> b0,
                                                 v3<double> b1) {
  if (capA && capB) return seg_seg(a0, a1, b0, b1);
  if (capA) {
    v3<double> d = a1 - a0;
    double s = sat(dot(b0 - a0, d) / fmax(1e-30, norm2(d)));
    return {s, 0, a0 + d * s, b0};
  }
  if (capB) {
    v3<double> d = b1 - b0;
    double t = sat(dot(a0 - b0, d) / fmax(1e-30, norm2(d)));
    return {0, t, a0, b0 + d * t};
  }
  return {0, 0, a0, b0};
}

// Verlet-list build: per shape, partners (either side) closer than target +
// skin, written to the shape's slab of S.slab slots; pcount gets the true
// count, and a count above the slab raises flag 8 (the host widens the slab
// and rebuilds).  kPairLanes lanes per shape (four shapes per wavefront): lane
// c of a shape's segment looks up cells c, c + kPairLanes, ... of the shape's
// box, kPairLanes cells at a time, and the segment then walks those cells'
// bucket entries kPairLanes at a time, one entry per lane per round; a
// round's partners are appended in lane order.  Every overlapping pair is
// kept only in the cell that holds the low corner of the two boxes' overlap
// (once per partner); the segment's first lane then sorts the list and drops
// the duplicates that cells hashing to one bucket produce, so the list, and
// every contact sum over it, does not depend on the order of the bucket fill.
constexpr int kPairLanes = 16;                            // lanes per shape
constexpr int kPairShapes = kBlock / kPairLanes;          // shapes per workgroup
constexpr int kPairSub = kBlock / kPairShapes;            // workgroups per table block
template <class T>
__global__ __launch_bounds__(kBlock) void k_pairs(Lat<T> L, const Inst* __restrict__ I, const Blk* __restrict__ tc,
                                                  int slab_in,
                                                  const double* __restrict__ x, const float* __restrict__ box,
                                                  const unsigned* __restrict__ start, const int* __restrict__ entries,
                                                  unsigned* __restrict__ pcount, int* __restrict__ plist,
                                                  int* __restrict__ flags) {
  const Blk bk = tc[blockIdx.x / kPairSub];
  const int lane = threadIdx.x % kWave, sl = lane % kPairLanes, seg = lane / kPairLanes;
  const int li = (blockIdx.x % kPairSub) * kPairShapes + int(threadIdx.x) / kPairLanes;
  // a whole wavefront returns together (its four shapes lie past the block)
  if ((blockIdx.x % kPairSub) * kPairShapes + int(threadIdx.x) / kWave * (kWave / kPairLanes) >= bk.count) return;
  const bool real = li < bk.count;
  const int s = bk.first + (real ? li : 0);
  const Inst& in = I[bk.inst];
  const Scalars& S = in.S;
  float lo[3], hi[3];
  int c0[3], c1[3];
  const unsigned slab = unsigned(slab_in);
  const size_t base = size_t(s) * slab;
  const unsigned long long segmask = ((1ull << kPairLanes) - 1ull) << (seg * kPairLanes);
  unsigned n = 0;
  const bool boxed = real && load_cells(box, s, S.inv_cell, lo, hi, c0, c1);
  v3<double> sa{0, 0, 0}, sb{0, 0, 0};
  double sr = 0;
  int s0 = 0, s1 = 0, nx = 1, ny = 1, ncell = 0;
  if (boxed) {
    shape_geom(L, in, x, s, sa, sb, sr, s0, s1);
    nx = c1[0] - c0[0] + 1;
    ny = c1[1] - c0[1] + 1;
    ncell = nx * ny * (c1[2] - c0[2] + 1);
  }
  const bool scap = s - in.s0 >= in.N;
  int cmax = ncell;  // wavefront-uniform bounds for the loops below
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) cmax = max(cmax, __shfl_xor(cmax, o, kWave));
  for (int cb = 0; cb < cmax; cb += kPairLanes) {
    const int k = cb + sl;
    int ci = 0, cj = 0, ck = 0;
    unsigned e0 = 0, e1 = 0;
    if (k < ncell) {
      ci = c0[0] + k % nx;
      cj = c0[1] + (k / nx) % ny;
      ck = c0[2] + k / (nx * ny);
      const unsigned h = cell_hash(ci, cj, ck, bk.inst, S.table_mask);
      e0 = start[h];
      e1 = max(e0, min(start[h + 1], (unsigned)S.entry_cap));
    }
    const unsigned len = e1 - e0;
    // the segment's bucket entries over these cells, shared out evenly: item
    // t (running count over the lanes' cells) is taken by lane t % kPairLanes
    // in round t / kPairLanes, which finds the item's cell by a binary search
    // over the lanes' running counts (the segment's longest bucket no longer
    // sets the number of rounds)
    unsigned excl = len;
#pragma unroll
    for (int d = 1; d < kPairLanes; d <<= 1) {
      const unsigned v = unsigned(__shfl_up(int(excl), d, kPairLanes));
      if (sl >= d) excl += v;
    }
    const unsigned total = unsigned(__shfl(int(excl), kPairLanes - 1, kPairLanes));
    excl -= len;
    unsigned rounds = (total + kPairLanes - 1) / kPairLanes;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) rounds = max(rounds, unsigned(__shfl_xor(int(rounds), o, kWave)));
    const float4* box4 = reinterpret_cast<const float4*>(box);
    for (unsigned r = 0; r < rounds; r++) {
      const unsigned item = r * kPairLanes + sl;
      int own = 0;  // the last lane whose running count is at most item
#pragma unroll
      for (int step = kPairLanes / 2; step > 0; step >>= 1)
        if (unsigned(__shfl(int(excl), own + step, kPairLanes)) <= item) own += step;
      const unsigned oe = unsigned(__shfl(int(e0), own, kPairLanes)) + item -
                          unsigned(__shfl(int(excl), own, kPairLanes));
      const int oci = __shfl(ci, own, kPairLanes), ocj = __shfl(cj, own, kPairLanes),
                ock = __shfl(ck, own, kPairLanes);
      bool found = false;
      int o = -1;
      if (item < total) o = entries[oe];
      const float4 ol = o >= 0 ? box4[2 * o] : make_float4(0.f, 0.f, 0.f, 0.f);
      const float4 oh = o >= 0 ? box4[2 * o + 1] : ol;
      if (o >= 0 && o != s && __float_as_int(ol.w) == bk.inst) {
        const float olo[3] = {ol.x, ol.y, ol.z}, ohi[3] = {oh.x, oh.y, oh.z};
        if (olo[0] <= hi[0] && lo[0] <= ohi[0] && olo[1] <= hi[1] && lo[1] <= ohi[1] && olo[2] <= hi[2] &&
            lo[2] <= ohi[2] && cell_of(fmaxf(lo[0], olo[0]), S.inv_cell) == oci &&
            cell_of(fmaxf(lo[1], olo[1]), S.inv_cell) == ocj && cell_of(fmaxf(lo[2], olo[2]), S.inv_cell) == ock) {
          v3<double> oa, ob;
          double orr;
          int o0, o1;
          shape_geom(L, in, x, o, oa, ob, orr, o0, o1);
          const bool low = s < o, ocap = o - in.s0 >= in.N;
          closest cp = low ? shape_closest(scap, sa, sb, ocap, oa, ob) : shape_closest(ocap, oa, ob, scap, sa, sb);
          const double dist = norm(cp.pb - cp.pa);
          found = dist < sr + orr + S.margin + S.skin && !excluded(L, s0, s1, o0, o1);
        }
      }
      const unsigned long long m = __ballot(found) & segmask;
      if (found) {
        const unsigned pos = n + unsigned(__popcll(m & ((1ull << lane) - 1ull)));
        if (pos < slab) plist[base + pos] = o;
      }
      n += unsigned(__popcll(m));
    }
  }
  // the segment's slab writes are read below by its first lane
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  if (sl != 0 || !real) return;
  pcount[s] = n;
  if (n > slab) {
    atomicOr(flags, 8);
    return;
  }
  // partners ascending: the contact sums do not depend on bucket order
  for (unsigned i = 1; i < n; i++) {
    const int v = plist[base + i];
    unsigned j = i;
    while (j > 0 && plist[base + j - 1] > v) {
      plist[base + j] = plist[base + j - 1];
      j--;
    }
    plist[base + j] = v;
  }
  // a shape covering two cells that hash to one bucket meets a partner twice
  unsigned u = 0;
  for (unsigned i = 0; i < n; i++)
    if (u == 0 || plist[base + u - 1] != plist[base + i]) plist[base + u++] = plist[base + i];
  pcount[s] = u;
}

// Node-ordered gradient records of the device loop.  Each beam end owns one
// slot; the slots of the 64 nodes of a gather block are interleaved (slot u of
// lane l at base + 64 u + l, u in CSR order of the node's incident beams), so
// that the gather reads one coalesced column per incident beam, with no index
// chain.  rec holds per slot the end's share of the chord force (-f on end a,
// f on end b) and the end's rotational gradient ([6][S]); cap the capsule
// contact force on the end ([3][S]).  Slots past a node's degree stay zero.
struct EllOut {
  double* rec;
  double* cap;
  long S;
};
template <class T>
__device__ __forceinline__ void ell_cap(Lat<T> L, EllOut eo, int k, v3<double> f0, v3<double> f1) {
  const long sa = L.eslot[2 * k], sb = L.eslot[2 * k + 1];
  eo.cap[sa] = f0.x;
  eo.cap[eo.S + sa] = f0.y;
  eo.cap[2 * eo.S + sa] = f0.z;
  eo.cap[sb] = f1.x;
  eo.cap[eo.S + sb] = f1.y;
  eo.cap[2 * eo.S + sb] = f1.z;
}

// k_contact: MC/src/sim.cpp:159-165 over the Verlet list.  Each shape
// accumulates its own force; each pair's energy is counted by its lower shape.
// partials [4][ncols]: E_contact, max_pen, contacts, sum of per-shape energy
// changes; one column per wavefront (64 consecutive shapes), column col.
template <class T>
__device__ __forceinline__ void contact_walk(Lat<T> L, const Inst& in, const double* __restrict__ x,
                                             const unsigned* __restrict__ pcount, const int* __restrict__ plist,
                                             int slab, double* __restrict__ cf_sph, double* __restrict__ cf_cap,
                                             TermBufs ct, double* __restrict__ part, int s, bool on, size_t col,
                                             size_t ncols, EllOut eo = EllOut{nullptr, nullptr, 0}) {
  const int N = L.N, B = L.B;
  const Scalars& S = in.S;
  double v[4] = {0, 0, 0, 0};
  if (on) {
    v3<double> f0{0, 0, 0}, f1{0, 0, 0};
    const size_t e0 = size_t(s) * slab, e1 = e0 + min(pcount[s], unsigned(slab));
    const bool scap = s - in.s0 >= in.N;
    if (e1 > e0) {
      v3<double> sa, sb;
      double sr;
      int s0, s1;
      shape_geom(L, in, x, s, sa, sb, sr, s0, s1);
      for (size_t e = e0; e < e1; e++) {
        const int o = plist[e];
        v3<double> oa, ob;
        double orr;
        int o0, o1;
        shape_geom(L, in, x, o, oa, ob, orr, o0, o1);
        const bool low = s < o, ocap = o - in.s0 >= in.N;
        closest cp = low ? shape_closest(scap, sa, sb, ocap, oa, ob) : shape_closest(ocap, oa, ob, scap, sa, sb);
        v3<double> delta = cp.pb - cp.pa;
        double dist = norm(delta);
        double target = (low ? sr + orr : orr + sr) + S.margin;
        if (dist >= target) continue;
        // instance-local shape indices (the fallback normal of coincident
        // shapes depends on them)
        int ia = (low ? s : o) - in.s0, ib = (low ? o : s) - in.s0;
        v3<double> n = dist > 1e-12 ? over(delta, dist) : unit(v3<double>{0.37 + 0.001 * ia, 0.73, -0.41 + 0.001 * ib});
        double pen = target - dist, mag = S.kc * pen;
        v3<double> fv = low ? n * mag : n * (-mag);
        double u = low ? cp.s : cp.t;
        if (scap) {
          f0 = f0 + fv * (1 - u);
          f1 = f1 + fv * u;
        } else {
          f0 = f0 + fv;
        }
        if (low) {
          v[0] += 0.5 * S.kc * pen * pen;
          v[1] = fmax(v[1], pen);
          v[2] += 1;
        }
      }
    }
    if (scap) {
      const int k = in.b0 + (s - in.s0 - in.N);
      if (cf_cap) {
        cf_cap[k] = f0.x;
        cf_cap[B + k] = f0.y;
        cf_cap[2 * B + k] = f0.z;
        cf_cap[3 * B + k] = f1.x;
        cf_cap[4 * B + k] = f1.y;
        cf_cap[5 * B + k] = f1.z;
      }
      if (eo.cap) ell_cap(L, eo, k, f0, f1);
    } else {
      const int i = in.n0 + (s - in.s0);
      cf_sph[i] = f0.x;
      cf_sph[N + i] = f0.y;
      cf_sph[2 * N + i] = f0.z;
    }
    if (ct.out) ct.out[s] = v[0];
    if (ct.ref) v[3] = v[0] - ct.ref[s];
  }
  v[0] = wave_sum(v[0]);
  v[1] = wave_max(v[1]);
  v[2] = wave_sum(v[2]);
  v[3] = wave_sum(v[3]);
  if ((threadIdx.x & (kWave - 1)) == 0)
#pragma unroll
    for (int k = 0; k < 4; k++) part[size_t(k) * ncols + col] = v[k];
}
// contact_walk over blocks of NT shapes, P blocks: one partial column per
// wavefront (64 consecutive shapes of the instance), so that the fold over
// the columns does not depend on the workgroup size
template <class T, int NT = kBlock>
__device__ __forceinline__ void contact_body(Lat<T> L, const Inst& in, const double* __restrict__ x,
                                             const unsigned* __restrict__ pcount, const int* __restrict__ plist,
                                             int slab, double* __restrict__ cf_sph, double* __restrict__ cf_cap,
                                             TermBufs ct, double* __restrict__ part, int s, bool on, int blk, int P) {
  constexpr int W = NT / kWave;
  contact_walk<T>(L, in, x, pcount, plist, slab, cf_sph, cf_cap, ct, part, s, on,
                  size_t(blk) * W + threadIdx.x / kWave, size_t(P) * W);
}
// Contact shapes of a chunk of 64 whose Verlet lists are long: one block,
// kContactLanes lanes per shape.  Lane c of a shape's group evaluates partner slots c,
// c + kContactLanes, ... with contact_body's arithmetic; after each round the
// group adds the round's pairs up in ascending slot order from the lanes'
// factors (fv, u, pen, shuffled), so that every shape's force and energy sums
// are contact_walk's sequential walk, operation for operation.  The per-shape
// totals are folded over the chunk's 64 shapes by one wavefront as in
// contact_walk (the chunk's partial column blk).  A wavefront's time is the
// longest partner list of its 16 shapes over kContactLanes instead of the
// longest of 64.  Chunks with at most kContactSerial partners per shape stay
// with contact_walk (one lane per shape), where lane groups would cost four
// times the instructions for little to share.
constexpr int kContactLanes = kBlock / kWave;
constexpr int kContactSerial = 8;  // default; MMC_CONTACT_SERIAL overrides (tests)
template <class T>
__device__ __forceinline__ void contact_group(Lat<T> L, const Inst& in, const double* __restrict__ x,
                                              const unsigned* __restrict__ pcount, const int* __restrict__ plist,
                                              int slab, double* __restrict__ cf_sph, double* __restrict__ cf_cap,
                                              TermBufs ct, double* __restrict__ part, int first, int count, int blk,
                                              int P, EllOut eo = EllOut{nullptr, nullptr, 0}) {
  static_assert(kBlock == kWave * kContactLanes, "a block holds 64 shapes");
  const int N = L.N, B = L.B;
  const Scalars& S = in.S;
  const int g = threadIdx.x / kContactLanes, c = threadIdx.x % kContactLanes;
  const int gbase = (threadIdx.x & (kWave - 1)) & ~(kContactLanes - 1);
  const bool on = g < count;
  const int s = first + (on ? g : 0);
  const size_t e0 = size_t(s) * slab;
  const int n = on ? int(min(pcount[s], unsigned(slab))) : 0;
  const bool scap = s - in.s0 >= in.N;
  double v[4] = {0, 0, 0, 0};
  v3<double> f0{0, 0, 0}, f1{0, 0, 0};
  v3<double> sa{0, 0, 0}, sb{0, 0, 0};
  double sr = 0;
  int s0 = 0, s1 = 0;
  if (n > c) shape_geom(L, in, x, s, sa, sb, sr, s0, s1);
  // rounds of kContactLanes slots; the trip count is the wavefront's longest
  // list, so that the shuffles below see every lane
  // the partner of the lane's next slot and its pose are loaded while the
  // current one is evaluated (one dependent load chain per round overlaps the
  // previous round's arithmetic)
  int onext = c < n ? plist[e0 + c] : -1;
  v3<double> anext{0, 0, 0}, bnext{0, 0, 0};
  double rnext = 0;
  {
    int o0, o1;
    if (onext >= 0) shape_geom(L, in, x, onext, anext, bnext, rnext, o0, o1);
  }
  for (int k = 0; __ballot(k < n) != 0; k += kContactLanes) {
    const int j = k + c;
    const int o = onext;
    const v3<double> oa = anext, ob = bnext;
    const double orr = rnext;
    onext = j + kContactLanes < n ? plist[e0 + j + kContactLanes] : -1;
    if (onext >= 0) {
      int o0, o1;
      shape_geom(L, in, x, onext, anext, bnext, rnext, o0, o1);
    }
    int flag = 0;  // 1: in contact, 2: the pair's lower shape
    v3<double> fv{0, 0, 0};
    double u = 0, pen = 0;
    if (j < n) {
      const bool low = s < o, ocap = o - in.s0 >= in.N;
      closest cp = low ? shape_closest(scap, sa, sb, ocap, oa, ob) : shape_closest(ocap, oa, ob, scap, sa, sb);
      v3<double> delta = cp.pb - cp.pa;
      double dist = norm(delta);
      double target = (low ? sr + orr : orr + sr) + S.margin;
      if (dist < target) {
        int ia = (low ? s : o) - in.s0, ib = (low ? o : s) - in.s0;
        v3<double> nv =
            dist > 1e-12 ? over(delta, dist) : unit(v3<double>{0.37 + 0.001 * ia, 0.73, -0.41 + 0.001 * ib});
        pen = target - dist;
        double mag = S.kc * pen;
        fv = low ? nv * mag : nv * (-mag);
        u = low ? cp.s : cp.t;
        flag = low ? 3 : 1;
      }
    }
    if (__ballot(flag != 0) == 0) continue;
#pragma unroll
    for (int q = 0; q < kContactLanes; q++) {
      const int src = gbase + q;
      const int fq = __shfl(flag, src);
      const v3<double> fvq{__shfl(fv.x, src), __shfl(fv.y, src), __shfl(fv.z, src)};
      const double uq = __shfl(u, src), penq = __shfl(pen, src);
      if (fq & 1) {
        if (scap) {
          f0 = f0 + fvq * (1 - uq);
          f1 = f1 + fvq * uq;
        } else {
          f0 = f0 + fvq;
        }
        if (fq & 2) {
          v[0] += 0.5 * S.kc * penq * penq;
          v[1] = fmax(v[1], penq);
          v[2] += 1;
        }
      }
    }
  }
  __shared__ double cv[4][kWave];
  if (c == 0) {
    if (on) {
      if (scap) {
        const int kb = in.b0 + (s - in.s0 - in.N);
        if (cf_cap) {
          cf_cap[kb] = f0.x;
          cf_cap[B + kb] = f0.y;
          cf_cap[2 * B + kb] = f0.z;
          cf_cap[3 * B + kb] = f1.x;
          cf_cap[4 * B + kb] = f1.y;
          cf_cap[5 * B + kb] = f1.z;
        }
        if (eo.cap) ell_cap(L, eo, kb, f0, f1);
      } else {
        const int i = in.n0 + (s - in.s0);
        cf_sph[i] = f0.x;
        cf_sph[N + i] = f0.y;
        cf_sph[2 * N + i] = f0.z;
      }
      if (ct.out) ct.out[s] = v[0];
      if (ct.ref) v[3] = v[0] - ct.ref[s];
    }
#pragma unroll
    for (int r = 0; r < 4; r++) cv[r][g] = v[r];
  }
  __syncthreads();
  if (threadIdx.x < kWave) {
    double w[4];
#pragma unroll
    for (int r = 0; r < 4; r++) w[r] = cv[r][threadIdx.x];
    w[0] = wave_sum(w[0]);
    w[1] = wave_max(w[1]);
    w[2] = wave_sum(w[2]);
    w[3] = wave_sum(w[3]);
    if (threadIdx.x == 0)
#pragma unroll
      for (int r = 0; r < 4; r++) part[size_t(r) * P + blk] = w[r];
  }
  __syncthreads();  // cv is free for the block's next chunk
}
// Chunks of the device loop's contact table (64 shapes each) whose longest
// Verlet list exceeds `serial` partners: flag[j] = 1 and j appended to list
// (count reset by the caller).  One wavefront per chunk.  Run after every list
// build and table change; the flags only route work (both routes give the
// same sums), so a stale flag costs time, not results.
__global__ __launch_bounds__(kBlock) void k_classify(const Blk* __restrict__ tc, int nchunks,
                                                     const unsigned* __restrict__ pcount, int slab, int serial,
                                                     int* __restrict__ flag, int* __restrict__ list,
                                                     unsigned* __restrict__ count) {
  const int j = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave, l = threadIdx.x & (kWave - 1);
  if (j >= nchunks) return;
  const Blk bk = tc[j];
  const double n = l < bk.count ? double(min(pcount[bk.first + l], unsigned(slab))) : 0.0;
  const bool heavy = wave_max(n) > double(serial);
  if (l == 0) {
    flag[j] = heavy ? 1 : 0;
    if (heavy) list[atomicAdd(count, 1u)] = j;
  }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_contact(Lat<T> L, const Inst* __restrict__ I, const Blk* __restrict__ tc,
                                                    const double* __restrict__ x, const unsigned* __restrict__ pcount,
                                                    const int* __restrict__ plist, int slab,
                                                    double* __restrict__ cf_sph, double* __restrict__ cf_cap,
                                                    TermBufs ct, double* __restrict__ part) {
  const Blk bk = tc[blockIdx.x];
  contact_body<T>(L, I[bk.inst], x, pcount, plist, slab, cf_sph, cf_cap, ct, part, bk.first + threadIdx.x,
                  int(threadIdx.x) < bk.count, blockIdx.x, gridDim.x);
}


// ---------------------------------------------------------------------------
// k_beam: per-beam energy, gradient record and trial state, MC/src/sim.cpp:139-141.
// rec [9][B]: f (force on b, -f on a), rotational gradient of a, of b.
// partials [4][P]: energy, elastic, plastic, sum of per-beam energy changes.
// ---------------------------------------------------------------------------
template <class T, int NT = kBlock>
__device__ __forceinline__ void beam_body(Lat<T> L, Material M, const double* __restrict__ x,
                                                 const double* __restrict__ R, double* __restrict__ rec,
                                                 double* __restrict__ trial, TermBufs bt,
                                                 double* __restrict__ part, int k, bool on, int blk, int P,
                                                 EllOut eo = EllOut{nullptr, nullptr, 0}) {
  const int N = L.N, B = L.B;
  double v[4] = {0, 0, 0, 0};
  if (on) {
    const int ia = L.ba[k], ib = L.bb[k];
    v3<double> d = node_x<T>(x, N, ib) - node_x<T>(x, N, ia);
    double Ln = norm(d);
    double out[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double tr[MMC_TRIAL_STRIDE];
#pragma unroll
    for (int c = 0; c < MMC_TRIAL_STRIDE; c++) tr[c] = 0;
    if (Ln < 1e-10) {
      v[0] = 1e8 * ((1e-10 - Ln) * (1e-10 - Ln));
    } else {
      const double L0 = L.L0[k], r = L.rad[k], A = L.area[k], I = L.inert[k], J = 2 * I;
      double pl[7];
#pragma unroll
      for (int c = 0; c < 7; c++) pl[c] = L.plastic[c * B + k];
      v3<double> t = over(d, Ln);
      double eps = (Ln - L0) / L0;
      ret1 ax = return1(eps, pl[0], pl[1], M.E, M.sy, M.H);
      double uax = A * L0 * (ax.energy + ax.hard), dax = A * L0 * M.sy * ax.dg;
      v3<double> f = t * (A * ax.force);
      // shear: material directors against the chord, in FP64 for either
      // element type (the mismatch 2t - da - db cancels unit vectors to a few
      // percent of their length)
      quat<double> Ra = node_R<double>(R, N, ia), Rb = node_R<double>(R, N, ib);
      v3<double> e0{L.e0[k], L.e0[B + k], L.e0[2 * B + k]};
      v3<double> da = qrot(Ra, e0), db = qrot(Rb, e0);
      double cs = 0.5 * M.shear * M.G * A * L0;
      v3<double> ma = da - t, mb = db - t;
      double us = 0.5 * cs * (norm2(ma) + norm2(mb));
      v3<double> gt = ((t * 2.0 - da) - db) * cs;
      v3<double> trans = over(gt - t * dot(gt, t), Ln);
      f = f + trans;
      v3<double> ga = -(cross(da, t) * cs), gb = -(cross(db, t) * cs);
      // bending + torsion from the relative rotation (formed in FP64); its
      // logarithm, the beam axis in the frame of a and the moment transport in T
      quat<double> Q = qnormalize(qmul(qconj(Ra), Rb));
      quat<T> QT{T(Q.w), T(Q.x), T(Q.y), T(Q.z)}, RaT{T(Ra.w), T(Ra.x), T(Ra.y), T(Ra.z)};
      v3<T> tT = cvt<T>(t);
      v3<T> phiT = qlog(QT), axT = unit(qrotinv(RaT, tT));
      v3<double> phi = cvt<double>(phiT), axis = cvt<double>(axT);
      v3<double> rotp{pl[2], pl[3], pl[4]};
      double tw = dot(phi, axis), ptw = dot(rotp, axis);
      v3<double> bend = phi - axis * tw, pb = rotp - axis * ptw;
      double kb = M.E * I / L0, kt = M.G * J / L0, my = M.sy * I / r;
      double ty = (M.sy / sqrt(3.0)) * J / r;
      ret3 br = return3(bend, pb, pl[5], kb, my, M.hb_frac * kb);
      ret1 to = return1(tw, ptw, pl[6], kt, ty, M.ht_frac * kt);
      v3<double> nrotp = br.p + axis * to.p;
      v3<double> mom = br.force + axis * to.force;
      double ur = br.energy + br.hard + to.energy + to.hard, dr = my * br.dg + ty * to.dg;
      v3<double> world = cvt<double>(qrot(RaT, jlinvt(phiT, cvt<T>(mom))));
      ga = ga - world;
      gb = gb + world;
      v[0] = (uax + dax) + us + (ur + dr);
      v[1] = uax + us + ur;
      v[2] = dax + dr;
      out[0] = f.x;
      out[1] = f.y;
      out[2] = f.z;
      out[3] = ga.x;
      out[4] = ga.y;
      out[5] = ga.z;
      out[6] = gb.x;
      out[7] = gb.y;
      out[8] = gb.z;
      if (trial) {
        tr[0] = ax.p;
        tr[1] = ax.a;
        tr[2] = nrotp.x;
        tr[3] = nrotp.y;
        tr[4] = nrotp.z;
        tr[5] = br.a;
        tr[6] = to.a;
        tr[7] = eps;
        tr[8] = ax.force;
        tr[9] = mom.x;
        tr[10] = mom.y;
        tr[11] = mom.z;
        tr[12] = uax + us + ur;
        tr[13] = dax + dr;
        tr[14] = ax.yielded;
        tr[15] = br.yielded;
        tr[16] = to.yielded;
      }
    }
    if (rec)
#pragma unroll
      for (int c = 0; c < 9; c++) rec[c * B + k] = out[c];
    if (eo.rec) {
      const long sa = L.eslot[2 * k], sb = L.eslot[2 * k + 1];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        eo.rec[c * eo.S + sa] = -out[c];
        eo.rec[c * eo.S + sb] = out[c];
        eo.rec[(3 + c) * eo.S + sa] = out[3 + c];
        eo.rec[(3 + c) * eo.S + sb] = out[6 + c];
      }
    }
    if (trial)
#pragma unroll
      for (int c = 0; c < MMC_TRIAL_STRIDE; c++) trial[c * B + k] = tr[c];
    if (bt.out) bt.out[k] = v[0];
    if (bt.ref) v[3] = v[0] - bt.ref[k];
  }
  // one partial column per wavefront (64 consecutive beams of the instance):
  // the fold over the columns does not depend on the workgroup size
  constexpr int W = NT / kWave;
#pragma unroll
  for (int c = 0; c < 4; c++) v[c] = wave_sum(v[c]);
  if ((threadIdx.x & (kWave - 1)) == 0)
#pragma unroll
    for (int c = 0; c < 4; c++) part[size_t(c) * P * W + size_t(blk) * W + threadIdx.x / kWave] = v[c];
}
template <class T>
__global__ __launch_bounds__(kBlock) void k_beam(Lat<T> L, const Inst* __restrict__ I, const Blk* __restrict__ tb,
                                                 const double* __restrict__ x, const double* __restrict__ R,
                                                 double* __restrict__ rec, double* __restrict__ trial, TermBufs bt,
                                                 double* __restrict__ part) {
  const Blk bk = tb[blockIdx.x];
  beam_body<T>(L, I[bk.inst].M, x, R, rec, trial, bt, part, bk.first + threadIdx.x, int(threadIdx.x) < bk.count,
               blockIdx.x, gridDim.x);
}


constexpr int kMaxHist = 16;  // longest L-BFGS history of the device loop

// ---------------------------------------------------------------------------
// k_gather: per-node assembly + pack, MC/src/sim.cpp:139-143,164,168-172, in
// element-parallel form: blocks of 6 waves over 64 nodes (the 64-node table),
// wave d assembles DOF d of its block's nodes.  partials [3][P] over the
// 64-node blocks: sum of squares and max |g| over unlocked DOFs, g.p (the
// search direction's directional derivative at this point, when p is given).
// ---------------------------------------------------------------------------
constexpr int kIncUnroll = 8;  // incident beams whose loads are issued together
constexpr int kGRows = 3 + 4 * kMaxHist + 5;  // gradient rows, then pair products

// Gradient component d of node i (before the lock), with the per-component
// arithmetic of the reference's assembly in a fixed order: incident beams in
// CSR order (chord force with the side's sign, then the capsule contact force
// of that side; rotational gradient of the side), the sphere contact force,
// gravity, centering, the two platens.  Split in two: gather_load issues the
// loads of the first kIncUnroll incident beams (indices, then values, every
// load before the first use; an index past the node's list re-reads its last
// entry and its value is not used), gather_sum adds them up.
struct IncVals {
  int e0, e1;
  int code[kIncUnroll];
  double val[kIncUnroll], cap[kIncUnroll];
};
template <class T>
__device__ __forceinline__ void gather_load(Lat<T> L, const double* __restrict__ rec,
                                            const double* __restrict__ cf_cap, int d, int i, IncVals& v) {
  const int B = L.B;
  v.e0 = L.inc_off[i];
  v.e1 = L.inc_off[i + 1];
  const int elast = max(v.e1 - 1, 0);
#pragma unroll
  for (int u = 0; u < kIncUnroll; u++) v.code[u] = L.inc[min(v.e0 + u, elast)];
#pragma unroll
  for (int u = 0; u < kIncUnroll; u++) {
    const int k = v.code[u] >> 1, side = v.code[u] & 1;
    if (d < 3) {
      v.val[u] = rec[size_t(d) * B + k];
      v.cap[u] = cf_cap[size_t(3 * side + d) * B + k];
    } else {
      v.val[u] = rec[size_t(3 + 3 * side + (d - 3)) * B + k];
      v.cap[u] = 0.0;
    }
  }
}
template <class T>
__device__ __forceinline__ double gather_sum(Lat<T> L, const Inst& in, double top, const double* __restrict__ x,
                                             const double* __restrict__ rec, const double* __restrict__ cf_sph,
                                             const double* __restrict__ cf_cap, double cen0, double cen1, int d,
                                             int i, const IncVals& v) {
  const int N = L.N, B = L.B;
  const Scalars& S = in.S;
  double gs = 0.0;
#pragma unroll
  for (int u = 0; u < kIncUnroll; u++) {
    if (v.e0 + u < v.e1) {
      if (d < 3) {
        gs = (v.code[u] & 1) ? gs + v.val[u] : gs - v.val[u];
        gs = gs + v.cap[u];
      } else {
        gs = gs + v.val[u];
      }
    }
  }
  for (int e = v.e0 + kIncUnroll; e < v.e1; e++) {
    const int c = L.inc[e], k = c >> 1, side = c & 1;
    if (d < 3) {
      const double f = rec[size_t(d) * B + k];
      gs = side ? gs + f : gs - f;
      gs = gs + cf_cap[size_t(3 * side + d) * B + k];
    } else {
      gs = gs + rec[size_t(3 + 3 * side + (d - 3)) * B + k];
    }
  }
  if (d >= 3) return gs;
  gs = gs + cf_sph[size_t(d) * N + i];
  if (d == 1) {
    gs += L.mass[i] * S.gravity;
    const double y = x[N + i];
    const double pb = S.jr - y;
    if (pb > 0) gs -= S.kc * pb;
    const double pt = y + S.jr - top;
    if (pt > 0) gs += S.kc * pt;
  } else {
    gs += S.centering * (d == 0 ? cen0 : cen1) / in.N;
  }
  return S.ts * gs;
}

// Gradient component d of node i from the node-ordered records (EllOut): the
// same terms in the same order as gather_load + gather_sum (incident beams in
// CSR order: the end's chord-force share, then its capsule contact force; the
// rotational gradient of the end), so the sums are bit for bit those of the
// CSR walk.  The md slots of the node's block (md: the block's largest degree)
// are loaded before the first use, without waiting for the node's degree.
constexpr int kEllUnroll = 8;
template <class T>
__device__ __forceinline__ double gather_ell(Lat<T> L, const Inst& in, double top, const double* __restrict__ x,
                                             const double* __restrict__ erec, const double* __restrict__ ecap,
                                             long S, const double* __restrict__ cf_sph, double cen0, double cen1,
                                             int d, int i, int deg, long sb, int md) {
  const int N = L.N;
  const Scalars& Sc = in.S;
  double r[kEllUnroll], c[kEllUnroll];
  const double* rr = erec + d * S + sb;
  const double* cc = ecap + (d < 3 ? d : 0) * S + sb;
#pragma unroll
  for (int u = 0; u < kEllUnroll; u++) {
    r[u] = 0.0;
    c[u] = 0.0;
    if (u < md) {
      r[u] = rr[long(u) * kWave];
      if (d < 3) c[u] = cc[long(u) * kWave];
    }
  }
  double gs = 0.0;
#pragma unroll
  for (int u = 0; u < kEllUnroll; u++) {
    if (u < deg) {
      gs = gs + r[u];
      if (d < 3) gs = gs + c[u];
    }
  }
  for (int u = kEllUnroll; u < deg; u++) {
    gs = gs + rr[long(u) * kWave];
    if (d < 3) gs = gs + cc[long(u) * kWave];
  }
  if (d >= 3) return gs;
  gs = gs + cf_sph[size_t(d) * N + i];
  if (d == 1) {
    gs += L.mass[i] * Sc.gravity;
    const double y = x[N + i];
    const double pb = Sc.jr - y;
    if (pb > 0) gs -= Sc.kc * pb;
    const double pt = y + Sc.jr - top;
    if (pt > 0) gs += Sc.kc * pt;
  } else {
    gs += Sc.centering * (d == 0 ? cen0 : cen1) / in.N;
  }
  return Sc.ts * gs;
}

// The six gradient components of node i at once (gather_sum's arithmetic per
// component, in the same order): the incidence list is walked once and every
// record load of the six components is issued before the first use.
template <class T>
__device__ __forceinline__ void gather_node(Lat<T> L, const Inst& in, double top, const double* __restrict__ x,
                                            const double* __restrict__ rec, const double* __restrict__ cf_sph,
                                            const double* __restrict__ cf_cap, double cen0, double cen1, int i,
                                            double (&g)[6]) {
  const int N = L.N, B = L.B;
  const Scalars& S = in.S;
  const int e0 = L.inc_off[i], e1 = L.inc_off[i + 1];
  const int elast = max(e1 - 1, 0);
  int code[kIncUnroll];
#pragma unroll
  for (int u = 0; u < kIncUnroll; u++) code[u] = L.inc[min(e0 + u, elast)];
  double val[6][kIncUnroll], cap[3][kIncUnroll];
#pragma unroll
  for (int u = 0; u < kIncUnroll; u++) {
    const int k = code[u] >> 1, side = code[u] & 1;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      val[d][u] = rec[size_t(d) * B + k];
      cap[d][u] = cf_cap[size_t(3 * side + d) * B + k];
      val[3 + d][u] = rec[size_t(3 + 3 * side + d) * B + k];
    }
  }
  const double sph[3] = {cf_sph[i], cf_sph[N + i], cf_sph[2 * N + i]};
  const double mass = L.mass[i], y = x[N + i];
#pragma unroll
  for (int d = 0; d < 6; d++) {
    double gs = 0.0;
#pragma unroll
    for (int u = 0; u < kIncUnroll; u++) {
      if (e0 + u < e1) {
        if (d < 3) {
          gs = (code[u] & 1) ? gs + val[d][u] : gs - val[d][u];
          gs = gs + cap[d][u];
        } else {
          gs = gs + val[d][u];
        }
      }
    }
    for (int e = e0 + kIncUnroll; e < e1; e++) {
      const int c = L.inc[e], k = c >> 1, side = c & 1;
      if (d < 3) {
        const double f = rec[size_t(d) * B + k];
        gs = side ? gs + f : gs - f;
        gs = gs + cf_cap[size_t(3 * side + d) * B + k];
      } else {
        gs = gs + rec[size_t(3 + 3 * side + (d - 3)) * B + k];
      }
    }
    if (d >= 3) {
      g[d] = gs;
      continue;
    }
    gs = gs + sph[d];
    if (d == 1) {
      gs += mass * S.gravity;
      const double pb = S.jr - y;
      if (pb > 0) gs -= S.kc * pb;
      const double pt = y + S.jr - top;
      if (pt > 0) gs += S.kc * pt;
    } else {
      gs += S.centering * (d == 0 ? cen0 : cen1) / in.N;
    }
    g[d] = S.ts * gs;
  }
}

// Node centroid (x, z) of the instance from the state kernel's partial rows
// 5 and 6 (every block reduces them in the same order).
template <int NT>
__device__ __forceinline__ void centroid(const Inst& in, const double* __restrict__ spart, int Ps, double& c0,
                                         double& c1) {
  double t[2];
  reduce_columns<2, NT>(spart + 5 * Ps, Ps, in.pn0, in.pn1, t, 0u);
  c0 = t[0] * (1.0 / in.N);
  c1 = t[1] * (1.0 / in.N);
}

// Block totals of NR per-thread rows (row 1 a maximum, the others sums) over
// the NW waves of an element-parallel block; thread r < NR returns row r's
// total.  Within a wave the first 64 rows are reduced by a butterfly that
// halves the rows at every step (each lane keeps the half whose row bit
// matches its lane bit and sends the other half across), so that lane l ends
// with row l after 63 exchanges instead of 6 per row; rows past 64 are summed
// one by one.  Fixed order throughout.
// One step of that butterfly: rows [0, 2H) of a, combined with a partner lane
// whose bit H differs and whose higher bits agree: lane ^ H across rows (H =
// 32, 16: permlane swaps, which also do the selection), lane ^ 15 and lane ^ 7
// (row mirrors, H = 8, 4) and lane ^ 2, lane ^ 1 (quad permutes) within a
// row.  The partners of the six steps span all 64 lanes, so lane l still ends
// with the total of row l.
template <int H, int NA>
__device__ __forceinline__ void halve_rows(double (&a)[NA], int lane) {
  if constexpr (H >= 16) {
#pragma unroll
    for (int r = 0; r < H; r++) {
      double x = a[r], y = a[r + H];
      lane_swap_f64<H>(x, y);
      a[r] = x + y;
    }
  } else {
    constexpr int C = H == 8 ? kDppXor15 : (H == 4 ? kDppXor7 : (H == 2 ? kDppXor2 : kDppXor1));
    const bool up = lane & H;
#pragma unroll
    for (int r = 0; r < H; r++) {
      const double send = up ? a[r] : a[r + H];
      const double keep = up ? a[r + H] : a[r];
      a[r] = keep + dpp_f64<C>(send);
    }
  }
}
template <int NW, int NA>
__device__ __forceinline__ double block_rows(double (&a)[NA], int NR, double (*sh)[NA]) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const double mx = wave_max(a[1]);
  if constexpr (NA < kWave) {
#pragma unroll
    for (int r = 0; r < NA; r++) {
      if (r < NR) {
        const double s = r == 1 ? mx : wave_sum(a[r]);
        if (lane == 0) sh[w][r] = s;
      }
    }
  } else {
    a[1] = 0.0;
    halve_rows<32>(a, lane);
    halve_rows<16>(a, lane);
    halve_rows<8>(a, lane);
    halve_rows<4>(a, lane);
    halve_rows<2>(a, lane);
    halve_rows<1>(a, lane);
    sh[w][lane] = lane == 1 ? mx : a[0];
#pragma unroll
    for (int r = kWave; r < NA; r++) {
      if (r < NR) {
        const double s = wave_sum(a[r]);
        if (lane == 0) sh[w][r] = s;
      }
    }
  }
  __syncthreads();
  double tot = 0.0;
  if (int(threadIdx.x) < NR) {
    const int r = threadIdx.x;
    tot = sh[0][r];
#pragma unroll
    for (int k = 1; k < NW; k++) tot = r == 1 ? fmax(tot, sh[k][r]) : tot + sh[k][r];
  }
  return tot;
}

template <class T>
__global__ __launch_bounds__(kElemThreads) void k_gather(Lat<T> L, const Inst* __restrict__ I,
                                                         const Blk* __restrict__ td, const double* __restrict__ tops,
                                                         const double* __restrict__ x, const double* __restrict__ rec,
                                                         const double* __restrict__ cf_sph,
                                                         const double* __restrict__ cf_cap,
                                                         const double* __restrict__ spart, int Ps,
                                                         double* __restrict__ g, const double* __restrict__ p,
                                                         double* __restrict__ part) {
  const Blk bk = td[blockIdx.x];
  const Inst& in = I[bk.inst];
  const int d = threadIdx.x / kWave, l = threadIdx.x % kWave, i = bk.first + l;
  const bool on = l < bk.count;
  const int ic = on ? i : bk.first;
  const size_t e = size_t(d) * L.N + ic;
  const bool locked = (L.lock[ic] >> d) & 1u;
  const double pv = p ? p[e] : 0.0;
  IncVals iv;
  gather_load<T>(L, rec, cf_cap, d, ic, iv);
  double c0, c1;
  centroid<kElemThreads>(in, spart, Ps, c0, c1);
  const double gv = gather_sum<T>(L, in, tops[bk.inst], x, rec, cf_sph, cf_cap, c0, c1, d, ic, iv);
  const double gd = locked ? 0.0 : gv;
  if (on) g[e] = gd;
  double a[4] = {on && !locked ? gd * gd : 0.0, on && !locked ? fabs(gd) : 0.0, on ? gd * pv : 0.0, 0.0};
  __shared__ double sh[kElemThreads / kWave][4];
  const double t = block_rows<kElemThreads / kWave>(a, 3, sh);
  if (threadIdx.x < 3) part[size_t(threadIdx.x) * gridDim.x + blockIdx.x] = t;
}


// Record layout (doubles): 0 energy, 1 elastic, 2 plastic, 3 contact, 4 reaction,
// 5 max_pen, 6 rms, 7 max, 8 contacts, 9 finite, 10 flags, 11 largest node
// displacement since the contact list was built, 12 energy change against the
// reference record summed per term (when per-term energies were kept),
// 13 centering energy, 14 g.p
constexpr int kRec = 16;

// The instance constants a record needs (the decision kernels read them
// ahead of the decision).
struct RecConst {
  int N;
  double centering, n_free;
};
__device__ __forceinline__ RecConst rec_const(const Inst& in) { return {in.N, in.S.centering, in.n_free}; }
// The record of instance in from its 20 reduced partial rows: 9 of the state
// kernel, 4 of the contact shapes, 4 of the beams, 3 of the gradient (fl: the
// instance's flag word).
__device__ __forceinline__ void record_from_rows(const RecConst& in, const double* __restrict__ rows, int fl,
                                                 const double* __restrict__ refrec, double* __restrict__ out) {
  const int N = in.N;
  const double centering = in.centering, n_free = in.n_free;
  const double* s = rows;
  const double* c = rows + 9;
  const double* b = rows + 13;
  const double* g = rows + 17;
  double cx = s[5] * (1.0 / N), cz = s[6] * (1.0 / N);
  double ecen = 0.5 * centering * (cx * cx + cz * cz);
  double energy = b[0] + s[0] + ecen + c[0] + s[1];
  bool fin = isfinite(energy) && !(fl & 1);
  out[0] = fin ? energy : INFINITY;
  out[1] = b[1];
  out[2] = b[2];
  out[3] = c[0] + s[1];
  out[4] = s[2];
  out[5] = fmax(s[3], c[1]);
  out[6] = fin ? (n_free > 0 ? sqrt(g[0] / n_free) : 0.0) : 0.0;
  out[7] = fin ? g[1] : 0.0;
  out[8] = s[4] + c[2];
  out[9] = fin ? 1.0 : 0.0;
  out[10] = (double)fl;
  out[11] = s[7];
  out[12] = refrec ? (s[8] + c[3] + b[3]) + (ecen - refrec[13]) : 0.0;
  out[13] = ecen;
  out[14] = g[2];
  out[15] = 0.0;
}

// Row pass of the record reductions: 20 record rows (9 of the state kernel,
// 4 of the contact shapes, 4 of the beams, 3 of the gradient) and, with
// history on, the 4 nold + 5 pair-product rows of the gather kernel.  Every
// row is folded in one fixed order, whichever block and wave folds it: lane l
// folds columns l, l + 64, ... in ascending order, then the wave folds its 64
// lane values (wave_sum / wave_max, a fixed butterfly).  The loads of the first RD * 64 columns of
// every row of a wave (RD = kDecColRounds unless a pass names its own) are
// issued together (one memory latency for instances of up to that many
// columns per row).
constexpr int kDecWaves = kBlock / kWave;
constexpr int kDecColRounds = 5;

// Row j of an instance's partials: 0..8 state kernel, 9..12 contact shapes,
// 13..16 beams, 17.. gradient rows (3, then the pair products); mx: folded by
// maximum
struct DecRow {
  const double* p;
  int n;
  bool mx;
};
struct DecParts {
  const double *spart, *cpart, *bpart, *gpart;
  int Ps, Pc, Pb, Pg;
};
__device__ __forceinline__ bool row_max(int j) { return j == 3 || j == 7 || j == 10 || j == 18; }
// (the beam and contact rows hold one column per wavefront of the kBlock-item
// blocks: Pc and Pb count blocks)
constexpr int kBW = kBlock / kWave;
__device__ __forceinline__ DecRow decide_row(int j, const Inst& in, const DecParts& P) {
  if (j < 9) return {P.spart + size_t(j) * P.Ps + in.pn0, in.pn1 - in.pn0, row_max(j)};
  if (j < 13)
    return {P.cpart + size_t(j - 9) * P.Pc * kBW + in.pc0 * kBW, (in.pc1 - in.pc0) * kBW, row_max(j)};
  if (j < 17) return {P.bpart + size_t(j - 13) * P.Pb * kBW + in.pb0 * kBW, (in.pb1 - in.pb0) * kBW, false};
  return {P.gpart + size_t(j - 17) * P.Pg + in.pd0, in.pd1 - in.pd0, row_max(j)};
}
// Folds rows jbase + w + q js (w the wave, q < NQ) that lie in [0, R): lane 0
// of the folding wave calls put(j, total) (returns after a block barrier);
// row(j) describes row j.  The first
// loads carry no branches (an index past a row's end reads the row's last
// column, or dummy for a row outside [0, R) or an empty one, and the value is
// replaced by the fold's identity), so that all of them are in flight
// together.
// RD: columns per lane loaded in the first round (RD x 64 columns per row).
template <int NQ, int RD = kDecColRounds, class Row, class Put>
__device__ __forceinline__ void rows_pass(Row row, const double* dummy, int R, int jbase, int js, Put put) {
  constexpr int kCols = RD * kWave;
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  double x[NQ][RD];
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int j = jbase + wv + q * js;
    const DecRow rw = row(j < R ? j : 0);
    const bool any = j < R && rw.n > 0;
    const double* base = any ? rw.p : dummy;
    const int last = any ? rw.n - 1 : 0;
#pragma unroll
    for (int u = 0; u < RD; u++) x[q][u] = base[min(lane + u * kWave, last)];
  }
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int j = jbase + wv + q * js;
    const DecRow rw = row(j < R ? j : 0);
    const double id = rw.mx ? -INFINITY : 0.0;
    double a = id;
#pragma unroll
    for (int u = 0; u < RD; u++) {
      const double e = (j < R && lane + u * kWave < rw.n) ? x[q][u] : id;
      a = rw.mx ? fmax(a, e) : a + e;
    }
    if (j < R) {  // (wave-uniform)
#pragma unroll 4
      for (int col = lane + kCols; col < rw.n; col += kWave) a = rw.mx ? fmax(a, rw.p[col]) : a + rw.p[col];
      const double t = rw.mx ? wave_max(a) : wave_sum(a);
      if (lane == 0) put(j, t);
    }
  }
  __syncthreads();
}

// Reduces the partial columns of instance in (node blocks of spart and gpart,
// shape blocks of cpart, beam blocks of bpart) into its record, in the order
// of the device loop's decision kernel.
__device__ __forceinline__ void final_body(const Inst& in, const double* __restrict__ spart, int Ps,
                                           const double* __restrict__ cpart, int Pc,
                                           const double* __restrict__ bpart, int Pb,
                                           const double* __restrict__ gpart, int Pg,
                                           int* __restrict__ iflag, const double* __restrict__ refrec,
                                           double* __restrict__ out) {
  constexpr int NQ = 20 / kDecWaves;
  __shared__ double rows[20];
  const DecParts P{spart, cpart, bpart, gpart, Ps, Pc, Pb, Pg};
  rows_pass<NQ>([&](int j) { return decide_row(j, in, P); }, spart + in.pn0, 20, 0, kDecWaves,
                [&](int j, double t) { rows[j] = t; });
  if (threadIdx.x == 0) {
    const int fl = iflag[0];
    iflag[0] = 0;
    record_from_rows(rec_const(in), rows, fl, refrec, out);
  }
}
// one block per instance; records at out + inst * kRec
__global__ __launch_bounds__(kBlock) void k_final(const Inst* __restrict__ I, const double* __restrict__ spart, int Ps,
                                                  const double* __restrict__ cpart, int Pc,
                                                  const double* __restrict__ bpart, int Pb,
                                                  const double* __restrict__ gpart, int Pg,
                                                  int* __restrict__ iflags, const double* __restrict__ refrec,
                                                  double* __restrict__ out) {
  const int b = blockIdx.x;
  final_body(I[b], spart, Ps, cpart, Pc, bpart, Pb, gpart, Pg, iflags + b, refrec ? refrec + b * kRec : nullptr,
             out + b * kRec);
}


// ---------------------------------------------------------------------------
// beamPotential, MC/src/sim.cpp:468-509 (objectivity check)
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(kBlock) void k_beam_potential(Lat<T> L, const Inst* __restrict__ I,
                                                           const Blk* __restrict__ tb, const double* __restrict__ x,
                                                           const double* __restrict__ R, double* __restrict__ part) {
  const int N = L.N, B = L.B;
  const Blk bk = tb[blockIdx.x];
  const int k = bk.first + threadIdx.x;
  const Material M = I[bk.inst].M;
  double v[1] = {0};
  if (int(threadIdx.x) < bk.count) {
    const int ia = L.ba[k], ib = L.bb[k];
    v3<double> d = node_x<T>(x, N, ib) - node_x<T>(x, N, ia);
    double Ln = fmax(1e-12, norm(d));
    v3<double> t = over(d, Ln);
    const double L0 = L.L0[k], r = L.rad[k], A = L.area[k], I = L.inert[k], J = 2 * I;
    double pl[7];
    for (int c = 0; c < 7; c++) pl[c] = L.plastic[c * B + k];
    double eps = (Ln - L0) / L0;
    ret1 ax = return1(eps, pl[0], pl[1], M.E, M.sy, M.H);
    double total = A * L0 * (ax.energy + ax.hard);
    quat<double> Ra = node_R<double>(R, N, ia), Rb = node_R<double>(R, N, ib);
    v3<double> e0{L.e0[k], L.e0[B + k], L.e0[2 * B + k]};
    v3<double> da = qrot(Ra, e0), db = qrot(Rb, e0);
    double cs = 0.5 * M.shear * M.G * A * L0;
    total += 0.5 * cs * (norm2(da - t) + norm2(db - t));
    quat<double> Q = qnormalize(qmul(qconj(Ra), Rb));
    v3<double> phi = qlog(Q), axis = unit(qrotinv(Ra, t));
    v3<double> rotp{pl[2], pl[3], pl[4]};
    double tw = dot(phi, axis), ptw = dot(rotp, axis);
    v3<double> bend = phi - axis * tw, pb = rotp - axis * ptw;
    double kb = M.E * I / L0, kt = M.G * J / L0, my = M.sy * I / r, ty = (M.sy / sqrt(3.0)) * J / r;
    ret3 br = return3(bend, pb, pl[5], kb, my, M.hb_frac * kb);
    ret1 to = return1(tw, ptw, pl[6], kt, ty, M.ht_frac * kt);
    v[0] = total + (br.energy + br.hard + to.energy + to.hard);
  }
  block_reduce<1>(v, 0u);
  if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

__global__ __launch_bounds__(kBlock) void k_sum1(const double* __restrict__ part, int P, double* __restrict__ out) {
  double s[1];
  reduce_partials<1>(part, P, s, 0u);
  if (threadIdx.x == 0) out[0] = s[0];
}

// ---------------------------------------------------------------------------
// L-BFGS vector kernels (thread per node, 6 DOF planes), MC/src/sim.cpp:43-46,
// 184-191, 205-206.  The direction p = -H g of the two-loop recursion is applied
// in the compact form of Byrd, Nocedal & Schnabel (Math. Prog. 63, 1994, eq.
// 2.6): H g = gamma g + sum_j a_j s_j + c_j y_j, with the small coefficient
// system solved from the pair inner products.  Two vector passes per iteration
// (pair + dots, direction) replace the recursion's 2m dependent passes.
// ---------------------------------------------------------------------------
struct DirCoef {
  double cg;
  int k;
  int slot[kMaxHist];
  double cs[kMaxHist], cy[kMaxHist];
};
struct PairArgs {
  int t, nold;
  int slot[kMaxHist];
};

__device__ __forceinline__ void write_dot_partial(double v, double* part) {
  double t[1] = {v};
  block_reduce<1>(t, 0u);
  if (threadIdx.x == 0) part[blockIdx.x] = t[0];
}

// Direction epilogue: zero locked DOFs, clamp per node (clampDirection,
// MC/src/sim.cpp:46).
__device__ __forceinline__ void finish_node(double (&q)[6], unsigned lk, double mt, double mr) {
#pragma unroll
  for (int d = 0; d < 6; d++)
    if ((lk >> d) & 1u) q[d] = 0.0;
  double lt = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
  if (lt > mt) {
    double s = mt / lt;
    q[0] *= s;
    q[1] *= s;
    q[2] *= s;
  }
  double lr = sqrt(q[3] * q[3] + q[4] * q[4] + q[5] * q[5]);
  if (lr > mr) {
    double s = mr / lr;
    q[3] *= s;
    q[4] *= s;
    q[5] *= s;
  }
#pragma unroll
  for (int d = 0; d < 6; d++)
    if ((lk >> d) & 1u) q[d] = 0.0;
}

// p = -(cg g + sum_j cs_j s_j + cy_j y_j), locked and clamped; partial g.p
__device__ __forceinline__ void direction_body(int N, size_t n6, const double* __restrict__ g,
                                                      const double* __restrict__ Sh, const double* __restrict__ Yh,
                                                      const DirCoef& C, double* __restrict__ p,
                                                      const uint8_t* __restrict__ lock, double mt, double mr,
                                                      double* __restrict__ part) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double dv = 0;
  if (i < N) {
    double gv[6], q[6];
#pragma unroll
    for (int d = 0; d < 6; d++) {
      gv[d] = g[d * N + i];
      q[d] = C.cg * gv[d];
    }
    for (int j = 0; j < C.k; j++) {
      const double* s = Sh + n6 * C.slot[j];
      const double* y = Yh + n6 * C.slot[j];
#pragma unroll
      for (int d = 0; d < 6; d++) q[d] += C.cs[j] * s[d * N + i] + C.cy[j] * y[d * N + i];
    }
#pragma unroll
    for (int d = 0; d < 6; d++) q[d] = -q[d];
    finish_node(q, lock[i], mt, mr);
#pragma unroll
    for (int d = 0; d < 6; d++) {
      p[d * N + i] = q[d];
      dv += gv[d] * q[d];
    }
  }
  write_dot_partial(dv, part);
}
// Element-parallel form for the device loop: blocks of 6 waves, wave d holds
// DOF plane d of 64 nodes; the per-node clamp reads the node's other planes
// from LDS.  Same arithmetic per element as direction_body.  Writes p and
// returns this element's direction (0 past the block's nodes) and g.p term.
constexpr int kDirNodes = kElemNodes;
template <int MH>
__device__ __forceinline__ double direction_elem(size_t e, bool on, bool locked, double gv, double cg, int k,
                                                 const DirCoef& C, const double* __restrict__ Sh,
                                                 const double* __restrict__ Yh, size_t n6, double mt, double mr,
                                                 double* __restrict__ p, double& dv) {
  __shared__ double qs[6][kDirNodes];
  const int d = threadIdx.x / kDirNodes, l = threadIdx.x % kDirNodes;
  double q = 0.0;
  if (on) {
    // every history load issued before the first use
    double hs[MH], hy[MH];
#pragma unroll
    for (int j = 0; j < MH; j++) {
      if (j < k) {
        hs[j] = Sh[n6 * C.slot[j] + e];
        hy[j] = Yh[n6 * C.slot[j] + e];
      }
    }
    q = cg * gv;
#pragma unroll
    for (int j = 0; j < MH; j++)
      if (j < k) q += C.cs[j] * hs[j] + C.cy[j] * hy[j];
    q = -q;
    if (locked) q = 0.0;
  }
  qs[d][l] = q;
  __syncthreads();
  dv = 0.0;
  if (on) {
    const int b = d < 3 ? 0 : 3;
    const double a0 = qs[b][l], a1 = qs[b + 1][l], a2 = qs[b + 2][l];
    const double len = sqrt(a0 * a0 + a1 * a1 + a2 * a2), lim = d < 3 ? mt : mr;
    if (len > lim) q *= lim / len;
    if (locked) q = 0.0;
    p[e] = q;
    dv = gv * q;
  }
  return q;
}

__global__ __launch_bounds__(kBlock) void k_direction(int N, size_t n6, const double* __restrict__ g,
                                                      const double* __restrict__ Sh, const double* __restrict__ Yh,
                                                      DirCoef C, double* __restrict__ p,
                                                      const uint8_t* __restrict__ lock, double mt, double mr,
                                                      double* __restrict__ part) {
  direction_body(N, n6, g, Sh, Yh, C, p, lock, mt, mr, part);
}


// s = zc - z, y = gc - g into history slot t.  Partials [4*nold + 5][P]: for
// each retained pair j: s_j.y, y_j.y, s_j.gc, y_j.gc; then s.y, s.s, y.y,
// s.gc, y.gc.
__device__ __forceinline__ void pair_dots_body(int N, size_t n6, const double* __restrict__ z,
                                                      const double* __restrict__ zc, const double* __restrict__ g,
                                                      const double* __restrict__ gc, double* __restrict__ Sh,
                                                      double* __restrict__ Yh, const PairArgs& A, double* __restrict__ part,
                                                      int i, bool on) {
  const int P = gridDim.x;
  double sv[6] = {0, 0, 0, 0, 0, 0}, yv[6] = {0, 0, 0, 0, 0, 0}, gv[6] = {0, 0, 0, 0, 0, 0};
  if (on) {
    double* st = Sh + n6 * A.t;
    double* yt = Yh + n6 * A.t;
#pragma unroll
    for (int d = 0; d < 6; d++) {
      gv[d] = gc[d * N + i];
      sv[d] = zc[d * N + i] - z[d * N + i];
      yv[d] = gv[d] - g[d * N + i];
      st[d * N + i] = sv[d];
      yt[d * N + i] = yv[d];
    }
  }
  // all 4*nold + 5 partial dots in registers (unrolled over the maximum
  // history), then reduced 16 at a time
  constexpr int KM = 4 * kMaxHist + 5;
  double acc[KM];
#pragma unroll
  for (int r = 0; r < KM; r++) acc[r] = 0.0;
  if (on) {
#pragma unroll
    for (int jj = 0; jj < kMaxHist; jj++) {
      if (jj < A.nold) {
        const double* sj = Sh + n6 * A.slot[jj];
        const double* yj = Yh + n6 * A.slot[jj];
#pragma unroll
        for (int d = 0; d < 6; d++) {
          const double a = sj[d * N + i], b = yj[d * N + i];
          acc[4 * jj + 0] += a * yv[d];
          acc[4 * jj + 1] += b * yv[d];
          acc[4 * jj + 2] += a * gv[d];
          acc[4 * jj + 3] += b * gv[d];
        }
      }
    }
  }
  // the new pair's products go right after the retained pairs' rows
  double w[5] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int d = 0; d < 6; d++) {
    w[0] += sv[d] * yv[d];
    w[1] += sv[d] * sv[d];
    w[2] += yv[d] * yv[d];
    w[3] += sv[d] * gv[d];
    w[4] += yv[d] * gv[d];
  }
#pragma unroll
  for (int jj = 0; jj <= kMaxHist; jj++)
    if (jj == A.nold)
#pragma unroll
      for (int c = 0; c < 5; c++) acc[4 * jj + c] = w[c];
  const int K = 4 * A.nold + 5;
#pragma unroll
  for (int c0 = 0; c0 < KM; c0 += 16) {
    if (c0 >= K) break;
    double v[16];
#pragma unroll
    for (int r = 0; r < 16; r++) v[r] = (c0 + r < KM) ? acc[c0 + r] : 0.0;
    block_reduce<16>(v, 0u);
    if (threadIdx.x == 0)
      for (int r = 0; r < 16 && c0 + r < K; r++) part[(c0 + r) * P + blockIdx.x] = v[r];
  }
}
__global__ __launch_bounds__(kBlock) void k_pair_dots(int N, size_t n6, const double* __restrict__ z,
                                                      const double* __restrict__ zc, const double* __restrict__ g,
                                                      const double* __restrict__ gc, double* __restrict__ Sh,
                                                      double* __restrict__ Yh, PairArgs A, double* __restrict__ part) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  pair_dots_body(N, n6, z, zc, g, gc, Sh, Yh, A, part, i, i < N);
}


// Sums K rows of block partials [K][P] into out[K] (one block, fixed order).
__global__ __launch_bounds__(kBlock) void k_sum_rows(const double* __restrict__ part, int P, int K,
                                                     double* __restrict__ out) {
  for (int r = 0; r < K; r++) {
    double v[1];
    reduce_partials<1>(part + size_t(r) * P, P, v, 0u);
    if (threadIdx.x == 0) out[r] = v[0];
  }
}

// ---------------------------------------------------------------------------
// Device-resident minimiser, MC/src/sim.cpp:174-208.  A control block holds
// the loop state; every kernel of one L-BFGS iteration is predicated on it, so
// the iteration can be captured once as a hipGraph and replayed without host
// round trips.  One replay (one line-search trial) runs four kernels:
//   k_dstate_g  p = -H g (compact form) and g.p          phase DIRECTION, or
//               p = -g                                   phase STEEPEST,
//               then the trial state z + alpha p         (alpha = 1), or the
//               stored p at the current alpha            phase SEARCH
//   k_elem_g    beams and contact shapes of the trial
//   k_gather_g  gradient, the trial's curvature pair and its products
//   k_decide_g  a fresh direction that is not a descent direction (g.p >= 0)
//               discards the trial (steepest descent next); otherwise
//               Armijo: accept (pair update, convergence test, next
//               coefficients, phase DIRECTION), halve (phase SEARCH: the next
//               replay is a backtrack), steepest retry, or failure
// A fresh direction's trial is evaluated before its descent check is known
// (the check needs g.p over the whole lattice); the rare non-descent
// direction costs one discarded trial and saves a launch on every other one.
// Only the deciding block writes the control block; the wide kernels read it,
// and kernel boundaries order the two.
// ---------------------------------------------------------------------------
enum { PH_DIRECTION = 0, PH_STEEPEST = 2, PH_SEARCH = 4 };
enum { ST_RUNNING = 0, ST_CONVERGED = 1, ST_FAILED = 2, ST_ITERCAP = 3, ST_HOLD = 4 };

constexpr int kMaxSlots = kMaxHist + 1;  // history slots of the device loop: m + 1 (one spare)
struct Ctl {
  int status, phase, par, it, iterations, backtracks, evals, retry;
  int bt, count, spare, m, max_it, max_ls, rel_escape, ls_mode;
  int wolfe, pad;
  double step, dg, top, armijo, min_step, tol_rms, tol_max, rel_tol, pen_tol, ts, skin_half, wolfe_eps;
  double cur[kRec], cand[kRec];
  DirCoef coef;
  int order[kMaxHist];              // history slots, oldest first (count of them)
  double SY[kMaxSlots][kMaxSlots];  // by slot: s_i . y_j (i not newer than j)
  double YY[kMaxSlots][kMaxSlots];  // by slot: y_i . y_j
  double u[kMaxSlots], v[kMaxSlots];  // s_j . g and y_j . g of the current gradient
};

// Device-loop history bookkeeping.  The loop keeps m + 1 slots: `order` lists
// the count retained pairs oldest first, `spare` is a free slot.  Every trial
// writes its (s, y) into the spare slot and computes its products with the
// pairs that survive an acceptance (all of them, or all but the oldest when
// the history is full), so that the products are ready when the trial is
// accepted (the surviving pairs are order[off..count), off = 1 when the
// history is full); a rejected trial's slot is simply overwritten by the next
// one.

// MC/src/sim.cpp:205-206 on the device loop's slots: accept the new pair if
// s.y > 1e-12 sqrt(s.s y.y), else clear the history.  d holds the products in
// the layout of pair_dots_body.  One wavefront (all 64 lanes call it, c in
// LDS): lane jj writes the products of retained pair jj (the retained pairs
// are order[off..count), off = 1 when the history is full and the oldest pair
// is evicted); the new pair goes into the spare slot, and the next spare is
// the evicted slot or the lowest free one.
__device__ inline void history_update_wave(Ctl& c, const double* d) {
  const int l = threadIdx.x & (kWave - 1);
  const int off = c.count == c.m ? 1 : 0, nold = c.count - off, t = c.spare;
  const double* nw = d + 4 * nold;
  const double sy = nw[0], ss = nw[1], yy = nw[2];
  if (!(sy > 1e-12 * sqrt(fmax(1e-30, ss * yy)))) {
    if (l == 0) c.count = 0;
    return;
  }
  // every read of the old order precedes the first write (one wavefront:
  // its LDS accesses complete in program order, and the fence keeps the
  // compiler from moving a write above another lane's read)
  const int evicted = off ? c.order[0] : -1;
  const int j = l < nold ? c.order[l + off] : 0;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (l < nold) {
    c.SY[j][t] = d[4 * l + 0];
    c.YY[j][t] = c.YY[t][j] = d[4 * l + 1];
    c.u[j] = d[4 * l + 2];
    c.v[j] = d[4 * l + 3];
    c.order[l] = j;
  }
  unsigned used = (l < nold ? 1u << j : 0u) | (l == 0 ? 1u << t : 0u);
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) used |= unsigned(__shfl_xor(int(used), o, kWave));
  if (l == 0) {
    c.SY[t][t] = sy;
    c.YY[t][t] = yy;
    c.u[t] = nw[3];
    c.v[t] = nw[4];
    c.order[nold] = t;
    c.count = nold + 1;
    c.spare = evicted >= 0 ? evicted : __builtin_ctz(~used);
  }
}

// Coefficients of H g = gamma g + sum_j a_j s_j + c_j y_j for the retained
// history (oldest first from slot head): Byrd, Nocedal & Schnabel 1994, eq. 2.6:
// w = R^{-1} u, a = R^{-T}((D + gamma Y'Y) w - gamma v), c = -gamma w, with R
// the upper triangle of S'Y and D its diagonal.  gamma is the reference's H0
// (MC/src/sim.cpp:186); an empty history gives p = -0.1 g.
__host__ __device__ inline void compact_coefficients(int m, int count, int head, const double (*SY)[kMaxHist],
                                                     const double (*YY)[kMaxHist], const double* u, const double* v,
                                                     DirCoef& C) {
  const int k = count;
  C.k = k;
  if (k == 0) {
    C.cg = 0.10;
    return;
  }
  int sl[kMaxHist];
  for (int i = 0; i < k; i++) sl[i] = (head + i) % m;
  const double sy = SY[sl[k - 1]][sl[k - 1]], yy = YY[sl[k - 1]][sl[k - 1]];
  const double gamma = (sy > 1e-20 && yy > 1e-20) ? fmin(1e4, fmax(1e-7, sy / yy)) : 0.10;
  double w[kMaxHist], t[kMaxHist], a[kMaxHist];
  for (int i = k - 1; i >= 0; i--) {
    double acc = u[sl[i]];
    for (int j = i + 1; j < k; j++) acc -= SY[sl[i]][sl[j]] * w[j];
    w[i] = acc / SY[sl[i]][sl[i]];
  }
  for (int i = 0; i < k; i++) {
    double acc = SY[sl[i]][sl[i]] * w[i] - gamma * v[sl[i]];
    for (int j = 0; j < k; j++) acc += gamma * YY[sl[i]][sl[j]] * w[j];
    t[i] = acc;
  }
  for (int i = 0; i < k; i++) {
    double acc = t[i];
    for (int j = 0; j < i; j++) acc -= SY[sl[j]][sl[i]] * a[j];
    a[i] = acc / SY[sl[i]][sl[i]];
  }
  C.cg = gamma;
  for (int i = 0; i < k; i++) {
    C.slot[i] = sl[i];
    C.cs[i] = a[i];
    C.cy[i] = -gamma * w[i];
  }
}

// Curvature-pair bookkeeping after the pair inner products d (layout of
// k_pair_dots), MC/src/sim.cpp:205-206.  Shared by both minimiser drivers.
__host__ __device__ inline void history_update(int m, int& count, int& head, double (*SY)[kMaxHist],
                                               double (*YY)[kMaxHist], double* u, double* v, const PairArgs& A,
                                               const double* d) {
  const double* nw = d + 4 * A.nold;
  const double sy = nw[0], ss = nw[1], yy = nw[2];
  if (sy > 1e-12 * sqrt(fmax(1e-30, ss * yy))) {
    for (int jj = 0; jj < A.nold; jj++) {
      const int j = A.slot[jj];
      SY[j][A.t] = d[4 * jj + 0];
      YY[j][A.t] = YY[A.t][j] = d[4 * jj + 1];
      u[j] = d[4 * jj + 2];
      v[j] = d[4 * jj + 3];
    }
    SY[A.t][A.t] = sy;
    YY[A.t][A.t] = yy;
    u[A.t] = nw[3];
    v[A.t] = nw[4];
    if (count == m)
      head = (head + 1) % m;  // the new pair took the oldest pair's slot
    else
      count++;
  } else {
    count = 0;
    head = 0;
  }
}

// Slot of the next pair and the retained pairs whose products it needs.
__host__ __device__ inline PairArgs pair_slots(int m, int count, int head) {
  PairArgs A{};
  A.t = count == m ? head % m : (head + count) % m;
  A.nold = 0;
  for (int j = 0; j < count; j++) {
    const int sj = (head + j) % m;
    if (sj != A.t) A.slot[A.nold++] = sj;
  }
  return A;
}

// compact_coefficients for one wavefront on the device loop's slots (all 64
// lanes call it; lane l holds history row l, oldest first).  The two triangular solves run by columns:
// each step broadcasts one solved component and every lane updates its own
// residual, so the dependent chain is k steps instead of k(k+1)/2.  Each lane
// reads its rows of S'Y and Y'Y (and its column of S'Y) into registers first,
// so a step is a division, a lane read and a multiply-add.  The division by
// the lane's own diagonal element d goes through its reciprocal, computed
// before the chain: q0 = r (1/d), then one correction q0 + (r - q0 d)(1/d),
// which is the correctly rounded quotient r / d (Markstein, IBM J. Res. Dev.
// 34, 1990: a reciprocal within half an ulp and a quotient within one ulp;
// no overflow or underflow here), so the result is that of the division,
// bit for bit, at three dependent operations per step.
template <int MH = kMaxHist>
__device__ void compact_coefficients_wave(const Ctl& c, DirCoef& C) {
  const int k = c.count, l = threadIdx.x & (kWave - 1);
  if (k == 0) {
    if (l == 0) {
      C.k = 0;
      C.cg = 0.10;
    }
    return;
  }
  const bool on = l < k;
  const int sl = on ? c.order[l] : 0;
  const int last = c.order[k - 1];
  const double sy = c.SY[last][last], yy = c.YY[last][last];
  const double gamma = (sy > 1e-20 && yy > 1e-20) ? fmin(1e4, fmax(1e-7, sy / yy)) : 0.10;
  const double dl = on ? c.SY[sl][sl] : 1.0;
  const double rdl = 1.0 / dl;
  auto quot = [&](double x) {
    const double q0 = x * rdl;
    return fma(fma(-q0, dl, x), rdl, q0);
  };
  // order[] entries past count are stale but valid slot numbers
  double syr[MH], syc[MH], yyr[MH];
#pragma unroll
  for (int i = 0; i < MH; i++) {
    const int oi = c.order[i];
    syr[i] = c.SY[sl][oi];
    syc[i] = c.SY[oi][sl];
    yyr[i] = c.YY[sl][oi];
  }
  // w = R^{-1} u, R the upper triangle of S'Y (R_ij = s_i . y_j, i <= j)
  double r = on ? c.u[sl] : 0.0, w = 0.0;
#pragma unroll
  for (int i = MH - 1; i >= 0; i--) {
    if (i < k) {
      const double wi = read_lane(quot(r), i);
      if (l == i) w = wi;
      if (l < i) r -= syr[i] * wi;
    }
  }
  // t = (D + gamma Y'Y) w - gamma v (the lane reads first, all independent of
  // the accumulation; lanes past k hold w = 0 and are not added)
  double t = on ? c.SY[sl][sl] * w - gamma * c.v[sl] : 0.0;
  double wr[MH];
#pragma unroll
  for (int j = 0; j < MH; j++) wr[j] = read_lane(w, j);
#pragma unroll
  for (int j = 0; j < MH; j++) t = (on && j < k) ? t + gamma * yyr[j] * wr[j] : t;
  // a = R^{-T} t
  double a = 0.0;
  r = t;
#pragma unroll
  for (int i = 0; i < MH; i++) {
    if (i < k) {
      const double ai = read_lane(quot(r), i);
      if (l == i) a = ai;
      if (on && l > i) r -= syc[i] * ai;
    }
  }
  if (on) {
    C.slot[l] = sl;
    C.cs[l] = a;
    C.cy[l] = -gamma * w;
  }
  if (l == 0) {
    C.k = k;
    C.cg = gamma;
  }
}

// A replay evaluates a trial right after a fresh direction (phase DIRECTION or
// STEEPEST) or on a backtrack (phase SEARCH).
__device__ __forceinline__ bool ctl_trial_phase(const Ctl* C) {
  return C->status == ST_RUNNING &&
         (C->phase == PH_SEARCH || C->phase == PH_DIRECTION || C->phase == PH_STEEPEST);
}

// Direction and trial state of one trial, element-parallel (blocks of 6
// waves over 64 nodes, wave d on DOF plane d).  A fresh direction (phase
// DIRECTION: the control block's compact-form coefficients; STEEPEST: -g) is
// written to p, its g.p terms go to partial row 9, and the trial is z + p;
// phase SEARCH takes the stored p at the control block's step.  Wave 0 then
// evaluates the pose of the block's nodes (state_node: partial rows 0..8 as
// k_state).  partials [10][Pd].
template <class T, int MH>
__global__ __launch_bounds__(kElemThreads) void k_dstate_g(Lat<T> L, const Inst* __restrict__ I,
                                                           const Blk* __restrict__ td, const Ctl* __restrict__ Cs,
                                                           double* __restrict__ z0, double* __restrict__ z1,
                                                           const double* __restrict__ g0, const double* __restrict__ g1,
                                                           const double* __restrict__ Sh, const double* __restrict__ Yh,
                                                           size_t n6, double* __restrict__ p, double* __restrict__ x,
                                                           double* __restrict__ R, const double* __restrict__ xref,
                                                           double* __restrict__ nb0, double* __restrict__ nb1,
                                                           double* __restrict__ part) {
  const Blk bk = td[blockIdx.x];
  const Ctl* C = Cs + bk.inst;
  if (!ctl_trial_phase(C)) return;
  const Inst& in = I[bk.inst];
  const int N = L.N;
  const int d = threadIdx.x / kWave, l = threadIdx.x % kWave, i = bk.first + l;
  const bool on = l < bk.count;
  const int ic = on ? i : bk.first;
  const size_t e = size_t(d) * N + ic;
  const int par = C->par, phase = C->phase;
  const bool locked = (L.lock[ic] >> d) & 1u;
  const double zv = (par ? z1 : z0)[e];
  // wave 0 evaluates the block's node poses below: its inputs are loaded now,
  // so that their latency overlaps the direction
  NodeBase nbase{};
  if (d == 0) nbase = node_base(L, xref, ic);
  double q, dv = 0.0, step;
  if (phase == PH_SEARCH) {
    step = C->step;
    q = p[e];
  } else {
    step = 1.0;
    const double gv = (par ? g1 : g0)[e];
    const bool steep = phase == PH_STEEPEST;
    q = direction_elem<MH>(e, on, on && locked, gv, steep ? 1.0 : C->coef.cg, steep ? 0 : C->coef.k, C->coef, Sh, Yh,
                       n6, in.mt, in.mr, p, dv);
  }
  // trial DOF value, as state_body
  double zz = zv + step * q;
  if (locked) zz = 0.0;
  __shared__ double zs[6][kWave];
  zs[d][l] = zz;
  if (on) (par ? z0 : z1)[e] = zz;  // the candidate buffer
  __syncthreads();
  double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (d == 0 && on) {
    Scalars S = in.S;
    S.top = C->top;
    const double qn[6] = {zs[0][l], zs[1][l], zs[2][l], zs[3][l], zs[4][l], zs[5][l]};
    TermBufs nt{par ? nb1 : nb0, par ? nb0 : nb1};
    double sv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    state_node<T>(L, S, qn, nbase, x, R, xref != nullptr, nt, i, sv);
#pragma unroll
    for (int k = 0; k < 9; k++) v[k] = sv[k];
  }
  v[9] = dv;
  block_reduce<10, kElemThreads>(v, (1u << 3) | (1u << 7));
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 10; k++) part[size_t(k) * gridDim.x + blockIdx.x] = v[k];
}
// Location of the Verlet pair list, read by the graph kernels at run time so
// that a list reallocated with a wider slab needs no new graph.
struct ListRef {
  const int* plist;
  int slab, pad;
};

// An instance's block ranges in the device loop's block tables.  The loop's
// tables list only the instances still running (compacted as a batch drains);
// each instance keeps its own blocks in their own order, so its partials, and
// every fold of them, do not depend on which other instances are listed.
struct GRng {
  int pd0, pd1;  // blocks of the 64-node table
  int pb0, pb1;  // partial columns (wavefronts) of the beams
  int pc0, pc1;  // partial columns (wavefronts) of the contact shapes
  int pad0, pad1;
};

// Node centroid (x, z) of one instance from the state kernel's partial rows 5
// and 6, by one wavefront, in exactly the order of reduce_columns<2, NT>
// (virtual thread t = 64 w + lane sums columns c0 + t, c0 + t + NT, ...;
// the lanes of virtual wave w by wave_sum, the wave totals in ascending w).
template <int NT>
__device__ __forceinline__ void centroid_wave(const double* __restrict__ part, int P, int c0, int c1, int n,
                                              double* __restrict__ out) {
  const int lane = threadIdx.x & (kWave - 1);
  double tot[2];
#pragma unroll
  for (int k = 0; k < 2; k++) {
    double a = 0.0;
    for (int w = 0; w < NT / kWave; w++) {
      double s = 0.0;
      for (int i = c0 + w * kWave + lane; i < c1; i += NT) s += part[size_t(k) * P + i];
      s = wave_sum(s);
      a = w == 0 ? s : a + s;
    }
    tot[k] = a;
  }
  if (lane == 0) {
    out[0] = tot[0] * (1.0 / n);
    out[1] = tot[1] * (1.0 / n);
  }
}

// beams and contact shapes of one trial in a single launch (independent work,
// one launch gap fewer per trial): blocks [0, Pb) evaluate 256 beams each;
// the next ceil(Pc / 4) blocks four chunks of 64 contact shapes each, one per
// wavefront (contact_walk); the next H blocks the chunks flagged by
// k_classify (long lists, contact_group); the last |act| blocks one running
// instance each, the node centroid of its trial for k_gather_g (once per
// instance instead of once per gather block).  The partials have one column
// per 64 items.
template <class T>
__global__ __launch_bounds__(kBlock) void k_elem_g(const Ctl* __restrict__ Cs, Lat<T> L, const Inst* __restrict__ I,
                                                   const Blk* __restrict__ tb, const Blk* __restrict__ tc,
                                                   const double* __restrict__ x, const double* __restrict__ R,
                                                   double* __restrict__ rec, double* __restrict__ ub0,
                                                   double* __restrict__ ub1, double* __restrict__ bpart, int Pb,
                                                   const unsigned* __restrict__ pcount,
                                                   const ListRef* __restrict__ lr, double* __restrict__ cf_sph,
                                                   double* __restrict__ cf_cap, double* __restrict__ cb0,
                                                   double* __restrict__ cb1, double* __restrict__ cpart, int Pc,
                                                   const int* __restrict__ cflag, const int* __restrict__ hlist,
                                                   const unsigned* __restrict__ hcount, int H,
                                                   const int* __restrict__ act, const GRng* __restrict__ G,
                                                   const double* __restrict__ dspart, int Pd,
                                                   double* __restrict__ cent, EllOut eo) {
  const int Pcb = (Pc + kBW - 1) / kBW;  // blocks of four contact chunks
  const int b = blockIdx.x;
  if (b < Pb) {
    const Blk bk = tb[b];
    const Ctl* C = Cs + bk.inst;
    if (!ctl_trial_phase(C)) return;
    const int par = C->par;
    TermBufs bt{par ? ub1 : ub0, par ? ub0 : ub1};
    beam_body<T>(L, I[bk.inst].M, x, R, nullptr, nullptr, bt, bpart, bk.first + threadIdx.x,
                 int(threadIdx.x) < bk.count, b, Pb, eo);
  } else if (b < Pb + Pcb) {
    // wavefront w walks chunk 4 (b - Pb) + w, one lane per shape, unless the
    // chunk is flagged for the lane groups below (no barriers on this path)
    const int j = (b - Pb) * kBW + int(threadIdx.x) / kWave, l = threadIdx.x & (kWave - 1);
    if (j >= Pc || cflag[j]) return;
    const Blk bk = tc[j];
    const Ctl* C = Cs + bk.inst;
    if (!ctl_trial_phase(C)) return;
    const int par = C->par;
    TermBufs ct{par ? cb1 : cb0, par ? cb0 : cb1};
    contact_walk<T>(L, I[bk.inst], x, pcount, lr->plist, lr->slab, cf_sph, nullptr, ct, cpart, bk.first + l,
                    l < bk.count, size_t(j), size_t(Pc), eo);
  } else if (b < Pb + Pcb + H) {
    // H blocks share the flagged chunks (block-uniform loop)
    const unsigned nh = *hcount;
    for (unsigned q = b - Pb - Pcb; q < nh; q += H) {
      const int j = hlist[q];
      const Blk bk = tc[j];
      const Ctl* C = Cs + bk.inst;
      if (!ctl_trial_phase(C)) continue;
      const int par = C->par;
      TermBufs ct{par ? cb1 : cb0, par ? cb0 : cb1};
      contact_group<T>(L, I[bk.inst], x, pcount, lr->plist, lr->slab, cf_sph, nullptr, ct, cpart, bk.first,
                       bk.count, j, Pc, eo);
    }
  } else {
    const int k = act[b - Pb - Pcb - H];
    if (threadIdx.x < kWave && ctl_trial_phase(Cs + k))
      centroid_wave<kElemThreads>(dspart + 5 * size_t(Pd), Pd, G[k].pd0, G[k].pd1, I[k].N, cent + 2 * k);
  }
}

// ---------------------------------------------------------------------------
// Rows of the device loop's decision: 0..8 the state rows of k_dstate_g,
// 9..12 contact shapes, 13..16 beams, 17..19 gradient (sum g^2, max |g|, g.p),
// 20 the fresh direction's g.p (k_dstate_g row 9), 21.. the curvature-pair
// products (gather rows 3..).
// ---------------------------------------------------------------------------
constexpr int kGDecRows = 21 + 4 * kMaxHist + 5;
struct GraphParts {
  const double *spart, *cpart, *bpart, *gpart;  // [10][Pd], [4][Pc], [4][Pb], [kGRows][Pd]
  int Pd, Pc, Pb;                               // Pc, Pb: wavefront columns
};
__device__ __forceinline__ DecRow graph_row(int j, const GRng& g, const GraphParts& P) {
  if (j < 9) return {P.spart + size_t(j) * P.Pd + g.pd0, g.pd1 - g.pd0, row_max(j)};
  if (j < 13) return {P.cpart + size_t(j - 9) * P.Pc + g.pc0, g.pc1 - g.pc0, row_max(j)};
  if (j < 17) return {P.bpart + size_t(j - 13) * P.Pb + g.pb0, g.pb1 - g.pb0, false};
  if (j == 20) return {P.spart + size_t(9) * P.Pd + g.pd0, g.pd1 - g.pd0, false};
  return {P.gpart + size_t(j < 20 ? j - 17 : j - 18) * P.Pd + g.pd0, g.pd1 - g.pd0, row_max(j)};
}

// The deciding block copies the control block to LDS (every load of the copy
// issued before the first store, so that the copy costs one memory latency),
// runs its scalar logic there, and copies it back.
static_assert(sizeof(Ctl) % 8 == 0, "Ctl is copied as 8-byte words");
constexpr int kCtlWords = int(sizeof(Ctl) / 8);
template <int NT>
__device__ __forceinline__ void ctl_load(Ctl& sc, const Ctl* C) {
  constexpr int RD = (kCtlWords + NT - 1) / NT;
  double cr[RD];
#pragma unroll
  for (int q = 0; q < RD; q++)
    cr[q] = reinterpret_cast<const double*>(C)[min(int(threadIdx.x) + q * NT, kCtlWords - 1)];
#pragma unroll
  for (int q = 0; q < RD; q++) {
    const int e = threadIdx.x + q * NT;
    if (e < kCtlWords) reinterpret_cast<double*>(&sc)[e] = cr[q];
  }
}
template <int NT>
__device__ __forceinline__ void ctl_store(Ctl* C, const Ctl& sc) {
  constexpr int RD = (kCtlWords + NT - 1) / NT;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < RD; q++) {
    const int e = threadIdx.x + q * NT;
    if (e < kCtlWords) reinterpret_cast<double*>(C)[e] = reinterpret_cast<const double*>(&sc)[e];
  }
}

// The decision on the folded rows (MC/src/sim.cpp:178-183,191-207), by one
// block of NT threads: the descent check of a fresh direction (:191; a
// non-descent direction discards the trial and steepest descent follows),
// the candidate record, Armijo (the reference's total-energy test or the
// per-term change; the mixed policy adds the approximate-Wolfe test of Hager
// & Zhang, SIAM J. Optim. 16, 2005, for changes below its energy noise), and
// on acceptance the curvature-pair update, the swap of current and candidate
// buffers, the next convergence test and the next coefficients.
// decide_core runs it on a control block in LDS (every thread of the block
// calls it; returns after a barrier); decide_block copies the control block
// in from global memory and back.
template <int NT, int MH = kMaxHist>
__device__ void decide_core(Ctl& sc, const RecConst& in, const double* rows, int fl, int* __restrict__ iflag) {
  static_assert(NT >= 2 * kWave, "the decision runs on two wavefronts");
  __shared__ int accepted;
  // The pair update and the next direction's coefficients do not depend on
  // the acceptance test, only on whether it passes: wavefront 1 computes them
  // on a copy of the control block (spec) while thread 0 decides, and they
  // are kept on acceptance, the common outcome.  Coefficients solved after an
  // acceptance that ends the solve (converged, iteration cap) are unused.
  __shared__ Ctl spec;
  constexpr int kHistWord = int(offsetof(Ctl, coef) / 8);  // coef, order, SY, YY, u, v to the end
  static_assert(offsetof(Ctl, coef) % 8 == 0, "the history tail is copied as 8-byte words");
  __syncthreads();
  for (int w = threadIdx.x; w < kCtlWords; w += NT)
    reinterpret_cast<double*>(&spec)[w] = reinterpret_cast<const double*>(&sc)[w];
  __syncthreads();
  if (threadIdx.x == 0) {
    Ctl* c = &sc;
    *iflag = 0;
    c->evals++;
    accepted = 0;
    bool discard = false;
    if (c->phase != PH_SEARCH) {
      const double dg = rows[20];
      if (c->phase == PH_DIRECTION && !(dg < 0)) {
        c->count = 0;
        c->phase = PH_STEEPEST;
        discard = true;
      } else {
        c->dg = dg;
        c->step = 1.0;
        c->bt = 0;
        c->phase = PH_SEARCH;
      }
    }
    if (!discard) {
      record_from_rows(in, rows, fl, c->cur, c->cand);
      const double* r = c->cand;
      if (r[11] >= c->skin_half) {  // contact list stale at this trial: rebuild, then redo it
        c->status = ST_HOLD;
      } else {
        c->backtracks++;
        const double dE = c->ls_mode ? r[12] : r[0] - c->cur[0];
        const double bound = c->armijo * c->step * c->dg;
        bool ok = r[9] != 0.0 && dE <= bound;
        if (!ok && c->wolfe && r[9] != 0.0 && fabs(dE) <= c->wolfe_eps * fabs(c->cur[0]))
          ok = r[14] <= (2.0 * c->armijo - 1.0) * c->dg;
        if (ok) {
          accepted = 1;
          c->par ^= 1;
          for (int k = 0; k < kRec; k++) c->cur[k] = c->cand[k];
          c->it++;
          c->retry = 0;
          if (c->it >= c->max_it) {
            c->status = ST_ITERCAP;
          } else {
            c->iterations = c->it + 1;
            const double* rc = c->cur;
            const double rel = rc[6] / fmax(1e-6, c->ts * fmax(1.0, rc[4]));
            const bool conv = (rc[6] <= c->tol_rms && rc[7] <= c->tol_max) || (c->rel_escape && rel < c->rel_tol);
            if (conv && rc[5] <= c->pen_tol)
              c->status = ST_CONVERGED;
            else
              c->phase = PH_DIRECTION;
          }
        } else {
          c->step *= 0.5;
          c->bt++;
          if (c->bt >= c->max_ls || c->step < c->min_step) {
            if (c->retry == 0) {  // steepest-descent retry
              c->retry = 1;
              c->count = 0;
              c->phase = PH_STEEPEST;
            } else {
              c->status = ST_FAILED;
            }
          }
        }
      }
    }
  } else if (threadIdx.x >= kWave && threadIdx.x < 2 * kWave) {
    // pair update from the products in rows 21.. and the coefficients, by one
    // wavefront: its LDS accesses complete in program order, and the fence
    // keeps the compiler from moving a lane's reads of the updated history
    // (other lanes' writes) above the update
    if (spec.m > 0) history_update_wave(spec, rows + 21);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    compact_coefficients_wave<MH>(spec, spec.coef);
  }
  __syncthreads();
  if (accepted) {
    for (int w = kHistWord + threadIdx.x; w < kCtlWords; w += NT)
      reinterpret_cast<double*>(&sc)[w] = reinterpret_cast<const double*>(&spec)[w];
    if (threadIdx.x == 0) {
      sc.count = spec.count;
      sc.spare = spec.spare;
    }
  }
  __syncthreads();
}
template <int NT, int MH = kMaxHist>
__device__ void decide_block(Ctl* __restrict__ C, const RecConst& in, const double* rows, int fl,
                             int* __restrict__ iflag) {
  __shared__ Ctl sc;
  ctl_load<NT>(sc, C);
  decide_core<NT, MH>(sc, in, rows, fl, iflag);
  ctl_store<NT>(C, sc);
}

// Gradient assembly of the trial (element-parallel, as k_gather; the node
// centroid of the centering term comes from k_elem_g's cent), then
// (history on) the trial's curvature pair s = zc - z, y = gc - g written into
// the spare history slot with all its products (partial rows 3..: for each
// retained pair j, oldest first: s_j.y, y_j.y, s_j.gc, y_j.gc; then s.y, s.s,
// y.y, s.gc, y.gc), computed by the thread that assembled the element.  The
// history elements are loaded before the assembly so that their latency
// overlaps it.
// MH: the history capacity the kernel is built for (the arrays below live in
// registers; the common m <= 10 gets the smaller build, with half the
// register footprint of the m <= 16 one).
// FD: the trial's decision is taken here, by the instance's last gather block
// to finish (the arrival as in k_decide_g), instead of by k_decide_g after
// the gather: one launch fewer per trial.  The rows the earlier kernels of the
// trial wrote (state, contact, beam rows and the fresh direction's g.p) are
// folded by the instance's first blocks before they arrive, the gather rows
// by the last block; every row by one wave in k_decide_g's order, so the
// choice (the loop's tuner makes it) changes no result.
struct FuseArgs {
  const GRng* G;                         // block ranges per instance
  const double *dspart, *cpart, *bpart;  // the state, contact and beam rows
  int Pc, Pb;                            // their wavefront columns
  int* iflags;                           // per-instance evaluation flags
  unsigned* count;                       // per-instance arrival counters (0 between trials)
  double* drows;                         // [n_inst][kGDecRows] rows folded before the arrival
};
// rows of the decision known before the gather: 0..16 and 20; the gather's:
// 17..19, then 21..
constexpr int kEarlyRows = 18;
__device__ __forceinline__ int early_row(int jj) { return jj < 17 ? jj : 20; }
__device__ __forceinline__ int gather_row(int jj) { return jj < 3 ? 17 + jj : 18 + jj; }
template <class T, int MH, bool FD>
__global__ __launch_bounds__(kElemThreads) void k_gather_g(Lat<T> L, const Inst* __restrict__ I,
                                                           const Blk* __restrict__ td,
                                                           Ctl* __restrict__ Cs, const double* __restrict__ x,
                                                           const double* __restrict__ erec,
                                                           const double* __restrict__ cf_sph,
                                                           const double* __restrict__ ecap, long eS,
                                                           const double* __restrict__ cent,
                                                           const double* __restrict__ z0, const double* __restrict__ z1,
                                                           double* __restrict__ g0, double* __restrict__ g1,
                                                           const double* __restrict__ p, double* __restrict__ Sh,
                                                           double* __restrict__ Yh, size_t n6,
                                                           double* __restrict__ part, FuseArgs fa) {
  const Blk bk = td[blockIdx.x];
  const Ctl* C = Cs + bk.inst;
  if (!ctl_trial_phase(C)) return;
  const Inst& in = I[bk.inst];
  const int Pd = gridDim.x;
  const int d = threadIdx.x / kWave, l = threadIdx.x % kWave, i = bk.first + l;
  const bool on = l < bk.count;
  const int ic = on ? i : bk.first;
  const size_t e = size_t(d) * L.N + ic;
  const int par = C->par, m = C->m, count = C->count, spare = C->spare;
  const int off = count == m ? 1 : 0, nold = m > 0 ? count - off : 0;
  double* gnew = par ? g0 : g1;
  const double* gcur = par ? g1 : g0;
  const double* zcur = par ? z1 : z0;
  const double* zcand = par ? z0 : z1;
  constexpr int NA = 3 + 4 * MH + 5;  // gradient rows, then pair products
  // rows past NA stay zero: the row butterfly of block_rows takes 64 rows
  constexpr int NP = NA <= kWave ? kWave : NA;
  double hs[MH], hy[MH];
#pragma unroll
  for (int jj = 0; jj < MH; jj++) {
    hs[jj] = 0.0;
    hy[jj] = 0.0;
    if (jj < nold) {
      const size_t o = size_t(C->order[jj + off]) * n6 + e;
      hs[jj] = Sh[o];
      hy[jj] = Yh[o];
    }
  }
  const double zcv = zcand[e], zv = zcur[e], gov = gcur[e], pv = p[e];
  const bool locked = (L.lock[ic] >> d) & 1u;
  const int deg = L.inc_off[ic + 1] - L.inc_off[ic];
  const double c0 = cent[2 * bk.inst], c1 = cent[2 * bk.inst + 1];
  const double gv = gather_ell<T>(L, in, C->top, x, erec, ecap, eS, cf_sph, c0, c1, d, ic, deg,
                                  long(bk.pad >> 5) + l, bk.pad & 31);
  const double gd = locked ? 0.0 : gv;
  if (on) gnew[e] = gd;
  double a[NP];
#pragma unroll
  for (int r = 0; r < NP; r++) a[r] = 0.0;
  a[0] = on && !locked ? gd * gd : 0.0;
  a[1] = on && !locked ? fabs(gd) : 0.0;
  a[2] = on ? gd * pv : 0.0;
  int NR = 3;
  if (m > 0) {
    const double sv = on ? zcv - zv : 0.0, yv = on ? gd - gov : 0.0, gc = on ? gd : 0.0;
    if (on) {
      Sh[size_t(spare) * n6 + e] = sv;
      Yh[size_t(spare) * n6 + e] = yv;
    }
#pragma unroll
    for (int jj = 0; jj < MH; jj++) {
      if (jj < nold) {
        a[3 + 4 * jj + 0] = hs[jj] * yv;
        a[3 + 4 * jj + 1] = hy[jj] * yv;
        a[3 + 4 * jj + 2] = hs[jj] * gc;
        a[3 + 4 * jj + 3] = hy[jj] * gc;
      }
    }
    const double w[5] = {sv * yv, sv * sv, yv * yv, sv * gc, yv * gc};
#pragma unroll
    for (int jj = 0; jj <= MH; jj++)
      if (jj == nold)
#pragma unroll
        for (int c = 0; c < 5; c++) a[3 + 4 * jj + c] = w[c];
    NR = 3 + 4 * nold + 5;
  }
  __shared__ double sh[kElemThreads / kWave][NP];
  const double t = block_rows<kElemThreads / kWave>(a, NR, sh);
  if (int(threadIdx.x) < NR) part[size_t(threadIdx.x) * Pd + blockIdx.x] = t;
  if constexpr (FD) {
    constexpr int NW = kElemThreads / kWave;
    const int fl = fa.iflags[bk.inst];
    const RecConst rc = rec_const(in);
    const GRng g = fa.G[bk.inst];
    const int q = blockIdx.x - g.pd0, nblk = g.pd1 - g.pd0;
    const GraphParts P{fa.dspart, fa.cpart, fa.bpart, part, Pd, fa.Pc, fa.Pb};
    double* drow = fa.drows + size_t(bk.inst) * kGDecRows;
    if (q * NW < kEarlyRows)  // (block-uniform) 3 rows per wave cover them with one block
      rows_pass<(kEarlyRows + NW - 1) / NW, 16>([&](int jj) { return graph_row(early_row(jj), g, P); },
                                                 P.spart + g.pd0, kEarlyRows, q * NW, nblk * NW,
                                                 [&](int jj, double v) { drow[early_row(jj)] = v; });
    __shared__ int last;
    __syncthreads();
    if (threadIdx.x == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      last = atomicAdd(fa.count + bk.inst, 1u) == unsigned(g.pd1 - g.pd0 - 1);
      if (last) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        fa.count[bk.inst] = 0;
      }
    }
    __syncthreads();
    if (!last) return;
    // the gather rows at this trial's history (3, then 4 per retained pair and
    // 5 for the new one), then the early rows from drows
    constexpr int NG = (3 + 4 * (MH - 1) + 5 + NW - 1) / NW;
    __shared__ double rows[kGDecRows];
    const int RG = 3 + (m > 0 ? 4 * nold + 5 : 0);
    rows_pass<NG, kDecColRounds>([&](int jj) { return graph_row(gather_row(jj), g, P); }, P.spart + g.pd0, RG, 0, NW,
                                 [&](int jj, double v) { rows[gather_row(jj)] = v; });
    if (int(threadIdx.x) < kEarlyRows) rows[early_row(threadIdx.x)] = drow[early_row(threadIdx.x)];
    __syncthreads();
    decide_block<kElemThreads, MH>(Cs + bk.inst, rc, rows, fl, fa.iflags + bk.inst);
  }
}

// The decision of one trial: D blocks per instance, each folding its
// share of the decision rows (rows_pass) into drows, so that the row loads are
// spread over several compute units; the block that arrives last at the
// instance's counter reads the folded rows and decides (decide_core on the
// control block, which every block reads at its start), and resets the
// counter for the next trial.  Arrival: a block barrier completes
// the block's row stores, one thread releases them (device-scope release
// fence: L2 write-back, for the other XCDs) and counts the arrival; only the
// last arriver acquires (invalidates its cached lines) before reading.
// D: blocks per instance, chosen by the loop's tuner (any D gives the same
// rows: each row is folded by one wavefront in one fixed order).  D = 1: the
// instance's one block folds every row into LDS and decides, with no arrival
// and no fence (on batches of many lattices each block's L2 write-back is the
// decision's largest cost).  MH: the history capacity the decision's
// coefficient solve is built for.
template <int D, int MH>
__global__ __launch_bounds__(kBlock) void k_decide_g(Ctl* __restrict__ Cs, const Inst* __restrict__ I,
                                                     const int* __restrict__ act, const GRng* __restrict__ G,
                                                     const double* __restrict__ dspart, int Pd,
                                                     const double* __restrict__ cpart, int Pc,
                                                     const double* __restrict__ bpart, int Pb,
                                                     const double* __restrict__ gpart, int* __restrict__ iflags,
                                                     double* drows, unsigned* dcount) {
  // rows at a full history of MH pairs (MH - 1 retained and the new one)
  constexpr int kRows = 21 + 4 * (MH - 1) + 5;
  constexpr int NQ = (kRows + D * kDecWaves - 1) / (D * kDecWaves);
  const int b = act[blockIdx.x / D], part = blockIdx.x % D;
  Ctl* C = Cs + b;
  const Inst& in = I[b];
  if (!ctl_trial_phase(C)) return;
  // what the deciding block needs besides the rows, read ahead (the release
  // fence below waits for these loads): the flag word, the instance
  // constants and the control block (only the deciding block writes it)
  const int fl = iflags[b];
  const RecConst rc = rec_const(in);
  constexpr int RD = (kCtlWords + kBlock - 1) / kBlock;
  double cr[RD];
#pragma unroll
  for (int q = 0; q < RD; q++)
    cr[q] = reinterpret_cast<const double*>(C)[min(int(threadIdx.x) + q * kBlock, kCtlWords - 1)];
  __shared__ int last;
  __shared__ double rows[kGDecRows];
  {
    const int m = C->m, count = C->count;
    const int R = 21 + (m > 0 ? 4 * (count - (count == m ? 1 : 0)) + 5 : 0);
    const GraphParts P{dspart, cpart, bpart, gpart, Pd, Pc, Pb};
    const GRng g = G[b];
    // first-round columns per lane: as many as 32 registers per row pair hold
    // (the beam and contact rows have one column per 64 items); with many
    // rows per wave (few blocks per lattice, chosen for batches of small
    // lattices) two, and one at one block per lattice (chosen for many small
    // lattices), so that the fold's registers leave room for more resident
    // blocks
    constexpr int RD = NQ <= 2 ? 16 : (NQ <= 4 ? 8 : (D == 1 ? 1 : 2));
    double* out = drows + size_t(b) * kGDecRows;
    // (rows_pass ends with a block barrier)
    if constexpr (D == 1)
      rows_pass<NQ, RD>([&](int j) { return graph_row(j, g, P); }, dspart + g.pd0, R, 0, kDecWaves,
                        [&](int j, double t) { rows[j] = t; });
    else
      rows_pass<NQ, RD>([&](int j) { return graph_row(j, g, P); }, dspart + g.pd0, R, part * kDecWaves,
                        D * kDecWaves, [&](int j, double t) { out[j] = t; });
  }
  if constexpr (D > 1) {
    if (threadIdx.x == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      last = atomicAdd(dcount + b, 1u) == unsigned(D - 1);
      if (last) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        dcount[b] = 0;
      }
    }
    __syncthreads();
    if (!last) return;
    if (int(threadIdx.x) < kGDecRows) rows[threadIdx.x] = drows[size_t(b) * kGDecRows + threadIdx.x];
  }
  __shared__ Ctl sc;
#pragma unroll
  for (int q = 0; q < RD; q++) {
    const int w = threadIdx.x + q * kBlock;
    if (w < kCtlWords) reinterpret_cast<double*>(&sc)[w] = cr[q];
  }
  __syncthreads();
  decide_core<kBlock, MH>(sc, rc, rows, fl, iflags + b);
  ctl_store<kBlock>(C, sc);
}

// ---------------------------------------------------------------------------
// Block solver: the whole minimisation of one lattice in one workgroup
// (MC/src/sim.cpp:174-211, the same phase machine as the graph-replayed
// loop), for lattices small enough that one workgroup holds their work.  One
// launch takes the running instances, one workgroup each, and every
// workgroup loops over trials until its instance has converged, failed, hit
// its iteration cap or found its contact list stale (HOLD: the host rebuilds
// the list and launches again), or until max_trials (a bound the phase
// machine never reaches; it only guarantees the exit).  The control block
// lives in LDS for the whole launch; the only synchronisation is the block
// barrier, so lattices of any mix of sizes run side by side without host
// round trips or graph replays.  A trial, as the graph loop's four kernels:
//   nodes    thread per node: fresh direction (its six DOFs, locks, clamps)
//            or the stored one, trial DOFs, pose, gravity, platens
//   beams    thread per beam (beam_body); contact shapes thread per shape
//            (contact_body)
//   gather   thread per node, its six DOFs in turn: gradient assembly, the
//            curvature pair into the spare slot and its products
//   decide   decide_core on the block's folded rows
// Partial sums are kept per thread across the passes over the instance's
// items and folded by the block in a fixed order: results are reproducible,
// and a lattice's results do not depend on the other lattices of the launch.
// ---------------------------------------------------------------------------
constexpr int kSolveThreads = 512, kSolveWaves = kSolveThreads / kWave;
// largest lattice (nodes) the block solver takes: one pass of the workgroup
// over the nodes
constexpr int kBlockEngineNodes = kSolveThreads;
constexpr int kSolveBudget = 64;  // trials per workgroup and launch

template <class T, int MH>
__global__ __launch_bounds__(kSolveThreads) void k_solve_block(
    Lat<T> L, const Inst* __restrict__ I, Ctl* __restrict__ Cs, const int* __restrict__ act, int max_trials,
    double* __restrict__ z0, double* __restrict__ z1, double* __restrict__ g0, double* __restrict__ g1,
    double* __restrict__ Sh, double* __restrict__ Yh, size_t n6, double* __restrict__ p, double* __restrict__ x,
    double* __restrict__ R, const double* __restrict__ xref, double* __restrict__ nb0, double* __restrict__ nb1,
    double* __restrict__ ub0, double* __restrict__ ub1, double* __restrict__ cb0, double* __restrict__ cb1,
    const unsigned* __restrict__ pcount, const ListRef* __restrict__ lr, double* __restrict__ rec,
    double* __restrict__ cf_sph, double* __restrict__ cf_cap, int* __restrict__ iflags) {
  const int b = act[blockIdx.x];
  const Inst& in = I[b];
  const int N = L.N, tid = threadIdx.x;
  constexpr int NA = 3 + 4 * MH + 5;  // gradient rows, then pair products
  constexpr int NP = NA <= kWave ? kWave : NA;
  __shared__ Ctl sc;
  __shared__ double rows[kGDecRows];
  __shared__ double partw[4][kSolveWaves];  // one column per wavefront
  __shared__ double shr[kSolveWaves][NP];
  __shared__ double qsh[6][kSolveThreads];  // raw directions, one column per thread
  ctl_load<kSolveThreads>(sc, Cs + b);
  __syncthreads();
  for (int t = 0; t < max_trials && sc.status == ST_RUNNING; t++) {
    const int par = sc.par, phase = sc.phase;
    const bool fresh = phase != PH_SEARCH, steep = phase == PH_STEEPEST;
    const double step = fresh ? 1.0 : sc.step;
    double* zcur = par ? z1 : z0;
    double* zcand = par ? z0 : z1;
    const double* gcur = par ? g1 : g0;
    double* gnew = par ? g0 : g1;
    // ---- nodes: direction and trial state ----
    {
      const int kk = steep ? 0 : sc.coef.k;
      const double cg = steep ? 1.0 : sc.coef.cg;
      Scalars S = in.S;
      S.top = sc.top;
      const TermBufs nt{par ? nb1 : nb0, par ? nb0 : nb1};
      double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int base = 0; base < in.N; base += kSolveThreads) {
        const int li = base + tid;
        if (li >= in.N) continue;
        const int i = in.n0 + li;
        const unsigned lk = L.lock[i];
        const NodeBase nbase = node_base(L, xref, i);
        double q[6];
        if (fresh) {
          // one DOF at a time (its history loads in flight together; more at
          // once spill), the raw direction parked in this thread's LDS column
#pragma unroll 1
          for (int d = 0; d < 6; d++) {
            const size_t e = size_t(d) * N + i;
            double hs[MH], hy[MH];
#pragma unroll
            for (int j = 0; j < MH; j++) {
              if (j < kk) {
                hs[j] = Sh[n6 * sc.coef.slot[j] + e];
                hy[j] = Yh[n6 * sc.coef.slot[j] + e];
              }
            }
            double v = cg * gcur[e];
#pragma unroll
            for (int j = 0; j < MH; j++)
              if (j < kk) v += sc.coef.cs[j] * hs[j] + sc.coef.cy[j] * hy[j];
            v = -v;
            qsh[d][tid] = ((lk >> d) & 1u) ? 0.0 : v;
          }
#pragma unroll
          for (int d = 0; d < 6; d++) q[d] = qsh[d][tid];
#pragma unroll
          for (int h = 0; h < 2; h++) {  // per-node clamps: translation, rotation
            const double len = sqrt(q[3 * h] * q[3 * h] + q[3 * h + 1] * q[3 * h + 1] + q[3 * h + 2] * q[3 * h + 2]);
            const double lim = h == 0 ? in.mt : in.mr;
#pragma unroll
            for (int c = 3 * h; c < 3 * h + 3; c++) {
              if (len > lim) q[c] *= lim / len;
              if ((lk >> c) & 1u) q[c] = 0.0;
            }
          }
#pragma unroll
          for (int d = 0; d < 6; d++) {
            const size_t e = size_t(d) * N + i;
            p[e] = q[d];
            acc[9] += gcur[e] * q[d];
          }
        } else {
#pragma unroll
          for (int d = 0; d < 6; d++) q[d] = p[size_t(d) * N + i];
        }
        double zz[6];
#pragma unroll
        for (int d = 0; d < 6; d++) {
          const size_t e = size_t(d) * N + i;
          zz[d] = zcur[e] + step * q[d];
          if ((lk >> d) & 1u) zz[d] = 0.0;
          zcand[e] = zz[d];
        }
        double sv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        state_node<T>(L, S, zz, nbase, x, R, xref != nullptr, nt, i, sv);
#pragma unroll
        for (int k = 0; k < 9; k++) acc[k] = (k == 3 || k == 7) ? fmax(acc[k], sv[k]) : acc[k] + sv[k];
      }
      block_reduce<10, kSolveThreads>(acc, (1u << 3) | (1u << 7));
      if (tid == 0) {
#pragma unroll
        for (int k = 0; k < 9; k++) rows[k] = acc[k];
        rows[20] = acc[9];
      }
    }
    // ---- beams, then contact shapes ----
    {
      const TermBufs bt{par ? ub1 : ub0, par ? ub0 : ub1};
      double acc[4] = {0, 0, 0, 0};
      for (int base = 0; base < in.B; base += kSolveThreads) {
        const int li = base + tid;
        const bool on = li < in.B;
        beam_body<T, kSolveThreads>(L, in.M, x, R, rec, nullptr, bt, &partw[0][0], in.b0 + (on ? li : 0), on, 0,
                                    1);
        __syncthreads();
        if (tid == 0)
#pragma unroll
          for (int c = 0; c < 4; c++)
#pragma unroll
            for (int w = 0; w < kSolveWaves; w++) acc[c] += partw[c][w];
        __syncthreads();
      }
      const TermBufs ct{par ? cb1 : cb0, par ? cb0 : cb1};
      const int* plist = lr->plist;
      const int slab = lr->slab;
      double accc[4] = {0, 0, 0, 0};
      for (int base = 0; base < in.NS; base += kSolveThreads) {
        const int li = base + tid;
        const bool on = li < in.NS;
        contact_body<T, kSolveThreads>(L, in, x, pcount, plist, slab, cf_sph, cf_cap, ct, &partw[0][0],
                                       in.s0 + (on ? li : 0), on, 0, 1);
        __syncthreads();
        if (tid == 0)
#pragma unroll
          for (int w = 0; w < kSolveWaves; w++) {
            accc[0] += partw[0][w];
            accc[1] = fmax(accc[1], partw[1][w]);
            accc[2] += partw[2][w];
            accc[3] += partw[3][w];
          }
        __syncthreads();
      }
      if (tid == 0)
#pragma unroll
        for (int c = 0; c < 4; c++) {
          rows[9 + c] = accc[c];
          rows[13 + c] = acc[c];
        }
    }
    __syncthreads();
    // ---- gather: gradient, curvature pair and its products ----
    {
      const int m = sc.m, count = sc.count, spare = sc.spare;
      const int off = count == m ? 1 : 0, nold = m > 0 ? count - off : 0;
      const double c0 = rows[5] * (1.0 / in.N), c1 = rows[6] * (1.0 / in.N);
      int slots[MH];
#pragma unroll
      for (int jj = 0; jj < MH; jj++) slots[jj] = jj < nold ? sc.order[jj + off] : 0;
      double a[NP];
#pragma unroll
      for (int r = 0; r < NP; r++) a[r] = 0.0;
      // pass 1: gradient assembly (gather_node: all six components from one
      // walk of the incidence list), norms, g.p, and the pair s, y written
      for (int base = 0; base < in.N; base += kSolveThreads) {
        const int li = base + tid;
        if (li >= in.N) continue;
        const int i = in.n0 + li;
        const unsigned lk = L.lock[i];
        double gval[6];
        gather_node<T>(L, in, sc.top, x, rec, cf_sph, cf_cap, c0, c1, i, gval);
#pragma unroll
        for (int d = 0; d < 6; d++) {
          const size_t e = size_t(d) * N + i;
          const bool locked = (lk >> d) & 1u;
          const double gd = locked ? 0.0 : gval[d];
          gnew[e] = gd;
          if (!locked) {
            a[0] += gd * gd;
            a[1] = fmax(a[1], fabs(gd));
          }
          a[2] += gd * p[e];
          if (m > 0) {
            Sh[size_t(spare) * n6 + e] = zcand[e] - zcur[e];
            Yh[size_t(spare) * n6 + e] = gd - gcur[e];
          }
        }
      }
      // pass 2: the pair's products with the retained pairs and itself (each
      // thread re-reads only what it wrote in pass 1)
      if (m > 0) {
        for (int base = 0; base < in.N; base += kSolveThreads) {
          const int li = base + tid;
          if (li >= in.N) continue;
          const int i = in.n0 + li;
#pragma unroll 1
          for (int d = 0; d < 6; d++) {
            const size_t e = size_t(d) * N + i;
            double hs[MH], hy[MH];
#pragma unroll
            for (int jj = 0; jj < MH; jj++) {
              hs[jj] = 0.0;
              hy[jj] = 0.0;
              if (jj < nold) {
                hs[jj] = Sh[size_t(slots[jj]) * n6 + e];
                hy[jj] = Yh[size_t(slots[jj]) * n6 + e];
              }
            }
            const double gd = gnew[e], sv = Sh[size_t(spare) * n6 + e], yv = Yh[size_t(spare) * n6 + e];
#pragma unroll
            for (int jj = 0; jj < MH; jj++) {
              if (jj < nold) {
                a[3 + 4 * jj + 0] += hs[jj] * yv;
                a[3 + 4 * jj + 1] += hy[jj] * yv;
                a[3 + 4 * jj + 2] += hs[jj] * gd;
                a[3 + 4 * jj + 3] += hy[jj] * gd;
              }
            }
            const double w[5] = {sv * yv, sv * sv, yv * yv, sv * gd, yv * gd};
#pragma unroll
            for (int jj = 0; jj <= MH; jj++)
              if (jj == nold)
#pragma unroll
                for (int c = 0; c < 5; c++) a[3 + 4 * jj + c] += w[c];
          }
        }
      }
      const int NR = m > 0 ? 3 + 4 * nold + 5 : 3;
      const double tt = block_rows<kSolveWaves>(a, NR, shr);
      if (tid < NR) rows[tid < 3 ? 17 + tid : 18 + tid] = tt;
    }
    __syncthreads();
    decide_core<kSolveThreads, MH>(sc, rec_const(in), rows, iflags[b], iflags + b);
  }
  ctl_store<kSolveThreads>(Cs + b, sc);
}

// list reference positions of the nodes of a node block table: xref = x
__global__ __launch_bounds__(kBlock) void k_copy_xref(const Blk* __restrict__ tn, int N, const double* __restrict__ x,
                                                      double* __restrict__ xref) {
  const Blk bk = tn[blockIdx.x];
  if (int(threadIdx.x) >= bk.count) return;
  const int i = bk.first + threadIdx.x;
#pragma unroll
  for (int d = 0; d < 3; d++) xref[size_t(d) * N + i] = x[size_t(d) * N + i];
}

__global__ __launch_bounds__(kBlock) void k_fill_zero(double* __restrict__ a, long n) {
  long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

// commit, MC/src/sim.cpp:212: base pose <- state(z); plastic <- trial.next,
// for the instances whose mask entry is set
__global__ __launch_bounds__(kBlock) void k_commit_nodes(const Blk* __restrict__ tn, const int* __restrict__ mask, int N,
                                                         const double* __restrict__ x, const double* __restrict__ R,
                                                         double* __restrict__ bpos, double* __restrict__ bquat) {
  const Blk bk = tn[blockIdx.x];
  if (!mask[bk.inst] || int(threadIdx.x) >= bk.count) return;
  const int i = bk.first + threadIdx.x;
  for (int d = 0; d < 3; d++) bpos[d * N + i] = x[d * N + i];
  for (int d = 0; d < 4; d++) bquat[d * N + i] = R[d * N + i];
}
__global__ __launch_bounds__(kBlock) void k_commit_beams(const Blk* __restrict__ tb, const int* __restrict__ mask, int B,
                                                         const double* __restrict__ trial,
                                                         double* __restrict__ plastic) {
  const Blk bk = tb[blockIdx.x];
  if (!mask[bk.inst] || int(threadIdx.x) >= bk.count) return;
  const int k = bk.first + threadIdx.x;
  for (int c = 0; c < 7; c++) plastic[c * B + k] = trial[c * B + k];
}

// diagnostics reductions, MC/src/sim.cpp:339-370.
// beam partials [6][Pb]: sum|epsp|, max|epsp|, sum|rotp|, max|rotp|, #plastic
// beams, #hinges; node partials [2][Pn]: min y (negated), max y
__global__ __launch_bounds__(kBlock) void k_stats_beams(const Blk* __restrict__ tb, int B,
                                                        const double* __restrict__ plastic, double* __restrict__ part) {
  const Blk bk = tb[blockIdx.x];
  const int i = bk.first + threadIdx.x, P = gridDim.x;
  double v[6] = {0, 0, 0, 0, 0, 0};
  if (int(threadIdx.x) < bk.count) {
    double pa = fabs(plastic[i]);
    double pr = sqrt(plastic[2 * B + i] * plastic[2 * B + i] + plastic[3 * B + i] * plastic[3 * B + i] +
                     plastic[4 * B + i] * plastic[4 * B + i]);
    v[0] = pa;
    v[1] = pa;
    v[2] = pr;
    v[3] = pr;
    v[4] = pa > 1e-5 ? 1 : 0;
    v[5] = pr > 1e-4 ? 1 : 0;
  }
  block_reduce<6>(v, (1u << 1) | (1u << 3));
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; k++) part[k * P + blockIdx.x] = v[k];
}
__global__ __launch_bounds__(kBlock) void k_stats_nodes(const Blk* __restrict__ tn, int N,
                                                        const double* __restrict__ bpos, double* __restrict__ part) {
  const Blk bk = tn[blockIdx.x];
  const int i = bk.first + threadIdx.x, P = gridDim.x;
  double v[2] = {-INFINITY, -INFINITY};
  if (int(threadIdx.x) < bk.count) {
    v[0] = -bpos[N + i];
    v[1] = bpos[N + i];
  }
  block_reduce<2>(v, 3u);
  if (threadIdx.x == 0)
    for (int k = 0; k < 2; k++) part[k * P + blockIdx.x] = v[k];
}
// one block per instance: out[8 * inst + k] in the order of mmc_state_stats
__global__ __launch_bounds__(kBlock) void k_stats_final(const Inst* __restrict__ I, const double* __restrict__ bpartial,
                                                        int Pb, const double* __restrict__ npartial, int Pn,
                                                        double* __restrict__ out) {
  const Inst& in = I[blockIdx.x];
  double b[6], n[2];
  reduce_columns<6>(bpartial, Pb, in.pb0, in.pb1, b, (1u << 1) | (1u << 3));
  reduce_columns<2>(npartial, Pn, in.pn0, in.pn1, n, 3u);
  if (threadIdx.x == 0) {
    double* o = out + 8 * blockIdx.x;
    for (int k = 0; k < 6; k++) o[k] = b[k];
    o[6] = n[0];
    o[7] = n[1];
  }
}

// z of every instance from its current buffer (after the device loop)
__global__ __launch_bounds__(kBlock) void k_take_current(const Blk* __restrict__ tn, const Ctl* __restrict__ Cs, int N,
                                                         const double* __restrict__ z0, const double* __restrict__ z1,
                                                         double* __restrict__ z) {
  const Blk bk = tn[blockIdx.x];
  if (int(threadIdx.x) >= bk.count) return;
  const int i = bk.first + threadIdx.x;
  const double* src = Cs[bk.inst].par ? z1 : z0;
  for (int d = 0; d < 6; d++) z[size_t(d) * N + i] = src[size_t(d) * N + i];
}
// z = 0 on the nodes of the masked instances
__global__ __launch_bounds__(kBlock) void k_zero_masked(const Blk* __restrict__ tn, const int* __restrict__ mask,
                                                        int N, double* __restrict__ z) {
  const Blk bk = tn[blockIdx.x];
  if (!mask[bk.inst] || int(threadIdx.x) >= bk.count) return;
  const int i = bk.first + threadIdx.x;
  for (int d = 0; d < 6; d++) z[size_t(d) * N + i] = 0.0;
}
// a masked instance's loop state at the start of a solve: iterate z = 0 in
// slot 0, its gradient g (evaluated there) in slot 0
__global__ __launch_bounds__(kBlock) void k_seed_masked(const Blk* __restrict__ tn, const int* __restrict__ mask,
                                                        int N, const double* __restrict__ g,
                                                        double* __restrict__ z0, double* __restrict__ g0) {
  const Blk bk = tn[blockIdx.x];
  if (!mask[bk.inst] || int(threadIdx.x) >= bk.count) return;
  const int i = bk.first + threadIdx.x;
  for (int d = 0; d < 6; d++) {
    z0[size_t(d) * N + i] = 0.0;
    g0[size_t(d) * N + i] = g[size_t(d) * N + i];
  }
}

}  // namespace

// ===========================================================================
// Host side: the context
// ===========================================================================
#define HIPCALL(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      throw HipFail(std::string(#expr) + ": " + hipGetErrorString(e_));                      \
    }                                                                                        \
  } while (0)

namespace {
struct HipFail : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct Fail : std::runtime_error {
  int code;
  Fail(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
std::string g_create_error;

int blocks(long n) { return int((n + kBlock - 1) / kBlock); }

template <class X>
X* dalloc(size_t n) {
  X* p = nullptr;
  if (n == 0) n = 1;
  HIPCALL(hipMalloc(&p, n * sizeof(X)));
  return p;
}
}  // namespace

// Device scalar scratch and pinned host readback sizes (doubles).  The largest
// readback is the pair inner products: 4 * kMaxHist + 5.
constexpr int kScal = 128;
constexpr int kHost = 256;
static_assert(4 * kMaxHist + 5 <= kScal && 128 + 4 * kMaxHist + 5 <= kHost, "scalar buffers too small");

// A context holds n_inst independent lattices (one for the single-lattice
// API) in one concatenated SoA with per-instance constants, block tables and
// control blocks; every launch covers all of them.
// ---------------------------------------------------------------------------
// Static scratch of the allocation-free entry points (mmc_beam_potential_noalloc):
// the block partials of the beam-potential reduction, one double per 256-beam
// block, so up to kStaticScratchDoubles blocks (16.7 M beams) per call.  Each
// GPU holds its own copy for the life of the process.  Host code never uses the
// array's name as a device address (it is a host-side stand-in): the address of
// the device copy is taken once per device with hipGetSymbolAddress, while that
// device is current, and kept in g_scratch_ptr.  Calls that overlap on one GPU
// share the array and race; a call overwrites every partial it later reads, so
// nothing needs clearing between calls.
// ---------------------------------------------------------------------------
constexpr int kStaticScratchDoubles = 1 << 16;  // 512 KiB
static __device__ __align__(256) double scratch[kStaticScratchDoubles];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
// Contexts may be driven from several host threads at once (a sharded batch).
// A graph is recorded by stream capture, and a capture in progress is
// invalidated by any operation that synchronises with its stream from
// elsewhere in the process: copies and memsets on the legacy default stream
// (which synchronise with every blocking stream) and frees (which
// synchronise the device).  Every capture, and every run-phase section that
// issues such operations (list builds, table uploads, kernel timing, capture
// of frames, lattice readback, set-up and tear-down), holds this lock.
std::recursive_mutex g_sync_mu;
using SyncLock = std::lock_guard<std::recursive_mutex>;
double* g_scratch_ptr[kMaxDevices] = {};

// Takes the device copy's address for `device` (current) once; throws on failure.
void static_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) throw Fail(MMC_ERR_CAPACITY, "device ordinal beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    throw HipFail(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = static_cast<double*>(ptr);
}

struct mmc_ctx {
  std::string err;
  int precision = MMC_FP64;
  int n_inst = 0;
  int N = 0, B = 0, NS = 0;  // totals over the instances
  std::vector<mmc_config> icfg;  // per instance (translation_scale = cell_size)
  std::vector<Inst> hinst;
  mmc_config cfg{};  // instance 0 (single-lattice API)
  hipStream_t stream = nullptr;
  int device = 0;  // the device current at creation; every launch goes there
  // device buffers
  double *bpos = nullptr, *bquat = nullptr, *mass = nullptr;
  uint8_t* lock = nullptr;
  int *ba = nullptr, *bb = nullptr;
  double *L0 = nullptr, *rad = nullptr, *area = nullptr, *inert = nullptr, *e0 = nullptr, *plastic = nullptr;
  // labels of the captured bodies (Node::layer, Beam::layer, Beam::family)
  int *nlayer = nullptr, *blayer = nullptr, *bfam = nullptr;
  // capture: previous committed poses (mmc_capture_mark), the bodies of the
  // last capture and its lattice table
  double *prev_x = nullptr, *prev_q = nullptr;
  mmc_snapshot* snap = nullptr;
  long snap_cap = 0;
  CapInst* capt = nullptr;
  int *inc_off = nullptr, *inc = nullptr, *adj_off = nullptr, *adj = nullptr;
  // node-ordered gradient records of the device loop (EllOut): slot of each
  // beam end, [6][eS] end records, [3][eS] capsule contact forces on the ends
  int* eslot = nullptr;
  double *erec = nullptr, *ecap = nullptr;
  long eS = 0;
  double *x = nullptr;
  void* R = nullptr;  // FP64 planes
  double *rec = nullptr, *trial = nullptr, *cf_sph = nullptr, *cf_cap = nullptr;
  float* box = nullptr;
  unsigned *cnt = nullptr, *start = nullptr;
  int* entries = nullptr;
  int entry_cap = 0;
  unsigned table = 0;
  int* flags = nullptr;  // [0]: list-build bits (4 entry overflow, 8 slab overflow); [4 + b]: instance b's bits
  // instances and block tables
  Inst* inst_d = nullptr;
  Blk *tn = nullptr, *tb = nullptr, *tc = nullptr, *td = nullptr;
  double* tops_d = nullptr;   // platen of every instance (host-driven evaluations)
  double* htops = nullptr;    // pinned staging of tops_d
  int* mask_d = nullptr;      // commit mask
  int* smask_d = nullptr;     // lattices started by batch_start
  std::vector<char> in_flight;  // started by batch_start, not yet finished
  int* hmask = nullptr;       // pinned staging of mask_d
  // Verlet contact list
  double* xref = nullptr;
  unsigned* pcount = nullptr;
  int* plist = nullptr;  // [NS][slab]
  int slab = 8;
  bool list_valid = false;
  int64_t rebuilds = 0;
  void* scan_tmp = nullptr;
  size_t scan_bytes = 0;
  double *spart = nullptr, *cpart = nullptr, *bpart = nullptr, *gpart = nullptr, *vpart = nullptr;
  double *stb = nullptr, *stn = nullptr, *stout = nullptr, *hstat = nullptr;
  int Ps = 0, Pc = 0, Pb = 0, Pd = 0;  // blocks of the node, shape, beam and 64-node tables
  double* record = nullptr;  // [3][n_inst][kRec]: slot 0 current, 1 candidate, 2 final
  double* hrec = nullptr;    // pinned readback of one record slot
  double* scal = nullptr;    // misc scalars (dots)
  double *z = nullptr, *g = nullptr, *p = nullptr, *zc = nullptr, *gc = nullptr, *zt = nullptr, *zin = nullptr;
  double *Sh = nullptr, *Yh = nullptr;  // [m + 1][6N]
  int hist_slots = 1;
  int max_hist = 0;  // longest history over the instances
  // device loop engine: the graph-replayed loop, or (MMC_DEVICE_LOOP=block,
  // lattices of up to kBlockEngineNodes nodes) the workgroup-per-lattice solver
  bool block_engine = false;
  template <class T>
  void solve_blocks(Lat<T> L, const std::vector<int>& set, int max_trials);
  double* host = nullptr;  // pinned readback
  // device-resident minimiser: control blocks (one per instance), their pinned
  // staging copies, the (current, candidate) buffer pairs selected by each
  // instance's Ctl::par, and the graph of kGraphIterations replays
  Ctl* ctl = nullptr;
  Ctl* hctl = nullptr;
  // block tables of the captured device loop: the instances of gset only
  std::vector<Blk> htd, htb, htc, htn;  // host copies of the full tables
  Blk *rtc = nullptr, *rtn = nullptr;    // shape and node tables of a partial list rebuild
  Blk *gtd = nullptr, *gtb = nullptr, *gtc = nullptr;
  GRng* grng = nullptr;  // [n_inst] ranges in those tables
  int* gact = nullptr;   // [n_inst] instances of the decision grid
  int gPd = 0, gPb = 0, gPc = 0;
  std::vector<int> gset;
  int64_t captures = 0;
  // launch shape of the loop (it changes no result): blocks per instance of
  // the decision kernel (0: the decision runs in the gather kernel's last
  // block), chosen by tune() at the first capture over a set of each size
  // class (sets of 1, 2, 3-4, 5-8, ... instances) and reused for later sets
  // of that class; and the timings the current choice was made on
  int dec_split = 8;
  // tuner timings (us): gather + decision at splits 1, 2, 4, 8, 16, then the
  // gather with the decision fused
  double tune_dec_us[6] = {0, 0, 0, 0, 0, 0};
  struct LoopShape {
    bool tuned = false;
    int split = 8;
    double us[6] = {0, 0, 0, 0, 0, 0};
  };
  LoopShape shapes[32];
  template <class T>
  void launch_dstate(Lat<T> Lt, const Ctl* C);
  template <class T>
  void launch_elem(Lat<T> Lt, const Ctl* C);
  template <class T>
  void launch_gather(Lat<T> Lt, Ctl* C);
  void launch_decide(Ctl* C);
  void graph_tables(const std::vector<int>& set);
  void tune();
  // host-clock seconds: initial evaluation, graph captures, replay loop,
  // minimize_graph in all, final evaluation (reported at destruction when
  // MMC_LOOP_TIMING is set)
  // [5]: contact list builds, [6]: of them partial, [7..9]: count + scan,
  // fill + pairs, reference positions + publication
  double loop_s[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  int64_t partial_builds = 0;
  int64_t loop_polls = 0;
  Ctl* tctl = nullptr;  // scratch control blocks of time_loop_kernel
  double *cal0 = nullptr, *cal1 = nullptr;  // buffers of the MMC_K_COPY timing
  double* dspart = nullptr;    // [10][Pd] partials of the trial state kernel
  double* cent = nullptr;      // [n_inst][2] node centroid (x, z) of the trial (k_elem_g)
  // contact chunks of the loop's table with long lists (k_classify)
  int* cflag = nullptr;
  int* hlist = nullptr;
  unsigned* hcount = nullptr;
  int contact_serial = kContactSerial;  // -1: every chunk to the lane groups
  void classify();
  double* drows = nullptr;     // [n_inst][kGDecRows] rows folded by the decision kernel's blocks
  unsigned* dcount = nullptr;  // [n_inst] blocks of the decision kernel done with their rows
  double *zbuf[2] = {nullptr, nullptr}, *gbuf[2] = {nullptr, nullptr};
  // per-item energies (nodes, beams, shapes) at the current iterate and at the
  // candidate, for the per-term energy change of the line search
  double *nbuf[2] = {nullptr, nullptr}, *ubuf[2] = {nullptr, nullptr}, *cbuf[2] = {nullptr, nullptr};
  bool terms_out0 = false;  // the next host-driven evaluation writes slot 0 of those
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  ListRef* lref = nullptr;  // device copy of (plist, slab) for the graph kernels
  int64_t graph_replays = 0;
  bool have_trial = false;
  int evaluations = 0;
  int64_t total_evaluations = 0;
  // kernel timer
  int timed = -1;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int64_t timed_launches = 0;
  double timed_ms = 0;
  bool timer_pending = false;
  void tick(int id) {
    if (id == timed) HIPCALL(hipEventRecord(ev0, stream));
  }
  void tock(int id) {
    if (id == timed) {
      HIPCALL(hipEventRecord(ev1, stream));
      timer_pending = true;
    }
  }
  // called after every stream synchronisation: at most one timed launch lies
  // between two synchronisations
  void harvest() {
    if (!timer_pending) return;
    float ms = 0;
    HIPCALL(hipEventElapsedTime(&ms, ev0, ev1));
    timed_ms += ms;
    timed_launches++;
    timer_pending = false;
  }

  Lat<double> lat64() const {
    return Lat<double>{N, B, bpos, bquat, mass, lock, ba, bb, L0, rad, area, inert, e0, plastic, inc_off, inc, adj_off, adj, eslot};
  }
  Lat<float> lat32() const {
    return Lat<float>{N, B, bpos, bquat, mass, lock, ba, bb, L0, rad, area, inert, e0, plastic, inc_off, inc, adj_off, adj,
                      eslot};
  }
  int* iflags() const { return flags + 4; }
  double* rec_slot(int s) const { return record + size_t(s) * n_inst * kRec; }
  ~mmc_ctx();
  void setup(const mmc_lattice* const* L, const mmc_config* c, int n, int prec);
  void setup_generated(const mmc_config* c, int n, int prec);
  void begin_setup(const mmc_config* c, int n, int prec, const std::vector<int>& nn, const std::vector<int>& nb);
  void upload_lattices(const mmc_lattice* const* L, std::vector<int>& off, std::vector<int>& list);
  void finish_setup(const std::vector<int>& off, const std::vector<int>& list);
  // envelope and mass of every lattice (mmc_lattice_info)
  struct LatInfo {
    double width = 0, height = 0, depth = 0, mass = 0;
  };
  std::vector<LatInfo> hlat;
  void ensure_entries(size_t need);
  void upload_inst() {
    if (inst_d) HIPCALL(hipMemcpy(inst_d, hinst.data(), hinst.size() * sizeof(Inst), hipMemcpyHostToDevice));
  }
  // the contact lists of every instance (tables tc, tn), or of some
  // instances only (their blocks of those tables in ctab, ntab)
  template <class T>
  void build_list(Lat<T> L, const Blk* ctab = nullptr, int cblocks = 0, const Blk* ntab = nullptr, int nblocks = 0);
  void rebuild() {
    if (precision == MMC_MIXED)
      build_list<float>(lat32());
    else
      build_list<double>(lat64());
  }
  void rebuild_held(const std::vector<int>& held);
  template <class T>
  void capture_iteration(Lat<T> L);
  void capture(const std::vector<int>& set);
  double time_loop_kernel(int kernel, int launches);
  void minimize_graph(const double* tops, const int32_t* active, mmc_opt_stats* st);
  void init_ctl(int b, double top, bool run);
  // per-lattice progression of a batch: each lattice starts a solve when the
  // caller says so and is finished when its solve ends, the others go on
  std::vector<double> atop;  // every instance's current platen target
  void batch_iterates(const int32_t* zmask);
  void batch_start(const double* tops, const int32_t* mask);
  void batch_advance(int min_done, int32_t* running);
  void batch_finish(const int32_t* mask, mmc_opt_stats* st, mmc_objective_out* fin);
  void minimize_host(double top, mmc_opt_stats& st);
  void final_evaluation(const double* tops, mmc_opt_stats* st, mmc_objective_out* fin);
  void evaluate(const double* zsrc, const double* psrc, double step, double* zdst, double* gdst, const double* tops,
                bool trials, double* rec_out);
  template <class T>
  void evaluate_t(Lat<T> L, const double* zsrc, const double* psrc, double step, double* zdst, double* gdst,
                  bool trials, double* rec_out);
  void set_tops(const double* tops) {
    for (int b = 0; b < n_inst; b++) htops[b] = tops[b];
    HIPCALL(hipMemcpyAsync(tops_d, htops, n_inst * sizeof(double), hipMemcpyHostToDevice, stream));
  }
  void read(void* dst, const void* src, size_t bytes) {
    HIPCALL(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
    HIPCALL(hipStreamSynchronize(stream));
    harvest();
  }
  void zero(double* a, long n) { k_fill_zero<<<blocks(n), kBlock, 0, stream>>>(a, n); }
  void need_single(const char* what) const {
    if (n_inst != 1) throw Fail(MMC_ERR_ARG, std::string(what) + " needs a single-lattice context");
  }
};

mmc_ctx::~mmc_ctx() {
  SyncLock sync_lock(g_sync_mu);
  if (std::getenv("MMC_LOOP_TIMING"))
    std::fprintf(stderr,
                 "[loop] %d instances: initial evaluations %.3f s, captures %lld (%.3f s), replay loop %.3f s "
                 "(%lld graph launches, %lld polls), minimize %.3f s, final evaluations %.3f s, "
                 "%lld contact list builds %.3f s (%lld partial, %.3f s; count + scan %.3f, fill + pairs %.3f, "
                 "rest %.3f); decision %d blocks per lattice\n",
                 n_inst, loop_s[0], (long long)captures, loop_s[1], loop_s[2], (long long)graph_replays,
                 (long long)loop_polls, loop_s[3], loop_s[4], (long long)rebuilds, loop_s[5],
                 (long long)partial_builds, loop_s[6], loop_s[7], loop_s[8], loop_s[9], dec_split);
  void* ptrs[] = {bpos, bquat, mass, lock, ba, bb, L0, rad, area, inert, e0, plastic, inc_off, inc, adj_off, adj, x, R,
                  rec, trial, cf_sph, cf_cap, box, cnt, start, entries, flags, scan_tmp, spart, cpart, bpart, gpart,
                  vpart, stb, stn, stout, record, scal, z, g, p, zc, gc, zt, zin, Sh, Yh, xref, pcount, plist,
                  inst_d, tn, tb, tc, td, tops_d, mask_d, smask_d, ctl, lref, dspart, cent, cflag, hlist, hcount, drows, dcount, gtd, gtb, gtc,
                  grng, gact, tctl, rtc, rtn, eslot, erec, ecap, cal0, cal1, nlayer, blayer, bfam, prev_x, prev_q,
                  snap, capt};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  void* pinned[] = {host, hctl, htops, hmask, hrec, hstat};
  for (void* q : pinned)
    if (q) (void)hipHostFree(q);
  for (int b = 0; b < 2; b++) {
    void* bufs[] = {zbuf[b], gbuf[b], nbuf[b], ubuf[b], cbuf[b]};
    for (void* q : bufs)
      if (q) (void)hipFree(q);
  }
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (stream) (void)hipStreamDestroy(stream);
}

namespace {
// Splits every instance's item range [first, first + count) into blocks of
// `per` items; range(b) gives instance b's items.  Returns the table and, for
// each instance, its block range.
template <class F>
std::vector<Blk> block_table(int n, int per, F range, std::vector<int>& b0, std::vector<int>& b1) {
  std::vector<Blk> t;
  b0.assign(n, 0);
  b1.assign(n, 0);
  for (int b = 0; b < n; b++) {
    int first, count;
    range(b, first, count);
    b0[b] = int(t.size());
    for (int o = 0; o < count; o += per) t.push_back(Blk{b, first + o, std::min(per, count - o), 0});
    b1[b] = int(t.size());
  }
  if (t.empty()) t.push_back(Blk{0, 0, 0, 0});
  return t;
}
}  // namespace

void mmc_ctx::setup(const mmc_lattice* const* Ls, const mmc_config* cs, int n, int prec) {
  SyncLock sync_lock(g_sync_mu);
  std::vector<int> nn(size_t(std::max(n, 0))), nb(size_t(std::max(n, 0)));
  for (int b = 0; b < n; b++) nn[size_t(b)] = Ls[b]->n_nodes, nb[size_t(b)] = Ls[b]->n_beams;
  begin_setup(cs, n, prec, nn, nb);
  std::vector<int> off, list;
  upload_lattices(Ls, off, list);
  finish_setup(off, list);
}

// Lattices built on the device (generate_lattices) straight into the SoA.
void mmc_ctx::setup_generated(const mmc_config* cs, int n, int prec) {
  SyncLock sync_lock(g_sync_mu);
  if (n <= 0) throw Fail(MMC_ERR_ARG, "a context needs at least one lattice");
  std::vector<int> nn(static_cast<size_t>(n)), nb(static_cast<size_t>(n));
  for (int b = 0; b < n; b++) {
    GenCounts k;
    try {
      k = gen_counts(cs[b]);
    } catch (const std::exception& e) {
      throw Fail(MMC_ERR_ARG, e.what());
    }
    nn[size_t(b)] = k.nodes, nb[size_t(b)] = k.beams;
  }
  begin_setup(cs, n, prec, nn, nb);
  std::vector<int> n0(static_cast<size_t>(n)), b0(static_cast<size_t>(n));
  long pairs = 0;
  for (int b = 0; b < n; b++) {
    n0[size_t(b)] = hinst[size_t(b)].n0, b0[size_t(b)] = hinst[size_t(b)].b0;
    pairs += gen_counts(icfg[size_t(b)]).connected;
  }
  bpos = dalloc<double>(3 * size_t(N));
  bquat = dalloc<double>(4 * size_t(N));
  mass = dalloc<double>(N);
  lock = dalloc<uint8_t>(N);
  ba = dalloc<int>(B);
  bb = dalloc<int>(B);
  L0 = dalloc<double>(B);
  rad = dalloc<double>(B);
  area = dalloc<double>(B);
  inert = dalloc<double>(B);
  e0 = dalloc<double>(3 * size_t(B));
  plastic = dalloc<double>(7 * size_t(B));
  HIPCALL(hipMemset(plastic, 0, 7 * size_t(B) * sizeof(double)));
  nlayer = dalloc<int>(N);
  blayer = dalloc<int>(B);
  bfam = dalloc<int>(B);
  inc_off = dalloc<int>(N + 1);
  inc = dalloc<int>(2 * size_t(B));
  adj_off = dalloc<int>(N + 1);
  adj = dalloc<int>(2 * size_t(pairs));
  std::vector<GenResult> res;
  try {
    res = generate_lattices(icfg, n0, b0, N, B,
                            GenOut{bpos, bquat, mass, lock, ba, bb, L0, rad, area, inert, e0, nlayer, blayer, bfam,
                                   inc_off, inc, adj_off, adj},
                            stream);
  } catch (const std::exception& e) {
    throw Fail(MMC_ERR_HIP, e.what());
  }
  for (int b = 0; b < n; b++) {
    Inst& in = hinst[size_t(b)];
    const GenResult& r = res[size_t(b)];
    in.n_free = 6.0 * in.N - (r.lock_a == r.lock_b ? 2 : 3);
    LatInfo& li = hlat[size_t(b)];
    li.mass = r.total_mass;
  }
  // the gather's record slots are laid out on the host from the incidence
  std::vector<int> off(size_t(N) + 1), list(2 * size_t(B));
  HIPCALL(hipMemcpy(off.data(), inc_off, (size_t(N) + 1) * sizeof(int), hipMemcpyDeviceToHost));
  HIPCALL(hipMemcpy(list.data(), inc, 2 * size_t(B) * sizeof(int), hipMemcpyDeviceToHost));
  finish_setup(off, list);
}

// Instances, constants and counts shared by both ways of supplying lattices.
void mmc_ctx::begin_setup(const mmc_config* cs, int n, int prec, const std::vector<int>& nn,
                          const std::vector<int>& nb) {
  if (n <= 0) throw Fail(MMC_ERR_ARG, "a context needs at least one lattice");
  precision = prec;
  n_inst = n;
  icfg.assign(cs, cs + n);
  max_hist = 0;
  long long tn_ = 0, tb_ = 0;
  for (int b = 0; b < n; b++) {
    mmc_config& c = icfg[b];
    c.translation_scale = c.cell_size;  // MC/src/sim.cpp:49
    if (nn[size_t(b)] <= 0 || nb[size_t(b)] <= 0) throw Fail(MMC_ERR_ARG, "lattice has no nodes or beams");
    if (c.history < 0 || c.history > kMaxHist) throw Fail(MMC_ERR_ARG, "history must be in [0, 16]");
    if (c.line_search < 0 || c.line_search > 2) throw Fail(MMC_ERR_ARG, "line_search must be 0, 1 or 2");
    max_hist = std::max(max_hist, c.history);
    tn_ += nn[size_t(b)];
    tb_ += nb[size_t(b)];
  }
  if (tn_ + tb_ > 0x3fffffffLL) throw Fail(MMC_ERR_CAPACITY, "batch exceeds 2^30 nodes + beams");
  cfg = icfg[0];
  N = int(tn_);
  B = int(tb_);
  NS = N + B;
  HIPCALL(hipGetDevice(&device));
  static_scratch_init(device);
  HIPCALL(hipStreamCreate(&stream));
  // ---- instances: ranges and constants ----
  hinst.assign(n, Inst{});
  {
    int n0 = 0, b0 = 0;
    for (int b = 0; b < n; b++) {
      Inst& in = hinst[b];
      const mmc_config& c = icfg[b];
      in.n0 = n0;
      in.N = nn[size_t(b)];
      in.b0 = b0;
      in.B = nb[size_t(b)];
      in.s0 = n0 + b0;  // shapes of instance b: its spheres, then its capsules
      in.NS = in.N + in.B;
      n0 += in.N;
      b0 += in.B;
      const double jr = c.joint_scale * c.max_radius;
      Scalars& S = in.S;
      S.ts = c.translation_scale;
      S.jr = jr;
      S.gravity = c.gravity;
      S.kc = c.contact_stiffness;
      S.margin = c.contact_margin;
      S.centering = c.centering_stiffness;
      S.cell = std::max(2.8 * jr, 0.72 * c.cell_size);  // MC/src/sim.cpp:151
      S.inv_cell = 1.0 / S.cell;
      S.skin = 0.5 * S.cell;
      const mmc_material& m = c.material;
      in.M = Material{m.young, m.young / (2 * (1 + m.nu)), m.yield, m.hardening, m.bend_hardening,
                      m.torsion_hardening, c.shear_correction, m.density};
      in.mt = c.max_translation_step / c.translation_scale;
      in.mr = c.max_rotation_step;
    }
  }
  hlat.assign(size_t(n), LatInfo{});
  for (int b = 0; b < n; b++) {
    const mmc_config& c = icfg[size_t(b)];
    LatInfo& li = hlat[size_t(b)];
    li.width = c.cells_x * c.cell_size;
    li.height = c.cells_y * c.cell_size;
    li.depth = c.cells_z * c.cell_size;
  }
}

// Lattices handed over on the host: staged into SoA planes, incidence and
// exclusion adjacency built here; off and list receive the incidence.
void mmc_ctx::upload_lattices(const mmc_lattice* const* Ls, std::vector<int>& off, std::vector<int>& list) {
  const int n = n_inst;
  for (int b = 0; b < n; b++) {
    LatInfo& li = hlat[size_t(b)];
    li.width = Ls[b]->width, li.height = Ls[b]->height, li.depth = Ls[b]->depth;
    li.mass = 0;
    for (int i = 0; i < Ls[b]->n_nodes; i++) li.mass += Ls[b]->mass[i];
  }
  // ---- host-side SoA staging of the concatenated lattice ----
  std::vector<double> h(7 * size_t(std::max(N, B)));
  auto up = [&](double* dst, const double* src, size_t cnt_) {
    HIPCALL(hipMemcpy(dst, src, cnt_ * sizeof(double), hipMemcpyHostToDevice));
  };
  // AoS field of k components of every instance -> SoA planes of the total
  auto soa = [&](int total, int k, bool nodes, auto get) {
    for (int b = 0; b < n; b++) {
      const Inst& in = hinst[b];
      const int off = nodes ? in.n0 : in.b0, cntb = nodes ? in.N : in.B;
      for (int i = 0; i < cntb; i++)
        for (int d = 0; d < k; d++) h[size_t(d) * total + off + i] = get(b, i, d);
    }
  };
  bpos = dalloc<double>(3 * size_t(N));
  soa(N, 3, true, [&](int b, int i, int d) { return Ls[b]->x[3 * size_t(i) + d]; });
  up(bpos, h.data(), 3 * size_t(N));
  bquat = dalloc<double>(4 * size_t(N));
  soa(N, 4, true, [&](int b, int i, int d) { return Ls[b]->q ? Ls[b]->q[4 * size_t(i) + d] : (d == 0 ? 1.0 : 0.0); });
  up(bquat, h.data(), 4 * size_t(N));
  mass = dalloc<double>(N);
  soa(N, 1, true, [&](int b, int i, int) { return Ls[b]->mass[i]; });
  up(mass, h.data(), N);
  std::vector<uint8_t> lk(N, 0);
  std::vector<int> ha(B), hb(B);
  for (int b = 0; b < n; b++) {
    Inst& in = hinst[b];
    const mmc_lattice* L = Ls[b];
    in.n_free = 0;
    for (int i = 0; i < in.N; i++)
      for (int d = 0; d < 6; d++) {
        if (L->locked && L->locked[6 * size_t(i) + d])
          lk[in.n0 + i] |= uint8_t(1u << d);
        else
          in.n_free += 1;
      }
    for (int k = 0; k < in.B; k++) {
      if (L->beam_a[k] < 0 || L->beam_a[k] >= in.N || L->beam_b[k] < 0 || L->beam_b[k] >= in.N)
        throw Fail(MMC_ERR_ARG, "beam endpoint out of range");
      ha[in.b0 + k] = in.n0 + L->beam_a[k];
      hb[in.b0 + k] = in.n0 + L->beam_b[k];
    }
  }
  lock = dalloc<uint8_t>(N);
  HIPCALL(hipMemcpy(lock, lk.data(), N, hipMemcpyHostToDevice));
  ba = dalloc<int>(B);
  bb = dalloc<int>(B);
  HIPCALL(hipMemcpy(ba, ha.data(), B * sizeof(int), hipMemcpyHostToDevice));
  HIPCALL(hipMemcpy(bb, hb.data(), B * sizeof(int), hipMemcpyHostToDevice));
  L0 = dalloc<double>(B);
  soa(B, 1, false, [&](int b, int i, int) { return Ls[b]->L0[i]; });
  up(L0, h.data(), B);
  rad = dalloc<double>(B);
  soa(B, 1, false, [&](int b, int i, int) { return Ls[b]->radius[i]; });
  up(rad, h.data(), B);
  area = dalloc<double>(B);
  soa(B, 1, false, [&](int b, int i, int) { return Ls[b]->area[i]; });
  up(area, h.data(), B);
  inert = dalloc<double>(B);
  soa(B, 1, false, [&](int b, int i, int) { return Ls[b]->inertia[i]; });
  up(inert, h.data(), B);
  e0 = dalloc<double>(3 * size_t(B));
  soa(B, 3, false, [&](int b, int i, int d) { return Ls[b]->e0[3 * size_t(i) + d]; });
  up(e0, h.data(), 3 * size_t(B));
  plastic = dalloc<double>(7 * size_t(B));
  soa(B, 7, false, [&](int b, int i, int d) { return Ls[b]->plastic ? Ls[b]->plastic[7 * size_t(i) + d] : 0.0; });
  up(plastic, h.data(), 7 * size_t(B));
  {
    std::vector<int> lab(std::max(N, B));
    auto labels = [&](int*& dst, int count, bool nodes, const int32_t* mmc_lattice::*field) {
      for (int b = 0; b < n; b++) {
        const Inst& in = hinst[b];
        const int32_t* src = Ls[b]->*field;
        const int off = nodes ? in.n0 : in.b0, cntb = nodes ? in.N : in.B;
        for (int i = 0; i < cntb; i++) lab[size_t(off + i)] = src ? src[i] : 0;
      }
      dst = dalloc<int>(count);
      HIPCALL(hipMemcpy(dst, lab.data(), size_t(count) * sizeof(int), hipMemcpyHostToDevice));
    };
    labels(nlayer, N, true, &mmc_lattice::node_layer);
    labels(blayer, B, false, &mmc_lattice::beam_layer);
    labels(bfam, B, false, &mmc_lattice::beam_family);
  }
  // ---- CSR beam incidence, ordered by beam index ----
  off.assign(size_t(N) + 1, 0);
  list.assign(2 * size_t(B), 0);
  for (int k = 0; k < B; k++) {
    off[ha[k] + 1]++;
    off[hb[k] + 1]++;
  }
  for (int i = 0; i < N; i++) off[i + 1] += off[i];
  {
    std::vector<int> cur(off.begin(), off.end() - 1);
    for (int k = 0; k < B; k++) {
      list[cur[ha[k]]++] = 2 * k;
      list[cur[hb[k]]++] = 2 * k + 1;
    }
  }
  inc_off = dalloc<int>(N + 1);
  inc = dalloc<int>(2 * size_t(B));
  HIPCALL(hipMemcpy(inc_off, off.data(), (N + 1) * sizeof(int), hipMemcpyHostToDevice));
  HIPCALL(hipMemcpy(inc, list.data(), 2 * size_t(B) * sizeof(int), hipMemcpyHostToDevice));
  // ---- exclusion adjacency from the connected pair sets ----
  std::vector<std::vector<int>> nb(N);
  for (int b = 0; b < n; b++) {
    const mmc_lattice* L = Ls[b];
    const int o = hinst[b].n0, Nb = hinst[b].N;
    for (int64_t c2 = 0; c2 < L->n_connected; c2++) {
      uint64_t key = L->connected[c2];
      int u = int(uint32_t(key >> 32)), v = int(uint32_t(key & 0xffffffffu));
      if (u < 0 || u >= Nb || v < 0 || v >= Nb) throw Fail(MMC_ERR_ARG, "connected pair out of range");
      nb[o + u].push_back(o + v);
      nb[o + v].push_back(o + u);
    }
  }
  std::vector<int> aoff(N + 1, 0), alist;
  for (int i = 0; i < N; i++) {
    std::sort(nb[i].begin(), nb[i].end());
    nb[i].erase(std::unique(nb[i].begin(), nb[i].end()), nb[i].end());
    aoff[i + 1] = aoff[i] + int(nb[i].size());
    alist.insert(alist.end(), nb[i].begin(), nb[i].end());
  }
  adj_off = dalloc<int>(N + 1);
  adj = dalloc<int>(alist.size());
  HIPCALL(hipMemcpy(adj_off, aoff.data(), (N + 1) * sizeof(int), hipMemcpyHostToDevice));
  if (!alist.empty()) HIPCALL(hipMemcpy(adj, alist.data(), alist.size() * sizeof(int), hipMemcpyHostToDevice));
}

// Block tables, record slots and work buffers (off, list: the beam incidence).
void mmc_ctx::finish_setup(const std::vector<int>& off, const std::vector<int>& list) {
  const int n = n_inst;
  // ---- block tables ----
  {
    std::vector<int> r0, r1;
    auto put = [&](Blk*& dst, const std::vector<Blk>& t) {
      dst = dalloc<Blk>(t.size());
      HIPCALL(hipMemcpy(dst, t.data(), t.size() * sizeof(Blk), hipMemcpyHostToDevice));
    };
    auto t = block_table(n, kBlock, [&](int b, int& f, int& c) { f = hinst[b].n0, c = hinst[b].N; }, r0, r1);
    Ps = int(t.size());
    for (int b = 0; b < n; b++) hinst[b].pn0 = r0[b], hinst[b].pn1 = r1[b];
    put(tn, t);
    htn = t;
    rtn = dalloc<Blk>(t.size());
    t = block_table(n, kBlock, [&](int b, int& f, int& c) { f = hinst[b].b0, c = hinst[b].B; }, r0, r1);
    Pb = int(t.size());
    for (int b = 0; b < n; b++) hinst[b].pb0 = r0[b], hinst[b].pb1 = r1[b];
    put(tb, t);
    htb = t;
    t = block_table(n, kBlock, [&](int b, int& f, int& c) { f = hinst[b].s0, c = hinst[b].NS; }, r0, r1);
    Pc = int(t.size());
    for (int b = 0; b < n; b++) hinst[b].pc0 = r0[b], hinst[b].pc1 = r1[b];
    put(tc, t);
    htc = t;
    rtc = dalloc<Blk>(t.size());
    t = block_table(n, kDirNodes, [&](int b, int& f, int& c) { f = hinst[b].n0, c = hinst[b].N; }, r0, r1);
    Pd = int(t.size());
    for (int b = 0; b < n; b++) hinst[b].pd0 = r0[b], hinst[b].pd1 = r1[b];
    // node-ordered record slots of the loop's gather blocks (64 nodes each):
    // block j's slots start at base_j, slot u of lane l at base_j + 64 u + l;
    // the block's largest degree and base ride in its table entry (pad)
    {
      std::vector<int> es(2 * size_t(B), 0);
      long base = 0;
      for (Blk& bk : t) {
        int md = 0;
        for (int l = 0; l < bk.count; l++) md = std::max(md, off[bk.first + l + 1] - off[bk.first + l]);
        if (md > 31) throw Fail(MMC_ERR_CAPACITY, "a node has more than 31 incident beams");
        if (base + long(md) * kWave >= (1L << 26)) throw Fail(MMC_ERR_CAPACITY, "too many beam ends for the gather");
        bk.pad = int(base << 5) | md;
        for (int l = 0; l < bk.count; l++) {
          const int i = bk.first + l;
          for (int u = 0; u < off[i + 1] - off[i]; u++) es[size_t(list[size_t(off[i] + u)])] = int(base + long(u) * kWave + l);
        }
        base += long(md) * kWave;
      }
      eS = std::max(base, 1L);
      eslot = dalloc<int>(2 * size_t(B));
      HIPCALL(hipMemcpy(eslot, es.data(), 2 * size_t(B) * sizeof(int), hipMemcpyHostToDevice));
      erec = dalloc<double>(6 * size_t(eS));
      ecap = dalloc<double>(3 * size_t(eS));
      HIPCALL(hipMemset(erec, 0, 6 * size_t(eS) * sizeof(double)));
      HIPCALL(hipMemset(ecap, 0, 3 * size_t(eS) * sizeof(double)));
    }
    put(td, t);
    htd = t;
    gtd = dalloc<Blk>(htd.size());
    gtb = dalloc<Blk>(htb.size() * kBW);  // the loop's element blocks may hold 64 items
    gtc = dalloc<Blk>(htc.size() * kBW);
    grng = dalloc<GRng>(n);
    gact = dalloc<int>(n);
  }
  // ---- evaluation buffers ----
  x = dalloc<double>(3 * size_t(N));
  R = dalloc<double>(4 * size_t(N));
  rec = dalloc<double>(9 * size_t(B));
  trial = dalloc<double>(size_t(MMC_TRIAL_STRIDE) * B);
  cf_sph = dalloc<double>(3 * size_t(N));
  cf_cap = dalloc<double>(6 * size_t(B));
  box = dalloc<float>(8 * size_t(NS));
  table = 1024;
  while (table < 2u * unsigned(NS)) table <<= 1;
  cnt = dalloc<unsigned>(table + 1);
  start = dalloc<unsigned>(table + 1);
  HIPCALL(hipMemset(cnt, 0, (table + 1) * sizeof(unsigned)));
  flags = dalloc<int>(4 + size_t(n));
  HIPCALL(hipMemset(flags, 0, (4 + size_t(n)) * sizeof(int)));
  xref = dalloc<double>(3 * size_t(N));
  pcount = dalloc<unsigned>(size_t(NS));
  HIPCALL(hipMemset(pcount, 0, size_t(NS) * sizeof(unsigned)));
  HIPCALL(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, cnt, start, table + 1, stream));
  scan_tmp = dalloc<char>(scan_bytes);
  for (Inst& in : hinst) in.S.table_mask = table - 1;
  ensure_entries(size_t(NS) * 16);
  slab = 8;
  plist = dalloc<int>(size_t(NS) * slab);
  spart = dalloc<double>(9 * size_t(Ps));
  cpart = dalloc<double>(4 * size_t(kBW) * size_t(Pc));  // one column per wavefront
  bpart = dalloc<double>(4 * size_t(kBW) * size_t(Pb));
  gpart = dalloc<double>(size_t(kGRows) * size_t(Pd));  // gradient rows, then pair rows
  vpart = dalloc<double>((4 * size_t(kMaxHist) + 8) * size_t(std::max(Ps, Pd)));
  stb = dalloc<double>(6 * size_t(Pb));
  stn = dalloc<double>(2 * size_t(Ps));
  stout = dalloc<double>(8 * size_t(n));
  record = dalloc<double>(3 * size_t(n) * kRec);
  scal = dalloc<double>(kScal);
  size_t n6 = 6 * size_t(N);
  z = dalloc<double>(n6);
  g = dalloc<double>(n6);
  p = dalloc<double>(n6);
  zc = dalloc<double>(n6);
  gc = dalloc<double>(n6);
  zt = dalloc<double>(n6);
  zin = dalloc<double>(n6);
  hist_slots = max_hist + 1;  // the device loop keeps one spare slot
  {
    int max_n = 0;
    for (const Inst& in : hinst) max_n = std::max(max_n, in.N);
    // the graph-replayed loop unless MMC_DEVICE_LOOP=block selects the block
    // solver (only for lattices it takes in one pass over the nodes)
    const char* eng = std::getenv("MMC_DEVICE_LOOP");
    block_engine = eng && std::strcmp(eng, "block") == 0 && max_n <= kBlockEngineNodes;
  }
  Sh = dalloc<double>(n6 * hist_slots);
  Yh = dalloc<double>(n6 * hist_slots);
  HIPCALL(hipHostMalloc((void**)&host, kHost * sizeof(double), hipHostMallocDefault));
  HIPCALL(hipHostMalloc((void**)&hctl, size_t(n) * sizeof(Ctl), hipHostMallocDefault));
  HIPCALL(hipHostMalloc((void**)&htops, size_t(n) * sizeof(double), hipHostMallocDefault));
  HIPCALL(hipHostMalloc((void**)&hmask, size_t(n) * sizeof(int), hipHostMallocDefault));
  HIPCALL(hipHostMalloc((void**)&hrec, size_t(n) * kRec * sizeof(double), hipHostMallocDefault));
  HIPCALL(hipHostMalloc((void**)&hstat, size_t(n) * 8 * sizeof(double), hipHostMallocDefault));
  ctl = dalloc<Ctl>(n);
  dspart = dalloc<double>(10 * size_t(Pd));
  cent = dalloc<double>(2 * size_t(n));
  cflag = dalloc<int>(htc.size() * kBW);
  hlist = dalloc<int>(htc.size() * kBW);
  hcount = dalloc<unsigned>(1);
  if (const char* cs = std::getenv("MMC_CONTACT_SERIAL")) contact_serial = std::atoi(cs);
  HIPCALL(hipMemset(cflag, 0, htc.size() * kBW * sizeof(int)));
  HIPCALL(hipMemset(hcount, 0, sizeof(unsigned)));
  HIPCALL(hipMemset(cent, 0, 2 * size_t(n) * sizeof(double)));
  drows = dalloc<double>(size_t(n) * kGDecRows);
  dcount = dalloc<unsigned>(n);
  HIPCALL(hipMemset(dcount, 0, size_t(n) * sizeof(unsigned)));
  lref = dalloc<ListRef>(1);
  tops_d = dalloc<double>(n);
  mask_d = dalloc<int>(n);
  smask_d = dalloc<int>(n);
  inst_d = dalloc<Inst>(n);
  upload_inst();  // the instances are complete here (table mask, entry capacity)
  for (int b = 0; b < 2; b++) {
    zbuf[b] = dalloc<double>(n6);
    gbuf[b] = dalloc<double>(n6);
    nbuf[b] = dalloc<double>(size_t(N));
    ubuf[b] = dalloc<double>(size_t(B));
    cbuf[b] = dalloc<double>(size_t(NS));
  }
  HIPCALL(hipStreamSynchronize(stream));
}

void mmc_ctx::ensure_entries(size_t need) {
  if (entries && size_t(entry_cap) >= need) return;
  if (entries) HIPCALL(hipFree(entries));
  size_t cap = std::max<size_t>(need, 1024);
  if (cap > size_t(0x7fffffff)) throw Fail(MMC_ERR_CAPACITY, "contact grid entry capacity exceeds 2^31");
  entries = dalloc<int>(cap);
  entry_cap = int(cap);
  for (Inst& in : hinst) in.S.entry_cap = entry_cap;
  upload_inst();
}

// Contact list rebuild at the current positions x of every instance (rare:
// only when some node has moved skin/2 since the last build).  Two host
// synchronisations size the grid entries and the pair list exactly.
template <class T>
void mmc_ctx::build_list(Lat<T> Lt, const Blk* ctab, int cblocks, const Blk* ntab, int nblocks) {
  SyncLock sync_lock(g_sync_mu);
  const auto t_build = std::chrono::steady_clock::now();
  const bool part = ctab != nullptr;
  if (!part) {
    ctab = tc;
    cblocks = Pc;
  }
  k_count<T><<<cblocks, kBlock, 0, stream>>>(Lt, inst_d, ctab, x, box, cnt, iflags());
  HIPCALL(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, cnt, start, table + 1, stream));
  unsigned total = 0;
  read(&total, start + table, sizeof(unsigned));
  const auto t_scan = std::chrono::steady_clock::now();
  loop_s[7] += std::chrono::duration<double>(t_scan - t_build).count();
  ensure_entries(size_t(total) + 1024);
  tick(MMC_K_FILL);
  k_fill<<<cblocks, kBlock, 0, stream>>>(inst_d, ctab, box, cnt, start, entries, flags);
  tock(MMC_K_FILL);
  k_pairs<T><<<cblocks * kPairSub, kBlock, 0, stream>>>(Lt, inst_d, ctab, slab, x, box, start, entries,
                                                        pcount, plist, flags);
  int fl = 0;
  read(&fl, flags, sizeof(int));
  const auto t_pairs = std::chrono::steady_clock::now();
  loop_s[8] += std::chrono::duration<double>(t_pairs - t_scan).count();
  if (fl & 8) {
    // some shape has more partners than slots: widen the slab and enumerate
    // every instance again (the wider slab holds no list yet)
    int cleared = fl & ~8;
    HIPCALL(hipMemcpy(flags, &cleared, sizeof(int), hipMemcpyHostToDevice));
    HIPCALL(hipFree(plist));
    plist = nullptr;
    if (slab >= (1 << 16)) throw Fail(MMC_ERR_CAPACITY, "contact partner slab exceeds 65536 per shape");
    slab *= 2;
    plist = dalloc<int>(size_t(NS) * slab);
    build_list<T>(Lt);
    return;
  }
  if (part)
    k_copy_xref<<<nblocks, kBlock, 0, stream>>>(ntab, N, x, xref);
  else
    HIPCALL(hipMemcpyAsync(xref, x, 3 * size_t(N) * sizeof(double), hipMemcpyDeviceToDevice, stream));
  HIPCALL(hipGetLastError());
  // publish the list location for the graph kernels (stream-ordered before
  // any later replay; staged through pinned memory)
  ListRef* hl = reinterpret_cast<ListRef*>(host + kHost - 4);
  hl->plist = plist;
  hl->slab = slab;
  hl->pad = 0;
  HIPCALL(hipMemcpyAsync(lref, hl, sizeof(ListRef), hipMemcpyHostToDevice, stream));
  classify();
  HIPCALL(hipStreamSynchronize(stream));
  list_valid = true;
  rebuilds++;
  const auto t_end = std::chrono::steady_clock::now();
  const double dt = std::chrono::duration<double>(t_end - t_build).count();
  loop_s[5] += dt;
  loop_s[9] += std::chrono::duration<double>(t_end - t_pairs).count();
  if (part) {
    loop_s[6] += dt;
    partial_builds++;
  }
}

// One objective evaluation of every instance at its platen (tops_d), records
// into rec_out[inst * kRec] and the pinned hrec.
template <class T>
void mmc_ctx::evaluate_t(Lat<T> Lt, const double* zsrc, const double* psrc, double step, double* zdst, double* gdst,
                         bool trials, double* rec_out) {
  tick(MMC_K_STATE);
  const bool terms = terms_out0;
  terms_out0 = false;
  k_state<T><<<Ps, kBlock, 0, stream>>>(Lt, inst_d, tn, tops_d, zsrc, psrc, step, psrc ? zdst : nullptr, x,
                                        (double*)R, list_valid ? xref : nullptr,
                                        TermBufs{nullptr, terms ? nbuf[0] : nullptr}, spart);
  tock(MMC_K_STATE);
  bool built = false;
  if (!list_valid) {
    build_list<T>(Lt);
    built = true;
  }
  tick(MMC_K_BEAM);
  k_beam<T><<<Pb, kBlock, 0, stream>>>(Lt, inst_d, tb, x, (const double*)R, rec, trials ? trial : nullptr,
                                       TermBufs{nullptr, terms ? ubuf[0] : nullptr}, bpart);
  tock(MMC_K_BEAM);
  for (;;) {
    tick(MMC_K_CONTACT);
    k_contact<T><<<Pc, kBlock, 0, stream>>>(Lt, inst_d, tc, x, pcount, plist, slab, cf_sph, cf_cap,
                                            TermBufs{nullptr, terms ? cbuf[0] : nullptr}, cpart);
    tock(MMC_K_CONTACT);
    tick(MMC_K_GATHER);
    k_gather<T><<<Pd, kElemThreads, 0, stream>>>(Lt, inst_d, td, tops_d, x, rec, cf_sph, cf_cap, spart, Ps, gdst,
                                                 nullptr, gpart);
    tock(MMC_K_GATHER);
    tick(MMC_K_FINAL);
    k_final<<<n_inst, kBlock, 0, stream>>>(inst_d, spart, Ps, cpart, Pc, bpart, Pb, gpart, Pd, iflags(), nullptr,
                                           rec_out);
    tock(MMC_K_FINAL);
    HIPCALL(hipGetLastError());
    HIPCALL(hipMemcpyAsync(hrec, rec_out, size_t(n_inst) * kRec * sizeof(double), hipMemcpyDeviceToHost, stream));
    HIPCALL(hipStreamSynchronize(stream));
    harvest();
    // a stale list (some node moved skin/2 or more since the build) cannot be
    // trusted: rebuild at these positions (the stale instances' lists only)
    // and redo the contact part
    std::vector<int> stale;
    for (int b = 0; b < n_inst; b++)
      if (hrec[b * kRec + 11] >= 0.5 * hinst[b].S.skin) stale.push_back(b);
    if (built || stale.empty()) break;
    rebuild_held(stale);
    built = true;
  }
  evaluations++;
  total_evaluations++;
}

void mmc_ctx::evaluate(const double* zsrc, const double* psrc, double step, double* zdst, double* gdst,
                       const double* tops, bool trials, double* rec_out) {
  set_tops(tops);
  if (precision == MMC_MIXED)
    evaluate_t<float>(lat32(), zsrc, psrc, step, zdst, gdst, trials, rec_out);
  else
    evaluate_t<double>(lat64(), zsrc, psrc, step, zdst, gdst, trials, rec_out);
}

namespace {
void fill_out(const double* r, mmc_objective_out* o) {
  if (!o) return;
  o->energy = r[0];
  o->elastic = r[1];
  o->plastic = r[2];
  o->contact = r[3];
  o->reaction = r[4];
  o->max_pen = r[5];
  o->rms = r[6];
  o->max = r[7];
  o->contacts = int32_t(r[8]);
  o->finite = int32_t(r[9]);
}
template <class F>
int guarded(mmc_ctx* c, F&& f) {
  if (!c) return MMC_ERR_ARG;
  try {
    f();
    return MMC_OK;
  } catch (const Fail& e) {
    c->err = e.what();
    return e.code;
  } catch (const HipFail& e) {
    c->err = e.what();
    return MMC_ERR_HIP;
  } catch (const std::exception& e) {
    c->err = e.what();
    return MMC_ERR_STATE;
  }
}
void aos_from_soa(const double* soa, double* aos, int n, int k) {
  for (int i = 0; i < n; i++)
    for (int d = 0; d < k; d++) aos[size_t(k) * i + d] = soa[size_t(d) * n + i];
}
void soa_from_aos(const double* aos, double* soa, int n, int k) {
  for (int i = 0; i < n; i++)
    for (int d = 0; d < k; d++) soa[size_t(d) * n + i] = aos[size_t(k) * i + d];
}
int create_ctx(const mmc_lattice* const* L, const mmc_config* cfgs, int32_t n, int32_t precision, mmc_ctx** out) {
  if (!L || !cfgs || !out || n <= 0 || (precision != MMC_FP64 && precision != MMC_MIXED)) {
    g_create_error = "mmc_create: invalid argument";
    return MMC_ERR_ARG;
  }
  for (int b = 0; b < n; b++)
    if (!L[b]) {
      g_create_error = "mmc_create: NULL lattice";
      return MMC_ERR_ARG;
    }
  *out = nullptr;
  mmc_ctx* c = new mmc_ctx();
  try {
    c->setup(L, cfgs, n, precision);
  } catch (const Fail& e) {
    g_create_error = e.what();
    delete c;
    return e.code;
  } catch (const std::exception& e) {
    g_create_error = e.what();
    delete c;
    return MMC_ERR_HIP;
  }
  *out = c;
  return MMC_OK;
}
}  // namespace

extern "C" {

void mmc_config_defaults(mmc_config* c) {
  std::memset(c, 0, sizeof(*c));
  c->cells_x = 4;
  c->cells_y = 6;
  c->cells_z = 4;
  c->elements_per_strut = 3;
  c->cell_size = .010;
  c->min_radius = .00082;
  c->max_radius = .00122;
  c->joint_scale = 1.18;
  c->trigger_scale = .38;
  c->imperfection = .00022;
  c->member_imperfection = .00022;
  c->material = mmc_material{1180, 26e6, .485, 5.2e6, 1.4e6, .06, .045};
  c->gravity = 9.81;
  c->shear_correction = .62;
  c->contact_stiffness = 1.2e6;
  c->contact_margin = 2e-5;
  c->centering_stiffness = 1.5;
  c->contact_samples = 3;
  c->load_steps = 52;
  c->unload_steps = 30;
  c->max_iterations = 130;
  c->history = 10;
  c->max_line_search = 22;
  c->max_adaptive_refinements = 7;
  c->fps = 24;
  c->frames = 83;
  c->threads = 0;
  c->max_strain = .62;
  c->step_time = .06;
  c->translation_scale = .010;
  c->max_translation_step = .0022;
  c->max_rotation_step = .30;
  c->gradient_rms_tolerance = 1.2e-3;
  c->gradient_max_tolerance = 1.5e-2;
  c->armijo = 1e-4;
  c->min_line_step = 1e-9;
  c->penetration_tolerance = .00012;
  c->seed = 0x494e445553545259ULL;
  c->lattice = MMC_LATTICE_BCC;
  c->relative_escape = 1;
  c->relative_tolerance = 1.5e-3;
}

int mmc_create(const mmc_lattice* L, const mmc_config* cfg, int32_t precision, mmc_ctx** out) {
  return create_ctx(&L, cfg, 1, precision, out);
}

int mmc_create_batch(const mmc_lattice* lattices, const mmc_config* cfgs, int32_t n, int32_t precision,
                     mmc_ctx** out) {
  if (!lattices || n <= 0) {
    g_create_error = "mmc_create_batch: invalid argument";
    return MMC_ERR_ARG;
  }
  std::vector<const mmc_lattice*> L(n);
  for (int b = 0; b < n; b++) L[b] = lattices + b;
  return create_ctx(L.data(), cfgs, n, precision, out);
}

void mmc_destroy(mmc_ctx* c) { delete c; }

const char* mmc_last_error(const mmc_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int mmc_instances(const mmc_ctx* c) { return c ? c->n_inst : 0; }

int mmc_create_generated(const mmc_config* cfgs, int32_t n, int32_t precision, mmc_ctx** out) {
  if (!cfgs || n <= 0 || !out) {
    g_create_error = "NULL or empty argument";
    return MMC_ERR_ARG;
  }
  *out = nullptr;
  if (precision != MMC_FP64 && precision != MMC_MIXED) {
    g_create_error = "unknown precision policy";
    return MMC_ERR_ARG;
  }
  auto* c = new mmc_ctx();
  try {
    c->setup_generated(cfgs, n, precision);
  } catch (const Fail& e) {
    g_create_error = e.what();
    delete c;
    return e.code;
  } catch (const std::exception& e) {
    g_create_error = e.what();
    delete c;
    return MMC_ERR_HIP;
  }
  *out = c;
  return MMC_OK;
}

int mmc_lattice_info_get(mmc_ctx* c, int32_t b, mmc_lattice_info* o) {
  return guarded(c, [&] {
    SyncLock sync_lock(g_sync_mu);
    if (!o || b < 0 || b >= c->n_inst) throw Fail(MMC_ERR_ARG, "lattice index out of range");
    const Inst& in = c->hinst[size_t(b)];
    const mmc_ctx::LatInfo& li = c->hlat[size_t(b)];
    int excl = 0;
    HIPCALL(hipMemcpy(&excl, c->adj_off + in.n0 + in.N, sizeof(int), hipMemcpyDeviceToHost));
    int first = 0;
    HIPCALL(hipMemcpy(&first, c->adj_off + in.n0, sizeof(int), hipMemcpyDeviceToHost));
    *o = mmc_lattice_info{in.N, in.B, int64_t(excl - first) / 2, li.width, li.height, li.depth, li.mass};
  });
}

int mmc_read_lattice(mmc_ctx* c, int32_t b, double* x, double* mass, uint8_t* locked, int32_t* beam_a,
                     int32_t* beam_b, double* L0, double* radius, double* area, double* inertia, double* e0,
                     int32_t* node_layer, int32_t* beam_layer, int32_t* beam_family, int32_t* adj_off,
                     int32_t* adj) {
  return guarded(c, [&] {
    SyncLock sync_lock(g_sync_mu);
    if (b < 0 || b >= c->n_inst) throw Fail(MMC_ERR_ARG, "lattice index out of range");
    const Inst& in = c->hinst[size_t(b)];
    const int N = in.N, B = in.B;
    // planes of this lattice's range, node-major on the host
    auto planes = [&](double* dst, const double* src, int total, int first, int cnt, int k) {
      if (!dst) return;
      std::vector<double> t(size_t(cnt) * k);
      for (int d = 0; d < k; d++)
        HIPCALL(hipMemcpy(t.data() + size_t(d) * cnt, src + size_t(d) * total + first, size_t(cnt) * sizeof(double),
                          hipMemcpyDeviceToHost));
      aos_from_soa(t.data(), dst, cnt, k);
    };
    auto ints = [&](int32_t* dst, const int* src, int first, int cnt, int shift) {
      if (!dst) return;
      HIPCALL(hipMemcpy(dst, src + first, size_t(cnt) * sizeof(int), hipMemcpyDeviceToHost));
      for (int i = 0; i < cnt; i++) dst[i] -= shift;
    };
    planes(x, c->bpos, c->N, in.n0, N, 3);
    planes(mass, c->mass, c->N, in.n0, N, 1);
    if (locked) {
      std::vector<uint8_t> lk(static_cast<size_t>(N));
      HIPCALL(hipMemcpy(lk.data(), c->lock + in.n0, size_t(N), hipMemcpyDeviceToHost));
      for (int i = 0; i < N; i++)
        for (int d = 0; d < 6; d++) locked[6 * size_t(i) + d] = uint8_t((lk[size_t(i)] >> d) & 1u);
    }
    ints(beam_a, c->ba, in.b0, B, in.n0);
    ints(beam_b, c->bb, in.b0, B, in.n0);
    planes(L0, c->L0, c->B, in.b0, B, 1);
    planes(radius, c->rad, c->B, in.b0, B, 1);
    planes(area, c->area, c->B, in.b0, B, 1);
    planes(inertia, c->inert, c->B, in.b0, B, 1);
    planes(e0, c->e0, c->B, in.b0, B, 3);
    ints(node_layer, c->nlayer, in.n0, N, 0);
    ints(beam_layer, c->blayer, in.b0, B, 0);
    ints(beam_family, c->bfam, in.b0, B, 0);
    std::vector<int> ao(size_t(N) + 1);
    HIPCALL(hipMemcpy(ao.data(), c->adj_off + in.n0, (size_t(N) + 1) * sizeof(int), hipMemcpyDeviceToHost));
    if (adj_off)
      for (int i = 0; i <= N; i++) adj_off[i] = ao[size_t(i)] - ao[0];
    if (adj && ao[size_t(N)] > ao[0]) ints(adj, c->adj, ao[0], ao[size_t(N)] - ao[0], in.n0);
  });
}

int mmc_objective(mmc_ctx* c, const double* zh, double top, int32_t trials, mmc_objective_out* out, double* grad) {
  return guarded(c, [&] {
    c->need_single("mmc_objective");
    if (!zh) throw Fail(MMC_ERR_ARG, "z is NULL");
    const int N = c->N;
    std::vector<double> s(6 * size_t(N));
    soa_from_aos(zh, s.data(), N, 6);
    HIPCALL(hipMemcpyAsync(c->zin, s.data(), s.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    c->evaluate(c->zin, nullptr, 0.0, nullptr, c->gc, &top, trials != 0, c->rec_slot(2));
    fill_out(c->hrec, out);
    const bool finite = c->hrec[9] != 0.0;
    if (trials) {
      HIPCALL(hipMemcpyAsync(c->zt, c->zin, s.size() * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
      c->have_trial = true;
    }
    if (grad) {
      c->read(s.data(), c->gc, s.size() * sizeof(double));
      if (!finite)
        std::fill(grad, grad + s.size(), 0.0);
      else
        aos_from_soa(s.data(), grad, N, 6);
    }
    HIPCALL(hipStreamSynchronize(c->stream));
  });
}

}  // extern "C"

namespace {
bool host_converged(const mmc_config& cf, const double* r) {
  double rms = r[6], mx = r[7];
  double rel = rms / std::max(1e-6, cf.translation_scale * std::max(1.0, r[4]));
  bool ok = (rms <= cf.gradient_rms_tolerance && mx <= cf.gradient_max_tolerance) ||
            (cf.relative_escape && rel < cf.relative_tolerance);
  return ok && r[5] <= cf.penetration_tolerance;
}
bool use_host_loop() {
  const char* e = std::getenv("MMC_HOST_LOOP");
  return e && e[0] == '1';
}
}  // namespace

// kGraphIterations replays of the phase machine as one graph (see Ctl); every
// launch covers all instances, each block gated on its own instance's phase.
constexpr int kGraphIterations = 8;
// graph launches between two reads of the status heads: trials grow from 8
// to 32 per read as a solve goes on
constexpr int kPollLaunches0 = 8 / kGraphIterations > 0 ? 8 / kGraphIterations : 1;
constexpr int kPollLaunchesMax = 32 / kGraphIterations > 0 ? 32 / kGraphIterations : 1;
template <class T>
void mmc_ctx::capture_iteration(Lat<T> Lt) {
  if (graph_exec) HIPCALL(hipGraphExecDestroy(graph_exec));
  if (graph) HIPCALL(hipGraphDestroy(graph));
  graph_exec = nullptr;
  graph = nullptr;
  HIPCALL(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  for (int u = 0; u < kGraphIterations; u++) {
    launch_dstate<T>(Lt, ctl);
    launch_elem<T>(Lt, ctl);
    launch_gather<T>(Lt, ctl);
    if (dec_split != 0) launch_decide(ctl);
  }
  HIPCALL(hipStreamEndCapture(stream, &graph));
  HIPCALL(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
}

// The contact lists of the instances of `held` only (their trial positions are
// in x): the other instances' lists and reference positions are left as they
// are, and stay exact for them.
void mmc_ctx::rebuild_held(const std::vector<int>& held) {
  SyncLock sync_lock(g_sync_mu);
  if (held.size() == size_t(n_inst)) {
    rebuild();
    return;
  }
  std::vector<Blk> c, nb;
  for (int i : held) {
    const Inst& in = hinst[size_t(i)];
    c.insert(c.end(), htc.begin() + in.pc0, htc.begin() + in.pc1);
    nb.insert(nb.end(), htn.begin() + in.pn0, htn.begin() + in.pn1);
  }
  HIPCALL(hipStreamSynchronize(stream));
  HIPCALL(hipMemcpy(rtc, c.data(), c.size() * sizeof(Blk), hipMemcpyHostToDevice));
  HIPCALL(hipMemcpy(rtn, nb.data(), nb.size() * sizeof(Blk), hipMemcpyHostToDevice));
  if (precision == MMC_MIXED)
    build_list<float>(lat32(), rtc, int(c.size()), rtn, int(nb.size()));
  else
    build_list<double>(lat64(), rtc, int(c.size()), rtn, int(nb.size()));
}

// One launch of the block solver over the instances of `set`.
template <class T>
void mmc_ctx::solve_blocks(Lat<T> Lt, const std::vector<int>& set, int max_trials) {
  HIPCALL(hipMemcpyAsync(gact, set.data(), set.size() * sizeof(int), hipMemcpyHostToDevice, stream));
  const size_t n6 = 6 * size_t(N);
  const int nb = int(set.size());
  if (max_hist <= 10)
    k_solve_block<T, 10><<<nb, kSolveThreads, 0, stream>>>(
        Lt, inst_d, ctl, gact, max_trials, zbuf[0], zbuf[1], gbuf[0], gbuf[1], Sh, Yh, n6, p, x, (double*)R, xref,
        nbuf[0], nbuf[1], ubuf[0], ubuf[1], cbuf[0], cbuf[1], pcount, lref, rec, cf_sph, cf_cap, iflags());
  else
    k_solve_block<T, kMaxHist><<<nb, kSolveThreads, 0, stream>>>(
        Lt, inst_d, ctl, gact, max_trials, zbuf[0], zbuf[1], gbuf[0], gbuf[1], Sh, Yh, n6, p, x, (double*)R, xref,
        nbuf[0], nbuf[1], ubuf[0], ubuf[1], cbuf[0], cbuf[1], pcount, lref, rec, cf_sph, cf_cap, iflags());
  HIPCALL(hipGetLastError());
  // the set is staged from host memory that the caller may reuse
  HIPCALL(hipStreamSynchronize(stream));
}

// Block tables of the device loop for the instances of `set` (ascending):
// each instance's blocks of the full 64-node table, in order, its beams in
// blocks of 256 and its contact shapes in blocks of 64 from the instance's
// first item; ranges of the beam and shape partials in columns of 64 items.
void mmc_ctx::graph_tables(const std::vector<int>& set) {
  std::vector<Blk> d, b, c;
  std::vector<GRng> r(size_t(n_inst), GRng{0, 0, 0, 0, 0, 0, 0, 0});
  auto split = [&](std::vector<Blk>& t, int inst, int first, int count, int E) {
    for (int o = 0; o < count; o += E) t.push_back(Blk{inst, first + o, std::min(E, count - o), 0});
  };
  for (int i : set) {
    const Inst& in = hinst[size_t(i)];
    GRng& g = r[size_t(i)];
    g.pd0 = int(d.size());
    d.insert(d.end(), htd.begin() + in.pd0, htd.begin() + in.pd1);
    g.pd1 = int(d.size());
    g.pb0 = int(b.size()) * kBW;
    split(b, i, in.b0, in.B, kBlock);
    g.pb1 = int(b.size()) * kBW;
    g.pc0 = int(c.size());
    split(c, i, in.s0, in.NS, kWave);
    g.pc1 = int(c.size());
  }
  // partial columns and table entries within what setup allocated (kBW
  // entries and columns per block of the kBlock tables)
  if (b.size() > htb.size() || c.size() > size_t(kBW) * htc.size())
    throw Fail(MMC_ERR_STATE, "device loop tables exceed their buffers");
  gPd = int(d.size());
  gPb = int(b.size());
  gPc = int(c.size());
  // the stream is idle here (the caller synchronised it): plain copies
  if (gPd) HIPCALL(hipMemcpy(gtd, d.data(), d.size() * sizeof(Blk), hipMemcpyHostToDevice));
  if (gPb) HIPCALL(hipMemcpy(gtb, b.data(), b.size() * sizeof(Blk), hipMemcpyHostToDevice));
  if (gPc) HIPCALL(hipMemcpy(gtc, c.data(), c.size() * sizeof(Blk), hipMemcpyHostToDevice));
  HIPCALL(hipMemcpy(grng, r.data(), r.size() * sizeof(GRng), hipMemcpyHostToDevice));
  HIPCALL(hipMemcpy(gact, set.data(), set.size() * sizeof(int), hipMemcpyHostToDevice));
  gset = set;
  if (list_valid) classify();
}

// helper blocks of k_elem_g for the flagged contact chunks
constexpr int kHelperBlocks = 256;

// The gather kernel of the device loop: the build for the longest history,
// with the decision fused (dec_split 0) or not.
template <class T>
void mmc_ctx::launch_gather(Lat<T> Lt, Ctl* C) {
  const size_t n6 = 6 * size_t(N);
  const FuseArgs fa{grng, dspart, cpart, bpart, gPc, gPb * kBW, iflags(), dcount, drows};
  auto go = [&](auto kern) {
    kern<<<gPd, kElemThreads, 0, stream>>>(Lt, inst_d, gtd, C, x, erec, cf_sph, ecap, eS, cent, zbuf[0], zbuf[1],
                                           gbuf[0], gbuf[1], p, Sh, Yh, n6, gpart, fa);
  };
  if (max_hist <= 10) {
    if (dec_split == 0)
      go(k_gather_g<T, 10, true>);
    else
      go(k_gather_g<T, 10, false>);
  } else {
    if (dec_split == 0)
      go(k_gather_g<T, kMaxHist, true>);
    else
      go(k_gather_g<T, kMaxHist, false>);
  }
}

// The direction kernel, built for the longest history of the instances.
template <class T>
void mmc_ctx::launch_dstate(Lat<T> Lt, const Ctl* C) {
  const size_t n6 = 6 * size_t(N);
  auto go = [&](auto kern) {
    kern<<<gPd, kElemThreads, 0, stream>>>(Lt, inst_d, gtd, C, zbuf[0], zbuf[1], gbuf[0], gbuf[1], Sh, Yh, n6, p, x,
                                           (double*)R, xref, nbuf[0], nbuf[1], dspart);
  };
  if (max_hist <= 10)
    go(k_dstate_g<T, 10>);
  else
    go(k_dstate_g<T, kMaxHist>);
}

template <class T>
void mmc_ctx::launch_elem(Lat<T> Lt, const Ctl* C) {
  const int H = std::min(gPc, kHelperBlocks);
  const int nb = gPb + (gPc + kBW - 1) / kBW + H + int(gset.size());
  k_elem_g<T><<<nb, kBlock, 0, stream>>>(C, Lt, inst_d, gtb, gtc, x, (const double*)R, rec, ubuf[0], ubuf[1], bpart,
                                         gPb, pcount, lref, cf_sph, cf_cap, cbuf[0], cbuf[1], cpart, gPc, cflag, hlist,
                                         hcount, H, gact, grng, dspart, gPd, cent, EllOut{erec, ecap, eS});
}

// Flags the chunks of the loop's contact table whose lists are long
// (k_classify); stream-ordered after the list build or table change that
// calls it, before any later replay.
void mmc_ctx::classify() {
  if (gPc == 0) return;
  HIPCALL(hipMemsetAsync(hcount, 0, sizeof(unsigned), stream));
  const int per = kBlock / kWave;
  k_classify<<<(gPc + per - 1) / per, kBlock, 0, stream>>>(gtc, gPc, pcount, slab, contact_serial, cflag, hlist,
                                                           hcount);
  HIPCALL(hipGetLastError());
}

void mmc_ctx::launch_decide(Ctl* C) {
  const int na = int(gset.size());
  const int cw = gPc, bw = gPb * kBW;  // columns of the partials (one per 64 items)
  auto go = [&](auto kern, int d) {
    kern<<<na * d, kBlock, 0, stream>>>(C, inst_d, gact, grng, dspart, gPd, cpart, cw, bpart, bw, gpart, iflags(),
                                        drows, dcount);
  };
  const bool small = max_hist <= 10;
  switch (dec_split) {
    case 1:
      small ? go(k_decide_g<1, 10>, 1) : go(k_decide_g<1, kMaxHist>, 1);
      break;
    case 2:
      small ? go(k_decide_g<2, 10>, 2) : go(k_decide_g<2, kMaxHist>, 2);
      break;
    case 4:
      small ? go(k_decide_g<4, 10>, 4) : go(k_decide_g<4, kMaxHist>, 4);
      break;
    case 16:
      small ? go(k_decide_g<16, 10>, 16) : go(k_decide_g<16, kMaxHist>, 16);
      break;
    default:
      small ? go(k_decide_g<8, 10>, 8) : go(k_decide_g<8, kMaxHist>, 8);
  }
}

// Launch-shape tuner of the device loop, run at the first capture over a set
// of each size class (a batch's graph is captured again over fewer instances
// as it drains; few instances favour many blocks each, many instances one),
// on the current state: the gather and decision of a trial as two kernels,
// the decision at 1, 2, 4, 8 and 16 blocks per instance, and as one (the
// decision in the gather's last block per instance), each timed with HIP
// events over back-to-back launches (time_loop_kernel).  Of the splits, the
// most blocks within 3% of the fastest; the fused form when it beats that.
// The choice changes no result (one fixed fold order per row), so the tuning
// is free to differ between contexts and size classes.  A tune in the middle
// of a solve changes no state: the timed kernels read a scratch copy of the
// control blocks and write only what the loop rewrites before its next use
// (the candidate gradient, the spare history slot, partials, folded rows,
// the arrival counters, left at zero, and the evaluation flags, zero between
// two polls).  MMC_TUNE=0 keeps the split 8; MMC_DECIDE_BLOCKS fixes the
// shape (0: fused; tests).
void mmc_ctx::tune() {
  const char* env = std::getenv("MMC_TUNE");
  const char* fd = std::getenv("MMC_DECIDE_BLOCKS");
  if (fd) {
    const int d = std::atoi(fd);
    if (d != 0 && d != 1 && d != 2 && d != 4 && d != 8 && d != 16)
      throw Fail(MMC_ERR_ARG, "MMC_DECIDE_BLOCKS must be 0, 1, 2, 4, 8 or 16");
    dec_split = d;
    return;
  }
  if (env && std::strcmp(env, "0") == 0) return;
  constexpr int kLaunches = 20;
  constexpr int kSplits = 5;
  const int splits[kSplits] = {1, 2, 4, 8, 16};
  dec_split = 8;
  const double gather = 1e3 * time_loop_kernel(MMC_K_GATHER_G, kLaunches);
  int best = 0;
  for (int k = 0; k < kSplits; k++) {
    dec_split = splits[k];
    tune_dec_us[k] = gather + 1e3 * time_loop_kernel(MMC_K_DECIDE_G, kLaunches);
    if (tune_dec_us[k] < tune_dec_us[best]) best = k;
  }
  // the most blocks within 3% of the fastest: the timed launches find the
  // partial rows in cache, while in the loop the other kernels have just
  // written them, and more blocks spread those colder loads
  for (int k = kSplits - 1; k > best; k--)
    if (tune_dec_us[k] <= 1.03 * tune_dec_us[best]) {
      best = k;
      break;
    }
  dec_split = 0;
  tune_dec_us[kSplits] = 1e3 * time_loop_kernel(MMC_K_GATHER_G, kLaunches);
  if (tune_dec_us[kSplits] >= tune_dec_us[best]) dec_split = splits[best];
}

void mmc_ctx::capture(const std::vector<int>& set) {
  SyncLock sync_lock(g_sync_mu);
  HIPCALL(hipStreamSynchronize(stream));
  graph_tables(set);
  // size class: ceil(log2(instances))
  int cls = 0;
  while ((size_t(1) << cls) < set.size()) cls++;
  LoopShape& shape = shapes[cls];
  if (!shape.tuned) {
    tune();
    shape.tuned = true;
    shape.split = dec_split;
    std::copy(tune_dec_us, tune_dec_us + 6, shape.us);
  } else {
    dec_split = shape.split;
    std::copy(shape.us, shape.us + 6, tune_dec_us);
  }
  captures++;
  if (precision == MMC_MIXED)
    capture_iteration<float>(lat32());
  else
    capture_iteration<double>(lat64());
}

// One device-loop kernel timed on its own: the kernels run against a scratch
// copy of the control blocks (tctl) set to a trial of every captured
// instance, and write only buffers that the loop treats as scratch between
// iterations (the candidate z and gradient, the direction, the spare history
// slot, poses, records and partials).  The trial kernels only read the
// control blocks: `launches` back-to-back launches between two events.  The
// decision kernel (and the gather kernel with the decision fused) rewrites
// them, so each of its launches follows a restore kernel; the restores alone
// are timed the same way and subtracted.
namespace {
__global__ __launch_bounds__(kBlock) void k_copy_words(double* __restrict__ dst, const double* __restrict__ src,
                                                       long n) {
  const long i = long(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) dst[i] = src[i];
}
}  // namespace
double mmc_ctx::time_loop_kernel(int kernel, int launches) {
  SyncLock sync_lock(g_sync_mu);
  if (gset.empty()) throw Fail(MMC_ERR_STATE, "the device loop has not run yet (no mmc_minimize)");
  if (!ev0) {
    HIPCALL(hipEventCreate(&ev0));
    HIPCALL(hipEventCreate(&ev1));
  }
  HIPCALL(hipStreamSynchronize(stream));
  std::vector<Ctl> h(static_cast<size_t>(n_inst));
  HIPCALL(hipMemcpy(h.data(), ctl, h.size() * sizeof(Ctl), hipMemcpyDeviceToHost));
  for (Ctl& c : h) {
    c.status = ST_RUNNING;
    c.phase = kernel == MMC_K_DSTATE_G ? PH_DIRECTION : PH_SEARCH;
    c.step = 1.0;
    c.bt = 0;
    c.dg = -1.0;
    // full history (m - 1 retained pairs): the most rows a gather or a
    // decision has; the decision then folds rows the gather may not have
    // written since setup (scratch values, only timed)
    if ((kernel == MMC_K_GATHER_G || kernel == MMC_K_DECIDE_G) && c.m > 0) c.count = c.m;
  }
  if (!tctl) tctl = dalloc<Ctl>(2 * size_t(n_inst));  // [0, n): scratch, [n, 2n): its initial state
  Ctl* tinit = tctl + n_inst;
  HIPCALL(hipMemcpy(tctl, h.data(), h.size() * sizeof(Ctl), hipMemcpyHostToDevice));
  HIPCALL(hipMemcpy(tinit, h.data(), h.size() * sizeof(Ctl), hipMemcpyHostToDevice));
  const long words = long(n_inst) * long(sizeof(Ctl) / 8);
  const int wblocks = int((words + kBlock - 1) / kBlock);
  auto loop = [&](auto Lt, auto* tag, bool restore, bool work) {
    using T = std::remove_pointer_t<decltype(tag)>;
    HIPCALL(hipEventRecord(ev0, stream));
    for (int i = 0; i < launches; i++) {
      if (restore) k_copy_words<<<wblocks, kBlock, 0, stream>>>((double*)tctl, (const double*)tinit, words);
      if (!work) continue;
      switch (kernel) {
        case MMC_K_DSTATE_G:
          launch_dstate<T>(Lt, tctl);
          break;
        case MMC_K_ELEM_G:
          launch_elem<T>(Lt, tctl);
          break;
        case MMC_K_GATHER_G:
          launch_gather<T>(Lt, tctl);
          break;
        default:
          launch_decide(tctl);
      }
    }
    HIPCALL(hipEventRecord(ev1, stream));
    HIPCALL(hipEventSynchronize(ev1));
    HIPCALL(hipGetLastError());
    float ms = 0;
    HIPCALL(hipEventElapsedTime(&ms, ev0, ev1));
    return double(ms);
  };
  auto run = [&](bool restore, bool work) {
    return precision == MMC_MIXED ? loop(lat32(), (float*)nullptr, restore, work)
                                  : loop(lat64(), (double*)nullptr, restore, work);
  };
  // the decision rewrites the control blocks: in its own kernel, or in the
  // gather's last blocks when fused (the decision kernel then has no launch)
  const bool decides = kernel == MMC_K_DECIDE_G || (kernel == MMC_K_GATHER_G && dec_split == 0);
  if (kernel == MMC_K_DECIDE_G && dec_split == 0) return 0.0;
  if (!decides) return run(false, true) / launches;
  const double both = run(true, true), restores = run(true, false);
  return std::max(0.0, both - restores) / launches;
}

namespace {
__global__ __launch_bounds__(kBlock) void k_release_hold(Ctl* __restrict__ Cs, int n) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < n && Cs[b].status == ST_HOLD) Cs[b].status = ST_RUNNING;
}
}  // namespace

// Simulator::minimize, MC/src/sim.cpp:174-211, for every active instance,
// driven from the device: the host replays the captured graph in batches and
// reads the 16-word status heads of the control blocks after each batch.  It
// steps in only to rebuild a stale contact list (status HOLD) and for the
// initial evaluation.  Inactive instances (active[b] == 0) are left alone.
// Control block of instance b at the start of a solve at platen top, its
// current record from the evaluation at z = 0 (hrec); run = false marks it as
// not part of the solve.
void mmc_ctx::init_ctl(int b, double top, bool run) {
  Ctl& h = hctl[b];
  std::memset(&h, 0, sizeof(Ctl));
  const mmc_config& cf = icfg[b];
  std::memcpy(h.cur, hrec + b * kRec, kRec * sizeof(double));
  h.m = cf.history;
  h.max_it = cf.max_iterations;
  h.max_ls = cf.max_line_search;
  h.rel_escape = cf.relative_escape;
  h.top = top;
  h.armijo = cf.armijo;
  h.min_step = cf.min_line_step;
  h.tol_rms = cf.gradient_rms_tolerance;
  h.tol_max = cf.gradient_max_tolerance;
  h.rel_tol = cf.relative_tolerance;
  h.pen_tol = cf.penetration_tolerance;
  h.ts = cf.translation_scale;
  h.skin_half = 0.5 * hinst[b].S.skin;
  h.ls_mode = cf.line_search != 0;
  h.wolfe = precision == MMC_MIXED || cf.line_search == 2;
  h.wolfe_eps = 1e-6;
  h.phase = PH_DIRECTION;
  h.coef.cg = 0.10;
  if (!run) {
    h.status = ST_CONVERGED;  // not part of this solve
  } else if (cf.max_iterations <= 0) {
    h.status = ST_ITERCAP;
  } else {
    h.iterations = 1;
    h.status = host_converged(cf, h.cur) ? ST_CONVERGED : ST_RUNNING;
  }
}

void mmc_ctx::minimize_graph(const double* tops, const int32_t* active, mmc_opt_stats* st) {
  const int n = n_inst;
  auto clock = [] { return std::chrono::steady_clock::now(); };
  auto secs = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
  const auto t_init = clock();
  zero(zbuf[0], long(6 * size_t(N)));
  terms_out0 = true;
  evaluate(zbuf[0], nullptr, 0.0, nullptr, gbuf[0], tops, false, rec_slot(0));
  loop_s[0] += secs(t_init, clock());
  bool any = false;
  for (int b = 0; b < n; b++) {
    init_ctl(b, tops[b], !active || active[b]);
    any = any || hctl[b].status == ST_RUNNING;
  }
  // every lattice's solve starts here: none is in flight for mmc_batch_*
  std::fill(in_flight.begin(), in_flight.end(), 0);
  HIPCALL(hipMemcpyAsync(ctl, hctl, size_t(n) * sizeof(Ctl), hipMemcpyHostToDevice, stream));
  if (block_engine) {
    // launches of kSolveBudget trials per workgroup: a lattice held for a
    // list rebuild waits at most one budget for the others; the total is
    // bounded as the graph loop's replays are
    int max_it = 0, max_ls = 0;
    for (const mmc_config& cf : icfg) max_it = std::max(max_it, cf.max_iterations), max_ls = std::max(max_ls, cf.max_line_search);
    const int64_t guard = int64_t(max_it + 1) * (max_ls + 3) * 2 + 64;
    int64_t trials = 0;
    for (;;) {
      std::vector<int> set;
      int hold = 0;
      for (int b = 0; b < n; b++) {
        if (hctl[b].status == ST_RUNNING || hctl[b].status == ST_HOLD) set.push_back(b);
        hold += hctl[b].status == ST_HOLD;
      }
      if (set.empty()) break;
      if (hold > 0) {
        // the held trials' positions are still in x: rebuild there and redo them
        std::vector<int> held;
        for (int b = 0; b < n; b++)
          if (hctl[b].status == ST_HOLD) held.push_back(b);
        rebuild_held(held);
        k_release_hold<<<(n + kBlock - 1) / kBlock, kBlock, 0, stream>>>(ctl, n);
      }
      if (trials > guard) throw Fail(MMC_ERR_STATE, "device minimiser exceeded its trial bound");
      trials += kSolveBudget;
      const auto t_run = clock();
      if (precision == MMC_MIXED)
        solve_blocks<float>(lat32(), set, kSolveBudget);
      else
        solve_blocks<double>(lat64(), set, kSolveBudget);
      HIPCALL(hipMemcpy2DAsync(hctl, sizeof(Ctl), ctl, sizeof(Ctl), 16 * sizeof(int), n, hipMemcpyDeviceToHost,
                               stream));
      HIPCALL(hipStreamSynchronize(stream));
      loop_s[2] += secs(t_run, clock());
      loop_polls++;
    }
  } else {
  // the loop's graph covers the running instances only; as a batch drains it
  // is captured again over the remaining ones (whenever they are at most
  // half of those captured), so that the replays of the last few lattices do
  // not dispatch every other lattice's blocks
  auto running_set = [&] {
    std::vector<int> set;
    for (int b = 0; b < n; b++)
      if (hctl[b].status == ST_RUNNING || hctl[b].status == ST_HOLD) set.push_back(b);
    return set;
  };
  {
    std::vector<int> set = running_set();
    const auto t_cap = clock();
    if (!set.empty() && (set != gset || !graph_exec)) capture(set);
    loop_s[1] += secs(t_cap, clock());
  }
  const auto t_loop = clock();
  // every replay advances each running phase machine by at least one step;
  // bound the total so that a logic error cannot spin forever
  int max_it = 0, max_ls = 0;
  for (const mmc_config& cf : icfg) max_it = std::max(max_it, cf.max_iterations), max_ls = std::max(max_ls, cf.max_line_search);
  const int64_t guard = int64_t(max_it + 1) * (max_ls + 3) * 2 + 64;
  int64_t replays = 0;
  int batch = kPollLaunches0;
  auto heads = [&] {
    HIPCALL(hipMemcpy2DAsync(hctl, sizeof(Ctl), ctl, sizeof(Ctl), 16 * sizeof(int), n, hipMemcpyDeviceToHost,
                             stream));
    HIPCALL(hipStreamSynchronize(stream));
    int running = 0, hold = 0;
    for (int b = 0; b < n; b++) {
      running += hctl[b].status == ST_RUNNING;
      hold += hctl[b].status == ST_HOLD;
    }
    return std::make_pair(running, hold);
  };
  auto state = any ? heads() : std::make_pair(0, 0);
  while (state.first + state.second > 0) {
    if (state.second > 0) {
      // the held trials' positions are still in x: rebuild there and redo them
      std::vector<int> held;
      for (int b = 0; b < n; b++)
        if (hctl[b].status == ST_HOLD) held.push_back(b);
      rebuild_held(held);
      k_release_hold<<<(n + kBlock - 1) / kBlock, kBlock, 0, stream>>>(ctl, n);
    }
    for (int r = 0; r < batch; r++) HIPCALL(hipGraphLaunch(graph_exec, stream));
    replays += int64_t(batch) * kGraphIterations;
    graph_replays += batch;
    batch = std::min(kPollLaunchesMax, batch * 2);
    state = heads();
    if (replays > guard) throw Fail(MMC_ERR_STATE, "device minimiser exceeded its replay bound");
    loop_polls++;
    const int left = state.first + state.second;
    if (left > 0 && 2 * size_t(left) <= gset.size()) {
      const auto t_cap = clock();
      capture(running_set());
      loop_s[1] += secs(t_cap, clock());
    }
  }
  loop_s[2] += secs(t_loop, clock());
  }
  HIPCALL(hipMemcpyAsync(hctl, ctl, size_t(n) * sizeof(Ctl), hipMemcpyDeviceToHost, stream));
  HIPCALL(hipStreamSynchronize(stream));
  for (int b = 0; b < n; b++) {
    const Ctl& h = hctl[b];
    st[b] = mmc_opt_stats{};
    st[b].iterations = h.iterations;
    st[b].backtracks = h.backtracks;
    st[b].converged = h.status == ST_CONVERGED;
    st[b].evaluations = h.evals;
    evaluations += h.evals;
    total_evaluations += h.evals;
  }
  // the current iterates; Armijo steps decrease the energy strictly, so the
  // reference's best-iterate fallback (MC/src/sim.cpp:209) selects them (the
  // approximate-Wolfe acceptances of the mixed path admit changes within the
  // energy's rounding noise, where "best" carries no information)
  k_take_current<<<Ps, kBlock, 0, stream>>>(tn, ctl, N, zbuf[0], zbuf[1], z);
  HIPCALL(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Per-lattice progression of a batch (no reference counterpart).  The rounds
// of minimize_graph start every lattice's solve together and end it when the
// slowest has finished; here each lattice starts (batch_start) when its
// driver has its next target and is finished (batch_finish) when its own
// solve has ended, while the others go on in the same graph replays.  A
// lattice's own sequence of operations is the round's: its iterate set to 0,
// the evaluation there, the phase machine, the final evaluation at its
// iterate, the commit; the other lattices' evaluations in between only
// rewrite what every trial recomputes (poses, records, contact forces), so
// its results are those of a round, bit for bit.  Line search 0 only (the
// per-term line search keeps per-item energies of the current iterate that
// an evaluation would rewrite).
// z for an evaluation: the in-flight lattices' current iterates, 0 for the
// others (not started, or finished and committed) and for those in zmask
void mmc_ctx::batch_iterates(const int32_t* zmask) {
  const int n = n_inst;
  for (int b = 0; b < n; b++) hmask[b] = !in_flight[size_t(b)] || (zmask && zmask[b]);
  HIPCALL(hipMemcpyAsync(mask_d, hmask, size_t(n) * sizeof(int), hipMemcpyHostToDevice, stream));
  k_take_current<<<Ps, kBlock, 0, stream>>>(tn, ctl, N, zbuf[0], zbuf[1], z);
  k_zero_masked<<<Ps, kBlock, 0, stream>>>(tn, mask_d, N, z);
  HIPCALL(hipGetLastError());
}

void mmc_ctx::batch_start(const double* tops, const int32_t* mask) {
  if (block_engine) throw Fail(MMC_ERR_STATE, "per-lattice progression runs on the graph loop (MMC_DEVICE_LOOP=block)");
  const int n = n_inst;
  const bool first = atop.empty();
  if (first) {
    atop.assign(size_t(n), 0.0);
    in_flight.assign(size_t(n), 0);
  }
  std::vector<int> sm(size_t(n), 0);
  int started = 0;
  for (int b = 0; b < n; b++) {
    if (!mask[b]) continue;
    if (in_flight[size_t(b)]) throw Fail(MMC_ERR_STATE, "batch_start of a lattice whose solve is in flight");
    if (icfg[b].line_search != 0) throw Fail(MMC_ERR_ARG, "per-lattice progression supports line_search 0 only");
    atop[size_t(b)] = tops[b];
    sm[size_t(b)] = 1;
    started++;
  }
  if (!started && !first) return;
  const auto t0 = std::chrono::steady_clock::now();
  batch_iterates(mask);
  terms_out0 = false;
  evaluate(z, nullptr, 0.0, nullptr, gc, atop.data(), false, rec_slot(0));
  HIPCALL(hipMemcpyAsync(smask_d, sm.data(), size_t(n) * sizeof(int), hipMemcpyHostToDevice, stream));
  k_seed_masked<<<Ps, kBlock, 0, stream>>>(tn, smask_d, N, gc, zbuf[0], gbuf[0]);
  HIPCALL(hipGetLastError());
  if (first) {
    // every control block, the others marked as not solving
    for (int b = 0; b < n; b++) init_ctl(b, atop[size_t(b)], sm[size_t(b)] != 0);
    HIPCALL(hipMemcpyAsync(ctl, hctl, size_t(n) * sizeof(Ctl), hipMemcpyHostToDevice, stream));
  } else {
    for (int b = 0; b < n; b++)
      if (sm[size_t(b)]) {
        init_ctl(b, atop[size_t(b)], true);
        HIPCALL(hipMemcpyAsync(ctl + b, hctl + b, sizeof(Ctl), hipMemcpyHostToDevice, stream));
      }
  }
  HIPCALL(hipStreamSynchronize(stream));
  for (int b = 0; b < n; b++)
    if (sm[size_t(b)]) in_flight[size_t(b)] = 1;
  loop_s[0] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Graph replays until at least min_done of the in-flight lattices (or all of
// them) have ended their solves; running[b] = 1 for the in-flight lattices
// still solving.
void mmc_ctx::batch_advance(int min_done, int32_t* running) {
  const int n = n_inst;
  if (in_flight.empty()) in_flight.assign(size_t(n), 0);
  auto heads = [&] {
    HIPCALL(hipMemcpy2DAsync(hctl, sizeof(Ctl), ctl, sizeof(Ctl), 16 * sizeof(int), n, hipMemcpyDeviceToHost,
                             stream));
    HIPCALL(hipStreamSynchronize(stream));
  };
  auto live = [&](int b) { return hctl[b].status == ST_RUNNING || hctl[b].status == ST_HOLD; };
  int flying = 0;
  for (int b = 0; b < n; b++) flying += in_flight[size_t(b)] != 0;
  const int need = std::max(1, std::min(min_done, flying));
  auto ended = [&] {
    int e = 0;
    for (int b = 0; b < n; b++) e += in_flight[size_t(b)] && !live(b);
    return e;
  };
  auto running_set = [&] {
    std::vector<int> set;
    for (int b = 0; b < n; b++)
      if (in_flight[size_t(b)] && live(b)) set.push_back(b);
    return set;
  };
  heads();
  if (flying > 0 && ended() < need) {
    const auto t_loop = std::chrono::steady_clock::now();
    {
      const std::vector<int> set = running_set();
      const bool covered =
          graph_exec != nullptr && std::includes(gset.begin(), gset.end(), set.begin(), set.end());
      if (!set.empty() && (!covered || 2 * set.size() <= gset.size())) {
        const auto t_cap = std::chrono::steady_clock::now();
        capture(set);
        loop_s[1] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_cap).count();
      }
    }
    int max_it = 0, max_ls = 0;
    for (const mmc_config& cf : icfg) max_it = std::max(max_it, cf.max_iterations), max_ls = std::max(max_ls, cf.max_line_search);
    const int64_t guard = int64_t(max_it + 1) * (max_ls + 3) * 2 + 64;
    int64_t replays = 0;
    int batch = kPollLaunches0;
    for (;;) {
      std::vector<int> held;
      for (int b = 0; b < n; b++)
        if (in_flight[size_t(b)] && hctl[b].status == ST_HOLD) held.push_back(b);
      if (!held.empty()) {
        // the held trials' positions are still in x: rebuild there and redo them
        rebuild_held(held);
        k_release_hold<<<(n + kBlock - 1) / kBlock, kBlock, 0, stream>>>(ctl, n);
      }
      if (running_set().empty()) break;
      for (int r = 0; r < batch; r++) HIPCALL(hipGraphLaunch(graph_exec, stream));
      replays += int64_t(batch) * kGraphIterations;
      graph_replays += batch;
      batch = std::min(kPollLaunchesMax, batch * 2);
      heads();
      loop_polls++;
      if (ended() >= need) break;
      if (replays > guard) throw Fail(MMC_ERR_STATE, "device minimiser exceeded its replay bound");
      const std::vector<int> set = running_set();
      if (!set.empty() && 2 * set.size() <= gset.size()) {
        const auto t_cap = std::chrono::steady_clock::now();
        capture(set);
        loop_s[1] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_cap).count();
      }
    }
    loop_s[2] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_loop).count();
  }
  for (int b = 0; b < n; b++) running[b] = in_flight[size_t(b)] && live(b);
}

// The final evaluation of the masked lattices (in flight, solves ended) at
// their iterates, with trial states for mmc_commit_batch; statistics and
// outputs into stats[b], fin[b] for the masked b only.
void mmc_ctx::batch_finish(const int32_t* mask, mmc_opt_stats* so, mmc_objective_out* fo) {
  const int n = n_inst;
  if (in_flight.empty()) in_flight.assign(size_t(n), 0);
  HIPCALL(hipMemcpyAsync(hctl, ctl, size_t(n) * sizeof(Ctl), hipMemcpyDeviceToHost, stream));
  HIPCALL(hipStreamSynchronize(stream));
  std::vector<mmc_opt_stats> s(static_cast<size_t>(n));
  std::vector<mmc_objective_out> f(static_cast<size_t>(n));
  for (int b = 0; b < n; b++) {
    if (!mask[b]) continue;
    if (!in_flight[size_t(b)] || hctl[b].status == ST_RUNNING || hctl[b].status == ST_HOLD)
      throw Fail(MMC_ERR_STATE, "batch_finish of a lattice whose solve has not ended");
    const Ctl& h = hctl[b];
    s[size_t(b)].iterations = h.iterations;
    s[size_t(b)].backtracks = h.backtracks;
    s[size_t(b)].converged = h.status == ST_CONVERGED;
    s[size_t(b)].evaluations = h.evals;
    evaluations += h.evals;
    total_evaluations += h.evals;
  }
  const auto t0 = std::chrono::steady_clock::now();
  batch_iterates(nullptr);
  final_evaluation(atop.data(), s.data(), f.data());
  HIPCALL(hipStreamSynchronize(stream));
  loop_s[4] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int b = 0; b < n; b++) {
    if (!mask[b]) continue;
    so[b] = s[size_t(b)];
    if (fo) fo[b] = f[size_t(b)];
    in_flight[size_t(b)] = 0;
  }
}

// The same loop driven from the host (one synchronisation per trial): kept as
// a cross-check of the device phase machine (MMC_HOST_LOOP=1, one lattice).
void mmc_ctx::minimize_host(double top, mmc_opt_stats& st) {
  const int P = Ps;
  const size_t n6 = 6 * size_t(N);
  const mmc_config& cf = cfg;
  const double mt = cf.max_translation_step / cf.translation_scale, mr = cf.max_rotation_step;
  zero(z, long(n6));
  double cur[kRec], cand[kRec];
  evaluate(z, nullptr, 0.0, nullptr, g, &top, false, rec_slot(0));
  std::memcpy(cur, hrec, sizeof cur);
  const int m = cf.history;
  int count = 0, head = 0;
  static thread_local double SY[kMaxHist][kMaxHist], YY[kMaxHist][kMaxHist], uu[kMaxHist], vv[kMaxHist];
  auto direction = [&](const DirCoef& C) {
    tick(MMC_K_DIRECTION);
    k_direction<<<P, kBlock, 0, stream>>>(N, n6, g, Sh, Yh, C, p, lock, mt, mr, vpart);
    tock(MMC_K_DIRECTION);
    k_sum_rows<<<1, kBlock, 0, stream>>>(vpart, P, 1, scal);
    read(host + 32, scal, sizeof(double));
    return host[32];
  };
  DirCoef steepest{};
  steepest.cg = 1.0;
  auto line_search = [&](double dg) {
    double step = 1.0;
    for (int bt = 0; bt < cf.max_line_search; bt++) {
      evaluate(z, p, step, zc, gc, &top, false, rec_slot(1));
      std::memcpy(cand, hrec, sizeof cand);
      st.backtracks++;
      if (cand[9] != 0.0 && cand[0] <= cur[0] + cf.armijo * step * dg) return true;
      step *= 0.5;
      if (step < cf.min_line_step) break;
    }
    return false;
  };
  for (int it = 0; it < cf.max_iterations; it++) {
    st.iterations = it + 1;
    if (host_converged(cf, cur)) {
      st.converged = 1;
      break;
    }
    DirCoef C{};
    compact_coefficients(m, count, head, SY, YY, uu, vv, C);
    double dg = direction(C);
    if (!(dg < 0)) {  // MC/src/sim.cpp:191
      count = head = 0;
      dg = direction(steepest);
    }
    bool ok = line_search(dg);
    if (!ok) {  // steepest-descent retry, MC/src/sim.cpp:193-203
      count = head = 0;
      dg = direction(steepest);
      ok = line_search(dg);
    }
    if (!ok) break;
    if (m > 0) {  // curvature pair, MC/src/sim.cpp:205-206
      PairArgs A = pair_slots(m, count, head);
      tick(MMC_K_PAIR);
      k_pair_dots<<<P, kBlock, 0, stream>>>(N, n6, z, zc, g, gc, Sh, Yh, A, vpart);
      tock(MMC_K_PAIR);
      const int K = 4 * A.nold + 5;
      k_sum_rows<<<1, kBlock, 0, stream>>>(vpart, P, K, scal);
      read(host + 128, scal, K * sizeof(double));
      history_update(m, count, head, SY, YY, uu, vv, A, host + 128);
    }
    std::swap(z, zc);
    std::swap(g, gc);
    std::memcpy(cur, cand, sizeof cur);
  }
}

// The reference's final objective(z, top, true) (MC/src/sim.cpp:210) at every
// instance's iterate z: trial states for the commit, the reported record and
// the convergence verdict on it.
void mmc_ctx::final_evaluation(const double* tops, mmc_opt_stats* st, mmc_objective_out* fin) {
  evaluate(z, nullptr, 0.0, nullptr, gc, tops, true, rec_slot(2));
  HIPCALL(hipMemcpyAsync(zt, z, 6 * size_t(N) * sizeof(double), hipMemcpyDeviceToDevice, stream));
  have_trial = true;
  for (int b = 0; b < n_inst; b++) {
    const double* f2 = hrec + b * kRec;
    st[b].rms = f2[6];
    st[b].max = f2[7];
    st[b].energy = f2[0];
    st[b].converged = host_converged(icfg[b], f2) ? 1 : 0;
    if (fin) fill_out(f2, fin + b);
  }
}

extern "C" {

// Simulator::minimize, MC/src/sim.cpp:174-211.  Every vector stays on the
// device; the final objective (with trial states) is evaluated at the last
// iterate, which is left in z.
int mmc_minimize(mmc_ctx* c, double top, mmc_opt_stats* stats, mmc_objective_out* fin, double* z_out) {
  return guarded(c, [&] {
    c->need_single("mmc_minimize");
    const size_t n6 = 6 * size_t(c->N);
    c->evaluations = 0;
    mmc_opt_stats st{};
    if (use_host_loop()) {
      if (c->cfg.line_search != 0) throw Fail(MMC_ERR_ARG, "the host-driven loop supports line_search 0 only");
      c->minimize_host(top, st);
    } else {
      c->minimize_graph(&top, nullptr, &st);
    }
    c->final_evaluation(&top, &st, fin);
    st.evaluations = c->evaluations;
    if (stats) *stats = st;
    if (z_out) {
      std::vector<double> s(n6);
      c->read(s.data(), c->z, n6 * sizeof(double));
      aos_from_soa(s.data(), z_out, c->N, 6);
    }
    HIPCALL(hipStreamSynchronize(c->stream));
  });
}

// mmc_minimize for every active instance of a batch context at once.
int mmc_minimize_batch(mmc_ctx* c, const double* tops, const int32_t* active, mmc_opt_stats* stats,
                       mmc_objective_out* fin) {
  return guarded(c, [&] {
    if (!tops || !stats) throw Fail(MMC_ERR_ARG, "NULL argument");
    c->evaluations = 0;
    const auto t0 = std::chrono::steady_clock::now();
    c->minimize_graph(tops, active, stats);
    const auto t1 = std::chrono::steady_clock::now();
    c->final_evaluation(tops, stats, fin);
    HIPCALL(hipStreamSynchronize(c->stream));
    c->loop_s[3] += std::chrono::duration<double>(t1 - t0).count();
    c->loop_s[4] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
  });
}

int mmc_batch_start(mmc_ctx* c, const double* tops, const int32_t* mask) {
  return guarded(c, [&] {
    if (!tops || !mask) throw Fail(MMC_ERR_ARG, "NULL argument");
    c->batch_start(tops, mask);
  });
}

int mmc_batch_advance(mmc_ctx* c, int32_t min_done, int32_t* running) {
  return guarded(c, [&] {
    if (!running) throw Fail(MMC_ERR_ARG, "NULL argument");
    c->batch_advance(min_done, running);
  });
}

int mmc_batch_finish(mmc_ctx* c, const int32_t* mask, mmc_opt_stats* stats, mmc_objective_out* fin) {
  return guarded(c, [&] {
    if (!mask || !stats) throw Fail(MMC_ERR_ARG, "NULL argument");
    c->batch_finish(mask, stats, fin);
  });
}

// commit of the instances whose mask entry is set (all when mask is NULL)
int mmc_commit_batch(mmc_ctx* c, const int32_t* mask) {
  return guarded(c, [&] {
    if (!c->have_trial) throw Fail(MMC_ERR_STATE, "missing constitutive trial states");
    for (int b = 0; b < c->n_inst; b++) c->hmask[b] = mask ? (mask[b] != 0) : 1;
    HIPCALL(hipMemcpyAsync(c->mask_d, c->hmask, c->n_inst * sizeof(int), hipMemcpyHostToDevice, c->stream));
    // recompute state(z) of the trial evaluation in FP64 (for either precision
    // policy the committed pose is FP64), then copy it into the base pose
    k_state<double><<<c->Ps, kBlock, 0, c->stream>>>(c->lat64(), c->inst_d, c->tn, c->tops_d, c->zt, nullptr, 0.0,
                                                     nullptr, c->x, c->zin, nullptr, TermBufs{nullptr, nullptr},
                                                     c->spart);
    k_commit_nodes<<<c->Ps, kBlock, 0, c->stream>>>(c->tn, c->mask_d, c->N, c->x, c->zin, c->bpos, c->bquat);
    k_commit_beams<<<c->Pb, kBlock, 0, c->stream>>>(c->tb, c->mask_d, c->B, c->trial, c->plastic);
    HIPCALL(hipGetLastError());
    HIPCALL(hipStreamSynchronize(c->stream));
    c->have_trial = false;
  });
}

int mmc_commit(mmc_ctx* c) { return mmc_commit_batch(c, nullptr); }

int mmc_read_state(mmc_ctx* c, double* xh, double* qh, double* ph) {
  return guarded(c, [&] {
    c->need_single("mmc_read_state");
    std::vector<double> s(7 * size_t(std::max(c->N, c->B)));
    if (xh) {
      c->read(s.data(), c->bpos, 3 * size_t(c->N) * sizeof(double));
      aos_from_soa(s.data(), xh, c->N, 3);
    }
    if (qh) {
      c->read(s.data(), c->bquat, 4 * size_t(c->N) * sizeof(double));
      aos_from_soa(s.data(), qh, c->N, 4);
    }
    if (ph) {
      c->read(s.data(), c->plastic, 7 * size_t(c->B) * sizeof(double));
      aos_from_soa(s.data(), ph, c->B, 7);
    }
  });
}

int mmc_read_trial(mmc_ctx* c, double* t) {
  return guarded(c, [&] {
    c->need_single("mmc_read_trial");
    if (!t) throw Fail(MMC_ERR_ARG, "trial is NULL");
    std::vector<double> s(size_t(MMC_TRIAL_STRIDE) * c->B);
    c->read(s.data(), c->trial, s.size() * sizeof(double));
    aos_from_soa(s.data(), t, c->B, MMC_TRIAL_STRIDE);
  });
}

// state statistics of every instance (out[n_inst])
int mmc_read_stats_batch(mmc_ctx* c, mmc_state_stats* o) {
  return guarded(c, [&] {
    if (!o) throw Fail(MMC_ERR_ARG, "NULL argument");
    k_stats_beams<<<c->Pb, kBlock, 0, c->stream>>>(c->tb, c->B, c->plastic, c->stb);
    k_stats_nodes<<<c->Ps, kBlock, 0, c->stream>>>(c->tn, c->N, c->bpos, c->stn);
    k_stats_final<<<c->n_inst, kBlock, 0, c->stream>>>(c->inst_d, c->stb, c->Pb, c->stn, c->Ps, c->stout);
    c->read(c->hstat, c->stout, size_t(c->n_inst) * 8 * sizeof(double));
    for (int b = 0; b < c->n_inst; b++) {
      const double* r = c->hstat + 8 * b;
      o[b].sum_axial = r[0];
      o[b].max_axial = r[1];
      o[b].sum_rot = r[2];
      o[b].max_rot = r[3];
      o[b].plastic_beams = int32_t(r[4]);
      o[b].plastic_hinges = int32_t(r[5]);
      o[b].y_min = -r[6];
      o[b].y_max = r[7];
    }
  });
}

int mmc_read_stats(mmc_ctx* c, mmc_state_stats* o) {
  if (c && c->n_inst != 1) {
    c->err = "mmc_read_stats needs a single-lattice context";
    return MMC_ERR_ARG;
  }
  return mmc_read_stats_batch(c, o);
}

int mmc_capture_mark(mmc_ctx* c) {
  return guarded(c, [&] {
    SyncLock sync_lock(g_sync_mu);
    if (!c->prev_x) {
      c->prev_x = dalloc<double>(3 * size_t(c->N));
      c->prev_q = dalloc<double>(4 * size_t(c->N));
    }
    launch_capture_mark(c->bpos, c->bquat, c->prev_x, c->prev_q, c->N, c->stream);
    HIPCALL(hipGetLastError());
  });
}

int mmc_capture(mmc_ctx* c, const double* tops, const int32_t* mask, const int32_t* velocities, mmc_snapshot* out) {
  return guarded(c, [&] {
    SyncLock sync_lock(g_sync_mu);
    if (!tops || !out) throw Fail(MMC_ERR_ARG, "NULL argument");
    std::vector<CapInst> ci;
    long total = 0;
    int per = 0;
    for (int b = 0; b < c->n_inst; b++) {
      if (mask && !mask[b]) continue;
      const bool vel = velocities && velocities[b];
      if (vel && !c->prev_x) throw Fail(MMC_ERR_STATE, "no previous poses for the velocities (mmc_capture_mark)");
      const Inst& in = c->hinst[size_t(b)];
      const long bodies = long(in.B) + in.N + MMC_MACHINE_BODIES;
      ci.push_back(CapInst{total, in.n0, in.N, in.b0, in.B, vel ? 1 : 0, tops[b],
                           1.0 / std::max(1e-6, c->icfg[size_t(b)].step_time), in.S.jr});
      total += bodies;
      per = std::max(per, int((bodies + 255) / 256));
    }
    if (ci.empty()) return;
    if (ci.size() > 65535) throw Fail(MMC_ERR_CAPACITY, "a capture takes at most 65,535 lattices");
    if (total > c->snap_cap) {
      if (c->snap) HIPCALL(hipFree(c->snap));
      c->snap = nullptr;
      c->snap = dalloc<mmc_snapshot>(size_t(total));
      c->snap_cap = total;
    }
    if (!c->capt) c->capt = dalloc<CapInst>(size_t(c->n_inst));
    HIPCALL(hipStreamSynchronize(c->stream));
    HIPCALL(hipMemcpy(c->capt, ci.data(), ci.size() * sizeof(CapInst), hipMemcpyHostToDevice));
    CapArgs a{c->capt, int(ci.size()), c->N, c->B, per, c->bpos, c->bquat, c->prev_x, c->prev_q, c->ba, c->bb,
              c->rad, c->plastic, c->nlayer, c->blayer, c->bfam, c->snap};
    launch_capture(a, c->stream);
    HIPCALL(hipGetLastError());
    c->read(out, c->snap, size_t(total) * sizeof(mmc_snapshot));
  });
}

int mmc_kernel_timer(mmc_ctx* c, int32_t kernel) {
  return guarded(c, [&] {
    if (kernel > MMC_K_PAIR) throw Fail(MMC_ERR_ARG, "unknown kernel id");
    if (!c->ev0) {
      HIPCALL(hipEventCreate(&c->ev0));
      HIPCALL(hipEventCreate(&c->ev1));
    }
    c->timed = kernel < 0 ? -1 : kernel;
    c->timer_pending = false;
    c->timed_launches = 0;
    c->timed_ms = 0;
  });
}

int mmc_kernel_timer_read(mmc_ctx* c, int64_t* launches, double* mean_ms) {
  return guarded(c, [&] {
    if (launches) *launches = c->timed_launches;
    if (mean_ms) *mean_ms = c->timed_launches ? c->timed_ms / double(c->timed_launches) : 0.0;
  });
}

int64_t mmc_evaluations(const mmc_ctx* c) { return c ? c->total_evaluations : 0; }

int mmc_time_kernel(mmc_ctx* c, int32_t kernel, int32_t launches, double* mean_ms) {
  return guarded(c, [&] {
    SyncLock sync_lock(g_sync_mu);
    if (launches <= 0 || !mean_ms) throw Fail(MMC_ERR_ARG, "launches must be positive");
    if (!c->ev0) {
      HIPCALL(hipEventCreate(&c->ev0));
      HIPCALL(hipEventCreate(&c->ev1));
    }
    if (kernel >= MMC_K_DSTATE_G && kernel <= MMC_K_DECIDE_G) {
      *mean_ms = c->time_loop_kernel(kernel, launches);
      return;
    }
    if (kernel == MMC_K_COPY) {
      constexpr long words = 1L << 21;  // 16 MiB
      if (!c->cal0) {
        c->cal0 = dalloc<double>(words);
        c->cal1 = dalloc<double>(words);
        HIPCALL(hipMemset(c->cal0, 0, words * sizeof(double)));
      }
      HIPCALL(hipEventRecord(c->ev0, c->stream));
      for (int i = 0; i < launches; i++)
        k_copy_words<<<unsigned((words + kBlock - 1) / kBlock), kBlock, 0, c->stream>>>(c->cal1, c->cal0, words);
      HIPCALL(hipEventRecord(c->ev1, c->stream));
      HIPCALL(hipEventSynchronize(c->ev1));
      HIPCALL(hipGetLastError());
      float ms = 0;
      HIPCALL(hipEventElapsedTime(&ms, c->ev0, c->ev1));
      *mean_ms = double(ms) / launches;
      return;
    }
    auto run = [&](auto Lt, auto* tag) {
      using T = std::remove_pointer_t<decltype(tag)>;
      double* R = static_cast<double*>(c->R);
      HIPCALL(hipEventRecord(c->ev0, c->stream));
      for (int i = 0; i < launches; i++) {
        switch (kernel) {
          case MMC_K_BEAM:
            k_beam<T><<<c->Pb, kBlock, 0, c->stream>>>(Lt, c->inst_d, c->tb, c->x, R, c->rec, nullptr,
                                                       TermBufs{nullptr, nullptr}, c->bpart);
            break;
          case MMC_K_CONTACT:
            k_contact<T><<<c->Pc, kBlock, 0, c->stream>>>(Lt, c->inst_d, c->tc, c->x, c->pcount, c->plist, c->slab,
                                                          c->cf_sph, c->cf_cap, TermBufs{nullptr, nullptr},
                                                          c->cpart);
            break;
          case MMC_K_GATHER:
            k_gather<T><<<c->Pd, kElemThreads, 0, c->stream>>>(Lt, c->inst_d, c->td, c->tops_d, c->x, c->rec,
                                                               c->cf_sph, c->cf_cap, c->spart, c->Ps, c->gc, nullptr,
                                                               c->gpart);
            break;
          case MMC_K_STATE:
            k_state<T><<<c->Ps, kBlock, 0, c->stream>>>(Lt, c->inst_d, c->tn, c->tops_d, c->z, nullptr, 0.0,
                                                        nullptr, c->x, R, nullptr, TermBufs{nullptr, nullptr},
                                                        c->spart);
            break;
          default:
            throw Fail(MMC_ERR_ARG, "kernel id cannot be timed in isolation");
        }
      }
      HIPCALL(hipEventRecord(c->ev1, c->stream));
      HIPCALL(hipEventSynchronize(c->ev1));
      HIPCALL(hipGetLastError());
      float ms = 0;
      HIPCALL(hipEventElapsedTime(&ms, c->ev0, c->ev1));
      *mean_ms = double(ms) / launches;
    };
    if (c->precision == MMC_MIXED)
      run(c->lat32(), (float*)nullptr);
    else
      run(c->lat64(), (double*)nullptr);
  });
}

int mmc_loop_shape(mmc_ctx* c, int32_t* decide_blocks, double* decide_us) {
  return guarded(c, [&] {
    if (decide_blocks) *decide_blocks = c->dec_split;
    if (decide_us)
      for (int k = 0; k < 6; k++) decide_us[k] = c->tune_dec_us[k];
  });
}

int mmc_beam_potential(mmc_ctx* c, const double* xh, const double* qh, double* out) {
  return guarded(c, [&] {
    c->need_single("mmc_beam_potential");
    if (!xh || !qh || !out) throw Fail(MMC_ERR_ARG, "NULL argument");
    const int N = c->N;
    std::vector<double> s(4 * size_t(N));
    soa_from_aos(xh, s.data(), N, 3);
    HIPCALL(hipMemcpyAsync(c->x, s.data(), 3 * size_t(N) * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCALL(hipStreamSynchronize(c->stream));
    soa_from_aos(qh, s.data(), N, 4);
    HIPCALL(hipMemcpyAsync(c->zin, s.data(), 4 * size_t(N) * sizeof(double), hipMemcpyHostToDevice, c->stream));
    k_beam_potential<double><<<c->Pb, kBlock, 0, c->stream>>>(c->lat64(), c->inst_d, c->tb, c->x, c->zin,
                                                              c->bpart);
    k_sum1<<<1, kBlock, 0, c->stream>>>(c->bpart, c->Pb, c->scal);
    c->read(c->host + 48, c->scal, sizeof(double));
    *out = c->host[48];
  });
}

int mmc_beam_potential_noalloc(mmc_ctx* c, const double* x, const double* q, double* energy, void* stream) {
  return guarded(c, [&] {
    c->need_single("mmc_beam_potential_noalloc");
    if (!x || !q || !energy) throw Fail(MMC_ERR_ARG, "NULL argument");
    if (c->Pb > kStaticScratchDoubles)
      throw Fail(MMC_ERR_CAPACITY, "mmc_beam_potential_noalloc: " + std::to_string(c->B) +
                                       " beams need " + std::to_string(c->Pb) + " partials; the static scratch holds " +
                                       std::to_string(kStaticScratchDoubles) + " (16,777,216 beams)");
    int dev = -1;
    HIPCALL(hipGetDevice(&dev));
    if (dev != c->device)
      throw Fail(MMC_ERR_ARG, "mmc_beam_potential_noalloc: device " + std::to_string(dev) +
                                  " is current, the context lives on device " + std::to_string(c->device));
    double* part = g_scratch_ptr[c->device];  // taken at mmc_create on this device
    if (!part) throw Fail(MMC_ERR_STATE, "static scratch not initialised on this device");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    k_beam_potential<double><<<c->Pb, kBlock, 0, st>>>(c->lat64(), c->inst_d, c->tb, x, q, part);
    k_sum1<<<1, kBlock, 0, st>>>(part, c->Pb, energy);
    HIPCALL(hipGetLastError());
  });
}

}  // extern "C"
