// kernels.hpp — device data layout and kernels of the gfx950 equilibrium solver.
//
// Layout in HBM (all structure-of-arrays, one plane per component):
//   nodes   base position bx|by|bz (FP64), base orientation bq w|x|y|z (FP64),
//           mass, lock bitmask (bit d locks DOF d)
//   state   x (3 planes FP64) and R (4 planes, element type T) of the current
//           evaluation point, produced once per objective evaluation
//   beams   a, b, L0, radius, area, I, e0 (3 planes), plastic (7 planes FP64)
//   vectors z, grad, L-BFGS s_h / y_h: 6 DOF planes of N doubles each
//   records per-beam element output (9 planes: chord force f, rotational
//           gradients of ends a and b), per-shape contact forces
// Gradients are assembled per node through a CSR beam-incidence list in a fixed
// order: no floating-point atomics anywhere, so every evaluation is
// bit-reproducible.  Reductions are two-level (wave shuffle + LDS per block,
// fixed-order sum of block partials in the consumer kernel).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "devmath.hpp"

namespace mmc {

constexpr int kBlock = 256;
constexpr int kWave = 64;
// element-parallel blocks over nodes: 6 waves (one per DOF) of 64 nodes
constexpr int kElemNodes = kWave, kElemThreads = 6 * kElemNodes;

struct Material {
  double E, G, sy, H, hb_frac, ht_frac, shear, density;
};

// Read-only lattice description on the device.
template <class T>
struct Lat {
  int N, B;
  const double* bpos;   // [3][N]
  const double* bquat;  // [4][N]
  const double* mass;   // [N]
  const uint8_t* lock;  // [N]
  const int* ba;        // [B]
  const int* bb;        // [B]
  const double* L0;     // [B]
  const double* rad;    // [B]
  const double* area;   // [B]
  const double* inert;  // [B]
  const double* e0;     // [3][B]
  const double* plastic;  // [7][B]
  const int* inc_off;     // [N+1]
  const int* inc;         // beam*2 + side
  const int* adj_off;     // [N+1]
  const int* adj;         // sorted neighbour lists (contact exclusion)
  const int* eslot;       // [2B] slot of each beam end (beam*2 + side) in the node-ordered records
};

// Per-evaluation scalars handed between kernels.
struct Scalars {
  double ts;          // translation scale
  double jr;          // joint sphere radius
  double gravity;
  double kc;          // contact stiffness
  double margin;
  double centering;
  double top;         // platen position
  double cell;        // broadphase cell
  double inv_cell;
  double skin;        // Verlet skin of the contact pair list
  int slab;           // partner slots per shape in the pair list
  unsigned table_mask;
  int entry_cap;
};

// One lattice of a batch (a single lattice is a batch of one).  All lattices
// share one concatenated SoA; an instance owns contiguous node, beam and shape
// ranges (its shapes are its spheres, then its capsules), and contiguous block
// ranges in every block table, so that every block partial belongs to exactly
// one instance.
struct Inst {
  Scalars S;                  // per-instance constants (S.top is unused: the platen is in the control block)
  Material M;
  double n_free, mt, mr;      // free DOFs; direction clamps in z units
  int n0, N, b0, B, s0, NS;   // node, beam, shape ranges
  int pn0, pn1;               // blocks of kBlock nodes
  int pb0, pb1;               // blocks of kBlock beams
  int pc0, pc1;               // blocks of kBlock shapes
  int pd0, pd1;               // blocks of kWave nodes (element-parallel vector kernels)
};
// One entry of a block table: the block covers items [first, first + count)
// of instance inst.
struct Blk {
  int inst, first, count, pad;
};

// ---- block reduction helpers (wave64 butterfly + LDS across waves) ----
// Lane exchanges without the LDS crossbar: DPP moves inside rows of 16 lanes
// (quad permutes for lane ^ 1 and lane ^ 2, the half-row mirror for lane ^ 7,
// the row mirror for lane ^ 15) and, across rows, gfx950's permlane swaps
// (permlane16: rows 0 <-> 1 and 2 <-> 3; permlane32: the two 32-lane halves).
// Callers run them with all 64 lanes active.
constexpr int kDppXor1 = 0xB1;   // quad_perm:[1,0,3,2]
constexpr int kDppXor2 = 0x4E;   // quad_perm:[2,3,0,1]
constexpr int kDppXor7 = 0x141;  // row_half_mirror
constexpr int kDppXor15 = 0x140; // row_mirror
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = unsigned(__builtin_amdgcn_update_dpp(0, int(b), CTRL, 0xF, 0xF, false));
  const unsigned hi = unsigned(__builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, 0xF, 0xF, false));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}
// (x, y) -> (x', y'), W = 16 or 32: lanes whose bit W is clear get x' = x, y'
// = x of lane + W; lanes whose bit W is set get x' = y of lane - W, y' = y.
// x' + y' is then, per lane, the sum of row r (bit clear: x) or row r + W
// (bit set: y) over the lane and its partner lane ^ W.
template <int W>
__device__ __forceinline__ void lane_swap_f64(double& x, double& y) {
  const long long bx = __double_as_longlong(x), by = __double_as_longlong(y);
  unsigned xl = unsigned(bx), xh = unsigned(bx >> 32), yl = unsigned(by), yh = unsigned(by >> 32);
  if constexpr (W == 32) {
    const auto l = __builtin_amdgcn_permlane32_swap(xl, yl, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(xh, yh, false, false);
    xl = l[0], yl = l[1], xh = h[0], yh = h[1];
  } else {
    static_assert(W == 16, "lane_swap_f64: W is 16 or 32");
    const auto l = __builtin_amdgcn_permlane16_swap(xl, yl, false, false);
    const auto h = __builtin_amdgcn_permlane16_swap(xh, yh, false, false);
    xl = l[0], yl = l[1], xh = h[0], yh = h[1];
  }
  x = __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(xh) << 32) | xl));
  y = __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(yh) << 32) | yl));
}
// v of lane `lane` (wave-uniform) in every lane: two scalar lane reads, no LDS
__device__ __forceinline__ double read_lane(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(b), lane));
  const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(b >> 32), lane));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}
// Sum over the wave, the same value in every lane: the 16-lane row sums by
// DPP butterfly, then the four rows in order ((r0 + r1) + (r2 + r3)).
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<kDppXor1>(v);
  v += dpp_f64<kDppXor2>(v);
  v += dpp_f64<kDppXor7>(v);
  v += dpp_f64<kDppXor15>(v);
  return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}
__device__ __forceinline__ double wave_max(double v) {
  v = fmax(v, dpp_f64<kDppXor1>(v));
  v = fmax(v, dpp_f64<kDppXor2>(v));
  v = fmax(v, dpp_f64<kDppXor7>(v));
  v = fmax(v, dpp_f64<kDppXor15>(v));
  return fmax(fmax(read_lane(v, 0), read_lane(v, 16)), fmax(read_lane(v, 32), read_lane(v, 48)));
}
// Reduces K values per thread over a block of NT threads; mode bit k set = max,
// else sum.  Result valid in thread 0.  Fixed combination order: deterministic.
template <int K, int NT = kBlock>
__device__ __forceinline__ void block_reduce(double (&v)[K], unsigned maxmask) {
  __shared__ double sh[K][NT / kWave];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = ((maxmask >> k) & 1) ? wave_max(v[k]) : wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; k++) sh[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      double a = sh[k][0];
      for (int i = 1; i < NT / kWave; i++) a = ((maxmask >> k) & 1) ? fmax(a, sh[k][i]) : a + sh[k][i];
      v[k] = a;
    }
  }
  __syncthreads();
}
// Sum (or max) of P block partials laid out [K][P], evaluated identically by
// every block of NT threads that calls it (all threads participate; result
// broadcast).
template <int K, int NT = kBlock>
__device__ __forceinline__ void reduce_partials(const double* part, int P, double (&out)[K], unsigned maxmask) {
  double v[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    bool mx = (maxmask >> k) & 1;
    double a = mx ? -INFINITY : 0.0;
    for (int i = threadIdx.x; i < P; i += NT) a = mx ? fmax(a, part[k * P + i]) : a + part[k * P + i];
    v[k] = a;
  }
  block_reduce<K, NT>(v, maxmask);
  __shared__ double bc[K];
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; k++) bc[k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++) out[k] = bc[k];
  __syncthreads();
}

// Sum (or max) of the partial columns [c0, c1) of K rows with row stride P,
// evaluated identically by every block of NT threads that calls it.
template <int K, int NT = kBlock>
__device__ __forceinline__ void reduce_columns(const double* part, int P, int c0, int c1, double (&out)[K],
                                               unsigned maxmask) {
  double v[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    bool mx = (maxmask >> k) & 1;
    double a = mx ? -INFINITY : 0.0;
    for (int i = c0 + threadIdx.x; i < c1; i += NT) a = mx ? fmax(a, part[k * P + i]) : a + part[k * P + i];
    v[k] = a;
  }
  block_reduce<K, NT>(v, maxmask);
  __shared__ double bc[K];
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; k++) bc[k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++) out[k] = bc[k];
  __syncthreads();
}

}  // namespace mmc
