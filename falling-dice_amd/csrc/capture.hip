// capture.hip — Simulator::capture (projects/metamaterial-compression/
// industry-grade/src/sim.cpp:215-310, MC/ below) on the device: every strut,
// joint and machine body of a committed frame, written in the byte layout of
// the META3 cache's DiskSnapshot (MC/src/main.cpp:19-22) by one thread per
// body, for many lattices per launch.
//
// The arithmetic is the reference's, operation for operation, with no fused
// multiply-adds (this file is built with -ffp-contract=off): positions, half
// extents, velocities and strut orientations (alignX, MC/src/sim.cpp:41, and
// quatFromAxes, MC/include/core.hpp:43, from sqrt, division, products and
// sums, all correctly rounded) equal the reference's bit for bit; angular
// velocities go through atan2 (quatLog, MC/include/core.hpp:52), whose device
// and host implementations may differ in the last place.
#include "capture.hpp"

namespace mmc {
namespace {

struct V3 {
  double x, y, z;
};
struct Q4 {
  double w, x, y, z;
};
__device__ __forceinline__ V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 scale(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// normalize, MC/include/core.hpp:39: division by the length as multiplication
// by its reciprocal; (0, 1, 0) for a degenerate or non-finite vector
__device__ __forceinline__ V3 normalize(V3 v) {
  const double l2 = dot(v, v);
  return (isfinite(l2) && l2 > 1e-30) ? scale(v, 1.0 / sqrt(l2)) : V3{0, 1, 0};
}
__device__ __forceinline__ Q4 qnormalized(Q4 q) {
  const double n = sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
  if (n < 1e-30) return Q4{1, 0, 0, 0};
  return {q.w / n, q.x / n, q.y / n, q.z / n};
}
__device__ __forceinline__ Q4 qmul(Q4 a, Q4 b) {
  return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
          a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x, a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}
// quatLog, MC/include/core.hpp:52
__device__ __forceinline__ V3 qlog(Q4 q) {
  q = qnormalized(q);
  if (q.w < 0) q = Q4{-q.w, -q.x, -q.y, -q.z};
  const double s = sqrt(q.x * q.x + q.y * q.y + q.z * q.z);
  if (s < 1e-12) return {2 * q.x, 2 * q.y, 2 * q.z};
  const double a = 2 * atan2(s, fmin(1.0, fmax(-1.0, q.w)));
  return scale(V3{q.x, q.y, q.z}, a / s);
}
// quatFromAxes, MC/include/core.hpp:43: the rotation with columns x, y, z
__device__ Q4 from_axes(V3 X, V3 Y, V3 Z) {
  const double m00 = X.x, m01 = Y.x, m02 = Z.x, m10 = X.y, m11 = Y.y, m12 = Z.y, m20 = X.z, m21 = Y.z, m22 = Z.z;
  Q4 q;
  const double tr = m00 + m11 + m22;
  if (tr > 0) {
    const double s = sqrt(tr + 1) * 2;
    q = {0.25 * s, (m21 - m12) / s, (m02 - m20) / s, (m10 - m01) / s};
  } else if (m00 > m11 && m00 > m22) {
    const double s = sqrt(1 + m00 - m11 - m22) * 2;
    q = {(m21 - m12) / s, 0.25 * s, (m01 + m10) / s, (m02 + m20) / s};
  } else if (m11 > m22) {
    const double s = sqrt(1 + m11 - m00 - m22) * 2;
    q = {(m02 - m20) / s, (m01 + m10) / s, 0.25 * s, (m12 + m21) / s};
  } else {
    const double s = sqrt(1 + m22 - m00 - m11) * 2;
    q = {(m10 - m01) / s, (m02 + m20) / s, (m12 + m21) / s, 0.25 * s};
  }
  return qnormalized(q);
}
// alignX, MC/src/sim.cpp:41 (orthogonal, MC/include/core.hpp:47)
__device__ Q4 align_x(V3 d) {
  const V3 X = normalize(d);
  const V3 h = fabs(X.z) < .8 ? V3{0, 0, 1} : V3{0, 1, 0};
  V3 Y = normalize(cross(h, X));
  const V3 Z = normalize(cross(X, Y));
  Y = normalize(cross(Z, X));
  return from_axes(X, Y, Z);
}

__device__ __forceinline__ V3 load3(const double* p, int stride, int i) {
  return {p[i], p[size_t(stride) + i], p[2 * size_t(stride) + i]};
}
__device__ __forceinline__ Q4 load4(const double* p, int stride, int i) {
  return {p[i], p[size_t(stride) + i], p[2 * size_t(stride) + i], p[3 * size_t(stride) + i]};
}

struct Motion {
  V3 v, w;
};
// nodeVelocity / nodeOmega, MC/src/sim.cpp:225-233, of the global node i
__device__ Motion node_motion(const CapArgs& A, const CapInst& c, int i) {
  if (!c.velocities) return Motion{V3{0, 0, 0}, V3{0, 0, 0}};
  const V3 x = load3(A.x, A.N, i), px = load3(A.px, A.N, i);
  const Q4 R = load4(A.q, A.N, i), pR = load4(A.pq, A.N, i);
  const Q4 dq = qnormalized(qmul(R, Q4{pR.w, -pR.x, -pR.y, -pR.z}));
  return Motion{scale(sub(x, px), c.inv_dt), scale(qlog(dq), c.inv_dt)};
}

__device__ __forceinline__ void put(mmc_snapshot& s, V3 p, Q4 q, V3 v, V3 w, V3 half) {
  s.p[0] = p.x, s.p[1] = p.y, s.p[2] = p.z;
  s.q[0] = q.w, s.q[1] = q.x, s.q[2] = q.y, s.q[3] = q.z;
  s.v[0] = v.x, s.v[1] = v.y, s.v[2] = v.z;
  s.w[0] = w.x, s.w[1] = w.y, s.w[2] = w.z;
  s.half[0] = half.x, s.half[1] = half.y, s.half[2] = half.z;
}

// Grid: blocks_per_inst x n; thread r of lattice blockIdx.y writes body r of
// that lattice: beams [0, B), nodes [B, B + N), machine bodies after them.
__global__ __launch_bounds__(256) void k_capture(CapArgs A) {
  const CapInst c = A.inst[blockIdx.y];
  const int r = int(blockIdx.x) * 256 + int(threadIdx.x);
  if (r >= c.B + c.N + MMC_MACHINE_BODIES) return;
  mmc_snapshot s;
  if (r < c.B) {
    // strut, MC/src/sim.cpp:236-265
    const int k = c.b0 + r;
    const int na = A.a[k], nb = A.b[k];
    const V3 pa = load3(A.x, A.N, na), pb = load3(A.x, A.N, nb);
    const V3 d = sub(pb, pa);
    const double L = fmax(1e-8, sqrt(dot(d, d)));
    const double pAxial = fabs(A.plastic[k]);
    const V3 rotp = load3(A.plastic + 2 * size_t(A.B), A.B, k);
    const double pRot = sqrt(dot(rotp, rotp));
    const double severity = fmax(pAxial / 0.10, pRot / 0.42);
    int material = 0;
    if (severity > 0.28) material = 1;
    if (severity > 0.62) material = 2;
    if (severity > 1.05) material = 3;
    const Motion ma = node_motion(A, c, na), mb = node_motion(A, c, nb);
    const double rad = A.radius[k];
    put(s, scale(add(pa, pb), 0.5), align_x(d), scale(add(ma.v, mb.v), 0.5), scale(add(ma.w, mb.w), 0.5),
        V3{0.5 * L, rad, rad});
    s.material = material;
    s.id = r;
    s.kind = 0;  // BodyKind::Strut
    s.module_class = A.beam_family[k];
    s.ring = A.beam_layer[k];
    s.slot = r;
  } else if (r < c.B + c.N) {
    // joint, MC/src/sim.cpp:267-281
    const int ni = r - c.B, i = c.n0 + ni;
    const Motion m = node_motion(A, c, i);
    put(s, load3(A.x, A.N, i), load4(A.q, A.N, i), m.v, m.w, V3{c.jr, c.jr, c.jr});
    s.material = 4;
    s.id = r;
    s.kind = 1;  // BodyKind::Joint
    s.module_class = A.node_layer[i];
    s.ring = A.node_layer[i];
    s.slot = ni;
  } else {
    // rigid test machine, MC/src/sim.cpp:284-307: platens at y = 0 and y =
    // top, base, the load frame's two columns and crosshead, the top fixture
    const int m = r - c.B - c.N;
    V3 p, half;
    int material = 6, kind = 3;  // BodyKind::Machine
    switch (m) {
      case 0: p = {0, -0.0030, 0}, half = {0.026, 0.0030, 0.026}, material = 5, kind = 2; break;
      case 1: p = {0, c.top + 0.0030, 0}, half = {0.026, 0.0030, 0.026}, material = 5, kind = 2; break;
      case 2: p = {0, -0.0095, 0}, half = {0.036, 0.0035, 0.036}; break;
      case 3: p = {-0.037, 0.054, -0.049}, half = {0.0026, 0.061, 0.0026}; break;
      case 4: p = {0.037, 0.054, -0.049}, half = {0.0026, 0.061, 0.0026}; break;
      case 5: p = {0, 0.111, -0.049}, half = {0.042, 0.0032, 0.0036}; break;
      default: p = {0, c.top + 0.0080, -0.001}, half = {0.0105, 0.0025, 0.0105}, material = 7; break;
    }
    put(s, p, Q4{1, 0, 0, 0}, V3{0, 0, 0}, V3{0, 0, 0}, half);
    s.material = material;
    s.id = r;
    s.kind = kind;
    s.module_class = -1;
    s.ring = -1;
    s.slot = -1;
  }
  A.out[c.out0 + r] = s;
}

__global__ __launch_bounds__(256) void k_capture_mark(const double* __restrict__ x, const double* __restrict__ q,
                                                      double* __restrict__ px, double* __restrict__ pq, long N) {
  const long i = long(blockIdx.x) * 256 + threadIdx.x;
  if (i >= N) return;
  for (int d = 0; d < 3; d++) px[d * N + i] = x[d * N + i];
  for (int d = 0; d < 4; d++) pq[d * N + i] = q[d * N + i];
}

}  // namespace

void launch_capture(const CapArgs& a, hipStream_t stream) {
  if (a.n == 0 || a.blocks_per_inst == 0) return;
  k_capture<<<dim3(unsigned(a.blocks_per_inst), unsigned(a.n)), 256, 0, stream>>>(a);
}

void launch_capture_mark(const double* x, const double* q, double* px, double* pq, long N, hipStream_t stream) {
  if (N <= 0) return;
  k_capture_mark<<<unsigned((N + 255) / 256), 256, 0, stream>>>(x, q, px, pq, N);
}

}  // namespace mmc
