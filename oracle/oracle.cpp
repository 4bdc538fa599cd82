// oracle.cpp — FP64 CPU oracle of the beam-lattice compression solver.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, the smoke check and the benchmark's
// cpu_baseline leg load this library, and only as the checker; the product
// (falling-dice_amd/) never links or calls it.
//
// It restates, single-threaded and in FP64, the reference solver of
// projects/metamaterial-compression/industry-grade (MC/ below):
//   math helpers        MC/include/core.hpp:30-53
//   return maps         MC/src/sim.cpp:5-8
//   contact geometry    MC/src/sim.cpp:9-39
//   L-BFGS helpers      MC/src/sim.cpp:42-46
//   lattice (BCC)       MC/src/sim.cpp:49-135
//   state/objective     MC/src/sim.cpp:136-172
//   minimize/commit     MC/src/sim.cpp:174-213
//   diagnostics/run     MC/src/sim.cpp:312-467
//   checks              MC/src/sim.cpp:468-577
// plus a simple-cubic lattice generator that the reference does not have.
//
// Every floating-point expression keeps the reference's operation order
// (including v/s evaluated as v*(1/s) for vectors), and the contact broadphase
// keeps the reference's hash-grid container and insertion sequence, so that the
// pinning tests can compare this oracle with the reference compiled from source
// bit for bit.  Pinning: tests/test_oracle_pin.py (against oracle/_ref and the
// golden fixtures in tests/golden/) and tests/test_oracle_math.py (closed forms).
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace {

constexpr double kPi = 3.1415926535897932384626433832795;
constexpr double kInf = std::numeric_limits<double>::infinity();

double clampd(double x, double lo, double hi) { return std::max(lo, std::min(hi, x)); }
double sat(double x) { return clampd(x, 0.0, 1.0); }
double mixd(double a, double b, double t) { return a * (1.0 - t) + b * t; }
double sqr(double x) { return x * x; }

struct V3 {
  double x = 0, y = 0, z = 0;
};
V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 scl(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
V3 divs(V3 a, double s) { return scl(a, 1.0 / s); }  // vector / scalar = times reciprocal
double dotv(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 crossv(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double len2(V3 a) { return dotv(a, a); }
double len(V3 a) { return std::sqrt(len2(a)); }
V3 unit(V3 v) {
  double l2 = len2(v);
  return (std::isfinite(l2) && l2 > 1e-30) ? divs(v, std::sqrt(l2)) : V3{0, 1, 0};
}
V3 vmin(V3 a, V3 b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
V3 vmax(V3 a, V3 b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }
void addto(V3& a, V3 b) { a = add(a, b); }
void subfrom(V3& a, V3 b) { a = sub(a, b); }

struct Q4 {
  double w = 1, x = 0, y = 0, z = 0;
};
Q4 qnorm(Q4 q) {
  double n = std::sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
  if (n < 1e-30) return Q4{};
  return {q.w / n, q.x / n, q.y / n, q.z / n};
}
Q4 qconj(Q4 q) { return {q.w, -q.x, -q.y, -q.z}; }
Q4 qmul(Q4 a, Q4 b) {
  return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
          a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x, a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}
// v' = v + w*t + u x t with t = 2 u x v  (MC/include/core.hpp:40)
V3 qrot(Q4 q, V3 v) {
  V3 u{q.x, q.y, q.z};
  V3 t = scl(crossv(u, v), 2.0);
  return add(add(v, scl(t, q.w)), crossv(u, t));
}
V3 qrotinv(Q4 q, V3 v) { return qrot(qconj(q), v); }
Q4 qaxisangle(V3 axis, double angle) {
  axis = unit(axis);
  double s = std::sin(0.5 * angle);
  return qnorm({std::cos(0.5 * angle), axis.x * s, axis.y * s, axis.z * s});
}
// MC/include/core.hpp:51
Q4 qexp(V3 v) {
  double a = len(v);
  if (a < 1e-12) return qnorm({1 - 0.125 * a * a, 0.5 * v.x, 0.5 * v.y, 0.5 * v.z});
  return qaxisangle(divs(v, a), a);
}
// MC/include/core.hpp:52
V3 qlog(Q4 q) {
  q = qnorm(q);
  if (q.w < 0) q = {-q.w, -q.x, -q.y, -q.z};
  double s = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z);
  if (s < 1e-12) return {2 * q.x, 2 * q.y, 2 * q.z};
  double a = 2 * std::atan2(s, clampd(q.w, -1, 1));
  return scl(V3{q.x, q.y, q.z}, a / s);
}
// J_l^{-T}(p) v = v + 1/2 p x v + A p x (p x v)   (MC/include/core.hpp:53)
V3 jlinvt(V3 p, V3 v) {
  double a = len(p);
  double A = (a < 1e-7) ? (1.0 / 12.0) : (1 / (a * a) - (1 + std::cos(a)) / (2 * a * std::sin(a)));
  return add(add(v, scl(crossv(p, v), 0.5)), scl(crossv(p, crossv(p, v)), A));
}

uint64_t mix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}
// Reference RNG stream (MC/include/core.hpp:46): state advances by mix64.
struct Stream {
  uint64_t s;
  explicit Stream(uint64_t seed) : s(mix64(seed)) {}
  uint32_t next32() {
    s = mix64(s);
    return uint32_t(s >> 32);
  }
  double uniform() { return (double(next32()) + 0.5) / 4294967296.0; }
  double symmetric() { return 2 * uniform() - 1; }
};
uint64_t node_pair(int a, int b) {
  if (a > b) std::swap(a, b);
  return (uint64_t(uint32_t(a)) << 32) | uint32_t(b);
}
int64_t grid_key(int x, int y, int z) {
  constexpr int64_t B = 1LL << 20, M = (1LL << 21) - 1;
  return ((int64_t(x) + B) & M) | (((int64_t(y) + B) & M) << 21) | (((int64_t(z) + B) & M) << 42);
}
struct GridHash {
  size_t operator()(int64_t k) const noexcept { return size_t(mix64(uint64_t(k))); }
};
V3 perpendicular(V3 n) {
  V3 h = std::abs(n.z) < 0.8 ? V3{0, 0, 1} : V3{0, 1, 0};
  return unit(crossv(h, n));
}

// ---- return maps (MC/src/sim.cpp:5-8) ----
struct Ret1 {
  double force = 0, p = 0, a = 0, dg = 0, energy = 0, hard = 0;
  bool yielded = false;
};
Ret1 ret1(double total, double pc, double ac, double k, double y, double H) {
  Ret1 r;
  double trial = k * (total - pc);
  double f = std::abs(trial) - (y + H * ac);
  r.p = pc;
  r.a = ac;
  if (f > 0) {
    double sgn = trial >= 0 ? 1 : -1;
    r.dg = f / std::max(1e-20, k + H);
    r.p = pc + sgn * r.dg;
    r.a = ac + r.dg;
    r.yielded = true;
  }
  double el = total - r.p;
  r.force = k * el;
  r.energy = 0.5 * k * (el * el);
  r.hard = 0.5 * H * (r.a * r.a);
  return r;
}
struct Ret3 {
  V3 force, p;
  double a = 0, dg = 0, energy = 0, hard = 0;
  bool yielded = false;
};
Ret3 ret3(V3 t, V3 pc, double ac, double k, double y, double H) {
  Ret3 r;
  V3 trial = scl(sub(t, pc), k);
  double n = len(trial);
  double f = n - (y + H * ac);
  r.p = pc;
  r.a = ac;
  if (f > 0 && n > 1e-20) {
    V3 dir = divs(trial, n);
    r.dg = f / std::max(1e-20, k + H);
    r.p = add(pc, scl(dir, r.dg));
    r.a = ac + r.dg;
    r.yielded = true;
  }
  V3 el = sub(t, r.p);
  r.force = scl(el, k);
  r.energy = 0.5 * k * len2(el);
  r.hard = 0.5 * H * (r.a * r.a);
  return r;
}

// ---- contact geometry (MC/src/sim.cpp:9-39) ----
struct Shape {
  V3 a, b;
  double r = 0;
  int n0 = -1, n1 = -1, beam = -1;
  bool capsule = false;
};
struct Closest {
  double s = 0, t = 0;
  V3 pa, pb;
};
// Segment-segment closest points (Ericson, Real-Time Collision Detection 5.1.9).
Closest seg_seg(V3 p1, V3 q1, V3 p2, V3 q2) {
  V3 d1 = sub(q1, p1), d2 = sub(q2, p2), r = sub(p1, p2);
  double a = dotv(d1, d1), e = dotv(d2, d2), f = dotv(d2, r), s = 0, t = 0;
  if (a <= 1e-20 && e <= 1e-20) return {0, 0, p1, p2};
  if (a <= 1e-20) {
    s = 0;
    t = sat(f / e);
  } else {
    double c = dotv(d1, r);
    if (e <= 1e-20) {
      t = 0;
      s = sat(-c / a);
    } else {
      double b = dotv(d1, d2), den = a * e - b * b;
      s = std::abs(den) > 1e-20 ? sat((b * f - c * e) / den) : 0;
      t = (b * s + f) / e;
      if (t < 0) {
        t = 0;
        s = sat(-c / a);
      } else if (t > 1) {
        t = 1;
        s = sat((b - c) / a);
      }
    }
  }
  return {s, t, add(p1, scl(d1, s)), add(p2, scl(d2, t))};
}
Closest shape_pair(const Shape& A, const Shape& B) {
  if (A.capsule && B.capsule) return seg_seg(A.a, A.b, B.a, B.b);
  if (A.capsule) {
    V3 d = sub(A.b, A.a);
    double s = sat(dotv(sub(B.a, A.a), d) / std::max(1e-30, len2(d)));
    return {s, 0, add(A.a, scl(d, s)), B.a};
  }
  if (B.capsule) {
    V3 d = sub(B.b, B.a);
    double t = sat(dotv(sub(A.a, B.a), d) / std::max(1e-30, len2(d)));
    return {0, t, A.a, add(B.a, scl(d, t))};
  }
  return {0, 0, A.a, B.a};
}
void spread(std::vector<V3>& g, const Shape& s, double u, V3 v) {
  if (s.capsule) {
    addto(g[size_t(s.n0)], scl(v, 1 - u));
    addto(g[size_t(s.n1)], scl(v, u));
  } else {
    addto(g[size_t(s.n0)], v);
  }
}

struct Plastic {
  double epsp = 0, alpha = 0;
  V3 rotp;
  double alphaB = 0, alphaT = 0;
};
struct Trial {
  Plastic next;
  double strain = 0, stress = 0;
  V3 moment;
  double elastic = 0, dissipation = 0;
  bool axial = false, bend = false, torsion = false;
};
struct Node {
  V3 x0, x;
  Q4 R;
  double mass = 0;
  int layer = 0;
};
struct Beam {
  int a = -1, b = -1, layer = 0, family = 0;
  double L0 = 0, radius = 0.001, area = 0, I = 0, J = 0;
  V3 e0{1, 0, 0};
  Plastic plastic;
};
struct Eval {
  double energy = 0, elastic = 0, plastic = 0, contact = 0, reaction = 0, maxPen = 0;
  int contacts = 0;
  std::vector<double> grad;
  std::vector<Trial> trial;
  // line_search 1 only: per-item energies (nodes [0,N), beams [N,N+B), shapes
  // from N+B, a pair's energy held by its lower shape) and the centering energy
  std::vector<double> terms;
  double ecen = 0;
};
struct Stats {
  int iterations = 0, backtracks = 0, evaluations = 0;
  double rms = kInf, max = kInf, energy = kInf;
  bool converged = false;
};

// thrown by run() when the frame observer asks to stop: not an equilibrium failure
struct RunStopped {};

}  // namespace

struct orc {
  mmc_config cfg;
  std::vector<Node> nodes;
  std::vector<Beam> beams;
  std::unordered_set<uint64_t> connected;
  std::vector<uint64_t> connected_list;  // insertion order, for export
  std::vector<uint8_t> locked;
  std::vector<V3> baseX;
  std::vector<Q4> baseR;
  double width = 0, height = 0, depth = 0, area = 0, mass = 0, top0 = 0;
  double loadW = 0, recW = 0, prevDisp = 0, prevForce = 0, tare = 0;
  Eval last;  // final objective of the last minimize
  std::vector<double> lastZ;
  bool have_last = false;
  std::string err;
  int evaluations = 0;
  // optional per-frame observer of run(): returns nonzero to end the run there
  int (*frame_cb)(const mmc_frame*, int32_t, void*) = nullptr;
  void* frame_cb_user = nullptr;

  double jr() const { return cfg.joint_scale * cfg.max_radius; }
  void add_connected(uint64_t k) {
    if (connected.insert(k).second) connected_list.push_back(k);
  }

  // ---- graded BCC lattice, MC/src/sim.cpp:51-132 ----
  void build_bcc() {
    int nx = cfg.cells_x, ny = cfg.cells_y, nz = cfg.cells_z;
    int corners = (nx + 1) * (ny + 1) * (nz + 1);
    auto corner = [&](int i, int j, int k) { return (j * (nz + 1) + k) * (nx + 1) + i; };
    auto centre = [&](int i, int j, int k) { return corners + (j * nz + k) * nx + i; };
    double h = cfg.cell_size;
    width = nx * h;
    height = ny * h;
    depth = nz * h;
    area = width * depth;
    double J = jr();
    V3 origin{-0.5 * width, J, -0.5 * depth};
    Stream rng(cfg.seed);
    for (int j = 0; j <= ny; j++)
      for (int k = 0; k <= nz; k++)
        for (int i = 0; i <= nx; i++) {
          Node n;
          n.x0 = add(origin, V3{i * h, j * h, k * h});
          double yf = double(j) / ny, env = std::sin(kPi * yf);
          double sg = ((i + 2 * k + 3 * j) & 1) ? 1.0 : -1.0;
          n.x0.x += cfg.imperfection * env * (0.55 * sg + 0.45 * rng.symmetric());
          n.x0.z += cfg.imperfection * env * (0.35 * sg + 0.65 * rng.symmetric());
          n.x = n.x0;
          n.layer = j;
          nodes.push_back(n);
        }
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++)
        for (int i = 0; i < nx; i++) {
          Node n;
          n.x0 = add(origin, V3{(i + 0.5) * h, (j + 0.5) * h, (k + 0.5) * h});
          double yf = (j + 0.5) / double(ny), env = std::sin(kPi * yf);
          double sg = ((i + k + j) & 1) ? 1.0 : -1.0;
          double sx = (j & 1) ? -1.0 : 1.0, sz = ((j / 2) & 1) ? -1.0 : 1.0;
          n.x0.x += h * (0.082 * (1 - 0.32 * yf) * sx + 0.018 * env * sg) +
                    1.2 * cfg.imperfection * env * rng.symmetric();
          n.x0.z += h * (0.032 * (0.65 + 0.35 * (1 - yf)) * sz - 0.012 * env * sg) +
                    1.2 * cfg.imperfection * env * rng.symmetric();
          n.x = n.x0;
          n.layer = j;
          nodes.push_back(n);
        }
    std::unordered_set<uint64_t> struts;
    auto strut_radius = [&](int a, int b, int layer, double scale) {
      V3 m = scl(add(nodes[size_t(a)].x0, nodes[size_t(b)].x0), 0.5);
      double y = sat((m.y - J) / height);
      double radial = sat(std::sqrt(m.x * m.x + m.z * m.z) / (0.72 * std::max(width, depth)));
      double grade = 0.58 + 0.42 * std::pow(1 - y, 0.72);
      grade *= 0.94 + 0.12 * radial;
      if (layer == 0)
        grade *= cfg.trigger_scale;
      else if (layer == 1)
        grade *= std::sqrt(cfg.trigger_scale);
      else if (layer == 2)
        grade *= 0.88;
      return scale * mixd(cfg.min_radius, cfg.max_radius, sat(grade));
    };
    auto strut = [&](int a, int b, int layer, int family, double scale) {
      uint64_t key = node_pair(a, b);
      if (a == b || !struts.insert(key).second) return;
      V3 pa = nodes[size_t(a)].x0, pb = nodes[size_t(b)].x0;
      V3 tangent = unit(sub(pb, pa));
      uint64_t hsh = mix64(cfg.seed ^ key ^ uint64_t(family + 31) * 0x9e3779b97f4a7c15ULL);
      V3 helper{double(int((hsh >> 0) & 1023) - 511), double(int((hsh >> 10) & 1023) - 511),
                double(int((hsh >> 20) & 1023) - 511)};
      V3 ax1 = crossv(tangent, helper);
      ax1 = len2(ax1) < 1e-12 ? perpendicular(tangent) : unit(ax1);
      V3 ax2 = unit(crossv(tangent, ax1));
      double lf = layer == 0 ? 2.60 : (layer == 1 ? 1.60 : 1.0);
      double amp = cfg.member_imperfection * lf * (0.55 + 0.90 * double((hsh >> 32) & 65535) / 65535.0);
      int segs = std::clamp(cfg.elements_per_strut, 2, 5);
      std::vector<int> chain{a};
      for (int s = 1; s < segs; s++) {
        double t = double(s) / segs;
        Node mid;
        mid.x0 = add(add(add(scl(pa, 1 - t), scl(pb, t)), scl(ax1, std::sin(kPi * t) * amp)),
                     scl(ax2, 0.22 * std::sin(2 * kPi * t) * amp));
        mid.x = mid.x0;
        mid.layer = layer;
        chain.push_back(int(nodes.size()));
        nodes.push_back(mid);
      }
      chain.push_back(b);
      double r = strut_radius(a, b, layer, scale);
      for (size_t s = 0; s + 1 < chain.size(); s++) push_beam(chain[s], chain[s + 1], layer, family, r);
      add_connected(key);
    };
    const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                           {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++)
        for (int i = 0; i < nx; i++) {
          int c = centre(i, j, k);
          for (int q = 0; q < 8; q++) strut(c, corner(i + off[q][0], j + off[q][1], k + off[q][2]), j, q, 1.0);
        }
    for (int face : {0, ny}) {
      int lay = std::clamp(face, 0, ny - 1);
      for (int k = 0; k <= nz; k++)
        for (int i = 0; i < nx; i++) strut(corner(i, face, k), corner(i + 1, face, k), lay, 30, 0.62);
      for (int i = 0; i <= nx; i++)
        for (int k = 0; k < nz; k++) strut(corner(i, face, k), corner(i, face, k + 1), lay, 31, 0.62);
    }
    top0 = J + height + J;
  }

  void push_beam(int u, int v, int layer, int family, double r) {
    Beam e;
    e.a = u;
    e.b = v;
    e.layer = layer;
    e.family = family;
    V3 d = sub(nodes[size_t(v)].x0, nodes[size_t(u)].x0);
    e.L0 = len(d);
    e.e0 = unit(d);
    e.radius = r;
    e.area = kPi * r * r;
    e.I = 0.25 * kPi * std::pow(r, 4);
    e.J = 2 * e.I;
    beams.push_back(e);
    add_connected(node_pair(u, v));
  }

  // ---- simple cubic lattice (no reference counterpart) ----
  // Nodes on the (nx+1)(ny+1)(nz+1) grid in the reference's corner order; one
  // beam per cell edge.  Imperfection: Gaussian jitter of std position_sigma
  // (x,z everywhere, y on interior layers) by Box-Muller on the reference RNG,
  // or the reference's corner-node imperfection when position_sigma == 0.
  void build_cubic() {
    int nx = cfg.cells_x, ny = cfg.cells_y, nz = cfg.cells_z;
    double h = cfg.cell_size;
    width = nx * h;
    height = ny * h;
    depth = nz * h;
    area = width * depth;
    double J = jr();
    V3 origin{-0.5 * width, J, -0.5 * depth};
    Stream rng(cfg.seed);
    auto gauss2 = [&](double& g0, double& g1) {
      double u1 = rng.uniform(), u2 = rng.uniform();
      double rr = std::sqrt(-2.0 * std::log(u1));
      g0 = rr * std::cos(2 * kPi * u2);
      g1 = rr * std::sin(2 * kPi * u2);
    };
    for (int j = 0; j <= ny; j++)
      for (int k = 0; k <= nz; k++)
        for (int i = 0; i <= nx; i++) {
          Node n;
          n.x0 = add(origin, V3{i * h, j * h, k * h});
          if (cfg.position_sigma > 0) {
            double g0, g1, g2, g3;
            gauss2(g0, g1);
            gauss2(g2, g3);
            n.x0.x += cfg.position_sigma * g0;
            n.x0.z += cfg.position_sigma * g1;
            if (j > 0 && j < ny) n.x0.y += cfg.position_sigma * g2;
          } else {
            double yf = double(j) / ny, env = std::sin(kPi * yf);
            double sg = ((i + 2 * k + 3 * j) & 1) ? 1.0 : -1.0;
            n.x0.x += cfg.imperfection * env * (0.55 * sg + 0.45 * rng.symmetric());
            n.x0.z += cfg.imperfection * env * (0.35 * sg + 0.65 * rng.symmetric());
          }
          n.x = n.x0;
          n.layer = j;
          nodes.push_back(n);
        }
    auto id = [&](int i, int j, int k) { return (j * (nz + 1) + k) * (nx + 1) + i; };
    std::vector<double> radii;
    for (int j = 0; j <= ny; j++)
      for (int k = 0; k <= nz; k++)
        for (int i = 0; i <= nx; i++) {
          int c = id(i, j, k);
          int nb[3] = {i < nx ? id(i + 1, j, k) : -1, j < ny ? id(i, j + 1, k) : -1, k < nz ? id(i, j, k + 1) : -1};
          for (int f = 0; f < 3; f++) {
            if (nb[f] < 0) continue;
            double r;
            if (cfg.radius_mode == 1) {
              r = cfg.min_radius + (cfg.max_radius - cfg.min_radius) * rng.uniform();
            } else {
              V3 m = scl(add(nodes[size_t(c)].x0, nodes[size_t(nb[f])].x0), 0.5);
              r = mixd(cfg.min_radius, cfg.max_radius, sat((m.y - J) / height));
            }
            push_beam(c, nb[f], std::min(j, ny - 1), f, r);
          }
        }
    top0 = J + height + J;
  }

  // MC/src/sim.cpp:134
  void distribute_mass() {
    mass = 0;
    for (Node& n : nodes) n.mass = 0;
    for (const Beam& e : beams) {
      double m = cfg.material.density * e.area * e.L0;
      nodes[size_t(e.a)].mass += 0.5 * m;
      nodes[size_t(e.b)].mass += 0.5 * m;
      mass += m;
    }
    double J = jr(), jm = cfg.material.density * (4.0 / 3.0) * kPi * std::pow(J, 3) * 0.28;
    for (Node& n : nodes) {
      n.mass += jm;
      mass += jm;
    }
    locked.assign(nodes.size() * 6, 0);
    int a = -1, b = -1;
    double da = kInf, db = kInf;
    for (int i = 0; i < int(nodes.size()); i++) {
      if (nodes[size_t(i)].layer != 0) continue;
      V3 p = nodes[size_t(i)].x0;
      double qa = sqr(p.x + 0.5 * width) + sqr(p.z + 0.5 * depth);
      double qb = sqr(p.x - 0.5 * width) + sqr(p.z + 0.5 * depth);
      if (qa < da) {
        da = qa;
        a = i;
      }
      if (qb < db) {
        db = qb;
        b = i;
      }
    }
    if (a < 0 || b < 0) throw std::runtime_error("failed to create anti-drift constraints");
    locked[size_t(6 * a)] = 1;
    locked[size_t(6 * a + 2)] = 1;
    locked[size_t(6 * b + 2)] = 1;
  }

  void init_base() {
    baseX.resize(nodes.size());
    baseR.resize(nodes.size());
    for (size_t i = 0; i < nodes.size(); i++) {
      baseX[i] = nodes[i].x;
      baseR[i] = nodes[i].R;
    }
  }

  // MC/src/sim.cpp:136
  void state(const std::vector<double>& z, std::vector<V3>& x, std::vector<Q4>& R) const {
    x.resize(nodes.size());
    R.resize(nodes.size());
    double ts = cfg.translation_scale;
    for (size_t i = 0; i < nodes.size(); i++) {
      size_t o = 6 * i;
      double t[3], r[3];
      for (int d = 0; d < 3; d++) {
        t[d] = locked[o + size_t(d)] ? 0.0 : z[o + size_t(d)];
        r[d] = locked[o + size_t(d) + 3] ? 0.0 : z[o + size_t(d) + 3];
      }
      x[i] = add(baseX[i], scl(V3{t[0], t[1], t[2]}, ts));
      R[i] = qnorm(qmul(qexp(V3{r[0], r[1], r[2]}), baseR[i]));
    }
  }

  // MC/src/sim.cpp:138-172
  Eval objective(const std::vector<double>& z, double top, bool trials) {
    evaluations++;
    Eval o;
    size_t N = nodes.size();
    o.grad.assign(N * 6, 0);
    if (trials) o.trial.resize(beams.size());
    const bool per_term = cfg.line_search == 1;
    const size_t NB = N + beams.size();
    if (per_term) o.terms.assign(2 * NB, 0.0);  // N + B items, then N + B shapes
    std::vector<V3> x, gx(N), gr(N);
    std::vector<Q4> R;
    state(z, x, R);
    const mmc_material& M = cfg.material;
    double E = M.young, G = E / (2 * (1 + M.nu)), sy = M.yield, H = M.hardening;
    for (size_t bi = 0; bi < beams.size(); bi++) {
      const Beam& e = beams[bi];
      size_t ia = size_t(e.a), ib = size_t(e.b);
      V3 d = sub(x[ib], x[ia]);
      double L = len(d);
      if (L < 1e-10) {
        o.energy += 1e8 * sqr(1e-10 - L);
        if (per_term) o.terms[N + bi] = 1e8 * sqr(1e-10 - L);
        continue;
      }
      V3 t = divs(d, L);
      Trial tr;
      double eps = (L - e.L0) / e.L0;
      Ret1 ax = ret1(eps, e.plastic.epsp, e.plastic.alpha, E, sy, H);
      tr.strain = eps;
      tr.stress = ax.force;
      tr.next = e.plastic;
      tr.next.epsp = ax.p;
      tr.next.alpha = ax.a;
      tr.axial = ax.yielded;
      double uax = e.area * e.L0 * (ax.energy + ax.hard), dax = e.area * e.L0 * sy * ax.dg;
      o.energy += uax + dax;
      o.elastic += uax;
      o.plastic += dax;
      double Nf = e.area * ax.force;
      subfrom(gx[ia], scl(t, Nf));
      addto(gx[ib], scl(t, Nf));
      // shear: directors against the chord
      V3 da = qrot(R[ia], e.e0), db = qrot(R[ib], e.e0);
      double cs = 0.5 * cfg.shear_correction * G * e.area * e.L0;
      V3 ma = sub(da, t), mb = sub(db, t);
      double us = 0.5 * cs * (len2(ma) + len2(mb));
      o.energy += us;
      o.elastic += us;
      V3 gt = scl(sub(sub(scl(t, 2), da), db), cs);
      V3 tr_ = divs(sub(gt, scl(t, dotv(gt, t))), L);
      subfrom(gx[ia], tr_);
      addto(gx[ib], tr_);
      subfrom(gr[ia], scl(crossv(da, t), cs));
      subfrom(gr[ib], scl(crossv(db, t), cs));
      // bending + torsion from the relative rotation
      Q4 Q = qnorm(qmul(qconj(R[ia]), R[ib]));
      V3 phi = qlog(Q), axis = unit(qrotinv(R[ia], t));
      double tw = dotv(phi, axis), ptw = dotv(e.plastic.rotp, axis);
      V3 bend = sub(phi, scl(axis, tw)), pb = sub(e.plastic.rotp, scl(axis, ptw));
      double kb = E * e.I / e.L0, kt = G * e.J / e.L0, my = sy * e.I / e.radius;
      double ty = (sy / std::sqrt(3.0)) * e.J / e.radius;
      double hb = M.bend_hardening * kb, ht = M.torsion_hardening * kt;
      Ret3 br = ret3(bend, pb, e.plastic.alphaB, kb, my, hb);
      Ret1 to = ret1(tw, ptw, e.plastic.alphaT, kt, ty, ht);
      tr.next.rotp = add(br.p, scl(axis, to.p));
      tr.next.alphaB = br.a;
      tr.next.alphaT = to.a;
      tr.moment = add(br.force, scl(axis, to.force));
      tr.bend = br.yielded;
      tr.torsion = to.yielded;
      double ur = br.energy + br.hard + to.energy + to.hard, dr = my * br.dg + ty * to.dg;
      o.energy += ur + dr;
      o.elastic += ur;
      o.plastic += dr;
      tr.elastic = uax + us + ur;
      tr.dissipation = dax + dr;
      if (per_term) o.terms[N + bi] = ((uax + dax) + us) + (ur + dr);
      V3 world = qrot(R[ia], jlinvt(phi, tr.moment));
      subfrom(gr[ia], world);
      addto(gr[ib], world);
      if (trials) o.trial[bi] = tr;
    }
    for (size_t i = 0; i < N; i++) {
      o.energy += nodes[i].mass * cfg.gravity * x[i].y;
      gx[i].y += nodes[i].mass * cfg.gravity;
      if (per_term) o.terms[i] = nodes[i].mass * cfg.gravity * x[i].y;
    }
    V3 cen;
    for (const V3& p : x) addto(cen, p);
    cen = divs(cen, double(N));
    o.energy += 0.5 * cfg.centering_stiffness * (sqr(cen.x) + sqr(cen.z));
    o.ecen = 0.5 * cfg.centering_stiffness * (sqr(cen.x) + sqr(cen.z));
    for (V3& g : gx) {
      g.x += cfg.centering_stiffness * cen.x / double(N);
      g.z += cfg.centering_stiffness * cen.z / double(N);
    }
    double J = jr();
    contact(x, gx, o);
    double kc = cfg.contact_stiffness;
    for (size_t i = 0; i < N; i++) {
      double pb = J - x[i].y;
      if (pb > 0) {
        double m = kc * pb, en = 0.5 * kc * pb * pb;
        o.energy += en;
        o.contact += en;
        o.maxPen = std::max(o.maxPen, pb);
        o.contacts++;
        gx[i].y -= m;
        if (per_term) o.terms[i] += en;
      }
      double pt = x[i].y + J - top;
      if (pt > 0) {
        double m = kc * pt, en = 0.5 * kc * pt * pt;
        o.energy += en;
        o.contact += en;
        o.reaction += m;
        o.maxPen = std::max(o.maxPen, pt);
        o.contacts++;
        gx[i].y += m;
        if (per_term) o.terms[i] += en;
      }
    }
    double ts = cfg.translation_scale;
    for (size_t i = 0; i < N; i++) {
      double* g = &o.grad[6 * i];
      g[0] = ts * gx[i].x;
      g[1] = ts * gx[i].y;
      g[2] = ts * gx[i].z;
      g[3] = gr[i].x;
      g[4] = gr[i].y;
      g[5] = gr[i].z;
    }
    for (size_t i = 0; i < o.grad.size(); i++)
      if (locked[i]) o.grad[i] = 0;
    if (!std::isfinite(o.energy)) {
      o.energy = kInf;
      std::fill(o.grad.begin(), o.grad.end(), 0.0);
    }
    return o;
  }

  bool excluded(const Shape& a, const Shape& b) const {
    if (a.beam >= 0 && a.beam == b.beam) return true;
    int ea[2] = {a.n0, a.n1}, eb[2] = {b.n0, b.n1};
    int na = a.capsule ? 2 : 1, nb = b.capsule ? 2 : 1;
    for (int i = 0; i < na; i++)
      for (int j = 0; j < nb; j++)
        if (ea[i] == eb[j] || connected.count(node_pair(ea[i], eb[j]))) return true;
    return false;
  }

  // Sphere/capsule self-contact, MC/src/sim.cpp:148-165.  The hash grid uses the
  // reference's container, hash and insertion sequence so that the pair visiting
  // order (and hence the floating-point accumulation order) is the reference's.
  void contact(const std::vector<V3>& x, std::vector<V3>& gx, Eval& o) const {
    double J = jr();
    std::vector<Shape> shapes;
    shapes.reserve(nodes.size() + beams.size());
    for (int i = 0; i < int(nodes.size()); i++) shapes.push_back({x[size_t(i)], x[size_t(i)], J, i, -1, -1, false});
    for (int bi = 0; bi < int(beams.size()); bi++) {
      const Beam& e = beams[size_t(bi)];
      shapes.push_back({x[size_t(e.a)], x[size_t(e.b)], e.radius, e.a, e.b, bi, true});
    }
    double cell = std::max(2.8 * J, 0.72 * cfg.cell_size);
    std::unordered_map<int64_t, std::vector<int>, GridHash> grid;
    grid.reserve(shapes.size() * 3);
    for (int si = 0; si < int(shapes.size()); si++) {
      const Shape& s = shapes[size_t(si)];
      double pad = s.r + cfg.contact_margin;
      V3 lo = sub(vmin(s.a, s.b), V3{pad, pad, pad}), hi = add(vmax(s.a, s.b), V3{pad, pad, pad});
      int i0 = int(std::floor(lo.x / cell)), j0 = int(std::floor(lo.y / cell)), k0 = int(std::floor(lo.z / cell));
      int i1 = int(std::floor(hi.x / cell)), j1 = int(std::floor(hi.y / cell)), k1 = int(std::floor(hi.z / cell));
      for (int j = j0; j <= j1; j++)
        for (int k = k0; k <= k1; k++)
          for (int i = i0; i <= i1; i++) grid[grid_key(i, j, k)].push_back(si);
    }
    std::unordered_set<uint64_t> seen;
    seen.reserve(shapes.size() * 8);
    double kc = cfg.contact_stiffness;
    for (const auto& kv : grid) {
      const std::vector<int>& list = kv.second;
      for (size_t ia = 0; ia < list.size(); ia++)
        for (size_t ib = ia + 1; ib < list.size(); ib++) {
          int ai = list[ia], bi = list[ib];
          if (!seen.insert(node_pair(ai, bi)).second) continue;
          const Shape &A = shapes[size_t(ai)], &B = shapes[size_t(bi)];
          if (excluded(A, B)) continue;
          Closest cp = shape_pair(A, B);
          V3 delta = sub(cp.pb, cp.pa);
          double dist = len(delta), target = A.r + B.r + cfg.contact_margin;
          if (dist >= target) continue;
          V3 n = dist > 1e-12 ? divs(delta, dist) : unit(V3{0.37 + 0.001 * ai, 0.73, -0.41 + 0.001 * bi});
          double pen = target - dist, mag = kc * pen, en = 0.5 * kc * pen * pen;
          o.energy += en;
          o.contact += en;
          o.maxPen = std::max(o.maxPen, pen);
          o.contacts++;
          if (!o.terms.empty()) o.terms[nodes.size() + beams.size() + size_t(std::min(ai, bi))] += en;
          spread(gx, A, cp.s, scl(n, mag));
          spread(gx, B, cp.t, scl(n, -mag));
        }
    }
  }

  // ---- L-BFGS helpers, MC/src/sim.cpp:43-46 ----
  static double vdot(const std::vector<double>& a, const std::vector<double>& b) {
    double r = 0;
    for (size_t i = 0; i < a.size(); i++) r += a[i] * b[i];
    return r;
  }
  double vrms(const std::vector<double>& a) const {
    double s = 0;
    size_t n = 0;
    for (size_t i = 0; i < a.size(); i++)
      if (!locked[i]) {
        s += a[i] * a[i];
        n++;
      }
    return n ? std::sqrt(s / double(n)) : 0;
  }
  double vmaxabs(const std::vector<double>& a) const {
    double m = 0;
    for (size_t i = 0; i < a.size(); i++)
      if (!locked[i]) m = std::max(m, std::abs(a[i]));
    return m;
  }
  void clamp_direction(std::vector<double>& p) const {
    double mt = cfg.max_translation_step / cfg.translation_scale;
    for (size_t i = 0; i < p.size() / 6; i++) {
      double* q = &p[6 * i];
      V3 t{q[0], q[1], q[2]};
      double lt = len(t);
      if (lt > mt) {
        t = scl(t, mt / lt);
        q[0] = t.x;
        q[1] = t.y;
        q[2] = t.z;
      }
      V3 r{q[3], q[4], q[5]};
      double lr = len(r);
      if (lr > cfg.max_rotation_step) {
        r = scl(r, cfg.max_rotation_step / lr);
        q[3] = r.x;
        q[4] = r.y;
        q[5] = r.z;
      }
    }
    for (size_t i = 0; i < p.size(); i++)
      if (locked[i]) p[i] = 0;
  }
  bool converged(const Eval& o, double rms, double mx) const {
    double rel = rms / std::max(1e-6, cfg.translation_scale * std::max(1.0, o.reaction));
    bool ok = (rms <= cfg.gradient_rms_tolerance && mx <= cfg.gradient_max_tolerance) ||
              (cfg.relative_escape && rel < cfg.relative_tolerance);
    return ok && o.maxPen <= cfg.penetration_tolerance;
  }
  std::vector<double> steepest(const std::vector<double>& g) const {
    std::vector<double> p(g.size());
    for (size_t i = 0; i < g.size(); i++) p[i] = -g[i];
    for (size_t i = 0; i < p.size(); i++)
      if (locked[i]) p[i] = 0;
    clamp_direction(p);
    return p;
  }

  // MC/src/sim.cpp:174-211
  Stats minimize(double top, std::vector<double>& z, Eval& fin) {
    Stats st;
    int ev0 = evaluations;
    struct Pair {
      std::vector<double> s, y;
      double rho;
    };
    std::vector<Pair> hist;
    Eval cur = objective(z, top, false), best = cur;
    std::vector<double> bestZ = z;
    auto search = [&](const std::vector<double>& p, double dg, std::vector<double>& zc, Eval& cand) {
      double step = 1.0;
      for (int bt = 0; bt < cfg.max_line_search; bt++) {
        for (size_t i = 0; i < z.size(); i++) zc[i] = locked[i] ? 0.0 : z[i] + step * p[i];
        cand = objective(zc, top, false);
        st.backtracks++;
        bool accept;
        if (cfg.line_search == 1 || cfg.line_search == 2) {
          // Armijo on the energy change summed per term: free of the
          // cancellation of two nearly equal totals
          double acc = 0;
          for (size_t i = 0; i < cand.terms.size(); i++) acc += cand.terms[i] - cur.terms[i];
          const double dE = acc + (cand.ecen - cur.ecen);
          accept = std::isfinite(cand.energy) && dE <= cfg.armijo * step * dg;
          // line_search 2: the approximate-Wolfe test of Hager & Zhang (SIAM
          // J. Optim. 16, 2005, eq. 4.1, epsilon 1e-6 |E|) when the change is
          // within the energy's resolution: the slope along p at the trial,
          // phi'(step) = g(zc).p, against (2 armijo - 1) phi'(0)
          if (!accept && cfg.line_search == 2 && std::isfinite(cand.energy) &&
              std::abs(dE) <= 1e-6 * std::abs(cur.energy))
            accept = vdot(cand.grad, p) <= (2.0 * cfg.armijo - 1.0) * dg;
        } else {
          accept = std::isfinite(cand.energy) && cand.energy <= cur.energy + cfg.armijo * step * dg;
        }
        if (accept) return true;
        step *= 0.5;
        if (step < cfg.min_line_step) break;
      }
      return false;
    };
    for (int it = 0; it < cfg.max_iterations; it++) {
      st.iterations = it + 1;
      st.rms = vrms(cur.grad);
      st.max = vmaxabs(cur.grad);
      st.energy = cur.energy;
      if (converged(cur, st.rms, st.max)) {
        st.converged = true;
        break;
      }
      // two-loop recursion
      std::vector<double> p = cur.grad, alpha(hist.size());
      for (int h = int(hist.size()) - 1; h >= 0; h--) {
        alpha[size_t(h)] = hist[size_t(h)].rho * vdot(hist[size_t(h)].s, p);
        for (size_t i = 0; i < p.size(); i++) p[i] -= alpha[size_t(h)] * hist[size_t(h)].y[i];
      }
      double H0 = 0.10;
      if (!hist.empty()) {
        double sy = vdot(hist.back().s, hist.back().y), yy = vdot(hist.back().y, hist.back().y);
        if (sy > 1e-20 && yy > 1e-20) H0 = clampd(sy / yy, 1e-7, 1e4);
      }
      for (double& v : p) v *= H0;
      for (size_t h = 0; h < hist.size(); h++) {
        double beta = hist[h].rho * vdot(hist[h].y, p);
        for (size_t i = 0; i < p.size(); i++) p[i] += hist[h].s[i] * (alpha[h] - beta);
      }
      for (double& v : p) v = -v;
      for (size_t i = 0; i < p.size(); i++)
        if (locked[i]) p[i] = 0;
      clamp_direction(p);
      double dg = vdot(cur.grad, p);
      if (!(dg < 0)) {
        hist.clear();
        p = steepest(cur.grad);
        dg = vdot(cur.grad, p);
      }
      std::vector<double> zc(z.size());
      Eval cand;
      bool ok = search(p, dg, zc, cand);
      if (!ok) {
        hist.clear();
        p = steepest(cur.grad);
        dg = vdot(cur.grad, p);
        ok = search(p, dg, zc, cand);
      }
      if (!ok) break;
      std::vector<double> sv(z.size()), yv(z.size());
      for (size_t i = 0; i < z.size(); i++) {
        sv[i] = zc[i] - z[i];
        yv[i] = cand.grad[i] - cur.grad[i];
      }
      double sy = vdot(sv, yv), ss = vdot(sv, sv), yy = vdot(yv, yv);
      if (sy > 1e-12 * std::sqrt(std::max(1e-30, ss * yy))) {
        if (int(hist.size()) == cfg.history) hist.erase(hist.begin());
        hist.push_back({std::move(sv), std::move(yv), 1 / sy});
      } else {
        hist.clear();
      }
      z = std::move(zc);
      cur = std::move(cand);
      if (cur.energy < best.energy) {
        best = cur;
        bestZ = z;
      }
    }
    if (!st.converged && best.energy < cur.energy) {
      z = bestZ;
      cur = best;
    }
    fin = objective(z, top, true);
    st.rms = vrms(fin.grad);
    st.max = vmaxabs(fin.grad);
    st.energy = fin.energy;
    st.converged = converged(fin, st.rms, st.max);
    st.evaluations = evaluations - ev0;
    return st;
  }

  // MC/src/sim.cpp:212
  void commit(const std::vector<double>& z, const Eval& o) {
    std::vector<V3> x;
    std::vector<Q4> R;
    state(z, x, R);
    if (o.trial.size() != beams.size()) throw std::runtime_error("missing constitutive trial states");
    for (size_t i = 0; i < nodes.size(); i++) {
      nodes[i].x = x[i];
      nodes[i].R = R[i];
      baseX[i] = x[i];
      baseR[i] = R[i];
    }
    for (size_t i = 0; i < beams.size(); i++) beams[i].plastic = o.trial[i].next;
  }
  double top_at(double s, bool loading) const {
    double low = top0 - cfg.max_strain * height;
    return loading ? mixd(top0, low, sat(s)) : mixd(low, top0, sat(s));
  }

  // MC/src/sim.cpp:312-372
  mmc_frame diagnostics(double time, double top, bool loading, const Eval& o, const Stats& s) const {
    mmc_frame d;
    std::memset(&d, 0, sizeof(d));
    d.time = time;
    d.platen_y = top;
    d.displacement = top0 - top;
    d.engineering_strain = d.displacement / std::max(1e-12, height);
    d.reaction_force = std::max(0.0, o.reaction - tare);
    d.engineering_stress = d.reaction_force / std::max(1e-12, area);
    d.loading_work = loadW;
    d.recovered_work = recW;
    d.dissipated_work = std::max(0.0, loadW - recW);
    d.specific_energy_absorption = d.dissipated_work / std::max(1e-12, mass);
    d.elastic_energy = o.elastic;
    d.plastic_dissipation = o.plastic;
    d.contact_energy = o.contact;
    d.residual_rms = s.rms;
    d.residual_max = s.max;
    d.max_penetration = o.maxPen;
    d.active_contacts = o.contacts;
    d.iterations = s.iterations;
    d.converged = s.converged;
    d.loading = loading;
    double sa = 0, ma = 0, sr = 0, mr = 0;
    int pbm = 0, phg = 0;
    for (const Beam& e : beams) {
      double pa = std::abs(e.plastic.epsp), pr = len(e.plastic.rotp);
      sa += pa;
      ma = std::max(ma, pa);
      sr += pr;
      mr = std::max(mr, pr);
      if (pa > 1e-5) ++pbm;
      if (pr > 1e-4) ++phg;
    }
    if (!beams.empty()) {
      d.mean_plastic_strain = sa / double(beams.size());
      d.max_plastic_strain = ma;
      d.mean_plastic_rotation = sr / double(beams.size());
      d.max_plastic_rotation = mr;
    }
    d.plastic_beams = pbm;
    d.plastic_hinges = phg;
    double ylo = kInf, yhi = -kInf;
    for (const Node& n : nodes) {
      ylo = std::min(ylo, n.x.y);
      yhi = std::max(yhi, n.x.y);
    }
    double hcur = std::max(0.0, yhi - ylo);
    d.permanent_set = std::max(0.0, 1.0 - hcur / std::max(1e-12, height));
    return d;
  }

  // MC/src/sim.cpp:374-467
  void run(std::vector<mmc_frame>& out, bool verbose) {
    loadW = recW = prevDisp = prevForce = tare = 0;
    double time = 0, current = top0;
    bool first = true;
    int accepted = 0;
    std::function<void(double, bool, int)> advance = [&](double target, bool loading, int refinement) {
      std::vector<double> z(nodes.size() * 6, 0.0);
      Eval fin;
      Stats st = minimize(target, z, fin);
      bool ok = st.converged && fin.maxPen <= cfg.penetration_tolerance && std::isfinite(fin.energy);
      if (!ok && refinement < cfg.max_adaptive_refinements && std::abs(target - current) > 2.5e-5) {
        double mid = 0.5 * (current + target);
        advance(mid, loading, refinement + 1);
        advance(target, loading, refinement + 1);
        return;
      }
      if (!ok) {
        char msg[256];
        std::snprintf(msg, sizeof msg, "quasi-static equilibrium failed at platen=%g m, residual=%g, penetration=%g mm after %d iterations",
                      target, st.rms, fin.maxPen * 1000.0, st.iterations);
        throw std::runtime_error(msg);
      }
      commit(z, fin);
      if (first) tare = fin.reaction;
      double F = std::max(0.0, fin.reaction - tare), disp = top0 - target;
      if (!first) {
        double dd = disp - prevDisp, w = 0.5 * (prevForce + F) * dd;
        if (dd >= 0.0)
          loadW += std::max(0.0, w);
        else
          recW += std::max(0.0, -w);
      }
      prevDisp = disp;
      prevForce = F;
      current = target;
      out.push_back(diagnostics(time, target, loading, fin, st));
      if (frame_cb && frame_cb(&out.back(), int32_t(out.size()), frame_cb_user)) throw RunStopped{};
      if (verbose) {
        const mmc_frame& d = out.back();
        std::printf("[oracle] %3d %s strain=%.4f force=%.3f N residual=%.2e pen=%.3f mm iter=%d%s\n", accepted,
                    loading ? "load  " : "unload", d.engineering_strain, d.reaction_force, d.residual_rms,
                    d.max_penetration * 1000.0, d.iterations, refinement ? " [REFINED]" : " [OK]");
        std::fflush(stdout);
      }
      accepted++;
      first = false;
      time += cfg.step_time;
    };
    advance(top0, true, 0);
    for (int s = 1; s <= cfg.load_steps; s++) advance(top_at(double(s) / std::max(1, cfg.load_steps), true), true, 0);
    for (int s = 1; s <= cfg.unload_steps; s++)
      advance(top_at(double(s) / std::max(1, cfg.unload_steps), false), false, 0);
  }

  // MC/src/sim.cpp:468-509
  double beam_potential(const std::vector<V3>& x, const std::vector<Q4>& R) const {
    const mmc_material& M = cfg.material;
    double E = M.young, G = E / (2.0 * (1.0 + M.nu)), sy = M.yield, H = M.hardening, total = 0;
    for (const Beam& e : beams) {
      V3 d = sub(x[size_t(e.b)], x[size_t(e.a)]);
      double L = std::max(1e-12, len(d));
      V3 t = divs(d, L);
      double eps = (L - e.L0) / e.L0;
      Ret1 ax = ret1(eps, e.plastic.epsp, e.plastic.alpha, E, sy, H);
      total += e.area * e.L0 * (ax.energy + ax.hard);
      V3 da = qrot(R[size_t(e.a)], e.e0), db = qrot(R[size_t(e.b)], e.e0);
      double cs = 0.5 * cfg.shear_correction * G * e.area * e.L0;
      total += 0.5 * cs * (len2(sub(da, t)) + len2(sub(db, t)));
      Q4 Q = qnorm(qmul(qconj(R[size_t(e.a)]), R[size_t(e.b)]));
      V3 phi = qlog(Q), axis = unit(qrotinv(R[size_t(e.a)], t));
      double tw = dotv(phi, axis), ptw = dotv(e.plastic.rotp, axis);
      V3 bend = sub(phi, scl(axis, tw)), pb = sub(e.plastic.rotp, scl(axis, ptw));
      double kb = E * e.I / e.L0, kt = G * e.J / e.L0, my = sy * e.I / e.radius;
      double ty = (sy / std::sqrt(3.0)) * e.J / e.radius;
      Ret3 br = ret3(bend, pb, e.plastic.alphaB, kb, my, M.bend_hardening * kb);
      Ret1 to = ret1(tw, ptw, e.plastic.alphaT, kt, ty, M.torsion_hardening * kt);
      total += br.energy + br.hard + to.energy + to.hard;
    }
    return total;
  }

  // MC/src/sim.cpp:511-536
  bool rigid_rotation_test(double* error) const {
    size_t N = nodes.size();
    std::vector<V3> x(N);
    std::vector<Q4> R(N);
    for (size_t i = 0; i < N; i++) {
      const Node& n = nodes[i];
      double yf = sat((n.x0.y - jr()) / std::max(1e-12, height));
      x[i] = n.x0;
      x[i].y -= 0.018 * height * yf;
      x[i].x += 0.00018 * std::sin(2.0 * kPi * yf + 0.17 * double(i));
      R[i] = qaxisangle(V3{0, 0, 1}, 0.035 * yf);
    }
    double e0 = beam_potential(x, R);
    Q4 Q = qaxisangle(unit(V3{0.37, 0.81, -0.44}), 1.137);
    V3 shift{0.081, -0.024, 0.053};
    std::vector<V3> xr(N);
    std::vector<Q4> Rr(N);
    for (size_t i = 0; i < N; i++) {
      xr[i] = add(qrot(Q, x[i]), shift);
      Rr[i] = qnorm(qmul(Q, R[i]));
    }
    double e1 = beam_potential(xr, Rr);
    double rel = std::abs(e1 - e0) / std::max({1e-12, std::abs(e0), std::abs(e1)});
    if (error) *error = rel;
    return rel < 1e-9;
  }

  // MC/src/sim.cpp:538-577
  bool gradient_check(double* error) {
    size_t N = nodes.size();
    std::vector<double> z(N * 6, 0.0);
    Stream rng(cfg.seed ^ 0x475241444348454bULL);
    for (size_t i = 0; i < N; i++) {
      size_t o = 6 * i;
      z[o + 0] = 0.003 * rng.symmetric();
      z[o + 1] = 0.100 + 0.003 * rng.symmetric();
      z[o + 2] = 0.003 * rng.symmetric();
      z[o + 3] = 0.004 * rng.symmetric();
      z[o + 4] = 0.004 * rng.symmetric();
      z[o + 5] = 0.004 * rng.symmetric();
      for (int d = 0; d < 6; d++)
        if (locked[o + size_t(d)]) z[o + size_t(d)] = 0.0;
    }
    double top = top0 + 0.020;
    Eval a = objective(z, top, false);
    double num = 0, den = 0;
    int checked = 0;
    for (size_t k = 0; k < z.size() && checked < 42; k++) {
      if (locked[k]) continue;
      if ((mix64(k * 0x9e3779b97f4a7c15ULL + cfg.seed) & 15ULL) > 1ULL) continue;
      double h = (k % 6 < 3) ? 2e-6 : 1e-6;
      std::vector<double> zp = z, zm = z;
      zp[k] += h;
      zm[k] -= h;
      double fp = objective(zp, top, false).energy, fm = objective(zm, top, false).energy;
      double numeric = (fp - fm) / (2.0 * h), analytic = a.grad[k];
      num += sqr(numeric - analytic);
      den += sqr(numeric) + sqr(analytic);
      ++checked;
    }
    double rel = std::sqrt(num / std::max(1e-30, den));
    if (error) *error = rel;
    return checked >= 24 && rel < 0.075;
  }
};

namespace {
void trial_row(const Trial& t, double* o) {
  o[0] = t.next.epsp;
  o[1] = t.next.alpha;
  o[2] = t.next.rotp.x;
  o[3] = t.next.rotp.y;
  o[4] = t.next.rotp.z;
  o[5] = t.next.alphaB;
  o[6] = t.next.alphaT;
  o[7] = t.strain;
  o[8] = t.stress;
  o[9] = t.moment.x;
  o[10] = t.moment.y;
  o[11] = t.moment.z;
  o[12] = t.elastic;
  o[13] = t.dissipation;
  o[14] = t.axial;
  o[15] = t.bend;
  o[16] = t.torsion;
}
void fill_out(const Eval& e, const orc* h, mmc_objective_out* out) {
  if (!out) return;
  out->energy = e.energy;
  out->elastic = e.elastic;
  out->plastic = e.plastic;
  out->contact = e.contact;
  out->reaction = e.reaction;
  out->max_pen = e.maxPen;
  out->contacts = e.contacts;
  out->finite = std::isfinite(e.energy) ? 1 : 0;
  out->rms = h->vrms(e.grad);
  out->max = h->vmaxabs(e.grad);
}
}  // namespace

extern "C" {

orc* orc_create(const mmc_config* cfg) {
  orc* h = new orc();
  h->cfg = *cfg;
  h->cfg.translation_scale = h->cfg.cell_size;  // MC/src/sim.cpp:49
  try {
    if (cfg->lattice == MMC_LATTICE_CUBIC)
      h->build_cubic();
    else
      h->build_bcc();
    h->distribute_mass();
    h->init_base();
  } catch (const std::exception& e) {
    h->err = e.what();
  }
  return h;
}

orc* orc_create_lattice(const mmc_config* cfg, const mmc_lattice* L) {
  orc* h = new orc();
  h->cfg = *cfg;
  h->cfg.translation_scale = h->cfg.cell_size;
  h->width = L->width;
  h->height = L->height;
  h->depth = L->depth;
  h->area = L->width * L->depth;
  for (int i = 0; i < L->n_nodes; i++) {
    Node n;
    n.x0 = V3{L->x[3 * i], L->x[3 * i + 1], L->x[3 * i + 2]};
    n.x = n.x0;
    if (L->q) n.R = Q4{L->q[4 * i], L->q[4 * i + 1], L->q[4 * i + 2], L->q[4 * i + 3]};
    n.layer = L->node_layer ? L->node_layer[i] : 0;
    h->nodes.push_back(n);
  }
  for (int b = 0; b < L->n_beams; b++) {
    Beam e;
    e.a = L->beam_a[b];
    e.b = L->beam_b[b];
    e.layer = L->beam_layer ? L->beam_layer[b] : 0;
    e.family = L->beam_family ? L->beam_family[b] : 0;
    e.L0 = L->L0[b];
    e.radius = L->radius[b];
    e.area = L->area[b];
    e.I = L->inertia[b];
    e.J = 2 * e.I;
    e.e0 = V3{L->e0[3 * b], L->e0[3 * b + 1], L->e0[3 * b + 2]};
    if (L->plastic) {
      const double* p = L->plastic + 7 * b;
      e.plastic.epsp = p[0];
      e.plastic.alpha = p[1];
      e.plastic.rotp = V3{p[2], p[3], p[4]};
      e.plastic.alphaB = p[5];
      e.plastic.alphaT = p[6];
    }
    h->beams.push_back(e);
  }
  for (int64_t c = 0; c < L->n_connected; c++) h->add_connected(L->connected[c]);
  h->top0 = h->jr() + h->height + h->jr();
  try {
    h->distribute_mass();
    h->init_base();
  } catch (const std::exception& e) {
    h->err = e.what();
  }
  return h;
}

void orc_destroy(orc* h) { delete h; }
const char* orc_last_error(orc* h) { return h->err.c_str(); }

void orc_info(orc* h, int32_t* nn, int32_t* nb, int64_t* nc, double* env) {
  if (nn) *nn = int32_t(h->nodes.size());
  if (nb) *nb = int32_t(h->beams.size());
  if (nc) *nc = int64_t(h->connected_list.size());
  if (env) {
    env[0] = h->width;
    env[1] = h->height;
    env[2] = h->depth;
    env[3] = h->area;
    env[4] = h->mass;
    env[5] = h->top0;
  }
}

void orc_lattice(orc* h, double* x, double* q, double* mass, uint8_t* locked, int32_t* a, int32_t* b, double* L0,
                 double* radius, double* area, double* inertia, double* e0, double* plastic, uint64_t* connected,
                 int32_t* node_layer, int32_t* beam_layer, int32_t* beam_family) {
  for (size_t i = 0; i < h->nodes.size(); i++) {
    const Node& n = h->nodes[i];
    if (x) {
      x[3 * i] = h->baseX[i].x;
      x[3 * i + 1] = h->baseX[i].y;
      x[3 * i + 2] = h->baseX[i].z;
    }
    if (q) {
      q[4 * i] = h->baseR[i].w;
      q[4 * i + 1] = h->baseR[i].x;
      q[4 * i + 2] = h->baseR[i].y;
      q[4 * i + 3] = h->baseR[i].z;
    }
    if (mass) mass[i] = n.mass;
    if (node_layer) node_layer[i] = n.layer;
  }
  if (locked) std::memcpy(locked, h->locked.data(), h->locked.size());
  for (size_t k = 0; k < h->beams.size(); k++) {
    const Beam& e = h->beams[k];
    if (a) a[k] = e.a;
    if (b) b[k] = e.b;
    if (L0) L0[k] = e.L0;
    if (radius) radius[k] = e.radius;
    if (area) area[k] = e.area;
    if (inertia) inertia[k] = e.I;
    if (e0) {
      e0[3 * k] = e.e0.x;
      e0[3 * k + 1] = e.e0.y;
      e0[3 * k + 2] = e.e0.z;
    }
    if (plastic) {
      double* p = plastic + 7 * k;
      p[0] = e.plastic.epsp;
      p[1] = e.plastic.alpha;
      p[2] = e.plastic.rotp.x;
      p[3] = e.plastic.rotp.y;
      p[4] = e.plastic.rotp.z;
      p[5] = e.plastic.alphaB;
      p[6] = e.plastic.alphaT;
    }
    if (beam_layer) beam_layer[k] = e.layer;
    if (beam_family) beam_family[k] = e.family;
  }
  if (connected) std::memcpy(connected, h->connected_list.data(), h->connected_list.size() * sizeof(uint64_t));
}

double orc_objective(orc* h, const double* z, double top, int32_t trials, mmc_objective_out* out, double* grad,
                     double* trial) {
  std::vector<double> zz(z, z + h->nodes.size() * 6);
  Eval e = h->objective(zz, top, trials != 0);
  fill_out(e, h, out);
  if (grad) std::memcpy(grad, e.grad.data(), e.grad.size() * sizeof(double));
  if (trial && trials)
    for (size_t k = 0; k < e.trial.size(); k++) trial_row(e.trial[k], trial + MMC_TRIAL_STRIDE * k);
  return e.energy;
}

// The line search's energy change from z0 to z1 in both forms: out[0] the
// per-term sum (line_search 1), out[1] the difference of the totals.
void orc_energy_change(orc* h, const double* z0, const double* z1, double top, double* out) {
  const size_t n = h->nodes.size() * 6;
  const int keep = h->cfg.line_search;
  h->cfg.line_search = 1;
  Eval a = h->objective(std::vector<double>(z0, z0 + n), top, false);
  Eval b = h->objective(std::vector<double>(z1, z1 + n), top, false);
  h->cfg.line_search = keep;
  double acc = 0;
  for (size_t i = 0; i < b.terms.size(); i++) acc += b.terms[i] - a.terms[i];
  out[0] = acc + (b.ecen - a.ecen);
  out[1] = b.energy - a.energy;
}

int orc_minimize(orc* h, double top, double* z, mmc_opt_stats* st, mmc_objective_out* fin) {
  std::vector<double> zz(z, z + h->nodes.size() * 6);
  Eval f;
  Stats s = h->minimize(top, zz, f);
  std::memcpy(z, zz.data(), zz.size() * sizeof(double));
  h->last = f;
  h->have_last = true;
  h->lastZ = zz;
  if (st) {
    st->iterations = s.iterations;
    st->backtracks = s.backtracks;
    st->rms = s.rms;
    st->max = s.max;
    st->energy = s.energy;
    st->converged = s.converged;
    st->evaluations = s.evaluations;
  }
  if (fin) {
    fin->energy = f.energy;
    fin->elastic = f.elastic;
    fin->plastic = f.plastic;
    fin->contact = f.contact;
    fin->reaction = f.reaction;
    fin->max_pen = f.maxPen;
    fin->contacts = f.contacts;
    fin->finite = std::isfinite(f.energy);
    fin->rms = s.rms;
    fin->max = s.max;
  }
  return 0;
}

int orc_commit(orc* h) {
  if (!h->have_last) {
    h->err = "missing constitutive trial states";
    return MMC_ERR_STATE;
  }
  try {
    h->commit(h->lastZ, h->last);
  } catch (const std::exception& e) {
    h->err = e.what();
    return MMC_ERR_STATE;
  }
  h->have_last = false;
  return 0;
}

int orc_run(orc* h, mmc_frame* frames, int32_t cap, int32_t* n, int32_t verbose) {
  std::vector<mmc_frame> out;
  int rc = 0;
  try {
    h->run(out, verbose != 0);
  } catch (const RunStopped&) {
  } catch (const std::exception& e) {
    h->err = e.what();
    rc = MMC_ERR_EQUILIBRIUM;
  }
  int m = std::min<int>(cap, int(out.size()));
  if (frames) std::memcpy(frames, out.data(), size_t(m) * sizeof(mmc_frame));
  if (n) *n = int32_t(out.size());
  return rc;
}

void orc_set_frame_callback(orc* h, int (*cb)(const mmc_frame*, int32_t, void*), void* user) {
  h->frame_cb = cb;
  h->frame_cb_user = user;
}

double orc_beam_potential(orc* h, const double* x, const double* q) {
  size_t N = h->nodes.size();
  std::vector<V3> xx(N);
  std::vector<Q4> RR(N);
  for (size_t i = 0; i < N; i++) {
    xx[i] = V3{x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    RR[i] = Q4{q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
  }
  return h->beam_potential(xx, RR);
}
int orc_rigid_rotation_test(orc* h, double* err) { return h->rigid_rotation_test(err) ? 1 : 0; }
int orc_gradient_check(orc* h, double* err) { return h->gradient_check(err) ? 1 : 0; }

void orc_read_state(orc* h, double* x, double* q, double* plastic) {
  orc_lattice(h, x, q, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, plastic,
              nullptr, nullptr, nullptr, nullptr);
}
void orc_set_plastic(orc* h, const double* p) {
  for (size_t k = 0; k < h->beams.size(); k++) {
    const double* s = p + 7 * k;
    Plastic& q = h->beams[k].plastic;
    q.epsp = s[0];
    q.alpha = s[1];
    q.rotp = V3{s[2], s[3], s[4]};
    q.alphaB = s[5];
    q.alphaT = s[6];
  }
}
void orc_set_pose(orc* h, const double* x, const double* q) {
  for (size_t i = 0; i < h->nodes.size(); i++) {
    h->baseX[i] = V3{x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    h->baseR[i] = Q4{q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
    h->nodes[i].x = h->baseX[i];
    h->nodes[i].R = h->baseR[i];
  }
}

namespace {
// quatFromAxes, MC/include/core.hpp:43: the rotation with columns X, Y, Z
Q4 from_axes(V3 X, V3 Y, V3 Z) {
  double m00 = X.x, m01 = Y.x, m02 = Z.x, m10 = X.y, m11 = Y.y, m12 = Z.y, m20 = X.z, m21 = Y.z, m22 = Z.z;
  Q4 q;
  double tr = m00 + m11 + m22;
  if (tr > 0) {
    double s = std::sqrt(tr + 1) * 2;
    q = {0.25 * s, (m21 - m12) / s, (m02 - m20) / s, (m10 - m01) / s};
  } else if (m00 > m11 && m00 > m22) {
    double s = std::sqrt(1 + m00 - m11 - m22) * 2;
    q = {(m21 - m12) / s, 0.25 * s, (m01 + m10) / s, (m02 + m20) / s};
  } else if (m11 > m22) {
    double s = std::sqrt(1 + m11 - m00 - m22) * 2;
    q = {(m02 - m20) / s, (m01 + m10) / s, 0.25 * s, (m12 + m21) / s};
  } else {
    double s = std::sqrt(1 + m22 - m00 - m11) * 2;
    q = {(m10 - m01) / s, (m02 + m20) / s, (m12 + m21) / s, 0.25 * s};
  }
  return qnorm(q);
}
// alignX, MC/src/sim.cpp:41, with orthogonal, MC/include/core.hpp:47
Q4 align_x(V3 d) {
  V3 X = unit(d);
  V3 h = std::abs(X.z) < .8 ? V3{0, 0, 1} : V3{0, 1, 0};
  V3 Y = unit(crossv(h, X));
  V3 Z = unit(crossv(X, Y));
  Y = unit(crossv(Z, X));
  return from_axes(X, Y, Z);
}
void put(mmc_snapshot& s, V3 p, Q4 q, V3 v, V3 w, V3 half, int material, int id, int kind, int module_class,
         int ring, int slot) {
  const double vals[16] = {p.x, p.y, p.z, q.w, q.x, q.y, q.z, v.x, v.y, v.z, w.x, w.y, w.z, half.x, half.y, half.z};
  std::memcpy(s.p, vals, 3 * sizeof(double));
  std::memcpy(s.q, vals + 3, 4 * sizeof(double));
  std::memcpy(s.v, vals + 7, 3 * sizeof(double));
  std::memcpy(s.w, vals + 10, 3 * sizeof(double));
  std::memcpy(s.half, vals + 13, 3 * sizeof(double));
  s.material = material;
  s.id = id;
  s.kind = kind;
  s.module_class = module_class;
  s.ring = ring;
  s.slot = slot;
}
}  // namespace

// Simulator::capture(top, prevX, prevR), MC/src/sim.cpp:215-310: struts (beam
// order), joints (node order), then the rigid machine; velocities from the
// previous poses prev_x (3N) / prev_q (4N), zero when they are NULL.
void orc_capture(orc* h, double top, const double* prev_x, const double* prev_q, mmc_snapshot* out) {
  const std::vector<Node>& nodes = h->nodes;
  const std::vector<Beam>& beams = h->beams;
  const double dt = std::max(1e-6, h->cfg.step_time);
  auto vel = [&](int i) {
    if (!prev_x) return V3{};
    V3 px{prev_x[3 * size_t(i)], prev_x[3 * size_t(i) + 1], prev_x[3 * size_t(i) + 2]};
    return divs(sub(nodes[size_t(i)].x, px), dt);
  };
  auto omega = [&](int i) {
    if (!prev_q) return V3{};
    const double* r = prev_q + 4 * size_t(i);
    Q4 dq = qnorm(qmul(nodes[size_t(i)].R, qconj(Q4{r[0], r[1], r[2], r[3]})));
    return divs(qlog(dq), dt);
  };
  size_t o = 0;
  for (size_t bi = 0; bi < beams.size(); bi++) {
    const Beam& e = beams[bi];
    V3 a = nodes[size_t(e.a)].x, b = nodes[size_t(e.b)].x, d = sub(b, a);
    double L = std::max(1e-8, len(d));
    double pAxial = std::abs(e.plastic.epsp), pRot = len(e.plastic.rotp);
    double severity = std::max(pAxial / 0.10, pRot / 0.42);
    int material = 0;
    if (severity > 0.28) material = 1;
    if (severity > 0.62) material = 2;
    if (severity > 1.05) material = 3;
    put(out[o++], scl(add(a, b), 0.5), align_x(d), scl(add(vel(e.a), vel(e.b)), 0.5),
        scl(add(omega(e.a), omega(e.b)), 0.5), V3{0.5 * L, e.radius, e.radius}, material, int(bi), 0, e.family,
        e.layer, int(bi));
  }
  const double J = h->jr();
  for (size_t ni = 0; ni < nodes.size(); ni++)
    put(out[o++], nodes[ni].x, nodes[ni].R, vel(int(ni)), omega(int(ni)), V3{J, J, J}, 4,
        int(beams.size() + ni), 1, nodes[ni].layer, nodes[ni].layer, int(ni));
  int id = int(beams.size() + nodes.size());
  auto box = [&](V3 p, V3 half, int material, int kind) {
    put(out[o++], p, Q4{}, V3{}, V3{}, half, material, id++, kind, -1, -1, -1);
  };
  box({0, -0.0030, 0}, {0.026, 0.0030, 0.026}, 5, 2);
  box({0, top + 0.0030, 0}, {0.026, 0.0030, 0.026}, 5, 2);
  box({0, -0.0095, 0}, {0.036, 0.0035, 0.036}, 6, 3);
  const double columnY = 0.054;
  box({-0.037, columnY, -0.049}, {0.0026, 0.061, 0.0026}, 6, 3);
  box({0.037, columnY, -0.049}, {0.0026, 0.061, 0.0026}, 6, 3);
  box({0, 0.111, -0.049}, {0.042, 0.0032, 0.0036}, 6, 3);
  box({0, top + 0.0080, -0.001}, {0.0105, 0.0025, 0.0105}, 7, 3);
}

void orc_scalar_return(double total, double pc, double ac, double k, double y, double H, double* o) {
  Ret1 r = ret1(total, pc, ac, k, y, H);
  o[0] = r.force;
  o[1] = r.p;
  o[2] = r.a;
  o[3] = r.dg;
  o[4] = r.energy;
  o[5] = r.hard;
  o[6] = r.yielded;
}
void orc_vector_return(const double* t, const double* pc, double ac, double k, double y, double H, double* o) {
  Ret3 r = ret3(V3{t[0], t[1], t[2]}, V3{pc[0], pc[1], pc[2]}, ac, k, y, H);
  o[0] = r.force.x;
  o[1] = r.force.y;
  o[2] = r.force.z;
  o[3] = r.p.x;
  o[4] = r.p.y;
  o[5] = r.p.z;
  o[6] = r.a;
  o[7] = r.dg;
  o[8] = r.energy;
  o[9] = r.hard;
  o[10] = r.yielded;
  o[11] = 0;
}
void orc_closest_segments(const double* p1, const double* q1, const double* p2, const double* q2, double* o) {
  Closest c = seg_seg(V3{p1[0], p1[1], p1[2]}, V3{q1[0], q1[1], q1[2]}, V3{p2[0], p2[1], p2[2]}, V3{q2[0], q2[1], q2[2]});
  o[0] = c.s;
  o[1] = c.t;
  o[2] = c.pa.x;
  o[3] = c.pa.y;
  o[4] = c.pa.z;
  o[5] = c.pb.x;
  o[6] = c.pb.y;
  o[7] = c.pb.z;
}
void orc_quat_exp(const double* v, double* q) {
  Q4 r = qexp(V3{v[0], v[1], v[2]});
  q[0] = r.w;
  q[1] = r.x;
  q[2] = r.y;
  q[3] = r.z;
}
void orc_quat_log(const double* q, double* v) {
  V3 r = qlog(Q4{q[0], q[1], q[2], q[3]});
  v[0] = r.x;
  v[1] = r.y;
  v[2] = r.z;
}
void orc_jl_inv_t(const double* p, const double* v, double* o) {
  V3 r = jlinvt(V3{p[0], p[1], p[2]}, V3{v[0], v[1], v[2]});
  o[0] = r.x;
  o[1] = r.y;
  o[2] = r.z;
}

}  // extern "C"
