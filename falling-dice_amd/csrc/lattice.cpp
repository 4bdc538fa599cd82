// lattice.cpp — host lattice generators (BCC reference family and simple cubic).
// See lattice.hpp.  Reference: projects/metamaterial-compression/industry-grade/
// src/sim.cpp:51-135 and include/core.hpp:46-47 (MC/ below).
#include "lattice.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <unordered_set>

namespace meta {
namespace {

constexpr double kPi = 3.1415926535897932384626433832795;

struct P3 {
  double x, y, z;
};
P3 operator+(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
P3 operator-(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
P3 operator*(P3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dotp(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
P3 crossp(P3 a, P3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
P3 unitp(P3 v) {
  double l2 = dotp(v, v);
  return (std::isfinite(l2) && l2 > 1e-30) ? v * (1.0 / std::sqrt(l2)) : P3{0, 1, 0};
}
double clamp01(double v) { return std::max(0.0, std::min(1.0, v)); }

// splitmix64 finaliser and the RNG stream of MC/include/core.hpp:46
uint64_t splitmix(uint64_t v) {
  v += 0x9e3779b97f4a7c15ULL;
  v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
  v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
  return v ^ (v >> 31);
}
class Rng {
 public:
  explicit Rng(uint64_t seed) : s_(splitmix(seed)) {}
  double uniform() {
    s_ = splitmix(s_);
    return (double(uint32_t(s_ >> 32)) + 0.5) / 4294967296.0;
  }
  double signed_uniform() { return 2 * uniform() - 1; }

 private:
  uint64_t s_;
};
uint64_t pair_key(int u, int v) {
  if (u > v) std::swap(u, v);
  return (uint64_t(uint32_t(u)) << 32) | uint32_t(v);
}

class Builder {
 public:
  Builder(const mmc_config& c, HostLattice& L) : c_(c), L_(L) {}

  void bcc();
  void cubic();
  void finish();

 private:
  const mmc_config& c_;
  HostLattice& L_;
  std::vector<P3> pos_;
  std::unordered_set<uint64_t> excl_;

  int add_node(P3 p, int layer) {
    pos_.push_back(p);
    L_.node_layer.push_back(layer);
    return int(pos_.size()) - 1;
  }
  void exclude(uint64_t key) {
    if (excl_.insert(key).second) L_.connected.push_back(key);
  }
  void add_beam(int u, int v, int layer, int family, double r) {
    P3 d = pos_[size_t(v)] - pos_[size_t(u)];
    P3 t = unitp(d);
    L_.a.push_back(u);
    L_.b.push_back(v);
    L_.layer.push_back(layer);
    L_.family.push_back(family);
    L_.L0.push_back(std::sqrt(dotp(d, d)));
    L_.e0.insert(L_.e0.end(), {t.x, t.y, t.z});
    L_.radius.push_back(r);
    L_.area.push_back(kPi * r * r);
    L_.inertia.push_back(0.25 * kPi * std::pow(r, 4));
    exclude(pair_key(u, v));
  }
};

// Graded BCC accordion lattice, MC/src/sim.cpp:51-132.
void Builder::bcc() {
  const int nx = c_.cells_x, ny = c_.cells_y, nz = c_.cells_z;
  const double h = c_.cell_size, jr = L_.joint_radius;
  const int corners = (nx + 1) * (ny + 1) * (nz + 1);
  auto corner = [&](int i, int j, int k) { return (j * (nz + 1) + k) * (nx + 1) + i; };
  auto centre = [&](int i, int j, int k) { return corners + (j * nz + k) * nx + i; };
  const P3 origin{-0.5 * L_.width, jr, -0.5 * L_.depth};
  Rng rng(c_.seed);
  for (int j = 0; j <= ny; j++)
    for (int k = 0; k <= nz; k++)
      for (int i = 0; i <= nx; i++) {
        P3 p = origin + P3{i * h, j * h, k * h};
        double yf = double(j) / ny, env = std::sin(kPi * yf), sg = ((i + 2 * k + 3 * j) & 1) ? 1.0 : -1.0;
        p.x += c_.imperfection * env * (0.55 * sg + 0.45 * rng.signed_uniform());
        p.z += c_.imperfection * env * (0.35 * sg + 0.65 * rng.signed_uniform());
        add_node(p, j);
      }
  for (int j = 0; j < ny; j++)
    for (int k = 0; k < nz; k++)
      for (int i = 0; i < nx; i++) {
        P3 p = origin + P3{(i + 0.5) * h, (j + 0.5) * h, (k + 0.5) * h};
        double yf = (j + 0.5) / double(ny), env = std::sin(kPi * yf), sg = ((i + k + j) & 1) ? 1.0 : -1.0;
        double sx = (j & 1) ? -1.0 : 1.0, sz = ((j / 2) & 1) ? -1.0 : 1.0;
        p.x += h * (0.082 * (1 - 0.32 * yf) * sx + 0.018 * env * sg) + 1.2 * c_.imperfection * env * rng.signed_uniform();
        p.z += h * (0.032 * (0.65 + 0.35 * (1 - yf)) * sz - 0.012 * env * sg) +
               1.2 * c_.imperfection * env * rng.signed_uniform();
        add_node(p, j);
      }
  std::unordered_set<uint64_t> struts;
  auto graded_radius = [&](int u, int v, int layer, double scale) {
    P3 m = (pos_[size_t(u)] + pos_[size_t(v)]) * 0.5;
    double y = clamp01((m.y - jr) / L_.height);
    double radial = clamp01(std::sqrt(m.x * m.x + m.z * m.z) / (0.72 * std::max(L_.width, L_.depth)));
    double grade = 0.58 + 0.42 * std::pow(1 - y, 0.72);
    grade *= 0.94 + 0.12 * radial;
    if (layer == 0)
      grade *= c_.trigger_scale;
    else if (layer == 1)
      grade *= std::sqrt(c_.trigger_scale);
    else if (layer == 2)
      grade *= 0.88;
    double t = clamp01(grade);
    return scale * (c_.min_radius * (1.0 - t) + c_.max_radius * t);
  };
  auto strut = [&](int u, int v, int layer, int family, double scale) {
    const uint64_t key = pair_key(u, v);
    if (u == v || !struts.insert(key).second) return;
    const P3 pu = pos_[size_t(u)], pv = pos_[size_t(v)];
    const P3 tangent = unitp(pv - pu);
    const uint64_t hs = splitmix(c_.seed ^ key ^ uint64_t(family + 31) * 0x9e3779b97f4a7c15ULL);
    const P3 helper{double(int((hs >> 0) & 1023) - 511), double(int((hs >> 10) & 1023) - 511),
                    double(int((hs >> 20) & 1023) - 511)};
    P3 n1 = crossp(tangent, helper);
    if (dotp(n1, n1) < 1e-12) {
      P3 up = std::abs(tangent.z) < 0.8 ? P3{0, 0, 1} : P3{0, 1, 0};
      n1 = unitp(crossp(up, tangent));
    } else {
      n1 = unitp(n1);
    }
    const P3 n2 = unitp(crossp(tangent, n1));
    const double lf = layer == 0 ? 2.60 : (layer == 1 ? 1.60 : 1.0);
    const double amp = c_.member_imperfection * lf * (0.55 + 0.90 * double((hs >> 32) & 65535) / 65535.0);
    const int segs = std::clamp(c_.elements_per_strut, 2, 5);
    std::vector<int> chain{u};
    for (int s = 1; s < segs; s++) {
      double t = double(s) / segs;
      P3 p = ((pu * (1 - t) + pv * t) + n1 * (std::sin(kPi * t) * amp)) + n2 * (0.22 * std::sin(2 * kPi * t) * amp);
      chain.push_back(add_node(p, layer));
    }
    chain.push_back(v);
    const double r = graded_radius(u, v, layer, scale);
    for (size_t s = 0; s + 1 < chain.size(); s++) add_beam(chain[s], chain[s + 1], layer, family, r);
    exclude(key);
  };
  static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
  for (int j = 0; j < ny; j++)
    for (int k = 0; k < nz; k++)
      for (int i = 0; i < nx; i++)
        for (int q = 0; q < 8; q++)
          strut(centre(i, j, k), corner(i + off[q][0], j + off[q][1], k + off[q][2]), j, q, 1.0);
  for (int face : {0, ny}) {
    const int lay = std::clamp(face, 0, ny - 1);
    for (int k = 0; k <= nz; k++)
      for (int i = 0; i < nx; i++) strut(corner(i, face, k), corner(i + 1, face, k), lay, 30, 0.62);
    for (int i = 0; i <= nx; i++)
      for (int k = 0; k < nz; k++) strut(corner(i, face, k), corner(i, face, k + 1), lay, 31, 0.62);
  }
}

// Simple cubic lattice: grid nodes in the reference's corner order, one beam
// per cell edge emitted node by node (x, y, z edge), radius graded linearly in
// y between min_radius and max_radius at the beam midpoint, or drawn uniformly
// per beam (radius_mode 1).  Imperfection: Gaussian jitter of std
// position_sigma (Box-Muller on the reference RNG; y only on interior layers),
// or the reference's corner-node pattern (MC/src/sim.cpp:65-67) when sigma = 0.
void Builder::cubic() {
  const int nx = c_.cells_x, ny = c_.cells_y, nz = c_.cells_z;
  const double h = c_.cell_size, jr = L_.joint_radius;
  const P3 origin{-0.5 * L_.width, jr, -0.5 * L_.depth};
  Rng rng(c_.seed);
  auto normal_pair = [&](double& g0, double& g1) {
    double u1 = rng.uniform(), u2 = rng.uniform();
    double rr = std::sqrt(-2.0 * std::log(u1));
    g0 = rr * std::cos(2 * kPi * u2);
    g1 = rr * std::sin(2 * kPi * u2);
  };
  for (int j = 0; j <= ny; j++)
    for (int k = 0; k <= nz; k++)
      for (int i = 0; i <= nx; i++) {
        P3 p = origin + P3{i * h, j * h, k * h};
        if (c_.position_sigma > 0) {
          double g0, g1, g2, g3;
          normal_pair(g0, g1);
          normal_pair(g2, g3);
          p.x += c_.position_sigma * g0;
          p.z += c_.position_sigma * g1;
          if (j > 0 && j < ny) p.y += c_.position_sigma * g2;
        } else {
          double yf = double(j) / ny, env = std::sin(kPi * yf), sg = ((i + 2 * k + 3 * j) & 1) ? 1.0 : -1.0;
          p.x += c_.imperfection * env * (0.55 * sg + 0.45 * rng.signed_uniform());
          p.z += c_.imperfection * env * (0.35 * sg + 0.65 * rng.signed_uniform());
        }
        add_node(p, j);
      }
  auto id = [&](int i, int j, int k) { return (j * (nz + 1) + k) * (nx + 1) + i; };
  for (int j = 0; j <= ny; j++)
    for (int k = 0; k <= nz; k++)
      for (int i = 0; i <= nx; i++) {
        const int c = id(i, j, k);
        const int nbr[3] = {i < nx ? id(i + 1, j, k) : -1, j < ny ? id(i, j + 1, k) : -1,
                            k < nz ? id(i, j, k + 1) : -1};
        for (int f = 0; f < 3; f++) {
          if (nbr[f] < 0) continue;
          double r;
          if (c_.radius_mode == 1) {
            r = c_.min_radius + (c_.max_radius - c_.min_radius) * rng.uniform();
          } else {
            P3 m = (pos_[size_t(c)] + pos_[size_t(nbr[f])]) * 0.5;
            double t = clamp01((m.y - jr) / L_.height);
            r = c_.min_radius * (1.0 - t) + c_.max_radius * t;
          }
          add_beam(c, nbr[f], std::min(j, ny - 1), f, r);
        }
      }
}

// distributeMass, MC/src/sim.cpp:134: half of each beam's mass to each end plus
// a joint-sphere share; locks x,z of the bottom node nearest (-w/2,-d/2) and z
// of the one nearest (+w/2,-d/2).
void Builder::finish() {
  const size_t N = pos_.size();
  L_.n_nodes = int(N);
  L_.n_beams = int(L_.a.size());
  L_.x.resize(3 * N);
  for (size_t i = 0; i < N; i++) {
    L_.x[3 * i] = pos_[i].x;
    L_.x[3 * i + 1] = pos_[i].y;
    L_.x[3 * i + 2] = pos_[i].z;
  }
  L_.mass.assign(N, 0.0);
  L_.total_mass = 0;
  for (int k = 0; k < L_.n_beams; k++) {
    double m = c_.material.density * L_.area[size_t(k)] * L_.L0[size_t(k)];
    L_.mass[size_t(L_.a[size_t(k)])] += 0.5 * m;
    L_.mass[size_t(L_.b[size_t(k)])] += 0.5 * m;
    L_.total_mass += m;
  }
  const double jm = c_.material.density * (4.0 / 3.0) * kPi * std::pow(L_.joint_radius, 3) * 0.28;
  for (size_t i = 0; i < N; i++) {
    L_.mass[i] += jm;
    L_.total_mass += jm;
  }
  L_.locked.assign(6 * N, 0);
  int ia = -1, ib = -1;
  double da = std::numeric_limits<double>::infinity(), db = da;
  for (size_t i = 0; i < N; i++) {
    if (L_.node_layer[i] != 0) continue;
    const P3 p = pos_[i];
    double qa = (p.x + 0.5 * L_.width) * (p.x + 0.5 * L_.width) + (p.z + 0.5 * L_.depth) * (p.z + 0.5 * L_.depth);
    double qb = (p.x - 0.5 * L_.width) * (p.x - 0.5 * L_.width) + (p.z + 0.5 * L_.depth) * (p.z + 0.5 * L_.depth);
    if (qa < da) {
      da = qa;
      ia = int(i);
    }
    if (qb < db) {
      db = qb;
      ib = int(i);
    }
  }
  if (ia < 0 || ib < 0) throw std::runtime_error("failed to create anti-drift constraints");
  L_.locked[size_t(6 * ia)] = 1;
  L_.locked[size_t(6 * ia + 2)] = 1;
  L_.locked[size_t(6 * ib + 2)] = 1;
}

}  // namespace

mmc_lattice HostLattice::view() const {
  mmc_lattice v{};
  v.n_nodes = n_nodes;
  v.n_beams = n_beams;
  v.x = x.data();
  v.q = q.empty() ? nullptr : q.data();
  v.mass = mass.data();
  v.locked = locked.data();
  v.beam_a = a.data();
  v.beam_b = b.data();
  v.L0 = L0.data();
  v.radius = radius.data();
  v.area = area.data();
  v.inertia = inertia.data();
  v.e0 = e0.data();
  v.plastic = plastic.empty() ? nullptr : plastic.data();
  v.n_connected = int64_t(connected.size());
  v.connected = connected.data();
  v.node_layer = node_layer.data();
  v.beam_layer = layer.data();
  v.beam_family = family.data();
  v.width = width;
  v.height = height;
  v.depth = depth;
  return v;
}

HostLattice build_lattice(const mmc_config& c) {
  if (c.cells_x < 1 || c.cells_y < 1 || c.cells_z < 1) throw std::runtime_error("cell counts must be positive");
  if (!(c.cell_size > 0) || !(c.max_radius > 0)) throw std::runtime_error("cell size and radius must be positive");
  HostLattice L;
  L.width = c.cells_x * c.cell_size;
  L.height = c.cells_y * c.cell_size;
  L.depth = c.cells_z * c.cell_size;
  L.area_ref = L.width * L.depth;
  L.joint_radius = c.joint_scale * c.max_radius;
  Builder b(c, L);
  if (c.lattice == MMC_LATTICE_CUBIC)
    b.cubic();
  else if (c.lattice == MMC_LATTICE_BCC)
    b.bcc();
  else
    throw std::runtime_error("unknown lattice family");
  b.finish();
  L.top0 = L.joint_radius + L.height + L.joint_radius;
  return L;
}

}  // namespace meta
