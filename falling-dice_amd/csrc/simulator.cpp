// simulator.cpp — the Simulator facade over the C-ABI (see simulator.hpp) and
// the mmc_sim_* C entry points.  Reference: projects/metamaterial-compression/
// industry-grade/src/sim.cpp (MC/src/sim.cpp below).
#include "simulator.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <iomanip>
#include <iostream>
#include <mutex>
#include <thread>
#include <limits>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

namespace meta {
namespace {

constexpr double kPi = 3.1415926535897932384626433832795;
constexpr double kInf = std::numeric_limits<double>::infinity();

Vec3 add(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
Vec3 mul(Vec3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot3(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
Vec3 cross3(Vec3 a, Vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
Vec3 unit3(Vec3 v) {
  double l2 = dot3(v, v);
  return (std::isfinite(l2) && l2 > 1e-30) ? mul(v, 1.0 / std::sqrt(l2)) : Vec3{0, 1, 0};
}
Quat qn(Quat q) {
  double n = std::sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
  if (n < 1e-30) return Quat{};
  return {q.w / n, q.x / n, q.y / n, q.z / n};
}
Quat qm(Quat a, Quat b) {
  return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
          a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x, a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}
Vec3 qr(Quat q, Vec3 v) {
  Vec3 u{q.x, q.y, q.z};
  Vec3 t = mul(cross3(u, v), 2.0);
  return add(add(v, mul(t, q.w)), cross3(u, t));
}
Quat q_axis_angle(Vec3 axis, double angle) {
  axis = unit3(axis);
  double s = std::sin(0.5 * angle);
  return qn({std::cos(0.5 * angle), axis.x * s, axis.y * s, axis.z * s});
}
uint64_t splitmix(uint64_t v) {
  v += 0x9e3779b97f4a7c15ULL;
  v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
  v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
  return v ^ (v >> 31);
}

}  // namespace

mmc_config to_c(const SimulationConfig& s, const SimulationExtensions& e) {
  mmc_config c;
  mmc_config_defaults(&c);
  c.cells_x = s.cellsX;
  c.cells_y = s.cellsY;
  c.cells_z = s.cellsZ;
  c.elements_per_strut = s.elementsPerStrut;
  c.cell_size = s.cellSize;
  c.min_radius = s.minRadius;
  c.max_radius = s.maxRadius;
  c.joint_scale = s.jointScale;
  c.trigger_scale = s.triggerScale;
  c.imperfection = s.imperfection;
  c.member_imperfection = s.memberImperfection;
  c.material = mmc_material{s.material.density, s.material.young, s.material.nu, s.material.yield,
                            s.material.hardening, s.material.bendHardening, s.material.torsionHardening};
  c.gravity = s.gravity;
  c.shear_correction = s.shearCorrection;
  c.contact_stiffness = s.contactStiffness;
  c.contact_margin = s.contactMargin;
  c.centering_stiffness = s.centeringStiffness;
  c.contact_samples = s.contactSamples;
  c.load_steps = s.loadSteps;
  c.unload_steps = s.unloadSteps;
  c.max_iterations = s.maxIterations;
  c.history = s.history;
  c.max_line_search = s.maxLineSearch;
  c.max_adaptive_refinements = s.maxAdaptiveRefinements;
  c.fps = s.fps;
  c.frames = s.frames;
  c.threads = s.threads;
  c.max_strain = s.maxStrain;
  c.step_time = s.stepTime;
  c.translation_scale = s.translationScale;
  c.max_translation_step = s.maxTranslationStep;
  c.max_rotation_step = s.maxRotationStep;
  c.gradient_rms_tolerance = s.gradientRmsTolerance;
  c.gradient_max_tolerance = s.gradientMaxTolerance;
  c.armijo = s.armijo;
  c.min_line_step = s.minLineStep;
  c.penetration_tolerance = s.penetrationTolerance;
  c.seed = s.seed;
  c.lattice = e.lattice;
  c.relative_escape = e.relativeEscape ? 1 : 0;
  c.relative_tolerance = e.relativeTolerance;
  c.position_sigma = e.positionSigma;
  c.radius_mode = e.radiusMode;
  c.line_search = e.lineSearch;
  return c;
}

void from_c(const mmc_config& c, SimulationConfig* out, SimulationExtensions* ext) {
  SimulationConfig s;
  s.cellsX = c.cells_x;
  s.cellsY = c.cells_y;
  s.cellsZ = c.cells_z;
  s.elementsPerStrut = c.elements_per_strut;
  s.cellSize = c.cell_size;
  s.minRadius = c.min_radius;
  s.maxRadius = c.max_radius;
  s.jointScale = c.joint_scale;
  s.triggerScale = c.trigger_scale;
  s.imperfection = c.imperfection;
  s.memberImperfection = c.member_imperfection;
  s.material.density = c.material.density;
  s.material.young = c.material.young;
  s.material.nu = c.material.nu;
  s.material.yield = c.material.yield;
  s.material.hardening = c.material.hardening;
  s.material.bendHardening = c.material.bend_hardening;
  s.material.torsionHardening = c.material.torsion_hardening;
  s.gravity = c.gravity;
  s.shearCorrection = c.shear_correction;
  s.contactStiffness = c.contact_stiffness;
  s.contactMargin = c.contact_margin;
  s.centeringStiffness = c.centering_stiffness;
  s.contactSamples = c.contact_samples;
  s.loadSteps = c.load_steps;
  s.unloadSteps = c.unload_steps;
  s.maxIterations = c.max_iterations;
  s.history = c.history;
  s.maxLineSearch = c.max_line_search;
  s.maxAdaptiveRefinements = c.max_adaptive_refinements;
  s.fps = c.fps;
  s.frames = c.frames;
  s.threads = c.threads;
  s.maxStrain = c.max_strain;
  s.stepTime = c.step_time;
  s.translationScale = c.translation_scale;
  s.maxTranslationStep = c.max_translation_step;
  s.maxRotationStep = c.max_rotation_step;
  s.gradientRmsTolerance = c.gradient_rms_tolerance;
  s.gradientMaxTolerance = c.gradient_max_tolerance;
  s.armijo = c.armijo;
  s.minLineStep = c.min_line_step;
  s.penetrationTolerance = c.penetration_tolerance;
  s.seed = c.seed;
  if (out) *out = s;
  if (ext) {
    ext->lattice = c.lattice;
    ext->relativeEscape = c.relative_escape != 0;
    ext->relativeTolerance = c.relative_tolerance;
    ext->positionSigma = c.position_sigma;
    ext->radiusMode = c.radius_mode;
    ext->lineSearch = c.line_search;
  }
}

// ---------------------------------------------------------------------------
// Extensions handed to the next Simulator of a thread, and per instance
// ---------------------------------------------------------------------------
namespace {
thread_local std::unique_ptr<SimulationExtensions> t_next_ext;
std::mutex g_ext_mu;
std::unordered_map<const Simulator*, SimulationExtensions> g_ext;

SimulationExtensions take_next_extensions() {
  SimulationExtensions e;
  if (t_next_ext) {
    e = *t_next_ext;
    t_next_ext.reset();
  }
  return e;
}
}  // namespace

void setNextSimulationExtensions(const SimulationExtensions& ext) {
  t_next_ext = std::make_unique<SimulationExtensions>(ext);
}

SimulationExtensions simulationExtensions(const Simulator& sim) {
  std::lock_guard<std::mutex> lock(g_ext_mu);
  auto it = g_ext.find(&sim);
  return it == g_ext.end() ? SimulationExtensions{} : it->second;
}

// ---------------------------------------------------------------------------
// Simulator: the reference's members, the hot path through a Driver
// ---------------------------------------------------------------------------
// MC/src/sim.cpp:49: the lattice is built on the host (build, distributeMass)
// into the reference's members; the device context exists while run() or a
// check needs it.
Simulator::Simulator(SimulationConfig cfg) : cfg_(std::move(cfg)) {
  const SimulationExtensions ext = take_next_extensions();
  cfg_.translationScale = cfg_.cellSize;
  const HostLattice L = build_lattice(to_c(cfg_, ext));
  {
    std::lock_guard<std::mutex> lock(g_ext_mu);
    g_ext[this] = ext;
  }
  const size_t N = size_t(L.n_nodes), B = size_t(L.n_beams);
  nodes_.resize(N);
  baseX_.resize(N);
  baseR_.assign(N, Quat{});
  for (size_t i = 0; i < N; i++) {
    Node& n = nodes_[i];
    n.x0 = n.x = baseX_[i] = Vec3{L.x[3 * i], L.x[3 * i + 1], L.x[3 * i + 2]};
    n.R = Quat{};
    n.mass = L.mass[i];
    n.layer = L.node_layer[i];
  }
  beams_.resize(B);
  for (size_t k = 0; k < B; k++) {
    Beam& e = beams_[k];
    e.a = L.a[k];
    e.b = L.b[k];
    e.layer = L.layer[k];
    e.family = L.family[k];
    e.L0 = L.L0[k];
    e.radius = L.radius[k];
    e.area = L.area[k];
    e.I = L.inertia[k];
    e.J = 2 * e.I;
    e.e0 = Vec3{L.e0[3 * k], L.e0[3 * k + 1], L.e0[3 * k + 2]};
  }
  connected_.insert(L.connected.begin(), L.connected.end());
  locked_ = L.locked;
  width_ = L.width;
  height_ = L.height;
  depth_ = L.depth;
  area_ = L.area_ref;
  mass_ = L.total_mass;
  top0_ = L.top0;
}

// The members as the device takes them: base pose and plastic state as
// committed, the initial positions for the objectivity test, the exclusion
// set in the construction's insertion order.
HostLattice Simulator::committedLattice() const {
  const SimulationExtensions ext = simulationExtensions(*this);
  HostLattice L = build_lattice(to_c(cfg_, ext));  // connectivity, masses, locks, envelope
  const size_t N = nodes_.size(), B = beams_.size();
  if (size_t(L.n_nodes) != N || size_t(L.n_beams) != B) throw std::runtime_error("lattice size mismatch");
  L.x0 = L.x;
  L.q.resize(4 * N);
  for (size_t i = 0; i < N; i++) {
    const Vec3& x = baseX_[i];
    const Quat& q = baseR_[i];
    L.x[3 * i] = x.x, L.x[3 * i + 1] = x.y, L.x[3 * i + 2] = x.z;
    L.q[4 * i] = q.w, L.q[4 * i + 1] = q.x, L.q[4 * i + 2] = q.y, L.q[4 * i + 3] = q.z;
  }
  L.plastic.resize(7 * B);
  for (size_t k = 0; k < B; k++) {
    const PlasticState& p = beams_[k].plastic;
    double* o = &L.plastic[7 * k];
    o[0] = p.epsp, o[1] = p.alpha, o[2] = p.rotp.x, o[3] = p.rotp.y, o[4] = p.rotp.z, o[5] = p.alphaB,
    o[6] = p.alphaT;
  }
  return L;
}

// MC/src/sim.cpp:374-467 on the device; the committed state of every accepted
// increment is copied back into the members (also when an equilibrium failure
// ends the schedule, as the reference's own commits leave them).
SimulationResult Simulator::run() {
  const SimulationExtensions ext = simulationExtensions(*this);
  Driver d(to_c(cfg_, ext), ext.precision, committedLattice());
  d.setVerbose(ext.verbose);
  d.setCaptureFrames(ext.captureFrames);
  auto copy_back = [&] {
    const std::vector<Node>& n = d.nodes();
    const std::vector<Beam>& b = d.beams();
    for (size_t i = 0; i < nodes_.size(); i++) {
      nodes_[i].x = baseX_[i] = n[i].x;
      nodes_[i].R = baseR_[i] = n[i].R;
    }
    for (size_t k = 0; k < beams_.size(); k++) beams_[k].plastic = b[k].plastic;
    loadingWork_ = d.loadingWork();
    recoveredWork_ = d.recoveredWork();
    prevDisp_ = d.prevDisp();
    prevForce_ = d.prevForce();
    reactionOffset_ = d.reactionOffset();
  };
  SimulationResult r;
  try {
    r = d.run();
  } catch (...) {
    copy_back();
    throw;
  }
  copy_back();
  return r;
}

namespace {
// MC/src/sim.cpp:645-888 through mmc_export_stl: the union of the capsules and
// joint spheres at the initial positions x0
void export_stl_of(const std::vector<Vec3>& x0, const std::vector<int>& a, const std::vector<int>& b,
                   const std::vector<double>& radius, double jr, const std::filesystem::path& path, int segments,
                   mmc_stl_report* rep) {
  if (x0.empty() || a.empty()) throw std::runtime_error("cannot mesh empty lattice");
  std::vector<double> x(3 * x0.size());
  for (size_t i = 0; i < x0.size(); i++) x[3 * i] = x0[i].x, x[3 * i + 1] = x0[i].y, x[3 * i + 2] = x0[i].z;
  mmc_lattice v{};
  v.n_nodes = int32_t(x0.size());
  v.n_beams = int32_t(a.size());
  v.x = x.data();
  v.beam_a = a.data();
  v.beam_b = b.data();
  v.radius = radius.data();
  const std::string p = path.string();
  if (mmc_export_stl(&v, jr, segments, p.c_str(), rep) != MMC_OK) throw std::runtime_error(mmc_stl_last_error());
}
}  // namespace

void Simulator::exportSTL(const std::filesystem::path& path, int radialSegments) const {
  if (nodes_.empty() || beams_.empty()) throw std::runtime_error("cannot mesh empty lattice");
  std::vector<Vec3> x0(nodes_.size());
  for (size_t i = 0; i < nodes_.size(); i++) x0[i] = nodes_[i].x0;
  std::vector<int> a(beams_.size()), b(beams_.size());
  std::vector<double> r(beams_.size());
  for (size_t k = 0; k < beams_.size(); k++) a[k] = beams_[k].a, b[k] = beams_[k].b, r[k] = beams_[k].radius;
  export_stl_of(x0, a, b, r, cfg_.jointScale * cfg_.maxRadius, path, radialSegments, nullptr);
}

void Driver::exportSTL(const std::filesystem::path& path, int radialSegments, mmc_stl_report* report) const {
  const std::vector<double>& xs = lat_.x0.empty() ? lat_.x : lat_.x0;
  std::vector<Vec3> x0(size_t(lat_.n_nodes));
  for (size_t i = 0; i < x0.size(); i++) x0[i] = Vec3{xs[3 * i], xs[3 * i + 1], xs[3 * i + 2]};
  export_stl_of(x0, lat_.a, lat_.b, lat_.radius, cfg_.jointScale * cfg_.maxRadius, path, radialSegments, report);
}

// MC/src/sim.cpp:511-536 and :538-577, on the committed state, through a
// device context made for the check.
bool Simulator::rigidRotationTest(double* error) const {
  const SimulationExtensions ext = simulationExtensions(*this);
  Driver d(to_c(cfg_, ext), ext.precision, committedLattice());
  return d.rigidRotationTest(error);
}

bool Simulator::gradientCheck(double* relativeError) const {
  const SimulationExtensions ext = simulationExtensions(*this);
  Driver d(to_c(cfg_, ext), ext.precision, committedLattice());
  return d.gradientCheck(relativeError);
}

// ---------------------------------------------------------------------------
// Driver
// ---------------------------------------------------------------------------
Driver::Driver(const mmc_config& c, int precision) : c_(c) {
  c_.translation_scale = c_.cell_size;  // MC/src/sim.cpp:49
  lat_ = build_lattice(c_);
  create(precision);
}

Driver::Driver(const mmc_config& c, int precision, HostLattice lat) : c_(c), lat_(std::move(lat)) {
  c_.translation_scale = c_.cell_size;
  create(precision);
}

void Driver::create(int precision) {
  from_c(c_, &cfg_, &ext_);
  ext_.precision = precision;
  mmc_lattice v = lat_.view();
  int rc = mmc_create(&v, &c_, precision, &ctx_);
  if (rc != MMC_OK) throw std::runtime_error(mmc_last_error(nullptr));
}

Driver::~Driver() { mmc_destroy(ctx_); }

void Driver::check(int rc) const {
  if (rc != MMC_OK) throw std::runtime_error(mmc_last_error(ctx_));
}

// Host mirror of the committed lattice state (read from the device on demand).
void Driver::sync() const {
  if (synced_) return;
  const int N = lat_.n_nodes, B = lat_.n_beams;
  std::vector<double> x(3 * size_t(N)), q(4 * size_t(N)), p(7 * size_t(B));
  check(mmc_read_state(ctx_, x.data(), q.data(), p.data()));
  nodes_.resize(size_t(N));
  beams_.resize(size_t(B));
  for (int i = 0; i < N; i++) {
    Node& n = nodes_[size_t(i)];
    const std::vector<double>& x0 = lat_.x0.empty() ? lat_.x : lat_.x0;
    n.x0 = Vec3{x0[3 * size_t(i)], x0[3 * size_t(i) + 1], x0[3 * size_t(i) + 2]};
    n.x = Vec3{x[3 * size_t(i)], x[3 * size_t(i) + 1], x[3 * size_t(i) + 2]};
    n.R = Quat{q[4 * size_t(i)], q[4 * size_t(i) + 1], q[4 * size_t(i) + 2], q[4 * size_t(i) + 3]};
    n.mass = lat_.mass[size_t(i)];
    n.layer = lat_.node_layer[size_t(i)];
  }
  for (int k = 0; k < B; k++) {
    Beam& e = beams_[size_t(k)];
    size_t s = size_t(k);
    e.a = lat_.a[s];
    e.b = lat_.b[s];
    e.layer = lat_.layer[s];
    e.family = lat_.family[s];
    e.L0 = lat_.L0[s];
    e.radius = lat_.radius[s];
    e.area = lat_.area[s];
    e.I = lat_.inertia[s];
    e.J = 2 * e.I;
    e.e0 = Vec3{lat_.e0[3 * s], lat_.e0[3 * s + 1], lat_.e0[3 * s + 2]};
    const double* pp = &p[7 * s];
    e.plastic = PlasticState{pp[0], pp[1], Vec3{pp[2], pp[3], pp[4]}, pp[5], pp[6]};
  }
  synced_ = true;
}

const std::vector<Node>& Driver::nodes() const {
  sync();
  return nodes_;
}
const std::vector<Beam>& Driver::beams() const {
  sync();
  return beams_;
}

// MC/src/sim.cpp:213
double Driver::topY(double s, bool loading) const {
  double low = lat_.top0 - cfg_.maxStrain * lat_.height;
  double t = std::max(0.0, std::min(1.0, s));
  return loading ? lat_.top0 * (1.0 - t) + low * t : low * (1.0 - t) + lat_.top0 * t;
}

// MC/src/sim.cpp:215-310, computed on the device from the committed state
// (mmc_capture): the velocities are those against the poses stored by the
// mmc_capture_mark before this increment's minimisation, or zero on the
// schedule's first frame (the reference's prevX = prevR = nullptr).
static_assert(sizeof(Snapshot) == sizeof(mmc_snapshot), "Snapshot and mmc_snapshot differ in size");
static_assert(offsetof(Snapshot, q) == offsetof(mmc_snapshot, q) && offsetof(Snapshot, v) == offsetof(mmc_snapshot, v) &&
                  offsetof(Snapshot, w) == offsetof(mmc_snapshot, w) &&
                  offsetof(Snapshot, half) == offsetof(mmc_snapshot, half) &&
                  offsetof(Snapshot, material) == offsetof(mmc_snapshot, material) &&
                  offsetof(Snapshot, slot) == offsetof(mmc_snapshot, slot),
              "Snapshot and mmc_snapshot differ in layout");
std::vector<Snapshot> Driver::capture(double top, bool velocities) const {
  std::vector<Snapshot> out(size_t(lat_.n_beams) + size_t(lat_.n_nodes) + MMC_MACHINE_BODIES);
  const int32_t one = 1, vel = velocities ? 1 : 0;
  check(mmc_capture(ctx_, &top, &one, &vel, reinterpret_cast<mmc_snapshot*>(out.data())));
  return out;
}

// MC/src/sim.cpp:312-372; plastic and height reductions come from the device.
namespace {
// MC/src/sim.cpp:312-372 from the final objective record, the optimiser
// statistics and the device reductions over the committed state.
FrameDiagnostics make_diagnostics(const HostLattice& lat, double time, double top, bool loading,
                                  const mmc_objective_out& o, const mmc_opt_stats& s, const mmc_state_stats& st,
                                  double reactionOffset, double loadingWork, double recoveredWork) {
  FrameDiagnostics d;
  const double height = lat.height;
  d.time = time;
  d.platenY = top;
  d.displacement = lat.top0 - top;
  d.engineeringStrain = d.displacement / std::max(1e-12, height);
  d.reactionForce = std::max(0.0, o.reaction - reactionOffset);
  d.engineeringStress = d.reactionForce / std::max(1e-12, lat.area_ref);
  d.loadingWork = loadingWork;
  d.recoveredWork = recoveredWork;
  d.dissipatedWork = std::max(0.0, loadingWork - recoveredWork);
  d.specificEnergyAbsorption = d.dissipatedWork / std::max(1e-12, lat.total_mass);
  d.elasticEnergy = o.elastic;
  d.plasticDissipation = o.plastic;
  d.contactEnergy = o.contact;
  d.residualRms = s.rms;
  d.residualMax = s.max;
  d.maxPenetration = o.max_pen;
  d.activeContacts = o.contacts;
  d.iterations = s.iterations;
  d.converged = s.converged != 0;
  d.loading = loading;
  const double B = double(lat.n_beams);
  if (lat.n_beams > 0) {
    d.meanPlasticStrain = st.sum_axial / B;
    d.maxPlasticStrain = st.max_axial;
    d.meanPlasticRotation = st.sum_rot / B;
    d.maxPlasticRotation = st.max_rot;
  }
  d.plasticBeams = st.plastic_beams;
  d.plasticHinges = st.plastic_hinges;
  const double hcur = std::max(0.0, st.y_max - st.y_min);
  d.permanentSet = std::max(0.0, 1.0 - hcur / std::max(1e-12, height));
  return d;
}
}  // namespace

FrameDiagnostics Driver::diagnostics(double time, double top, bool loading, const mmc_objective_out& o,
                                        const mmc_opt_stats& s) const {
  mmc_state_stats st{};
  check(mmc_read_stats(ctx_, &st));
  return make_diagnostics(lat_, time, top, loading, o, s, st, reactionOffset_, loadingWork_, recoveredWork_);
}

// MC/src/sim.cpp:379-388
void Driver::begin() {
  loadingWork_ = recoveredWork_ = prevDisp_ = prevForce_ = reactionOffset_ = 0.0;
  time_ = 0.0;
  currentTop_ = lat_.top0;
  first_ = true;
  acceptedSteps_ = 0;
  cursor_ = 0;
}

// One equilibrium increment with adaptive bisection, MC/src/sim.cpp:390-454.
void Driver::advanceTo(double targetTop, bool loading, int refinement, SimulationResult& result) {
  // the committed poses before this increment, for the frame's velocities
  // (MC/src/sim.cpp:391-397), kept on the device
  if (ext_.captureFrames) check(mmc_capture_mark(ctx_));
  mmc_opt_stats stats{};
  mmc_objective_out fin{};
  check(mmc_minimize(ctx_, targetTop, &stats, &fin, nullptr));
  const bool acceptable = stats.converged && fin.max_pen <= cfg_.penetrationTolerance && std::isfinite(fin.energy);
  if (!acceptable && refinement < cfg_.maxAdaptiveRefinements && std::abs(targetTop - currentTop_) > 2.5e-5) {
    const double midpoint = 0.5 * (currentTop_ + targetTop);
    advanceTo(midpoint, loading, refinement + 1, result);
    advanceTo(targetTop, loading, refinement + 1, result);
    return;
  }
  if (!acceptable) {
    std::ostringstream message;
    message << "quasi-static equilibrium failed at platen=" << targetTop << " m, residual=" << stats.rms
            << ", penetration=" << fin.max_pen * 1000.0 << " mm after " << stats.iterations << " iterations";
    throw std::runtime_error(message.str());
  }
  check(mmc_commit(ctx_));
  synced_ = false;
  if (first_) reactionOffset_ = fin.reaction;
  const double correctedForce = std::max(0.0, fin.reaction - reactionOffset_);
  const double displacement = lat_.top0 - targetTop;
  if (!first_) {
    const double dd = displacement - prevDisp_;
    const double work = 0.5 * (prevForce_ + correctedForce) * dd;
    if (dd >= 0.0)
      loadingWork_ += std::max(0.0, work);
    else
      recoveredWork_ += std::max(0.0, -work);
  }
  prevDisp_ = displacement;
  prevForce_ = correctedForce;
  currentTop_ = targetTop;
  if (ext_.captureFrames)
    result.frames.push_back(capture(targetTop, !first_));
  result.diagnostics.push_back(diagnostics(time_, targetTop, loading, fin, stats));
  const FrameDiagnostics& d = result.diagnostics.back();
  if (ext_.verbose)
    std::cout << "[equilibrium] " << std::setw(3) << acceptedSteps_ << " " << (loading ? "load  " : "unload")
              << " strain=" << std::fixed << std::setprecision(4) << d.engineeringStrain
              << " force=" << std::setprecision(3) << d.reactionForce << " N"
              << " residual=" << std::scientific << std::setprecision(2) << d.residualRms
              << " pen=" << std::fixed << std::setprecision(3) << d.maxPenetration * 1000.0 << " mm"
              << " iter=" << d.iterations << (refinement ? " [REFINED]" : " [OK]") << '\n';
  acceptedSteps_++;
  first_ = false;
  time_ += cfg_.stepTime;
}

// Next scheduled increment, MC/src/sim.cpp:456-465.  Returns false once the
// schedule is complete.
bool Driver::advance(SimulationResult& result) {
  const int L = cfg_.loadSteps, U = cfg_.unloadSteps;
  if (cursor_ > L + U) return false;
  const int k = cursor_++;
  if (k == 0)
    advanceTo(lat_.top0, true, 0, result);  // gravity/contact equilibrium and load-cell tare
  else if (k <= L)
    advanceTo(topY(double(k) / std::max(1, L), true), true, 0, result);
  else
    advanceTo(topY(double(k - L) / std::max(1, U), false), false, 0, result);
  return cursor_ <= L + U;
}

// MC/src/sim.cpp:374-467
SimulationResult Driver::run() {
  SimulationResult result;
  result.frames.reserve(size_t(cfg_.loadSteps + cfg_.unloadSteps + 24));
  result.diagnostics.reserve(size_t(cfg_.loadSteps + cfg_.unloadSteps + 24));
  begin();
  while (advance(result)) {
  }
  return result;
}

// MC/src/sim.cpp:511-536; the potential is evaluated on the device.
bool Driver::rigidRotationTest(double* error) const {
  const int N = lat_.n_nodes;
  const double jr = cfg_.jointScale * cfg_.maxRadius, height = lat_.height;
  std::vector<Vec3> x(static_cast<size_t>(N));
  std::vector<Quat> R(static_cast<size_t>(N));
  for (int i = 0; i < N; ++i) {
    const std::vector<double>& xi = lat_.x0.empty() ? lat_.x : lat_.x0;  // Node::x0
    Vec3 x0{xi[3 * size_t(i)], xi[3 * size_t(i) + 1], xi[3 * size_t(i) + 2]};
    double yf = std::max(0.0, std::min(1.0, (x0.y - jr) / std::max(1e-12, height)));
    x[size_t(i)] = x0;
    x[size_t(i)].y -= 0.018 * height * yf;
    x[size_t(i)].x += 0.00018 * std::sin(2.0 * kPi * yf + 0.17 * double(i));
    R[size_t(i)] = q_axis_angle({0, 0, 1}, 0.035 * yf);
  }
  auto potential = [&](const std::vector<Vec3>& xs, const std::vector<Quat>& Rs) {
    std::vector<double> xa(3 * size_t(N)), qa(4 * size_t(N));
    for (int i = 0; i < N; i++) {
      xa[3 * size_t(i)] = xs[size_t(i)].x;
      xa[3 * size_t(i) + 1] = xs[size_t(i)].y;
      xa[3 * size_t(i) + 2] = xs[size_t(i)].z;
      qa[4 * size_t(i)] = Rs[size_t(i)].w;
      qa[4 * size_t(i) + 1] = Rs[size_t(i)].x;
      qa[4 * size_t(i) + 2] = Rs[size_t(i)].y;
      qa[4 * size_t(i) + 3] = Rs[size_t(i)].z;
    }
    double e = 0;
    check(mmc_beam_potential(ctx_, xa.data(), qa.data(), &e));
    return e;
  };
  const double e0 = potential(x, R);
  const Quat Q = q_axis_angle(unit3(Vec3{0.37, 0.81, -0.44}), 1.137);
  const Vec3 shift{0.081, -0.024, 0.053};
  std::vector<Vec3> xr(static_cast<size_t>(N));
  std::vector<Quat> Rr(static_cast<size_t>(N));
  for (int i = 0; i < N; ++i) {
    xr[size_t(i)] = add(qr(Q, x[size_t(i)]), shift);
    Rr[size_t(i)] = qn(qm(Q, R[size_t(i)]));
  }
  const double e1 = potential(xr, Rr);
  const double rel = std::abs(e1 - e0) / std::max({1e-12, std::abs(e0), std::abs(e1)});
  if (error) *error = rel;
  return rel < 1e-9;
}

// MC/src/sim.cpp:538-577; every objective evaluation runs on the device.
bool Driver::gradientCheck(double* relativeError) const {
  const int N = lat_.n_nodes;
  std::vector<double> z(6 * size_t(N), 0.0);
  // the reference RNG stream (MC/include/core.hpp:46) seeded with seed ^ "GRADCHEK"
  uint64_t state = splitmix(cfg_.seed ^ 0x475241444348454bULL);
  auto uniform = [&]() {
    state = splitmix(state);
    return (double(uint32_t(state >> 32)) + 0.5) / 4294967296.0;
  };
  auto signed_uniform = [&]() { return 2 * uniform() - 1; };
  for (int i = 0; i < N; ++i) {
    const size_t b = 6 * size_t(i);
    z[b + 0] = 0.003 * signed_uniform();
    z[b + 1] = 0.100 + 0.003 * signed_uniform();  // lift off the lower platen
    z[b + 2] = 0.003 * signed_uniform();
    z[b + 3] = 0.004 * signed_uniform();
    z[b + 4] = 0.004 * signed_uniform();
    z[b + 5] = 0.004 * signed_uniform();
    for (int d = 0; d < 6; ++d)
      if (lat_.locked[b + size_t(d)]) z[b + size_t(d)] = 0.0;
  }
  const double top = lat_.top0 + 0.020;
  std::vector<double> grad(z.size());
  mmc_objective_out o{};
  check(mmc_objective(ctx_, z.data(), top, 0, &o, grad.data()));
  double num = 0, den = 0;
  int checked = 0;
  for (size_t k = 0; k < z.size() && checked < 42; ++k) {
    if (lat_.locked[k]) continue;
    if ((splitmix(k * 0x9e3779b97f4a7c15ULL + cfg_.seed) & 15ULL) > 1ULL) continue;
    const double h = (k % 6 < 3) ? 2e-6 : 1e-6;
    std::vector<double> zp = z, zm = z;
    zp[k] += h;
    zm[k] -= h;
    mmc_objective_out op{}, om{};
    check(mmc_objective(ctx_, zp.data(), top, 0, &op, nullptr));
    check(mmc_objective(ctx_, zm.data(), top, 0, &om, nullptr));
    const double numeric = (op.energy - om.energy) / (2.0 * h), analytic = grad[k];
    num += (numeric - analytic) * (numeric - analytic);
    den += numeric * numeric + analytic * analytic;
    ++checked;
  }
  const double rel = std::sqrt(num / std::max(1e-30, den));
  if (relativeError) *relativeError = rel;
  return checked >= 24 && rel < 0.075;
}

// ---------------------------------------------------------------------------
// BatchSimulator
// ---------------------------------------------------------------------------
BatchSimulator::BatchSimulator(std::vector<mmc_config> cfgs, int precision, bool generate, int shards)
    : c_(std::move(cfgs)) {
  if (c_.empty()) throw std::runtime_error("a batch needs at least one lattice");
  const size_t n = c_.size();
  cfgs_.resize(n);
  lats_.resize(n);
  for (size_t b = 0; b < n; b++) {
    c_[b].translation_scale = c_[b].cell_size;  // MC/src/sim.cpp:49
    from_c(c_[b], &cfgs_[b], nullptr);
  }
  if (shards <= 0) {
    const char* e = std::getenv("MMC_BATCH_SHARDS");
    shards = e ? std::max(1, std::atoi(e)) : 1;
  }
  shards = std::min<int>(shards, int(n));
  if (shards > 1) {
    shard_lattices_.resize(size_t(shards));
    for (size_t b = 0; b < n; b++) shard_lattices_[b % size_t(shards)].push_back(int(b));
    for (const std::vector<int>& idx : shard_lattices_) {
      std::vector<mmc_config> sub;
      for (int b : idx) sub.push_back(c_[size_t(b)]);
      shards_.push_back(std::make_unique<BatchSimulator>(std::move(sub), precision, generate, 1));
      const BatchSimulator& sh = *shards_.back();
      for (size_t k = 0; k < idx.size(); k++) lats_[size_t(idx[k])] = sh.lats_[k];
    }
    return;
  }
  if (generate) {
    // built on the device: the host keeps each lattice's counts, envelope,
    // mass and platen (MC/src/sim.cpp:56,131,134)
    int rc = mmc_create_generated(c_.data(), int32_t(n), precision, &ctx_);
    if (rc != MMC_OK) throw std::runtime_error(mmc_last_error(nullptr));
    for (size_t b = 0; b < n; b++) {
      mmc_lattice_info info{};
      check(mmc_lattice_info_get(ctx_, int32_t(b), &info));
      HostLattice& L = lats_[b];
      L.n_nodes = info.n_nodes;
      L.n_beams = info.n_beams;
      L.width = info.width;
      L.height = info.height;
      L.depth = info.depth;
      L.area_ref = L.width * L.depth;
      L.joint_radius = c_[b].joint_scale * c_[b].max_radius;
      L.total_mass = info.mass;
      L.top0 = L.joint_radius + L.height + L.joint_radius;
    }
    return;
  }
  std::vector<mmc_lattice> views(n);
  for (size_t b = 0; b < n; b++) {
    lats_[b] = build_lattice(c_[b]);
    views[b] = lats_[b].view();
  }
  int rc = mmc_create_batch(views.data(), c_.data(), int32_t(n), precision, &ctx_);
  if (rc != MMC_OK) throw std::runtime_error(mmc_last_error(nullptr));
}

BatchSimulator::~BatchSimulator() { mmc_destroy(ctx_); }

void BatchSimulator::check(int rc) const {
  if (rc != MMC_OK) throw std::runtime_error(mmc_last_error(ctx_));
}

namespace {
// One lattice's schedule (MC/src/sim.cpp:374-467) as an explicit state
// machine: the recursion of advanceTo becomes a stack of pending targets.
struct Schedule {
  struct Target {
    double top;
    bool loading;
    int refinement;
  };
  double loadingWork = 0, recoveredWork = 0, prevDisp = 0, prevForce = 0, reactionOffset = 0;
  double time = 0, currentTop = 0;
  bool first = true, done = false;
  int cursor = 0;
  std::vector<Target> stack;  // back() is solved next
};

// The next scheduled increment of a lattice whose target stack is empty,
// MC/src/sim.cpp:456-465; false when its schedule is complete.
bool push_next(Schedule& s, const SimulationConfig& cf, const HostLattice& L) {
  if (!s.stack.empty()) return true;
  const int Ls = cf.loadSteps, Us = cf.unloadSteps;
  if (s.cursor > Ls + Us) return false;
  const int k = s.cursor++;
  const double low = L.top0 - cf.maxStrain * L.height;
  auto topY = [&](double f, bool loading) {
    const double t = std::max(0.0, std::min(1.0, f));
    return loading ? L.top0 * (1.0 - t) + low * t : low * (1.0 - t) + L.top0 * t;
  };
  if (k == 0)
    s.stack.push_back({L.top0, true, 0});
  else if (k <= Ls)
    s.stack.push_back({topY(double(k) / std::max(1, Ls), true), true, 0});
  else
    s.stack.push_back({topY(double(k - Ls) / std::max(1, Us), false), false, 0});
  return true;
}

// The outcome of a solve at the lattice's current target: 1 to commit, 0 to
// solve again (bisection, MC/src/sim.cpp:403-407: the midpoint first, then the
// target again), -1 for an equilibrium failure (MC/src/sim.cpp:412-418).
int judge(Schedule& s, const SimulationConfig& cf, const mmc_opt_stats& so, const mmc_objective_out& fo,
          BatchSimulator::Outcome& out) {
  const Schedule::Target t = s.stack.back();
  const bool acceptable = so.converged && fo.max_pen <= cf.penetrationTolerance && std::isfinite(fo.energy);
  if (acceptable) return 1;
  if (t.refinement < cf.maxAdaptiveRefinements && std::abs(t.top - s.currentTop) > 2.5e-5) {
    s.stack.pop_back();
    s.stack.push_back({t.top, t.loading, t.refinement + 1});
    s.stack.push_back({0.5 * (s.currentTop + t.top), t.loading, t.refinement + 1});
    return 0;
  }
  std::ostringstream message;
  message << "quasi-static equilibrium failed at platen=" << t.top << " m, residual=" << so.rms
          << ", penetration=" << fo.max_pen * 1000.0 << " mm after " << so.iterations << " iterations";
  out.ok = false;
  out.error = message.str();
  s.done = true;
  return -1;
}

// After the commit of a solve: work accounting (MC/src/sim.cpp:420-433) and
// the increment's diagnostics row.
void record(Schedule& s, const SimulationConfig& cf, const HostLattice& L, const mmc_objective_out& fo,
            const mmc_opt_stats& so, const mmc_state_stats& st, BatchSimulator::Outcome& out) {
  const Schedule::Target t = s.stack.back();
  s.stack.pop_back();
  if (s.first) s.reactionOffset = fo.reaction;
  const double correctedForce = std::max(0.0, fo.reaction - s.reactionOffset);
  const double displacement = L.top0 - t.top;
  if (!s.first) {
    const double dd = displacement - s.prevDisp;
    const double work = 0.5 * (s.prevForce + correctedForce) * dd;
    if (dd >= 0.0)
      s.loadingWork += std::max(0.0, work);
    else
      s.recoveredWork += std::max(0.0, -work);
  }
  s.prevDisp = displacement;
  s.prevForce = correctedForce;
  s.currentTop = t.top;
  out.result.diagnostics.push_back(make_diagnostics(L, s.time, t.top, t.loading, fo, so, st, s.reactionOffset,
                                                    s.loadingWork, s.recoveredWork));
  s.first = false;
  s.time += cf.stepTime;
}

// Per-lattice progression (opt-in, MMC_BATCH_PER_LATTICE=1) needs line search
// 0 and the graph-replayed loop.  Rounds are the default: on the measured
// batches one lattice's schedule dominates (1,024 designs of 6^3: 190,722
// iterations against a median of 14,253), and its trials are cheaper in a
// round's tail, where the graph covers only the last few lattices, than
// interleaved with every other lattice's solves.
bool per_lattice(const std::vector<mmc_config>& cfgs) {
  const char* r = std::getenv("MMC_BATCH_PER_LATTICE");
  if (!r || std::string(r) != "1") return false;
  const char* e = std::getenv("MMC_DEVICE_LOOP");
  if (e && std::string(e) == "block") return false;
  for (const mmc_config& c : cfgs)
    if (c.line_search != 0) return false;
  return true;
}
}  // namespace

std::vector<BatchSimulator::Outcome> BatchSimulator::run(bool verbose) {
  if (!shards_.empty()) return run_shards(verbose);
  if (observer_) check(mmc_capture_mark(ctx_));
  return per_lattice(c_) ? run_per_lattice(verbose) : run_rounds(verbose);
}

// Every shard on a host thread of its own (each shard's context has its own
// stream); frames and outcomes land at the lattices' indices in the batch.
std::vector<BatchSimulator::Outcome> BatchSimulator::run_shards(bool verbose) {
  const size_t K = shards_.size();
  std::vector<Outcome> out(lats_.size());
  std::vector<std::exception_ptr> err(K);
  std::mutex mu;  // the observer is called from every shard's thread
  for (size_t k = 0; k < K; k++) {
    if (observer_) {
      const std::vector<int>& idx = shard_lattices_[k];
      shards_[k]->setFrameObserver([this, &mu, &idx](int lattice, int frame, const mmc_snapshot* b, int64_t c) {
        std::lock_guard<std::mutex> lock(mu);
        observer_(idx[size_t(lattice)], frame, b, c);
      });
    } else {
      shards_[k]->setFrameObserver(nullptr);
    }
  }
  std::vector<std::thread> threads;
  for (size_t k = 0; k < K; k++)
    threads.emplace_back([&, k] {
      try {
        std::vector<Outcome> r = shards_[k]->run(verbose && k == 0);
        const std::vector<int>& idx = shard_lattices_[k];
        for (size_t j = 0; j < idx.size(); j++) out[size_t(idx[j])] = std::move(r[j]);
      } catch (...) {
        err[k] = std::current_exception();
      }
    });
  for (std::thread& t : threads) t.join();
  for (const std::exception_ptr& e : err)
    if (e) std::rethrow_exception(e);
  return out;
}

int64_t BatchSimulator::rounds() const {
  int64_t r = rounds_;
  for (const auto& sh : shards_) r = std::max(r, sh->rounds());
  return r;
}

int64_t BatchSimulator::evaluations() const {
  if (shards_.empty()) return mmc_evaluations(ctx_);
  int64_t e = 0;
  for (const auto& sh : shards_) e += sh->evaluations();
  return e;
}

void BatchSimulator::observe(const std::vector<int32_t>& mask, const std::vector<double>& tops,
                             const std::vector<int>& frames, const std::vector<int32_t>& first) {
  const int n = instances();
  std::vector<int32_t> vel(size_t(n), 0);
  int64_t total = 0;
  for (int b = 0; b < n; b++)
    if (mask[size_t(b)]) {
      vel[size_t(b)] = !first[size_t(b)];
      total += int64_t(lats_[size_t(b)].n_beams) + lats_[size_t(b)].n_nodes + MMC_MACHINE_BODIES;
    }
  if (total == 0) return;
  bodies_.resize(size_t(total));
  check(mmc_capture(ctx_, tops.data(), mask.data(), vel.data(), bodies_.data()));
  check(mmc_capture_mark(ctx_));
  int64_t o = 0;
  for (int b = 0; b < n; b++)
    if (mask[size_t(b)]) {
      const int64_t count = int64_t(lats_[size_t(b)].n_beams) + lats_[size_t(b)].n_nodes + MMC_MACHINE_BODIES;
      observer_(b, frames[size_t(b)], bodies_.data() + o, count);
      o += count;
    }
}

// Rounds: every lattice with a target solves in one mmc_minimize_batch; the
// round ends when the slowest has finished.
std::vector<BatchSimulator::Outcome> BatchSimulator::run_rounds(bool verbose) {
  const int n = instances();
  auto last = std::chrono::steady_clock::now();
  const size_t nz = size_t(n);
  std::vector<Outcome> out(nz);
  std::vector<Schedule> sch(nz);
  for (int b = 0; b < n; b++) sch[size_t(b)].currentTop = lats_[size_t(b)].top0;
  std::vector<double> tops(nz);
  std::vector<int32_t> active(nz), mask(nz);
  std::vector<mmc_opt_stats> stats(nz);
  std::vector<mmc_objective_out> fin(nz);
  std::vector<mmc_state_stats> st(nz);
  std::vector<int32_t> first(nz);
  std::vector<int> nframe(nz);
  const auto t_run = std::chrono::steady_clock::now();
  for (;;) {
    int any = 0;
    for (int b = 0; b < n; b++) {
      Schedule& s = sch[size_t(b)];
      active[size_t(b)] = 0;
      if (s.done) continue;
      if (!push_next(s, cfgs_[size_t(b)], lats_[size_t(b)])) {
        s.done = true;
        continue;
      }
      tops[size_t(b)] = s.stack.back().top;
      active[size_t(b)] = 1;
      any++;
    }
    if (!any) break;
    rounds_++;
    if (verbose && std::chrono::steady_clock::now() - last > std::chrono::seconds(10)) {
      last = std::chrono::steady_clock::now();
      size_t frames = 0;
      for (const Outcome& o : out) frames += o.result.diagnostics.size();
      std::cerr << "[batch] round " << rounds_ << ": " << any << " of " << n << " lattices running, " << frames
                << " frames" << std::endl;
    }
    auto t_min = std::chrono::steady_clock::now();
    check(mmc_minimize_batch(ctx_, tops.data(), active.data(), stats.data(), fin.data()));
    phase_s_[0] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_min).count();
    int commits = 0;
    for (int b = 0; b < n; b++) {
      mask[size_t(b)] = 0;
      if (!active[size_t(b)]) continue;
      if (judge(sch[size_t(b)], cfgs_[size_t(b)], stats[size_t(b)], fin[size_t(b)], out[size_t(b)]) == 1) {
        mask[size_t(b)] = 1;
        commits++;
      }
    }
    if (!commits) continue;
    auto t_commit = std::chrono::steady_clock::now();
    check(mmc_commit_batch(ctx_, mask.data()));
    check(mmc_read_stats_batch(ctx_, st.data()));
    phase_s_[1] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_commit).count();
    for (int b = 0; b < n; b++) {
      first[size_t(b)] = sch[size_t(b)].first;
      if (mask[size_t(b)])
        record(sch[size_t(b)], cfgs_[size_t(b)], lats_[size_t(b)], fin[size_t(b)], stats[size_t(b)], st[size_t(b)],
               out[size_t(b)]);
      nframe[size_t(b)] = int(out[size_t(b)].result.diagnostics.size()) - 1;
    }
    if (observer_) observe(mask, tops, nframe, first);
  }
  phase_s_[2] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_run).count();
  if (verbose)
    std::cerr << "[batch] " << rounds_ << " rounds, " << phase_s_[2] << " s: minimize " << phase_s_[0]
              << " s, commit + statistics " << phase_s_[1] << " s" << std::endl;
  return out;
}

// Per lattice: every lattice starts its next solve as soon as its last one is
// committed (or bisected); the device loop runs until a sixteenth of the
// lattices in flight have ended their solves, which are then finished,
// judged, committed and restarted together.
std::vector<BatchSimulator::Outcome> BatchSimulator::run_per_lattice(bool verbose) {
  const int n = instances();
  auto last = std::chrono::steady_clock::now();
  const size_t nz = size_t(n);
  std::vector<Outcome> out(nz);
  std::vector<Schedule> sch(nz);
  for (int b = 0; b < n; b++) sch[size_t(b)].currentTop = lats_[size_t(b)].top0;
  std::vector<double> tops(nz);
  std::vector<int32_t> start(nz), flying(nz, 0), running(nz), done(nz), mask(nz);
  std::vector<mmc_opt_stats> stats(nz);
  std::vector<mmc_objective_out> fin(nz);
  std::vector<mmc_state_stats> st(nz);
  std::vector<int32_t> first(nz);
  std::vector<int> nframe(nz);
  std::vector<double> ctop(nz);
  const auto t_run = std::chrono::steady_clock::now();
  // the lattices of `start` whose schedules go on begin their next solves
  auto launch = [&] {
    int k = 0;
    for (int b = 0; b < n; b++) {
      if (!start[size_t(b)]) continue;
      Schedule& s = sch[size_t(b)];
      if (s.done || !push_next(s, cfgs_[size_t(b)], lats_[size_t(b)])) {
        s.done = true;
        start[size_t(b)] = 0;
        continue;
      }
      tops[size_t(b)] = s.stack.back().top;
      flying[size_t(b)] = 1;
      k++;
    }
    auto t0 = std::chrono::steady_clock::now();
    check(mmc_batch_start(ctx_, tops.data(), start.data()));
    phase_s_[0] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return k;
  };
  for (int b = 0; b < n; b++) start[size_t(b)] = 1;
  int in_flight = launch();
  while (in_flight > 0) {
    rounds_++;
    if (verbose && std::chrono::steady_clock::now() - last > std::chrono::seconds(10)) {
      last = std::chrono::steady_clock::now();
      size_t frames = 0;
      for (const Outcome& o : out) frames += o.result.diagnostics.size();
      std::cerr << "[batch] step " << rounds_ << ": " << in_flight << " of " << n << " lattices solving, " << frames
                << " frames" << std::endl;
    }
    auto t_min = std::chrono::steady_clock::now();
    check(mmc_batch_advance(ctx_, std::max(1, in_flight / 16), running.data()));
    for (int b = 0; b < n; b++) done[size_t(b)] = flying[size_t(b)] && !running[size_t(b)];
    check(mmc_batch_finish(ctx_, done.data(), stats.data(), fin.data()));
    phase_s_[0] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_min).count();
    int commits = 0;
    for (int b = 0; b < n; b++) {
      mask[size_t(b)] = 0;
      start[size_t(b)] = 0;
      if (!done[size_t(b)]) continue;
      flying[size_t(b)] = 0;
      in_flight--;
      const int j = judge(sch[size_t(b)], cfgs_[size_t(b)], stats[size_t(b)], fin[size_t(b)], out[size_t(b)]);
      if (j == 1) {
        mask[size_t(b)] = 1;
        commits++;
      }
      start[size_t(b)] = j >= 0;
    }
    if (commits) {
      auto t_commit = std::chrono::steady_clock::now();
      check(mmc_commit_batch(ctx_, mask.data()));
      check(mmc_read_stats_batch(ctx_, st.data()));
      phase_s_[1] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_commit).count();
      for (int b = 0; b < n; b++) {
        first[size_t(b)] = sch[size_t(b)].first;
        if (mask[size_t(b)]) {
          // the committed target, for the frame's moving platen
          ctop[size_t(b)] = sch[size_t(b)].stack.back().top;
          record(sch[size_t(b)], cfgs_[size_t(b)], lats_[size_t(b)], fin[size_t(b)], stats[size_t(b)],
                 st[size_t(b)], out[size_t(b)]);
        }
        nframe[size_t(b)] = int(out[size_t(b)].result.diagnostics.size()) - 1;
      }
      if (observer_) observe(mask, ctop, nframe, first);
    }
    in_flight += launch();
  }
  phase_s_[2] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_run).count();
  if (verbose)
    std::cerr << "[batch] per lattice, " << rounds_ << " steps, " << phase_s_[2] << " s: solves " << phase_s_[0]
              << " s, commit + statistics " << phase_s_[1] << " s" << std::endl;
  return out;
}

}  // namespace meta

// ===========================================================================
// C entry points of the facade
// ===========================================================================
struct mmc_sim {
  std::unique_ptr<meta::Driver> sim;
  std::string err;
};
struct mmc_batch {
  std::unique_ptr<meta::BatchSimulator> sim;
  std::string err;
};

namespace {
std::string g_sim_create_error;
mmc_frame to_frame(const meta::FrameDiagnostics& d) {
  mmc_frame f;
  std::memset(&f, 0, sizeof f);
  f.time = d.time;
  f.platen_y = d.platenY;
  f.displacement = d.displacement;
  f.engineering_strain = d.engineeringStrain;
  f.engineering_stress = d.engineeringStress;
  f.reaction_force = d.reactionForce;
  f.loading_work = d.loadingWork;
  f.recovered_work = d.recoveredWork;
  f.dissipated_work = d.dissipatedWork;
  f.specific_energy_absorption = d.specificEnergyAbsorption;
  f.elastic_energy = d.elasticEnergy;
  f.plastic_dissipation = d.plasticDissipation;
  f.contact_energy = d.contactEnergy;
  f.residual_rms = d.residualRms;
  f.residual_max = d.residualMax;
  f.max_penetration = d.maxPenetration;
  f.mean_plastic_strain = d.meanPlasticStrain;
  f.max_plastic_strain = d.maxPlasticStrain;
  f.mean_plastic_rotation = d.meanPlasticRotation;
  f.max_plastic_rotation = d.maxPlasticRotation;
  f.permanent_set = d.permanentSet;
  f.active_contacts = d.activeContacts;
  f.plastic_beams = d.plasticBeams;
  f.plastic_hinges = d.plasticHinges;
  f.iterations = d.iterations;
  f.converged = d.converged ? 1u : 0u;
  f.loading = d.loading ? 1u : 0u;
  return f;
}
template <class F>
int sim_guard(mmc_sim* s, F&& f) {
  if (!s || !s->sim) return MMC_ERR_ARG;
  try {
    f();
    return MMC_OK;
  } catch (const std::exception& e) {
    s->err = e.what();
    return MMC_ERR_EQUILIBRIUM;
  }
}
}  // namespace

extern "C" {

int mmc_sim_create(const mmc_config* cfg, int32_t precision, mmc_sim** out) {
  if (!cfg || !out) return MMC_ERR_ARG;
  *out = nullptr;
  auto* s = new mmc_sim();
  try {
    s->sim = std::make_unique<meta::Driver>(*cfg, precision);
    s->sim->setCaptureFrames(false);
    s->sim->setVerbose(false);
  } catch (const std::exception& e) {
    s->err = e.what();
    g_sim_create_error = e.what();
    *out = s;
    return MMC_ERR_ARG;
  }
  *out = s;
  return MMC_OK;
}

void mmc_sim_destroy(mmc_sim* s) { delete s; }

const char* mmc_sim_last_error(const mmc_sim* s) { return s ? s->err.c_str() : g_sim_create_error.c_str(); }

int mmc_sim_info(const mmc_sim* s, int32_t* nn, int32_t* nb, double* height, double* mass, double* area) {
  if (!s || !s->sim) return MMC_ERR_ARG;
  const meta::HostLattice& L = s->sim->lattice();
  if (nn) *nn = L.n_nodes;
  if (nb) *nb = L.n_beams;
  if (height) *height = s->sim->initialHeight();
  if (mass) *mass = s->sim->mass();
  if (area) *area = s->sim->area();
  return MMC_OK;
}

int mmc_sim_lattice(const mmc_sim* s, double* x, double* mass, uint8_t* locked, int32_t* a, int32_t* b, double* L0,
                    double* radius, double* e0) {
  if (!s || !s->sim) return MMC_ERR_ARG;
  const meta::HostLattice& L = s->sim->lattice();
  if (x) std::memcpy(x, L.x.data(), L.x.size() * sizeof(double));
  if (mass) std::memcpy(mass, L.mass.data(), L.mass.size() * sizeof(double));
  if (locked) std::memcpy(locked, L.locked.data(), L.locked.size());
  if (a) std::memcpy(a, L.a.data(), L.a.size() * sizeof(int32_t));
  if (b) std::memcpy(b, L.b.data(), L.b.size() * sizeof(int32_t));
  if (L0) std::memcpy(L0, L.L0.data(), L.L0.size() * sizeof(double));
  if (radius) std::memcpy(radius, L.radius.data(), L.radius.size() * sizeof(double));
  if (e0) std::memcpy(e0, L.e0.data(), L.e0.size() * sizeof(double));
  return MMC_OK;
}

int mmc_sim_run(mmc_sim* s, mmc_frame* frames, int32_t cap, int32_t* n, int32_t verbose) {
  return sim_guard(s, [&] {
    s->sim->setVerbose(verbose != 0);
    meta::SimulationResult r;
    std::string failure;
    try {
      r = s->sim->run();
    } catch (const std::exception& e) {
      failure = e.what();
    }
    (void)r;
    if (!failure.empty()) throw std::runtime_error(failure);
    int m = std::min<int>(cap, int(r.diagnostics.size()));
    for (int i = 0; i < m && frames; i++) frames[i] = to_frame(r.diagnostics[size_t(i)]);
    if (n) *n = int32_t(r.diagnostics.size());
  });
}

int mmc_sim_begin(mmc_sim* s) {
  return sim_guard(s, [&] { s->sim->begin(); });
}

int mmc_sim_advance(mmc_sim* s, mmc_frame* frames, int32_t cap, int32_t* n, int32_t* done) {
  return sim_guard(s, [&] {
    meta::SimulationResult r;
    bool more = s->sim->advance(r);
    int m = std::min<int>(cap, int(r.diagnostics.size()));
    for (int i = 0; i < m && frames; i++) frames[i] = to_frame(r.diagnostics[size_t(i)]);
    if (n) *n = int32_t(r.diagnostics.size());
    if (done) *done = more ? 0 : 1;
  });
}

int mmc_sim_rigid_rotation_test(mmc_sim* s, double* error, int32_t* pass) {
  return sim_guard(s, [&] {
    bool ok = s->sim->rigidRotationTest(error);
    if (pass) *pass = ok;
  });
}

int mmc_sim_gradient_check(mmc_sim* s, double* error, int32_t* pass) {
  return sim_guard(s, [&] {
    bool ok = s->sim->gradientCheck(error);
    if (pass) *pass = ok;
  });
}

mmc_ctx* mmc_sim_context(mmc_sim* s) { return (s && s->sim) ? s->sim->context() : nullptr; }

int mmc_sim_export_stl(mmc_sim* s, const char* path, int32_t radial_segments, mmc_stl_report* report) {
  if (!path) return MMC_ERR_ARG;
  return sim_guard(s, [&] { s->sim->exportSTL(path, radial_segments, report); });
}

static int batch_create(const mmc_config* cfgs, int32_t n, int32_t precision, mmc_batch** out, bool generate) {
  if (!cfgs || n <= 0 || !out) return MMC_ERR_ARG;
  *out = nullptr;
  auto* s = new mmc_batch();
  try {
    s->sim = std::make_unique<meta::BatchSimulator>(std::vector<mmc_config>(cfgs, cfgs + n), precision, generate);
  } catch (const std::exception& e) {
    s->err = e.what();
    g_sim_create_error = e.what();
    *out = s;
    return MMC_ERR_ARG;
  }
  *out = s;
  return MMC_OK;
}

int mmc_batch_create(const mmc_config* cfgs, int32_t n, int32_t precision, mmc_batch** out) {
  return batch_create(cfgs, n, precision, out, false);
}

int mmc_batch_create_generated(const mmc_config* cfgs, int32_t n, int32_t precision, mmc_batch** out) {
  return batch_create(cfgs, n, precision, out, true);
}

void mmc_batch_destroy(mmc_batch* s) { delete s; }

const char* mmc_batch_last_error(const mmc_batch* s) { return s ? s->err.c_str() : g_sim_create_error.c_str(); }

int mmc_batch_info(const mmc_batch* s, int32_t i, int32_t* nn, int32_t* nb, double* height, double* mass,
                   double* area) {
  if (!s || !s->sim || i < 0 || i >= s->sim->instances()) return MMC_ERR_ARG;
  const meta::HostLattice& L = s->sim->lattice(i);
  if (nn) *nn = L.n_nodes;
  if (nb) *nb = L.n_beams;
  if (height) *height = L.height;
  if (mass) *mass = L.total_mass;
  if (area) *area = L.area_ref;
  return MMC_OK;
}

int mmc_batch_run(mmc_batch* s, mmc_frame* frames, int32_t cap, int32_t* n_frames, int32_t* status,
                  int32_t verbose) {
  if (!s || !s->sim) return MMC_ERR_ARG;
  try {
    auto res = s->sim->run(verbose != 0);
    for (size_t b = 0; b < res.size(); b++) {
      const auto& d = res[b].result.diagnostics;
      const int m = std::min<int>(cap, int(d.size()));
      for (int i = 0; i < m && frames; i++) frames[b * size_t(cap) + size_t(i)] = to_frame(d[size_t(i)]);
      if (n_frames) n_frames[b] = int32_t(d.size());
      if (status) status[b] = res[b].ok ? MMC_OK : MMC_ERR_EQUILIBRIUM;
      if (!res[b].ok && s->err.empty()) s->err = "lattice " + std::to_string(b) + ": " + res[b].error;
    }
    return MMC_OK;
  } catch (const std::exception& e) {
    s->err = e.what();
    return MMC_ERR_HIP;
  }
}

int mmc_batch_set_frame_observer(mmc_batch* s, mmc_batch_frame_fn fn, void* user) {
  if (!s || !s->sim) return MMC_ERR_ARG;
  if (fn)
    s->sim->setFrameObserver([fn, user](int lattice, int frame, const mmc_snapshot* bodies, int64_t count) {
      fn(int32_t(lattice), int32_t(frame), bodies, count, user);
    });
  else
    s->sim->setFrameObserver(nullptr);
  return MMC_OK;
}

int64_t mmc_batch_evaluations(const mmc_batch* s) { return (s && s->sim) ? s->sim->evaluations() : 0; }

mmc_ctx* mmc_batch_context(mmc_batch* s) { return (s && s->sim) ? s->sim->context() : nullptr; }

int mmc_lattice_build(const mmc_config* cfg, mmc_lattice* view, void** handle) {
  if (!cfg || !view || !handle) return MMC_ERR_ARG;
  *handle = nullptr;
  try {
    mmc_config c = *cfg;
    c.translation_scale = c.cell_size;
    auto* L = new meta::HostLattice(meta::build_lattice(c));
    *view = L->view();
    *handle = L;
  } catch (const std::exception& e) {
    g_sim_create_error = e.what();
    return MMC_ERR_ARG;
  }
  return MMC_OK;
}

void mmc_lattice_free(void* handle) { delete static_cast<meta::HostLattice*>(handle); }

}  // extern "C"
