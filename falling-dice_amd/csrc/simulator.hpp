// simulator.hpp — the reference's solver API (projects/metamaterial-compression/
// industry-grade/include/sim.hpp, MC/ below) over the gfx950 hot path.
//
// meta::Simulator and every struct it exposes keep the reference's exact
// layout (same members, types, order and defaults), so MC/src/main.cpp and
// MC/src/render.cpp compiled unchanged against the reference's own headers link
// against libmmc.so in place of MC/src/sim.cpp: the constructor, run(),
// exportSTL(), rigidRotationTest() and gradientCheck() are defined here, and
// the reference header's inline accessors and implicit destructor work on the
// same members.  The equilibrium hot path (objective, minimize, commit) runs on
// the GPU through the C-ABI of include/mmc.h; run() drives it through a
// Driver (below) and copies the committed state back into the members.
//
// What the reference's SimulationConfig cannot carry (the cubic lattice
// family, the precision policy, the parity-form line search, ...) is a
// SimulationExtensions value handed to the next Simulator constructed on the
// calling thread (setNextSimulationExtensions); without one every Simulator
// behaves as the reference's.
#pragma once
#include <cstdint>
#include <filesystem>
#include <functional>
#include <limits>
#include <memory>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/mmc.h"
#include "lattice.hpp"

namespace meta {

// MC/include/core.hpp:38,40 (layout: three and four doubles)
struct Vec3 {
  double x = 0, y = 0, z = 0;
};
struct Quat {
  double w = 1, x = 0, y = 0, z = 0;
};

// MC/include/sim.hpp:6-15
enum class BodyKind : int32_t { Strut = 0, Joint = 1, Platen = 2, Machine = 3 };
struct Snapshot {
  Vec3 p;
  Quat q;
  Vec3 v, w, half;
  int material = 0, id = 0, kind = 0, moduleClass = -1, ring = -1, slot = -1;
};
struct FrameDiagnostics {
  double time = 0, platenY = 0, displacement = 0, engineeringStrain = 0, engineeringStress = 0, reactionForce = 0;
  double loadingWork = 0, recoveredWork = 0, dissipatedWork = 0, specificEnergyAbsorption = 0;
  double elasticEnergy = 0, plasticDissipation = 0, contactEnergy = 0, residualRms = 0, residualMax = 0,
         maxPenetration = 0;
  double meanPlasticStrain = 0, maxPlasticStrain = 0, meanPlasticRotation = 0, maxPlasticRotation = 0,
         permanentSet = 0;
  int activeContacts = 0, plasticBeams = 0, plasticHinges = 0, iterations = 0;
  bool converged = false, loading = true;
};
struct SimulationResult {
  std::vector<std::vector<Snapshot>> frames;
  std::vector<FrameDiagnostics> diagnostics;
};

// MC/include/sim.hpp:16-26: same names, order and defaults
struct MaterialConfig {
  std::string name = "FFF TPU 95A provisional isotropic baseline";
  double density = 1180, young = 26e6, nu = .485, yield = 5.2e6, hardening = 1.4e6, bendHardening = .06,
         torsionHardening = .045;
};
struct SimulationConfig {
  int cellsX = 4, cellsY = 6, cellsZ = 4;
  double cellSize = .010, minRadius = .00082, maxRadius = .00122, jointScale = 1.18, triggerScale = .38,
         imperfection = .00022, memberImperfection = .00022;
  int elementsPerStrut = 3;
  MaterialConfig material;
  double gravity = 9.81, shearCorrection = .62, contactStiffness = 1.2e6, contactMargin = 2e-5,
         centeringStiffness = 1.5;
  int contactSamples = 3, loadSteps = 52, unloadSteps = 30, maxIterations = 130, history = 10, maxLineSearch = 22;
  double maxStrain = .62, stepTime = .06, translationScale = .010, maxTranslationStep = .0022, maxRotationStep = .30;
  double gradientRmsTolerance = 1.2e-3, gradientMaxTolerance = 1.5e-2, armijo = 1e-4, minLineStep = 1e-9,
         penetrationTolerance = .00012;
  int maxAdaptiveRefinements = 7;
  int fps = 24, frames = 83, threads = 0;
  uint64_t seed = 0x494e445553545259ULL;
};

// MC/include/sim.hpp:27-33
struct Node {
  Vec3 x0, x;
  Quat R;
  double mass = 0;
  int layer = 0;
};
struct PlasticState {
  double epsp = 0, alpha = 0;
  Vec3 rotp{0, 0, 0};
  double alphaB = 0, alphaT = 0;
};
struct Beam {
  int a = -1, b = -1, layer = 0, family = 0;
  double L0 = 0, radius = .001, area = 0, I = 0, J = 0;
  Vec3 e0{1, 0, 0};
  PlasticState plastic;
};
struct TrialBeam {
  PlasticState next;
  double strain = 0, stress = 0;
  Vec3 moment{0, 0, 0};
  double elastic = 0, dissipation = 0;
  bool axialPlastic = false, bendPlastic = false, torsionPlastic = false;
};
struct ObjectiveDiag {
  double elastic = 0, plastic = 0, contact = 0, reaction = 0, maxPen = 0;
  int contacts = 0;
};
struct Objective {
  double energy = 0;
  std::vector<double> grad;
  ObjectiveDiag d;
  std::vector<TrialBeam> trial;
};
struct OptStats {
  int iterations = 0, backtracks = 0;
  double rms = std::numeric_limits<double>::infinity(), max = std::numeric_limits<double>::infinity(),
         energy = std::numeric_limits<double>::infinity();
  bool converged = false;
};

// MC/include/sim.hpp:35-54: public API and data members of the reference.
class Simulator {
 public:
  explicit Simulator(SimulationConfig cfg);
  SimulationResult run();
  // MC/src/sim.cpp:645-888 on the device (mmc_export_stl)
  void exportSTL(const std::filesystem::path& path, int radialSegments = 12) const;
  bool rigidRotationTest(double* error = nullptr) const;
  bool gradientCheck(double* relativeError = nullptr) const;
  const std::vector<Node>& nodes() const { return nodes_; }
  const std::vector<Beam>& beams() const { return beams_; }
  double initialHeight() const { return height_; }
  double mass() const { return mass_; }
  double area() const { return area_; }

 private:
  SimulationConfig cfg_;
  std::vector<Node> nodes_;
  std::vector<Beam> beams_;
  std::unordered_set<uint64_t> connected_;
  std::vector<uint8_t> locked_;
  std::vector<Vec3> baseX_;
  std::vector<Quat> baseR_;
  double width_ = 0, height_ = 0, depth_ = 0, area_ = 0, mass_ = 0, top0_ = 0;
  double loadingWork_ = 0, recoveredWork_ = 0, prevDisp_ = 0, prevForce_ = 0, reactionOffset_ = 0;

  // the committed lattice as the device takes it (base pose, plastic state)
  HostLattice committedLattice() const;
};

// ---------------------------------------------------------------------------
// Additive configuration (no counterpart in the reference's SimulationConfig)
// ---------------------------------------------------------------------------
struct SimulationExtensions {
  int lattice = MMC_LATTICE_BCC;     // MMC_LATTICE_BCC (the reference's) or MMC_LATTICE_CUBIC
  bool relativeEscape = true;        // the relative-residual escape of MC/src/sim.cpp:177
  double relativeTolerance = 1.5e-3;
  double positionSigma = 0;          // cubic: Gaussian node jitter (m); 0 = corner-node imperfection
  int radiusMode = 0;                // cubic: 0 linear grading along y, 1 per-beam uniform
  int lineSearch = 0;                // mmc_config::line_search (0 = the reference's Armijo test)
  int precision = MMC_FP64;          // MMC_FP64 or MMC_MIXED
  bool captureFrames = true;         // per-frame snapshots in SimulationResult::frames
  bool verbose = true;               // per-step "[equilibrium]" line on stdout
};
// Applies to the next Simulator constructed on the calling thread (once).
void setNextSimulationExtensions(const SimulationExtensions& ext);
// The extensions a Simulator was constructed with (the reference's behaviour
// when none were set).  The entry of a destroyed Simulator stays until another
// Simulator is constructed at the same address (a few dozen bytes each).
SimulationExtensions simulationExtensions(const Simulator& sim);

mmc_config to_c(const SimulationConfig& c, const SimulationExtensions& e);
void from_c(const mmc_config& c, SimulationConfig* s, SimulationExtensions* e);

// ---------------------------------------------------------------------------
// Driver: one lattice's device context and the reference schedule around it
// (MC/src/sim.cpp:374-467) — tare, load, unload, adaptive bisection, work
// accounting, diagnostics and snapshot capture — stepwise (begin/advance) or
// whole (run).  Simulator::run, the mmc_sim_* C entry points and the CLI use
// it; it has no counterpart in the reference (whose Simulator does this on
// its own members).
// ---------------------------------------------------------------------------
class Driver {
 public:
  // A lattice built from the configuration (mmc_lattice_build's construction).
  Driver(const mmc_config& c, int precision);
  // A given lattice (Simulator::run hands over its committed state).
  Driver(const mmc_config& c, int precision, HostLattice lat);
  ~Driver();
  Driver(const Driver&) = delete;
  Driver& operator=(const Driver&) = delete;

  SimulationResult run();
  // Stepwise schedule (run() is begin() followed by advance() until it returns
  // false): scheduled increment 0 is the tare, then loadSteps loading and
  // unloadSteps unloading increments (MC/src/sim.cpp:456-465).
  void begin();
  bool advance(SimulationResult& result);
  bool rigidRotationTest(double* error = nullptr) const;
  bool gradientCheck(double* relativeError = nullptr) const;
  void exportSTL(const std::filesystem::path& path, int radialSegments, mmc_stl_report* report = nullptr) const;
  // committed state mirrored from the device (read on demand)
  const std::vector<Node>& nodes() const;
  const std::vector<Beam>& beams() const;
  double initialHeight() const { return lat_.height; }
  double mass() const { return lat_.total_mass; }
  double area() const { return lat_.area_ref; }
  void setVerbose(bool v) { ext_.verbose = v; }
  void setCaptureFrames(bool v) { ext_.captureFrames = v; }
  mmc_ctx* context() const { return ctx_; }
  const HostLattice& lattice() const { return lat_; }
  // work accounting of the schedule so far (MC/src/sim.cpp:420-433)
  double loadingWork() const { return loadingWork_; }
  double recoveredWork() const { return recoveredWork_; }
  double prevDisp() const { return prevDisp_; }
  double prevForce() const { return prevForce_; }
  double reactionOffset() const { return reactionOffset_; }

 private:
  SimulationConfig cfg_;
  SimulationExtensions ext_;
  mmc_config c_;
  HostLattice lat_;
  mmc_ctx* ctx_ = nullptr;
  mutable std::vector<Node> nodes_;
  mutable std::vector<Beam> beams_;
  mutable bool synced_ = false;
  double loadingWork_ = 0, recoveredWork_ = 0, prevDisp_ = 0, prevForce_ = 0, reactionOffset_ = 0;
  double time_ = 0, currentTop_ = 0;
  bool first_ = true;
  int acceptedSteps_ = 0, cursor_ = 0;

  void create(int precision);
  void advanceTo(double targetTop, bool loading, int refinement, SimulationResult& result);
  void check(int rc) const;
  void sync() const;
  double topY(double s, bool loading) const;
  std::vector<Snapshot> capture(double top, bool velocities) const;
  FrameDiagnostics diagnostics(double time, double top, bool loading, const mmc_objective_out& o,
                               const mmc_opt_stats& s) const;
};

// Many independent lattices solved together (no reference counterpart): one
// batched device context, and per lattice the schedule of Simulator::run
// (tare, load, unload, adaptive bisection, work accounting, diagnostics).
// Every round solves the next target of every unfinished lattice in one
// device minimisation; a lattice whose equilibrium fails stops with the
// reference's error text while the others continue.
class BatchSimulator {
 public:
  // generate: the lattices are built on the device (mmc_create_generated);
  // otherwise on the host (mmc_lattice_build) and uploaded.  shards: the
  // lattices are split over that many independent contexts (lattice b to
  // shard b mod shards), each with its own stream and host thread, so that a
  // shard's rounds wait only for that shard's slowest lattice; 0 reads
  // MMC_BATCH_SHARDS (default 1).  A lattice's results do not depend on the
  // other lattices of its context, so they do not depend on the sharding.
  BatchSimulator(std::vector<mmc_config> cfgs, int precision, bool generate = false, int shards = 0);
  ~BatchSimulator();
  BatchSimulator(const BatchSimulator&) = delete;
  BatchSimulator& operator=(const BatchSimulator&) = delete;

  struct Outcome {
    SimulationResult result;  // diagnostics only (no snapshots)
    bool ok = true;
    std::string error;
  };
  // verbose: one progress line on stderr about every 10 s (rounds, lattices
  // still running, frames so far)
  std::vector<Outcome> run(bool verbose = false);
  int instances() const { return int(lats_.size()); }
  const HostLattice& lattice(int i) const { return lats_[size_t(i)]; }
  // the device context (the first shard's when sharded)
  mmc_ctx* context() const { return shards_.empty() ? ctx_ : shards_[0]->context(); }
  int64_t rounds() const;
  int shards() const { return shards_.empty() ? 1 : int(shards_.size()); }
  // objective evaluations over every shard
  int64_t evaluations() const;
  // Per-frame bodies of every lattice (mmc_batch_set_frame_observer); empty:
  // no capture.
  using FrameObserver = std::function<void(int lattice, int frame, const mmc_snapshot* bodies, int64_t count)>;
  void setFrameObserver(FrameObserver f) { observer_ = std::move(f); }

 private:
  FrameObserver observer_;
  std::vector<std::unique_ptr<BatchSimulator>> shards_;  // empty: one context (ctx_)
  std::vector<std::vector<int>> shard_lattices_;        // the lattices of each shard
  std::vector<Outcome> run_shards(bool verbose);
  std::vector<mmc_snapshot> bodies_;  // host staging of one capture
  // The frames of the lattices in `mask` just committed (velocities unless
  // it is a lattice's first frame), handed to the observer; then the
  // committed poses of every lattice become the previous poses.
  void observe(const std::vector<int32_t>& mask, const std::vector<double>& tops, const std::vector<int>& frames,
               const std::vector<int32_t>& first);
  std::vector<SimulationConfig> cfgs_;
  std::vector<mmc_config> c_;
  std::vector<HostLattice> lats_;
  mmc_ctx* ctx_ = nullptr;
  int64_t rounds_ = 0;
  double phase_s_[3] = {0, 0, 0};  // seconds in minimize, in commit + statistics, in run
  void check(int rc) const;
  std::vector<Outcome> run_rounds(bool verbose);
  std::vector<Outcome> run_per_lattice(bool verbose);
};

}  // namespace meta
