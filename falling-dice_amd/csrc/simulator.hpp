// simulator.hpp — C++20 facade with the public API of the reference's
// meta::Simulator (projects/metamaterial-compression/industry-grade/include/
// sim.hpp:6-54, MC/ below).  Construction, the load/unload schedule with
// adaptive bisection, diagnostics, snapshot capture and the verification checks
// run on the host; the equilibrium hot path (objective, minimize, commit) runs
// on the GPU through the C-ABI of include/mmc.h.
#pragma once
#include <cstdint>
#include <filesystem>
#include <memory>
#include <string>
#include <vector>

#include "../../include/mmc.h"
#include "lattice.hpp"

namespace meta {

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

// MC/include/sim.hpp:16-26, same names and defaults; the fields after the
// marker are additive (no reference counterpart).
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
  // --- additive ---
  int lattice = MMC_LATTICE_BCC;
  bool relativeEscape = true;
  double relativeTolerance = 1.5e-3;
  double positionSigma = 0;
  int radiusMode = 0;
  int lineSearch = 0;  // 0 total-energy Armijo (the reference's), 1 per-term energy change; 2 per-term with the approximate-Wolfe slope test below 1e-6 |E|
  int precision = MMC_FP64;
  bool captureFrames = true;  // per-frame snapshots for the META3 cache
  bool verbose = true;        // per-step "[equilibrium]" line on stdout
};

mmc_config to_c(const SimulationConfig& c);
SimulationConfig from_c(const mmc_config& c);

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

class Simulator {
 public:
  explicit Simulator(SimulationConfig cfg);
  ~Simulator();
  Simulator(const Simulator&) = delete;
  Simulator& operator=(const Simulator&) = delete;

  SimulationResult run();
  // Stepwise schedule (run() is begin() followed by advance() until it returns
  // false): scheduled increment 0 is the tare, then loadSteps loading and
  // unloadSteps unloading increments (MC/src/sim.cpp:456-465).
  void begin();
  bool advance(SimulationResult& result);
  // Manufacturing export is outside this build (the reference's STL path,
  // MC/src/sim.cpp:579-888); throws.
  void exportSTL(const std::filesystem::path& path, int radialSegments = 12) const;
  bool rigidRotationTest(double* error = nullptr) const;
  bool gradientCheck(double* relativeError = nullptr) const;
  const std::vector<Node>& nodes() const;
  const std::vector<Beam>& beams() const;
  double initialHeight() const { return lat_.height; }
  double mass() const { return lat_.total_mass; }
  double area() const { return lat_.area_ref; }
  const SimulationConfig& config() const { return cfg_; }
  void setVerbose(bool v) { cfg_.verbose = v; }
  mmc_ctx* context() const { return ctx_; }
  const HostLattice& lattice() const { return lat_; }

 private:
  SimulationConfig cfg_;
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

  void advanceTo(double targetTop, bool loading, int refinement, SimulationResult& result);

  void check(int rc) const;
  void sync() const;
  double topY(double s, bool loading) const;
  std::vector<Snapshot> capture(double top, const std::vector<Vec3>* prevX, const std::vector<Quat>* prevR) const;
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
  BatchSimulator(std::vector<SimulationConfig> cfgs, int precision);
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
  mmc_ctx* context() const { return ctx_; }
  int64_t rounds() const { return rounds_; }

 private:
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
