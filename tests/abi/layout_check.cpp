// Layout check of the facade (falling-dice_amd/csrc/simulator.hpp) against the
// reference's solver header (MC/include/sim.hpp): both in one translation
// unit, the reference's under namespace refmeta.  Every struct of the API must
// have the same size and alignment, every public member the same offset and
// default value, and meta::Simulator the same size: then MC/src/main.cpp,
// compiled against the reference's own headers, can link libmmc.so in place
// of MC/src/sim.cpp.  Built and run by tests/test_dropin.py (CPU).
#include <algorithm>
#include <array>
#include <atomic>
#include <bit>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <limits>
#include <memory>
#include <numeric>
#include <optional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

#define meta refmeta
#include "sim.hpp"  // the reference's (MC/include on the include path)
#undef meta
#include "simulator.hpp"  // the facade's (falling-dice_amd/csrc on the include path)

#define SAME_TYPE(T)                                                   \
  static_assert(sizeof(meta::T) == sizeof(refmeta::T), "size of " #T); \
  static_assert(alignof(meta::T) == alignof(refmeta::T), "alignment of " #T);
#define SAME_FIELD(T, f) \
  static_assert(offsetof(meta::T, f) == offsetof(refmeta::T, f), "offset of " #T "::" #f);

SAME_TYPE(Vec3)
SAME_TYPE(Quat)
SAME_TYPE(Snapshot)
SAME_TYPE(FrameDiagnostics)
SAME_TYPE(SimulationResult)
SAME_TYPE(MaterialConfig)
SAME_TYPE(SimulationConfig)
SAME_TYPE(Node)
SAME_TYPE(PlasticState)
SAME_TYPE(Beam)
SAME_TYPE(TrialBeam)
SAME_TYPE(ObjectiveDiag)
SAME_TYPE(Objective)
SAME_TYPE(OptStats)
SAME_TYPE(Simulator)

#define VEC3_FIELDS(X, T) X(T, x) X(T, y) X(T, z)
#define QUAT_FIELDS(X, T) X(T, w) X(T, x) X(T, y) X(T, z)
#define SNAPSHOT_FIELDS(X, T) \
  X(T, p) X(T, q) X(T, v) X(T, w) X(T, half) X(T, material) X(T, id) X(T, kind) X(T, moduleClass) X(T, ring) X(T, slot)
#define FRAME_FIELDS(X, T)                                                                                        \
  X(T, time) X(T, platenY) X(T, displacement) X(T, engineeringStrain) X(T, engineeringStress) X(T, reactionForce) \
  X(T, loadingWork) X(T, recoveredWork) X(T, dissipatedWork) X(T, specificEnergyAbsorption) X(T, elasticEnergy)  \
  X(T, plasticDissipation) X(T, contactEnergy) X(T, residualRms) X(T, residualMax) X(T, maxPenetration)          \
  X(T, meanPlasticStrain) X(T, maxPlasticStrain) X(T, meanPlasticRotation) X(T, maxPlasticRotation)              \
  X(T, permanentSet) X(T, activeContacts) X(T, plasticBeams) X(T, plasticHinges) X(T, iterations)                \
  X(T, converged) X(T, loading)
#define RESULT_FIELDS(X, T) X(T, frames) X(T, diagnostics)
#define MATERIAL_FIELDS(X, T) \
  X(T, density) X(T, young) X(T, nu) X(T, yield) X(T, hardening) X(T, bendHardening) X(T, torsionHardening)
#define CONFIG_NUMERIC_FIELDS(X, T)                                                                                   \
  X(T, cellsX) X(T, cellsY) X(T, cellsZ) X(T, cellSize) X(T, minRadius) X(T, maxRadius) X(T, jointScale)              \
  X(T, triggerScale) X(T, imperfection) X(T, memberImperfection) X(T, elementsPerStrut) X(T, gravity)                 \
  X(T, shearCorrection) X(T, contactStiffness) X(T, contactMargin) X(T, centeringStiffness) X(T, contactSamples)      \
  X(T, loadSteps) X(T, unloadSteps) X(T, maxIterations) X(T, history) X(T, maxLineSearch) X(T, maxStrain)             \
  X(T, stepTime) X(T, translationScale) X(T, maxTranslationStep) X(T, maxRotationStep) X(T, gradientRmsTolerance)     \
  X(T, gradientMaxTolerance) X(T, armijo) X(T, minLineStep) X(T, penetrationTolerance) X(T, maxAdaptiveRefinements)   \
  X(T, fps) X(T, frames) X(T, threads) X(T, seed)
#define NODE_FIELDS(X, T) X(T, x0) X(T, x) X(T, R) X(T, mass) X(T, layer)
#define PLASTIC_FIELDS(X, T) X(T, epsp) X(T, alpha) X(T, rotp) X(T, alphaB) X(T, alphaT)
#define BEAM_FIELDS(X, T) \
  X(T, a) X(T, b) X(T, layer) X(T, family) X(T, L0) X(T, radius) X(T, area) X(T, I) X(T, J) X(T, e0) X(T, plastic)
#define TRIAL_FIELDS(X, T)                                                                                      \
  X(T, next) X(T, strain) X(T, stress) X(T, moment) X(T, elastic) X(T, dissipation) X(T, axialPlastic) \
  X(T, bendPlastic) X(T, torsionPlastic)
#define DIAG_FIELDS(X, T) X(T, elastic) X(T, plastic) X(T, contact) X(T, reaction) X(T, maxPen) X(T, contacts)
#define OBJECTIVE_FIELDS(X, T) X(T, energy) X(T, grad) X(T, d) X(T, trial)
#define OPT_FIELDS(X, T) X(T, iterations) X(T, backtracks) X(T, rms) X(T, max) X(T, energy) X(T, converged)

VEC3_FIELDS(SAME_FIELD, Vec3)
QUAT_FIELDS(SAME_FIELD, Quat)
SNAPSHOT_FIELDS(SAME_FIELD, Snapshot)
FRAME_FIELDS(SAME_FIELD, FrameDiagnostics)
RESULT_FIELDS(SAME_FIELD, SimulationResult)
MATERIAL_FIELDS(SAME_FIELD, MaterialConfig)
SAME_FIELD(MaterialConfig, name)
CONFIG_NUMERIC_FIELDS(SAME_FIELD, SimulationConfig)
SAME_FIELD(SimulationConfig, material)
NODE_FIELDS(SAME_FIELD, Node)
PLASTIC_FIELDS(SAME_FIELD, PlasticState)
BEAM_FIELDS(SAME_FIELD, Beam)
TRIAL_FIELDS(SAME_FIELD, TrialBeam)
DIAG_FIELDS(SAME_FIELD, ObjectiveDiag)
OBJECTIVE_FIELDS(SAME_FIELD, Objective)
OPT_FIELDS(SAME_FIELD, OptStats)

// defaults (run time: the member initialisers)
static int failures = 0;
#define SAME_DEFAULT(T, f)                                                      \
  if (!(meta::T{}.f == refmeta::T{}.f)) {                                       \
    std::printf("default of %s::%s differs\n", #T, #f);                         \
    failures++;                                                                 \
  }

int main() {
  FRAME_FIELDS(SAME_DEFAULT, FrameDiagnostics)
  MATERIAL_FIELDS(SAME_DEFAULT, MaterialConfig)
  SAME_DEFAULT(MaterialConfig, name)
  CONFIG_NUMERIC_FIELDS(SAME_DEFAULT, SimulationConfig)
  SAME_DEFAULT(Node, mass)
  SAME_DEFAULT(Node, layer)
  SAME_DEFAULT(PlasticState, epsp) SAME_DEFAULT(PlasticState, alpha) SAME_DEFAULT(PlasticState, alphaB)
  SAME_DEFAULT(PlasticState, alphaT) SAME_DEFAULT(PlasticState, rotp.x) SAME_DEFAULT(PlasticState, rotp.y)
  SAME_DEFAULT(PlasticState, rotp.z) SAME_DEFAULT(Beam, e0.x) SAME_DEFAULT(Beam, e0.y) SAME_DEFAULT(Beam, e0.z)
  SAME_DEFAULT(Node, R.w) SAME_DEFAULT(Node, R.x) SAME_DEFAULT(Node, R.y) SAME_DEFAULT(Node, R.z)
  SAME_DEFAULT(Beam, a) SAME_DEFAULT(Beam, b) SAME_DEFAULT(Beam, layer) SAME_DEFAULT(Beam, family)
  SAME_DEFAULT(Beam, L0) SAME_DEFAULT(Beam, radius) SAME_DEFAULT(Beam, area) SAME_DEFAULT(Beam, I) SAME_DEFAULT(Beam, J)
  SAME_DEFAULT(Snapshot, material) SAME_DEFAULT(Snapshot, id) SAME_DEFAULT(Snapshot, kind)
  SAME_DEFAULT(Snapshot, moduleClass) SAME_DEFAULT(Snapshot, ring) SAME_DEFAULT(Snapshot, slot)
  SAME_DEFAULT(OptStats, iterations) SAME_DEFAULT(OptStats, backtracks) SAME_DEFAULT(OptStats, rms)
  SAME_DEFAULT(OptStats, max) SAME_DEFAULT(OptStats, energy) SAME_DEFAULT(OptStats, converged)
  std::printf("sizeof(Simulator) %zu, sizeof(SimulationConfig) %zu: %s\n", sizeof(meta::Simulator),
              sizeof(meta::SimulationConfig), failures ? "DEFAULTS DIFFER" : "layouts and defaults equal");
  return failures ? 1 : 0;
}
