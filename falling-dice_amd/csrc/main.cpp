// main.cpp — command-line front end with the reference's interface
// (projects/metamaterial-compression/industry-grade/src/main.cpp, MC/src/main.cpp
// below): --mode sim|validate|convergence with the same options, the same
// output files (compression.meta3, compression_metrics.csv, summary.json,
// VALIDATION.txt, convergence.csv) in the same byte formats, and the same
// stdout/stderr lines and exit codes.  The equilibrium solve runs on the GPU.
// STL export runs on the GPU too (--mode stl, and --mode sim without
// --skip-stl).  Rendering is not part of this build: --mode render|hero exit
// with an error, and --mode all runs the simulation, STL and validation.
// Additive options (no reference counterpart) are listed by --help.
#include <array>
#include <cmath>
#include <limits>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>

#include "simulator.hpp"

namespace meta {
namespace {

constexpr uint32_t kCacheVersion = 3;

// META3 cache records, MC/src/main.cpp:9-31
struct CacheHeader {
  char magic[8]{};
  uint32_t version = kCacheVersion;
  uint32_t frames = 0;
  uint32_t bodies = 0;
  uint32_t fps = 24;
  uint32_t snapshotBytes = 0;
  uint32_t diagnosticsBytes = 0;
};
struct DiskSnapshot {
  double p[3], q[4], v[3], w[3], half[3];
  int32_t material, id, kind, moduleClass, ring, slot;
};
static_assert(sizeof(CacheHeader) == 32 && sizeof(DiskSnapshot) == 152, "META3 layout");
static_assert(sizeof(mmc_frame) == 192, "DiskDiagnostics layout");

DiskSnapshot to_disk(const Snapshot& s) {
  return {{s.p.x, s.p.y, s.p.z}, {s.q.w, s.q.x, s.q.y, s.q.z}, {s.v.x, s.v.y, s.v.z}, {s.w.x, s.w.y, s.w.z},
          {s.half.x, s.half.y, s.half.z}, s.material, s.id, s.kind, s.moduleClass, s.ring, s.slot};
}
mmc_frame to_disk(const FrameDiagnostics& d) {
  mmc_frame f;
  std::memset(&f, 0, sizeof f);
  f = mmc_frame{d.time, d.platenY, d.displacement, d.engineeringStrain, d.engineeringStress, d.reactionForce,
                d.loadingWork, d.recoveredWork, d.dissipatedWork, d.specificEnergyAbsorption, d.elasticEnergy,
                d.plasticDissipation, d.contactEnergy, d.residualRms, d.residualMax, d.maxPenetration,
                d.meanPlasticStrain, d.maxPlasticStrain, d.meanPlasticRotation, d.maxPlasticRotation,
                d.permanentSet, d.activeContacts, d.plasticBeams, d.plasticHinges, d.iterations,
                uint32_t(d.converged), uint32_t(d.loading)};
  return f;
}

// MC/src/main.cpp:73-93
void save_cache(const std::filesystem::path& path, const SimulationResult& result, int fps) {
  if (result.frames.empty()) throw std::runtime_error("cannot cache empty simulation");
  const size_t bodies = result.frames.front().size();
  for (const auto& frame : result.frames)
    if (frame.size() != bodies) throw std::runtime_error("cache requires stable body count");
  std::filesystem::create_directories(path.parent_path());
  std::ofstream out(path, std::ios::binary);
  if (!out) throw std::runtime_error("cannot write cache " + path.string());
  CacheHeader h;
  std::memcpy(h.magic, "META3", 5);
  h.frames = uint32_t(result.frames.size());
  h.bodies = uint32_t(bodies);
  h.fps = uint32_t(fps);
  h.snapshotBytes = sizeof(DiskSnapshot);
  h.diagnosticsBytes = sizeof(mmc_frame);
  out.write(reinterpret_cast<const char*>(&h), sizeof(h));
  for (size_t f = 0; f < result.frames.size(); ++f) {
    mmc_frame d = to_disk(result.diagnostics[f]);
    out.write(reinterpret_cast<const char*>(&d), sizeof(d));
    for (const Snapshot& s : result.frames[f]) {
      DiskSnapshot q = to_disk(s);
      out.write(reinterpret_cast<const char*>(&q), sizeof(q));
    }
  }
}

// MC/src/main.cpp:112-128
void write_metrics(const std::filesystem::path& path, const SimulationResult& result) {
  std::ofstream out(path);
  if (!out) throw std::runtime_error("cannot write metrics " + path.string());
  out << "frame,time,loading,platen_y_m,displacement_m,engineering_strain,engineering_stress_pa,"
         "reaction_force_n,loading_work_j,recovered_work_j,dissipated_work_j,sea_j_per_kg,"
         "elastic_energy_j,incremental_plastic_dissipation_j,contact_energy_j,residual_rms,residual_max,"
         "max_penetration_m,mean_plastic_strain,max_plastic_strain,mean_plastic_rotation_rad,"
         "max_plastic_rotation_rad,permanent_set,active_contacts,plastic_beams,plastic_hinges,iterations,converged\n";
  for (size_t f = 0; f < result.diagnostics.size(); ++f) {
    const auto& d = result.diagnostics[f];
    out << f << ',' << std::setprecision(14) << d.time << ',' << (d.loading ? 1 : 0) << ',' << d.platenY << ','
        << d.displacement << ',' << d.engineeringStrain << ',' << d.engineeringStress << ',' << d.reactionForce << ','
        << d.loadingWork << ',' << d.recoveredWork << ',' << d.dissipatedWork << ',' << d.specificEnergyAbsorption
        << ',' << d.elasticEnergy << ',' << d.plasticDissipation << ',' << d.contactEnergy << ',' << d.residualRms
        << ',' << d.residualMax << ',' << d.maxPenetration << ',' << d.meanPlasticStrain << ','
        << d.maxPlasticStrain << ',' << d.meanPlasticRotation << ',' << d.maxPlasticRotation << ','
        << d.permanentSet << ',' << d.activeContacts << ',' << d.plasticBeams << ',' << d.plasticHinges << ','
        << d.iterations << ',' << (d.converged ? 1 : 0) << '\n';
  }
}

// MC/src/main.cpp:130-135
struct Summary {
  double peakForce = 0, peakStress = 0, loadingWork = 0, recoveredWork = 0, dissipatedWork = 0, sea = 0;
  double maxPen = 0, permanentSet = 0, maxPlasticStrain = 0, maxPlasticRotation = 0;
  int unconverged = 0, frames = 0, peakForceFrame = 0;
};
Summary summarize(const SimulationResult& r) {
  Summary s;
  s.frames = int(r.diagnostics.size());
  for (int i = 0; i < int(r.diagnostics.size()); ++i) {
    const auto& d = r.diagnostics[size_t(i)];
    if (d.reactionForce > s.peakForce) {
      s.peakForce = d.reactionForce;
      s.peakForceFrame = i;
    }
    s.peakStress = std::max(s.peakStress, d.engineeringStress);
    s.maxPen = std::max(s.maxPen, d.maxPenetration);
    s.maxPlasticStrain = std::max(s.maxPlasticStrain, d.maxPlasticStrain);
    s.maxPlasticRotation = std::max(s.maxPlasticRotation, d.maxPlasticRotation);
    if (!d.converged) ++s.unconverged;
  }
  if (!r.diagnostics.empty()) {
    const auto& f = r.diagnostics.back();
    s.loadingWork = f.loadingWork;
    s.recoveredWork = f.recoveredWork;
    s.dissipatedWork = f.dissipatedWork;
    s.sea = f.specificEnergyAbsorption;
    s.permanentSet = f.permanentSet;
  }
  return s;
}

// MC/src/main.cpp:137-155
void write_summary(const std::filesystem::path& path, const SimulationResult& r, const SimulationConfig& cfg,
                   double massKg) {
  Summary s = summarize(r);
  std::ofstream out(path);
  if (!out) throw std::runtime_error("cannot write summary");
  out << "{\n"
      << "  \"solver\": \"incremental quasi-static 6-DOF corotational beam/contact energy minimization\",\n"
      << "  \"material\": \"" << cfg.material.name << "\",\n"
      << "  \"cells\": [" << cfg.cellsX << ", " << cfg.cellsY << ", " << cfg.cellsZ << "],\n"
      << "  \"frames\": " << s.frames << ",\n"
      << "  \"mass_kg\": " << std::setprecision(12) << massKg << ",\n"
      << "  \"peak_force_n\": " << s.peakForce << ",\n"
      << "  \"peak_stress_pa\": " << s.peakStress << ",\n"
      << "  \"loading_work_j\": " << s.loadingWork << ",\n"
      << "  \"recovered_work_j\": " << s.recoveredWork << ",\n"
      << "  \"dissipated_work_j\": " << s.dissipatedWork << ",\n"
      << "  \"specific_energy_absorption_j_per_kg\": " << s.sea << ",\n"
      << "  \"permanent_set\": " << s.permanentSet << ",\n"
      << "  \"max_penetration_m\": " << s.maxPen << ",\n"
      << "  \"unconverged_steps\": " << s.unconverged << "\n"
      << "}\n";
}

struct Args {
  std::string mode = "all";
  std::filesystem::path output = "output", cache;
  int threads = 0;
  bool quick = false, noUnload = false, skipStl = false;
  bool strictValidation = false;  // exit 1 when VALIDATION.txt holds a FAIL (additive)
  int stlQuality = 14;
  SimulationConfig sim;
  SimulationExtensions ext;  // the additive options
};

void usage() {
  std::cout << "metamaterial_compression --mode sim|render|hero|stl|validate|convergence|all [options]\n"
               "gfx950 build: sim, stl, validate, convergence (all = sim + stl + validate; render/hero not built)\n"
               "additive options: --lattice bcc|cubic --cell-size M --min-radius M --max-radius M --nu V\n"
               "  --density V --seed N --history N --grad-tol V --grad-rms-tol V --grad-max-tol V\n"
               "  --no-relative-escape --position-sigma M --radius-mode 0|1 --line-search 0|1|2 --imperfection M\n"
               "  --member-imperfection M --joint-scale V --precision fp64|mixed\n"
               "  --strict-validation (exit 1 when VALIDATION.txt holds a FAIL)\n";
}

// MC/src/main.cpp:165-181 plus the additive options
Args parse_args(int argc, char** argv) {
  Args a;
  for (int n = 1; n < argc; ++n) {
    std::string o = argv[n];
    auto value = [&]() {
      if (++n >= argc) throw std::runtime_error("missing value for " + o);
      return std::string(argv[n]);
    };
    if (o == "--mode") a.mode = value();
    else if (o == "--output") a.output = value();
    else if (o == "--cache") a.cache = value();
    else if (o == "--threads") a.threads = std::stoi(value());
    else if (o == "--width" || o == "--height" || o == "--spp" || o == "--depth" || o == "--frame-start" ||
             o == "--frame-end" || o == "--frame-step" || o == "--subframes" || o == "--hero-width" ||
             o == "--hero-height" || o == "--hero-spp")
      (void)std::stoi(value());  // render options: accepted, unused here
    else if (o == "--stl-quality") a.stlQuality = std::stoi(value());
    else if (o == "--cells-x") a.sim.cellsX = std::stoi(value());
    else if (o == "--cells-y") a.sim.cellsY = std::stoi(value());
    else if (o == "--cells-z") a.sim.cellsZ = std::stoi(value());
    else if (o == "--load-steps") a.sim.loadSteps = std::stoi(value());
    else if (o == "--unload-steps") a.sim.unloadSteps = std::stoi(value());
    else if (o == "--max-strain") a.sim.maxStrain = std::stod(value());
    else if (o == "--max-iterations") a.sim.maxIterations = std::stoi(value());
    else if (o == "--young") a.sim.material.young = std::stod(value());
    else if (o == "--yield") a.sim.material.yield = std::stod(value());
    else if (o == "--hardening") a.sim.material.hardening = std::stod(value());
    else if (o == "--contact-stiffness") a.sim.contactStiffness = std::stod(value());
    else if (o == "--quick") a.quick = true;
    else if (o == "--no-unload") a.noUnload = true;
    else if (o == "--no-denoise") {
    } else if (o == "--skip-stl") a.skipStl = true;
    else if (o == "--strict-validation") a.strictValidation = true;
    // ---- additive ----
    else if (o == "--lattice") {
      std::string v = value();
      if (v == "bcc") a.ext.lattice = MMC_LATTICE_BCC;
      else if (v == "cubic") a.ext.lattice = MMC_LATTICE_CUBIC;
      else throw std::runtime_error("unknown lattice " + v);
    } else if (o == "--cell-size") a.sim.cellSize = std::stod(value());
    else if (o == "--min-radius") a.sim.minRadius = std::stod(value());
    else if (o == "--max-radius") a.sim.maxRadius = std::stod(value());
    else if (o == "--nu") a.sim.material.nu = std::stod(value());
    else if (o == "--density") a.sim.material.density = std::stod(value());
    else if (o == "--seed") a.sim.seed = std::stoull(value(), nullptr, 0);
    else if (o == "--history") a.sim.history = std::stoi(value());
    else if (o == "--grad-tol") a.sim.gradientRmsTolerance = a.sim.gradientMaxTolerance = std::stod(value());
    else if (o == "--grad-rms-tol") a.sim.gradientRmsTolerance = std::stod(value());
    else if (o == "--grad-max-tol") a.sim.gradientMaxTolerance = std::stod(value());
    else if (o == "--no-relative-escape") a.ext.relativeEscape = false;
    else if (o == "--position-sigma") a.ext.positionSigma = std::stod(value());
    else if (o == "--radius-mode") a.ext.radiusMode = std::stoi(value());
    else if (o == "--line-search") a.ext.lineSearch = std::stoi(value());
    else if (o == "--imperfection") a.sim.imperfection = std::stod(value());
    else if (o == "--member-imperfection") a.sim.memberImperfection = std::stod(value());
    else if (o == "--joint-scale") a.sim.jointScale = std::stod(value());
    else if (o == "--precision") {
      std::string v = value();
      if (v == "fp64") a.ext.precision = MMC_FP64;
      else if (v == "mixed") a.ext.precision = MMC_MIXED;
      else throw std::runtime_error("unknown precision " + v);
    } else if (o == "--help") {
      usage();
      std::exit(0);
    } else
      throw std::runtime_error("unknown option " + o);
  }
  if (a.quick) {
    a.sim.cellsX = 3;
    a.sim.cellsY = 4;
    a.sim.cellsZ = 3;
    a.sim.loadSteps = 24;
    a.sim.unloadSteps = 14;
    a.sim.maxIterations = 95;
    a.stlQuality = 10;
  }
  if (a.noUnload) a.sim.unloadSteps = 0;
  a.sim.threads = a.threads;
  a.sim.frames = a.sim.loadSteps + a.sim.unloadSteps + 1;
  if (a.cache.empty()) a.cache = a.output / "compression.meta3";
  return a;
}

// MC/src/main.cpp:198
bool finite_result(const SimulationResult& r) {
  for (const auto& d : r.diagnostics) {
    const double values[] = {d.time, d.platenY, d.displacement, d.engineeringStrain, d.engineeringStress,
                             d.reactionForce, d.loadingWork, d.recoveredWork, d.dissipatedWork,
                             d.specificEnergyAbsorption, d.elasticEnergy, d.plasticDissipation, d.contactEnergy,
                             d.residualRms, d.residualMax, d.maxPenetration, d.meanPlasticStrain,
                             d.maxPlasticStrain, d.meanPlasticRotation, d.maxPlasticRotation, d.permanentSet};
    for (double v : values)
      if (!std::isfinite(v)) return false;
  }
  auto fin = [](const Vec3& v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); };
  for (const auto& f : r.frames)
    for (const auto& s : f)
      if (!fin(s.p) || !fin(s.v) || !fin(s.w) || !std::isfinite(s.q.w) || !std::isfinite(s.q.x) ||
          !std::isfinite(s.q.y) || !std::isfinite(s.q.z))
        return false;
  return true;
}

// MC/src/main.cpp:200-208
// Returns false when any check of the report is FAIL (the reference's exit
// code ignores that; --strict-validation turns it into exit code 1).
bool write_validation(const std::filesystem::path& path, const Simulator& sim, const SimulationResult* result) {
  double objectivity = std::numeric_limits<double>::infinity(), gradient = objectivity;
  bool obj = sim.rigidRotationTest(&objectivity), grad = sim.gradientCheck(&gradient);
  std::ofstream out(path);
  if (!out) throw std::runtime_error("cannot write validation report");
  out << "Metamaterial compression engineering validation\n"
      << "===============================================\n"
      << "rigid_motion_objectivity_error=" << std::setprecision(12) << objectivity << " " << (obj ? "PASS" : "FAIL")
      << '\n'
      << "finite_difference_gradient_relative_error=" << gradient << " " << (grad ? "PASS" : "FAIL") << '\n'
      << "nodes=" << sim.nodes().size() << '\n'
      << "beams=" << sim.beams().size() << '\n'
      << "mass_kg=" << sim.mass() << '\n'
      << "reference_area_m2=" << sim.area() << '\n';
  if (result) {
    Summary s = summarize(*result);
    bool finite = finite_result(*result);
    double convergedFraction =
        result->diagnostics.empty() ? 0 : 1.0 - double(s.unconverged) / result->diagnostics.size();
    bool penetration = s.maxPen < 0.00025;
    bool energy = s.loadingWork + 1e-10 >= s.recoveredWork && s.dissipatedWork >= -1e-10;
    out << "finite_state=" << (finite ? "PASS" : "FAIL") << '\n'
        << "converged_step_fraction=" << convergedFraction << '\n'
        << "peak_force_n=" << s.peakForce << '\n'
        << "loading_work_j=" << s.loadingWork << '\n'
        << "recovered_work_j=" << s.recoveredWork << '\n'
        << "dissipated_work_j=" << s.dissipatedWork << '\n'
        << "specific_energy_absorption_j_per_kg=" << s.sea << '\n'
        << "permanent_set=" << s.permanentSet << '\n'
        << "max_contact_penetration_mm=" << s.maxPen * 1000 << " " << (penetration ? "PASS" : "WARN") << '\n'
        << "energy_accounting=" << (energy ? "PASS" : "FAIL") << '\n'
        << "overall_numerical_validation=" << ((obj && grad && finite && energy) ? "PASS" : "FAIL") << '\n';
    return obj && grad && finite && energy;
  }
  return obj && grad;
}

// MC/src/main.cpp:210-228
void run_convergence(const Args& args) {
  std::filesystem::create_directories(args.output);
  struct Case {
    const char* name;
    int steps, iterations;
  };
  std::array<Case, 3> cases{{{"coarse", 10, 500}, {"medium", 14, 500}, {"fine", 20, 600}}};
  std::ofstream csv(args.output / "convergence.csv");
  csv << "case,load_steps,unload_steps,max_iterations,peak_force_n,loading_work_j,dissipated_work_j,permanent_set,"
         "max_penetration_m,unconverged_steps\n";
  for (const Case& k : cases) {
    SimulationConfig c = args.sim;
    c.cellsX = 2;
    c.cellsY = 4;
    c.cellsZ = 2;
    c.maxStrain = .50;
    c.loadSteps = k.steps;
    c.unloadSteps = std::max(7, int(std::lround(.70 * k.steps)));
    c.maxIterations = k.iterations;
    c.contactStiffness = 1.0e6;
    c.frames = c.loadSteps + c.unloadSteps + 1;
    std::cout << "\n[convergence] running " << k.name << "\n";
    setNextSimulationExtensions(args.ext);
    Simulator sim(c);
    SimulationResult r = sim.run();
    Summary s = summarize(r);
    csv << k.name << ',' << k.steps << ',' << c.unloadSteps << ',' << k.iterations << ',' << std::setprecision(12)
        << s.peakForce << ',' << s.loadingWork << ',' << s.dissipatedWork << ',' << s.permanentSet << ','
        << s.maxPen << ',' << s.unconverged << '\n';
    save_cache(args.output / (std::string("convergence_") + k.name + ".meta3"), r, c.fps);
  }
}

}  // namespace
}  // namespace meta

int main(int argc, char** argv) {
  using namespace meta;
  try {
    Args args = parse_args(argc, argv);
    std::filesystem::create_directories(args.output);
    if (args.mode == "convergence") {
      run_convergence(args);
      return 0;
    }
    if (args.mode == "render" || args.mode == "hero")
      throw std::runtime_error("--mode " + args.mode + " is not part of the gfx950 equilibrium solver build");
    SimulationResult result;
    std::unique_ptr<Simulator> sim;
    if (args.mode == "sim" || args.mode == "all") {
      setNextSimulationExtensions(args.ext);
      sim = std::make_unique<Simulator>(args.sim);
      result = sim->run();
      save_cache(args.cache, result, args.sim.fps);
      write_metrics(args.output / "compression_metrics.csv", result);
      write_summary(args.output / "summary.json", result, args.sim, sim->mass());
      if (!args.skipStl) sim->exportSTL(args.output / "graded_bcc_energy_absorber.stl", args.stlQuality);
      if (!write_validation(args.output / "VALIDATION.txt", *sim, &result) && args.strictValidation)
        throw std::runtime_error("validation failed (VALIDATION.txt)");
    } else if (args.mode == "validate" || args.mode == "stl") {
      setNextSimulationExtensions(args.ext);
      sim = std::make_unique<Simulator>(args.sim);
      if (args.mode == "stl") sim->exportSTL(args.output / "graded_bcc_energy_absorber.stl", args.stlQuality);
      const bool ok = write_validation(args.output / "VALIDATION.txt", *sim, nullptr);
      if (!ok && args.strictValidation) throw std::runtime_error("validation failed (VALIDATION.txt)");
      return 0;
    } else {
      throw std::runtime_error("unknown mode " + args.mode);
    }
    if (!result.diagnostics.empty()) {
      Summary s = summarize(result);
      std::cout << "[final] peak_force=" << s.peakForce << " N loading_work=" << s.loadingWork
                << " J recovered=" << s.recoveredWork << " J dissipated=" << s.dissipatedWork << " J SEA=" << s.sea
                << " J/kg permanent_set=" << s.permanentSet << " max_penetration=" << s.maxPen * 1000
                << " mm unconverged=" << s.unconverged << '\n';
    }
    return 0;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << '\n';
    return 1;
  }
}
