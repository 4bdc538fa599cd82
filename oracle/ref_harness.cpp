// ref_harness.cpp — exposes the reference solver, compiled from its own sources,
// through the oracle's C API with the prefix ref_ (TEST INFRASTRUCTURE ONLY).
//
// Built by oracle/Makefile from the sources where they lie under
// /root/reference/projects/metamaterial-compression/industry-grade (MC/): this
// translation unit includes MC/src/sim.cpp so that the private hot path
// (objective/minimize/commit, MC/include/sim.hpp:48-53) can be driven directly.
// The output goes to oracle/_ref/ only; no reference source enters the repository.
//
// A cubic lattice is injected by replacing the nodes, beams and exclusion set
// of a constructed Simulator and re-running its distributeMass()
// (MC/src/sim.cpp:134), which pins everything downstream of lattice generation
// on the lattice family the benchmark configurations use.
#include "core.hpp"
#define private public
#include "sim.hpp"
#undef private
#include "sim.cpp"

#include <cstring>
#include <sstream>

#include "../oracle/oracle.h"

using namespace meta;

struct ref_h {
  std::unique_ptr<Simulator> sim;
  std::vector<uint64_t> connected_list;
  Objective last;
  std::vector<double> lastZ;
  bool have_last = false;
  std::string err;
};

namespace {
SimulationConfig to_ref(const mmc_config* c) {
  SimulationConfig s;
  s.cellsX = c->cells_x;
  s.cellsY = c->cells_y;
  s.cellsZ = c->cells_z;
  s.cellSize = c->cell_size;
  s.minRadius = c->min_radius;
  s.maxRadius = c->max_radius;
  s.jointScale = c->joint_scale;
  s.triggerScale = c->trigger_scale;
  s.imperfection = c->imperfection;
  s.memberImperfection = c->member_imperfection;
  s.elementsPerStrut = c->elements_per_strut;
  s.material.density = c->material.density;
  s.material.young = c->material.young;
  s.material.nu = c->material.nu;
  s.material.yield = c->material.yield;
  s.material.hardening = c->material.hardening;
  s.material.bendHardening = c->material.bend_hardening;
  s.material.torsionHardening = c->material.torsion_hardening;
  s.gravity = c->gravity;
  s.shearCorrection = c->shear_correction;
  s.contactStiffness = c->contact_stiffness;
  s.contactMargin = c->contact_margin;
  s.centeringStiffness = c->centering_stiffness;
  s.contactSamples = c->contact_samples;
  s.loadSteps = c->load_steps;
  s.unloadSteps = c->unload_steps;
  s.maxIterations = c->max_iterations;
  s.history = c->history;
  s.maxLineSearch = c->max_line_search;
  s.maxStrain = c->max_strain;
  s.stepTime = c->step_time;
  s.translationScale = c->translation_scale;
  s.maxTranslationStep = c->max_translation_step;
  s.maxRotationStep = c->max_rotation_step;
  s.gradientRmsTolerance = c->gradient_rms_tolerance;
  s.gradientMaxTolerance = c->gradient_max_tolerance;
  s.armijo = c->armijo;
  s.minLineStep = c->min_line_step;
  s.penetrationTolerance = c->penetration_tolerance;
  s.maxAdaptiveRefinements = c->max_adaptive_refinements;
  s.fps = c->fps;
  s.frames = c->frames;
  s.threads = c->threads;
  s.seed = c->seed;
  return s;
}
void trial_row(const TrialBeam& t, double* o) {
  const double v[MMC_TRIAL_STRIDE] = {t.next.epsp, t.next.alpha, t.next.rotp.x, t.next.rotp.y, t.next.rotp.z,
                                      t.next.alphaB, t.next.alphaT, t.strain, t.stress, t.moment.x, t.moment.y,
                                      t.moment.z, t.elastic, t.dissipation, double(t.axialPlastic),
                                      double(t.bendPlastic), double(t.torsionPlastic)};
  std::memcpy(o, v, sizeof v);
}
void fill(const Objective& o, const Simulator& s, mmc_objective_out* out) {
  if (!out) return;
  out->energy = o.energy;
  out->elastic = o.d.elastic;
  out->plastic = o.d.plastic;
  out->contact = o.d.contact;
  out->reaction = o.d.reaction;
  out->max_pen = o.d.maxPen;
  out->contacts = o.d.contacts;
  out->finite = std::isfinite(o.energy);
  out->rms = vrms(o.grad, s.locked_);
  out->max = vmax(o.grad, s.locked_);
}
mmc_frame to_frame(const FrameDiagnostics& d) {
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
  f.converged = d.converged;
  f.loading = d.loading;
  return f;
}
// The reference keeps connected_ in an unordered_set; export it sorted.
void collect_connected(ref_h* h) {
  h->connected_list.assign(h->sim->connected_.begin(), h->sim->connected_.end());
  std::sort(h->connected_list.begin(), h->connected_list.end());
}
}  // namespace

extern "C" {

ref_h* ref_create(const mmc_config* cfg) {
  ref_h* h = new ref_h();
  try {
    h->sim = std::make_unique<Simulator>(to_ref(cfg));
    collect_connected(h);
  } catch (const std::exception& e) {
    h->err = e.what();
  }
  return h;
}

ref_h* ref_create_lattice(const mmc_config* cfg, const mmc_lattice* L) {
  ref_h* h = new ref_h();
  try {
    SimulationConfig small = to_ref(cfg);
    small.cellsX = small.cellsY = small.cellsZ = 1;
    h->sim = std::make_unique<Simulator>(small);
    Simulator& s = *h->sim;
    s.cfg_ = to_ref(cfg);
    s.cfg_.translationScale = s.cfg_.cellSize;
    s.nodes_.clear();
    s.beams_.clear();
    s.connected_.clear();
    for (int i = 0; i < L->n_nodes; i++) {
      Node n;
      n.x0 = Vec3{L->x[3 * i], L->x[3 * i + 1], L->x[3 * i + 2]};
      n.x = n.x0;
      if (L->q) n.R = Quat{L->q[4 * i], L->q[4 * i + 1], L->q[4 * i + 2], L->q[4 * i + 3]};
      n.layer = L->node_layer ? L->node_layer[i] : 0;
      s.nodes_.push_back(n);
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
      e.e0 = Vec3{L->e0[3 * b], L->e0[3 * b + 1], L->e0[3 * b + 2]};
      if (L->plastic) {
        const double* p = L->plastic + 7 * b;
        e.plastic.epsp = p[0];
        e.plastic.alpha = p[1];
        e.plastic.rotp = Vec3{p[2], p[3], p[4]};
        e.plastic.alphaB = p[5];
        e.plastic.alphaT = p[6];
      }
      s.beams_.push_back(e);
    }
    for (int64_t c = 0; c < L->n_connected; c++) s.connected_.insert(L->connected[c]);
    s.width_ = L->width;
    s.height_ = L->height;
    s.depth_ = L->depth;
    s.area_ = s.width_ * s.depth_;
    double jr = s.cfg_.jointScale * s.cfg_.maxRadius;
    s.top0_ = jr + s.height_ + jr;
    s.distributeMass();
    s.baseX_.resize(s.nodes_.size());
    s.baseR_.resize(s.nodes_.size());
    for (size_t i = 0; i < s.nodes_.size(); i++) {
      s.baseX_[i] = s.nodes_[i].x;
      s.baseR_[i] = s.nodes_[i].R;
    }
    collect_connected(h);
  } catch (const std::exception& e) {
    h->err = e.what();
  }
  return h;
}

void ref_destroy(ref_h* h) { delete h; }
const char* ref_last_error(ref_h* h) { return h->err.c_str(); }

void ref_info(ref_h* h, int32_t* nn, int32_t* nb, int64_t* nc, double* env) {
  Simulator& s = *h->sim;
  if (nn) *nn = int32_t(s.nodes_.size());
  if (nb) *nb = int32_t(s.beams_.size());
  if (nc) *nc = int64_t(h->connected_list.size());
  if (env) {
    env[0] = s.width_;
    env[1] = s.height_;
    env[2] = s.depth_;
    env[3] = s.area_;
    env[4] = s.mass_;
    env[5] = s.top0_;
  }
}

void ref_lattice(ref_h* h, double* x, double* q, double* mass, uint8_t* locked, int32_t* a, int32_t* b, double* L0,
                 double* radius, double* area, double* inertia, double* e0, double* plastic, uint64_t* connected,
                 int32_t* node_layer, int32_t* beam_layer, int32_t* beam_family) {
  Simulator& s = *h->sim;
  for (size_t i = 0; i < s.nodes_.size(); i++) {
    const Node& n = s.nodes_[i];
    if (x) {
      x[3 * i] = s.baseX_[i].x;
      x[3 * i + 1] = s.baseX_[i].y;
      x[3 * i + 2] = s.baseX_[i].z;
    }
    if (q) {
      q[4 * i] = s.baseR_[i].w;
      q[4 * i + 1] = s.baseR_[i].x;
      q[4 * i + 2] = s.baseR_[i].y;
      q[4 * i + 3] = s.baseR_[i].z;
    }
    if (mass) mass[i] = n.mass;
    if (node_layer) node_layer[i] = n.layer;
  }
  if (locked) std::memcpy(locked, s.locked_.data(), s.locked_.size());
  for (size_t k = 0; k < s.beams_.size(); k++) {
    const Beam& e = s.beams_[k];
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

double ref_objective(ref_h* h, const double* z, double top, int32_t trials, mmc_objective_out* out, double* grad,
                     double* trial) {
  Simulator& s = *h->sim;
  std::vector<double> zz(z, z + s.nodes_.size() * 6);
  Objective o = s.objective(zz, top, trials != 0);
  fill(o, s, out);
  if (grad) std::memcpy(grad, o.grad.data(), o.grad.size() * sizeof(double));
  if (trial && trials)
    for (size_t k = 0; k < o.trial.size(); k++) trial_row(o.trial[k], trial + MMC_TRIAL_STRIDE * k);
  return o.energy;
}

int ref_minimize(ref_h* h, double top, double* z, mmc_opt_stats* st, mmc_objective_out* fin) {
  Simulator& s = *h->sim;
  std::vector<double> zz(z, z + s.nodes_.size() * 6);
  Objective f;
  OptStats o = s.minimize(top, zz, f);
  std::memcpy(z, zz.data(), zz.size() * sizeof(double));
  h->last = f;
  h->lastZ = zz;
  h->have_last = true;
  if (st) {
    st->iterations = o.iterations;
    st->backtracks = o.backtracks;
    st->rms = o.rms;
    st->max = o.max;
    st->energy = o.energy;
    st->converged = o.converged;
    st->evaluations = 0;
  }
  if (fin) {
    fill(f, s, fin);
    fin->rms = o.rms;
    fin->max = o.max;
  }
  return 0;
}

int ref_commit(ref_h* h) {
  if (!h->have_last) {
    h->err = "missing constitutive trial states";
    return MMC_ERR_STATE;
  }
  try {
    h->sim->commit(h->lastZ, h->last);
  } catch (const std::exception& e) {
    h->err = e.what();
    return MMC_ERR_STATE;
  }
  h->have_last = false;
  return 0;
}

int ref_run(ref_h* h, mmc_frame* frames, int32_t cap, int32_t* n, int32_t verbose) {
  std::ostringstream sink;
  std::streambuf* old = nullptr;
  if (!verbose) old = std::cout.rdbuf(sink.rdbuf());
  int rc = 0;
  SimulationResult r;
  try {
    r = h->sim->run();
  } catch (const std::exception& e) {
    h->err = e.what();
    rc = MMC_ERR_EQUILIBRIUM;
  }
  if (old) std::cout.rdbuf(old);
  int m = std::min<int>(cap, int(r.diagnostics.size()));
  for (int i = 0; i < m; i++) frames[i] = to_frame(r.diagnostics[size_t(i)]);
  if (n) *n = int32_t(r.diagnostics.size());
  return rc;
}

double ref_beam_potential(ref_h* h, const double* x, const double* q) {
  Simulator& s = *h->sim;
  size_t N = s.nodes_.size();
  std::vector<Vec3> xx(N);
  std::vector<Quat> RR(N);
  for (size_t i = 0; i < N; i++) {
    xx[i] = Vec3{x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    RR[i] = Quat{q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
  }
  return s.beamPotential(xx, RR);
}
int ref_export_stl(ref_h* h, const char* path, int32_t segments) {
  try {
    h->sim->exportSTL(path, segments);
  } catch (const std::exception& e) {
    h->err = e.what();
    return MMC_ERR_STATE;
  }
  return 0;
}

void ref_capture(ref_h* h, double top, const double* px, const double* pq, mmc_snapshot* out) {
  const size_t N = h->sim->nodes_.size();
  std::vector<Vec3> X;
  std::vector<Quat> R;
  if (px) {
    X.resize(N);
    for (size_t i = 0; i < N; i++) X[i] = Vec3{px[3 * i], px[3 * i + 1], px[3 * i + 2]};
  }
  if (pq) {
    R.resize(N);
    for (size_t i = 0; i < N; i++) R[i] = Quat{pq[4 * i], pq[4 * i + 1], pq[4 * i + 2], pq[4 * i + 3]};
  }
  std::vector<Snapshot> s = h->sim->capture(top, px ? &X : nullptr, pq ? &R : nullptr);
  for (size_t k = 0; k < s.size(); k++) {
    mmc_snapshot& o = out[k];
    const Snapshot& a = s[k];
    const double v[16] = {a.p.x, a.p.y, a.p.z, a.q.w, a.q.x, a.q.y, a.q.z, a.v.x, a.v.y, a.v.z,
                          a.w.x, a.w.y, a.w.z, a.half.x, a.half.y, a.half.z};
    std::memcpy(o.p, v, 3 * sizeof(double));
    std::memcpy(o.q, v + 3, 4 * sizeof(double));
    std::memcpy(o.v, v + 7, 3 * sizeof(double));
    std::memcpy(o.w, v + 10, 3 * sizeof(double));
    std::memcpy(o.half, v + 13, 3 * sizeof(double));
    o.material = a.material;
    o.id = a.id;
    o.kind = a.kind;
    o.module_class = a.moduleClass;
    o.ring = a.ring;
    o.slot = a.slot;
  }
}
int ref_rigid_rotation_test(ref_h* h, double* err) { return h->sim->rigidRotationTest(err) ? 1 : 0; }
int ref_gradient_check(ref_h* h, double* err) { return h->sim->gradientCheck(err) ? 1 : 0; }
void ref_read_state(ref_h* h, double* x, double* q, double* plastic) {
  ref_lattice(h, x, q, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, plastic,
              nullptr, nullptr, nullptr, nullptr);
}
void ref_set_plastic(ref_h* h, const double* p) {
  Simulator& s = *h->sim;
  for (size_t k = 0; k < s.beams_.size(); k++) {
    const double* v = p + 7 * k;
    PlasticState& q = s.beams_[k].plastic;
    q.epsp = v[0];
    q.alpha = v[1];
    q.rotp = Vec3{v[2], v[3], v[4]};
    q.alphaB = v[5];
    q.alphaT = v[6];
  }
}
void ref_set_pose(ref_h* h, const double* x, const double* q) {
  Simulator& s = *h->sim;
  for (size_t i = 0; i < s.nodes_.size(); i++) {
    s.baseX_[i] = Vec3{x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    s.baseR_[i] = Quat{q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
    s.nodes_[i].x = s.baseX_[i];
    s.nodes_[i].R = s.baseR_[i];
  }
}
void ref_scalar_return(double total, double pc, double ac, double k, double y, double H, double* o) {
  ScalarReturn r = scalarReturn(total, pc, ac, k, y, H);
  const double v[7] = {r.force, r.p, r.a, r.dg, r.energy, r.hard, double(r.yielded)};
  std::memcpy(o, v, sizeof v);
}
void ref_vector_return(const double* t, const double* pc, double ac, double k, double y, double H, double* o) {
  VectorReturn r = vectorReturn(Vec3{t[0], t[1], t[2]}, Vec3{pc[0], pc[1], pc[2]}, ac, k, y, H);
  const double v[12] = {r.force.x, r.force.y, r.force.z, r.p.x, r.p.y, r.p.z, r.a, r.dg, r.energy, r.hard,
                        double(r.yielded), 0};
  std::memcpy(o, v, sizeof v);
}
void ref_closest_segments(const double* p1, const double* q1, const double* p2, const double* q2, double* o) {
  ClosestPair c = closestSegments(Vec3{p1[0], p1[1], p1[2]}, Vec3{q1[0], q1[1], q1[2]}, Vec3{p2[0], p2[1], p2[2]},
                                  Vec3{q2[0], q2[1], q2[2]});
  const double v[8] = {c.s, c.t, c.pa.x, c.pa.y, c.pa.z, c.pb.x, c.pb.y, c.pb.z};
  std::memcpy(o, v, sizeof v);
}
void ref_quat_exp(const double* v, double* q) {
  Quat r = quatExp(Vec3{v[0], v[1], v[2]});
  q[0] = r.w;
  q[1] = r.x;
  q[2] = r.y;
  q[3] = r.z;
}
void ref_quat_log(const double* q, double* v) {
  Vec3 r = quatLog(Quat{q[0], q[1], q[2], q[3]});
  v[0] = r.x;
  v[1] = r.y;
  v[2] = r.z;
}
void ref_jl_inv_t(const double* p, const double* v, double* o) {
  Vec3 r = leftJacobianInverseTransposeApply(Vec3{p[0], p[1], p[2]}, Vec3{v[0], v[1], v[2]});
  o[0] = r.x;
  o[1] = r.y;
  o[2] = r.z;
}

}  // extern "C"
