// generate.hpp — lattice construction on the device: Simulator::build and
// distributeMass (projects/metamaterial-compression/industry-grade/src/
// sim.cpp:49-135) for the graded BCC family, and the simple cubic family, for
// many lattices at once, written straight into a context's SoA planes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/mmc.h"

namespace mmc {

// Closed-form sizes of the lattice a configuration builds (no device work).
struct GenCounts {
  int nodes = 0, beams = 0;
  long connected = 0;  // exclusion pairs
  long draws = 0;      // values taken from the configuration's RNG stream
};
GenCounts gen_counts(const mmc_config& c);

// Device arrays of the concatenated lattices (planes of the totals N, B).
struct GenOut {
  double *x, *q, *mass;  // x[3][N], q[4][N] (identity), mass[N]
  uint8_t* lock;         // bit d locks DOF d
  int *a, *b;            // global node indices
  double *L0, *radius, *area, *inertia, *e0;  // e0[3][B]
  int *node_layer, *beam_layer, *beam_family;
  int *inc_off, *inc;    // CSR beam incidence over the global nodes, ordered by beam (entry 2k + end)
  int *adj_off, *adj;    // CSR exclusion adjacency, sorted (adj sized by the connected total x 2)
};

// What the host keeps of each lattice.
struct GenResult {
  double total_mass = 0;  // distributeMass's mass_ (beams in order, then joints)
  int lock_a = -1, lock_b = -1;  // the anti-drift nodes (local indices)
};

// Builds the lattices of cfgs (n0[i], b0[i]: their first node and beam in
// the concatenation; N, B: totals) into `out` on `stream`, and returns each
// lattice's mass and lock nodes.  Throws std::runtime_error on failure.
std::vector<GenResult> generate_lattices(const std::vector<mmc_config>& cfgs, const std::vector<int>& n0,
                                         const std::vector<int>& b0, int N, int B, const GenOut& out,
                                         hipStream_t stream);

}  // namespace mmc
