// lattice.hpp — host-side lattice construction for the Simulator facade.
//
// Two families behind one call:
//  * graded BCC "accordion" lattice, arithmetic-for-arithmetic the construction
//    of projects/metamaterial-compression/industry-grade/src/sim.cpp:51-132
//    (same RNG stream, same libm calls), so node positions, radii and the
//    self-contact exclusion set equal the reference's bit for bit;
//  * simple cubic lattice with one beam per cell edge (the benchmark lattices).
// Mass lumping and anti-drift locks follow MC/src/sim.cpp:134 for both.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mmc.h"

namespace meta {

struct HostLattice {
  int n_nodes = 0, n_beams = 0;
  std::vector<double> x;         // 3N base positions
  std::vector<int32_t> node_layer;
  std::vector<double> mass;      // N
  std::vector<uint8_t> locked;   // 6N
  std::vector<int32_t> a, b, layer, family;
  std::vector<double> L0, radius, area, inertia, e0;  // e0: 3B
  std::vector<uint64_t> connected;  // node pairs excluded from self-contact (insertion order)
  double width = 0, height = 0, depth = 0, area_ref = 0, total_mass = 0, top0 = 0, joint_radius = 0;
  // committed state after construction (all empty for a new lattice): base
  // orientations (4N, w,x,y,z; empty = identity), plastic states (7B; empty =
  // zero) and the initial positions when x holds a committed base pose (3N;
  // empty = x)
  std::vector<double> q, plastic, x0;

  // A view for mmc_create.
  mmc_lattice view() const;
};

// Throws std::runtime_error on an invalid configuration.
HostLattice build_lattice(const mmc_config& cfg);

}  // namespace meta
