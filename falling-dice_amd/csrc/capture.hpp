// capture.hpp — device side of Simulator::capture (projects/metamaterial-
// compression/industry-grade/src/sim.cpp:215-310): the renderable bodies of a
// committed frame computed where the state lives.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/mmc.h"

namespace mmc {

// One lattice of a capture launch: where its bodies go in the output, its
// node and beam ranges in the concatenated SoA, the platen, and its constants.
struct CapInst {
  long out0;          // first body of the lattice in the output
  int n0, N, b0, B;   // node and beam ranges
  int velocities;     // 0: zero velocities (the reference's first frame)
  double top;         // platen of the frame (the moving platen's two bodies)
  double inv_dt;      // 1 / max(1e-6, step_time)
  double jr;          // joint radius joint_scale * max_radius
};

// Device arrays the kernel reads: committed poses and plastic state in SoA
// planes of the totals, the previous poses of mmc_capture_mark, beam topology
// and the per-item labels of the lattice.
struct CapArgs {
  const CapInst* inst;  // [n] lattices taking part
  int n, N, B;          // lattices; node and beam totals (plane strides)
  int blocks_per_inst;  // grid x
  const double *x, *q;  // committed pose: x[3][N], q[4][N] (w, x, y, z)
  const double *px, *pq;  // previous pose
  const int *a, *b;
  const double *radius, *plastic;  // radius[B], plastic[7][B]
  const int *node_layer, *beam_layer, *beam_family;
  mmc_snapshot* out;
};

void launch_capture(const CapArgs& a, hipStream_t stream);
void launch_capture_mark(const double* x, const double* q, double* px, double* pq, long N, hipStream_t stream);

}  // namespace mmc
