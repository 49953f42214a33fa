// Host-side launcher of the one-cell-per-lane persistent solver
// (lane_kernels.hip), called by p3d_chem_solve.
#pragma once
#include <stdint.h>

#include "../../include/prism3d_amd.h"

namespace p3d {

// Structure of the network the lane kernel was generated for.
bool lane_network_matches(int ns, int nr, const int *rtype, const int *reactants,
                          const int *products, int e_index, const int *roles7);

struct LaneArgs {
  int64_t n;
  double *y;
  const double *nH, *T, *G0, *AV, *zeta, *fH2, *fCO;
  int32_t *conv, *stats;
  // device tables of the network handle
  const int *rtype;
  const double *alpha, *beta, *gamma, *post_w;
  p3d_solve_opts opts;
  void *stream;
};

// Slot-memory entries per lane (doubles).
int lane_slot_doubles();

// Fraction of lanes that must have found the work queue empty before the
// cells still integrating are handed to the one-cell-per-workgroup kernel.
constexpr double kLaneRescueFrac = 0.95;

// Device work queue of one launch: head counter, number of lanes that found
// it empty, and the list of cells handed over (count + indices).
struct LaneQueue {
  unsigned long long *head;
  unsigned *exited;
  unsigned *n_rescue;
  int64_t *rescue;
  unsigned exit_threshold;
};

// Launches the persistent kernel over `n_waves` wavefronts compiled for
// `occupancy` (2, 3 or 4) wavefronts per SIMD; ws holds
// n_waves * 64 * lane_slot_doubles() doubles.  The queue counters are reset
// on the stream first.
int lane_launch(const LaneArgs &a, int64_t n_waves, int occupancy, double *ws,
                const LaneQueue &q);

}  // namespace p3d
