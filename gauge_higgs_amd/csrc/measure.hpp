// Launch table of the measurement kernels (observables.hpp:36-101): Wilson
// loops and Polyakov loops traced in a probe irrep. They depend only on the
// link storage (precision, N, D), not on the scalar's representation.
#pragma once
#include "ops.hpp"

namespace ghx {

// Probe irrep R' of a measurement. kind 1: fundamental, 2: adjoint,
// 3: Young rows (at most N-1 <= 2 rows, rows[0] + 1 <= kProbeMaxH).
struct Probe {
  int kind;
  int nrows;
  int rows[2];
};
constexpr int kProbeMaxH = 24;

struct MeasOps {
  // Per reduction block: sum over its sites x and planes mu < nu of
  // Re chi_R'(W_{Ra x Rb}(x; mu, nu)), into partials[block].
  void (*wilson)(const LaunchCtx&, const void* U, Probe p, int Ra, int Rb);
  // Per reduction block: sum over its transverse sites (x_tdir = 0) of
  // Re chi_R'(prod_t U_tdir), into partials[block].
  void (*polyakov)(const LaunchCtx&, const void* U, Probe p, int tdir);
};

bool meas_ops(int prec, int N, int D, MeasOps* o);

}  // namespace ghx
