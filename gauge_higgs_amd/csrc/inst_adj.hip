// Adjoint-scalar instantiations (rep_adjoint.hpp).
#include "launch.hpp"
namespace ghx {
bool ops_adj(int prec, int N, int D, Ops* o) {
  if (N == 2) return pick<2, RepAdj>(prec, D, o);
  if (N == 3) return pick<3, RepAdj>(prec, D, o);
  return false;
}
}  // namespace ghx
