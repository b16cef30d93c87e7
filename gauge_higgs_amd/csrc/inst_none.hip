// Pure-gauge instantiations (GaugeHMC path, hmc.hpp).
#include "launch.hpp"
namespace ghx {
bool ops_none(int prec, int N, int D, Ops* o) {
  if (N == 2) return pick<2, RepNone>(prec, D, o);
  if (N == 3) return pick<3, RepNone>(prec, D, o);
  return false;
}
}  // namespace ghx
