// Fundamental-scalar instantiations (rep_fundamental.hpp).
#include "launch.hpp"
namespace ghx {
bool ops_fund(int prec, int N, int D, Ops* o) {
  if (N == 2) return pick<2, RepFund>(prec, D, o);
  if (N == 3) return pick<3, RepFund>(prec, D, o);
  return false;
}
}  // namespace ghx
