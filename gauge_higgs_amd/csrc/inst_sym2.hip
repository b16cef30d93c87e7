// Two-index symmetric irrep {2} (SU(3) sextet d=6, SU(2) spin-1 d=3) through
// the direct U Phi U^T form.
#include "launch.hpp"
namespace ghx {
bool ops_sym2(int prec, int N, int D, Ops* o) {
  if (N == 2) return pick<2, RepSym2>(prec, D, o);
  if (N == 3) return pick<3, RepSym2>(prec, D, o);
  return false;
}
}  // namespace ghx
