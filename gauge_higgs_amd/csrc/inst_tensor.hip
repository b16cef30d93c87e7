// Small tensor-irrep instantiations (rep_general.hpp matrix-free path):
// SU(3) {2} (sextet, d=6) and {1,1} (d=3); SU(2) {2} (d=3) and {3} (d=4).
#include "launch.hpp"
namespace ghx {
template <typename R, int N> using RepT2x6 = RepTensor<R, N, 2, 6>;
template <typename R, int N> using RepT2x3 = RepTensor<R, N, 2, 3>;
template <typename R, int N> using RepT3x4 = RepTensor<R, N, 3, 4>;
bool ops_tensor(int prec, int N, int D, int nbox, int d, Ops* o) {
  if (N == 3 && nbox == 2 && d == 6) return pick<3, RepT2x6>(prec, D, o);
  if (N == 3 && nbox == 2 && d == 3) return pick<3, RepT2x3>(prec, D, o);
  if (N == 2 && nbox == 2 && d == 3) return pick<2, RepT2x3>(prec, D, o);
  if (N == 2 && nbox == 3 && d == 4) return pick<2, RepT3x4>(prec, D, o);
  return false;
}
}  // namespace ghx
