// Multi-invariant quartic potential tables (scalar_invariants.hpp:20-208).
#pragma once
#include <complex>
#include <vector>

#include "rep_build.hpp"

namespace ghx {

// Matrix of M -> sum_a [T_a, [T_a, M]] on vec(M), index i*d + j; d^2 x d^2.
std::vector<std::complex<double>> casimir_superop(const RepInfo& r);
// Distinct eigenvalues (channel Casimirs) in increasing order.
std::vector<double> casimir_channels(const RepInfo& r);
// A = sum_c 2 f_c P_c on vec(M); d^2 x d^2, row (p*d+q), column (i*d+j).
std::vector<std::complex<double>> multi_superop(const RepInfo& r, const std::vector<double>& lambda, const std::vector<double>& f);
// The same superoperator on the real space of Hermitian d x d matrices, in the
// orthonormal basis E_k (Tr E_k E_l = delta_kl): k < d the diagonal units,
// then for each pair i < j (row-major) S = (e_ij + e_ji)/sqrt2 followed by
// K = i(e_ij - e_ji)/sqrt2. Entry [k * d^2 + l] = Tr(E_k A(E_l)), real because
// A maps Hermitian matrices to Hermitian ones (it is a real polynomial in the
// Casimir superoperator M -> sum_a [T_a, [T_a, M]]).
std::vector<double> hermitian_superop(int d, const std::vector<std::complex<double>>& A);

}  // namespace ghx
