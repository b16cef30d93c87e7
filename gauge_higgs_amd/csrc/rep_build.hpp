// Host-side construction of the scalar representation: generators T^a_R and,
// for tensor irreps, the orthonormal real basis W of the Young-symmetrized
// subspace of V^{(x)n}. The basis must be the reference's exactly (the scalar
// field lives in it), so the enumeration follows rep_general.hpp:338-411:
// single-row diagrams use normalised sorted multisets in odometer order,
// general diagrams symmetrize every tuple 0..N^n-1 in index order with
// c = b a (rows then signed columns) and keep the modified-Gram-Schmidt
// survivors above 1e-9.
#pragma once
#include <complex>
#include <stdexcept>
#include <string>
#include <vector>

namespace ghx {

struct RepInfo {
  int kind = 0;  // RepKind
  int N = 0, d = 0, nality = 0, nbox = 0;
  bool real = false;
  long dimT = 0;
  std::vector<int> rows;
  std::vector<double> W;                 // d x dimT, row-major (tensor kind)
  std::vector<std::complex<double>> T;   // (N^2-1) x d x d
  std::string name;
};

// Parse "", "none", "fund", "adj", "r1,r2,...[:real]". Throws std::runtime_error
// with the reference's messages for invalid diagrams.
RepInfo build_rep(int N, const std::string& spec);

// Young rows of a "r1,r2,...[:real]" spec, validated with the reference's
// messages (rep_general.hpp:58-69, 338-345); *real receives the suffix flag.
std::vector<int> parse_young_rows(int N, std::string spec, bool* real);
// Dimension of the SU(N) irrep with these rows (hook-content formula).
long young_dimension(int N, const std::vector<int>& rows);

// Fundamental generators in the reference order (algebra.hpp:24-50), N x N each.
std::vector<std::complex<double>> fundamental_generators(int N);

}  // namespace ghx
