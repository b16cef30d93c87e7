// Symmetric-power engine for a large tensor irrep R = Young rows (k1[, k2]).
//
// The irrep spanned by the reference's basis W (rep_general.hpp:338-411) sits,
// up to a unitary intertwiner, inside S = Sym^k1(V) (x) Sym^k2(V): the slots of
// row 1 symmetrised, then those of row 2. With the orthonormal monomial bases
// |m> = sqrt(m!/k!) sum_{tuples of type m} |t> of Sym^k, the map
// Yc = B^T W^T : C^d -> S is an isometry up to a scale c (Schur), and
//   D^(R)(U) = (1/c) Yc^T (Sym^k1(U) (x) Sym^k2(U)) Yc.
// Sym^k(U) in that basis is diag(sqrt(m!)) C_k diag(1/sqrt(n!)), where column n
// of C_k holds the coefficients of prod_j (sum_i U_ij x_i)^(n_j); C_k follows
// from C_(k-1) one linear factor at a time:
//   C_k[m][n] = sum_i U[i][j] C_(k-1)[m - e_i][n - e_j],  j = first index with n_j > 0.
// A tensor apply costs n one-slot passes over N^n entries; this costs
// sum_k s_k^2 N + s1 s2 (s1 + s2) per apply (s_k = dim Sym^k), about 12x fewer
// operations for SU(3) {6,1} (d = 48, N^n = 2187, dim S = 84).
#pragma once
#include <vector>

#include "rep_build.hpp"

namespace ghx {

constexpr int kSymMaxS = 36;      // largest dim Sym^k of the engine (SU(3): k <= 7)
constexpr int kSymMaxDimS = 256;  // largest dim S
constexpr int kSymMaxD = 64;      // largest irrep: the per-site kernel holds one component per lane of a wave

struct SymBuild {
  bool ok = false;
  int N = 0, k1 = 0, k2 = 0, s1 = 0, s2 = 1, dimS = 0, d = 0;
  double invc = 0.0;
  std::vector<double> Yc;       // dimS x d, row-major (real: W and the monomial bases are real)
  std::vector<int> lvl_off;     // offset of level k in the per-level tables, k = 0..k1+1
  std::vector<int> colj, colsrc;  // per (level k, column n): j and the index of n - e_j at level k-1
  std::vector<int> rowsrc;        // per (level k, row m) x N: index of m - e_i at level k-1, or -1
  std::vector<double> fnorm;      // per (level k, m): sqrt(m!)
  double max_err = 0.0;           // host check against the tensor apply on generic links
};

// Builds the engine for a tensor irrep and checks D^(R)(U) against the tensor
// apply W U^(x n) W^T on three generic links; ok = false when the irrep has
// more than two rows, exceeds the size limits or fails the check (1e-10).
SymBuild build_sym(const RepInfo& r);

// The device form. The unnormalised C_k are built in LDS by a program of
// passes, 64 entries per pass (one per lane), each entry
//   L[dst] = Us[0][j] L[s0] + Us[1][j] L[s1] + Us[2][j] L[s2]
// over LDS complex slots: slot 0 holds zero (a missing source), slot 1 takes
// the writes of idle lanes, Us (N rows x 4 columns from slot kSymLdsUs) holds
// the link with a fourth column e_0 (j = 3: a plain copy, used to save C_k2).
// The normalisations diag(sqrt(m!)) are moved out of the build:
//   S1 (x) S2 = F (C1 (x) C2) F^-1,  F = diag(sqrt(a!) sqrt(b!)) over q = (a, b),
// so the forward hop lifts V = F^-1 Yc phi and projects with Yc F / c, the
// adjoint lifts F Yc phi and projects with Yc F^-1 / c.
// SU(3) irreps with k1 <= kSymPairMaxK (dim Sym^k1 <= 28) may run two links
// per wave, C_k built a column per lane (sym_link_pair_kernel).
constexpr int kSymPairMaxK = 6;
// The apply keeps a row of C1 V in registers when dim Sym^k2 <= kSymRowS2.
constexpr int kSymRowS2 = 4;
constexpr int kSymLdsUs = 4;
constexpr int kSymLdsBase = 16;
struct SymProgram {
  std::vector<int> prog;           // 4 ints per entry: s0, s1, s2, (dst << 2) | j
  int npass = 0;                   // even (padded with an idle pass)
  int lds_one = 0;                 // level 0's entry (1)
  int lds_S1 = 0, lds_S2 = 0, lds_V = 0, lds_T1 = 0, lds_W = 0, lds_total = 0;
  int lds_Cb = 0;                  // the second level buffer: level k lives at lds_S1 when k1 - k is even, else here
  // layout of one link's region in the paired kernel (cx units): every level
  // in place in C (the last with row stride s1p), the saved C_k2, then V, which
  // the row-blocked apply overwrites with its result (T1 = V); the two-stage
  // apply keeps T1 apart
  int pair_C = 0, pair_S2 = 0, pair_V = 0, pair_T1 = 0, pair_total = 0;
  // row strides of C1 and C2 in LDS: C_k1 is written with an odd row stride
  // (s1 + 1 when s1 is even), so that lanes reading one column of its rows hit
  // different LDS banks; C2 shares it when k2 == k1 and keeps s2 when copied
  int s1p = 0, s2p = 0;
  std::vector<double> pre;         // [2][dimS]: 1/F (forward), F (adjoint)
  std::vector<double> Yp;          // [2][ng][d][4], q = 4 g + i: Yc F / c (forward), Yc F^-1 / c (adjoint); zero for q >= dimS
  int ng = 0;                      // groups of 4 components (dimS rounded up)
  // The projections as sparse rows. Every basis vector of the irrep is a weight
  // vector (the Young symmetrizer keeps a tuple's colour content, and vectors of
  // different content are orthogonal, so Gram-Schmidt never mixes them), and a
  // monomial pair q = (a, b) has the content a + b: row l of Yc is nonzero only
  // on the q that split the content of l, at most dim Sym^k2 of them ({6,1}:
  // 111 nonzeros of 48 x 84). Entry k of row l of direction dir is
  // (ysq, ysv)[(dir ynz + k) d + l], rows padded with (0, 0.0) to ynz entries.
  std::vector<int> ysq;
  std::vector<double> ysv;
  int ynz = 0;
  bool ok = false;                 // the program, run on the host, reproduces D(U) and D(U)^dag (1e-10)
  double max_err = 0.0;
};
SymProgram sym_program(const SymBuild& s);  // ok = false for N > 3

}  // namespace ghx
