// Host check of the symmetric-power engine (gauge_higgs_amd/csrc/symrep.cpp):
// for each irrep given as "N:rows", builds the engine (which checks
// Yc^T (Sym^k1 U (x) Sym^k2 U) Yc / c against the tensor apply W U^(x n) W^T
// on generic links) and its device program (run here on the host and checked
// against the engine's matrices for D(U) and D(U)^dag). Prints one line per
// irrep: spec ok_build err_build ok_program err_program npass lds_total d
// ynz s2 nnz (nonzeros per projection row at most, dim Sym^k2, nonzeros of
// the forward projection).
#include <cstdio>
#include <cstdlib>
#include <string>

#include "symrep.hpp"

int main(int argc, char** argv) {
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    const size_t c = a.find(':');
    const int N = std::atoi(a.substr(0, c).c_str());
    const ghx::RepInfo r = ghx::build_rep(N, a.substr(c + 1));
    const ghx::SymBuild s = ghx::build_sym(r);
    const ghx::SymProgram p = s.ok ? ghx::sym_program(s) : ghx::SymProgram{};
    int nnz = 0;
    for (int e = 0; e < p.ynz * r.d; ++e) nnz += p.ysv.empty() ? 0 : p.ysv[size_t(e)] != 0.0;
    std::printf("%s %d %.3e %d %.3e %d %d %d %d %d %d\n", a.c_str(), int(s.ok), s.max_err, int(p.ok), p.max_err, p.npass, p.lds_total, r.d,
                p.ynz, s.s2, nnz);
    std::fflush(stdout);
  }
  return 0;
}
