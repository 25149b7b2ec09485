// Driver that compiles TSEMLib's own header-only multi-index numbering
// (include/Multiindex.h, include/Correspondence.h) and prints, for a given
// truncation order M, every (number, multi-index) pair and index2number of
// each. Used to pin the oracle's Correspondence bit-exactly.
#include <iostream>
#include <vector>
#include <cstdlib>
#include "Correspondence.h"

int main(int argc, char** argv) {
  int M = argc > 1 ? std::atoi(argv[1]) : 8;
  Correspondence<3> c(M);
  std::cout << M << ' ' << c.n_index() << '\n';
  for (int i = 0; i < c.n_index(); ++i) {
    Multiindex<3> a = c.number2index(i);
    std::cout << i << ' ' << a.index[0] << ' ' << a.index[1] << ' ' << a.index[2] << ' ' << c.index2number(a) << '\n';
  }
  for (int o = 0; o <= M; ++o) std::cout << c.n_index_begin(o) << ' ' << c.n_index_end(o) << '\n';
  return 0;
}
