// ORACLE -- test infrastructure only. Driver for oracle/_ref/refcheck, a program
// linked from the reference's own src/state.cpp (compiled where it lies under the
// reference tree by oracle/Makefile, target "ref"). It lets tests compare this
// project's word packing with the reference's, on the same inputs:
//   refcheck expand E N          <- nWords hex words on stdin  -> N digits (expandBits)
//   refcheck compact E N         <- N digits on stdin          -> hex words (compactBits)
#include "common.h"
#include "state.h"

#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

int main(int argc, char** argv) {
  if (argc < 2) { return 2; }
  std::string cmd = argv[1];
  if ((cmd == "expand" || cmd == "compact") && argc == 4) {
    u32 E = u32(strtoul(argv[2], nullptr, 10));
    u32 N = u32(strtoul(argv[3], nullptr, 10));
    if (cmd == "expand") {
      std::vector<u32> words((E - 1) / 32 + 1);
      for (u32& w : words) {
        if (scanf("%x", &w) != 1) { return 3; }
      }
      for (int d : expandBits(words, N, E)) { printf("%d\n", d); }
    } else {
      std::vector<int> digits(N);
      for (int& d : digits) {
        if (scanf("%d", &d) != 1) { return 3; }
      }
      for (u32 w : compactBits(digits, E)) { printf("%08x\n", w); }
    }
    return 0;
  }
  return 2;
}
