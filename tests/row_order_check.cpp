// Host-side check of the fused row pass's chunked block order (chunkedItem, kfixed.h), built by
// tests/test_row_order.py with hipcc for the host only. For every block count from 8 to 4200 and
// every group of the launch (8 * (ceil(nb/8) + 1) groups): each block is stored once (by a
// normal group or a redo group) and published once (by a normal group or a first visit), each
// wait is on the publisher of the predecessor block (nb - 1 with the wrap-around for block 0's
// redo), that publisher is a lower-index group on the same XCD (same index mod 8), and a redo's
// own first visit is one of the first eight groups, which never wait. Prints "OK" or the first
// failures.
#include <cstdio>
#include <vector>

#include "kfixed.h"

int main() {
  int bad = 0;
  auto fail = [&](int nb, int i, const char* what) {
    if (++bad <= 10) { std::printf("nb=%d group=%d: %s\n", nb, i, what); }
  };
  for (int nb = 8; nb <= 4200; ++nb) {
    const int groups = 8 * ((nb + 7) / 8 + 1);
    std::vector<int> stored(nb, 0), publisher(nb, -1), firstVisit(nb, -1);
    std::vector<FusedItem> items(groups);
    for (int i = 0; i < groups; ++i) {
      const FusedItem it = chunkedItem(nb, i);
      items[i] = it;
      if (it.none) { continue; }
      if (it.lb < 0 || it.lb >= nb) {
        fail(nb, i, "block out of range");
        continue;
      }
      if (!it.redo) {
        if (publisher[it.lb] >= 0) { fail(nb, i, "block published twice"); }
        publisher[it.lb] = i;
      }
      if (it.first) {
        firstVisit[it.lb] = i;
      } else {
        ++stored[it.lb];
      }
    }
    for (int b = 0; b < nb; ++b) {
      if (stored[b] != 1) { fail(nb, b, "block not stored exactly once"); }
      if (publisher[b] < 0) { fail(nb, b, "block never published"); }
    }
    for (int i = 0; i < groups; ++i) {
      const FusedItem& it = items[i];
      if (it.none || it.first) { continue; }
      if (it.pred < 0 || it.pred >= nb) {
        fail(nb, i, "predecessor out of range");
        continue;
      }
      const bool wrap = it.redo && it.lb == 0;
      if (it.wrap != wrap || it.pred != (wrap ? nb - 1 : it.lb - 1)) { fail(nb, i, "wrong predecessor"); }
      const int p = publisher[it.pred];
      if (p < 0 || p >= i || p % 8 != i % 8) { fail(nb, i, "predecessor not an earlier group of the same XCD"); }
      if (it.redo && (firstVisit[it.lb] < 0 || firstVisit[it.lb] >= 8)) { fail(nb, i, "redo without an early first visit"); }
    }
  }
  if (bad) {
    std::printf("FAIL %d\n", bad);
    return 1;
  }
  std::printf("OK\n");
  return 0;
}
