// Row passes compiled per (W, R): k_rowx<W, R, F0, F1, F2>, T = W/F0 threads.
// The host picks "rowx<W>r<R>" when it exists.
#include "kfixed.h"

extern "C" const OwlKernelEntry owl_kernel_table_row_b[] = {
    {"rowx512r4", (const void*)&k_rowx<512, 4, 4, 8, 16>, 128},
    {"rowx1024r4", (const void*)&k_rowx<1024, 4, 4, 16, 16>, 256},
    {nullptr, nullptr, 0},
};
