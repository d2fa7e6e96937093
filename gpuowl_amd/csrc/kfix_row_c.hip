// Row passes compiled per (W, R): k_rowx<W, R, F0, F1, F2>, T = W/F0 threads.
// The host picks "rowx<W>r<R>" when it exists.
#include "kfixed.h"

extern "C" const OwlKernelEntry owl_kernel_table_row_c[] = {
    {"rowx256r4", (const void*)&k_rowx<256, 4, 4, 4, 16>, 64},
    {"rowx256r8", (const void*)&k_rowx<256, 8, 2, 8, 16>, 128},
    {nullptr, nullptr, 0},
};
