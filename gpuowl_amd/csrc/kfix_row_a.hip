// Row passes compiled per (W, R): k_rowx<W, R, F0, F1, F2>, T = W/F0 threads.
// The host picks "rowx<W>r<R>" when it exists.
#include "kfixed.h"

extern "C" const OwlKernelEntry owl_kernel_table_row_a[] = {
    {"rowx256r2", (const void*)&k_rowx<256, 2, 2, 8, 16>, 128},
    {"rowx512r2", (const void*)&k_rowx<512, 2, 2, 16, 16>, 256},
    {"rowx1024r2", (const void*)&k_rowx<1024, 2, 4, 16, 16>, 256},
    {"rowx2048r1", (const void*)&k_rowx<2048, 1, 8, 16, 16>, 256},
    {"rowx2048r2", (const void*)&k_rowx<2048, 2, 8, 16, 16>, 256},
    {"rowx4096r1", (const void*)&k_rowx<4096, 1, 16, 16, 16>, 256},
    {nullptr, nullptr, 0},
};
