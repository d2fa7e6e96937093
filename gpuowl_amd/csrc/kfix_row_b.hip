// Row passes compiled per (W, R): k_rowx<W, R, F0, F1, F2>, T = W/F0 threads.
// The host picks "rowx<W>r<R>" when it exists; jobs that share launches (batches, the server)
// take "rowx<W>r<R>b" where one is compiled: fewer threads per group, measured faster when a
// launch's groups already fill the device, slower for one job alone.
#include "kfixed.h"

extern "C" const OwlKernelEntry owl_kernel_table_row_b[] = {
    {"rowx512r4", (const void*)&k_rowx<512, 4, 2, 16, 16>, 256},
    {"rowx512r4b", (const void*)&k_rowx<512, 4, 4, 8, 16>, 128},
    {"rowx1024r4", (const void*)&k_rowx<1024, 4, 4, 16, 16>, 256},
    {nullptr, nullptr, 0},
};
