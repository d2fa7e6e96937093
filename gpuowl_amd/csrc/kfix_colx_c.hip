// Two-region column passes with two column pairs per group, compiled for the H whose
// k_colx needs more than half of a CU's LDS (H >= 2304), so that one group runs per CU:
// k_colx<H, T, 16, C1, C2, 2>. The host picks "colx<H>p2" when it exists, (H/16) % R == 0,
// the one-pair flag is not set and W/4 groups per job still give every CU a group.
#include "kfixed.h"

extern "C" const OwlKernelEntry owl_kernel_table_colx_c[] = {
    {"colx2304p2", (const void*)&k_colx<2304, 512, 16, 16, 9, 2>, 512},
    {"colx2560p2", (const void*)&k_colx<2560, 512, 16, 16, 10, 2>, 512},
    {"colx2816p2", (const void*)&k_colx<2816, 512, 16, 16, 11, 2>, 512},
    {"colx3072p2", (const void*)&k_colx<3072, 512, 16, 16, 12, 2>, 512},
    {"colx3328p2", (const void*)&k_colx<3328, 512, 16, 16, 13, 2>, 512},
    {"colx3584p2", (const void*)&k_colx<3584, 512, 16, 16, 14, 2>, 512},
    {"colx3840p2", (const void*)&k_colx<3840, 512, 16, 16, 15, 2>, 512},
    {"colx4096p2", (const void*)&k_colx<4096, 512, 16, 16, 16, 2>, 512},
    {nullptr, nullptr, 0},
};
