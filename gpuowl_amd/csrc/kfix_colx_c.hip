// Two-region column passes with several column pairs per group (k_colx<H, T, 16, C1, C2, true>),
// compiled for the H whose k_colx needs more than half of a CU's LDS (H >= 2304), so that one
// group runs per CU. The host runs "colx<H>m" in place of "colx<H>" when a launch has enough
// groups to give every CU one with fewer groups per job (see Gpu::colPairsPerGroup).
#include "kfixed.h"

extern "C" const OwlKernelEntry owl_kernel_table_colx_c[] = {
    {"colx2304m", (const void*)&k_colx<2304, 512, 16, 16, 9, true>, 512},
    {"colx2560m", (const void*)&k_colx<2560, 512, 16, 16, 10, true>, 512},
    {"colx2816m", (const void*)&k_colx<2816, 512, 16, 16, 11, true>, 512},
    {"colx3072m", (const void*)&k_colx<3072, 512, 16, 16, 12, true>, 512},
    {"colx3328m", (const void*)&k_colx<3328, 512, 16, 16, 13, true>, 512},
    {"colx3584m", (const void*)&k_colx<3584, 512, 16, 16, 14, true>, 512},
    {"colx3840m", (const void*)&k_colx<3840, 512, 16, 16, 15, true>, 512},
    {"colx4096m", (const void*)&k_colx<4096, 512, 16, 16, 16, true>, 512},
    {nullptr, nullptr, 0},
};
