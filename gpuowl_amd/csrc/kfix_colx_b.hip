// Two-region column passes compiled per H (H >= 2048): k_colx<H, T, 16, C1, C2> (forward plan
// [16, C1, C2]). The host picks "colx<H>" when it exists and (H/16) % R == 0.
#include "kfixed.h"

extern "C" const OwlKernelEntry owl_kernel_table_colx_b[] = {
    {"colx2048", (const void*)&k_colx<2048, 256, 16, 16, 8>, 256},
    {"colx2304", (const void*)&k_colx<2304, 512, 16, 16, 9>, 512},
    {"colx2560", (const void*)&k_colx<2560, 512, 16, 16, 10>, 512},
    {"colx2816", (const void*)&k_colx<2816, 512, 16, 16, 11>, 512},
    {"colx3072", (const void*)&k_colx<3072, 512, 16, 16, 12>, 512},
    {"colx3328", (const void*)&k_colx<3328, 512, 16, 16, 13>, 512},
    {"colx3584", (const void*)&k_colx<3584, 512, 16, 16, 14>, 512},
    {"colx3840", (const void*)&k_colx<3840, 512, 16, 16, 15>, 512},
    {"colx4096", (const void*)&k_colx<4096, 512, 16, 16, 16>, 512},
    {nullptr, nullptr, 0},
};
