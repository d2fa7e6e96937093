// Two-region column passes compiled per H (H < 2048): k_colx<H, T, 16, C1, C2> (forward plan
// [16, C1, C2]). The host picks "colx<H>" when it exists and (H/16) % R == 0.
#include "kfixed.h"

extern "C" const OwlKernelEntry owl_kernel_table_colx_a[] = {
    {"colx512", (const void*)&k_colx<512, 64, 16, 16, 2>, 64},
    {"colx768", (const void*)&k_colx<768, 128, 16, 16, 3>, 128},
    {"colx1024", (const void*)&k_colx<1024, 128, 16, 16, 4>, 128},
    {"colx1152", (const void*)&k_colx<1152, 256, 16, 8, 9>, 256},
    {"colx1280", (const void*)&k_colx<1280, 256, 16, 16, 5>, 256},
    {"colx1408", (const void*)&k_colx<1408, 256, 16, 8, 11>, 256},
    {"colx1536", (const void*)&k_colx<1536, 256, 16, 16, 6>, 256},
    {"colx1664", (const void*)&k_colx<1664, 256, 16, 8, 13>, 256},
    {"colx1792", (const void*)&k_colx<1792, 256, 16, 16, 7>, 256},
    {"colx1920", (const void*)&k_colx<1920, 256, 16, 8, 15>, 256},
    {nullptr, nullptr, 0},
};
