// One-region column passes compiled per H (H < 2048): k_coly<H, C1, C2>, T = H/C2 threads
// (forward plan [16, C1, C2]). The host picks "coly<H>" when it exists.
#include "kfixed.h"

extern "C" const OwlKernelEntry owl_kernel_table_coly_a[] = {
    {"coly512", (const void*)&k_coly<512, 4, 8>, 64},
    {"coly768", (const void*)&k_coly<768, 4, 12>, 64},
    {"coly1024", (const void*)&k_coly<1024, 8, 8>, 128},
    {"coly1152", (const void*)&k_coly<1152, 8, 9>, 128},
    {"coly1280", (const void*)&k_coly<1280, 8, 10>, 128},
    {"coly1408", (const void*)&k_coly<1408, 8, 11>, 128},
    {"coly1536", (const void*)&k_coly<1536, 8, 12>, 128},
    {"coly1664", (const void*)&k_coly<1664, 8, 13>, 128},
    {"coly1792", (const void*)&k_coly<1792, 8, 14>, 128},
    {"coly1920", (const void*)&k_coly<1920, 8, 15>, 128},
    {nullptr, nullptr, 0},
};
