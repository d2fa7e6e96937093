// One-region column passes compiled per H (H >= 2048): k_coly<H, C1, C2>, T = H/C2 threads
// (forward plan [16, C1, C2]). The host picks "coly<H>" when it exists.
#include "kfixed.h"

extern "C" const OwlKernelEntry owl_kernel_table_coly_b[] = {
    {"coly2048", (const void*)&k_coly<2048, 16, 8>, 256},
    {"coly2304", (const void*)&k_coly<2304, 16, 9>, 256},
    {"coly2560", (const void*)&k_coly<2560, 16, 10>, 256},
    {"coly2816", (const void*)&k_coly<2816, 16, 11>, 256},
    {"coly3072", (const void*)&k_coly<3072, 16, 12>, 256},
    {"coly3328", (const void*)&k_coly<3328, 16, 13>, 256},
    {"coly3584", (const void*)&k_coly<3584, 16, 14>, 256},
    {"coly3840", (const void*)&k_coly<3840, 16, 15>, 256},
    {"coly4096", (const void*)&k_coly<4096, 16, 16>, 256},
    {nullptr, nullptr, 0},
};
