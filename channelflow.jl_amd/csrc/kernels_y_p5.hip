// Wall-normal kernels of the layouts L = 32 lanes of K = 8 (step kernels): one of the
// translation units of kernels_y.hip (the layouts compile in parallel).
#define CF_Y_PART 5
#include "kernels_y.hip"
