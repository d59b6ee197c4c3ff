// Wall-normal kernels of the layouts L = 16 lanes of K = 4, and of K = 8 (step kernels): one of the
// translation units of kernels_y.hip (the layouts compile in parallel).
#define CF_Y_PART 2
#include "kernels_y.hip"
