// Wall-normal kernels of the layouts L = 64 lanes of K = 8: one of the
// translation units of kernels_y.hip (the layouts compile in parallel).
#define CF_Y_PART 6
#include "kernels_y.hip"
