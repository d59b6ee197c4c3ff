// Wall-normal kernels of the layouts L = 4 and 8 lanes of K = 4 coefficients: one of the
// translation units of kernels_y.hip (the layouts compile in parallel).
#define CF_Y_PART 1
#include "kernels_y.hip"
