// Wall-normal kernels of the layout L = 64 lanes of K = 6 coefficients: one of
// the translation units of kernels_y.hip (the layouts compile in parallel).
#define CF_Y_PART 7
#include "kernels_y.hip"
