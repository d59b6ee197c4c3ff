// Gaussian noise on the padded physical grid for random_state
// (src/initialization.jl:32-51): epsilon * N(0, 1) samples for u, v, w.
//
// Stream: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC 2011), key = (seed bits 0-31, seed bits
// 32-63).  Counter words, for the pair q = element index / 2 of component c:
//   (q bits 0-31, q bits 32-63, c, 0).
// One counter block gives 128 bits: words 0,1 -> u1 = ((w0 << 21) ^ (w1 >> 11)
// masked to 53 bits, + 1) / 2^53 in (0, 1], words 2,3 -> u2 the same way;
// Box-Muller turns them into the normals of elements 2q (cos) and 2q + 1
// (sin), each times epsilon.  Element index = x + Nxp (z + Nzp y) of a
// component, the reference's evaluation order (physicalfield.jl:39-43).
#include "cf_device.h"
#include "cf_kernels.h"

namespace cf {

struct Philox4 {
    unsigned c[4];
};

__device__ __forceinline__ Philox4 philox4x32_10(Philox4 ctr, unsigned k0, unsigned k1) {
    constexpr unsigned M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const unsigned hi0 = __umulhi(M0, ctr.c[0]), lo0 = M0 * ctr.c[0];
        const unsigned hi1 = __umulhi(M1, ctr.c[2]), lo1 = M1 * ctr.c[2];
        Philox4 n;
        n.c[0] = hi1 ^ ctr.c[1] ^ k0;
        n.c[1] = lo1;
        n.c[2] = hi0 ^ ctr.c[3] ^ k1;
        n.c[3] = lo0;
        ctr = n;
        k0 += W0;
        k1 += W1;
    }
    return ctr;
}

// uniform in (0, 1] from 53 random bits
__device__ __forceinline__ double u01(unsigned hi, unsigned lo) {
    const unsigned long long b = ((unsigned long long)hi << 21) ^ (unsigned long long)(lo >> 11);
    return ((double)(b & ((1ull << 53) - 1)) + 1.0) * (1.0 / 9007199254740992.0);
}

__global__ void __launch_bounds__(256) gaussian_kernel(double* out, long long n, double eps, unsigned k0, unsigned k1) {
    // one thread: two consecutive elements of one component
    const long long pair = (long long)blockIdx.x * 256 + threadIdx.x;
    const int comp = blockIdx.y;
    const long long i = 2 * pair;
    if (i >= n) return;
    Philox4 ctr{{(unsigned)(pair & 0xFFFFFFFFu), (unsigned)(pair >> 32), (unsigned)comp, 0u}};
    const Philox4 r = philox4x32_10(ctr, k0, k1);
    const double u1 = u01(r.c[0], r.c[1]), u2 = u01(r.c[2], r.c[3]);
    const double rad = sqrt(-2.0 * log(u1));
    double s, c;
    sincospi(2.0 * u2, &s, &c);
    double* o = out + (size_t)comp * n;
    o[i] = eps * rad * c;
    if (i + 1 < n) o[i + 1] = eps * rad * s;
}

hipError_t launch_gaussian(double* out3, long long n, double eps, unsigned long long seed, hipStream_t st) {
    const long long pairs = (n + 1) / 2;
    dim3 grid((unsigned)((pairs + 255) / 256), 3);
    hipLaunchKernelGGL(gaussian_kernel, grid, dim3(256), 0, st, out3, n, eps, (unsigned)(seed & 0xFFFFFFFFull),
                       (unsigned)(seed >> 32));
    return hipGetLastError();
}

}  // namespace cf
