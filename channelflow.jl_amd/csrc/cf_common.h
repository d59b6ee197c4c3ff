// Host/device shared definitions of the MI355X CNRK2 library.
#pragma once
#include <hip/hip_runtime.h>

namespace cf {

constexpr int MAX_FACTORS = 16;

// per-degree constant table of the wall-normal solves (see build_cheb_table)
constexpr int CHEB_TAB_ARRAYS = 4;
constexpr int CHEB_TAB_MAX = 512;  // entries per array: K L <= 512

// FFT of length n factored into radices; tw[t] = exp(-2 pi i t/n) (device).
// Per pass f (radix R, span Ns = product of earlier radices): reciprocal
// multipliers q(x) = (x * mul) >> 32 with mul = floor(2^32/d) + 1 for the
// divisors d = n/R (mul_m) and d = Ns (mul_ns); exact for x*d < 2^32.
struct FFTPlan {
    int n;
    int nf;
    int radix[MAX_FACTORS];
    unsigned long long mul_m[MAX_FACTORS], mul_ns[MAX_FACTORS];
    const double2* tw;
};

__host__ __device__ inline unsigned long long recip_mul(unsigned d) { return (1ull << 32) / d + 1ull; }
__host__ __device__ __forceinline__ unsigned fast_div(unsigned x, unsigned long long mul) {
    return (unsigned)(((unsigned long long)x * mul) >> 32);
}

// Grid sizes (src/grids.jl:22-74): resolved (Nx, Ny, Nz), padded Nxp, Nzp,
// half spectrum Nxh; P = Ny - 1; B = Nxh*Nz modes; S = B*Ny coefficients.
struct GridDims {
    int Nx, Ny, Nz, Nxh, Nxp, Nzp, P;
    long long B, S, Sz;  // Sz = Nxh*Nzp*Ny (z-padded x-half spectrum)
    double alpha, beta;  // 2 pi / Lx, 2 pi / Lz
};

// Per-mode constants of one Stokes solver bank (one CNRK2 stage), used by
// the influence/tau correction in re-solve form.  Index with MC_*.
enum {
    MC_A11 = 0, MC_A12, MC_A21, MC_A22,  // inverse 2x2 influence matrix (row-major)
    MC_SIG1, MC_SIG2,                    // auxiliary tau residuals sigma0
    MC_D0P, MC_D0M,                      // wall amplitudes of the auxiliary pressure p0
    MC_VP1, MC_VP0, MC_VM1, MC_VM0,      // v+(P-1), v+(P), v-(P-1), v-(P)
    MC_PPT, MC_PMT,                      // p+(P), p-(P)
    MC_COUNT = 16
};

}  // namespace cf
