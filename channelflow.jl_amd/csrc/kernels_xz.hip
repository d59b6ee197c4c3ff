// Periodic-direction pencil kernels.
//
// z pass: for every (k, y) line, the l spectrum is embedded in / truncated
// from the 3/2-padded length Nzp (negative-l block at the end, Nyquist
// removed; ffts.jl:242-293) and transformed by a Stockham FFT in LDS.  A
// block owns KT consecutive k of one y plane, so every HBM row it touches is
// KT contiguous complex values.
//
// x pass: for every (zp, y) line the Nxh retained coefficients are padded to
// Nxp/2+1, inverse real-transformed (brfft), the rotational product u x w is
// formed on the padded points (nonlinear.jl:187-189) and forward
// real-transformed (rfft), keeping k < Nxh.  Real transforms of even length
// use a half-length complex FFT with the standard pack/unpack.
#include "cf_device.h"
#include "cf_kernels.h"

namespace cf {

constexpr int ZNT = 256;
constexpr int XNT = 256;

// z inverse (dir=+1): (k,l,y) -> (k,zp,y); z forward (dir=-1): (k,zp,y) -> (k,l,y)
template <int DIR>
__global__ void __launch_bounds__(ZNT) zpass_kernel(ZArgs a) {
    extern __shared__ double2 smem[];
    const GridDims& gd = a.gd;
    const int KT = a.KT, ld = a.ld;
    double2* bufA = smem;
    double2* bufB = smem + KT * ld;
    const int y = blockIdx.y, f = blockIdx.z;
    const int k0 = blockIdx.x * KT;
    const int tid = threadIdx.x;
    const int Nz = gd.Nz, Nzp = gd.Nzp, Nxh = gd.Nxh;
    const int npos = (Nz >> 1) + 1, nneg = Nz - npos;
    const bool xeven = (gd.Nx % 2 == 0), zeven = (Nz % 2 == 0);
    if (DIR > 0) {
        const double2* in = a.in[f] + (size_t)Nxh * Nz * y;
        for (int idx = tid; idx < KT * Nzp; idx += ZNT) {
            int z = idx / KT, kk = idx - z * KT;
            int k = k0 + kk;
            double2 v = czero();
            if (k < Nxh && !(xeven && k == gd.Nx / 2)) {
                int l = -1;
                if (z < npos) l = (zeven && z == Nz / 2) ? -1 : z;
                else if (z >= Nzp - nneg) l = npos + (z - (Nzp - nneg));
                if (l >= 0) v = in[k + (size_t)Nxh * l];
            }
            bufA[kk * ld + z] = v;
        }
    } else {
        const double2* in = a.in[f] + (size_t)Nxh * Nzp * y;
        for (int idx = tid; idx < KT * Nzp; idx += ZNT) {
            int z = idx / KT, kk = idx - z * KT;
            int k = k0 + kk;
            bufA[kk * ld + z] = (k < Nxh) ? in[k + (size_t)Nxh * z] : czero();
        }
    }
    __syncthreads();
    double2* res = lds_fft(bufA, bufB, a.plan, KT, ld, DIR, tid, ZNT);
    if (DIR > 0) {
        double2* out = a.out[f] + (size_t)Nxh * Nzp * y;
        for (int idx = tid; idx < KT * Nzp; idx += ZNT) {
            int z = idx / KT, kk = idx - z * KT;
            int k = k0 + kk;
            if (k < Nxh) out[k + (size_t)Nxh * z] = res[kk * ld + z];
        }
    } else {
        double2* out = a.out[f] + (size_t)Nxh * Nz * y;
        for (int idx = tid; idx < KT * Nz; idx += ZNT) {
            int l = idx / KT, kk = idx - l * KT;
            int k = k0 + kk;
            int z = (l < npos) ? l : Nzp - nneg + (l - npos);
            if (k < Nxh) out[k + (size_t)Nxh * l] = res[kk * ld + z];
        }
    }
}

// ---------------------------------------------------------------------------
// x pass helpers.  Even Nxp = 2M: complex FFT of length M; odd: length Nxp.
// ---------------------------------------------------------------------------
struct XCtx {
    const XArgs& a;
    int Nxp, Nxh, M, LPB, ld, tid;
    bool even;
    long long gl0, nlines;
    double2 *bufA, *bufB;
    __device__ __forceinline__ XCtx(const XArgs& a_, double2* s) : a(a_) {
        Nxp = a.gd.Nxp;
        Nxh = a.gd.Nxh;
        even = (Nxp % 2 == 0);
        M = even ? Nxp / 2 : Nxp;
        LPB = a.LPB;
        ld = a.ld;
        tid = threadIdx.x;
        gl0 = (long long)blockIdx.x * LPB;
        nlines = (long long)a.gd.Nzp * a.gd.Ny;
        bufA = s;
        bufB = s + LPB * ld;
    }
    // inverse real transform of LPB lines of `in` (Nxh complex per line)
    // into phys[line * Nxp + x]
    __device__ __forceinline__ void c2r(const double2* __restrict__ in, double* phys) {
        const bool xeven = (a.gd.Nx % 2 == 0);
        // stage X_k (k = 0..M) in bufB (zero padding, Nyquist plane and Im(DC) removed)
        const int nX = even ? M + 1 : Nxp / 2 + 1;
        for (int idx = tid; idx < LPB * nX; idx += XNT) {
            int ln = idx / nX, k = idx - ln * nX;
            long long gl = gl0 + ln;
            double2 v = czero();
            if (gl < nlines && k < Nxh && !(xeven && k == a.gd.Nx / 2)) v = in[(size_t)gl * Nxh + k];
            if (k == 0) v.y = 0.0;
            if (even && k == M) v.y = 0.0;
            bufB[ln * ld + k] = v;
        }
        __syncthreads();
        if (even) {
            for (int idx = tid; idx < LPB * M; idx += XNT) {
                int ln = idx / M, k = idx - ln * M;
                double2 Xk = bufB[ln * ld + k];
                double2 Xc = cconj(bufB[ln * ld + (M - k)]);
                double2 w = cconj(a.twx[k]);  // exp(+2 pi i k / Nxp)
                bufA[ln * ld + k] = cadd(cadd(Xk, Xc), cmuli(cmul(w, csub(Xk, Xc))));
            }
        } else {
            for (int idx = tid; idx < LPB * M; idx += XNT) {
                int ln = idx / M, k = idx - ln * M;
                double2 v;
                if (k <= Nxp / 2) v = bufB[ln * ld + k];
                else v = cconj(bufB[ln * ld + (Nxp - k)]);
                bufA[ln * ld + k] = v;
            }
        }
        __syncthreads();
        double2* res = lds_fft(bufA, bufB, a.plan, LPB, ld, +1, tid, XNT);
        for (int idx = tid; idx < LPB * M; idx += XNT) {
            int ln = idx / M, j = idx - ln * M;
            double2 z = res[ln * ld + j];
            if (even) {
                phys[ln * Nxp + 2 * j] = z.x;
                phys[ln * Nxp + 2 * j + 1] = z.y;
            } else {
                phys[ln * Nxp + j] = z.x;
            }
        }
        __syncthreads();
    }
    // forward real transform of phys lines -> out (Nxh complex per line)
    __device__ __forceinline__ void r2c(const double* phys, double2* __restrict__ out) {
        for (int idx = tid; idx < LPB * M; idx += XNT) {
            int ln = idx / M, j = idx - ln * M;
            bufA[ln * ld + j] = even ? cmk(phys[ln * Nxp + 2 * j], phys[ln * Nxp + 2 * j + 1]) : cmk(phys[ln * Nxp + j], 0.0);
        }
        __syncthreads();
        double2* res = lds_fft(bufA, bufB, a.plan, LPB, ld, -1, tid, XNT);
        for (int idx = tid; idx < LPB * Nxh; idx += XNT) {
            int ln = idx / Nxh, k = idx - ln * Nxh;
            long long gl = gl0 + ln;
            if (gl >= nlines) continue;
            double2 X;
            if (even) {
                double2 Zk = res[ln * ld + (k % M)];
                double2 Zc = cconj(res[ln * ld + ((M - k) % M)]);
                double2 w = a.twx[k];  // exp(-2 pi i k / Nxp)
                X = csub(cscale(cadd(Zk, Zc), 0.5), cscale(cmuli(cmul(w, csub(Zk, Zc))), 0.5));
            } else {
                X = res[ln * ld + k];
            }
            out[(size_t)gl * Nxh + k] = X;
        }
        __syncthreads();
    }
};

// Rotational x pass for even Nxp: the six fields of LPB lines are one batch
// of 6*LPB length-M FFTs; the product is formed from the transformed batch
// in LDS and packed directly as the input of the 3*LPB forward FFTs.
__device__ __forceinline__ void xpass_rotational_even(const XArgs& a, double2* smem) {
    const int Nxp = a.gd.Nxp, Nxh = a.gd.Nxh, M = Nxp / 2, LPB = a.LPB, ld = a.ld, tid = threadIdx.x;
    const long long gl0 = (long long)blockIdx.x * LPB, nlines = (long long)a.gd.Nzp * a.gd.Ny;
    const bool xeven = (a.gd.Nx % 2 == 0);
    double2* bufA = smem;
    double2* bufB = smem + 6 * LPB * ld;
    const int nX = M + 1;
    // stage X_k, k = 0..M (padding, Nyquist plane and Im of DC/Nyquist removed)
    for (int idx = tid; idx < 6 * LPB * nX; idx += XNT) {
        int q = idx / nX, k = idx - q * nX;
        int f = q / LPB, ln = q - f * LPB;
        long long gl = gl0 + ln;
        double2 v = czero();
        if (gl < nlines && k < Nxh && !(xeven && k == a.gd.Nx / 2)) v = a.in[f][(size_t)gl * Nxh + k];
        if (k == 0 || k == M) v.y = 0.0;
        bufB[q * ld + k] = v;
    }
    __syncthreads();
    for (int idx = tid; idx < 6 * LPB * M; idx += XNT) {
        int q = idx / M, k = idx - q * M;
        double2 Xk = bufB[q * ld + k];
        double2 Xc = cconj(bufB[q * ld + (M - k)]);
        double2 w = cconj(a.twx[k]);  // exp(+2 pi i k / Nxp)
        bufA[q * ld + k] = cadd(cadd(Xk, Xc), cmuli(cmul(w, csub(Xk, Xc))));
    }
    __syncthreads();
    double2* res = lds_fft(bufA, bufB, a.plan, 6 * LPB, ld, +1, tid, XNT);
    double2* nxt = (res == bufA) ? bufB : bufA;
    // u x w at x = 2j (real parts) and x = 2j+1 (imaginary parts)
    for (int idx = tid; idx < LPB * M; idx += XNT) {
        int ln = idx / M, j = idx - ln * M;
        double2 u0 = res[(0 * LPB + ln) * ld + j], u1 = res[(1 * LPB + ln) * ld + j], u2 = res[(2 * LPB + ln) * ld + j];
        double2 w0 = res[(3 * LPB + ln) * ld + j], w1 = res[(4 * LPB + ln) * ld + j], w2 = res[(5 * LPB + ln) * ld + j];
        nxt[(0 * LPB + ln) * ld + j] = cmk(u1.x * w2.x - u2.x * w1.x, u1.y * w2.y - u2.y * w1.y);
        nxt[(1 * LPB + ln) * ld + j] = cmk(u2.x * w0.x - u0.x * w2.x, u2.y * w0.y - u0.y * w2.y);
        nxt[(2 * LPB + ln) * ld + j] = cmk(u0.x * w1.x - u1.x * w0.x, u0.y * w1.y - u1.y * w0.y);
    }
    __syncthreads();
    double2* r2 = lds_fft(nxt, res, a.plan, 3 * LPB, ld, -1, tid, XNT);
    for (int idx = tid; idx < 3 * LPB * Nxh; idx += XNT) {
        int q = idx / Nxh, k = idx - q * Nxh;
        int c = q / LPB, ln = q - c * LPB;
        long long gl = gl0 + ln;
        if (gl >= nlines) continue;
        double2 Zk = r2[q * ld + k];
        double2 Zc = cconj(r2[q * ld + ((M - k) % M)]);
        double2 w = a.twx[k];  // exp(-2 pi i k / Nxp)
        a.out[c][(size_t)gl * Nxh + k] = csub(cscale(cadd(Zk, Zc), 0.5), cscale(cmuli(cmul(w, csub(Zk, Zc))), 0.5));
    }
}

__global__ void __launch_bounds__(XNT) xpass_kernel(XArgs a) {
    extern __shared__ double2 smem[];
    if (a.mode == X_ROTATIONAL && a.gd.Nxp % 2 == 0) {
        xpass_rotational_even(a, smem);
        return;
    }
    XCtx c(a, smem);
    double* phys = reinterpret_cast<double*>(smem + 2 * a.LPB * a.ld);
    const int Nxp = c.Nxp, LPB = c.LPB;
    if (a.mode == X_ROTATIONAL) {
        for (int f = 0; f < 6; ++f) c.c2r(a.in[f], phys + (size_t)f * LPB * Nxp);
        for (int idx = c.tid; idx < LPB * Nxp; idx += XNT) {
            double u0 = phys[0 * LPB * Nxp + idx], u1 = phys[1 * LPB * Nxp + idx], u2 = phys[2 * LPB * Nxp + idx];
            double w0 = phys[3 * LPB * Nxp + idx], w1 = phys[4 * LPB * Nxp + idx], w2 = phys[5 * LPB * Nxp + idx];
            phys[0 * LPB * Nxp + idx] = u1 * w2 - u2 * w1;
            phys[1 * LPB * Nxp + idx] = u2 * w0 - u0 * w2;
            phys[2 * LPB * Nxp + idx] = u0 * w1 - u1 * w0;
        }
        __syncthreads();
        for (int f = 0; f < 3; ++f) c.r2c(phys + (size_t)f * LPB * Nxp, a.out[f]);
    } else if (a.mode == X_INVERSE) {
        for (int f = 0; f < a.nfields; ++f) {
            c.c2r(a.in[f], phys);
            for (int idx = c.tid; idx < LPB * Nxp; idx += XNT) {
                int ln = idx / Nxp, x = idx - ln * Nxp;
                long long gl = c.gl0 + ln;
                if (gl < c.nlines) a.rout[f][(size_t)gl * Nxp + x] = phys[idx];
            }
            __syncthreads();
        }
    } else {  // X_FORWARD
        for (int f = 0; f < a.nfields; ++f) {
            for (int idx = c.tid; idx < LPB * Nxp; idx += XNT) {
                int ln = idx / Nxp, x = idx - ln * Nxp;
                long long gl = c.gl0 + ln;
                phys[idx] = (gl < c.nlines) ? a.rin[f][(size_t)gl * Nxp + x] : 0.0;
            }
            __syncthreads();
            c.r2c(phys, a.out[f]);
        }
    }
}

size_t z_lds_bytes(const ZArgs& a) { return 2 * (size_t)a.KT * a.ld * sizeof(double2); }

hipError_t launch_zpass(int dir, const ZArgs& a, hipStream_t st) {
    dim3 grid((unsigned)((a.gd.Nxh + a.KT - 1) / a.KT), (unsigned)a.gd.Ny, (unsigned)a.nfields);
    size_t lds = z_lds_bytes(a);
    if (dir > 0) hipLaunchKernelGGL(zpass_kernel<+1>, grid, dim3(ZNT), lds, st, a);
    else hipLaunchKernelGGL(zpass_kernel<-1>, grid, dim3(ZNT), lds, st, a);
    return hipGetLastError();
}

size_t x_lds_bytes(const XArgs& a) {
    if (a.mode == X_ROTATIONAL && a.gd.Nxp % 2 == 0) return 2 * 6 * (size_t)a.LPB * a.ld * sizeof(double2);
    int nphys = (a.mode == X_ROTATIONAL) ? 6 : 1;
    return 2 * (size_t)a.LPB * a.ld * sizeof(double2) + (size_t)nphys * a.LPB * a.gd.Nxp * sizeof(double);
}

hipError_t launch_xpass(const XArgs& a, hipStream_t st) {
    long long nlines = (long long)a.gd.Nzp * a.gd.Ny;
    dim3 grid((unsigned)((nlines + a.LPB - 1) / a.LPB));
    hipLaunchKernelGGL(xpass_kernel, grid, dim3(XNT), x_lds_bytes(a), st, a);
    return hipGetLastError();
}

}  // namespace cf
