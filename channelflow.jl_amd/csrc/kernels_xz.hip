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
// use a half-length complex FFT with the standard pack/unpack.  For the
// rotating form the vorticity component Omega_y = dV_x/dz - dV_z/dx is
// assembled here in k-space from the z pass's i beta l V_x and from V_z, so
// the wall-normal kernels transform five fields instead of six.
#include "cf_device.h"
#include "cf_kernels.h"

namespace cf {

constexpr int ZNT = 256;
constexpr int XNT = 256;
constexpr int XLOADS = 2;  // x pass: pack iterations whose HBM loads are issued together

// z inverse (dir=+1): (k,l,y) -> (k,zp,y); z forward (dir=-1): (k,zp,y) -> (k,l,y)
// 1-D grid of ntiles * Ny * nfields blocks, k tile fastest in logical order.
// C = double2: FP64 fields; C = float2: FP32 fields of the mixed step (in-place
// plans only, twiddles rounded from the FP64 table).
// lb: the block's logical index within its instance (after the XCD remap)
template <int DIR, int PPT, typename C>
__device__ __forceinline__ void zpass_body(const ZArgs& a, unsigned lb) {
    using T = creal_t<C>;
    extern __shared__ double2 smem_raw[];
    C* smem = reinterpret_cast<C*>(smem_raw);
    const GridDims& gd = a.gd;
    const int KT = a.KT, ld = a.ld;
    C* bufA = smem;
    C* bufB = smem + KT * ld;
    const unsigned ntiles = (unsigned)((gd.Nxh + KT - 1) / KT);
    const unsigned tile = lb % ntiles, plane = lb / ntiles;
    const int y = (int)(plane % (unsigned)gd.Ny), f = (int)(plane / (unsigned)gd.Ny);
    const int k0 = (int)tile * KT;
    const int tid = threadIdx.x;
    const int Nz = gd.Nz, Nzp = gd.Nzp, Nxh = gd.Nxh;
    const int npos = (Nz >> 1) + 1, nneg = Nz - npos;
    const bool xeven = (gd.Nx % 2 == 0), zeven = (Nz % 2 == 0);
    const C zero = cplx<C>(T(0), T(0));
    if constexpr (PPT > 0)  // in-place mode: the second buffer holds the twiddle table
        for (int i = tid; i < Nzp; i += ZNT) bufB[i] = from_c64<C>(a.plan.tw[i]);
    // Loads are issued in batches of ZLOADS per thread before their LDS
    // stores, so each thread keeps several HBM requests in flight (a store
    // right after each load would wait for every load in turn).
    constexpr int ZLOADS = 4;
    const int total = KT * Nzp;
    if (DIR > 0) {
        const C* in = reinterpret_cast<const C*>(a.in[f]) + (size_t)Nxh * Nz * y;
        const bool dz = (a.dz_mask >> f) & 1u;  // emit d/dz: multiply by i beta l
        for (int base = tid; base < total; base += ZLOADS * ZNT) {
            C v[ZLOADS];
            int lw[ZLOADS];  // wavenumber label l of the loaded value (dz scale)
#pragma unroll
            for (int b = 0; b < ZLOADS; ++b) {
                const int idx = base + b * ZNT;
                v[b] = zero;
                lw[b] = 0;
                if (idx < total) {
                    int z = (int)fast_div((unsigned)idx, a.mul_kt), kk = idx - z * KT;
                    int k = k0 + kk;
                    if (k < Nxh && !(xeven && k == gd.Nx / 2)) {
                        int l = -1;
                        if (z < npos) l = (zeven && z == Nz / 2) ? -1 : z;
                        else if (z >= Nzp - nneg) l = npos + (z - (Nzp - nneg));
                        if (l >= 0) {
                            v[b] = in[k + (size_t)Nxh * l];
                            lw[b] = l <= Nz / 2 ? l : l - Nz;
                        }
                    }
                }
            }
#pragma unroll
            for (int b = 0; b < ZLOADS; ++b) {
                const int idx = base + b * ZNT;
                if (idx < total) {
                    int z = (int)fast_div((unsigned)idx, a.mul_kt), kk = idx - z * KT;
                    bufA[kk * ld + z] = dz ? cmuli(cscale(v[b], T(gd.beta * lw[b]))) : v[b];
                }
            }
        }
    } else {
        const C* in = reinterpret_cast<const C*>(a.in[f]) + (size_t)Nxh * Nzp * y;
        for (int base = tid; base < total; base += ZLOADS * ZNT) {
            C v[ZLOADS];
#pragma unroll
            for (int b = 0; b < ZLOADS; ++b) {
                const int idx = base + b * ZNT;
                v[b] = zero;
                if (idx < total) {
                    int z = (int)fast_div((unsigned)idx, a.mul_kt), kk = idx - z * KT;
                    int k = k0 + kk;
                    if (k < Nxh) v[b] = in[k + (size_t)Nxh * z];
                }
            }
#pragma unroll
            for (int b = 0; b < ZLOADS; ++b) {
                const int idx = base + b * ZNT;
                if (idx < total) {
                    int z = (int)fast_div((unsigned)idx, a.mul_kt), kk = idx - z * KT;
                    bufA[kk * ld + z] = v[b];
                }
            }
        }
    }
    __syncthreads();
    C* res = bufA;
    if constexpr (PPT > 0) lds_fft_inplace<PPT, false>(bufA, a.plan, (const C*)bufB, KT, ld, DIR, tid, ZNT);  // bufB: twiddles
    else res = lds_fft(bufA, bufB, a.plan, KT, ld, DIR, tid, ZNT);
    if (DIR > 0) {
        C* out = reinterpret_cast<C*>(a.out[f]) + (size_t)Nxh * Nzp * y;
        for (int idx = tid; idx < KT * Nzp; idx += ZNT) {
            int z = (int)fast_div((unsigned)idx, a.mul_kt), kk = idx - z * KT;
            int k = k0 + kk;
            if (k < Nxh) out[k + (size_t)Nxh * z] = res[kk * ld + z];
        }
    } else {
        C* out = reinterpret_cast<C*>(a.out[f]) + (size_t)Nxh * Nz * y;
        for (int idx = tid; idx < KT * Nz; idx += ZNT) {
            int l = (int)fast_div((unsigned)idx, a.mul_kt), kk = idx - l * KT;
            int k = k0 + kk;
            int z = (l < npos) ? l : Nzp - nneg + (l - npos);
            if (k < Nxh) out[k + (size_t)Nxh * l] = res[kk * ld + z];
        }
    }
}

template <int DIR, int PPT, typename C>
__global__ void __launch_bounds__(ZNT) zpass_kernel(ZArgs a) {
    zpass_body<DIR, PPT, C>(a, xcd_logical(blockIdx.x, gridDim.x));
}

// Batched ensemble launch: logical blocks (XCD-remapped over the whole grid)
// [start[i], start[i+1]) belong to instance i with arguments tab[i].
template <int DIR, int PPT>
__global__ void __launch_bounds__(ZNT) zpass_batch_kernel(const ZArgs* __restrict__ tab,
                                                          const unsigned* __restrict__ start, int n) {
    const unsigned lb = xcd_logical(blockIdx.x, gridDim.x);
    const int i = batch_instance(start, n, lb);
    zpass_body<DIR, PPT, double2>(tab[i], lb - start[i]);
}

// ---------------------------------------------------------------------------
// x pass helpers.  Even Nxp = 2M: complex FFT of length M; odd: length Nxp.
// ---------------------------------------------------------------------------
// Pointwise arithmetic shared by every x pass (odd and even Nxp, six-line,
// split and direct-load rotating passes, ensembles), each multiply-add written
// as an explicit fma: every variant then rounds identically whatever the
// surrounding code lets the compiler contract, so any tiling gives the same
// bits.
__device__ __forceinline__ double xfma(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float xfma(float a, float b, float c) { return fmaf(a, b, c); }

// packed input of the half-length inverse transform of a real line:
// Z_k = (X_k + conj X_{M-k}) + i conj(t) (X_k - conj X_{M-k}), t = exp(-2 pi i k / Nxp)
template <typename C>
__device__ __forceinline__ C x_pack(C xk, C xm, C t) {
    using T = creal_t<C>;
    const T sx = xk.x + xm.x, sy = xk.y - xm.y;  // X_k + conj X_{M-k}
    const T dx = xk.x - xm.x, dy = xk.y + xm.y;  // X_k - conj X_{M-k}
    const T px = xfma(t.x, dx, t.y * dy), py = xfma(t.x, dy, -(t.y * dx));  // conj(t) d
    return cplx<C>(sx - py, sy + px);
}

// coefficient k of a real line from the half-length forward transform Z:
// X_k = ((Z_k + conj Z_{M-k}) - i t (Z_k - conj Z_{M-k})) / 2
template <typename C>
__device__ __forceinline__ C x_unpack(C zk, C zm, C t) {
    using T = creal_t<C>;
    const T sx = zk.x + zm.x, sy = zk.y - zm.y;
    const T dx = zk.x - zm.x, dy = zk.y + zm.y;
    const T px = xfma(t.x, dx, -(t.y * dy)), py = xfma(t.x, dy, t.y * dx);  // t d
    return cplx<C>(T(0.5) * (sx + py), T(0.5) * (sy - px));
}

// v - i s w (Omega_y = field 4 - i alpha k V_z)
template <typename C>
__device__ __forceinline__ C x_minus_i(C v, C w, creal_t<C> s) {
    return cplx<C>(xfma(s, w.y, v.x), xfma(-s, w.x, v.y));
}

// V x Omega on the paired points (x = 2j in .x, x = 2j+1 in .y)
template <typename C>
__device__ __forceinline__ void x_cross(C u0, C u1, C u2, C w0, C w1, C w2, C& n0, C& n1, C& n2) {
    n0 = cplx<C>(xfma(u1.x, w2.x, -(u2.x * w1.x)), xfma(u1.y, w2.y, -(u2.y * w1.y)));
    n1 = cplx<C>(xfma(u2.x, w0.x, -(u0.x * w2.x)), xfma(u2.y, w0.y, -(u0.y * w2.y)));
    n2 = cplx<C>(xfma(u0.x, w1.x, -(u1.x * w0.x)), xfma(u0.y, w1.y, -(u1.y * w0.y)));
}

// u gx + v gy + w gz on the paired points
template <typename C>
__device__ __forceinline__ C x_dot3(C u, C v, C w, C gx, C gy, C gz) {
    return cplx<C>(xfma(w.x, gz.x, xfma(v.x, gy.x, u.x * gx.x)), xfma(w.y, gz.y, xfma(v.y, gy.y, u.y * gx.y)));
}

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
    // into phys[line * Nxp + x]; with `sub`, of in - i alpha k sub (the
    // rotating form's Omega_y line)
    __device__ __forceinline__ void c2r(const double2* __restrict__ in, double* phys, bool dx = false,
                                        const double2* __restrict__ sub = nullptr) {
        const bool xeven = (a.gd.Nx % 2 == 0);
        // stage X_k (k = 0..M) in bufB (zero padding, Nyquist plane and Im(DC) removed)
        const int nX = even ? M + 1 : Nxp / 2 + 1;
        for (int idx = tid; idx < LPB * nX; idx += XNT) {
            int ln = idx / nX, k = idx - ln * nX;
            long long gl = gl0 + ln;
            double2 v = czero();
            if (gl < nlines && k < Nxh && !(xeven && k == a.gd.Nx / 2)) {
                v = in[(size_t)gl * Nxh + k];
                if (sub) v = x_minus_i(v, sub[(size_t)gl * Nxh + k], a.gd.alpha * k);
            }
            if (k == 0) v.y = 0.0;
            if (even && k == M) v.y = 0.0;
            if (dx) v = cmuli(cscale(v, a.gd.alpha * k));
            bufB[ln * ld + k] = v;
        }
        __syncthreads();
        if (even) {
            for (int idx = tid; idx < LPB * M; idx += XNT) {
                int ln = idx / M, k = idx - ln * M;
                bufA[ln * ld + k] = x_pack(bufB[ln * ld + k], bufB[ln * ld + (M - k)], a.twx[k]);
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
                X = x_unpack(res[ln * ld + (k % M)], res[ln * ld + ((M - k) % M)], a.twx[k]);
            } else {
                X = res[ln * ld + k];
            }
            out[(size_t)gl * Nxh + k] = X;
        }
        __syncthreads();
    }
};

// Nonlinear x pass for even Nxp.  Each wavefront owns one (zp, y) line and a
// private LDS region: the fields the form needs are one batch of length-M
// complex FFTs (half-length real transforms), the pointwise products are
// formed in LDS and packed directly as the input of the forward FFTs.
//   rotating   in V, Omega (6)             out V x Omega (3)          (nonlinear.jl:187-189)
//   convective in V, dV/dy, dV/dz (9);     out (V . grad) V (3)       (nonlinear.jl:124-133)
//              dV/dx is formed here as i alpha k V
//   divergence in V (3)                    out V_i V_j, i <= j (6)    (nonlinear.jl:157-160)
// Only wave-level ordering is needed after the block loads the twiddles.
// A block has a.waves wavefronts (64 * a.waves threads).
//   gradient   in V, dV/dy, dV/dz (9);     out dV_i/dx_j on the padded points (9 real fields, no forward FFT)
__host__ __device__ constexpr int x_fft_lines_in(int mode) {
    return (mode == X_CONVECTIVE || mode == X_GRADIENT) ? 12 : ((mode == X_DIVERGENCE || mode == X_ROT_SPLIT) ? 3 : 6);
}
__host__ __device__ constexpr int x_fields_in(int mode) {
    return (mode == X_CONVECTIVE || mode == X_GRADIENT) ? 9 : (mode == X_DIVERGENCE ? 3 : 6);
}
__host__ __device__ constexpr int x_fields_out(int mode) { return mode == X_DIVERGENCE ? 6 : (mode == X_GRADIENT ? 0 : 3); }
__host__ __device__ constexpr int x_buffer_lines(int mode) {
    return x_fft_lines_in(mode) > x_fields_out(mode) ? x_fft_lines_in(mode) : x_fields_out(mode);
}
// LDS slots of one wave's region: one buffer (in place) or two
__host__ __device__ constexpr int x_region_slots(int mode, int ppt, int ld) { return (ppt > 0 ? 1 : 2) * x_buffer_lines(mode) * ld; }

// retained coefficient X_k of field f on line gl for the half-length inverse
// transform (padding, Nyquist plane and Im of DC / of k = M removed)
template <typename C>
__device__ __forceinline__ C x_coef(const XArgs& a, int f, long long gl, int k, int M) {
    C v = cplx<C>(0, 0);
    if (k < a.gd.Nxh && !((a.gd.Nx % 2 == 0) && k == a.gd.Nx / 2))
        v = reinterpret_cast<const C*>(a.in[f])[(size_t)gl * a.gd.Nxh + k];
    if (k == 0 || k == M) v.y = 0;
    return v;
}

// Coefficient X_k of FFT line f.  Rotating forms: line 4 is Omega_y =
// dV_x/dz - dV_z/dx, formed here as field 4 (i beta l V_x, emitted by the z
// inverse) minus i alpha k times field 2 (V_z).
template <int MODE, typename C>
__device__ __forceinline__ C x_line_coef(const XArgs& a, int f, long long gl, int k, int M) {
    C v = x_coef<C>(a, f, gl, k, M);
    if constexpr (MODE == X_ROTATIONAL || MODE == X_ROT_SPLIT || MODE == X_ROT_ASYNC) {
        if (f == 4) v = x_minus_i(v, x_coef<C>(a, 2, gl, k, M), creal_t<C>(a.gd.alpha * k));
    }
    return v;
}

template <int MODE, int PPT, typename C>
__device__ __forceinline__ void xpass_wave(const XArgs& a, C* smem, unsigned block) {
    using T = creal_t<C>;
    constexpr int NIN = x_fields_in(MODE), NFFT = x_fft_lines_in(MODE), NOUT = x_fields_out(MODE);
    constexpr int NB = x_buffer_lines(MODE);
    const int Nxp = a.gd.Nxp, Nxh = a.gd.Nxh, M = Nxp / 2, ld = a.ld;
    const long long nlines = (long long)a.gd.Nzp * a.gd.Ny;
    const bool xeven = (a.gd.Nx % 2 == 0);
    // twiddles: FFT table (M) and pack/unpack table exp(-2 pi i k / Nxp), k <= M
    C* tw = smem;
    C* twx = smem + M;
    for (int i = threadIdx.x; i < M; i += blockDim.x) tw[i] = from_c64<C>(a.plan.tw[i]);
    for (int i = threadIdx.x; i <= M; i += blockDim.x) twx[i] = from_c64<C>(a.twx[i]);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long gl = (long long)block * a.waves + wave;
    if (gl >= nlines) return;  // whole wavefront leaves; no block barrier follows
    C* bufA = smem + 2 * M + 1 + (size_t)wave * x_region_slots(MODE, PPT, ld);
    if constexpr (MODE == X_ROT_SPLIT) {
        // rotating form, three LDS lines: velocity lines first, their physical
        // values kept in registers (points j = lane + 64 i; 3 M <= 64 PPT
        // bounds i), then the vorticity lines in the same buffer
        static_assert(PPT > 0, "the split rotating pass is in place");
        constexpr int JM = (PPT + 2) / 3;
        C u[3][JM];
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            // XLOADS iterations' HBM loads in flight before their LDS stores
            for (int base = lane; base < 3 * M; base += XLOADS * 64) {
                C Xk[XLOADS], Xm[XLOADS];
#pragma unroll
                for (int b = 0; b < XLOADS; ++b) {
                    const int idx = base + b * 64;
                    Xk[b] = Xm[b] = cplx<C>(0, 0);
                    if (idx < 3 * M) {
                        const int q = (int)fast_div((unsigned)idx, a.mul_m), k = idx - q * M;
                        Xk[b] = x_line_coef<MODE, C>(a, 3 * half + q, gl, k, M);
                        Xm[b] = x_line_coef<MODE, C>(a, 3 * half + q, gl, M - k, M);
                    }
                }
#pragma unroll
                for (int b = 0; b < XLOADS; ++b) {
                    const int idx = base + b * 64;
                    if (idx < 3 * M) {
                        const int q = (int)fast_div((unsigned)idx, a.mul_m), k = idx - q * M;
                        bufA[q * ld + k] = x_pack(Xk[b], Xm[b], twx[k]);
                    }
                }
            }
            wave_sync();
            lds_fft_inplace<PPT, true>(bufA, a.plan, (const C*)tw, 3, ld, +1, lane, 64);
            if (half == 0) {
#pragma unroll
                for (int i = 0; i < JM; ++i) {
                    const int j = lane + 64 * i;
#pragma unroll
                    for (int q = 0; q < 3; ++q) u[q][i] = j < M ? bufA[q * ld + j] : cplx<C>(0, 0);
                }
                wave_sync();  // every lane has read before the vorticity lines overwrite
            }
        }
        // V x Omega at x = 2j (real parts) and 2j+1 (imaginary parts), in place
#pragma unroll
        for (int i = 0; i < JM; ++i) {
            const int j = lane + 64 * i;
            if (j < M) {
                const C w0 = bufA[0 * ld + j], w1 = bufA[1 * ld + j], w2 = bufA[2 * ld + j];
                C n0, n1, n2;
                x_cross(u[0][i], u[1][i], u[2][i], w0, w1, w2, n0, n1, n2);
                bufA[0 * ld + j] = n0;
                bufA[1 * ld + j] = n1;
                bufA[2 * ld + j] = n2;
            }
        }
        wave_sync();
        lds_fft_inplace<PPT, true>(bufA, a.plan, (const C*)tw, 3, ld, -1, lane, 64);
        for (int idx = lane; idx < 3 * Nxh; idx += 64) {
            const int c = (int)fast_div((unsigned)idx, a.mul_nxh), k = idx - c * Nxh;
            reinterpret_cast<C*>(a.out[c])[(size_t)gl * Nxh + k] =
                x_unpack(bufA[c * ld + k], bufA[c * ld + ((M - k) % M)], twx[k]);
        }
        return;
    } else if constexpr (PPT > 0) {
      if constexpr (MODE == X_ROT_ASYNC) {
        // Raw coefficients of the six fields straight from HBM into the six
        // lines (global_load_lds: no registers, every load in flight at
        // once); padding k = Nxh..M and the x Nyquist k = Nx/2 are zero.
        static_assert(sizeof(C) == 16, "direct LDS loads move 16-byte elements");
        const int nload = xeven ? a.gd.Nx / 2 : Nxh;  // retained coefficients, Nyquist excluded
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            const C* src = reinterpret_cast<const C*>(a.in[f]) + (size_t)gl * Nxh;
            for (int k0 = 0; k0 < nload; k0 += 64)
                if (k0 + lane < nload)
                    __builtin_amdgcn_global_load_lds((const void*)(src + k0 + lane),
                                                     (__attribute__((address_space(3))) void*)(bufA + f * ld + k0),
                                                     16, 0, 0);
        }
        for (int idx = lane; idx < 6 * (M + 1 - nload); idx += 64) {
            const int f = idx / (M + 1 - nload), k = nload + idx - f * (M + 1 - nload);
            bufA[f * ld + k] = cplx<C>(0, 0);
        }
        __builtin_amdgcn_s_waitcnt(0);  // the direct loads have landed
        wave_sync();
        // pack pairs (k, M - k) in place, each pair by one lane; line 4 first
        // (Omega_y = X_4 - i alpha k X_2 reads line 2 before it is packed)
        // the same values (and operation order) as x_coef / x_line_coef
        auto coef = [&](int q, int k) {
            C v = bufA[q * ld + k];
            if (k == 0 || k == M) v.y = 0;
            return v;
        };
        auto line_coef = [&](int q, int k) {
            C v = coef(q, k);
            if (q == 4) v = x_minus_i(v, coef(2, k), T(a.gd.alpha * k));
            return v;
        };
        auto pack_pair = [&](int q, int k) {
            const C xk = line_coef(q, k), xm = line_coef(q, M - k);
            bufA[q * ld + k] = x_pack(xk, xm, twx[k]);
            if (M - k != k && M - k < M) bufA[q * ld + (M - k)] = x_pack(xm, xk, twx[M - k]);
        };
        const int npair = M / 2 + 1;  // k = 0..M/2
        for (int k = lane; k < npair; k += 64) pack_pair(4, k);
        wave_sync();
        for (int idx = lane; idx < 5 * npair; idx += 64) {
            const int j = idx / npair, k = idx - j * npair;
            pack_pair(j < 4 ? j : 5, k);
        }
        wave_sync();
      } else {
        // pack straight from HBM into the single buffer; FFT line q reads
        // field src(q), times i alpha k for the convective x derivatives
        for (int base = lane; base < NFFT * M; base += XLOADS * 64) {
            C Xk[XLOADS], Xm[XLOADS];  // loads of XLOADS iterations in flight together
#pragma unroll
            for (int b = 0; b < XLOADS; ++b) {
                const int idx = base + b * 64;
                Xk[b] = Xm[b] = cplx<C>(0, 0);
                if (idx < NFFT * M) {
                    const int q = (int)fast_div((unsigned)idx, a.mul_m), k = idx - q * M;
                    int f = q;
                    if constexpr (MODE == X_CONVECTIVE || MODE == X_GRADIENT) f = q < 3 ? q : q - 3;
                    Xk[b] = x_line_coef<MODE, C>(a, f, gl, k, M);
                    Xm[b] = x_line_coef<MODE, C>(a, f, gl, M - k, M);
                }
            }
#pragma unroll
            for (int b = 0; b < XLOADS; ++b) {
                const int idx = base + b * 64;
                if (idx < NFFT * M) {
                    const int q = (int)fast_div((unsigned)idx, a.mul_m), k = idx - q * M;
                    C xk = Xk[b], xm = Xm[b];
                    if constexpr (MODE == X_CONVECTIVE || MODE == X_GRADIENT) {
                        if (q >= 3 && q < 6) {
                            xk = cmuli(cscale(xk, T(a.gd.alpha * k)));
                            xm = cmuli(cscale(xm, T(a.gd.alpha * (M - k))));
                        }
                    }
                    bufA[q * ld + k] = x_pack(xk, xm, twx[k]);
                }
            }
        }
        wave_sync();
      }
        lds_fft_inplace<PPT, true>(bufA, a.plan, (const C*)tw, NFFT, ld, +1, lane, 64);
        if constexpr (MODE == X_GRADIENT) {
            // line 3 + 3 d + i holds dV_i/dx_d (d = x, y, z): values at x = 2m, 2m+1
            for (int idx = lane; idx < 9 * M; idx += 64) {
                const int c = (int)fast_div((unsigned)idx, a.mul_m), m = idx - c * M;
                const int i = c / 3, d = c - 3 * i;
                const C v = bufA[(3 + 3 * d + i) * ld + m];
                double* o = a.rout[c] + (size_t)gl * Nxp + 2 * m;
                o[0] = v.x;
                o[1] = v.y;
            }
            return;
        }
        // products at x = 2j (real parts) and 2j+1 (imaginary parts), in place:
        // each lane reads all lines at its j before writing them
        for (int j = lane; j < M; j += 64) {
            if constexpr (MODE == X_ROTATIONAL || MODE == X_ROT_ASYNC) {
                C u0 = bufA[0 * ld + j], u1 = bufA[1 * ld + j], u2 = bufA[2 * ld + j];
                C w0 = bufA[3 * ld + j], w1 = bufA[4 * ld + j], w2 = bufA[5 * ld + j];
                C n0, n1, n2;
                x_cross(u0, u1, u2, w0, w1, w2, n0, n1, n2);
                bufA[0 * ld + j] = n0;
                bufA[1 * ld + j] = n1;
                bufA[2 * ld + j] = n2;
            } else if constexpr (MODE == X_CONVECTIVE) {
                C u = bufA[0 * ld + j], v = bufA[1 * ld + j], w = bufA[2 * ld + j];
                C n[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    n[i] = x_dot3(u, v, w, bufA[(3 + i) * ld + j], bufA[(6 + i) * ld + j], bufA[(9 + i) * ld + j]);
                }
#pragma unroll
                for (int i = 0; i < 3; ++i) bufA[i * ld + j] = n[i];
            } else {
                C u = bufA[0 * ld + j], v = bufA[1 * ld + j], w = bufA[2 * ld + j];
                bufA[0 * ld + j] = cplx<C>(u.x * u.x, u.y * u.y);
                bufA[1 * ld + j] = cplx<C>(u.x * v.x, u.y * v.y);
                bufA[2 * ld + j] = cplx<C>(u.x * w.x, u.y * w.y);
                bufA[3 * ld + j] = cplx<C>(v.x * v.x, v.y * v.y);
                bufA[4 * ld + j] = cplx<C>(v.x * w.x, v.y * w.y);
                bufA[5 * ld + j] = cplx<C>(w.x * w.x, w.y * w.y);
            }
        }
        wave_sync();
        lds_fft_inplace<PPT, true>(bufA, a.plan, (const C*)tw, NOUT, ld, -1, lane, 64);
        for (int idx = lane; idx < NOUT * Nxh; idx += 64) {
            int c = (int)fast_div((unsigned)idx, a.mul_nxh), k = idx - c * Nxh;
            reinterpret_cast<C*>(a.out[c])[(size_t)gl * Nxh + k] =
                x_unpack(bufA[c * ld + k], bufA[c * ld + ((M - k) % M)], twx[k]);
        }
        return;
    }
    C* bufB = bufA + NB * ld;
    const int nX = M + 1;
    // stage X_k, k = 0..M (padding, Nyquist plane and Im of DC/Nyquist removed)
    for (int idx = lane; idx < NIN * nX; idx += 64) {
        int f = (int)fast_div((unsigned)idx, a.mul_nx), k = idx - f * nX;
        C v = cplx<C>(0, 0);
        if (k < Nxh && !(xeven && k == a.gd.Nx / 2)) {
            v = reinterpret_cast<const C*>(a.in[f])[(size_t)gl * Nxh + k];
            if constexpr (MODE == X_ROTATIONAL)  // Omega_y = field 4 - i alpha k V_z
                if (f == 4) v = x_minus_i(v, reinterpret_cast<const C*>(a.in[2])[(size_t)gl * Nxh + k], T(a.gd.alpha * k));
        }
        if (k == 0 || k == M) v.y = 0;
        bufB[f * ld + k] = v;
    }
    wave_sync();
    // pack for the half-length inverse transform; FFT line q reads staged
    // field src(q), multiplied by i alpha k for the convective x derivatives
    for (int idx = lane; idx < NFFT * M; idx += 64) {
        int q = (int)fast_div((unsigned)idx, a.mul_m), k = idx - q * M;
        int f = q;
        bool dx = false;
        if constexpr (MODE == X_CONVECTIVE || MODE == X_GRADIENT) {
            dx = (q >= 3 && q < 6);
            f = q < 3 ? q : q - 3;
        }
        C Xk = bufB[f * ld + k];
        C Xm = bufB[f * ld + (M - k)];
        if (dx) {
            Xk = cmuli(cscale(Xk, T(a.gd.alpha * k)));
            Xm = cmuli(cscale(Xm, T(a.gd.alpha * (M - k))));
        }
        bufA[q * ld + k] = x_pack(Xk, Xm, twx[k]);
    }
    wave_sync();
    C* res = lds_fft<true>(bufA, bufB, a.plan, (const C*)tw, NFFT, ld, +1, lane, 64);
    if constexpr (MODE == X_GRADIENT) {
        for (int idx = lane; idx < 9 * M; idx += 64) {
            const int c = (int)fast_div((unsigned)idx, a.mul_m), m = idx - c * M;
            const int i = c / 3, d = c - 3 * i;
            const C v = res[(3 + 3 * d + i) * ld + m];
            double* o = a.rout[c] + (size_t)gl * Nxp + 2 * m;
            o[0] = v.x;
            o[1] = v.y;
        }
        return;
    }
    C* nxt = (res == bufA) ? bufB : bufA;
    // pointwise products at x = 2j (real parts) and x = 2j+1 (imaginary parts)
    for (int j = lane; j < M; j += 64) {
        if constexpr (MODE == X_ROTATIONAL) {
            x_cross(res[0 * ld + j], res[1 * ld + j], res[2 * ld + j], res[3 * ld + j], res[4 * ld + j], res[5 * ld + j],
                    nxt[0 * ld + j], nxt[1 * ld + j], nxt[2 * ld + j]);
        } else if constexpr (MODE == X_CONVECTIVE) {
            C u = res[0 * ld + j], v = res[1 * ld + j], w = res[2 * ld + j];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                nxt[i * ld + j] = x_dot3(u, v, w, res[(3 + i) * ld + j], res[(6 + i) * ld + j], res[(9 + i) * ld + j]);
            }
        } else {
            C u = res[0 * ld + j], v = res[1 * ld + j], w = res[2 * ld + j];
            nxt[0 * ld + j] = cplx<C>(u.x * u.x, u.y * u.y);
            nxt[1 * ld + j] = cplx<C>(u.x * v.x, u.y * v.y);
            nxt[2 * ld + j] = cplx<C>(u.x * w.x, u.y * w.y);
            nxt[3 * ld + j] = cplx<C>(v.x * v.x, v.y * v.y);
            nxt[4 * ld + j] = cplx<C>(v.x * w.x, v.y * w.y);
            nxt[5 * ld + j] = cplx<C>(w.x * w.x, w.y * w.y);
        }
    }
    wave_sync();
    C* r2 = lds_fft<true>(nxt, res, a.plan, (const C*)tw, NOUT, ld, -1, lane, 64);
    for (int idx = lane; idx < NOUT * Nxh; idx += 64) {
        int c = (int)fast_div((unsigned)idx, a.mul_nxh), k = idx - c * Nxh;
        reinterpret_cast<C*>(a.out[c])[(size_t)gl * Nxh + k] = x_unpack(r2[c * ld + k], r2[c * ld + ((M - k) % M)], twx[k]);
    }
}

// one kernel per (form, in-place width) so each gets its own register budget
template <int MODE, int PPT, typename C>
__global__ void __launch_bounds__(XNT) xwave_kernel(XArgs a) {
    extern __shared__ double2 smem[];
    xpass_wave<MODE, PPT>(a, reinterpret_cast<C*>(smem), blockIdx.x);
}

// Batched ensemble launch (FP64); every instance of the launch has the same
// wavefronts per block (the block size) and in-place width.
template <int MODE, int PPT>
__global__ void __launch_bounds__(XNT) xwave_batch_kernel(const XArgs* __restrict__ tab,
                                                          const unsigned* __restrict__ start, int n) {
    extern __shared__ double2 smem[];
    const int i = batch_instance(start, n, blockIdx.x);
    xpass_wave<MODE, PPT>(tab[i], smem, blockIdx.x - start[i]);
}

__global__ void __launch_bounds__(XNT) xpass_kernel(XArgs a) {
    extern __shared__ double2 smem[];
    XCtx c(a, smem);
    double* phys = reinterpret_cast<double*>(smem + 2 * a.LPB * a.ld);
    const int Nxp = c.Nxp, LPB = c.LPB;
    if (a.mode == X_CONVECTIVE) {  // odd Nxp: phys lines V, dV/dx, dV/dy, dV/dz
        for (int f = 0; f < 3; ++f) c.c2r(a.in[f], phys + (size_t)f * LPB * Nxp);
        for (int f = 0; f < 3; ++f) c.c2r(a.in[f], phys + (size_t)(3 + f) * LPB * Nxp, true);
        for (int f = 3; f < 9; ++f) c.c2r(a.in[f], phys + (size_t)(3 + f) * LPB * Nxp);
        for (int idx = c.tid; idx < LPB * Nxp; idx += XNT) {
            double u = phys[idx], v = phys[1 * LPB * Nxp + idx], w = phys[2 * LPB * Nxp + idx];
            double n[3];
#pragma unroll
            for (int i = 0; i < 3; ++i)
                n[i] = xfma(w, phys[(9 + i) * LPB * Nxp + idx],
                            xfma(v, phys[(6 + i) * LPB * Nxp + idx], u * phys[(3 + i) * LPB * Nxp + idx]));
#pragma unroll
            for (int i = 0; i < 3; ++i) phys[i * LPB * Nxp + idx] = n[i];
        }
        __syncthreads();
        for (int f = 0; f < 3; ++f) c.r2c(phys + (size_t)f * LPB * Nxp, a.out[f]);
    } else if (a.mode == X_DIVERGENCE) {  // odd Nxp: products into phys lines 3..8
        for (int f = 0; f < 3; ++f) c.c2r(a.in[f], phys + (size_t)f * LPB * Nxp);
        for (int idx = c.tid; idx < LPB * Nxp; idx += XNT) {
            double u = phys[idx], v = phys[1 * LPB * Nxp + idx], w = phys[2 * LPB * Nxp + idx];
            phys[3 * LPB * Nxp + idx] = u * u;
            phys[4 * LPB * Nxp + idx] = u * v;
            phys[5 * LPB * Nxp + idx] = u * w;
            phys[6 * LPB * Nxp + idx] = v * v;
            phys[7 * LPB * Nxp + idx] = v * w;
            phys[8 * LPB * Nxp + idx] = w * w;
        }
        __syncthreads();
        for (int f = 0; f < 6; ++f) c.r2c(phys + (size_t)(3 + f) * LPB * Nxp, a.out[f]);
    } else if (a.mode == X_ROTATIONAL) {
        for (int f = 0; f < 6; ++f) c.c2r(a.in[f], phys + (size_t)f * LPB * Nxp, false, f == 4 ? a.in[2] : nullptr);
        for (int idx = c.tid; idx < LPB * Nxp; idx += XNT) {
            double u0 = phys[0 * LPB * Nxp + idx], u1 = phys[1 * LPB * Nxp + idx], u2 = phys[2 * LPB * Nxp + idx];
            double w0 = phys[3 * LPB * Nxp + idx], w1 = phys[4 * LPB * Nxp + idx], w2 = phys[5 * LPB * Nxp + idx];
            phys[0 * LPB * Nxp + idx] = xfma(u1, w2, -(u2 * w1));
            phys[1 * LPB * Nxp + idx] = xfma(u2, w0, -(u0 * w2));
            phys[2 * LPB * Nxp + idx] = xfma(u0, w1, -(u1 * w0));
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

size_t z_lds_bytes(const ZArgs& a) {
    // in place: one buffer plus the twiddle table (Nzp entries); else two buffers
    const size_t elem = a.mixed ? sizeof(float2) : sizeof(double2);
    return (a.ppt > 0 ? (size_t)a.KT * a.ld + a.gd.Nzp : 2 * (size_t)a.KT * a.ld) * elem;
}

// in-place instantiations (points per thread); the host picks the smallest that fits
int z_inplace_ppt(const FFTPlan& plan, int KT) {
    for (int p : {8, 16})
        if (fft_inplace_fits(plan, KT, p, ZNT)) return p;
    return 0;
}

template <int DIR>
static void launch_z(const ZArgs& a, dim3 grid, size_t lds, hipStream_t st) {
    if (a.mixed) {
        if (a.ppt == 8) hipLaunchKernelGGL((zpass_kernel<DIR, 8, float2>), grid, dim3(ZNT), lds, st, a);
        else hipLaunchKernelGGL((zpass_kernel<DIR, 16, float2>), grid, dim3(ZNT), lds, st, a);
        return;
    }
    switch (a.ppt) {
        case 8: hipLaunchKernelGGL((zpass_kernel<DIR, 8, double2>), grid, dim3(ZNT), lds, st, a); break;
        case 16: hipLaunchKernelGGL((zpass_kernel<DIR, 16, double2>), grid, dim3(ZNT), lds, st, a); break;
        default: hipLaunchKernelGGL((zpass_kernel<DIR, 0, double2>), grid, dim3(ZNT), lds, st, a); break;
    }
}

hipError_t launch_zpass(int dir, const ZArgs& a, hipStream_t st) {
    if (a.ppt != 0 && a.ppt != 8 && a.ppt != 16) return hipErrorInvalidValue;
    if (a.mixed && a.ppt == 0) return hipErrorInvalidValue;  // FP32 fields: in-place plans only
    dim3 grid((unsigned)((a.gd.Nxh + a.KT - 1) / a.KT) * (unsigned)a.gd.Ny * (unsigned)a.nfields);
    size_t lds = z_lds_bytes(a);
    if (dir > 0) launch_z<+1>(a, grid, lds, st);
    else launch_z<-1>(a, grid, lds, st);
    return hipGetLastError();
}

unsigned z_blocks(const ZArgs& a) {
    return (unsigned)((a.gd.Nxh + a.KT - 1) / a.KT) * (unsigned)a.gd.Ny * (unsigned)a.nfields;
}

hipError_t launch_zpass_batch(int dir, int ppt, const ZArgs* tab, const unsigned* start, int n, unsigned nblocks,
                              size_t lds, hipStream_t st) {
#define CF_ZB(D, P)                                                                                            \
    {                                                                                                          \
        if (lds > 64 * 1024)                                                                                   \
            (void)hipFuncSetAttribute((const void*)zpass_batch_kernel<D, P>,                                   \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                   \
        hipLaunchKernelGGL((zpass_batch_kernel<D, P>), dim3(nblocks), dim3(ZNT), lds, st, tab, start, n);      \
        return hipGetLastError();                                                                              \
    }
    if (dir > 0) {
        if (ppt == 8) CF_ZB(+1, 8)
        if (ppt == 16) CF_ZB(+1, 16)
        if (ppt == 0) CF_ZB(+1, 0)
    } else {
        if (ppt == 8) CF_ZB(-1, 8)
        if (ppt == 16) CF_ZB(-1, 16)
        if (ppt == 0) CF_ZB(-1, 0)
    }
#undef CF_ZB
    return hipErrorInvalidValue;
}

static bool x_wave_path(const XArgs& a) {
    return a.gd.Nxp % 2 == 0 && (a.mode == X_ROTATIONAL || a.mode == X_CONVECTIVE || a.mode == X_DIVERGENCE ||
                                 a.mode == X_GRADIENT || a.mode == X_ROT_SPLIT || a.mode == X_ROT_ASYNC);
}

int x_phys_lines(int mode) {
    return mode == X_ROTATIONAL ? 6 : (mode == X_CONVECTIVE || mode == X_GRADIENT) ? 12 : mode == X_DIVERGENCE ? 9 : 1;
}

static size_t x_wave_lds(const XArgs& a, int waves) {
    const size_t elem = a.mixed ? sizeof(float2) : sizeof(double2);
    return ((size_t)a.gd.Nxp + 1 + (size_t)waves * x_region_slots(a.mode, a.ppt, a.ld)) * elem;
}

int x_inplace_ppt(const XArgs& a) {
    if (a.gd.Nxp % 2 != 0) return 0;
    if (a.mode == X_ROT_SPLIT) {  // in place only
        for (int p : {5, 9})
            if (fft_inplace_fits(a.plan, x_buffer_lines(a.mode), p, 64)) return p;
        return 0;
    }
    for (int p : {9, 18})
        if (fft_inplace_fits(a.plan, x_buffer_lines(a.mode), p, 64)) return p;
    return 0;
}

int x_max_waves(const XArgs& a) {
    for (int w = XNT / 64; w >= 1; --w)
        if (x_wave_lds(a, w) <= 160 * 1024) return w;
    return 0;
}

size_t x_lds_bytes(const XArgs& a) {
    if (x_wave_path(a)) return x_wave_lds(a, a.waves);
    int nphys = x_phys_lines(a.mode);
    return 2 * (size_t)a.LPB * a.ld * sizeof(double2) + (size_t)nphys * a.LPB * a.gd.Nxp * sizeof(double);
}

unsigned x_blocks(const XArgs& a) {
    const long long nlines = (long long)a.gd.Nzp * a.gd.Ny;
    return (unsigned)((nlines + a.waves - 1) / a.waves);
}

hipError_t launch_xpass_batch(int mode, int ppt, int waves, const XArgs* tab, const unsigned* start, int n,
                              unsigned nblocks, size_t lds, hipStream_t st) {
    if (waves < 1 || waves > XNT / 64) return hipErrorInvalidValue;
    const dim3 block(64 * waves);
#define CF_XB(M, P)                                                                                            \
    {                                                                                                          \
        if (lds > 64 * 1024)                                                                                   \
            (void)hipFuncSetAttribute((const void*)xwave_batch_kernel<M, P>,                                   \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                   \
        hipLaunchKernelGGL((xwave_batch_kernel<M, P>), dim3(nblocks), block, lds, st, tab, start, n);          \
        return hipGetLastError();                                                                              \
    }
#define CF_XBM(M)                        \
    {                                    \
        if (ppt == 9) CF_XB(M, 9)        \
        if (ppt == 18) CF_XB(M, 18)      \
        if (ppt == 0) CF_XB(M, 0)        \
    }
    if (mode == X_ROT_SPLIT) {
        if (ppt == 5) CF_XB(X_ROT_SPLIT, 5)
        if (ppt == 9) CF_XB(X_ROT_SPLIT, 9)
    }
    if (mode == X_ROT_ASYNC) {
        if (ppt == 9) CF_XB(X_ROT_ASYNC, 9)
        if (ppt == 18) CF_XB(X_ROT_ASYNC, 18)
    }
    if (mode == X_ROTATIONAL) CF_XBM(X_ROTATIONAL)
    if (mode == X_CONVECTIVE) CF_XBM(X_CONVECTIVE)
#undef CF_XBM
#undef CF_XB
    return hipErrorInvalidValue;
}

hipError_t launch_xpass(const XArgs& a, hipStream_t st) {
    long long nlines = (long long)a.gd.Nzp * a.gd.Ny;
    const bool wave = x_wave_path(a);
    if (wave && (a.waves < 1 || a.waves > XNT / 64)) return hipErrorInvalidValue;
    if (wave && a.mode == X_ROT_SPLIT && a.ppt != 5 && a.ppt != 9) return hipErrorInvalidValue;
    if (wave && a.mode != X_ROT_SPLIT && a.ppt != 0 && a.ppt != 9 && a.ppt != 18) return hipErrorInvalidValue;
    if (a.mixed && (!wave || a.ppt == 0)) return hipErrorInvalidValue;  // FP32 fields: in-place wave path only
    const int per_block = wave ? a.waves : a.LPB;
    dim3 grid((unsigned)((nlines + per_block - 1) / per_block));
    size_t lds = x_lds_bytes(a);
    if (!wave) {
        if (lds > 64 * 1024) (void)hipFuncSetAttribute((const void*)xpass_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(xpass_kernel, grid, dim3(XNT), lds, st, a);
        return hipGetLastError();
    }
    const dim3 block(64 * a.waves);
#define CF_XWAVE(MODE, PPT, C)                                                                                   \
    {                                                                                                            \
        if (lds > 64 * 1024)                                                                                     \
            (void)hipFuncSetAttribute((const void*)xwave_kernel<MODE, PPT, C>,                                   \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                     \
        hipLaunchKernelGGL((xwave_kernel<MODE, PPT, C>), grid, block, lds, st, a);                               \
        return hipGetLastError();                                                                                \
    }
// one statement (braced): every path returns the launch status
#define CF_XMODE(MODE)                                    \
    {                                                     \
        if (a.mixed) {                                    \
            if (a.ppt == 9) CF_XWAVE(MODE, 9, float2)     \
            CF_XWAVE(MODE, 18, float2)                    \
        }                                                 \
        switch (a.ppt) {                                  \
            case 9: CF_XWAVE(MODE, 9, double2)            \
            case 18: CF_XWAVE(MODE, 18, double2)          \
            default: CF_XWAVE(MODE, 0, double2)           \
        }                                                 \
    }
    if (a.mode == X_ROT_SPLIT) {
        if (a.mixed) {
            if (a.ppt == 5) CF_XWAVE(X_ROT_SPLIT, 5, float2)
            CF_XWAVE(X_ROT_SPLIT, 9, float2)
        }
        if (a.ppt == 5) CF_XWAVE(X_ROT_SPLIT, 5, double2)
        CF_XWAVE(X_ROT_SPLIT, 9, double2)
    }
    if (a.mode == X_ROT_ASYNC) {  // FP64, in place only
        if (a.mixed || (a.ppt != 9 && a.ppt != 18)) return hipErrorInvalidValue;
        if (a.ppt == 9) CF_XWAVE(X_ROT_ASYNC, 9, double2)
        CF_XWAVE(X_ROT_ASYNC, 18, double2)
    }
    if (a.mode == X_ROTATIONAL) CF_XMODE(X_ROTATIONAL)
    if (a.mode == X_CONVECTIVE) CF_XMODE(X_CONVECTIVE)
    if (a.mode == X_DIVERGENCE) CF_XMODE(X_DIVERGENCE)
    if (a.mode == X_GRADIENT && !a.mixed) {
        switch (a.ppt) {
            case 9: CF_XWAVE(X_GRADIENT, 9, double2)
            case 18: CF_XWAVE(X_GRADIENT, 18, double2)
            default: CF_XWAVE(X_GRADIENT, 0, double2)
        }
    }
    return hipErrorInvalidValue;
#undef CF_XMODE
#undef CF_XWAVE
}

}  // namespace cf
