// Wall-normal (Chebyshev) pencil kernels.
//
// A block owns M = NT/L consecutive Fourier modes s = ix + Nxh*iz of the
// reference (k,l,n) layout; a group of L lanes owns one mode, lane t holding
// Chebyshev coefficients n = Kt..Kt+K-1 in registers (K = 4, or 6 / 8 when
// Ny exceeds 256 / 384).  HBM rows (fixed n, M
// consecutive modes) are moved through LDS, which also stages the DCT-I
// (REDFT00) as a length-2P Stockham FFT of the even extension.
//
// Reference: CNRK2 stage (src/timesteppers/cnrk2.jl:170-205), Stokes solve
// (src/solvers/stokes.jl:81-108, batchedinfluence.jl:129-159, meanmode.jl:103-150),
// Chebyshev transforms (src/ffts.jl:65-82,151-166, transforms/chebyshev.jl:45-66),
// curl (src/fields/operators.jl:172-190), pressure (src/pressure.jl:196-285).
#include "cf_device.h"
#include "cf_kernels.h"

// This file is compiled as several translation units so that the layouts
// build in parallel: part 0 (this file on its own) holds the dispatchers and
// the reduction kernels, parts 1-7 (kernels_y_p<n>.hip) each instantiate the
// kernels of some (L, K) layouts; the list is at the end of the file.
#ifndef CF_Y_PART
#define CF_Y_PART 0
#endif

namespace cf {

constexpr int YNT = 256;  // threads per block of every y kernel

struct ModeInfo {
    long long s;
    int ix, iz, l;
    double kx, kz, kappa2;
    bool valid, nyq, mean;
};

__device__ __forceinline__ ModeInfo mode_info(const GridDims& gd, long long s) {
    ModeInfo mi;
    mi.s = s;
    mi.valid = s < gd.B;
    const unsigned ss = mi.valid ? (unsigned)s : 0u;  // B = Nxh*Nz < 2^31
    mi.iz = (int)(ss / (unsigned)gd.Nxh);
    mi.ix = (int)ss - mi.iz * gd.Nxh;
    mi.l = mi.iz <= (gd.Nz >> 1) ? mi.iz : mi.iz - gd.Nz;
    mi.kx = gd.alpha * mi.ix;
    mi.kz = gd.beta * mi.l;
    mi.kappa2 = mi.kx * mi.kx + mi.kz * mi.kz;
    mi.nyq = ((gd.Nx % 2 == 0) && mi.ix == gd.Nx / 2) || ((gd.Nz % 2 == 0) && mi.iz == gd.Nz / 2);
    mi.mean = (ss == 0);
    return mi;
}

// LDS slots (double2) of the per-degree table: CHEB_TAB_ARRAYS arrays of K L doubles
template <int L, int K>
__host__ __device__ constexpr int y_table_slots() { return CHEB_TAB_ARRAYS * K * L / 2; }

// ---------------------------------------------------------------------------
// Per-wavefront tile.  Wave w of block b owns the W = 64/L consecutive modes
// starting at b*M + w*W, and a private LDS region holding two FFT lines per
// mode (the DCT-I staging) that also serves as the solve's stash.  Spectral
// rows move between HBM and the group registers directly (lane t of a mode
// reads/writes degrees Kt..Kt+K-1; the W modes of one row are contiguous), so
// no block-level barrier is needed anywhere after the twiddle table load.
// ---------------------------------------------------------------------------
template <int L, int K>
struct Tile {
    static constexpr int M = YNT / L;  // modes per block
    static constexpr int W = 64 / L;   // modes per wavefront
    const GridDims& gd;
    long long s, sw0;  // this group's mode, the wave's first mode
    int lane, mw, ld;
    double2 *bufA, *bufB, *stash;
    const double2* tw;
    // region: LDS slots per wave (y_region_slots); smem starts with the
    // twiddle table of the DCT plan (ntw entries) and the per-degree table
    // (y_table_slots<L>), see load_tables
    const double* tab;
    // block: this block's index within its launch (or within its instance
    // of a batched ensemble launch)
    __device__ __forceinline__ Tile(const GridDims& g_, double2* smem, int ntw, int ld_, int region, unsigned block)
        : gd(g_), ld(ld_) {
        const int tid = threadIdx.x, wave = tid >> 6;
        lane = tid & 63;
        mw = lane / L;
        sw0 = (long long)block * M + wave * W;
        s = sw0 + mw;
        tw = smem;
        tab = reinterpret_cast<const double*>(smem + ntw);
        bufA = smem + ntw + y_table_slots<L, K>() + (size_t)wave * region;
        bufB = bufA + W * ld;
        stash = bufA + mw * 3 * K * L;
    }
    // Non-Nyquist-first order (YOrder): the wave's first slot sw0 becomes a
    // mode index.  W divides nxr, so the wave's W slots lie in one z row and
    // its modes stay consecutive; a wave past the last non-Nyquist mode is
    // empty (modes >= B: zero loads, no stores).
    __device__ __forceinline__ void order_modes(long long nreg, int nxr, int zskip) {
        if (sw0 >= nreg) {
            sw0 = gd.B;
        } else {
            const unsigned q = (unsigned)sw0 / (unsigned)nxr;
            const int iz = (int)q + ((int)q >= zskip ? 1 : 0);
            sw0 = (long long)iz * gd.Nxh + ((unsigned)sw0 - q * (unsigned)nxr);
        }
        s = sw0 + mw;
    }
    __device__ __forceinline__ void order_modes(const YOrder& o) { order_modes(o.nreg, o.nxr, o.zskip); }
    // spectral field column of this group's mode (zero beyond P / absent mode)
    __device__ __forceinline__ CVec<K> load(const double2* __restrict__ src, const Group<L, K>& g) const {
        CVec<K> v;
        const bool ok = s < gd.B;
#pragma unroll
        for (int e = 0; e < K; ++e) {
            int n = K * g.t + e;
            v.v[e] = (ok && n < gd.Ny) ? src[s + gd.B * n] : czero();
        }
        return v;
    }
    __device__ __forceinline__ void store(double2* __restrict__ dst, const Group<L, K>& g, const CVec<K>& v) const {
        const bool ok = s < gd.B;
#pragma unroll
        for (int e = 0; e < K; ++e) {
            int n = K * g.t + e;
            if (ok && n < gd.Ny) dst[s + gd.B * n] = v.v[e];
        }
    }
    // forward Chebyshev transform of y-physical rows (values at y_j, j=0..P):
    // REDFT00, times norm, endpoint 1/2, Nyquist modes zeroed (normalize_forward!).
    // SC: storage of the rows (double2, or float2 widened on load); the
    // transform itself runs in FP64.
    template <typename SC>
    __device__ __forceinline__ CVec<K> dct_forward(const SC* __restrict__ src, const FFTPlan& plan, double norm,
                                                const Group<L, K>& g, bool zero_mode) {
        const int P = gd.P;
        // YLOADS rows' HBM loads in flight before their LDS stores
        constexpr int YLOADS = 4;
        for (int base = lane; base < (P + 1) * W; base += YLOADS * 64) {
            SC r[YLOADS];
#pragma unroll
            for (int b = 0; b < YLOADS; ++b) {
                const int idx = base + b * 64;
                const int n = idx / W, m = idx - n * W;
                const long long sm = sw0 + m;
                r[b] = (idx < (P + 1) * W && sm < gd.B) ? src[sm + gd.B * n] : from_c64<SC>(czero());
            }
#pragma unroll
            for (int b = 0; b < YLOADS; ++b) {
                const int idx = base + b * 64;
                if (idx < (P + 1) * W) {
                    const int n = idx / W, m = idx - n * W;
                    const double2 v = to_c64(r[b]);
                    bufA[m * ld + n] = v;
                    if (n >= 1 && n <= P - 1) bufA[m * ld + 2 * P - n] = v;
                }
            }
        }
        wave_sync();
        double2* res = lds_fft<true, true>(bufA, bufB, plan, tw, W, ld, -1, lane, 64);
        CVec<K> v;
#pragma unroll
        for (int e = 0; e < K; ++e) {
            int n = K * g.t + e;
            double w = (n == 0 || n == P) ? 0.5 * norm : norm;
            v.v[e] = (n <= P && !zero_mode) ? cscale(res[mw * ld + n], w) : czero();
        }
        wave_sync();
        return v;
    }
    // inverse: coefficient registers -> values at y_j (endpoints doubled,
    // REDFT00, x1/2: InverseFFT!'s chebyshev + copy_to_padded! scale).
    // Computed in FP64; SC = float2 rounds the stored values.  INPLACE: one
    // LDS line per mode, transformed in place (8 points per lane).
    template <typename SC, bool INPLACE = false>
    __device__ __forceinline__ void dct_inverse(SC* __restrict__ dst, const FFTPlan& plan, const Group<L, K>& g,
                                                const CVec<K>& a) {
        const int P = gd.P;
#pragma unroll
        for (int e = 0; e < K; ++e) {
            int n = K * g.t + e;
            if (n <= P) {
                double2 x = a.v[e];
                bufA[mw * ld + n] = (n == 0 || n == P) ? cscale(x, 2.0) : x;
                if (n >= 1 && n <= P - 1) bufA[mw * ld + 2 * P - n] = x;
            }
        }
        wave_sync();
        double2* res = bufA;
        if constexpr (INPLACE) lds_fft_inplace<8, true>(bufA, plan, tw, W, ld, -1, lane, 64);
        else res = lds_fft<true, true>(bufA, bufB, plan, tw, W, ld, -1, lane, 64);
        for (int idx = lane; idx < (P + 1) * W; idx += 64) {
            int n = idx / W, m = idx - n * W;
            long long sm = sw0 + m;
            if (sm < gd.B) dst[sm + gd.B * n] = from_c64<SC>(cscale(res[m * ld + n], 0.5));
        }
        wave_sync();
    }
};

// LDS slots of one wave's region: two FFT lines per mode, or the solve's
// stash of three coefficient vectors (K L slots each) per mode.
template <int L, int K>
__host__ __device__ constexpr int y_region_slots(int ld) {
    return (2 * (64 / L) * ld > 3 * K * 64) ? 2 * (64 / L) * ld : 3 * K * 64;
}

// Block prologue of every y kernel: the DCT twiddle table and the
// per-degree table into LDS.
template <int L, int K>
__device__ __forceinline__ void load_tables(double2* smem, const YArgs& a) {
    for (int i = threadIdx.x; i < a.dct.n; i += YNT) smem[i] = a.dct.tw[i];
    double* tab = reinterpret_cast<double*>(smem + a.dct.n);
    constexpr int NT = K * L;
    for (int i = threadIdx.x; i < CHEB_TAB_ARRAYS * NT; i += YNT) {
        const int arr = i / NT, n = i - arr * NT;
        tab[i] = a.cheb_tab[arr * a.ntab + n];
    }
    __syncthreads();
}

// Lane-private LDS stash of a coefficient vector (each lane writes and reads
// back only its own four slots, so no barrier is needed between the two).
template <int L, int K>
__device__ __forceinline__ void stash(double2* slot, const Group<L, K>& g, const CVec<K>& v) {
#pragma unroll
    for (int e = 0; e < K; ++e) slot[K * g.t + e] = v.v[e];
}
template <int L, int K>
__device__ __forceinline__ CVec<K> unstash(const double2* slot, const Group<L, K>& g) {
    CVec<K> v;
#pragma unroll
    for (int e = 0; e < K; ++e) v.v[e] = slot[K * g.t + e];
    return v;
}

// coefficient of degree n (broadcast to the group)
template <int L, int K>
__device__ __forceinline__ double2 coef_at(const Group<L, K>& g, const CVec<K>& a, int n) {
    double2 v = czero();
#pragma unroll
    for (int e = 0; e < K; ++e)
        if (K * g.t + e == n) v = a.v[e];
    return g.sum(v);
}

template <int L, int K>
__device__ __forceinline__ void load_fac(const double* __restrict__ f, long long s, int Ny, const Group<L, K>& g,
                                         double out[K]) {
#pragma unroll
    for (int e = 0; e < K; ++e) {
        int n = K * g.t + e;
        out[e] = (n < Ny) ? f[s * Ny + n] : 0.0;
    }
}

// Total velocity V = U + Ub e_x at (0,0) and the y-physical fields each
// nonlinear form needs, at the Lobatto points, into out6:
//   rotating   V, Omega_x, Omega_z of Omega = curl V (nonlinear.jl:173-193,
//              operators.jl:172-190); out6[4] is not written: Omega_y =
//              dV_x/dz - dV_z/dx is formed by the periodic passes (the z
//              inverse emits i beta l V_x from V_x into z field 4, the x pass
//              subtracts i alpha k V_z), one Chebyshev transform fewer
//   convective V, dV/dy            (nonlinear.jl:113-140; x and z derivatives
//                                   are formed by the periodic passes)
//   divergence V                   (nonlinear.jl:146-167)
// get(c) returns U component c: from registers in the stage kernel, from HBM
// in the prologue (loaded where first needed, which shortens live ranges).
template <int L, int K, typename SC, bool INPLACE = false, typename GetU>
__device__ __forceinline__ void inverse_six(Tile<L, K>& tile, const YArgs& a, const Group<L, K>& g, const ModeInfo& mi,
                                            GetU get) {
    SC* out6[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) out6[i] = reinterpret_cast<SC*>(a.out6[i]);
    CVec<K> V0 = get(0);
    if (mi.mean && mi.valid) {
#pragma unroll
        for (int e = 0; e < K; ++e) {
            int n = K * g.t + e;
            if (n < a.gd.Ny) V0.v[e] = cadd(V0.v[e], cmk(a.base[n], 0.0));
        }
    }
    if (a.form != FORM_ROTATING) {
        tile.template dct_inverse<SC, INPLACE>(out6[0], a.dct, g, V0);
        if (a.form == FORM_CONVECTIVE) tile.template dct_inverse<SC, INPLACE>(out6[3], a.dct, g, cheb_deriv(g, V0));
        const CVec<K> V1 = get(1);
        tile.template dct_inverse<SC, INPLACE>(out6[1], a.dct, g, V1);
        if (a.form == FORM_CONVECTIVE) tile.template dct_inverse<SC, INPLACE>(out6[4], a.dct, g, cheb_deriv(g, V1));
        const CVec<K> V2 = get(2);
        tile.template dct_inverse<SC, INPLACE>(out6[2], a.dct, g, V2);
        if (a.form == FORM_CONVECTIVE) tile.template dct_inverse<SC, INPLACE>(out6[5], a.dct, g, cheb_deriv(g, V2));
        return;
    }
    // order chosen so each velocity component dies as early as possible
    const CVec<K> V1 = get(1);
    CVec<K> W;
    tile.template dct_inverse<SC, INPLACE>(out6[0], a.dct, g, V0);
    {
        CVec<K> d1 = cheb_deriv(g, V0);
#pragma unroll
        for (int e = 0; e < K; ++e) W.v[e] = csub(cmuli(cscale(V1.v[e], mi.kx)), d1.v[e]);  // i k V2 - dy V1
    }
    tile.template dct_inverse<SC, INPLACE>(out6[5], a.dct, g, W);
    const CVec<K> V2 = get(2);
    tile.template dct_inverse<SC, INPLACE>(out6[2], a.dct, g, V2);
    {
        CVec<K> d3 = cheb_deriv(g, V2);
#pragma unroll
        for (int e = 0; e < K; ++e) W.v[e] = csub(d3.v[e], cmuli(cscale(V1.v[e], mi.kz)));  // dy V3 - i l V2
    }
    tile.template dct_inverse<SC, INPLACE>(out6[3], a.dct, g, W);
    tile.template dct_inverse<SC, INPLACE>(out6[1], a.dct, g, V1);
}

// ---------------------------------------------------------------------------
// Stokes solve of one mode given R (re-solve form of the influence + tau
// correction; see the derivation in the library documentation).  Returns
// U[3], P.  mean/nyq modes handled separately by the caller.
// ---------------------------------------------------------------------------
template <int L, int K>
__device__ __forceinline__ void solve_nonzero(const Group<L, K>& g, const YArgs& a, const Bank& bk, const ModeInfo& mi,
                                              const double2* Rs0, const double2* Rs1, const double2* Rs2,
                                              CVec<K> Uout[3], CVec<K>& Pout) {
    const int P = a.gd.P, Ny = a.gd.Ny;
    const double* mc = bk.mconst + mi.s * MC_COUNT;
    const double shift = bk.lambda + bk.nu * mi.kappa2;
    double fp[K], fv[K];
    load_fac(a.facP, mi.s, Ny, g, fp);
    // r = D Ry + i k Rx + i l Rz
    CVec<K> r;
    {
        CVec<K> R1 = unstash(Rs1, g);
        r = cheb_deriv(g, R1);
        CVec<K> R0 = unstash(Rs0, g), R2 = unstash(Rs2, g);
#pragma unroll
        for (int e = 0; e < K; ++e)
            r.v[e] = cadd(r.v[e], cadd(cmuli(cscale(R0.v[e], mi.kx)), cmuli(cscale(R2.v[e], mi.kz))));
    }
    double2 p1P, v1a, v1b, bp, bm;
    {
        CVec<K> p1 = helm_solve(g, P, mi.kappa2, fp, false, r, czero(), czero());
        p1P = coef_at(g, p1, P);
        CVec<K> q = cheb_deriv(g, p1);
        CVec<K> R1 = unstash(Rs1, g);
#pragma unroll
        for (int e = 0; e < K; ++e) q.v[e] = csub(q.v[e], R1.v[e]);
        load_fac(bk.facV, mi.s, Ny, g, fv);
        CVec<K> v1 = helm_solve(g, P, shift, fv, false, q, czero(), czero());
        bp = cscale(wall_deriv_plus(g, v1), -1.0);
        bm = cscale(wall_deriv_minus(g, v1), -1.0);
        v1a = coef_at(g, v1, P - 1);
        v1b = coef_at(g, v1, P);
    }
    double2 dp = cadd(cscale(bp, mc[MC_A11]), cscale(bm, mc[MC_A12]));
    double2 dm = cadd(cscale(bp, mc[MC_A21]), cscale(bm, mc[MC_A22]));
    double2 ra, rb;
    {
        CVec<K> R1 = unstash(Rs1, g);
        ra = coef_at(g, R1, P - 1);
        rb = coef_at(g, R1, P);
    }
    double2 p2P = cadd(p1P, cadd(cscale(dp, mc[MC_PPT]), cscale(dm, mc[MC_PMT])));
    double2 v2a = cadd(v1a, cadd(cscale(dp, mc[MC_VP1]), cscale(dm, mc[MC_VM1])));
    double2 v2b = cadd(v1b, cadd(cscale(dp, mc[MC_VP0]), cscale(dm, mc[MC_VM0])));
    double2 s1 = cscale(csub(cadd(cscale(v2a, shift), cscale(p2P, 2.0 * P)), ra), 1.0 / (1.0 - mc[MC_SIG1]));
    double2 s2 = cscale(csub(cscale(v2b, shift), rb), 1.0 / (1.0 - mc[MC_SIG2]));
    // final pressure: rhs r + s1 D T_{P-1} + s2 D T_P, wall data per parity
#pragma unroll
    for (int e = 0; e < K; ++e) {
        int n = K * g.t + e;
        if (n % 2 == 0) {
            if (n < P - 1) r.v[e] = cfma(n == 0 ? (double)(P - 1) : 2.0 * (P - 1), s1, r.v[e]);
        } else {
            if (n < P) r.v[e] = cfma(2.0 * P, s2, r.v[e]);
        }
    }
    double2 de = cadd(cscale(cadd(dp, dm), 0.5), cscale(s1, 0.5 * (mc[MC_D0P] + mc[MC_D0M])));
    double2 dd = cadd(cscale(csub(dp, dm), 0.5), cscale(s2, 0.5 * (mc[MC_D0P] - mc[MC_D0M])));
    Pout = helm_solve(g, P, mi.kappa2, fp, false, r, de, dd);
    __builtin_amdgcn_sched_barrier(0);
    // the three velocity solves share the operator: one fused solve
    CVec<K> q[3];
    {
        const CVec<K> R0 = unstash(Rs0, g), R2 = unstash(Rs2, g);
#pragma unroll
        for (int e = 0; e < K; ++e) {
            q[0].v[e] = csub(cmuli(cscale(Pout.v[e], mi.kx)), R0.v[e]);
            q[2].v[e] = csub(cmuli(cscale(Pout.v[e], mi.kz)), R2.v[e]);
        }
    }
    q[1] = cheb_deriv(g, Pout);
    {
        const CVec<K> R1 = unstash(Rs1, g);
#pragma unroll
        for (int e = 0; e < K; ++e) q[1].v[e] = csub(q[1].v[e], R1.v[e]);
    }
    const double2 zero3[3] = {czero(), czero(), czero()};
    helm_solve_n<3>(g, P, shift, fv, false, q, zero3, zero3, Uout);
}

// mean mode (meanmode.jl:103-150); returns the new gradients
template <int L, int K>
__device__ __forceinline__ void solve_mean(const Group<L, K>& g, const YArgs& a, const Bank& bk, const double2* Rs0,
                                           const double2* Rs1, const double2* Rs2, CVec<K> Uout[3], CVec<K>& Pout,
                                           double gout[2]) {
    const int P = a.gd.P, Ny = a.gd.Ny;
    const CVec<K> R[3] = {unstash(Rs0, g), unstash(Rs1, g), unstash(Rs2, g)};
    // pressure: integrate Ry through degree P-1, zero bulk mean
    {
        double2 prev3 = g.up1(R[1].v[K - 1]);
        double2 next0 = g.down1(R[1].v[0]);
        CVec<K> r = R[1];
#pragma unroll
        for (int e = 0; e < K; ++e) {
            int n = K * g.t + e;
            double2 rm = (e == 0) ? prev3 : r.v[e - 1];  // r_{n-1}
            double2 rp = (e == K - 1) ? next0 : r.v[e + 1];  // r_{n+1}
            double2 lower = (n == 1) ? rm : cscale(rm, 0.5);
            double2 upper = (n + 1 < P) ? cscale(rp, 0.5) : czero();
            Pout.v[e] = (n >= 1 && n <= P) ? cscale(csub(lower, upper), 1.0 / n) : czero();
        }
        double2 m = bulk_mean(g, Pout);  // p_0 = 0 at this point
        if (g.t == 0) Pout.v[0] = cscale(m, -1.0);
    }
#pragma unroll
    for (int e = 0; e < K; ++e) Uout[1].v[e] = czero();
    double fv[K];
    load_fac(bk.facV, 0, Ny, g, fv);
    for (int c = 0; c < 2; ++c) {
        const CVec<K>& src = R[c == 0 ? 0 : 2];
        CVec<K> f;
#pragma unroll
        for (int e = 0; e < K; ++e) f.v[e] = cscale(src.v[e], -1.0);
        CVec<K> u = helm_solve(g, P, bk.lambda, fv, false, f, czero(), czero());
        double grad;
        if (a.bulk) {
            double target = c == 0 ? a.c0 : a.c1;
            grad = (target - bulk_mean(g, u).x) / bk.resp_mean;
        } else {
            grad = c == 0 ? a.c0 : a.c1;
        }
#pragma unroll
        for (int e = 0; e < K; ++e) {
            int n = K * g.t + e;
            if (n < Ny) u.v[e] = cadd(u.v[e], cmk(grad * bk.resp[n], 0.0));
        }
        Uout[c == 0 ? 0 : 2] = u;
        gout[c] = grad;
    }
}

template <int L, int K>
__device__ __forceinline__ void zero_vec(CVec<K>& v) {
#pragma unroll
    for (int e = 0; e < K; ++e) v.v[e] = czero();
}

// Laplacian D^2 a - kappa^2 a
template <int L, int K>
__device__ __forceinline__ CVec<K> laplacian(const Group<L, K>& g, const CVec<K>& a, double kappa2) {
    CVec<K> d = cheb_deriv(g, cheb_deriv(g, a));
#pragma unroll
    for (int e = 0; e < K; ++e) d.v[e] = cfma(-kappa2, a.v[e], d.v[e]);
    return d;
}

// ---------------------------------------------------------------------------
// Full CNRK2 stage: forward DCT of the nonlinear term, Q update, RHS,
// Stokes solve, state write, inverse transforms for the next stage.
// ---------------------------------------------------------------------------
// SC: storage of the y-physical fields (double2; float2 in the mixed step).
template <int L, int K, typename SC>
__device__ __forceinline__ void ystage_body(const YArgs& a, unsigned block) {
    extern __shared__ double2 smem[];
    load_tables<L, K>(smem, a);
    Tile<L, K> tile(a.gd, smem, a.dct.n, a.ld, y_region_slots<L, K>(a.ld), block);
    if (a.order.on) tile.order_modes(a.order);
    Group<L, K> g(tile.tab);
    ModeInfo mi = mode_info(a.gd, tile.s);
    const int Ny = a.gd.Ny;
    const bool meanmode = mi.mean && mi.valid;

    // Per component: forward DCT of the nonlinear term, Q = N (stage 1) or
    // A_j Q + N, R_c = lambda U_c + nu Lap U_c - d_c P + (B_j/C_j) Q_c.
    double gold0 = a.c0, gold1 = a.c1;
    CVec<K> Pold = tile.load(a.P, g);
    CVec<K> R[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        CVec<K> Qc;
        if (a.Ncoef[c]) {
            Qc = tile.load(a.Ncoef[c], g);
        } else {
            Qc = tile.dct_forward(reinterpret_cast<const SC*>(a.Nphys[c]), a.dct, a.norm, g, mi.nyq);
#pragma unroll
            for (int e = 0; e < K; ++e) Qc.v[e] = cscale(Qc.v[e], a.nl_sign);
        }
        if (a.Fcoef[c]) {  // forcing(tstage, U, N): N += F (cnrk2.jl:175)
            const CVec<K> F = tile.load(a.Fcoef[c], g);
#pragma unroll
            for (int e = 0; e < K; ++e) Qc.v[e] = cadd(Qc.v[e], F.v[e]);
        }
        double bN = 0.0;
        if (a.bulk && meanmode && c != 1) bN = bulk_mean(g, Qc).x;
        if (a.stage != 0) {
            CVec<K> Qo = tile.load(a.Q[c], g);
#pragma unroll
            for (int e = 0; e < K; ++e) Qc.v[e] = cfma(a.Aj, Qo.v[e], Qc.v[e]);
        }
        if (a.store_Q) tile.store(a.Q[c], g, Qc);
        CVec<K> U = tile.load(a.U[c], g);
        if (a.bulk && meanmode && c != 1) {  // _bulkpressuregradient, cnrk2.jl:98-107
            double2 sh = csub(wall_deriv_plus(g, U), wall_deriv_minus(g, U));
            double gv = a.nu * sh.x / 2.0 + (c == 0 ? a.nu * a.curv_mean : 0.0) + bN;
            if (c == 0) gold0 = gv; else gold1 = gv;
        }
        CVec<K> lap = laplacian(g, U, mi.kappa2);
        CVec<K> dP;
        if (c == 1) {
            dP = cheb_deriv(g, Pold);
        } else {
            double kk = (c == 0) ? mi.kx : mi.kz;
#pragma unroll
            for (int e = 0; e < K; ++e) dP.v[e] = cmuli(cscale(Pold.v[e], kk));
        }
#pragma unroll
        for (int e = 0; e < K; ++e) {
            double2 x = cscale(U.v[e], a.lambda);
            x = cfma(a.nu, lap.v[e], x);
            x = csub(x, dP.v[e]);
            R[c].v[e] = cfma(a.weight, Qc.v[e], x);
            // materialise R_c here: otherwise the RHS arithmetic is sunk into
            // the solve branches below and every input of all three
            // components stays live until then
            asm volatile("" : "+v"(R[c].v[e].x), "+v"(R[c].v[e].y));
        }
    }
    if (meanmode) {
#pragma unroll
        for (int e = 0; e < K; ++e) {
            int n = K * g.t + e;
            if (n < Ny) R[0].v[e] = cadd(R[0].v[e], cmk(2.0 * a.nu * a.curv[n], 0.0));
        }
        if (g.t == 0) {
            R[0].v[0] = csub(R[0].v[0], cmk(gold0, 0.0));
            R[2].v[0] = csub(R[2].v[0], cmk(gold1, 0.0));
        }
    }
    double2* st = tile.stash;
    stash(st, g, R[0]);
    stash(st + K * L, g, R[1]);
    stash(st + 2 * K * L, g, R[2]);
    CVec<K> Un[3], Pn;
    if (meanmode) {
        double gnew[2] = {0.0, 0.0};
        solve_mean(g, a, a.bank, st, st + K * L, st + 2 * K * L, Un, Pn, gnew);
        if (g.t == 0 && a.grad_out) {
            a.grad_out[0] = gnew[0];
            a.grad_out[1] = gnew[1];
        }
    } else if (mi.nyq || !mi.valid) {
#pragma unroll
        for (int c = 0; c < 3; ++c) zero_vec<L, K>(Un[c]);
        zero_vec<L, K>(Pn);
    } else {
        solve_nonzero(g, a, a.bank, mi, st, st + K * L, st + 2 * K * L, Un, Pn);
    }
    wave_sync();  // the stash aliases the wave's FFT lines reused below
    tile.store(a.U[0], g, Un[0]);
    tile.store(a.U[1], g, Un[1]);
    tile.store(a.U[2], g, Un[2]);
    tile.store(a.P, g, Pn);
    if (a.out6[0]) inverse_six<L, K, SC>(tile, a, g, mi, [&](int c) { return Un[c]; });
}

// Nyquist modes of a stage under the non-Nyquist-first order, elementwise
// (blk-th of order.nyq_blocks blocks).  The result equals ystage_body's for
// such a mode: the transformed nonlinear term is zero there
// (normalize_forward!), so Q = A_j Q_old + nl_sign*0 (+ F), and U, P and the
// y-physical outputs are zero.  The x Nyquist column ix = Nx/2 (every iz)
// comes first, then the z Nyquist row iz = Nz/2 (ix < nxr).
template <typename SC>
__device__ __forceinline__ void ystage_nyquist(const YArgs& a, unsigned blk) {
    const YOrder& o = a.order;
    const GridDims& gd = a.gd;
    const unsigned nnyq = (unsigned)o.nnyq;
    const unsigned total = nnyq * (unsigned)gd.Ny;
    const unsigned ncol = (gd.Nx % 2 == 0) ? (unsigned)gd.Nz : 0u;
    const bool six = a.out6[0] != nullptr, vort = a.form != FORM_DIVERGENCE;
    for (unsigned e = blk * YNT + threadIdx.x; e < total; e += o.nyq_blocks * YNT) {
        const unsigned n = e / nnyq, q = e - n * nnyq;
        const long long s = q < ncol ? (long long)q * gd.Nxh + gd.Nx / 2 : (long long)(gd.Nz / 2) * gd.Nxh + (q - ncol);
        const long long i = s + gd.B * n;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double2 Qc = a.Ncoef[c] ? a.Ncoef[c][i] : cscale(czero(), a.nl_sign);
            if (a.Fcoef[c]) Qc = cadd(Qc, a.Fcoef[c][i]);
            if (a.stage != 0) Qc = cfma(a.Aj, a.Q[c][i], Qc);
            if (a.store_Q) a.Q[c][i] = Qc;
            a.U[c][i] = czero();
        }
        a.P[i] = czero();
        if (six) {
#pragma unroll
            for (int f = 0; f < 6; ++f)
                if (f < 3 || vort) reinterpret_cast<SC*>(a.out6[f])[i] = from_c64<SC>(czero());
        }
    }
}

template <int L, int K, typename SC>
// two waves per SIMD (<= 256 registers per lane): one wave per SIMD leaves
// the wave-serial scans and LDS FFTs unhidden.  K = 6 and 8 (Ny above 256)
// keep 1.5 and 2 times the coefficients per lane and run one wave per SIMD,
// whose larger register budget holds them without scratch (at two waves
// K = 6 spills 160 bytes per lane).
__global__ void __launch_bounds__(YNT) __attribute__((amdgpu_waves_per_eu(K >= 6 ? 1 : 2))) ystage_kernel(YArgs a) {
    if (a.order.on) {  // block-uniform branches
        if (blockIdx.x >= a.order.nyq_first) {
            ystage_nyquist<SC>(a, blockIdx.x - a.order.nyq_first);
            return;
        }
        // z rows of Nxh modes leave the blocks' row segments unaligned to
        // cache lines: neighbouring blocks share lines, so they run on one XCD
        ystage_body<L, K, SC>(a, a.order.xcd ? xcd_logical(blockIdx.x, a.order.nyq_first) : blockIdx.x);
        return;
    }
    ystage_body<L, K, SC>(a, blockIdx.x);
}

// prologue: state U -> six y-physical fields (velocity and vorticity)
template <int L, int K, typename SC>
__device__ __forceinline__ void yprologue_body(const YArgs& a, unsigned block) {
    extern __shared__ double2 smem[];
    load_tables<L, K>(smem, a);
    Tile<L, K> tile(a.gd, smem, a.dct.n, a.ld, y_region_slots<L, K>(a.ld), block);
    Group<L, K> g(tile.tab);
    ModeInfo mi = mode_info(a.gd, tile.s);
    inverse_six<L, K, SC>(tile, a, g, mi, [&](int c) { return tile.load(a.U[c], g); });
}

// The single-problem prologue keeps its own copy of the body: passing the
// byval kernel argument by reference to yprologue_body makes the compiler
// materialise part of it in scratch memory.  Under the non-Nyquist-first
// order (a.order.on) the launch has only the regular blocks: the periodic
// passes never read the Nyquist modes' y-physical values, so none are
// written.
// INPLACE: the DCT lines are transformed in place (one line per mode, half
// the LDS of the two-buffer form, so more blocks per CU); the host picks it
// where the plan fits 8 points per lane (fft_inplace_fits)
template <int L, int K, typename SC, bool INPLACE = false>
__global__ void __launch_bounds__(YNT) yprologue_kernel(YArgs a) {
    extern __shared__ double2 smem[];
    load_tables<L, K>(smem, a);
    // the order's fields are read one by one (a reference to the byval
    // argument, or a copy of the whole struct, lands in scratch)
    const bool ordered = a.order.on != 0;
    const unsigned block = (ordered && a.order.xcd) ? xcd_logical(blockIdx.x, gridDim.x) : blockIdx.x;
    Tile<L, K> tile(a.gd, smem, a.dct.n, a.ld, INPLACE ? (64 / L) * a.ld : y_region_slots<L, K>(a.ld), block);
    if (ordered) tile.order_modes(a.order.nreg, a.order.nxr, a.order.zskip);
    Group<L, K> g(tile.tab);
    ModeInfo mi = mode_info(a.gd, tile.s);
    if constexpr (L == 4) {  // all loads first: the 4-lane build stays within 128 registers so
        CVec<K> U[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) U[c] = tile.load(a.U[c], g);
        inverse_six<L, K, SC, INPLACE>(tile, a, g, mi, [&](int c) { return U[c]; });
    } else {
        inverse_six<L, K, SC, INPLACE>(tile, a, g, mi, [&](int c) { return tile.load(a.U[c], g); });
    }
}

// Batched ensemble launches: instance i owns blocks [start[i], start[i+1])
// of the grid and its own argument block tab[i] (grid, factors, fields).
// All instances of one launch share the group size L.
template <int L, int K>
__global__ void __launch_bounds__(YNT) __attribute__((amdgpu_waves_per_eu(2)))
ystage_batch_kernel(const YArgs* __restrict__ tab, const unsigned* __restrict__ start, int n) {
    const int i = batch_instance(start, n, blockIdx.x);
    ystage_body<L, K, double2>(tab[i], blockIdx.x - start[i]);
}

template <int L, int K>
__global__ void __launch_bounds__(YNT) yprologue_batch_kernel(const YArgs* __restrict__ tab,
                                                              const unsigned* __restrict__ start, int n) {
    const int i = batch_instance(start, n, blockIdx.x);
    yprologue_body<L, K, double2>(tab[i], blockIdx.x - start[i]);
}

// Stokes solve! of a bank: R (3 spectral) -> U (3), P (stokes.jl:81-108)
template <int L, int K>
__global__ void __launch_bounds__(YNT) ysolve_kernel(YArgs a) {
    extern __shared__ double2 smem[];
    load_tables<L, K>(smem, a);
    Tile<L, K> tile(a.gd, smem, a.dct.n, a.ld, y_region_slots<L, K>(a.ld), blockIdx.x);
    Group<L, K> g(tile.tab);
    ModeInfo mi = mode_info(a.gd, tile.s);
    CVec<K> R[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) R[c] = tile.load(a.Rin[c], g);
    CVec<K> Un[3], Pn;
    double gnew[2] = {0.0, 0.0};
    double2* st = tile.stash;
    stash(st, g, R[0]);
    stash(st + K * L, g, R[1]);
    stash(st + 2 * K * L, g, R[2]);
    if (mi.mean && mi.valid) {
        solve_mean(g, a, a.bank, st, st + K * L, st + 2 * K * L, Un, Pn, gnew);
        if (g.t == 0 && a.grad_out) {
            a.grad_out[0] = gnew[0];
            a.grad_out[1] = gnew[1];
        }
    } else if (mi.nyq || !mi.valid) {
        for (int c = 0; c < 3; ++c) zero_vec<L, K>(Un[c]);
        zero_vec<L, K>(Pn);
    } else {
        solve_nonzero(g, a, a.bank, mi, st, st + K * L, st + 2 * K * L, Un, Pn);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) tile.store(a.U[c], g, Un[c]);
    tile.store(a.P, g, Pn);
}

// R = -Laplacian(U) (projection source, pressure.jl:125)
template <int L, int K>
__global__ void __launch_bounds__(YNT) yneglap_kernel(YArgs a) {
    extern __shared__ double2 smem[];
    load_tables<L, K>(smem, a);
    Tile<L, K> tile(a.gd, smem, a.dct.n, a.ld, y_region_slots<L, K>(a.ld), blockIdx.x);
    Group<L, K> g(tile.tab);
    ModeInfo mi = mode_info(a.gd, tile.s);
    for (int c = 0; c < 3; ++c) {
        CVec<K> U = tile.load(a.U[c], g);
        CVec<K> lap = laplacian(g, U, mi.kappa2);
#pragma unroll
        for (int e = 0; e < K; ++e) lap.v[e] = cscale(lap.v[e], -1.0);
        tile.store(a.Rout[c], g, lap);
    }
}

// pressure reconstruction (_pressure_poisson, pressure.jl:196-285):
// A (3 spectral accelerations), v = U[1] -> P
template <int L, int K>
__global__ void __launch_bounds__(YNT) ypressure_kernel(YArgs a) {
    extern __shared__ double2 smem[];
    load_tables<L, K>(smem, a);
    Tile<L, K> tile(a.gd, smem, a.dct.n, a.ld, y_region_slots<L, K>(a.ld), blockIdx.x);
    Group<L, K> g(tile.tab);
    ModeInfo mi = mode_info(a.gd, tile.s);
    const int P = a.gd.P, Ny = a.gd.Ny;
    CVec<K> A[3];
    for (int c = 0; c < 3; ++c) A[c] = tile.load(a.Rin[c], g);
    CVec<K> v = tile.load(a.U[1], g);
    CVec<K> Pn;
    if (mi.mean && mi.valid) {
        double2 prev3 = g.up1(A[1].v[K - 1]);
        double2 next0 = g.down1(A[1].v[0]);
#pragma unroll
        for (int e = 0; e < K; ++e) {
            int n = K * g.t + e;
            double2 rm = (e == 0) ? prev3 : A[1].v[e - 1];
            double2 rp = (e == K - 1) ? next0 : A[1].v[e + 1];
            double2 lower = (n == 1) ? rm : cscale(rm, 0.5);
            double2 upper = (n + 1 < Ny - 1) ? cscale(rp, 0.5) : czero();
            Pn.v[e] = (n >= 1 && n <= P) ? cscale(csub(lower, upper), 1.0 / n) : czero();
        }
        double2 m = bulk_mean(g, Pn);
        if (g.t == 0) Pn.v[0] = cscale(m, -1.0);
    } else if (mi.nyq || !mi.valid) {
        zero_vec<L, K>(Pn);
    } else {
        // div A
        CVec<K> r = cheb_deriv(g, A[1]);
#pragma unroll
        for (int e = 0; e < K; ++e)
            r.v[e] = cadd(r.v[e], cadd(cmuli(cscale(A[0].v[e], mi.kx)), cmuli(cscale(A[2].v[e], mi.kz))));
        // wall data P_y(+-1) = A_y(+-1) + nu v_yy(+-1)
        double2 up = czero(), lo = czero();
#pragma unroll
        for (int e = 0; e < K; ++e) {
            int n = K * g.t + e;
            if (n < Ny) {
                double nn = n;
                double2 term = cfma(a.nu * (nn * nn * (nn * nn - 1.0) / 3.0), v.v[e], A[1].v[e]);
                up = cadd(up, term);
                lo = (n % 2 == 0) ? cadd(lo, term) : csub(lo, term);
            }
        }
        up = g.sum(up);
        lo = g.sum(lo);
        double fp[K];
        load_fac(a.facPN, mi.s, Ny, g, fp);
        // Neumann parity data: even (g+ - g-)/2, odd (g+ + g-)/2
        Pn = helm_solve(g, P, mi.kappa2, fp, true, r, cscale(csub(up, lo), 0.5), cscale(cadd(up, lo), 0.5));
    }
    tile.store(a.P, g, Pn);
}

// Batched Chebyshev transforms of spectral fields (no group logic).
// dir < 0: y-physical values -> normalised coefficients (normalize_forward!)
// dir > 0: coefficients -> values at the Lobatto points
template <int L, int K>
__global__ void __launch_bounds__(YNT) ydct_kernel(YArgs a, int dir) {
    extern __shared__ double2 smem[];
    load_tables<L, K>(smem, a);
    Tile<L, K> tile(a.gd, smem, a.dct.n, a.ld, y_region_slots<L, K>(a.ld), blockIdx.x);
    Group<L, K> g(tile.tab);
    ModeInfo mi = mode_info(a.gd, tile.s);
    for (int c = 0; c < a.nfields; ++c) {
        if (dir < 0) {
            CVec<K> v = tile.dct_forward(a.Nphys[c], a.dct, a.norm, g, mi.nyq);
#pragma unroll
            for (int e = 0; e < K; ++e) v.v[e] = cscale(v.v[e], a.nl_sign);
            tile.store(a.Rout[c], g, v);
        } else {
            CVec<K> v = tile.load(a.Rin[c], g);
            tile.dct_inverse(a.out6[c], a.dct, g, v);
        }
    }
}

// Divergence form (nonlinear.jl:146-167, div! of the flux tensor): the six
// y-physical products T = (VxVx, VxVy, VxVz, VyVy, VyVz, VzVz) are forward
// transformed and N_i = -(i k T_ix + d/dy T_iy + i l T_iz) is written to Rout.
template <int L, int K>
__global__ void __launch_bounds__(YNT) ydivergence_kernel(YArgs a) {
    extern __shared__ double2 smem[];
    load_tables<L, K>(smem, a);
    Tile<L, K> tile(a.gd, smem, a.dct.n, a.ld, y_region_slots<L, K>(a.ld), blockIdx.x);
    Group<L, K> g(tile.tab);
    ModeInfo mi = mode_info(a.gd, tile.s);
    CVec<K> N[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) zero_vec<L, K>(N[c]);
    // tensor entry (i, j) for products p = 0..5
    const int pi[6] = {0, 0, 0, 1, 1, 2}, pj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int p = 0; p < 6; ++p) {
        CVec<K> T = tile.dct_forward(a.Nphys[p], a.dct, a.norm, g, mi.nyq);
        CVec<K> dT = cheb_deriv(g, T);
        // contribution to N_i from d_j T_ij and, off the diagonal, to N_j from d_i T_ij
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            if (side == 1 && pi[p] == pj[p]) continue;
            const int i = side == 0 ? pi[p] : pj[p], j = side == 0 ? pj[p] : pi[p];
#pragma unroll
            for (int e = 0; e < K; ++e) {
                double2 d = (j == 1) ? dT.v[e] : cmuli(cscale(T.v[e], j == 0 ? mi.kx : mi.kz));
                N[i].v[e] = csub(N[i].v[e], d);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) tile.store(a.Rout[c], g, N[c]);
}

// Volume inner products (postprocessing.jl:30-60, 87-91).  Per mode,
// Re(u^H M v) with the exact Chebyshev mass matrix
// M_mn = (mu(m+n) + mu(|m-n|)) / 2, mu(k) = 1/(1-k^2) for even k (0 for odd),
// weighted 1 for k = 0 and 2 for k > 0 (conjugate pairs), Nyquist planes
// excluded.  Y_DOT: sum over the three components of U (U) and V (Rin).
// Y_DISSIPATION: u = v = curl U (operators.jl:172-190); the caller scales by nu.
// Each block writes its partial sum to partial[blockIdx.x].
template <int L, int K>
__global__ void __launch_bounds__(YNT) ydot_kernel(YArgs a, int curl) {
    extern __shared__ double2 smem[];
    load_tables<L, K>(smem, a);
    Tile<L, K> tile(a.gd, smem, a.dct.n, a.ld, y_region_slots<L, K>(a.ld), blockIdx.x);
    Group<L, K> g(tile.tab);
    ModeInfo mi = mode_info(a.gd, tile.s);
    const int Ny = a.gd.Ny;
    double2* vline = tile.stash;  // this group's copy of v, indexed by degree
    CVec<K> W0, W1, W2;
    if (curl) {
        CVec<K> U0 = tile.load(a.U[0], g), U1 = tile.load(a.U[1], g), U2 = tile.load(a.U[2], g);
        CVec<K> d2 = cheb_deriv(g, U2), d0 = cheb_deriv(g, U0);
#pragma unroll
        for (int e = 0; e < K; ++e) {
            W0.v[e] = csub(d2.v[e], cmuli(cscale(U1.v[e], mi.kz)));
            W1.v[e] = cmuli(csub(cscale(U0.v[e], mi.kz), cscale(U2.v[e], mi.kx)));
            W2.v[e] = csub(cmuli(cscale(U1.v[e], mi.kx)), d0.v[e]);
        }
    }
    double acc = 0.0;
#pragma unroll 1
    for (int c = 0; c < 3; ++c) {
        CVec<K> u, v;
        if (curl) {
            u = (c == 0) ? W0 : (c == 1) ? W1 : W2;
            v = u;
        } else {
            u = tile.load(a.U[c], g);
            v = tile.load(a.Rin[c], g);
        }
        stash(vline, g, v);
        wave_sync();
#pragma unroll
        for (int e = 0; e < K; ++e) {
            const int m = K * g.t + e;
            double2 s = czero();
            if (m < Ny)
                for (int n = m & 1; n < Ny; n += 2) {
                    const double kp = m + n, km = m - n;
                    const double Mmn = 0.5 * (1.0 / (1.0 - kp * kp) + 1.0 / (1.0 - km * km));
                    s = cfma(Mmn, vline[n], s);
                }
            acc += u.v[e].x * s.x + u.v[e].y * s.y;  // Re(conj(u_m) (M v)_m)
        }
        wave_sync();
    }
    acc = g.sum(acc);
    const bool use = mi.valid && !mi.nyq;
    double* part = reinterpret_cast<double*>(smem);  // twiddle area is free now
    __syncthreads();
    constexpr int M = YNT / L;
    const int mloc = threadIdx.x / L;
    if (g.t == 0) part[mloc] = use ? (mi.ix == 0 ? 1.0 : 2.0) * acc : 0.0;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = 0.0;
        for (int i = 0; i < M; ++i) b += part[i];
        a.partial[blockIdx.x] = b;
    }
}

// Plane statistics of the TOTAL velocity V = U + Ub e_x (examples/mkm590/
// statistics.jl:26-43) at the Lobatto points.  By Parseval on the padded
// grid, the x-z plane mean is the (0,0) mode's y-physical value and the
// plane covariance of a, b is sum over the other modes of
// w_k Re(a(k,l,y) conj(b(k,l,y))) with w = 1 (k = 0), 2 (k > 0), Nyquist
// planes excluded.  partial[block * Ny * 9 + y * 9 + q] (PLANE_STATS_Q
// quantities: means u v w, then uu vv ww uv uw vw).
template <int L, int K>
__global__ void __launch_bounds__(YNT) ystats_kernel(YArgs a) {
    extern __shared__ double2 smem[];
    load_tables<L, K>(smem, a);
    Tile<L, K> tile(a.gd, smem, a.dct.n, a.ld, y_region_slots<L, K>(a.ld), blockIdx.x);
    Group<L, K> g(tile.tab);
    ModeInfo mi = mode_info(a.gd, tile.s);
    const int Ny = a.gd.Ny, P = a.gd.P;
    constexpr int M = YNT / L, W = 64 / L, Q = PLANE_STATS_Q;
    double2 vals[3][K];  // values at y = Kt+e (group layout after the inverse transform)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        CVec<K> U = tile.load(a.U[c], g);
        if (c == 0 && mi.mean && mi.valid) {
#pragma unroll
            for (int e = 0; e < K; ++e) {
                int n = K * g.t + e;
                if (n < Ny) U.v[e] = cadd(U.v[e], cmk(a.base[n], 0.0));
            }
        }
        // inverse Chebyshev transform in the wave's lines, result read back in group layout
#pragma unroll
        for (int e = 0; e < K; ++e) {
            int n = K * g.t + e;
            if (n <= P) {
                double2 x = U.v[e];
                tile.bufA[tile.mw * tile.ld + n] = (n == 0 || n == P) ? cscale(x, 2.0) : x;
                if (n >= 1 && n <= P - 1) tile.bufA[tile.mw * tile.ld + 2 * P - n] = x;
            }
        }
        wave_sync();
        double2* res = lds_fft<true>(tile.bufA, tile.bufB, a.dct, tile.tw, W, tile.ld, -1, tile.lane, 64);
#pragma unroll
        for (int e = 0; e < K; ++e) {
            int n = K * g.t + e;
            vals[c][e] = (n <= P) ? cscale(res[tile.mw * tile.ld + n], 0.5) : czero();
        }
        wave_sync();
    }
    // this mode's contributions at each y of the lane: 9 quantities x K points
    const bool use = mi.valid && !mi.nyq;
    const double wk = (mi.ix == 0) ? 1.0 : 2.0;
    const int pi[6] = {0, 1, 2, 0, 0, 1}, pj[6] = {0, 1, 2, 1, 2, 2};
    __syncthreads();  // tables and wave lines are free: reuse LDS as the block's [Ny][Q] accumulator
    double* red = reinterpret_cast<double*>(smem);
    for (int idx = threadIdx.x; idx < Ny * Q; idx += YNT) red[idx] = 0.0;
    __syncthreads();
    const int mloc = threadIdx.x / L;
    // the block's modes add in a fixed order (deterministic sums)
    for (int m = 0; m < M; ++m) {
        if (mloc == m) {
#pragma unroll
            for (int e = 0; e < K; ++e) {
                const int y = K * g.t + e;
                if (y >= Ny) continue;
#pragma unroll
                for (int c = 0; c < 3; ++c) red[y * Q + c] += (use && mi.mean) ? vals[c][e].x : 0.0;
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    const double2 u = vals[pi[k]][e], v = vals[pj[k]][e];
                    red[y * Q + 3 + k] += (use && !mi.mean) ? wk * (u.x * v.x + u.y * v.y) : 0.0;
                }
            }
        }
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < Ny * Q; idx += YNT) a.partial[(size_t)blockIdx.x * Ny * Q + idx] = red[idx];
}

#if CF_Y_PART == 0
__global__ void __launch_bounds__(256) reduce_columns_kernel(const double* part, int nblocks, int nq, double* out) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= nq) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += part[(size_t)b * nq + q];
    out[q] = s;
}

hipError_t launch_reduce_columns(const double* part, int nblocks, int nq, double* out, hipStream_t st) {
    hipLaunchKernelGGL(reduce_columns_kernel, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, st, part, nblocks, nq, out);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) reduce_kernel(const double* part, int n, double* out) {
    __shared__ double s[256];
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) v += part[i];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

hipError_t launch_reduce(const double* part, int n, double* out, hipStream_t st) {
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(256), 0, st, part, n, out);
    return hipGetLastError();
}
#endif

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
int y_group_lanes(int Ny);
int y_group_k(int Ny);
size_t y_lds_bytes(int Ny, int ld);

template <typename K>
static void allow_lds(K kernel, size_t lds) {
    if (lds > 64 * 1024) (void)hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

// Non-Nyquist-first order of a Y_STAGE launch (YOrder); returns its grid,
// or 0 (plain order) when the grid has no Nyquist modes or a wave's W modes
// would straddle z rows.
template <int L>
static unsigned y_stage_order(YArgs& a) {
    constexpr int M = YNT / L, W = 64 / L;
    const GridDims& gd = a.gd;
    const bool xe = gd.Nx % 2 == 0, ze = gd.Nz % 2 == 0;
    const int nxr = xe ? gd.Nx / 2 : gd.Nxh, nzr = ze ? gd.Nz - 1 : gd.Nz;
    if ((!xe && !ze) || nxr <= 0 || nzr <= 0 || nxr % W != 0) return 0;
    YOrder& o = a.order;
    o.on = 1;
    // XCD-contiguous order when a block's row segment (M modes, 16 B each) is
    // at least one 128 B line.  Measured on MI355X: 128x97x128 (M = 8) 905 ->
    // 930 steps/s, HBM fetch 1.56 -> 1.11 of the algorithmic bytes; at
    // 256x193x256 (M = 4) the same order lost 6 %.
    o.xcd = (M * 16 >= 128) ? 1 : 0;
    o.nxr = nxr;
    o.zskip = ze ? gd.Nz / 2 : gd.Nz;
    o.nreg = (long long)nxr * nzr;
    o.nnyq = (int)(gd.B - o.nreg);
    o.nyq_first = (unsigned)((o.nreg + M - 1) / M);
    const long long elems = (long long)o.nnyq * gd.Ny;  // about four per thread
    o.nyq_blocks = (unsigned)((elems + 4 * YNT - 1) / (4 * YNT));
    if (o.nyq_blocks == 0) o.nyq_blocks = 1;
    return o.nyq_first + o.nyq_blocks;
}

template <int L, int K>
static size_t y_lds_lk(int Ny, int ld) {
    return (2 * (size_t)(Ny - 1) + (size_t)CHEB_TAB_ARRAYS * K * L / 2 + (YNT / 64) * (size_t)y_region_slots<L, K>(ld)) *
           sizeof(double2);
}

template <int L, int K>
hipError_t launch_y_step(int which, const YArgs& a, hipStream_t st);
template <int L, int K>
hipError_t launch_y(int which, const YArgs& a, hipStream_t st);
template <int L>
hipError_t launch_y_batch(int which, const YArgs* tab, const unsigned* start, int n, unsigned nblocks, size_t lds,
                          hipStream_t st);

#if CF_Y_PART != 0
template <int L, int K>
hipError_t launch_y_step(int which, const YArgs& a, hipStream_t st) {
    constexpr int M = YNT / L;
    const dim3 grid((unsigned)((a.gd.B + M - 1) / M));
    const size_t lds = y_lds_lk<L, K>(a.gd.Ny, a.ld);
    YArgs b = a;
    if (which == Y_STAGE) {
        allow_lds(ystage_kernel<L, K, double2>, lds);
        allow_lds(ystage_kernel<L, K, float2>, lds);
        const unsigned nb = y_stage_order<L>(b);
        const dim3 g2 = nb ? dim3(nb) : grid;
        if (a.mixed) hipLaunchKernelGGL((ystage_kernel<L, K, float2>), g2, dim3(YNT), lds, st, b);
        else hipLaunchKernelGGL((ystage_kernel<L, K, double2>), g2, dim3(YNT), lds, st, b);
    } else if (which == Y_PROLOGUE) {
        // regular (non-Nyquist) blocks only, see yprologue_kernel
        const dim3 g2 = y_stage_order<L>(b) ? dim3(b.order.nyq_first) : grid;
        if (fft_inplace_fits(a.dct, 64 / L, 8, 64)) {
            const size_t ldsi = (2 * (size_t)(a.gd.Ny - 1) + (size_t)CHEB_TAB_ARRAYS * K * L / 2 +
                                 (YNT / 64) * (size_t)(64 / L) * a.ld) * sizeof(double2);
            allow_lds(yprologue_kernel<L, K, double2, true>, ldsi);
            allow_lds(yprologue_kernel<L, K, float2, true>, ldsi);
            if (a.mixed) hipLaunchKernelGGL((yprologue_kernel<L, K, float2, true>), g2, dim3(YNT), ldsi, st, b);
            else hipLaunchKernelGGL((yprologue_kernel<L, K, double2, true>), g2, dim3(YNT), ldsi, st, b);
            return hipGetLastError();
        }
        allow_lds(yprologue_kernel<L, K, double2>, lds);
        allow_lds(yprologue_kernel<L, K, float2>, lds);
        if (a.mixed) hipLaunchKernelGGL((yprologue_kernel<L, K, float2>), g2, dim3(YNT), lds, st, b);
        else hipLaunchKernelGGL((yprologue_kernel<L, K, double2>), g2, dim3(YNT), lds, st, b);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int L, int K>
hipError_t launch_y(int which, const YArgs& a, hipStream_t st) {
    if (which == Y_STAGE || which == Y_PROLOGUE) return launch_y_step<L, K>(which, a, st);
    constexpr int M = YNT / L;
    dim3 grid((unsigned)((a.gd.B + M - 1) / M));
    size_t lds = y_lds_lk<L, K>(a.gd.Ny, a.ld);
    allow_lds(ysolve_kernel<L, K>, lds);
    allow_lds(yneglap_kernel<L, K>, lds);
    allow_lds(ypressure_kernel<L, K>, lds);
    allow_lds(ydct_kernel<L, K>, lds);
    allow_lds(ydivergence_kernel<L, K>, lds);
    allow_lds(ydot_kernel<L, K>, lds);
    allow_lds(ystats_kernel<L, K>, lds);
    switch (which) {
        case Y_SOLVE: hipLaunchKernelGGL((ysolve_kernel<L, K>), grid, dim3(YNT), lds, st, a); break;
        case Y_NEGLAP: hipLaunchKernelGGL((yneglap_kernel<L, K>), grid, dim3(YNT), lds, st, a); break;
        case Y_PRESSURE: hipLaunchKernelGGL((ypressure_kernel<L, K>), grid, dim3(YNT), lds, st, a); break;
        case Y_DCT_FWD: hipLaunchKernelGGL((ydct_kernel<L, K>), grid, dim3(YNT), lds, st, a, -1); break;
        case Y_DCT_INV: hipLaunchKernelGGL((ydct_kernel<L, K>), grid, dim3(YNT), lds, st, a, +1); break;
        case Y_DIVERGENCE: hipLaunchKernelGGL((ydivergence_kernel<L, K>), grid, dim3(YNT), lds, st, a); break;
        case Y_DOT: hipLaunchKernelGGL((ydot_kernel<L, K>), grid, dim3(YNT), lds, st, a, 0); break;
        case Y_DISSIPATION: hipLaunchKernelGGL((ydot_kernel<L, K>), grid, dim3(YNT), lds, st, a, 1); break;
        case Y_PLANE_STATS: hipLaunchKernelGGL((ystats_kernel<L, K>), grid, dim3(YNT), lds, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int L>
hipError_t launch_y_batch(int which, const YArgs* tab, const unsigned* start, int n, unsigned nblocks, size_t lds,
                                 hipStream_t st) {
    if (which == Y_STAGE) {
        allow_lds(ystage_batch_kernel<L, 4>, lds);
        hipLaunchKernelGGL((ystage_batch_kernel<L, 4>), dim3(nblocks), dim3(YNT), lds, st, tab, start, n);
    } else if (which == Y_PROLOGUE) {
        allow_lds(yprologue_batch_kernel<L, 4>, lds);
        hipLaunchKernelGGL((yprologue_batch_kernel<L, 4>), dim3(nblocks), dim3(YNT), lds, st, tab, start, n);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

#endif  // CF_Y_PART != 0

#if CF_Y_PART == 0
unsigned y_blocks(const GridDims& gd) {
    const int L = y_group_lanes(gd.Ny);
    if (!L) return 0;
    const long long M = YNT / L;
    return (unsigned)((gd.B + M - 1) / M);
}

// Group layout of a wall-normal size: K = 4 coefficients per lane and the
// smallest L with 4 L >= Ny; above 256 points L = 64 with K = 6 (up to 384)
// or K = 8 (up to 512).
int y_group_lanes(int Ny) {
    int need = (Ny + 3) / 4;
    if (need <= 4) return 4;
    if (need <= 8) return 8;
    if (need <= 16) return 16;
    if (need <= 32) return 32;
    if (need <= 64) return 64;
    if (Ny <= 512) return 64;
    return 0;
}
int y_group_k(int Ny) { return Ny <= 256 ? 4 : (Ny <= 384 ? 6 : (Ny <= 512 ? 8 : 0)); }

// DCT twiddle table (2P entries) followed by one region per wavefront
size_t y_lds_bytes(int Ny, int ld) {
    return y_lds_bytes_layout(Ny, ld, y_group_lanes(Ny), y_group_k(Ny));
}

// Layouts built: K = 4 with L = 4..64 (the defaults up to 256 points), K = 6
// with L = 64 (the default up to 384) and L = 32, and K = 8 with L = 16, 32,
// 64 (the default up to 512); the non-default ones are alternatives the tuner
// may pick for the step's stage and prologue kernels.
size_t y_lds_bytes_layout(int Ny, int ld, int L, int K) {
    if (K * L < Ny) return 0;
    if (K == 4) {
        switch (L) {
            case 4: return y_lds_lk<4, 4>(Ny, ld);
            case 8: return y_lds_lk<8, 4>(Ny, ld);
            case 16: return y_lds_lk<16, 4>(Ny, ld);
            case 32: return y_lds_lk<32, 4>(Ny, ld);
            case 64: return y_lds_lk<64, 4>(Ny, ld);
            default: return 0;
        }
    }
    if (K == 6) return L == 64 ? y_lds_lk<64, 6>(Ny, ld) : (L == 32 ? y_lds_lk<32, 6>(Ny, ld) : 0);
    if (K == 8) {
        switch (L) {
            case 16: return y_lds_lk<16, 8>(Ny, ld);
            case 32: return y_lds_lk<32, 8>(Ny, ld);
            case 64: return y_lds_lk<64, 8>(Ny, ld);
            default: return 0;
        }
    }
    return 0;
}

hipError_t launch_ykernel_batch(int which, int L, const YArgs* tab, const unsigned* start, int n, unsigned nblocks,
                                size_t lds, hipStream_t st) {
    switch (L) {
        case 4: return launch_y_batch<4>(which, tab, start, n, nblocks, lds, st);
        case 8: return launch_y_batch<8>(which, tab, start, n, nblocks, lds, st);
        case 16: return launch_y_batch<16>(which, tab, start, n, nblocks, lds, st);
        case 32: return launch_y_batch<32>(which, tab, start, n, nblocks, lds, st);
        case 64: return launch_y_batch<64>(which, tab, start, n, nblocks, lds, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ykernel(int which, const YArgs& a, hipStream_t st) {
    if ((which == Y_STAGE || which == Y_PROLOGUE) && a.yk != 0 &&
        !(a.yk == y_group_k(a.gd.Ny) && a.yl == y_group_lanes(a.gd.Ny))) {
        if (!y_lds_bytes_layout(a.gd.Ny, a.ld, a.yl, a.yk)) return hipErrorInvalidValue;
        if (a.yk == 6 && a.yl == 32) return launch_y_step<32, 6>(which, a, st);
        if (a.yk == 6 && a.yl == 64) return launch_y_step<64, 6>(which, a, st);
        if (a.yk == 8 && a.yl == 16) return launch_y_step<16, 8>(which, a, st);
        if (a.yk == 8 && a.yl == 32) return launch_y_step<32, 8>(which, a, st);
        if (a.yk == 8 && a.yl == 64) return launch_y_step<64, 8>(which, a, st);
        return hipErrorInvalidValue;
    }
    if (y_group_k(a.gd.Ny) == 6) return launch_y<64, 6>(which, a, st);
    if (y_group_k(a.gd.Ny) == 8) return launch_y<64, 8>(which, a, st);
    switch (y_group_lanes(a.gd.Ny)) {
        case 4: return launch_y<4, 4>(which, a, st);
        case 8: return launch_y<8, 4>(which, a, st);
        case 16: return launch_y<16, 4>(which, a, st);
        case 32: return launch_y<32, 4>(which, a, st);
        case 64: return launch_y<64, 4>(which, a, st);
        default: return hipErrorInvalidValue;
    }
}

#endif  // CF_Y_PART == 0

// Layouts and the translation unit that instantiates each: the default
// layouts (K = 4 on L = 4..64; K = 6 and 8 on L = 64 above 256 and 384
// points) with every kernel, the alternative layouts (K = 8 on 16 and 32
// lanes, K = 6 on 32) with the step's two kernels only, and the ensemble's
// batched kernels on K = 4.
#if CF_Y_PART == 1
template hipError_t launch_y<4, 4>(int, const YArgs&, hipStream_t);
template hipError_t launch_y<8, 4>(int, const YArgs&, hipStream_t);
template hipError_t launch_y_batch<4>(int, const YArgs*, const unsigned*, int, unsigned, size_t, hipStream_t);
template hipError_t launch_y_batch<8>(int, const YArgs*, const unsigned*, int, unsigned, size_t, hipStream_t);
#elif CF_Y_PART == 2
template hipError_t launch_y<16, 4>(int, const YArgs&, hipStream_t);
template hipError_t launch_y_step<16, 8>(int, const YArgs&, hipStream_t);
template hipError_t launch_y_batch<16>(int, const YArgs*, const unsigned*, int, unsigned, size_t, hipStream_t);
#elif CF_Y_PART == 3
template hipError_t launch_y<32, 4>(int, const YArgs&, hipStream_t);
template hipError_t launch_y_batch<32>(int, const YArgs*, const unsigned*, int, unsigned, size_t, hipStream_t);
#elif CF_Y_PART == 4
template hipError_t launch_y<64, 4>(int, const YArgs&, hipStream_t);
template hipError_t launch_y_batch<64>(int, const YArgs*, const unsigned*, int, unsigned, size_t, hipStream_t);
#elif CF_Y_PART == 5
template hipError_t launch_y_step<32, 8>(int, const YArgs&, hipStream_t);
#elif CF_Y_PART == 6
template hipError_t launch_y<64, 8>(int, const YArgs&, hipStream_t);
template hipError_t launch_y_step<64, 8>(int, const YArgs&, hipStream_t);
#elif CF_Y_PART == 7
template hipError_t launch_y<64, 6>(int, const YArgs&, hipStream_t);
template hipError_t launch_y_step<64, 6>(int, const YArgs&, hipStream_t);
template hipError_t launch_y_step<32, 6>(int, const YArgs&, hipStream_t);
#endif

}  // namespace cf
