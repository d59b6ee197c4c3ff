// Device-side building blocks for the CNRK2 step on gfx950 (CDNA4, wave64).
//
//  * complex FP64 arithmetic on double2 (and FP32 on float2 for the mixed step)
//  * lane-group cross-lane primitives (DPP / permlane) and scans used by the
//    wall-normal (Chebyshev) recurrences: one mode is held by a group of L
//    lanes, lane t owning coefficients n = 4t .. 4t+3
//  * a batched Stockham autosort FFT staged in LDS (mixed radix 2/3/4/5/8 +
//    generic odd radix) used by the x, z and Chebyshev transforms
#pragma once
#include <hip/hip_runtime.h>

#include "cf_common.h"

namespace cf {

// ---------------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 cmk(double r, double i) { return make_double2(r, i); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return cmk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return cmk(a.x - b.x, a.y - b.y); }
// products written as explicit fused multiply-adds: a product rounds the same
// way in every kernel it is inlined into, whatever else the compiler contracts
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return cmk(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cconj(double2 a) { return cmk(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return cmk(a.x * s, a.y * s); }
__device__ __forceinline__ double2 cmuli(double2 a) { return cmk(-a.y, a.x); }    // i*a
__device__ __forceinline__ double2 cmulni(double2 a) { return cmk(a.y, -a.x); }   // -i*a
__device__ __forceinline__ double2 cfma(double s, double2 a, double2 b) { return cmk(fma(s, a.x, b.x), fma(s, a.y, b.y)); }
__device__ __forceinline__ double2 czero() { return cmk(0.0, 0.0); }

// FP32 complex (the y-physical / periodic-pass storage of the mixed step)
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cmuli(float2 a) { return make_float2(-a.y, a.x); }
__device__ __forceinline__ float2 cmulni(float2 a) { return make_float2(a.y, -a.x); }
__device__ __forceinline__ float2 cfma(float s, float2 a, float2 b) { return make_float2(fmaf(s, a.x, b.x), fmaf(s, a.y, b.y)); }

// complex element type C (double2 or float2) -> its real type; construction
template <typename C> struct cplx_traits;
template <> struct cplx_traits<double2> { using real = double; };
template <> struct cplx_traits<float2> { using real = float; };
template <typename C> using creal_t = typename cplx_traits<C>::real;
template <typename C>
__device__ __forceinline__ C cplx(creal_t<C> r, creal_t<C> i) {
    C z;
    z.x = r;
    z.y = i;
    return z;
}
// storage conversions (round to nearest when narrowing)
__device__ __forceinline__ double2 to_c64(double2 a) { return a; }
__device__ __forceinline__ double2 to_c64(float2 a) { return make_double2((double)a.x, (double)a.y); }
template <typename C> __device__ __forceinline__ C from_c64(double2 a);
template <> __device__ __forceinline__ double2 from_c64<double2>(double2 a) { return a; }
template <> __device__ __forceinline__ float2 from_c64<float2>(double2 a) { return make_float2((float)a.x, (float)a.y); }

// ---------------------------------------------------------------------------
// cross-lane moves.  A wave64 is four DPP rows of 16 lanes.  Within a row:
// DPP row shifts, quad permutes and mirrors; across rows: row_bcast:15/31,
// row_newbcast:0 (64-bit DPP), v_permlane16_swap / v_permlane32_swap and
// readlane.  All are VALU operations (no LDS round trip).
// ---------------------------------------------------------------------------
namespace dpp {
enum : int {
    QUAD_XOR1 = 0xB1,        // quad_perm [1,0,3,2]
    QUAD_XOR2 = 0x4E,        // quad_perm [2,3,0,1]
    ROW_SHL = 0x100,         // + d: lane r <- r+d within the row
    ROW_SHR = 0x110,         // + d: lane r <- r-d within the row
    ROW_MIRROR = 0x140,      // lane r <- 15-r
    ROW_HALF_MIRROR = 0x141, // lane r <- 7-r within each half row
    ROW_BCAST15 = 0x142,     // row k+1 <- lane 15 of row k (rows selected by the row mask)
    ROW_BCAST31 = 0x143,     // rows 2,3 <- lane 31 (rows selected by the row mask)
    ROW_NEWBCAST0 = 0x150,   // every lane <- first lane of its row (64-bit)
};
// Lanes whose source lies outside the row, or whose row is not in ROWMASK,
// receive an unspecified value (callers select them away).
template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ double mov(double x) {
    return __longlong_as_double(__builtin_amdgcn_mov_dpp(__double_as_longlong(x), CTRL, ROWMASK, 0xF, false));
}
template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ double2 mov(double2 x) {
    return make_double2(mov<CTRL, ROWMASK>(x.x), mov<CTRL, ROWMASK>(x.y));
}
// Lanes whose source lies outside the row receive 0 (the additive identity).
template <int CTRL>
__device__ __forceinline__ double mov0(double x) {
    return __longlong_as_double(__builtin_amdgcn_mov_dpp(__double_as_longlong(x), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double2 mov0(double2 x) { return make_double2(mov0<CTRL>(x.x), mov0<CTRL>(x.y)); }
// Lanes whose source lies outside the row receive 1 (the multiplicative identity).
template <int CTRL>
__device__ __forceinline__ double mov1(double x) {
    return __longlong_as_double(
        __builtin_amdgcn_update_dpp(0x3FF0000000000000ll, __double_as_longlong(x), CTRL, 0xF, 0xF, false));
}
// value held by the odd row of this lane's row pair (rows 0/1, 2/3), same column
__device__ __forceinline__ double odd_row(double x) {
    int lo = __double2loint(x), hi = __double2hiint(x);
    auto l = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    auto h = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    return __hiloint2double((int)h[1], (int)l[1]);
}
__device__ __forceinline__ void row_pair(double x, double& even, double& odd) {
    int lo = __double2loint(x), hi = __double2hiint(x);
    auto l = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    auto h = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    even = __hiloint2double((int)h[0], (int)l[0]);
    odd = __hiloint2double((int)h[1], (int)l[1]);
}
__device__ __forceinline__ void wave_halves(double x, double& lower, double& upper) {
    int lo = __double2loint(x), hi = __double2hiint(x);
    auto l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    auto h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    lower = __hiloint2double((int)h[0], (int)l[0]);
    upper = __hiloint2double((int)h[1], (int)l[1]);
}
__device__ __forceinline__ double readlane(double x, int lane) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), lane),
                            __builtin_amdgcn_readlane(__double2loint(x), lane));
}
__device__ __forceinline__ double2 readlane(double2 x, int lane) { return make_double2(readlane(x.x, lane), readlane(x.y, lane)); }
__device__ __forceinline__ double2 odd_row(double2 x) { return make_double2(odd_row(x.x), odd_row(x.y)); }
}  // namespace dpp

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Group of L lanes (power of two, 4 <= L <= 64) owning one mode; t = index
// within the group, r = index within the DPP row.  Groups of L <= 16 lie in
// one row; L = 32 is a row pair, L = 64 the whole wave.  Every lane of a group
// must call a group operation together (values of other groups are never used).
//
// Each lane holds K consecutive Chebyshev coefficients (K = 4, 6 or 8, even:
// element e of a lane has the parity of e), so a group covers degrees below
// K L.  tab: per-degree constants (see build_cheb_table), NT = K L entries
// per array.
template <int L, int K = 4>
struct Group {
    static_assert(K % 2 == 0, "parity of a coefficient must follow its element index");
    static constexpr int NT = K * L;
    int lane, t, r;
    const double* tab;
    __device__ __forceinline__ explicit Group(const double* tab_ = nullptr) : tab(tab_) {
        lane = lane_id();
        t = lane & (L - 1);
        r = lane & 15;
    }

    // ---- neighbours: group lane t-1 (undefined at t = 0), t+1 (undefined at t = L-1)
    template <typename V>
    __device__ __forceinline__ V up1(V x) const {
        V in_row = dpp::mov<dpp::ROW_SHR + 1>(x);
        if constexpr (L <= 16) return in_row;
        else {
            V across = dpp::mov<dpp::ROW_BCAST15, (L == 32 ? 0xA : 0xE)>(x);
            return r == 0 ? across : in_row;
        }
    }
    template <typename V>
    __device__ __forceinline__ V down1(V x) const {
        V in_row = dpp::mov<dpp::ROW_SHL + 1>(x);
        if constexpr (L <= 16) return in_row;
        else {
            V across = dpp::mov<dpp::ROW_NEWBCAST0>(dpp::odd_row(x));  // first lane of the next row (rows 0, 2)
            if constexpr (L == 64) {
                V mid = dpp::readlane(x, 32);  // lane 31 <- lane 32
                if (lane == 31) across = mid;
            }
            return r == 15 ? across : in_row;
        }
    }

    // ---- all-reduce over the group
    __device__ __forceinline__ double sum(double v) const {
        v += dpp::mov<dpp::QUAD_XOR1>(v);
        v += dpp::mov<dpp::QUAD_XOR2>(v);
        if constexpr (L >= 8) v += dpp::mov<dpp::ROW_HALF_MIRROR>(v);
        if constexpr (L >= 16) v += dpp::mov<dpp::ROW_MIRROR>(v);
        if constexpr (L >= 32) {
            double e, o;
            dpp::row_pair(v, e, o);
            v = e + o;
        }
        if constexpr (L >= 64) {
            double lo, hi;
            dpp::wave_halves(v, lo, hi);
            v = lo + hi;
        }
        return v;
    }
    __device__ __forceinline__ double2 sum(double2 v) const { return make_double2(sum(v.x), sum(v.y)); }

    // ---- inclusive suffix sum over group lanes t..L-1
    template <int D>
    __device__ __forceinline__ bool suffix_ok() const {
        return (L <= 16) ? (t + D < L) : (r + D < 16);
    }
    template <int D>
    __device__ __forceinline__ bool prefix_ok() const {
        return ((L <= 16) ? t : r) >= D;
    }
    // Steps inside a DPP row.  For L >= 16 a group spans whole rows, so the
    // lanes whose source falls outside the row are exactly those that must
    // combine with the identity, which bound-controlled moves deliver.
    template <int D>
    __device__ __forceinline__ void suffix_sum_step(double2& v) const {
        if constexpr (D < L && D < 16) {
            if constexpr (L >= 16) {
                v = cadd(v, dpp::mov0<dpp::ROW_SHL + D>(v));
            } else {
                double2 o = dpp::mov<dpp::ROW_SHL + D>(v);
                if (suffix_ok<D>()) v = cadd(v, o);
            }
        }
    }
    __device__ __forceinline__ double2 suffix_sum(double2 v) const {
        suffix_sum_step<1>(v);
        suffix_sum_step<2>(v);
        suffix_sum_step<4>(v);
        suffix_sum_step<8>(v);
        if constexpr (L >= 32) {  // even rows add the next row's total (its first lane)
            double2 o = dpp::mov<dpp::ROW_NEWBCAST0>(dpp::odd_row(v));
            if ((lane & 16) == 0) v = cadd(v, o);
        }
        if constexpr (L == 64) {  // lower half adds the upper half's total (lane 32)
            double2 o = dpp::readlane(v, 32);
            if (lane < 32) v = cadd(v, o);
        }
        return v;
    }

    // ---- affine compositions x -> A + T x.  Suffix: lane t ends with
    // f_t o f_{t+1} o ... o f_{L-1}; prefix: f_t o f_{t-1} o ... o f_0.
    // Combining with a later/earlier composite (oT, oA): (T oT, A + T oA).
    // NR right-hand sides share the multiplicative part T (one operator):
    // A[0..NR) are scanned together, each step moving T once.
    template <int D, int NR>
    __device__ __forceinline__ void suffix_affine_step(double& T, double2* A) const {
        if constexpr (D < L && D < 16) {
            if constexpr (L >= 16) {
                double oT = dpp::mov1<dpp::ROW_SHL + D>(T);
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    double2 oA = dpp::mov0<dpp::ROW_SHL + D>(A[r]);
                    A[r] = cfma(T, oA, A[r]);
                }
                T *= oT;
            } else {
                double oT = dpp::mov<dpp::ROW_SHL + D>(T);
                double2 oA[NR];
#pragma unroll
                for (int r = 0; r < NR; ++r) oA[r] = dpp::mov<dpp::ROW_SHL + D>(A[r]);
                if (suffix_ok<D>()) {
#pragma unroll
                    for (int r = 0; r < NR; ++r) A[r] = cfma(T, oA[r], A[r]);
                    T *= oT;
                }
            }
        }
    }
    template <int NR>
    __device__ __forceinline__ void suffix_affine_n(double& T, double2* A) const {
        suffix_affine_step<1, NR>(T, A);
        suffix_affine_step<2, NR>(T, A);
        suffix_affine_step<4, NR>(T, A);
        suffix_affine_step<8, NR>(T, A);
        if constexpr (L >= 32) {
            double oT = dpp::mov<dpp::ROW_NEWBCAST0>(dpp::odd_row(T));
            double2 oA[NR];
#pragma unroll
            for (int r = 0; r < NR; ++r) oA[r] = dpp::mov<dpp::ROW_NEWBCAST0>(dpp::odd_row(A[r]));
            if ((lane & 16) == 0) {
#pragma unroll
                for (int r = 0; r < NR; ++r) A[r] = cfma(T, oA[r], A[r]);
                T *= oT;
            }
        }
        if constexpr (L == 64) {
            double oT = dpp::readlane(T, 32);
            double2 oA[NR];
#pragma unroll
            for (int r = 0; r < NR; ++r) oA[r] = dpp::readlane(A[r], 32);
            if (lane < 32) {
#pragma unroll
                for (int r = 0; r < NR; ++r) A[r] = cfma(T, oA[r], A[r]);
                T *= oT;
            }
        }
    }
    __device__ __forceinline__ void suffix_affine(double& T, double2& A) const { suffix_affine_n<1>(T, &A); }
    template <int D, int NR>
    __device__ __forceinline__ void prefix_affine_step(double& T, double2* A) const {
        if constexpr (D < L && D < 16) {
            if constexpr (L >= 16) {
                double oT = dpp::mov1<dpp::ROW_SHR + D>(T);
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    double2 oA = dpp::mov0<dpp::ROW_SHR + D>(A[r]);
                    A[r] = cfma(T, oA, A[r]);
                }
                T *= oT;
            } else {
                double oT = dpp::mov<dpp::ROW_SHR + D>(T);
                double2 oA[NR];
#pragma unroll
                for (int r = 0; r < NR; ++r) oA[r] = dpp::mov<dpp::ROW_SHR + D>(A[r]);
                if (prefix_ok<D>()) {
#pragma unroll
                    for (int r = 0; r < NR; ++r) A[r] = cfma(T, oA[r], A[r]);
                    T *= oT;
                }
            }
        }
    }
    template <int NR>
    __device__ __forceinline__ void prefix_affine_n(double& T, double2* A) const {
        prefix_affine_step<1, NR>(T, A);
        prefix_affine_step<2, NR>(T, A);
        prefix_affine_step<4, NR>(T, A);
        prefix_affine_step<8, NR>(T, A);
        if constexpr (L >= 32) {  // odd rows take the previous row's total (its lane 15)
            double oT = dpp::mov<dpp::ROW_BCAST15, 0xA>(T);
            double2 oA[NR];
#pragma unroll
            for (int r = 0; r < NR; ++r) oA[r] = dpp::mov<dpp::ROW_BCAST15, 0xA>(A[r]);
            if (lane & 16) {
#pragma unroll
                for (int r = 0; r < NR; ++r) A[r] = cfma(T, oA[r], A[r]);
                T *= oT;
            }
        }
        if constexpr (L == 64) {  // upper half takes the lower half's total (lane 31)
            double oT = dpp::mov<dpp::ROW_BCAST31, 0xC>(T);
            double2 oA[NR];
#pragma unroll
            for (int r = 0; r < NR; ++r) oA[r] = dpp::mov<dpp::ROW_BCAST31, 0xC>(A[r]);
            if (lane & 32) {
#pragma unroll
                for (int r = 0; r < NR; ++r) A[r] = cfma(T, oA[r], A[r]);
                T *= oT;
            }
        }
    }
    __device__ __forceinline__ void prefix_affine(double& T, double2& A) const { prefix_affine_n<1>(T, &A); }
};

// Per-degree constants of the Chebyshev tau rows and of the bulk mean,
// built once per block (LDS) so that the solves need no FP64 divisions:
//   [0, nt)    c1(n) = c_{n-2} / (4 n (n-1))      (n >= 2; c_0 = 2, else 1)
//   [nt, 2nt)  c2(n) = -1 / (2 (n^2 - 1))        (n >= 2)
//   [2nt, 3nt) c3(n) = 1 / (4 n (n+1))           (n >= 2)
//   [3nt, 4nt) bw(n) = 1 / (1 - n^2), even n; 0 for odd n
// (tau rows: docs/src/spatial.md:448-470; bulk mean: meanmode.jl:13-19)
// The host builds the table once per problem (build_cheb_table in setup.cpp);
// every y block copies it into LDS.

// ---------------------------------------------------------------------------
// Chebyshev coefficient vectors in the blocked group layout: lane t holds
// v[e] = a_{Kt+e}.  Entries above degree P must be zero.
// ---------------------------------------------------------------------------
template <int K = 4>
struct CVec {
    double2 v[K];
};

// d/dy (backward recurrence b_n = b_{n+2} + 2(n+1) a_{n+1}, b_0 halved;
// _batchdiff!, batchedinfluence.jl:32-43) as two parity suffix sums of 2m a_m:
// b_n is the sum of 2m a_m over m > n of the other parity.
template <int L, int K>
__device__ __forceinline__ CVec<K> cheb_deriv(const Group<L, K>& g, const CVec<K>& a) {
    const int n0 = K * g.t;
    double2 c[K];
#pragma unroll
    for (int e = 0; e < K; ++e) c[e] = cscale(a.v[e], 2.0 * (n0 + e));
    double2 E = c[0], O = c[1];
#pragma unroll
    for (int e = 2; e < K; e += 2) {
        E = cadd(E, c[e]);
        O = cadd(O, c[e + 1]);
    }
    // exclusive suffix sums over lanes of E and O
    double2 accE = csub(g.suffix_sum(E), E), accO = csub(g.suffix_sum(O), O);
    CVec<K> b;
#pragma unroll
    for (int e = K - 1; e >= 0; --e) {
        if (e & 1) {
            b.v[e] = accE;
            accO = cadd(c[e], accO);
        } else {
            b.v[e] = accO;
            accE = cadd(c[e], accE);
        }
    }
    if (g.t == 0) b.v[0] = cscale(b.v[0], 0.5);
    return b;
}

// Wall values / derivatives / bulk mean as group reductions.
template <int L, int K>
__device__ __forceinline__ double2 wall_deriv_plus(const Group<L, K>& g, const CVec<K>& a) {  // sum n^2 a_n
    double2 s = czero();
#pragma unroll
    for (int e = 0; e < K; ++e) {
        double n = K * g.t + e;
        s = cfma(n * n, a.v[e], s);
    }
    return g.sum(s);
}
template <int L, int K>
__device__ __forceinline__ double2 wall_deriv_minus(const Group<L, K>& g, const CVec<K>& a) {  // sum (-1)^(n+1) n^2 a_n
    double2 s = czero();
#pragma unroll
    for (int e = 0; e < K; ++e) {
        double n = K * g.t + e;
        s = cfma(((e & 1) ? 1.0 : -1.0) * n * n, a.v[e], s);
    }
    return g.sum(s);
}
template <int L, int K>
__device__ __forceinline__ double2 bulk_mean(const Group<L, K>& g, const CVec<K>& a) {  // a_0 + sum_{even m>=2} a_m/(1-m^2)
    double2 s = czero();
    const double* bw = g.tab + 3 * Group<L, K>::NT;
#pragma unroll
    for (int e = 0; e < K; e += 2) s = cfma(bw[K * g.t + e], a.v[e], s);
    return g.sum(s);
}

// ---------------------------------------------------------------------------
// Chebyshev-tau Helmholtz solve (t0 D^2 - t1) a = f on rows 0..P-2 plus wall
// rows, by the quasi-tridiagonal even/odd elimination.  `fac` are this mode's
// elimination factors s_j (host-computed), indexed by degree n = 2j + parity.
// Data are the parity wall data: Dirichlet even = (g+ + g-)/2, odd =
// (g+ - g-)/2; Neumann even = (g+ - g-)/2, odd = (g+ + g-)/2.
// Both parities are processed together: parity p lives in elements p, p+2.
// ---------------------------------------------------------------------------
struct HelmCoef {
    double lo, up;  // row coefficients for unknown j (analytic in n, t1)
};
__device__ __forceinline__ double cfac_(int m) { return m == 0 ? 2.0 : 1.0; }

// NR right-hand sides f[r] with their parity wall data (de[r], dd[r]) share
// the operator: its elimination coefficients, the multiplicative parts of
// the scans and the wall-row weights are computed once, and the NR
// independent recurrences interleave (more independent work per scan step).
template <int NR, int L, int K>
__device__ __forceinline__ void helm_solve_n(const Group<L, K>& g, int P, double t1, const double fac[K], bool neumann,
                                             const CVec<K>* f, const double2* de, const double2* dd, CVec<K>* out) {
    const int n0 = K * g.t;
    constexpr int NT = Group<L, K>::NT;
    // per element: elimination coefficients; T = -up*s (backward), alpha = -lo*s
    double T[K], alpha[K];
    bool rows[K];
#pragma unroll
    for (int e = 0; e < K; ++e) {
        const int n = n0 + e;
        const bool row = (n >= 2) && (n <= P);
        rows[e] = row;
        const double s = row ? fac[e] : 0.0;
        const double en2 = (n + 2 <= P - 2) ? 1.0 : 0.0;
        const double lo = -t1 * g.tab[n];
        const double up = -t1 * en2 * g.tab[2 * NT + n];
        T[e] = row ? -up * s : 0.0;
        alpha[e] = row ? -lo * s : (n < 2 ? 1.0 : 0.0);  // j=0: identity (gamma_0 = 1)
    }
    // A = s * rhs of the 3-term tau stencil in coefficients n - 2, n, n + 2
    // (elements e -+ 2 of this lane, or of the previous / next lane)
    double2 A[NR][K];
    {
        const bool hasprev = g.t > 0, hasnext = g.t + 1 < L;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            double2 fm2[K], fp2[K];
            const double2 a2 = g.up1(f[r].v[K - 2]), a3 = g.up1(f[r].v[K - 1]);
            const double2 b0 = g.down1(f[r].v[0]), b1 = g.down1(f[r].v[1]);
            fm2[0] = hasprev ? a2 : czero();
            fm2[1] = hasprev ? a3 : czero();
            fp2[K - 2] = hasnext ? b0 : czero();
            fp2[K - 1] = hasnext ? b1 : czero();
#pragma unroll
            for (int e = 2; e < K; ++e) {
                fm2[e] = f[r].v[e - 2];
                fp2[e - 2] = f[r].v[e];
            }
#pragma unroll
            for (int e = 0; e < K; ++e) {
                const int n = n0 + e;
                const double c1 = g.tab[n], c2 = g.tab[NT + n], c3 = g.tab[2 * NT + n];
                const double2 Fm = (n - 2 >= 0 && n - 2 <= P - 2) ? fm2[e] : czero();
                const double2 F0 = (n <= P - 2) ? f[r].v[e] : czero();
                const double2 Fp = (n + 2 <= P - 2) ? fp2[e] : czero();
                double2 rhs = czero();
                if (rows[e]) {
                    rhs = cscale(Fm, c1);
                    rhs = cfma(c2, F0, rhs);
                    rhs = cfma(c3, Fp, rhs);
                }
                A[r][e] = cscale(rhs, rows[e] ? fac[e] : 0.0);
            }
        }
    }
    // ---- backward sweep: beta_j = A_j + T_j beta_{j+1}, per parity ----
    // lane composite over its K/2 elements of each parity p (e = p, p+2, ...):
    // f_p o f_{p+2} o ... o f_{p+K-2}
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        double cT = T[K - 2 + p];
        double2 cA[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) cA[r] = A[r][K - 2 + p];
#pragma unroll
        for (int e = K - 4 + p; e >= 0; e -= 2) {
#pragma unroll
            for (int r = 0; r < NR; ++r) cA[r] = cfma(T[e], cA[r], A[r][e]);
            cT = T[e] * cT;
        }
        // inclusive suffix scan of composites (f_t o f_{t+1} o ...); the carry
        // is beta at the first j of the next lane: the A part at lane t+1
        g.template suffix_affine_n<NR>(cT, cA);
        const bool ok = g.t + 1 < L;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            double2 carry = g.down1(cA[r]);
            if (!ok) carry = czero();
            A[r][K - 2 + p] = cfma(T[K - 2 + p], carry, A[r][K - 2 + p]);  // A now holds beta
#pragma unroll
            for (int e = K - 4 + p; e >= 0; e -= 2) A[r][e] = cfma(T[e], A[r][e + 2], A[r][e]);
        }
    }
    // ---- forward sweep: delta_j = beta_j + alpha_j delta_{j-1}, gamma_j = prod alpha ----
    double gm[K];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        // lane composite g_{p+K-2} o ... o g_{p+2} o g_p: x -> beta + alpha x per element
        double fT = alpha[p];
        double2 fA[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) fA[r] = A[r][p];
#pragma unroll
        for (int e = p + 2; e < K; e += 2) {
#pragma unroll
            for (int r = 0; r < NR; ++r) fA[r] = cfma(alpha[e], fA[r], A[r][e]);
            fT = alpha[e] * fT;
        }
        g.template prefix_affine_n<NR>(fT, fA);  // (later) o (earlier)
        const bool ok = g.t >= 1;
        double carryT = g.up1(fT);
        if (!ok) carryT = 1.0;
        gm[p] = alpha[p] * carryT;
#pragma unroll
        for (int e = p + 2; e < K; e += 2) gm[e] = alpha[e] * gm[e - 2];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            double2 carryA = g.up1(fA[r]);
            if (!ok) carryA = czero();
            A[r][p] = cfma(alpha[p], carryA, A[r][p]);  // A now holds delta
#pragma unroll
            for (int e = p + 2; e < K; e += 2) A[r][e] = cfma(alpha[e], A[r][e - 2], A[r][e]);
        }
    }
    // ---- wall row: x0 = (data - sum w delta) / sum w gamma ----
    double swg[2] = {0.0, 0.0};
#pragma unroll
    for (int e = 0; e < K; ++e) {
        const int n = n0 + e;
        const double w = (n <= P) ? (neumann ? (double)n * n : 1.0) : 0.0;
        swg[e & 1] += w * gm[e];
    }
    double inv[2];
#pragma unroll
    for (int p = 0; p < 2; ++p) inv[p] = 1.0 / g.sum(swg[p]);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        double2 swd[2] = {czero(), czero()};
#pragma unroll
        for (int e = 0; e < K; ++e) {
            const int n = n0 + e;
            const double w = (n <= P) ? (neumann ? (double)n * n : 1.0) : 0.0;
            swd[e & 1] = cfma(w, A[r][e], swd[e & 1]);
        }
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const double2 s1 = g.sum(swd[p]);
            const double2 dat = p == 0 ? de[r] : dd[r];
            const double2 x0 = cscale(csub(dat, s1), inv[p]);
#pragma unroll
            for (int e = p; e < K; e += 2) {
                const int n = n0 + e;
                out[r].v[e] = (n <= P) ? cfma(gm[e], x0, A[r][e]) : czero();
            }
        }
    }
}

template <int L, int K>
__device__ __forceinline__ CVec<K> helm_solve(const Group<L, K>& g, int P, double t1, const double fac[K], bool neumann,
                                              const CVec<K>& f, double2 data_even, double2 data_odd) {
    CVec<K> out;
    helm_solve_n<1>(g, P, t1, fac, neumann, &f, &data_even, &data_odd, &out);
    return out;
}

// ---------------------------------------------------------------------------
// Batched Stockham FFT in LDS.
// Lines of length n at stride `ld` (complex elements) in buffers a, b; the
// result ends in *a or *b (returned).  sign = -1 forward, +1 backward.
// tw[t] = exp(-2 pi i t / n), t < n (global memory, read-only).
// ---------------------------------------------------------------------------
template <int R, typename C>
__device__ __forceinline__ void dft_small(C* v, int sign) {
    using T = creal_t<C>;
    if constexpr (R == 2) {
        C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 3) {
        const T c = T(-0.5), s = T(0.86602540378443864676);  // cos(2pi/3), sin(2pi/3)
        C a = v[0], b = v[1], d = v[2];
        C t = cadd(b, d), u = csub(b, d);
        C m = cfma(c, t, a);
        C su = cscale(u, sign < 0 ? -s : s);  // forward: v1 = m - i s u
        C isu = cmuli(su);
        v[0] = cadd(a, t);
        v[1] = cadd(m, isu);
        v[2] = csub(m, isu);
    } else if constexpr (R == 4) {
        C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        C t2 = cadd(v[1], v[3]), t3 = csub(v[1], v[3]);
        C jt3 = (sign < 0) ? cmulni(t3) : cmuli(t3);
        v[0] = cadd(t0, t2);
        v[1] = cadd(t1, jt3);
        v[2] = csub(t0, t2);
        v[3] = csub(t1, jt3);
    } else if constexpr (R == 5) {
        const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
        const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);
        C a = v[0];
        C b1 = cadd(v[1], v[4]), b2 = cadd(v[2], v[3]);
        C d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
        C m1 = cadd(a, cadd(cscale(b1, c1), cscale(b2, c2)));
        C m2 = cadd(a, cadd(cscale(b1, c2), cscale(b2, c1)));
        // forward: -i*(s1 d1 + s2 d2), -i*(s2 d1 - s1 d2)
        C n1 = cadd(cscale(d1, s1), cscale(d2, s2));
        C n2 = csub(cscale(d1, s2), cscale(d2, s1));
        C j1 = (sign < 0) ? cmulni(n1) : cmuli(n1);
        C j2 = (sign < 0) ? cmulni(n2) : cmuli(n2);
        v[0] = cadd(a, cadd(b1, b2));
        v[1] = cadd(m1, j1);
        v[4] = csub(m1, j1);
        v[2] = cadd(m2, j2);
        v[3] = csub(m2, j2);
    } else if constexpr (R == 8) {
        const T h = T(0.70710678118654752440);
        // radix-8 = 2 x radix-4 with twiddles
        C e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        dft_small<4>(e, sign);
        dft_small<4>(o, sign);
        // twiddles W8^k, k=0..3 (forward W8 = exp(-i pi/4))
        C w1 = cplx<C>(h, sign < 0 ? -h : h);
        C w2 = cplx<C>(T(0), sign < 0 ? T(-1) : T(1));
        C w3 = cplx<C>(-h, sign < 0 ? -h : h);
        o[1] = cmul(o[1], w1);
        o[2] = cmul(o[2], w2);
        o[3] = cmul(o[3], w3);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[k] = cadd(e[k], o[k]);
            v[k + 4] = csub(e[k], o[k]);
        }
    }
}

template <int R, typename C>
__device__ __forceinline__ void stockham_pass_fixed(const C* src, C* dst, int n, int nlines, int ld, int Ns,
                                                    unsigned long long mul_m, unsigned long long mul_ns,
                                                    const C* tw, int sign, int tid, int nthr) {
    const int m = n / R;
    const int twstep = n / (Ns * R);
    const int total = nlines * m;
    for (int idx = tid; idx < total; idx += nthr) {
        int line = (int)fast_div((unsigned)idx, mul_m);
        int j = idx - line * m;
        int jq = (int)fast_div((unsigned)j, mul_ns);
        int jm = j - jq * Ns;
        const C* s = src + (size_t)line * ld;
        C* d = dst + (size_t)line * ld;
        C v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            C x = s[j + r * m];
            if (r > 0) {
                C w = tw[r * jm * twstep];  // r*jm*twstep < n
                if (sign > 0) w = cconj(w);
                x = cmul(x, w);
            }
            v[r] = x;
        }
        dft_small<R>(v, sign);
        int idxD = jq * Ns * R + jm;
#pragma unroll
        for (int r = 0; r < R; ++r) d[idxD + r * Ns] = v[r];
    }
}

// generic odd radix (rare sizes): O(R^2) per butterfly, reads LDS directly
template <typename C>
__device__ __forceinline__ void stockham_pass_generic(const C* src, C* dst, int n, int R, int nlines, int ld, int Ns,
                                                      const C* tw, int sign, int tid, int nthr) {
    const int m = n / R;
    const int twstep = n / (Ns * R);
    const int total = nlines * m;
    for (int idx = tid; idx < total; idx += nthr) {
        int line = idx / m;
        int j = idx - line * m;
        const C* s = src + (size_t)line * ld;
        C* d = dst + (size_t)line * ld;
        int jm = j % Ns;
        int idxD = (j / Ns) * Ns * R + jm;
        for (int q = 0; q < R; ++q) {
            C acc = cplx<C>(0, 0);
            for (int r = 0; r < R; ++r) {
                C x = s[j + r * m];
                // twiddle (r * jm / (Ns R)) and DFT_R phase (r q / R), combined:
                // exp(-2 pi i r (jm + q Ns) / (Ns R))
                int t = (int)(((long long)r * (jm + (long long)q * Ns) * twstep) % n);
                C w = tw[t];
                if (sign > 0) w = cconj(w);
                acc = cadd(acc, cmul(x, w));
            }
            d[idxD + q * Ns] = acc;
        }
    }
}
// the same pass as an out-of-line call: keeps the rare path's registers out
// of kernels whose register budget is tight elsewhere
template <typename C>
__device__ __noinline__ void stockham_pass_generic_call(const C* src, C* dst, int n, int R, int nlines, int ld, int Ns,
                                                        const C* tw, int sign, int tid, int nthr) {
    stockham_pass_generic(src, dst, n, R, nlines, ld, Ns, tw, sign, tid, nthr);
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2): block b runs on XCD b % 8 as that XCD's (b / 8)-th block.  The
// returned logical index gives XCD x a contiguous range of logical blocks,
// so neighbouring tiles that share cache lines share an L2.
constexpr int NXCD = 8;
__device__ __forceinline__ unsigned xcd_logical(unsigned b, unsigned nblocks) {
    unsigned q = nblocks / NXCD, r = nblocks % NXCD, x = b % NXCD, slot = b / NXCD;
    return (x < r) ? x * (q + 1) + slot : r * (q + 1) + (x - r) * q + slot;
}

// Instance of a batched (ensemble) launch owning block b: the largest i < n
// with start[i] <= b (start is ascending, start[0] = 0).  Uniform per block.
__device__ __forceinline__ int batch_instance(const unsigned* __restrict__ start, int n, unsigned b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Ordering of LDS accesses among the lanes of one wavefront: LDS operations
// of a wave are performed in issue order, so a compiler-level barrier with
// wavefront-scope fences is all a wave-private exchange needs.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Run all passes; returns the buffer holding the result.  WAVE = false: the
// lines are shared by the block and every thread of the block must call
// (passes end in __syncthreads); WAVE = true: the lines are private to one
// wavefront, tid/nthr are its lane/64 and passes end in wave_sync().
// tw: twiddle table exp(-2 pi i t/n) (global or LDS).
// GENERIC_INLINE: the generic-radix pass is inlined (no call: a call forces
// registers live across it to be saved, which at a tight register budget
// means scratch) rather than called out of line.
template <bool WAVE = false, bool GENERIC_INLINE = false, typename C>
__device__ __forceinline__ C* lds_fft(C* a, C* b, const FFTPlan& plan, const C* tw, int nlines, int ld, int sign, int tid,
                                      int nthr) {
    int Ns = 1;
    C* src = a;
    C* dst = b;
    for (int f = 0; f < plan.nf; ++f) {
        int R = plan.radix[f];
        const unsigned long long mm = plan.mul_m[f], mn = plan.mul_ns[f];
        switch (R) {
            case 2: stockham_pass_fixed<2>(src, dst, plan.n, nlines, ld, Ns, mm, mn, tw, sign, tid, nthr); break;
            case 3: stockham_pass_fixed<3>(src, dst, plan.n, nlines, ld, Ns, mm, mn, tw, sign, tid, nthr); break;
            case 4: stockham_pass_fixed<4>(src, dst, plan.n, nlines, ld, Ns, mm, mn, tw, sign, tid, nthr); break;
            case 5: stockham_pass_fixed<5>(src, dst, plan.n, nlines, ld, Ns, mm, mn, tw, sign, tid, nthr); break;
            case 8: stockham_pass_fixed<8>(src, dst, plan.n, nlines, ld, Ns, mm, mn, tw, sign, tid, nthr); break;
            default:
                if constexpr (GENERIC_INLINE) stockham_pass_generic(src, dst, plan.n, R, nlines, ld, Ns, tw, sign, tid, nthr);
                else stockham_pass_generic_call(src, dst, plan.n, R, nlines, ld, Ns, tw, sign, tid, nthr);
                break;
        }
        if constexpr (WAVE) wave_sync();
        else __syncthreads();
        Ns *= R;
        C* t = src;
        src = dst;
        dst = t;
    }
    return src;
}
// In-place variant: one buffer.  Each pass first gathers every butterfly of
// this thread into registers, synchronises, then scatters the outputs.  PPT
// bounds the points per thread (nlines * n <= PPT * nthr), so a thread holds
// at most ceil(PPT / R) butterflies of radix R.  The plan must contain only
// radices 2, 3, 4, 5 and 8 (fft_inplace_fits).
template <int R, int PPT, bool WAVE, typename C>
__device__ __forceinline__ void stockham_pass_inplace(C* buf, int n, int nlines, int ld, int Ns,
                                                      unsigned long long mul_m, unsigned long long mul_ns, const C* tw,
                                                      int sign, int tid, int nthr) {
    constexpr int NB = (PPT + R - 1) / R;
    const int m = n / R;
    const int twstep = n / (Ns * R);
    const int total = nlines * m;
    C v[NB][R];
    int dst[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int idx = tid + b * nthr;
        dst[b] = -1;
        if (idx < total) {
            int line = (int)fast_div((unsigned)idx, mul_m);
            int j = idx - line * m;
            int jq = (int)fast_div((unsigned)j, mul_ns);
            int jm = j - jq * Ns;
            const C* s = buf + (size_t)line * ld;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                C x = s[j + r * m];
                if (r > 0) {
                    C w = tw[r * jm * twstep];
                    if (sign > 0) w = cconj(w);
                    x = cmul(x, w);
                }
                v[b][r] = x;
            }
            dft_small<R>(v[b], sign);
            dst[b] = line * ld + jq * Ns * R + jm;
        }
    }
    if constexpr (WAVE) wave_sync();
    else __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; ++b)
        if (dst[b] >= 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) buf[dst[b] + r * Ns] = v[b][r];
        }
    if constexpr (WAVE) wave_sync();
    else __syncthreads();
}

template <int PPT, bool WAVE, typename C>
__device__ __forceinline__ void lds_fft_inplace(C* buf, const FFTPlan& plan, const C* tw, int nlines, int ld, int sign,
                                                int tid, int nthr) {
    int Ns = 1;
    for (int f = 0; f < plan.nf; ++f) {
        const int R = plan.radix[f];
        const unsigned long long mm = plan.mul_m[f], mn = plan.mul_ns[f];
        switch (R) {
            case 2: stockham_pass_inplace<2, PPT, WAVE>(buf, plan.n, nlines, ld, Ns, mm, mn, tw, sign, tid, nthr); break;
            case 3: stockham_pass_inplace<3, PPT, WAVE>(buf, plan.n, nlines, ld, Ns, mm, mn, tw, sign, tid, nthr); break;
            case 4: stockham_pass_inplace<4, PPT, WAVE>(buf, plan.n, nlines, ld, Ns, mm, mn, tw, sign, tid, nthr); break;
            case 5: stockham_pass_inplace<5, PPT, WAVE>(buf, plan.n, nlines, ld, Ns, mm, mn, tw, sign, tid, nthr); break;
            default: stockham_pass_inplace<8, PPT, WAVE>(buf, plan.n, nlines, ld, Ns, mm, mn, tw, sign, tid, nthr); break;
        }
        Ns *= R;
    }
}

__host__ __device__ inline bool fft_inplace_fits(const FFTPlan& plan, int nlines, int PPT, int nthr) {
    for (int f = 0; f < plan.nf; ++f) {
        const int R = plan.radix[f];
        if (R != 2 && R != 3 && R != 4 && R != 5 && R != 8) return false;
    }
    return (long long)nlines * plan.n <= (long long)PPT * nthr;
}

__device__ __forceinline__ double2* lds_fft(double2* a, double2* b, const FFTPlan& plan, int nlines, int ld, int sign, int tid,
                                            int nthr) {
    return lds_fft<false, false, double2>(a, b, plan, plan.tw, nlines, ld, sign, tid, nthr);
}

}  // namespace cf
