// Learned subgrid closure: the box filter that defines the coarse (filtered)
// field, and the BF16/MFMA multilayer perceptron that predicts the subgrid
// force from the local velocity gradient.
//
// Filter (fine spectral -> coarse spectral, same Lx, Lz): for every retained
// coarse mode (k, l, n) the fine coefficient of the same wavenumbers times the
// top-hat transfer functions sinc(alpha k dx / 2) sinc(beta l dz / 2) (box
// widths dx, dz); coarse Nyquist planes zero; wall-normal Chebyshev series
// truncated to the coarse degrees.
#include "cf_device.h"
#include "cf_kernels.h"

namespace cf {

__device__ __forceinline__ double sinc_half(double w) {  // sin(w/2)/(w/2), 1 at 0
    const double h = 0.5 * w;
    return (h == 0.0) ? 1.0 : sin(h) / h;
}

__global__ void __launch_bounds__(256) filter_kernel(FilterArgs a) {
    const long long Sc = (long long)a.c.Nxh * a.c.Nz * a.c.Ny;
    const int comp = blockIdx.y;
    const double2* __restrict__ src = a.in[comp];
    double2* __restrict__ dst = a.out[comp];
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < Sc; idx += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(idx % a.c.Nxh);
        const long long r = idx / a.c.Nxh;
        const int iz = (int)(r % a.c.Nz);
        const int n = (int)(r / a.c.Nz);
        const int l = (iz <= a.c.Nz / 2) ? iz : iz - a.c.Nz;
        const bool nyq = ((a.c.Nx % 2 == 0) && k == a.c.Nx / 2) || ((a.c.Nz % 2 == 0) && iz == a.c.Nz / 2);
        double2 v = make_double2(0.0, 0.0);
        if (!nyq) {
            const int izf = (l >= 0) ? l : l + a.f.Nz;
            const double2 u = src[k + (size_t)a.f.Nxh * (izf + (size_t)a.f.Nz * n)];
            const double w = sinc_half(a.c.alpha * k * a.dx) * sinc_half(a.c.beta * l * a.dz);
            v = make_double2(u.x * w, u.y * w);
        }
        dst[idx] = v;
    }
}

hipError_t launch_filter(const FilterArgs& a, hipStream_t st) {
    const long long Sc = (long long)a.c.Nxh * a.c.Nz * a.c.Ny;
    const unsigned blocks = (unsigned)std::min<long long>((Sc + 255) / 256, 65535);
    hipLaunchKernelGGL(filter_kernel, dim3(blocks, 3), dim3(256), 0, st, a);
    return hipGetLastError();
}


// ---------------------------------------------------------------------------
// Subgrid-force MLP: 10 -> 64 -> 64 -> 64 -> 3, GELU (erf form) after the
// hidden layers.  Weights and activations bfloat16 on the matrix cores with
// FP32 accumulation: v_mfma_f32_16x16x32_bf16 (gfx950) for every product
// whose reduction length is a multiple of 32, v_mfma_f32_16x16x16_bf16 for
// the 16-wide ones (the padded input layer and the output-layer gradient);
// biases and master weights FP32.  theta layout: [W1 (64x16) | b1 | W2 (64x64) | b2 | W3 | b3 | W4 (16x64) | b4 (16)];
// padded input columns 10..15 and output rows 3..15 stay zero.
//
// MFMA lane layouts (lane l, lr = l % 16, lg = l / 16):
//   16x16x16: A (M x K): A[lr][4 lg + j], j < 4      B (K x N): B[4 lg + j][lr]
//   16x16x32: A (M x K): A[lr][8 lg + j], j < 8      B (K x N): B[8 lg + j][lr]
//   D (M x N): D[4 lg + r][lr], r < 4 (both)
// ---------------------------------------------------------------------------
namespace {
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MIN = 10, MK = 16, MH = 64, MOUT = 3, MKO = 16;
constexpr int OW1 = 0, OB1 = OW1 + MH * MK, OW2 = OB1 + MH, OB2 = OW2 + MH * MH, OW3 = OB2 + MH, OB3 = OW3 + MH * MH,
              OW4 = OB3 + MH, OB4 = OW4 + MKO * MH, NPAR = OB4 + MKO;
static_assert(NPAR == MLP_NPARAM, "parameter layout");

__device__ __forceinline__ unsigned short f2bf(float f) {  // round to nearest even (finite values)
    const unsigned u = __float_as_uint(f);
    return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * expf(-0.5f * z * z) * 0.39894228040143268f;
}
__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x4 ld4(const unsigned short* p) { return *reinterpret_cast<const bf16x4*>(p); }
// 8 consecutive bf16 (16-byte aligned: every row stride and offset used is a
// multiple of 8 elements)
__device__ __forceinline__ bf16x8 ld8(const unsigned short* p) { return *reinterpret_cast<const bf16x8*>(p); }
__device__ __forceinline__ bf16x8 ld8s(const unsigned short* p, int stride) {
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (short)p[j * stride];
    return v;
}
__device__ __forceinline__ bf16x4 ld4s(const unsigned short* p, int stride) {
    bf16x4 v;
    v[0] = (short)p[0];
    v[1] = (short)p[stride];
    v[2] = (short)p[2 * stride];
    v[3] = (short)p[3 * stride];
    return v;
}
__device__ __forceinline__ void copy_weights(unsigned short* W, const unsigned short* __restrict__ wbf) {
    for (int i = threadIdx.x; i < NPAR; i += blockDim.x) W[i] = wbf[i];
}
// normalised features of padded point p: dV_i/dx_j (9) and wall distance 1 - |y|
__device__ __forceinline__ void point_features(const double* const* g, long long p, long long plane, int Ny,
                                               const float* norm, float* x) {
    const int j = (int)(p / plane);
    const double y = cos(M_PI * j / (Ny - 1));
#pragma unroll
    for (int c = 0; c < 9; ++c) x[c] = ((float)g[c][p] - norm[c]) * norm[MIN + c];
    x[9] = ((float)(1.0 - fabs(y)) - norm[9]) * norm[MIN + 9];
}

// one hidden layer for a wave: rows x 64 outputs from rows x K inputs (bf16)
// in `in` (row stride K), weights W (64 x K), bias b; writes z (fp32, stride
// 64, if zs) and bf16(gelu(z)) into `out` (stride 64)
template <int RB, int KS>
__device__ __forceinline__ void hidden_layer(const unsigned short* in, const unsigned short* W, const float* b, float* zs,
                                             unsigned short* out, int lr, int lg) {
    constexpr int K = KS * 16;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int ob = 0; ob < 4; ++ob) {
            f32x4 c = {0.f, 0.f, 0.f, 0.f};
            if constexpr (KS % 2 == 0) {
#pragma unroll
                for (int ks = 0; ks < KS / 2; ++ks)
                    c = mfma32(ld8(in + (rb * 16 + lr) * K + ks * 32 + 8 * lg), ld8(W + (ob * 16 + lr) * K + ks * 32 + 8 * lg),
                               c);
            } else {
#pragma unroll
                for (int ks = 0; ks < KS; ++ks)
                    c = mfma16(ld4(in + (rb * 16 + lr) * K + ks * 16 + 4 * lg),
                               ld4(W + (ob * 16 + lr) * K + ks * 16 + 4 * lg), c);
            }
            const float bias = b[ob * 16 + lr];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = rb * 16 + 4 * lg + r, col = ob * 16 + lr;
                const float z = c[r] + bias;
                if (zs) zs[row * MH + col] = z;
                out[row * MH + col] = f2bf(gelu(z));
            }
        }
}
}  // namespace

// Inference: a wave owns MLP_INFER_RB row blocks of 16 consecutive padded
// points; a block holds the bf16 weights (21 KB) and MLP_INFER_WAVES waves of
// two activation buffers (16 RB x 64 bf16 each).  Each wave's layers are a
// serial chain of matrix products and GELUs, so more and shorter waves run
// faster: 16 points per wave in blocks of eight waves (53 KB) fit three
// blocks, 24 waves, in a CU.  Measured at 32x33x32 (76 k points): 64 points
// per wave in blocks of four (86 KB, one block per CU) 67 us; in blocks of
// two 48 us; 32 points per wave in blocks of four about 25 us less per step
// than that; 16 points per wave in blocks of eight, 8 us less again.
constexpr int MLP_INFER_RB = 1, MLP_INFER_WAVES = 8, MLP_INFER_PW = 16 * MLP_INFER_RB;
__global__ void __launch_bounds__(64 * MLP_INFER_WAVES) mlp_infer_kernel(MLPInferArgs a) {
    extern __shared__ double2 smem_raw[];
    constexpr int RB = MLP_INFER_RB, PW = MLP_INFER_PW;
    unsigned short* W = reinterpret_cast<unsigned short*>(smem_raw);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lg = lane >> 4;
    unsigned short* bufA = W + NPAR + (NPAR & 3 ? 4 - (NPAR & 3) : 0) + wave * 2 * PW * MH;
    unsigned short* bufB = bufA + PW * MH;
    copy_weights(W, a.wbf);
    __syncthreads();
    const long long p0 = ((long long)blockIdx.x * MLP_INFER_WAVES + wave) * PW;
    if (lane < PW) {
        float x[MIN];
        const long long p = p0 + lane;
        if (p < a.npts) point_features(a.grad, p, a.plane, a.Ny, a.norm, x);
        else
            for (int c = 0; c < MIN; ++c) x[c] = 0.f;
#pragma unroll
        for (int c = 0; c < MK; ++c) bufA[lane * MK + c] = c < MIN ? f2bf(x[c]) : (unsigned short)0;
    }
    wave_sync();
    hidden_layer<RB, 1>(bufA, W + OW1, a.theta + OB1, nullptr, bufB, lr, lg);
    wave_sync();
    hidden_layer<RB, 4>(bufB, W + OW2, a.theta + OB2, nullptr, bufA, lr, lg);
    wave_sync();
    hidden_layer<RB, 4>(bufA, W + OW3, a.theta + OB3, nullptr, bufB, lr, lg);
    wave_sync();
    // output layer: column lr < 3 is a force component
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
            c = mfma32(ld8(bufB + (rb * 16 + lr) * MH + ks * 32 + 8 * lg), ld8(W + OW4 + lr * MH + ks * 32 + 8 * lg), c);
        if (lr < MOUT) {
            const float bias = a.theta[OB4 + lr], scale = a.norm[2 * MIN + lr];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const long long p = p0 + rb * 16 + 4 * lg + r;
                if (p < a.npts) a.out[lr][p] = (double)((c[r] + bias) * scale);
            }
        }
    }
}

// Raw training features (10 per point) and targets (3 per point from the
// fields in a.out, if T is given)
__global__ void __launch_bounds__(256) mlp_features_kernel(MLPInferArgs a, float* X, float* T) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.npts) return;
    const int j = (int)(p / a.plane);
    const double y = cos(M_PI * j / (a.Ny - 1));
#pragma unroll
    for (int c = 0; c < 9; ++c) X[p * MIN + c] = (float)a.grad[c][p];
    X[p * MIN + 9] = (float)(1.0 - fabs(y));
    if (T)
#pragma unroll
        for (int c = 0; c < MOUT; ++c) T[p * MOUT + c] = (float)a.out[c][p];
}

// Training: forward and backward of a wave's 32 batch samples; the wave
// writes its full parameter-gradient partial and its loss partial.
__global__ void __launch_bounds__(128) mlp_train_kernel(MLPTrainArgs a) {
    extern __shared__ double2 smem_raw[];
    unsigned short* W = reinterpret_cast<unsigned short*>(smem_raw);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lg = lane >> 4;
    constexpr int R = 32;  // samples per wave
    // per-wave region: Z1..Z3 (fp32 R x 64), H1..H3 (bf16 R x 64), X (bf16 R x 16), G (bf16 R x 64)
    float* Z = reinterpret_cast<float*>(smem_raw + (NPAR * 2 + 15) / 16) + wave * (3 * R * MH + (3 * R * MH + R * MK + R * MH) / 2);
    unsigned short* Hb = reinterpret_cast<unsigned short*>(Z + 3 * R * MH);
    unsigned short* X = Hb + 3 * R * MH;
    unsigned short* G = X + R * MK;
    float *Z1 = Z, *Z2 = Z + R * MH, *Z3 = Z + 2 * R * MH;
    unsigned short *H1 = Hb, *H2 = Hb + R * MH, *H3 = Hb + 2 * R * MH;
    copy_weights(W, a.wbf);
    __syncthreads();
    const long long w = (long long)blockIdx.x * 2 + wave;  // global wave = partial row
    const long long b0 = w * R;
    if (b0 >= a.batch) return;  // a wave past the batch owns no partial row (no block barrier follows)
    // inputs: lanes 0..31 load one sample each
    float tl[MOUT] = {0.f, 0.f, 0.f};
    if (lane < R) {
        const long long b = b0 + lane;
        float x[MIN];
        if (b < a.batch) {
            const long long s = a.idx[b];
#pragma unroll
            for (int c = 0; c < MIN; ++c) x[c] = (a.X[s * MIN + c] - a.norm[c]) * a.norm[MIN + c];
        } else {
#pragma unroll
            for (int c = 0; c < MIN; ++c) x[c] = 0.f;
        }
#pragma unroll
        for (int c = 0; c < MK; ++c) X[lane * MK + c] = c < MIN ? f2bf(x[c]) : (unsigned short)0;
    }
    wave_sync();
    hidden_layer<2, 1>(X, W + OW1, a.theta + OB1, Z1, H1, lr, lg);
    wave_sync();
    hidden_layer<2, 4>(H1, W + OW2, a.theta + OB2, Z2, H2, lr, lg);
    wave_sync();
    hidden_layer<2, 4>(H2, W + OW3, a.theta + OB3, Z3, H3, lr, lg);
    wave_sync();
    float* part = a.part + w * (long long)(NPAR + 1);
    // output layer and dY = 2 (y - t) / batch (columns >= 3 zero)
    float loss = 0.f;
    float dbias = 0.f;
    const float inv = 2.0f / (float)a.batch;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
        f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
            c = mfma32(ld8(H3 + (rb * 16 + lr) * MH + ks * 32 + 8 * lg), ld8(W + OW4 + lr * MH + ks * 32 + 8 * lg), c);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = rb * 16 + 4 * lg + r;
            const long long b = b0 + row;
            float d = 0.f;
            if (lr < MOUT && b < a.batch) {
                const long long s = a.idx[b];
                const float t = a.T[s * MOUT + lr] / a.norm[2 * MIN + lr];
                const float e = c[r] + a.theta[OB4 + lr] - t;
                loss += e * e;
                d = inv * e;
            }
            dbias += d;
            G[row * MKO + lr] = f2bf(d);
        }
    }
    // bias gradient of the output layer and the loss: sums over the 4 lane groups
    dbias += __shfl_xor(dbias, 16);
    dbias += __shfl_xor(dbias, 32);
    loss += __shfl_xor(loss, 1);
    loss += __shfl_xor(loss, 2);
    loss += __shfl_xor(loss, 4);
    loss += __shfl_xor(loss, 8);
    loss += __shfl_xor(loss, 16);
    loss += __shfl_xor(loss, 32);
    if (lg == 0) part[OB4 + lr] = dbias;
    if (lane == 0) part[NPAR] = loss / (float)a.batch;
    wave_sync();
    // dW4[o][k] = sum_p dY[p][o] H3[p][k]
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        f32x4 c = {0.f, 0.f, 0.f, 0.f};
        c = mfma32(ld8s(G + 8 * lg * MKO + lr, MKO), ld8s(H3 + 8 * lg * MH + nb * 16 + lr, MH), c);
#pragma unroll
        for (int r = 0; r < 4; ++r) part[OW4 + (4 * lg + r) * MH + nb * 16 + lr] = c[r];
    }
    // back through layers 3, 2, 1
    const unsigned short* Wl[3] = {W + OW4, W + OW3, W + OW2};
    const float* Zl[3] = {Z3, Z2, Z1};
    const unsigned short* Hin[3] = {H2, H1, X};
    const int OWl[3] = {OW3, OW2, OW1}, OBl[3] = {OB3, OB2, OB1};
#pragma unroll
    for (int L = 0; L < 3; ++L) {
        // dH[p][k] = sum_o dZnext[p][o] W[o][k] (G: R x KN, W: KN x 64)
        f32x4 dh[2][4];
        const int KN = (L == 0) ? MKO : MH;
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                f32x4 c = {0.f, 0.f, 0.f, 0.f};
                if (KN == MKO)
                    c = mfma16(ld4(G + (rb * 16 + lr) * KN + 4 * lg), ld4s(Wl[L] + 4 * lg * MH + nb * 16 + lr, MH), c);
                else
                    for (int ks = 0; ks < KN / 32; ++ks)
                        c = mfma32(ld8(G + (rb * 16 + lr) * KN + ks * 32 + 8 * lg),
                                   ld8s(Wl[L] + (ks * 32 + 8 * lg) * MH + nb * 16 + lr, MH), c);
                dh[rb][nb] = c;
            }
        wave_sync();  // every lane has read G before it is overwritten
        // dZ = dH * gelu'(Z); bias gradient from the unrounded dZ; G <- bf16(dZ)
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            float db = 0.f;
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = rb * 16 + 4 * lg + r, col = nb * 16 + lr;
                    const float dz = dh[rb][nb][r] * gelu_grad(Zl[L][row * MH + col]);
                    db += dz;
                    G[row * MH + col] = f2bf(dz);
                }
            db += __shfl_xor(db, 16);
            db += __shfl_xor(db, 32);
            if (lg == 0) part[OBl[L] + nb * 16 + lr] = db;
        }
        wave_sync();
        // dW[o][k] = sum_p dZ[p][o] Hin[p][k]  (Hin: R x KI)
        const int KI = (L == 2) ? MK : MH;
#pragma unroll
        for (int ob = 0; ob < 4; ++ob)
            for (int nb = 0; nb < KI / 16; ++nb) {
                f32x4 c = {0.f, 0.f, 0.f, 0.f};
                c = mfma32(ld8s(G + 8 * lg * MH + ob * 16 + lr, MH), ld8s(Hin[L] + 8 * lg * KI + nb * 16 + lr, KI), c);
#pragma unroll
                for (int r = 0; r < 4; ++r) part[OWl[L] + (ob * 16 + 4 * lg + r) * KI + nb * 16 + lr] = c[r];
            }
    }
}

// Two-level fixed-order reduction of the per-wave partials:
// chunk[c][j] = sum_{w in chunk c} part[w][j], then grad[j] = sum_c chunk[c][j]
// (grad[NPAR] = loss).  Chunks of MLP_RCHUNK waves run in parallel.
constexpr int MLP_RCHUNK = 32;
__global__ void __launch_bounds__(256) mlp_reduce_chunks_kernel(const float* __restrict__ part, int nw,
                                                                float* __restrict__ chunk) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > NPAR) return;
    const int c = blockIdx.y, w0 = c * MLP_RCHUNK, w1 = min(nw, w0 + MLP_RCHUNK);
    float s = 0.f;
    for (int w = w0; w < w1; ++w) s += part[(long long)w * (NPAR + 1) + j];
    chunk[(long long)c * (NPAR + 1) + j] = s;
}
__global__ void __launch_bounds__(256) mlp_reduce_kernel(const float* __restrict__ chunk, int nc, float* __restrict__ grad) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > NPAR) return;
    float s = 0.f;
    for (int c = 0; c < nc; ++c) s += chunk[(long long)c * (NPAR + 1) + j];
    grad[j] = s;
}

__global__ void __launch_bounds__(256) mlp_adam_kernel(float* theta, unsigned short* wbf, float* m, float* v,
                                                       const float* __restrict__ grad, float lr, float c1, float c2) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= NPAR) return;
    const float g = grad[j];
    const float mj = 0.9f * m[j] + 0.1f * g;
    const float vj = 0.999f * v[j] + 0.001f * g * g;
    m[j] = mj;
    v[j] = vj;
    const float th = theta[j] - lr * (mj / c1) / (sqrtf(vj / c2) + 1e-8f);
    theta[j] = th;
    wbf[j] = f2bf(th);
}

size_t mlp_infer_lds() { return (size_t)(NPAR + 4) * 2 + MLP_INFER_WAVES * 2 * MLP_INFER_PW * MH * 2; }
size_t mlp_train_lds() {
    const size_t wave = 3 * 32 * MH * 4 + (3 * 32 * MH + 32 * MK + 32 * MH) * 2;
    return ((size_t)(NPAR * 2 + 15) / 16) * 16 + 2 * wave;
}

hipError_t launch_mlp_infer(const MLPInferArgs& a, hipStream_t st) {
    const size_t lds = mlp_infer_lds();
    (void)hipFuncSetAttribute((const void*)mlp_infer_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    constexpr int pts = MLP_INFER_PW * MLP_INFER_WAVES;
    const unsigned blocks = (unsigned)((a.npts + pts - 1) / pts);
    hipLaunchKernelGGL(mlp_infer_kernel, dim3(blocks), dim3(64 * MLP_INFER_WAVES), lds, st, a);
    return hipGetLastError();
}

hipError_t launch_mlp_features(const MLPInferArgs& a, float* X, float* T, hipStream_t st) {
    const unsigned blocks = (unsigned)((a.npts + 255) / 256);
    hipLaunchKernelGGL(mlp_features_kernel, dim3(blocks), dim3(256), 0, st, a, X, T);
    return hipGetLastError();
}

int mlp_train_waves(long long batch) { return (int)((batch + 31) / 32); }
// work floats: wave partials, then the chunk sums
long long mlp_work_floats(long long batch) {
    const long long nw = mlp_train_waves(batch), nc = (nw + MLP_RCHUNK - 1) / MLP_RCHUNK;
    return (nw + nc) * (NPAR + 1);
}

hipError_t launch_mlp_grad(const MLPTrainArgs& a, float* grad, hipStream_t st) {
    const size_t lds = mlp_train_lds();
    (void)hipFuncSetAttribute((const void*)mlp_train_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    const int nw = mlp_train_waves(a.batch), nc = (nw + MLP_RCHUNK - 1) / MLP_RCHUNK;
    float* chunk = a.part + (long long)nw * (NPAR + 1);
    hipLaunchKernelGGL(mlp_train_kernel, dim3((unsigned)((nw + 1) / 2)), dim3(128), lds, st, a);
    hipLaunchKernelGGL(mlp_reduce_chunks_kernel, dim3((NPAR + 1 + 255) / 256, (unsigned)nc), dim3(256), 0, st, a.part, nw,
                       chunk);
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3((NPAR + 1 + 255) / 256), dim3(256), 0, st, chunk, nc, grad);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(float* theta, unsigned short* wbf, float* m, float* v, const float* grad, float lr, int step,
                           hipStream_t st) {
    const float c1 = 1.0f - powf(0.9f, (float)step), c2 = 1.0f - powf(0.999f, (float)step);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((NPAR + 255) / 256), dim3(256), 0, st, theta, wbf, m, v, grad, lr, c1, c2);
    return hipGetLastError();
}

// Plane-mean part of a spectral force: every mode but (k, l) = (0, 0) zeroed
__global__ void __launch_bounds__(256) keep_mean_kernel(double2* F0, double2* F1, double2* F2, long long S, long long B) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= S || i % B == 0) return;
    const double2 z = make_double2(0.0, 0.0);
    F0[i] = z;
    F1[i] = z;
    F2[i] = z;
}

hipError_t launch_keep_mean(double2* F0, double2* F1, double2* F2, long long S, long long B, hipStream_t st) {
    if (S <= 0 || B <= 0 || (S + 255) / 256 > 0x7FFFFFFFll) return hipErrorInvalidValue;
    hipLaunchKernelGGL(keep_mean_kernel, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, st, F0, F1, F2, S, B);
    return hipGetLastError();
}

}  // namespace cf
