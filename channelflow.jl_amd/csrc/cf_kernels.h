// Kernel argument blocks and launchers (host-visible, no torch types).
#pragma once
#include <hip/hip_runtime.h>

#include "cf_common.h"

namespace cf {

// Per-bank (one implicit shift lambda, one viscosity) solver constants.
struct Bank {
    const double* facV;    // (B, Ny) elimination factors of nu D^2 - (lambda + nu kappa^2)
    const double* mconst;  // (B, MC_COUNT) influence/tau constants
    const double* resp;    // (Ny) mean-mode unit-gradient response c
    double resp_mean;      // bulk mean of c
    double lambda, nu;
};

enum { Y_STAGE = 0, Y_PROLOGUE, Y_SOLVE, Y_NEGLAP, Y_PRESSURE, Y_DCT_FWD, Y_DCT_INV, Y_DIVERGENCE, Y_DOT, Y_DISSIPATION, Y_PLANE_STATS };

// nonlinear forms (src/nonlinear.jl:113-193); values match CF_FORM_* of the C ABI
enum { FORM_ROTATING = 0, FORM_CONVECTIVE = 1, FORM_DIVERGENCE = 2 };

// Mode order of a Y_STAGE launch.  on = 0: block b owns modes b*M .. b*M+M-1.
// on = 1: the blocks below nyq_first own the nreg non-Nyquist modes (nxr per
// z row, z row zskip left out), so a launch whose regular blocks fill whole
// rounds of the machine has no partly filled last round; blocks from
// nyq_first on (nyq_blocks of them) handle the nnyq Nyquist modes
// elementwise.  xcd = 1: the regular blocks run in XCD-contiguous order
// (xcd_logical), so blocks sharing the cache lines of unaligned row segments
// share one L2.
struct YOrder {
    int on, xcd, nxr, zskip, nnyq;
    unsigned nyq_first, nyq_blocks;
    long long nreg;
};

struct YArgs {
    GridDims gd;
    FFTPlan dct;  // length 2P
    int ld;       // LDS line stride (complex elements)
    double norm;  // forward normalisation 1/(Nxp Nzp P)
    // CNRK2 stage
    int stage, store_Q;
    double lambda, nu, Aj, weight;
    // mean-flow constraint: bulk=1 -> (c0, c1) perturbation bulk targets;
    // bulk=0 -> prescribed uniform pressure gradient
    int bulk;
    double c0, c1;
    const double* base;  // Chebyshev coefficients of Ub (Ny)
    const double* curv;  // Ub'' coefficients (Ny)
    double curv_mean;    // bulk mean of Ub''
    const double* facP;  // (B, Ny) factors of D^2 - kappa^2 (Dirichlet; kappa^2=1 placeholder at (0,0))
    const double* facPN; // (B, Ny) factors of D^2 - kappa^2 with Neumann rows (pressure init)
    const double* cheb_tab;  // CHEB_TAB_ARRAYS x ntab per-degree constants (build_cheb_table layout)
    int ntab;                // entries per array (>= K L of the launch)
    Bank bank;
    int form;                 // FORM_*: which y-physical fields the epilogue emits
    int mixed;                // Y_STAGE / Y_PROLOGUE: Nphys and out6 hold FP32 complex (float2)
    double nl_sign;           // sign applied to the transformed nonlinear term
    const double2* Nphys[6];  // y-physical nonlinear term (Y_DIVERGENCE: the six products)
    const double2* Ncoef[3];  // nonlinear term already in coefficients (skip the DCT) or null
    const double2* Fcoef[3];  // forcing acceleration in coefficients, added to N (or null)
    double2* U[3];
    double2* P;
    double2* Q[3];
    double2* out6[6];  // y-physical outputs: (V, Omega), (V, dV/dy) or V by form; or inverse transforms
    const double2* Rin[3];
    double2* Rout[3];
    double* grad_out;  // 2 doubles (device) or null
    int nfields;
    double* partial;   // Y_DOT / Y_DISSIPATION: one partial sum per block (device)
    YOrder order;      // Y_STAGE: set by the launcher (zero: plain order)
    // Y_STAGE / Y_PROLOGUE: group layout, L lanes of K coefficients per mode
    // (yk = 0: the default layout of Ny; other kernels always use it)
    int yl, yk;
};

struct ZArgs {
    GridDims gd;
    FFTPlan plan;  // length Nzp
    int KT, ld, nfields;
    int ppt;           // > 0: in-place FFT in one LDS buffer with this many points per thread; 0: two buffers
    unsigned dz_mask;  // bit f: inverse pass multiplies field f by i*beta*l (z derivative)
    unsigned long long mul_kt;  // recip_mul(KT): index splits without integer division
    int mixed;                  // 1: fields are FP32 complex (float2) in the same buffers; in-place plans only
    const double2* in[9];
    double2* out[9];
};

// X_ROT_SPLIT: the rotating form's products with the velocity lines
// transformed first and held in registers, so the LDS region holds three
// lines instead of six (more wavefronts per CU)
// X_ROT_ASYNC: the rotating pass of X_ROTATIONAL (six LDS lines, in place)
// whose HBM loads go straight to LDS (global_load_lds), all six fields in
// flight at once; FP64 only
enum { X_ROTATIONAL = 0, X_INVERSE, X_FORWARD, X_CONVECTIVE, X_DIVERGENCE, X_GRADIENT, X_ROT_SPLIT, X_ROT_ASYNC };

struct XArgs {
    GridDims gd;
    FFTPlan plan;         // length Nxp/2 (even Nxp) or Nxp (odd)
    const double2* twx;   // exp(-2 pi i t / Nxp), t < Nxp
    int LPB, ld, mode, nfields;
    int waves;             // even Nxp: wavefronts (lines) per block, 1..4
    int ppt;               // even Nxp: in-place FFT points per lane (one LDS buffer), 0: two buffers
    unsigned long long mul_m, mul_nxh, mul_nx;  // recip_mul of Nxp/2, Nxh and Nxp/2+1 (index splits)
    int mixed;             // 1: in/out are FP32 complex (float2); even Nxp, in-place wave path only
    const double2* in[9];  // (k, zp, y) complex, Nxh per line
    double2* out[6];
    const double* rin[3];  // (x, zp, y) real
    double* rout[9];       // (x, zp, y) real outputs (X_INVERSE; X_GRADIENT: dV_i/dx_j at 3 i + j)
};

// Box filter of a fine spectral vector field onto a coarse grid (same box)
struct FilterArgs {
    GridDims f, c;   // fine and coarse grids
    double dx, dz;   // top-hat widths
    const double2* in[3];
    double2* out[3];
};
hipError_t launch_filter(const FilterArgs& a, hipStream_t st);

// Subgrid-force MLP (10 -> 64 -> 64 -> 64 -> 3): parameter count of theta
constexpr int MLP_NPARAM = 10448;
// norm: in_mean[10], in_scale[10], out_scale[3]
struct MLPInferArgs {
    const float* theta;          // FP32 parameters (biases read here)
    const unsigned short* wbf;   // bfloat16 copy of theta (weights read here)
    const float* norm;
    const double* grad[9];       // dV_i/dx_j on the padded points
    long long npts, plane;       // points, points per y plane (Nxp Nzp)
    int Ny;
    double* out[3];              // predicted force on the padded points
};
struct MLPTrainArgs {
    const float* theta;
    const unsigned short* wbf;
    const float* norm;
    const float* X;              // raw features, 10 per sample
    const float* T;              // raw targets, 3 per sample
    const int* idx;              // batch -> sample
    long long batch;
    float* part;                 // per-wave partials: waves x (MLP_NPARAM + 1)
};
hipError_t launch_mlp_infer(const MLPInferArgs& a, hipStream_t st);
hipError_t launch_mlp_features(const MLPInferArgs& a, float* X, float* T, hipStream_t st);  // targets from a.out
int mlp_train_waves(long long batch);
long long mlp_work_floats(long long batch);  // partial buffer of launch_mlp_grad
hipError_t launch_mlp_grad(const MLPTrainArgs& a, float* grad, hipStream_t st);  // grad: MLP_NPARAM + 1 (loss last)
hipError_t launch_mlp_adam(float* theta, unsigned short* wbf, float* m, float* v, const float* grad, float lr, int step,
                           hipStream_t st);

// zeroes every Fourier mode of three spectral fields but (k, l) = (0, 0):
// S coefficients per field, B modes per y (the (0,0) mode sits at n B)
hipError_t launch_keep_mean(double2* F0, double2* F1, double2* F2, long long S, long long B, hipStream_t st);

hipError_t launch_ykernel(int which, const YArgs& a, hipStream_t st);
unsigned y_blocks(const GridDims& gd);  // blocks of a y launch (partials of Y_DOT)
// out[0] = sum of part[0..n) in a fixed order (one block)
hipError_t launch_reduce(const double* part, int n, double* out, hipStream_t st);
// out[q] = sum over b < nblocks of part[b * nq + q], q < nq, in a fixed order
hipError_t launch_reduce_columns(const double* part, int nblocks, int nq, double* out, hipStream_t st);
// epsilon * N(0,1) samples into out3[3][n] (Philox4x32-10 + Box-Muller, counter = element)
hipError_t launch_gaussian(double* out3, long long n, double eps, unsigned long long seed, hipStream_t st);
constexpr int PLANE_STATS_Q = 9;  // per y: means of u, v, w; covariances uu vv ww uv uw vw
int y_group_lanes(int Ny);  // lanes L per mode of the wall-normal kernels (0: Ny unsupported)
int y_group_k(int Ny);      // coefficients K per lane (4; 6 above 256 points, 8 above 384)
// LDS bytes of a y launch in layout (L, K); 0 when (L, K) is not built or does
// not cover Ny coefficients
size_t y_lds_bytes_layout(int Ny, int ld, int L, int K);
size_t y_lds_bytes(int Ny, int ld);
hipError_t launch_zpass(int dir, const ZArgs& a, hipStream_t st);
size_t z_lds_bytes(const ZArgs& a);
int z_inplace_ppt(const FFTPlan& plan, int KT);  // 0: the plan needs two buffers
hipError_t launch_xpass(const XArgs& a, hipStream_t st);
size_t x_lds_bytes(const XArgs& a);
int x_phys_lines(int mode);  // physical lines per (zp, y) line of the odd-Nxp x pass
int x_max_waves(const XArgs& a);  // even Nxp: wavefronts per block that fit the LDS (0: none)
int x_inplace_ppt(const XArgs& a);  // even Nxp: in-place points per lane for the mode (0: two buffers)

// Batched ensemble launches (FP64): instance i of n owns blocks
// [start[i], start[i+1]) of an nblocks grid and reads its arguments from
// tab[i] (device memory).  All instances of one launch share the template
// parameters named in the call; lds is the largest instance's need.
hipError_t launch_ykernel_batch(int which, int L, const YArgs* tab, const unsigned* start, int n, unsigned nblocks,
                                size_t lds, hipStream_t st);  // Y_STAGE or Y_PROLOGUE
hipError_t launch_zpass_batch(int dir, int ppt, const ZArgs* tab, const unsigned* start, int n, unsigned nblocks,
                              size_t lds, hipStream_t st);
hipError_t launch_xpass_batch(int mode, int ppt, int waves, const XArgs* tab, const unsigned* start, int n,
                              unsigned nblocks, size_t lds, hipStream_t st);  // rotating (six-line or split) / convective, even Nxp
unsigned z_blocks(const ZArgs& a);  // grid of one instance's z pass
unsigned x_blocks(const XArgs& a);  // grid of one instance's even-Nxp x pass

// Fixed device scratch of the entry points that allocate nothing (32 MiB on
// every GPU the library runs on, held until the process ends; scratch.hip).
constexpr long long SCRATCH_BYTES = 32ll << 20;
constexpr long long SCRATCH_DOUBLES = SCRATCH_BYTES / 8;
// its device address on `device` (looked up once per device, then cached)
hipError_t scratch_address(int device, double** out);

}  // namespace cf
