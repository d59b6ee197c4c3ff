/*
 * channelflow_hip.h — C ABI of the MI355X (gfx950) CNRK2 step library
 * (libcfhip.so).  Drop-in device backend for ChannelFlow.jl's hot path:
 * the three-stage CNRK2 `step!` of plane Couette/Poiseuille flow.
 *
 * Conventions (identical to the reference's storage):
 *   - spectral field: (k, l, n) column-major, k fastest, interleaved
 *     complex128, size Nxh*Nz*Ny with Nxh = Nx/2+1
 *     (src/fields/spectralfield.jl:16-29, src/grids.jl:71-74)
 *   - padded physical field: (x, z, y) column-major float64, size
 *     Nxp*Nzp*Ny, Nxp = ceil(3Nx/2), Nzp = ceil(3Nz/2)
 *     (src/fields/physicalfield.jl:13-57, src/grids.jl:49-63)
 *   - all field pointers are DEVICE pointers; `stream` is a hipStream_t
 *     (NULL = default stream).  Calls are asynchronous on that stream
 *     unless stated otherwise.
 *   - status: 0 = success; CF_ERR_ARGUMENT mirrors Julia's ArgumentError,
 *     CF_ERR_DIMENSION DimensionMismatch; all validation happens before any
 *     work is enqueued; cf_last_error() describes the last failure.
 *   - a problem handle is not thread-safe (one trajectory per handle;
 *     src/timesteppers/cnrk2.jl:147); distinct handles may run concurrently
 *     on distinct streams.
 */
#ifndef CHANNELFLOW_HIP_H
#define CHANNELFLOW_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define CF_OK 0
#define CF_ERR_ARGUMENT 1
#define CF_ERR_DIMENSION 2
#define CF_ERR_DEVICE 3
#define CF_ERR_UNSUPPORTED 4

/* nonlinear forms (src/nonlinear.jl:214-224) */
#define CF_FORM_ROTATING 0
#define CF_FORM_CONVECTIVE 1
#define CF_FORM_DIVERGENCE 2

/* arithmetic of the step (cf_config.precision).  MIXED: the y-physical
 * fields, the x/z FFTs and the nonlinear products are FP32; the state U, P,
 * the accumulator Q, the Chebyshev transforms' arithmetic, the Helmholtz-tau
 * and influence-matrix solves and the mean mode stay FP64.  Only the time
 * step (cf_step, cf_advance, cf_advance_to) uses it; every other call is FP64. */
#define CF_PRECISION_FP64 0
#define CF_PRECISION_MIXED 1

/* mean-flow constraint (src/problem.jl:81-87) */
#define CF_CONSTRAINT_PRESSURE_GRADIENT 0
#define CF_CONSTRAINT_BULK_VELOCITY 1

typedef struct cf_problem cf_problem;

typedef struct cf_config {
    int Nx, Ny, Nz;          /* Grid(Nx, Ny, Nz, Lx, Lz), src/grids.jl:27-31 */
    double Lx, Lz;
    double nu, dt;           /* CNRK2(grid, baseflow, nu, dt), cnrk2.jl:58-83 */
    const double* baseflow;  /* HOST: Ny base-velocity values at the Lobatto
                                points y_j = cos(pi j/(Ny-1)) (problem.jl:105) */
    int form;                /* CF_FORM_*: nonlinear form of the problem */
    int constraint;          /* CF_CONSTRAINT_*: the problem default */
    double constraint_values[2];
    int device;              /* HIP device ordinal */
    int precision;           /* CF_PRECISION_*; MIXED needs even Nxp and a
                                rotating or convective form */
} cf_config;

/* ChannelFlowProblem(grid, profile, nu, dt; form, pressuregradient|bulkvelocity)
 * (src/problem.jl:69-107; CouetteFlow/PoiseuilleFlow 122-144) followed by
 * adapt(CuArray, problem) (ext/ChannelFlowCUDAExt.jl:18-32): factors all three
 * stage Stokes solvers on the host in FP64 and uploads them. */
int cf_problem_create(const cf_config* cfg, cf_problem** out);
int cf_problem_destroy(cf_problem* h);
const char* cf_last_error(void);

/* sizes: out[0..7] = Nxh, Nz, Ny (spectral), Nxp, Nzp, Ny (padded physical),
 * spectral complex count, physical real count */
int cf_problem_sizes(const cf_problem* h, long long out[8]);

/* step!(scheme, nlterm, U, P, t; forcing=NoForcing(), pressuregradient | bulkvelocity)
 * (src/timesteppers/cnrk2.jl:149-207).  Advances U = (U0, U1, U2) and stage
 * pressure P in place by dt; *t_out = t + dt.  constraint = -1 uses the
 * problem's own constraint, otherwise CF_CONSTRAINT_* with values[2]
 * (pressure gradient (dPdx, dPdz) or TOTAL bulk velocity (Ubulk, Wbulk)). */
int cf_step(cf_problem* h, double* U0, double* U1, double* U2, double* P, double t, int constraint,
            const double* values, double* t_out, void* stream);

/* nsteps consecutive step! calls (same semantics, one enqueue) */
int cf_advance(cf_problem* h, double* U0, double* U1, double* U2, double* P, double t, int nsteps,
               int constraint, const double* values, double* t_out, void* stream);

/* Integrate from t to T as Flows.jl does with the problem's constant step
 * (ext/ChannelFlowFlowsExt.jl:36-52): step! with dt while a full step fits,
 * then one shortened step of the remaining duration (T - t - n dt > 1e-10 dt)
 * whose stage solvers are factored for that duration and cached on the
 * handle.  *t_out = T. */
int cf_advance_to(cf_problem* h, double* U0, double* U1, double* U2, double* P, double t, double T, int constraint,
                  const double* values, double* t_out, void* stream);

/* cf_advance_to that allocates no device memory (no counterpart in the
 * reference).  For callers that cannot call hipMalloc: code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation, e.g. a
 * Newton or periodic-orbit search that varies T on every call.  Same
 * arguments and results (bit for bit) as cf_advance_to.  The shortened final
 * step's three stage solver banks are built on the host and copied into a
 * fixed 32 MiB device array that every GPU the library runs on holds until
 * the process ends; the call then waits for `stream` once (the copies).
 * Grids whose three banks exceed it (3 x 8 x (B (Ny + 16) + Ny) bytes,
 * B = (Nx/2+1) Nz; e.g. 128x97x128 needs 22.6 MiB) return CF_ERR_UNSUPPORTED
 * before anything is enqueued.  All calls on one GPU share that array: calls
 * that overlap on one GPU (other streams, other handles) race. */
int cf_advance_to_noalloc(cf_problem* h, double* U0, double* U1, double* U2, double* P, double t, double T,
                          int constraint, const double* values, double* t_out, void* stream);

/* Measurement variant of cf_step (no counterpart in the reference): one step
 * with HIP events recorded on `stream` around every launch; waits for the
 * step, then ms[0..4] = device milliseconds spent in the prologue (y), the
 * z-inverse passes, the x passes, the z-forward passes and the y stages. */
int cf_step_timed(cf_problem* h, double* U0, double* U1, double* U2, double* P, double t, int constraint,
                  const double* values, double* t_out, double ms[5], void* stream);

/* Velocity-gradient tensor of the total velocity V = U + Ub e_x at the padded
 * physical points (the derivatives of operators.jl:14-83 and _batchdiff!,
 * batchedinfluence.jl:32-43, each followed by InverseFFT!, ffts.jl:151-166):
 * out (DEVICE) holds 9 padded physical fields, field 3 i + j = dV_i/dx_j with
 * x_j = (x, y, z).  Needs an even padded Nx.  Uses the problem's workspaces. */
int cf_velocity_gradient(cf_problem* h, const double* U0, const double* U1, const double* U2, double* out, void* stream);

/* Box filter for the learned closure's data (no counterpart in the
 * reference; the filtered-DNS reference of the closure workload): spectral
 * U of `fine` -> spectral Ubar of `coarse` (same Lx, Lz; coarse strictly
 * coarser in x and z, not finer in y).  Each retained coarse mode takes the
 * fine coefficient of the same wavenumbers times the top-hat transfer
 * functions sin(alpha k dx/2)/(alpha k dx/2) * sin(beta l dz/2)/(beta l dz/2);
 * coarse Nyquist planes are zero; Chebyshev degrees above the coarse Ny-1
 * are dropped. */
int cf_box_filter(cf_problem* fine, cf_problem* coarse, const double* U0, const double* U1, const double* U2, double* V0,
                  double* V1, double* V2, double dx, double dz, void* stream);

/* Learned subgrid closure (no counterpart in the reference; the closure
 * workload's network).  MLP 10 -> 64 -> 64 -> 64 -> 3, GELU (erf form) after
 * the hidden layers, bfloat16 weights and activations on the matrix cores
 * with FP32 accumulation.  theta: CF_MLP_NPARAM floats, layout
 * [W1 64x16 | b1 64 | W2 64x64 | b2 64 | W3 64x64 | b3 64 | W4 16x64 | b4 16]
 * (row-major, output index first; input columns 10..15 of W1 and output
 * rows 3..15 of W4 zero); wbf: its bfloat16 copy (bits, round to nearest
 * even); norm: in_mean[10], in_scale[10], out_scale[3].  Inputs of a point:
 * dV_i/dx_j (9, as cf_velocity_gradient) and the wall distance 1 - |y|,
 * normalised (x - mean) * scale; output y * out_scale.  All pointers are
 * DEVICE pointers; the grid is the padded (Nxp, Nzp, Ny) point grid. */
#define CF_MLP_NPARAM 10448
/* raw features X (10 per point) and, if target3 is given, raw targets T
 * (3 per point) from 9 gradient fields and 3 target force fields */
int cf_mlp_features(const double* grad9, const double* target3, int Nxp, int Nzp, int Ny, float* X, float* T, void* stream);
/* predicted force (3 padded physical fields) from 9 gradient fields */
int cf_mlp_infer(const float* theta, const unsigned short* wbf, const float* norm, const double* grad9, int Nxp, int Nzp,
                 int Ny, double* force3, void* stream);
/* floats of the work buffer cf_mlp_grad needs for a batch */
long long cf_mlp_work_floats(long long batch);
/* gradient of L = (1/batch) sum_b sum_{i<3} (y_bi - T_{idx_b,i}/out_scale_i)^2
 * over the samples idx[0..batch) of (X, T): grad[0..CF_MLP_NPARAM) and the
 * loss at grad[CF_MLP_NPARAM]; deterministic (fixed-order reduction) */
int cf_mlp_grad(const float* theta, const unsigned short* wbf, const float* norm, const float* X, const float* T,
                const int* idx, long long batch, float* work, float* grad, void* stream);
/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8; step counts from 1) on
 * theta, refreshing wbf */
int cf_mlp_adam(float* theta, unsigned short* wbf, float* m, float* v, const float* grad, float lr, int step, void* stream);

/* forcing(tstage, U, F) (cnrk2.jl:173-175, forcing.jl:3-11).  During each
 * stage of cf_step / cf_advance / cf_advance_to / cf_step_timed, after the
 * nonlinear term is enqueued, the library calls fn(user, tstage, U0, U1, U2,
 * F0, F1, F2, stream) on the host with tstage = t + (0, 1/3, 3/4) dt.  fn
 * must ENQUEUE on `stream` device work that overwrites F0..F2 (spectral
 * fields owned by the caller, given here) with the forcing acceleration; it
 * must not read U on the host, and must not call entry points on the same
 * handle that use its workspaces (transforms, nonlinear term, gradient,
 * projection, diagnostics): the stage's nonlinear term is live in them.  A
 * second handle on the same grid serves such work.  The library then adds F into the nonlinear
 * term N before the bulk-gradient evaluation and the Q update, as the
 * reference's forcing adds into N.  fn returns 0 on success; a nonzero
 * return stops the step with CF_ERR_ARGUMENT (earlier stages may already
 * be enqueued).  fn = NULL restores NoForcing.  Problems with a forcing
 * cannot join an ensemble. */
typedef int (*cf_forcing_fn)(void* user, double tstage, const double* U0, const double* U1, const double* U2,
                             double* F0, double* F1, double* F2, void* stream);
int cf_set_forcing(cf_problem* h, cf_forcing_fn fn, void* user, double* F0, double* F1, double* F2);

/* The learned closure as a device-side forcing (no counterpart in the
 * reference, whose forcing is a host callable; for callers that want the
 * closure's step replayed from a graph, without a host callback per stage).
 * While registered, every stage of cf_step / cf_advance / cf_advance_to /
 * cf_step_timed computes F0..F2 on the device before the stage kernel, with
 * the kernels a host forcing would enqueue: cf_velocity_gradient of the stage
 * velocity into grad9 (9 padded physical fields), cf_mlp_infer into force3
 * (3 padded physical fields) and cf_forward3 of force3 into F0..F2 (spectral),
 * all on `helper`'s workspaces; mean_only != 0 then zeroes every mode but
 * (k, l) = (0, 0).  These kernels run on a stream of the problem's own,
 * forked from the step's stream and joined back before each stage kernel
 * (a parallel branch of the step's graph).  helper: another problem on the same grid and device
 * (the stepping problem's workspaces hold the stage's nonlinear term); it
 * and every buffer must stay alive while registered.  theta, wbf, norm as
 * for cf_mlp_infer; they are read at every stage, so training between steps
 * takes effect.  The plain step stays graph-replayed (one graph per field
 * set, constraint and closure).  theta = NULL removes the closure.  Exclusive
 * with cf_set_forcing; a problem with a closure cannot join an ensemble. */
int cf_set_closure(cf_problem* h, cf_problem* helper, const float* theta, const unsigned short* wbf, const float* norm,
                   double* grad9, double* force3, double* F0, double* F1, double* F2, int mean_only);

/* nonlinear(tstage, U, N) (cnrk2.jl:121-126,172): the step's nonlinear term
 * supplied by the caller in place of the problem's form.  At each stage of
 * cf_step / cf_advance / cf_advance_to / cf_step_timed the library calls
 * fn(user, tstage, U0, U1, U2, N0, N1, N2, stream) on the host with
 * tstage = t + (0, 1/3, 3/4) dt, before the forcing callback.  fn must
 * ENQUEUE on `stream` device work that OVERWRITES N0..N2 (spectral fields
 * owned by the caller, given here) with the signed nonlinear acceleration of
 * the total velocity U + Ub e_x, the term the stage adds to Q as is.  The
 * same restrictions as for the forcing callback apply (no host reads of U,
 * no calls on this handle that use its workspaces).  The step then runs no
 * transforms of its own.  A nonzero return stops the step with
 * CF_ERR_ARGUMENT.  fn = NULL restores the problem's form.  Such a problem
 * cannot join an ensemble. */
int cf_set_nonlinear(cf_problem* h, cf_forcing_fn fn, void* user, double* N0, double* N1, double* N2);

/* Tile autotuner (no counterpart in the reference; a performance setting):
 * times the nonlinear term's z passes for each tile width and its x pass for
 * each variant and wavefront count on the problem's workspaces, on `stream`,
 * and keeps the fastest for the FP64 step (and, for a mixed-precision
 * problem, the fastest FP32 x pass of its step).  Results of every call are
 * identical for any tiling.  Synchronises the stream. */
int cf_problem_tune(cf_problem* h, void* stream);
/* current tiles: out[0] z tile width, out[1] z in-place points per thread
 * (0: two buffers), out[2] x wavefronts per block */
int cf_problem_tiles(const cf_problem* h, int out[3]);

/* Tile choice of the FP64 step, to read and to set (no counterpart in the
 * reference; lets a caller keep a tuning across runs, and the tests run every
 * variant): tiles[0] z tile width (1..min(32, Nx/2+1)), tiles[1] z in-place
 * points per thread (read only; set from the width), tiles[2] x pass variant
 * of the rotating form with even padded Nx (0 six LDS lines, 1 split: three
 * lines, velocity held in registers, 2 six lines with direct-to-LDS loads;
 * other forms: 0), tiles[3] x wavefronts per block (1..the variant's LDS
 * limit), tiles[4] Chebyshev coefficients per lane of the wall-normal stage
 * and prologue kernels: the default (4 up to Ny = 256, 6 up to 384, 8
 * above) or an alternative where built: 8 on fewer lanes for
 * 64 < Ny <= 384, 6 on 32 lanes for 97 <= Ny <= 192 and on 64 lanes for
 * 193 <= Ny <= 256.
 * The x and z choices give identical results; the wall-normal layout changes
 * the summation order of its scans (results agree to roundoff).  An
 * unavailable choice returns CF_ERR_ARGUMENT and changes nothing. */
int cf_problem_get_tiles(const cf_problem* h, int tiles[5]);
int cf_problem_set_tiles(cf_problem* h, const int tiles[5]);

/* Uniform pressure gradient (dPdx, dPdz) returned by the last Stokes solve of
 * the last step (stokes.jl:101-107).  Synchronises the stream; HOST output. */
int cf_last_gradient(cf_problem* h, double out[2], void* stream);

/* ForwardFFT! (src/ffts.jl:65-82): padded physical u -> resolved spectrum U */
int cf_forward(cf_problem* h, const double* u, double* U, void* stream);
/* InverseFFT! (src/ffts.jl:151-166): resolved spectrum U -> padded physical u */
int cf_inverse(cf_problem* h, const double* U, double* u, void* stream);
/* the same for a vector field (FFT!/IFFT! of VectorField, ffts.jl:198-227):
 * u holds 3 contiguous padded physical fields; one launch per pass for all three */
int cf_forward3(cf_problem* h, const double* u, double* U0, double* U1, double* U2, void* stream);
int cf_inverse3(cf_problem* h, const double* U0, const double* U1, const double* U2, double* u, void* stream);

/* NonLinearTerm of the problem's form, V = U + Ub e_x: rotating N = FFT(V x curl V)
 * (src/nonlinear.jl:173-193), convective N = -FFT((V.grad)V) (113-140) or
 * divergence N = -div FFT(V V) (146-167); overwrites N0..N2 */
int cf_nonlinear(cf_problem* h, const double* U0, const double* U1, const double* U2, double* N0, double* N1,
                 double* N2, void* stream);

/* solve!(StokesSolver(grid, nu, lambda), U, P, R; pressuregradient|bulkvelocity,
 * baseflow=problem base) (src/solvers/stokes.jl:81-108).  constraint/values as
 * in cf_step (-1 = zero pressure gradient).  grads (HOST, may be NULL)
 * receives (dPdx, dPdz) and synchronises the stream. */
int cf_stokes_solve(cf_problem* h, double nu, double lambda, const double* R0, const double* R1, const double* R2,
                    double* U0, double* U1, double* U2, double* P, int constraint, const double* values,
                    double* grads, void* stream);

/* project!(U, problem) (src/pressure.jl:117-135), in place */
int cf_project(cf_problem* h, double* U0, double* U1, double* U2, void* stream);

/* pressure(U, problem) (src/pressure.jl:176-185): stage pressure of a
 * divergence-free no-slip velocity (modified pressure for the rotational form) */
int cf_pressure(cf_problem* h, const double* U0, const double* U1, const double* U2, double* P, void* stream);

/* dot(U::VectorField, V::VectorField) for spectral fields (postprocessing.jl:30-60):
 * the exact volume-averaged inner product <U . V> through the Chebyshev mass
 * matrix, Nyquist planes excluded.  HOST output; synchronises the stream. */
int cf_dot(cf_problem* h, const double* U0, const double* U1, const double* U2, const double* V0, const double* V1,
           const double* V2, double* out, void* stream);

/* dissipation_rate(U, nu) (postprocessing.jl:87-91): nu <|curl U|^2>.  HOST
 * output; synchronises the stream. */
int cf_dissipation(cf_problem* h, const double* U0, const double* U1, const double* U2, double nu, double* out,
                   void* stream);

/* power_input(U, nu; pressuregradient) (postprocessing.jl:113-128), from the
 * zero Fourier mode of U0 and U2.  grad may be NULL (zero).  HOST output;
 * synchronises the stream. */
int cf_power_input(cf_problem* h, const double* U0, const double* U2, double nu, const double grad[2], double* out,
                   void* stream);

/* Plane statistics of the TOTAL velocity U + Ub e_x at the Lobatto points y_j
 * (the x-z averages of examples/mkm590/statistics.jl:26-43 sample!, computed
 * exactly from the spectral coefficients by Parseval on the padded grid).
 * HOST out[Ny][9]: plane means U, V, W, then plane covariances
 * uu, vv, ww, uv, uw, vw.  Synchronises the stream; time averaging (the
 * parallel-variance merge of samples) is left to the caller. */
int cf_plane_statistics(cf_problem* h, const double* U0, const double* U1, const double* U2, double* out, void* stream);

/* Gaussian noise epsilon * N(0,1) for u, v, w on the padded physical grid
 * (initialization.jl:37-42): out (DEVICE) holds 3 fields of Nxp*Nzp*Ny
 * doubles in (x, z, y) order.  Stream: Philox4x32-10 keyed by the 64-bit seed,
 * counter = (element pair, component); Box-Muller.  Julia's Xoshiro stream is
 * not reproduced. */
int cf_gaussian_noise(cf_problem* h, double epsilon, unsigned long long seed, double* out, void* stream);

/* random_state(problem, epsilon; seed) (initialization.jl:32-51) on the
 * device: noise (cf_gaussian_noise) -> FFT -> project! -> pressure, into
 * U0..U2 and P. */
int cf_random_state(cf_problem* h, double epsilon, unsigned long long seed, double* U0, double* U1, double* U2, double* P,
                    void* stream);

/* Ensembles.  The reference advances one problem per step! call and runs
 * concurrent trajectories as independent problems (cnrk2.jl:147,
 * docs/src/gpu.md:42); an ensemble advances n such problems together, each
 * exactly as cf_advance would (bit-identical results), with every kernel
 * class of the step issued as one launch per group of instances that share
 * its tile parameters, and the whole ensemble step captured once into a HIP
 * graph.  problems: n distinct handles (FP64 precision, rotating or
 * convective form, even padded Nx) on one device; fields: 4n DEVICE
 * pointers, (U0, U1, U2, P) of instance i at 4i..4i+3, fixed for the life of
 * the ensemble.  tune != 0 picks each grid's z tile width and x wavefronts per
 * block by timing the candidates on the device (workspaces only; results do
 * not depend on the tiling).  Each instance uses its own problem's
 * constraint.  Synchronises at creation. */
typedef struct cf_ensemble cf_ensemble;
int cf_ensemble_create(cf_problem* const* problems, double* const* fields, int n, int tune, cf_ensemble** out);
/* nsteps ensemble steps (graph launches) on `stream` */
int cf_ensemble_advance(cf_ensemble* e, int nsteps, void* stream);
/* launches per ensemble step; tiles (may be NULL): 2n ints, z tile width and
 * x wavefronts per block of each instance */
int cf_ensemble_info(const cf_ensemble* e, int* launches, int* tiles);
int cf_ensemble_destroy(cf_ensemble* e);

#ifdef __cplusplus
}
#endif
#endif /* CHANNELFLOW_HIP_H */
