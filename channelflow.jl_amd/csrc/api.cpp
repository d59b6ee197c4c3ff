// C ABI of the MI355X CNRK2 library (include/channelflow_hip.h).
//
// A cf_problem owns the factored solver banks of the three CNRK2 stages
// (src/timesteppers/cnrk2.jl:58-83), the transform plans and the device
// workspaces.  One step (cnrk2.jl:149-207) is 13 kernel launches:
//   prologue (y):            U -> (V, Omega) at the Lobatto points
//   3 x [ z inverse -> x pass (u x w) -> z forward -> y stage ]
// where the y stage does the forward Chebyshev transform, Q update, RHS
// assembly, Stokes solve and the inverse transforms for the next stage.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <algorithm>
#include <array>
#include <atomic>
#include <list>
#include <vector>

#include "../../include/channelflow_hip.h"
#include "cf_host.h"
#include "cf_kernels.h"

using namespace cf;

static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPTRY(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) return fail(CF_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static const double CNRK2_A[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
static const double CNRK2_B[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};
static const double CNRK2_C[3] = {1.0 / 6.0, 5.0 / 24.0, 1.0 / 8.0};

struct DevBank {
    double* facV = nullptr;
    double* mconst = nullptr;
    double* resp = nullptr;
    double resp_mean = 0, lambda = 0, nu = 0;
    Bank view() const {
        Bank b;
        b.facV = facV;
        b.mconst = mconst;
        b.resp = resp;
        b.resp_mean = resp_mean;
        b.lambda = lambda;
        b.nu = nu;
        return b;
    }
};

struct Constraint {
    int bulk;
    double c0, c1;
};

struct cf_problem {
    int device = 0;
    GridDims gd{};
    double Lx = 0, Lz = 0, nu = 0, dt = 0;
    int form = 0, constraint = 0;
    double cvals[2] = {0, 0};
    std::vector<double> base, curv;
    double base_mean = 0, curv_mean = 0;
    // device data
    std::vector<void*> allocs;
    double* d_base = nullptr;
    double* d_curv = nullptr;
    double* d_facP = nullptr;
    double* d_tab = nullptr;  // per-degree tau / bulk-mean constants
    DevBank stage[3];
    std::list<std::pair<std::pair<double, double>, DevBank>> extra;  // stable references
    FFTPlan pz{}, py{}, px{};
    double2* d_twx = nullptr;
    double2* six = nullptr;   // 6 S
    double2* zbuf = nullptr;  // 6 Sz (9 Sz for the convective form)
    double2* Q = nullptr;     // 3 S
    double2* R = nullptr;     // 3 S
    double* d_grad = nullptr;
    double* d_scalar = nullptr;  // reduction result of the diagnostics
    double* d_noise = nullptr;   // 3 padded physical fields, allocated by the first cf_random_state
    int KT = 1, zld = 0, LPB = 1, xld = 0, yld = 0;
    int xwaves = 4;  // even Nxp: wavefronts per block of the nonlinear x pass
    int zppt = 0;    // z pass in-place points per thread (0: two buffers)
    int xppt = 0;    // even-Nxp x pass in-place points per lane (0: two buffers)
    // forcing(tstage, U, F) hook: host callback and the caller's F fields
    cf_forcing_fn force_fn = nullptr;
    void* force_user = nullptr;
    double2* F[3] = {nullptr, nullptr, nullptr};
    // nonlinear(tstage, U, N) hook replacing the built-in form: host callback
    // and the caller's N fields (coefficients)
    cf_forcing_fn nl_fn = nullptr;
    void* nl_user = nullptr;
    double2* NL[3] = {nullptr, nullptr, nullptr};
    // mixed-precision step: FP32 y-physical fields and periodic passes
    int mixed = 0;
    int KT32 = 1, zppt32 = 0, xppt32 = 0, xwaves32 = 0;
    // velocity-gradient x pass (12 FFT lines): in-place width, wavefronts per block
    int xpptg = 0, xwavesg = 0;
    // rotating form, FP64: x pass variant in use (0 six-line X_ROTATIONAL,
    // 1 split X_ROT_SPLIT with three LDS lines, 2 X_ROT_ASYNC: six lines with
    // direct-to-LDS loads), the split pass's in-place width and wavefronts
    // per block, and the async pass's wavefronts per block (its width is xppt)
    int xsplit = 0, xppts = 0, xwavess = 0, xwavesa = 0;
    // largest wavefronts per block that fit the LDS, per variant (0: unavailable)
    int xmaxw[3] = {0, 0, 0};
    // the same for the mixed step's FP32 x pass (largest wavefronts per
    // block of the six-line and the split variant in xmaxw32)
    int xsplit32 = 0, xppts32 = 0, xwavess32 = 0;
    int xmaxw32[2] = {0, 0};
    // group layout (lanes, coefficients per lane) of the step's wall-normal
    // stage and prologue kernels
    int yl = 0, yk = 0;
    unsigned long long serial = 0;  // creation number, unique in the process
    // learned closure registered as a device-side forcing (cf_set_closure):
    // the helper problem whose workspaces compute it, the network and the
    // caller's buffers; all null when none is registered
    struct ClosureCfg {
        cf_problem* helper = nullptr;
        unsigned long long helper_serial = 0;  // tells a helper from a later one at the same address
        const float* theta = nullptr;
        const unsigned short* wbf = nullptr;
        const float* norm = nullptr;
        double* grad9 = nullptr;
        double* force3 = nullptr;
        double2* F[3] = {nullptr, nullptr, nullptr};
        int mean_only = 0;
        bool active() const { return theta != nullptr; }
        bool operator==(const ClosureCfg& o) const {
            return helper == o.helper && helper_serial == o.helper_serial && theta == o.theta && wbf == o.wbf &&
                   norm == o.norm && grad9 == o.grad9 && force3 == o.force3 && F[0] == o.F[0] && F[1] == o.F[1] &&
                   F[2] == o.F[2] && mean_only == o.mean_only;
        }
    } closure;
    // one-step graphs of the plain step (no host callbacks), one per field
    // pointer set and constraint, least recently used dropped past GRAPH_CACHE;
    // cleared when the tiling changes
    struct StepGraph {
        const void* f[4];
        Constraint c;
        ClosureCfg cl;  // the closure the graph computes (none: all null)
        hipGraphExec_t exec;
        unsigned long long used;
    };
    std::vector<StepGraph> graphs;
    unsigned long long graph_clock = 0;
    hipStream_t cap = nullptr;  // capture stream (created with the first graph)
    // the registered closure's kernels run on `side`, forked from the step's
    // stream after each stage's inputs are final and joined back before the
    // stage kernel (created with the first closure registration)
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev) (void)hipSetDevice(dev);
        else prev = -1;
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

template <typename T>
static int dalloc(cf_problem* h, T** p, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T) + 16);
    if (e != hipSuccess) return fail(CF_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
    h->allocs.push_back(q);
    *p = (T*)q;
    return 0;
}

template <typename T>
static int upload(cf_problem* h, T** p, const std::vector<T>& v) {
    if (int rc = dalloc(h, p, v.size())) return rc;
    HIPTRY(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

static int make_plan(cf_problem* h, int n, FFTPlan& plan) {
    std::vector<int> radix;
    plan_factors(n, radix);
    if ((int)radix.size() > MAX_FACTORS) return fail(CF_ERR_UNSUPPORTED, "FFT length has too many factors");
    plan.n = n;
    plan.nf = (int)radix.size();
    int Ns = 1;
    for (int i = 0; i < plan.nf; ++i) {
        plan.radix[i] = radix[i];
        plan.mul_m[i] = recip_mul((unsigned)(n / radix[i]));
        plan.mul_ns[i] = recip_mul((unsigned)Ns);
        Ns *= radix[i];
    }
    if (n == 1) plan.nf = 0;
    std::vector<double> tw;
    twiddles(n, tw);
    double* d = nullptr;
    if (int rc = upload(h, &d, tw)) return rc;
    plan.tw = (const double2*)d;
    return 0;
}

static int make_bank(cf_problem* h, double nu, double lambda, DevBank& db) {
    HostBank hb;
    std::string err;
    if (!build_bank(h->gd, nu, lambda, hb, err)) return fail(CF_ERR_ARGUMENT, err);
    if (int rc = upload(h, &db.facV, hb.facV)) return rc;
    if (int rc = upload(h, &db.mconst, hb.mconst)) return rc;
    if (int rc = upload(h, &db.resp, hb.resp)) return rc;
    db.resp_mean = hb.resp_mean;
    db.lambda = lambda;
    db.nu = nu;
    return 0;
}

// Banks beyond the three stage banks (shortened final steps of cf_advance_to,
// cf_stokes_solve at other (nu, lambda)) are cached most-recent first, at
// most EXTRA_BANKS of them: a caller that varies the span continuously
// (Newton or periodic-orbit searches) keeps a bounded footprint.
static const size_t EXTRA_BANKS = 6;

static void free_bank(cf_problem* h, DevBank& db) {
    for (void* p : {(void*)db.facV, (void*)db.mconst, (void*)db.resp}) {
        auto it = std::find(h->allocs.begin(), h->allocs.end(), p);
        if (it == h->allocs.end()) continue;
        (void)hipFree(p);
        h->allocs.erase(it);
    }
    db = DevBank{};
}

static int get_bank(cf_problem* h, double nu, double lambda, DevBank** out) {
    for (int j = 0; j < 3; ++j)
        if (h->stage[j].nu == nu && h->stage[j].lambda == lambda) {
            *out = &h->stage[j];
            return 0;
        }
    for (auto it = h->extra.begin(); it != h->extra.end(); ++it)
        if (it->first.first == nu && it->first.second == lambda) {
            h->extra.splice(h->extra.begin(), h->extra, it);  // list nodes keep their addresses
            *out = &h->extra.front().second;
            return 0;
        }
    if (h->extra.size() >= EXTRA_BANKS) {
        // work enqueued earlier, on any stream, may still read the banks
        // about to be freed
        HIPTRY(hipDeviceSynchronize());
        while (h->extra.size() >= EXTRA_BANKS) {
            free_bank(h, h->extra.back().second);
            h->extra.pop_back();
        }
    }
    DevBank db;
    if (int rc = make_bank(h, nu, lambda, db)) return rc;
    h->extra.push_front({{nu, lambda}, db});
    *out = &h->extra.front().second;
    return 0;
}

static void drop_graphs(cf_problem* h) {
    for (auto& g : h->graphs) (void)hipGraphExecDestroy(g.exec);
    h->graphs.clear();
}

// the tiling changed: cached step graphs hold the old launch shapes (they may
// still run, so the device is drained first)
static int invalidate_graphs(cf_problem* h) {
    if (h->graphs.empty()) return 0;
    if (hipDeviceSynchronize() != hipSuccess) return fail(CF_ERR_DEVICE, "hipDeviceSynchronize");
    drop_graphs(h);
    return 0;
}

static void free_all(cf_problem* h) {
    drop_graphs(h);
    if (h->cap) (void)hipStreamDestroy(h->cap);
    h->cap = nullptr;
    if (h->side) (void)hipStreamDestroy(h->side);
    if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
    if (h->ev_join) (void)hipEventDestroy(h->ev_join);
    h->side = nullptr;
    h->ev_fork = h->ev_join = nullptr;
    for (void* p : h->allocs) (void)hipFree(p);
    h->allocs.clear();
}

extern "C" const char* cf_last_error(void) { return g_err.c_str(); }

extern "C" int cf_problem_create(const cf_config* cfg, cf_problem** out) {
    if (!cfg || !out) return fail(CF_ERR_ARGUMENT, "null argument");
    *out = nullptr;
    if (cfg->Ny % 2 == 0 || cfg->Ny < 5) return fail(CF_ERR_ARGUMENT, "Gibson's tau correction requires odd Ny >= 5");
    if (!(cfg->Nx > 0 && cfg->Nz > 0)) return fail(CF_ERR_ARGUMENT, "Nx and Nz must be positive");
    if (!(std::isfinite(cfg->Lx) && std::isfinite(cfg->Lz) && cfg->Lx > 0 && cfg->Lz > 0))
        return fail(CF_ERR_ARGUMENT, "Lx and Lz must be finite and positive");
    if (!(std::isfinite(cfg->dt) && cfg->dt > 0)) return fail(CF_ERR_ARGUMENT, "dt must be finite and positive");
    if (!(std::isfinite(cfg->nu) && cfg->nu > 0)) return fail(CF_ERR_ARGUMENT, "nu must be finite and positive");
    if (!cfg->baseflow) return fail(CF_ERR_DIMENSION, "baseflow must have one value per wall-normal point");
    if (cfg->constraint != CF_CONSTRAINT_PRESSURE_GRADIENT && cfg->constraint != CF_CONSTRAINT_BULK_VELOCITY)
        return fail(CF_ERR_ARGUMENT, "unknown mean-flow constraint");
    if (!(std::isfinite(cfg->constraint_values[0]) && std::isfinite(cfg->constraint_values[1])))
        return fail(CF_ERR_ARGUMENT, "mean-flow constraint values must be finite");
    if (cfg->form != CF_FORM_ROTATING && cfg->form != CF_FORM_CONVECTIVE && cfg->form != CF_FORM_DIVERGENCE)
        return fail(CF_ERR_ARGUMENT, "unknown nonlinear form");
    if (cfg->precision != CF_PRECISION_FP64 && cfg->precision != CF_PRECISION_MIXED)
        return fail(CF_ERR_ARGUMENT, "unknown precision");
    if (cfg->precision == CF_PRECISION_MIXED && (cfg->form == CF_FORM_DIVERGENCE || ((3 * cfg->Nx + 1) / 2) % 2 != 0))
        return fail(CF_ERR_UNSUPPORTED, "mixed precision needs an even padded Nx and the rotating or convective form");
    if (y_group_lanes(cfg->Ny) == 0) return fail(CF_ERR_UNSUPPORTED, "Ny above 511 is not supported");
    // device indexing: mode counts and padded line counts in 32 bits
    if ((long long)(cfg->Nx / 2 + 1) * cfg->Nz * cfg->Ny >= (1ll << 31) ||
        (long long)((3 * cfg->Nx + 1) / 2) * ((3 * cfg->Nz + 1) / 2) * cfg->Ny >= (1ll << 31))
        return fail(CF_ERR_UNSUPPORTED, "grid too large for one device");

    cf_problem* h = new cf_problem();
    static std::atomic<unsigned long long> created{0};
    h->serial = ++created;
    h->device = cfg->device;
    DeviceGuard dg(h->device);
    GridDims& gd = h->gd;
    gd.Nx = cfg->Nx;
    gd.Ny = cfg->Ny;
    gd.Nz = cfg->Nz;
    gd.Nxh = cfg->Nx / 2 + 1;
    gd.Nxp = (3 * cfg->Nx + 1) / 2;
    gd.Nzp = (3 * cfg->Nz + 1) / 2;
    gd.P = cfg->Ny - 1;
    gd.B = (long long)gd.Nxh * gd.Nz;
    gd.S = gd.B * gd.Ny;
    gd.Sz = (long long)gd.Nxh * gd.Nzp * gd.Ny;
    gd.alpha = 2.0 * M_PI / cfg->Lx;
    gd.beta = 2.0 * M_PI / cfg->Lz;
    h->Lx = cfg->Lx;
    h->Lz = cfg->Lz;
    h->nu = cfg->nu;
    h->dt = cfg->dt;
    h->form = cfg->form;
    h->constraint = cfg->constraint;
    h->cvals[0] = cfg->constraint_values[0];
    h->cvals[1] = cfg->constraint_values[1];
    int Ny = gd.Ny;
    for (int j = 0; j < Ny; ++j)
        if (!std::isfinite(cfg->baseflow[j])) {
            delete h;
            return fail(CF_ERR_ARGUMENT, "baseflow values must be finite");
        }
    h->base.resize(Ny);
    h->curv.resize(Ny);
    host_chebcoeffs(Ny, cfg->baseflow, h->base.data());  // cnrk2.jl:76
    std::vector<double> d1(Ny);
    host_diff(d1.data(), h->base.data(), Ny);  // diff2!, cnrk2.jl:77
    host_diff(h->curv.data(), d1.data(), Ny);
    h->base_mean = host_bulkmean(h->base.data(), Ny);
    h->curv_mean = host_bulkmean(h->curv.data(), Ny);

    int rc = 0;
    auto bail = [&](int code) {
        free_all(h);
        delete h;
        return code;
    };
    if ((rc = upload(h, &h->d_base, h->base))) return bail(rc);
    if ((rc = upload(h, &h->d_curv, h->curv))) return bail(rc);
    std::vector<double> facP;
    build_pressure_factors(gd, facP);
    if ((rc = upload(h, &h->d_facP, facP))) return bail(rc);
    {
        std::vector<double> tab;
        build_cheb_table(tab);
        if ((rc = upload(h, &h->d_tab, tab))) return bail(rc);
    }
    for (int j = 0; j < 3; ++j)
        if ((rc = make_bank(h, h->nu, 1.0 / (CNRK2_C[j] * h->dt), h->stage[j]))) return bail(rc);
    // plans
    if ((rc = make_plan(h, gd.Nzp, h->pz))) return bail(rc);
    if ((rc = make_plan(h, 2 * gd.P, h->py))) return bail(rc);
    int M = (gd.Nxp % 2 == 0) ? gd.Nxp / 2 : gd.Nxp;
    if ((rc = make_plan(h, M, h->px))) return bail(rc);
    {
        std::vector<double> tw;
        twiddles(gd.Nxp, tw);
        double* d = nullptr;
        if ((rc = upload(h, &d, tw))) return bail(rc);
        h->d_twx = (double2*)d;
    }
    // launch geometry (LDS <= 64 KB per block)
    const size_t LDS_MAX = 64 * 1024;
    h->yld = 2 * gd.P + 1;  // odd stride: the two lines of a wave fall in different LDS banks
    h->zld = gd.Nzp + 1;
    // z pass: one LDS buffer of ~32 KB (several blocks per CU) when the
    // in-place FFT applies, else two buffers within LDS_MAX
    h->KT = (int)std::max<size_t>(1, std::min<size_t>(16, (32 * 1024) / ((size_t)h->zld * sizeof(double2))));
    h->KT = std::min(h->KT, gd.Nxh);
    h->zppt = z_inplace_ppt(h->pz, h->KT);
    if (!h->zppt) {
        h->KT = (int)std::max<size_t>(1, std::min<size_t>(16, LDS_MAX / (2 * (size_t)h->zld * sizeof(double2))));
        h->KT = std::min(h->KT, gd.Nxh);
    }
    h->xld = M + 1;
    {
        // odd Nxp: two FFT lines plus the form's physical lines per (zp, y) line
        const int xmode = h->form == CF_FORM_CONVECTIVE ? X_CONVECTIVE : h->form == CF_FORM_DIVERGENCE ? X_DIVERGENCE : X_ROTATIONAL;
        size_t per_line = 2 * (size_t)h->xld * sizeof(double2) + (size_t)x_phys_lines(xmode) * gd.Nxp * sizeof(double);
        h->LPB = (int)std::max<size_t>(1, std::min<size_t>(8, LDS_MAX / per_line));
    }
    if (y_lds_bytes(gd.Ny, h->yld) > 160 * 1024) return bail(fail(CF_ERR_UNSUPPORTED, "wall-normal tile exceeds LDS"));
    h->yl = y_group_lanes(gd.Ny);
    h->yk = y_group_k(gd.Ny);
    if (gd.Nxp % 2 == 0) {
        XArgs xa;
        std::memset(&xa, 0, sizeof(xa));
        xa.gd = gd;
        xa.ld = h->xld;
        xa.mode = h->form == CF_FORM_CONVECTIVE ? X_CONVECTIVE : h->form == CF_FORM_DIVERGENCE ? X_DIVERGENCE : X_ROTATIONAL;
        xa.plan = h->px;
        h->xppt = x_inplace_ppt(xa);
        xa.ppt = h->xppt;
        h->xwaves = x_max_waves(xa);
        if (h->xwaves < 1) return bail(fail(CF_ERR_UNSUPPORTED, "streamwise line exceeds LDS"));
        if (xa.mode == X_ROTATIONAL) {
            XArgs xs = xa;
            xs.mode = X_ROT_SPLIT;
            xs.ppt = 0;
            h->xppts = x_inplace_ppt(xs);
            xs.ppt = h->xppts;
            h->xwavess = h->xppts ? x_max_waves(xs) : 0;
            h->xsplit = h->xwavess > 0 ? 1 : 0;
            h->xwavesa = h->xppt ? h->xwaves : 0;  // same LDS lines as the six-line in-place pass
        }
        h->xmaxw[0] = h->xwaves;
        h->xmaxw[1] = h->xwavess;
        h->xmaxw[2] = h->xwavesa;
        {
            XArgs xg = xa;
            xg.mode = X_GRADIENT;
            xg.ppt = 0;
            h->xpptg = x_inplace_ppt(xg);
            xg.ppt = h->xpptg;
            h->xwavesg = x_max_waves(xg);
        }
        if (cfg->precision == CF_PRECISION_MIXED) {
            xa.mixed = 1;
            xa.ppt = 0;
            h->xppt32 = x_inplace_ppt(xa);
            xa.ppt = h->xppt32;
            h->xwaves32 = x_max_waves(xa);
            if (xa.mode == X_ROTATIONAL) {
                XArgs xs = xa;
                xs.mode = X_ROT_SPLIT;
                xs.ppt = 0;
                h->xppts32 = x_inplace_ppt(xs);
                xs.ppt = h->xppts32;
                h->xwavess32 = h->xppts32 ? x_max_waves(xs) : 0;
                h->xsplit32 = h->xwavess32 > 0 ? 1 : 0;
            }
            h->xmaxw32[0] = h->xwaves32;
            h->xmaxw32[1] = h->xwavess32;
        }
    }
    if (cfg->precision == CF_PRECISION_MIXED) {
        // FP32 z tiles: twice the k per ~32 KB buffer
        h->KT32 = (int)std::max<size_t>(1, std::min<size_t>(32, (32 * 1024) / ((size_t)h->zld * sizeof(float2))));
        h->KT32 = std::min(h->KT32, gd.Nxh);
        while (h->KT32 > 1 && !z_inplace_ppt(h->pz, h->KT32)) --h->KT32;
        h->zppt32 = z_inplace_ppt(h->pz, h->KT32);
        if (!h->zppt32 || !h->xppt32 || h->xwaves32 < 1)
            return bail(fail(CF_ERR_UNSUPPORTED, "mixed precision needs FFT lengths with radices 2, 3, 4, 5 and 8"));
        h->mixed = 1;
    }
    // workspaces
    if ((rc = dalloc(h, &h->six, 6 * (size_t)gd.S))) return bail(rc);
    // 9 fields: the convective form and the velocity gradient need V, dV/dy, dV/dz
    if ((rc = dalloc(h, &h->zbuf, 9 * (size_t)gd.Sz))) return bail(rc);
    if ((rc = dalloc(h, &h->Q, 3 * (size_t)gd.S))) return bail(rc);
    if ((rc = dalloc(h, &h->R, 3 * (size_t)gd.S))) return bail(rc);
    if ((rc = dalloc(h, &h->d_grad, 2))) return bail(rc);
    if ((rc = dalloc(h, &h->d_scalar, 2))) return bail(rc);
    if (hipMemset(h->d_grad, 0, 2 * sizeof(double)) != hipSuccess) return bail(fail(CF_ERR_DEVICE, "hipMemset"));
    *out = h;
    return CF_OK;
}

extern "C" int cf_problem_destroy(cf_problem* h) {
    if (!h) return CF_OK;
    DeviceGuard dg(h->device);
    (void)hipDeviceSynchronize();
    free_all(h);
    delete h;
    return CF_OK;
}

extern "C" int cf_problem_sizes(const cf_problem* h, long long out[8]) {
    if (!h || !out) return fail(CF_ERR_ARGUMENT, "null argument");
    out[0] = h->gd.Nxh;
    out[1] = h->gd.Nz;
    out[2] = h->gd.Ny;
    out[3] = h->gd.Nxp;
    out[4] = h->gd.Nzp;
    out[5] = h->gd.Ny;
    out[6] = h->gd.S;
    out[7] = (long long)h->gd.Nxp * h->gd.Nzp * h->gd.Ny;
    return CF_OK;
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
static YArgs base_yargs(cf_problem* h) {
    YArgs a;
    std::memset(&a, 0, sizeof(a));
    a.gd = h->gd;
    a.dct = h->py;
    a.ld = h->yld;
    a.norm = 1.0 / ((double)h->gd.Nxp * h->gd.Nzp * h->gd.P);
    a.nu = h->nu;
    a.base = h->d_base;
    a.curv = h->d_curv;
    a.curv_mean = h->curv_mean;
    a.facP = h->d_facP;
    a.facPN = h->d_facP;
    a.form = h->form;
    a.nl_sign = 1.0;
    a.cheb_tab = h->d_tab;
    a.ntab = CHEB_TAB_MAX;
    return a;
}

static double2* fld(double2* base, long long n, int i) { return base + (size_t)n * i; }

static ZArgs make_zargs(cf_problem* h, int nf, double2* const* in, double2* const* out, unsigned dz_mask, bool mixed) {
    ZArgs z;
    std::memset(&z, 0, sizeof(z));
    z.gd = h->gd;
    z.plan = h->pz;
    z.KT = mixed ? h->KT32 : h->KT;
    z.ld = h->zld;
    z.nfields = nf;
    z.ppt = mixed ? h->zppt32 : h->zppt;
    z.mixed = mixed ? 1 : 0;
    z.dz_mask = dz_mask;
    z.mul_kt = recip_mul((unsigned)z.KT);
    for (int i = 0; i < nf; ++i) {
        z.in[i] = in[i];
        z.out[i] = out[i];
    }
    return z;
}

static int zpass(cf_problem* h, int dir, int nf, double2* const* in, double2* const* out, hipStream_t st,
                 unsigned dz_mask = 0, bool mixed = false) {
    HIPTRY(launch_zpass(dir, make_zargs(h, nf, in, out, dz_mask, mixed), st));
    return 0;
}

static XArgs base_xargs(cf_problem* h, bool mixed = false) {
    XArgs x;
    std::memset(&x, 0, sizeof(x));
    x.gd = h->gd;
    x.plan = h->px;
    x.twx = h->d_twx;
    x.LPB = h->LPB;
    x.ld = h->xld;
    x.waves = mixed ? h->xwaves32 : h->xwaves;
    x.ppt = mixed ? h->xppt32 : h->xppt;
    x.mixed = mixed ? 1 : 0;
    x.mul_m = recip_mul((unsigned)std::max(1, h->gd.Nxp / 2));
    x.mul_nxh = recip_mul((unsigned)h->gd.Nxh);
    x.mul_nx = recip_mul((unsigned)(std::max(1, h->gd.Nxp / 2) + 1));
    return x;
}

// y-physical fields of the form in six (V, Omega | V, dV/dy | V) -> y-physical
// nonlinear term in six[0..2] (rotating, convective: V x Omega, (V.grad)V) or
// the six flux products V_i V_j in six[0..5] (divergence)
struct Timer;
static int tmark(Timer* tm, int c, hipStream_t st);
// y-physical field i of the six buffer in the step's storage (mixed: FP32
// complex fields packed contiguously, S elements apart)
static double2* six_field(cf_problem* h, int i, bool mixed) {
    return mixed ? (double2*)((float2*)h->six + (size_t)h->gd.S * i) : fld(h->six, h->gd.S, i);
}

// The three launches of the nonlinear term's periodic work: z inverse of the
// form's y-physical fields, x pass (products), z forward of the products.
// mixed: FP32 complex fields (z buffer fields Sz elements apart).
struct NonlinearArgs {
    ZArgs zin, zout;
    XArgs x;
};

static NonlinearArgs nonlinear_args(cf_problem* h, bool mixed) {
    const GridDims& gd = h->gd;
    double2* six[6];
    double2* zb[9];
    for (int i = 0; i < 6; ++i) six[i] = six_field(h, i, mixed);
    const int nzb = h->form == CF_FORM_CONVECTIVE ? 9 : 6;
    for (int i = 0; i < nzb; ++i) zb[i] = mixed ? (double2*)((float2*)h->zbuf + (size_t)gd.Sz * i) : fld(h->zbuf, gd.Sz, i);
    NonlinearArgs na;
    na.x = base_xargs(h, mixed);
    XArgs& x = na.x;
    if (h->form == CF_FORM_ROTATING) {
        // z field 4 is i beta l V_x (dV_x/dz); the x pass forms Omega_y from
        // it and V_z, so the y kernels emit five y-physical fields
        double2* zin[6] = {six[0], six[1], six[2], six[3], six[0], six[5]};
        na.zin = make_zargs(h, 6, zin, zb, 0x10u, mixed);
        x.mode = X_ROTATIONAL;
        if (!mixed && h->xsplit == 2) {
            x.mode = X_ROT_ASYNC;
            x.waves = h->xwavesa;
        } else if (mixed ? h->xsplit32 : h->xsplit) {
            x.mode = X_ROT_SPLIT;
            x.ppt = mixed ? h->xppts32 : h->xppts;
            x.waves = mixed ? h->xwavess32 : h->xwavess;
        }
        for (int i = 0; i < 6; ++i) x.in[i] = zb[i];
        for (int i = 0; i < 3; ++i) x.out[i] = zb[i];  // in place: a wavefront reads its line first
    } else if (h->form == CF_FORM_CONVECTIVE) {
        // V, dV/dy and (from V again) dV/dz
        double2* zin[9] = {six[0], six[1], six[2], six[3], six[4], six[5], six[0], six[1], six[2]};
        na.zin = make_zargs(h, 9, zin, zb, 0x1C0u, mixed);
        x.mode = X_CONVECTIVE;
        for (int i = 0; i < 9; ++i) x.in[i] = zb[i];
        for (int i = 0; i < 3; ++i) x.out[i] = zb[i];
    } else {
        na.zin = make_zargs(h, 3, six, zb, 0, mixed);
        x.mode = X_DIVERGENCE;
        for (int i = 0; i < 3; ++i) x.in[i] = zb[i];
        for (int i = 0; i < 6; ++i) x.out[i] = zb[i];
    }
    const int nout = h->form == CF_FORM_DIVERGENCE ? 6 : 3;
    na.zout = make_zargs(h, nout, zb, six, 0, mixed);
    return na;
}

static int enqueue_nonlinear_phys(cf_problem* h, hipStream_t st, Timer* tm = nullptr, bool mixed = false) {
    const NonlinearArgs na = nonlinear_args(h, mixed);
    HIPTRY(launch_zpass(+1, na.zin, st));
    if (int rc = tmark(tm, 1, st)) return rc;
    HIPTRY(launch_xpass(na.x, st));
    if (int rc = tmark(tm, 2, st)) return rc;
    HIPTRY(launch_zpass(-1, na.zout, st));
    if (int rc = tmark(tm, 3, st)) return rc;
    return 0;
}

// the transformed nonlinear term's sign (nonlinear.jl:139, 166, 192) and, for
// the divergence form, its assembly in coefficient space into Ncoef[0..2]
static int enqueue_nonlinear_coef(cf_problem* h, double2* const Ncoef[3], hipStream_t st) {
    YArgs a = base_yargs(h);
    for (int i = 0; i < 6; ++i) a.Nphys[i] = fld(h->six, h->gd.S, i);
    for (int i = 0; i < 3; ++i) a.Rout[i] = Ncoef[i];
    HIPTRY(launch_ykernel(Y_DIVERGENCE, a, st));
    return 0;
}

static YArgs prologue_args(cf_problem* h, double2* const U[3], bool mixed) {
    YArgs a = base_yargs(h);
    a.yl = h->yl;
    a.yk = h->yk;
    a.mixed = mixed ? 1 : 0;
    for (int i = 0; i < 3; ++i) a.U[i] = U[i];
    for (int i = 0; i < 6; ++i) a.out6[i] = six_field(h, i, mixed);
    return a;
}

static int enqueue_prologue(cf_problem* h, double2* const U[3], hipStream_t st, bool mixed = false) {
    HIPTRY(launch_ykernel(Y_PROLOGUE, prologue_args(h, U, mixed), st));
    return 0;
}

static int resolve_constraint(cf_problem* h, int kind, const double* values, Constraint& c) {
    if (kind == -1) {
        kind = h->constraint;
        values = h->cvals;
    }
    if (kind != CF_CONSTRAINT_PRESSURE_GRADIENT && kind != CF_CONSTRAINT_BULK_VELOCITY)
        return fail(CF_ERR_ARGUMENT, "specify either pressuregradient or bulkvelocity");
    double v0 = values ? values[0] : 0.0, v1 = values ? values[1] : 0.0;
    if (!(std::isfinite(v0) && std::isfinite(v1))) return fail(CF_ERR_ARGUMENT, "mean-flow constraint values must be finite");
    c.bulk = (kind == CF_CONSTRAINT_BULK_VELOCITY);
    // StokesSolver: bulk target is the perturbation mean (stokes.jl:96-99)
    c.c0 = c.bulk ? v0 - h->base_mean : v0;
    c.c1 = v1;
    return 0;
}

static int check_fields(const void* const* f, int n) {
    for (int i = 0; i < n; ++i) {
        if (!f[i]) return fail(CF_ERR_ARGUMENT, "null field pointer");
        for (int j = 0; j < i; ++j)
            if (f[i] == f[j]) return fail(CF_ERR_ARGUMENT, "fields must have distinct storage");
    }
    return 0;
}

// Optional per-kernel-class timing with HIP events recorded on the launch
// stream around every launch (measurement only).
struct Timer {
    static constexpr int NCLASS = 5;  // prologue, z inverse, x pass, z forward, y stage
    std::vector<hipEvent_t> ev;
    std::vector<int> cls;
    int mark(int c, hipStream_t st) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return fail(CF_ERR_DEVICE, "hipEventCreate");
        if (hipEventRecord(e, st) != hipSuccess) return fail(CF_ERR_DEVICE, "hipEventRecord");
        ev.push_back(e);
        cls.push_back(c);
        return 0;
    }
    ~Timer() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};
static int tmark(Timer* tm, int c, hipStream_t st) { return tm ? tm->mark(c, st) : 0; }

// Arguments of CNRK2 stage j (cnrk2.jl:168-204) for step length dt with
// solver bank `bank`; Ncoef: the divergence form's coefficient-space term.
static YArgs stage_args(cf_problem* h, double2* const U[3], double2* P, int j, const Constraint& c, double dt,
                        const DevBank* bank, bool mixed, double2* const Ncoef[3], bool forced = false) {
    const GridDims& gd = h->gd;
    YArgs a = base_yargs(h);
    a.yl = h->yl;
    a.yk = h->yk;
    a.mixed = mixed ? 1 : 0;
    a.nl_sign = (h->form == CF_FORM_CONVECTIVE) ? -1.0 : 1.0;
    a.stage = j;
    a.store_Q = (j < 2);
    a.lambda = 1.0 / (CNRK2_C[j] * dt);
    a.Aj = CNRK2_A[j];
    a.weight = CNRK2_B[j] / CNRK2_C[j];
    a.bulk = c.bulk;
    a.c0 = c.c0;
    a.c1 = c.c1;
    a.bank = bank->view();
    for (int i = 0; i < 3; ++i) {
        a.Nphys[i] = six_field(h, i, mixed);
        a.Ncoef[i] = (h->form == CF_FORM_DIVERGENCE) ? Ncoef[i] : nullptr;
        a.Fcoef[i] = forced ? h->F[i] : nullptr;
        a.U[i] = U[i];
        a.Q[i] = fld(h->Q, gd.S, i);
    }
    a.P = P;
    a.grad_out = h->d_grad;
    if (j < 2)
        for (int i = 0; i < 6; ++i) a.out6[i] = six_field(h, i, mixed);
    return a;
}

static const double STAGE_T[3] = {0.0, 1.0 / 3.0, 3.0 / 4.0};  // stage times, cnrk2.jl:170

static int enqueue_closure(cf_problem* h, double2* const U[3], hipStream_t st);
static int forward_fields(cf_problem* h, int nf, const double* u, double* const* U, hipStream_t st);

// nsteps CNRK2 steps of length dt with the stage solver banks `banks`
// (default: the problem's dt and its three stage banks)
static int enqueue_steps(cf_problem* h, double2* const U[3], double2* P, int nsteps, const Constraint& c, hipStream_t st,
                         Timer* tm = nullptr, double dt = 0.0, DevBank* const* banks = nullptr, double t0 = 0.0) {
    const GridDims& gd = h->gd;
    const bool mixed = h->mixed != 0;
    const bool forced = h->force_fn != nullptr || h->closure.active();
    const bool custom = h->nl_fn != nullptr;  // caller's nonlinear(tstage, U, N) in coefficients
    // a registered closure (exclusive with a host forcing) computes F on the
    // side stream while this stream computes the nonlinear term: in a captured
    // graph the two form parallel branches of each stage
    const bool closure = h->closure.active() && !h->force_fn;
    if (!(dt > 0.0)) dt = h->dt;
    for (int step = 0; step < nsteps; ++step) {
        if (int rc = tmark(tm, -1, st)) return rc;
        if (!custom)
            if (int rc = enqueue_prologue(h, U, st, mixed)) return rc;
        if (int rc = tmark(tm, 0, st)) return rc;
        for (int j = 0; j < 3; ++j) {
            const double ts = t0 + (double)step * dt + STAGE_T[j] * dt;
            double2* Ncoef[3] = {fld(h->R, gd.S, 0), fld(h->R, gd.S, 1), fld(h->R, gd.S, 2)};
            if (closure) {  // fork: U of this stage is final on `st`
                HIPTRY(hipEventRecord(h->ev_fork, st));
                HIPTRY(hipStreamWaitEvent(h->side, h->ev_fork, 0));
                if (int rc = enqueue_closure(h, U, h->side)) return rc;
                HIPTRY(hipEventRecord(h->ev_join, h->side));
            }
            if (custom) {  // nonlinear(tstage, U, N) overwrites N (cnrk2.jl:121-126,173)
                for (int i = 0; i < 3; ++i) Ncoef[i] = h->NL[i];
                if (h->nl_fn(h->nl_user, ts, (const double*)U[0], (const double*)U[1], (const double*)U[2],
                             (double*)h->NL[0], (double*)h->NL[1], (double*)h->NL[2], st) != 0)
                    return fail(CF_ERR_ARGUMENT, "nonlinear callback failed");
            } else {
                if (int rc = enqueue_nonlinear_phys(h, st, tm, mixed)) return rc;
                if (h->form == CF_FORM_DIVERGENCE)
                    if (int rc = enqueue_nonlinear_coef(h, Ncoef, st)) return rc;
            }
            if (h->force_fn) {
                if (h->force_fn(h->force_user, ts, (const double*)U[0], (const double*)U[1], (const double*)U[2],
                                (double*)h->F[0], (double*)h->F[1], (double*)h->F[2], st) != 0)
                    return fail(CF_ERR_ARGUMENT, "forcing callback failed");
            } else if (closure) {  // join: F is final before the stage kernel reads it
                HIPTRY(hipStreamWaitEvent(st, h->ev_join, 0));
            }
            YArgs a = stage_args(h, U, P, j, c, dt, banks ? banks[j] : &h->stage[j], mixed, Ncoef, forced);
            if (custom) {  // N arrives in coefficients; no y-physical fields are needed
                for (int i = 0; i < 3; ++i) a.Ncoef[i] = h->NL[i];
                for (int i = 0; i < 6; ++i) a.out6[i] = nullptr;
            }
            HIPTRY(launch_ykernel(Y_STAGE, a, st));
            if (int rc = tmark(tm, 4, st)) return rc;
        }
    }
    return 0;
}

// Plain steps (no host callbacks) replayed from a cached one-step graph: one
// hipGraphLaunch per step instead of the step's 13 kernel launches.  The graph
// holds kernel nodes only (no memset or copy nodes) with the arguments of
// these field pointers, this constraint and the registered closure (whose
// kernels it includes).
static constexpr size_t GRAPH_CACHE = 4;

static int graph_steps(cf_problem* h, double2* const U[3], double2* P, int nsteps, const Constraint& c, hipStream_t st) {
    const void* key[4] = {U[0], U[1], U[2], P};
    cf_problem::StepGraph* g = nullptr;
    for (auto& e : h->graphs)
        if (std::equal(key, key + 4, e.f) && e.c.bulk == c.bulk && e.c.c0 == c.c0 && e.c.c1 == c.c1 &&
            e.cl == h->closure)
            g = &e;
    if (!g) {
        if (!h->cap && hipStreamCreateWithFlags(&h->cap, hipStreamNonBlocking) != hipSuccess)
            return fail(CF_ERR_DEVICE, "hipStreamCreate");
        HIPTRY(hipStreamBeginCapture(h->cap, hipStreamCaptureModeRelaxed));
        const int rc = enqueue_steps(h, U, P, 1, c, h->cap);
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(h->cap, &graph);
        hipGraphExec_t exec = nullptr;
        hipError_t ei = hipSuccess;
        if (!rc && ec == hipSuccess) ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        if (rc) return rc;
        if (ec != hipSuccess) return fail(CF_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
        if (ei != hipSuccess) return fail(CF_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        if (h->graphs.size() >= GRAPH_CACHE) {
            // the evicted graph may still run on some stream: wait for the device
            auto lru = std::min_element(h->graphs.begin(), h->graphs.end(),
                                        [](const auto& a, const auto& b) { return a.used < b.used; });
            HIPTRY(hipDeviceSynchronize());
            (void)hipGraphExecDestroy(lru->exec);
            h->graphs.erase(lru);
        }
        cf_problem::StepGraph e;
        std::copy(key, key + 4, e.f);
        e.c = c;
        e.cl = h->closure;
        e.exec = exec;
        h->graphs.push_back(e);
        g = &h->graphs.back();
    }
    g->used = ++h->graph_clock;
    for (int s = 0; s < nsteps; ++s) HIPTRY(hipGraphLaunch(g->exec, st));
    return 0;
}

// nominal steps of the problem's dt: graph replay when no host callback runs
// inside the step, else launched one by one
static int run_steps(cf_problem* h, double2* const U[3], double2* P, int nsteps, const Constraint& c, hipStream_t st,
                     double t0, bool graph) {
    // a caller capturing its own graph on `st` records the kernels themselves
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (graph && st && hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) graph = false;
    if (graph && nsteps > 0 && !h->force_fn && !h->nl_fn) return graph_steps(h, U, P, nsteps, c, st);
    return enqueue_steps(h, U, P, nsteps, c, st, nullptr, 0.0, nullptr, t0);
}

extern "C" int cf_advance(cf_problem* h, double* U0, double* U1, double* U2, double* P, double t, int nsteps,
                          int constraint, const double* values, double* t_out, void* stream) {
    if (!h) return fail(CF_ERR_ARGUMENT, "null problem");
    const void* f[4] = {U0, U1, U2, P};
    if (int rc = check_fields(f, 4)) return rc;
    if (nsteps < 0) return fail(CF_ERR_ARGUMENT, "nsteps must be nonnegative");
    if (!std::isfinite(t)) return fail(CF_ERR_ARGUMENT, "t must be finite");
    Constraint c;
    if (int rc = resolve_constraint(h, constraint, values, c)) return rc;
    DeviceGuard dg(h->device);
    double2* U[3] = {(double2*)U0, (double2*)U1, (double2*)U2};
    if (int rc = run_steps(h, U, (double2*)P, nsteps, c, (hipStream_t)stream, t, true)) return rc;
    if (t_out) *t_out = t + nsteps * h->dt;
    return CF_OK;
}

// Flows.step! adapter semantics (ext/ChannelFlowFlowsExt.jl:36-52): nominal
// steps of dt from t while they fit in T, then one shortened step of the
// remaining duration with stage solvers built for it (cached per duration)
// doubles of one stage bank laid out in the fixed scratch: facV (B Ny),
// mconst (B MC_COUNT), resp (Ny), each segment 256-byte aligned
static long long scratch_bank_doubles(const GridDims& gd) {
    auto up = [](long long n) { return (n + 31) / 32 * 32; };
    return up(gd.B * (long long)gd.Ny) + up(gd.B * (long long)MC_COUNT) + up(gd.Ny);
}

// cf_advance_to; noalloc: the shortened step's three stage banks are built on
// the host and copied into the fixed device scratch instead of the handle's
// bank cache (no device allocation)
static int advance_to(cf_problem* h, double* U0, double* U1, double* U2, double* P, double t, double T, int constraint,
                      const double* values, double* t_out, void* stream, bool noalloc) {
    if (!h) return fail(CF_ERR_ARGUMENT, "null problem");
    const void* f[4] = {U0, U1, U2, P};
    if (int rc = check_fields(f, 4)) return rc;
    if (!(std::isfinite(t) && std::isfinite(T) && T >= t)) return fail(CF_ERR_ARGUMENT, "need finite t <= T");
    Constraint c;
    if (int rc = resolve_constraint(h, constraint, values, c)) return rc;
    DeviceGuard dg(h->device);
    hipStream_t st = (hipStream_t)stream;
    double2* U[3] = {(double2*)U0, (double2*)U1, (double2*)U2};
    const double span = T - t, tol = 1e-10 * h->dt;
    long long n = (long long)std::floor((span + tol) / h->dt);
    double rem = span - (double)n * h->dt;
    if (rem <= tol) rem = 0.0;
    if (n > (1ll << 30)) return fail(CF_ERR_ARGUMENT, "too many steps");
    // everything that can fail is checked before the first step is enqueued
    const long long per_bank = scratch_bank_doubles(h->gd);
    HostBank hb[3];
    double* scratch = nullptr;
    if (rem > 0.0 && noalloc) {
        if (3 * per_bank > SCRATCH_DOUBLES)
            return fail(CF_ERR_UNSUPPORTED, "the shortened step's solver banks (" + std::to_string(3 * per_bank * 8 >> 20) +
                                                " MiB) exceed the fixed device scratch of " +
                                                std::to_string(SCRATCH_BYTES >> 20) + " MiB; use cf_advance_to");
        for (int j = 0; j < 3; ++j) {
            std::string err;
            if (!build_bank(h->gd, h->nu, 1.0 / (CNRK2_C[j] * rem), hb[j], err)) return fail(CF_ERR_ARGUMENT, err);
        }
        HIPTRY(scratch_address(h->device, &scratch));
    }
    // instantiating a graph allocates: the allocation-free form launches one by one
    if (int rc = run_steps(h, U, (double2*)P, (int)n, c, st, t, !noalloc)) return rc;
    if (rem > 0.0) {
        DevBank* b[3];
        DevBank sb[3];
        if (noalloc) {
            // stream order: the copies follow every earlier use of the scratch on
            // this stream; the wait lets the host banks go out of scope
            for (int j = 0; j < 3; ++j) {
                double* base = scratch + j * per_bank;
                const long long o1 = (h->gd.B * (long long)h->gd.Ny + 31) / 32 * 32;
                const long long o2 = o1 + (h->gd.B * (long long)MC_COUNT + 31) / 32 * 32;
                HIPTRY(hipMemcpyAsync(base, hb[j].facV.data(), hb[j].facV.size() * sizeof(double),
                                      hipMemcpyHostToDevice, st));
                HIPTRY(hipMemcpyAsync(base + o1, hb[j].mconst.data(), hb[j].mconst.size() * sizeof(double),
                                      hipMemcpyHostToDevice, st));
                HIPTRY(hipMemcpyAsync(base + o2, hb[j].resp.data(), hb[j].resp.size() * sizeof(double),
                                      hipMemcpyHostToDevice, st));
                sb[j].facV = base;
                sb[j].mconst = base + o1;
                sb[j].resp = base + o2;
                sb[j].resp_mean = hb[j].resp_mean;
                sb[j].lambda = hb[j].lambda;
                sb[j].nu = hb[j].nu;
                b[j] = &sb[j];
            }
            HIPTRY(hipStreamSynchronize(st));
        } else {
            for (int j = 0; j < 3; ++j)
                if (int rc = get_bank(h, h->nu, 1.0 / (CNRK2_C[j] * rem), &b[j])) return rc;
        }
        if (int rc = enqueue_steps(h, U, (double2*)P, 1, c, st, nullptr, rem, b, t + (double)n * h->dt)) return rc;
    }
    if (t_out) *t_out = T;
    return CF_OK;
}

extern "C" int cf_advance_to(cf_problem* h, double* U0, double* U1, double* U2, double* P, double t, double T,
                             int constraint, const double* values, double* t_out, void* stream) {
    return advance_to(h, U0, U1, U2, P, t, T, constraint, values, t_out, stream, false);
}

extern "C" int cf_advance_to_noalloc(cf_problem* h, double* U0, double* U1, double* U2, double* P, double t, double T,
                                     int constraint, const double* values, double* t_out, void* stream) {
    return advance_to(h, U0, U1, U2, P, t, T, constraint, values, t_out, stream, true);
}

static int velocity_gradient_on(cf_problem* h, double2* const U[3], double* out, hipStream_t st);

// Velocity-gradient tensor of the total velocity V = U + Ub e_x on the padded
// grid: the convective prologue (V, dV/dy), the z pass with dV/dz, and the
// gradient x pass (dV/dx = i alpha k V) writing 9 real fields.
extern "C" int cf_velocity_gradient(cf_problem* h, const double* U0, const double* U1, const double* U2, double* out,
                                    void* stream) {
    if (!h) return fail(CF_ERR_ARGUMENT, "null problem");
    const void* f[4] = {U0, U1, U2, out};
    if (int rc = check_fields(f, 4)) return rc;
    const GridDims& gd = h->gd;
    if (gd.Nxp % 2 != 0 || h->xwavesg < 1) return fail(CF_ERR_UNSUPPORTED, "velocity gradient needs an even padded Nx");
    DeviceGuard dg(h->device);
    double2* U[3] = {(double2*)U0, (double2*)U1, (double2*)U2};
    return velocity_gradient_on(h, U, out, (hipStream_t)stream);
}

// the gradient's three launches on h's workspaces (checked by the caller)
static int velocity_gradient_on(cf_problem* h, double2* const U[3], double* out, hipStream_t st) {
    const GridDims& gd = h->gd;
    YArgs a = prologue_args(h, U, false);
    a.form = FORM_CONVECTIVE;
    HIPTRY(launch_ykernel(Y_PROLOGUE, a, st));
    double2* six[6];
    double2* zb[9];
    for (int i = 0; i < 6; ++i) six[i] = fld(h->six, gd.S, i);
    for (int i = 0; i < 9; ++i) zb[i] = fld(h->zbuf, gd.Sz, i);
    double2* zin[9] = {six[0], six[1], six[2], six[3], six[4], six[5], six[0], six[1], six[2]};
    if (int rc = zpass(h, +1, 9, zin, zb, st, 0x1C0u)) return rc;
    XArgs x = base_xargs(h);
    x.mode = X_GRADIENT;
    x.ppt = h->xpptg;
    x.waves = h->xwavesg;
    const size_t nphys = (size_t)gd.Nxp * gd.Nzp * gd.Ny;
    for (int i = 0; i < 9; ++i) {
        x.in[i] = zb[i];
        x.rout[i] = out + nphys * i;
    }
    HIPTRY(launch_xpass(x, st));
    return CF_OK;
}

// Box filter (closure data): fine problem's spectral U -> coarse problem's
// spectral Ubar, both problems on the same box.
extern "C" int cf_box_filter(cf_problem* fine, cf_problem* coarse, const double* U0, const double* U1, const double* U2,
                             double* V0, double* V1, double* V2, double dx, double dz, void* stream) {
    if (!fine || !coarse) return fail(CF_ERR_ARGUMENT, "null problem");
    const void* f[6] = {U0, U1, U2, V0, V1, V2};
    if (int rc = check_fields(f, 6)) return rc;
    if (fine->device != coarse->device) return fail(CF_ERR_ARGUMENT, "problems must share a device");
    if (std::fabs(fine->Lx - coarse->Lx) > 1e-12 * fine->Lx || std::fabs(fine->Lz - coarse->Lz) > 1e-12 * fine->Lz)
        return fail(CF_ERR_ARGUMENT, "fine and coarse grids must share Lx and Lz");
    const GridDims &F = fine->gd, &Cg = coarse->gd;
    if (!(Cg.Nx / 2 < (F.Nx + 1) / 2 && Cg.Nz / 2 < (F.Nz + 1) / 2 && Cg.Ny <= F.Ny))
        return fail(CF_ERR_DIMENSION, "the coarse grid must be strictly coarser in x and z and not finer in y");
    if (!(std::isfinite(dx) && std::isfinite(dz) && dx >= 0 && dz >= 0)) return fail(CF_ERR_ARGUMENT, "bad filter widths");
    DeviceGuard dg(fine->device);
    FilterArgs a;
    std::memset(&a, 0, sizeof(a));
    a.f = F;
    a.c = Cg;
    a.dx = dx;
    a.dz = dz;
    a.in[0] = (const double2*)U0;
    a.in[1] = (const double2*)U1;
    a.in[2] = (const double2*)U2;
    a.out[0] = (double2*)V0;
    a.out[1] = (double2*)V1;
    a.out[2] = (double2*)V2;
    HIPTRY(launch_filter(a, (hipStream_t)stream));
    return CF_OK;
}

// ---------------------------------------------------------------------------
// Subgrid-force MLP entry points (learned closure)
// ---------------------------------------------------------------------------
static MLPInferArgs mlp_args(const float* theta, const unsigned short* wbf, const float* norm, const double* grad9, int Nxp,
                             int Nzp, int Ny) {
    MLPInferArgs a;
    std::memset(&a, 0, sizeof(a));
    a.theta = theta;
    a.wbf = wbf;
    a.norm = norm;
    a.plane = (long long)Nxp * Nzp;
    a.npts = a.plane * Ny;
    a.Ny = Ny;
    for (int c = 0; c < 9; ++c) a.grad[c] = grad9 + (size_t)a.npts * c;
    return a;
}

static int mlp_dims_ok(int Nxp, int Nzp, int Ny) {
    if (Nxp < 1 || Nzp < 1 || Ny < 2) return fail(CF_ERR_ARGUMENT, "bad point grid");
    if ((long long)Nxp * Nzp * Ny >= (1ll << 40)) return fail(CF_ERR_UNSUPPORTED, "point grid too large");
    return 0;
}

extern "C" int cf_mlp_features(const double* grad9, const double* target3, int Nxp, int Nzp, int Ny, float* X, float* T,
                               void* stream) {
    if (!grad9 || !X || (target3 && !T)) return fail(CF_ERR_ARGUMENT, "null argument");
    if (int rc = mlp_dims_ok(Nxp, Nzp, Ny)) return rc;
    MLPInferArgs a = mlp_args(nullptr, nullptr, nullptr, grad9, Nxp, Nzp, Ny);
    if (target3)  // the target fields travel in the output slots
        for (int c = 0; c < 3; ++c) a.out[c] = const_cast<double*>(target3) + (size_t)a.npts * c;
    HIPTRY(launch_mlp_features(a, X, target3 ? T : nullptr, (hipStream_t)stream));
    return CF_OK;
}

extern "C" int cf_mlp_infer(const float* theta, const unsigned short* wbf, const float* norm, const double* grad9, int Nxp,
                            int Nzp, int Ny, double* force3, void* stream) {
    if (!theta || !wbf || !norm || !grad9 || !force3) return fail(CF_ERR_ARGUMENT, "null argument");
    if (int rc = mlp_dims_ok(Nxp, Nzp, Ny)) return rc;
    MLPInferArgs a = mlp_args(theta, wbf, norm, grad9, Nxp, Nzp, Ny);
    for (int c = 0; c < 3; ++c) a.out[c] = force3 + (size_t)a.npts * c;
    HIPTRY(launch_mlp_infer(a, (hipStream_t)stream));
    return CF_OK;
}

extern "C" long long cf_mlp_work_floats(long long batch) {
    return batch > 0 ? mlp_work_floats(batch) : 0;
}

extern "C" int cf_mlp_grad(const float* theta, const unsigned short* wbf, const float* norm, const float* X, const float* T,
                           const int* idx, long long batch, float* work, float* grad, void* stream) {
    if (!theta || !wbf || !norm || !X || !T || !idx || !work || !grad) return fail(CF_ERR_ARGUMENT, "null argument");
    if (batch < 1 || batch > (1ll << 30)) return fail(CF_ERR_ARGUMENT, "batch must be in 1..2^30");
    MLPTrainArgs a;
    std::memset(&a, 0, sizeof(a));
    a.theta = theta;
    a.wbf = wbf;
    a.norm = norm;
    a.X = X;
    a.T = T;
    a.idx = idx;
    a.batch = batch;
    a.part = work;
    HIPTRY(launch_mlp_grad(a, grad, (hipStream_t)stream));
    return CF_OK;
}

extern "C" int cf_mlp_adam(float* theta, unsigned short* wbf, float* m, float* v, const float* grad, float lr, int step,
                           void* stream) {
    if (!theta || !wbf || !m || !v || !grad) return fail(CF_ERR_ARGUMENT, "null argument");
    if (step < 1 || !(lr > 0.0f)) return fail(CF_ERR_ARGUMENT, "need step >= 1 and lr > 0");
    HIPTRY(launch_mlp_adam(theta, wbf, m, v, grad, lr, step, (hipStream_t)stream));
    return CF_OK;
}

// The registered closure's force of one stage into F, on the helper's
// workspaces: velocity gradient of U, the network at every padded point, the
// forward transform of its three outputs, and for a plane-mean closure every
// mode but (0, 0) zeroed.  The same kernels, in the same order, as the host
// forcing that calls cf_velocity_gradient, cf_mlp_infer and cf_forward3.
static int enqueue_closure(cf_problem* h, double2* const U[3], hipStream_t st) {
    const cf_problem::ClosureCfg& c = h->closure;
    cf_problem* hp = c.helper;
    const GridDims& gd = hp->gd;
    if (int rc = velocity_gradient_on(hp, U, c.grad9, st)) return rc;
    MLPInferArgs a = mlp_args(c.theta, c.wbf, c.norm, c.grad9, gd.Nxp, gd.Nzp, gd.Ny);
    for (int i = 0; i < 3; ++i) a.out[i] = c.force3 + (size_t)a.npts * i;
    HIPTRY(launch_mlp_infer(a, st));
    double* Fs[3] = {(double*)c.F[0], (double*)c.F[1], (double*)c.F[2]};
    if (int rc = forward_fields(hp, 3, c.force3, Fs, st)) return rc;
    if (c.mean_only) HIPTRY(launch_keep_mean(c.F[0], c.F[1], c.F[2], gd.S, gd.B, st));
    return 0;
}

extern "C" int cf_set_closure(cf_problem* h, cf_problem* helper, const float* theta, const unsigned short* wbf,
                              const float* norm, double* grad9, double* force3, double* F0, double* F1, double* F2,
                              int mean_only) {
    if (!h) return fail(CF_ERR_ARGUMENT, "null problem");
    cf_problem::ClosureCfg c;
    if (theta) {
        if (h->force_fn) return fail(CF_ERR_ARGUMENT, "a host forcing is registered on this problem");
        if (!helper || helper == h) return fail(CF_ERR_ARGUMENT, "the closure needs a helper problem of its own");
        if (!wbf || !norm) return fail(CF_ERR_ARGUMENT, "null network argument");
        const void* f[5] = {grad9, force3, F0, F1, F2};
        if (int rc = check_fields(f, 5)) return rc;
        const GridDims &a = h->gd, &b = helper->gd;
        if (helper->device != h->device) return fail(CF_ERR_ARGUMENT, "the helper must share the problem's device");
        if (a.Nx != b.Nx || a.Ny != b.Ny || a.Nz != b.Nz || h->Lx != helper->Lx || h->Lz != helper->Lz)
            return fail(CF_ERR_DIMENSION, "the helper must have the problem's grid");
        if (b.Nxp % 2 != 0 || helper->xwavesg < 1)
            return fail(CF_ERR_UNSUPPORTED, "the closure's velocity gradient needs an even padded Nx");
        if (int rc = mlp_dims_ok(b.Nxp, b.Nzp, b.Ny)) return rc;
        c.helper = helper;
        c.helper_serial = helper->serial;
        c.theta = theta;
        c.wbf = wbf;
        c.norm = norm;
        c.grad9 = grad9;
        c.force3 = force3;
        c.F[0] = (double2*)F0;
        c.F[1] = (double2*)F1;
        c.F[2] = (double2*)F2;
        c.mean_only = mean_only ? 1 : 0;
        DeviceGuard dg(h->device);
        if (!h->side && hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking) != hipSuccess)
            return fail(CF_ERR_DEVICE, "hipStreamCreate");
        if (!h->ev_fork && hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) != hipSuccess)
            return fail(CF_ERR_DEVICE, "hipEventCreate");
        if (!h->ev_join && hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming) != hipSuccess)
            return fail(CF_ERR_DEVICE, "hipEventCreate");
    }
    h->closure = c;
    if (!h->force_fn)
        for (int i = 0; i < 3; ++i) h->F[i] = c.F[i];
    return CF_OK;
}

extern "C" int cf_set_forcing(cf_problem* h, cf_forcing_fn fn, void* user, double* F0, double* F1, double* F2) {
    if (!h) return fail(CF_ERR_ARGUMENT, "null problem");
    if (fn && h->closure.active()) return fail(CF_ERR_ARGUMENT, "a closure is registered on this problem");
    if (fn) {
        const void* f[3] = {F0, F1, F2};
        if (int rc = check_fields(f, 3)) return rc;
    }
    h->force_fn = fn;
    h->force_user = fn ? user : nullptr;
    h->F[0] = fn ? (double2*)F0 : nullptr;
    h->F[1] = fn ? (double2*)F1 : nullptr;
    h->F[2] = fn ? (double2*)F2 : nullptr;
    return CF_OK;
}

extern "C" int cf_set_nonlinear(cf_problem* h, cf_forcing_fn fn, void* user, double* N0, double* N1, double* N2) {
    if (!h) return fail(CF_ERR_ARGUMENT, "null problem");
    if (fn) {
        const void* f[3] = {N0, N1, N2};
        if (int rc = check_fields(f, 3)) return rc;
    }
    h->nl_fn = fn;
    h->nl_user = fn ? user : nullptr;
    h->NL[0] = fn ? (double2*)N0 : nullptr;
    h->NL[1] = fn ? (double2*)N1 : nullptr;
    h->NL[2] = fn ? (double2*)N2 : nullptr;
    return CF_OK;
}

extern "C" int cf_step(cf_problem* h, double* U0, double* U1, double* U2, double* P, double t, int constraint,
                       const double* values, double* t_out, void* stream) {
    return cf_advance(h, U0, U1, U2, P, t, 1, constraint, values, t_out, stream);
}

extern "C" int cf_step_timed(cf_problem* h, double* U0, double* U1, double* U2, double* P, double t, int constraint,
                             const double* values, double* t_out, double ms[5], void* stream) {
    if (!h || !ms) return fail(CF_ERR_ARGUMENT, "null argument");
    const void* f[4] = {U0, U1, U2, P};
    if (int rc = check_fields(f, 4)) return rc;
    Constraint c;
    if (int rc = resolve_constraint(h, constraint, values, c)) return rc;
    DeviceGuard dg(h->device);
    hipStream_t st = (hipStream_t)stream;
    double2* U[3] = {(double2*)U0, (double2*)U1, (double2*)U2};
    Timer tm;
    if (int rc = enqueue_steps(h, U, (double2*)P, 1, c, st, &tm, 0.0, nullptr, t)) return rc;
    HIPTRY(hipEventSynchronize(tm.ev.back()));
    for (int i = 0; i < Timer::NCLASS; ++i) ms[i] = 0.0;
    for (size_t i = 1; i < tm.ev.size(); ++i) {
        float e = 0.f;
        HIPTRY(hipEventElapsedTime(&e, tm.ev[i - 1], tm.ev[i]));
        if (tm.cls[i] >= 0) ms[tm.cls[i]] += e;
    }
    if (t_out) *t_out = t + h->dt;
    return CF_OK;
}

extern "C" int cf_last_gradient(cf_problem* h, double out[2], void* stream) {
    if (!h || !out) return fail(CF_ERR_ARGUMENT, "null argument");
    DeviceGuard dg(h->device);
    HIPTRY(hipMemcpyAsync(out, h->d_grad, 2 * sizeof(double), hipMemcpyDeviceToHost, (hipStream_t)stream));
    HIPTRY(hipStreamSynchronize((hipStream_t)stream));
    return CF_OK;
}

// nf (1..3) fields per launch: one x, one z and one y launch for all of them
static int forward_fields(cf_problem* h, int nf, const double* u, double* const* U, hipStream_t st) {
    const GridDims& gd = h->gd;
    const size_t nphys = (size_t)gd.Nxp * gd.Nzp * gd.Ny;
    XArgs x = base_xargs(h);
    x.mode = X_FORWARD;
    x.nfields = nf;
    double2* zi[3];
    double2* zo[3];
    YArgs a = base_yargs(h);
    a.nfields = nf;
    for (int c = 0; c < nf; ++c) {
        x.rin[c] = u + nphys * c;
        x.out[c] = fld(h->zbuf, gd.Sz, c);
        zi[c] = fld(h->zbuf, gd.Sz, c);
        zo[c] = fld(h->six, gd.S, c);
        a.Nphys[c] = zo[c];
        a.Rout[c] = (double2*)U[c];
    }
    HIPTRY(launch_xpass(x, st));
    if (int rc = zpass(h, -1, nf, zi, zo, st)) return rc;
    HIPTRY(launch_ykernel(Y_DCT_FWD, a, st));
    return CF_OK;
}

static int inverse_fields(cf_problem* h, int nf, double* const* U, double* u, hipStream_t st) {
    const GridDims& gd = h->gd;
    const size_t nphys = (size_t)gd.Nxp * gd.Nzp * gd.Ny;
    YArgs a = base_yargs(h);
    a.nfields = nf;
    double2* zi[3];
    double2* zo[3];
    XArgs x = base_xargs(h);
    x.mode = X_INVERSE;
    x.nfields = nf;
    for (int c = 0; c < nf; ++c) {
        a.Rin[c] = (const double2*)U[c];
        a.out6[c] = fld(h->six, gd.S, c);
        zi[c] = a.out6[c];
        zo[c] = fld(h->zbuf, gd.Sz, c);
        x.in[c] = zo[c];
        x.rout[c] = u + nphys * c;
    }
    HIPTRY(launch_ykernel(Y_DCT_INV, a, st));
    if (int rc = zpass(h, +1, nf, zi, zo, st)) return rc;
    HIPTRY(launch_xpass(x, st));
    return CF_OK;
}

extern "C" int cf_forward(cf_problem* h, const double* u, double* U, void* stream) {
    if (!h || !u || !U) return fail(CF_ERR_ARGUMENT, "null argument");
    DeviceGuard dg(h->device);
    double* Us[1] = {U};
    return forward_fields(h, 1, u, Us, (hipStream_t)stream);
}

extern "C" int cf_forward3(cf_problem* h, const double* u, double* U0, double* U1, double* U2, void* stream) {
    if (!h || !u) return fail(CF_ERR_ARGUMENT, "null argument");
    const void* f[3] = {U0, U1, U2};
    if (int rc = check_fields(f, 3)) return rc;
    DeviceGuard dg(h->device);
    double* Us[3] = {U0, U1, U2};
    return forward_fields(h, 3, u, Us, (hipStream_t)stream);
}

extern "C" int cf_inverse(cf_problem* h, const double* U, double* u, void* stream) {
    if (!h || !u || !U) return fail(CF_ERR_ARGUMENT, "null argument");
    DeviceGuard dg(h->device);
    double* Us[1] = {const_cast<double*>(U)};
    return inverse_fields(h, 1, Us, u, (hipStream_t)stream);
}

extern "C" int cf_inverse3(cf_problem* h, const double* U0, const double* U1, const double* U2, double* u, void* stream) {
    if (!h || !u) return fail(CF_ERR_ARGUMENT, "null argument");
    const void* f[3] = {U0, U1, U2};
    if (int rc = check_fields(f, 3)) return rc;
    DeviceGuard dg(h->device);
    double* Us[3] = {const_cast<double*>(U0), const_cast<double*>(U1), const_cast<double*>(U2)};
    return inverse_fields(h, 3, Us, u, (hipStream_t)stream);
}

// nonlinear term of U into spectral N (3 fields)
static int enqueue_nonlinear_spec(cf_problem* h, double2* const U[3], double2* const N[3], hipStream_t st) {
    if (int rc = enqueue_prologue(h, U, st)) return rc;
    if (int rc = enqueue_nonlinear_phys(h, st)) return rc;
    if (h->form == CF_FORM_DIVERGENCE) return enqueue_nonlinear_coef(h, N, st);
    YArgs a = base_yargs(h);
    a.nl_sign = (h->form == CF_FORM_CONVECTIVE) ? -1.0 : 1.0;
    a.nfields = 3;
    for (int i = 0; i < 3; ++i) {
        a.Nphys[i] = fld(h->six, h->gd.S, i);
        a.Rout[i] = N[i];
    }
    HIPTRY(launch_ykernel(Y_DCT_FWD, a, st));
    return 0;
}

extern "C" int cf_nonlinear(cf_problem* h, const double* U0, const double* U1, const double* U2, double* N0, double* N1,
                            double* N2, void* stream) {
    if (!h) return fail(CF_ERR_ARGUMENT, "null problem");
    const void* f[6] = {U0, U1, U2, N0, N1, N2};
    if (int rc = check_fields(f, 6)) return rc;
    DeviceGuard dg(h->device);
    double2* U[3] = {(double2*)U0, (double2*)U1, (double2*)U2};
    double2* N[3] = {(double2*)N0, (double2*)N1, (double2*)N2};
    return enqueue_nonlinear_spec(h, U, N, (hipStream_t)stream);
}

static int enqueue_solve(cf_problem* h, DevBank* bk, const double2* const R[3], double2* const U[3], double2* P,
                         const Constraint& c, hipStream_t st) {
    YArgs a = base_yargs(h);
    a.bank = bk->view();
    a.nu = bk->nu;
    a.bulk = c.bulk;
    a.c0 = c.c0;
    a.c1 = c.c1;
    for (int i = 0; i < 3; ++i) {
        a.Rin[i] = R[i];
        a.U[i] = U[i];
    }
    a.P = P;
    a.grad_out = h->d_grad;
    HIPTRY(launch_ykernel(Y_SOLVE, a, st));
    return 0;
}

extern "C" int cf_stokes_solve(cf_problem* h, double nu, double lambda, const double* R0, const double* R1,
                               const double* R2, double* U0, double* U1, double* U2, double* P, int constraint,
                               const double* values, double* grads, void* stream) {
    if (!h) return fail(CF_ERR_ARGUMENT, "null problem");
    const void* f[7] = {U0, U1, U2, P, R0, R1, R2};
    if (int rc = check_fields(f, 7)) return rc;
    if (!(std::isfinite(nu) && nu > 0)) return fail(CF_ERR_ARGUMENT, "nu must be finite and positive");
    if (!(std::isfinite(lambda) && lambda >= 0)) return fail(CF_ERR_ARGUMENT, "lambda must be finite and nonnegative");
    Constraint c;
    if (constraint == -1) {
        c.bulk = 0;
        c.c0 = c.c1 = 0.0;
    } else if (int rc = resolve_constraint(h, constraint, values, c)) {
        return rc;
    }
    DeviceGuard dg(h->device);
    DevBank* bk = nullptr;
    if (int rc = get_bank(h, nu, lambda, &bk)) return rc;
    const double2* R[3] = {(const double2*)R0, (const double2*)R1, (const double2*)R2};
    double2* U[3] = {(double2*)U0, (double2*)U1, (double2*)U2};
    hipStream_t st = (hipStream_t)stream;
    if (int rc = enqueue_solve(h, bk, R, U, (double2*)P, c, st)) return rc;
    if (grads) {
        HIPTRY(hipMemcpyAsync(grads, h->d_grad, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPTRY(hipStreamSynchronize(st));
    }
    return CF_OK;
}

extern "C" int cf_project(cf_problem* h, double* U0, double* U1, double* U2, void* stream) {
    if (!h) return fail(CF_ERR_ARGUMENT, "null problem");
    const void* f[3] = {U0, U1, U2};
    if (int rc = check_fields(f, 3)) return rc;
    DeviceGuard dg(h->device);
    hipStream_t st = (hipStream_t)stream;
    DevBank* bk = nullptr;
    if (int rc = get_bank(h, 1.0, 0.0, &bk)) return rc;
    double2* U[3] = {(double2*)U0, (double2*)U1, (double2*)U2};
    double2* R[3];
    for (int i = 0; i < 3; ++i) R[i] = fld(h->R, h->gd.S, i);
    YArgs a = base_yargs(h);
    for (int i = 0; i < 3; ++i) {
        a.U[i] = U[i];
        a.Rout[i] = R[i];
    }
    HIPTRY(launch_ykernel(Y_NEGLAP, a, st));
    Constraint c;
    if (h->constraint == CF_CONSTRAINT_BULK_VELOCITY) {
        if (int rc = resolve_constraint(h, CF_CONSTRAINT_BULK_VELOCITY, h->cvals, c)) return rc;
    } else {
        c.bulk = 0;
        c.c0 = c.c1 = 0.0;
    }
    const double2* Rc[3] = {R[0], R[1], R[2]};
    return enqueue_solve(h, bk, Rc, U, h->six, c, st);
}

extern "C" int cf_pressure(cf_problem* h, const double* U0, const double* U1, const double* U2, double* P,
                           void* stream) {
    if (!h) return fail(CF_ERR_ARGUMENT, "null problem");
    const void* f[4] = {U0, U1, U2, P};
    if (int rc = check_fields(f, 4)) return rc;
    DeviceGuard dg(h->device);
    hipStream_t st = (hipStream_t)stream;
    double2* U[3] = {(double2*)U0, (double2*)U1, (double2*)U2};
    double2* A[3];
    for (int i = 0; i < 3; ++i) A[i] = fld(h->R, h->gd.S, i);
    if (int rc = enqueue_nonlinear_spec(h, U, A, st)) return rc;
    YArgs a = base_yargs(h);
    for (int i = 0; i < 3; ++i) a.Rin[i] = A[i];
    a.U[1] = U[1];
    a.P = (double2*)P;
    HIPTRY(launch_ykernel(Y_PRESSURE, a, st));
    return CF_OK;
}

// ---------------------------------------------------------------------------
// diagnostics (postprocessing.jl:30-128)
// ---------------------------------------------------------------------------
static double host_bulkmean_real(const double2* m, int Ny) {
    std::vector<double> re((size_t)Ny);
    for (int n = 0; n < Ny; ++n) re[n] = m[n].x;
    return host_bulkmean(re.data(), Ny);
}

static int enqueue_dot(cf_problem* h, int which, double2* const U[3], const double2* const V[3], double scale, double* out,
                       hipStream_t st) {
    YArgs a = base_yargs(h);
    for (int i = 0; i < 3; ++i) {
        a.U[i] = U[i];
        a.Rin[i] = V ? V[i] : nullptr;
    }
    double* part = reinterpret_cast<double*>(h->R);  // >= one double per block
    a.partial = part;
    HIPTRY(launch_ykernel(which, a, st));
    HIPTRY(launch_reduce(part, (int)y_blocks(h->gd), h->d_scalar, st));
    double r = 0.0;
    HIPTRY(hipMemcpyAsync(&r, h->d_scalar, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPTRY(hipStreamSynchronize(st));
    *out = scale * r;
    return CF_OK;
}

extern "C" int cf_dot(cf_problem* h, const double* U0, const double* U1, const double* U2, const double* V0, const double* V1,
                      const double* V2, double* out, void* stream) {
    if (!h || !out) return fail(CF_ERR_ARGUMENT, "null argument");
    const void* fu[3] = {U0, U1, U2};
    const void* fv[3] = {V0, V1, V2};
    if (int rc = check_fields(fu, 3)) return rc;
    if (int rc = check_fields(fv, 3)) return rc;
    DeviceGuard dg(h->device);
    double2* U[3] = {(double2*)U0, (double2*)U1, (double2*)U2};
    const double2* V[3] = {(const double2*)V0, (const double2*)V1, (const double2*)V2};
    return enqueue_dot(h, Y_DOT, U, V, 1.0, out, (hipStream_t)stream);
}

extern "C" int cf_dissipation(cf_problem* h, const double* U0, const double* U1, const double* U2, double nu, double* out,
                              void* stream) {
    if (!h || !out) return fail(CF_ERR_ARGUMENT, "null argument");
    const void* fu[3] = {U0, U1, U2};
    if (int rc = check_fields(fu, 3)) return rc;
    if (!std::isfinite(nu)) return fail(CF_ERR_ARGUMENT, "nu must be finite");
    DeviceGuard dg(h->device);
    double2* U[3] = {(double2*)U0, (double2*)U1, (double2*)U2};
    return enqueue_dot(h, Y_DISSIPATION, U, nullptr, nu, out, (hipStream_t)stream);
}

extern "C" int cf_power_input(cf_problem* h, const double* U0, const double* U2, double nu, const double grad[2],
                              double* out, void* stream) {
    if (!h || !U0 || !U2 || !out) return fail(CF_ERR_ARGUMENT, "null argument");
    if (!std::isfinite(nu)) return fail(CF_ERR_ARGUMENT, "nu must be finite");
    const double g[2] = {grad ? grad[0] : 0.0, grad ? grad[1] : 0.0};
    if (!(std::isfinite(g[0]) && std::isfinite(g[1]))) return fail(CF_ERR_ARGUMENT, "pressure gradient must be finite");
    DeviceGuard dg(h->device);
    hipStream_t st = (hipStream_t)stream;
    const int Ny = h->gd.Ny;
    // zero Fourier mode column: element (0, 0, n) at n * B
    std::vector<double2> col(2 * (size_t)Ny);
    HIPTRY(hipMemcpy2DAsync(col.data(), sizeof(double2), U0, (size_t)h->gd.B * sizeof(double2), sizeof(double2), Ny,
                            hipMemcpyDeviceToHost, st));
    HIPTRY(hipMemcpy2DAsync(col.data() + Ny, sizeof(double2), U2, (size_t)h->gd.B * sizeof(double2), sizeof(double2), Ny,
                            hipMemcpyDeviceToHost, st));
    HIPTRY(hipStreamSynchronize(st));
    double input = 0.0;
    for (int c = 0; c < 2; ++c) {
        const double2* m = col.data() + (size_t)c * Ny;
        // wall values and the one-sided derivatives of the mean profile
        double up_r = 0, up_i = 0, lo_r = 0, lo_i = 0, dup_r = 0, dup_i = 0, dlo_r = 0, dlo_i = 0;
        for (int n = 0; n < Ny; ++n) {
            const double sgn = (n % 2 == 0) ? 1.0 : -1.0, n2 = (double)n * n;
            up_r += m[n].x;
            up_i += m[n].y;
            lo_r += sgn * m[n].x;
            lo_i += sgn * m[n].y;
            dup_r += n2 * m[n].x;
            dup_i += n2 * m[n].y;
            dlo_r -= sgn * n2 * m[n].x;  // T_n'(-1) = (-1)^(n+1) n^2
            dlo_i -= sgn * n2 * m[n].y;
        }
        // nu/2 Re(conj(u(+1)) u'(+1) - conj(u(-1)) u'(-1))
        input += 0.5 * nu * ((up_r * dup_r + up_i * dup_i) - (lo_r * dlo_r + lo_i * dlo_i));
        input -= g[c] * host_bulkmean_real(m, Ny);
    }
    *out = input;
    return CF_OK;
}

/* plane statistics of the total velocity at the Lobatto points */
extern "C" int cf_plane_statistics(cf_problem* h, const double* U0, const double* U1, const double* U2, double* out,
                                   void* stream) {
    if (!h || !out) return fail(CF_ERR_ARGUMENT, "null argument");
    const void* fu[3] = {U0, U1, U2};
    if (int rc = check_fields(fu, 3)) return rc;
    DeviceGuard dg(h->device);
    hipStream_t st = (hipStream_t)stream;
    const int Ny = h->gd.Ny, nq = Ny * PLANE_STATS_Q;
    YArgs a = base_yargs(h);
    a.U[0] = (double2*)U0;
    a.U[1] = (double2*)U1;
    a.U[2] = (double2*)U2;
    double* part = reinterpret_cast<double*>(h->R);  // nblocks * nq doubles <= 6 S
    a.partial = part;
    HIPTRY(launch_ykernel(Y_PLANE_STATS, a, st));
    double* red = reinterpret_cast<double*>(h->Q);  // nq doubles (Q is free between steps)
    HIPTRY(launch_reduce_columns(part, (int)y_blocks(h->gd), nq, red, st));
    std::vector<double> tmp((size_t)nq);
    HIPTRY(hipMemcpyAsync(tmp.data(), red, nq * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPTRY(hipStreamSynchronize(st));
    // out: Ny x 9, columns U V W uu vv ww uv uw vw
    std::memcpy(out, tmp.data(), nq * sizeof(double));
    return CF_OK;
}

// ---------------------------------------------------------------------------
// random_state on the device (src/initialization.jl:32-51)
// ---------------------------------------------------------------------------
extern "C" int cf_gaussian_noise(cf_problem* h, double epsilon, unsigned long long seed, double* out, void* stream) {
    if (!h || !out) return fail(CF_ERR_ARGUMENT, "null argument");
    if (!(std::isfinite(epsilon) && epsilon >= 0)) return fail(CF_ERR_ARGUMENT, "epsilon must be finite and non-negative");
    DeviceGuard dg(h->device);
    const long long n = (long long)h->gd.Nxp * h->gd.Nzp * h->gd.Ny;
    HIPTRY(launch_gaussian(out, n, epsilon, seed, (hipStream_t)stream));
    return CF_OK;
}

extern "C" int cf_random_state(cf_problem* h, double epsilon, unsigned long long seed, double* U0, double* U1, double* U2,
                               double* P, void* stream) {
    if (!h) return fail(CF_ERR_ARGUMENT, "null problem");
    const void* f[4] = {U0, U1, U2, P};
    if (int rc = check_fields(f, 4)) return rc;
    if (!(std::isfinite(epsilon) && epsilon >= 0)) return fail(CF_ERR_ARGUMENT, "epsilon must be finite and non-negative");
    DeviceGuard dg(h->device);
    const long long n = (long long)h->gd.Nxp * h->gd.Nzp * h->gd.Ny;
    if (!h->d_noise)
        if (int rc = dalloc(h, &h->d_noise, 3 * (size_t)n)) return rc;
    if (int rc = cf_gaussian_noise(h, epsilon, seed, h->d_noise, stream)) return rc;
    if (int rc = cf_forward3(h, h->d_noise, U0, U1, U2, stream)) return rc;
    if (int rc = cf_project(h, U0, U1, U2, stream)) return rc;
    return cf_pressure(h, U0, U1, U2, P, stream);
}

// Tile autotuner of one problem (FP64 step): times the z passes of the
// nonlinear term for each in-place tile width and the x pass for each
// wavefront count on the problem's workspaces, and keeps the fastest.  The
// results of every pass do not depend on the tiling.
extern "C" int cf_problem_tune(cf_problem* h, void* stream) {
    if (!h) return fail(CF_ERR_ARGUMENT, "null problem");
    DeviceGuard dg(h->device);
    if (int rc = invalidate_graphs(h)) return rc;
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t a, b;
    HIPTRY(hipEventCreate(&a));
    HIPTRY(hipEventCreate(&b));
    auto timed = [&](auto&& fn, float* best) -> int {
        *best = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
            HIPTRY(hipEventRecord(a, st));
            for (int k = 0; k < 3; ++k)
                if (int rc = fn()) return rc;
            HIPTRY(hipEventRecord(b, st));
            HIPTRY(hipEventSynchronize(b));
            float t = 0;
            HIPTRY(hipEventElapsedTime(&t, a, b));
            *best = std::min(*best, t);
        }
        return 0;
    };
    int rc = 0;
    // z passes: both directions of the form's nonlinear term
    const int keepKT = h->KT, keepPPT = h->zppt;
    int bestKT = keepKT, bestPPT = keepPPT;
    float best = 1e30f;
    for (int kt = 1; kt <= std::min(32, h->gd.Nxh) && !rc; ++kt) {
        const int ppt = z_inplace_ppt(h->pz, kt);
        if (!ppt && !(kt == keepKT && keepPPT == 0)) continue;
        if (z_lds_bytes([&] {
                ZArgs z;
                std::memset(&z, 0, sizeof(z));
                z.gd = h->gd;
                z.KT = kt;
                z.ld = h->zld;
                z.ppt = ppt;
                return z;
            }()) > 64 * 1024)
            continue;
        h->KT = kt;
        h->zppt = ppt;
        const NonlinearArgs na = nonlinear_args(h, false);
        float ms;
        rc = timed([&] {
            HIPTRY(launch_zpass(+1, na.zin, st));
            HIPTRY(launch_zpass(-1, na.zout, st));
            return 0;
        }, &ms);
        if (!rc && ms < best) {
            best = ms;
            bestKT = kt;
            bestPPT = ppt;
        }
    }
    h->KT = bestKT;
    h->zppt = bestPPT;
    // x pass: wavefronts per block (even padded Nx), and for the rotating
    // form the six-line or the split (three-line) variant
    if (!rc && h->gd.Nxp % 2 == 0) {
        const int keepSplit = h->xsplit;
        int bestW[3] = {h->xwaves, h->xwavess, h->xwavesa}, bestSplit = keepSplit;
        int* waves[3] = {&h->xwaves, &h->xwavess, &h->xwavesa};
        const int* maxv = h->xmaxw;
        best = 1e30f;
        for (int split = 0; split < 3 && !rc; ++split) {
            if (maxv[split] < 1 || (split > 0 && h->form != CF_FORM_ROTATING)) continue;
            h->xsplit = split;
            for (int w = 1; w <= maxv[split] && !rc; ++w) {
                *waves[split] = w;
                const NonlinearArgs na = nonlinear_args(h, false);
                float ms;
                rc = timed([&] {
                    HIPTRY(launch_xpass(na.x, st));
                    return 0;
                }, &ms);
                if (!rc && ms < best) {
                    best = ms;
                    bestSplit = split;
                    bestW[split] = w;
                }
            }
        }
        h->xsplit = rc ? keepSplit : bestSplit;
        h->xwaves = bestW[0];
        h->xwavess = bestW[1];
        h->xwavesa = bestW[2];
    }
    // group layout of the step's stage and prologue kernels: the default or,
    // where built, K = 8 (on half the lanes where the default is K = 4, on
    // the same 64 lanes where it is K = 6) or K = 6 on 32 lanes (where the
    // default is 64 lanes of 4 and Ny <= 192);
    // timed on stage-2 launches over the workspaces (U in R, P in the z
    // buffer; no gradient output)
    if (!rc) {
        const int L0 = y_group_lanes(h->gd.Ny), K0 = y_group_k(h->gd.Ny);
        int cand[3][2] = {{L0, K0}, {K0 == 6 ? L0 : L0 / 2, 8}, {32, 6}};
        int ncand = ((K0 == 4 && L0 >= 32) || K0 == 6) ? 2 : 1;
        if (K0 == 4 && L0 == 64 && h->gd.Ny <= 192) ncand = 3;  // 32 lanes of 6 also cover Ny
        const int keepL = h->yl, keepK = h->yk;
        int bestL = keepL, bestK = keepK;
        best = 1e30f;
        const bool mixed = h->mixed != 0;
        for (int ci = 0; ci < ncand && !rc; ++ci) {
            const size_t lds = y_lds_bytes_layout(h->gd.Ny, h->yld, cand[ci][0], cand[ci][1]);
            if (lds == 0 || lds > 160 * 1024) continue;
            h->yl = cand[ci][0];
            h->yk = cand[ci][1];
            double2* U[3] = {fld(h->R, h->gd.S, 0), fld(h->R, h->gd.S, 1), fld(h->R, h->gd.S, 2)};
            double2* Ncoef[3] = {fld(h->Q, h->gd.S, 0), fld(h->Q, h->gd.S, 1), fld(h->Q, h->gd.S, 2)};
            const Constraint c0{0, 0.0, 0.0};
            YArgs a = stage_args(h, U, h->zbuf, 2, c0, h->dt, &h->stage[2], mixed, Ncoef);
            a.grad_out = nullptr;
            float ms;
            rc = timed([&] {
                HIPTRY(launch_ykernel(Y_STAGE, a, st));
                return 0;
            }, &ms);
            if (!rc && ms < best) {
                best = ms;
                bestL = h->yl;
                bestK = h->yk;
            }
        }
        h->yl = rc ? keepL : bestL;
        h->yk = rc ? keepK : bestK;
    }
    // mixed step: its FP32 x pass (six-line or split, wavefronts per block)
    if (!rc && h->mixed && h->gd.Nxp % 2 == 0) {
        const int keepSplit = h->xsplit32;
        int bestW[2] = {h->xwaves32, h->xwavess32}, bestSplit = keepSplit;
        int* waves[2] = {&h->xwaves32, &h->xwavess32};
        best = 1e30f;
        for (int split = 0; split < 2 && !rc; ++split) {
            if (h->xmaxw32[split] < 1 || (split > 0 && h->form != CF_FORM_ROTATING)) continue;
            h->xsplit32 = split;
            for (int w = 1; w <= h->xmaxw32[split] && !rc; ++w) {
                *waves[split] = w;
                const NonlinearArgs na = nonlinear_args(h, true);
                float ms;
                rc = timed([&] {
                    HIPTRY(launch_xpass(na.x, st));
                    return 0;
                }, &ms);
                if (!rc && ms < best) {
                    best = ms;
                    bestSplit = split;
                    bestW[split] = w;
                }
            }
        }
        h->xsplit32 = rc ? keepSplit : bestSplit;
        h->xwaves32 = bestW[0];
        h->xwavess32 = bestW[1];
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}

extern "C" int cf_problem_tiles(const cf_problem* h, int out[3]) {
    if (!h || !out) return fail(CF_ERR_ARGUMENT, "null argument");
    out[0] = h->KT;
    out[1] = h->zppt;
    out[2] = h->xsplit == 2 ? h->xwavesa : h->xsplit ? h->xwavess : h->xwaves;  // of the x pass variant in use
    return CF_OK;
}

extern "C" int cf_problem_get_tiles(const cf_problem* h, int tiles[5]) {
    if (!h || !tiles) return fail(CF_ERR_ARGUMENT, "null argument");
    tiles[0] = h->KT;
    tiles[1] = h->zppt;
    tiles[2] = h->xsplit;
    tiles[3] = h->gd.Nxp % 2 ? 0 : h->xsplit == 2 ? h->xwavesa : h->xsplit ? h->xwavess : h->xwaves;
    tiles[4] = h->yk;
    return CF_OK;
}

extern "C" int cf_problem_set_tiles(cf_problem* h, const int tiles[5]) {
    if (!h || !tiles) return fail(CF_ERR_ARGUMENT, "null argument");
    // z tile: an in-place width (one LDS buffer) or, for plans that need two
    // buffers, a width whose two buffers fit 64 KB
    const int kt = tiles[0];
    if (kt < 1 || kt > std::min(32, h->gd.Nxh)) return fail(CF_ERR_ARGUMENT, "z tile width out of range");
    ZArgs z;
    std::memset(&z, 0, sizeof(z));
    z.gd = h->gd;
    z.KT = kt;
    z.ld = h->zld;
    z.ppt = z_inplace_ppt(h->pz, kt);
    if (!z.ppt && kt > 16) return fail(CF_ERR_ARGUMENT, "two-buffer z tiles hold at most 16 columns");
    if (z_lds_bytes(z) > 64 * 1024) return fail(CF_ERR_ARGUMENT, "z tile exceeds 64 KB of LDS");
    const int var = tiles[2], w = tiles[3];
    if (h->gd.Nxp % 2 == 0) {
        if (var < 0 || var > 2 || (var > 0 && h->form != CF_FORM_ROTATING) || h->xmaxw[var] < 1)
            return fail(CF_ERR_ARGUMENT, "x pass variant unavailable for this problem");
        if (w < 1 || w > h->xmaxw[var]) return fail(CF_ERR_ARGUMENT, "x wavefronts per block out of range");
    } else if (var != 0) {
        return fail(CF_ERR_ARGUMENT, "odd padded Nx has one x pass");
    }
    // wall-normal layout: K coefficients per lane on the fewest lanes covering Ny
    const int yk = tiles[4];
    int yl = 0;
    if (yk == y_group_k(h->gd.Ny)) {
        yl = y_group_lanes(h->gd.Ny);
    } else if (yk == 6 || yk == 8) {
        yl = 16;
        while (yl * yk < h->gd.Ny) yl *= 2;
        const size_t lds = yl <= 64 ? y_lds_bytes_layout(h->gd.Ny, h->yld, yl, yk) : 0;
        if (lds == 0 || lds > 160 * 1024) return fail(CF_ERR_ARGUMENT, "wall-normal layout unavailable for this Ny");
    } else {
        return fail(CF_ERR_ARGUMENT, "wall-normal coefficients per lane must be 4, 6 or 8");
    }
    DeviceGuard dg(h->device);
    if (int rc = invalidate_graphs(h)) return rc;
    h->KT = kt;
    h->zppt = z.ppt;
    if (h->gd.Nxp % 2 == 0) {
        h->xsplit = var;
        (var == 2 ? h->xwavesa : var == 1 ? h->xwavess : h->xwaves) = w;
    }
    h->yl = yl;
    h->yk = yk;
    return CF_OK;
}

// ---------------------------------------------------------------------------
// Ensembles: many independent problems advanced together.  Each kernel class
// of the step is one launch per group of instances that share its template
// parameters (y: group lanes L; z: in-place width; x: form, in-place width and
// wavefronts per block); a block finds its instance from a prefix table of
// block counts and reads that instance's argument block from device memory.
// One ensemble step is captured once into a HIP graph and replayed.
// ---------------------------------------------------------------------------
namespace {

enum { B_YPRO = 0, B_ZINV, B_X, B_ZFWD, B_YSTAGE };

struct BatchLaunch {
    int cls;        // B_*
    int tp, waves;  // template parameter (L or in-place width), x wavefronts per block
    int mode;       // x form
    const void* tab;
    const unsigned* start;
    int n;
    unsigned nblocks;
    size_t lds;
};

// per-instance tile choices (the tunables): z tile width and x wavefronts per block
struct Tiles {
    int KT, zppt, xwaves;
};

}  // namespace

struct cf_ensemble {
    int device = 0;
    std::vector<cf_problem*> probs;
    std::vector<std::array<double2*, 4>> fields;  // U0, U1, U2, P per instance
    std::vector<Tiles> tiles;
    std::vector<void*> allocs;
    std::vector<BatchLaunch> step;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static void ens_free(cf_ensemble* e) {
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    if (e->cap) (void)hipStreamDestroy(e->cap);
    for (void* p : e->allocs) (void)hipFree(p);
    e->exec = nullptr;
    e->graph = nullptr;
    e->cap = nullptr;
    e->allocs.clear();
}

static hipError_t run_batch(const BatchLaunch& b, hipStream_t st) {
    switch (b.cls) {
        case B_YPRO: return launch_ykernel_batch(Y_PROLOGUE, b.tp, (const YArgs*)b.tab, b.start, b.n, b.nblocks, b.lds, st);
        case B_YSTAGE: return launch_ykernel_batch(Y_STAGE, b.tp, (const YArgs*)b.tab, b.start, b.n, b.nblocks, b.lds, st);
        case B_ZINV: return launch_zpass_batch(+1, b.tp, (const ZArgs*)b.tab, b.start, b.n, b.nblocks, b.lds, st);
        case B_ZFWD: return launch_zpass_batch(-1, b.tp, (const ZArgs*)b.tab, b.start, b.n, b.nblocks, b.lds, st);
        case B_X: return launch_xpass_batch(b.mode, b.tp, b.waves, (const XArgs*)b.tab, b.start, b.n, b.nblocks, b.lds, st);
        default: return hipErrorInvalidValue;
    }
}

// Upload one group's argument table and block prefix; append its launch.
template <typename A>
static int ens_group(cf_ensemble* e, int cls, int tp, int waves, int mode, const std::vector<A>& args,
                     const std::vector<unsigned>& blocks, const std::vector<size_t>& lds, std::vector<BatchLaunch>& out) {
    const int n = (int)args.size();
    std::vector<unsigned> start(n + 1, 0);
    size_t need = 0;
    for (int i = 0; i < n; ++i) {
        start[i + 1] = start[i] + blocks[i];
        need = std::max(need, lds[i]);
    }
    if ((unsigned long long)start[n] >= (1ull << 31)) return fail(CF_ERR_UNSUPPORTED, "ensemble launch too large");
    void* t = nullptr;
    void* s = nullptr;
    HIPTRY(hipMalloc(&t, sizeof(A) * n));
    e->allocs.push_back(t);
    HIPTRY(hipMalloc(&s, sizeof(unsigned) * (n + 1)));
    e->allocs.push_back(s);
    HIPTRY(hipMemcpy(t, args.data(), sizeof(A) * n, hipMemcpyHostToDevice));
    HIPTRY(hipMemcpy(s, start.data(), sizeof(unsigned) * (n + 1), hipMemcpyHostToDevice));
    out.push_back(BatchLaunch{cls, tp, waves, mode, t, (const unsigned*)s, n, start[n], need});
    return 0;
}

// Group instances by key and emit one launch per group, in first-seen order.
template <typename A, typename KeyFn, typename InfoFn>
static int ens_class(cf_ensemble* e, int cls, const std::vector<A>& all, KeyFn key, InfoFn info,
                     std::vector<BatchLaunch>& out) {
    std::vector<std::array<int, 3>> keys;
    for (const A& a : all) {
        std::array<int, 3> k = key(a);
        bool seen = false;
        for (auto& q : keys) seen = seen || (q == k);
        if (!seen) keys.push_back(k);
    }
    for (auto& k : keys) {
        std::vector<A> args;
        std::vector<unsigned> blocks;
        std::vector<size_t> lds;
        for (const A& a : all)
            if (key(a) == k) {
                unsigned b;
                size_t l;
                info(a, b, l);
                args.push_back(a);
                blocks.push_back(b);
                lds.push_back(l);
            }
        if (int rc = ens_group(e, cls, k[0], k[1], k[2], args, blocks, lds, out)) return rc;
    }
    return 0;
}

// Launch tables of one ensemble step for the current tile choices.
static int ens_build(cf_ensemble* e, std::vector<BatchLaunch>& step) {
    const int n = (int)e->probs.size();
    std::vector<YArgs> pro(n);
    std::vector<ZArgs> zin(n), zout(n);
    std::vector<XArgs> xs(n);
    std::vector<std::array<YArgs, 3>> stages(n);
    for (int i = 0; i < n; ++i) {
        cf_problem* h = e->probs[i];
        double2* U[3] = {e->fields[i][0], e->fields[i][1], e->fields[i][2]};
        Constraint c;
        if (int rc = resolve_constraint(h, -1, nullptr, c)) return rc;
        pro[i] = prologue_args(h, U, false);
        NonlinearArgs na = nonlinear_args(h, false);
        const Tiles& t = e->tiles[i];
        for (ZArgs* z : {&na.zin, &na.zout}) {
            z->KT = t.KT;
            z->ppt = t.zppt;
            z->mul_kt = recip_mul((unsigned)t.KT);
        }
        na.x.waves = t.xwaves;
        zin[i] = na.zin;
        zout[i] = na.zout;
        xs[i] = na.x;
        double2* Nc[3] = {nullptr, nullptr, nullptr};
        for (int j = 0; j < 3; ++j) stages[i][j] = stage_args(h, U, e->fields[i][3], j, c, h->dt, &h->stage[j], false, Nc);
    }
    auto ykey = [](const YArgs& a) { return std::array<int, 3>{y_group_lanes(a.gd.Ny), 0, 0}; };
    auto yinfo = [](const YArgs& a, unsigned& b, size_t& l) {
        b = y_blocks(a.gd);
        l = y_lds_bytes(a.gd.Ny, a.ld);
    };
    auto zkey = [](const ZArgs& a) { return std::array<int, 3>{a.ppt, 0, 0}; };
    auto zinfo = [](const ZArgs& a, unsigned& b, size_t& l) {
        b = z_blocks(a);
        l = z_lds_bytes(a);
    };
    auto xkey = [](const XArgs& a) { return std::array<int, 3>{a.ppt, a.waves, a.mode}; };
    auto xinfo = [](const XArgs& a, unsigned& b, size_t& l) {
        b = x_blocks(a);
        l = x_lds_bytes(a);
    };
    step.clear();
    if (int rc = ens_class(e, B_YPRO, pro, ykey, yinfo, step)) return rc;
    for (int j = 0; j < 3; ++j) {
        std::vector<YArgs> st(n);
        for (int i = 0; i < n; ++i) st[i] = stages[i][j];
        if (int rc = ens_class(e, B_ZINV, zin, zkey, zinfo, step)) return rc;
        if (int rc = ens_class(e, B_X, xs, xkey, xinfo, step)) return rc;
        if (int rc = ens_class(e, B_ZFWD, zout, zkey, zinfo, step)) return rc;
        if (int rc = ens_class(e, B_YSTAGE, st, ykey, yinfo, step)) return rc;
    }
    return 0;
}

// Autotuner: per distinct grid, time the batched z inverse pass for each
// in-place tile width KT (and the x pass for each wavefront count) on the
// instances' own workspaces (results do not depend on the tiling: every
// line is transformed by the same butterflies), keep the fastest.
static int ens_time_class(cf_ensemble* e, int cls, float* ms) {
    std::vector<BatchLaunch> step;
    if (int rc = ens_build(e, step)) return rc;
    hipEvent_t a, b;
    HIPTRY(hipEventCreate(&a));
    HIPTRY(hipEventCreate(&b));
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
        HIPTRY(hipEventRecord(a, e->cap));
        for (const BatchLaunch& L : step)
            if (L.cls == cls) HIPTRY(run_batch(L, e->cap));
        HIPTRY(hipEventRecord(b, e->cap));
        HIPTRY(hipEventSynchronize(b));
        float t = 0;
        HIPTRY(hipEventElapsedTime(&t, a, b));
        best = std::min(best, t);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = best;
    return 0;
}

static int ens_tune(cf_ensemble* e) {
    const int n = (int)e->probs.size();
    std::vector<std::array<int, 3>> grids;
    for (cf_problem* h : e->probs) {
        std::array<int, 3> g{h->gd.Nx, h->gd.Ny, h->gd.Nz};
        bool seen = false;
        for (auto& q : grids) seen = seen || (q == g);
        if (!seen) grids.push_back(g);
    }
    for (auto& g : grids) {
        std::vector<int> members;
        for (int i = 0; i < n; ++i)
            if (e->probs[i]->gd.Nx == g[0] && e->probs[i]->gd.Ny == g[1] && e->probs[i]->gd.Nz == g[2]) members.push_back(i);
        cf_problem* h = e->probs[members[0]];
        // z tile width: in-place widths only (one LDS buffer), 1..16 k per tile
        Tiles keep = e->tiles[members[0]];
        float best = 1e30f;
        for (int kt = 1; kt <= std::min(16, h->gd.Nxh); kt *= 2) {
            const int ppt = z_inplace_ppt(h->pz, kt);
            if (!ppt) continue;
            for (int i : members) {
                e->tiles[i].KT = kt;
                e->tiles[i].zppt = ppt;
            }
            float ms;
            if (int rc = ens_time_class(e, B_ZINV, &ms)) return rc;
            if (ms < best) {
                best = ms;
                keep.KT = kt;
                keep.zppt = ppt;
            }
        }
        for (int i : members) {
            e->tiles[i].KT = keep.KT;
            e->tiles[i].zppt = keep.zppt;
        }
        // x wavefronts per block (LDS permitting)
        best = 1e30f;
        int wkeep = e->tiles[members[0]].xwaves;
        for (int w = 1; w <= h->xmaxw[0]; w *= 2) {
            for (int i : members) e->tiles[i].xwaves = w;
            float ms;
            if (int rc = ens_time_class(e, B_X, &ms)) return rc;
            if (ms < best) {
                best = ms;
                wkeep = w;
            }
        }
        for (int i : members) e->tiles[i].xwaves = wkeep;
    }
    // the tables built while timing are replaced by the final build
    ens_free(e);
    return 0;
}

extern "C" int cf_ensemble_create(cf_problem* const* problems, double* const* fields, int n, int tune, cf_ensemble** out) {
    if (!problems || !fields || !out || n < 1) return fail(CF_ERR_ARGUMENT, "need n >= 1 problems and their fields");
    *out = nullptr;
    std::vector<const void*> all;
    for (int i = 0; i < n; ++i) {
        cf_problem* h = problems[i];
        if (!h) return fail(CF_ERR_ARGUMENT, "null problem");
        for (int j = 0; j < i; ++j)
            if (problems[j] == h) return fail(CF_ERR_ARGUMENT, "each instance needs its own problem (workspaces)");
        if (h->device != problems[0]->device) return fail(CF_ERR_ARGUMENT, "ensemble problems must share a device");
        if (h->mixed) return fail(CF_ERR_UNSUPPORTED, "ensembles run the FP64 step");
        if (h->force_fn) return fail(CF_ERR_UNSUPPORTED, "a forced problem cannot join an ensemble (host callback)");
        if (h->closure.active()) return fail(CF_ERR_UNSUPPORTED, "a problem with a closure cannot join an ensemble");
        if (h->nl_fn) return fail(CF_ERR_UNSUPPORTED, "a problem with a nonlinear callback cannot join an ensemble");
        if (h->form == CF_FORM_DIVERGENCE || h->gd.Nxp % 2 != 0)
            return fail(CF_ERR_UNSUPPORTED, "ensembles need the rotating or convective form and an even padded Nx");
        if (y_group_k(h->gd.Ny) != 4) return fail(CF_ERR_UNSUPPORTED, "ensemble members need Ny <= 256");
        for (int f = 0; f < 4; ++f) all.push_back(fields[4 * i + f]);
    }
    if (int rc = check_fields(all.data(), (int)all.size())) return rc;
    DeviceGuard dg(problems[0]->device);
    cf_ensemble* e = new cf_ensemble();
    e->device = problems[0]->device;
    auto bail = [&](int code) {
        if (e->cap) {
            hipGraph_t g = nullptr;
            hipStreamCaptureStatus cs;
            if (hipStreamIsCapturing(e->cap, &cs) == hipSuccess && cs == hipStreamCaptureStatusActive)
                (void)hipStreamEndCapture(e->cap, &g);
            if (g) (void)hipGraphDestroy(g);
        }
        ens_free(e);
        delete e;
        return code;
    };
    for (int i = 0; i < n; ++i) {
        cf_problem* h = problems[i];
        e->probs.push_back(h);
        e->fields.push_back({(double2*)fields[4 * i], (double2*)fields[4 * i + 1], (double2*)fields[4 * i + 2],
                             (double2*)fields[4 * i + 3]});
        e->tiles.push_back(Tiles{h->KT, h->zppt, h->xwaves});
        if (!h->zppt) e->tiles.back().KT = h->KT;  // two-buffer z plans keep the problem's width
    }
    if (hipStreamCreateWithFlags(&e->cap, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(CF_ERR_DEVICE, "hipStreamCreate"));
    if (tune) {
        if (int rc = ens_tune(e)) return bail(rc);
        if (hipStreamCreateWithFlags(&e->cap, hipStreamNonBlocking) != hipSuccess)
            return bail(fail(CF_ERR_DEVICE, "hipStreamCreate"));
    }
    if (int rc = ens_build(e, e->step)) return bail(rc);
    if (hipStreamBeginCapture(e->cap, hipStreamCaptureModeRelaxed) != hipSuccess)
        return bail(fail(CF_ERR_DEVICE, "hipStreamBeginCapture"));
    for (const BatchLaunch& L : e->step) {
        hipError_t err = run_batch(L, e->cap);
        if (err != hipSuccess) return bail(fail(CF_ERR_DEVICE, std::string("ensemble launch: ") + hipGetErrorString(err)));
    }
    if (hipStreamEndCapture(e->cap, &e->graph) != hipSuccess) return bail(fail(CF_ERR_DEVICE, "hipStreamEndCapture"));
    if (hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0) != hipSuccess)
        return bail(fail(CF_ERR_DEVICE, "hipGraphInstantiate"));
    *out = e;
    return CF_OK;
}

extern "C" int cf_ensemble_advance(cf_ensemble* e, int nsteps, void* stream) {
    if (!e) return fail(CF_ERR_ARGUMENT, "null ensemble");
    if (nsteps < 0) return fail(CF_ERR_ARGUMENT, "nsteps must be nonnegative");
    DeviceGuard dg(e->device);
    for (int s = 0; s < nsteps; ++s) HIPTRY(hipGraphLaunch(e->exec, (hipStream_t)stream));
    return CF_OK;
}

extern "C" int cf_ensemble_info(const cf_ensemble* e, int* launches, int* tiles) {
    if (!e) return fail(CF_ERR_ARGUMENT, "null ensemble");
    if (launches) *launches = (int)e->step.size();
    if (tiles)
        for (size_t i = 0; i < e->tiles.size(); ++i) {
            tiles[2 * i] = e->tiles[i].KT;
            tiles[2 * i + 1] = e->tiles[i].xwaves;
        }
    return CF_OK;
}

extern "C" int cf_ensemble_destroy(cf_ensemble* e) {
    if (!e) return CF_OK;
    DeviceGuard dg(e->device);
    if (e->cap) (void)hipStreamSynchronize(e->cap);
    (void)hipDeviceSynchronize();
    ens_free(e);
    delete e;
    return CF_OK;
}
