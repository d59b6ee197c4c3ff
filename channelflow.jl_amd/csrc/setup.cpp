// Host-side setup of the device solver banks (FP64): Chebyshev-tau
// Helmholtz elimination factors, influence/tau constants per Fourier mode,
// mean-mode responses, FFT plans and twiddle tables.
//
// Follows the constructors of the reference: BatchedInfluenceSolver
// (src/solvers/batchedinfluence.jl:45-98), MeanModeSolver
// (src/solvers/meanmode.jl:38-59), CNRK2 (src/timesteppers/cnrk2.jl:58-83).
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "cf_host.h"

namespace cf {

// --------------------------------------------------------------------------
// Chebyshev-tau Helmholtz (t0 D^2 - t1) a = f, rows 0..P-2 + two wall rows,
// via the quasi-tridiagonal even/odd elimination that the device kernel
// also uses (same factors s_n, indexed by degree n >= 2).
// --------------------------------------------------------------------------
static double cfac(int m) { return m == 0 ? 2.0 : 1.0; }

void HostHelm::setup(int P_, double t0_, double t1_) {
    P = P_;
    t0 = t0_;
    t1 = t1_;
    fac.assign(P + 1, 0.0);
    alpha.assign(P + 3, 0.0);
    for (int p = 0; p < 2; ++p) {
        int J = (P - p) / 2;
        double anext = 0.0;  // alpha_{J+1}
        for (int j = J; j >= 1; --j) {
            int n = 2 * j + p;
            double nn = n;
            double en = (n <= P - 2) ? 1.0 : 0.0, en2 = (n + 2 <= P - 2) ? 1.0 : 0.0;
            double lo = -t1 * cfac(n - 2) / (4.0 * nn * (nn - 1.0));
            double d = t0 + t1 * en / (2.0 * (nn * nn - 1.0));
            double up = -t1 * en2 / (4.0 * nn * (nn + 1.0));
            double s = 1.0 / (d + up * anext);
            fac[n] = s;
            alpha[n] = -lo * s;
            anext = alpha[n];
        }
    }
}

void HostHelm::solve(double* out, const double* f, double gp, double gm, bool neumann) const {
    auto F = [&](int m) { return (m >= 0 && m <= P - 2) ? f[m] : 0.0; };
    for (int p = 0; p < 2; ++p) {
        int J = (P - p) / 2;
        std::vector<double> beta(J + 2, 0.0), delta(J + 1, 0.0), gam(J + 1, 0.0);
        for (int j = J; j >= 1; --j) {
            int n = 2 * j + p;
            double nn = n;
            double en2 = (n + 2 <= P - 2) ? 1.0 : 0.0;
            double up = -t1 * en2 / (4.0 * nn * (nn + 1.0));
            double rhs = cfac(n - 2) * F(n - 2) / (4.0 * nn * (nn - 1.0)) - F(n) / (2.0 * (nn * nn - 1.0)) +
                         F(n + 2) / (4.0 * nn * (nn + 1.0));
            beta[j] = fac[n] * (rhs - up * beta[j + 1]);
        }
        gam[0] = 1.0;
        for (int j = 1; j <= J; ++j) {
            int n = 2 * j + p;
            gam[j] = alpha[n] * gam[j - 1];
            delta[j] = alpha[n] * delta[j - 1] + beta[j];
        }
        double g;
        if (!neumann) g = (p == 0) ? 0.5 * (gp + gm) : 0.5 * (gp - gm);
        else g = (p == 0) ? 0.5 * (gp - gm) : 0.5 * (gp + gm);
        double swd = 0.0, swg = 0.0;
        for (int j = 0; j <= J; ++j) {
            int n = 2 * j + p;
            double w = neumann ? (double)n * n : 1.0;
            swd += w * delta[j];
            swg += w * gam[j];
        }
        double x0 = (g - swd) / swg;
        for (int j = 0; j <= J; ++j) out[2 * j + p] = gam[j] * x0 + delta[j];
    }
}

// derivative coefficients (backward recurrence, b_0 halved)
static void hdiff(double* b, const double* a, int Ny) {
    b[Ny - 1] = 0.0;
    if (Ny >= 2) b[Ny - 2] = 2.0 * (Ny - 1) * a[Ny - 1];
    for (int n = Ny - 3; n >= 0; --n) b[n] = b[n + 2] + 2.0 * (n + 1) * a[n + 1];
    b[0] *= 0.5;
}
static double hwall_dplus(const double* a, int Ny) {
    double s = 0.0;
    for (int n = 0; n < Ny; ++n) s += (double)n * n * a[n];
    return s;
}
static double hwall_dminus(const double* a, int Ny) {
    double s = 0.0;
    for (int n = 0; n < Ny; ++n) s += ((n % 2) ? 1.0 : -1.0) * (double)n * n * a[n];
    return s;
}
double host_bulkmean(const double* a, int Ny) {
    double v = a[0];
    for (int n = 2; n <= Ny - 1; n += 2) v += a[n] / (1.0 - (double)n * n);
    return v;
}

// --------------------------------------------------------------------------
// pressure factors (stage independent): D^2 - kappa^2 (kappa^2 := 1 at (0,0))
// --------------------------------------------------------------------------
void build_pressure_factors(const GridDims& gd, std::vector<double>& facP) {
    int Ny = gd.Ny, P = gd.P;
    facP.assign((size_t)gd.B * Ny, 0.0);
    HostHelm h;
    for (long long s = 0; s < gd.B; ++s) {
        int ix = (int)(s % gd.Nxh), iz = (int)(s / gd.Nxh);
        int l = iz <= (gd.Nz >> 1) ? iz : iz - gd.Nz;
        double k = gd.alpha * ix, ll = gd.beta * l;
        double kap2 = k * k + ll * ll;
        h.setup(P, 1.0, kap2 == 0.0 ? 1.0 : kap2);
        std::memcpy(&facP[(size_t)s * Ny], h.fac.data(), sizeof(double) * Ny);
    }
}

// --------------------------------------------------------------------------
// A bank: one (nu, lambda) Stokes solver for all modes.
// Influence responses and auxiliary tau problem as in the reference
// constructor; only the scalars the device re-solve needs are kept.
// --------------------------------------------------------------------------
bool build_bank(const GridDims& gd, double nu, double lambda, HostBank& hb, std::string& err) {
    int Ny = gd.Ny, P = gd.P;
    hb.nu = nu;
    hb.lambda = lambda;
    hb.facV.assign((size_t)gd.B * Ny, 0.0);
    hb.mconst.assign((size_t)gd.B * MC_COUNT, 0.0);
    std::vector<double> zero(Ny, 0.0), work(Ny), dw(Ny), pp(Ny), pm(Ny), vp(Ny), vm(Ny), p0(Ny), v0(Ny);
    HostHelm hp, hv;
    for (long long s = 0; s < gd.B; ++s) {
        int ix = (int)(s % gd.Nxh), iz = (int)(s / gd.Nxh);
        int l = iz <= (gd.Nz >> 1) ? iz : iz - gd.Nz;
        double k = gd.alpha * ix, ll = gd.beta * l;
        double kap2 = k * k + ll * ll;
        double shift = lambda + nu * kap2;
        hp.setup(P, 1.0, kap2 == 0.0 ? 1.0 : kap2);
        hv.setup(P, nu, shift);
        std::memcpy(&hb.facV[(size_t)s * Ny], hv.fac.data(), sizeof(double) * Ny);
        // homogeneous responses (batchedinfluence.jl:64-74)
        hp.solve(pp.data(), zero.data(), 1.0, 0.0, false);
        hdiff(work.data(), pp.data(), Ny);
        hv.solve(vp.data(), work.data(), 0.0, 0.0, false);
        hp.solve(pm.data(), zero.data(), 0.0, 1.0, false);
        hdiff(work.data(), pm.data(), Ny);
        hv.solve(vm.data(), work.data(), 0.0, 0.0, false);
        // influence matrix inverse (76-88)
        double A = hwall_dplus(vp.data(), Ny), Bq = hwall_dplus(vm.data(), Ny);
        double C = hwall_dminus(vp.data(), Ny), D = hwall_dminus(vm.data(), Ny);
        double scale = std::fmax(std::fmax(std::fabs(A), std::fabs(Bq)), std::fmax(std::fabs(C), std::fabs(D)));
        A /= scale; Bq /= scale; C /= scale; D /= scale;
        double det = A * D - Bq * C;
        bool nonzero = !(kap2 == 0.0);
        if (nonzero && !(std::isfinite(det) && std::fabs(det) > 2.220446049250313e-16 * std::fmax(std::fabs(A * D), std::fabs(Bq * C)))) {
            err = "singular influence matrix for system " + std::to_string(s + 1);
            return false;
        }
        double* mc = &hb.mconst[(size_t)s * MC_COUNT];
        if (!nonzero) continue;  // the mean slot is solved by the mean-mode path
        mc[MC_A11] = D / (det * scale);
        mc[MC_A12] = -Bq / (det * scale);
        mc[MC_A21] = -C / (det * scale);
        mc[MC_A22] = A / (det * scale);
        // auxiliary tau problem: source (T_{P-1} + T_P)' (93-105)
        std::fill(work.begin(), work.end(), 0.0);
        work[P - 1] = work[P] = 1.0;
        hdiff(dw.data(), work.data(), Ny);
        hp.solve(p0.data(), dw.data(), 0.0, 0.0, false);
        hdiff(dw.data(), p0.data(), Ny);
        hv.solve(v0.data(), dw.data(), 0.0, 0.0, false);
        double upper = -hwall_dplus(v0.data(), Ny), lower = -hwall_dminus(v0.data(), Ny);
        double d0p = mc[MC_A11] * upper + mc[MC_A12] * lower;
        double d0m = mc[MC_A21] * upper + mc[MC_A22] * lower;
        for (int n = 0; n < Ny; ++n) {
            p0[n] += d0p * pp[n] + d0m * pm[n];
            v0[n] += d0p * vp[n] + d0m * vm[n];
        }
        hdiff(dw.data(), p0.data(), Ny);
        mc[MC_SIG1] = shift * v0[P - 1] + dw[P - 1];
        mc[MC_SIG2] = shift * v0[P];
        mc[MC_D0P] = d0p;
        mc[MC_D0M] = d0m;
        mc[MC_VP1] = vp[P - 1];
        mc[MC_VP0] = vp[P];
        mc[MC_VM1] = vm[P - 1];
        mc[MC_VM0] = vm[P];
        mc[MC_PPT] = pp[P];
        mc[MC_PMT] = pm[P];
    }
    // mean-mode unit-gradient response c: (nu D^2 - lambda) c = 1, c(+-1) = 0
    hb.resp.assign(Ny, 0.0);
    {
        HostHelm h;
        h.setup(P, nu, lambda);
        std::vector<double> e0(Ny, 0.0);
        e0[0] = 1.0;
        h.solve(hb.resp.data(), e0.data(), 0.0, 0.0, false);
        hb.resp_mean = host_bulkmean(hb.resp.data(), Ny);
        if (!(std::isfinite(hb.resp_mean) && hb.resp_mean != 0.0)) {
            err = "the uniform-gradient response has no finite nonzero bulk mean";
            return false;
        }
    }
    return true;
}

// --------------------------------------------------------------------------
// FFT plans
// --------------------------------------------------------------------------
// Pass order: odd radices first, then 2, 4, 8.  A Stockham pass of radix R
// after spans Ns writes runs of Ns consecutive points Ns*R apart; the first
// pass (Ns = 1) thus writes at stride R, which for R = 8 puts eight lanes of
// a 16-lane LDS access on one bank group (16-byte points, 64 banks).  An odd
// first radix writes at a stride coprime to the bank count.
void plan_factors(int n, std::vector<int>& radix) {
    radix.clear();
    int m = n;
    int c8 = 0, c4 = 0, c2 = 0;
    while (m % 8 == 0 && m > 8) { ++c8; m /= 8; }
    while (m % 4 == 0) { ++c4; m /= 4; }
    while (m % 2 == 0) { ++c2; m /= 2; }
    while (m % 3 == 0) { radix.push_back(3); m /= 3; }
    while (m % 5 == 0) { radix.push_back(5); m /= 5; }
    for (int p = 7; p * p <= m; p += 2)
        while (m % p == 0) { radix.push_back(p); m /= p; }
    if (m > 1) radix.push_back(m);
    radix.insert(radix.end(), c2, 2);
    radix.insert(radix.end(), c4, 4);
    radix.insert(radix.end(), c8, 8);
    if (radix.empty()) radix.push_back(1);
}

void build_cheb_table(std::vector<double>& tab) {
    const int nt = CHEB_TAB_MAX;
    tab.assign((size_t)CHEB_TAB_ARRAYS * nt, 0.0);
    for (int n = 0; n < nt; ++n) {
        const double nn = n;
        if (n >= 2) {
            tab[n] = (n == 2 ? 2.0 : 1.0) / (4.0 * nn * (nn - 1.0));
            tab[nt + n] = -1.0 / (2.0 * (nn * nn - 1.0));
            tab[2 * nt + n] = 1.0 / (4.0 * nn * (nn + 1.0));
        }
        if (n % 2 == 0) tab[3 * nt + n] = 1.0 / (1.0 - nn * nn);
    }
}

void twiddles(int n, std::vector<double>& tw) {
    tw.resize(2 * (size_t)std::max(n, 1));
    const long double PI = 3.141592653589793238462643383279502884L;
    for (int t = 0; t < n; ++t) {
        long double a = -2.0L * PI * (long double)t / (long double)n;
        tw[2 * t] = (double)cosl(a);
        tw[2 * t + 1] = (double)sinl(a);
    }
}

// Lobatto Chebyshev coefficients of values at y_j (DCT-I / P, ends halved)
void host_chebcoeffs(int Ny, const double* v, double* c) {
    int P = Ny - 1;
    const long double PI = 3.141592653589793238462643383279502884L;
    for (int k = 0; k <= P; ++k) {
        long double acc = v[0] + ((k % 2) ? -v[P] : v[P]);
        for (int j = 1; j < P; ++j) acc += 2.0L * v[j] * cosl(PI * (long double)j * k / P);
        c[k] = (double)(acc / P);
    }
    c[0] *= 0.5;
    c[P] *= 0.5;
}

void host_diff(double* b, const double* a, int Ny) { hdiff(b, a, Ny); }

}  // namespace cf
