// Host-side helpers of the CNRK2 library (setup of device banks and plans).
#pragma once
#include <string>
#include <vector>

#include "cf_common.h"

namespace cf {

struct HostHelm {
    int P = 0;
    double t0 = 1.0, t1 = 0.0;
    std::vector<double> fac, alpha;  // by degree n (n >= 2)
    void setup(int P, double t0, double t1);
    void solve(double* out, const double* f, double gp, double gm, bool neumann) const;
};

struct HostBank {
    double nu = 0, lambda = 0;
    std::vector<double> facV, mconst, resp;
    double resp_mean = 0;
};

void build_pressure_factors(const GridDims& gd, std::vector<double>& facP);
bool build_bank(const GridDims& gd, double nu, double lambda, HostBank& hb, std::string& err);
void plan_factors(int n, std::vector<int>& radix);
void twiddles(int n, std::vector<double>& tw);
// per-degree constants of the Chebyshev tau rows and the bulk mean, 4 arrays
// of CHEB_TAB_MAX entries: c_{n-2}/(4n(n-1)), -1/(2(n^2-1)), 1/(4n(n+1)) for
// n >= 2 (else 0) and 1/(1-n^2) for even n (else 0)
void build_cheb_table(std::vector<double>& tab);
void host_chebcoeffs(int Ny, const double* v, double* c);
void host_diff(double* b, const double* a, int Ny);
double host_bulkmean(const double* a, int Ny);

}  // namespace cf
