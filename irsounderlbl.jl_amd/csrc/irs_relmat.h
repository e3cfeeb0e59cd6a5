// Relaxation-matrix line mixing: host data of an uploaded relaxation matrix,
// the per-layer term builder (W(T), Y(T), eigenmodes) and the GPU term sum.
#pragma once

#include <stdint.h>

#include <complex>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/irs.h"

// One line-mixing term on the internal grid: w(z) = w((nu - c) f + i y) enters as
// a Re w + b Im w for |nu - c| <= cutoff.  VP_Y dispersive lines have a = 0;
// VP_W poles a + ib = conj-free amplitude times f/sqrt(pi); their no-coupling
// baseline enters with a = -S(T) f/sqrt(pi), b = 0.  fp64 = 1 for the VP_W terms,
// whose full and baseline sums nearly cancel.
struct LmTerm {
    double c, f, y, a, b;
    int32_t fp64, pad;
};

struct RelmatBandHost {
    int li, lf, isot, molecule;
    int64_t line0, n;                  // lines [line0, line0 + n)
    int table;                         // W0/B0 table index, -1: none
    std::vector<int32_t> row_ptr;      // coupling pairs of line i: [row_ptr[i], row_ptr[i+1])
    std::vector<int32_t> col;
    std::vector<double> w0r, b0r;
};

struct RelmatHost {
    std::vector<RelmatBandHost> bands;
    std::vector<int16_t> Ji;
    std::vector<int8_t> branch;
    std::vector<double> nu, gV_air, E_lower, Dipo0, PopuT0, DipoT;
    std::vector<float> n_air, shift;
};

// Builds the host copy and every band's T-independent coupling table
// (line_mixing.jl:530-578).  Returns an empty string or the error text.
std::string relmat_build(const irs_relmat *rm, RelmatHost &out);

// Terms of one (T, p) layer: per band, mode 1 = VP_Y dispersive lines with Y(T)
// (line_mixing.jl:712-738,1193-1221), mode 2 = VP_W eigenmodes and their
// no-coupling baseline (line_mixing.jl:892-970,1030-1091), 0 = skipped; sorted by
// centre.  Returns an empty string or the error text.
std::string relmat_terms(const RelmatHost &rm, const uint8_t *band_mode, double T, double p_atm,
                         double keep_threshold, std::vector<LmTerm> &out);

// Sum-rule-corrected W (n x n, row-major, original line order) and PopuT of one band.
void relmat_build_W(const RelmatHost &rm, const RelmatBandHost &b, double T, std::vector<double> &W,
                    std::vector<double> &rho);

// General complex eigen-decomposition of an n x n row-major matrix (Householder
// Hessenberg reduction, shifted QR to Schur form, back-substitution): eigenvalues
// and right eigenvectors (columns of V, row-major).  false if QR did not converge.
bool complex_eig(int n, std::vector<std::complex<double>> &A, std::vector<std::complex<double>> &w,
                 std::vector<std::complex<double>> &V);

// Device side: terms of unit u (u = profile * n_layers + layer) are
// terms[off[u] .. off[u+1]), sorted by c.
struct LmTermArgs {
    const LmTerm *terms;
    const int64_t *off;
    int64_t n;                     // internal grid points
    int n_units;
    double an, sn, den;            // grid node i = (an + i sn)/den
    double cutoff, x_far;
    double *out64;                 // mode A: [n_units][n] perturbation
    float *tau;                    // mode B: tau[u][i] += max(tau_lm[u][i] + coef[u] * dsigma, 0)
    const float *tau_lm;
    const double *coef;
};
hipError_t launch_lm_terms(const LmTermArgs &a, hipStream_t s);
