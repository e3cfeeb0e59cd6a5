// Every-pair FP64 line sum: the reference kernel's semantics on the device
// (src/CrossSections/voigt.jl:243-277 and the VP_Y term of
// src/CrossSections/line_mixing.jl:1166-1186, combined as :1351-1369), for the
// "plain FP64" figure beside the FP32 engine.  One thread per (grid point,
// layer) loops over the species of the layer and over every line whose
// unshifted centre lies within the widest window (+ the pressure shift bound),
// tests |nu - nu0| <= cutoff on the shifted FP64 centre, and evaluates
// H = y / (sqrt(pi) (x^2 + y^2)) for |x| > x_far, else Re w(x + iy) (FP64
// Faddeeva, ~1e-14), minus the pedestal (2 - (dnu/c)^2) Hc.  The per-species
// sum is clamped at 0 (twice on the line-mixing CO2 list) and times coef added
// to the FP64 layer optical depth.  Lines of a wave share one index range, so
// the FP64 records arrive as broadcast loads.
#include "irs_devmath.h"
#include "irs_internal.h"

namespace {

__device__ __forceinline__ int lower_key64(const int32_t *k, int n, int64_t v) {
    int a = 0, b = n;
    while (a < b) {
        const int m = (a + b) >> 1;
        if ((int64_t)k[m] < v) a = m + 1; else b = m;
    }
    return a;
}

__global__ void __launch_bounds__(256) linesum64_kernel(const Linesum64Args a) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int k = blockIdx.y;
    const int64_t ib = (int64_t)blockIdx.x * 256;                    // first point of the block
    const int64_t ie = min(ib + 255, a.n_hi - 1);
    const double nu = (a.an + (double)i * a.sn) / a.den;
    double tau = 0.0;
    for (int s = 0; s < a.n_species; ++s) {
        const ItemDev it = a.items[(int64_t)k * a.n_species + s];
        if (!it.active || it.n == 0) continue;
        const bool lm = it.lm != 0;
        const double Cw = lm ? fmax(it.cutoff, it.lm_cutoff) : it.cutoff;
        const int64_t W = (int64_t)ceil(Cw / a.dnu) + it.M + 1;
        const int32_t *keys = a.keys + it.key_off;
        // one line range per block (block-uniform): every line that can reach the block
        const int j0 = lower_key64(keys, it.n, ib - W), j1 = lower_key64(keys, it.n, ie + W + 1);
        const double *r = a.rec64 + (it.rec_off) * 8;
        double sv = 0.0, sd = 0.0;
        if (i < a.n_hi) {
            const double ic2 = 1.0 / (it.cutoff * it.cutoff);
            for (int j = j0; j < j1; ++j) {
                const double *q = r + (int64_t)j * 8;     // {c, f, y, Snorm (0: rejected), Hc, YSn, fd, yd}
                const double d = nu - q[0];
                const double ad = fabs(d);
                if (ad <= it.cutoff && q[3] != 0.0) {
                    const double x = d * q[1], y = q[2];
                    double H;
                    if (fabs(x) > it.x_far) H = y * IRS_INV_SQRT_PI / (x * x + y * y);
                    else { double wi; w64(x, y, H, wi); }
                    sv += q[3] * (H - (2.0 - d * d * ic2) * q[4]);
                }
                if (lm && q[5] != 0.0 && ad <= it.lm_cutoff) {
                    const double x = d * q[6], y = q[7];
                    double im;
                    if (fabs(x) > it.x_far) im = x * IRS_INV_SQRT_PI / (x * x + y * y);
                    else { double wr; w64(x, y, wr, im); }
                    sd += q[5] * im;
                }
            }
        }
        double sigma = fmax(sv, 0.0);                    // voigt.jl:276
        if (lm) sigma = fmax(sigma + sd, 0.0);           // line_mixing.jl:1368
        tau += sigma * it.coef;
    }
    if (i < a.n_hi) a.tau[(int64_t)k * a.n_hi + i] = tau;
}

}  // namespace

hipError_t launch_linesum64(const Linesum64Args &a, hipStream_t s) {
    if (a.n_hi <= 0 || a.n_layers <= 0) return hipSuccess;
    hipLaunchKernelGGL(linesum64_kernel, dim3((unsigned)((a.n_hi + 255) / 256), (unsigned)a.n_layers), dim3(256), 0,
                       s, a);
    return hipGetLastError();
}
