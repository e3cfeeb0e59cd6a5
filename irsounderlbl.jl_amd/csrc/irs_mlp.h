// Internal declarations of the emulator MLP and emulator kernels (k_mlp.hip, k_emu.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GEMM_GELU  1    // C32 <- Z, C16 <- bf16(gelu(Z))
#define GEMM_DGELU 2    // v *= gelu'(Zin[m][n]) before storing

struct GemmArgs {
    const uint16_t *A, *B;       // bf16
    int M, N, K, lda, ldb, ldc;
    float alpha;
    const float *bias;           // [N] or null
    const float *Zin;            // [M][ldc] for GEMM_DGELU
    float *C32;                  // optional f32 output
    uint16_t *C16;               // optional bf16 output
    int flags;
    int splitk;                  // > 1: K in splitk slices, partial products into Cpart [splitk][M][ldc],
    float *Cpart;                // summed in slice order into C32 (no bias / epilogue flags)
};

hipError_t launch_gemm_bf16(const GemmArgs &g, bool ta, bool tb, hipStream_t s);
hipError_t launch_f32_to_bf16(const float *x, uint16_t *y, int64_t n, hipStream_t s);
hipError_t launch_gather_batch(const float *X, const float *T, const int32_t *idx, int B, int n_in, int ld,
                               int n_out, uint16_t *xb, float *tb, hipStream_t s);
hipError_t launch_mse_grad(const float *Y, const float *T, int64_t n, float scale, uint16_t *dY, float *partial,
                           hipStream_t s);
hipError_t launch_colsum_bf16(const uint16_t *dZ, int M, int N, float *db, float *scratch, int64_t scratch_floats,
                              hipStream_t s);
hipError_t launch_adam(float *p, const float *g, float *m, float *v, uint16_t *w16, int64_t n, float lr, float b1,
                       float b2, float eps, int step, hipStream_t s);
hipError_t launch_gelu_tangent(const float *Z, const float *dZ, int64_t n, uint16_t *dH, hipStream_t s);
hipError_t launch_sum_partials(const float *partial, int n, float scale, float *out, hipStream_t s);
hipError_t launch_broadcast_row(const float *row, int B, int N, float *out, hipStream_t s);

// Emulated forward model (k_emu.hip): decode of the three per-layer targets
// (t_k, g_k, m_k) from per-layer PCA coefficients fused with the channel RTE.
struct EmuRteArgs {
    const float *coef;           // [n_prof][L][n_pc] PCA coefficients a_{k,m}
    const float *basis;          // [L][n_pc][3 n_ch]: the layer's basis over (t, g, m)
    const float *mean;           // [L][3 n_ch]
    const float *inv_w;          // [L][3 n_ch] decode weights
    const double *ch_nu;         // [n_ch]
    const double *T_lev, *T_ave; // [n_prof][L+1], [n_prof][L]
    const double *scal;          // [n_prof][3] = {T_sfc, mu, eps}
    int n_prof, L, n_pc, n_ch;
    double *R, *BT;              // [n_prof][n_ch]
    float *t_out;                // optional [n_prof][L][n_ch] decoded level transmittances T_k
};

hipError_t launch_emu_coef(const float *y, const float *phi, const float *A, float scale, int64_t r0, int64_t rows,
                           int L, int n_phi, int n_pc, float *coef, hipStream_t s);
hipError_t launch_emu_rte(const EmuRteArgs &a, hipStream_t s);

// Spot-check correction: BT[q][c] -= sum_s e[q][s] W[s][c], R = B(nu_c, BT) (at most
// EMU_MAX_SPOT spot channels).
#define EMU_MAX_SPOT 1024
hipError_t launch_emu_correct(const float *e, const float *W, int n_spot, int64_t n_ch, int n_prof,
                              const double *ch_nu, double *R, double *BT, hipStream_t s);

// Emulator Jacobian outputs (all multiplied by dBT/dR)
struct EmuAdjArgs {
    double *BT;                  // [n_prof][n_ch]
    float *dt, *dg, *dm;         // dBT/dt_k, dBT/dg_k, dBT/dm_k [n_prof][L][n_ch]
    float *dTave;                // dBT/dT_ave,k (Planck)       [n_prof][L][n_ch]
    float *dTlev;                // dBT/dT_lev,j (Planck)       [n_prof][L+1][n_ch]
    float *dTsfc;                // dBT/dT_sfc                  [n_prof][n_ch]
    const float *Jcoef;          // d a_{k,m} / d feature_i     [n_prof][L][n_pc][n_feat]
    float *dfeat;                // dBT/dfeature                [n_prof][L][n_feat][n_ch]
};

hipError_t launch_emu_adjoint(const EmuRteArgs &a, const EmuAdjArgs &j, int n_feat, hipStream_t s);
