// C ABI of the emulator: the device-resident MLP (BF16 MFMA training and
// inference) and the emulated forward model (PCA decode + channel RTE).
// Neither has a counterpart in the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/irs.h"
#include "irs_mlp.h"

namespace {

template <typename T>
hipError_t dalloc(T **p, size_t n) {
    *p = nullptr;
    return hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
}

}  // namespace

struct irs_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int n_in = 0, n_in_pad = 64, H = 512, n_out = 0, max_batch = 0;
    int dims[5] = {};
    int64_t off_W[4] = {}, off_b[4] = {}, n_params = 0;
    float *P = nullptr, *G = nullptr, *Mo = nullptr, *Vo = nullptr;
    uint16_t *P16 = nullptr;
    uint16_t *X16 = nullptr, *H16[3] = {}, *dY16 = nullptr, *dZ16[3] = {};
    float *Z[3] = {}, *Y = nullptr, *T = nullptr, *partial = nullptr, *loss = nullptr, *tang = nullptr;
    float *Xin = nullptr;            // staging for host inputs [max_batch][n_in]
    float *Xall = nullptr, *Tall = nullptr;
    int64_t n_data = 0;
    int32_t *idx = nullptr;
    int64_t idx_cap = 0;
    int step = 0;
    float *Gpart = nullptr;          // split-K partial products of the weight gradients
};

// Split-K slices of a weight-gradient GEMM (few 64 x 64 tiles, K = batch): enough
// slices that tiles x slices reaches 256 workgroups, each slice at least 128 deep.
#define IRS_GPART_FLOATS (2 * 256 * 64 * 64)
static int weight_grad_splits(int M, int N, int K) {
    const int tiles = (M / 64) * (N / 64);
    int S = 1;
    while (tiles * S < 256 && K % (2 * S) == 0 && (K / (2 * S)) % 32 == 0 && K / (2 * S) >= 128 &&
           (int64_t)2 * S * M * N <= IRS_GPART_FLOATS)
        S *= 2;
    return S;
}

struct irs_emu {
    irs_mlp *mlp = nullptr;
    int L = 0, n_pc = 0, n_ch = 0, n_feat = 0, n_local = 0, n_phi = 0;
    std::vector<float> x_mean, x_std;  // [L][n_feat]
    std::vector<float> A_host;         // [L][n_phi][n_pc] regression of the coefficients on phi(z)
    float coef_scale = 1.0f;           // a = A phi + coef_scale MLP(z)
    float *A = nullptr;                // device copy of A_host
    float *basis = nullptr;            // [L][n_pc][3 n_ch]
    float *mean = nullptr;             // [L][3 n_ch]
    float *inv_w = nullptr;            // [L][3 n_ch]
    double *ch_nu = nullptr;           // [n_ch]
    float *coef = nullptr, *phi = nullptr;
    double *prof = nullptr;            // T_lev, T_ave, scal
    double *R = nullptr, *BT = nullptr;
    int64_t cap_rows = 0, cap_phi = 0, cap_prof = 0, cap_out = 0;
    std::string err;
    // profile-context features (irs_emu_set_context): proj [n_ctx][n_vec] applied to
    // [T levels, ln q_H2O levels, ln q_O3 levels, T_sfc, 1/mu] - mean
    int n_ctx = 0, n_vec = 0;
    std::vector<double> ctx_mean, ctx_proj;
    // spot-check correction (irs_emu_set_correction): spot blocks of corr_bn[b] channels,
    // W [n_spot][n_ch] on the device, the leave-one-block-out maps on the host
    std::vector<int> corr_bn;
    std::vector<int64_t> corr_c0;
    int corr_ns = 0;
    double corr_w_raw = 1.0, corr_w_lbo = 0.0;   // fallback statistic max(w_raw max|e|, w_lbo residual)
    float *corr_W = nullptr, *corr_e = nullptr;
    int64_t cap_corr_e = 0;
    std::vector<float> corr_lbo;
};

namespace {

#define MCHK(expr)                                                                          \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            m->err = std::string(#expr) + ": " + hipGetErrorString(e_);                     \
            return IRS_EHIP;                                                                \
        }                                                                                   \
    } while (0)

int mfail(irs_mlp *m, int code, const char *msg) {
    m->err = msg;
    return code;
}

// Forward pass of B rows already staged in X16; keeps Z (f32) and H (bf16) of
// the hidden layers; output Y (f32).
int mlp_forward_staged(irs_mlp *m, int B) {
    const uint16_t *in = m->X16;
    int ld_in = m->n_in_pad;
    for (int l = 0; l < 4; ++l) {
        GemmArgs g{};
        g.A = in;
        g.B = m->P16 + m->off_W[l];
        g.M = B;
        g.N = m->dims[l + 1];
        g.K = m->dims[l];
        g.lda = ld_in;
        g.ldb = m->dims[l + 1];
        g.ldc = m->dims[l + 1];
        g.alpha = 1.0f;
        g.bias = m->P + m->off_b[l];
        if (l < 3) {
            g.C32 = m->Z[l];
            g.C16 = m->H16[l];
            g.flags = GEMM_GELU;
            in = m->H16[l];
            ld_in = m->dims[l + 1];
        } else {
            g.C32 = m->Y;
        }
        MCHK(launch_gemm_bf16(g, false, false, m->stream));
    }
    return IRS_OK;
}

// Backward of the MSE loss for B rows (targets in T); gradient into G.
int mlp_backward(irs_mlp *m, int B) {
    const int nout = m->n_out;
    const int nblk = (int)(((int64_t)B * nout + 255) / 256);
    MCHK(launch_mse_grad(m->Y, m->T, (int64_t)B * nout, 2.0f / ((float)B * nout), m->dY16, m->partial,
                         m->stream));
    MCHK(launch_sum_partials(m->partial, nblk, 1.0f / ((float)B * nout), m->loss, m->stream));
    const uint16_t *dcur = m->dY16;            // gradient wrt the output of layer l (bf16)
    for (int l = 3; l >= 0; --l) {
        const uint16_t *act = l == 0 ? m->X16 : m->H16[l - 1];
        const int din = m->dims[l], dout = m->dims[l + 1];
        // dW_l = act^T . dcur      [din x dout]
        GemmArgs gw{};
        gw.A = act;
        gw.B = dcur;
        gw.M = din;
        gw.N = dout;
        gw.K = B;
        gw.lda = din;
        gw.ldb = dout;
        gw.ldc = dout;
        gw.alpha = 1.0f;
        gw.C32 = m->G + m->off_W[l];
        gw.splitk = weight_grad_splits(din, dout, B);
        gw.Cpart = m->Gpart;
        MCHK(launch_gemm_bf16(gw, true, false, m->stream));
        MCHK(launch_colsum_bf16(dcur, B, dout, m->G + m->off_b[l], m->Gpart, IRS_GPART_FLOATS, m->stream));
        if (l > 0) {
            // dZ_{l-1} = (dcur . W_l^T) * gelu'(Z_{l-1})     [B x din]
            GemmArgs gx{};
            gx.A = dcur;
            gx.B = m->P16 + m->off_W[l];
            gx.M = B;
            gx.N = din;
            gx.K = dout;
            gx.lda = dout;
            gx.ldb = dout;
            gx.ldc = din;
            gx.alpha = 1.0f;
            gx.Zin = m->Z[l - 1];
            gx.C16 = m->dZ16[l - 1];
            gx.flags = GEMM_DGELU;
            MCHK(launch_gemm_bf16(gx, false, true, m->stream));
            dcur = m->dZ16[l - 1];
        }
    }
    return IRS_OK;
}

int stage_inputs(irs_mlp *m, const float *x, int64_t rows, int B) {
    // host rows -> Xin (f32) -> X16 (bf16, padded); rows beyond `rows` are zero
    MCHK(hipMemsetAsync(m->Xin, 0, sizeof(float) * (size_t)B * m->n_in, m->stream));
    MCHK(hipMemcpyAsync(m->Xin, x, sizeof(float) * (size_t)rows * m->n_in, hipMemcpyHostToDevice, m->stream));
    MCHK(launch_gather_batch(m->Xin, nullptr, nullptr, B, m->n_in, m->n_in_pad, 0, m->X16, nullptr, m->stream));
    return IRS_OK;
}

}  // namespace

extern "C" {

int irs_mlp_create(int device, int n_in, int n_hidden, int n_out, int max_batch, uint64_t seed, irs_mlp **out) {
    if (!out) return IRS_EINVAL;
    *out = nullptr;
    if (n_in < 1 || n_in > 64 || n_hidden < 64 || n_hidden % 64 || n_out < 64 || n_out % 64 || max_batch < 64 ||
        max_batch % 64)
        return IRS_EINVAL;
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || device < 0 || device >= nd) return IRS_ENODEV;
    if (hipSetDevice(device) != hipSuccess) return IRS_ENODEV;
    irs_mlp *m = new irs_mlp();
    m->device = device;
    m->n_in = n_in;
    m->H = n_hidden;
    m->n_out = n_out;
    m->max_batch = max_batch;
    const int d[5] = {m->n_in_pad, n_hidden, n_hidden, n_hidden, n_out};
    std::memcpy(m->dims, d, sizeof d);
    int64_t o = 0;
    for (int l = 0; l < 4; ++l) {
        m->off_W[l] = o;
        o += (int64_t)d[l] * d[l + 1];
        m->off_b[l] = o;
        o += d[l + 1];
    }
    m->n_params = o;
    bool ok = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && dalloc(&m->P, o) == hipSuccess && dalloc(&m->G, o) == hipSuccess && dalloc(&m->Mo, o) == hipSuccess &&
         dalloc(&m->Vo, o) == hipSuccess && dalloc(&m->P16, o) == hipSuccess;
    const size_t B = max_batch;
    ok = ok && dalloc(&m->X16, B * m->n_in_pad) == hipSuccess && dalloc(&m->Xin, B * n_in) == hipSuccess &&
         dalloc(&m->dY16, B * n_out) == hipSuccess && dalloc(&m->Y, B * n_out) == hipSuccess &&
         dalloc(&m->T, B * n_out) == hipSuccess && dalloc(&m->partial, (B * n_out + 255) / 256) == hipSuccess &&
         dalloc(&m->loss, 1) == hipSuccess && dalloc(&m->tang, B * n_hidden) == hipSuccess &&
         dalloc(&m->Gpart, IRS_GPART_FLOATS) == hipSuccess;
    for (int l = 0; l < 3 && ok; ++l)
        ok = dalloc(&m->H16[l], B * n_hidden) == hipSuccess && dalloc(&m->dZ16[l], B * n_hidden) == hipSuccess &&
             dalloc(&m->Z[l], B * n_hidden) == hipSuccess;
    if (!ok) {
        irs_mlp_destroy(m);
        return IRS_ENOMEM;
    }
    // He-uniform weights (GELU), zero biases; padded input rows of W1 stay zero
    std::vector<float> p(o, 0.0f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < 4; ++l) {
        const int fan_in = l == 0 ? n_in : d[l];
        const float lim = std::sqrt(6.0f / (float)fan_in) * (l == 3 ? 0.5f : 1.0f);
        std::uniform_real_distribution<float> U(-lim, lim);
        for (int r = 0; r < d[l]; ++r)
            for (int c = 0; c < d[l + 1]; ++c)
                p[m->off_W[l] + (int64_t)r * d[l + 1] + c] = (l == 0 && r >= n_in) ? 0.0f : U(rng);
    }
    if (irs_mlp_set_params(m, p.data()) != IRS_OK) {
        irs_mlp_destroy(m);
        return IRS_EHIP;
    }
    *out = m;
    return IRS_OK;
}

void irs_mlp_destroy(irs_mlp *m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    void *ptrs[] = {m->P, m->G, m->Mo, m->Vo, m->P16, m->X16, m->dY16, m->Y, m->T, m->partial, m->loss, m->tang,
                    m->Xin, m->Xall, m->Tall, m->idx, m->H16[0], m->H16[1], m->H16[2], m->dZ16[0], m->dZ16[1],
                    m->dZ16[2], m->Z[0], m->Z[1], m->Z[2], m->Gpart};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

const char *irs_mlp_last_error(const irs_mlp *m) { return m ? m->err.c_str() : "null mlp"; }

int64_t irs_mlp_n_params(const irs_mlp *m) { return m ? m->n_params : -1; }

int irs_mlp_input_pad(const irs_mlp *m) { return m ? m->n_in_pad : -1; }

int irs_mlp_get_params(irs_mlp *m, float *p) {
    if (!m || !p) return IRS_EINVAL;
    (void)hipSetDevice(m->device);
    MCHK(hipMemcpyAsync(p, m->P, sizeof(float) * m->n_params, hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    return IRS_OK;
}

int irs_mlp_set_params(irs_mlp *m, const float *p) {
    if (!m || !p) return IRS_EINVAL;
    (void)hipSetDevice(m->device);
    MCHK(hipMemcpyAsync(m->P, p, sizeof(float) * m->n_params, hipMemcpyHostToDevice, m->stream));
    MCHK(hipMemsetAsync(m->Mo, 0, sizeof(float) * m->n_params, m->stream));
    MCHK(hipMemsetAsync(m->Vo, 0, sizeof(float) * m->n_params, m->stream));
    MCHK(launch_f32_to_bf16(m->P, m->P16, m->n_params, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    m->step = 0;
    return IRS_OK;
}

int irs_mlp_forward(irs_mlp *m, const float *x, int64_t n, float *y) {
    if (!m || !x || !y || n < 0) return IRS_EINVAL;
    (void)hipSetDevice(m->device);
    for (int64_t r0 = 0; r0 < n; r0 += m->max_batch) {
        const int64_t rows = std::min<int64_t>(m->max_batch, n - r0);
        const int B = (int)((rows + 63) / 64 * 64);
        int rc = stage_inputs(m, x + r0 * m->n_in, rows, B);
        if (rc) return rc;
        rc = mlp_forward_staged(m, B);
        if (rc) return rc;
        MCHK(hipMemcpyAsync(y + r0 * m->n_out, m->Y, sizeof(float) * rows * m->n_out, hipMemcpyDeviceToHost,
                            m->stream));
    }
    MCHK(hipStreamSynchronize(m->stream));
    return IRS_OK;
}

int irs_mlp_loss_grad(irs_mlp *m, const float *x, const float *t, int B, float *grad, float *loss) {
    if (!m || !x || !t || !grad || !loss) return IRS_EINVAL;
    if (B < 64 || B % 64 || B > m->max_batch) return mfail(m, IRS_EINVAL, "batch must be a multiple of 64 <= max_batch");
    (void)hipSetDevice(m->device);
    int rc = stage_inputs(m, x, B, B);
    if (rc) return rc;
    MCHK(hipMemcpyAsync(m->T, t, sizeof(float) * (size_t)B * m->n_out, hipMemcpyHostToDevice, m->stream));
    rc = mlp_forward_staged(m, B);
    if (rc) return rc;
    rc = mlp_backward(m, B);
    if (rc) return rc;
    MCHK(hipMemcpyAsync(grad, m->G, sizeof(float) * m->n_params, hipMemcpyDeviceToHost, m->stream));
    MCHK(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    return IRS_OK;
}

int irs_mlp_set_data(irs_mlp *m, const float *X, const float *T, int64_t n) {
    if (!m || !X || !T || n < 1) return IRS_EINVAL;
    (void)hipSetDevice(m->device);
    if (m->Xall) (void)hipFree(m->Xall);
    if (m->Tall) (void)hipFree(m->Tall);
    m->Xall = m->Tall = nullptr;
    m->n_data = 0;
    MCHK(dalloc(&m->Xall, (size_t)n * m->n_in));
    MCHK(dalloc(&m->Tall, (size_t)n * m->n_out));
    MCHK(hipMemcpyAsync(m->Xall, X, sizeof(float) * (size_t)n * m->n_in, hipMemcpyHostToDevice, m->stream));
    MCHK(hipMemcpyAsync(m->Tall, T, sizeof(float) * (size_t)n * m->n_out, hipMemcpyHostToDevice, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    m->n_data = n;
    return IRS_OK;
}

int irs_mlp_train(irs_mlp *m, int steps, int batch, float lr, uint64_t seed, float *losses) {
    if (!m || steps < 0) return IRS_EINVAL;
    if (batch < 64 || batch % 64 || batch > m->max_batch) return mfail(m, IRS_EINVAL, "batch must be a multiple of 64 <= max_batch");
    if (m->n_data < 1) return mfail(m, IRS_EINVAL, "no training data (irs_mlp_set_data)");
    (void)hipSetDevice(m->device);
    // minibatch indices: shuffled epochs, drawn on the host once per call
    const int64_t need = (int64_t)steps * batch;
    if (need > m->idx_cap) {
        if (m->idx) (void)hipFree(m->idx);
        m->idx = nullptr;
        m->idx_cap = 0;
        MCHK(dalloc(&m->idx, (size_t)need));
        m->idx_cap = need;
    }
    std::vector<int32_t> idx((size_t)need);
    std::vector<int32_t> perm((size_t)m->n_data);
    for (int64_t i = 0; i < m->n_data; ++i) perm[i] = (int32_t)i;
    std::mt19937_64 rng(seed);
    int64_t pos = m->n_data;
    for (int64_t i = 0; i < need; ++i) {
        if (pos == m->n_data) {
            std::shuffle(perm.begin(), perm.end(), rng);
            pos = 0;
        }
        idx[i] = perm[pos++];
    }
    MCHK(hipMemcpyAsync(m->idx, idx.data(), sizeof(int32_t) * need, hipMemcpyHostToDevice, m->stream));
    std::vector<float> lh(steps, 0.0f);
    float *dloss = nullptr;
    MCHK(dalloc(&dloss, (size_t)std::max(steps, 1)));
    int rc = IRS_OK;
    for (int s = 0; s < steps && rc == IRS_OK; ++s) {
        hipError_t e = launch_gather_batch(m->Xall, m->Tall, m->idx + (int64_t)s * batch, batch, m->n_in,
                                           m->n_in_pad, m->n_out, m->X16, m->T, m->stream);
        if (e != hipSuccess) { rc = mfail(m, IRS_EHIP, hipGetErrorString(e)); break; }
        rc = mlp_forward_staged(m, batch);
        if (rc) break;
        rc = mlp_backward(m, batch);
        if (rc) break;
        m->step += 1;
        e = launch_adam(m->P, m->G, m->Mo, m->Vo, m->P16, m->n_params, lr, 0.9f, 0.999f, 1e-8f, m->step, m->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dloss + s, m->loss, sizeof(float), hipMemcpyDeviceToDevice, m->stream);
        if (e != hipSuccess) { rc = mfail(m, IRS_EHIP, hipGetErrorString(e)); break; }
    }
    if (rc == IRS_OK && steps > 0) {
        hipError_t e = hipMemcpyAsync(lh.data(), dloss, sizeof(float) * steps, hipMemcpyDeviceToHost, m->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
        if (e != hipSuccess) rc = mfail(m, IRS_EHIP, hipGetErrorString(e));
    } else {
        (void)hipStreamSynchronize(m->stream);
    }
    (void)hipFree(dloss);
    if (rc == IRS_OK && losses)
        for (int s = 0; s < steps; ++s) losses[s] = lh[s];
    return rc;
}

// dy/dx for every row: J[r][o][i], forward-mode tangents through the network
// (bf16 tangents, f32 accumulation).
int irs_mlp_input_jacobian(irs_mlp *m, const float *x, int64_t n, float *J) {
    if (!m || !x || !J || n < 0) return IRS_EINVAL;
    (void)hipSetDevice(m->device);
    std::vector<float> w1((size_t)m->n_in * m->H);
    MCHK(hipMemcpyAsync(w1.data(), m->P + m->off_W[0], sizeof(float) * w1.size(), hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    float *row = nullptr;
    MCHK(dalloc(&row, (size_t)m->H));
    std::vector<float> dy((size_t)m->max_batch * m->n_out);
    int rc = IRS_OK;
    for (int64_t r0 = 0; r0 < n && rc == IRS_OK; r0 += m->max_batch) {
        const int64_t rows = std::min<int64_t>(m->max_batch, n - r0);
        const int B = (int)((rows + 63) / 64 * 64);
        rc = stage_inputs(m, x + r0 * m->n_in, rows, B);
        if (rc) break;
        rc = mlp_forward_staged(m, B);
        if (rc) break;
        for (int i = 0; i < m->n_in && rc == IRS_OK; ++i) {
            hipError_t e = hipMemcpyAsync(row, w1.data() + (size_t)i * m->H, sizeof(float) * m->H,
                                          hipMemcpyHostToDevice, m->stream);
            if (e == hipSuccess) e = launch_broadcast_row(row, B, m->H, m->tang, m->stream);
            if (e == hipSuccess) e = launch_gelu_tangent(m->Z[0], m->tang, (int64_t)B * m->H, m->dZ16[0], m->stream);
            for (int l = 1; l < 4 && e == hipSuccess; ++l) {
                GemmArgs g{};
                g.A = m->dZ16[l - 1];
                g.B = m->P16 + m->off_W[l];
                g.M = B;
                g.N = m->dims[l + 1];
                g.K = m->dims[l];
                g.lda = m->dims[l];
                g.ldb = m->dims[l + 1];
                g.ldc = m->dims[l + 1];
                g.alpha = 1.0f;
                g.C32 = l < 3 ? m->tang : m->Y;
                e = launch_gemm_bf16(g, false, false, m->stream);
                if (e == hipSuccess && l < 3)
                    e = launch_gelu_tangent(m->Z[l], m->tang, (int64_t)B * m->H, m->dZ16[l], m->stream);
            }
            if (e == hipSuccess)
                e = hipMemcpyAsync(dy.data(), m->Y, sizeof(float) * rows * m->n_out, hipMemcpyDeviceToHost, m->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
            if (e != hipSuccess) { rc = mfail(m, IRS_EHIP, hipGetErrorString(e)); break; }
            for (int64_t r = 0; r < rows; ++r)
                for (int o = 0; o < m->n_out; ++o)
                    J[((r0 + r) * m->n_out + o) * m->n_in + i] = dy[(size_t)r * m->n_out + o];
        }
    }
    (void)hipFree(row);
    return rc;
}

// ---- emulator ----------------------------------------------------------------
int irs_emu_create(irs_mlp *mlp, int n_layers, int n_pc, int n_channels, int n_feat, int n_local,
                   const float *x_mean, const float *x_std, const float *A, float coef_scale, const float *basis,
                   const float *mean, const float *inv_w, const double *ch_nu, irs_emu **out) {
    if (!out) return IRS_EINVAL;
    *out = nullptr;
    if (!mlp || !x_mean || !x_std || !A || !basis || !mean || !inv_w || !ch_nu || n_layers < 1 || n_pc < 1 ||
        n_channels < 1 || n_pc != mlp->n_out || n_local < 0 || n_local > n_feat || n_feat < mlp->n_in ||
        n_feat > 24 || !(coef_scale > 0.0f))
        return IRS_EINVAL;
    (void)hipSetDevice(mlp->device);
    irs_emu *e = new irs_emu();
    e->mlp = mlp;
    e->L = n_layers;
    e->n_pc = n_pc;
    e->n_ch = n_channels;
    e->n_feat = n_feat;
    e->n_local = n_local;
    e->n_phi = 1 + n_local + n_local * (n_local + 1) / 2 + (n_feat - n_local);
    e->coef_scale = coef_scale;
    const size_t W = (size_t)3 * n_channels, nA = (size_t)n_layers * e->n_phi * n_pc;
    e->x_mean.assign(x_mean, x_mean + (size_t)n_layers * n_feat);
    e->x_std.assign(x_std, x_std + (size_t)n_layers * n_feat);
    e->A_host.assign(A, A + nA);
    bool ok = dalloc(&e->A, nA) == hipSuccess && dalloc(&e->basis, (size_t)n_layers * n_pc * W) == hipSuccess &&
              dalloc(&e->mean, (size_t)n_layers * W) == hipSuccess &&
              dalloc(&e->inv_w, (size_t)n_layers * W) == hipSuccess && dalloc(&e->ch_nu, n_channels) == hipSuccess;
    ok = ok && hipMemcpy(e->A, A, sizeof(float) * nA, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(e->basis, basis, sizeof(float) * (size_t)n_layers * n_pc * W, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(e->mean, mean, sizeof(float) * (size_t)n_layers * W, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(e->inv_w, inv_w, sizeof(float) * (size_t)n_layers * W, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(e->ch_nu, ch_nu, sizeof(double) * n_channels, hipMemcpyHostToDevice) == hipSuccess;
    for (size_t i = 0; ok && i < e->x_std.size(); ++i) ok = e->x_std[i] > 0.0f;
    if (!ok) {
        irs_emu_destroy(e);
        return IRS_ENOMEM;
    }
    *out = e;
    return IRS_OK;
}

void irs_emu_destroy(irs_emu *e) {
    if (!e) return;
    (void)hipSetDevice(e->mlp->device);
    (void)hipStreamSynchronize(e->mlp->stream);
    void *ptrs[] = {e->A, e->basis, e->mean, e->inv_w, e->ch_nu, e->coef, e->phi, e->prof, e->R, e->BT, e->corr_W,
                    e->corr_e};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete e;
}

const char *irs_emu_last_error(const irs_emu *e) { return e ? e->err.c_str() : "null emulator"; }

}  // extern "C"

namespace {

#define ECHK(expr)                                                                          \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            e->err = std::string(#expr) + ": " + hipGetErrorString(e_);                     \
            return IRS_EHIP;                                                                \
        }                                                                                   \
    } while (0)

template <typename T>
hipError_t grow(T **p, int64_t &cap, int64_t need) {
    if (need <= cap) return hipSuccess;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    cap = 0;
    hipError_t r = dalloc(p, (size_t)need);
    if (r == hipSuccess) cap = need;
    return r;
}

// Standardised features z [rows][n_feat] and the regression terms phi(z) [rows][n_phi]:
// 1, z_i (i < n_local), z_a z_b (a <= b < n_local), z_i (n_local <= i < n_feat).
void emu_inputs(const irs_emu *e, const float *features, int64_t rows, std::vector<float> &z, std::vector<float> &phi) {
    const int nf = e->n_feat, nl = e->n_local;
    z.resize((size_t)rows * nf);
    phi.resize((size_t)rows * e->n_phi);
    for (int64_t r = 0; r < rows; ++r) {
        const int l = (int)(r % e->L);
        float *zr = z.data() + r * nf, *pr = phi.data() + r * e->n_phi;
        for (int i = 0; i < nf; ++i)
            zr[i] = (features[r * nf + i] - e->x_mean[(size_t)l * nf + i]) / e->x_std[(size_t)l * nf + i];
        int j = 0;
        pr[j++] = 1.0f;
        for (int i = 0; i < nl; ++i) pr[j++] = zr[i];
        for (int a = 0; a < nl; ++a)
            for (int b = a; b < nl; ++b) pr[j++] = zr[a] * zr[b];
        for (int i = nl; i < nf; ++i) pr[j++] = zr[i];
    }
}

// coefficients a [rows][n_pc] = A phi + coef_scale MLP(z[:, :n_in]) on the device
int emu_coefficients(irs_emu *e, const float *features, int64_t rows, std::vector<float> *z_out = nullptr) {
    irs_mlp *m = e->mlp;
    int64_t cap = e->cap_rows;
    ECHK(grow(&e->coef, cap, rows * e->n_pc));
    e->cap_rows = cap;
    cap = e->cap_phi;
    ECHK(grow(&e->phi, cap, rows * e->n_phi));
    e->cap_phi = cap;
    std::vector<float> z, phi;
    emu_inputs(e, features, rows, z, phi);
    std::vector<float> zin;
    const float *zm = z.data();
    if (m->n_in != e->n_feat) {
        zin.resize((size_t)rows * m->n_in);
        for (int64_t r = 0; r < rows; ++r)
            for (int i = 0; i < m->n_in; ++i) zin[r * m->n_in + i] = z[r * e->n_feat + i];
        zm = zin.data();
    }
    ECHK(hipMemcpyAsync(e->phi, phi.data(), sizeof(float) * phi.size(), hipMemcpyHostToDevice, m->stream));
    for (int64_t r0 = 0; r0 < rows; r0 += m->max_batch) {
        const int64_t nr = std::min<int64_t>(m->max_batch, rows - r0);
        const int B = (int)((nr + 63) / 64 * 64);
        int rc = stage_inputs(m, zm + r0 * m->n_in, nr, B);
        if (rc) {
            e->err = m->err;
            return rc;
        }
        rc = mlp_forward_staged(m, B);
        if (rc) {
            e->err = m->err;
            return rc;
        }
        ECHK(launch_emu_coef(m->Y, e->phi, e->A, e->coef_scale, r0, nr, e->L, e->n_phi, e->n_pc, e->coef, m->stream));
    }
    // synchronous here: z and phi are pageable host memory freed on return
    ECHK(hipStreamSynchronize(m->stream));
    if (z_out) z_out->swap(z);
    return IRS_OK;
}

int emu_upload_profiles(irs_emu *e, const double *T_lev, const double *T_ave, const double *scal, int n_prof) {
    irs_mlp *m = e->mlp;
    const int L = e->L;
    const int64_t pd = (int64_t)n_prof * (2 * L + 1 + 3);
    int64_t cp = e->cap_prof;
    ECHK(grow(&e->prof, cp, pd));
    e->cap_prof = cp;
    double *dp = e->prof;
    ECHK(hipMemcpyAsync(dp, T_lev, sizeof(double) * n_prof * (L + 1), hipMemcpyHostToDevice, m->stream));
    ECHK(hipMemcpyAsync(dp + (size_t)n_prof * (L + 1), T_ave, sizeof(double) * n_prof * L, hipMemcpyHostToDevice,
                        m->stream));
    ECHK(hipMemcpyAsync(dp + (size_t)n_prof * (2 * L + 1), scal, sizeof(double) * n_prof * 3, hipMemcpyHostToDevice,
                        m->stream));
    return IRS_OK;
}

EmuRteArgs emu_args(const irs_emu *e, int n_prof) {
    EmuRteArgs a{};
    a.coef = e->coef;
    a.basis = e->basis;
    a.mean = e->mean;
    a.inv_w = e->inv_w;
    a.ch_nu = e->ch_nu;
    a.T_lev = e->prof;
    a.T_ave = e->prof + (size_t)n_prof * (e->L + 1);
    a.scal = e->prof + (size_t)n_prof * (2 * e->L + 1);
    a.n_prof = n_prof;
    a.L = e->L;
    a.n_pc = e->n_pc;
    a.n_ch = e->n_ch;
    return a;
}

}  // namespace

extern "C" {

// features [n_prof][L][n_feat] (raw), T_lev [n_prof][L+1], T_ave [n_prof][L],
// scal [n_prof][3] = {T_sfc, mu, eps}; R, BT [n_prof][n_ch].  Optional
// t_out [n_prof][L][n_ch] (decoded level transmittances along the path).
int irs_emu_forward(irs_emu *e, const float *features, const double *T_lev, const double *T_ave,
                    const double *scal, int n_prof, int source_function, double *R, double *BT, float *t_out) {
    if (!e || !features || !T_lev || !T_ave || !scal || !R || !BT || n_prof < 1) return IRS_EINVAL;
    if (source_function != IRS_SRC_CIM) {
        e->err = "the emulator reproduces the CIM source function only";
        return IRS_EINVAL;
    }
    irs_mlp *m = e->mlp;
    (void)hipSetDevice(m->device);
    const int L = e->L;
    int rc = emu_coefficients(e, features, (int64_t)n_prof * L);
    if (rc) return rc;
    rc = emu_upload_profiles(e, T_lev, T_ave, scal, n_prof);
    if (rc) return rc;
    int64_t co = e->cap_out;
    if (co < (int64_t)n_prof * e->n_ch) {
        if (e->R) (void)hipFree(e->R);
        if (e->BT) (void)hipFree(e->BT);
        e->R = e->BT = nullptr;
        e->cap_out = 0;
        ECHK(dalloc(&e->R, (size_t)n_prof * e->n_ch));
        ECHK(dalloc(&e->BT, (size_t)n_prof * e->n_ch));
        e->cap_out = (int64_t)n_prof * e->n_ch;
    }
    float *tdev = nullptr;
    if (t_out) ECHK(dalloc(&tdev, (size_t)n_prof * L * e->n_ch));
    EmuRteArgs a = emu_args(e, n_prof);
    a.R = e->R;
    a.BT = e->BT;
    a.t_out = tdev;
    hipError_t er = launch_emu_rte(a, m->stream);
    if (er == hipSuccess)
        er = hipMemcpyAsync(R, e->R, sizeof(double) * n_prof * e->n_ch, hipMemcpyDeviceToHost, m->stream);
    if (er == hipSuccess)
        er = hipMemcpyAsync(BT, e->BT, sizeof(double) * n_prof * e->n_ch, hipMemcpyDeviceToHost, m->stream);
    if (er == hipSuccess && t_out)
        er = hipMemcpyAsync(t_out, tdev, sizeof(float) * n_prof * L * e->n_ch, hipMemcpyDeviceToHost, m->stream);
    if (er == hipSuccess) er = hipStreamSynchronize(m->stream);
    (void)hipStreamSynchronize(m->stream);
    if (tdev) (void)hipFree(tdev);
    if (er != hipSuccess) {
        e->err = std::string("emulated forward model: ") + hipGetErrorString(er);
        return IRS_EHIP;
    }
    return IRS_OK;
}

}  // extern "C"

extern "C" {

// Emulator Jacobians for n_prof scenes: BT [n_prof][n_ch]; dBT/dfeature
// [n_prof][L][n_feat][n_ch] (raw features, through the MLP input Jacobian, the
// regression and the decode); the Planck terms dBT/dT_ave [n_prof][L][n_ch],
// dBT/dT_lev [n_prof][L+1][n_ch], dBT/dT_sfc [n_prof][n_ch].
int irs_emu_jacobian(irs_emu *e, const float *features, const double *T_lev, const double *T_ave,
                     const double *scal, int n_prof, double *BT, float *dfeat, float *dTave, float *dTlev,
                     float *dTsfc) {
    if (!e || !features || !T_lev || !T_ave || !scal || n_prof < 1 || !BT || !dfeat || !dTave || !dTlev || !dTsfc)
        return IRS_EINVAL;
    irs_mlp *m = e->mlp;
    (void)hipSetDevice(m->device);
    const int L = e->L, nf = e->n_feat, nl = e->n_local, nin = m->n_in, npc = e->n_pc, nch = e->n_ch;
    const int nphi = e->n_phi, nq = nl * (nl + 1) / 2;
    const int64_t rows = (int64_t)n_prof * L;
    std::vector<float> z;
    int rc = emu_coefficients(e, features, rows, &z);
    if (rc) return rc;
    // MLP input Jacobian on the standardized inputs
    std::vector<float> zin((size_t)rows * nin);
    for (int64_t r = 0; r < rows; ++r)
        for (int i = 0; i < nin; ++i) zin[r * nin + i] = z[r * nf + i];
    std::vector<float> Jm((size_t)rows * npc * nin);
    rc = irs_mlp_input_jacobian(m, zin.data(), rows, Jm.data());
    if (rc) {
        e->err = m->err;
        return rc;
    }
    // d a / d feature = (coef_scale dMLP/dz + A dphi/dz) / x_std
    std::vector<float> J((size_t)rows * npc * nf, 0.0f);
    for (int64_t r = 0; r < rows; ++r) {
        const int l = (int)(r % L);
        const float *zr = z.data() + r * nf;
        const float *Al = e->A_host.data() + (size_t)l * nphi * npc;
        for (int o = 0; o < npc; ++o) {
            float *Jr = J.data() + ((size_t)r * npc + o) * nf;
            for (int i = 0; i < nin; ++i) Jr[i] = e->coef_scale * Jm[((size_t)r * npc + o) * nin + i];
            for (int i = 0; i < nl; ++i) Jr[i] += Al[(size_t)(1 + i) * npc + o];
            int jj = 1 + nl;
            for (int a = 0; a < nl; ++a)
                for (int b = a; b < nl; ++b, ++jj) {
                    const float w = Al[(size_t)jj * npc + o];
                    Jr[a] += w * zr[b];
                    Jr[b] += w * zr[a];
                }
            for (int i = nl; i < nf; ++i) Jr[i] += Al[(size_t)(1 + nl + nq + (i - nl)) * npc + o];
            for (int i = 0; i < nf; ++i) Jr[i] /= e->x_std[(size_t)l * nf + i];
        }
    }
    rc = emu_upload_profiles(e, T_lev, T_ave, scal, n_prof);
    if (rc) return rc;
    std::vector<void *> bufs;
    auto get = [&](size_t bytes) -> void * {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(bytes, 4)) != hipSuccess) return nullptr;
        bufs.push_back(p);
        return p;
    };
    float *dJ = (float *)get(sizeof(float) * J.size());
    EmuAdjArgs j{};
    j.BT = (double *)get(sizeof(double) * n_prof * nch);
    j.dt = (float *)get(sizeof(float) * rows * nch);
    j.dg = (float *)get(sizeof(float) * rows * nch);
    j.dm = (float *)get(sizeof(float) * rows * nch);
    j.dTave = (float *)get(sizeof(float) * rows * nch);
    j.dTlev = (float *)get(sizeof(float) * (rows + n_prof) * nch);
    j.dTsfc = (float *)get(sizeof(float) * n_prof * nch);
    j.dfeat = (float *)get(sizeof(float) * rows * nf * nch);
    j.Jcoef = dJ;
    bool ok = true;
    for (void *p : bufs) ok = ok && p;
    hipError_t er = hipSuccess;
    if (ok) {
        er = hipMemcpyAsync(dJ, J.data(), sizeof(float) * J.size(), hipMemcpyHostToDevice, m->stream);
        EmuRteArgs a = emu_args(e, n_prof);
        if (er == hipSuccess) er = launch_emu_adjoint(a, j, nf, m->stream);
        if (er == hipSuccess)
            er = hipMemcpyAsync(BT, j.BT, sizeof(double) * n_prof * nch, hipMemcpyDeviceToHost, m->stream);
        if (er == hipSuccess)
            er = hipMemcpyAsync(dfeat, j.dfeat, sizeof(float) * rows * nf * nch, hipMemcpyDeviceToHost, m->stream);
        if (er == hipSuccess)
            er = hipMemcpyAsync(dTave, j.dTave, sizeof(float) * rows * nch, hipMemcpyDeviceToHost, m->stream);
        if (er == hipSuccess)
            er = hipMemcpyAsync(dTlev, j.dTlev, sizeof(float) * (rows + n_prof) * nch, hipMemcpyDeviceToHost,
                                m->stream);
        if (er == hipSuccess)
            er = hipMemcpyAsync(dTsfc, j.dTsfc, sizeof(float) * n_prof * nch, hipMemcpyDeviceToHost, m->stream);
        if (er == hipSuccess) er = hipStreamSynchronize(m->stream);
    }
    (void)hipStreamSynchronize(m->stream);
    for (void *p : bufs) (void)hipFree(p);
    if (!ok) {
        e->err = "device allocation failed";
        return IRS_ENOMEM;
    }
    if (er != hipSuccess) {
        e->err = std::string("emulator Jacobian: ") + hipGetErrorString(er);
        return IRS_EHIP;
    }
    return IRS_OK;
}

}  // extern "C"

// ---- emulator + spot check + fallback behind one call -------------------------------
extern "C" int irs_emu_set_context(irs_emu *e, int n_ctx, int n_vec, const double *mean, const double *proj) {
    if (!e || n_ctx < 0 || n_vec < 0 || (n_ctx > 0 && (!mean || !proj))) return IRS_EINVAL;
    if (e->n_local + n_ctx != e->n_feat) {
        e->err = "n_ctx must equal n_feat - n_local";
        return IRS_EINVAL;
    }
    e->n_ctx = n_ctx;
    e->n_vec = n_vec;
    e->ctx_mean.assign(mean, mean + (size_t)n_vec);
    e->ctx_proj.assign(proj, proj + (size_t)n_ctx * n_vec);
    return IRS_OK;
}

namespace {

// Raw emulator inputs of one scene [L][n_feat] (13 layer features, then the profile
// context): ln p_mid, T_ave, ln vmr_cg of H2O and O3 (floor 1e-30), CO2 vmr_cg in
// ppm, 1/mu, ln of the H2O and O3 columns above the layer (sum vmr_cg dp, + 1e-9 /
// + 1e-12), T_ave of the layer above (its own at the top), and the Curtis-Godson
// means of the path above: H2O-weighted ln p_mid and T_ave, O3-weighted T_ave and
// mass-weighted T_ave ((sum_{j>k} u_j x_j + eps x_k) / (sum_{j>k} u_j + eps), with
// eps 1e-9, 1e-9, 1e-12, 1e-6, so the layer's own value where nothing is above).
int emu_scene_features(const irs_emu *e, const irs_profile &pr, int T_method, float *out, double *T_ave,
                       double &T_sfc) {
    const int nl = pr.n_levels, L = nl - 1, ns = pr.n_species;
    std::vector<double> pmid(L), dp(L), vc((size_t)ns * L), pc((size_t)ns * L), tc((size_t)ns * L);
    int rc = irs_layer_properties(&pr, T_method, pmid.data(), dp.data(), T_ave, vc.data(), pc.data(), tc.data());
    if (rc) return rc;
    auto idx = [&](int mol) {
        for (int s = 0; s < ns; ++s)
            if (pr.species[s] == mol) return s;
        return -1;
    };
    const int sh = idx(1), so = idx(3), sc = idx(2);
    auto v = [&](int s, int k) { return s < 0 ? 0.0 : vc[(size_t)s * L + k]; };
    // sums over the layers above k: amounts u = vmr dp of H2O and O3 and the mass dp,
    // and their moments in ln p and T (Curtis-Godson means of the overlying path)
    std::vector<double> ah(L, 0.0), ao(L, 0.0), ahp(L, 0.0), aht(L, 0.0), aot(L, 0.0), am(L, 0.0), amt(L, 0.0);
    double acc_h = 0.0, acc_o = 0.0, acc_hp = 0.0, acc_ht = 0.0, acc_ot = 0.0, acc_m = 0.0, acc_mt = 0.0;
    for (int k = L - 1; k >= 0; --k) {
        ah[k] = acc_h;
        ao[k] = acc_o;
        ahp[k] = acc_hp;
        aht[k] = acc_ht;
        aot[k] = acc_ot;
        am[k] = acc_m;
        amt[k] = acc_mt;
        const double uh = v(sh, k) * dp[k], uo = v(so, k) * dp[k], lnp = std::log(pmid[k]);
        acc_h += uh;
        acc_o += uo;
        acc_hp += uh * lnp;
        acc_ht += uh * T_ave[k];
        acc_ot += uo * T_ave[k];
        acc_m += dp[k];
        acc_mt += dp[k] * T_ave[k];
    }
    T_sfc = pr.T_sfc > 0.0 ? pr.T_sfc : pr.T[0];
    // context vector
    std::vector<double> cx(e->n_ctx, 0.0);
    if (e->n_ctx > 0) {
        if (e->n_vec != 3 * nl + 2) return IRS_EINVAL;
        std::vector<double> vec((size_t)e->n_vec);
        auto lnq = [&](int s, int j) {
            return s < 0 ? std::log(1e-30) : std::log(std::max(pr.vmr[(size_t)s * nl + j], 1e-30));
        };
        for (int j = 0; j < nl; ++j) {
            vec[j] = pr.T[j];
            vec[nl + j] = lnq(sh, j);
            vec[2 * nl + j] = lnq(so, j);
        }
        vec[3 * nl] = T_sfc;
        vec[3 * nl + 1] = 1.0 / pr.mu;
        for (int c = 0; c < e->n_ctx; ++c) {
            double a = 0.0;
            for (int j = 0; j < e->n_vec; ++j) a += e->ctx_proj[(size_t)c * e->n_vec + j] * (vec[j] - e->ctx_mean[j]);
            cx[c] = a;
        }
    }
    for (int k = 0; k < L; ++k) {
        float *f = out + (size_t)k * e->n_feat;
        const double lnp = std::log(pmid[k]), Tk = T_ave[k];
        const double loc[13] = {lnp, Tk, std::log(std::max(v(sh, k), 1e-30)),
                                std::log(std::max(v(so, k), 1e-30)), v(sc, k) * 1e6, 1.0 / pr.mu,
                                std::log(ah[k] + 1e-9), std::log(ao[k] + 1e-12), T_ave[k + 1 < L ? k + 1 : k],
                                (ahp[k] + 1e-9 * lnp) / (ah[k] + 1e-9), (aht[k] + 1e-9 * Tk) / (ah[k] + 1e-9),
                                (aot[k] + 1e-12 * Tk) / (ao[k] + 1e-12), (amt[k] + 1e-6 * Tk) / (am[k] + 1e-6)};
        for (int i = 0; i < e->n_local; ++i) f[i] = (float)loc[i];
        for (int c = 0; c < e->n_ctx; ++c) f[e->n_local + c] = (float)cx[c];
    }
    return IRS_OK;
}

}  // namespace

extern "C" int irs_emu_set_correction(irs_emu *e, int n_blocks, const int64_t *block_c0, const int32_t *block_n,
                                      const float *W, const float *lbo_W, double w_raw, double w_lbo) {
    if (!e || n_blocks < 0) return IRS_EINVAL;
    irs_mlp *m = e->mlp;
    (void)hipSetDevice(m->device);
    if (e->corr_W) (void)hipFree(e->corr_W);
    e->corr_W = nullptr;
    e->corr_bn.clear();
    e->corr_c0.clear();
    e->corr_lbo.clear();
    e->corr_ns = 0;
    if (n_blocks == 0) return IRS_OK;
    if (!block_c0 || !block_n || !W || !(w_raw >= 0.0) || !(w_lbo >= 0.0) || (w_lbo > 0.0 && (!lbo_W || n_blocks < 2))) {
        e->err = "irs_emu_set_correction: blocks and W are required, the weights are >= 0, and a leave-one-block-out "
                 "weight needs lbo_W and two or more blocks";
        return IRS_EINVAL;
    }
    int64_t ns = 0, n_lbo = 0;
    for (int b = 0; b < n_blocks; ++b) {
        if (block_n[b] < 1 || block_c0[b] < 0 || block_c0[b] + block_n[b] > e->n_ch) return IRS_EINVAL;
        ns += block_n[b];
    }
    for (int b = 0; b < n_blocks; ++b) n_lbo += (ns - block_n[b]) * block_n[b];
    if (ns > EMU_MAX_SPOT) {
        e->err = "irs_emu_set_correction: more than EMU_MAX_SPOT (1024) spot channels";
        return IRS_EINVAL;
    }
    hipError_t er = dalloc(&e->corr_W, (size_t)ns * e->n_ch);
    if (er == hipSuccess)
        er = hipMemcpy(e->corr_W, W, sizeof(float) * (size_t)ns * e->n_ch, hipMemcpyHostToDevice);
    if (er != hipSuccess) {
        if (e->corr_W) (void)hipFree(e->corr_W);
        e->corr_W = nullptr;
        e->err = std::string("irs_emu_set_correction: ") + hipGetErrorString(er);
        return IRS_EHIP;
    }
    e->corr_bn.assign(block_n, block_n + n_blocks);
    e->corr_c0.assign(block_c0, block_c0 + n_blocks);
    if (w_lbo > 0.0) e->corr_lbo.assign(lbo_W, lbo_W + n_lbo);
    e->corr_ns = (int)ns;
    e->corr_w_raw = w_raw;
    e->corr_w_lbo = w_lbo;
    return IRS_OK;
}

// Trained emulator file (written by the Python package, ChannelODEmulator.export_binary),
// little-endian, no padding:
//   "IRSEMU01", int32 n_in n_hidden n_out L n_pc n_ch n_feat n_local n_ctx n_vec n_blocks
//   corrected, float coef_scale, int64 n_params, float params[n_params],
//   float x_mean[L n_feat] x_std[L n_feat] A[L n_phi n_pc] basis[L n_pc 3 n_ch]
//   mean[L 3 n_ch] inv_w[L 3 n_ch], double ch_nu[n_ch] ctx_mean[n_vec] ctx_proj[n_ctx n_vec],
//   double threshold w_raw w_lbo, int64 block_c0[n_blocks], int32 block_n[n_blocks],
//   and when corrected: float W[n_spot n_ch] lbo[sum_b (n_spot - n_b) n_b].
extern "C" int irs_emu_load(const char *path, int device, int max_batch, irs_mlp **mlp_out, irs_emu **emu_out,
                            double *threshold, int max_blocks, int64_t *block_c0, int32_t *block_n,
                            int *n_blocks_out) {
    if (!path || !mlp_out || !emu_out || !threshold || !n_blocks_out || max_blocks < 0 ||
        (max_blocks > 0 && (!block_c0 || !block_n)))
        return IRS_EINVAL;
    *mlp_out = nullptr;
    *emu_out = nullptr;
    FILE *f = std::fopen(path, "rb");
    if (!f) return IRS_ENOTFOUND;
    bool ok = true;
    auto rd = [&](void *dst, size_t bytes) { ok = ok && std::fread(dst, 1, bytes, f) == bytes; };
    char magic[8] = {};
    rd(magic, 8);
    ok = ok && std::memcmp(magic, "IRSEMU01", 8) == 0;
    int32_t h[12] = {};
    rd(h, sizeof h);
    const int n_in = h[0], n_hidden = h[1], n_out = h[2], L = h[3], n_pc = h[4], n_ch = h[5], n_feat = h[6],
              n_local = h[7], n_ctx = h[8], n_vec = h[9], nb = h[10], corrected = h[11];
    ok = ok && n_in > 0 && n_hidden > 0 && n_out == n_pc && L > 0 && L <= 64 && n_pc > 0 && n_ch > 0 &&
         n_local >= 0 && n_local <= n_feat && n_feat <= 24 && n_ctx == n_feat - n_local && n_vec >= 0 &&
         nb >= 0 && nb <= max_blocks && (corrected == 0 || corrected == 1) && (corrected == 0 || nb > 0);
    float coef_scale = 0.0f;
    int64_t n_params = 0;
    rd(&coef_scale, sizeof coef_scale);
    rd(&n_params, sizeof n_params);
    const int n_phi = 1 + n_local + n_local * (n_local + 1) / 2 + (n_feat - n_local);
    const size_t W3 = ok ? (size_t)3 * n_ch : 0;
    std::vector<float> params, xm, xs, A, basis, mean, inv_w, Wc, lbo;
    std::vector<double> nu, cm, cp;
    double thr[3] = {};
    std::vector<int64_t> c0;
    std::vector<int32_t> bn;
    auto grab = [&](auto &v, size_t n) {
        if (!ok || n > ((size_t)1 << 34)) { ok = false; return; }
        v.resize(n);
        rd(v.data(), n * sizeof(v[0]));
    };
    if (ok) {
        grab(params, (size_t)n_params);
        grab(xm, (size_t)L * n_feat);
        grab(xs, (size_t)L * n_feat);
        grab(A, (size_t)L * n_phi * n_pc);
        grab(basis, (size_t)L * n_pc * W3);
        grab(mean, (size_t)L * W3);
        grab(inv_w, (size_t)L * W3);
        grab(nu, (size_t)n_ch);
        grab(cm, (size_t)n_vec);
        grab(cp, (size_t)n_ctx * n_vec);
        rd(thr, sizeof thr);
        grab(c0, (size_t)nb);
        grab(bn, (size_t)nb);
    }
    int64_t ns = 0, n_lbo = 0;
    for (int b = 0; ok && b < nb; ++b) {
        ok = bn[b] > 0 && c0[b] >= 0 && c0[b] + bn[b] <= n_ch;
        ns += ok ? bn[b] : 0;
    }
    for (int b = 0; ok && b < nb; ++b) n_lbo += (ns - bn[b]) * bn[b];
    if (ok && corrected) {
        grab(Wc, (size_t)ns * n_ch);
        grab(lbo, (size_t)n_lbo);
    }
    std::fclose(f);
    if (!ok) return IRS_EINVAL;
    irs_mlp *m = nullptr;
    int rc = irs_mlp_create(device, n_in, n_hidden, n_out, max_batch, 0, &m);
    if (rc) return rc;
    if (irs_mlp_n_params(m) != n_params) {
        irs_mlp_destroy(m);
        return IRS_EINVAL;
    }
    irs_emu *e = nullptr;
    rc = irs_mlp_set_params(m, params.data());
    if (!rc)
        rc = irs_emu_create(m, L, n_pc, n_ch, n_feat, n_local, xm.data(), xs.data(), A.data(), coef_scale,
                            basis.data(), mean.data(), inv_w.data(), nu.data(), &e);
    if (!rc) rc = irs_emu_set_context(e, n_ctx, n_vec, cm.data(), cp.data());
    if (!rc && corrected)
        rc = irs_emu_set_correction(e, nb, c0.data(), bn.data(), Wc.data(), lbo.empty() ? nullptr : lbo.data(),
                                    thr[1], thr[2]);
    if (rc) {
        if (e) irs_emu_destroy(e);
        irs_mlp_destroy(m);
        return rc;
    }
    *threshold = thr[0];
    for (int b = 0; b < nb; ++b) {
        block_c0[b] = c0[b];
        block_n[b] = bn[b];
    }
    *n_blocks_out = nb;
    *mlp_out = m;
    *emu_out = e;
    return IRS_OK;
}

namespace {

// Fallback statistic of the corrected emulator for one scene from its spot errors
// es[n_spot] (blocks in order): max(w_raw max|e|, w_lbo r) with r the largest
// leave-one-block-out residual |e_b - e_{S\b} M_b| (M_b [n_spot - n_b][n_b] row-major,
// the map fitted on calibration scenes from the other blocks' errors to block b's).
double emu_spot_stat(const irs_emu *e, const float *es) {
    const int nb = (int)e->corr_bn.size(), ns = e->corr_ns;
    double raw = 0.0, res = 0.0;
    for (int i = 0; i < ns; ++i) raw = std::max(raw, (double)std::fabs(es[i]));
    if (e->corr_w_lbo > 0.0) {
        std::vector<double> pred;
        const float *M = e->corr_lbo.data();
        int off = 0;
        for (int b = 0; b < nb; ++b) {
            const int n = e->corr_bn[b], no = ns - n;
            pred.assign((size_t)n, 0.0);
            int j = 0;
            for (int i = 0; i < ns; ++i) {
                if (i >= off && i < off + n) continue;
                const double x = es[i];
                const float *row = M + (size_t)j * n;
                for (int c = 0; c < n; ++c) pred[c] += x * (double)row[c];
                ++j;
            }
            for (int c = 0; c < n; ++c) res = std::max(res, std::fabs((double)es[off + c] - pred[c]));
            M += (size_t)no * n;
            off += n;
        }
    }
    return std::max(e->corr_w_raw * raw, e->corr_w_lbo * res);
}

}  // namespace

extern "C" int irs_forward_emulated(irs_ctx *ctx, irs_emu *e, const irs_profile *profiles, int n_prof,
                                    const irs_sounder *snd, const irs_fm_options *opt, int n_blocks,
                                    const int64_t *block_c0, const int32_t *block_n, const int32_t *block_lists,
                                    double threshold, double *R, double *BT, uint8_t *fell_back, double *spot_err) {
    if (!ctx || !e || !profiles || n_prof < 1 || !snd || !R || !BT || !fell_back || n_blocks < 0 ||
        (n_blocks > 0 && (!block_c0 || !block_n || !block_lists)))
        return IRS_EINVAL;
    if (e->n_local != 13 || e->n_local + e->n_ctx != e->n_feat) {
        e->err = "the emulator needs 13 layer features and its profile context (irs_emu_set_context)";
        return IRS_EINVAL;
    }
    irs_fm_options o;
    if (opt) o = *opt; else irs_default_options(&o);
    const int L = profiles[0].n_levels - 1, ns = profiles[0].n_species;
    if (L != e->L) {
        e->err = "profiles and emulator have different numbers of layers";
        return IRS_EINVAL;
    }
    const int64_t nch = e->n_ch;
    if (snd->n_channels != nch) {
        e->err = "sounder and emulator have different channels";
        return IRS_EINVAL;
    }
    if (e->corr_ns > 0) {
        bool ok = n_blocks == (int)e->corr_bn.size();
        for (int b = 0; ok && b < n_blocks; ++b) ok = block_n[b] == e->corr_bn[b] && block_c0[b] == e->corr_c0[b];
        if (!ok) {
            e->err = "the spot blocks differ from those of the emulator's correction (irs_emu_set_correction)";
            return IRS_EINVAL;
        }
    }
    // 1. emulator
    std::vector<float> X((size_t)n_prof * L * e->n_feat);
    std::vector<double> tlev((size_t)n_prof * (L + 1)), tave((size_t)n_prof * L), scal((size_t)n_prof * 3);
    for (int q = 0; q < n_prof; ++q) {
        const irs_profile &pr = profiles[q];
        if (pr.n_levels - 1 != L || pr.n_species != ns) return IRS_EINVAL;
        double ts = 0.0;
        int rc = emu_scene_features(e, pr, o.T_method, X.data() + (size_t)q * L * e->n_feat,
                                    tave.data() + (size_t)q * L, ts);
        if (rc) return rc;
        for (int k = 0; k <= L; ++k) tlev[(size_t)q * (L + 1) + k] = pr.T[k];
        scal[3 * q] = ts;
        scal[3 * q + 1] = pr.mu;
        scal[3 * q + 2] = pr.eps_sfc;
    }
    int rc = irs_emu_forward(e, X.data(), tlev.data(), tave.data(), scal.data(), n_prof, o.source_function, R, BT,
                             nullptr);
    if (rc) return rc;
    // 2. spot check: LBL of each block of consecutive channels on its own sub-grid,
    //    with the line lists that reach it
    std::vector<double> err((size_t)n_prof, 0.0);
    std::vector<irs_profile> sub(profiles, profiles + n_prof);
    const bool corr = e->corr_ns > 0;
    const int ns_spot = e->corr_ns;
    std::vector<float> es(corr ? (size_t)n_prof * ns_spot : 0);
    std::vector<double> spot_R(corr ? (size_t)n_prof * ns_spot : 0), spot_BT(spot_R.size());
    std::vector<int64_t> spot_ch;
    for (int b = 0; b < n_blocks; ++b) {
        const int64_t c0 = block_c0[b];
        const int32_t nb = block_n[b];
        if (c0 < 0 || nb < 1 || c0 + nb > nch) return IRS_EINVAL;
        irs_sounder sb = *snd;
        sb.nu_min = snd->nu_min + (double)c0 * snd->dnu;
        sb.nu_max = snd->nu_min + (double)(c0 + nb - 1) * snd->dnu;
        sb.n_channels = nb;
        for (auto &p : sub) p.linelist = block_lists + (size_t)b * ns;
        std::vector<double> rb((size_t)n_prof * nb), bb((size_t)n_prof * nb);
        rc = irs_forward(ctx, sub.data(), n_prof, &sb, &o, rb.data(), bb.data());
        if (rc) return rc;
        const size_t off = spot_ch.size();
        for (int c = 0; c < nb; ++c) spot_ch.push_back(c0 + c);
        for (int q = 0; q < n_prof; ++q)
            for (int c = 0; c < nb; ++c) {
                const double d = BT[(size_t)q * nch + c0 + c] - bb[(size_t)q * nb + c];
                err[q] = std::max(err[q], std::fabs(d));
                if (corr) {
                    const size_t i = (size_t)q * ns_spot + off + c;
                    es[i] = (float)d;
                    spot_R[i] = rb[(size_t)q * nb + c];
                    spot_BT[i] = bb[(size_t)q * nb + c];
                }
            }
    }
    // 2b. correction of every channel from the spot errors (the map fitted on
    //     calibration scenes), then the spot channels take their LBL values
    if (corr) {
        irs_mlp *m = e->mlp;
        (void)hipSetDevice(m->device);
        for (int q = 0; q < n_prof; ++q) err[q] = emu_spot_stat(e, es.data() + (size_t)q * ns_spot);
        if (e->cap_corr_e < (int64_t)es.size()) {
            if (e->corr_e) (void)hipFree(e->corr_e);
            e->corr_e = nullptr;
            e->cap_corr_e = 0;
            ECHK(dalloc(&e->corr_e, es.size()));
            e->cap_corr_e = (int64_t)es.size();
        }
        hipError_t er = hipMemcpyAsync(e->corr_e, es.data(), sizeof(float) * es.size(), hipMemcpyHostToDevice,
                                       m->stream);
        if (er == hipSuccess)
            er = launch_emu_correct(e->corr_e, e->corr_W, ns_spot, nch, n_prof, e->ch_nu, e->R, e->BT, m->stream);
        if (er == hipSuccess)
            er = hipMemcpyAsync(R, e->R, sizeof(double) * n_prof * nch, hipMemcpyDeviceToHost, m->stream);
        if (er == hipSuccess)
            er = hipMemcpyAsync(BT, e->BT, sizeof(double) * n_prof * nch, hipMemcpyDeviceToHost, m->stream);
        if (er == hipSuccess) er = hipStreamSynchronize(m->stream);
        if (er != hipSuccess) {
            (void)hipStreamSynchronize(m->stream);
            e->err = std::string("spot-check correction: ") + hipGetErrorString(er);
            return IRS_EHIP;
        }
        for (int q = 0; q < n_prof; ++q)
            for (int i = 0; i < ns_spot; ++i) {
                R[(size_t)q * nch + spot_ch[i]] = spot_R[(size_t)q * ns_spot + i];
                BT[(size_t)q * nch + spot_ch[i]] = spot_BT[(size_t)q * ns_spot + i];
            }
    }
    // 3. full LBL for the scenes over the threshold
    std::vector<irs_profile> fb;
    std::vector<int> fq;
    for (int q = 0; q < n_prof; ++q) {
        fell_back[q] = err[q] > threshold ? 1 : 0;
        if (spot_err) spot_err[q] = err[q];
        if (fell_back[q]) {
            fb.push_back(profiles[q]);
            fq.push_back(q);
        }
    }
    if (!fb.empty()) {
        std::vector<double> rf(fb.size() * nch), bf(fb.size() * nch);
        rc = irs_forward(ctx, fb.data(), (int)fb.size(), snd, &o, rf.data(), bf.data());
        if (rc) return rc;
        for (size_t i = 0; i < fq.size(); ++i) {
            std::memcpy(R + (size_t)fq[i] * nch, rf.data() + i * nch, sizeof(double) * nch);
            std::memcpy(BT + (size_t)fq[i] * nch, bf.data() + i * nch, sizeof(double) * nch);
        }
    }
    return IRS_OK;
}
