// Per-layer channel optical depths along the viewing path (the emulator's
// training target); tau_j are vertical layer optical depths, mu the cosine of
// the profile's zenith angle.
//
//   T^ch_k(c)   = sum_taps ILS(c) * exp(-sum_{j>=k} tau_j / mu)  level-to-space, k = 0..L
//   t^ch_k(c)   = sum_taps ILS(c) * exp(-tau_k / mu)             layer alone
//   tau^ch_k(c) = ln(T^ch_{k+1} / T^ch_k)          where T^ch_{k+1} >= floor
//               = -ln t^ch_k                       otherwise (the layer sits under an opaque
//                                                  column and does not reach the top)
//
// With these, prod_{j>=k} exp(-tau^ch_j) = T^ch_k wherever the column above
// transmits, so a channel-space RTE reproduces the ILS-convolved level
// transmittances of the scene it was computed on.
//
// A second pass gives the channel weight of the CIM source correction
// (src/Solver/schwarzschild.jl:130-209: the layer adds
// (B(T_top) - B(T_ave)) (T_{k+1} - T_k) f_cim(s_k), s_k = tau_k / mu):
//   G_k = sum_taps ILS * (T_{k+1} - T_k) f_cim(s_k),   g_k = G_k / T^ch_{k+1}
// (or, under an opaque column, the layer-alone sum_taps ILS (1 - e^{-s_k}) f_cim(s_k)),
// so that the channel RTE  sum_k B(T_ave)(T^ch_{k+1} - T^ch_k)
// + (B(T_top) - B(T_ave)) g_k T^ch_{k+1}  matches the ILS-convolved line-by-line
// radiance of the scene up to the variation of B across the ILS.  With raw_G
// the pass stores G_k itself (the emulator's level-transmittance target).
//
// One workgroup per (channel, profile); thread t walks tap points t, t+256, ...
// of the channel's window, forms the top-down optical depth in FP32 and keeps
// one accumulator per level and per layer in registers (static indices up to
// IRS_CHOD_MAXL layers), then a wave-shuffle + LDS reduction.
#include <algorithm>

#include "irs_devmath.h"
#include "irs_internal.h"

#define IRS_CHOD_MAXL 64

template <int MAXL>
__global__ void __launch_bounds__(256) chod_kernel(const ChodArgs a) {
    __shared__ float red[4][2 * MAXL + 1];
    const int c = blockIdx.x, q = blockIdx.y, t = threadIdx.x;
    const int L = a.n_layers;
    const float *tau = a.tau + (int64_t)q * L * a.n;
    const int64_t ic = a.ch_idx[c];
    const float imu = (float)(1.0 / a.scal[3 * q + 1]);
    float lev[MAXL + 1], lay[MAXL];
#pragma unroll
    for (int k = 0; k < MAXL; ++k) lev[k] = lay[k] = 0.0f;
    lev[MAXL] = 0.0f;
    const int64_t nt = 2 * a.half + 1;
    for (int64_t k0 = t; k0 < nt; k0 += 256) {
        const int64_t i = ic + k0 - a.half;
        if (i < 0 || i >= a.n) continue;
        const float w = (float)a.taps[k0];
        float cum = 0.0f;
        lev[MAXL] += w;                         // TOA: transmittance 1
#pragma unroll
        for (int k = MAXL - 1; k >= 0; --k) {
            if (k < L) {
                const float tk = tau[(int64_t)k * a.n + i] * imu;
                cum += tk;
                lev[k] += w * __expf(-cum);
                lay[k] += w * __expf(-tk);
            }
        }
    }
    // reduce 2*MAXL+1 values: lev[0..L-1], lev[MAXL] (top), lay[0..L-1]
    const int wv = t >> 6, ln = t & 63;
#pragma unroll
    for (int k = 0; k <= MAXL; ++k) {
        if (k < L || k == MAXL) {
            float v = lev[k];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            if (ln == 0) red[wv][k] = v;
        }
    }
#pragma unroll
    for (int k = 0; k < MAXL; ++k) {
        if (k < L) {
            float v = lay[k];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            if (ln == 0) red[wv][MAXL + 1 + k] = v;
        }
    }
    __syncthreads();
    if (t < L) {
        const int k = t;
        const double sc = a.dnu;
        const double up = sc * ((double)red[0][k + 1 == L ? MAXL : k + 1] + red[1][k + 1 == L ? MAXL : k + 1] +
                                red[2][k + 1 == L ? MAXL : k + 1] + red[3][k + 1 == L ? MAXL : k + 1]);
        const double here = sc * ((double)red[0][k] + red[1][k] + red[2][k] + red[3][k]);
        const double alone = sc * ((double)red[0][MAXL + 1 + k] + red[1][MAXL + 1 + k] + red[2][MAXL + 1 + k] +
                                   red[3][MAXL + 1 + k]);
        double od;
        if (up >= a.floor && here > 0.0) od = log(up / here);
        else od = alone > 0.0 ? -log(alone) : (double)a.od_max;
        od = fmin(fmax(od, 0.0), (double)a.od_max);
        a.out[((int64_t)q * L + k) * a.n_ch + c] = (float)od;
        if (a.lev_out) {
            a.lev_out[((int64_t)q * (L + 1) + k) * a.n_ch + c] = (float)here;
            if (k == L - 1) a.lev_out[((int64_t)q * (L + 1) + L) * a.n_ch + c] = (float)up;
        }
    }
}

__device__ __forceinline__ float cim_f32(float s, float e) {   // f_cim, schwarzschild.jl:59-64
    return s < 0.06f ? s * (1.0f / 6.0f) : 1.0f - 2.0f * (e / expm1f(-s) + 1.0f / s);
}

template <int MAXL>
__global__ void __launch_bounds__(256) chod_cim_kernel(const ChodArgs a) {
    __shared__ float red[4][2 * MAXL];
    const int c = blockIdx.x, q = blockIdx.y, t = threadIdx.x;
    const int L = a.n_layers;
    const float *tau = a.tau + (int64_t)q * L * a.n;
    const int64_t ic = a.ch_idx[c];
    const float imu = (float)(1.0 / a.scal[3 * q + 1]);
    float G[MAXL], Ga[MAXL];
#pragma unroll
    for (int k = 0; k < MAXL; ++k) G[k] = Ga[k] = 0.0f;
    const int64_t nt = 2 * a.half + 1;
    for (int64_t k0 = t; k0 < nt; k0 += 256) {
        const int64_t i = ic + k0 - a.half;
        if (i < 0 || i >= a.n) continue;
        const float w = (float)a.taps[k0];
        float cum = 0.0f;
#pragma unroll
        for (int k = MAXL - 1; k >= 0; --k) {
            if (k < L) {
                const float sk = tau[(int64_t)k * a.n + i] * imu;
                const float e = __expf(-sk);
                const float top = __expf(-cum);
                cum += sk;
                const float f = cim_f32(sk, e);
                const float u = 1.0f - e;
                G[k] += w * (top * u) * f;
                Ga[k] += w * u * f;
            }
        }
    }
    const int wv = t >> 6, ln = t & 63;
#pragma unroll
    for (int k = 0; k < MAXL; ++k) {
        if (k < L) {
            float v = G[k], va = Ga[k];
            for (int o = 32; o > 0; o >>= 1) {
                v += __shfl_xor(v, o, 64);
                va += __shfl_xor(va, o, 64);
            }
            if (ln == 0) {
                red[wv][k] = v;
                red[wv][MAXL + k] = va;
            }
        }
    }
    __syncthreads();
    if (t < L) {
        const int k = t;
        const double sc = a.dnu;
        const double g = sc * ((double)red[0][k] + red[1][k] + red[2][k] + red[3][k]);
        const double ga = sc * ((double)red[0][MAXL + k] + red[1][MAXL + k] + red[2][MAXL + k] + red[3][MAXL + k]);
        const double up = (double)a.lev_in[((int64_t)q * (L + 1) + k + 1) * a.n_ch + c];
        const double v = a.raw_G ? g : (up >= a.floor ? g / up : ga);
        a.g_out[((int64_t)q * L + k) * a.n_ch + c] = (float)fmin(fmax(v, 0.0), 1.0);
    }
}

// ---- equally spaced exact-node channels (idx_c = idx_0 + c S) -------------
// The same sums in three passes per profile, so that every per-point quantity
// is formed once instead of once per channel whose window covers the point:
//   1. chod_rows_kernel: one thread per grid point writes the rows [n_rows][n]
//      (FP32): k < L the level-to-space transmittance exp(-sum_{j>=k} s_j),
//      L the top of the atmosphere (1), L+1+k the layer alone exp(-s_k), and with
//      the CIM pass 2L+1+k = T_{k+1} (1 - e^{-s_k}) f_cim(s_k) and
//      3L+1+k = (1 - e^{-s_k}) f_cim(s_k);
//   2. chod_band_kernel: the tap sums of every row at every channel, the banded
//      reduction of the channel ILS (k = r + S m: for one residue r the G channels
//      of a workgroup slide along one strided sequence of the row, so one tap and
//      one new value per row feed G FMAs; NB rows share each tap);
//   3. chod_final_kernel: the per-(channel, layer) epilogue of chod_kernel and
//      chod_cim_kernel on those sums.
// With m_out, pass 2 also runs with the moment taps (k - half) dnu ILS_k over
// the level rows: M_k = ILS (x) [(nu - nu_c)(T_{k+1} - T_k)], the first-order
// term of Planck across the ILS, B(nu) = B(nu_c) + B'(nu_c)(nu - nu_c) + ..., that
// the channel-centre RTE leaves out.
__global__ void __launch_bounds__(256) chod_rows_kernel(const ChodArgs a, int q, float *rows, int cim) {
    const int L = a.n_layers;
    const int64_t n = a.n;
    const float *tau = a.tau + (int64_t)q * L * n;
    const float imu = (float)(1.0 / a.scal[3 * q + 1]);
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        float cum = 0.0f, top = 1.0f;
        rows[(int64_t)L * n + i] = 1.0f;
        for (int k = L - 1; k >= 0; --k) {
            const float sk = tau[(int64_t)k * n + i] * imu;
            const float e = __expf(-sk);
            cum += sk;
            const float lev = __expf(-cum);
            rows[(int64_t)k * n + i] = lev;
            rows[(int64_t)(L + 1 + k) * n + i] = e;
            if (cim) {
                const float f = cim_f32(sk, e), u = 1.0f - e;
                rows[(int64_t)(2 * L + 1 + k) * n + i] = (top * u) * f;
                rows[(int64_t)(3 * L + 1 + k) * n + i] = u * f;
            }
            top = lev;
        }
    }
}

struct BandArgs {
    const float *rows;             // [n_rows][n]
    int n_rows;
    int64_t n;
    const double *taps;            // [2 half + 1]
    int64_t half, idx0;
    int S, mc;                     // channel stride in grid steps; rows per work item
    int64_t n_ch;
    double *sums;                  // [n_rows][n_ch], tap sums (not yet scaled by dnu)
};

template <int G, int NB>
__global__ void __launch_bounds__(256) chod_band_kernel(const BandArgs a) {
    __shared__ float red[4][NB * G];
    const int n_groups = (int)((a.n_ch + G - 1) / G);
    const int per_xcd = (n_groups + 7) / 8;
    const int b = blockIdx.x;
    const int grp = (b % 8) * per_xcd + b / 8;                 // XCD-contiguous channel groups
    if (grp >= n_groups) return;                              // block-uniform
    const int r0 = blockIdx.y * NB;
    // element offsets fit 32 bits (the launcher checks n_rows n < 2^31)
    const unsigned n32 = (unsigned)a.n;
    unsigned roff[NB];
#pragma unroll
    for (int bb = 0; bb < NB; ++bb) roff[bb] = (unsigned)min(r0 + bb, a.n_rows - 1) * n32;
    const int64_t c0 = (int64_t)grp * G;
    const int S = a.S, mc = a.mc;
    const int64_t nt = 2 * a.half + 1;
    const int64_t base = a.idx0 - a.half + c0 * S;            // grid index of (r = 0, m = 0), channel c0
    const int64_t rows0 = (nt + S - 1) / S;
    const int64_t items = (int64_t)S * ((rows0 + mc - 1) / mc);
    float acc[NB][G];
#pragma unroll
    for (int bb = 0; bb < NB; ++bb)
#pragma unroll
        for (int g = 0; g < G; ++g) acc[bb][g] = 0.0f;
    for (int64_t w = threadIdx.x; w < items; w += 256) {
        const int r = (int)(w % S);
        const int m0 = (int)((w / S) * mc);
        const int mr = (int)((nt - r + S - 1) / S);             // rows of residue r
        if (m0 >= mr) continue;
        const int m1 = min(m0 + mc, mr);
        const int64_t jb = base + r;                            // grid index of (m, g): jb + S (m + g)
        const double *tp = a.taps + r + (int64_t)S * m0;
        float win[NB][G];
        if (jb + (int64_t)S * m0 >= 0 && jb + (int64_t)S * (m1 + G - 1) < a.n) {
            unsigned o = (unsigned)(jb + (int64_t)S * m0);
#pragma unroll
            for (int g = 0; g < G; ++g) {
#pragma unroll
                for (int bb = 0; bb < NB; ++bb) win[bb][g] = a.rows[roff[bb] + o];
                o += S;
            }
            for (int m = m0; m < m1; m += G) {
#pragma unroll
                for (int u = 0; u < G; ++u) {
                    if (m + u < m1) {
                        const float t = (float)tp[(unsigned)(S * (m + u - m0))];
#pragma unroll
                        for (int bb = 0; bb < NB; ++bb) {
#pragma unroll
                            for (int g = 0; g < G; ++g) acc[bb][g] = fmaf(win[bb][(u + g) % G], t, acc[bb][g]);
                            win[bb][u] = a.rows[roff[bb] + o];             // row m + u + G (past m1: unused)
                        }
                        o += S;
                    }
                }
            }
        } else {
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int64_t j = jb + (int64_t)S * (m0 + g);
                const bool in = j >= 0 && j < a.n;
#pragma unroll
                for (int bb = 0; bb < NB; ++bb) win[bb][g] = in ? a.rows[roff[bb] + (unsigned)j] : 0.0f;
            }
            for (int m = m0; m < m1; m += G) {
#pragma unroll
                for (int u = 0; u < G; ++u) {
                    const int mm = m + u;
                    if (mm < m1) {
                        const float t = (float)a.taps[r + (int64_t)S * mm];
                        const int64_t j = jb + (int64_t)S * (mm + G);
                        const bool in = j >= 0 && j < a.n;
#pragma unroll
                        for (int bb = 0; bb < NB; ++bb) {
#pragma unroll
                            for (int g = 0; g < G; ++g) acc[bb][g] = fmaf(win[bb][(u + g) % G], t, acc[bb][g]);
                            win[bb][u] = in ? a.rows[roff[bb] + (unsigned)j] : 0.0f;
                        }
                    }
                }
            }
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int bb = 0; bb < NB; ++bb)
#pragma unroll
        for (int g = 0; g < G; ++g) {
            float v = acc[bb][g];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            if (lane == 0) red[wv][bb * G + g] = v;
        }
    __syncthreads();
    if (threadIdx.x < NB * G) {
        const int bb = threadIdx.x / G, g = threadIdx.x % G;
        const int64_t c = c0 + g;
        const int row = r0 + bb;
        if (c < a.n_ch && row < a.n_rows)
            a.sums[(int64_t)row * a.n_ch + c] = (double)red[0][threadIdx.x] + red[1][threadIdx.x] +
                                                red[2][threadIdx.x] + red[3][threadIdx.x];
    }
}

__global__ void __launch_bounds__(256) moment_taps_kernel(const double *taps, int64_t half, double dnu, double *mt) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i <= 2 * half) mt[i] = taps[i] * (double)(i - half) * dnu;
}

__global__ void __launch_bounds__(256) chod_final_kernel(const ChodArgs a, int q, const double *sums,
                                                         const double *msums) {
    const int L = a.n_layers;
    const int64_t id = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (id >= (int64_t)L * a.n_ch) return;
    const int k = (int)(id / a.n_ch);
    const int64_t c = id % a.n_ch, nc = a.n_ch;
    const double sc = a.dnu;
    const double here = sc * sums[(int64_t)k * nc + c];
    const double up = sc * sums[(int64_t)(k + 1) * nc + c];          // row L: the top
    const double alone = sc * sums[(int64_t)(L + 1 + k) * nc + c];
    double od;
    if (up >= a.floor && here > 0.0) od = log(up / here);
    else od = alone > 0.0 ? -log(alone) : (double)a.od_max;
    od = fmin(fmax(od, 0.0), (double)a.od_max);
    a.out[((int64_t)q * L + k) * nc + c] = (float)od;
    if (a.lev_out) {
        a.lev_out[((int64_t)q * (L + 1) + k) * nc + c] = (float)here;
        if (k == L - 1) a.lev_out[((int64_t)q * (L + 1) + L) * nc + c] = (float)up;
    }
    if (a.g_out) {
        const double g = sc * sums[(int64_t)(2 * L + 1 + k) * nc + c];
        const double ga = sc * sums[(int64_t)(3 * L + 1 + k) * nc + c];
        const double upf = (double)(float)up;                            // as stored in lev_out
        const double v = a.raw_G ? g : (upf >= a.floor ? g / upf : ga);
        a.g_out[((int64_t)q * L + k) * nc + c] = (float)fmin(fmax(v, 0.0), 1.0);
    }
    if (a.m_out) {
        const double M = sc * (msums[(int64_t)(k + 1) * nc + c] - msums[(int64_t)k * nc + c]);
        a.m_out[((int64_t)q * L + k) * nc + c] = up >= a.floor ? (float)(M / up) : 0.0f;
    }
}

#define CHOD_BAND_G 16
#define CHOD_BAND_NB 2

int64_t chod_banded_rows(int n_layers, bool cim) { return cim ? 4 * (int64_t)n_layers + 1 : 2 * (int64_t)n_layers + 1; }

hipError_t launch_chod_banded(const ChodArgs &a, int n_prof, int64_t idx0, int stride, float *rows, double *sums,
                              double *msums, hipStream_t s) {
    if (a.n_layers > IRS_CHOD_MAXL || a.n_layers < 1 || stride < 1 || a.n_ch <= 0) return hipErrorInvalidValue;
    if (a.g_out && !a.lev_out) return hipErrorInvalidValue;
    if (a.m_out && !msums) return hipErrorInvalidValue;
    const int L = a.n_layers;
    const int cim = a.g_out ? 1 : 0;
    if (chod_banded_rows(L, cim) * a.n >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
    BandArgs b{};
    b.rows = rows;
    b.n_rows = (int)chod_banded_rows(L, cim);
    b.n = a.n;
    b.taps = a.taps;
    b.half = a.half;
    b.idx0 = idx0;
    b.S = stride;
    b.n_ch = a.n_ch;
    b.sums = sums;
    const int64_t rows0 = (2 * a.half + 1 + stride - 1) / stride;
    const int64_t chunks = std::max<int64_t>(1, (rows0 + 47) / 48);
    b.mc = (int)((rows0 + chunks - 1) / chunks);
    const int n_groups = (int)((a.n_ch + CHOD_BAND_G - 1) / CHOD_BAND_G);
    const dim3 bgrid((unsigned)(((n_groups + 7) / 8) * 8), (unsigned)((b.n_rows + CHOD_BAND_NB - 1) / CHOD_BAND_NB));
    const unsigned rgrid = (unsigned)std::min<int64_t>((a.n + 255) / 256, 16384);
    const unsigned fgrid = (unsigned)(((int64_t)L * a.n_ch + 255) / 256);
    BandArgs bm = b;                                       // moment taps over the level rows 0..L
    double *mtaps = msums ? msums + (int64_t)(L + 1) * a.n_ch : nullptr;
    if (a.m_out) {
        bm.n_rows = L + 1;
        bm.taps = mtaps;
        bm.sums = msums;
        hipLaunchKernelGGL(moment_taps_kernel, dim3((unsigned)((2 * a.half + 1 + 255) / 256)), dim3(256), 0, s, a.taps,
                           a.half, a.dnu, mtaps);
    }
    const dim3 mgrid(bgrid.x, (unsigned)((L + 1 + CHOD_BAND_NB - 1) / CHOD_BAND_NB));
    for (int q = 0; q < n_prof; ++q) {
        hipLaunchKernelGGL(chod_rows_kernel, dim3(rgrid), dim3(256), 0, s, a, q, rows, cim);
        hipLaunchKernelGGL((chod_band_kernel<CHOD_BAND_G, CHOD_BAND_NB>), bgrid, dim3(256), 0, s, b);
        if (a.m_out) hipLaunchKernelGGL((chod_band_kernel<CHOD_BAND_G, CHOD_BAND_NB>), mgrid, dim3(256), 0, s, bm);
        hipLaunchKernelGGL(chod_final_kernel, dim3(fgrid), dim3(256), 0, s, a, q, (const double *)sums,
                           (const double *)msums);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_chod(const ChodArgs &a, int n_prof, hipStream_t s) {
    if (a.n_layers > IRS_CHOD_MAXL || a.n_layers < 1) return hipErrorInvalidValue;
    if (a.g_out && !(a.lev_out && a.lev_in == a.lev_out)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(chod_kernel<IRS_CHOD_MAXL>, dim3((unsigned)a.n_ch, (unsigned)n_prof), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !a.g_out) return e;
    hipLaunchKernelGGL(chod_cim_kernel<IRS_CHOD_MAXL>, dim3((unsigned)a.n_ch, (unsigned)n_prof), dim3(256), 0, s, a);
    return hipGetLastError();
}
