// FP32 tiled Voigt (+ VP_Y dispersive) line sum: tau[layer][nu] for a batch
// of profiles.
//
// Semantics restated from voigt_cross_section_kernel! (src/CrossSections/
// voigt.jl:243-277), lm_dispersive_kernel! (src/CrossSections/line_mixing.jl:
// 1166-1186), compute_voigt_lm_cross_sections (:1351-1369) and the tau
// accumulation of forward_model (src/Sensor/sounder.jl:283-306):
//   sigma_s(nu) = max(sum_{|nu-nu0|<=cut} Snorm (H(x,y) - (2 - (dnu/cut)^2) Hc), 0)
//   sigma_CO2   = max(max(sigma_v, 0) + sum_{|nu-nu0|<=lm_cut} YSn Im w, 0)   (VP_Y)
//   tau         = sum_s coef_s sigma_s         (coef folded into the records)
//
// Design (MI355X): one workgroup = one spectral tile of 256*P grid points of
// one (profile, layer); thread t owns points i0 + t + 256 p, so each wave
// covers 64 consecutive points per p (coalesced stores, coherent branches).
// Lines are read with wave-uniform indices (scalar loads, SGPR operands).  Per
// species, the lines near the tile are split by six binary searches on the
// sorted unshifted-centre keys into
//   EDGE  : window boundary crosses the tile   -> per-pair window test + pedestal
//   FAST  : far wing for every point of tile   -> A/(du^2+g^2): 3 VALU + 1 rcp
//   CORE  : may have |x| <= x_far in the tile  -> far wing, asymptotic series of
//           w(z) for |z| >= 8, or exact FP32 w(z), chosen per 64-point slice
// The interior (FAST + CORE) lines' pedestal, a quadratic in the grid offset,
// is summed per tile in FP64 (three moments) instead of per pair.  Partial
// sums run over 16- or 32-line chunks and are folded into Kahan-compensated
// FP32 accumulators, combined in FP64 before the clamp.  Offsets are formed from the
// integer grid index and the line centre split into an integer index plus an
// FP32 fraction (centre computed in FP64), so du carries one FP32 rounding.
#include "irs_devmath.h"
#include "irs_internal.h"

__device__ __forceinline__ void kahan_add(float &s, float &c, float v) {
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

__device__ __forceinline__ int lower_key(const int32_t *k, int n, int64_t v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((int64_t)k[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ int upper_key(const int32_t *k, int n, int64_t v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((int64_t)k[mid] <= v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// First line with key >= v (lower bound): one load from the list's key table
// when v lies in its range, else a binary search of the keys.
__device__ __forceinline__ int first_ge(const LinesumArgs &a, const ItemDev &it, const int32_t *k, int n,
                                        int64_t v) {
    const int64_t r = v - a.ktab_lo;
    if (a.ktab && r >= 0 && r < a.ktab_n) return a.ktab[it.ktab_off + r];
    return lower_key(k, n, v);
}

// Line records read with a wave-uniform index go through the constant address
// space, so that they are fetched by scalar loads into SGPRs (the records are
// written by the prep kernel and only read here).
#define IRS_CONST __attribute__((address_space(4)))
typedef float irs_f4 __attribute__((ext_vector_type(4)));
typedef const IRS_CONST irs_f4 *crec;
__device__ __forceinline__ crec as_crec(const float4 *p) { return (crec)p; }
__device__ __forceinline__ float4 ld(crec p, int64_t j) {
    const irs_f4 v = p[j];
    return make_float4(v.x, v.y, v.z, v.w);
}

// FAST: far wing, window covers the whole tile, pedestal handled per tile.
// Line records stream through LDS in chunks of LS_CHUNK lines: each thread
// loads LS_CHUNK/256 records with coalesced global loads one chunk ahead
// (double buffer) and stages {c, g^2, A, B} with c the line centre relative to
// the tile origin in grid steps, formed as float(cn - i0) + frac (one FP32
// rounding: |error| <= ulp(|c|)/2, i.e. ~1e-7 of |du| for every FAST pair,
// since |du| >= |c| - tile width and |du| > core half-width).  Every wave then
// reads the chunk with broadcast LDS reads; a pair costs sub, fma, rcp, fma
// (+ mul, fma for the dispersive term).  Partial sums run over 32 lines and
// are folded into the Kahan accumulators.
#define LS_CHUNK 512
#define LS_PER_THREAD (LS_CHUNK / IRS_LS_THREADS)

template <int P, bool LM>
__device__ __forceinline__ void fast_line(const float4 q, const float (&tp)[P], float (&part)[P],
                                          float (&partd)[P]) {
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const float du = tp[p] - q.x;
        const float r = irs_rcpf(fmaf(du, du, q.y));
        part[p] = fmaf(q.z, r, part[p]);
        if (LM) partd[p] = fmaf(q.w * du, r, partd[p]);
    }
}

template <int P, bool LM>
__device__ __forceinline__ void fast_range(const float4 *__restrict__ r0p, const float4 *__restrict__ r1p,
                                           int jlo, int jhi, int t, int i0, float4 *s_rec,
                                           float (&acc)[P], float (&cmp)[P], float (&accd)[P],
                                           float (&cmpd)[P]) {
    const int n = jhi - jlo;
    if (n <= 0) return;
    const int nch = (n + LS_CHUNK - 1) / LS_CHUNK;
    float tp[P];
#pragma unroll
    for (int p = 0; p < P; ++p) tp[p] = (float)(t + 256 * p);
    float4 g[LS_PER_THREAD];
    float gb[LS_PER_THREAD];
#pragma unroll
    for (int u = 0; u < LS_PER_THREAD; ++u) {
        const int j = jlo + t + u * IRS_LS_THREADS;
        g[u] = j < jhi ? r0p[j] : make_float4(__int_as_float(i0 + 1000000), 0.0f, 1.0f, 0.0f);
        gb[u] = (LM && j < jhi) ? r1p[j].x : 0.0f;
    }
    for (int c = 0; c < nch; ++c) {
        float4 *buf = s_rec + (c & 1) * LS_CHUNK;
#pragma unroll
        for (int u = 0; u < LS_PER_THREAD; ++u)
            buf[t + u * IRS_LS_THREADS] = make_float4((float)(__float_as_int(g[u].x) - i0) + g[u].y, g[u].z,
                                                      g[u].w, gb[u]);
        __syncthreads();
        if (c + 1 < nch) {
#pragma unroll
            for (int u = 0; u < LS_PER_THREAD; ++u) {
                const int j = jlo + (c + 1) * LS_CHUNK + t + u * IRS_LS_THREADS;
                g[u] = j < jhi ? r0p[j] : make_float4(__int_as_float(i0 + 1000000), 0.0f, 1.0f, 0.0f);
                gb[u] = (LM && j < jhi) ? r1p[j].x : 0.0f;
            }
        }
        const int cnt = min(LS_CHUNK, n - c * LS_CHUNK);
        for (int k0 = 0; k0 < cnt; k0 += 32) {
            const int ke = min(k0 + 32, cnt);
            float part[P], partd[P];
#pragma unroll
            for (int p = 0; p < P; ++p) { part[p] = 0.0f; partd[p] = 0.0f; }
            int k = k0;
            for (; k + 2 <= ke; k += 2) {
                const float4 qa = buf[k], qb = buf[k + 1];
                fast_line<P, LM>(qa, tp, part, partd);
                fast_line<P, LM>(qb, tp, part, partd);
            }
            for (; k < ke; ++k) fast_line<P, LM>(buf[k], tp, part, partd);
#pragma unroll
            for (int p = 0; p < P; ++p) {
                kahan_add(acc[p], cmp[p], part[p]);
                if (LM) kahan_add(accd[p], cmpd[p], partd[p]);
            }
        }
    }
    __syncthreads();
}

// CORE: interior lines whose Voigt core (|x| <= x_far) may reach the tile.
// For each line and each 64-point slice of a wave, the slice's distance from
// the line centre (a wave-uniform integer, compared with the line's KO and KI
// from its mid-core record) selects the form evaluated for the slice:
//   far   : every point has |x| > x_far            -> Lorentz wing A/(du^2 + g^2)
//   mid   : no point has |z| < IRS_Z1              -> asymptotic series of w(z) where
//                                                     |x| <= x_far, Lorentz wing beyond
//   inner : some point may have |z| < IRS_Z1       -> as mid, and exact FP32 w(z) at the
//                                                     points with |z| < IRS_Z1
// The series (see midcore_coeffs in the prep kernel) is a degree-8 polynomial
// in v = 1/(x^2 + y^2) after one reciprocal, within 3.3e-7 (Re w) and 2.6e-7
// (Im w) relative of w(z) for |z| >= 8; the exact branch is taken only by waves
// holding a point with |z| < 8.  The pedestal of these lines is part of the
// per-tile moments.
// 1 + k2 v + k3 v^2 + ... + k9 v^8 (Horner in FP32); the caller forms v = 1/(x^2 + Y).
__device__ __forceinline__ float series_core(float v, const float4 k0, const float4 k1, float k2) {
    float q = fmaf(k1.z, v, k1.y);
    q = fmaf(q, v, k1.x);
    q = fmaf(q, v, k0.w);
    q = fmaf(q, v, k0.z);
    q = fmaf(q, v, k0.y);
    q = fmaf(q, v, k0.x);
    q = fmaf(q, v, k2);
    return fmaf(q, v, 1.0f);
}

// Voigt term of one CORE pair outside the far form: Sn Re w(x + iy) (the series
// where |x| <= x_far, the Lorentz wing beyond it: both share v).  `pin` is set
// where the point needs the exact w(z) (|x| <= x_far and |z| < IRS_Z1); the
// returned value there is the series value, which the caller replaces.
__device__ __forceinline__ float core_voigt(float du, const float4 q1, crec c, float xfar, bool &pin) {
    const float4 k0 = ld(c, 0), k1 = ld(c, 1), k2 = ld(c, 2);
    const float x = du * q1.z;
    const float zz = fmaf(x, x, k1.w);
    const float v = irs_rcpf(zz);
    const bool in_core = fabsf(x) <= xfar;
    pin = in_core && zz < (float)(IRS_Z1 * IRS_Z1);
    return k2.x * v * (in_core ? series_core(v, k0, k1, 1.5f) : 1.0f);
}

// Dispersive term of one CORE pair outside the far form: YSn Im w(x_d + i y_d),
// with `pin` as for core_voigt.
__device__ __forceinline__ float core_disp(float du, const float4 q2, crec c, float xfar, bool &pin) {
    const float4 k0 = ld(c, 3), k1 = ld(c, 4), k2 = ld(c, 5);
    const float x = du * q2.y;
    const float zz = fmaf(x, x, k1.w);
    const float v = irs_rcpf(zz);
    const bool in_core = fabsf(x) <= xfar;
    pin = in_core && zz < (float)(IRS_Z1 * IRS_Z1);
    return (k2.x * x) * v * (in_core ? series_core(v, k0, k1, 0.5f) : 1.0f);
}

// Exact inner-core values: Sn Re w (Voigt, type 0) or YSn Im w (dispersive, type 1),
// one w(z) evaluation site for both.
__device__ __forceinline__ float inner_value(float du, const float4 q1, const float4 q2, int type) {
    float wr, wi;
    w32_core(du * (type ? q2.y : q1.z), type ? q2.z : q1.w, wr, wi);
    return type ? q2.w * wi : q2.x * wr;
}

// Queues the inner points m (a contiguous lane range) of slice p as one task of
// at most 32 lanes, or two when the range is wider.
__device__ __forceinline__ void queue_inner(unsigned long long m, int j, int p, int type, int lane, int &qn, int &qj,
                                            int &qi) {
    const int n = __popcll(m);
    if (n == 0) return;
    const int a = __ffsll((long long)m) - 1;
    const int n0 = min(n, 32);
    if (lane == qn) { qj = j; qi = p | (a << 4) | (n0 << 10) | (type << 17); }
    ++qn;
    if (n > 32) {
        if (lane == qn) { qj = j; qi = p | ((a + 32) << 4) | ((n - 32) << 10) | (type << 17); }
        ++qn;
    }
}

// Inner-core work is packed: a line's points with |z| < IRS_Z1 span ~30 grid
// points, so evaluating them where they fall would run w(z) over whole 64-lane
// slices (1.5 slices per line on average).  Instead the slice's inner points
// (a contiguous lane range [a, a + n), found by ballot) are zeroed in the main
// pass and queued as a task (line, slice, a, n, type) in two registers; full
// queues are flushed two tasks per wave evaluation (one per half-wave), each
// lane adding its value with an LDS atomic into the wave's private slot array
// [type][P][64] (in the otherwise idle FAST staging buffer), which the lanes
// fold into their compensated sums at the end of the range (a slice with more
// than 32 inner points makes two tasks).  w(z) then has one code site in the
// kernel instead of one per slice and term.  The deferred slots need
// P * (LM ? 2 : 1) * 4 * 64 floats of the staging buffer.
#define CORE_DEFER(P, LM) ((P) * ((LM) ? 2 : 1) * 4 * 64 <= 2 * LS_CHUNK * 4)

template <int P, bool LM>
__device__ __forceinline__ void core_range(crec r0p, crec r1p, crec r2p, crec r3p,
                                           int jlo, int jhi, int tl, float xfar, float *s_inner,
                                           float (&acc)[P], float (&cmp)[P],
                                           float (&accd)[P], float (&cmpd)[P]) {
    constexpr bool DEFER = CORE_DEFER(P, LM);
    const int lane = threadIdx.x & 63;
    const int tw = uni(tl & ~63);                      // first point of this wave's slices
    float *iv = s_inner + (threadIdx.x >> 6) * (P * 64 * (LM ? 2 : 1));
    if (DEFER) {
        __syncthreads();                               // the staging buffer is free
#pragma unroll
        for (int p = 0; p < P * (LM ? 2 : 1); ++p) iv[p * 64 + lane] = 0.0f;
    }
    int qn = 0, qj = 0, qi = 0;                        // task queue (lane q holds task q)
    int j = jlo;
    while (j < jhi) {
        const int je = min(j + 16, jhi);
        float part[P], partd[P];
#pragma unroll
        for (int p = 0; p < P; ++p) { part[p] = 0.0f; partd[p] = 0.0f; }
        for (; j < je && (!DEFER || qn <= 64 - 4 * P); ++j) {
            const float4 q0 = ld(r0p, j), q1 = ld(r1p, j), q2 = ld(r2p, j), c2 = ld(r3p + 6 * (int64_t)j, 2);
            const float4 c5 = LM ? ld(r3p + 6 * (int64_t)j, 5) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            const crec c = r3p + 6 * (int64_t)j;
            const int cidx = __float_as_int(q0.x);
            const int KO = __float_as_int(c2.y), KI = __float_as_int(c2.z);
            const int KOd = LM ? __float_as_int(c5.y) : -1, KId = LM ? __float_as_int(c5.z) : -1;
            const float base = (float)(tl - cidx);
            const int dw = tw - cidx;
#pragma unroll
            for (int p = 0; p < P; ++p) {
                // lower bound of |du| over the slice: du in (d0 - 1, d0 + 63]
                const int d0 = dw + 256 * p;
                const int mind = d0 >= 1 ? d0 - 1 : (d0 + 63 < 0 ? -(d0 + 63) : 0);
                const float du = (base + (float)(256 * p)) - q0.y;
                const bool farV = mind > KO, farD = !LM || mind > KOd;
                float r = 0.0f;
                if (farV || farD) r = irs_rcpf(fmaf(du, du, q0.z));
                if (farV) {
                    part[p] = fmaf(q0.w, r, part[p]);
                } else {
                    bool pin;
                    float h = core_voigt(du, q1, c, xfar, pin);
                    if (mind <= KI) {
                        if (DEFER) {
                            queue_inner(__ballot(pin), j, p, 0, lane, qn, qj, qi);
                            if (pin) h = 0.0f;
                        } else if (pin) {
                            h = inner_value(du, q1, q2, 0);
                        }
                    }
                    part[p] += h;
                }
                if (LM) {
                    if (farD) {
                        partd[p] = fmaf(q1.x * du, r, partd[p]);
                    } else {
                        bool pin;
                        float d = core_disp(du, q2, c, xfar, pin);
                        if (mind <= KId) {
                            if (DEFER) {
                                queue_inner(__ballot(pin), j, p, 1, lane, qn, qj, qi);
                                if (pin) d = 0.0f;
                            } else if (pin) {
                                d = inner_value(du, q1, q2, 1);
                            }
                        }
                        partd[p] += d;
                    }
                }
            }
        }
#pragma unroll
        for (int p = 0; p < P; ++p) {
            kahan_add(acc[p], cmp[p], part[p]);
            if (LM) kahan_add(accd[p], cmpd[p], partd[p]);
        }
        if (DEFER && (qn > 64 - 4 * P || j >= jhi)) {
            // flush: tasks e (lanes 0-31) and e + 1 (lanes 32-63)
            const int h = lane >> 5, l = lane & 31;
            for (int e = 0; e < qn; e += 2) {
                const int e1 = min(e + 1, qn - 1);
                const int jt = h ? __builtin_amdgcn_readlane(qj, e1) : __builtin_amdgcn_readlane(qj, e);
                const int it = h ? __builtin_amdgcn_readlane(qi, e1) : __builtin_amdgcn_readlane(qi, e);
                const int pt = it & 15, at = (it >> 4) & 63, nt = (it >> 10) & 127, ty = it >> 17;
                if ((h == 0 || e + 1 < qn) && l < nt) {
                    const float4 q0 = ld(r0p, jt), q1 = ld(r1p, jt), q2 = ld(r2p, jt);
                    const int d = at + l;                              // lane of the point in slice pt
                    const float du = (float)(tw + d + 256 * pt - __float_as_int(q0.x)) - q0.y;
                    atomicAdd(&iv[(ty * P + pt) * 64 + d], inner_value(du, q1, q2, ty));
                }
            }
            qn = 0;
        }
    }
    if (DEFER) {
        __syncthreads();
#pragma unroll
        for (int p = 0; p < P; ++p) {
            kahan_add(acc[p], cmp[p], iv[p * 64 + lane]);
            if (LM) kahan_add(accd[p], cmpd[p], iv[(P + p) * 64 + lane]);
        }
        __syncthreads();                               // before the staging buffer is reused
    }
}

// EDGE: far wing with per-pair window test and pedestal.
template <int P, bool LM>
__device__ __forceinline__ void edge_range(crec r0p, crec r1p,
                                           int jlo, int jhi, int tl, float C2, int CiD, float CfD, float iC2,
                                           float (&acc)[P], float (&cmp)[P], float (&accd)[P],
                                           float (&cmpd)[P]) {
    for (int j0 = jlo; j0 < jhi; j0 += 16) {
        const int je = min(j0 + 16, jhi);
        float part[P], partd[P];
#pragma unroll
        for (int p = 0; p < P; ++p) { part[p] = 0.0f; partd[p] = 0.0f; }
        for (int j = j0; j < je; ++j) {
            const float4 q0 = ld(r0p, j), q1 = ld(r1p, j);
            const float base = (float)(tl - __float_as_int(q0.x));
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const float du = (base + (float)(256 * p)) - q0.y;
                const float d2 = du * du;
                const float r = irs_rcpf(d2 + q0.z);
                const float v = fmaf(q0.w, r, -q1.y * fmaf(-d2, iC2, 2.0f));
                part[p] += d2 <= C2 ? v : 0.0f;
                if (LM)
                    partd[p] += in_window(tl - __float_as_int(q0.x) + 256 * p, q0.y, CiD, CfD) ? q1.x * du * r
                                                                                                 : 0.0f;
            }
        }
#pragma unroll
        for (int p = 0; p < P; ++p) {
            kahan_add(acc[p], cmp[p], part[p]);
            if (LM) kahan_add(accd[p], cmpd[p], partd[p]);
        }
    }
}

// EDGE lines by local expansion.  An EDGE line lies about one cutoff from the
// tile (the expansion is used when every EDGE line is at least
// EDGE_EXPAND_MIN grid steps from every slice centre), so on a 64-point slice
// its far wing and pedestal are, to within 5 (32/d)^4 of its Lorentz value at
// slice distance d (2e-8 at d = 4000; 2e-11 for a 25 cm^-1 cutoff on a
// 0.001 cm^-1 grid), the cubic in the offset delta from the slice centre
//   A/((d+delta)^2 + g^2) - P (2 - (d+delta)^2/C^2)
//     = sum_k a_k delta^k,  a_0 = A/D - P (2 - d^2/C^2),  a_1 = -2dA/D^2 + 2Pd/C^2,
//       a_2 = A (3d^2 - g^2)/D^3 + P/C^2,  a_3 = -4dA (d^2 - g^2)/D^4     (D = d^2 + g^2)
// and the dispersive wing B (d+delta)/((d+delta)^2 + g^2) the cubic
//       b_0 = Bd/D,  b_1 = B (g^2 - d^2)/D^2,  b_2 = Bd (d^2 - 3g^2)/D^3,
//       b_3 = -B (d^4 - 6d^2g^2 + g^4)/D^4.
// Lanes take one line each: per slice of this wave, a line whose window covers
// the whole slice adds its coefficients to the lane's partial sums, a line
// whose window misses the slice adds nothing, and a line whose window boundary
// falls inside the slice is evaluated pair by pair with the per-point test (as
// in edge_range) in a wave-uniform loop over those lines.  The coefficient
// sums are reduced across the wave and every point evaluates its slice's cubic.
#define EDGE_EXPAND_MIN 4000
template <int P, bool LM>
__device__ __forceinline__ void edge_expand(const float4 *__restrict__ r0p, const float4 *__restrict__ r1p,
                                            int jl0, int jl1, int jr0, int jr1, int t, int tl, float C, float Clm,
                                            float C2, int CiD, float CfD, float iC2, float (&acc)[P], float (&cmp)[P],
                                            float (&accd)[P], float (&cmpd)[P]) {
    const int lane = t & 63;
    const int tw = uni(tl & ~63);
    const int nL = jl1 - jl0, ntot = nL + (jr1 - jr0);
    if (ntot <= 0) return;
    const float inV = C - 31.55f, outV = C + 31.55f, inD = Clm - 31.55f, outD = Clm + 31.55f;
    const float dl = (float)lane - 31.5f;
    // one slice at a time, so that only its eight coefficient sums are live
    // (the lines' two records are re-read per slice from cache)
#pragma unroll
    for (int p = 0; p < P; ++p) {
        float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f, b0 = 0.0f, b1 = 0.0f, b2 = 0.0f, b3 = 0.0f;
        for (int g0 = 0; g0 < ntot; g0 += 64) {
            const int v = g0 + lane;
            const bool ok = v < ntot;
            const int j = ok ? (v < nL ? jl0 + v : jr0 + (v - nL)) : (nL > 0 ? jl0 : jr0);
            const float4 q0 = r0p[j];
            const float4 q1 = r1p[j];
            const int cidx = __float_as_int(q0.x);
            const float A = ok ? q0.w : 0.0f, Pd = ok ? q1.y : 0.0f, B = ok ? q1.x : 0.0f, g2 = q0.z;
            const float d = (float)(tw + 256 * p - cidx) + (31.5f - q0.y);   // slice centre - line centre
            const float ad = fabsf(d);
            const bool fullV = ad <= inV, bndV = ok && !fullV && ad < outV;
            const float d2 = d * d;
            const float iD = irs_rcpf(d2 + g2);
            const float iD2 = iD * iD;
            if (fullV) {
                const float L = A * iD;
                a0 += L - Pd * fmaf(-d2, iC2, 2.0f);
                a1 += 2.0f * d * (Pd * iC2 - L * iD);
                a2 = fmaf(L * iD2, 3.0f * d2 - g2, fmaf(Pd, iC2, a2));
                a3 = fmaf(-4.0f * d * L * iD2 * iD, d2 - g2, a3);
            }
            bool bndD = false;
            if (LM) {
                const bool fullD = ad <= inD;
                bndD = ok && !fullD && ad < outD;
                if (fullD) {
                    const float Bi = B * iD;
                    b0 = fmaf(Bi, d, b0);
                    b1 = fmaf(Bi * iD, g2 - d2, b1);
                    b2 = fmaf(Bi * iD2 * d, d2 - 3.0f * g2, b2);
                    b3 = fmaf(-Bi * iD2 * iD, fmaf(d2, d2 - 6.0f * g2, g2 * g2), b3);
                }
            }
            // lines whose window boundary falls in this slice: pair by pair
            unsigned long long mV = __ballot(bndV), mD = LM ? __ballot(bndD) : 0ull;
            unsigned long long m = mV | mD;
            while (m) {
                const int l = __builtin_ctzll(m);
                m &= m - 1;
                const int jj = uni(__shfl(j, l, 64));
                const float4 e0 = r0p[jj], e1 = r1p[jj];
                const float du = ((float)(tl - __float_as_int(e0.x)) + (float)(256 * p)) - e0.y;
                const float dd2 = du * du;
                const float r = irs_rcpf(dd2 + e0.z);
                if ((mV >> l) & 1ull) {
                    const float w = fmaf(e0.w, r, -e1.y * fmaf(-dd2, iC2, 2.0f));
                    kahan_add(acc[p], cmp[p], dd2 <= C2 ? w : 0.0f);
                }
                if (LM && ((mD >> l) & 1ull)) {
                    const bool in = in_window(tl - __float_as_int(e0.x) + 256 * p, e0.y, CiD, CfD);
                    kahan_add(accd[p], cmpd[p], in ? e1.x * du * r : 0.0f);
                }
            }
        }
        // wave sums of the coefficients, then the cubic at each point
        for (int o = 32; o > 0; o >>= 1) {
            a0 += __shfl_xor(a0, o, 64);
            a1 += __shfl_xor(a1, o, 64);
            a2 += __shfl_xor(a2, o, 64);
            a3 += __shfl_xor(a3, o, 64);
        }
        kahan_add(acc[p], cmp[p], fmaf(fmaf(fmaf(a3, dl, a2), dl, a1), dl, a0));
        if (LM) {
            for (int o = 32; o > 0; o >>= 1) {
                b0 += __shfl_xor(b0, o, 64);
                b1 += __shfl_xor(b1, o, 64);
                b2 += __shfl_xor(b2, o, 64);
                b3 += __shfl_xor(b3, o, 64);
            }
            kahan_add(accd[p], cmpd[p], fmaf(fmaf(fmaf(b3, dl, b2), dl, b1), dl, b0));
        }
    }
}

// FAST lines at least FAST_EXPAND_R grid steps from the tile, by local
// expansion: on a 64-point slice at distance d >= FAST_EXPAND_R + 32 from the
// line centre, A/((d+delta)^2 + g^2) = A sum_k h_k delta^k with
//   h_0 = 1/D, h_1 = -2d h_0/D, h_k = -(2d h_{k-1} + h_{k-2})/D   (D = d^2 + g^2)
// (the series of 1/u for the quadratic u(delta)), and the dispersive wing
// B (d+delta)/u = B sum_k (d h_k + h_{k-1}) delta^k.  Degree 6: the first
// omitted term is below 8 (32/d)^7 <= 2.3e-8 of the line's value at d = 532.
// Lanes take one line each and add A h_k (and the dispersive coefficients) for
// one slice of the wave at a time; a wave reduction and one polynomial per
// point follow.  The lines' windows cover the tile (they are interior lines) and
// their pedestal is in the per-tile moments.
#define FAST_EXPAND_R 500
#define FAST_EXPAND_K 7
template <int P, bool LM>
__device__ __forceinline__ void fast_expand(const float4 *__restrict__ r0p, const float4 *__restrict__ r1p,
                                            int jl0, int jl1, int jr0, int jr1, int t, int tl, float (&acc)[P],
                                            float (&cmp)[P], float (&accd)[P], float (&cmpd)[P]) {
    const int lane = t & 63;
    const int tw = uni(tl & ~63);
    const int nL = max(jl1 - jl0, 0), ntot = nL + max(jr1 - jr0, 0);
    if (ntot <= 0) return;
    const float dl = (float)lane - 31.5f;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        float a[FAST_EXPAND_K], b[FAST_EXPAND_K];
#pragma unroll
        for (int k = 0; k < FAST_EXPAND_K; ++k) a[k] = b[k] = 0.0f;
        for (int g0 = 0; g0 < ntot; g0 += 64) {
            const int v = g0 + lane;
            const bool ok = v < ntot;
            const int j = ok ? (v < nL ? jl0 + v : jr0 + (v - nL)) : (nL > 0 ? jl0 : jr0);
            const float4 q0 = r0p[j];
            const float A = ok ? q0.w : 0.0f;
            const float B = (LM && ok) ? r1p[j].x : 0.0f;
            const float d = (float)(tw + 256 * p - __float_as_int(q0.x)) + (31.5f - q0.y);   // slice centre - line
            const float iD = irs_rcpf(fmaf(d, d, q0.z));
            const float d2 = -2.0f * d;
            float hm = 0.0f, h = iD;                                 // h_{k-1}, h_k
#pragma unroll
            for (int k = 0; k < FAST_EXPAND_K; ++k) {
                a[k] = fmaf(A, h, a[k]);
                if (LM) b[k] = fmaf(B, fmaf(d, h, hm), b[k]);
                const float hn = fmaf(d2, h, -hm) * iD;
                hm = h;
                h = hn;
            }
        }
        float c = a[FAST_EXPAND_K - 1];
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
        float val = c;
#pragma unroll
        for (int k = FAST_EXPAND_K - 2; k >= 0; --k) {
            float ck = a[k];
            for (int o = 32; o > 0; o >>= 1) ck += __shfl_xor(ck, o, 64);
            val = fmaf(val, dl, ck);
        }
        kahan_add(acc[p], cmp[p], val);
        if (LM) {
            float e = b[FAST_EXPAND_K - 1];
            for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
            float vd = e;
#pragma unroll
            for (int k = FAST_EXPAND_K - 2; k >= 0; --k) {
                float ek = b[k];
                for (int o = 32; o > 0; o >>= 1) ek += __shfl_xor(ek, o, 64);
                vd = fmaf(vd, dl, ek);
            }
            kahan_add(accd[p], cmpd[p], vd);
        }
    }
}

// Degenerate small-cutoff case (window narrower than tile + core): exact
// branch, window test and pedestal for every pair.
template <int P, bool LM>
__device__ __forceinline__ void general_range(crec r0p, crec r1p, crec r2p, int jlo, int jhi, int tl,
                                              float xfar, float C2, int CiD, float CfD, float iC2, float (&acc)[P],
                                              float (&cmp)[P], float (&accd)[P], float (&cmpd)[P]) {
    for (int j = jlo; j < jhi; ++j) {
        const float4 q0 = ld(r0p, j);
        const float4 q1 = ld(r1p, j);
        const float4 q2 = ld(r2p, j);
        const float base = (float)(tl - __float_as_int(q0.x));
        const bool dline = LM && (q1.x != 0.0f || q2.w != 0.0f);
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const float du = (base + (float)(256 * p)) - q0.y;
            const float d2 = du * du;
            const float x = du * q1.z;
            float h;
            if (fabsf(x) > xfar) {
                h = q0.w * irs_rcpf(d2 + q0.z);
            } else {
                float wr, wi;
                w32(x, q1.w, wr, wi);
                h = q2.x * wr;
            }
            const float v = h - q1.y * fmaf(-d2, iC2, 2.0f);
            if (d2 <= C2) kahan_add(acc[p], cmp[p], v);
            if (dline) {
                const float xd = du * q2.y;
                float dd;
                if (fabsf(xd) > xfar) {
                    dd = q1.x * du * irs_rcpf(d2 + q0.z);
                } else {
                    float wr, wi;
                    w32(xd, q2.z, wr, wi);
                    dd = q2.w * wi;
                }
                if (in_window(tl - __float_as_int(q0.x) + 256 * p, q0.y, CiD, CfD)) kahan_add(accd[p], cmpd[p], dd);
            }
        }
    }
}

__device__ __forceinline__ double block_sum(double v, double *red, int t) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

// FAR: lines at least D grid steps from every coarse node of the tile's
// stencil span [Sl, Sh] (nodes Sl + n*HC, n = 0..NC-1) and whose windows cover
// the span.  Their far-wing sum (A/(du^2+g^2), and B du/(du^2+g^2) for VP_Y
// lines) is a smooth function of the grid position there: it is evaluated at
// the NC coarse nodes only and added to every point of the tile by four-point
// (cubic) Lagrange interpolation.  For a Lorentz wing at distance >= D the
// interpolation error is below 2.8 (HC/D)^4 of the line's contribution (7e-7 at
// D = 1.5 cm^-1, HC = 0.032 cm^-1); the pedestal of these lines stays in the
// per-tile moments.  The far lines stream through LDS in chunks of LS_CHUNK
// ({centre index relative to node 0, fraction, g^2, A} and B).  Thread t
// evaluates the four consecutive nodes 4g .. 4g+3 (g = t % NG, NG = ceil(NC/4))
// for the chunk's lines k = s, s+S, ... (s = t / NG, S = 256 / NG slices), so
// one LDS record and one index conversion serve four nodes; the offset to node
// 4g+m is formed by m FP32 additions of HC (relative error of du <= 4 ulp).  The S partial
// sums per node are combined in FP64 through the (then idle) staging buffer.
#define FAR_MAXNC 132
#define FAR_LDS (2 * FAR_MAXNC)

template <int P, bool LM>
__device__ __forceinline__ void far_coarse(const float4 *__restrict__ r0p, const float4 *__restrict__ r1p, int jl0,
                                           int jl1, int jr0, int jr1, int t, int64_t i0, int64_t Sl, int HC,
                                           int NC, float *s_far, float4 *s_rec, float (&acc)[P], float (&cmp)[P],
                                           float (&accd)[P], float (&cmpd)[P]) {
    const int NG = (NC + 3) >> 2;                      // node groups, <= 33
    const int S = IRS_LS_THREADS / NG;                 // >= 7
    const int g = t % NG, sl = t / NG;
    const bool act = t < S * NG;
    const int nL = jl1 - jl0, ntot = nL + (jr1 - jr0);
    const float xrel = (float)(4 * g * HC);           // offset of node 4g from Sl
    const float fhc = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int((float)HC)));
    float a[4], c[4], ad[4], cd[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) a[m] = c[m] = ad[m] = cd[m] = 0.0f;
    for (int base = 0; base < ntot; base += LS_CHUNK) {
        const int cnt = min(LS_CHUNK, ntot - base);
        // stage the chunk: {centre index - Sl, fraction, g^2, A} and B; the index
        // difference is below 2^24 in magnitude, so it is exact as a float
#pragma unroll
        for (int u = 0; u < LS_PER_THREAD; ++u) {
            const int k = t + u * IRS_LS_THREADS;
            if (k < cnt) {
                const int v = base + k;
                const int j = v < nL ? jl0 + v : jr0 + (v - nL);
                const float4 q = r0p[j];
                s_rec[k] = make_float4((float)(__float_as_int(q.x) - (int)Sl), q.y, q.z, q.w);
                if (LM) s_rec[LS_CHUNK + k].x = r1p[j].x;
            }
        }
        __syncthreads();
        if (act) {
            float part[4], partd[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) part[m] = partd[m] = 0.0f;
            int n = 0;
            for (int k = sl; k < cnt; k += S) {
                const float4 q = s_rec[k];
                const float b = LM ? s_rec[LS_CHUNK + k].x : 0.0f;
                float du = (xrel - q.x) - q.y;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    if (m) du += fhc;
                    const float r = irs_rcpf(fmaf(du, du, q.z));
                    part[m] = fmaf(q.w, r, part[m]);
                    if (LM) partd[m] = fmaf(b * du, r, partd[m]);
                }
                if (++n == 32) {
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        kahan_add(a[m], c[m], part[m]);
                        if (LM) kahan_add(ad[m], cd[m], partd[m]);
                        part[m] = partd[m] = 0.0f;
                    }
                    n = 0;
                }
            }
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                kahan_add(a[m], c[m], part[m]);
                if (LM) kahan_add(ad[m], cd[m], partd[m]);
            }
        }
        __syncthreads();
    }
    // per-(slice, node) partial sums in the staging buffer: [S][4 NG] twice (<= 2048 floats)
    float *red = (float *)s_rec, *redd = red + S * 4 * NG;
    if (act) {
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            red[sl * 4 * NG + 4 * g + m] = a[m] - c[m];
            redd[sl * 4 * NG + 4 * g + m] = ad[m] - cd[m];
        }
    }
    __syncthreads();
    float *F = s_far, *Fd = s_far + FAR_MAXNC;
    if (t < NC) {
        double v = 0.0, vd = 0.0;
        for (int q = 0; q < S; ++q) {
            v += (double)red[q * 4 * NG + t];
            vd += (double)redd[q * 4 * NG + t];
        }
        F[t] = (float)v;
        Fd[t] = (float)vd;
    }
    __syncthreads();
    const float ihc = 1.0f / (float)HC;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int off = (int)(i0 - Sl) + t + 256 * p;          // grid offset from node 0
        const int n0 = off / HC;                               // node left of the point, >= 1
        const float u = (float)(off - n0 * HC) * ihc;          // in [0, 1)
        const float wm = -u * (u - 1.0f) * (u - 2.0f) * (1.0f / 6.0f);
        const float w0 = (u + 1.0f) * (u - 1.0f) * (u - 2.0f) * 0.5f;
        const float w1 = -(u + 1.0f) * u * (u - 2.0f) * 0.5f;
        const float w2 = (u + 1.0f) * u * (u - 1.0f) * (1.0f / 6.0f);
        const float v = wm * F[n0 - 1] + w0 * F[n0] + w1 * F[n0 + 1] + w2 * F[n0 + 2];
        kahan_add(acc[p], cmp[p], v);
        if (LM) {
            const float vd = wm * Fd[n0 - 1] + w0 * Fd[n0] + w1 * Fd[n0 + 1] + w2 * Fd[n0 + 2];
            kahan_add(accd[p], cmpd[p], vd);
        }
    }
    __syncthreads();
}

template <int P, bool LM>
__device__ __forceinline__ void species_tile(const LinesumArgs &a, const ItemDev &it, int t, int64_t i0,
                                             int64_t i1, float (&tau)[P], int *s_rng, double *s_red,
                                             float4 *s_rec, float *s_far) {
    const int32_t *keys = a.keys + it.key_off;
    const int n = it.n;
    const double dnu = a.dnu;
    const double Cd = it.cutoff / dnu;
    const double Cmx = LM ? fmax((double)it.C, (double)it.Clm) : (double)it.C;
    const double Cmn = LM ? fmin((double)it.C, (double)it.Clm) : (double)it.C;
    const int64_t Cmax = (int64_t)ceil(Cmx), Cmin = (int64_t)floor(Cmn);
    const int64_t M = it.M;
    const double numax = a.nu_lo + (double)(i1 + Cmax + M) * dnu;
    const int64_t K = (int64_t)ceil(it.kcore_per_nu * numax) + 2;
    const int64_t vA = i0 - Cmax - M - 1;
    const int64_t vB = i1 - 1 - Cmin + M + 1;
    const int64_t vC = i0 - K - M - 1;
    const int64_t vD = i1 - 1 + K + M + 1;
    const int64_t vE = i0 + Cmin - M - 2;
    const int64_t vF = i1 - 1 + Cmax + M + 1;
    const bool nondeg = vA <= vB && vB <= vC && vC <= vD && vD <= vE && vE <= vF;
    // coarse far-wing split (see far_coarse); HC divides the tile width
    const int HC = a.far_hc;
    const int NC = HC > 0 ? (IRS_LS_THREADS * P) / HC + 3 : 0;
    const int64_t Sl = i0 - HC, Sh = Sl + (int64_t)(NC - 1) * HC, Dd = a.far_d;
    const int64_t vB2 = Sh - Cmin + M + 1;
    const int64_t vG = Sl - Dd - M - 1;
    const int64_t vH = Sh + Dd + M;
    const int64_t vE2 = Sl + Cmin - M - 2;
    const bool far = nondeg && HC > 0 && NC <= FAR_MAXNC && vA <= vB2 && vB2 <= vG && vG < vC && vD < vH &&
                     vH <= vE2 && vE2 <= vF;
    // FAST lines beyond FAST_EXPAND_R of the tile go to fast_expand: keys <= vL
    // (centre <= i0 - FAST_EXPAND_R - 3) and keys >= vR (centre >= i1 + FAST_EXPAND_R + 2)
    const int64_t vL = i0 - FAST_EXPAND_R - M - 2;
    const int64_t vR = i1 - 1 + FAST_EXPAND_R + M + 1;
    if (t < 12) {
        int r;
        switch (t) {
            case 0: r = first_ge(a, it, keys, n, vA); break;
            case 1: r = first_ge(a, it, keys, n, vB); break;
            case 2: r = first_ge(a, it, keys, n, vC); break;
            case 3: r = first_ge(a, it, keys, n, vD + 1); break;      // first key > vD
            case 4: r = first_ge(a, it, keys, n, vE + 1); break;
            case 5: r = first_ge(a, it, keys, n, vF + 1); break;
            case 6: r = far ? first_ge(a, it, keys, n, vB2) : 0; break;
            case 7: r = far ? first_ge(a, it, keys, n, vG + 1) : 0; break;
            case 8: r = far ? first_ge(a, it, keys, n, vH) : 0; break;
            case 9: r = far ? first_ge(a, it, keys, n, vE2 + 1) : 0; break;
            case 10: r = first_ge(a, it, keys, n, vL + 1); break;
            default: r = first_ge(a, it, keys, n, vR); break;
        }
        s_rng[t] = r;
    }
    __syncthreads();
    const int jA = uni(s_rng[0]), jC = uni(s_rng[2]);
    const int jD = uni(s_rng[3]), jF = uni(s_rng[5]);
    // interior ranges: [jB, jE) without the far split, [jB2, jE2) with it
    const int jB = far ? uni(s_rng[6]) : uni(s_rng[1]);
    const int jE = far ? uni(s_rng[9]) : uni(s_rng[4]);
    const int jG = far ? uni(s_rng[7]) : jB;
    const int jH = far ? uni(s_rng[8]) : jE;
    // FAST ranges [jG, jC) and [jD, jH): far parts [jG, jLf) and [jRf, jH) by expansion.
    // far_hc == 0 (far_wing_cm = 0) turns every shortcut off: every pair on the grid.
    const bool shortcuts = HC > 0;
    // (only when there are enough of them: the per-slice reductions cost about
    // as much as 128 lines evaluated pair by pair)
#define FAST_EXPAND_MINLINES 128
    int jLf = shortcuts ? min(max(uni(s_rng[10]), jG), jC) : jG;
    int jRf = shortcuts ? max(min(uni(s_rng[11]), jH), jD) : jH;
    if ((jLf - jG) + (jH - jRf) < FAST_EXPAND_MINLINES) { jLf = jG; jRf = jH; }
    __syncthreads();

    const float4 *r0p = a.rec0 + it.rec_off;
    const float4 *r1p = a.rec1 + it.rec_off;
    const float4 *r2p = a.rec2 + it.rec_off;
    const int64_t imid = i0 + 128 * P;
    const int tl = (int)(i0 + t);
    const float C2 = (float)((double)it.C * (double)it.C);
    // the dispersive window Clm = lm_cutoff / dnu as integer part + FP32 fraction (in_window)
    const double Cld = fmin(it.lm_cutoff / dnu, 1.0e9);
    const int CiD = (int)floor(Cld);
    const float CfD = (float)(Cld - (double)CiD);
    const double iC2d = 1.0 / (Cd * Cd);
    const float iC2 = (float)iC2d;
    const float xfar = (float)it.x_far;

    float acc[P], cmp[P], accd[P], cmpd[P];
#pragma unroll
    for (int p = 0; p < P; ++p) { acc[p] = cmp[p] = accd[p] = cmpd[p] = 0.0f; }

    double q0 = 0.0, q1 = 0.0, q2 = 0.0;
    if (nondeg) {
        // pedestal moments of the interior (FAR + FAST + CORE) lines about the
        // tile centre, from the FP64 prefix sums of P, P c, P c^2
        if (jE > jB) {
            const double *P0 = a.pre + it.rec_off, *P1 = P0 + a.n_rec, *P2 = P1 + a.n_rec;
            const double S0 = P0[jE - 1] - (jB > 0 ? P0[jB - 1] : 0.0);
            const double S1 = P1[jE - 1] - (jB > 0 ? P1[jB - 1] : 0.0);
            const double S2 = P2[jE - 1] - (jB > 0 ? P2[jB - 1] : 0.0);
            const double m = (double)imid;
            q0 = S0;
            q1 = S1 - m * S0;
            q2 = S2 - 2.0 * m * S1 + m * m * S0;
        }
        if (far && (jG > jB || jE > jH))
            far_coarse<P, LM>(r0p, r1p, jB, jG, jH, jE, t, i0, Sl, HC, NC, s_far, s_rec, acc, cmp, accd, cmpd);
        fast_expand<P, LM>(r0p, r1p, jG, jLf, jRf, jH, t, tl, acc, cmp, accd, cmpd);
        fast_range<P, LM>(r0p, r1p, jLf, jC, t, (int)i0, s_rec, acc, cmp, accd, cmpd);
        core_range<P, LM>(as_crec(r0p), as_crec(r1p), as_crec(r2p), as_crec(a.rec3 + 6 * it.rec_off), jC, jD, tl,
                          xfar, (float *)s_rec, acc, cmp, accd, cmpd);
        fast_range<P, LM>(r0p, r1p, jD, jRf, t, (int)i0, s_rec, acc, cmp, accd, cmpd);
        // EDGE lines: local expansion when every one of them is far from the tile
        const bool expand = shortcuts && Cmn - (double)(256 * P + HC + M + 34) >= (double)EDGE_EXPAND_MIN;
        if (expand) {
            edge_expand<P, LM>(r0p, r1p, jA, jB, jE, jF, t, tl, it.C, it.Clm, C2, CiD, CfD, iC2, acc, cmp, accd,
                               cmpd);
        } else {
            edge_range<P, LM>(as_crec(r0p), as_crec(r1p), jA, jB, tl, C2, CiD, CfD, iC2, acc, cmp, accd, cmpd);
            edge_range<P, LM>(as_crec(r0p), as_crec(r1p), jE, jF, tl, C2, CiD, CfD, iC2, acc, cmp, accd, cmpd);
        }
        if (a.work && t == 0) {
            unsigned long long *w = a.work + 8 * (blockIdx.x % IRS_CNT_SLOTS);
            const unsigned long long edge = (unsigned long long)((jB - jA) + (jF - jE));
            const unsigned long long fexp = (unsigned long long)((jLf - jG) + (jH - jRf));
            const unsigned long long slices = (unsigned long long)((i1 - i0 + 63) / 64);
            const unsigned long long fine =
                (unsigned long long)((jC - jLf) + (jRf - jD) + (jD - jC)) * (unsigned long long)(i1 - i0) +
                (expand ? 0ull : edge * (unsigned long long)(i1 - i0));
            const unsigned long long coarse = far ? (unsigned long long)((jG - jB) + (jE - jH)) * NC : 0ull;
            const unsigned long long expd = ((expand ? edge : 0ull) + fexp) * slices;
            if (expd) atomicAdd(&w[0], expd);
            atomicAdd(&w[1], fine);
            if (coarse) atomicAdd(&w[2], coarse);
            if (LM) {
                atomicAdd(&w[3], fine);
                if (coarse) atomicAdd(&w[4], coarse);
            }
        }
    } else {
        general_range<P, LM>(as_crec(r0p), as_crec(r1p), as_crec(r2p), jA, jF, tl, xfar, C2, CiD, CfD, iC2, acc, cmp, accd, cmpd);
        if (a.work && t == 0) {
            unsigned long long *w = a.work + 8 * (blockIdx.x % IRS_CNT_SLOTS);
            atomicAdd(&w[1], (unsigned long long)(jF - jA) * (unsigned long long)(i1 - i0));
            if (LM) atomicAdd(&w[3], (unsigned long long)(jF - jA) * (unsigned long long)(i1 - i0));
        }
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const double v = (double)(i0 + t + 256 * p - imid);
        const double ped = (2.0 - v * v * iC2d) * q0 + 2.0 * v * iC2d * q1 - iC2d * q2;
        double sv = ((double)acc[p] - (double)cmp[p]) - ped;
        sv = sv > 0.0 ? sv : 0.0;
        if (LM) {
            sv += (double)accd[p] - (double)cmpd[p];
            sv = sv > 0.0 ? sv : 0.0;
        }
        tau[p] += (float)sv;
    }
}

// One work unit: the tile `tile` of (profile, layer), all species, written to
// out[profile][layer][i0:i1).
template <int P>
__device__ __forceinline__ void linesum_unit(const LinesumArgs &a, int tile, int layer, int prof, int *s_rng,
                                             double *s_red, float4 *s_rec, float *s_far) {
    const int t = threadIdx.x;
    const int64_t i0 = (int64_t)tile * (IRS_LS_THREADS * P);
    const int64_t i1 = min(i0 + (int64_t)(IRS_LS_THREADS * P), a.n_hi);
    float tau[P];
#pragma unroll
    for (int p = 0; p < P; ++p) tau[p] = 0.0f;
    const ItemDev *items = a.items + ((int64_t)prof * a.n_layers + layer) * a.n_species;
    for (int s = 0; s < a.n_species; ++s) {
        const ItemDev it = items[s];
        if (!it.active || it.n == 0) continue;
        if (it.lm) species_tile<P, true>(a, it, t, i0, i1, tau, s_rng, s_red, s_rec, s_far);
        else species_tile<P, false>(a, it, t, i0, i1, tau, s_rng, s_red, s_rec, s_far);
    }
    float *out = a.out + ((int64_t)prof * a.n_layers + layer) * a.n_hi;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int64_t i = i0 + t + 256 * p;
        if (i < a.n_hi) out[i] = tau[p];
    }
}

// One workgroup per unit: blockIdx = (tile rank, layer, profile).
// Occupancy: the register budget is capped for 32 / P resident workgroups per CU
// (P = 4: eight waves per SIMD, 64 VGPRs, with spills; the LDS footprint of a
// workgroup, ~17.5 KB, still allows eight per CU).  The extra waves hide the
// latency of the scalar record loads, LDS reads and spill traffic: on full IASI
// 8 / 6 / 5 / 4 waves per SIMD measured 26.9 / 28.9 / 31.2 / 35.3 ms.
#define IRS_LS_MINB(P) (32 / (P))
template <int P>
__global__ void __launch_bounds__(IRS_LS_THREADS, IRS_LS_MINB(P)) linesum_kernel(const LinesumArgs a) {
    __shared__ int s_rng[12];
    __shared__ double s_red[12];
    __shared__ float4 s_rec[2 * LS_CHUNK];
    __shared__ float s_far[FAR_LDS];
    const int tile = a.tile_order ? a.tile_order[blockIdx.x] : (int)blockIdx.x;
    linesum_unit<P>(a, tile, blockIdx.y, blockIdx.z, s_rng, s_red, s_rec, s_far);
}

// Persistent form: a grid sized to the resident capacity of the device pulls
// units from a queue counter profile by profile, and within a profile in order
// of decreasing tile cost (tile_order, heaviest tiles first, every layer of a
// tile together).  The resident units then span some tens of tiles of each of
// one profile's layers, whose neighbouring tiles read the same line records
// through L2; a tile-major order across profiles would spread them over two
// tiles of every (profile, layer) item with no such reuse.  The unit index is
// fetched by one lane and broadcast through LDS, so the loop exit is
// block-uniform and every wave of every block reaches it once the queue is
// past the last unit.  The counter must be zero at launch.
template <int P>
__global__ void __launch_bounds__(IRS_LS_THREADS, IRS_LS_MINB(P)) linesum_persistent_kernel(const LinesumArgs a) {
    __shared__ int s_rng[12];
    __shared__ double s_red[12];
    __shared__ float4 s_rec[2 * LS_CHUNK];
    __shared__ float s_far[FAR_LDS];
    __shared__ unsigned s_unit;
    const unsigned per_prof = (unsigned)a.tiles * (unsigned)a.n_layers;
    const unsigned total = per_prof * (unsigned)a.n_prof;
    for (;;) {
        if (threadIdx.x == 0) s_unit = atomicAdd(a.queue, 1u);
        __syncthreads();
        const unsigned u = s_unit;
        __syncthreads();
        if (u >= total) break;
        const unsigned prof = u / per_prof, rem = u - prof * per_prof;
        const unsigned rank = rem / (unsigned)a.n_layers;
        const int tile = a.tile_order ? a.tile_order[rank] : (int)rank;
        const int layer = (int)(rem - rank * (unsigned)a.n_layers);
        linesum_unit<P>(a, tile, layer, (int)prof, s_rng, s_red, s_rec, s_far);
    }
}

// Zeroes the queue counter and the prep work counters (a kernel rather than a
// memset node, so the sequence can be captured in a graph).
__global__ void reset_counters_kernel(unsigned *queue, unsigned long long *counters, int n_counters) {
    const int t = threadIdx.x;
    if (t == 0 && queue) *queue = 0u;
    if (counters)
        for (int i = t; i < n_counters; i += 64) counters[i] = 0ull;
}

hipError_t launch_reset_counters(unsigned *queue, unsigned long long *counters, int n_counters, hipStream_t s) {
    hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, s, queue, counters, n_counters);
    return hipGetLastError();
}

template <int P>
static hipError_t launch_ls_p(const LinesumArgs &a, int persistent_blocks, hipStream_t s) {
    if (persistent_blocks > 0)
        hipLaunchKernelGGL(linesum_persistent_kernel<P>, dim3(persistent_blocks), dim3(IRS_LS_THREADS), 0, s, a);
    else
        hipLaunchKernelGGL(linesum_kernel<P>, dim3(a.tiles, a.n_layers, a.n_prof), dim3(IRS_LS_THREADS), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_linesum(const LinesumArgs &a, int P, int persistent_blocks, hipStream_t s) {
    if (persistent_blocks > 0 && !a.queue) return hipErrorInvalidValue;
    if (!a.rec3 || !a.pre) return hipErrorInvalidValue;
    if (P == 16) return launch_ls_p<16>(a, persistent_blocks, s);
    if (P == 8) return launch_ls_p<8>(a, persistent_blocks, s);
    return launch_ls_p<4>(a, persistent_blocks, s);
}

// Resident workgroups of the persistent kernel on the current device.
int linesum_persistent_capacity(int P) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    hipError_t e;
    if (P == 16) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, linesum_persistent_kernel<16>, IRS_LS_THREADS, 0);
    else if (P == 8) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, linesum_persistent_kernel<8>, IRS_LS_THREADS, 0);
    else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, linesum_persistent_kernel<4>, IRS_LS_THREADS, 0);
    if (e != hipSuccess) return 0;
    return (per_cu > 1 ? per_cu : 1) * cus;
}
