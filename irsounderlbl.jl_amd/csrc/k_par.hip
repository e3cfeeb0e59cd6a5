// HITRAN .par ingest on the device (parse_hitran_par / load_hitran_par,
// src/HITRAN/linelist.jl:90-103, src/HITRAN/fetch.jl:12-61): the file's bytes
// are scanned for record boundaries, one thread parses one record's fixed-width
// fields, the kept records are compacted in file order and sorted stably by
// wavenumber (HITRANLinelist, linelist.jl:51-56).
//
// Fields (1-based inclusive columns of the reference): mol 1:2 (Int8), iso 3:3
// (Int8), nu 4:15, S 16:25, A 26:35 (Float64), gamma_air 36:40, gamma_self
// 41:45 (Float32), E'' 46:55 (Float64), n_air 56:59, delta_air 60:67 (Float32).
// A record shorter than 67 characters, or one whose fields do not parse, is
// skipped, as the reference's try/catch does.  Numbers are read with the
// grammar [space][sign]digits[.digits][(e|E)[sign]digits][space] and rounded
// to nearest-even exactly (multi-word integer arithmetic), the result strtod /
// strtof give; a field outside that grammar or outside the exact range
// (|decimal exponent| > 60 for Float64, outside the normal range for Float32)
// is handed to the host parser (status PAR_HOST), which applies the full rules.
#include <stdint.h>

#include "irs_internal.h"

#define PAR_SKIP 0
#define PAR_OK 1
#define PAR_HOST 2

namespace {

__device__ __forceinline__ bool is_space(unsigned char c) {
    return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}

// little-endian multi-word unsigned integers
constexpr int NW = 6;
struct Big {
    uint64_t w[NW];
};

__device__ __forceinline__ void big_set(Big &a, uint64_t v) {
#pragma unroll
    for (int i = 0; i < NW; ++i) a.w[i] = 0;
    a.w[0] = v;
}

__device__ __forceinline__ void big_mul_small(Big &a, uint64_t m) {
    uint64_t carry = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        const __uint128_t p = (__uint128_t)a.w[i] * m + carry;
        a.w[i] = (uint64_t)p;
        carry = (uint64_t)(p >> 64);
    }
}

__device__ __forceinline__ int big_bitlen(const Big &a) {
    for (int i = NW - 1; i >= 0; --i)
        if (a.w[i]) return 64 * i + 64 - __clzll((long long)a.w[i]);
    return 0;
}

__device__ __forceinline__ void big_shl(Big &a, int s) {
    const int ws = s / 64, bs = s % 64;
    for (int i = NW - 1; i >= 0; --i) {
        uint64_t v = 0;
        const int j = i - ws;
        if (j >= 0) {
            v = a.w[j] << bs;
            if (bs && j - 1 >= 0) v |= a.w[j - 1] >> (64 - bs);
        }
        a.w[i] = v;
    }
}

__device__ __forceinline__ void big_shr1(Big &a) {
#pragma unroll
    for (int i = 0; i < NW; ++i) a.w[i] = (a.w[i] >> 1) | (i + 1 < NW ? a.w[i + 1] << 63 : 0);
}

__device__ __forceinline__ bool big_ge(const Big &a, const Big &b) {
    for (int i = NW - 1; i >= 0; --i)
        if (a.w[i] != b.w[i]) return a.w[i] > b.w[i];
    return true;
}

__device__ __forceinline__ void big_sub(Big &a, const Big &b) {
    uint64_t borrow = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        const uint64_t x = a.w[i], y = b.w[i];
        const uint64_t d = x - y - borrow;
        borrow = (x < y) || (x - y < borrow) ? 1 : 0;
        a.w[i] = d;
    }
}

__device__ __forceinline__ bool big_zero(const Big &a) {
    uint64_t o = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) o |= a.w[i];
    return o == 0;
}

// a >> s (result < 2^64 assumed)
__device__ __forceinline__ uint64_t big_shr_u64(const Big &a, int s) {
    const int ws = s / 64, bs = s % 64;
    uint64_t v = a.w[ws] >> bs;
    if (bs && ws + 1 < NW) v |= a.w[ws + 1] << (64 - bs);
    return v;
}

__device__ __forceinline__ bool big_any_below(const Big &a, int s) {   // any of bits [0, s) set
    const int ws = s / 64, bs = s % 64;
    for (int i = 0; i < ws; ++i)
        if (a.w[i]) return true;
    return bs ? (a.w[ws] & ((1ull << bs) - 1)) != 0 : false;
}

// m * 10^e10 rounded to nearest-even with P significand bits: q * 2^ex,
// q in [2^(P-1), 2^P) (m > 0).  Needs |e10| <= 60 and m < 10^19.
__device__ void dec_to_bin(uint64_t m, int e10, int P, uint64_t &q_out, int &ex_out) {
    Big N;
    big_set(N, m);
    if (e10 >= 0) {
        for (int i = 0; i < e10; ++i) big_mul_small(N, 10);
        const int b = big_bitlen(N);
        if (b <= P) {
            q_out = N.w[0] << (P - b);
            ex_out = b - P;
            return;
        }
        const int sh = b - P;
        uint64_t q = big_shr_u64(N, sh);
        const bool guard = (big_shr_u64(N, sh - 1) & 1ull) != 0;
        const bool sticky = big_any_below(N, sh - 1);
        if (guard && (sticky || (q & 1ull))) ++q;
        int ex = sh;
        if (q == (1ull << P)) {
            q >>= 1;
            ++ex;
        }
        q_out = q;
        ex_out = ex;
        return;
    }
    Big D;
    big_set(D, 1);
    for (int i = 0; i < -e10; ++i) big_mul_small(D, 10);
    const int bd = big_bitlen(D), bm = 64 - __clzll((long long)m);
    // numerator m 2^k with a quotient of P + 2 or P + 3 bits
    const int k = P + 2 + bd - bm;
    big_shl(N, k);
    const int qbits = big_bitlen(N) - bd + 1;
    Big Ds = D;
    big_shl(Ds, qbits - 1);
    uint64_t Q = 0;
    for (int i = qbits - 1; i >= 0; --i) {
        if (big_ge(N, Ds)) {
            big_sub(N, Ds);
            Q |= 1ull << i;
        }
        big_shr1(Ds);
    }
    const int b = 64 - __clzll((long long)Q);
    const int sh = b - P;
    uint64_t q = Q >> sh;
    const bool guard = ((Q >> (sh - 1)) & 1ull) != 0;
    const bool sticky = (Q & ((1ull << (sh - 1)) - 1)) != 0 || !big_zero(N);
    if (guard && (sticky || (q & 1ull))) ++q;
    int ex = sh - k;
    if (q == (1ull << P)) {
        q >>= 1;
        ++ex;
    }
    q_out = q;
    ex_out = ex;
}

// returns PAR_OK with the value, PAR_SKIP (not a number) or PAR_HOST
template <typename T>
__device__ int parse_real(const unsigned char *s, int w, T &out) {
    int i = 0;
    while (i < w && is_space(s[i])) ++i;
    bool neg = false;
    if (i < w && (s[i] == '+' || s[i] == '-')) neg = s[i++] == '-';
    uint64_t m = 0;
    int nd = 0, e10 = 0, ndig = 0;
    bool any = false;
    while (i < w && s[i] >= '0' && s[i] <= '9') {
        any = true;
        if (m || s[i] != '0') {
            if (nd >= 19) return PAR_HOST;
            m = m * 10 + (s[i] - '0');
            ++nd;
        }
        ++i;
    }
    if (i < w && s[i] == '.') {
        ++i;
        while (i < w && s[i] >= '0' && s[i] <= '9') {
            any = true;
            if (m || s[i] != '0') {
                if (nd >= 19) return PAR_HOST;
                m = m * 10 + (s[i] - '0');
                ++nd;
            }
            --e10;
            ++i;
        }
    }
    if (!any) return PAR_HOST;                 // "inf", "nan", "0x...", "" and the like: the host decides
    if (i < w && (s[i] == 'e' || s[i] == 'E')) {
        ++i;
        bool eneg = false;
        if (i < w && (s[i] == '+' || s[i] == '-')) eneg = s[i++] == '-';
        if (!(i < w && s[i] >= '0' && s[i] <= '9')) return PAR_HOST;
        int ev = 0;
        while (i < w && s[i] >= '0' && s[i] <= '9') {
            if (ev < 100000) ev = ev * 10 + (s[i] - '0');
            ++i;
        }
        e10 += eneg ? -ev : ev;
    }
    while (i < w && is_space(s[i])) ++i;
    if (i != w) return PAR_HOST;
    (void)ndig;
    if (m == 0) {
        out = neg ? (T)-0.0 : (T)0.0;
        return PAR_OK;
    }
    const bool f32 = sizeof(T) == 4;
    if (f32) {
        // normal float range for every m < 10^nd: 10^e10 >= 1e-37 and 10^(nd + e10) <= 1e38
        if (e10 < -37 || nd + e10 > 38 || e10 > 60) return PAR_HOST;
    } else if (e10 < -60 || e10 > 60) {
        return PAR_HOST;
    }
    uint64_t q;
    int ex;
    dec_to_bin(m, e10, f32 ? 24 : 53, q, ex);
    const double v = ldexp((double)q, ex);     // exact: q has at most 53 bits, the result is normal
    out = (T)(neg ? -v : v);
    return PAR_OK;
}

__device__ int parse_int8(const unsigned char *s, int w, int8_t &out) {
    int i = 0;
    while (i < w && is_space(s[i])) ++i;
    bool neg = false;
    if (i < w && (s[i] == '+' || s[i] == '-')) neg = s[i++] == '-';
    int v = 0;
    bool any = false;
    while (i < w && s[i] >= '0' && s[i] <= '9') {
        any = true;
        v = v * 10 + (s[i] - '0');
        if (v > 1000) v = 1000;
        ++i;
    }
    while (i < w && is_space(s[i])) ++i;
    if (!any || i != w) return PAR_HOST;
    if (neg) v = -v;
    if (v < -128 || v > 127) return PAR_SKIP;   // the reference's parse throws: record skipped
    out = (int8_t)v;
    return PAR_OK;
}

}  // namespace

// newline counts per segment of PAR_SEG bytes
#define PAR_SEG 64
__global__ void __launch_bounds__(256) par_count_kernel(const unsigned char *b, int64_t n, int64_t *blk) {
    __shared__ int s[256];
    const int64_t seg = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int c = 0;
    const int64_t o = seg * PAR_SEG;
    for (int i = 0; i < PAR_SEG; ++i)
        if (o + i < n && b[o + i] == '\n') ++c;
    s[threadIdx.x] = c;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) blk[blockIdx.x] = s[0];
}

// newline positions in order: blk_off = exclusive scan of the block counts
__global__ void __launch_bounds__(256) par_newlines_kernel(const unsigned char *b, int64_t n, const int64_t *blk_off,
                                                           int64_t *nl) {
    __shared__ int s[256];
    const int64_t seg = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t o = seg * PAR_SEG;
    int c = 0;
    for (int i = 0; i < PAR_SEG; ++i)
        if (o + i < n && b[o + i] == '\n') ++c;
    s[threadIdx.x] = c;
    __syncthreads();
    // inclusive Hillis-Steele scan (256 entries)
    for (int st = 1; st < 256; st <<= 1) {
        const int v = threadIdx.x >= st ? s[threadIdx.x - st] : 0;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    int64_t p = blk_off[blockIdx.x] + s[threadIdx.x] - c;
    for (int i = 0; i < PAR_SEG; ++i)
        if (o + i < n && b[o + i] == '\n') nl[p++] = o + i;
}

// one thread per record r: [start, end) = after newline r-1 .. newline r (or the end)
__global__ void __launch_bounds__(256) par_parse_kernel(const unsigned char *b, int64_t n, const int64_t *nl,
                                                        int64_t n_nl, int64_t n_rec, ParFields f, uint8_t *status,
                                                        unsigned long long *n_host) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= n_rec) return;
    const int64_t start = r == 0 ? 0 : nl[r - 1] + 1;
    int64_t end = r < n_nl ? nl[r] : n;
    if (end > start && b[end - 1] == '\r') --end;
    uint8_t st = PAR_SKIP;
    if (end - start >= 67) {
        const unsigned char *s = b + start;
        bool ascii = true;
        for (int i = 0; i < 67; ++i) ascii = ascii && s[i] < 0x80;
        if (ascii) {
            int8_t mol = 0, iso = 0;
            double nu = 0, S = 0, A = 0, E = 0;
            float ga = 0, gs = 0, na = 0, da = 0;
            int rs[10];
            rs[0] = parse_int8(s + 0, 2, mol);
            rs[1] = parse_int8(s + 2, 1, iso);
            rs[2] = parse_real<double>(s + 3, 12, nu);
            rs[3] = parse_real<double>(s + 15, 10, S);
            rs[4] = parse_real<double>(s + 25, 10, A);
            rs[5] = parse_real<float>(s + 35, 5, ga);
            rs[6] = parse_real<float>(s + 40, 5, gs);
            rs[7] = parse_real<double>(s + 45, 10, E);
            rs[8] = parse_real<float>(s + 55, 4, na);
            rs[9] = parse_real<float>(s + 59, 8, da);
            bool skip = false, host = false;
            for (int i = 0; i < 10; ++i) {
                skip = skip || rs[i] == PAR_SKIP;
                host = host || rs[i] == PAR_HOST;
            }
            if (skip) {
                st = PAR_SKIP;
            } else if (host) {
                st = PAR_HOST;
                atomicAdd(n_host, 1ull);
            } else {
                st = PAR_OK;
                f.mol[r] = mol;
                f.iso[r] = iso;
                f.nu[r] = nu;
                f.S[r] = S;
                f.A[r] = A;
                f.gamma_air[r] = ga;
                f.gamma_self[r] = gs;
                f.elower[r] = E;
                f.n_air[r] = na;
                f.delta_air[r] = da;
            }
        }
    }
    status[r] = st;
}

// records kept (status OK and nu in [nu_min, nu_max]) per block of 256 records
__global__ void __launch_bounds__(256) par_keep_count_kernel(const uint8_t *status, const double *nu, int64_t n_rec,
                                                             double nu_min, double nu_max, int64_t *blk) {
    __shared__ int s[256];
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    s[threadIdx.x] = (r < n_rec && status[r] == PAR_OK && nu[r] >= nu_min && nu[r] <= nu_max) ? 1 : 0;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) blk[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(256) par_compact_kernel(const uint8_t *status, ParFields f, int64_t n_rec,
                                                          double nu_min, double nu_max, const int64_t *blk_off,
                                                          ParFields g) {
    __shared__ int s[256];
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int keep = (r < n_rec && status[r] == PAR_OK && f.nu[r] >= nu_min && f.nu[r] <= nu_max) ? 1 : 0;
    s[threadIdx.x] = keep;
    __syncthreads();
    for (int st = 1; st < 256; st <<= 1) {
        const int v = threadIdx.x >= st ? s[threadIdx.x - st] : 0;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    if (!keep) return;
    const int64_t p = blk_off[blockIdx.x] + s[threadIdx.x] - 1;
    g.mol[p] = f.mol[r];
    g.iso[p] = f.iso[r];
    g.nu[p] = f.nu[r];
    g.S[p] = f.S[r];
    g.A[p] = f.A[r];
    g.gamma_air[p] = f.gamma_air[r];
    g.gamma_self[p] = f.gamma_self[r];
    g.elower[p] = f.elower[r];
    g.n_air[p] = f.n_air[r];
    g.delta_air[p] = f.delta_air[r];
}

// total order of the wavenumbers as unsigned keys (NaN last)
__device__ __forceinline__ uint64_t sort_key(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__global__ void __launch_bounds__(256) par_keys_kernel(const double *nu, int64_t n, int64_t n_pad, uint64_t *key,
                                                       uint32_t *idx, unsigned long long *descents) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n_pad) return;
    key[i] = i < n ? sort_key(nu[i]) : ~0ull;
    idx[i] = (uint32_t)i;
    if (i + 1 < n && sort_key(nu[i]) > sort_key(nu[i + 1])) atomicAdd(descents, 1ull);
}

// one bitonic compare-exchange stage on (key, idx): ties ordered by idx, so the
// result is the stable order
__global__ void __launch_bounds__(256) par_bitonic_kernel(uint64_t *key, uint32_t *idx, int64_t n_pad, int64_t k,
                                                          int64_t j) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n_pad) return;
    const int64_t p = i ^ j;
    if (p <= i) return;
    const bool up = (i & k) == 0;
    const uint64_t ki = key[i], kp = key[p];
    const uint32_t ii = idx[i], ip = idx[p];
    const bool gt = ki > kp || (ki == kp && ii > ip);
    if (gt == up) {
        key[i] = kp;
        key[p] = ki;
        idx[i] = ip;
        idx[p] = ii;
    }
}

__global__ void __launch_bounds__(256) par_gather_kernel(ParFields g, const uint32_t *idx, int64_t n, ParFields h) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = idx[i];
    h.mol[i] = g.mol[r];
    h.iso[i] = g.iso[r];
    h.nu[i] = g.nu[r];
    h.S[i] = g.S[r];
    h.A[i] = g.A[r];
    h.gamma_air[i] = g.gamma_air[r];
    h.gamma_self[i] = g.gamma_self[r];
    h.elower[i] = g.elower[r];
    h.n_air[i] = g.n_air[r];
    h.delta_air[i] = g.delta_air[r];
}

hipError_t launch_par_count(const unsigned char *b, int64_t n, int64_t *blk, int64_t n_blk, hipStream_t s) {
    hipLaunchKernelGGL(par_count_kernel, dim3((unsigned)n_blk), dim3(256), 0, s, b, n, blk);
    return hipGetLastError();
}

hipError_t launch_par_newlines(const unsigned char *b, int64_t n, const int64_t *blk_off, int64_t n_blk, int64_t *nl,
                               hipStream_t s) {
    hipLaunchKernelGGL(par_newlines_kernel, dim3((unsigned)n_blk), dim3(256), 0, s, b, n, blk_off, nl);
    return hipGetLastError();
}

hipError_t launch_par_parse(const unsigned char *b, int64_t n, const int64_t *nl, int64_t n_nl, int64_t n_rec,
                            const ParFields &f, uint8_t *status, unsigned long long *n_host, hipStream_t s) {
    if (n_rec <= 0) return hipSuccess;
    hipLaunchKernelGGL(par_parse_kernel, dim3((unsigned)((n_rec + 255) / 256)), dim3(256), 0, s, b, n, nl, n_nl,
                       n_rec, f, status, n_host);
    return hipGetLastError();
}

hipError_t launch_par_keep_count(const uint8_t *status, const double *nu, int64_t n_rec, double nu_min, double nu_max,
                                 int64_t *blk, hipStream_t s) {
    if (n_rec <= 0) return hipSuccess;
    hipLaunchKernelGGL(par_keep_count_kernel, dim3((unsigned)((n_rec + 255) / 256)), dim3(256), 0, s, status, nu,
                       n_rec, nu_min, nu_max, blk);
    return hipGetLastError();
}

hipError_t launch_par_compact(const uint8_t *status, const ParFields &f, int64_t n_rec, double nu_min, double nu_max,
                              const int64_t *blk_off, const ParFields &g, hipStream_t s) {
    if (n_rec <= 0) return hipSuccess;
    hipLaunchKernelGGL(par_compact_kernel, dim3((unsigned)((n_rec + 255) / 256)), dim3(256), 0, s, status, f, n_rec,
                       nu_min, nu_max, blk_off, g);
    return hipGetLastError();
}

// stable sort of g (n records) by wavenumber into h; key and idx hold n_pad
// (a power of two >= n) entries.  *sorted_out: 1 when g was already in order
// (h then untouched: the caller keeps g).
hipError_t launch_par_sort(const ParFields &g, int64_t n, int64_t n_pad, uint64_t *key, uint32_t *idx,
                           unsigned long long *descents, const ParFields &h, int *sorted_out, hipStream_t s) {
    *sorted_out = 1;
    if (n <= 1) return hipSuccess;
    hipError_t e = hipMemsetAsync(descents, 0, sizeof(unsigned long long), s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(par_keys_kernel, dim3((unsigned)((n_pad + 255) / 256)), dim3(256), 0, s, g.nu, n, n_pad, key,
                       idx, descents);
    unsigned long long d = 0;
    e = hipMemcpyAsync(&d, descents, sizeof(d), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    if (d == 0) return hipSuccess;
    *sorted_out = 0;
    for (int64_t k = 2; k <= n_pad; k <<= 1)
        for (int64_t j = k >> 1; j > 0; j >>= 1)
            hipLaunchKernelGGL(par_bitonic_kernel, dim3((unsigned)((n_pad + 255) / 256)), dim3(256), 0, s, key, idx,
                               n_pad, k, j);
    hipLaunchKernelGGL(par_gather_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, idx, n, h);
    return hipGetLastError();
}
