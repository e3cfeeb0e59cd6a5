// HITRAN .par ingest, host side: the device parse of k_par.hip, the host parser
// for the records the device hands over, and the parsed-linelist cache
// (load_linelist, src/HITRAN/cache.jl:77-198): the whole file is parsed once and
// stored as a structure-of-arrays binary file whose name is keyed by the .par's
// absolute path, size, modification time and the cache version; a window
// [nu_min, nu_max] is applied in memory after loading.
#include <sys/stat.h>
#include <sys/types.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/irs.h"
#include "irs_internal.h"

namespace {

struct HostLines {
    std::vector<int8_t> mol, iso;
    std::vector<double> nu, S, A, elower;
    std::vector<float> gamma_air, gamma_self, n_air, delta_air;
    void resize(size_t n) {
        mol.resize(n), iso.resize(n), nu.resize(n), S.resize(n), A.resize(n), elower.resize(n);
        gamma_air.resize(n), gamma_self.resize(n), n_air.resize(n), delta_air.resize(n);
    }
    size_t size() const { return nu.size(); }
    irs_par_lines view() {
        irs_par_lines o{};
        o.mol = mol.data(), o.iso = iso.data(), o.nu = nu.data(), o.S = S.data(), o.A = A.data();
        o.elower = elower.data(), o.gamma_air = gamma_air.data(), o.gamma_self = gamma_self.data();
        o.n_air = n_air.data(), o.delta_air = delta_air.data();
        return o;
    }
};

bool c_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }

std::string trim(const char *s, int w) {
    int a = 0, b = w;
    while (a < b && c_space(s[a])) ++a;
    while (b > a && c_space(s[b - 1])) --b;
    return std::string(s + a, s + b);
}

// parse(Int8, field): surrounding white space, an optional sign, an optional
// 0x / 0o / 0b base prefix, digits; out of range or malformed -> false
bool host_int8(const char *s, int w, int8_t &out) {
    std::string t = trim(s, w);
    size_t i = 0;
    bool neg = false;
    if (i < t.size() && (t[i] == '+' || t[i] == '-')) neg = t[i++] == '-';
    int base = 10;
    if (i + 1 < t.size() && t[i] == '0' && (t[i + 1] == 'x' || t[i + 1] == 'o' || t[i + 1] == 'b')) {
        base = t[i + 1] == 'x' ? 16 : (t[i + 1] == 'o' ? 8 : 2);
        i += 2;
    }
    if (i >= t.size()) return false;
    long v = 0;
    for (; i < t.size(); ++i) {
        const char c = t[i];
        int d = -1;
        if (c >= '0' && c <= '9') d = c - '0';
        else if (c >= 'a' && c <= 'f') d = 10 + c - 'a';
        else if (c >= 'A' && c <= 'F') d = 10 + c - 'A';
        if (d < 0 || d >= base) return false;
        v = v * base + d;
        if (v > 1000) v = 1000;
    }
    if (neg) v = -v;
    if (v < -128 || v > 127) return false;
    out = (int8_t)v;
    return true;
}

// parse(Float64 / Float32, field): surrounding white space, then the C
// library's strtod / strtof over the whole rest
template <typename T>
bool host_real(const char *s, int w, T &out) {
    std::string t = trim(s, w);
    if (t.empty()) return false;
    char *end = nullptr;
    errno = 0;
    const T v = sizeof(T) == 4 ? (T)std::strtof(t.c_str(), &end) : (T)std::strtod(t.c_str(), &end);
    if (end != t.c_str() + t.size()) return false;
    out = v;
    return true;
}

// one record [s, s + len) (newline and a trailing CR removed): false = skipped
bool host_record(const char *s, int64_t len, int8_t &mol, int8_t &iso, double &nu, double &S, double &A, float &ga,
                 float &gs, double &E, float &na, float &da) {
    if (len < 67) return false;
    for (int i = 0; i < 67; ++i)
        if ((unsigned char)s[i] >= 0x80) return false;
    return host_int8(s, 2, mol) && host_int8(s + 2, 1, iso) && host_real(s + 3, 12, nu) && host_real(s + 15, 10, S) &&
           host_real(s + 25, 10, A) && host_real(s + 35, 5, ga) && host_real(s + 40, 5, gs) &&
           host_real(s + 45, 10, E) && host_real(s + 55, 4, na) && host_real(s + 59, 8, da);
}

struct DevScratch {
    std::vector<void *> ptrs;
    ~DevScratch() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    T *get(size_t n) {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(n * sizeof(T), 16)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (T *)p;
    }
};

#define PCHK(expr)                                                                           \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) return ctx_fail(ctx, IRS_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int alloc_fields(DevScratch &ds, ParFields &f, size_t n) {
    f.mol = ds.get<int8_t>(n), f.iso = ds.get<int8_t>(n), f.nu = ds.get<double>(n), f.S = ds.get<double>(n);
    f.A = ds.get<double>(n), f.elower = ds.get<double>(n), f.gamma_air = ds.get<float>(n);
    f.gamma_self = ds.get<float>(n), f.n_air = ds.get<float>(n), f.delta_air = ds.get<float>(n);
    return (f.mol && f.iso && f.nu && f.S && f.A && f.elower && f.gamma_air && f.gamma_self && f.n_air &&
            f.delta_air) ? IRS_OK : IRS_ENOMEM;
}

int parse_bytes(irs_ctx *ctx, const char *bytes, int64_t n_bytes, double nu_min, double nu_max, int64_t capacity,
                const irs_par_lines *out, int64_t *n_lines, int64_t *n_skipped) {
    hipStream_t st = ctx_stream(ctx);
    *n_lines = 0;
    if (n_skipped) *n_skipped = 0;
    if (n_bytes <= 0) return IRS_OK;
    DevScratch ds;
    const int64_t seg = 256 * 64;
    const int64_t n_blk = (n_bytes + seg - 1) / seg;
    unsigned char *db = ds.get<unsigned char>((size_t)n_bytes);
    int64_t *dblk = ds.get<int64_t>((size_t)n_blk);
    if (!db || !dblk) return ctx_fail(ctx, IRS_ENOMEM, "device allocation failed");
    PCHK(hipMemcpyAsync(db, bytes, (size_t)n_bytes, hipMemcpyHostToDevice, st));
    PCHK(launch_par_count(db, n_bytes, dblk, n_blk, st));
    std::vector<int64_t> blk((size_t)n_blk);
    PCHK(hipMemcpyAsync(blk.data(), dblk, sizeof(int64_t) * n_blk, hipMemcpyDeviceToHost, st));
    PCHK(hipStreamSynchronize(st));
    int64_t n_nl = 0;
    for (auto &v : blk) {
        const int64_t c = v;
        v = n_nl;
        n_nl += c;
    }
    int64_t *dnl = ds.get<int64_t>((size_t)std::max<int64_t>(n_nl, 1));
    if (!dnl) return ctx_fail(ctx, IRS_ENOMEM, "device allocation failed");
    PCHK(hipMemcpyAsync(dblk, blk.data(), sizeof(int64_t) * n_blk, hipMemcpyHostToDevice, st));
    PCHK(launch_par_newlines(db, n_bytes, dblk, n_blk, dnl, st));
    const int64_t n_rec = n_nl + (bytes[n_bytes - 1] != '\n' ? 1 : 0);
    ParFields f{};
    if (alloc_fields(ds, f, (size_t)n_rec)) return ctx_fail(ctx, IRS_ENOMEM, "device allocation failed");
    uint8_t *dstat = ds.get<uint8_t>((size_t)n_rec);
    unsigned long long *dcnt = ds.get<unsigned long long>(2);
    if (!dstat || !dcnt) return ctx_fail(ctx, IRS_ENOMEM, "device allocation failed");
    PCHK(hipMemsetAsync(dcnt, 0, 2 * sizeof(unsigned long long), st));
    PCHK(launch_par_parse(db, n_bytes, dnl, n_nl, n_rec, f, dstat, dcnt, st));
    unsigned long long n_host = 0;
    PCHK(hipMemcpyAsync(&n_host, dcnt, sizeof(n_host), hipMemcpyDeviceToHost, st));
    PCHK(hipStreamSynchronize(st));
    std::vector<uint8_t> stat((size_t)n_rec);
    if (n_host > 0) {
        // the records outside the device grammar: the host parser decides
        std::vector<int64_t> nl((size_t)n_nl);
        PCHK(hipMemcpyAsync(stat.data(), dstat, (size_t)n_rec, hipMemcpyDeviceToHost, st));
        if (n_nl) PCHK(hipMemcpyAsync(nl.data(), dnl, sizeof(int64_t) * n_nl, hipMemcpyDeviceToHost, st));
        PCHK(hipStreamSynchronize(st));
        for (int64_t r = 0; r < n_rec; ++r) {
            if (stat[r] != 2) continue;
            const int64_t a = r == 0 ? 0 : nl[r - 1] + 1;
            int64_t b = r < n_nl ? nl[r] : n_bytes;
            if (b > a && bytes[b - 1] == '\r') --b;
            int8_t mol, iso;
            double nu, S, A, E;
            float ga, gs, na, da;
            if (host_record(bytes + a, b - a, mol, iso, nu, S, A, ga, gs, E, na, da)) {
                stat[r] = 1;
                PCHK(hipMemcpy(f.mol + r, &mol, 1, hipMemcpyHostToDevice));
                PCHK(hipMemcpy(f.iso + r, &iso, 1, hipMemcpyHostToDevice));
                PCHK(hipMemcpy(f.nu + r, &nu, 8, hipMemcpyHostToDevice));
                PCHK(hipMemcpy(f.S + r, &S, 8, hipMemcpyHostToDevice));
                PCHK(hipMemcpy(f.A + r, &A, 8, hipMemcpyHostToDevice));
                PCHK(hipMemcpy(f.gamma_air + r, &ga, 4, hipMemcpyHostToDevice));
                PCHK(hipMemcpy(f.gamma_self + r, &gs, 4, hipMemcpyHostToDevice));
                PCHK(hipMemcpy(f.elower + r, &E, 8, hipMemcpyHostToDevice));
                PCHK(hipMemcpy(f.n_air + r, &na, 4, hipMemcpyHostToDevice));
                PCHK(hipMemcpy(f.delta_air + r, &da, 4, hipMemcpyHostToDevice));
            } else {
                stat[r] = 0;
            }
        }
        PCHK(hipMemcpy(dstat, stat.data(), (size_t)n_rec, hipMemcpyHostToDevice));
    }
    // keep (parsed and inside the window), in file order
    const int64_t n_blk2 = (n_rec + 255) / 256;
    int64_t *dblk2 = ds.get<int64_t>((size_t)n_blk2);
    if (!dblk2) return ctx_fail(ctx, IRS_ENOMEM, "device allocation failed");
    PCHK(launch_par_keep_count(dstat, f.nu, n_rec, nu_min, nu_max, dblk2, st));
    std::vector<int64_t> blk2((size_t)n_blk2);
    PCHK(hipMemcpyAsync(blk2.data(), dblk2, sizeof(int64_t) * n_blk2, hipMemcpyDeviceToHost, st));
    PCHK(hipStreamSynchronize(st));
    int64_t n_keep = 0;
    for (auto &v : blk2) {
        const int64_t c = v;
        v = n_keep;
        n_keep += c;
    }
    if (n_skipped) {
        // records not parsed (short or malformed), whatever the window
        if (n_host == 0) PCHK(hipMemcpy(stat.data(), dstat, (size_t)n_rec, hipMemcpyDeviceToHost));
        int64_t ok = 0;
        for (int64_t r = 0; r < n_rec; ++r) ok += stat[r] == 1;
        *n_skipped = n_rec - ok;
    }
    if (n_keep > capacity)
        return ctx_fail(ctx, IRS_EINVAL, "irs_par_parse: " + std::to_string(n_keep) + " lines exceed the capacity " +
                                             std::to_string(capacity));
    if (n_keep == 0) return IRS_OK;
    if (n_keep > (int64_t)UINT32_MAX) return ctx_fail(ctx, IRS_EINVAL, "irs_par_parse: more than 2^32 lines");
    PCHK(hipMemcpyAsync(dblk2, blk2.data(), sizeof(int64_t) * n_blk2, hipMemcpyHostToDevice, st));
    ParFields g{}, h{};
    if (alloc_fields(ds, g, (size_t)n_keep)) return ctx_fail(ctx, IRS_ENOMEM, "device allocation failed");
    PCHK(launch_par_compact(dstat, f, n_rec, nu_min, nu_max, dblk2, g, st));
    int64_t n_pad = 1;
    while (n_pad < n_keep) n_pad <<= 1;
    uint64_t *dkey = ds.get<uint64_t>((size_t)n_pad);
    uint32_t *didx = ds.get<uint32_t>((size_t)n_pad);
    unsigned long long *ddesc = dcnt + 1;
    if (!dkey || !didx || alloc_fields(ds, h, (size_t)n_keep))
        return ctx_fail(ctx, IRS_ENOMEM, "device allocation failed");
    int sorted = 1;
    PCHK(launch_par_sort(g, n_keep, n_pad, dkey, didx, ddesc, h, &sorted, st));
    const ParFields &src = sorted ? g : h;
    const size_t n = (size_t)n_keep;
    PCHK(hipMemcpyAsync(out->mol, src.mol, n, hipMemcpyDeviceToHost, st));
    PCHK(hipMemcpyAsync(out->iso, src.iso, n, hipMemcpyDeviceToHost, st));
    PCHK(hipMemcpyAsync(out->nu, src.nu, 8 * n, hipMemcpyDeviceToHost, st));
    PCHK(hipMemcpyAsync(out->S, src.S, 8 * n, hipMemcpyDeviceToHost, st));
    PCHK(hipMemcpyAsync(out->A, src.A, 8 * n, hipMemcpyDeviceToHost, st));
    PCHK(hipMemcpyAsync(out->elower, src.elower, 8 * n, hipMemcpyDeviceToHost, st));
    PCHK(hipMemcpyAsync(out->gamma_air, src.gamma_air, 4 * n, hipMemcpyDeviceToHost, st));
    PCHK(hipMemcpyAsync(out->gamma_self, src.gamma_self, 4 * n, hipMemcpyDeviceToHost, st));
    PCHK(hipMemcpyAsync(out->n_air, src.n_air, 4 * n, hipMemcpyDeviceToHost, st));
    PCHK(hipMemcpyAsync(out->delta_air, src.delta_air, 4 * n, hipMemcpyDeviceToHost, st));
    PCHK(hipStreamSynchronize(st));
    *n_lines = n_keep;
    return IRS_OK;
}

// ---- cache file ---------------------------------------------------------------
// "IRSLL001", uint32 version, uint32 0, uint64 n, then nu, S, A, E'' (f64), gamma_air,
// gamma_self, n_air, delta_air (f32), mol, iso (i8), each n entries.
const char CACHE_MAGIC[8] = {'I', 'R', 'S', 'L', 'L', '0', '0', '1'};

uint64_t fnv1a(const std::string &s) {
    uint64_t h = 1469598103934665603ull;
    for (unsigned char c : s) {
        h ^= c;
        h *= 1099511628211ull;
    }
    return h;
}

bool read_cache(const std::string &path, HostLines &L) {
    FILE *fp = std::fopen(path.c_str(), "rb");
    if (!fp) return false;
    char magic[8];
    uint32_t ver = 0, pad = 0;
    uint64_t n = 0;
    bool ok = std::fread(magic, 1, 8, fp) == 8 && std::memcmp(magic, CACHE_MAGIC, 8) == 0 &&
              std::fread(&ver, 4, 1, fp) == 1 && std::fread(&pad, 4, 1, fp) == 1 && std::fread(&n, 8, 1, fp) == 1 &&
              ver == IRS_LINELIST_CACHE_VERSION && n < ((uint64_t)1 << 40);
    if (ok) {
        L.resize((size_t)n);
        auto rd = [&](void *p, size_t sz) { return n == 0 || std::fread(p, sz, (size_t)n, fp) == (size_t)n; };
        ok = rd(L.nu.data(), 8) && rd(L.S.data(), 8) && rd(L.A.data(), 8) && rd(L.elower.data(), 8) &&
             rd(L.gamma_air.data(), 4) && rd(L.gamma_self.data(), 4) && rd(L.n_air.data(), 4) &&
             rd(L.delta_air.data(), 4) && rd(L.mol.data(), 1) && rd(L.iso.data(), 1);
        char extra;
        ok = ok && std::fread(&extra, 1, 1, fp) == 0;            // exactly the expected size
    }
    std::fclose(fp);
    return ok;
}

bool mkdirs(const std::string &dir) {
    if (dir.empty()) return false;
    std::string cur;
    for (size_t i = 0; i <= dir.size(); ++i) {
        if (i == dir.size() || dir[i] == '/') {
            cur = dir.substr(0, i);
            if (!cur.empty()) {
                struct stat sb;
                if (::stat(cur.c_str(), &sb) != 0 && ::mkdir(cur.c_str(), 0777) != 0 && errno != EEXIST) return false;
            }
        }
    }
    struct stat sb;
    return ::stat(dir.c_str(), &sb) == 0 && S_ISDIR(sb.st_mode);
}

bool write_cache(const std::string &path, const HostLines &L) {
    const std::string tmp = path + ".tmp." + std::to_string((long long)::getpid());
    FILE *fp = std::fopen(tmp.c_str(), "wb");
    if (!fp) return false;
    const uint32_t ver = IRS_LINELIST_CACHE_VERSION, pad = 0;
    const uint64_t n = L.size();
    auto wr = [&](const void *p, size_t sz) { return n == 0 || std::fwrite(p, sz, (size_t)n, fp) == (size_t)n; };
    bool ok = std::fwrite(CACHE_MAGIC, 1, 8, fp) == 8 && std::fwrite(&ver, 4, 1, fp) == 1 &&
              std::fwrite(&pad, 4, 1, fp) == 1 && std::fwrite(&n, 8, 1, fp) == 1 && wr(L.nu.data(), 8) &&
              wr(L.S.data(), 8) && wr(L.A.data(), 8) && wr(L.elower.data(), 8) && wr(L.gamma_air.data(), 4) &&
              wr(L.gamma_self.data(), 4) && wr(L.n_air.data(), 4) && wr(L.delta_air.data(), 4) &&
              wr(L.mol.data(), 1) && wr(L.iso.data(), 1);
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::remove(tmp.c_str());
        return false;
    }
    return true;
}

bool read_file(const std::string &path, std::vector<char> &buf) {
    FILE *fp = std::fopen(path.c_str(), "rb");
    if (!fp) return false;
    buf.clear();
    char tmp[1 << 16];
    size_t r;
    while ((r = std::fread(tmp, 1, sizeof tmp, fp)) > 0) buf.insert(buf.end(), tmp, tmp + r);
    const bool ok = !std::ferror(fp);
    std::fclose(fp);
    return ok;
}

}  // namespace

extern "C" {

int irs_par_parse(irs_ctx *ctx, const char *bytes, int64_t n_bytes, double nu_min, double nu_max, int64_t capacity,
                  const irs_par_lines *out, int64_t *n_lines, int64_t *n_skipped) {
    if (!ctx) return IRS_EINVAL;
    if (!n_lines || (n_bytes > 0 && !bytes) || n_bytes < 0 || capacity < 0 || (capacity > 0 && !out))
        return ctx_fail(ctx, IRS_EINVAL, "null argument");
    (void)hipSetDevice(ctx_device(ctx));
    return parse_bytes(ctx, bytes, n_bytes, nu_min, nu_max, capacity, out, n_lines, n_skipped);
}

int irs_par_load(irs_ctx *ctx, const char *par_path, const char *cache_dir, int rebuild, double nu_min, double nu_max,
                 int64_t capacity, const irs_par_lines *out, int64_t *n_lines) {
    if (!ctx) return IRS_EINVAL;
    if (!par_path || !n_lines || capacity < 0 || (capacity > 0 && !out))
        return ctx_fail(ctx, IRS_EINVAL, "null argument");
    (void)hipSetDevice(ctx_device(ctx));
    *n_lines = 0;
    struct stat sb;
    if (::stat(par_path, &sb) != 0 || !S_ISREG(sb.st_mode))
        return ctx_fail(ctx, IRS_ENOTFOUND, std::string("File not found: ") + par_path);
    HostLines L;
    bool have = false;
    std::string cpath;
    if (cache_dir && *cache_dir) {
        char abs[PATH_MAX];
        const std::string ap = ::realpath(par_path, abs) ? std::string(abs) : std::string(par_path);
        const long long mt = (long long)sb.st_mtim.tv_sec * 1000000000ll + sb.st_mtim.tv_nsec;
        const std::string key = ap + '\0' + std::to_string((long long)sb.st_size) + '\0' + std::to_string(mt) + '\0' +
                                std::to_string(IRS_LINELIST_CACHE_VERSION);
        char hex[32];
        std::snprintf(hex, sizeof hex, "%016llx", (unsigned long long)fnv1a(key));
        cpath = std::string(cache_dir) + "/linelist_" + hex + ".irsll";
        if (!rebuild) have = read_cache(cpath, L);
    }
    if (!have) {
        std::vector<char> buf;
        if (!read_file(par_path, buf)) return ctx_fail(ctx, IRS_ENOTFOUND, std::string("cannot read ") + par_path);
        const int64_t cap = (int64_t)buf.size() / 67 + 1;
        L.resize((size_t)cap);
        irs_par_lines v = L.view();
        int64_t n = 0;
        int rc = parse_bytes(ctx, buf.data(), (int64_t)buf.size(), -INFINITY, INFINITY, cap, &v, &n, nullptr);
        if (rc) return rc;
        L.resize((size_t)n);
        if (!cpath.empty()) {
            if (!mkdirs(cache_dir) || !write_cache(cpath, L))
                return ctx_fail(ctx, IRS_EINVAL, "cannot write the linelist cache " + cpath);
        }
    }
    // the window, in memory (filter_linelist, linelist.jl:63-67)
    int64_t k = 0;
    for (size_t i = 0; i < L.size(); ++i)
        if (L.nu[i] >= nu_min && L.nu[i] <= nu_max) ++k;
    if (k > capacity)
        return ctx_fail(ctx, IRS_EINVAL, "irs_par_load: " + std::to_string(k) + " lines exceed the capacity " +
                                             std::to_string(capacity));
    k = 0;
    for (size_t i = 0; i < L.size(); ++i) {
        if (!(L.nu[i] >= nu_min && L.nu[i] <= nu_max)) continue;
        out->mol[k] = L.mol[i], out->iso[k] = L.iso[i], out->nu[k] = L.nu[i], out->S[k] = L.S[i];
        out->A[k] = L.A[i], out->elower[k] = L.elower[i], out->gamma_air[k] = L.gamma_air[i];
        out->gamma_self[k] = L.gamma_self[i], out->n_air[k] = L.n_air[i], out->delta_air[k] = L.delta_air[i];
        ++k;
    }
    *n_lines = k;
    return IRS_OK;
}

}  // extern "C"
