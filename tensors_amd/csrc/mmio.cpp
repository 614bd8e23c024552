// MatrixMarket coordinate files on the host, the on-disk format of
// MatrixCSR::LoadFromMatrixMarket / WriteToMatrixMarket
// (MatrixCSR/MatrixMarket.hpp:3-306).
//
// Reading: the header (banner, size line) is checked with the reference's rules
// (MatrixMarket.hpp:21-137): "%%MatrixMarket matrix coordinate <field>
// <symmetry>", field "complex" refused, any other field read as one real value
// per entry except "pattern" (value 1); symmetry "general" or "symmetric"
// (then every off-diagonal entry is mirrored when the CSR is built, as the
// reference passes symmetrize to its triples constructor, :190-196);
// "skew-symmetric" / "hermitian" refused.  The entry lines are parsed in
// parallel from a memory-mapped file: the data section is cut into line-aligned
// chunks, each thread counts and then parses its chunk into the positions that
// follow the chunks before it, so the triples come out in file order exactly as
// a sequential reader yields them.  Indices are returned zero-based.
//
// Writing: "%%MatrixMarket matrix coordinate real general", the size line, then
// one "i j value" line per nonzero, rows in order (%.17E for FP64 as the
// reference, %.8E for FP32, one digit more than the reference's %.7E so that
// every float reads back exactly).  The rows are formatted in parallel and
// written in order; every row is written (the reference's four-thread writer
// joins only two of its four buffers, :283-295).

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <charconv>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"

namespace tg {
namespace {

struct MappedFile {
    const char* data = nullptr;
    size_t size = 0;
    int fd = -1;
    ~MappedFile() {
        if (data && size) munmap((void*)data, size);
        if (fd >= 0) close(fd);
    }
    int open_read(const char* path) {
        fd = ::open(path, O_RDONLY);
        if (fd < 0) return fail(TG_ERR_INVALID, "MatrixMarket: file %s not found", path);
        struct stat sb;
        if (fstat(fd, &sb) != 0) return fail(TG_ERR_INVALID, "MatrixMarket: cannot stat %s", path);
        size = (size_t)sb.st_size;
        if (size == 0) return fail(TG_ERR_INVALID, "MatrixMarket: %s is empty", path);
        void* p = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
        if (p == MAP_FAILED) {
            size = 0;
            return fail(TG_ERR_INVALID, "MatrixMarket: cannot map %s", path);
        }
        data = (const char*)p;
        return TG_OK;
    }
};

struct Header {
    int32_t m = 0, n = 0;
    int64_t entries = 0;
    int field = TG_MM_REAL;
    int symmetric = 0;
    size_t data_off = 0;  // first byte after the size line
};

std::string lower(std::string s) {
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return s;
}

// Next whitespace-separated word on the current line starting at *pos.
std::string word(const char* d, size_t size, size_t* pos) {
    size_t p = *pos;
    while (p < size && (d[p] == ' ' || d[p] == '\t' || d[p] == '\r')) ++p;
    const size_t b = p;
    while (p < size && !std::isspace((unsigned char)d[p])) ++p;
    *pos = p;
    return std::string(d + b, p - b);
}

int parse_header(const char* d, size_t size, const char* path, Header* h) {
    size_t pos = 0;
    if (lower(word(d, size, &pos)) != "%%matrixmarket")
        return fail(TG_ERR_INVALID, "MatrixMarket: %s is not a MatrixMarket file", path);
    if (lower(word(d, size, &pos)) != "matrix")
        return fail(TG_ERR_INVALID, "MatrixMarket: second word in %s is not \"matrix\"", path);
    if (lower(word(d, size, &pos)) != "coordinate")
        return fail(TG_ERR_INVALID, "MatrixMarket: %s is not in coordinate format (dense files are not loaded)", path);
    const std::string field = lower(word(d, size, &pos));
    if (field == "complex")
        return fail(TG_ERR_DTYPE, "MatrixMarket: %s holds complex values; real types only", path);
    h->field = field == "pattern" ? TG_MM_PATTERN : field == "integer" ? TG_MM_INTEGER : TG_MM_REAL;
    const std::string symm = lower(word(d, size, &pos));
    if (symm == "symmetric")
        h->symmetric = 1;
    else if (symm == "general")
        h->symmetric = 0;
    else if (symm == "skew-symmetric" || symm == "hermitian")
        return fail(TG_ERR_INVALID, "MatrixMarket: symmetry \"%s\" of %s is not supported", symm.c_str(), path);
    else
        return fail(TG_ERR_INVALID, "MatrixMarket: symmetry \"%s\" of %s is not valid", symm.c_str(), path);
    // rest of the banner line, then comment (and blank) lines
    while (pos < size && d[pos] != '\n') ++pos;
    for (;;) {
        if (pos >= size) return fail(TG_ERR_INVALID, "MatrixMarket: %s has no size line", path);
        ++pos;  // past '\n'
        size_t q = pos;
        while (q < size && (d[q] == ' ' || d[q] == '\t' || d[q] == '\r')) ++q;
        if (q < size && d[q] != '%' && d[q] != '\n') break;
        pos = q;
        while (pos < size && d[pos] != '\n') ++pos;
    }
    long long v[3];
    for (int t = 0; t < 3; ++t) {
        while (pos < size && (d[pos] == ' ' || d[pos] == '\t')) ++pos;
        auto r = std::from_chars(d + pos, d + size, v[t]);
        if (r.ec != std::errc()) return fail(TG_ERR_INVALID, "MatrixMarket: bad size line in %s", path);
        pos = (size_t)(r.ptr - d);
    }
    if (v[0] < 0 || v[0] > INT32_MAX) return fail(TG_ERR_INVALID, "MatrixMarket: invalid row count in %s", path);
    if (v[1] < 0 || v[1] > INT32_MAX) return fail(TG_ERR_INVALID, "MatrixMarket: invalid column count in %s", path);
    if (v[2] < 0) return fail(TG_ERR_INVALID, "MatrixMarket: invalid nonzero count in %s", path);
    h->m = (int32_t)v[0];
    h->n = (int32_t)v[1];
    h->entries = v[2];
    while (pos < size && d[pos] != '\n') ++pos;
    h->data_off = pos < size ? pos + 1 : size;
    return TG_OK;
}

// A data line: not blank and not a comment.
inline bool entry_line(const char* d, size_t b, size_t e) {
    while (b < e && (d[b] == ' ' || d[b] == '\t' || d[b] == '\r')) ++b;
    return b < e && d[b] != '%';
}

int thread_count(int32_t requested) {
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    return requested > 0 ? std::min<int>(requested, 256) : std::min(hw, 64);
}

}  // namespace
}  // namespace tg

using namespace tg;

extern "C" int tg_mm_info(const char* path, int32_t* m, int32_t* n, int64_t* entries, int32_t* field,
                          int32_t* symmetric) {
    if (!path || !m || !n || !entries || !field || !symmetric) return fail(TG_ERR_INVALID, "tg_mm_info: null argument");
    MappedFile f;
    int rc = f.open_read(path);
    if (rc) return rc;
    Header h;
    rc = parse_header(f.data, f.size, path, &h);
    if (rc) return rc;
    *m = h.m;
    *n = h.n;
    *entries = h.entries;
    *field = h.field;
    *symmetric = h.symmetric;
    return TG_OK;
}

extern "C" int tg_mm_read(const char* path, int64_t entries, int32_t* I, int32_t* J, double* V, int32_t threads) {
    if (!path || entries < 0 || (entries > 0 && (!I || !J))) return fail(TG_ERR_INVALID, "tg_mm_read: bad argument");
    MappedFile f;
    int rc = f.open_read(path);
    if (rc) return rc;
    Header h;
    rc = parse_header(f.data, f.size, path, &h);
    if (rc) return rc;
    if (h.entries != entries)
        return fail(TG_ERR_INVALID, "tg_mm_read: %s holds %lld entries, the buffers %lld", path,
                    (long long)h.entries, (long long)entries);
    const char* d = f.data;
    const size_t size = f.size, off = h.data_off;
    // line-aligned chunks of the data section
    const int T = (int)std::max<int64_t>(1, std::min<int64_t>(thread_count(threads), (int64_t)(size - off) / 65536 + 1));
    std::vector<size_t> cut(T + 1);
    cut[0] = off;
    cut[T] = size;
    for (int t = 1; t < T; ++t) {
        size_t p = off + (size - off) * (size_t)t / (size_t)T;
        p = std::max(p, cut[t - 1]);
        while (p < size && d[p - 1] != '\n') ++p;
        cut[t] = p;
    }
    // pass 1: entry lines per chunk
    std::vector<int64_t> count(T + 1, 0);
    auto for_lines = [&](int t, auto&& fn) {
        size_t b = cut[t];
        while (b < cut[t + 1]) {
            const char* nl = (const char*)memchr(d + b, '\n', cut[t + 1] - b);
            const size_t e = nl ? (size_t)(nl - d) : cut[t + 1];
            if (entry_line(d, b, e) && !fn(b, e)) return;
            b = e + 1;
        }
    };
    {
        std::vector<std::thread> pool;
        for (int t = 0; t < T; ++t)
            pool.emplace_back([&, t] {
                int64_t c = 0;
                for_lines(t, [&](size_t, size_t) { return ++c, true; });
                count[t + 1] = c;
            });
        for (auto& th : pool) th.join();
    }
    for (int t = 0; t < T; ++t) count[t + 1] += count[t];
    if (count[T] < entries)
        return fail(TG_ERR_INVALID, "tg_mm_read: %s announces %lld entries but holds %lld", path,
                    (long long)entries, (long long)count[T]);
    // pass 2: parse into the chunk's positions (entries beyond the announced count are ignored)
    const bool pattern = h.field == TG_MM_PATTERN;
    std::vector<int64_t> bad(T, -1);  // first bad entry per chunk
    {
        std::vector<std::thread> pool;
        for (int t = 0; t < T; ++t)
            pool.emplace_back([&, t] {
                int64_t k = count[t];
                for_lines(t, [&](size_t b, size_t e) {
                    if (k >= entries) return false;
                    const char* p = d + b;
                    const char* end = d + e;
                    auto skip = [&] {
                        while (p < end && (*p == ' ' || *p == '\t')) ++p;
                    };
                    long long i = 0, j = 0;
                    double v = 1.0;
                    skip();
                    auto r = std::from_chars(p, end, i);
                    bool ok = r.ec == std::errc();
                    p = r.ptr;
                    skip();
                    r = std::from_chars(p, end, j);
                    ok = ok && r.ec == std::errc();
                    p = r.ptr;
                    if (!pattern && ok) {
                        skip();
                        if (p < end && *p == '+') ++p;  // from_chars takes no leading '+'
                        auto rv = std::from_chars(p, end, v);
                        ok = rv.ec == std::errc();
                    }
                    if (!ok || i < 1 || i > h.m || j < 1 || j > h.n) {
                        bad[t] = k;
                        return false;
                    }
                    I[k] = (int32_t)(i - 1);
                    J[k] = (int32_t)(j - 1);
                    if (V) V[k] = v;
                    ++k;
                    return true;
                });
            });
        for (auto& th : pool) th.join();
    }
    for (int t = 0; t < T; ++t)
        if (bad[t] >= 0)
            return fail(TG_ERR_INVALID, "tg_mm_read: entry %lld of %s is malformed or out of range",
                        (long long)bad[t] + 1, path);
    return TG_OK;
}

namespace {

template <typename RP, typename S>
int mm_write_t(const char* path, int32_t m, int32_t n, const RP* rp, const int32_t* ci, const S* vals, int T) {
    const int64_t nnz = (int64_t)rp[m] - (int64_t)rp[0];
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return fail(TG_ERR_INVALID, "tg_mm_write: cannot open %s for writing", path);
    std::fprintf(fp, "%%%%MatrixMarket matrix coordinate %s general\n%d %d %lld\n", vals ? "real" : "pattern", m, n,
                 (long long)nnz);
    // rows cut into T ranges of about equal nonzero counts, formatted in parallel
    std::vector<int32_t> r0(T + 1);
    r0[0] = 0;
    r0[T] = m;
    for (int t = 1; t < T; ++t) {
        const int64_t target = (int64_t)rp[0] + nnz * t / T;
        r0[t] = (int32_t)(std::lower_bound(rp, rp + m + 1, (RP)target) - rp);
        r0[t] = std::max(r0[t], r0[t - 1]);
    }
    std::vector<std::string> text(T);
    std::vector<std::thread> pool;
    for (int t = 0; t < T; ++t)
        pool.emplace_back([&, t] {
            std::string& s = text[t];
            char line[96];
            for (int32_t i = r0[t]; i < r0[t + 1]; ++i)
                for (RP l = rp[i]; l < rp[i + 1]; ++l) {
                    int len;
                    if (!vals)
                        len = std::snprintf(line, sizeof line, "%d %d\n", i + 1, ci[l] + 1);
                    else if constexpr (sizeof(S) == 8)
                        len = std::snprintf(line, sizeof line, "%d %d %.17E\n", i + 1, ci[l] + 1, (double)vals[l]);
                    else
                        len = std::snprintf(line, sizeof line, "%d %d %.8E\n", i + 1, ci[l] + 1, (double)vals[l]);
                    s.append(line, (size_t)len);
                }
        });
    for (auto& th : pool) th.join();
    bool ok = true;
    for (int t = 0; t < T; ++t) ok = ok && std::fwrite(text[t].data(), 1, text[t].size(), fp) == text[t].size();
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? TG_OK : fail(TG_ERR_INVALID, "tg_mm_write: writing %s failed", path);
}

}  // namespace

extern "C" int tg_mm_write(const char* path, int32_t m, int32_t n, const void* rp, tg_dtype rp_type,
                           const int32_t* ci, const void* vals, tg_dtype val_type, int32_t threads) {
    if (!path || m < 0 || n < 0 || !rp) return fail(TG_ERR_INVALID, "tg_mm_write: bad argument");
    if (rp_type != TG_I32 && rp_type != TG_I64) return fail(TG_ERR_DTYPE, "tg_mm_write: rp_type must be I32/I64");
    if (vals && val_type != TG_F64 && val_type != TG_F32)
        return fail(TG_ERR_DTYPE, "tg_mm_write: val_type must be F64/F32");
    const int T = thread_count(threads);
    const bool f64 = val_type == TG_F64;
    if (rp_type == TG_I64) {
        const int64_t* r = (const int64_t*)rp;
        if (r[m] > r[0] && !ci) return fail(TG_ERR_INVALID, "tg_mm_write: ci is null");
        return f64 ? mm_write_t(path, m, n, r, ci, (const double*)vals, T)
                   : mm_write_t(path, m, n, r, ci, (const float*)vals, T);
    }
    const int32_t* r = (const int32_t*)rp;
    if (r[m] > r[0] && !ci) return fail(TG_ERR_INVALID, "tg_mm_write: ci is null");
    return f64 ? mm_write_t(path, m, n, r, ci, (const double*)vals, T)
               : mm_write_t(path, m, n, r, ci, (const float*)vals, T);
}
