// dhip_io.hpp -- the reference's on-disk formats for CSR matrices and vectors (host code).
//
//   MatrixMarket   MatrixCL::exportMTX = operator<< (num/spmat.h:1084-1089, 1213-1226):
//                  "%%MatrixMarket matrix coordinate real general", "rows cols nnz", then one
//                  line "row+1 col+1 value" per entry in CSR order, the value in scientific
//                  notation with 15 digits after the point.
//   plain          out(std::ostream&, const MatrixCL&) / in() (num/spmat.h:1228-1255):
//                  "rows cols nnz", the row pointers, the column indices, the values (stream
//                  default precision: 6 significant digits), each item followed by a blank,
//                  one line per array.
//   vector text    VecDescBaseCL::Write(os, false) = operator<< of VectorBaseCL
//                  (misc/problem.h:928-944, num/spmat.h:1183-1188): the count, six blanks,
//                  every value with 15 significant digits followed by a blank, a newline.
//   vector binary  VecDescBaseCL::Write(os, true): the count as a native size_t, then the raw
//                  doubles (misc/problem.h:933-936).
//
// iostream formats a double through the C library's printf conversions, so the same
// conversions (%.15e, %g, %.15g) give the reference's bytes.  Large arrays are formatted by
// up to 16 threads into per-thread buffers that are emitted in order, to a file or to a
// caller's buffer.
#pragma once

#include <algorithm>
#include <cerrno>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace dhipio {

enum { IO_OK = 0, IO_E_OPEN = 1, IO_E_WRITE = 2, IO_E_FORMAT = 3, IO_E_COUNT = 4, IO_E_RANGE = 5, IO_E_SPACE = 6,
       IO_E_NOMEM = 7 };

enum { FMT_MARKET = 0, FMT_PLAIN = 1 };

// bytes per item that no formatted item exceeds (an int32 index is at most 11 characters,
// %.15e at most 23, %g at most 13, %.15g at most 22, each plus its separator)
constexpr int64_t MTX_LINE_MAX = 48, PLAIN_ITEM_MAX = 14, VEC_ITEM_MAX = 24, HEAD_MAX = 128;

inline int64_t matrix_bound(int format, int64_t nrows, int64_t nnz) {
    return format == FMT_MARKET ? HEAD_MAX + MTX_LINE_MAX * nnz
                                : HEAD_MAX + 12 * (nrows + 1) + (12 + PLAIN_ITEM_MAX) * nnz;
}
inline int64_t vec_bound(int64_t n, bool binary) { return binary ? 8 + 8 * n : HEAD_MAX + VEC_ITEM_MAX * n; }

inline int io_threads(int64_t items) {
    const int64_t per = 1 << 16;  // below this many items per thread, fewer threads
    unsigned hw = std::thread::hardware_concurrency();
    int64_t t = hw ? (int64_t)hw : 1;
    t = std::min<int64_t>(t, 16);
    t = std::min<int64_t>(t, (items + per - 1) / per);
    return (int)std::max<int64_t>(t, 1);
}

// destination of formatted bytes: a file or a caller's buffer of cap bytes
struct Sink {
    FILE *f = nullptr;
    char *buf = nullptr;
    int64_t cap = 0, len = 0;
    int put(const char *p, size_t k) {
        if (f) return std::fwrite(p, 1, k, f) == k ? IO_OK : IO_E_WRITE;
        if (len + (int64_t)k > cap) return IO_E_SPACE;
        std::memcpy(buf + len, p, k);
        len += (int64_t)k;
        return IO_OK;
    }
    int put(const std::string &s) { return s.empty() ? IO_OK : put(s.data(), s.size()); }
};

template <class... A>
inline void append(std::string &s, const char *fmt, A... a) {
    char tmp[64];
    const int k = std::snprintf(tmp, sizeof tmp, fmt, a...);
    if (k <= 0) return;
    if ((size_t)k < sizeof tmp) {
        s.append(tmp, (size_t)k);
        return;
    }
    std::vector<char> big((size_t)k + 1);
    std::snprintf(big.data(), big.size(), fmt, a...);
    s.append(big.data(), (size_t)k);
}

// Formats the items [b_t, b_{t+1}) of every part t with fmt(lo, hi, out) in parallel and emits
// the parts in order.  A part whose thread cannot be started is formatted by the caller; an
// allocation failure inside a worker is reported, never thrown across the thread.
template <class F>
inline int write_parts(Sink &out, const std::vector<int64_t> &bounds, F &&fmt) {
    const int parts = (int)bounds.size() - 1;
    std::vector<std::string> buf(parts);
    std::vector<char> failed(parts, 0);
    auto run = [&](int t) {
        try {
            fmt(bounds[t], bounds[t + 1], buf[t]);
        } catch (...) {
            failed[t] = 1;
        }
    };
    std::vector<std::thread> th;
    int started = 0;
    if (parts > 1) {
        try {
            th.reserve(parts);
            for (; started < parts; ++started) th.emplace_back(run, started);
        } catch (...) {
            // fewer threads than parts: the rest runs here
        }
    }
    for (int t = started; t < parts; ++t) run(t);
    for (auto &x : th) x.join();
    for (int t = 0; t < parts; ++t)
        if (failed[t]) return IO_E_NOMEM;
    for (auto &s : buf)
        if (int rc = out.put(s)) return rc;
    return IO_OK;
}

inline std::vector<int64_t> even_bounds(int64_t count) {
    const int parts = io_threads(count);
    std::vector<int64_t> b(parts + 1);
    for (int t = 0; t <= parts; ++t) b[t] = count * t / parts;
    return b;
}

// ---------------------------------------------------------------------------
// writers
// ---------------------------------------------------------------------------
inline int format_matrix(Sink &out, int format, int64_t nrows, int64_t ncols, const int32_t *rp, const int32_t *col,
                         const double *val) {
    const int64_t nnz = rp[nrows];
    std::string head;
    int rc;
    if (format == FMT_MARKET) {
        append(head, "%%%%MatrixMarket matrix coordinate real general\n%lld %lld %lld\n", (long long)nrows,
               (long long)ncols, (long long)nnz);
        if ((rc = out.put(head))) return rc;
        // parts by rows, balanced by entries
        const int parts = io_threads(nnz);
        std::vector<int64_t> b(parts + 1, nrows);
        b[0] = 0;
        for (int t = 1; t < parts; ++t) {
            const int64_t target = nnz * t / parts;
            b[t] = std::upper_bound(rp, rp + nrows + 1, (int32_t)target) - rp - 1;
            b[t] = std::max(b[t], b[t - 1]);
        }
        return write_parts(out, b, [&](int64_t r0, int64_t r1, std::string &s) {
            s.reserve((size_t)(rp[r1] - rp[r0]) * 32);
            for (int64_t r = r0; r < r1; ++r)
                for (int32_t k = rp[r]; k < rp[r + 1]; ++k)
                    append(s, "%lld %lld %.15e\n", (long long)(r + 1), (long long)col[k] + 1, val[k]);
        });
    }
    append(head, "%lld %lld %lld\n", (long long)nrows, (long long)ncols, (long long)nnz);
    if ((rc = out.put(head))) return rc;
    if ((rc = write_parts(out, even_bounds(nrows + 1), [&](int64_t a, int64_t b, std::string &s) {
             for (int64_t i = a; i < b; ++i) append(s, "%d ", rp[i]);
         })))
        return rc;
    if ((rc = out.put("\n", 1))) return rc;
    if ((rc = write_parts(out, even_bounds(nnz), [&](int64_t a, int64_t b, std::string &s) {
             for (int64_t i = a; i < b; ++i) append(s, "%d ", col[i]);
         })))
        return rc;
    if ((rc = out.put("\n", 1))) return rc;
    if ((rc = write_parts(out, even_bounds(nnz), [&](int64_t a, int64_t b, std::string &s) {
             for (int64_t i = a; i < b; ++i) append(s, "%g ", val[i]);
         })))
        return rc;
    return out.put("\n", 1);
}

inline int format_vec(Sink &out, const double *x, int64_t n, bool binary) {
    int rc;
    if (binary) {
        const size_t cnt = (size_t)n;
        if ((rc = out.put((const char *)&cnt, sizeof cnt))) return rc;
        return n > 0 ? out.put((const char *)x, sizeof(double) * (size_t)n) : IO_OK;
    }
    std::string head;
    append(head, "%lld      ", (long long)n);
    if ((rc = out.put(head))) return rc;
    if ((rc = write_parts(out, even_bounds(n), [&](int64_t a, int64_t b, std::string &s) {
             s.reserve((size_t)(b - a) * 24);
             for (int64_t i = a; i < b; ++i) append(s, "%.15g ", x[i]);
         })))
        return rc;
    return out.put("\n", 1);
}

// to a file
template <class F>
inline int to_file(const char *path, F &&fmt) {
    FILE *f = std::fopen(path, "wb");
    if (!f) return IO_E_OPEN;
    Sink out;
    out.f = f;
    const int rc = fmt(out);
    const int rc2 = std::fclose(f) == 0 ? IO_OK : IO_E_WRITE;
    return rc ? rc : rc2;
}

// ---------------------------------------------------------------------------
// readers (from memory; a file is read whole first)
// ---------------------------------------------------------------------------
inline int slurp(const char *path, std::string &s) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return IO_E_OPEN;
    char tmp[1 << 16];
    size_t k;
    while ((k = std::fread(tmp, 1, sizeof tmp, f)) > 0) s.append(tmp, k);
    std::fclose(f);
    return IO_OK;
}

// tokens of a NUL-terminated text (std::string storage is)
struct Scanner {
    const char *p, *end;
    void blanks() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool int64(long long &v) {
        blanks();
        if (p >= end) return false;
        char *q;
        errno = 0;
        v = std::strtoll(p, &q, 10);
        if (q == p || errno) return false;
        p = q;
        return true;
    }
    bool real(double &v) {
        blanks();
        if (p >= end) return false;
        char *q;
        errno = 0;
        v = std::strtod(p, &q);
        if (q == p) return false;
        // beyond the double range (e.g. DBL_MAX written with 15 digits): the largest finite
        // value of that sign, as operator>> stores it
        if (errno == ERANGE && std::fabs(v) == HUGE_VAL) v = std::copysign(DBL_MAX, v);
        p = q;
        return true;
    }
    void skip_line() {
        while (p < end && *p != '\n') ++p;
        if (p < end) ++p;
    }
};

// Header of either matrix format: sizes only.  Counts that the text is too short to hold are
// rejected before anything is allocated for them.
inline int read_matrix_header(Scanner &S, int format, long long &nr, long long &nc, long long &nz) {
    if (format == FMT_MARKET) {
        static const char magic[] = "%%MatrixMarket matrix coordinate";
        if ((size_t)(S.end - S.p) < sizeof magic - 1 || std::strncmp(S.p, magic, sizeof magic - 1) != 0)
            return IO_E_FORMAT;
        const char *eol = (const char *)std::memchr(S.p, '\n', (size_t)(S.end - S.p));
        const std::string head(S.p, eol ? eol : S.end);
        if (head.find(" real ") == std::string::npos && head.find(" integer ") == std::string::npos)
            return IO_E_FORMAT;
        if (head.find(" general") == std::string::npos) return IO_E_FORMAT;
        while (S.p < S.end && *S.p == '%') S.skip_line();
    }
    if (!S.int64(nr) || !S.int64(nc) || !S.int64(nz) || nr < 0 || nc < 0 || nz < 0) return IO_E_FORMAT;
    if (nr >= (1ll << 31) || nc >= (1ll << 31) || nz >= (1ll << 31)) return IO_E_RANGE;
    const long long left = (long long)(S.end - S.p);
    // an entry takes at least 6 bytes ("1 1 0\n"); a plain item at least 2 ("0 ")
    if (format == FMT_MARKET ? nz > left / 6 + 1 : (nr + 1 + 2 * nz) > left / 2 + 1) return IO_E_FORMAT;
    return IO_OK;
}

// A matrix from text; with rp == nullptr only the sizes (nnz = the stored count).  MatrixMarket
// entries may come in any order: the result is CSR with the columns of a row ascending and
// repeated (row, col) pairs summed in file order.  The plain format is read as written.
inline int parse_matrix(const std::string &s, int format, int64_t *nrows, int64_t *ncols, int64_t *nnz, int32_t *rp,
                        int32_t *col, double *val) {
    Scanner S{s.data(), s.data() + s.size()};
    long long nr, nc, nz;
    int rc;
    if ((rc = read_matrix_header(S, format, nr, nc, nz))) return rc;
    *nrows = nr;
    *ncols = nc;
    *nnz = nz;
    if (!rp) return IO_OK;
    if (format == FMT_PLAIN) {
        long long v;
        for (long long i = 0; i <= nr; ++i) {
            if (!S.int64(v) || v < 0 || v > nz || (i > 0 && v < rp[i - 1]) || (i == 0 && v != 0)) return IO_E_FORMAT;
            rp[i] = (int32_t)v;
        }
        if (rp[nr] != nz) return IO_E_FORMAT;
        for (long long k = 0; k < nz; ++k) {
            if (!S.int64(v) || v < 0 || v >= nc) return IO_E_FORMAT;
            col[k] = (int32_t)v;
        }
        for (long long k = 0; k < nz; ++k)
            if (!S.real(val[k])) return IO_E_FORMAT;
        return IO_OK;
    }
    std::vector<int32_t> ri(nz), ci(nz);
    std::vector<double> vv(nz);
    for (long long k = 0; k < nz; ++k) {
        long long i, j;
        if (!S.int64(i) || !S.int64(j) || !S.real(vv[k])) return IO_E_FORMAT;
        if (i < 1 || i > nr || j < 1 || j > nc) return IO_E_RANGE;
        ri[k] = (int32_t)(i - 1);
        ci[k] = (int32_t)(j - 1);
    }
    // counting sort by row (stable: file order within a row), then per-row sort by column
    std::vector<int64_t> cnt(nr + 1, 0);
    for (long long k = 0; k < nz; ++k) ++cnt[ri[k] + 1];
    for (long long i = 0; i < nr; ++i) cnt[i + 1] += cnt[i];
    std::vector<int64_t> order(nz);
    {
        std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
        for (long long k = 0; k < nz; ++k) order[pos[ri[k]]++] = k;
    }
    int64_t out = 0;
    rp[0] = 0;
    for (long long i = 0; i < nr; ++i) {
        auto b = order.begin() + cnt[i], e = order.begin() + cnt[i + 1];
        std::stable_sort(b, e, [&](int64_t x, int64_t y) { return ci[x] < ci[y]; });
        for (auto it = b; it != e; ++it) {
            if (out > rp[i] && col[out - 1] == ci[*it]) {
                val[out - 1] += vv[*it];
            } else {
                col[out] = ci[*it];
                val[out] = vv[*it];
                ++out;
            }
        }
        rp[i + 1] = (int32_t)out;
    }
    *nnz = out;
    return IO_OK;
}

// the count stored at the start of a vector text or binary image
inline int parse_vec_count(const std::string &s, bool binary, int64_t *n) {
    if (binary) {
        size_t cnt = 0;
        if (s.size() < sizeof cnt) return IO_E_FORMAT;
        std::memcpy(&cnt, s.data(), sizeof cnt);
        *n = (int64_t)cnt;
        return IO_OK;
    }
    Scanner S{s.data(), s.data() + s.size()};
    long long cnt;
    if (!S.int64(cnt) || cnt < 0) return IO_E_FORMAT;
    *n = cnt;
    return IO_OK;
}

// VecDescBaseCL::Read: the stored count must equal n
inline int parse_vec(const std::string &s, double *x, int64_t n, bool binary) {
    int64_t cnt;
    int rc = parse_vec_count(s, binary, &cnt);
    if (rc) return rc;
    if (cnt != n) return IO_E_COUNT;
    if (binary) {
        if ((int64_t)s.size() < 8 + 8 * n) return IO_E_FORMAT;
        if (n > 0) std::memcpy(x, s.data() + 8, sizeof(double) * (size_t)n);
        return IO_OK;
    }
    Scanner S{s.data(), s.data() + s.size()};
    long long skip;
    S.int64(skip);
    for (int64_t i = 0; i < n; ++i)
        if (!S.real(x[i])) return IO_E_FORMAT;
    return IO_OK;
}

inline const char *io_message(int rc) {
    switch (rc) {
    case IO_E_OPEN: return "cannot open file";
    case IO_E_WRITE: return "write failed";
    case IO_E_FORMAT: return "data does not hold the expected format";
    case IO_E_COUNT: return "Number of Unknowns does not match, wrong FE-type?";
    case IO_E_RANGE: return "index or size out of range";
    case IO_E_SPACE: return "buffer too small";
    case IO_E_NOMEM: return "host allocation failed";
    default: return "ok";
    }
}

}  // namespace dhipio
