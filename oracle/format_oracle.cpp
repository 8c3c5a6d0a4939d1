// format_oracle.cpp -- the reference's matrix and vector file formats, written and read with
// C++ iostreams as the reference does.  TEST INFRASTRUCTURE ONLY: tests/ compare the library's
// files with these byte for byte; no product code links or loads this.
//
//   orc_fmt_mtx      operator<<(std::ostream&, const SparseMatBaseCL<double>&), num/spmat.h:1213-1226,
//                    the body of MatrixCL::exportMTX (num/spmat.h:1084-1089)
//   orc_fmt_plain    out(std::ostream&, const SparseMatBaseCL<double>&), num/spmat.h:1228-1240
//   orc_read_plain   in(std::istream&, SparseMatBaseCL<double>&), num/spmat.h:1244-1255
//   orc_fmt_vec      VecDescBaseCL::Write, misc/problem.h:928-944, with operator<< of
//                    VectorBaseCL, num/spmat.h:1183-1188
//   orc_read_vec     VecDescBaseCL::Read, misc/problem.h:947-975, with in() of VectorBaseCL
//                    (num/spmat.h:1191-1205); returns 1 when the stored count differs from n
#include <cstdint>
#include <fstream>
#include <limits>

extern "C" {

int orc_fmt_mtx(const char *path, int64_t nrows, int64_t ncols, const int64_t *rp, const int64_t *col,
                const double *val) {
    std::ofstream os(path);
    if (!os) return 2;
    const size_t M = (size_t)nrows;
    os << "%%MatrixMarket matrix coordinate " << "real" << " general\n" << (size_t)nrows << ' ' << (size_t)ncols
       << ' ' << (size_t)rp[nrows] << '\n';
    os.precision(15);
    for (size_t row = 0; row < M; ++row)
        for (size_t k = (size_t)rp[row], e = (size_t)rp[row + 1]; k < e; ++k)
            os << row + 1 << ' ' << (size_t)col[k] + 1 << ' ' << std::scientific << val[k] << '\n';
    os << std::flush;
    return os ? 0 : 2;
}

int orc_fmt_plain(const char *path, int64_t nrows, int64_t ncols, const int64_t *rp, const int64_t *col,
                  const double *val) {
    std::ofstream os(path);
    if (!os) return 2;
    const size_t nnz = (size_t)rp[nrows];
    os << (size_t)nrows << ' ' << (size_t)ncols << ' ' << nnz << '\n';
    for (size_t row = 0; row <= (size_t)nrows; ++row) os << (size_t)rp[row] << ' ';
    os << '\n';
    for (size_t k = 0; k < nnz; ++k) os << (size_t)col[k] << ' ';
    os << '\n';
    for (size_t k = 0; k < nnz; ++k) os << val[k] << ' ';
    os << std::endl;
    return os ? 0 : 2;
}

// sizes with rp == nullptr, else the arrays
int orc_read_plain(const char *path, int64_t *nrows, int64_t *ncols, int64_t *nnz, int64_t *rp, int64_t *col,
                   double *val) {
    std::ifstream is(path);
    if (!is) return 2;
    size_t numrows, numcols, numnz;
    is >> numrows >> numcols >> numnz;
    *nrows = (int64_t)numrows;
    *ncols = (int64_t)numcols;
    *nnz = (int64_t)numnz;
    if (!rp) return is ? 0 : 3;
    size_t v;
    for (size_t row = 0; row <= numrows; ++row) {
        is >> v;
        rp[row] = (int64_t)v;
    }
    for (size_t k = 0; k < numnz; ++k) {
        is >> v;
        col[k] = (int64_t)v;
    }
    for (size_t k = 0; k < numnz; ++k) is >> val[k];
    return is ? 0 : 3;
}

int orc_fmt_vec(const char *path, const double *x, int64_t n, int binary) {
    std::ofstream os(path);
    if (!os) return 2;
    if (binary) {
        size_t numUnk = (size_t)n;
        os.write((const char *)&numUnk, sizeof(size_t));
        os.write((const char *)x, sizeof(double) * (size_t)n);
    } else {
        os.precision(std::numeric_limits<double>::digits10);
        os << (size_t)n << "      ";
        for (int64_t i = 0; i < n; ++i) os << x[i] << ' ';
        os << std::endl;
    }
    return os ? 0 : 2;
}

int orc_read_vec(const char *path, double *x, int64_t n, int binary) {
    std::ifstream is(path);
    if (!is) return 2;
    size_t readUnk = 0;
    if (binary) is.read((char *)&readUnk, sizeof(size_t));
    else {
        is >> readUnk;
        is.seekg(0);
    }
    if ((int64_t)readUnk != n) return 1;
    if (binary) is.read((char *)x, sizeof(double) * readUnk);
    else {
        size_t s;
        is >> s;
        for (size_t i = 0; i < s; ++i) is >> x[i];
    }
    return is ? 0 : 3;
}

}  // extern "C"
