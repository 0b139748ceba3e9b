// lx_matio.cpp -- dense matrices in the Octave text format the reference
// writes and reads around its refinement solver (A.dat, RHS.dat:
// src/itref/itref_setup.c:98-121, 155-199), so that external matrices can
// drive the solve.
//
// Writer, as fprint_dense2mm / print_dense2mm (src/matutil/matfprint.c:94-124): four header
// lines "# name: <name>", "# type: matrix", "# rows: <m>", "# columns: <n>",
// then the m x n values in row-major order, one per line, printed "%.16e \n"
// (17 significant digits: every double reads back exactly).
// Reader, as read_mm2{d,s}dense (src/matutil/matfread.c:13-27): lines starting
// with '#' are skipped, every other line gives one value in file order (%lf /
// %f) until count values are stored; a line that does not parse stores the
// last value read again, since sscanf leaves the variable as it was (0 before
// the first value).  itref_setup.c stores those values straight into its
// column-major arrays, which transposes a non-symmetric file;
// linompss_read_ddense2mm instead checks the header and places value (i, j)
// at A[j * lda + i].
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "linompss.h"

namespace {

template <typename T>
int fprint_dense2mm(const char *fname, const char *name, int m, int n, const T *A, int lda)
{
    if (!fname) return -1;
    if (!name) return -2;
    if (m < 0) return -3;
    if (n < 0) return -4;
    if (!A && m > 0 && n > 0) return -5;
    if (lda < (m > 1 ? m : 1)) return -6;
    FILE *f = fopen(fname, "w");
    if (!f) return LINOMPSS_ERR_IO;
    int bad = fprintf(f, "# name: %s\n# type: matrix\n# rows: %i\n# columns: %i\n", name, m, n) < 0;
    for (int i = 0; i < m && !bad; ++i)
        for (int j = 0; j < n && !bad; ++j) bad = fprintf(f, "%.16e \n", (double)A[(int64_t)j * lda + i]) < 0;
    if (fclose(f) != 0) bad = 1;
    return bad ? LINOMPSS_ERR_IO : 0;
}

template <typename T>
T parse(const char *s, T prev)
{
    char *end = nullptr;
    const T v = sizeof(T) == 8 ? (T)strtod(s, &end) : (T)strtof(s, &end);
    return end == s ? prev : v;  // nothing parsed: the variable keeps its value
}

template <typename T>
int64_t read_mm2dense(const char *fname, int64_t count, T *out)
{
    if (!fname) return -1;
    if (count < 0) return -2;
    if (!out && count > 0) return -3;
    FILE *f = fopen(fname, "r");
    if (!f) return LINOMPSS_ERR_IO;
    char buf[1024];
    T el = T(0);
    int64_t i = 0;
    while (i < count && fgets(buf, sizeof buf, f)) {
        if (buf[0] == '#') continue;
        el = parse<T>(buf, el);
        out[i++] = el;
    }
    fclose(f);
    return i;
}

// "# rows: m" / "# columns: n" of the header; -1 where absent.
int header_dims(FILE *f, int *m, int *n)
{
    char buf[1024];
    *m = *n = -1;
    long pos = ftell(f);
    while (fgets(buf, sizeof buf, f)) {
        if (buf[0] != '#') break;
        if (!strncmp(buf, "# rows:", 7)) *m = atoi(buf + 7);
        else if (!strncmp(buf, "# columns:", 10)) *n = atoi(buf + 10);
        pos = ftell(f);
    }
    fseek(f, pos, SEEK_SET);  // back to the first value line
    return (*m >= 0 && *n >= 0) ? 0 : LINOMPSS_ERR_IO;
}

template <typename T>
int read_dense2mm(const char *fname, int m, int n, T *A, int lda)
{
    if (!fname) return -1;
    if (m < 0) return -2;
    if (n < 0) return -3;
    if (!A && m > 0 && n > 0) return -4;
    if (lda < (m > 1 ? m : 1)) return -5;
    FILE *f = fopen(fname, "r");
    if (!f) return LINOMPSS_ERR_IO;
    int fm = 0, fn = 0;
    if (header_dims(f, &fm, &fn) || fm != m || fn != n) {
        fclose(f);
        return LINOMPSS_ERR_IO;
    }
    char buf[1024];
    T el = T(0);
    int64_t k = 0;
    const int64_t tot = (int64_t)m * n;
    while (k < tot && fgets(buf, sizeof buf, f)) {
        if (buf[0] == '#') continue;
        el = parse<T>(buf, el);
        A[(k % n) * (int64_t)lda + k / n] = el;  // value k is (k / n, k % n), row-major
        ++k;
    }
    fclose(f);
    return k == tot ? 0 : LINOMPSS_ERR_IO;
}

}  // namespace

extern "C" int linompss_fprint_ddense2mm(const char *fname, const char *name, int m, int n, const double *A, int lda)
{
    return fprint_dense2mm<double>(fname, name, m, n, A, lda);
}
extern "C" int linompss_fprint_sdense2mm(const char *fname, const char *name, int m, int n, const float *A, int lda)
{
    return fprint_dense2mm<float>(fname, name, m, n, A, lda);
}
extern "C" int64_t linompss_read_mm2ddense(const char *fname, int64_t count, double *out)
{
    return read_mm2dense<double>(fname, count, out);
}
extern "C" int64_t linompss_read_mm2sdense(const char *fname, int64_t count, float *out)
{
    return read_mm2dense<float>(fname, count, out);
}
extern "C" int linompss_dense2mm_dims(const char *fname, int *m, int *n)
{
    if (!fname) return -1;
    if (!m || !n) return -2;
    FILE *f = fopen(fname, "r");
    if (!f) return LINOMPSS_ERR_IO;
    const int rc = header_dims(f, m, n);
    fclose(f);
    return rc;
}
extern "C" int linompss_read_ddense2mm(const char *fname, int m, int n, double *A, int lda)
{
    return read_dense2mm<double>(fname, m, n, A, lda);
}
extern "C" int linompss_read_sdense2mm(const char *fname, int m, int n, float *A, int lda)
{
    return read_dense2mm<float>(fname, m, n, A, lda);
}
