// lx_ompss_api.cpp -- the reference's own BLAS-3 and tile-task symbols, with
// its ompssblas_t / ompsslapack_t descriptors, over this library's entry
// points (host code).
//
// The reference passes BLAS flags as 8-byte unions by value
// (src/fp/fpmatr.h:14-35): its MKL build stores a CBLAS enumeration in the
// int member (fpmatr.c:14-27: CblasNoTrans = 111 ... CblasRight = 142), its
// Fortran-BLAS build points the char * member at "N", "T", "L", "U", "R"
// (fpmatr.c:36-47); the LAPACK union stores 'U' / 'L' in its char member or
// points at "U" / "L".  A descriptor is decoded as an enumeration when its
// low 32 bits hold one of the CBLAS values, else as the string it points to,
// so a caller built against either flavour links unchanged.
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "linompss.h"

namespace {

// One letter per flag: N/T (trans), U/L (uplo), N/U (diag), L/R (side); 0 when
// the descriptor is neither a CBLAS value nor a readable string.
char blas_flag(ompssblas_t d)
{
    uint64_t v = 0;
    std::memcpy(&v, &d, sizeof d < sizeof v ? sizeof d : sizeof v);
    switch ((int32_t)(uint32_t)v) {
    case 111: return 'N';  // CblasNoTrans
    case 112: return 'T';  // CblasTrans
    case 113: return 'T';  // CblasConjTrans (real data)
    case 121: return 'U';  // CblasUpper
    case 122: return 'L';  // CblasLower
    case 131: return 'N';  // CblasNonUnit
    case 132: return 'U';  // CblasUnit
    case 141: return 'L';  // CblasLeft
    case 142: return 'R';  // CblasRight
    default: break;
    }
    if (v < 4096 || !d.s) return 0;  // not a CBLAS value and not a plausible pointer
    return (char)std::toupper((unsigned char)d.s[0]);
}

char lapack_flag(ompsslapack_t d)
{
    uint64_t v = 0;
    std::memcpy(&v, &d, sizeof d < sizeof v ? sizeof d : sizeof v);
    if (v < 256) return (char)std::toupper((int)(v & 0xff));  // the char member (upper bytes zero)
    return d.s ? (char)std::toupper((unsigned char)d.s[0]) : 0;
}

ompssblas_t blas_str(const char *s)
{
    ompssblas_t d;
    std::memset(&d, 0, sizeof d);
    d.s = const_cast<char *>(s);
    return d;
}

ompsslapack_t lapack_str(const char *s)
{
    ompsslapack_t d;
    std::memset(&d, 0, sizeof d);
    d.s = const_cast<char *>(s);
    return d;
}

// The reference's void entry points have no return value; argument and device
// errors are reported the way BLAS's xerbla reports them.
void report(const char *fn, int rc)
{
    if (rc != 0) std::fprintf(stderr, "linompss: %s returned %d\n", fn, rc);
}

template <typename T>
void potrf_tile(const char *fn, ompsslapack_t uplo, int n, T *A, int lda)
{
    const char u = lapack_flag(uplo);
    if (u != 'L' && u != 'U') return report(fn, -1);
    if (n < 0) return report(fn, -2);
    if (lda < (n > 1 ? n : 1)) return report(fn, -4);
    if (n == 0) return;
    const int b = n < 512 ? n : 512;
    auto chol = [&](T *P, int ld) {
        if constexpr (sizeof(T) == 8) return ompss_dchol_rl(n, b, 0, P, ld);
        else return ompss_schol_rl(n, b, 0, P, ld);
    };
    if (u == 'L') {
        const int rc = chol(A, lda);  // info > 0 is discarded, as the reference discards dpotrf's
        return report(fn, rc < 0 ? rc : 0);
    }
    // A = U^T U: factor the lower triangle of A^T, write U back
    std::vector<T> L((size_t)n * n);
    for (int j = 0; j < n; ++j)
        for (int i = j; i < n; ++i) L[(size_t)j * n + i] = A[(size_t)i * lda + j];
    const int rc = chol(L.data(), n);
    if (rc < 0) return report(fn, rc);
    for (int j = 0; j < n; ++j)
        for (int i = j; i < n; ++i) A[(size_t)i * lda + j] = L[(size_t)j * n + i];
}

}  // namespace

// ---- descriptor instances of the reference (fpmatr.c:36-47, the Fortran-BLAS
// flavour; OMPSSBLAS_ROWMAJOR / COLMAJOR are zero there as here) ------------
extern "C" {
ompssblas_t OMPSSBLAS_ROWMAJOR = blas_str(nullptr);
ompssblas_t OMPSSBLAS_COLMAJOR = blas_str(nullptr);
ompssblas_t OMPSSBLAS_TRANSP = blas_str("T");
ompssblas_t OMPSSBLAS_NTRANSP = blas_str("N");
ompssblas_t OMPSSBLAS_UPPERTRIANG = blas_str("U");
ompssblas_t OMPSSBLAS_LOWERTRIANG = blas_str("L");
ompssblas_t OMPSSBLAS_NTRIANG = blas_str("x");
ompssblas_t OMPSSBLAS_DIAGUNIT = blas_str("U");
ompssblas_t OMPSSBLAS_NDIAGUNIT = blas_str("N");
ompssblas_t OMPSSBLAS_LEFT = blas_str("L");
ompssblas_t OMPSSBLAS_RIGHT = blas_str("R");
ompsslapack_t OMPSSLAPACK_UPPERTRIANG = lapack_str("U");
ompsslapack_t OMPSSLAPACK_LOWERTRIANG = lapack_str("L");
}

// ---- blocked BLAS-3 (src/gemm/ompss_mm.h:23,27; src/syrk/ompss_syrk.h:20,22;
// src/trsm/ompss_trsm.h:20,23) --------------------------------------------------
extern "C" int ompss_dgemm(ompssblas_t transa, ompssblas_t transb, int b, int d, int c, int m, int n, int k,
                           double alpha, const double *A, int lda, const double *B, int ldb, double beta, double *C,
                           int ldc)
{
    const char ta = blas_flag(transa), tb = blas_flag(transb);
    if (ta != 'N' && ta != 'T') return -1;
    if (tb != 'N' && tb != 'T') return -2;
    return linompss_dgemm(ta, tb, b, d, c, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

extern "C" int ompss_sgemm(ompssblas_t transa, ompssblas_t transb, int b, int d, int c, int m, int n, int k,
                           float alpha, const float *A, int lda, const float *B, int ldb, float beta, float *C,
                           int ldc)
{
    const char ta = blas_flag(transa), tb = blas_flag(transb);
    if (ta != 'N' && ta != 'T') return -1;
    if (tb != 'N' && tb != 'T') return -2;
    return linompss_sgemm(ta, tb, b, d, c, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

extern "C" void ompss_dsyrk(ompssblas_t uplo, ompssblas_t trans, int b, int n, int k, double alpha, double *A,
                            int lda, double beta, double *C, int ldc)
{
    const char u = blas_flag(uplo), t = blas_flag(trans);
    if (u != 'L' && u != 'U') return report("ompss_dsyrk", -1);
    if (t != 'N' && t != 'T') return report("ompss_dsyrk", -2);
    report("ompss_dsyrk", linompss_dsyrk(u, t, b, n, k, alpha, A, lda, beta, C, ldc));
}

extern "C" void ompss_ssyrk(ompssblas_t uplo, ompssblas_t trans, int b, int n, int k, float alpha, float *A, int lda,
                            float beta, float *C, int ldc)
{
    const char u = blas_flag(uplo), t = blas_flag(trans);
    if (u != 'L' && u != 'U') return report("ompss_ssyrk", -1);
    if (t != 'N' && t != 'T') return report("ompss_ssyrk", -2);
    report("ompss_ssyrk", linompss_ssyrk(u, t, b, n, k, alpha, A, lda, beta, C, ldc));
}

extern "C" void ompss_dtrsm(ompssblas_t side, ompssblas_t uplo, ompssblas_t trans, ompssblas_t diag, int m, int n,
                            int b, double alpha, double *A, int lda, double *B, int ldb)
{
    const char s = blas_flag(side), u = blas_flag(uplo), t = blas_flag(trans), g = blas_flag(diag);
    if (s != 'L' && s != 'R') return report("ompss_dtrsm", -1);
    if (u != 'L' && u != 'U') return report("ompss_dtrsm", -2);
    if (t != 'N' && t != 'T') return report("ompss_dtrsm", -3);
    if (g != 'N' && g != 'U') return report("ompss_dtrsm", -4);
    report("ompss_dtrsm", linompss_dtrsm(s, u, t, g, m, n, b, alpha, A, lda, B, ldb));
}

extern "C" void ompss_strsm(ompssblas_t side, ompssblas_t uplo, ompssblas_t trans, ompssblas_t diag, int m, int n,
                            int b, float alpha, float *A, int lda, float *B, int ldb)
{
    const char s = blas_flag(side), u = blas_flag(uplo), t = blas_flag(trans), g = blas_flag(diag);
    if (s != 'L' && s != 'R') return report("ompss_strsm", -1);
    if (u != 'L' && u != 'U') return report("ompss_strsm", -2);
    if (t != 'N' && t != 'T') return report("ompss_strsm", -3);
    if (g != 'N' && g != 'U') return report("ompss_strsm", -4);
    report("ompss_strsm", linompss_strsm(s, u, t, g, m, n, b, alpha, A, lda, B, ldb));
}

// ---- tile tasks of libbblas (src/blocked_blas/tasks_potrf.h:23, tasks_trsm.h:24,
// tasks_syrk.h:21, task_gemm.h:27 and their s twins): one tile operation on
// host memory, run on the device and returned synchronously (the reference
// spawns an OmpSs task; p, its priority, has no meaning here) ------------------
extern "C" void task_dpotrf(ompsslapack_t uplo, int n, double *A, int lda, int p)
{
    (void)p;
    potrf_tile<double>("task_dpotrf", uplo, n, A, lda);
}

extern "C" void task_spotrf(ompsslapack_t uplo, int n, float *A, int lda, int p)
{
    (void)p;
    potrf_tile<float>("task_spotrf", uplo, n, A, lda);
}

extern "C" void task_dtrsm(ompssblas_t side, ompssblas_t uplo, ompssblas_t trans, ompssblas_t diag, int bm, int bn,
                           double alpha, double *A, int lda, double *B, int ldb, int p)
{
    (void)p;
    ompss_dtrsm(side, uplo, trans, diag, bm, bn, bm > 0 ? bm : 1, alpha, A, lda, B, ldb);
}

extern "C" void task_strsm(ompssblas_t side, ompssblas_t uplo, ompssblas_t trans, ompssblas_t diag, int bm, int bn,
                           float alpha, float *A, int lda, float *B, int ldb, int p)
{
    (void)p;
    ompss_strsm(side, uplo, trans, diag, bm, bn, bm > 0 ? bm : 1, alpha, A, lda, B, ldb);
}

extern "C" void task_dsyrk(ompssblas_t uplo, ompssblas_t trans, int n, int k, double alpha, double *A, int lda,
                           double beta, double *C, int ldc, int p)
{
    (void)p;
    ompss_dsyrk(uplo, trans, n > 0 ? n : 1, n, k, alpha, A, lda, beta, C, ldc);
}

extern "C" void task_ssyrk(ompssblas_t uplo, ompssblas_t trans, int n, int k, float alpha, float *A, int lda,
                           float beta, float *C, int ldc, int p)
{
    (void)p;
    ompss_ssyrk(uplo, trans, n > 0 ? n : 1, n, k, alpha, A, lda, beta, C, ldc);
}

extern "C" void task_dgemm(ompssblas_t transa, ompssblas_t transb, int bm, int bn, int bk, double alpha, double *A,
                           int lda, double *B, int ldb, double beta, double *C, int ldc, int p)
{
    (void)p;
    report("task_dgemm", ompss_dgemm(transa, transb, 1, 1, 1, bm, bn, bk, alpha, A, lda, B, ldb, beta, C, ldc));
}

extern "C" void task_sgemm(ompssblas_t transa, ompssblas_t transb, int bm, int bn, int bk, float alpha, float *A,
                           int lda, float *B, int ldb, float beta, float *C, int ldc, int p)
{
    (void)p;
    report("task_sgemm", ompss_sgemm(transa, transb, 1, 1, 1, bm, bn, bk, alpha, A, lda, B, ldb, beta, C, ldc));
}
