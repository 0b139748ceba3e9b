// lx_spdsolv_cli.cpp -- the reference's dspdsolv / sspdsolv program
// (src/spdsolv/spdsolv.c:36-63) over liblinompss's C-ABI.
//
//   dspdsolv m n b [reps] [check]              (spdsolv_config.c:14-40)
//
// m is the order (padded up to a multiple of b, with the reference's
// warning), n the number of right-hand sides, b the tile.  Setup
// (spdsolv_setup.c:12-34): A = GENMAT_SPD(m), X = GENMAT(m, n), B = A X, all
// drawn and multiplied on the device.  Each repetition solves A Y = B in
// place in HBM (timed: factor + forward + backward sweeps until completed,
// SPDSOLV_MAIN, spdsolv_main.c:10-20).  check > 0 prints the reference's
// "residual norm" | ||X||_inf - ||Y||_inf | / ||Y||_inf (spdsolv_check.c:7-20,
// MAT_NORMDIFF, densutil.c:134-143) and exits 1 above 1e-6 (or for a NaN).
// Exit 1 also for a usage error, 2 when the matrices cannot be set up, 3
// when the solve reports an error.  Writes ompss-spdsolv.stats.
//
// One difference: every repetition solves the original system (A and B are
// restored from device copies before each one, outside the timed region);
// the reference re-solves its own output from the second repetition on.
//
// Built once per precision: -DLX_CLI_DTYPE=0|1 (FP64 / FP32).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "linompss.h"
#include "lx_cli_log.hpp"

#ifndef LX_CLI_DTYPE
#define LX_CLI_DTYPE 0
#endif
typedef std::conditional<LX_CLI_DTYPE == 0, double, float>::type fp_t;

// LAPACK lange('i'): the largest row sum of |X| (m x n, leading dimension m)
static double norm_inf(int m, int n, const std::vector<fp_t> &X)
{
    std::vector<double> rows((size_t)m, 0.0);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) rows[i] += std::fabs((double)X[(size_t)j * m + i]);
    double mx = 0.0;
    for (double r : rows) mx = std::max(mx, r);
    return mx;
}

int main(int argc, char *argv[])
{
    // spdsolv_config.c:14-40
    if (argc < 4) {
        std::printf("usage: %s m n b [reps] [check]\n", argv[0]);
        return 1;
    }
    int m = std::atoi(argv[1]);
    const int n = std::atoi(argv[2]), b = std::atoi(argv[3]);
    if (b > 0 && m % b) {
        m = ((m + b - 1) / b) * b;
        std::printf("warn: padding m to %i\n", m);
    }
    int reps = 1, check = 0;
    if (argc > 4) {
        reps = std::atoi(argv[4]);
        if (argc > 5) check = std::atoi(argv[5]);
    }

    const size_t esz = sizeof(fp_t);
    const size_t abytes = esz * (size_t)std::max(m, 1) * (size_t)std::max(m, 1);
    const size_t bbytes = esz * (size_t)std::max(m, 1) * (size_t)std::max(n, 1);
    const long seed = lx_cli_seed();
    void *dA = nullptr, *dA0 = nullptr, *dB = nullptr, *dB0 = nullptr, *dX = nullptr;
    bool ok = m > 0 && n > 0 && b > 0 && linompss_device_count() > 0;
    if (ok) {
        dA = linompss_dev_alloc(abytes);
        dA0 = linompss_dev_alloc(abytes);
        dB = linompss_dev_alloc(bbytes);
        dB0 = linompss_dev_alloc(bbytes);
        dX = linompss_dev_alloc(bbytes);
        ok = dA && dA0 && dB && dB0 && dX;
    }
    // spdsolv_setup.c:12-34: GENMAT_SPD and GENMAT both seed srand48 with
    // time(NULL), so both start from the same seed
    ok = ok && !linompss_gen_drand48_spd(LX_CLI_DTYPE, m, seed, dA, m) &&
         !linompss_gen_drand48(LX_CLI_DTYPE, m, n, seed, dX, m) &&
         !linompss_gemm_dev(LX_CLI_DTYPE, 'N', 'N', m, n, m, 1.0, dA, m, dX, m, 0.0, dB, m, nullptr) &&
         !linompss_memcpy(dA0, dA, abytes, 3) && !linompss_memcpy(dB0, dB, bbytes, 3);
    if (!ok) {
        std::fprintf(stderr, "err: allocating matrix\n");
        return 2;
    }

    lx_cli_log log(reps, "spdsolv");
    for (int r = 0; r < reps; r++) {
        if (r > 0 && (linompss_memcpy(dA, dA0, abytes, 3) || linompss_memcpy(dB, dB0, bbytes, 3))) {
            std::fprintf(stderr, "err: restoring matrices\n");
            return 3;
        }
        log.start_rep();
        const int info = linompss_posv_dev(LX_CLI_DTYPE, m, n, b, dA, m, dB, m, nullptr, nullptr);
        log.end_rep();
        if (info < 0) {
            std::fprintf(stderr, "err: solve returned %d\n", info);
            return 3;
        }
        linompss_exec_stats st;
        if (linompss_last_exec_stats(&st) == 0) log.threads = st.grid;
    }

    // spdsolv_check.c:7-20
    int ret = 0;
    if (check != 0) {
        std::vector<fp_t> X((size_t)m * n), Y((size_t)m * n);
        if (linompss_memcpy(X.data(), dX, esz * X.size(), 2) || linompss_memcpy(Y.data(), dB, esz * Y.size(), 2)) {
            std::fprintf(stderr, "err: reading the solution\n");
            ret = 1;
        } else {
            const double ex = norm_inf(m, n, X), ey = norm_inf(m, n, Y);
            const double norm2 = std::fabs(ex - ey) / ey;
            std::printf("check: residual norm = %.16e\n", norm2);
            ret = !(norm2 <= lx_cli_tol(1e-6));
        }
    }

    log.dump();
    for (void *p : {dA, dA0, dB, dB0, dX}) linompss_dev_free(p);
    return ret;
}
