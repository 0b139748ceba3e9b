// lx_chol_cli.cpp -- the reference's dchol_{rl,ll} / schol_{rl,ll} programs
// (src/cholesky/dense/chol.c:40-84) over liblinompss's C-ABI.
//
//   dchol_rl n ts [bs] [reps] [check]          (chol_config.c:17-47)
//
// n is the order, ts the tile, bs the sub-block (unused by the RL/LL
// factors, as in the reference), reps the timed repetitions, check > 0 the
// factor check.  The matrix is GENMAT_SPD (chol_setup.c:16-38) drawn on the
// device; each repetition factors it in place in HBM (timed: the factor
// launch until it completed) and the check compares L L^T with the original
// (chol_check.c:12-42): "norm of difference" above 1e-4, or NaN, exits 1.
// Exit 1 also for a usage error, 2 when the matrices cannot be set up, 3 when
// the factor reports an error.  Writes ompss-chol.stats in the current
// directory.
//
// One difference: every repetition factors the original matrix (it is
// restored from a device copy before each one, outside the timed region);
// the reference refactors its own output from the second repetition on.
//
// Built once per precision and variant: -DLX_CLI_DTYPE=0|1 (FP64 / FP32),
// -DLX_CLI_VARIANT=0|1 (right / left looking), as the reference picks them
// at build time (-DDOUBLE_PRECISION, -DUSE_RL, ...).
#include <cstdio>
#include <cstdlib>

#include "linompss.h"
#include "lx_cli_log.hpp"

#ifndef LX_CLI_DTYPE
#define LX_CLI_DTYPE 0
#endif
#ifndef LX_CLI_VARIANT
#define LX_CLI_VARIANT 0
#endif

int main(int argc, char *argv[])
{
    // chol_config.c:17-47
    if (argc < 3) {
        std::printf("usage: %s n ts [bs] [reps] [check]\n", argv[0]);
        return 1;
    }
    const int m = std::atoi(argv[1]), ts = std::atoi(argv[2]);
    int bs = ts, reps = 1, check = 0;
    if (argc > 3) {
        bs = std::atoi(argv[3]);
        if (argc > 4) {
            reps = std::atoi(argv[4]);
            if (argc > 5) check = std::atoi(argv[5]);
        }
    }
    (void)bs;
    const size_t esz = LX_CLI_DTYPE == 0 ? sizeof(double) : sizeof(float);
    const size_t bytes = esz * (size_t)(m > 0 ? m : 1) * (size_t)(m > 0 ? m : 1);

    // chol_setup.c:16-38: A = GENMAT_SPD(m, m); Aorig a copy when checking
    void *dA = m > 0 && linompss_device_count() > 0 ? linompss_dev_alloc(bytes) : nullptr;
    void *dA0 = dA && (check || reps > 1) ? linompss_dev_alloc(bytes) : nullptr;
    if (!dA || ((check || reps > 1) && !dA0) || linompss_gen_drand48_spd(LX_CLI_DTYPE, m, lx_cli_seed(), dA, m) ||
        (dA0 && linompss_memcpy(dA0, dA, bytes, 3))) {
        std::fprintf(stderr, "err: allocating matrix\n");
        return 2;
    }

    lx_cli_log log(reps, "chol");
    for (int r = 0; r < reps; r++) {
        if (r > 0 && linompss_memcpy(dA, dA0, bytes, 3)) {
            std::fprintf(stderr, "err: restoring matrix\n");
            return 3;
        }
        log.start_rep();
        const int info = linompss_chol_dev(LX_CLI_DTYPE, LX_CLI_VARIANT, m, ts, dA, m, nullptr, nullptr);
        log.end_rep();
        if (info < 0) {
            std::fprintf(stderr, "err: factor returned %d\n", info);
            return 3;
        }
        linompss_exec_stats st;
        if (linompss_last_exec_stats(&st) == 0) log.threads = st.grid;
    }

    // chol_check.c:12-42
    int ret = 0;
    if (check > 0) {
        std::printf("check: potrf...");
        std::fflush(nullptr);
        double norm_diff = 0.0;
        const int rc = linompss_relerr_lower_dev(LX_CLI_DTYPE, m, dA, m, dA0, m, &norm_diff);
        if (rc) {
            std::fprintf(stderr, "err: check returned %d\n", rc);
            ret = 1;
        } else {
            std::printf("norm of difference: %22.16g\n", norm_diff);
            ret = !(norm_diff <= lx_cli_tol(0.0001));  // a NaN fails too
        }
    }

    log.dump();
    linompss_dev_free(dA);
    if (dA0) linompss_dev_free(dA0);
    return ret;
}
