// lx_cli_log.hpp -- the run log of the CLI front ends: the reference's
// ompss-<app>.stats file (src/util/file_log.c:9-51).  Repetitions are timed
// with gettimeofday, accumulated in whole microseconds, and written as the
// mean in milliseconds:
//   <nodename> <sysname> <release> <machine>
//   time: %.5f ms
//   reps: %d
//   threads: %d
// "threads" is the number of persistent device workgroups that ran the last
// repetition (the reference means its OmpSs worker threads; its assignment
// is commented out, file_log.c:14-15, so it always writes 0 there).
#pragma once

#include <sys/time.h>
#include <sys/utsname.h>

#include <cstdio>

struct lx_cli_log {
    char name[32];
    unsigned long elapsed_us = 0;
    timeval start{}, end{};
    int reps = 0;
    int threads = 0;
    utsname info{};

    lx_cli_log(int r, const char *app) : reps(r)
    {
        uname(&info);
        std::snprintf(name, sizeof name, "ompss-%s.stats", app);
    }
    void start_rep() { gettimeofday(&start, nullptr); }
    void end_rep()
    {
        gettimeofday(&end, nullptr);
        elapsed_us += (unsigned long)((end.tv_sec - start.tv_sec) * 1e6 + (end.tv_usec - start.tv_usec));
    }
    int dump() const
    {
        FILE *f = std::fopen(name, "w");
        if (!f) return 1;
        std::fprintf(f, "%s %s %s %s\n", info.nodename, info.sysname, info.release, info.machine);
        std::fprintf(f, "time: %.5f ms\n", elapsed_us / ((double)reps * 1e3));
        std::fprintf(f, "reps: %d\n", reps);
        std::fprintf(f, "threads: %d\n", threads);
        std::fclose(f);
        return 0;
    }
};

// The matrices' seed: srand48(time(NULL)) in the reference
// (genmat.c:56,112); LINOMPSS_SEED overrides it for reproducible runs.
#include <cstdlib>
#include <ctime>
inline long lx_cli_seed()
{
    const char *s = std::getenv("LINOMPSS_SEED");
    return s && *s ? std::strtol(s, nullptr, 10) : (long)std::time(nullptr);
}

// The pass threshold of the check (chol_check.c:36: 1e-4; spdsolv_check.c:17:
// 1e-6); LINOMPSS_CHECK_TOL tightens or loosens it.
inline double lx_cli_tol(double dflt)
{
    const char *s = std::getenv("LINOMPSS_CHECK_TOL");
    return s && *s ? std::strtod(s, nullptr) : dflt;
}
