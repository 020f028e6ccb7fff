// speed.cpp -- 1D speed mode of the harness (the tcfft_half_speed /
// cufft_half_speed binaries of the reference, speed.cpp:17-66).
//
//   <backend>_half_speed [-n N] [-b BATCH] [-m MAX_ITER] [-t SECONDS]
//
// Wall-clock time of doit(iter) (iter in-place transforms and one device
// synchronisation) for iter = 1, 2, 4, ... until one round takes longer than
// t_min seconds (4 in the reference; -t changes it, an addition) or iter
// exceeds MAX_ITER.  stdout: "n: %d, n_batch: %d, iter: %d, time per iter:
// %lf\n" (speed.cpp:64), seconds per transform of the whole batch.
#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "harness.h"
#include "tcfft_oracle.h"

int main(int argc, char **argv) {
    int n = 65536, n_batch = 1, max_times = 1 << 30;
    double t_min = 4.0;
    int c;
    while ((c = getopt(argc, argv, "n:b:m:t:")) != -1) {
        if (c == 'n') n = std::atoi(optarg);
        else if (c == 'b') n_batch = std::atoi(optarg);
        else if (c == 'm') max_times = std::atoi(optarg);
        else if (c == 't') t_min = std::atof(optarg);
        else std::printf("unkown option %c\n", optopt);
    }
    if (n < 1 || n_batch < 1) {
        std::fprintf(stderr, "invalid size n=%d batch=%d\n", n, n_batch);
        return 2;
    }
    std::vector<double> data((size_t)2 * n * n_batch);
    oracle_generate_data(data.data(), n, n_batch, 42);
    if (!setup(data.data(), n, n_batch)) return 1;
    double run_time = 0.0;
    int iter = 0;
    for (int i = 1; i <= max_times; i <<= 1) {
        const auto t0 = std::chrono::steady_clock::now();
        doit(i);
        run_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        iter = i;
        if (run_time > t_min) break;
        if (i > max_times / 2) break;  // the next doubling would pass MAX_ITER (and overflow at 2^30)
    }
    std::vector<double> result(data.size());
    finalize(result.data());
    std::printf("n: %d, n_batch: %d, iter: %d, time per iter: %lf\n", n, n_batch, iter, run_time / iter);
    return 0;
}
