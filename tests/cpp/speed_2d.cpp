// speed_2d.cpp -- 2D speed mode of the harness (the tcfft_half_2d_speed /
// cufft_half_2d_speed binaries of the reference, speed_2d.cpp:17-69).
//
//   <backend>_half_2d_speed [-x NX] [-y NY] [-b BATCH] [-m MAX_ITER] [-t SECONDS]
//
// Timing loop as in speed.cpp.  stdout: "nx: %d, ny: %d, n_batch: %d,
// iter: %d, time per iter: %lf\n" (speed_2d.cpp:67).
#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "harness.h"
#include "tcfft_oracle.h"

int main(int argc, char **argv) {
    int nx = 256, ny = 256, n_batch = 1, max_times = 1 << 30;
    double t_min = 4.0;
    int c;
    while ((c = getopt(argc, argv, "x:y:b:m:t:")) != -1) {
        if (c == 'x') nx = std::atoi(optarg);
        else if (c == 'y') ny = std::atoi(optarg);
        else if (c == 'b') n_batch = std::atoi(optarg);
        else if (c == 'm') max_times = std::atoi(optarg);
        else if (c == 't') t_min = std::atof(optarg);
        else std::printf("unkown option %c\n", optopt);
    }
    if (nx < 1 || ny < 1 || n_batch < 1) {
        std::fprintf(stderr, "invalid size nx=%d ny=%d batch=%d\n", nx, ny, n_batch);
        return 2;
    }
    std::vector<double> data((size_t)2 * nx * ny * n_batch);
    oracle_generate_data(data.data(), nx * ny, n_batch, 42);
    if (!setup_2d(data.data(), nx, ny, n_batch)) return 1;
    double run_time = 0.0;
    int iter = 0;
    for (int i = 1; i <= max_times; i <<= 1) {
        const auto t0 = std::chrono::steady_clock::now();
        doit_2d(i);
        run_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        iter = i;
        if (run_time > t_min) break;
        if (i > max_times / 2) break;  // the next doubling would pass MAX_ITER (and overflow at 2^30)
    }
    std::vector<double> result(data.size());
    finalize_2d(result.data());
    std::printf("nx: %d, ny: %d, n_batch: %d, iter: %d, time per iter: %lf\n", nx, ny, n_batch, iter,
                run_time / iter);
    return 0;
}
