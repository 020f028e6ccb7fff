// accuracy_2d.cpp -- 2D accuracy mode of the harness (the
// tcfft_half_2d_accuracy / cufft_half_2d_accuracy binaries of the
// reference, accuracy_2d.cpp:43-84).
//
//   <backend>_half_2d_accuracy [-x NX] [-y NY] [-b BATCH] [-s SEED]
//
// Images [b][x][y], y contiguous.  Inputs, standard and statistic as in
// accuracy.cpp, over nx*ny points per image (the generator fills each image
// as one line of nx*ny points, doit_2d.h:7-19; the statistic divides by
// nx*ny*batch, accuracy_2d.cpp:25-41).  stdout "%e\n"; stderr "rel_l2 %e".
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "harness.h"
#include "tcfft_oracle.h"

int main(int argc, char **argv) {
    int nx = 256, ny = 256, n_batch = 1, seed = 42;  // accuracy_2d.cpp:45 defaults
    int c;
    while ((c = getopt(argc, argv, "x:y:b:s:")) != -1) {
        if (c == 'x') nx = std::atoi(optarg);
        else if (c == 'y') ny = std::atoi(optarg);
        else if (c == 'b') n_batch = std::atoi(optarg);
        else if (c == 's') seed = std::atoi(optarg);
        else std::printf("unkown option %c\n", optopt);
    }
    if (nx < 1 || ny < 1 || n_batch < 1) {
        std::fprintf(stderr, "invalid size nx=%d ny=%d batch=%d\n", nx, ny, n_batch);
        return 2;
    }
    const size_t words = (size_t)2 * nx * ny * n_batch;
    std::vector<double> data(words), standard(words), tested(words);
    oracle_generate_data(data.data(), nx * ny, n_batch, seed);
    oracle_fft_2d(data.data(), standard.data(), nx, ny, n_batch, -1);
    if (!setup_2d(data.data(), nx, ny, n_batch)) return 1;
    doit_2d(1);
    finalize_2d(tested.data());
    std::printf("%e\n", oracle_get_error(tested.data(), standard.data(), nx * ny, n_batch));
    std::fprintf(stderr, "rel_l2 %e\n", oracle_rel_l2(tested.data(), standard.data(), words));
    return 0;
}
