// accuracy.cpp -- 1D accuracy mode of the harness (the tcfft_half_accuracy /
// cufft_half_accuracy binaries of the reference, accuracy.cpp:43-81).
//
//   <backend>_half_accuracy [-n N] [-b BATCH] [-s SEED]
//
// Inputs: the reference's generator (glibc srand/rand, U[0, 1e-4] per
// component, doit.h:7-19), restated in the FP64 checker library.  Standard:
// the checker's FP64 transform of the same double data (FFTW double in the
// reference, accuracy.cpp:11-23).  stdout: the reference's statistic, the
// mean over components of min(1, |t - s| / |s|) (accuracy.cpp:25-41), as
// "%e\n"; stderr: the relative L2 error over the whole batch as
// "rel_l2 %e".
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "harness.h"
#include "tcfft_oracle.h"

int main(int argc, char **argv) {
    int n = 65536, n_batch = 1, seed = 42;  // accuracy.cpp:45 defaults
    int c;
    while ((c = getopt(argc, argv, "n:b:s:")) != -1) {
        if (c == 'n') n = std::atoi(optarg);
        else if (c == 'b') n_batch = std::atoi(optarg);
        else if (c == 's') seed = std::atoi(optarg);
        else std::printf("unkown option %c\n", optopt);  // the reference's message, verbatim
    }
    if (n < 1 || n_batch < 1) {
        std::fprintf(stderr, "invalid size n=%d batch=%d\n", n, n_batch);
        return 2;
    }
    const size_t words = (size_t)2 * n * n_batch;
    std::vector<double> data(words), standard(words), tested(words);
    oracle_generate_data(data.data(), n, n_batch, seed);
    oracle_fft(data.data(), standard.data(), n, n_batch, -1);
    if (!setup(data.data(), n, n_batch)) return 1;
    doit(1);
    finalize(tested.data());
    std::printf("%e\n", oracle_get_error(tested.data(), standard.data(), n, n_batch));
    std::fprintf(stderr, "rel_l2 %e\n", oracle_rel_l2(tested.data(), standard.data(), words));
    return 0;
}
