// cli_1d.cpp -- the reference's 1D harness flow on the drop-in header.
//
// Mirrors accuracy.cpp / speed.cpp (flags -n -b -s / -n -b -m, same output
// lines) over the backend seam of tcfft_doit_half.cpp: host-side digit
// reversal by plan.radices, double -> half staging, tcfftExec in place,
// synchronize, half -> double.  The FP64 standard and the error statistic
// come from the checker library (oracle/), standing in for FFTW double.
//
//   cli_1d accuracy [-n N] [-b B] [-s SEED] [-r]   -> "%e" (accuracy.cpp:79), -r adds rel-L2
//   cli_1d speed    [-n N] [-b B] [-m MAX] [-t S]  -> "n: .., n_batch: .., iter: .., time per iter: .."
//   cli_1d sweep                                  -> accuracy lines for the reference's test.sh sizes
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unistd.h>
#include <vector>

#include "tcfft_half.h"
#include "tcfft_oracle.h"

namespace {

struct Backend {  // setup / doit / finalize seam (doit.h:2-4)
    tcfftHandle plan{};
    half *dev = nullptr;
    std::vector<half> host;
    int n = 0, batch = 0;

    bool setup(const std::vector<double> &data, int n_, int batch_) {
        n = n_;
        batch = batch_;
        tcfftCreate(&plan, n, batch);
        if (!plan.impl) return false;
        std::vector<int> rev(n);
        tcfft_gen_rev(n, plan.radices, plan.n_radices, rev.data());
        host.resize((size_t)2 * n * batch);
        for (int b = 0; b < batch; ++b)
            for (int i = 0; i < n; ++i) {
                host[2 * ((size_t)b * n + i)] = __float2half((float)data[2 * ((size_t)b * n + rev[i])]);
                host[2 * ((size_t)b * n + i) + 1] = __float2half((float)data[2 * ((size_t)b * n + rev[i]) + 1]);
            }
        if (hipMalloc(&dev, host.size() * sizeof(half)) != hipSuccess) return false;
        return hipMemcpy(dev, host.data(), host.size() * sizeof(half), hipMemcpyHostToDevice) == hipSuccess;
    }
    void doit(int iter) {
        for (int i = 0; i < iter; ++i) tcfftExec(plan, dev);
        (void)hipDeviceSynchronize();
    }
    void finalize(std::vector<double> &out) {
        (void)hipMemcpy(host.data(), dev, host.size() * sizeof(half), hipMemcpyDeviceToHost);
        out.resize(host.size());
        for (size_t i = 0; i < host.size(); ++i) out[i] = __half2float(host[i]);
    }
    ~Backend() {
        if (dev) (void)hipFree(dev);
        tcfftDestroy(&plan);
    }
};

int accuracy(int n, int batch, int seed, bool rel) {
    std::vector<double> data((size_t)2 * n * batch), standard(data.size()), tested;
    oracle_generate_data(data.data(), n, batch, seed);
    oracle_fft(data.data(), standard.data(), n, batch, -1);
    Backend be;
    if (!be.setup(data, n, batch)) return 1;
    be.doit(1);
    be.finalize(tested);
    std::printf("%e\n", oracle_get_error(tested.data(), standard.data(), n, batch));
    if (rel) std::printf("rel_l2 %e\n", oracle_rel_l2(tested.data(), standard.data(), tested.size()));
    return 0;
}

int speed(int n, int batch, int max_times, double t_min) {
    std::vector<double> data((size_t)2 * n * batch);
    oracle_generate_data(data.data(), n, batch, 42);
    Backend be;
    if (!be.setup(data, n, batch)) return 1;
    double run_time = 0;
    int iter = 0;
    for (int i = 1; i <= max_times; i <<= 1) {
        auto t0 = std::chrono::steady_clock::now();
        be.doit(i);
        run_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        iter = i;
        if (run_time > t_min) break;
    }
    std::printf("n: %d, n_batch: %d, iter: %d, time per iter: %lf\n", n, batch, iter, run_time / iter);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s accuracy|speed|sweep [-n N] [-b B] [-s SEED] [-m MAX] [-t SEC] [-r]\n", argv[0]);
        return 2;
    }
    std::string mode = argv[1];
    int n = 65536, batch = 1, seed = 42, max_times = 1 << 30;
    double t_min = 4.0;
    bool rel = false;
    optind = 2;
    int c;
    while ((c = getopt(argc, argv, "n:b:s:m:t:r")) != -1) {
        switch (c) {
        case 'n': n = std::atoi(optarg); break;
        case 'b': batch = std::atoi(optarg); break;
        case 's': seed = std::atoi(optarg); break;
        case 'm': max_times = std::atoi(optarg); break;
        case 't': t_min = std::atof(optarg); break;
        case 'r': rel = true; break;
        default: return 2;
        }
    }
    if (mode == "accuracy") return accuracy(n, batch, seed, rel);
    if (mode == "speed") return speed(n, batch, max_times, t_min);
    if (mode == "sweep") {  // the accuracy sizes of the reference's test.sh
        const int small[] = {256, 512, 1024};
        const int large[] = {131072, 262144, 524288};
        for (int s : small) {
            std::printf("Accuracy test: tcFFT-MI355X, FFT size: %d batchsize:32\n", s);
            if (accuracy(s, 32, 42, true)) return 1;
        }
        for (int s : large) {
            std::printf("Accuracy test: tcFFT-MI355X, FFT size: %d batchsize:1\n", s);
            if (accuracy(s, 1, 42, true)) return 1;
        }
        return 0;
    }
    std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
}
