// cli_2d.cpp -- the reference's 2D harness flow on the drop-in header.
//
// Mirrors accuracy_2d.cpp / speed_2d.cpp (flags -x -y -b -s / -x -y -b -m,
// same output lines) over the seam of tcfft_doit_half_2d.cpp: per-axis
// digit reversal by radices_x / radices_y, half staging, tcfftExec in place.
//
//   cli_2d accuracy [-x NX] [-y NY] [-b B] [-s SEED] [-r]
//   cli_2d speed    [-x NX] [-y NY] [-b B] [-m MAX] [-t S]
//   cli_2d sweep    -> accuracy lines for {256,512} x {256,512,1024} plus 1024 x 1024
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <unistd.h>
#include <vector>

#include "tcfft_half_2d.h"
#include "tcfft_oracle.h"

namespace {

struct Backend2D {
    tcfftHandle plan{};
    half *dev = nullptr;
    std::vector<half> host;

    bool setup(const std::vector<double> &data, int nx, int ny, int batch) {
        tcfftCreate(&plan, nx, ny, batch);
        if (!plan.impl) return false;
        std::vector<int> rx(nx), ry(ny);
        tcfft_gen_rev(nx, plan.radices_x, plan.n_radices_x, rx.data());
        tcfft_gen_rev(ny, plan.radices_y, plan.n_radices_y, ry.data());
        host.resize((size_t)2 * nx * ny * batch);
        for (int b = 0; b < batch; ++b)
            for (int x = 0; x < nx; ++x)
                for (int y = 0; y < ny; ++y) {
                    size_t dst = ((size_t)b * nx + x) * ny + y;
                    size_t src = ((size_t)b * nx + rx[x]) * ny + ry[y];
                    host[2 * dst] = __float2half((float)data[2 * src]);
                    host[2 * dst + 1] = __float2half((float)data[2 * src + 1]);
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
    ~Backend2D() {
        if (dev) (void)hipFree(dev);
        tcfftDestroy(&plan);
    }
};

int accuracy(int nx, int ny, int batch, int seed, bool rel) {
    std::vector<double> data((size_t)2 * nx * ny * batch), standard(data.size()), tested;
    oracle_generate_data(data.data(), nx * ny, batch, seed);
    oracle_fft_2d(data.data(), standard.data(), nx, ny, batch, -1);
    Backend2D be;
    if (!be.setup(data, nx, ny, batch)) return 1;
    be.doit(1);
    be.finalize(tested);
    std::printf("%e\n", oracle_get_error(tested.data(), standard.data(), nx * ny, batch));
    if (rel) std::printf("rel_l2 %e\n", oracle_rel_l2(tested.data(), standard.data(), tested.size()));
    return 0;
}

int speed(int nx, int ny, int batch, int max_times, double t_min) {
    std::vector<double> data((size_t)2 * nx * ny * batch);
    oracle_generate_data(data.data(), nx * ny, batch, 42);
    Backend2D be;
    if (!be.setup(data, nx, ny, batch)) return 1;
    double run_time = 0;
    int iter = 0;
    for (int i = 1; i <= max_times; i <<= 1) {
        auto t0 = std::chrono::steady_clock::now();
        be.doit(i);
        run_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        iter = i;
        if (run_time > t_min) break;
    }
    std::printf("nx: %d, ny: %d, n_batch: %d, iter: %d, time per iter: %lf\n", nx, ny, batch, iter, run_time / iter);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s accuracy|speed|sweep [-x NX] [-y NY] [-b B] [-s SEED] [-m MAX] [-t SEC] [-r]\n",
                     argv[0]);
        return 2;
    }
    std::string mode = argv[1];
    int nx = 256, ny = 256, batch = 1, seed = 42, max_times = 1 << 30;
    double t_min = 4.0;
    bool rel = false;
    optind = 2;
    int c;
    while ((c = getopt(argc, argv, "x:y:b:s:m:t:r")) != -1) {
        switch (c) {
        case 'x': nx = std::atoi(optarg); break;
        case 'y': ny = std::atoi(optarg); break;
        case 'b': batch = std::atoi(optarg); break;
        case 's': seed = std::atoi(optarg); break;
        case 'm': max_times = std::atoi(optarg); break;
        case 't': t_min = std::atof(optarg); break;
        case 'r': rel = true; break;
        default: return 2;
        }
    }
    if (mode == "accuracy") return accuracy(nx, ny, batch, seed, rel);
    if (mode == "speed") return speed(nx, ny, batch, max_times, t_min);
    if (mode == "sweep") {
        const int xs[] = {256, 512, 1024};
        const int ys[] = {256, 512, 1024};
        for (int x : xs)
            for (int y : ys) {
                if (x == 1024 && y != 1024) continue;
                std::printf("Accuracy test: tcFFT-MI355X, FFT size: %dx%d batchsize:1\n", x, y);
                if (accuracy(x, y, 1, 42, true)) return 1;
            }
        return 0;
    }
    std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
}
