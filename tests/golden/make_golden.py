"""Write the FP64 golden vectors for the FFT path (run once; output committed).

Inputs are complex values whose real and imaginary parts are exactly
representable in IEEE binary16 (uniform[-1,1) rounded to fp16), so the same
vectors drive the fp16 GPU path and the FP64 checker without a quantisation
term.  Expected outputs come from numpy.fft (pocketfft, FP64), an
implementation independent of the oracle's C restatement:
  forward  X[k] = sum_n x[n] exp(-2 pi i n k / N)         (FFTW_FORWARD sign)
  adjoint  Y[k] = sum_n x[n] exp(+2 pi i n k / N) = N * ifft(x)

Usage: python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def fp16_complex(rng, shape):
    re = rng.uniform(-1.0, 1.0, size=shape).astype(np.float16)
    im = rng.uniform(-1.0, 1.0, size=shape).astype(np.float16)
    return re.astype(np.float64) + 1j * im.astype(np.float64)


def main():
    rng = np.random.default_rng(20261202)
    # x is stored as its fp16 parts (exact); outputs as complex128
    one_d = {16: 4, 256: 2, 512: 1, 1024: 1, 4096: 1}
    for n, batch in one_d.items():
        x = fp16_complex(rng, (batch, n))
        extra = {"adjoint": np.fft.ifft(x, axis=-1) * n} if n <= 256 else {}
        np.savez(os.path.join(HERE, f"fft1d_n{n}.npz"),
                 x_re=x.real.astype(np.float16), x_im=x.imag.astype(np.float16),
                 forward=np.fft.fft(x, axis=-1), **extra)
    for nx, ny in ((16, 32), (32, 32)):
        x = fp16_complex(rng, (1, nx, ny))
        extra = {"adjoint": np.fft.ifft2(x, axes=(-2, -1)) * (nx * ny)} if nx * ny <= 512 else {}
        np.savez(os.path.join(HERE, f"fft2d_{nx}x{ny}.npz"),
                 x_re=x.real.astype(np.float16), x_im=x.imag.astype(np.float16),
                 forward=np.fft.fft2(x, axes=(-2, -1)), **extra)


if __name__ == "__main__":
    main()
