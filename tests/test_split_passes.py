"""Split-precision pass lists (tcfft_query_split, host only): every pass is
run here as the column DFT + twiddle + store its descriptor states, in FP64
with numpy, and the sequence must give numpy's own FFT of the shape.  This
pins the decomposition and the index arithmetic the GPU passes implement
(a wrong Qlo, twiddle modulus or store order fails), without a device."""
import numpy as np
import pytest

import tcfft_amd


def apply_pass(x, p, sign):
    """One pass on x of shape (signals, R*Q): the R-point DFT down each of
    the Q columns, output k of column q = q_hi*Qlo + q_lo times
    W_M^{q_hi k}, stored at q_hi*R*Qlo + k*Qlo + q_lo."""
    R, Q, Qlo = p["R"], 1 << p["q_log2"], 1 << p["qlo_log2"]
    y = x.reshape(x.shape[0], R, Q)
    Y = np.fft.fft(y, axis=1) if sign < 0 else np.fft.ifft(y, axis=1) * R
    q = np.arange(Q)
    qhi, qlo = q >> p["qlo_log2"], q & (Qlo - 1)
    k = np.arange(R)
    if p["m_log2"]:
        Y = Y * np.exp(sign * 2j * np.pi * np.outer(k, qhi) / float(1 << p["m_log2"]))[None]
    addr = qhi[None, :] * R * Qlo + k[:, None] * Qlo + qlo[None, :]
    out = np.empty_like(x)
    out[:, addr.ravel()] = Y.reshape(x.shape[0], -1)
    return out


def run_model(x, passes, sign):
    total = x.size
    for p in passes:
        sig_len = p["R"] << p["q_log2"]
        signals = total // sig_len
        assert p["units"] == signals * (1 << p["q_log2"]) // 16
        x = apply_pass(x.reshape(signals, sig_len), p, sign)
    return x


@pytest.mark.parametrize("n", [1 << k for k in range(8, 23) if k != 16])
@pytest.mark.parametrize("sign", [-1, 1])
def test_1d_pass_list_computes_the_dft(n, sign):
    batch = 2 if n <= (1 << 20) else 1
    passes = tcfft_amd.query_split(1, n, 0, batch)
    rng = np.random.default_rng(n)
    x = rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))
    got = run_model(x.copy(), passes, sign).reshape(batch, n)
    ref = np.fft.fft(x, axis=-1) if sign < 0 else np.fft.ifft(x, axis=-1) * n
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-12


@pytest.mark.parametrize("nx,ny", [(2, 256), (4, 512), (8, 256), (16, 256), (256, 256), (512, 1024),
                                   (1024, 1024), (4096, 256), (32, 8192)])
def test_2d_pass_list_computes_the_dft(nx, ny):
    batch = 2
    passes = tcfft_amd.query_split(2, nx, ny, batch)
    rng = np.random.default_rng(nx * ny)
    x = rng.standard_normal((batch, nx, ny)) + 1j * rng.standard_normal((batch, nx, ny))
    for sign in (-1, 1):
        got = run_model(x.copy().reshape(-1), passes, sign).reshape(batch, nx, ny)
        ref = np.fft.fft2(x) if sign < 0 else np.fft.ifft2(x) * (nx * ny)
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-12


def test_pass_structure_every_size():
    """Radices 16..256 (2..8 only for a lone short strided axis), 2 passes up
    to 2^16 points, 3 up to 2^24, 4 above; the first pass of a contiguous
    axis is the row store (Qlo = 1) and only the last pass of an axis has no
    twiddle; 65536 points keep their own kernel (no list)."""
    for k in range(8, 28):
        passes = tcfft_amd.query_split(1, 1 << k, 0, 1)
        if k == 16:
            assert passes == []
            continue
        assert len(passes) == (2 if k <= 16 else 3 if k <= 24 else 4), k
        assert sum(p["R"].bit_length() - 1 for p in passes) == k
        assert all(16 <= p["R"] <= 256 for p in passes)
        assert passes[0]["qlo_log2"] == 0 and all(p["qlo_log2"] >= 4 for p in passes[1:])
        assert [p["m_log2"] > 0 for p in passes] == [True] * (len(passes) - 1) + [False]
    p2 = tcfft_amd.query_split(2, 4, 512, 3)
    assert [p["R"] for p in p2] == [16, 32, 4] and p2[-1]["m_log2"] == 0 and p2[-1]["qlo_log2"] == 9


def test_query_split_rejects_bad_shapes():
    for args in ((1, 100, 0, 1), (1, 128, 0, 1), (1, 1 << 28, 0, 1), (2, 256, 128, 1), (3, 256, 256, 1),
                 (1, 1 << 20, 0, 4096)):
        with pytest.raises(tcfft_amd.TcfftError):
            tcfft_amd.query_split(*args)
