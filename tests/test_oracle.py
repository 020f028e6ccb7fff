"""Pins for the FP64 checker (oracle/).  CPU only.

Each oracle function is checked against something other than its own formula:
closed forms, brute force, numpy's pocketfft golden vectors, an independent
model of glibc's generator, and invariants chosen so that a dropped twiddle,
a wrong sign or a transposed index fails at least one test.
"""
import os

import numpy as np
import pytest

import oracle

RNG = np.random.default_rng(7)


def crand(shape, rng=RNG):
    return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)


# ---------------------------------------------------------------- gen_rev
def closed_form_rev(n, radices):
    """Position p = sum_i p_i * (r_0...r_{i-1}) holds index
    p_{m-1} + r_{m-1} * (p_{m-2} + r_{m-2} * (... + r_1 * p_0))."""
    m = len(radices)
    out = np.empty(n, dtype=np.int64)
    for p in range(n):
        digits = []
        q = p
        for r in radices:
            digits.append(q % r)
            q //= r
        idx = 0
        for i in range(0, m):  # p_0 is the most significant index digit
            idx = idx * radices[i] + digits[i]
        out[p] = idx
    return out


@pytest.mark.parametrize("radices", [[16, 16], [16, 16, 2], [16, 16, 4], [2, 4, 8], [16, 2, 16], [4, 4, 4, 4], [8]])
def test_gen_rev_closed_form(radices):
    n = int(np.prod(radices))
    got = oracle.gen_rev(n, radices)
    assert sorted(got.tolist()) == list(range(n))
    np.testing.assert_array_equal(got, closed_form_rev(n, radices))


def test_gen_rev_is_involution_for_symmetric_lists():
    rev = oracle.gen_rev(4096, [16, 16, 16])
    np.testing.assert_array_equal(rev[rev], np.arange(4096))


# ---------------------------------------------------------- generate_data
class GlibcRandom:
    """Independent model of glibc random_r TYPE_3 (x[i] = x[i-3] + x[i-31])."""

    def __init__(self, seed):
        r = [0] * 34
        r[0] = seed if seed != 0 else 1
        for i in range(1, 31):
            hi, lo = divmod(r[i - 1], 127773)
            word = 16807 * lo - 2836 * hi
            if word < 0:
                word += 2147483647
            r[i] = word
        for i in range(31, 34):
            r[i] = r[i - 31]
        self.r = r
        for _ in range(34, 344):
            self._step()

    def _step(self):
        v = (self.r[-31] + self.r[-3]) & 0xFFFFFFFF
        self.r.append(v)
        return v >> 1

    def rand(self):
        return self._step()


def test_glibc_model_known_answer():
    g = GlibcRandom(1)
    assert [g.rand() for _ in range(3)] == [1804289383, 846930886, 1681692777]


@pytest.mark.parametrize("seed", [42, 7])
def test_generate_data_matches_glibc_model(seed):
    n, b = 64, 3
    got = oracle.generate_data(n, b, seed)
    g = GlibcRandom(seed)
    rand_max = np.float32(2147483647)
    exp = np.empty((b, n), dtype=np.complex128)
    for i in range(b):
        for j in range(n):
            re = np.float32(np.float32(0.0001) * np.float32(g.rand())) / rand_max
            im = np.float32(np.float32(0.0001) * np.float32(g.rand())) / rand_max
            exp[i, j] = complex(float(np.float32(re)), float(np.float32(im)))
    np.testing.assert_array_equal(got, exp)
    assert got.real.min() >= 0 and got.real.max() <= 1e-4


# --------------------------------------------------------------- get_error
def test_get_error_hand_cases():
    std = np.array([[1 + 2j, -4 + 0.5j]])
    tst = np.array([[1.1 + 2j, -4 + 5.0j]])
    # components: |0.1/1| = .1, 0, 0, |4.5/0.5| = 9 -> clamped 1
    assert oracle.get_error(tst, std) == pytest.approx((0.1 + 0 + 0 + 1.0) / 2)
    # standard 0: x/0 = inf and 0/0 = nan both count as 1
    assert oracle.get_error(np.array([[1 + 0j]]), np.array([[0 + 0j]])) == pytest.approx(2.0)
    assert oracle.get_error(std, std) == 0.0


def test_rel_l2_hand_cases():
    assert oracle.rel_l2(np.array([3 + 0j]), np.array([0 + 4j])) == pytest.approx(5 / 4)
    assert oracle.rel_l2(np.array([1 + 1j, 2j]), np.array([1 + 1j, 2j])) == 0.0


# ------------------------------------------------------------ dft / dit
@pytest.mark.parametrize("n", [1, 2, 8, 16, 48])
def test_dft_ld_closed_forms(n):
    imp = np.zeros(n, complex)
    imp[0] = 1
    np.testing.assert_allclose(oracle.dft_ld(imp), np.ones(n), atol=1e-15)
    np.testing.assert_allclose(oracle.dft_ld(np.ones(n)), np.eye(1, n, 0).ravel() * n, atol=1e-12)
    k0 = 3 % n
    tone = np.exp(2j * np.pi * k0 * np.arange(n) / n)
    exp = np.zeros(n, complex)
    exp[k0] = n
    np.testing.assert_allclose(oracle.dft_ld(tone, -1), exp, atol=1e-12)
    # the adjoint sign sends the same tone to bin -k0
    exp2 = np.zeros(n, complex)
    exp2[(-k0) % n] = n
    np.testing.assert_allclose(oracle.dft_ld(tone, +1), exp2, atol=1e-12)


LISTS = [[16, 16], [16, 16, 2], [16, 16, 4], [2, 4, 8], [16, 2, 16], [4, 16, 2], [2, 2, 2, 2, 2], [8, 8, 8], [16, 16, 16]]


@pytest.mark.parametrize("radices", LISTS)
@pytest.mark.parametrize("sign", [-1, +1])
def test_dit_equals_brute_force(radices, sign):
    n = int(np.prod(radices))
    x = crand((2, n))
    rev = oracle.gen_rev(n, radices)
    got = oracle.dit(x[:, rev], radices, sign)
    for b in range(2):
        ref = oracle.dft_ld(x[b], sign)
        assert oracle.rel_l2(got[b], ref) < 1e-13


@pytest.mark.parametrize("sign", [-1, +1])
def test_fft_matches_golden(golden_dir, sign):
    for n in (16, 256, 512, 1024, 4096):
        d = np.load(os.path.join(golden_dir, f"fft1d_n{n}.npz"))
        x = d["x_re"].astype(np.float64) + 1j * d["x_im"].astype(np.float64)
        key = "forward" if sign < 0 else "adjoint"
        if key not in d:
            continue
        got = oracle.fft(x, sign)
        assert oracle.rel_l2(got, d[key]) < 1e-14, n


@pytest.mark.parametrize("n", [2 ** 13, 2 ** 16, 2 ** 17])
def test_fft_matches_numpy_large(n):
    x = crand((1, n))
    assert oracle.rel_l2(oracle.fft(x), np.fft.fft(x, axis=-1)) < 1e-13
    assert oracle.rel_l2(oracle.fft(x, +1), np.fft.ifft(x, axis=-1) * n) < 1e-13


def test_fft_invariants():
    n = 1024
    x = crand((1, n))
    y = crand((1, n))
    X = oracle.fft(x)
    # Parseval
    assert np.sum(abs(X) ** 2) == pytest.approx(n * np.sum(abs(x) ** 2), rel=1e-12)
    # linearity
    a, b = 0.3 - 1.1j, 2.0 + 0.5j
    assert oracle.rel_l2(oracle.fft(a * x + b * y), a * X + b * oracle.fft(y)) < 1e-13
    # transforming twice gives n * x[-m]
    XX = oracle.fft(X)
    np.testing.assert_allclose(XX[0], n * x[0][(-np.arange(n)) % n], rtol=0, atol=1e-9)
    # forward then adjoint is n * identity
    np.testing.assert_allclose(oracle.fft(X, +1), n * x, atol=1e-9)
    # shift theorem
    s = 5
    Xs = oracle.fft(np.roll(x, s, axis=-1))
    np.testing.assert_allclose(Xs, X * np.exp(-2j * np.pi * s * np.arange(n) / n), atol=1e-10)
    # Hermitian symmetry of a real input
    R = oracle.fft(x.real.astype(complex))
    np.testing.assert_allclose(R[0][(-np.arange(n)) % n], np.conj(R[0]), atol=1e-10)


# -------------------------------------------------------------------- 2D
@pytest.mark.parametrize("shape", [(16, 32), (32, 32)])
def test_fft_2d_matches_golden(golden_dir, shape):
    d = np.load(os.path.join(golden_dir, f"fft2d_{shape[0]}x{shape[1]}.npz"))
    x = d["x_re"].astype(np.float64) + 1j * d["x_im"].astype(np.float64)
    assert oracle.rel_l2(oracle.fft_2d(x), d["forward"]) < 1e-14
    if "adjoint" in d:
        assert oracle.rel_l2(oracle.fft_2d(x, +1), d["adjoint"]) < 1e-14


def test_fft_2d_separable_and_dit_2d():
    nx, ny = 64, 128
    x = crand((2, nx, ny))
    ref = np.fft.fft(np.fft.fft(x, axis=-1), axis=-2)
    assert oracle.rel_l2(oracle.fft_2d(x), ref) < 1e-13
    rx, ry = [16, 4], [16, 8]
    revx, revy = oracle.gen_rev(nx, rx), oracle.gen_rev(ny, ry)
    got = oracle.dit_2d(x[:, revx][:, :, revy], rx, ry)
    assert oracle.rel_l2(got, ref) < 1e-13


# --------------------------------------------------------- plan tables
REF_1D = {256: [16, 16], 512: [16, 16, 2], 1024: [16, 16, 4], 131072: [16, 16, 2, 16, 16],
          262144: [16, 16, 2, 16, 16, 2], 524288: [16, 16, 4, 16, 16, 2],
          16777216: [16] * 6, 33554432: [16, 16, 2, 16, 16, 16, 16],
          67108864: [16, 16, 2, 16, 16, 2, 16, 16], 134217728: [16, 16, 2, 16, 16, 2, 16, 16, 2]}


def test_ref_plan_table_lists_ten_sizes():
    for n in (256, 512, 1024, 131072, 262144, 524288, 16777216, 33554432, 67108864, 134217728):
        ok, rad, mer = oracle.ref_plan_1d(n)
        assert ok and int(np.prod(rad)) == n
        assert rad == REF_1D[n]
        # pass radices: 256 << merging, product is n
        assert int(np.prod([256 << m for m in mer])) == n
    for n in (65536, 4096, 2 ** 20, 8192, 100):
        assert not oracle.ref_plan_1d(n)[0]


@pytest.mark.parametrize("n", [256, 512, 1024, 131072, 262144])
def test_ref_radix_lists_reproduce_the_dft(n):
    """The reference's contract: gen_rev(plan.radices) then the DIT passes give
    the natural-order DFT (tcfft_doit_half.cpp:38-46 + tcfftExec)."""
    ok, rad, _ = oracle.ref_plan_1d(n)
    x = crand((1, n))
    rev = oracle.gen_rev(n, rad)
    got = oracle.dit(x[:, rev], rad)
    assert oracle.rel_l2(got, np.fft.fft(x, axis=-1)) < 1e-13


def test_ref_plan_2d():
    ok, rx, ry, mer = oracle.ref_plan_2d(512, 1024)
    assert ok and rx == [16, 16, 2] and ry == [16, 16, 4] and mer == [2, 1]
    assert not oracle.ref_plan_2d(1024, 256)[0]   # layer_1[2] is never assigned
    assert not oracle.ref_plan_2d(128, 256)[0]
    ok, rx, ry, mer = oracle.ref_plan_2d(256, 512)
    x = crand((1, 256, 512))
    revx, revy = oracle.gen_rev(256, rx), oracle.gen_rev(512, ry)
    got = oracle.dit_2d(x[:, revx][:, :, revy], rx, ry)
    assert oracle.rel_l2(got, np.fft.fft2(x, axes=(-2, -1))) < 1e-13


def test_planted_mistakes_are_caught():
    """A sign flip or a dropped twiddle in a DIT stage must fail the brute-force pin."""
    n, rad = 256, [16, 16]
    x = crand((1, n))
    rev = oracle.gen_rev(n, rad)
    good = oracle.dit(x[:, rev], rad, -1)
    ref = oracle.dft_ld(x[0], -1)
    assert oracle.rel_l2(good[0], ref) < 1e-13
    # wrong sign
    assert oracle.rel_l2(oracle.dit(x[:, rev], rad, +1)[0], ref) > 0.5
    # dropped inter-stage twiddle: two independent DFT-16 stages = 2D DFT, not 1D
    two_d = np.fft.fft2(x[0].reshape(16, 16)).T.reshape(-1)
    assert oracle.rel_l2(two_d, ref) > 0.5
    # transposed digit order (natural input instead of digit-reversed)
    assert oracle.rel_l2(oracle.dit(x, rad)[0], ref) > 0.5


# ---------------------------------------------------------------- real input
@pytest.mark.parametrize("n", [16, 512, 8192])
def test_rfft_irfft_match_numpy(n):
    x = RNG.uniform(-1, 1, (3, n))
    X = oracle.rfft(x)
    assert X.shape == (3, n // 2 + 1)
    assert oracle.rel_l2(X, np.fft.rfft(x, axis=-1)) < 1e-13
    Y = crand((3, n // 2 + 1))
    y = oracle.irfft(Y, n)
    assert y.shape == (3, n)
    np.testing.assert_allclose(y, n * np.fft.irfft(Y, n, axis=-1), rtol=0, atol=1e-10 * n)
    # round trip and Hermitian symmetry of the full transform it truncates
    np.testing.assert_allclose(oracle.irfft(X, n), n * x, rtol=0, atol=1e-9 * n)
    full = oracle.fft(x.astype(np.complex128))
    np.testing.assert_allclose(full[:, 1:][:, ::-1], np.conj(full[:, 1:]), rtol=0, atol=1e-9 * n)


def test_rfft_closed_forms():
    n, k0 = 1024, 37
    t = np.arange(n)
    want = np.zeros(n // 2 + 1, dtype=np.complex128)
    want[k0] = n / 2
    np.testing.assert_allclose(oracle.rfft(np.cos(2 * np.pi * k0 * t / n)[None])[0], want, atol=1e-10)
    want = np.zeros(n // 2 + 1, dtype=np.complex128)
    want[n // 2] = n
    np.testing.assert_allclose(oracle.rfft(((-1.0) ** t)[None])[0], want, atol=1e-10)
    # the imaginary parts of bins 0 and n/2 do not reach the inverse
    Y = crand((1, n // 2 + 1))
    Z = Y.copy()
    Z[0, 0] += 5j
    Z[0, -1] -= 3j
    np.testing.assert_array_equal(oracle.irfft(Y, n), oracle.irfft(Z, n))


def untangle(Z, sign=-1):
    """The even/odd split the real-input kernels compute, restated in numpy:
    X[k] = ((Z[k] + conj Z[L-k]) - i W^k (Z[k] - conj Z[L-k])) / 2."""
    L = Z.shape[-1]
    k = np.arange(L + 1)
    Zk = Z[..., k % L]
    Zm = np.conj(Z[..., (L - k) % L])
    w = np.exp(sign * 2j * np.pi * k / (2 * L))
    return ((Zk + Zm) - 1j * w * (Zk - Zm)) / 2


def tangle(X):
    """Inverse pre-step: Z'[k] = (X[k] + conj X[L-k]) + i W^-k (X[k] - conj X[L-k])."""
    L = X.shape[-1] - 1
    X = X.copy()
    X[..., 0] = X[..., 0].real
    X[..., L] = X[..., L].real
    k = np.arange(L)
    a, b = X[..., k], np.conj(X[..., L - k])
    return (a + b) + 1j * np.exp(2j * np.pi * k / (2 * L)) * (a - b)


@pytest.mark.parametrize("n", [512, 8192])
def test_even_odd_split_against_oracle(n):
    """The kernels' formulas against the oracle; a flipped twiddle sign, a
    missing conjugate or a dropped factor 1/2 fails."""
    x = RNG.uniform(-1, 1, (2, n))
    z = x[:, 0::2] + 1j * x[:, 1::2]
    Z = oracle.fft(z)
    ref = oracle.rfft(x)
    assert oracle.rel_l2(untangle(Z), ref) < 1e-13
    assert oracle.rel_l2(untangle(Z, +1), ref) > 0.1
    L = n // 2
    k = np.arange(L + 1)
    no_conj = ((Z[:, k % L] + Z[:, (L - k) % L]) - 1j * np.exp(-2j * np.pi * k / n)
               * (Z[:, k % L] - Z[:, (L - k) % L])) / 2
    assert oracle.rel_l2(no_conj, ref) > 0.1
    assert oracle.rel_l2(2 * untangle(Z), ref) > 0.5
    X = crand((2, L + 1))
    y = oracle.irfft(X, n)
    zz = oracle.fft(tangle(X), +1)
    got = np.stack([zz.real, zz.imag], -1).reshape(2, n)
    np.testing.assert_allclose(got, y, rtol=0, atol=1e-9 * n)
