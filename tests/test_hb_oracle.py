"""The oracle's harmonic-balance transforms (oracle/hb.py) pinned by numpy's
FFT, closed forms and identities: the permuted forward transform equals
rfft / N in the reference's packing, the inverse undoes it, a cosine and a
sine of harmonic k land on harmonic k alone, the j w k Q term turns the
samples of q into those of dq/dt, and the real-equivalent preconditioner
block solves the complex system of its harmonic."""
import numpy as np
import pytest

from oracle import hb


def test_dft_matches_rfft_packing():
    rng = np.random.default_rng(3)
    N, n = 129, 7
    xt = rng.standard_normal((N, n))
    xf = hb.dft_permuted(xt)
    R = np.fft.rfft(xt, axis=0) / N          # e^{-2 pi i k s / N}
    H = (N - 1) // 2
    for k in range(H + 1):
        np.testing.assert_allclose(xf[:, 2 * k], R[k].real, rtol=0, atol=1e-14)
        np.testing.assert_allclose(xf[:, 2 * k + 1], R[k].imag if k else 0.0, rtol=0, atol=1e-14)
        if k:
            np.testing.assert_allclose(xf[:, 2 * (N - k)], R[k].real, rtol=0, atol=1e-14)
            np.testing.assert_allclose(xf[:, 2 * (N - k) + 1], -R[k].imag, rtol=0, atol=1e-14)


def test_round_trip():
    rng = np.random.default_rng(4)
    xt = rng.standard_normal((33, 5))
    np.testing.assert_allclose(hb.ift_permuted(hb.dft_permuted(xt)), xt, rtol=0, atol=1e-13)


@pytest.mark.parametrize("k0", [1, 5, 16])
def test_single_harmonics(k0):
    N = 33
    s = np.arange(N)
    c = hb.dft_permuted(np.cos(2 * np.pi * k0 * s / N)[:, None])[0]
    si = hb.dft_permuted(np.sin(2 * np.pi * k0 * s / N)[:, None])[0]
    want_c = np.zeros(2 * N)
    want_c[2 * k0] = want_c[2 * (N - k0)] = 0.5
    want_s = np.zeros(2 * N)
    want_s[2 * k0 + 1], want_s[2 * (N - k0) + 1] = -0.5, 0.5
    np.testing.assert_allclose(c, want_c, atol=1e-14)
    np.testing.assert_allclose(si, want_s, atol=1e-14)


def test_jwq_is_the_time_derivative():
    N, T = 65, 2e-9
    w = 2 * np.pi / T
    t = np.arange(N) * T / N
    q = np.stack([np.cos(3 * w * t) + 0.25 * np.sin(7 * w * t), 0.5 + np.sin(w * t)], axis=1)
    dq = np.stack([-3 * w * np.sin(3 * w * t) + 0.25 * 7 * w * np.cos(7 * w * t),
                   w * np.cos(w * t)], axis=1)
    ff = hb.add_jwq(np.zeros((2, 2 * N)), hb.dft_permuted(q), T)
    np.testing.assert_allclose(hb.ift_permuted(ff), dq, rtol=0, atol=1e-6 * w)


def test_block_matrix_solves_its_harmonic():
    rng = np.random.default_rng(5)
    n, T, k = 6, 1e-9, 3
    G = rng.standard_normal((n, n)) + 4 * np.eye(n)
    C = 1e-12 * rng.standard_normal((n, n))
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    z = np.linalg.solve(G + 1j * k * (2 * np.pi / T) * C, b)
    K = hb.block_jacobi_matrix(G, C, k, T)
    x = np.linalg.solve(K, np.concatenate([b.real, b.imag]))
    np.testing.assert_allclose(x[:n] + 1j * x[n:], z, rtol=1e-12)


def test_even_sample_count_rejected():
    with pytest.raises(ValueError):
        hb.dft_permuted(np.zeros((4, 1)))
