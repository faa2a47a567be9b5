"""Harmonic-balance transforms and HB vector assembly -- TEST INFRASTRUCTURE
ONLY (the oracle restatement the GPU HB path is checked against).

Layouts follow the reference's block vectors:

* time domain: ``xt[s, j]``, sample s (fast time ``s * T / N``) of unknown j
  -- one block per sample (N_LOA_HBLoader.C:565-600, ``bX.block(i)``);
* frequency domain, "permuted": ``xf[j, 2 * N]`` -- one block per unknown,
  holding (Re, Im) of harmonic k at positions (2k, 2k + 1) for k = 0..H and
  of harmonic -k at (2(N - k), 2(N - k) + 1); N = 2H + 1 samples
  (N_ANP_HB.C:1470, ``size_ = 2 * posfreq + 1``).

Parity: the transforms are pinned in tests/test_hb_oracle.py by numpy's FFT,
by closed forms (a cosine and a sine of harmonic k), by the inverse round
trip, and by the derivative identity of the jwk Q term on a sinusoid.
"""
import numpy as np


def _check(N):
    if N < 1 or N % 2 == 0:
        raise ValueError("HB sample count must be odd (2H + 1)")


def dft_permuted(xt):
    """Forward transform of every unknown's samples (computePermutedDFT,
    N_LAS_BlockSystemHelpers.C:512-562, with the FFTW R2HC packing of
    N_UTL_FFTW_Interface.C:66-97): r_k = sum_s x_s cos(2 pi k s / N),
    i_k = -sum_s x_s sin(2 pi k s / N); harmonic k gets (r_k, i_k) / N and
    harmonic -k its conjugate; (Re, Im) of harmonic 0 is (r_0 / N, 0)."""
    xt = np.asarray(xt, dtype=np.float64)
    N, n = xt.shape
    _check(N)
    H = (N - 1) // 2
    s = np.arange(N)
    xf = np.zeros((n, 2 * N))
    xf[:, 0] = xt.sum(axis=0) / N
    for k in range(1, H + 1):
        ang = 2.0 * np.pi * ((k * s) % N) / N
        r = (np.cos(ang)[:, None] * xt).sum(axis=0)
        im = -(np.sin(ang)[:, None] * xt).sum(axis=0)
        xf[:, 2 * k] = r / N
        xf[:, 2 * k + 1] = im / N
        xf[:, 2 * (N - k)] = r / N
        xf[:, 2 * (N - k) + 1] = -im / N
    return xf


def ift_permuted(xf):
    """Inverse transform (computePermutedIFT, N_LAS_BlockSystemHelpers.C:
    571-609, with the FFTW HC2R packing and 1/N scaling of
    N_UTL_FFTW_Interface.C:101-130, undone by the dftScalar factor N):
    x_s = Re_0 + 2 sum_{k=1..H} (Re_k cos(2 pi k s / N) - Im_k sin(2 pi k s / N)).
    Only harmonics 0..H are read (the negative ones are their conjugates)."""
    xf = np.asarray(xf, dtype=np.float64)
    n, L = xf.shape
    N = L // 2
    _check(N)
    H = (N - 1) // 2
    s = np.arange(N)
    xt = np.tile(xf[:, 0], (N, 1))
    for k in range(1, H + 1):
        ang = 2.0 * np.pi * ((k * s) % N) / N
        xt += 2.0 * (np.cos(ang)[:, None] * xf[None, :, 2 * k]
                     - np.sin(ang)[:, None] * xf[None, :, 2 * k + 1])
    return xt


def add_jwq(ff, qf, period):
    """F_f + j w k Q_f over every harmonic k = -H..H, w = 2 pi / period
    (N_LOA_HBLoader.C:733-775): for k = 1..H the real part gains
    -Im(Q_k) k w and the imaginary part Re(Q_k) k w; harmonic -k gets the
    conjugate's term (-Im(Q_k) k w, -Re(Q_k) k w) at its position; harmonic 0
    gets nothing (its factor is 0 * w)."""
    ff = np.array(ff, dtype=np.float64)
    qf = np.asarray(qf, dtype=np.float64)
    n, L = ff.shape
    N = L // 2
    H = (N - 1) // 2
    w = 2.0 * np.pi / period
    for k in range(1, H + 1):
        ff[:, 2 * k] += -qf[:, 2 * k + 1] * k * w
        ff[:, 2 * (N - k)] += -qf[:, 2 * k + 1] * k * w
        ff[:, 2 * k + 1] += qf[:, 2 * k] * k * w
        ff[:, 2 * (N - k) + 1] += -qf[:, 2 * k] * k * w
    return ff


def block_jacobi_matrix(G, C, k, period):
    """Real-equivalent form [G, -k w C; k w C, G] of harmonic k's
    preconditioner block (N_LAS_HBBlockJacobiPrecond.C:313-370; G and C are
    the sample averages of dF/dx and dQ/dx, dense here)."""
    w = 2.0 * np.pi / period
    G = np.asarray(G, dtype=np.float64)
    C = np.asarray(C, dtype=np.float64)
    return np.block([[G, -k * w * C], [k * w * C, G]])
