"""FP64 numpy restatement of the demodulation F-statistic atoms -- TEST
INFRASTRUCTURE ONLY (imported by tests/ and bench.py's checks, never by the
product).

In the reference the atoms come from lalpulsar.ComputeFstat(...,
FSTATQ_ATOMS_PER_DET) (pyfstat/core.py:671-673, :1359-1365); lalpulsar is not
part of the reference source and cannot be imported here, so this follows the
definition of the resampling-free "demodulation" sum (the matched filter of
a signal that is a pure tone across each SFT), as tcw_demod_atoms documents
it (include/tcw_hip.h):

  dtau  = t_start + Tsft/2 + delay - ref_time
  phase = dtau (f + fdot dtau / 2)                       cycles
  kappa = (f + fdot dtau) (1 + delay_rate) Tsft          observed bin
  X     = exp(-2 pi i phase) sum_k y_k K_k,   K_k = (-1)^kf sin(pi delta) / (pi (kappa - k))
  over k = kf-D+1 .. kf+D (kf = floor(kappa), delta = kappa - kf; delta = 0:
  bin kf alone), bins outside the stored band count as zero;
  atoms a2 = a^2, b2 = b^2, ab = a b, Fa = a X, Fb = b X.

Pinned by tests/test_demod.py: the closed form at bin centres, a brute-force
DFT of a sampled tone (the SFT bins of a pure tone are (-1)^k sinc(pi
(kappa - k)) e^{i 2 pi phase} up to O((kappa-k)/N)^2, so X = sum K^2 > 0), the
kernel's sum rule sum_k K_k^2 -> 1 as D grows, and E[2F] = 4 sum K^2 in
Gaussian noise.
"""

import numpy as np


def kernel(kappa, dterms):
    """(first bin index kf-D+1, weights K [2D]) of one observed bin kappa."""
    kf = np.floor(kappa)
    delta = kappa - kf
    k = kf - dterms + 1 + np.arange(2 * dterms)
    sign = -1.0 if int(kf) % 2 else 1.0
    if delta == 0.0:
        K = np.where(k == kf, sign, 0.0)
    else:
        K = sign * np.sin(np.pi * delta) / (np.pi * (kappa - k))
    return int(kf) - dterms + 1, K


def demod_atoms(t_start, bins, k_min, Tsft, delay, delay_rate, a, b, freq, fdot, ref_time,
                dterms=8):
    """Atoms [n_tmpl, n_sft]: returns (a2, b2, ab, Fa complex, Fb complex) as
    float64 / complex128 arrays.  bins: complex [n_sft, n_bins].  Vectorised
    over the SFTs; the same sum as kernel() per SFT."""
    t_start = np.asarray(t_start, dtype=np.float64)
    n_sft, n_bins = bins.shape
    freq = np.atleast_1d(np.asarray(freq, dtype=np.float64))
    fdot = np.atleast_1d(np.asarray(fdot, dtype=np.float64))
    X = np.zeros((freq.size, n_sft), dtype=np.complex128)
    dtau = t_start + 0.5 * Tsft + np.asarray(delay, dtype=np.float64) - ref_time
    dop = (1.0 + np.asarray(delay_rate, dtype=np.float64)) * Tsft
    i = np.arange(2 * dterms)
    rows = np.arange(n_sft)[:, None]
    for t in range(freq.size):
        phase = dtau * (freq[t] + 0.5 * fdot[t] * dtau)
        kappa = (freq[t] + fdot[t] * dtau) * dop
        kf = np.floor(kappa)
        delta = kappa - kf
        k = kf[:, None] - dterms + 1 + i[None, :]
        sign = np.where(kf.astype(np.int64) % 2 == 1, -1.0, 1.0)[:, None]
        on_bin = (delta == 0.0)[:, None]
        with np.errstate(divide="ignore", invalid="ignore"):
            K = np.where(on_bin, np.where(k == kf[:, None], sign, 0.0),
                         sign * np.sin(np.pi * delta)[:, None] / (np.pi * (kappa[:, None] - k)))
        j = k.astype(np.int64) - k_min
        ok = (j >= 0) & (j < n_bins)
        y = np.where(ok, bins[rows, np.clip(j, 0, n_bins - 1)], 0.0)
        s = np.sum(y * K, axis=1)
        frac = phase - np.floor(phase)
        X[t] = s * np.exp(-2j * np.pi * frac)
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    a2 = np.broadcast_to(a * a, X.shape)
    b2 = np.broadcast_to(b * b, X.shape)
    ab = np.broadcast_to(a * b, X.shape)
    return a2, b2, ab, a * X, b * X


def twoF_all_data(a2, b2, ab, Fa, Fb):
    """2F over all atoms of one template (sums along the last axis), FP64."""
    A, B, C = a2.sum(-1), b2.sum(-1), ab.sum(-1)
    Fa_s, Fb_s = Fa.sum(-1), Fb.sum(-1)
    D = A * B - C * C
    F = (B * np.abs(Fa_s) ** 2 + A * np.abs(Fb_s) ** 2
         - 2 * C * np.real(Fa_s * np.conj(Fb_s))) / D
    return 2 * F


def tone_bins(kappa, phase, n_bins, k_min):
    """Analytic SFT bins of a unit pure tone: (-1)^k sinc(pi (kappa - k)) e^{2 pi i phase}
    for stored bins k = k_min .. k_min+n_bins-1."""
    k = k_min + np.arange(n_bins)
    return ((-1.0) ** k) * np.sinc(kappa - k) * np.exp(2j * np.pi * phase)


def tone_sft_dft(f_obs, phase_mid, Tsft, n_samples, k_min, n_bins):
    """SFT bins of a sampled complex tone by brute-force DFT: samples
    x_n = exp(i (2 pi phase_mid + 2 pi f_obs (t_n - T/2))) at t_n = n T/N,
    bins X_k = (1/N) sum_n x_n exp(-2 pi i k n / N) (unit-amplitude normalisation),
    returned for k = k_min .. k_min+n_bins-1."""
    n = np.arange(n_samples)
    t = n * (Tsft / n_samples)
    x = np.exp(1j * (2 * np.pi * phase_mid + 2 * np.pi * f_obs * (t - 0.5 * Tsft)))
    k = k_min + np.arange(n_bins)
    return np.exp(-2j * np.pi * np.outer(k, n) / n_samples) @ x / n_samples
