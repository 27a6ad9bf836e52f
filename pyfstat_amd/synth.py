"""Seeded synthetic F-statistic atoms for the benchmark and test workloads.

There is no SFT data here, so atoms are drawn from a simple antenna-pattern
model (one sidereal day period), the shape lalpulsar's ComputeFstat atoms take
(pyfstat/core.py:671-673 requests them; the per-atom fields are those of
pyfstat/tcw_fstat_map_funcs.py:610-617):

    a(t) = 0.3 + 0.5 cos(W t + phi_a),   b(t) = 0.4 sin(W t + phi_b),
    W = 2 pi / 86164 s,
    a2 = a^2, b2 = b^2, ab = a b                       (float32)
    Fa = a (n1 + s(t)),  Fb = b (n2 + s(t))            (complex64)

with n1, n2 independent complex standard normals (re, im ~ N(0, 1/2)) and
s(t) = 0.8 e^{i psi} times the injected transient window (0 outside it).

Timestamps sit on the grid t_i = 1000000000 + i*1800; a random fraction of the
interior timestamps is dropped as gaps (first and last always kept, so the
binned atom count equals the nominal SFT count).  Every workload draws from
``numpy.random.Generator(PCG64(20261201 + k))`` with k its 0-based index in the
list of workloads below.
"""

import numpy as np

from .atoms import (
    TRANSIENT_EXPONENTIAL,
    TRANSIENT_RECTANGULAR,
    FstatAtomVector,
    MultiFstatAtomVector,
    transientWindowRange_t,
)

OMEGA_SID = 2.0 * np.pi / 86164.0
T_START = 1_000_000_000
TSFT = 1800
SIGNAL_AMP = 0.8
SEED_BASE = 20261201


def rng_for(k):
    return np.random.Generator(np.random.PCG64(SEED_BASE + k))


def gap_mask(rng, n_sft, gap_frac):
    """Boolean keep-mask over n_sft grid slots; first/last always kept."""
    keep = np.ones(n_sft, dtype=bool)
    n_drop = int(round(gap_frac * n_sft))
    n_drop = min(n_drop, max(n_sft - 2, 0))
    if n_drop > 0:
        drop = rng.choice(np.arange(1, n_sft - 1), size=n_drop, replace=False)
        keep[drop] = False
    return keep


def detector_atoms(rng, t, phi_a, phi_b, signal):
    """Float32 columns [7, len(t)] for timestamps t (float64 seconds).

    ``signal`` is a complex array s(t) (zeros where no transient).
    """
    ph = OMEGA_SID * t
    a = 0.3 + 0.5 * np.cos(ph + phi_a)
    b = 0.4 * np.sin(ph + phi_b)
    n = t.shape[0]
    noise = rng.standard_normal((4, n)) * np.sqrt(0.5)
    n1 = noise[0] + 1j * noise[1]
    n2 = noise[2] + 1j * noise[3]
    Fa = a * (n1 + signal)
    Fb = b * (n2 + signal)
    cols = np.empty((7, n), dtype=np.float32)
    cols[0] = a * a
    cols[1] = b * b
    cols[2] = a * b
    cols[3] = Fa.real
    cols[4] = Fa.imag
    cols[5] = Fb.real
    cols[6] = Fb.imag
    return cols


def rect_signal(t, t_on, duration, psi, amp=SIGNAL_AMP):
    on = (t >= t_on) & (t < t_on + duration)
    return np.where(on, amp * np.exp(1j * psi), 0.0)


def exp_signal(t, t_on, tau, psi, amp=SIGNAL_AMP):
    dt = t - t_on
    on = (dt >= 0) & (dt <= 3 * tau)
    return np.where(on, amp * np.exp(1j * psi) * np.exp(-np.clip(dt, 0, None) / tau), 0.0)


def grid_window(wtype, t0, n_t0, n_tau, tau_min=2 * TSFT, dt0=TSFT, dtau=TSFT):
    """Window range covering n_t0 start times and n_tau durations."""
    return transientWindowRange_t(
        type=wtype, t0=t0, t0Band=(n_t0 - 1) * dt0, dt0=dt0,
        tau=tau_min, tauBand=(n_tau - 1) * dtau, dtau=dtau,
    )


def multi_from_detectors(dets, Tsft=TSFT):
    """MultiFstatAtomVector from a list of (timestamps, cols[7, n])."""
    multi = MultiFstatAtomVector(len(dets))
    for X, (ts, cols) in enumerate(dets):
        multi.data[X] = FstatAtomVector.from_arrays(
            Tsft, ts, cols[0], cols[1], cols[2], cols[3] + 1j * cols[4],
            cols[5] + 1j * cols[6],
        )
    return multi


# ----------------------------------------------------------------------------
# Single-template workloads
# ----------------------------------------------------------------------------


def workload_rect_plain(n_sft=500, n_t0=500, n_tau=500, gap_frac=0.10,
                        transient_sfts=150, seed_index=0):
    """Rect window, 1 template, 1 detector (500 SFTs, 10% gaps, 150-SFT
    transient), 500 t0 x 500 tau in 1800 s steps."""
    rng = rng_for(seed_index)
    keep = gap_mask(rng, n_sft, gap_frac)
    idx = np.nonzero(keep)[0]
    t = T_START + idx.astype(np.float64) * TSFT
    phi_a, phi_b, psi = rng.uniform(0, 2 * np.pi, 3)
    start = int(rng.integers(0, max(n_sft - transient_sfts, 1)))
    sig = rect_signal(t, T_START + start * TSFT, transient_sfts * TSFT, psi)
    cols = detector_atoms(rng, t, phi_a, phi_b, sig)
    multi = multi_from_detectors([(t.astype(np.uint32), cols)])
    win = grid_window(TRANSIENT_RECTANGULAR, T_START, n_t0, n_tau)
    return multi, win


def workload_exp_plain(n_sft=4000, n_t0=2000, n_tau=2000, gap_frac=0.15,
                       tau_inj=2.0e5, seed_index=1):
    """Exp window, 1 template, 2 detectors (4000 SFTs each, independent 15%
    gaps, per-detector phases), injected exponential decay tau=2e5 s,
    2000 t0 x 2000 tau."""
    rng = rng_for(seed_index)
    psi = rng.uniform(0, 2 * np.pi)
    t_on = T_START + int(rng.integers(0, n_sft // 2)) * TSFT
    dets = []
    for _X in range(2):
        keep = gap_mask(rng, n_sft, gap_frac)
        idx = np.nonzero(keep)[0]
        t = T_START + idx.astype(np.float64) * TSFT
        phi_a, phi_b = rng.uniform(0, 2 * np.pi, 2)
        sig = exp_signal(t, t_on, tau_inj, psi)
        dets.append((t.astype(np.uint32), detector_atoms(rng, t, phi_a, phi_b, sig)))
    multi = multi_from_detectors(dets)
    win = grid_window(TRANSIENT_EXPONENTIAL, T_START, n_t0, n_tau)
    return multi, win


# ----------------------------------------------------------------------------
# Batched workloads (device-friendly packed form)
# ----------------------------------------------------------------------------


class SyntheticBatch:
    """Many single-detector templates, packed.

    ``ts`` uint32 [total], ``cols`` float32 [7, total], ``off`` int64 [T+1]
    (template t owns atoms off[t]:off[t+1]), ``windows`` list of
    transientWindowRange_t, ``n_sft`` int64 [T] (binned atom counts).
    """

    def __init__(self, ts, cols, off, windows, n_sft, TAtom=TSFT):
        self.ts = ts
        self.cols = cols
        self.off = off
        self.windows = windows
        self.n_sft = n_sft
        self.TAtom = TAtom

    @property
    def T(self):
        return len(self.windows)

    def cells(self):
        return np.array([(w.t0Band // w.dt0 + 1) * (w.tauBand // w.dtau + 1)
                         for w in self.windows], dtype=np.int64)

    def template(self, t):
        """MultiFstatAtomVector of template t."""
        a, b = int(self.off[t]), int(self.off[t + 1])
        return multi_from_detectors([(self.ts[a:b], self.cols[:, a:b])], self.TAtom)

    def subset(self, idx):
        idx = list(idx)
        ts, cols, off = [], [], [0]
        for t in idx:
            a, b = int(self.off[t]), int(self.off[t + 1])
            ts.append(self.ts[a:b])
            cols.append(self.cols[:, a:b])
            off.append(off[-1] + b - a)
        return SyntheticBatch(np.concatenate(ts), np.ascontiguousarray(np.concatenate(cols, 1)),
                              np.asarray(off, np.int64), [self.windows[t] for t in idx],
                              self.n_sft[idx], self.TAtom)


def workload_rect_batched(T=4096, n_min=1000, n_max=8000, gap_frac=0.10,
                          inj_frac=0.05, seed_index=2):
    """Rect window, T templates with N_sft ~ U[n_min, n_max] (10% random gaps
    each), per-template phases, 5% carrying a rect transient; grid
    N_t0 = N_tau = N_sft // 4 in 1800 s steps from the data start, tau >= 3600 s."""
    rng = rng_for(seed_index)
    n_sft = rng.integers(n_min, n_max + 1, size=T).astype(np.int64)
    inj = rng.random(T) < inj_frac
    phases = rng.uniform(0, 2 * np.pi, size=(T, 3))
    keeps = [gap_mask(rng, int(n), gap_frac) for n in n_sft]
    counts = np.array([k.sum() for k in keeps], dtype=np.int64)
    off = np.zeros(T + 1, dtype=np.int64)
    off[1:] = np.cumsum(counts)
    total = int(off[-1])
    ts = np.empty(total, dtype=np.uint32)
    cols = np.empty((7, total), dtype=np.float32)
    windows = []
    for t in range(T):
        idx = np.nonzero(keeps[t])[0]
        tt = T_START + idx.astype(np.float64) * TSFT
        n = int(n_sft[t])
        if inj[t]:
            length = int(rng.integers(max(n // 10, 2), max(n // 2, 3)))
            start = int(rng.integers(0, max(n - length, 1)))
            sig = rect_signal(tt, T_START + start * TSFT, length * TSFT, phases[t, 2])
        else:
            sig = np.zeros(tt.shape[0], dtype=np.complex128)
        a, b = int(off[t]), int(off[t + 1])
        ts[a:b] = tt.astype(np.uint32)
        cols[:, a:b] = detector_atoms(rng, tt, phases[t, 0], phases[t, 1], sig)
        nm = n // 4
        windows.append(grid_window(TRANSIENT_RECTANGULAR, T_START, nm, nm))
    return SyntheticBatch(ts, cols, off, windows, n_sft)


def workload_exp_shared(T=256, n_sft=8192, n_t0=1024, n_tau=512, gap_frac=0.10,
                        inj_frac=0.05, tau_inj=2.0e5, seed_index=3):
    """Exp window, T templates sharing one timestamp grid (n_sft slots, 10%
    gaps common to both detectors and all templates), 2 detectors per template
    merged per bin, n_t0 x n_tau grid in 1800 s steps.

    Returns (list of MultiFstatAtomVector, window range).
    """
    rng = rng_for(seed_index)
    keep = gap_mask(rng, n_sft, gap_frac)
    idx = np.nonzero(keep)[0]
    t = T_START + idx.astype(np.float64) * TSFT
    templates = []
    for _k in range(T):
        psi = rng.uniform(0, 2 * np.pi)
        if rng.random() < inj_frac:
            t_on = T_START + int(rng.integers(0, n_t0)) * TSFT
            sig = exp_signal(t, t_on, tau_inj, psi)
        else:
            sig = np.zeros(t.shape[0], dtype=np.complex128)
        dets = []
        for _X in range(2):
            phi_a, phi_b = rng.uniform(0, 2 * np.pi, 2)
            dets.append((t.astype(np.uint32), detector_atoms(rng, t, phi_a, phi_b, sig)))
        templates.append(multi_from_detectors(dets))
    win = grid_window(TRANSIENT_EXPONENTIAL, T_START, n_t0, n_tau)
    return templates, win


class MultiDetectorBatch:
    """Many templates of several detectors each, packed like the inputs of
    BatchedTransientMaps: ``ts`` uint32 [total], ``cols`` float32 [7, total],
    ``seg_off`` int64 [S+1] (detector segments), ``seg_begin`` int64 [T+1]
    (segments of each template), ``windows`` (one per template)."""

    def __init__(self, ts, cols, seg_off, seg_begin, windows, TAtom=TSFT, **info):
        self.ts = ts
        self.cols = cols
        self.seg_off = seg_off
        self.seg_begin = seg_begin
        self.windows = windows
        self.TAtom = TAtom
        self.info = info

    @property
    def T(self):
        return len(self.windows)


def workload_exp_refine(T=4096, n_sft=2048, n_t0=256, n_tau=256, gap_frac=0.10,
                        tau_range=(5.0, 150.0), seed_index=4):
    """Exp window, T templates of 2 detectors (n_sft grid slots each,
    independent 10% gaps), every template carrying an exponential transient
    that starts on the 1800 s grid within the first n_t0 slots and decays with
    tau ~ U[tau_range] * 1800 s (off the grid); n_t0 x n_tau map grid from the
    data start in 1800 s steps, tau >= 3600 s.  The refinement workload: grid
    maxima of these maps are the candidates."""
    rng = rng_for(seed_index)
    t_grid = T_START + np.arange(n_sft, dtype=np.float64) * TSFT
    ts, cols, seg_off, seg_begin = [], [], [0], [0]
    windows, t_on, tau_inj = [], [], []
    for _t in range(T):
        psi = rng.uniform(0, 2 * np.pi)
        on = T_START + int(rng.integers(0, n_t0)) * TSFT
        tau = float(rng.uniform(*tau_range)) * TSFT
        for _X in range(2):
            keep = gap_mask(rng, n_sft, gap_frac)
            tt = t_grid[keep]
            phi_a, phi_b = rng.uniform(0, 2 * np.pi, 2)
            sig = exp_signal(tt, on, tau, psi)
            ts.append(tt.astype(np.uint32))
            cols.append(detector_atoms(rng, tt, phi_a, phi_b, sig))
            seg_off.append(seg_off[-1] + tt.shape[0])
        seg_begin.append(seg_begin[-1] + 2)
        windows.append(grid_window(TRANSIENT_EXPONENTIAL, T_START, n_t0, n_tau))
        t_on.append(on)
        tau_inj.append(tau)
    return MultiDetectorBatch(np.concatenate(ts), np.ascontiguousarray(np.concatenate(cols, 1)),
                              np.asarray(seg_off, np.int64), np.asarray(seg_begin, np.int64),
                              windows, t_on=np.asarray(t_on), tau_inj=np.asarray(tau_inj))


# ---------------------------------------------------------------------------
# SFTs for the device atom producer (pyfstat_amd.demod)
# ---------------------------------------------------------------------------

SIDEREAL_DAY = 86164.0905
YEAR = 3.15576e7


class SyntheticSFTs:
    """Host arrays of a synthetic SFT set (the SFTSet constructor's arguments)
    and the injected signal's parameters."""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    def sft_set(self, device=None):
        from .demod import SFTSet

        return SFTSet(self.t_start, self.bins, self.k_min, self.Tsft, self.delay,
                      self.delay_rate, self.a, self.b, self.det_off, device=device)


def tone_sft_bins(kappa, phase, k_min, n_bins):
    """Bins k_min .. k_min+n_bins-1 of a unit complex tone across one SFT whose
    observed frequency is kappa bins and whose phase at the SFT midpoint is
    `phase` cycles: (-1)^k sinc(pi (kappa - k)) exp(2 pi i phase)."""
    k = k_min + np.arange(n_bins)
    frac = phase - np.floor(phase)
    return np.where(k % 2 == 1, -1.0, 1.0) * np.sinc(kappa - k) * np.exp(2j * np.pi * frac)


def synthetic_sfts(n_det=2, n_sft=400, f_band=(100.0, 100.05), Tsft=TSFT, gap_frac=0.10,
                   ref_time=None, inj=None, noise=True, dterms=8, seed_index=60):
    """SFTs of n_det detectors (n_sft slots each, independent gaps) at one sky
    position: antenna patterns a, b with sidereal modulation, an SSB delay of
    an orbital (480 s, one year) and a rotational (0.02 s, one sidereal day)
    term, unit-variance complex Gaussian bins (noise=True) covering every
    template frequency in f_band with dterms bins to spare, and optionally a
    signal: inj = dict(f, fdot, amp_a, amp_b (complex), window=(kind, t_on,
    length)) with kind 'rect' (length = duration) or 'exp' (length = tau); the
    signal in SFT alpha is env(t_mid) (amp_a a + amp_b b) times the tone."""
    rng = rng_for(seed_index)
    t_all = T_START + np.arange(n_sft, dtype=np.int64) * Tsft
    ref = float(T_START + n_sft * Tsft // 2) if ref_time is None else float(ref_time)
    ts, off = [], [0]
    for _X in range(n_det):
        keep = gap_mask(rng, n_sft, gap_frac)
        ts.append(t_all[keep])
        off.append(off[-1] + int(keep.sum()))
    t_start = np.concatenate(ts).astype(np.uint32)
    mid = t_start.astype(np.float64) + 0.5 * Tsft
    a = np.empty(t_start.size)
    b = np.empty(t_start.size)
    for X in range(n_det):
        sl = slice(off[X], off[X + 1])
        ph = 2 * np.pi * mid[sl] / SIDEREAL_DAY + 1.3 * X
        a[sl] = 0.45 + 0.35 * np.cos(ph) + 0.1 * np.cos(2 * ph)
        b[sl] = 0.40 * np.sin(ph) + 0.08 * np.sin(2 * ph + 0.4)
    w_orb, w_rot = 2 * np.pi / YEAR, 2 * np.pi / SIDEREAL_DAY
    delay = 480.0 * np.cos(w_orb * mid + 0.7) + 0.02 * np.cos(w_rot * mid + 0.2)
    rate = -480.0 * w_orb * np.sin(w_orb * mid + 0.7) - 0.02 * w_rot * np.sin(w_rot * mid + 0.2)
    k_lo = int(np.floor(f_band[0] * (1 - 1.2e-4) * Tsft)) - dterms - 2
    k_hi = int(np.ceil(f_band[1] * (1 + 1.2e-4) * Tsft)) + dterms + 2
    n_bins = k_hi - k_lo + 1
    bins = np.zeros((t_start.size, n_bins), dtype=np.complex128)
    if noise:
        bins += (rng.standard_normal(bins.shape) + 1j * rng.standard_normal(bins.shape)) / np.sqrt(2)
    if inj is not None:
        dtau = mid + delay - ref
        f, fd = float(inj["f"]), float(inj.get("fdot", 0.0))
        kind, t_on, length = inj.get("window", ("rect", T_START, n_sft * Tsft))
        rel = mid - t_on
        if kind == "rect":
            env = ((rel >= 0) & (rel < length)).astype(np.float64)
        else:
            env = np.where((rel >= 0) & (rel <= 3 * length), np.exp(-np.clip(rel, 0, None) / length), 0.0)
        amp = env * (complex(inj["amp_a"]) * a + complex(inj["amp_b"]) * b)
        for al in np.nonzero(amp != 0)[0]:
            phase = dtau[al] * (f + 0.5 * fd * dtau[al])
            kappa = (f + fd * dtau[al]) * (1 + rate[al]) * Tsft
            bins[al] += amp[al] * tone_sft_bins(kappa, phase, k_lo, n_bins)
    return SyntheticSFTs(t_start=t_start, bins=bins.astype(np.complex64), k_min=k_lo, Tsft=Tsft,
                         delay=delay, delay_rate=rate, a=a.astype(np.float32),
                         b=b.astype(np.float32), det_off=np.asarray(off, dtype=np.int64),
                         ref_time=ref, inj=inj, f_band=f_band)

