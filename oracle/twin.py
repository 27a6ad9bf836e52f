"""Plain-Python twin of the C oracle, for tiny inputs -- TEST INFRASTRUCTURE ONLY.

A second, independently written restatement of the same reference steps, used
only to cross-check ``tcw_oracle.c``:

* binned merge: lalpulsar mergeMultiFstatAtomsBinned (called at
  pyfstat/tcw_fstat_map_funcs.py:706), float32 accumulation in detector order;
* map: pyfstat/tcw_fstat_map_funcs.py:742-815 with the per-cell window logic of
  pyfstat/pyCUDAkernels/cudaTransientFstatRectWindow.cu:21-124 and
  pyfstat/pyCUDAkernels/cudaTransientFstatExpWindow.cu:27-134, evaluated in
  FP64 Python floats, with uint32 wrap-around emulated by ``& 0xFFFFFFFF``.
"""

import math

import numpy as np

M32 = 0xFFFFFFFF


def _to_int32(u):
    u &= M32
    return u - (1 << 32) if u & 0x80000000 else u


def merge(ts, cols, det_off, TAtom):
    tmin = min(int(ts[det_off[X]]) for X in range(len(det_off) - 1))
    tmax = max(int(ts[det_off[X + 1] - 1]) for X in range(len(det_off) - 1))
    n = (tmax - tmin) // TAtom + 1
    out = np.zeros((7, n), dtype=np.float32)
    for X in range(len(det_off) - 1):
        for k in range(det_off[X], det_off[X + 1]):
            i = (int(ts[k]) - tmin) // TAtom
            for c in range(7):
                out[c, i] = np.float32(out[c, i] + cols[c, k])
    return out, tmin


def _idx(t, t0_data, TAtom, N, minus_one):
    half = TAtom // 2
    q = ((t - t0_data + half) & M32) // TAtom
    if minus_one:
        q = (q - 1) & M32
    i = _to_int32(q)
    if i < 0:
        i = 0
    if i >= N:
        i = N - 1
    return i


def _F(A, B, C, Far, Fai, Fbr, Fbi):
    sumAB = A + B
    diffAB = A - B
    disc = math.sqrt(diffAB * diffAB + 4.0 * C * C)
    denom = sumAB - disc
    cond = (sumAB + disc) / denom if denom > 0 else math.inf
    DdInv = 1.0 / (A * B - C * C) if cond < 1e4 else 0.0
    if DdInv > 0:
        return DdInv * (B * (Far * Far + Fai * Fai) + A * (Fbr * Fbr + Fbi * Fbi)
                        - 2.0 * C * (Far * Fbr + Fai * Fbi))
    return 2.0


def transient_map(merged, TAtom, t0_data, win):
    """win = (type, t0, t0Band, dt0, tau, tauBand, dtau); returns F float64 and
    the effective window after the NONE conversion."""
    typ, t0w, t0Band, dt0, tauw, tauBand, dtau = [int(x) for x in win]
    N = merged.shape[1]
    if typ == 0:
        typ, t0w, t0Band, dt0, tauw, tauBand, dtau = 1, t0_data, 0, TAtom, N * TAtom, 0, TAtom
    Nt0 = t0Band // dt0 + 1
    Ntau = tauBand // dtau + 1
    x = [[float(v) for v in merged[c]] for c in range(7)]
    F = np.empty((Nt0, Ntau))
    for m in range(Nt0):
        t0 = (t0w + m * dt0) & M32
        i0 = _idx(t0, t0_data, TAtom, N, False)
        acc = [0.0] * 7
        last = i0
        for n in range(Ntau):
            tau = (tauw + n * dtau) & M32
            t1 = (t0 + (3 * tau if typ == 2 else tau)) & M32
            i1 = _idx(t1, t0_data, TAtom, N, True)
            if typ == 1:
                i = last
                while i <= i1:
                    for c in range(7):
                        acc[c] += x[c][i]
                    last = i1 + 1
                    i += 1
            else:
                acc = [0.0] * 7
                for i in range(i0, i1 + 1):
                    ti = (t0_data + i * TAtom) & M32
                    w = math.exp(-((ti - t0) & M32) / tau) if t0 <= ti <= t1 else 0.0
                    for c in range(7):
                        acc[c] += x[c][i] * (w * w if c < 3 else w)
            F[m, n] = _F(*acc)
    return F, (typ, t0w, t0Band, dt0, tauw, tauBand, dtau)
