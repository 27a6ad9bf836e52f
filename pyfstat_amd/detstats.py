"""Detection statistics built on batched transient maps: semi-coherent 2F
(per-segment sums), per-detector 2F at the multi-detector maximum (the input
of transient BSGL), and cumulative 2F -- each for many templates or many
durations in one device pass.

Reference call sites, each of which runs one map per template (or per
duration, or per detector) from Python:
  SemiCoherentSearch.get_semicoherent_twoF / _get_per_segment_twoF /
      get_semicoherent_single_IFO_twoFs   pyfstat/core.py:2150-2275
      (window from init_semicoherent_parameters, core.py:2005-2053)
  ComputeFstat.get_transient_detstats / get_transient_log10BSGL
                                          pyfstat/core.py:1412-1545
  ComputeFstat.calculate_twoF_cumulative  pyfstat/core.py:1547-1671

The log10BSGL formula itself (lalpulsar.ComputeBSGL) lives in lalpulsar; it
is called when lalpulsar is importable, else only the 2F inputs are returned.
"""

import numpy as np

from .atoms import TRANSIENT_RECTANGULAR, atoms_to_arrays, transientWindowRange_t
from .batched import BatchedTransientMaps


def _pack(multi_list, per_detector):
    """ts, cols, seg_off, seg_begin for the templates, followed (if
    per_detector) by one single-detector template per detector of each."""
    ts, cols, seg_off, seg_begin = [], [], [0], [0]
    TAtom = None
    dets = []
    packed = [atoms_to_arrays(m) for m in multi_list]
    for p in packed:
        if TAtom is None:
            TAtom = p.TAtom
        elif p.TAtom != TAtom:
            raise ValueError("all templates must share one TAtom")
        base = seg_off[-1]
        for X in range(p.numDet):
            seg_off.append(base + int(p.det_off[X + 1]))
        seg_begin.append(seg_begin[-1] + p.numDet)
        ts.append(p.ts)
        cols.append(p.cols)
        dets.append(p.numDet)
    if per_detector:
        for p in packed:
            for X in range(p.numDet):
                a, e = int(p.det_off[X]), int(p.det_off[X + 1])
                seg_off.append(seg_off[-1] + (e - a))
                seg_begin.append(seg_begin[-1] + 1)
                ts.append(p.ts[a:e])
                cols.append(p.cols[:, a:e])
    return (np.concatenate(ts), np.ascontiguousarray(np.concatenate(cols, axis=1)),
            np.asarray(seg_off, np.int64), np.asarray(seg_begin, np.int64), TAtom,
            np.asarray(dets, np.int64))


def _bsgl(twoF, twoFX, BSGLSetup):
    if BSGLSetup is None:
        return None
    import lalpulsar  # only where lalsuite is installed

    return np.array([lalpulsar.ComputeBSGL(float(f), np.asarray(fx, np.float32), BSGLSetup)
                     for f, fx in zip(twoF, twoFX)])


# ----------------------------------------------------------------------------
# semi-coherent 2F
# ----------------------------------------------------------------------------


def semicoherent_window(minStartTime, maxStartTime, nsegs, Tsft=1800):
    """(windowRange, tboundaries, Tcoh) of SemiCoherentSearch: nsegs equal
    segments of [minStartTime, maxStartTime] as a rectangular window range with
    N_t0 = nsegs start times and one duration Tcoh (core.py:2005-2019,
    2037-2053).  Raises RuntimeError if Tcoh < 2 Tsft (core.py:2047-2053)."""
    tb = np.linspace(minStartTime, maxStartTime, nsegs + 1)
    Tcoh = tb[1] - tb[0]
    if Tcoh < 2 * Tsft:
        raise RuntimeError(
            "Per-segment coherent time {} may not be < Tsft={}"
            " to avoid degenerate F-statistic computations".format(Tcoh, Tsft))
    w = transientWindowRange_t(
        type=TRANSIENT_RECTANGULAR,
        t0=int(tb[0]),
        t0Band=int(tb[-1] - tb[0] - Tcoh),
        dt0=int(Tcoh),
        tau=int(Tcoh),
        tauBand=0,
        dtau=1,
    )
    return w, tb, Tcoh


def semicoherent_twoF(multi_list, window, singleFstats=False, BSGLSetup=None,
                      use_graph=False):
    """Semi-coherent 2F for every template: per-segment 2F = 2 F_mn[:, 0] of
    the segment window map, summed with NaNs treated as zero (core.py:2199-2206,
    2268-2275); with singleFstats also the per-detector sums
    (core.py:2228-2256).  Returns a dict with "twoF" [T],
    "twoF_per_segment" [T, nsegs] and, if requested, "twoFX" [T, X] and
    "twoFX_per_segment" [T, X, nsegs] (X = detectors of the template, padded
    with NaN), and "log10BSGL" [T] when a lalpulsar BSGLSetup is given."""
    ts, cols, seg_off, seg_begin, TAtom, dets = _pack(multi_list, singleFstats)
    T = len(multi_list)
    nT = len(seg_begin) - 1
    b = BatchedTransientMaps(ts, cols, seg_off, seg_begin, [window] * nT, TAtom)
    try:
        b.run(use_graph=use_graph)
        F = b.F_host()
        nsegs = int(b.F_off[1] - b.F_off[0])
        per = 2.0 * F[: b.F_off[T]].reshape(T, nsegs).astype(np.float64)
        out = {"twoF_per_segment": per, "twoF": np.nan_to_num(per, nan=0.0).sum(axis=1)}
        if singleFstats:
            Xmax = int(dets.max())
            perX = np.full((T, Xmax, nsegs), np.nan)
            k = T
            for t in range(T):
                for X in range(int(dets[t])):
                    perX[t, X] = 2.0 * F[b.F_off[k]:b.F_off[k + 1]]
                    k += 1
            out["twoFX_per_segment"] = perX
            out["twoFX"] = np.where(np.isnan(perX).all(axis=2), np.nan,
                                    np.nan_to_num(perX, nan=0.0).sum(axis=2))
            bs = _bsgl(out["twoF"], out["twoFX"], BSGLSetup)
            if bs is not None:
                out["log10BSGL"] = np.nan_to_num(bs, nan=0.0)
        return out
    finally:
        b.close()


# ----------------------------------------------------------------------------
# transient statistics with per-detector 2F at the maximum
# ----------------------------------------------------------------------------


def transient_detstats(multi_list, windows, BtSG=False, singleFstats=False, BSGLSetup=None,
                       use_graph=False):
    """maxTwoF (NaN -> 0), lnBtSG, t0_ML/tau_ML (and the argmax indices m_ML,
    n_ML) for every template and, with
    singleFstats, twoFXatMaxTwoF [T, X]: each detector's 2F at the cell that
    maximises the multi-detector F (core.py:1447-1471, 1503-1545), all maps in
    one device pass."""
    if not isinstance(windows, (list, tuple)):
        windows = [windows] * len(multi_list)
    ts, cols, seg_off, seg_begin, TAtom, dets = _pack(multi_list, singleFstats)
    T = len(multi_list)
    wins = list(windows)
    if singleFstats:
        for t in range(T):
            wins += [windows[t]] * int(dets[t])
    b = BatchedTransientMaps(ts, cols, seg_off, seg_begin, wins, TAtom, BtSG=BtSG)
    try:
        b.run(use_graph=use_graph)
        st = b.stats()
        maxTwoF = np.nan_to_num(2.0 * st["maxF"][:T].astype(np.float64), nan=0.0)
        out = {
            "maxTwoF": maxTwoF,
            "t0_ML": st["t0_ML"][:T].copy(),
            "tau_ML": st["tau_ML"][:T].copy(),
            "m_ML": st["m_ML"][:T].copy(),
            "n_ML": st["n_ML"][:T].copy(),
        }
        if BtSG:
            out["lnBtSG"] = st["lnBtSG"][:T].copy()
            out["t0_MP"] = st["t0_MP"][:T].copy()
            out["tau_MP"] = st["tau_MP"][:T].copy()
        if singleFstats:
            import torch

            Xmax = int(dets.max())
            idx = []
            k = T
            for t in range(T):
                cell = int(st["m_ML"][t]) * int(st["N_tau"][t]) + int(st["n_ML"][t])
                for _X in range(int(dets[t])):
                    idx.append(int(b.F_off[k]) + cell)
                    k += 1
            vals = b.F_d[torch.as_tensor(idx, device=b.device)].cpu().numpy()
            twoFX = np.full((T, Xmax), np.nan)
            j = 0
            for t in range(T):
                n = int(dets[t])
                twoFX[t, :n] = 2.0 * vals[j:j + n]
                j += n
            out["twoFXatMaxTwoF"] = twoFX
            bs = _bsgl(maxTwoF, twoFX, BSGLSetup)
            if bs is not None:
                out["log10BSGL"] = bs
        return out
    finally:
        b.close()


# ----------------------------------------------------------------------------
# cumulative 2F
# ----------------------------------------------------------------------------


def cumulative_durations(tstart, tend, num_segments, minStartTime, maxStartTime, Tsft=1800):
    """(tstart, tend, durations) of ComputeFstat._set_up_cumulative_times
    (core.py:1547-1571): durations = linspace(2 Tsft, tend - tstart, num_segments)."""
    tstart = max(tstart, minStartTime) if tstart else minStartTime
    tend = min(tend, maxStartTime) if tend else maxStartTime
    return tstart, tend, np.linspace(2 * Tsft, tend - tstart, num_segments)


def twoF_cumulative(multi, tstart=None, tend=None, num_segments=1000, minStartTime=None,
                    maxStartTime=None, Tsft=1800):
    """Cumulative 2F over [tstart, tstart + duration] for the durations of
    calculate_twoF_cumulative (core.py:1574-1671), where each duration is a
    1x1 rectangular window map (t0 = int(tstart), tau = int(duration),
    core.py:1447-1449).

    A rectangular window's F depends on tau only through
    i_t1 = (t0 + tau - t0_data + TAtom/2) / TAtom - 1, which moves by one per
    TAtom of tau, so every duration's 1x1 map equals one cell of a single-row
    map with tau on a TAtom grid; that row is computed in one launch and the
    cells are picked out.  Returns (tstart, durations, twoFs)."""
    import torch

    from . import _hiplib  # noqa: F401  (fail loudly without the library)

    p = atoms_to_arrays(multi)
    TA = p.TAtom
    if minStartTime is None:
        minStartTime = p.t_min
    if maxStartTime is None:
        maxStartTime = p.t_max + TA
    tstart, tend, durations = cumulative_durations(tstart, tend, num_segments, minStartTime,
                                                   maxStartTime, Tsft)
    t0 = int(tstart)
    taus = durations.astype(np.int64)  # int(tend - tstart), UINT4
    if np.any(taus < 0) or t0 < 0 or t0 + int(taus.max()) >= 2**32:
        raise ValueError("cumulative windows must lie within [0, 2^32) s")
    t0_data = int(p.t_min)
    half = TA // 2
    # breakpoint class of each duration: floor((t0 + tau - t0_data + TA/2) / TA)
    x = t0 + taus - t0_data + half
    if np.any(x < 0):
        raise ValueError("cumulative windows start before the data")
    q = x // TA
    q0 = int(q.min())
    tau0 = int(taus[np.argmin(q)])
    n_cols = int(q.max()) - q0 + 1
    w = transientWindowRange_t(type=TRANSIENT_RECTANGULAR, t0=t0, t0Band=0, dt0=TA, tau=tau0,
                               tauBand=(n_cols - 1) * TA, dtau=TA)
    seg_off = np.asarray(p.det_off, dtype=np.int64)
    b = BatchedTransientMaps(p.ts, p.cols, seg_off, np.array([0, p.numDet], np.int64), [w], TA)
    try:
        b.run()
        idx = torch.as_tensor(q - q0, device=b.device)
        twoFs = 2.0 * b.F_d[idx].cpu().numpy().astype(np.float64)
    finally:
        b.close()
    twoFs = np.nan_to_num(twoFs, nan=0.0)
    return tstart, durations, twoFs
