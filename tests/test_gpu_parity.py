"""Parity of the "hip" map version against the FP64 oracle, on the GPU.

Every test calls the product path (libtcw_hip through the C ABI, via
pyfstat_amd.tcw_fstat_map_funcs / pyfstat_amd.batched) and compares with the
oracle on the same seeded inputs:
  F_mn          max relative error <= 1e-4 outside the cond ~ 1e4 band
  t0_ML/tau_ML  equal, or a near-tie (oracle F within 1e-6 of its maximum)
  lnBtSG        absolute error <= 1e-4 + 2e-7 * maxF  (the map is float32,
                tcw_fstat_map_funcs.py:134, so lnBtSG = maxF + log sum ... carries
                float32 rounding of maxF itself)
  t0_MP/tau_MP  equal, or a near-tie of the posterior sums (1e-6)
"""

import numpy as np
import pytest

import oracle
from pyfstat_amd import synth
from pyfstat_amd.atoms import (
    TRANSIENT_EXPONENTIAL,
    TRANSIENT_NONE,
    TRANSIENT_RECTANGULAR,
    transientWindowRange_t,
)
from tests.helpers import check_argmax, compare_maps, oracle_for

pytestmark = pytest.mark.gpu

LNB_ATOL = 1e-4


def lnb_tol(maxF):
    return LNB_ATOL + 2e-7 * abs(float(maxF))


def _posterior_ok(om, window, t0_MP, tau_MP):
    E = np.exp(om.F_mn - om.maxF)
    rows = E.sum(axis=1)
    cols = E.sum(axis=0)
    t0_o, tau_o = oracle.posterior_modes(om.F_mn, window, om.maxF)
    if t0_MP != pytest.approx(t0_o, abs=1e-6):
        m = int(round((t0_MP - window.t0) / (window.t0Band / om.F_mn.shape[0]) - 0.5))
        assert rows[m] >= rows.max() * (1 - 1e-6), (t0_MP, t0_o)
    if tau_MP != pytest.approx(tau_o, abs=1e-6):
        n = int(round((tau_MP - window.tau) / (window.tauBand / om.F_mn.shape[1]) - 0.5))
        assert cols[n] >= cols.max() * (1 - 1e-6), (tau_MP, tau_o)


def _check_full(multi, window, force_generic=False, BtSG=True):
    from pyfstat_amd import tcw_fstat_map_funcs as tcw

    print(f"case: window={window} generic={force_generic}", flush=True)
    fm = tcw.hip_compute_transient_fstat_map(multi, window, BtSG=BtSG,
                                             force_generic=force_generic)
    om = oracle_for(multi, window, BtSG=BtSG)
    errs = compare_maps(fm.F_mn, om)
    m, n = fm.get_maxF_idx()
    check_argmax(fm.F_mn, int(m), int(n), om)
    # fused argmax agrees with the returned map's own argmax
    w_eff = om.window_eff
    assert fm.t0_ML == w_eff[1] + int(m) * w_eff[3]
    assert fm.tau_ML == w_eff[4] + int(n) * w_eff[6]
    assert float(fm.maxF) == float(fm.F_mn.max())
    if BtSG:
        assert abs(fm.lnBtSG - om.lnBtSG) <= lnb_tol(om.maxF), (fm.lnBtSG, om.lnBtSG)
        _posterior_ok(om, window, fm.t0_MP, fm.tau_MP)
    return fm, om, errs


def test_rect_plain_500(hip_ready):
    """Rect window, 1 template, 500 SFTs (10% gaps), 500 x 500 grid."""
    multi, win = synth.workload_rect_plain()
    fm, om, errs = _check_full(multi, win)
    print(errs)
    # same map through the generic per-cell kernel
    fm2, _, _ = _check_full(multi, win, force_generic=True)
    assert np.max(np.abs(fm2.F_mn - fm.F_mn) / np.maximum(fm.F_mn, 1e-30)) < 1e-5


def test_rect_full_span_cell_is_full_F(hip_ready):
    """F_mn[0, -1] over the full span equals the all-data F
    (tests/test_tcw_fstat_map_funcs.py:124-127), and TRANSIENT_NONE gives it."""
    from pyfstat_amd import tcw_fstat_map_funcs as tcw

    multi, _ = synth.workload_rect_plain(n_sft=300, n_t0=10, n_tau=10)
    N = 300
    win = transientWindowRange_t(TRANSIENT_RECTANGULAR, synth.T_START, 0, 1800,
                                 N * 1800, 0, 1800)
    fm = tcw.hip_compute_transient_fstat_map(multi, win, BtSG=True)
    none = transientWindowRange_t(TRANSIENT_NONE, 123, 0, 0, 456, 0, 0)
    before = repr(none)
    fn = tcw.hip_compute_transient_fstat_map(multi, none, BtSG=True)
    assert repr(none) == before  # caller's window range untouched
    assert fn.F_mn.shape == (1, 1)
    assert fn.F_mn[0, 0] == pytest.approx(fm.F_mn[0, 0], rel=1e-6)
    om = oracle_for(multi, none)
    compare_maps(fn.F_mn, om)
    assert fn.t0_ML == synth.T_START
    assert fn.tau_ML == N * 1800


@pytest.mark.parametrize(
    "case",
    [
        # non-square grid (rect.cu:48 bounds rows by N_tauRange)
        dict(t0=0, t0Band=59 * 1800, dt0=1800, tau=3600, tauBand=149 * 1800, dtau=1800),
        dict(t0=0, t0Band=149 * 1800, dt0=1800, tau=3600, tauBand=19 * 1800, dtau=1800),
        # coarser regular grid, a = 2
        dict(t0=0, t0Band=60 * 3600, dt0=3600, tau=7200, tauBand=60 * 3600, dtau=3600),
        # dt0 != dtau and off-grid spacings: generic kernel
        dict(t0=0, t0Band=80 * 1800, dt0=1800, tau=3600, tauBand=40 * 3600, dtau=3600),
        dict(t0=500, t0Band=100 * 1000, dt0=1000, tau=2500, tauBand=90 * 1300, dtau=1300),
        # start times before the data (uint32 wrap-around of t0 - t0_data)
        dict(t0=-20 * 1800, t0Band=60 * 1800, dt0=1800, tau=3600, tauBand=60 * 1800, dtau=1800),
        # tau = 0 and t0 past the end of the data
        dict(t0=250 * 1800, t0Band=80 * 1800, dt0=1800, tau=0, tauBand=30 * 1800, dtau=1800),
    ],
)
def test_rect_window_variants(hip_ready, case):
    multi, _ = synth.workload_rect_plain(n_sft=300, n_t0=10, n_tau=10)
    c = dict(case)
    c["t0"] = synth.T_START + c["t0"]
    win = transientWindowRange_t(type=TRANSIENT_RECTANGULAR, **c)
    _check_full(multi, win)
    _check_full(multi, win, force_generic=True)


def test_exp_two_detectors(hip_ready):
    """Exp window, 2 interleaved detectors, merged per bin."""
    multi, _ = synth.workload_exp_plain(n_sft=600)
    win = synth.grid_window(TRANSIENT_EXPONENTIAL, synth.T_START, 200, 150)
    _check_full(multi, win)


@pytest.mark.parametrize(
    "case",
    [
        dict(t0=900, t0Band=100 * 1700, dt0=1700, tau=2000, tauBand=50 * 2100, dtau=2100),
        dict(t0=-10 * 1800, t0Band=40 * 1800, dt0=1800, tau=1800, tauBand=40 * 5400, dtau=5400),
    ],
)
def test_exp_window_variants(hip_ready, case):
    multi, _ = synth.workload_exp_plain(n_sft=400)
    c = dict(case)
    c["t0"] = synth.T_START + c["t0"]
    win = transientWindowRange_t(type=TRANSIENT_EXPONENTIAL, **c)
    _check_full(multi, win)


@pytest.mark.parametrize(
    "case",
    [
        # on the atom grid: the recurrence kernel
        dict(t0=0, t0Band=299 * 1800, dt0=1800, tau=3600, tauBand=199 * 1800, dtau=1800),
        dict(t0=37 * 1800, t0Band=150 * 3600, dt0=3600, tau=1800, tauBand=90 * 5400, dtau=5400),
        # rows past the end of the data, windows past it
        dict(t0=500 * 1800, t0Band=200 * 1800, dt0=1800, tau=7200, tauBand=300 * 1800, dtau=1800),
    ],
)
def test_exp_recurrence_vs_oracle_and_direct(hip_ready, case):
    from pyfstat_amd import tcw_fstat_map_funcs as tcw

    multi, _ = synth.workload_exp_plain(n_sft=700, seed_index=31)
    c = dict(case)
    c["t0"] = synth.T_START + c["t0"]
    win = transientWindowRange_t(type=TRANSIENT_EXPONENTIAL, **c)
    fm, om, errs = _check_full(multi, win)
    fd = tcw.hip_compute_transient_fstat_map(multi, win, BtSG=True, force_generic=True)
    rel = np.abs(fd.F_mn - fm.F_mn) / np.maximum(np.abs(fm.F_mn), 1e-30)
    near = np.abs(om.cond / 1e4 - 1) < 1e-3
    assert np.where(near, 0, rel).max() < 2e-5


def test_merge_offgrid_and_duplicate_bins(hip_ready):
    """Timestamps off the TAtom grid, several atoms per bin, 3 detectors."""
    from pyfstat_amd.atoms import FstatAtomVector, MultiFstatAtomVector

    rng = np.random.default_rng(7)
    multi = MultiFstatAtomVector(3)
    for X in range(3):
        n = 150 + 20 * X
        ts = np.sort(synth.T_START + rng.integers(0, 200 * 1800, size=n)).astype(np.uint32)
        ts[0] = synth.T_START + 7 * X
        cols = rng.standard_normal((7, n)).astype(np.float32)
        cols[0] = np.abs(cols[0])
        cols[1] = np.abs(cols[1])
        multi.data[X] = FstatAtomVector.from_arrays(1800, ts, cols[0], cols[1], cols[2],
                                                    cols[3] + 1j * cols[4],
                                                    cols[5] + 1j * cols[6])
    for wtype in (TRANSIENT_RECTANGULAR, TRANSIENT_EXPONENTIAL):
        win = synth.grid_window(wtype, synth.T_START, 120, 80)
        _check_full(multi, win)


def _long_multi(seed=31, n_bins=3000):
    """3 detectors over ~n_bins TAtom slots, 10-30% gaps, a second atom in one
    bin of seven, detector start offsets inside the first bin."""
    from pyfstat_amd.atoms import FstatAtomVector, MultiFstatAtomVector

    rng = np.random.default_rng(seed)
    multi = MultiFstatAtomVector(3)
    for X in range(3):
        slots = np.nonzero(synth.gap_mask(rng, n_bins, 0.1 + 0.1 * X))[0]
        dup = rng.choice(slots, size=len(slots) // 7, replace=False)
        t = np.sort(np.concatenate([slots * synth.TSFT, dup * synth.TSFT + 600]))
        t = (t + synth.T_START + 13 * X).astype(np.int64)
        cols = synth.detector_atoms(rng, t.astype(np.float64), 0.3 * X, 1.1 + X,
                                    np.zeros(t.size, np.complex128))
        multi.data[X] = FstatAtomVector.from_arrays(synth.TSFT, t, cols[0], cols[1], cols[2],
                                                    cols[3] + 1j * cols[4],
                                                    cols[5] + 1j * cols[6])
    return multi


@pytest.mark.parametrize("force_generic", [False, True])
def test_rect_long_multi_detector_merge(hip_ready, force_generic):
    """~3000 bins and ~2600-3000 atoms per detector: the merge and prefix scan
    run over several rounds of bins and several LDS chunks of atoms per
    detector; window starts and ends cover the whole span (dt0 = dtau = 37
    TAtom, so the regular-grid kernel takes it unless forced generic)."""
    multi = _long_multi()
    win = synth.grid_window(TRANSIENT_RECTANGULAR, synth.T_START, 75, 70, dt0=37 * synth.TSFT,
                            dtau=37 * synth.TSFT)
    _check_full(multi, win, force_generic=force_generic)


def test_batched_rect_long_templates(hip_ready):
    """Ragged batch of templates with 1500-2600 SFTs (up to three rounds of
    the prefix scan each) against the oracle."""
    from pyfstat_amd.batched import BatchedTransientMaps

    batch = synth.workload_rect_batched(T=3, n_min=1500, n_max=2600, seed_index=13)
    b = BatchedTransientMaps.from_synthetic(batch, BtSG=True)
    b.run()
    maps = b.maps()
    for t in range(batch.T):
        print(f"long batched template {t}", flush=True)
        om = oracle_for(batch.template(t), batch.windows[t])
        compare_maps(maps[t].F_mn, om)
        assert abs(maps[t].lnBtSG - om.lnBtSG) <= lnb_tol(om.maxF)
        _posterior_ok(om, batch.windows[t], maps[t].t0_MP, maps[t].tau_MP)
    b.close()


@pytest.mark.parametrize("amp", [0.0, 3.0])
@pytest.mark.parametrize("wtype", [TRANSIENT_RECTANGULAR, TRANSIENT_EXPONENTIAL])
def test_reference_test_configuration(hip_ready, wtype, amp):
    """The set-up of tests/test_tcw_fstat_map_funcs.py:15-127: one day of 1800 s
    SFTs from GPS 700000000, one detector, t0 over [start, start + day - 2 Tsft],
    tau over [2 Tsft, day] in Tsft steps, a transient with tau = 0.25 day at
    t0 = start + 0.25 day (none for amp = 0).  The reference compares its map
    with lalpulsar's search of the same data; here the whole map, ML point,
    lnBtSG and posterior modes are compared with the oracle.  (With one
    detector, windows of 2-3 SFTs sit just below the cond = 1e4 guard and
    reach F ~ 1e4 in noise, so the ML point is not where the injection is;
    that is the formula, and the oracle agrees.)"""
    Tstart, Tsft, day = 700000000, 1800, 86400
    rng = np.random.default_rng(101 + int(amp))
    t = Tstart + Tsft * np.arange(day // Tsft, dtype=np.int64)
    t_on, tau_inj = Tstart + day // 4, day // 4
    if amp == 0:
        sig = np.zeros(t.size, np.complex128)
    elif wtype == TRANSIENT_RECTANGULAR:
        sig = synth.rect_signal(t.astype(np.float64), t_on, tau_inj, 0.7, amp=amp)
    else:
        sig = synth.exp_signal(t.astype(np.float64), t_on, tau_inj, 0.7, amp=amp)
    cols = synth.detector_atoms(rng, t.astype(np.float64), 0.2, 1.3, sig)
    multi = synth.multi_from_detectors([(t, cols)], Tsft=Tsft)
    win = transientWindowRange_t(wtype, Tstart, day - 2 * Tsft, Tsft, 2 * Tsft,
                                 day - 2 * Tsft, Tsft)
    fm, om, _ = _check_full(multi, win)
    assert not np.isnan(fm.maxF) and not np.isnan(fm.lnBtSG)
    if wtype == TRANSIENT_RECTANGULAR:
        # first t0, longest tau: the all-data F (tests/test_tcw_fstat_map_funcs.py:124-127)
        full = oracle_for(multi, transientWindowRange_t(TRANSIENT_NONE, 0, 0, 0, 0, 0, 0))
        assert fm.F_mn[0, -1] == pytest.approx(full.F_mn[0, 0], rel=1e-4)


def test_lal_strict_single_atom(hip_ready):
    """lal-strict mode raises where lalpulsar raises XLAL_EDOM; default mode
    returns the F = 2 fallback there."""
    from pyfstat_amd import _hiplib
    from pyfstat_amd import tcw_fstat_map_funcs as tcw

    multi, win = synth.workload_rect_plain(n_sft=100, n_t0=100, n_tau=20)
    with pytest.raises(_hiplib.SingleAtomError):
        tcw.hip_compute_transient_fstat_map(multi, win, lal_strict=True)
    safe = synth.grid_window(TRANSIENT_RECTANGULAR, synth.T_START, 98, 20)
    tcw.hip_compute_transient_fstat_map(multi, safe, lal_strict=True)
    om = oracle_for(multi, win)
    assert om.bad_cell is not None
    fm = tcw.hip_compute_transient_fstat_map(multi, win)
    compare_maps(fm.F_mn, om)


def test_call_compute_hip(hip_ready):
    """The reference entry point with version="hip"."""
    from pyfstat_amd import tcw_fstat_map_funcs as tcw

    multi, win = synth.workload_rect_plain(n_sft=200, n_t0=50, n_tau=50)
    fm, timing = tcw.call_compute_transient_fstat_map("hip", hip_ready, multi, win, BtSG=True)
    assert timing > 0
    assert not np.isnan(fm.maxF) and not np.isnan(fm.lnBtSG)
    assert fm.F_mn.dtype == np.float32
    # the numpy methods of the container agree with the fused device values
    lnB = fm.lnBtSG
    assert fm.get_lnBtSG() == pytest.approx(lnB, abs=1e-5)
    t0_MP, tau_MP = fm.t0_MP, fm.tau_MP
    assert fm.get_t0_max_posterior(win) == pytest.approx(t0_MP)
    assert fm.get_tau_max_posterior(win) == pytest.approx(tau_MP)


def _batch_small(T=12, seed_index=11):
    return synth.workload_rect_batched(T=T, n_min=200, n_max=900, seed_index=seed_index)


@pytest.mark.parametrize("use_graph", [False, True])
def test_batched_rect_vs_oracle(hip_ready, use_graph):
    from pyfstat_amd.batched import BatchedTransientMaps

    batch = _batch_small()
    b = BatchedTransientMaps.from_synthetic(batch, BtSG=True)
    for rep in range(3 if use_graph else 1):
        b.F_d.fill_(-1.0)  # eager work between replays
        b.run(use_graph=use_graph)
        maps = b.maps()
        for t in range(batch.T):
            print(f"batched template {t} replay {rep}", flush=True)
            om = oracle_for(batch.template(t), batch.windows[t])
            compare_maps(maps[t].F_mn, om)
            m, n = maps[t].get_maxF_idx()
            check_argmax(maps[t].F_mn, int(m), int(n), om)
            assert abs(maps[t].lnBtSG - om.lnBtSG) <= lnb_tol(om.maxF)
            _posterior_ok(om, batch.windows[t], maps[t].t0_MP, maps[t].tau_MP)
    b.close()


def test_batched_matches_single(hip_ready):
    """The batched pass and per-template single calls give identical maps."""
    from pyfstat_amd import tcw_fstat_map_funcs as tcw
    from pyfstat_amd.batched import BatchedTransientMaps

    batch = _batch_small(T=6, seed_index=12)
    b = BatchedTransientMaps.from_synthetic(batch, BtSG=True)
    b.run()
    maps = b.maps()
    for t in range(batch.T):
        fm = tcw.hip_compute_transient_fstat_map(batch.template(t), batch.windows[t], BtSG=True)
        np.testing.assert_array_equal(fm.F_mn, maps[t].F_mn)
        assert fm.lnBtSG == maps[t].lnBtSG
        assert (fm.t0_ML, fm.tau_ML) == (maps[t].t0_ML, maps[t].tau_ML)
    b.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_batched_exp_direct_sum_side_stream(hip_ready, use_graph):
    """Exp batch with enough (template, column block) pairs for the direct
    top-chunk sum and enough workgroups that it runs on the plan's second
    stream beside the chunk walks: every map agrees with the single-map call
    (which takes neither), a sample with the oracle; graph replays and the
    eager run give the same bits."""
    from pyfstat_amd import tcw_fstat_map_funcs as tcw
    from pyfstat_amd.batched import BatchedTransientMaps

    multis, win = synth.workload_exp_shared(T=288, n_sft=700, n_t0=256, n_tau=64, seed_index=44)
    b = BatchedTransientMaps.from_multi_atoms(multis, [win] * len(multis), BtSG=True)
    b.run(use_graph=False)
    ref = [m.F_mn.copy() for m in b.maps()]
    for rep in range(3 if use_graph else 1):
        b.F_d.fill_(-1.0)  # eager work between replays
        b.run(use_graph=use_graph)
        maps = b.maps()
        for t in range(len(multis)):
            assert np.array_equal(maps[t].F_mn, ref[t]), (rep, t)
    for t in range(len(multis)):
        fm = tcw.hip_compute_transient_fstat_map(multis[t], win, BtSG=True)
        np.testing.assert_allclose(maps[t].F_mn, fm.F_mn, rtol=2e-6, atol=0)
        assert (maps[t].t0_ML, maps[t].tau_ML) == (fm.t0_ML, fm.tau_ML)
    for t in (0, 97, 287):
        print(f"exp side-stream batch template {t}", flush=True)
        om = oracle_for(multis[t], win)
        compare_maps(maps[t].F_mn, om)
        assert abs(maps[t].lnBtSG - om.lnBtSG) <= lnb_tol(om.maxF)
    # on a caller's stream right after queued work that doubles every atom:
    # the fork to the plan's second stream waits for it, so F doubles exactly
    # (power-of-two scaling commutes with every rounding; F = 2 fallback
    # cells stay 2)
    import torch

    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        b.cols_d.mul_(2.0)
        b.run(use_graph=use_graph, stream=s)
    s.synchronize()
    maps2 = b.maps()
    for t in range(len(multis)):
        want = np.where(ref[t] == 2.0, np.float32(2.0), np.float32(2.0) * ref[t])
        assert np.array_equal(maps2[t].F_mn, want), t
    b.close()


def test_batched_mixed_windows_multi_detector(hip_ready):
    """Ragged batch mixing rect (regular and generic) and exp windows, 1-3 detectors."""
    from pyfstat_amd.batched import hip_compute_transient_fstat_maps

    multis, wins = [], []
    rng = np.random.default_rng(5)
    for k in range(9):
        if k % 3 == 0:
            multi, _ = synth.workload_exp_plain(n_sft=int(rng.integers(150, 400)), seed_index=20 + k)
        else:
            multi, _ = synth.workload_rect_plain(n_sft=int(rng.integers(150, 400)), seed_index=20 + k)
        wtype = [TRANSIENT_RECTANGULAR, TRANSIENT_EXPONENTIAL][k % 2]
        if k % 4 == 1:
            win = transientWindowRange_t(wtype, synth.T_START + 900, 70 * 1300, 1300, 2000,
                                         40 * 1700, 1700)
        else:
            win = synth.grid_window(wtype, synth.T_START, 60 + 7 * k, 40 + 5 * k)
        multis.append(multi)
        wins.append(win)
    maps = hip_compute_transient_fstat_maps(multis, wins, BtSG=True)
    for k in range(len(multis)):
        om = oracle_for(multis[k], wins[k])
        compare_maps(maps[k].F_mn, om)
        assert abs(maps[k].lnBtSG - om.lnBtSG) <= lnb_tol(om.maxF)
