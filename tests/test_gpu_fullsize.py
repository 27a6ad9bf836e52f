"""GPU parity at the benchmark workloads' own sizes.

The FP64 oracle cannot afford whole direct-sum exponential maps of these sizes
inside a test, so exponential maps are checked row by row: one-row oracle
windows (t0 fixed, the full tau range) at seeded rows that include the first
and last rows, the boundaries of the recurrence's row chunks and the rows
whose windows run past the data end.  The fused maximum and lnBtSG are
checked against the returned map itself (first row-major argmax, float64
logsumexp), and the maximum's row against the oracle.  Rectangular maps of
the largest batched templates are checked whole.  Tolerances as in
tests/test_gpu_parity.py.
"""

import numpy as np
import pytest

import oracle
from pyfstat_amd import synth
from pyfstat_amd.atoms import (
    FstatAtomVector,
    MultiFstatAtomVector,
    atoms_to_arrays,
    transientWindowRange_t,
)
from tests.helpers import COND_BAND, COND_MAX, check_argmax, compare_maps

pytestmark = pytest.mark.gpu


def lnb_tol(maxF):
    return 1e-4 + 2e-7 * abs(float(maxF))


def _row_window(w, m):
    return transientWindowRange_t(w.type, w.t0 + m * w.dt0, 0, w.dt0, w.tau, w.tauBand, w.dtau)


def _check_rows(F, multi_or_merged, win, rows, TAtom=synth.TSFT):
    """Rows of the GPU map F against one-row oracle windows."""
    if isinstance(multi_or_merged, tuple):
        merged, t0d = multi_or_merged
    else:
        p = atoms_to_arrays(multi_or_merged)
        merged, t0d = oracle.merge_binned(p.ts, p.cols, p.det_off, p.TAtom)
    worst = 0.0
    for m in rows:
        om = oracle.compute_map(merged, TAtom, t0d, _row_window(win, m))
        rel = np.abs(F[m:m + 1].astype(np.float64) - om.F_mn) / np.maximum(np.abs(om.F_mn), 1e-30)
        near = np.abs(om.cond / COND_MAX - 1.0) < COND_BAND
        rel = np.where(near, 0.0, rel)
        assert np.all(np.isfinite(F[m])), m
        assert rel.max() <= 1e-4, (m, float(rel.max()), int(np.argmax(rel)))
        worst = max(worst, float(rel.max()))
    return worst, (merged, t0d)


def _fused_stats_consistent(F, maxF, t0_ML, tau_ML, lnBtSG, win):
    Fd = F.astype(np.float64)
    j = int(np.argmax(F))  # first occurrence in row-major order
    m, n = divmod(j, F.shape[1])
    assert np.float32(maxF) == F.flat[j]
    assert (t0_ML, tau_ML) == (win.t0 + m * win.dt0, win.tau + n * win.dtau)
    mx = Fd.flat[j]
    lnb = np.log(70.0 / F.size) + mx + np.log(np.exp(Fd - mx).sum())
    assert abs(lnBtSG - lnb) <= lnb_tol(mx), (lnBtSG, lnb)
    return m, n


def test_exp_plain_full_size(hip_ready):
    """The exponential single-map workload at full size: 2 detectors x 4000
    SFTs (15% gaps each), 2000 t0 x 2000 tau -- one large map, so the
    recurrence runs in many row chunks, continued above the map to the first
    atom no window reaches."""
    from pyfstat_amd import tcw_fstat_map_funcs as tcw

    multi, win = synth.workload_exp_plain()
    print("case: exp plain 2000 x 2000", flush=True)
    fm = tcw.hip_compute_transient_fstat_map(multi, win, BtSG=True)
    F = fm.F_mn
    assert F.shape == (2000, 2000)
    rng = np.random.default_rng(101)
    rows = {0, 1, 63, 64, 127, 128, 1023, 1024, 1998, 1999}
    rows |= set(int(x) for x in rng.choice(2000, size=22, replace=False))
    _, mt = _check_rows(F, multi, win, sorted(rows))
    m, n = _fused_stats_consistent(F, fm.maxF, fm.t0_ML, fm.tau_ML, fm.lnBtSG, win)
    om = oracle.compute_map(mt[0], synth.TSFT, mt[1], _row_window(win, m))
    # the maximum's row: the oracle agrees on the column (or a near-tie) and value
    n_o = int(np.argmax(om.F_mn[0]))
    assert n_o == n or abs(om.F_mn[0, n] - om.F_mn[0, n_o]) <= 1e-6 * abs(om.F_mn[0, n_o])
    assert abs(fm.maxF - om.F_mn[0, n]) <= 1e-4 * abs(om.F_mn[0, n])


def test_exp_plain_full_size_scale_invariance(hip_ready):
    """Scaling a2, b2, ab by 16 and Fa, Fb by 4 (powers of two: every
    rounding commutes with the scaling) leaves the full-size map unchanged."""
    from pyfstat_amd import tcw_fstat_map_funcs as tcw

    multi, win = synth.workload_exp_plain()
    fm = tcw.hip_compute_transient_fstat_map(multi, win, BtSG=False)
    scaled = MultiFstatAtomVector(multi.length)
    for X in range(multi.length):
        v = multi.data[X]
        scaled.data[X] = FstatAtomVector.from_arrays(
            v.TAtom, v.timestamp, v.a2_alpha * 16, v.b2_alpha * 16, v.ab_alpha * 16,
            v.Fa_alpha * 4, v.Fb_alpha * 4)
    fs = tcw.hip_compute_transient_fstat_map(scaled, win, BtSG=False)
    np.testing.assert_array_equal(fs.F_mn, fm.F_mn)


def test_rect_batched_largest_templates(hip_ready):
    """Batched rect maps of templates with 7000-8000 SFTs (10% gaps; maps of
    up to 2000 x 2000 cells, the prefix scan's eight 1024-bin rounds), whole
    maps of two of them against the oracle."""
    from pyfstat_amd.batched import BatchedTransientMaps

    batch = synth.workload_rect_batched(T=6, n_min=7000, n_max=8000, seed_index=31)
    b = BatchedTransientMaps.from_synthetic(batch, BtSG=True)
    b.run(use_graph=True)
    maps = b.maps()
    for t in (0, 5):
        print(f"case: rect batched large template {t}: n_sft {batch.n_sft[t]}", flush=True)
        a, e = int(batch.off[t]), int(batch.off[t + 1])
        merged, t0d = oracle.merge_binned(batch.ts[a:e], batch.cols[:, a:e],
                                          np.array([0, e - a]), batch.TAtom)
        assert merged.shape[1] >= 7000
        om = oracle.compute_map(merged, batch.TAtom, t0d, batch.windows[t])
        compare_maps(maps[t].F_mn, om)
        m, n = maps[t].get_maxF_idx()
        check_argmax(maps[t].F_mn, int(m), int(n), om)
        assert abs(maps[t].lnBtSG - oracle.lnBtSG(om.F_mn, om.maxF)) <= lnb_tol(om.maxF)
    b.close()


def test_exp_shared_full_shape(hip_ready):
    """The shared-timestamp exponential batch at full shape: 256 templates x
    8192 slots (2 detectors, 10% gaps), 1024 t0 x 512 tau each; two templates
    checked row by row against the oracle (chunk boundaries of the default
    128-row work items included), every template's fused statistics against
    its own map."""
    from pyfstat_amd.batched import BatchedTransientMaps

    multis, win = synth.workload_exp_shared()
    assert len(multis) == 256
    b = BatchedTransientMaps.from_multi_atoms(multis, [win] * len(multis), BtSG=True)
    print("case: exp shared 256 x 1024 x 512", flush=True)
    b.run(use_graph=True)
    maps = b.maps()
    rows = [0, 1, 127, 128, 255, 256, 511, 512, 767, 1022, 1023]
    for t in (3, 200):
        print(f"case: exp shared template {t}", flush=True)
        assert maps[t].F_mn.shape == (1024, 512)
        _check_rows(maps[t].F_mn, multis[t], win, rows)
    for t in range(0, 256, 17):
        mp = maps[t]
        _fused_stats_consistent(mp.F_mn, mp.maxF, mp.t0_ML, mp.tau_ML, mp.lnBtSG, win)
    b.close()


@pytest.mark.parametrize("k", [-40, -20, 20, 40])
def test_atom_scale_sweep(hip_ready, k):
    """a2, b2, ab x 2^k and Fa, Fb x 2^(k/2): F is invariant, so rect and exp
    maps must match the oracle's (unscaled) values across this range of atom
    magnitudes (float32 intermediates of the quotient stay in range for
    |k| <= 40 at these window lengths)."""
    from pyfstat_amd import tcw_fstat_map_funcs as tcw

    s2 = np.float32(2.0 ** k)
    s1 = np.float32(2.0 ** (k // 2))
    for kind in ("rect", "exp"):
        if kind == "rect":
            multi, win = synth.workload_rect_plain(n_sft=600, n_t0=150, n_tau=200, seed_index=90)
        else:
            multi, win = synth.workload_exp_plain(n_sft=500, n_t0=120, n_tau=150, seed_index=91)
        print(f"case: scale sweep k={k} {kind}", flush=True)
        om = oracle.compute_transient_fstat_map(atoms_to_arrays(multi), win, BtSG=True)
        scaled = MultiFstatAtomVector(multi.length)
        for X in range(multi.length):
            v = multi.data[X]
            scaled.data[X] = FstatAtomVector.from_arrays(
                v.TAtom, v.timestamp, v.a2_alpha * s2, v.b2_alpha * s2, v.ab_alpha * s2,
                v.Fa_alpha * s1, v.Fb_alpha * s1)
        fs = tcw.hip_compute_transient_fstat_map(scaled, win, BtSG=True)
        compare_maps(fs.F_mn, om)
        m, n = fs.get_maxF_idx()
        check_argmax(fs.F_mn, int(m), int(n), om)
        assert abs(fs.lnBtSG - om.lnBtSG) <= lnb_tol(om.maxF)
