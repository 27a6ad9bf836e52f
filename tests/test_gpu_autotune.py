"""Every rows-per-item value the autotuner may pick gives maps that match the
oracle (same tolerances as tests/test_gpu_parity.py), and the tuner itself
picks, applies and caches a value per kernel mode."""

import numpy as np
import pytest

import oracle
from pyfstat_amd import _hiplib, autotune, synth
from tests.helpers import check_argmax, compare_maps

pytestmark = pytest.mark.gpu

LNB_ATOL = 1e-4


def _work(mode):
    if mode in (_hiplib.MODE_RECT_FAST, _hiplib.MODE_RECT_GENERIC):
        w = synth.workload_rect_batched(T=4, n_min=300, n_max=900, seed_index=50)
        seg_begin = np.arange(w.T + 1, dtype=np.int64)
        return w.ts, w.cols, w.off, seg_begin, w.windows
    w = synth.workload_exp_refine(T=3, n_sft=500, n_t0=200, n_tau=60, seed_index=51)
    return w.ts, w.cols, w.seg_off, w.seg_begin, w.windows


def _oracle_maps(ts, cols, seg_off, seg_begin, windows):
    out = []
    for t in range(len(windows)):
        s0, s1 = int(seg_begin[t]), int(seg_begin[t + 1])
        a, e = int(seg_off[s0]), int(seg_off[s1])
        m, t0d = oracle.merge_binned(ts[a:e], cols[:, a:e], np.asarray(seg_off[s0:s1 + 1]) - a,
                                     synth.TSFT)
        om = oracle.compute_map(m, synth.TSFT, t0d, windows[t])
        om.lnBtSG = oracle.lnBtSG(om.F_mn, om.maxF)
        out.append(om)
    return out


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_every_candidate_rows_value_matches_oracle(hip_ready, mode):
    from pyfstat_amd.batched import BatchedTransientMaps

    ts, cols, seg_off, seg_begin, windows = _work(mode)
    generic = mode in (_hiplib.MODE_RECT_GENERIC, _hiplib.MODE_EXP_DIRECT)
    oms = _oracle_maps(ts, cols, seg_off, seg_begin, windows)
    b = BatchedTransientMaps(ts, cols, seg_off, seg_begin, windows, synth.TSFT, BtSG=True,
                             force_generic=generic)
    assert b.items_per_mode()[mode] > 0
    for rows in autotune.CANDIDATES[mode]:
        print(f"mode {mode} rows {rows}", flush=True)
        autotune.set_rows(b, mode, rows)
        b.run()
        maps = b.maps()
        for t, om in enumerate(oms):
            compare_maps(maps[t].F_mn, om)
            m, n = maps[t].get_maxF_idx()
            check_argmax(maps[t].F_mn, int(m), int(n), om)
            assert abs(maps[t].lnBtSG - om.lnBtSG) <= LNB_ATOL + 2e-7 * abs(om.maxF)
    b.close()


def test_rows_belong_to_one_plan(hip_ready):
    """Rows set on one batch change neither another batch's plan nor the
    single-map call; a value the library rejects keeps the previous plan."""
    from pyfstat_amd import tcw_fstat_map_funcs as tcw
    from pyfstat_amd.batched import BatchedTransientMaps

    ts, cols, seg_off, seg_begin, windows = _work(_hiplib.MODE_RECT_FAST)
    fresh = BatchedTransientMaps(ts, cols, seg_off, seg_begin, windows, synth.TSFT, BtSG=True)
    items0 = fresh.items_per_mode().copy()
    tuned = BatchedTransientMaps(ts, cols, seg_off, seg_begin, windows, synth.TSFT, BtSG=True)
    autotune.set_rows(tuned, _hiplib.MODE_RECT_FAST, 64)
    assert not np.array_equal(tuned.items_per_mode(), items0)
    with pytest.raises(ValueError, match="multiple of 16"):
        autotune.set_rows(tuned, _hiplib.MODE_RECT_FAST, 24)
    assert tuned.rows_by_mode[_hiplib.MODE_RECT_FAST] == 64
    other = BatchedTransientMaps(ts, cols, seg_off, seg_begin, windows, synth.TSFT, BtSG=True)
    np.testing.assert_array_equal(other.items_per_mode(), items0)
    # the single-map call on the first template still matches the oracle
    multi = synth.multi_from_detectors([(ts[seg_off[0]:seg_off[1]],
                                         cols[:, seg_off[0]:seg_off[1]])])
    fm = tcw.hip_compute_transient_fstat_map(multi, windows[0], BtSG=True)
    compare_maps(fm.F_mn, _oracle_maps(ts, cols, seg_off, seg_begin, windows[:1])[0])
    for b in (fresh, tuned, other):
        b.close()


def test_tune_picks_applies_and_caches(tmp_path, monkeypatch):
    from pyfstat_amd.batched import BatchedTransientMaps

    monkeypatch.setenv("PYFSTAT_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    ts, cols, seg_off, seg_begin, windows = _work(_hiplib.MODE_RECT_FAST)
    b = BatchedTransientMaps(ts, cols, seg_off, seg_begin, windows, synth.TSFT, BtSG=True)
    res = autotune.tune(b, reps=2)
    assert set(res) == {_hiplib.MODE_RECT_FAST}
    r = res[_hiplib.MODE_RECT_FAST]
    assert r["rows"] in autotune.CANDIDATES[_hiplib.MODE_RECT_FAST] and not r["cached"]
    assert (tmp_path / "tune.json").exists()
    again = autotune.tune(b, reps=2)
    assert again[_hiplib.MODE_RECT_FAST]["cached"]
    assert again[_hiplib.MODE_RECT_FAST]["rows"] == r["rows"]
    b.run()
    oms = _oracle_maps(ts, cols, seg_off, seg_begin, windows)
    for t, m in enumerate(b.maps()):
        compare_maps(m.F_mn, oms[t])
    b.close()
