"""Semi-coherent 2F, per-detector 2F at the maximum, and cumulative 2F on the
GPU against per-map oracle results (F relative 1e-4, as for the maps)."""

import numpy as np
import pytest

import oracle
from pyfstat_amd import detstats, synth
from pyfstat_amd.atoms import (
    TRANSIENT_RECTANGULAR,
    atoms_to_arrays,
    extract_singleIFOmultiFatoms_from_multiAtoms,
    transientWindowRange_t,
)

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def _oracle(multi, window):
    return oracle.compute_transient_fstat_map(atoms_to_arrays(multi), window, BtSG=False)


def _templates(n=3, n_sft=300):
    return [synth.workload_exp_plain(n_sft=n_sft, n_t0=10, n_tau=10, seed_index=80 + k)[0]
            for k in range(n)]


def test_semicoherent_twoF_vs_oracle(hip_ready):
    multis = _templates()
    tmin = synth.T_START
    tmax = synth.T_START + 300 * synth.TSFT
    w, tb, Tcoh = detstats.semicoherent_window(tmin, tmax, 12)
    out = detstats.semicoherent_twoF(multis, w, singleFstats=True)
    assert out["twoF_per_segment"].shape == (3, 12)
    for t, multi in enumerate(multis):
        print(f"semicoherent template {t}", flush=True)
        om = _oracle(multi, w)
        np.testing.assert_allclose(out["twoF_per_segment"][t], 2 * om.F_mn[:, 0], rtol=RTOL)
        assert out["twoF"][t] == pytest.approx(2 * om.F_mn[:, 0].sum(), rel=RTOL)
        for X in range(multi.length):
            omX = _oracle(extract_singleIFOmultiFatoms_from_multiAtoms(multi, X), w)
            np.testing.assert_allclose(out["twoFX_per_segment"][t, X], 2 * omX.F_mn[:, 0],
                                       rtol=RTOL)


def test_twoFX_at_multi_detector_maximum(hip_ready):
    multis = _templates()
    w = synth.grid_window(TRANSIENT_RECTANGULAR, synth.T_START, 40, 30)
    out = detstats.transient_detstats(multis, w, BtSG=True, singleFstats=True)
    for t, multi in enumerate(multis):
        om = oracle.compute_transient_fstat_map(atoms_to_arrays(multi), w, BtSG=True)
        assert out["maxTwoF"][t] == pytest.approx(2 * om.maxF, rel=RTOL)
        assert out["lnBtSG"][t] == pytest.approx(om.lnBtSG, abs=1e-4 + 2e-7 * om.maxF)
        m = (out["t0_ML"][t] - w.t0) // w.dt0
        n = (out["tau_ML"][t] - w.tau) // w.dtau
        for X in range(multi.length):
            omX = _oracle(extract_singleIFOmultiFatoms_from_multiAtoms(multi, X), w)
            assert out["twoFXatMaxTwoF"][t, X] == pytest.approx(2 * omX.F_mn[m, n], rel=RTOL)


def test_twoF_cumulative_vs_one_cell_maps(hip_ready):
    multi = _templates(1, n_sft=400)[0]
    p = atoms_to_arrays(multi)
    tstart, durations, twoFs = detstats.twoF_cumulative(multi, num_segments=60)
    assert len(twoFs) == 60 and tstart == p.t_min
    for d, f in zip(durations, twoFs):
        w = transientWindowRange_t(TRANSIENT_RECTANGULAR, int(tstart), 0, synth.TSFT, int(d), 0,
                                   synth.TSFT)
        om = _oracle(multi, w)
        assert f == pytest.approx(2 * om.F_mn[0, 0], rel=RTOL), d


def test_grid_results_table_vs_single_maps(hip_ready, tmp_path):
    """The grid-search table from one batched pass agrees with per-template
    single-map calls (t0_ML/tau_ML from the caller's window and the argmax,
    as grid_based_searches.py:1131-1135 does), and survives the text file."""
    from pyfstat_amd import gridio
    from pyfstat_amd import tcw_fstat_map_funcs as tcw

    multis = _templates()
    w = synth.grid_window(TRANSIENT_RECTANGULAR, synth.T_START, 40, 30)
    out = detstats.transient_detstats(multis, w, BtSG=True, singleFstats=True)
    params = {"F0": 30.0 + 0.1 * np.arange(3), "F1": np.zeros(3), "F2": np.zeros(3),
              "Alpha": np.full(3, 1.0), "Delta": np.full(3, 0.5)}
    table = gridio.results_table(params, out, w, detstat="lnBtSG", detector_names=["H1", "L1"])
    assert "twoFH1atMaxTwoF" in table.dtype.names and "t0_MP" in table.dtype.names
    for t, multi in enumerate(multis):
        fm = tcw.hip_compute_transient_fstat_map(multi, w, BtSG=True)
        m, n = fm.get_maxF_idx()
        assert table["t0_ML"][t] == w.t0 + int(m) * w.dt0
        assert table["tau_ML"][t] == w.tau + int(n) * w.dtau
        assert table["maxTwoF"][t] == pytest.approx(2 * fm.maxF, rel=1e-6)
        assert table["lnBtSG"][t] == pytest.approx(fm.lnBtSG, abs=1e-9)
        assert table["t0_MP"][t] == fm.t0_MP and table["tau_MP"][t] == fm.tau_MP
    path = tmp_path / "grid.txt"
    gridio.write_results(str(path), table, header_lines=["hip"])
    back = gridio.read_results(str(path))
    np.testing.assert_array_equal(back["t0_ML"], table["t0_ML"])


def test_maps_npz_round_trip(hip_ready, tmp_path):
    from pyfstat_amd import gridio
    from pyfstat_amd.batched import BatchedTransientMaps

    batch = synth.workload_rect_batched(T=5, n_min=200, n_max=600, seed_index=14)
    b = BatchedTransientMaps.from_synthetic(batch, BtSG=True)
    b.run()
    path = tmp_path / "maps.npz"
    gridio.write_maps_npz(str(path), b)
    maps, wins, st = gridio.read_maps_npz(str(path))
    ref = b.maps()
    for t in range(batch.T):
        np.testing.assert_array_equal(maps[t], ref[t].F_mn)
        assert (wins[t].t0, wins[t].dt0, wins[t].tauBand) == (
            batch.windows[t].t0, batch.windows[t].dt0, batch.windows[t].tauBand)
        assert st["maxF"][t] == ref[t].maxF
    b.close()
