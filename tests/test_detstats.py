"""Host-side set-up of the detection statistics built on batched maps
(CPU only): semi-coherent window, cumulative durations, template packing."""

import numpy as np
import pytest

from pyfstat_amd import detstats, synth
from pyfstat_amd.atoms import atoms_to_arrays


def test_semicoherent_window_matches_segment_boundaries():
    w, tb, Tcoh = detstats.semicoherent_window(1_000_000_000, 1_000_000_000 + 10 * 86400, 10)
    assert Tcoh == 86400.0
    assert (w.type, w.t0, w.dt0, w.tau, w.tauBand, w.dtau) == (1, 1_000_000_000, 86400, 86400,
                                                               0, 1)
    assert w.t0Band // w.dt0 + 1 == 10  # one map row per segment
    np.testing.assert_allclose(tb, 1_000_000_000 + 86400 * np.arange(11))
    with pytest.raises(RuntimeError):
        detstats.semicoherent_window(0, 3 * 1800, 2)


def test_cumulative_durations_convention():
    t0, t1, d = detstats.cumulative_durations(None, None, 5, 1000, 1000 + 100 * 1800)
    assert (t0, t1) == (1000, 1000 + 100 * 1800)
    np.testing.assert_allclose(d, np.linspace(3600, 100 * 1800, 5))
    t0, t1, d = detstats.cumulative_durations(500, 10**9, 3, 1000, 5000)
    assert (t0, t1) == (1000, 5000)


def test_pack_appends_single_detector_templates():
    multis = [synth.workload_exp_plain(n_sft=50, seed_index=70 + k)[0] for k in range(3)]
    ts, cols, seg_off, seg_begin, TA, dets = detstats._pack(multis, True)
    assert list(dets) == [2, 2, 2]
    assert len(seg_begin) == 3 + 6 + 1
    p0 = atoms_to_arrays(multis[0])
    # template 3 is detector 0 of template 0, template 4 its detector 1
    s3 = seg_begin[3]
    np.testing.assert_array_equal(ts[seg_off[s3]:seg_off[s3 + 1]], p0.ts[:p0.det_off[1]])
    s4 = seg_begin[4]
    np.testing.assert_array_equal(cols[:, seg_off[s4]:seg_off[s4 + 1]],
                                  p0.cols[:, p0.det_off[1]:p0.det_off[2]])
