"""Window set-up from ComputeFstat options (pyfstat/core.py:821-899 and
1440-1449), branch by branch.  CPU only."""

import logging

import numpy as np
import pytest

import oracle
from pyfstat_amd import atoms as at
from pyfstat_amd import synth
from pyfstat_amd.windows import set_window_times, transient_window_range

TSFT = 1800
T0 = 1000000000


def fields(w):
    return at.window_fields(w)


def test_type_strings_and_unknown_type():
    for s, t in (("rect", at.TRANSIENT_RECTANGULAR), ("exp", at.TRANSIENT_EXPONENTIAL)):
        assert transient_window_range(s, TSFT).type == t
    with pytest.raises(ValueError, match=r"Unknown window-type \(box\) passed as input, "
                                         r"\[none, rect, exp\] allows\."):
        transient_window_range("box", TSFT)


def test_none_spans_all_data():
    # core.py:847-853: t0 = minStartTime, tau = maxStartTime - minStartTime,
    # both bands 0, spacings Tsft; band options are ignored
    w = transient_window_range("none", TSFT, minStartTime=T0, maxStartTime=T0 + 86400,
                               t0Band=7200, tauBand=7200, tauMin=5000, dt0=900, dtau=900)
    assert fields(w) == (at.TRANSIENT_NONE, T0, 0, TSFT, 86400, 0, TSFT)
    with pytest.raises(ValueError):
        transient_window_range("none", TSFT)


def test_defaults_without_bands():
    # no t0Band/tauBand: bands 0, spacings Tsft, tau not set (0) until call time
    w = transient_window_range("rect", TSFT, minStartTime=T0, maxStartTime=T0 + 86400,
                               tauMin=5000, dt0=900, dtau=900)
    assert fields(w) == (at.TRANSIENT_RECTANGULAR, 0, 0, TSFT, 0, 0, TSFT)


def test_reference_test_configuration():
    # tests/test_core.py:528-531: rect, t0Band = tauBand = tauMin = 2 Tsft
    w = transient_window_range("rect", TSFT, minStartTime=T0, maxStartTime=T0 + 86400,
                               t0Band=2 * TSFT, tauBand=2 * TSFT, tauMin=2 * TSFT)
    assert fields(w) == (at.TRANSIENT_RECTANGULAR, 0, 2 * TSFT, TSFT, 2 * TSFT, 2 * TSFT, TSFT)
    set_window_times(w, minStartTime=T0, maxStartTime=T0 + 86400)
    # tauBand != 0: tau keeps tauMin
    assert fields(w) == (at.TRANSIENT_RECTANGULAR, T0, 2 * TSFT, TSFT, 2 * TSFT, 2 * TSFT, TSFT)


def test_spacing_only_with_its_band():
    # dt0 is read only when t0Band is given (core.py:866-867), dtau and tauMin
    # only when tauBand is given (core.py:877-892)
    w = transient_window_range("exp", TSFT, t0Band=3600, dt0=900, dtau=600, tauMin=7200)
    assert (w.t0Band, w.dt0, w.tauBand, w.dtau, w.tau) == (3600, 900, 0, TSFT, 0)
    w = transient_window_range("exp", TSFT, tauBand=3600, dt0=900, dtau=600, tauMin=7200)
    assert (w.t0Band, w.dt0, w.tauBand, w.dtau, w.tau) == (0, TSFT, 3600, 600, 7200)
    # a zero spacing keeps the default (`if self.dt0:`)
    w = transient_window_range("rect", TSFT, t0Band=3600, tauBand=3600, dt0=0, dtau=0)
    assert (w.dt0, w.dtau) == (TSFT, TSFT)


def test_tau_min_default_is_two_sfts():
    w = transient_window_range("rect", TSFT, tauBand=36000)
    assert w.tau == 2 * TSFT


def test_non_integer_bands_truncate_with_warning(caplog):
    with caplog.at_level(logging.WARNING, logger="pyfstat_amd.windows"):
        w = transient_window_range("rect", TSFT, t0Band=3600.9, tauBand=np.int64(7200),
                                   tauMin=3600.5)
    assert (w.t0Band, w.tauBand, w.tau) == (3600, 7200, 3600)
    msgs = " ".join(r.getMessage() for r in caplog.records)
    for name in ("t0Band=3600.9", "tauBand=7200", "tauMin=3600.5"):
        assert f"Casting non-integer {name} to int" in msgs


def test_set_window_times_rules():
    w = transient_window_range("rect", TSFT, t0Band=3600)
    # tauBand == 0: tau = tend - tstart
    set_window_times(w, tstart=T0 + 1800, tend=T0 + 90000)
    assert (w.t0, w.tau) == (T0 + 1800, 88200)
    # defaults from min/maxStartTime; a falsy tstart falls back as well
    set_window_times(w, tstart=0, minStartTime=T0, maxStartTime=T0 + 7200)
    assert (w.t0, w.tau) == (T0, 7200)
    with pytest.raises(ValueError, match="Need tstart"):
        set_window_times(w, tstart=None, tend=T0)
    # tauBand != 0: tau untouched
    w2 = transient_window_range("rect", TSFT, tauBand=3600, tauMin=5400)
    set_window_times(w2, tstart=T0, tend=T0 + 86400)
    assert (w2.t0, w2.tau) == (T0, 5400)


def test_helper_window_drives_the_map():
    """A search set up through the helper gives the map shape and axes the
    reference's ComputeFstat would (N_t0 = t0Band/dt0 + 1, tcw.py:775-780)."""
    n = 96
    multi, _ = synth.workload_rect_plain(n_sft=n, n_t0=4, n_tau=4)
    arrs = at.atoms_to_arrays(multi)
    tmin, tmax = int(arrs.ts.min()), int(arrs.ts.max())
    w = transient_window_range("rect", TSFT, minStartTime=tmin, maxStartTime=tmax + TSFT,
                               t0Band=(n - 2) * TSFT - 2 * TSFT, tauBand=10 * TSFT)
    set_window_times(w, minStartTime=tmin, maxStartTime=tmax + TSFT)
    om = oracle.compute_transient_fstat_map(arrs, w, BtSG=False)
    assert om.F_mn.shape == ((n - 4) + 1, 11)
    assert om.t0_ML == tmin + om.m_max * TSFT
    assert om.tau_ML == 2 * TSFT + om.n_max * TSFT
