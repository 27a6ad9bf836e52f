"""Grid-search output (pyfstat_amd.gridio) on the CPU: column order and
formats as TransientGridSearch writes them (grid_based_searches.py:1023-1042,
1232-1253, 471-501; core.py:123-126), the F_mn file-name convention
(:1201-1230), and text round trips."""

import numpy as np
import pytest

from pyfstat_amd import gridio
from pyfstat_amd.atoms import TRANSIENT_RECTANGULAR, transientWindowRange_t


def test_output_key_order():
    # plain transient search: detstat maxTwoF, no BtSG
    assert gridio.output_keys() == ["F0", "F1", "F2", "Alpha", "Delta", "twoF", "maxTwoF",
                                    "t0_ML", "tau_ML"]
    # everything switched on, in the order _set_output_keys appends them
    keys = gridio.output_keys(detector_names=["H1", "L1"], singleFstats=True,
                              twoFXatMaxTwoF=True, detstat="lnBtSG", BtSG=True)
    assert keys == ["F0", "F1", "F2", "Alpha", "Delta", "twoF", "twoFH1", "twoFL1", "maxTwoF",
                    "twoFH1atMaxTwoF", "twoFL1atMaxTwoF", "lnBtSG", "t0_ML", "tau_ML",
                    "t0_MP", "tau_MP"]
    # a non-transient grid keeps only twoF (+ the detstat)
    assert gridio.output_keys(transient=False, detstat="log10BSGL") == [
        "F0", "F1", "F2", "Alpha", "Delta", "twoF", "log10BSGL"]


def test_formats():
    keys = gridio.output_keys(detector_names=["H1"], singleFstats=True, detstat="lnBtSG",
                              BtSG=True)
    fm = dict(zip(keys, gridio.savetxt_formats(keys)))
    assert fm["F0"] == fm["Delta"] == "%.16g"
    assert fm["twoF"] == fm["twoFH1"] == fm["maxTwoF"] == fm["lnBtSG"] == "%.9g"
    assert fm["t0_ML"] == fm["tau_ML"] == fm["t0_MP"] == fm["tau_MP"] == "%d"
    with pytest.raises(KeyError):
        gridio.savetxt_formats(["F0", "unknown"])


def _table():
    T = 4
    det = {"maxTwoF": np.array([10.25, 20.5, 30.125, 7.0]),
           "m_ML": np.array([1, 2, 0, 9]), "n_ML": np.array([0, 3, 1, 4]),
           "lnBtSG": np.array([1.5, 2.5, 3.5, -0.25]),
           "t0_MP": np.array([1e9 + 900.0, 1e9 + 2700, 1e9, 1e9 + 4500]),
           "tau_MP": np.array([3600.0, 5400, 7200, 9000])}
    w = transientWindowRange_t(TRANSIENT_RECTANGULAR, 1000000000, 1800 * 10, 1800, 3600,
                               1800 * 5, 1800)
    params = {"F0": 30.0 + 1e-7 * np.arange(T), "F1": np.full(T, -1e-10),
              "F2": np.zeros(T), "Alpha": np.full(T, 1.2345678901234), "Delta": np.full(T, -0.5)}
    return gridio.results_table(params, det, w, twoF=np.array([1.0, 2, 3, 4]),
                                detstat="lnBtSG"), det, w


def test_results_table_values():
    d, det, w = _table()
    assert d.dtype.names == ("F0", "F1", "F2", "Alpha", "Delta", "twoF", "maxTwoF", "lnBtSG",
                             "t0_ML", "tau_ML", "t0_MP", "tau_MP")
    np.testing.assert_array_equal(d["t0_ML"], w.t0 + det["m_ML"] * w.dt0)
    np.testing.assert_array_equal(d["tau_ML"], w.tau + det["n_ML"] * w.dtau)
    np.testing.assert_array_equal(d["maxTwoF"], det["maxTwoF"])


def test_write_read_round_trip(tmp_path):
    d, _, _ = _table()
    path = tmp_path / "grid.txt"
    gridio.write_results(str(path), d, header_lines=["search: hip", "version: test"])
    text = path.read_text().splitlines()
    assert text[0] == "# search: hip" and text[1] == "# version: test"
    assert text[2] == "# " + " ".join(d.dtype.names)
    first = text[3].split(" ")
    assert first[0] == "%.16g" % d["F0"][0] and first[8] == "%d" % d["t0_ML"][0]
    back = gridio.read_results(str(path))
    assert back.dtype.names == d.dtype.names
    for k in d.dtype.names:
        np.testing.assert_allclose(back[k], d[k], rtol=1e-8, err_msg=k)


def test_nan_written_as_nan(tmp_path):
    # np.nan_to_num leaves a structured array unchanged, as in the reference
    d, _, _ = _table()
    d["twoF"][1] = np.nan
    path = tmp_path / "grid.txt"
    gridio.write_results(str(path), d)
    assert np.isnan(gridio.read_results(str(path))["twoF"][1])


def test_fstat_map_filename():
    p = {"F0": 30.000000123, "F1": -1e-10, "F2": 0.0, "Alpha": 1.5, "Delta": -0.25}
    assert gridio.transient_fstat_map_filename("out/tCW_", p) == \
        "out/tCW_30.000000123_1.5_-0.25_-1e-10_0.dat"
    row = [p[k] for k in gridio.DEFAULT_SEARCH_KEYS]
    assert gridio.transient_fstat_map_filename("out/tCW_", row) == \
        gridio.transient_fstat_map_filename("out/tCW_", p)
    d, _, _ = _table()
    assert gridio.transient_fstat_map_filename("x", d[0]).startswith("x30_1.2345678901234_")
    with pytest.raises(ValueError):
        gridio.transient_fstat_map_filename("x", 3.0)
