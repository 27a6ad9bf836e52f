"""The exponential-window recurrence with long row chunks: a work item walks
up to 1024 t0 rows, its atoms staged in LDS in passes (or read from the atom
table when they do not fit), and its BtSG row sums written pass by pass and
rescaled when a later pass raises the item's reference.  Each case has a
strong transient starting in the map's lowest rows, which the walk (from the
largest t0 down) reaches last, so the reference rises after the first
passes.  Maps, argmax, lnBtSG and the posterior modes against the FP64
oracle; tolerances as in tests/test_gpu_parity.py."""

import numpy as np
import pytest

from pyfstat_amd import _hiplib, autotune, synth
from pyfstat_amd.atoms import TRANSIENT_EXPONENTIAL, transientWindowRange_t
from tests.helpers import check_argmax, compare_maps, oracle_for
from tests.test_gpu_parity import _posterior_ok, lnb_tol

pytestmark = pytest.mark.gpu

TS = synth.TSFT


def _templates(n_sft, n_t0, seed, t_on_row=30, tau_sig=40 * TS, amp=4.0, dt0=TS):
    """Two templates on one grid (2 detectors, 10% gaps): a strong exp
    transient starting at t0 row t_on_row, and noise only."""
    rng = synth.rng_for(seed)
    keep = synth.gap_mask(rng, n_sft, 0.10)
    t = synth.T_START + np.nonzero(keep)[0].astype(np.float64) * TS
    out = []
    for k in range(2):
        sig = (synth.exp_signal(t, synth.T_START + t_on_row * dt0, tau_sig, 0.7, amp=amp)
               if k == 0 else np.zeros(t.shape[0], dtype=np.complex128))
        dets = []
        for _X in range(2):
            pa, pb = rng.uniform(0, 2 * np.pi, 2)
            dets.append((t.astype(np.uint32), synth.detector_atoms(rng, t, pa, pb, sig)))
        out.append(synth.multi_from_detectors(dets))
    return out


def _window(n_t0, n_tau, dt0=TS, dtau=TS):
    return transientWindowRange_t(TRANSIENT_EXPONENTIAL, synth.T_START, (n_t0 - 1) * dt0, dt0,
                                  2 * TS, (n_tau - 1) * dtau, dtau)


def _check(multis, win, rows):
    from pyfstat_amd.batched import BatchedTransientMaps

    b = BatchedTransientMaps.from_multi_atoms(multis, [win] * len(multis), BtSG=True)
    autotune.set_rows(b, _hiplib.MODE_EXP_FAST, rows)
    assert b.items_per_mode()[_hiplib.MODE_EXP_FAST] > 0
    b.run()
    maps = b.maps()
    for t, multi in enumerate(multis):
        print(f"case: {win} rows {rows} template {t}", flush=True)
        om = oracle_for(multi, win, BtSG=True)
        compare_maps(maps[t].F_mn, om)
        m, n = maps[t].get_maxF_idx()
        check_argmax(maps[t].F_mn, int(m), int(n), om)
        assert abs(maps[t].lnBtSG - om.lnBtSG) <= lnb_tol(om.maxF), (maps[t].lnBtSG, om.lnBtSG)
        _posterior_ok(om, win, maps[t].t0_MP, maps[t].tau_MP)
        if t == 0:
            assert om.maxF > 70.0  # far above the noise: the reference rises mid-walk
    b.close()


@pytest.mark.parametrize("rows", [1024, 256, 64])
def test_long_chunks_lds_passes(hip_ready, rows):
    """dtau = TAtom, 120 tau columns: LDS passes of 128 rows; 600 t0 rows as
    one chunk of 608 rows (5 passes), chunks of 256, and chunks of 64 (the
    recurrence from zero and the chunk scan between them)."""
    multis = _templates(700, 600, seed=120)
    _check(multis, _window(600, 120), rows)


def test_narrow_passes(hip_ready):
    """dtau = 2 TAtom and 140 tau columns: the leaving atoms of a row span
    6 * 139 atoms, which leaves LDS room for passes of 60 rows."""
    multis = _templates(900, 400, seed=121)
    _check(multis, _window(400, 140, dtau=2 * TS), 1024)


def test_table_and_lds_blocks(hip_ready):
    """dtau = 2 TAtom and 300 tau columns: the first column block's atoms do
    not fit in LDS (steps read the atom table), the second block's do."""
    multis = _templates(1500, 300, seed=122)
    _check(multis, _window(300, 300, dtau=2 * TS), 1024)


def test_two_atom_row_steps(hip_ready):
    """dt0 = 2 TAtom: two atom steps per row, read from the atom table."""
    multis = _templates(900, 300, seed=123, dt0=2 * TS)
    _check(multis, _window(300, 100, dt0=2 * TS), 1024)
