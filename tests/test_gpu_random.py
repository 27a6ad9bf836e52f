"""Seeded random sweep against the oracle: templates with 1-3 detectors,
gaps, several atoms in some bins and off-grid timestamps; rect and exp
windows on and off the atom grid, starting before, inside and after the
data.  Every case runs through the single-map path (and the generic rect
kernel) and all of them together through one ragged batch.  Tolerances as
in test_gpu_parity (F relative 1e-4 outside the cond ~ 1e4 band, argmax or a
near-tie, lnBtSG 1e-4 + 2e-7 maxF, posterior modes or a near-tie)."""

import numpy as np
import pytest

from pyfstat_amd import synth
from pyfstat_amd.atoms import (
    TRANSIENT_EXPONENTIAL,
    TRANSIENT_RECTANGULAR,
    FstatAtomVector,
    MultiFstatAtomVector,
    transientWindowRange_t,
)
from tests.helpers import compare_maps, oracle_for
from tests.test_gpu_parity import _check_full, _posterior_ok, lnb_tol

pytestmark = pytest.mark.gpu

N_CASES = 40
TA = synth.TSFT


def _case(seed):
    rng = np.random.default_rng(1000 + seed)
    ndet = int(rng.integers(1, 4))
    n_bins = int(rng.integers(40, 400))
    multi = MultiFstatAtomVector(ndet)
    for X in range(ndet):
        slots = np.nonzero(synth.gap_mask(rng, n_bins, float(rng.uniform(0.0, 0.4))))[0]
        t = slots * TA
        if seed % 3 == 0:  # a second atom in some bins
            dup = rng.choice(slots, size=max(1, len(slots) // 5), replace=False)
            t = np.sort(np.concatenate([t, dup * TA + int(rng.integers(1, TA))]))
        if seed % 4 == 1:  # timestamps off the grid (inside their bins)
            t = t + int(rng.integers(0, TA // 2))
        t = (t + synth.T_START + 7 * X).astype(np.int64)
        cols = synth.detector_atoms(rng, t.astype(np.float64), float(rng.uniform(0, 6)),
                                    float(rng.uniform(0, 6)), np.zeros(t.size, np.complex128))
        multi.data[X] = FstatAtomVector.from_arrays(TA, t, cols[0], cols[1], cols[2],
                                                    cols[3] + 1j * cols[4],
                                                    cols[5] + 1j * cols[6])
    wtype = TRANSIENT_RECTANGULAR if seed % 2 == 0 else TRANSIENT_EXPONENTIAL
    n_t0 = int(rng.integers(1, 50))
    n_tau = int(rng.integers(1, 50))
    if rng.random() < 0.6:  # on the atom grid
        dt0 = TA * int(rng.integers(1, 4))
        dtau = dt0 if rng.random() < 0.7 else TA * int(rng.integers(1, 4))
        t0 = synth.T_START + TA * int(rng.integers(-10, n_bins // 2))
        tau = TA * int(rng.integers(2, 20))
    else:  # anywhere
        dt0 = int(rng.integers(300, 4000))
        dtau = int(rng.integers(300, 4000))
        t0 = synth.T_START + int(rng.integers(-20000, n_bins * TA // 2))
        tau = int(rng.integers(TA, 40 * TA))
    win = transientWindowRange_t(wtype, int(t0), (n_t0 - 1) * dt0, dt0, tau, (n_tau - 1) * dtau,
                                 dtau)
    return multi, win


@pytest.mark.parametrize("seed", range(N_CASES))
def test_random_single(hip_ready, seed):
    multi, win = _case(seed)
    print(f"random case {seed}: window {win}", flush=True)
    _check_full(multi, win)
    if win.type == TRANSIENT_RECTANGULAR:
        _check_full(multi, win, force_generic=True)


def test_random_batch(hip_ready):
    from pyfstat_amd.batched import hip_compute_transient_fstat_maps

    cases = [_case(s) for s in range(N_CASES)]
    maps = hip_compute_transient_fstat_maps([c[0] for c in cases], [c[1] for c in cases],
                                            BtSG=True)
    for s, ((multi, win), fm) in enumerate(zip(cases, maps)):
        print(f"random batch case {s}", flush=True)
        om = oracle_for(multi, win)
        compare_maps(fm.F_mn, om)
        assert abs(fm.lnBtSG - om.lnBtSG) <= lnb_tol(om.maxF)
        _posterior_ok(om, win, fm.t0_MP, fm.tau_MP)


@pytest.mark.parametrize("n_atoms", [1, 2, 3])
@pytest.mark.parametrize("wtype", [TRANSIENT_RECTANGULAR, TRANSIENT_EXPONENTIAL])
def test_smallest_templates(hip_ready, n_atoms, wtype):
    """Templates of 1-3 atoms (one merged bin per atom): most windows hold a
    single atom (D = 0 exactly, F = 2 fallback) or none."""
    rng = np.random.default_rng(7 + n_atoms)
    t = (synth.T_START + TA * np.arange(n_atoms)).astype(np.int64)
    cols = synth.detector_atoms(rng, t.astype(np.float64), 0.4, 1.7,
                                np.zeros(n_atoms, np.complex128))
    multi = synth.multi_from_detectors([(t, cols)], Tsft=TA)
    win = transientWindowRange_t(wtype, synth.T_START - TA, 4 * TA, TA, TA, 4 * TA, TA)
    _check_full(multi, win)
    if wtype == TRANSIENT_RECTANGULAR:
        _check_full(multi, win, force_generic=True)
