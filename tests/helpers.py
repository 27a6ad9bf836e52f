"""Comparison helpers shared by the parity tests (HIP path vs FP64 oracle)."""

import numpy as np

import oracle

# tolerance on F (and 2F): max relative error vs the FP64 oracle
F_RTOL = 1e-4
# cells whose condition number sits within this relative distance of the 1e4
# guard may legitimately flip between the formula and the F = 2 fallback
COND_BAND = 1e-3
COND_MAX = 1e4


def oracle_for(multi, window, BtSG=True):
    from pyfstat_amd.atoms import atoms_to_arrays

    return oracle.compute_transient_fstat_map(atoms_to_arrays(multi), window, BtSG=BtSG)


def compare_maps(F_hip, om, rtol=F_RTOL):
    """Return a dict of error figures; assert the F tolerance outside the
    condition-threshold band."""
    F_hip = np.asarray(F_hip, dtype=np.float64)
    F_orc = om.F_mn
    assert F_hip.shape == F_orc.shape, (F_hip.shape, F_orc.shape)
    near = np.abs(om.cond / COND_MAX - 1.0) < COND_BAND
    denom = np.maximum(np.abs(F_orc), 1e-30)
    rel = np.abs(F_hip - F_orc) / denom
    rel_ok = np.where(near, 0.0, rel)
    worst = float(rel_ok.max()) if rel_ok.size else 0.0
    assert np.all(np.isfinite(F_hip))
    assert worst <= rtol, f"max rel err {worst:.3e} at {np.unravel_index(np.argmax(rel_ok), rel_ok.shape)}"
    return {"max_rel_err": worst, "near_threshold_cells": int(near.sum())}


def check_argmax(F_hip, m, n, om, tie_rtol=1e-6):
    """(m, n) must be the oracle argmax, or a documented near-tie: the oracle F
    there within tie_rtol of the oracle maximum."""
    if (m, n) == (om.m_max, om.n_max):
        return True
    assert abs(om.F_mn[m, n] - om.maxF) <= tie_rtol * abs(om.maxF), (
        (m, n), (om.m_max, om.n_max), om.F_mn[m, n], om.maxF)
    return False
