"""Exponential-window refinement on the CPU: the FP64 reference
(oracle/refine_ref.py) pinned against the C oracle's map, finite differences
and the t0 invariance; the product's support rule against a brute-force loop."""

import numpy as np
import pytest

import oracle
from oracle import refine_ref as rr
from pyfstat_amd import synth
from pyfstat_amd.atoms import atoms_to_arrays
from pyfstat_amd.refine import exp_window_support

TA = synth.TSFT


@pytest.fixture(scope="module")
def small_exp():
    multi, _ = synth.workload_exp_plain(n_sft=240, n_t0=40, n_tau=30, tau_inj=3.0e4)
    p = atoms_to_arrays(multi)
    merged, t0d = oracle.merge_binned(p.ts, p.cols, p.det_off, p.TAtom)
    win = synth.grid_window(oracle.TRANSIENT_EXPONENTIAL, t0d, 40, 30)
    om = oracle.compute_map(merged, p.TAtom, t0d, win)
    return merged, t0d, win, om


def test_forward_equals_map_cells(small_exp):
    """At a grid cell with the map's support, F is the C oracle's map value."""
    merged, t0d, win, om = small_exp
    N = merged.shape[1]
    rng = np.random.default_rng(0)
    for _ in range(60):
        m, n = int(rng.integers(0, om.F_mn.shape[0])), int(rng.integers(0, om.F_mn.shape[1]))
        t0 = win.t0 + m * win.dt0
        tau = win.tau + n * win.dtau
        sup = rr.support_bruteforce(t0, tau, t0d, N, TA)
        F = rr.evaluate(merged, float(t0d), TA, sup[0], sup[1], float(t0), float(tau))[0]
        assert F == pytest.approx(om.F_mn[m, n], rel=1e-12, abs=1e-12), (m, n)


def test_analytic_derivatives_match_finite_differences(small_exp):
    merged, t0d, win, om = small_exp
    N = merged.shape[1]
    m, n = om.m_max, om.n_max
    t0 = float(win.t0 + m * win.dt0)
    tau = float(win.tau + n * win.dtau)
    kl, kh = rr.support_bruteforce(int(t0), int(tau), t0d, N, TA)
    for tau_x in (tau, tau * 1.137, tau * 0.71):
        F, dt0, dtau, g = rr.evaluate(merged, float(t0d), TA, kl, kh, t0, tau_x)
        h = 1e-4 * tau_x
        Fp = rr.evaluate(merged, float(t0d), TA, kl, kh, t0, tau_x + h)[0]
        Fm = rr.evaluate(merged, float(t0d), TA, kl, kh, t0, tau_x - h)[0]
        fd = (Fp - Fm) / (2 * h)
        assert dtau == pytest.approx(fd, rel=1e-6, abs=1e-9 * abs(F) / tau_x)
        # atoms: central differences in FP64 on a few entries of every column
        base = merged.astype(np.float64)
        grad = rr.atom_grad(merged.shape, float(t0d), TA, kl, kh, t0, tau_x, g)
        for c in range(7):
            for k in (kl, (kl + kh) // 2, kh):
                e = 1e-5 * max(1.0, abs(base[c, k]))
                xp = base.copy()
                xp[c, k] += e
                xm = base.copy()
                xm[c, k] -= e
                Fp = rr.evaluate(xp, float(t0d), TA, kl, kh, t0, tau_x)[0]
                Fm = rr.evaluate(xm, float(t0d), TA, kl, kh, t0, tau_x)[0]
                assert grad[c, k] == pytest.approx((Fp - Fm) / (2 * e), rel=1e-5, abs=1e-9)
        assert np.all(grad[:, :kl] == 0) and np.all(grad[:, kh + 1:] == 0)


def test_t0_invariance_on_fixed_support(small_exp):
    """All weights share exp(t0/tau); F is homogeneous of degree 0 in
    (A, B, C; Fa, Fb) with weights (2; 1), so F ignores t0 on a fixed support."""
    merged, t0d, win, om = small_exp
    N = merged.shape[1]
    t0 = float(win.t0 + 7 * win.dt0)
    tau = float(win.tau + 11 * win.dtau)
    kl, kh = rr.support_bruteforce(int(t0), int(tau), t0d, N, TA)
    F, dt0, dtau, g = rr.evaluate(merged, float(t0d), TA, kl, kh, t0, tau)
    for d in (-900.0, 123.4, 1799.0):
        F2 = rr.evaluate(merged, float(t0d), TA, kl, kh, t0 + d, tau)[0]
        assert F2 == pytest.approx(F, rel=1e-12)
    assert abs(dt0) <= 1e-12 * (abs(F) + 1) / tau * 1e3
    assert abs(dtau) > 0  # tau does matter


def test_planted_gradient_error_is_caught(small_exp):
    """A dropped factor 2 on the A/B/C weights in dtau must fail the FD check."""
    merged, t0d, win, om = small_exp
    N = merged.shape[1]
    t0, tau = float(win.t0), float(win.tau + 9 * win.dtau)
    kl, kh = rr.support_bruteforce(int(t0), int(tau), t0d, N, TA)
    S, U = rr.window_sums(merged, float(t0d), TA, kl, kh, t0, tau)
    F, _, dtau, g = rr.evaluate(merged, float(t0d), TA, kl, kh, t0, tau)
    wrong = float((U * g).sum() / tau ** 2)  # p_c dropped
    h = 1e-4 * tau
    fd = (rr.evaluate(merged, float(t0d), TA, kl, kh, t0, tau + h)[0]
          - rr.evaluate(merged, float(t0d), TA, kl, kh, t0, tau - h)[0]) / (2 * h)
    assert dtau == pytest.approx(fd, rel=1e-6)
    assert wrong != pytest.approx(fd, rel=1e-3)


def test_support_rule_matches_bruteforce():
    rng = np.random.default_rng(1)
    t_data = 1_000_000_000
    N = 300
    for _ in range(400):
        t0 = t_data + int(rng.integers(0, 320 * TA))
        tau = int(rng.integers(1, 120 * TA))
        ref = rr.support_bruteforce(t0, tau, t_data, N, TA)
        if ref is None:
            with pytest.raises(ValueError):
                exp_window_support([t0], [tau], [t_data], [N], TA)
            continue
        kl, kh = exp_window_support([t0], [tau], [t_data], [N], TA)
        assert (int(kl[0]), int(kh[0])) == ref, (t0 - t_data, tau)
    with pytest.raises(ValueError):
        exp_window_support([t_data - 1], [3600], [t_data], [N], TA)


def test_scan_max_brackets_the_grid_maximum(small_exp):
    """scan_max_tau over [tau-dtau, tau+dtau] is at least the grid value."""
    merged, t0d, win, om = small_exp
    N = merged.shape[1]
    m, n = om.m_max, om.n_max
    t0 = int(win.t0 + m * win.dt0)
    tau = int(win.tau + n * win.dtau)
    kl, kh = rr.support_bruteforce(t0, tau, t0d, N, TA)
    Fmax, tmax = rr.scan_max_tau(merged, float(t0d), TA, kl, kh, float(t0),
                                 float(max(tau - win.dtau, TA)), float(tau + win.dtau), n=201)
    assert Fmax >= om.F_mn[m, n] - 1e-9


def test_c_refine_oracle_equals_numpy_reference(small_exp):
    """oracle.refine_eval_batch (C, OpenMP; the refinement's CPU baseline) is
    the numpy reference evaluate(), candidate by candidate, on two templates
    (the second one's atoms scaled, so a template mix-up shows)."""
    merged, t0d, win, om = small_exp
    N = merged.shape[1]
    cols = np.concatenate([merged, 3.0 * merged], axis=1).astype(np.float32)
    m_off = np.array([0, N, 2 * N])
    t_data = np.array([t0d, t0d + 5 * TA], dtype=np.uint32)
    rng = np.random.default_rng(7)
    ct, lo, hi, t0s, taus = [], [], [], [], []
    for i in range(80):
        t = i % 2
        m, n = int(rng.integers(0, om.F_mn.shape[0])), int(rng.integers(0, om.F_mn.shape[1]))
        t0 = int(t_data[t]) + m * win.dt0
        tau = win.tau + n * win.dtau
        sup = rr.support_bruteforce(t0, tau, int(t_data[t]), N, TA)
        if sup is None:
            continue
        ct.append(t)
        lo.append(sup[0])
        hi.append(sup[1])
        t0s.append(t0 + rng.uniform(-0.3, 0.3) * TA)
        taus.append(tau * rng.uniform(0.8, 1.25))
    F, dt0, dtau, g = oracle.refine_eval_batch(cols, m_off, t_data, TA, ct, lo, hi, t0s, taus)
    Fr, dt0r, dtaur, gr = rr.batch_evaluate(cols.astype(np.float64), m_off, t_data, TA, ct, lo,
                                            hi, np.array(t0s), np.array(taus))
    np.testing.assert_allclose(F, Fr, rtol=1e-12)
    np.testing.assert_allclose(dtau, dtaur, rtol=1e-9, atol=1e-12 * np.abs(Fr).max() / np.min(taus))
    np.testing.assert_allclose(g, gr, rtol=1e-9, atol=1e-12 * np.abs(gr).max())
    np.testing.assert_allclose(dt0, dt0r, atol=1e-9 * np.abs(Fr).max() / np.min(taus))
