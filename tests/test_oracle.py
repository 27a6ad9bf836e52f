"""Pins for the FP64 oracle (CPU only).

Each oracle function is checked against something that does not restate its
own formula:
  map (rect)      exact hand-computed fixtures (tests/golden/hand_cases.json,
                  exact rationals, windows read off the time axis); numpy
                  linear-algebra F = x^H M^-1 x on windows chosen by time
                  membership; full-span cell == all-data F; the plain-Python
                  twin on random tiny inputs incl. uint32 wrap-around cases
  map (exp)       numpy exp weights over the half-open time window + linalg;
                  the twin
  F normalisation E[F] = 2 in Gaussian noise (Monte Carlo), scale invariance,
                  permutation invariance, the cond = 1e4 guard at known cond
  merge           numpy per-bin float32 sums in detector order; two-detector
                  merge == pre-summed single detector
  lnBtSG          scipy.special.logsumexp
  posterior modes numpy row/column sums and argmax
"""

import json
import os

import numpy as np
import pytest
from scipy.special import logsumexp

import oracle
from oracle import twin

HERE = os.path.dirname(os.path.abspath(__file__))
TA = 1800
T0 = 1_000_000_000


def _random_atoms(rng, n, scale=1.0):
    ph = rng.uniform(0, 2 * np.pi, 2)
    t = np.arange(n)
    a = 0.3 + 0.5 * np.cos(2 * np.pi * t * TA / 86164 + ph[0])
    b = 0.4 * np.sin(2 * np.pi * t * TA / 86164 + ph[1])
    cols = np.empty((7, n), dtype=np.float32)
    cols[0], cols[1], cols[2] = a * a, b * b, a * b
    noise = rng.standard_normal((4, n)) * np.sqrt(0.5) * scale
    cols[3], cols[4] = a * noise[0], a * noise[1]
    cols[5], cols[6] = b * noise[2], b * noise[3]
    return cols


def _linalg_F(cols, idx, w=None):
    """F = x^H M^-1 x by numpy.linalg.solve over atoms idx with weights w."""
    c = cols[:, idx].astype(np.float64)
    if w is None:
        w = np.ones(len(idx))
    A, B, C = (c[0] * w * w).sum(), (c[1] * w * w).sum(), (c[2] * w * w).sum()
    x = np.array([[(c[3] * w).sum(), (c[4] * w).sum()], [(c[5] * w).sum(), (c[6] * w).sum()]])
    M = np.array([[A, C], [C, B]])
    ev = np.linalg.eigvalsh(M)
    if ev[0] <= 0 or ev[1] / ev[0] >= 1e4:
        return 2.0, ev[1] / ev[0] if ev[0] > 0 else np.inf
    y = np.linalg.solve(M, x)
    return float((x * y).sum()), ev[1] / ev[0]


def test_hand_cases_fixture():
    data = json.load(open(os.path.join(HERE, "golden", "hand_cases.json")))
    for case in data["cases"]:
        ts = np.array(case["timestamps"], dtype=np.uint32)
        cols = np.array(case["atoms"], dtype=np.float32).T.copy()
        merged, t0d = oracle.merge_binned(ts, cols, np.array([0, ts.size]), case["TAtom"])
        om = oracle.compute_map(merged, case["TAtom"], t0d, case["window"])
        np.testing.assert_allclose(om.F_mn, np.array(case["F"]), rtol=1e-12, err_msg=case["name"])
        Ft, _ = twin.transient_map(merged, case["TAtom"], t0d, case["window"])
        np.testing.assert_allclose(Ft, np.array(case["F"]), rtol=1e-12)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_rect_map_vs_linalg_on_time_windows(seed):
    rng = np.random.default_rng(seed)
    N = 60
    cols = _random_atoms(rng, N)
    win = (1, T0, 40 * TA, TA, 2 * TA, 30 * TA, TA)
    om = oracle.compute_map(cols, TA, T0, win)
    t = T0 + TA * np.arange(N)
    checked = 0
    for m in range(om.F_mn.shape[0]):
        for n in range(om.F_mn.shape[1]):
            a = T0 + m * TA
            b = a + 2 * TA + n * TA
            idx = np.nonzero((t >= a) & (t < b))[0]
            if idx.size < 2:
                continue
            F, cond = _linalg_F(cols, idx)
            if abs(cond / 1e4 - 1) < 1e-6:
                continue
            assert om.F_mn[m, n] == pytest.approx(F, rel=1e-9, abs=1e-12), (m, n)
            checked += 1
    assert checked > 800


@pytest.mark.parametrize("seed", [4, 5])
def test_exp_map_vs_explicit_weights(seed):
    rng = np.random.default_rng(seed)
    N = 80
    cols = _random_atoms(rng, N)
    win = (2, T0, 30 * TA, TA, TA, 15 * TA, TA)
    om = oracle.compute_map(cols, TA, T0, win)
    t = T0 + TA * np.arange(N)
    for m in range(om.F_mn.shape[0]):
        for n in range(om.F_mn.shape[1]):
            t0 = T0 + m * TA
            tau = TA + n * TA
            idx = np.nonzero((t >= t0) & (t < t0 + 3 * tau))[0]
            w = np.exp(-(t[idx] - t0) / tau)
            F, cond = _linalg_F(cols, idx, w)
            assert om.F_mn[m, n] == pytest.approx(F, rel=1e-9, abs=1e-12), (m, n)


def test_full_span_cell_and_none_window():
    """F_mn[0,-1] over the full span == all-data F (tests/test_tcw_fstat_map_funcs.py:124-127)."""
    rng = np.random.default_rng(9)
    N = 120
    cols = _random_atoms(rng, N)
    F_all, _ = _linalg_F(cols, np.arange(N))
    om = oracle.compute_map(cols, TA, T0, (1, T0, 0, TA, 2 * TA, (N - 2) * TA, TA))
    assert om.F_mn[0, -1] == pytest.approx(F_all, rel=1e-10)
    on = oracle.compute_map(cols, TA, T0, (0, 5, 5, 5, 5, 5, 5))
    assert on.F_mn.shape == (1, 1)
    assert on.F_mn[0, 0] == pytest.approx(F_all, rel=1e-10)
    assert on.window_eff == (1, T0, 0, TA, N * TA, 0, TA)


@pytest.mark.parametrize(
    "win",
    [
        (1, T0, 20 * TA, TA, 2 * TA, 25 * TA, TA),
        (1, T0 - 7 * TA, 20 * TA, TA, 2 * TA, 25 * TA, TA),  # t0 before the data: wrap
        (1, T0 + 33 * TA, 10 * TA, TA, 0, 6 * TA, TA),  # tau = 0, t0 past the data
        (1, T0 + 500, 20 * 1000, 1000, 2500, 19 * 1300, 1300),
        (1, 0xFFFFFFFF - 20 * TA, 10 * TA, TA, 2 * TA, 5 * TA, TA),  # t0 + tau wraps
        (2, T0, 20 * TA, TA, TA, 10 * TA, TA),
        (2, T0 - 4 * TA, 20 * TA, TA, 2 * TA, 10 * 2000, 2000),
        (2, T0 + 700, 15 * 1700, 1700, 900, 12 * 2300, 2300),
        (0, 0, 0, 0, 0, 0, 0),
    ],
)
def test_c_oracle_equals_python_twin(win):
    rng = np.random.default_rng(abs(hash(win)) % 2**32)
    N = 30
    cols = _random_atoms(rng, N)
    om = oracle.compute_map(cols, TA, T0, win)
    Ft, weff = twin.transient_map(cols, TA, T0, win)
    assert om.window_eff == weff
    np.testing.assert_allclose(om.F_mn, Ft, rtol=1e-11, atol=1e-12)


def test_noise_expectation_E_F_is_2():
    """x ~ CN(0,1) per atom, Fa = sum a x, Fb = sum b x  =>  cov(Fa,Fb) = M and
    E[F] = tr(M^-1 M) = 2 (E[2F] = 4), for any window."""
    rng = np.random.default_rng(11)
    N, K = 40, 4000
    t = np.arange(N)
    ph = (0.7, 1.9)
    a = 0.3 + 0.5 * np.cos(2 * np.pi * t * TA / 86164 + ph[0])
    b = 0.4 * np.sin(2 * np.pi * t * TA / 86164 + ph[1])
    win = (1, T0, 0, TA, 12 * TA, 0, TA)  # one cell: atoms 0..11
    vals = []
    for _ in range(K):
        x = (rng.standard_normal(N) + 1j * rng.standard_normal(N)) * np.sqrt(0.5)
        cols = np.empty((7, N), dtype=np.float32)
        cols[0], cols[1], cols[2] = a * a, b * b, a * b
        cols[3], cols[4] = (a * x).real, (a * x).imag
        cols[5], cols[6] = (b * x).real, (b * x).imag
        vals.append(oracle.compute_map(cols, TA, T0, win).F_mn[0, 0])
    vals = np.array(vals)
    # 2F ~ chi^2 with 4 degrees of freedom: F has mean 2 and variance 2
    assert vals.mean() == pytest.approx(2.0, abs=4 * np.sqrt(2.0 / K))
    assert vals.var() == pytest.approx(2.0, rel=0.15)


def test_scale_and_permutation_invariance():
    rng = np.random.default_rng(12)
    N = 50
    cols = _random_atoms(rng, N)
    win = (1, T0, 30 * TA, TA, 2 * TA, 15 * TA, TA)
    base = oracle.compute_map(cols, TA, T0, win).F_mn
    s = np.float32(4.0)  # a power of two: the scaled float32 atoms are exact
    sc = cols.copy()
    sc[:3] *= s * s
    sc[3:] *= s
    np.testing.assert_allclose(oracle.compute_map(sc, TA, T0, win).F_mn, base, rtol=1e-9)
    perm = rng.permutation(N)
    full = (1, T0, 0, TA, N * TA, 0, TA)
    assert oracle.compute_map(cols[:, perm], TA, T0, full).F_mn[0, 0] == pytest.approx(
        oracle.compute_map(cols, TA, T0, full).F_mn[0, 0], rel=1e-12)


def test_cond_guard_at_known_condition_number():
    """Two atoms with a2 = b2 = 1/2 and ab = c/2 give A = B = 1, C = c and
    cond = (1+c)/(1-c).  The adjacent float32 values of ab around the c where
    cond = 1e4 put one window just below the guard (F from the formula) and
    one just above (F = 2)."""
    c_star = (1e4 - 1) / (1e4 + 1)
    ab_star = np.float32(c_star / 2)
    below = np.nextafter(ab_star, np.float32(0))
    above = np.nextafter(ab_star, np.float32(1))
    for ab, expect_formula in ((below, True), (above, False)):
        cols = np.zeros((7, 2), dtype=np.float32)
        cols[0] = 0.5
        cols[1] = 0.5
        cols[2] = ab
        cols[3] = 0.25
        cols[5] = -0.5
        om = oracle.compute_map(cols, TA, T0, (1, T0, 0, TA, 2 * TA, 0, TA))
        C = 2 * np.float64(ab)
        cond = (1 + C) / (1 - C)
        assert (cond < 1e4) == expect_formula
        assert om.cond[0, 0] == pytest.approx(cond, rel=1e-9)
        if expect_formula:
            Mi = np.linalg.inv(np.array([[1.0, C], [C, 1.0]]))
            x = np.array([0.5, -1.0])
            assert om.F_mn[0, 0] == pytest.approx(x @ Mi @ x, rel=1e-9)
        else:
            assert om.F_mn[0, 0] == 2.0


def test_single_atom_and_empty_windows_give_2():
    rng = np.random.default_rng(13)
    cols = _random_atoms(rng, 10)
    cols[:, 4] = 0  # a gap bin
    om = oracle.compute_map(cols, TA, T0, (1, T0 + 4 * TA, 0, TA, TA, 0, TA))
    assert om.F_mn[0, 0] == 2.0  # empty (gap) window
    om = oracle.compute_map(cols, TA, T0, (1, T0 + 2 * TA, 0, TA, TA, 0, TA))
    assert om.F_mn[0, 0] == 2.0  # one atom: D = 0 up to rounding -> cond huge
    assert om.bad_cell == (0, 0)


def test_merge_matches_numpy_per_bin_sums():
    rng = np.random.default_rng(14)
    ts_list, col_list, off = [], [], [0]
    for X in range(3):
        n = 40 + 5 * X
        ts = np.sort(T0 + rng.integers(0, 60 * TA, size=n)).astype(np.uint32)
        ts_list.append(ts)
        col_list.append(rng.standard_normal((7, n)).astype(np.float32))
        off.append(off[-1] + n)
    ts = np.concatenate(ts_list)
    cols = np.concatenate(col_list, axis=1)
    merged, tmin = oracle.merge_binned(ts, cols, np.array(off), TA)
    tmax = max(t[-1] for t in ts_list)
    nb = (int(tmax) - int(tmin)) // TA + 1
    ref = np.zeros((7, nb), dtype=np.float32)
    for X in range(3):
        bins = (ts_list[X].astype(np.int64) - tmin) // TA
        for k, i in enumerate(bins):  # float32 adds in detector, then atom order
            ref[:, i] = (ref[:, i] + col_list[X][:, k]).astype(np.float32)
    np.testing.assert_array_equal(merged, ref)
    m2, _ = twin.merge(ts, cols, off, TA)
    np.testing.assert_array_equal(merged, m2)


def test_two_detectors_equal_presummed_detector():
    rng = np.random.default_rng(15)
    N = 40
    c1 = _random_atoms(rng, N)
    c2 = _random_atoms(rng, N)
    ts = (T0 + TA * np.arange(N)).astype(np.uint32)
    both = np.concatenate([c1, c2], axis=1)
    m2, t0d = oracle.merge_binned(np.concatenate([ts, ts]), both, np.array([0, N, 2 * N]), TA)
    np.testing.assert_array_equal(m2, (c1 + c2).astype(np.float32))


def test_lnBtSG_and_posteriors_vs_library_routines():
    rng = np.random.default_rng(16)
    F = rng.gamma(2.0, 1.0, size=(37, 23)) + 5 * (rng.random((37, 23)) < 0.01)
    lnb = oracle.lnBtSG(F)
    assert lnb == pytest.approx(np.log(70 / F.size) + logsumexp(F), rel=1e-13)
    win = (1, T0, 36 * TA, TA, 2 * TA, 22 * TA, TA)
    t0_MP, tau_MP = oracle.posterior_modes(F, win)
    m = int(np.argmax(logsumexp(F, axis=1)))
    n = int(np.argmax(logsumexp(F, axis=0)))
    assert t0_MP == pytest.approx(T0 + (m + 0.5) * 36 * TA / 37)
    assert tau_MP == pytest.approx(2 * TA + (n + 0.5) * 22 * TA / 23)
