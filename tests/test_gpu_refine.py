"""Batched merge and exponential-window refinement on the GPU, against the
oracle (oracle.merge_binned, oracle/refine_ref.py).

Tolerances (FP64 sums on both sides; the kernel's as Horner recurrences in
the ratio of consecutive weights, one exp a lane for the first weight, the
oracle's from np.exp per bin):
  merge            bit-exact (float32 sums in detector order)
  2F, lnBtSG       relative 1e-8
  dF/dtau, dF/dS   relative 1e-6 of the candidate's largest component
  dF/dt0           |.| <= 1e-9 |F| / tau  (zero in exact arithmetic)
  atom gradients   float32 output: 2e-6 of the largest entry
  2F at grid maxima vs the float32 map: relative 1e-4
"""

import numpy as np
import pytest

import oracle
from oracle import refine_ref as rr
from pyfstat_amd import synth

pytestmark = pytest.mark.gpu

TA = synth.TSFT


def _device():
    import torch

    return torch.device("cuda", 0)


def _merged(work):
    import torch

    from pyfstat_amd.refine import MergedAtoms

    ts_d = torch.from_numpy(work.ts.view(np.int32)).to(_device())
    cols_d = torch.from_numpy(work.cols).to(_device())
    return MergedAtoms(ts_d, work.seg_off, work.seg_begin, work.TAtom, work.ts.size, _device(),
                       cols_raw=cols_d), cols_d


def _oracle_merge(work, t):
    s0, s1 = int(work.seg_begin[t]), int(work.seg_begin[t + 1])
    a, e = int(work.seg_off[s0]), int(work.seg_off[s1])
    det_off = np.asarray(work.seg_off[s0:s1 + 1]) - a
    return oracle.merge_binned(work.ts[a:e], work.cols[:, a:e], det_off, work.TAtom)


@pytest.fixture(scope="module")
def small(hip_ready):
    work = synth.workload_exp_refine(T=5, n_sft=400, n_t0=40, n_tau=40, seed_index=40)
    merged, cols_d = _merged(work)
    return work, merged, cols_d


def test_merge_batch_bit_exact(small):
    work, merged, _ = small
    cols = merged.cols.cpu().numpy()
    ab = merged.atom_bin.cpu().numpy()
    for t in range(work.T):
        print(f"merge template {t}", flush=True)
        ref, t0d = _oracle_merge(work, t)
        assert int(merged.t_data[t]) == t0d
        a, e = int(merged.m_off[t]), int(merged.m_off[t + 1])
        assert e - a == ref.shape[1]
        assert np.array_equal(cols[:, a:e], ref)
        s0, s1 = int(work.seg_begin[t]), int(work.seg_begin[t + 1])
        j0, j1 = int(work.seg_off[s0]), int(work.seg_off[s1])
        want = a + (work.ts[j0:j1].astype(np.int64) - t0d) // TA
        assert np.array_equal(ab[j0:j1], want)


def test_merge_adjoint_gathers_bin_gradients(small):
    import torch

    work, merged, cols_d = small
    raw = cols_d.clone().requires_grad_(True)
    out = merged.merge(raw)
    assert torch.equal(out, merged.cols)
    R = torch.randn(out.shape, generator=torch.Generator().manual_seed(0)).to(out.device)
    (out * R).sum().backward()
    ab = merged.atom_bin
    assert torch.equal(raw.grad, R[:, ab])


def _problem(work, merged, per_template=4, skip=()):
    from pyfstat_amd.batched import BatchedTransientMaps
    from pyfstat_amd.refine import ExpRefineProblem, candidates_from_maps

    b = BatchedTransientMaps(work.ts, work.cols, work.seg_off, work.seg_begin, work.windows,
                             work.TAtom)
    b.run()
    ct, t0, tau, Fg = candidates_from_maps(b, per_template=per_template)
    keep = ~np.isin(ct, np.asarray(skip, dtype=np.int64))
    b.close()
    prob = ExpRefineProblem(merged, ct[keep], t0[keep], tau[keep])
    return prob, Fg[keep][prob.order]


def _oracle_eval(work, merged, prob, t0, tau):
    cols = merged.cols.cpu().numpy().astype(np.float64)
    return rr.batch_evaluate(cols, merged.m_off, merged.t_data, TA, prob.cand_tmpl, prob.k_lo,
                             prob.k_hi, t0, tau)


def test_forward_at_grid_maxima_equals_map(small):
    work, merged, _ = small
    prob, Fgrid = _problem(work, merged)
    twoF, lnB, saved = prob.forward_raw(merged.cols, prob.t0_init, prob.tau_init)
    twoF = twoF.cpu().numpy()
    np.testing.assert_allclose(twoF, 2 * Fgrid, rtol=1e-4)
    np.testing.assert_allclose(lnB.cpu().numpy(), np.log(70.0) + twoF / 2, rtol=1e-12)
    F, dt0, dtau, g = _oracle_eval(work, merged, prob, prob.t0_init.cpu().numpy(),
                                   prob.tau_init.cpu().numpy())
    np.testing.assert_allclose(twoF, 2 * F, rtol=1e-8)


def _check_forward(prob, merged, work, t0, tau):
    import torch

    t0_d = torch.from_numpy(t0).to(_device())
    tau_d = torch.from_numpy(tau).to(_device())
    twoF, lnB, saved = prob.forward_raw(merged.cols, t0_d, tau_d)
    twoF = twoF.cpu().numpy()
    saved = saved.cpu().numpy()
    F, dt0, dtau, g = _oracle_eval(work, merged, prob, t0, tau)
    np.testing.assert_allclose(twoF, 2 * F, rtol=1e-8)
    np.testing.assert_allclose(saved[:, 9], F, rtol=1e-8)
    scale = np.maximum(np.abs(g).max(axis=1), 1e-300)
    assert np.all(np.abs(saved[:, 2:9] - g) <= 1e-6 * scale[:, None])
    tau_scale = np.maximum(np.abs(dtau), 1e-12 * np.abs(F) / tau)
    assert np.all(np.abs(saved[:, 1] - dtau) <= 1e-6 * tau_scale + 1e-12 * np.abs(F) / tau)
    assert np.all(np.abs(saved[:, 0]) <= 1e-9 * np.abs(F) / tau)
    return twoF, saved


def test_forward_off_grid_vs_oracle(small):
    work, merged, _ = small
    prob, _ = _problem(work, merged)
    rng = np.random.default_rng(2)
    t0 = prob.t0_init.cpu().numpy() + rng.uniform(-0.4, 0.4, prob.n) * TA
    tau = prob.tau_init.cpu().numpy() * rng.uniform(0.8, 1.25, prob.n)
    L = prob.support_lengths
    # supports of up to 8 bins run one candidate per lane, up to 32 bins four
    # lanes each (both from registers), longer ones 16 lanes each from LDS
    assert (L <= 8).any() and ((L > 8) & (L <= 32)).any() and (L > 32).any()
    print(f"off-grid forward: {prob.n} candidates, {(L <= 32).sum()} short", flush=True)
    _check_forward(prob, merged, work, t0, tau)


def test_gpu_finite_differences(small):
    """Central differences of the GPU forward itself agree with its dF/dtau."""
    import torch

    work, merged, _ = small
    prob, _ = _problem(work, merged)
    tau = prob.tau_init * 1.05
    h = 1e-5 * tau
    F0, _, saved = prob.forward_raw(merged.cols, prob.t0_init, tau)
    Fp = prob.forward_raw(merged.cols, prob.t0_init, tau + h, want_saved=False)[0]
    Fm = prob.forward_raw(merged.cols, prob.t0_init, tau - h, want_saved=False)[0]
    fd = ((Fp - Fm) / (2 * h) / 2).cpu().numpy()
    an = saved[:, 1].cpu().numpy()
    ok = saved[:, 9].cpu().numpy() != 2.0
    err = np.abs(fd - an)[ok]
    assert np.all(err <= 1e-5 * np.abs(an)[ok] + 1e-9 * (F0.cpu().numpy()[ok] / tau.cpu().numpy()[ok]))
    # t0 differences vanish on the fixed support
    d = torch.full_like(prob.t0_init, 300.0)
    F1 = prob.forward_raw(merged.cols, prob.t0_init + d, tau, want_saved=False)[0]
    np.testing.assert_allclose(F1.cpu().numpy(), F0.cpu().numpy(), rtol=1e-11)


def test_autograd_backward_vs_oracle(small):
    import torch

    work, merged, _ = small
    prob, _ = _problem(work, merged, skip=(2,))  # template 2 without candidates
    assert prob.cand_off[3] == prob.cand_off[2]
    rng = np.random.default_rng(4)
    wa = rng.standard_normal(prob.n)
    wb = rng.standard_normal(prob.n)
    cols = merged.cols.clone().requires_grad_(True)
    t0 = prob.t0_init.clone().requires_grad_(True)
    tau = (prob.tau_init * 1.02).requires_grad_(True)
    twoF, lnB = prob(cols, t0, tau)
    loss = (torch.from_numpy(wa).to(_device()) * twoF).sum() + \
        (torch.from_numpy(wb).to(_device()) * lnB).sum()
    loss.backward()
    F, dt0, dtau, g = _oracle_eval(work, merged, prob, t0.detach().cpu().numpy(),
                                   tau.detach().cpu().numpy())
    gF = 2 * wa + wb
    np.testing.assert_allclose(tau.grad.cpu().numpy(), gF * dtau, rtol=1e-6,
                               atol=1e-9 * np.abs(gF * F / tau.detach().cpu().numpy()).max())
    assert np.all(np.abs(t0.grad.cpu().numpy()) <= 1e-8 * np.abs(gF * F) / prob.tau_init.cpu().numpy())
    want = np.zeros(merged.cols.shape)
    cols_h = merged.cols.cpu().numpy()
    t0_h, tau_h = t0.detach().cpu().numpy(), tau.detach().cpu().numpy()
    for i in range(prob.n):
        t = int(prob.cand_tmpl[i])
        a, e = int(merged.m_off[t]), int(merged.m_off[t + 1])
        want[:, a:e] += rr.atom_grad((7, e - a), float(merged.t_data[t]), TA, int(prob.k_lo[i]),
                                     int(prob.k_hi[i]), float(t0_h[i]), float(tau_h[i]), g[i], gF[i])
    got = cols.grad.cpu().numpy()
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 2e-6 * scale
    a, e = int(merged.m_off[2]), int(merged.m_off[3])
    assert np.all(got[:, a:e] == 0)
    assert cols_h.shape == got.shape


def test_backward_wider_cols(small):
    """cols with columns past the merged bins: their gradient is zero and the
    rest equals the gradient for the exact-width cols."""
    import torch

    work, merged, _ = small
    prob, _ = _problem(work, merged)
    nb = merged.total_bins
    wide = torch.full((7, nb + 37), 123.0, dtype=torch.float32, device=_device())
    wide[:, :nb] = merged.cols
    cw = wide.clone().requires_grad_(True)
    prob(cw, prob.t0_init, prob.tau_init)[0].sum().backward()
    c = merged.cols.clone().requires_grad_(True)
    prob(c, prob.t0_init, prob.tau_init)[0].sum().backward()
    assert torch.equal(cw.grad[:, :nb], c.grad)
    assert torch.all(cw.grad[:, nb:] == 0)


def test_raw_atom_gradients_through_merge(small):
    work, merged, cols_d = small
    prob, _ = _problem(work, merged)
    raw = cols_d.clone().requires_grad_(True)
    m = merged.merge(raw)
    mm = m.detach().clone().requires_grad_(True)
    prob(mm, prob.t0_init, prob.tau_init)[0].sum().backward()
    twoF = prob(m, prob.t0_init, prob.tau_init)[0]
    twoF.sum().backward()
    import torch

    assert torch.equal(raw.grad, mm.grad[:, merged.atom_bin])


def test_ascent_monotone_bounded_and_converged(small):
    import torch

    work, merged, _ = small
    prob, _ = _problem(work, merged)
    tau0 = prob.tau_init
    lo = torch.clamp(tau0 - TA, min=float(TA))
    hi = tau0 + TA
    F_init = prob.forward_raw(merged.cols, prob.t0_init, tau0, want_saved=False)[0]
    tau0_before = tau0.clone()
    tau, twoF, nacc = prob.ascent(merged.cols, tau_lo=lo, tau_hi=hi, n_steps=20)
    assert torch.equal(tau0, tau0_before)  # the start point is not overwritten
    # in place through the C ABI (tau_in == tau_out): the same bits
    tau_ip = tau0.clone()
    twoF_ip = torch.empty_like(twoF)
    assert prob.lib.tcw_refine_ascent(
        prob.plan, merged.cols.data_ptr(), merged.cols.shape[1], prob.t0_init.data_ptr(),
        tau_ip.data_ptr(), tau_ip.data_ptr(), lo.data_ptr(), hi.data_ptr(), 20, 0.02,
        twoF_ip.data_ptr(), None, torch.cuda.current_stream().cuda_stream) == 0
    assert torch.equal(tau_ip, tau) and torch.equal(twoF_ip, twoF)
    tau_h, twoF_h = tau.cpu().numpy(), twoF.cpu().numpy()
    # never below the start (the forward and ascent kernels share the evaluation;
    # allow last-bit differences of separately compiled copies)
    assert np.all(twoF_h >= F_init.cpu().numpy() * (1 - 1e-12))
    assert np.all(tau_h >= lo.cpu().numpy() * (1 - 1e-12)) and np.all(tau_h <= hi.cpu().numpy() * (1 + 1e-12))
    again = prob.forward_raw(merged.cols, prob.t0_init, tau, want_saved=False)[0].cpu().numpy()
    np.testing.assert_allclose(again, twoF_h, rtol=1e-12)
    cols = merged.cols.cpu().numpy().astype(np.float64)
    close = 0
    for i in range(prob.n):
        t = int(prob.cand_tmpl[i])
        sub = cols[:, merged.m_off[t]:merged.m_off[t + 1]]
        Fmax, _ = rr.scan_max_tau(sub, float(merged.t_data[t]), TA, int(prob.k_lo[i]),
                                  int(prob.k_hi[i]), float(prob.t0_init[i]), float(lo[i]),
                                  float(hi[i]), n=401)
        close += int(twoF_h[i] >= 2 * Fmax - 1e-3 * abs(2 * Fmax))
    print(f"ascent: {close}/{prob.n} within 1e-3 of the scanned maximum; "
          f"accepted steps {nacc.cpu().numpy().tolist()}", flush=True)
    assert close >= 0.8 * prob.n


def test_ascent_many_candidates_consistent(hip_ready):
    """4096 grid maxima of every support class, 20 ascent steps: each
    candidate's 2F is F at the tau it returns (the forward kernel there, and
    the C oracle), i.e. the lanes of a multi-lane group walked one tau
    together.  Groups whose lanes drift apart in the last bit of their shared
    sums take different proposals, and after a few steps report an F that no
    single tau gives."""
    import torch

    work = synth.workload_exp_refine(T=256, n_sft=1024, n_t0=128, n_tau=128, seed_index=43)
    merged, _ = _merged(work)
    prob, _ = _problem(work, merged, per_template=16)
    L = prob.support_lengths
    print(f"{prob.n} candidates, supports {L.min()}..{L.max()} bins", flush=True)
    assert (L <= 8).sum() > 100 and ((L > 8) & (L <= 32)).sum() > 100 and (L > 32).sum() > 100
    tau0 = prob.tau_init
    lo = torch.clamp(tau0 - TA, min=float(TA))
    hi = tau0 + TA
    tau, twoF, nacc = prob.ascent(merged.cols, tau_lo=lo, tau_hi=hi, n_steps=20)
    again = prob.forward_raw(merged.cols, prob.t0_init, tau, want_saved=False)[0]
    twoF_h = twoF.cpu().numpy()
    np.testing.assert_allclose(again.cpu().numpy(), twoF_h, rtol=1e-11)
    Fo, _, _, _ = oracle.refine_eval_batch(merged.cols.cpu().numpy(), merged.m_off, merged.t_data,
                                           TA, prob.cand_tmpl, prob.k_lo, prob.k_hi,
                                           prob.t0_init.cpu().numpy(), tau.cpu().numpy())
    np.testing.assert_allclose(twoF_h, 2 * Fo, rtol=1e-9)
    assert nacc.cpu().numpy().max() >= 10  # long walks: the drift needs many steps


def test_ascent_on_user_stream(small):
    """The ascent forks part of its work to the plan's own stream: on a torch
    stream of the caller's, it still starts after the work queued there (the
    start tau is written on that stream just before) and the caller's later
    work sees its results."""
    import torch

    work, merged, _ = small
    prob, _ = _problem(work, merged)
    ref_tau, ref_twoF, ref_nacc = prob.ascent(merged.cols, n_steps=10)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        start = torch.empty_like(prob.tau_init)
        for _ in range(20):  # queued work the ascent must wait for
            start.copy_(prob.tau_init * 0.5)
            start.mul_(2.0)
        tau, twoF, nacc = prob.ascent(merged.cols, tau=start, n_steps=10)
        twoF_after = twoF * 1.0  # later work on the same stream
    s.synchronize()
    assert torch.equal(tau, ref_tau) and torch.equal(twoF, ref_twoF)
    assert torch.equal(nacc, ref_nacc) and torch.equal(twoF_after, ref_twoF)


def test_support_class_boundaries(small):
    """Supports on both sides of each kernel class boundary (1, 8 | 9, 32 |
    33, ... bins: one lane, four lanes, 16 lanes from LDS), several per
    template, against the FP64 oracle; the ascent from the same start."""
    import torch

    from pyfstat_amd.refine import ExpRefineProblem

    work, merged, _ = small
    rng = np.random.default_rng(7)
    ct, k_lo, k_hi, t0, tau = [], [], [], [], []
    for t in range(work.T):
        N = int(merged.m_off[t + 1] - merged.m_off[t])
        for L in (1, 2, 7, 8, 9, 16, 31, 32, 33, 64, 100, min(300, N)):
            lo = int(rng.integers(0, N - L + 1))
            ct.append(t)
            k_lo.append(lo)
            k_hi.append(lo + L - 1)
            t0.append(float(merged.t_data[t]) + lo * TA + rng.uniform(-0.3, 0.3) * TA)
            tau.append(max(L, 1) * TA * rng.uniform(0.3, 1.5))
    prob = ExpRefineProblem(merged, ct, t0, tau, k_lo=k_lo, k_hi=k_hi)
    L = prob.support_lengths
    print(f"class boundaries: {prob.n} candidates, lengths {sorted(set(L.tolist()))}", flush=True)
    _check_forward(prob, merged, work, prob.t0_init.cpu().numpy(), prob.tau_init.cpu().numpy())
    F0 = prob.forward_raw(merged.cols, prob.t0_init, prob.tau_init, want_saved=False)[0]
    tau2, twoF, nacc = prob.ascent(merged.cols, n_steps=12)
    assert torch.all(twoF >= F0 * (1 - 1e-12))
    again = prob.forward_raw(merged.cols, prob.t0_init, tau2, want_saved=False)[0]
    np.testing.assert_allclose(again.cpu().numpy(), twoF.cpu().numpy(), rtol=1e-12)


def test_ascent_nonpositive_tau(small):
    """A start tau <= 0 (or NaN) gives 2F = NaN, tau passed through and no
    accepted step, in every kernel class; the other candidates are unchanged."""
    import torch

    work, merged, _ = small
    prob, _ = _problem(work, merged)
    L = prob.support_lengths
    pick = [int(np.nonzero(L <= 8)[0][0]), int(np.nonzero((L > 8) & (L <= 32))[0][0]),
            int(np.nonzero(L > 32)[0][0])]
    tau = prob.tau_init.clone()
    bad = torch.tensor([0.0, -5.0, float("nan")], dtype=torch.float64, device=tau.device)
    tau[pick] = bad
    tau_out, twoF, nacc = prob.ascent(merged.cols, tau=tau, n_steps=8)
    ref_tau, ref_twoF, ref_nacc = prob.ascent(merged.cols, n_steps=8)
    keep = np.ones(prob.n, bool)
    keep[pick] = False
    assert torch.all(torch.isnan(twoF[pick]))
    assert torch.equal(nacc[pick], torch.zeros(3, dtype=torch.int32, device=nacc.device))
    got = tau_out[pick].cpu().numpy()
    assert got[0] == 0.0 and got[1] == -5.0 and np.isnan(got[2])
    k = torch.from_numpy(keep).to(twoF.device)
    assert torch.equal(twoF[k], ref_twoF[k]) and torch.equal(tau_out[k], ref_tau[k])


def test_global_path_long_support(hip_ready):
    """Supports longer than the LDS staging capacity read atoms from HBM."""
    import torch

    from pyfstat_amd.refine import ExpRefineProblem

    work = synth.workload_exp_refine(T=2, n_sft=3000, n_t0=8, n_tau=8, seed_index=41)
    merged, _ = _merged(work)
    ct = np.array([0, 0, 1, 1])
    t0 = np.array([synth.T_START, synth.T_START + 3 * TA, synth.T_START, synth.T_START + TA])
    tau = np.array([900 * TA, 20 * TA, 950 * TA, 5 * TA])
    prob = ExpRefineProblem(merged, ct, t0, tau)
    assert prob.support_lengths.max() > 2176
    print(f"long supports: {prob.support_lengths.tolist()}", flush=True)
    _check_forward(prob, merged, work, prob.t0_init.cpu().numpy(),
                   prob.tau_init.cpu().numpy() * 1.01)
    tau2, twoF, _ = prob.ascent(merged.cols, n_steps=5)
    assert torch.all(twoF >= prob.forward_raw(merged.cols, prob.t0_init, prob.tau_init,
                                              want_saved=False)[0] * (1 - 1e-12))


def test_wide_lds_unions(hip_ready):
    """Support unions of ~1200 bins (two long candidates per template, well
    inside the LDS staging capacity): forward and ascent against the
    reference."""
    import torch

    from pyfstat_amd.refine import ExpRefineProblem

    work = synth.workload_exp_refine(T=2, n_sft=3000, n_t0=8, n_tau=8, seed_index=42)
    merged, _ = _merged(work)
    ct = np.array([0, 0, 1, 1])
    k_lo = np.array([0, 500, 100, 150])
    k_hi = np.array([600, 1199, 1250, 1300])
    t0 = np.array([float(merged.t_data[t]) + k * TA for t, k in zip(ct, k_lo)])
    tau = np.array([200.0, 250.0, 380.0, 300.0]) * TA
    prob = ExpRefineProblem(merged, ct, t0, tau, k_lo=k_lo, k_hi=k_hi)
    print(f"wide LDS unions: supports {prob.support_lengths.tolist()}", flush=True)
    _check_forward(prob, merged, work, prob.t0_init.cpu().numpy(),
                   prob.tau_init.cpu().numpy() * 1.01)
    F0 = prob.forward_raw(merged.cols, prob.t0_init, prob.tau_init, want_saved=False)[0]
    tau2, twoF, _ = prob.ascent(merged.cols, n_steps=6)
    assert torch.all(twoF >= F0 * (1 - 1e-12))
    again = prob.forward_raw(merged.cols, prob.t0_init, tau2, want_saved=False)[0]
    np.testing.assert_allclose(again.cpu().numpy(), twoF.cpu().numpy(), rtol=1e-12)


def test_problem_validation(small):
    from pyfstat_amd.refine import ExpRefineProblem

    work, merged, _ = small
    with pytest.raises(ValueError):
        ExpRefineProblem(merged, [0], [synth.T_START], [3600], k_lo=[5], k_hi=[4])
    with pytest.raises(ValueError):
        ExpRefineProblem(merged, [7], [synth.T_START], [3600])
    with pytest.raises(ValueError):
        ExpRefineProblem(merged, [0], [synth.T_START], [3600], k_lo=[0], k_hi=[10**6])
