#!/usr/bin/env python3
"""Benchmark of the batched transient F-statistic map on one MI355X.

Default workload: rectangular-window maps for 4096 templates with ragged
N_sft ~ U[1000, 8000] (10% random gaps each, 5% with an injected transient),
N_t0 = N_tau = N_sft // 4, lnBtSG and the posterior modes fused
(BtSG=True) -- 6.3e9 2F cells per step.  One step = the whole batch through
the device path (binned merge + FP64 prefix scan, map kernels with fused
reductions, per-template finalize), replayed from a hipGraph; the atoms are
resident in HBM before the timed region.

Prints ONE JSON line.  Besides the contract fields it carries:
  roofline      the map kernel: 4 B per 2F cell written / its event-timed
                duration, against MI355X HBM peak (8 TB/s)
  plain         the same kernels launched once per template from a Python loop
                with a sync per template (the structure of TransientGridSearch),
                on a sample of templates, and the batched/plain speedup
  cpu_baseline  the FP64 C oracle (OpenMP over templates) on a bounded sample
  parity        max relative 2F error, argmax and lnBtSG agreement against the
                oracle on a sample of templates
"""

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

METRIC = ("2F(t0,τ) cells/s on 1 GPU, batched vs plain at ≤1e-4 rel err; "
          "% HBM/MFMA roofline")
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


WORKLOADS = {
    "rect_batched": "rect window, 4096 templates, N_sft~U[1000,8000], 10% gaps, "
                    "N_t0=N_tau=N_sft//4",
    "rect_plain": "rect window, 1 template, 500 SFTs (10% gaps), 500 t0 x 500 tau",
    "exp_plain": "exp window, 1 template, 2 detectors x 4000 SFTs (15% gaps), "
                 "2000 t0 x 2000 tau",
    "exp_shared": "exp window, 256 templates x 8192 shared timestamps (2 detectors, 10% "
                  "gaps), 1024 t0 x 512 tau",
    "exp_refine": "exp-window refinement: 4096 templates x 2 detectors x 2048 SFTs (10% "
                  "gaps), 256 t0 x 256 tau maps, 16 grid maxima each = 65536 candidates, "
                  "20 ascent steps in tau on fixed supports",
    "demod_atoms": "F-statistic atoms from SFTs: 4096 templates (1024 f spaced 1/T_obs x 4 "
                   "fdot) x 2 detectors x 4000 SFTs (10% gaps), one sky position, dterms 8",
}

DEMOD_METRIC = "F-statistic atoms/s on 1 GPU (demodulation sum over 2*dterms SFT bins)"
# bytes written per atom: t (4) + a2, b2, ab, Re/Im Fa, Re/Im Fb (7 x 4)
DEMOD_ATOM_BYTES = 32.0

REFINE_METRIC = "refined candidate-steps/s on 1 GPU (2F + d2F/dtau per step)"
REFINE_WARM_S = 0.5

# HBM bytes per launch of the map kernel measured with rocprofv3 --pmc FETCH_SIZE
# and WRITE_SIZE (scaled by a 1 GiB copy in the same run) for the default-size
# workloads; the commands and raw counters are in the named summaries.
PROFILED_TRAFFIC = {
    "rect_batched": (1.017e9 + 25.43e9, "profiles/rect_fast_kernel.rect_batched.md"),
    "exp_shared": (0.156e9 + 0.540e9, "profiles/exp_fast_kernel.exp_shared.md"),
    "demod_atoms": (0.0207e9 + 0.9440e9, "profiles/demod_atoms_kernel.demod_atoms.md"),
}
FP64_VALU_PEAK_TFLOPS = 78.6
# What bounds the map kernel instead of HBM (rocprofv3 counters of the same
# kernel, summarised in the profile files named here).
LIMITER = {
    "rect_batched": "VALU issue: 43 VALU instructions per cell (24 FP64 add/mul/FMA; part of "
                    "the rest is lanes past N_tau in each template's last 256-column block, 8.6% "
                    "of the lane-cells here), the VALU issuing every SIMD cycle at the "
                    "~1.76 GHz the card holds under this load; "
                    "profiles/rect_fast_kernel.rect_batched.md",
    "exp_shared": "latency at two waves per SIMD (one row chunk per column: 2048 waves): "
                  "~56 VALU instructions per cell at ~59% VALU busy, FP64 dependency chains and "
                  "LDS broadcasts of the entering atoms; profiles/exp_fast_kernel.exp_shared.md",
}


def make_workload(args):
    """(BatchedTransientMaps constructor kwargs builder, batch-like object)."""
    from pyfstat_amd import synth

    if args.workload == "rect_batched":
        return synth.workload_rect_batched(T=args.templates)
    if args.workload == "rect_plain":
        multi, win = synth.workload_rect_plain()
        return ([multi], [win])
    if args.workload == "exp_plain":
        multi, win = synth.workload_exp_plain()
        return ([multi], [win])
    if args.workload == "exp_shared":
        multis, win = synth.workload_exp_shared()
        return (multis, [win] * len(multis))
    raise ValueError(args.workload)


def build_batch(work, BtSG):
    from pyfstat_amd.batched import BatchedTransientMaps

    if isinstance(work, tuple):
        return BatchedTransientMaps.from_multi_atoms(work[0], work[1], BtSG=BtSG)
    return BatchedTransientMaps.from_synthetic(work, BtSG=BtSG)


def plain_loop(b, batch, idx, BtSG):
    """The same kernels, one template at a time, sync per template."""
    import torch

    from pyfstat_amd import _hiplib

    lib = b.lib
    dev = b.device
    max_cells = int(max(b.F_off[t + 1] - b.F_off[t] for t in idx))
    F = torch.empty(max_cells, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)
    ts_base = b.ts_d.data_ptr()
    cols_base = b.cols_d.data_ptr()
    flags = _hiplib.TCW_FLAG_BTSG if BtSG else 0
    st = _hiplib.Stats()
    at = _hiplib.Atoms()
    cells = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in idx:
        a, e = int(batch.off[t]), int(batch.off[t + 1])
        at.numDet = 1
        at.TAtom = batch.TAtom
        at.ts = ts_base + 4 * a
        at.cols = cols_base + 4 * a
        at.col_stride = b.col_stride
        at.det_off[0] = 0
        at.det_off[1] = e - a
        at.t_min = int(batch.ts[a])
        at.t_max = int(batch.ts[e - 1])
        win = _hiplib.window_struct((batch.windows[t].type, batch.windows[t].t0,
                                     batch.windows[t].t0Band, batch.windows[t].dt0,
                                     batch.windows[t].tau, batch.windows[t].tauBand,
                                     batch.windows[t].dtau))
        _hiplib.check(lib.tcw_map(b.handle, ctypes.byref(at), ctypes.byref(win),
                                  ctypes.c_void_p(F.data_ptr()), ctypes.byref(st), flags,
                                  ctypes.c_void_p(stream.cuda_stream)), "tcw_map")
        cells += int(st.N_t0 * st.N_tau)
    torch.cuda.synchronize()
    return cells, time.perf_counter() - t0


class Units:
    """Independent pieces of a workload that the FP64 C oracle maps on the
    CPU: unit i -> (merged atoms float32 [7, N], t0_data, window, cells).

    rect_batched: whole templates (seeded order); rect_plain: the one template
    (the all-core figure runs one copy per core, as one lal process per core
    would); exp_plain / exp_shared: single t0 rows (1-row windows over the full
    tau range) of seeded-random (template, row) pairs -- the direct exp sums are
    too slow on the CPU for whole maps, so the rate is measured on rows and
    applies to the whole map (every row costs about the same on average).
    """

    def __init__(self, args, work, batch, seed=5):
        import oracle
        from pyfstat_amd.atoms import atoms_to_arrays, transientWindowRange_t

        self.kind = args.workload
        self._oracle = oracle
        self._Win = transientWindowRange_t
        self._merged = {}
        rng = np.random.default_rng(seed)
        if batch is not None:
            self.templates = batch.T
            self.windows = batch.windows
            self._atoms = lambda t: (batch.ts[batch.off[t]:batch.off[t + 1]],
                                     batch.cols[:, batch.off[t]:batch.off[t + 1]],
                                     np.array([0, batch.off[t + 1] - batch.off[t]]), batch.TAtom)
        else:
            multis, wins = work
            self.templates = len(multis)
            self.windows = wins
            packed = [None] * len(multis)

            def atoms(t):
                if packed[t] is None:
                    packed[t] = atoms_to_arrays(multis[t])
                p = packed[t]
                return p.ts, p.cols, p.det_off, p.TAtom
            self._atoms = atoms
        self.rows = self.kind in ("exp_plain", "exp_shared")
        if self.rows:
            n_t0 = [w.t0Band // w.dt0 + 1 for w in self.windows]
            pairs = [(t, m) for t in range(self.templates) for m in range(n_t0[t])]
            self.order = [pairs[i] for i in rng.permutation(len(pairs))]
        elif self.kind == "rect_plain":
            self.order = None
        else:
            self.order = [(int(t), None) for t in rng.permutation(self.templates)]
        self.n = None if self.order is None else len(self.order)
        self.with_lnBtSG = not self.rows

    def merged(self, t):
        if t not in self._merged:
            ts, cols, det_off, TA = self._atoms(t)
            self._merged[t] = (self._oracle.merge_binned(ts, cols, det_off, TA), TA)
        return self._merged[t]

    def __call__(self, i):
        t, m = (0, None) if self.order is None else self.order[i]
        (merged, t0d), TA = self.merged(t)
        w = self.windows[t]
        if m is not None:
            w = self._Win(w.type, w.t0 + m * w.dt0, 0, w.dt0, w.tau, w.tauBand, w.dtau)
        cells = (w.t0Band // w.dt0 + 1) * (w.tauBand // w.dtau + 1)
        return merged, t0d, w, TA, cells, (t, m)

    def describe(self, k):
        if self.rows:
            return f"{k} seeded-random t0 rows (1-row windows, full tau range)"
        if self.order is None:
            return f"{k} copies of the template"
        return f"{k} seeded-random templates"


# Cells one oracle call holds on the host (8 bytes each): samples larger than
# this are timed by repeating a capped sample instead.
CPU_CELL_CAP = 2.5e8


def _time_units(units, k, threads, reps=1):
    """Map units 0..k-1 on `threads` OpenMP threads, reps times; (cells, seconds)."""
    import oracle

    oracle.set_threads(threads)
    sample = [units(i) for i in range(k)]  # merged atoms prepared outside the clock
    t0 = time.perf_counter()
    for _ in range(reps):
        oracle.compute_map_batch([x[0] for x in sample], [x[3] for x in sample],
                                 [x[1] for x in sample], [x[2] for x in sample],
                                 with_lnBtSG=units.with_lnBtSG)
    dt = time.perf_counter() - t0
    return reps * sum(x[4] for x in sample), dt


def _sized_run(units, threads, budget_s):
    """Grow the sample until one timed run takes about budget_s; past
    CPU_CELL_CAP cells, repeat the capped sample instead.  Returns
    (units in the sample, repetitions, cells, seconds)."""
    cap = units.n if units.n is not None else 10**9
    k = min(threads, cap)
    cells, dt = _time_units(units, k, threads)
    for _ in range(4):
        if dt >= 0.7 * budget_s or k >= cap:
            break
        kmax = max(threads, int(CPU_CELL_CAP / max(cells / k, 1.0)))
        k_new = min(cap, kmax, max(k + 1, int(k * budget_s / max(dt, 1e-3)) + 1))
        if k_new <= k:
            break
        k = k_new
        cells, dt = _time_units(units, k, threads)
    reps = 1
    if dt < 0.7 * budget_s:
        reps = int(np.ceil(budget_s / max(dt, 1e-3)))
        cells, dt = _time_units(units, k, threads, reps)
    return k, reps, cells, dt


def cpu_baseline(units, budget_s):
    """FP64 C oracle (built here with -O3 -march=native), single core and all
    cores (OpenMP over units), on bounded seeded samples of the workload."""
    import oracle

    flags = oracle.select_native()
    cores = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp:
        cores = min(cores, int(omp))
    k1, r1, c1, t1 = _sized_run(units, 1, 0.3 * budget_s)
    log(f"cpu single core: {units.describe(k1)} x {r1}, {c1:.3e} cells in {t1:.2f} s")
    kn, rn, cn, tn = _sized_run(units, cores, 0.7 * budget_s)
    log(f"cpu {cores} cores: {units.describe(kn)} x {rn}, {cn:.3e} cells in {tn:.2f} s")
    frac = ""
    if units.rows:
        total_rows = units.n
        frac = f"; {kn}/{total_rows} = {kn / total_rows:.2%} of the workload's rows"
    if rn > 1:
        frac += f"; the sample mapped {rn} times"
    return {
        "value": cn / tn,
        "unit": "cells/s",
        "cores": cores,
        "kind": "port",
        "single_core": c1 / t1,
        "cpu_model": oracle.cpu_model(),
        "build": flags,
        "sample": f"{units.describe(kn)} ({cn:.3e} cells, {tn:.1f} s{frac}), FP64 C oracle "
                  f"(lal-equivalent restatement, not lalpulsar)"
                  f"{' with lnBtSG' if units.with_lnBtSG else ''}, OpenMP over the sample; "
                  f"single core: {units.describe(k1)}{f' x {r1}' if r1 > 1 else ''} "
                  f"({c1:.3e} cells, {t1:.1f} s)",
    }


def _err_figures(F_gpu, om_F, om_cond):
    mask = np.abs(om_cond / 1e4 - 1) < 1e-3
    rel = np.abs(F_gpu - om_F) / np.maximum(np.abs(om_F), 1e-30)
    return float(np.where(mask, 0, rel).max()), int(mask.sum())


def parity_sample(b, units, n, seed=3):
    """GPU maps against the FP64 oracle on a seeded sample.

    Whole templates (rect workloads): every cell, the fused argmax and lnBtSG.
    Rows (exp workloads, 1-row oracle windows): every cell of each row; the
    fused argmax and lnBtSG are then checked against the GPU map itself
    (argmax of the stored map, float64 logsumexp of it), since a full FP64
    direct-sum map of these sizes takes minutes on the CPU.
    """
    import oracle

    if n <= 0:
        return None
    st = b.stats()
    worst, worst_lnb, argmax_exact, near, checked = 0.0, 0.0, 0, 0, 0
    nunits = 1 if units.n is None else min(n, units.n)
    tmpl_seen = set()
    for i in range(nunits):
        merged, t0d, w, TA, cells, (t, m) = units(i)
        om = oracle.compute_map(merged, TA, t0d, w)
        F = b.F_d[int(b.F_off[t]):int(b.F_off[t + 1])].cpu().numpy()
        nt0, ntau = int(st[t]["N_t0"]), int(st[t]["N_tau"])
        F = F.reshape(nt0, ntau)
        if m is not None:
            F = F[m:m + 1]
        e, k = _err_figures(F, om.F_mn, om.cond)
        worst = max(worst, e)
        near += k
        checked += F.size
        if t in tmpl_seen:
            continue
        tmpl_seen.add(t)
        if m is None:
            lnb = oracle.lnBtSG(om.F_mn, om.maxF)
            worst_lnb = max(worst_lnb, abs(float(st[t]["lnBtSG"]) - lnb))
            argmax_exact += int((int(st[t]["m_ML"]), int(st[t]["n_ML"])) == (om.m_max, om.n_max))
        else:
            Fm = b.F_d[int(b.F_off[t]):int(b.F_off[t + 1])].cpu().numpy().astype(np.float64)
            j = int(np.argmax(Fm))
            argmax_exact += int((int(st[t]["m_ML"]), int(st[t]["n_ML"])) == divmod(j, ntau))
            mx = Fm[j]
            lnb = np.log(70.0 / Fm.size) + mx + np.log(np.exp(Fm - mx).sum())
            worst_lnb = max(worst_lnb, abs(float(st[t]["lnBtSG"]) - lnb))
    out = {
        "max_rel_err_2F": worst,
        "cells_checked": checked,
        "cells_near_cond_threshold": near,
        "max_abs_err_lnBtSG": worst_lnb,
    }
    if units.rows:
        out["rows"] = nunits
        out["argmax_consistent"] = f"{argmax_exact}/{len(tmpl_seen)}"
        out["lnBtSG_vs"] = "float64 logsumexp of the GPU map"
    else:
        out["templates"] = len(tmpl_seen)
        out["argmax_exact"] = f"{argmax_exact}/{len(tmpl_seen)}"
    return out


def refine_cpu_baseline(prob, merged, cols, TA, budget_s, rng):
    """The refinement's likelihood evaluation (2F + all derivatives at a
    candidate) in the FP64 C oracle (-O3 -march=native, built here), OpenMP
    over candidates: single core and all cores, on seeded-random candidates
    of the workload at their start points, repeated to about budget_s."""
    import oracle

    flags = oracle.select_native()
    cores = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp:
        cores = min(cores, int(omp))
    cols_h = cols.cpu().numpy()
    ct = np.asarray(prob.cand_tmpl)
    lo, hi = np.asarray(prob.k_lo), np.asarray(prob.k_hi)
    t0 = prob.t0_init.cpu().numpy()
    tau = prob.tau_init.cpu().numpy()

    def run(idx, threads, budget):
        oracle.set_threads(threads)
        args = (cols_h, merged.m_off, merged.t_data, TA, ct[idx], lo[idx], hi[idx], t0[idx],
                tau[idx])
        done, c0 = 0, time.perf_counter()
        while True:
            oracle.refine_eval_batch(*args)
            done += idx.size
            dt = time.perf_counter() - c0
            if dt >= budget:
                return done, dt

    perm = rng.permutation(prob.n)
    one = perm[:4096]
    d1, t1 = run(one, 1, 0.3 * budget_s)
    dn, tn = run(perm, cores, 0.7 * budget_s)
    bins_all = int(np.asarray(prob.support_lengths).sum())
    return {
        "value": dn / tn,
        "unit": "candidate-steps/s",
        "cores": cores,
        "kind": "port",
        "single_core": d1 / t1,
        "cpu_model": oracle.cpu_model(),
        "build": flags,
        "sample": f"all {prob.n} candidates ({bins_all} support bins) at their start points, "
                  f"{dn // prob.n} passes in {tn:.1f} s, FP64 C oracle evaluation (2F + all "
                  f"derivatives, direct exp per bin), OpenMP over candidates; single core: "
                  f"{one.size} seeded-random candidates, {d1 // one.size} passes in {t1:.1f} s",
    }


def run_refine(args):
    """The refinement workload: setup (maps, grid maxima, supports) untimed;
    one step = tcw_refine_ascent over every candidate (n_steps proposals, each
    a fused 2F + derivative evaluation on the candidate's support)."""
    import torch

    from oracle import refine_ref as rr
    from pyfstat_amd import synth
    from pyfstat_amd import tcw_fstat_map_funcs as tcw
    from pyfstat_amd.batched import BatchedTransientMaps
    from pyfstat_amd.refine import ExpRefineProblem, MergedAtoms, candidates_from_maps

    tcw.init_transient_fstat_map_features("hip")
    log("generating refinement workload ...")
    t_gen = time.perf_counter()
    work = synth.workload_exp_refine(T=args.templates)
    b = BatchedTransientMaps(work.ts, work.cols, work.seg_off, work.seg_begin, work.windows,
                             work.TAtom)
    b.run()
    ct, t0, tau, Fg = candidates_from_maps(b, per_template=16)
    merged = MergedAtoms.from_batch(b)
    b.close()
    prob = ExpRefineProblem(merged, ct, t0, tau)
    TA = work.TAtom
    lo = torch.clamp(prob.tau_init - TA, min=float(TA))
    hi = prob.tau_init + TA
    L = prob.support_lengths
    torch.cuda.synchronize()
    log(f"ready in {time.perf_counter() - t_gen:.1f} s: {prob.n} candidates, mean support "
        f"{L.mean():.1f} bins")
    edges = [0, 8, 32, 128, 512, int(L.max())]
    classes = {f"{lo + 1}-{hi}": [int(((L > lo) & (L <= hi)).sum()), int(L[(L > lo) & (L <= hi)].sum())]
               for lo, hi in zip(edges[:-1], edges[1:]) if hi > lo}
    log(f"support lengths (bins): candidates, total bins per class {classes}")
    n_steps = args.refine_steps
    # at least W untimed steps and REFINE_WARM_S seconds of them: the first
    # few hundred microseconds of ascent launches run while the engine clock
    # is still rising (10-15% slower per launch in kernel traces)
    t_warm = time.perf_counter()
    n_warm = 0
    while n_warm < args.warmup or time.perf_counter() - t_warm < REFINE_WARM_S:
        prob.ascent(merged.cols, tau_lo=lo, tau_hi=hi, n_steps=n_steps)
        n_warm += 1
        if n_warm % 20 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()
    log(f"warmed up with {n_warm} ascent steps in {time.perf_counter() - t_warm:.2f} s")
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        tau_r, twoF, nacc = prob.ascent(merged.cols, tau_lo=lo, tau_hi=hi, n_steps=n_steps)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    evals = prob.n * (n_steps + 1)  # the first evaluation at the start point
    log(f"ascent: {ms:.3f} ms per step of {prob.n} candidates x {n_steps} proposals")

    # forward alone and forward+backward through autograd, for the record
    cols = merged.cols
    reps = 10
    e0.record(stream)
    for _ in range(reps):
        prob.forward_raw(cols, prob.t0_init, prob.tau_init)
    e1.record(stream)
    torch.cuda.synchronize()
    fwd_ms = e0.elapsed_time(e1) / reps
    c = cols.clone().requires_grad_(True)
    tt = prob.tau_init.clone().requires_grad_(True)

    def fwd_bwd():
        c.grad = None  # as an optimizer's zero_grad(set_to_none=True)
        tt.grad = None
        f2, lnb = prob(c, prob.t0_init, tt)
        f2.sum().backward()

    fwd_bwd()  # warm the allocator for the gradient buffers
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        fwd_bwd()
    e1.record(stream)
    torch.cuda.synchronize()
    fb_ms = e0.elapsed_time(e1) / reps

    F_init = prob.forward_raw(cols, prob.t0_init, prob.tau_init, want_saved=False)[0]
    gain = (twoF - F_init).cpu().numpy()
    nacc_h = nacc.cpu().numpy()
    # parity: a sample against the FP64 reference, and the ascent against a scan
    rng = np.random.default_rng(11)
    idx = np.sort(rng.choice(prob.n, size=min(args.check_templates * 4, prob.n), replace=False))
    cols_h = cols.cpu().numpy().astype(np.float64)
    t0_h = prob.t0_init.cpu().numpy()
    tau_h = tau_r.cpu().numpy()
    lo_h, hi_h = lo.cpu().numpy(), hi.cpu().numpy()
    twoF_h = twoF.cpu().numpy()
    worst, scan_gap = 0.0, 0.0
    for i in idx:
        t = int(prob.cand_tmpl[i])
        sub = cols_h[:, merged.m_off[t]:merged.m_off[t + 1]]
        F, _, _, _ = rr.evaluate(sub, float(merged.t_data[t]), TA, int(prob.k_lo[i]),
                                 int(prob.k_hi[i]), float(t0_h[i]), float(tau_h[i]))
        worst = max(worst, abs(twoF_h[i] - 2 * F) / abs(2 * F))
        Fmax, _ = rr.scan_max_tau(sub, float(merged.t_data[t]), TA, int(prob.k_lo[i]),
                                  int(prob.k_hi[i]), float(t0_h[i]), float(lo_h[i]),
                                  float(hi_h[i]), n=401)
        scan_gap = max(scan_gap, (2 * Fmax - twoF_h[i]) / abs(2 * Fmax))
    # every candidate at its ascent result against the C oracle (FP64)
    import oracle

    Fo, _, dtau_o, _ = oracle.refine_eval_batch(cols.cpu().numpy(), merged.m_off, merged.t_data,
                                                TA, prob.cand_tmpl, prob.k_lo, prob.k_hi,
                                                t0_h, tau_h)
    worst_all = float(np.max(np.abs(twoF_h - 2 * Fo) / np.abs(2 * Fo)))
    cpu = None
    if not args.no_cpu_baseline:
        cpu = refine_cpu_baseline(prob, merged, cols, TA, args.cpu_seconds, rng)
        log(f"cpu: {cpu}")
    flops = 42.0 * float(L.sum()) * (n_steps + 1)
    achieved = flops / (ms / 1e3) / 1e12
    out = {
        "metric": REFINE_METRIC,
        "value": evals / (ms / 1e3),
        "unit": "candidate-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"exp_refine: {WORKLOADS['exp_refine']}",
            "candidates": prob.n,
            "ascent_steps": n_steps,
            "mean_support_bins": float(L.mean()),
            "support_classes": classes,
        },
        "roofline": {
            "bound": "valu",
            "kernel": "refine_ascent_kernel (16 or 32 lanes a candidate) + refine_ascent_regs_kernel (two streams)",
            "achieved": achieved,
            "peak": FP64_VALU_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_VALU_PEAK_TFLOPS,
            "traffic": None,
            "flop_model": "42 flops per support bin per evaluation",
        },
        "forward_ms": fwd_ms,
        "forward_backward_ms": fb_ms,
        "ascent": {
            "mean_2F_gain": float(gain.mean()),
            "min_2F_gain": float(gain.min()),
            "mean_accepted": float(nacc_h.mean()),
        },
        "parity": {
            "candidates": int(idx.size),
            "max_rel_err_2F": worst,
            "max_rel_gap_to_scanned_max": scan_gap,
            "all_candidates_max_rel_err_2F_vs_C_oracle": worst_all,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    prob.close()


def demod_templates(sft, n_f=1024, fdots=(0.0, -1e-11, -2e-11, -3e-11)):
    """(freq, fdot) of the demod workload: n_f frequencies spaced 1/T_obs from
    100.02 Hz, each with every fdot."""
    T_obs = float(sft.t_start.max()) + sft.Tsft - float(sft.t_start.min())
    f = 100.02 + np.arange(n_f) / T_obs
    ff, dd = np.meshgrid(f, np.asarray(fdots, dtype=np.float64), indexing="ij")
    return ff.ravel(), dd.ravel()


def demod_cpu_baseline(sft, freqs, fdots, dterms, budget_s, rng):
    """The demodulation sum in the FP64 C oracle (-O3 -march=native, built
    here), OpenMP over templates and SFTs: single core and all cores, on
    seeded-random templates of the workload, grown to about budget_s."""
    import oracle

    flags = oracle.select_native()
    cores = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp:
        cores = min(cores, int(omp))
    perm = rng.permutation(freqs.size)

    def run(k, threads, reps=1):
        oracle.set_threads(threads)
        idx = perm[:k]
        c0 = time.perf_counter()
        for _ in range(reps):
            oracle.demod_atoms(sft.t_start, sft.bins, sft.k_min, sft.Tsft, sft.delay,
                               sft.delay_rate, sft.a, sft.b, freqs[idx], fdots[idx],
                               sft.ref_time, dterms=dterms)
        return time.perf_counter() - c0

    def sized(threads, budget):
        """(templates, repetitions, seconds): the sample grows to the whole
        workload, which is then repeated to fill the budget."""
        k = min(freqs.size, max(threads, 4))
        dt = run(k, threads)
        while dt < 0.7 * budget and k < freqs.size:
            k = min(freqs.size, max(k + 1, int(k * budget / max(dt, 1e-3))))
            dt = run(k, threads)
        reps = 1
        if dt < 0.7 * budget:
            reps = int(np.ceil(budget / max(dt, 1e-3)))
            dt = run(k, threads, reps)
        return k, reps, dt

    k1, r1, t1 = sized(1, 0.3 * budget_s)
    kn, rn, tn = sized(cores, 0.7 * budget_s)
    n = sft.t_start.size
    return {
        "value": rn * kn * n / tn,
        "unit": "atoms/s",
        "cores": cores,
        "kind": "port",
        "single_core": r1 * k1 * n / t1,
        "cpu_model": oracle.cpu_model(),
        "build": flags,
        "sample": f"{kn} seeded-random templates x {n} SFTs ({kn * n:.3e} atoms) x {rn} in "
                  f"{tn:.1f} s, FP64 C restatement of the demodulation sum, OpenMP over "
                  f"templates and SFTs; single core: {k1} templates x {r1} "
                  f"({r1 * k1 * n:.3e} atoms, {t1:.1f} s)",
    }


def run_demod(args):
    """Atoms from SFTs: one step = tcw_demod_atoms over every (template, SFT)
    of the workload into preallocated device buffers; SFT bins, timing and
    antenna patterns resident in HBM before the timed region."""
    import torch

    from oracle import demod_ref as dr
    from pyfstat_amd import synth
    from pyfstat_amd import tcw_fstat_map_funcs as tcw
    from pyfstat_amd.demod import NC, demod_atoms_into

    tcw.init_transient_fstat_map_features("hip")
    dterms = 8
    log("generating SFTs ...")
    t_gen = time.perf_counter()
    sft = synth.synthetic_sfts(n_det=2, n_sft=4000, seed_index=80, dterms=dterms)
    sfts = sft.sft_set()
    freqs, fdots = demod_templates(sft, n_f=max(1, args.templates // 4))
    T, n = freqs.size, sfts.n_sft
    dev = sfts.device
    f_d = torch.from_numpy(freqs).to(dev)
    fd_d = torch.from_numpy(fdots).to(dev)
    ts_d = torch.empty(T * n, dtype=torch.int32, device=dev)
    cols_d = torch.empty((NC, T * n), dtype=torch.float32, device=dev)
    atoms = T * n
    torch.cuda.synchronize()
    log(f"ready in {time.perf_counter() - t_gen:.1f} s: {T} templates x {n} SFTs "
        f"({sfts.n_bins} bins each) = {atoms:.3e} atoms")

    def step(stream=None):
        demod_atoms_into(sfts, f_d, fd_d, sft.ref_time, ts_d, cols_d, dterms=dterms,
                         stream=stream)

    stream = torch.cuda.current_stream()
    if args.profile_mode:
        for _ in range(args.warmup + args.steps):
            step(stream)
        torch.cuda.synchronize()
        ref = torch.empty(2**28, dtype=torch.float32, device=dev)
        ref2 = torch.empty_like(ref)
        for _ in range(3):
            ref2.copy_(ref)  # 1 GiB read + 1 GiB write, the counters' yardstick
        torch.cuda.synchronize()
        log("profile mode done: demod steps + 3 x 1 GiB copies")
        return
    t_warm = time.perf_counter()
    n_warm = 0
    while n_warm < args.warmup or time.perf_counter() - t_warm < REFINE_WARM_S:
        step(stream)
        n_warm += 1
        if n_warm % 20 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()
    log(f"warmed up with {n_warm} steps in {time.perf_counter() - t_warm:.2f} s")
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        step(stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    # one launch bracketed by events, averaged: the kernel's own duration
    reps = max(3, min(args.steps, 20))
    launch_ms = 0.0
    for _ in range(reps):
        e0.record(stream)
        step(stream)
        e1.record(stream)
        e1.synchronize()
        launch_ms += e0.elapsed_time(e1)
    launch_ms /= reps
    log(f"demod: {ms:.4f} ms per step, {launch_ms:.4f} ms per bracketed launch")
    copy_gbps = measured_copy_gbps(dev)
    fill_gbps = measured_fill_gbps(dev, DEMOD_ATOM_BYTES * atoms)
    log(f"device copy {copy_gbps:.0f} GB/s, fill of the output size {fill_gbps:.0f} GB/s")

    # parity: seeded templates against the FP64 oracle
    rng = np.random.default_rng(13)
    idx = np.sort(rng.choice(T, size=min(max(args.check_templates, 1) * 2, T), replace=False))
    ts_h = ts_d.cpu().numpy().view(np.uint32).reshape(T, n)
    cols_h = cols_d.cpu().numpy().reshape(NC, T, n)
    a2, b2, ab, Fa, Fb = dr.demod_atoms(sft.t_start, sft.bins.astype(np.complex128), sft.k_min,
                                        sft.Tsft, sft.delay, sft.delay_rate, sft.a, sft.b,
                                        freqs[idx], fdots[idx], sft.ref_time, dterms=dterms)
    a32, b32 = sft.a.astype(np.float32), sft.b.astype(np.float32)
    exact = bool(np.array_equal(ts_h[idx], np.broadcast_to(sft.t_start, (idx.size, n)))
                 and np.array_equal(cols_h[0, idx], np.broadcast_to(a32 * a32, (idx.size, n)))
                 and np.array_equal(cols_h[1, idx], np.broadcast_to(b32 * b32, (idx.size, n)))
                 and np.array_equal(cols_h[2, idx], np.broadcast_to(a32 * b32, (idx.size, n))))
    worst = 0.0
    for r, t in enumerate(idx):
        scale = max(np.abs(Fa[r]).max(), np.abs(Fb[r]).max())
        for c, ref in ((3, Fa[r].real), (4, Fa[r].imag), (5, Fb[r].real), (6, Fb[r].imag)):
            worst = max(worst, float(np.abs(cols_h[c, t] - ref).max() / scale))
    twoF_gpu = dr.twoF_all_data(*(cols_h[c, idx].astype(np.float64) for c in range(3)),
                                cols_h[3, idx] + 1j * cols_h[4, idx].astype(np.float64),
                                cols_h[5, idx] + 1j * cols_h[6, idx].astype(np.float64))
    twoF_o = dr.twoF_all_data(a2, b2, ab, Fa, Fb)
    parity = {
        "templates": int(idx.size),
        "t_a2_b2_ab_bit_exact": exact,
        "max_err_Fa_Fb_rel_to_template_max": worst,
        "max_rel_err_twoF_all_data": float(np.max(np.abs(twoF_gpu - twoF_o) / twoF_o)),
        "mean_twoF_all_data": float(twoF_o.mean()),
    }
    log(f"parity: {parity}")
    cpu = None
    if not args.no_cpu_baseline:
        cpu = demod_cpu_baseline(sft, freqs, fdots, dterms, args.cpu_seconds, rng)
        log(f"cpu: {cpu}")
    achieved = DEMOD_ATOM_BYTES * atoms / (launch_ms / 1e3) / 1e9
    traffic, traffic_src = None, None
    if args.traffic_kib:
        traffic = (args.traffic_kib[0] + args.traffic_kib[1]) * 1024.0
        traffic_src = "--traffic-kib"
    elif args.templates == 4096:
        traffic, traffic_src = PROFILED_TRAFFIC["demod_atoms"]
    out = {
        "metric": DEMOD_METRIC,
        "value": atoms / (ms / 1e3),
        "unit": "atoms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"demod_atoms: {WORKLOADS['demod_atoms']}",
            "templates": T,
            "sfts": n,
            "bins_per_sft": sfts.n_bins,
            "atoms": atoms,
            "dterms": dterms,
            "arithmetic": "fp64 phase and Doppler bin, fp32 kernel weights, sum and rotation",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "demod_atoms_kernel",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "copy_GBps": copy_gbps,
            "frac_of_copy": achieved / copy_gbps,
            "fill_GBps": fill_gbps,
            "frac_of_fill": achieved / fill_gbps,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "algorithmic_bytes_per_launch": DEMOD_ATOM_BYTES * atoms,
            "launch_ms": launch_ms,
            "limiter": "the store pattern: 256 contiguous bytes per wave store into 8 column "
                       "arrays x 64 template rows per workgroup writes at ~5.1-5.5 TB/s (a "
                       "contiguous fill: 6.9); the compute alone takes 0.111 ms; "
                       "profiles/demod_atoms_kernel.demod_atoms.md",
        },
        "cpu_baseline": cpu,
        "parity": parity,
    }
    print(json.dumps(out), flush=True)


def measured_fill_gbps(device, nbytes, reps=5):
    """Write-only bandwidth: torch fill of nbytes of device memory (HIP events
    on the current stream, best of reps) -- the ceiling of a kernel whose HBM
    traffic is its output."""
    import torch

    dst = torch.empty(int(nbytes) // 4, dtype=torch.float32, device=device)
    dst.fill_(1.0)
    stream = torch.cuda.current_stream()
    best = None
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        dst.fill_(2.0)
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    n = dst.numel() * 4
    del dst
    torch.cuda.empty_cache()
    return n / (best / 1e3) / 1e9


def measured_copy_gbps(device, reps=5):
    """Device-to-device copy bandwidth (1 GiB read + 1 GiB written per copy),
    HIP events on the current stream, best of reps -- the achievable HBM
    figure the map kernels are also compared with."""
    import torch

    src = torch.empty(2**28, dtype=torch.float32, device=device)
    dst = torch.empty_like(src)
    dst.copy_(src)
    stream = torch.cuda.current_stream()
    best = None
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        dst.copy_(src)
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    del src, dst
    torch.cuda.empty_cache()
    return 2.0 * 2**30 / (best / 1e3) / 1e9


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="rect_batched", choices=sorted(WORKLOADS))
    ap.add_argument("--templates", type=int, default=4096)
    ap.add_argument("--refine-steps", type=int, default=20)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in rows per work item instead of tuning them "
                         "on this batch before the timed steps")
    ap.add_argument("--rows", nargs="*", default=None, metavar="MODE:ROWS",
                    help="fixed rows per work item per kernel mode (e.g. 0:192 3:96); "
                         "implies --no-autotune")
    ap.add_argument("--no-btsg", action="store_true")
    ap.add_argument("--plain-templates", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--check-templates", type=int, default=6)
    ap.add_argument("--profile-mode", action="store_true",
                    help="for rocprofv3: a few ungraphed steps plus a 1 GiB device copy "
                         "(reference for counter scaling); no baselines")
    ap.add_argument("--traffic-kib", type=float, nargs=2, default=None,
                    metavar=("FETCH_KIB", "WRITE_KIB"),
                    help="rocprofv3 FETCH_SIZE/WRITE_SIZE of the map kernel (KiB per "
                         "dispatch, already scaled) to report as roofline.traffic")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(templates are independent: a multi-GPU run would shard them by cost).",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "exp_refine":
        run_refine(args)
        return
    if args.workload == "demod_atoms":
        run_demod(args)
        return

    import torch

    from pyfstat_amd import tcw_fstat_map_funcs as tcw
    from pyfstat_amd.batched import BatchedTransientMaps

    features, _ = tcw.init_transient_fstat_map_features("hip")
    log("generating workload ...")
    t_gen = time.perf_counter()
    work = make_workload(args)
    BtSG = not args.no_btsg
    use_graph = not args.no_graph
    b = build_batch(work, BtSG)
    batch = work if not isinstance(work, tuple) else None
    cells = b.total_cells
    st0 = None
    log(f"workload ready in {time.perf_counter() - t_gen:.1f} s: T={b.T}, cells={cells:.3e}")
    log(f"plan: {cells} cells, scratch {b.scratch_bytes() / 2**30:.2f} GiB")
    tuned = None
    if args.rows:
        from pyfstat_amd import autotune

        for item in args.rows:
            mode, rows = (int(x) for x in item.split(":"))
            autotune.set_rows(b, mode, rows)
        tuned = {item.split(":")[0]: int(item.split(":")[1]) for item in args.rows}
        log(f"rows per item by mode (fixed): {tuned}")
    elif not args.no_autotune and not args.profile_mode:
        from pyfstat_amd import autotune

        res = autotune.tune(b, reps=3, log=log)
        tuned = {str(m): r["rows"] for m, r in res.items()}
        log(f"autotune: rows per item by mode {tuned}")

    if args.profile_mode:
        for _ in range(args.warmup + args.steps):
            b.run(use_graph=False)
        torch.cuda.synchronize()
        ref = torch.empty(2**28, dtype=torch.float32, device=b.device)
        ref2 = torch.empty_like(ref)
        for _ in range(3):
            ref2.copy_(ref)  # 1 GiB read + 1 GiB write, the counters' yardstick
        torch.cuda.synchronize()
        log("profile mode done: map steps + 3 x 1 GiB copies")
        b.close()
        return

    for _ in range(args.warmup):
        b.run(use_graph=use_graph)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        b.run(use_graph=use_graph, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms_per_step = e0.elapsed_time(e1) / args.steps
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step (wall {wall:.3f} s)")

    # per-kernel durations: HIP events around each stage on the launch stream
    reps = max(3, min(args.steps, 10))
    stage_ms = np.zeros(3)
    for _ in range(reps):
        stage_ms += np.array(b.run_timed(stream=stream))
    stage_ms /= reps
    log(f"stage/maps/finalize ms: {stage_ms.round(4).tolist()}")

    atoms_binned = int(b.stats()["numAtoms"].sum())
    copy_gbps = measured_copy_gbps(b.device)
    log(f"1 GiB device copy: {copy_gbps:.0f} GB/s")
    units = Units(args, work, batch)
    n_check = max(args.check_templates, 32) if units.rows else args.check_templates
    parity = parity_sample(b, units, n_check)
    if parity:
        log(f"parity: {parity}")

    plain = None
    if args.plain_templates > 0 and batch is not None:
        rng = np.random.default_rng(9)
        idx = sorted(rng.choice(batch.T, size=min(args.plain_templates, batch.T),
                                replace=False).tolist())
        plain_loop(b, batch, idx[:8], BtSG)  # warm the single-map path
        pc, pt = plain_loop(b, batch, idx, BtSG)
        plain_rate = pc / pt
        plain = {
            "value": plain_rate,
            "unit": "cells/s",
            "templates": len(idx),
            "ms_per_template": 1e3 * pt / len(idx),
            "batched_over_plain": (cells / (ms_per_step / 1e3)) / plain_rate,
        }
        log(f"plain: {plain}")

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(units, args.cpu_seconds)
        log(f"cpu: {cpu}")

    value = cells / (ms_per_step / 1e3)
    map_bytes = 4.0 * cells
    achieved = map_bytes / (stage_ms[1] / 1e3) / 1e9
    traffic, traffic_src = None, None
    if args.traffic_kib:
        traffic = (args.traffic_kib[0] + args.traffic_kib[1]) * 1024.0
        traffic_src = "--traffic-kib"
    elif PROFILED_TRAFFIC.get(args.workload) and args.templates == 4096:
        traffic, traffic_src = PROFILED_TRAFFIC[args.workload]
    job_bytes = 4.0 * cells + 28.0 * atoms_binned
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {WORKLOADS[args.workload]}",
            "templates": b.T,
            "cells": cells,
            "binned_atoms": atoms_binned,
            "BtSG": BtSG,
            "hipgraph": use_graph,
            "rows_per_item_by_mode": tuned,
            "arithmetic": "fp64 prefix scan of float32 merged bins, fp64 window differences/determinant/"
                          "normalisation, fp32 map",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "map kernels of the step (rect_fast_kernel / exp_fast_kernel)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "copy_GBps": copy_gbps,
            "frac_of_copy": achieved / copy_gbps,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "algorithmic_bytes_per_launch": map_bytes,
            "launch_ms": stage_ms[1],
            "stage_ms": stage_ms[0],
            "finalize_ms": stage_ms[2],
            "job_GBps": job_bytes / (ms_per_step / 1e3) / 1e9,
            "limiter": LIMITER.get(args.workload),
        },
        "cpu_baseline": cpu,
        "plain": plain,
        "parity": parity,
    }
    print(json.dumps(out), flush=True)
    b.close()


if __name__ == "__main__":
    main()
