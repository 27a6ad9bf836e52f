"""Differentiable exponential-window candidates: 2F and lnBtSG of continuous
(t0, tau) on a fixed atom support, with a hand-written backward, and a fused
gradient ascent that refines grid maxima to off-grid tau.

The reference evaluates candidates one 1x1 map at a time (MCMCTransientSearch,
pyfstat/mcmc_based_searches.py:3511-3516 -> pyfstat/core.py:1412-1471) and has
no derivatives.  Here every candidate of every template is evaluated in one
launch (libtcw_hip: tcw_refine_*), and ``ExpRefine.apply`` is a
``torch.autograd.Function`` whose backward is a HIP kernel, giving d(2F) and
d(lnBtSG) with respect to the merged atoms (all seven columns), t0 and tau.
``MergedAtoms.merge`` carries the atom gradients back through the binned
detector merge (each atom receives its bin's gradient).

Support.  A candidate at grid cell (t0, tau) uses the bins the map kernel
weights there: t0 <= t_k <= t0 + 3 tau within the index range i_t0 .. i_t1 of
pyfstat/pyCUDAkernels/cudaTransientFstatExpWindow.cu:27-45 (see
``exp_window_support``).  The support is then held fixed while (t0, tau)
move.  On a fixed support all weights share the factor exp(t0/tau), under
which F is invariant, so dF/dt0 = 0 identically; F depends on t0 only through
which bins the support holds, i.e. not between grid points.  Refinement
therefore moves tau; t0 keeps its grid value.
"""

import ctypes
import math

import numpy as np

from . import _hiplib

LN70 = math.log(70.0)


def _torch():
    import torch

    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(torch, device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


# ----------------------------------------------------------------------------
# support of a grid cell
# ----------------------------------------------------------------------------


def exp_window_support(t0, tau, t_data, N, TAtom):
    """Fixed support (k_lo, k_hi) of exponential-window cells, vectorised.

    Bins k with i_t0 <= k <= i_t1 and t0 <= t_k <= t1 = t0 + 3 tau, where
    t_k = t_data + k*TAtom, i_t0 = (t0 - t_data + TAtom/2)/TAtom and
    i_t1 = (t1 - t_data + TAtom/2)/TAtom - 1, both clamped to [0, N-1] -- the
    bins with non-zero weight in the exponential-window kernel
    (cudaTransientFstatExpWindow.cu:27-45, 82-102).  Integer seconds;
    requires t_data <= t0 and t1 < 2^32 (no uint32 wrap-around).  Raises
    ValueError for an empty support.
    """
    t0 = np.asarray(t0, dtype=np.int64)
    tau = np.asarray(tau, dtype=np.int64)
    t_data = np.asarray(t_data, dtype=np.int64)
    N = np.asarray(N, dtype=np.int64)
    TA = int(TAtom)
    t1 = t0 + 3 * tau
    if np.any(t0 < t_data) or np.any(t1 >= 2**32) or np.any(tau <= 0):
        raise ValueError("candidate windows must satisfy t_data <= t0, tau > 0, t0 + 3 tau < 2^32")
    half = TA // 2
    i_t0 = np.minimum((t0 - t_data + half) // TA, N - 1)
    i_t1 = np.clip((t1 - t_data + half) // TA - 1, 0, N - 1)
    k_lo = np.maximum(i_t0, -((t_data - t0) // TA))  # ceil((t0 - t_data)/TA)
    k_hi = np.minimum(i_t1, (t1 - t_data) // TA)
    if np.any(k_hi < k_lo):
        bad = int(np.nonzero(k_hi < k_lo)[0][0])
        raise ValueError(f"candidate {bad}: empty window support")
    return k_lo.astype(np.int32), k_hi.astype(np.int32)


# ----------------------------------------------------------------------------
# merged atoms
# ----------------------------------------------------------------------------


class MergedAtoms:
    """Binned detector merge of many templates on the device
    (lalpulsar.mergeMultiFstatAtomsBinned, pyfstat/tcw_fstat_map_funcs.py:706,
    per template), with the atom-to-bin map kept for the adjoint.

    ``cols`` is float32 [7, total_bins]; template t owns bins
    ``m_off[t]:m_off[t+1]`` and bin k of it starts at ``t_data[t] + k*TAtom``.
    """

    def __init__(self, ts_d, seg_off, seg_begin, TAtom, n_atoms, device, cols_raw=None):
        torch = _torch()
        from .tcw_fstat_map_funcs import _ensure_handle

        self.lib = _hiplib.load()
        self.handle = _ensure_handle()
        self.device = device
        self.TAtom = int(TAtom)
        self.seg_off = np.ascontiguousarray(seg_off, dtype=np.int64)
        self.seg_begin = np.ascontiguousarray(seg_begin, dtype=np.int64)
        self.T = len(self.seg_begin) - 1
        self.n_atoms = int(n_atoms)
        self.ts_d = ts_d
        ts_h = ts_d.cpu().numpy().view(np.uint32)
        self.seg_tfirst = ts_h[self.seg_off[:-1]].copy()
        self.seg_tlast = ts_h[self.seg_off[1:] - 1].copy()
        self.desc = _hiplib.BatchDesc()
        self.desc.T = self.T
        self.desc.TAtom = self.TAtom
        self.desc.seg_begin = self.seg_begin.ctypes.data
        self.desc.seg_off = self.seg_off.ctypes.data
        self.desc.seg_tfirst = self.seg_tfirst.ctypes.data
        self.desc.seg_tlast = self.seg_tlast.ctypes.data
        self.desc.windows = None
        self.m_off = np.zeros(self.T + 1, dtype=np.int64)
        self.t_data = np.zeros(self.T, dtype=np.uint32)
        _hiplib.check(self.lib.tcw_merge_batch(
            self.handle, ctypes.byref(self.desc), None, None, 0, None, 0,
            self.m_off.ctypes.data, self.t_data.ctypes.data, None, None), "tcw_merge_batch")
        self.total_bins = int(self.m_off[-1])
        self.N = np.diff(self.m_off)
        self.atom_bin = torch.empty(self.n_atoms, dtype=torch.int64, device=device)
        self.cols = None
        if cols_raw is not None:
            self.cols = self.merge(cols_raw)

    @classmethod
    def from_batch(cls, batch):
        """Merged atoms of a BatchedTransientMaps' device-resident templates."""
        return cls(batch.ts_d, batch.seg_off, batch.seg_begin, batch.TAtom,
                   batch.col_stride, batch.device, cols_raw=batch.cols_d)

    def merge(self, cols_raw):
        """Merged float32 [7, total_bins] of raw atoms [7, n_atoms]
        (differentiable: the backward gathers each atom's bin gradient)."""
        return _MergeFunction.apply(cols_raw, self)

    def _merge_forward(self, cols_raw):
        torch = _torch()
        if cols_raw.dtype != torch.float32 or cols_raw.dim() != 2 or cols_raw.shape[0] != 7:
            raise ValueError("raw atoms must be float32 [7, n_atoms]")
        if cols_raw.shape[1] != self.n_atoms:
            raise ValueError(f"expected {self.n_atoms} atoms, got {cols_raw.shape[1]}")
        cols_raw = cols_raw.contiguous()
        out = torch.empty((7, self.total_bins), dtype=torch.float32, device=self.device)
        _hiplib.check(self.lib.tcw_merge_batch(
            self.handle, ctypes.byref(self.desc), _ptr(self.ts_d), _ptr(cols_raw),
            cols_raw.shape[1], _ptr(out), self.total_bins, self.m_off.ctypes.data,
            self.t_data.ctypes.data, _ptr(self.atom_bin), _stream(torch, self.device)),
            "tcw_merge_batch")
        return out

    def _merge_backward(self, grad_merged):
        torch = _torch()
        g = grad_merged.contiguous()
        out = torch.empty((7, self.n_atoms), dtype=torch.float32, device=self.device)
        _hiplib.check(self.lib.tcw_gather_bins(
            _ptr(g), g.shape[1], _ptr(self.atom_bin), self.n_atoms, _ptr(out), self.n_atoms,
            _stream(torch, self.device)), "tcw_gather_bins")
        return out


def _merge_function():
    torch = _torch()

    class MergeFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, cols_raw, merged):
            ctx.merged = merged
            return merged._merge_forward(cols_raw)

        @staticmethod
        def backward(ctx, grad):
            return ctx.merged._merge_backward(grad.float()), None

    return MergeFunction


class _Lazy:
    """Build the torch.autograd.Function subclasses on first use (importing
    this module does not import torch)."""

    def __init__(self, factory):
        self.factory = factory
        self.cls = None

    def apply(self, *args):
        if self.cls is None:
            self.cls = self.factory()
        return self.cls.apply(*args)


_MergeFunction = _Lazy(_merge_function)


# ----------------------------------------------------------------------------
# candidates
# ----------------------------------------------------------------------------


class ExpRefineProblem:
    """Exponential-window candidates of many templates with fixed supports.

    Parameters
    ----------
    merged : MergedAtoms
    cand_tmpl : int [n] template of each candidate.
    t0, tau : [n] the cells the supports are taken from (integer seconds), or
        pass ``k_lo``/``k_hi`` explicitly.

    Candidates are kept sorted by template and, within a template, by support
    length (stable); ``order[i]`` is the caller's index of candidate i, and
    every array this class takes or returns is in that sorted order.
    """

    def __init__(self, merged, cand_tmpl, t0, tau, k_lo=None, k_hi=None):
        torch = _torch()
        self.merged = merged
        self.lib = merged.lib
        self.device = merged.device
        cand_tmpl = np.asarray(cand_tmpl, dtype=np.int64)
        if cand_tmpl.size and (cand_tmpl.min() < 0 or cand_tmpl.max() >= merged.T):
            raise ValueError("candidate template index out of range")
        t0 = np.asarray(t0, dtype=np.float64)
        tau = np.asarray(tau, dtype=np.float64)
        if k_lo is None:
            k_lo, k_hi = exp_window_support(np.rint(t0), np.rint(tau),
                                            merged.t_data[cand_tmpl].astype(np.int64),
                                            merged.N[cand_tmpl], merged.TAtom)
        k_lo = np.asarray(k_lo, dtype=np.int32)
        k_hi = np.asarray(k_hi, dtype=np.int32)
        # by template, then by support length: the 16-lane groups of a wave
        # then walk supports of similar length
        self.order = np.lexsort((k_hi.astype(np.int64) - k_lo, cand_tmpl))
        self.cand_tmpl = cand_tmpl[self.order]
        t0 = t0[self.order]
        tau = tau[self.order]
        self.k_lo = np.ascontiguousarray(k_lo[self.order], dtype=np.int32)
        self.k_hi = np.ascontiguousarray(k_hi[self.order], dtype=np.int32)
        self.n = int(self.cand_tmpl.size)
        self.cand_off = np.zeros(merged.T + 1, dtype=np.int64)
        np.add.at(self.cand_off, self.cand_tmpl + 1, 1)
        self.cand_off = np.cumsum(self.cand_off)
        self.t0_init = torch.from_numpy(t0).to(self.device)
        self.tau_init = torch.from_numpy(tau).to(self.device)
        desc = _hiplib.RefineDesc()
        desc.T = merged.T
        desc.n_cand = self.n
        desc.TAtom = merged.TAtom
        desc.m_off = merged.m_off.ctypes.data
        desc.t_data = merged.t_data.ctypes.data
        desc.cand_off = self.cand_off.ctypes.data
        desc.k_lo = self.k_lo.ctypes.data
        desc.k_hi = self.k_hi.ctypes.data
        plan = ctypes.c_void_p()
        _hiplib.check(self.lib.tcw_refine_plan_create(merged.handle, ctypes.byref(desc),
                                                      ctypes.byref(plan)),
                      "tcw_refine_plan_create")
        self.plan = plan

    @property
    def support_lengths(self):
        return (self.k_hi.astype(np.int64) - self.k_lo + 1)

    # -- raw kernel calls ---------------------------------------------------
    def _check_inputs(self, cols, t0, tau):
        torch = _torch()
        if cols.dtype != torch.float32 or cols.dim() != 2 or cols.shape[0] != 7:
            raise ValueError("cols must be float32 [7, total_bins]")
        if cols.shape[1] < self.merged.total_bins or not cols.is_contiguous():
            raise ValueError("cols must be contiguous with at least total_bins columns")
        for name, x in (("t0", t0), ("tau", tau)):
            if x.dtype != torch.float64 or x.shape != (self.n,) or not x.is_contiguous():
                raise ValueError(f"{name} must be contiguous float64 [{self.n}]")

    def forward_raw(self, cols, t0, tau, want_saved=True):
        """(twoF, lnBtSG, saved) without autograd; saved is [n, 10] float64
        (dF/dt0, dF/dtau, dF/dS for the 7 sums, F)."""
        torch = _torch()
        self._check_inputs(cols, t0, tau)
        twoF = torch.empty(self.n, dtype=torch.float64, device=self.device)
        lnB = torch.empty(self.n, dtype=torch.float64, device=self.device)
        saved = (torch.empty((self.n, _hiplib.TCW_REFINE_SAVED), dtype=torch.float64,
                             device=self.device) if want_saved else None)
        _hiplib.check(self.lib.tcw_refine_forward(
            self.plan, _ptr(cols), cols.shape[1], _ptr(t0), _ptr(tau), _ptr(twoF), _ptr(lnB),
            _ptr(saved), _stream(torch, self.device)), "tcw_refine_forward")
        return twoF, lnB, saved

    def backward_raw(self, t0, tau, saved, g_twoF, g_lnB, want_cols=True, want_t0=True,
                     want_tau=True, cols_shape=None):
        torch = _torch()
        ncols = cols_shape[1] if cols_shape is not None else self.merged.total_bins
        gc = None
        if want_cols:
            # the kernel writes every merged bin; only columns past them need zeros
            gc = torch.empty((7, ncols), dtype=torch.float32, device=self.device)
            if ncols > self.merged.total_bins:
                gc[:, self.merged.total_bins:].zero_()
        gt0 = torch.empty(self.n, dtype=torch.float64, device=self.device) if want_t0 else None
        gtau = torch.empty(self.n, dtype=torch.float64, device=self.device) if want_tau else None
        g2 = g_twoF.to(torch.float64).contiguous() if g_twoF is not None else None
        gl = g_lnB.to(torch.float64).contiguous() if g_lnB is not None else None
        _hiplib.check(self.lib.tcw_refine_backward(
            self.plan, _ptr(t0), _ptr(tau), _ptr(saved), _ptr(g2), _ptr(gl), _ptr(gc), ncols,
            _ptr(gt0), _ptr(gtau), _stream(torch, self.device)), "tcw_refine_backward")
        return gc, gt0, gtau

    # -- public -------------------------------------------------------------
    def __call__(self, cols, t0, tau):
        """(twoF, lnBtSG), differentiable in cols, t0 and tau."""
        return ExpRefine.apply(cols, t0, tau, self)

    def ascent(self, cols, tau=None, tau_lo=None, tau_hi=None, n_steps=20, step0=0.02):
        """n_steps of gradient ascent of F in log(tau) per candidate (all steps
        of a candidate in one kernel).  Returns (tau, twoF, n_accept) as new
        tensors; t0 stays at ``t0_init`` and the tau passed in is not changed."""
        torch = _torch()
        tau_in = self.tau_init if tau is None else tau
        self._check_inputs(cols, self.t0_init, tau_in)
        tau = torch.empty_like(tau_in)
        twoF = torch.empty(self.n, dtype=torch.float64, device=self.device)
        nacc = torch.empty(self.n, dtype=torch.int32, device=self.device)
        lo = tau_lo.to(torch.float64).contiguous() if tau_lo is not None else None
        hi = tau_hi.to(torch.float64).contiguous() if tau_hi is not None else None
        _hiplib.check(self.lib.tcw_refine_ascent(
            self.plan, _ptr(cols), cols.shape[1], _ptr(self.t0_init), _ptr(tau_in), _ptr(tau),
            _ptr(lo), _ptr(hi), int(n_steps), float(step0), _ptr(twoF), _ptr(nacc),
            _stream(torch, self.device)), "tcw_refine_ascent")
        return tau, twoF, nacc

    def close(self):
        if getattr(self, "plan", None):
            self.lib.tcw_refine_plan_destroy(self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _refine_function():
    torch = _torch()

    class ExpRefineFunction(torch.autograd.Function):
        """(cols, t0, tau) -> (2F, lnBtSG) per candidate; backward by
        tcw_refine_backward."""

        @staticmethod
        def forward(ctx, cols, t0, tau, problem):
            twoF, lnB, saved = problem.forward_raw(cols, t0, tau)
            ctx.problem = problem
            ctx.cols_shape = tuple(cols.shape)
            ctx.save_for_backward(t0, tau, saved)
            return twoF, lnB

        @staticmethod
        def backward(ctx, g_twoF, g_lnB):
            t0, tau, saved = ctx.saved_tensors
            need = ctx.needs_input_grad
            gc, gt0, gtau = ctx.problem.backward_raw(
                t0, tau, saved, g_twoF, g_lnB, want_cols=need[0], want_t0=need[1],
                want_tau=need[2], cols_shape=ctx.cols_shape)
            return gc, gt0, gtau, None

    return ExpRefineFunction


ExpRefine = _Lazy(_refine_function)


# ----------------------------------------------------------------------------
# candidates from grid maps
# ----------------------------------------------------------------------------


def candidates_from_maps(batch, per_template=16):
    """Grid maxima of batched exponential-window maps as refinement candidates.

    For each template, the ``per_template`` t0 rows with the largest row
    maximum give one candidate each, at that row's maximising tau.  Returns
    numpy arrays (cand_tmpl, t0, tau, F_grid) with t0/tau in integer seconds
    and F_grid the map value (F, not 2F).
    """
    torch = _torch()
    cand_tmpl, t0s, taus, Fs = [], [], [], []
    shapes = [batch._shape(t) for t in range(batch.T)]
    uniform = len(set(shapes)) == 1
    if uniform:
        nt0, ntau = shapes[0]
        k = min(per_template, nt0)
        F = batch.F_d[: batch.T * nt0 * ntau].view(batch.T, nt0, ntau)
        rmax, rarg = F.max(dim=2)
        top, rows = rmax.topk(k, dim=1)
        cols_ = torch.gather(rarg, 1, rows)
        rows = rows.cpu().numpy()
        cols_ = cols_.cpu().numpy()
        top = top.cpu().numpy()
        for t in range(batch.T):
            w = batch.windows[t]
            cand_tmpl.append(np.full(k, t, dtype=np.int64))
            t0s.append(w.t0 + rows[t].astype(np.int64) * w.dt0)
            taus.append(w.tau + cols_[t].astype(np.int64) * w.dtau)
            Fs.append(top[t])
    else:
        for t in range(batch.T):
            nt0, ntau = shapes[t]
            k = min(per_template, nt0)
            F = batch.F_d[batch.F_off[t]:batch.F_off[t + 1]].view(nt0, ntau)
            rmax, rarg = F.max(dim=1)
            top, rows = rmax.topk(k)
            w = batch.windows[t]
            rows_h = rows.cpu().numpy().astype(np.int64)
            cand_tmpl.append(np.full(k, t, dtype=np.int64))
            t0s.append(w.t0 + rows_h * w.dt0)
            taus.append(w.tau + rarg[rows].cpu().numpy().astype(np.int64) * w.dtau)
            Fs.append(top.cpu().numpy())
    return (np.concatenate(cand_tmpl), np.concatenate(t0s), np.concatenate(taus),
            np.concatenate(Fs).astype(np.float64))
