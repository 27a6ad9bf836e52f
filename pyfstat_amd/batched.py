"""Many transient F-statistic maps in one device pass.

TransientGridSearch evaluates one map per Doppler template in a Python loop
(pyfstat/grid_based_searches.py:1186-1189 -> pyfstat/core.py:1451-1457), each
call paying host packing, H2D, a kernel launch and a full D2H.  Here the
templates' atoms stay resident on the device, a plan (shapes, output offsets,
work items) is built once, and one launch sequence -- optionally captured in a
hipGraph -- produces all maps and their per-map statistics.
"""

import ctypes

import numpy as np

from . import _hiplib
from .atoms import atoms_to_arrays, window_fields

STATS_DTYPE = np.dtype(_hiplib.Stats)


def _torch():
    import torch

    return torch


class BatchedTransientMaps:
    """Device-resident batch of templates with a reusable plan.

    Parameters
    ----------
    ts : uint32 [total] timestamps of every detector segment, concatenated.
    cols : float32 [7, total] atom columns (a2, b2, ab, Fa_re, Fa_im, Fb_re, Fb_im).
    seg_off : int64 [S+1] atom offsets of the detector segments.
    seg_begin : int64 [T+1] segment range of each template.
    windows : T window ranges (ours or lalpulsar's).
    TAtom : common atom length in seconds.
    """

    def __init__(self, ts, cols, seg_off, seg_begin, windows, TAtom, BtSG=False,
                 lal_strict=False, device=None, force_generic=False, rows=None,
                 seg_tfirst=None, seg_tlast=None):
        from .tcw_fstat_map_funcs import _HipState, _ensure_handle

        torch = _torch()
        self.lib = _hiplib.load()
        self.handle = _ensure_handle()
        self.device = torch.device("cuda", _HipState.device if device is None else device)
        on_device = isinstance(ts, torch.Tensor)
        self.seg_off = np.ascontiguousarray(seg_off, dtype=np.int64)
        self.seg_begin = np.ascontiguousarray(seg_begin, dtype=np.int64)
        self.T = len(self.seg_begin) - 1
        if len(windows) != self.T:
            raise ValueError(f"{len(windows)} windows for {self.T} templates")
        self.windows = list(windows)
        S = len(self.seg_off) - 1
        if on_device:
            # atoms already on the device (ts as int32 bits, cols [7, n]): the
            # caller gives each segment's first and last timestamp and
            # guarantees ascending timestamps within segments
            if seg_tfirst is None or seg_tlast is None:
                raise ValueError("device atoms need seg_tfirst and seg_tlast")
            if ts.dtype != torch.int32 or cols.dtype != torch.float32 or cols.dim() != 2 \
                    or cols.shape[0] != 7 or not cols.is_contiguous() or not ts.is_contiguous():
                raise ValueError("device atoms: ts int32 [n] and contiguous cols float32 [7, n]")
            if ts.device != self.device or cols.device != self.device:
                raise ValueError(f"device atoms must be on {self.device}")
            if self.seg_off[-1] > min(ts.numel(), cols.shape[1]):
                raise ValueError("seg_off runs past the device atoms")
            self.seg_tfirst = np.ascontiguousarray(seg_tfirst, dtype=np.uint32)
            self.seg_tlast = np.ascontiguousarray(seg_tlast, dtype=np.uint32)
            if self.seg_tfirst.size != S or self.seg_tlast.size != S:
                raise ValueError("seg_tfirst / seg_tlast need one entry per segment")
        else:
            ts = np.ascontiguousarray(ts, dtype=np.uint32)
            cols = np.ascontiguousarray(cols, dtype=np.float32)
            self.seg_tfirst = ts[self.seg_off[:-1]].copy()
            self.seg_tlast = ts[self.seg_off[1:] - 1].copy()
            for s in range(S):
                seg = ts[self.seg_off[s]:self.seg_off[s + 1]]
                if seg.size > 1 and np.any(seg[1:] < seg[:-1]):
                    raise ValueError(f"segment {s}: timestamps are not in ascending order")
        self.win_arr = (_hiplib.Window * self.T)(
            *[_hiplib.window_struct(window_fields(w)) for w in self.windows])
        self.flags = 0
        if BtSG:
            self.flags |= _hiplib.TCW_FLAG_BTSG
        if lal_strict:
            self.flags |= _hiplib.TCW_FLAG_LAL_STRICT
        if force_generic:
            self.flags |= _hiplib.TCW_FLAG_GENERIC
        self.BtSG = bool(BtSG)
        self.lal_strict = bool(lal_strict)
        self.TAtom = int(TAtom)
        desc = _hiplib.BatchDesc()
        desc.T = self.T
        desc.TAtom = int(TAtom)
        desc.seg_begin = self.seg_begin.ctypes.data
        desc.seg_off = self.seg_off.ctypes.data
        desc.seg_tfirst = self.seg_tfirst.ctypes.data
        desc.seg_tlast = self.seg_tlast.ctypes.data
        desc.windows = ctypes.addressof(self.win_arr)
        self._desc = desc
        self.F_off = np.zeros(self.T + 1, dtype=np.int64)
        # rows (t0) per work item per kernel mode, 0 = the library's choice;
        # a setting of this batch's plan only
        self.rows_by_mode = np.zeros(_hiplib.MODE_COUNT, dtype=np.int32)
        for mode, r in (rows or {}).items():
            self.rows_by_mode[int(mode)] = int(r)
        self.plan = None
        self._make_plan()
        if on_device:
            self.ts_d = ts
            self.cols_d = cols
        else:
            self.ts_d = torch.from_numpy(ts.view(np.int32)).to(self.device)
            self.cols_d = torch.from_numpy(cols).to(self.device)
        self.col_stride = cols.shape[1]
        self.F_d = torch.empty(int(self.F_off[-1]), dtype=torch.float32, device=self.device)
        self.stats_d = torch.empty(self.T * STATS_DTYPE.itemsize, dtype=torch.uint8,
                                   device=self.device)

    def _make_plan(self):
        plan = ctypes.c_void_p()
        _hiplib.check(self.lib.tcw_batch_plan_rows(self.handle, ctypes.byref(self._desc),
                                                   self.flags, self.rows_by_mode.ctypes.data,
                                                   ctypes.byref(plan), self.F_off.ctypes.data),
                      "tcw_batch_plan_rows")
        self.plan = plan

    def set_rows(self, mode, rows):
        """Rows per work item for one kernel mode (0: the library's choice);
        takes effect at the next ``replan``."""
        self.rows_by_mode[int(mode)] = int(rows)

    def replan(self):
        """Rebuild the plan (work items, scratch, graph) with the current
        rows-per-item settings; atoms and outputs stay where they are.  A
        setting the library rejects raises ValueError and keeps no plan."""
        self.close()
        self._make_plan()

    def items_per_mode(self):
        """Work items per kernel mode (rect regular, rect generic, exp direct,
        exp on the atom grid)."""
        out = np.zeros(_hiplib.MODE_COUNT, dtype=np.int64)
        _hiplib.check(self.lib.tcw_plan_items(self.plan, out.ctypes.data), "tcw_plan_items")
        return out

    # ------------------------------------------------------------------
    @classmethod
    def from_device_atoms(cls, atoms, windows, **kw):
        """From atoms produced on the device (``demod.DeviceAtoms``): no host copy."""
        return cls(atoms.ts_d, atoms.cols_d, atoms.seg_off, atoms.seg_begin, windows,
                   atoms.TAtom, seg_tfirst=atoms.seg_tfirst, seg_tlast=atoms.seg_tlast, **kw)

    @classmethod
    def from_multi_atoms(cls, multi_list, windows, **kw):
        """From a list of MultiFstatAtomVector (one per template)."""
        ts, cols, seg_off, seg_begin = [], [], [0], [0]
        TAtom = None
        for multi in multi_list:
            packed = atoms_to_arrays(multi)
            if TAtom is None:
                TAtom = packed.TAtom
            elif packed.TAtom != TAtom:
                raise ValueError("all templates must share one TAtom")
            base = seg_off[-1]
            for X in range(packed.numDet):
                seg_off.append(base + int(packed.det_off[X + 1]))
            seg_begin.append(seg_begin[-1] + packed.numDet)
            ts.append(packed.ts)
            cols.append(packed.cols)
        return cls(np.concatenate(ts), np.concatenate(cols, axis=1), seg_off, seg_begin,
                   windows, TAtom, **kw)

    @classmethod
    def from_synthetic(cls, batch, **kw):
        """From a pyfstat_amd.synth.SyntheticBatch (one detector per template)."""
        seg_begin = np.arange(batch.T + 1, dtype=np.int64)
        return cls(batch.ts, batch.cols, batch.off, seg_begin, batch.windows, batch.TAtom, **kw)

    # ------------------------------------------------------------------
    @property
    def total_cells(self):
        return int(self.F_off[-1])

    def scratch_bytes(self):
        return int(self.lib.tcw_plan_scratch_bytes(self.plan))

    def run(self, use_graph=False, stream=None):
        """Launch all maps (asynchronous on the stream)."""
        torch = _torch()
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        _hiplib.check(self.lib.tcw_batch_run(
            self.handle, self.plan, ctypes.c_void_p(self.ts_d.data_ptr()),
            ctypes.c_void_p(self.cols_d.data_ptr()), self.col_stride,
            ctypes.c_void_p(self.F_d.data_ptr()), ctypes.c_void_p(self.stats_d.data_ptr()),
            1 if use_graph else 0, ctypes.c_void_p(stream.cuda_stream)), "tcw_batch_run")

    def run_timed(self, stream=None):
        """Run once without a graph; returns (stage_ms, maps_ms, finalize_ms)
        from HIP events recorded on the launch stream (synchronises)."""
        torch = _torch()
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        ms = (ctypes.c_double * 3)()
        _hiplib.check(self.lib.tcw_batch_run_timed(
            self.handle, self.plan, ctypes.c_void_p(self.ts_d.data_ptr()),
            ctypes.c_void_p(self.cols_d.data_ptr()), self.col_stride,
            ctypes.c_void_p(self.F_d.data_ptr()), ctypes.c_void_p(self.stats_d.data_ptr()),
            ms, ctypes.c_void_p(stream.cuda_stream)), "tcw_batch_run_timed")
        return tuple(ms)

    def stats(self):
        """Per-template statistics as a numpy structured array (synchronises)."""
        raw = self.stats_d.cpu().numpy()
        st = raw.view(STATS_DTYPE).copy()
        if self.lal_strict and np.any(st["single_atom"]):
            bad = int(np.nonzero(st["single_atom"])[0][0])
            raise _hiplib.SingleAtomError(
                f"template {bad}: encountered a single-atom Fstat-calculation. This is "
                "degenerate and cannot be computed!")
        return st

    def F_host(self):
        return self.F_d.cpu().numpy()

    def map_view(self, F_all, t):
        st_shape = self._shape(t)
        return F_all[self.F_off[t]:self.F_off[t + 1]].reshape(st_shape)

    def _shape(self, t):
        w = self.windows[t]
        if w.type == 0:
            return (1, 1)
        return (w.t0Band // w.dt0 + 1, w.tauBand // w.dtau + 1)

    def maps(self):
        """List of pyTransientFstatMap, F_mn as views of one host copy."""
        from .tcw_fstat_map_funcs import _map_from_stats

        st = self.stats()
        F_all = self.F_host()
        out = []
        for t in range(self.T):
            s = st[t]
            F_mn = F_all[self.F_off[t]:self.F_off[t + 1]].reshape(int(s["N_t0"]), int(s["N_tau"]))
            out.append(_map_from_stats(F_mn, _StatsRow(s), self.BtSG))
        return out

    def close(self):
        if getattr(self, "plan", None):
            self.lib.tcw_plan_destroy(self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _StatsRow:
    """Attribute access to one row of the stats structured array."""

    def __init__(self, row):
        self._row = row

    def __getattr__(self, name):
        return self._row[name]


def hip_compute_transient_fstat_maps(multiFstatAtoms_list, windowRanges, BtSG=False,
                                     lal_strict=False, use_graph=False):
    """Batched counterpart of hip_compute_transient_fstat_map: one
    pyTransientFstatMap per (atoms, windowRange) pair, all computed in one
    device pass."""
    _hiplib.load()  # fail loudly without the library, also for an empty batch
    multiFstatAtoms_list = list(multiFstatAtoms_list)
    if not isinstance(windowRanges, (list, tuple)):
        windowRanges = [windowRanges] * len(multiFstatAtoms_list)
    if len(windowRanges) != len(multiFstatAtoms_list):
        raise ValueError(f"{len(multiFstatAtoms_list)} atom sets but {len(windowRanges)} "
                         "window ranges")
    if not multiFstatAtoms_list:
        return []
    b = BatchedTransientMaps.from_multi_atoms(multiFstatAtoms_list, windowRanges, BtSG=BtSG,
                                              lal_strict=lal_strict)
    try:
        b.run(use_graph=use_graph)
        return b.maps()
    finally:
        b.close()
