"""F-statistic atoms from SFTs on the device, feeding the batched maps.

The reference obtains atoms from ``lalpulsar.ComputeFstat(...,
FSTATQ_ATOMS_PER_DET)`` once per Doppler point on the host
(``pyfstat/core.py:671-673``, ``:1359-1365``).  Here one call of
``tcw_demod_atoms`` produces the atoms of many (f, fdot) templates at one sky
position from the SFT bins resident on the device, in the layout
``tcw_batch_run`` reads (``tcw_demod_atoms_sky``: several sky positions in one
launch, :func:`demod_atoms_sky`); :func:`batch_from_sfts` chains the two so that a
search goes from SFTs to transient maps without a host round trip.

The sum is the resampling-free demodulation F-statistic (see
``include/tcw_hip.h`` for the formula): the caller provides SFT bins
normalised to unit noise variance, and per SFT the SSB delay and its rate at
the SFT midpoint and the antenna patterns a, b for the sky position (in a
lalsuite environment: ``lalpulsar.GetDetectorStates`` and
``lalpulsar.ComputeAMCoeffs``, once per sky point).
"""

import ctypes

import numpy as np

from . import _hiplib

NC = 7


def _torch():
    import torch

    return torch


class SFTSet:
    """Normalised SFT bins of one sky position, detectors concatenated.

    Parameters
    ----------
    t_start : uint32 [n_sft] SFT start times (GPS s), ascending per detector.
    bins : complex [n_sft, n_bins] bins k_min .. k_min+n_bins-1, divided by
        sqrt(S Tsft / 2) (unit noise variance).
    k_min : int, frequency of stored bin j is (k_min + j) / Tsft.
    Tsft : int, SFT length (s).
    delay, delay_rate : float [n_sft] SSB minus detector time at each SFT
        midpoint (s) and its time derivative.
    a, b : float [n_sft] antenna patterns at the midpoints.
    det_off : int [n_det+1] SFT ranges of the detectors.
    """

    def __init__(self, t_start, bins, k_min, Tsft, delay, delay_rate, a, b, det_off,
                 device=None):
        torch = _torch()
        from .tcw_fstat_map_funcs import _HipState, _ensure_handle

        self.lib = _hiplib.load()
        self.handle = _ensure_handle()
        self.device = torch.device("cuda", _HipState.device if device is None else device)
        t = np.ascontiguousarray(t_start, dtype=np.uint32)
        bins = np.asarray(bins)
        if bins.ndim != 2 or bins.shape[0] != t.size:
            raise ValueError("bins must be [n_sft, n_bins] with one row per SFT")
        self.det_off = np.ascontiguousarray(det_off, dtype=np.int64)
        if self.det_off.size < 2 or self.det_off[0] != 0 or self.det_off[-1] != t.size or np.any(
                np.diff(self.det_off) < 1):
            raise ValueError("det_off must run from 0 to n_sft with no empty detector")
        for X in range(self.det_off.size - 1):
            s = t[self.det_off[X]:self.det_off[X + 1]]
            if np.any(s[1:] <= s[:-1]):
                raise ValueError(f"detector {X}: SFT start times are not ascending")
        if int(Tsft) < 2:
            raise ValueError("Tsft must be at least 2 seconds.")
        if int(k_min) < 0 or int(k_min) + bins.shape[1] > 2**31 - 1:
            raise ValueError("SFT band must lie in bins 0 .. 2^31-1")
        self.t_start = t
        self.n_sft = int(t.size)
        self.n_bins = int(bins.shape[1])
        self.k_min = int(k_min)
        self.Tsft = int(Tsft)
        pair = np.empty((self.n_sft, self.n_bins, 2), dtype=np.float32)
        pair[..., 0] = bins.real
        pair[..., 1] = bins.imag
        dev = self.device
        self.t_d = torch.from_numpy(t.view(np.int32)).to(dev)
        self.bins_d = torch.from_numpy(pair).to(dev)
        self.delay_d = torch.from_numpy(np.ascontiguousarray(delay, dtype=np.float64)).to(dev)
        self.rate_d = torch.from_numpy(np.ascontiguousarray(delay_rate, dtype=np.float64)).to(dev)
        self.a_d = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)
        self.b_d = torch.from_numpy(np.ascontiguousarray(b, dtype=np.float32)).to(dev)
        for name, x in (("delay", self.delay_d), ("delay_rate", self.rate_d), ("a", self.a_d),
                        ("b", self.b_d)):
            if x.numel() != self.n_sft:
                raise ValueError(f"{name} must have one value per SFT")
        self.desc = _hiplib.SftSet(
            self.n_sft, self.t_d.data_ptr(), self.bins_d.data_ptr(), self.n_bins, self.k_min,
            self.Tsft, 0, self.delay_d.data_ptr(), self.rate_d.data_ptr(), self.a_d.data_ptr(),
            self.b_d.data_ptr())

    @property
    def n_det(self):
        return self.det_off.size - 1


class DeviceAtoms:
    """Atoms of T templates on the device in the batched map's layout:
    ``ts_d`` uint32 (as int32) [T*n_sft], ``cols_d`` float32 [7, T*n_sft];
    template t, detector X is segment t*n_det + X."""

    def __init__(self, sfts, ts_d, cols_d, T):
        self.ts_d = ts_d
        self.cols_d = cols_d
        self.T = T
        self.TAtom = sfts.Tsft
        n, D = sfts.n_sft, sfts.n_det
        self.seg_off = (np.arange(T, dtype=np.int64)[:, None] * n
                        + sfts.det_off[None, :-1]).ravel()
        self.seg_off = np.append(self.seg_off, T * n).astype(np.int64)
        self.seg_begin = np.arange(T + 1, dtype=np.int64) * D
        first = sfts.t_start[sfts.det_off[:-1]]
        last = sfts.t_start[sfts.det_off[1:] - 1]
        self.seg_tfirst = np.tile(first, T).astype(np.uint32)
        self.seg_tlast = np.tile(last, T).astype(np.uint32)

    def host(self):
        """(ts uint32 [T*n_sft], cols float32 [7, T*n_sft]) copied to the host."""
        return (self.ts_d.cpu().numpy().view(np.uint32), self.cols_d.cpu().numpy())


def demod_atoms_into(sfts, freq_d, fdot_d, ref_time, ts_d, cols_d, dterms=8, stream=None):
    """tcw_demod_atoms on device arrays the caller owns: freq_d, fdot_d
    float64 [T]; ts_d int32 [>= T*n_sft], cols_d float32 [7, >= T*n_sft]."""
    torch = _torch()
    T = int(freq_d.numel())
    if int(fdot_d.numel()) != T:
        raise ValueError("freq and fdot must have the same length")
    if ts_d.numel() < T * sfts.n_sft or cols_d.dim() != 2 or cols_d.shape[0] != NC:
        raise ValueError("output buffers too small for T * n_sft atoms")
    st = torch.cuda.current_stream(sfts.device) if stream is None else stream
    _hiplib.check(sfts.lib.tcw_demod_atoms(
        sfts.handle, ctypes.byref(sfts.desc), T, freq_d.data_ptr(), fdot_d.data_ptr(),
        float(ref_time), int(dterms), ts_d.data_ptr(), cols_d.data_ptr(), cols_d.shape[1],
        ctypes.c_void_p(st.cuda_stream)), "tcw_demod_atoms")


def demod_atoms(sfts, freq, fdot, ref_time, dterms=8, stream=None):
    """Atoms of the templates (freq[i], fdot[i]) at reference time ref_time
    (SSB GPS s) from ``sfts`` (an :class:`SFTSet`), on the device."""
    torch = _torch()
    f = torch.as_tensor(np.ascontiguousarray(freq, dtype=np.float64)).reshape(-1)
    fd = torch.as_tensor(np.ascontiguousarray(fdot, dtype=np.float64)).reshape(-1)
    if f.numel() != fd.numel():
        raise ValueError("freq and fdot must have the same length")
    T = int(f.numel())
    f = f.to(sfts.device)
    fd = fd.to(sfts.device)
    total = T * sfts.n_sft
    ts_d = torch.empty(max(total, 1), dtype=torch.int32, device=sfts.device)
    cols_d = torch.empty((NC, max(total, 1)), dtype=torch.float32, device=sfts.device)
    if T > 0:
        demod_atoms_into(sfts, f, fd, ref_time, ts_d, cols_d, dterms=dterms, stream=stream)
    out = DeviceAtoms(sfts, ts_d, cols_d, T)
    out._keep = (f, fd)  # inputs stay alive until the queued kernel has run
    return out


def demod_atoms_sky(sfts, delay, delay_rate, a, b, freq, fdot, ref_time, dterms=8,
                    stream=None):
    """Atoms of the templates (freq[i], fdot[i]) at each of n_sky sky
    positions in one launch (``tcw_demod_atoms_sky``): ``delay``,
    ``delay_rate``, ``a``, ``b`` are [n_sky, n_sft] (row s: sky point s's
    per-SFT values, SFTs in ``sfts``' order; ``sfts``' own delay and antenna
    patterns are not used).  Returns :class:`DeviceAtoms` of n_sky * T
    templates, sky point s's template i being template s * T + i."""
    torch = _torch()
    n = sfts.n_sft
    per_sky = []
    for name, x, dt in (("delay", delay, np.float64), ("delay_rate", delay_rate, np.float64),
                        ("a", a, np.float32), ("b", b, np.float32)):
        x = np.ascontiguousarray(x, dtype=dt)
        if x.ndim != 2 or x.shape[1] != n:
            raise ValueError(f"{name} must be [n_sky, n_sft]")
        per_sky.append(x)
    n_sky = per_sky[0].shape[0]
    if any(x.shape[0] != n_sky for x in per_sky):
        raise ValueError("delay, delay_rate, a and b must have the same number of sky points")
    f = np.ascontiguousarray(freq, dtype=np.float64).reshape(-1)
    fd = np.ascontiguousarray(fdot, dtype=np.float64).reshape(-1)
    if f.size != fd.size:
        raise ValueError("freq and fdot must have the same length")
    T = int(f.size)
    dev = sfts.device
    keep = [torch.from_numpy(x).to(dev) for x in per_sky + [f, fd]]
    total = n_sky * T * n
    ts_d = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
    cols_d = torch.empty((NC, max(total, 1)), dtype=torch.float32, device=dev)
    st = torch.cuda.current_stream(dev) if stream is None else stream
    _hiplib.check(sfts.lib.tcw_demod_atoms_sky(
        sfts.handle, ctypes.byref(sfts.desc), n_sky, *(x.data_ptr() for x in keep[:4]), T,
        keep[4].data_ptr(), keep[5].data_ptr(), float(ref_time), int(dterms), ts_d.data_ptr(),
        cols_d.data_ptr(), cols_d.shape[1], ctypes.c_void_p(st.cuda_stream)),
        "tcw_demod_atoms_sky")
    out = DeviceAtoms(sfts, ts_d, cols_d, n_sky * T)
    out._keep = keep  # inputs stay alive until the queued kernel has run
    return out


def batch_from_sfts(sfts, freq, fdot, ref_time, windows, dterms=8, **kw):
    """Batched transient maps (``BatchedTransientMaps``) of every template,
    its atoms produced on the device from ``sfts``; ``windows``: one window
    range per template (or one for all)."""
    from .batched import BatchedTransientMaps

    atoms = demod_atoms(sfts, freq, fdot, ref_time, dterms=dterms)
    if not isinstance(windows, (list, tuple)):
        windows = [windows] * atoms.T
    return BatchedTransientMaps.from_device_atoms(atoms, windows, **kw)
