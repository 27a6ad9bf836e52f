"""FP64 CPU oracle of the transient F-statistic map -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import
this package, and only to check results or to time them.  The product (``pyfstat_amd``) never
imports, links or runs anything here.

Two independent restatements of the reference algorithm live here:

* ``tcw_oracle.c`` (built by ``oracle/Makefile`` into ``oracle/build/``), bound
  below through ctypes: the FP64 oracle used for parity at real sizes and as
  the CPU baseline.  It follows lalpulsar's XLALComputeTransientFstatMap as the
  reference mirrors it (pyfstat/tcw_fstat_map_funcs.py:656-834,
  pyfstat/pyCUDAkernels/cudaTransientFstatRectWindow.cu:14-124,
  pyfstat/pyCUDAkernels/cudaTransientFstatExpWindow.cu:15-134).
* ``twin.py``: a plain-Python loop restatement of the same steps for tiny
  inputs, used to cross-check the C oracle.

``tcw_oracle.c`` also restates, for timing, the refinement evaluation of
``refine_ref.py`` and the demodulation atom sum of ``demod_ref.py``
(``demod_atoms`` below; checked against the numpy version in
tests/test_demod.py).

How each function is pinned (closed forms, invariants, brute force, library
routines) is listed in tests/test_oracle.py.  Parity unpinned: lalpulsar's
lookup-table exp (XLALFastNegExp) and its REAL4 accumulation order, which only
an import of lalpulsar could pin; the oracle uses exact exp and FP64 sums.
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libtcw_oracle.so")

TRANSIENT_NONE = 0
TRANSIENT_RECTANGULAR = 1
TRANSIENT_EXPONENTIAL = 2
TRANSIENT_LAST = 3


class _Window(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in
                ("type", "t0", "t0Band", "dt0", "tau", "tauBand", "dtau")]


class _Result(ctypes.Structure):
    _fields_ = [
        ("maxF", ctypes.c_double),
        ("m_max", ctypes.c_int64),
        ("n_max", ctypes.c_int64),
        ("N_t0", ctypes.c_int64),
        ("N_tau", ctypes.c_int64),
        ("t0", ctypes.c_uint32),
        ("dt0", ctypes.c_uint32),
        ("tau", ctypes.c_uint32),
        ("dtau", ctypes.c_uint32),
        ("bad_m", ctypes.c_int64),
        ("bad_n", ctypes.c_int64),
    ]


def build(native=False):
    """Compile tcw_oracle.c with the committed Makefile (``native``: the
    -march=native variant, for timing on the machine that builds it)."""
    subprocess.run(["make", "-s", "-C", _HERE] + (["native"] if native else []), check=True)


_lib = None
_variant = "generic"


def select_native():
    """Build (on this machine) and use the -O3 -march=native library from now
    on; returns its compiler flags for the record."""
    global _lib, _LIB_PATH, _variant
    build(native=True)
    _LIB_PATH = os.path.join(_HERE, "build", "native", "libtcw_oracle.so")
    _lib = None
    _variant = "native"
    _load()
    return "gcc -O3 -march=native -ffp-contract=off -fopenmp"


def set_threads(n):
    """OpenMP threads of the batch functions; returns the previous maximum."""
    return int(_load().orc_set_threads(int(n)))


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), for the record."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build(native=_variant == "native")
        lib = ctypes.CDLL(_LIB_PATH)
        lib.orc_set_threads.restype = ctypes.c_int
        lib.orc_set_threads.argtypes = [ctypes.c_int]
        vp = ctypes.c_void_p
        lib.orc_refine_eval_batch.restype = ctypes.c_int
        lib.orc_refine_eval_batch.argtypes = [ctypes.c_int64, vp, ctypes.c_int64, vp, vp,
                                              ctypes.c_uint32, vp, vp, vp, vp, vp, vp, vp,
                                              vp, vp]
        P = ctypes.POINTER
        lib.orc_merge_size.restype = ctypes.c_int64
        lib.orc_merge_size.argtypes = [ctypes.c_int, P(ctypes.c_int64),
                                       P(ctypes.c_uint32), ctypes.c_uint32,
                                       P(ctypes.c_uint32)]
        lib.orc_merge.restype = ctypes.c_int
        lib.orc_merge.argtypes = [ctypes.c_int, P(ctypes.c_int64), P(ctypes.c_uint32),
                                  P(ctypes.c_float), ctypes.c_int64, ctypes.c_uint32,
                                  ctypes.c_uint32, ctypes.c_int64, P(ctypes.c_float)]
        lib.orc_map_shape.restype = ctypes.c_int
        lib.orc_map_shape.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                      P(_Window), P(_Window), P(ctypes.c_int64),
                                      P(ctypes.c_int64)]
        lib.orc_map.restype = ctypes.c_int
        lib.orc_map.argtypes = [P(ctypes.c_float), ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_uint32, P(_Window), P(ctypes.c_double),
                                P(ctypes.c_double), P(_Result)]
        lib.orc_lnBtSG.restype = ctypes.c_double
        lib.orc_lnBtSG.argtypes = [P(ctypes.c_double), ctypes.c_int64, ctypes.c_int64,
                                   ctypes.c_double]
        lib.orc_posterior_modes.restype = None
        lib.orc_posterior_modes.argtypes = [P(ctypes.c_double), ctypes.c_int64,
                                            ctypes.c_int64, ctypes.c_double, P(_Window),
                                            P(ctypes.c_double), P(ctypes.c_double)]
        lib.orc_map_batch.restype = ctypes.c_int
        lib.orc_map_batch.argtypes = [ctypes.c_int64, P(ctypes.c_float), P(ctypes.c_int64),
                                      P(ctypes.c_uint32), P(ctypes.c_uint32),
                                      P(ctypes.c_uint32), P(_Window), P(ctypes.c_double),
                                      P(ctypes.c_int64), P(ctypes.c_double),
                                      P(ctypes.c_int64), P(ctypes.c_int64),
                                      P(ctypes.c_double)]
        lib.orc_demod_atoms.restype = ctypes.c_int
        lib.orc_demod_atoms.argtypes = [ctypes.c_int64, vp, vp, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_double, vp, vp, vp, vp, ctypes.c_int64, vp,
                                        vp, ctypes.c_double, ctypes.c_int32, vp, vp]
        _lib = lib
    return _lib


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def _win(w):
    if isinstance(w, (tuple, list)):
        return _Window(*[int(x) for x in w])
    return _Window(*[int(getattr(w, f)) for f in
                     ("type", "t0", "t0Band", "dt0", "tau", "tauBand", "dtau")])


def merge_binned(ts, cols, det_off, TAtom):
    """lalpulsar mergeMultiFstatAtomsBinned restated: float32 [7, N], t0_data.

    ts uint32 [total], cols float32 [7, total], det_off int64 [numDet+1].
    """
    ts = np.ascontiguousarray(ts, dtype=np.uint32)
    cols = np.ascontiguousarray(cols, dtype=np.float32)
    det_off = np.ascontiguousarray(det_off, dtype=np.int64)
    lib = _load()
    tmin = ctypes.c_uint32(0)
    n = lib.orc_merge_size(len(det_off) - 1, _ptr(det_off, ctypes.c_int64),
                           _ptr(ts, ctypes.c_uint32), int(TAtom), ctypes.byref(tmin))
    if n <= 0:
        raise ValueError("bad atoms for merge")
    out = np.empty((7, n), dtype=np.float32)
    rc = lib.orc_merge(len(det_off) - 1, _ptr(det_off, ctypes.c_int64),
                       _ptr(ts, ctypes.c_uint32), _ptr(cols, ctypes.c_float),
                       cols.shape[1], int(TAtom), tmin.value, n,
                       _ptr(out, ctypes.c_float))
    if rc:
        raise ValueError("merge failed")
    return out, int(tmin.value)


class OracleMap:
    """FP64 map result: F (float64 [N_t0, N_tau]), maxF, argmax, effective window."""

    def __init__(self, F, cond, res, win_eff):
        self.F_mn = F
        self.cond = cond
        self.maxF = float(res.maxF)
        self.m_max = int(res.m_max)
        self.n_max = int(res.n_max)
        self.t0_ML = int(res.t0) + self.m_max * int(res.dt0)
        self.tau_ML = int(res.tau) + self.n_max * int(res.dtau)
        self.bad_cell = None if res.bad_m < 0 else (int(res.bad_m), int(res.bad_n))
        self.window_eff = win_eff


def compute_map(merged, TAtom, t0_data, window):
    """FP64 map over merged atoms (float32 [7, N]) for a window range."""
    merged = np.ascontiguousarray(merged, dtype=np.float32)
    lib = _load()
    w = _win(window)
    eff = _Window()
    nt0 = ctypes.c_int64()
    ntau = ctypes.c_int64()
    N = merged.shape[1]
    if lib.orc_map_shape(N, int(TAtom), int(t0_data), ctypes.byref(w), ctypes.byref(eff),
                         ctypes.byref(nt0), ctypes.byref(ntau)):
        raise ValueError(f"bad window {window}")
    F = np.empty((nt0.value, ntau.value), dtype=np.float64)
    cond = np.empty_like(F)
    res = _Result()
    rc = lib.orc_map(_ptr(merged, ctypes.c_float), N, int(TAtom), int(t0_data),
                     ctypes.byref(w), _ptr(F, ctypes.c_double), _ptr(cond, ctypes.c_double),
                     ctypes.byref(res))
    if rc:
        raise ValueError("oracle map failed")
    win_eff = tuple(int(getattr(eff, f)) for f in
                    ("type", "t0", "t0Band", "dt0", "tau", "tauBand", "dtau"))
    return OracleMap(F, cond, res, win_eff)


def lnBtSG(F, maxF=None):
    F = np.ascontiguousarray(F, dtype=np.float64)
    if maxF is None:
        maxF = float(F.max())
    return float(_load().orc_lnBtSG(_ptr(F, ctypes.c_double), F.shape[0], F.shape[1],
                                    float(maxF)))


def posterior_modes(F, window, maxF=None):
    """(t0_MP, tau_MP) from the caller's window range (tcw.py:219-287)."""
    F = np.ascontiguousarray(F, dtype=np.float64)
    if maxF is None:
        maxF = float(F.max())
    w = _win(window)
    t0 = ctypes.c_double()
    tau = ctypes.c_double()
    _load().orc_posterior_modes(_ptr(F, ctypes.c_double), F.shape[0], F.shape[1],
                                float(maxF), ctypes.byref(w), ctypes.byref(t0),
                                ctypes.byref(tau))
    return t0.value, tau.value


def compute_transient_fstat_map(packed, window, BtSG=False):
    """Full oracle path from packed per-detector atoms: merge, map, stats.

    ``packed`` has ``ts``, ``cols``, ``det_off``, ``TAtom``
    (pyfstat_amd.atoms.PackedAtoms or any object with those fields).
    Returns an OracleMap with lnBtSG/t0_MP/tau_MP set when BtSG.
    """
    merged, t0_data = merge_binned(packed.ts, packed.cols, packed.det_off, packed.TAtom)
    om = compute_map(merged, packed.TAtom, t0_data, window)
    om.merged = merged
    om.t0_data = t0_data
    om.lnBtSG = float("nan")
    om.t0_MP = float("nan")
    om.tau_MP = float("nan")
    if BtSG:
        om.lnBtSG = lnBtSG(om.F_mn, om.maxF)
        om.t0_MP, om.tau_MP = posterior_modes(om.F_mn, window, om.maxF)
    return om


def compute_map_batch(merged_list, TAtoms, t0_datas, windows, with_lnBtSG=False):
    """Many maps at once, OpenMP over templates (CPU baseline).

    Returns (list of F arrays, maxF [T], m_max [T], n_max [T], lnBtSG [T] or None).
    """
    lib = _load()
    T = len(merged_list)
    sizes = np.array([m.shape[1] for m in merged_list], dtype=np.uint32)
    atoms_off = np.zeros(T + 1, dtype=np.int64)
    atoms_off[1:] = np.cumsum(7 * sizes.astype(np.int64))
    atoms = np.empty(int(atoms_off[-1]), dtype=np.float32)
    shapes = []
    wins = (_Window * T)()
    for t, m in enumerate(merged_list):
        atoms[atoms_off[t]:atoms_off[t + 1]] = np.ascontiguousarray(m, np.float32).ravel()
        w = _win(windows[t])
        wins[t] = w
        eff = _Window()
        nt0 = ctypes.c_int64()
        ntau = ctypes.c_int64()
        if lib.orc_map_shape(int(sizes[t]), int(TAtoms[t]), int(t0_datas[t]), ctypes.byref(w),
                             ctypes.byref(eff), ctypes.byref(nt0), ctypes.byref(ntau)):
            raise ValueError("bad window")
        shapes.append((nt0.value, ntau.value))
    F_off = np.zeros(T + 1, dtype=np.int64)
    F_off[1:] = np.cumsum([a * b for a, b in shapes])
    F = np.empty(int(F_off[-1]), dtype=np.float64)
    maxF = np.empty(T)
    m_max = np.empty(T, dtype=np.int64)
    n_max = np.empty(T, dtype=np.int64)
    lnB = np.empty(T) if with_lnBtSG else None
    TA = np.ascontiguousarray(TAtoms, dtype=np.uint32)
    T0 = np.ascontiguousarray(t0_datas, dtype=np.uint32)
    rc = lib.orc_map_batch(
        T, _ptr(atoms, ctypes.c_float), _ptr(atoms_off, ctypes.c_int64),
        _ptr(sizes, ctypes.c_uint32), _ptr(TA, ctypes.c_uint32), _ptr(T0, ctypes.c_uint32),
        wins,
        _ptr(F, ctypes.c_double), _ptr(F_off, ctypes.c_int64), _ptr(maxF, ctypes.c_double),
        _ptr(m_max, ctypes.c_int64), _ptr(n_max, ctypes.c_int64),
        _ptr(lnB, ctypes.c_double) if lnB is not None else None)
    if rc:
        raise ValueError("oracle batch failed")
    Fs = [F[F_off[t]:F_off[t + 1]].reshape(shapes[t]) for t in range(T)]
    return Fs, maxF, m_max, n_max, lnB


def refine_eval_batch(cols, m_off, t_data, TA, cand_tmpl, k_lo, k_hi, t0, tau):
    """C restatement of refine_ref.batch_evaluate (FP64, direct exp per bin),
    OpenMP over candidates.  cols float32 [7, n_bins] (merged layout, template
    t at m_off[t]..).  Returns F [n], dt0 [n], dtau [n], g [n, 7]."""
    cols = np.ascontiguousarray(cols, dtype=np.float32)
    m_off = np.ascontiguousarray(m_off, dtype=np.int64)
    t_data = np.ascontiguousarray(t_data, dtype=np.uint32)
    ct = np.ascontiguousarray(cand_tmpl, dtype=np.int64)
    lo = np.ascontiguousarray(k_lo, dtype=np.int64)
    hi = np.ascontiguousarray(k_hi, dtype=np.int64)
    t0 = np.ascontiguousarray(t0, dtype=np.float64)
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    n = ct.size
    F = np.empty(n)
    dt0 = np.empty(n)
    dtau = np.empty(n)
    g = np.empty((n, 7))
    rc = _load().orc_refine_eval_batch(n, cols.ctypes.data, cols.shape[1], m_off.ctypes.data,
                                       t_data.ctypes.data, int(TA), ct.ctypes.data,
                                       lo.ctypes.data, hi.ctypes.data, t0.ctypes.data,
                                       tau.ctypes.data, F.ctypes.data, dt0.ctypes.data,
                                       dtau.ctypes.data, g.ctypes.data)
    if rc:
        raise ValueError("oracle refine batch failed")
    return F, dt0, dtau, g


def demod_atoms(t_start, bins, k_min, Tsft, delay, delay_rate, a, b, freq, fdot, ref_time,
                dterms=8):
    """C restatement of demod_ref.demod_atoms (FP64, OpenMP over templates and
    SFTs); the same return values: (a2, b2, ab, Fa complex, Fb complex), each
    [n_tmpl, n_sft]."""
    t_start = np.ascontiguousarray(t_start, dtype=np.uint32)
    n = t_start.size
    bins = np.asarray(bins)
    pair = np.empty((n, bins.shape[1], 2), dtype=np.float32)
    pair[..., 0] = bins.real
    pair[..., 1] = bins.imag
    delay = np.ascontiguousarray(delay, dtype=np.float64)
    rate = np.ascontiguousarray(delay_rate, dtype=np.float64)
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    b32 = np.ascontiguousarray(b, dtype=np.float32)
    freq = np.ascontiguousarray(np.atleast_1d(freq), dtype=np.float64)
    fdot = np.ascontiguousarray(np.atleast_1d(fdot), dtype=np.float64)
    T = freq.size
    Fa = np.empty((T, n, 2))
    Fb = np.empty((T, n, 2))
    rc = _load().orc_demod_atoms(n, t_start.ctypes.data, pair.ctypes.data, pair.shape[1],
                                 int(k_min), float(Tsft), delay.ctypes.data, rate.ctypes.data,
                                 a32.ctypes.data, b32.ctypes.data, T, freq.ctypes.data,
                                 fdot.ctypes.data, float(ref_time), int(dterms), Fa.ctypes.data,
                                 Fb.ctypes.data)
    if rc:
        raise ValueError("oracle demod failed")
    ad, bd = a32.astype(np.float64), b32.astype(np.float64)
    shape = (T, n)
    return (np.broadcast_to(ad * ad, shape), np.broadcast_to(bd * bd, shape),
            np.broadcast_to(ad * bd, shape), Fa[..., 0] + 1j * Fa[..., 1],
            Fb[..., 0] + 1j * Fb[..., 1])
