"""Transient-CW F(t0, tau) maps with a MI355X "hip" backend.

Drop-in counterpart of pyfstat/tcw_fstat_map_funcs.py: the same public names,
call signatures, return values and exception classes, with a third map version
``"hip"`` beside ``"lal"`` and ``"pycuda"`` in ``fstatmap_versions``
(pyfstat/tcw_fstat_map_funcs.py:320-327):

* ``call_compute_transient_fstat_map(version, features, multiFstatAtoms,
  windowRange, BtSG)`` -> ``(FstatMap, timingFstatMap)``
  (pyfstat/tcw_fstat_map_funcs.py:495-544);
* ``init_transient_fstat_map_features(feature, cudaDeviceName)`` ->
  ``(features, gpu_context)`` (pyfstat/tcw_fstat_map_funcs.py:361-492); for
  "hip" the context is ``None`` (the device is held by libtcw_hip), so PyFstat's
  ``"cuda" in version`` finalizer (pyfstat/core.py:493,515) stays inert;
* ``pyTransientFstatMap`` with the fields ``F_mn`` (float32, F not 2F),
  ``maxF``, ``t0_ML``, ``tau_ML``, ``lnBtSG``, ``t0_MP``, ``tau_MP`` and the
  same methods (pyfstat/tcw_fstat_map_funcs.py:50-317).

The "hip" version computes everything the reference splits between the GPU
and numpy -- detector merge, map, max/argmax, lnBtSG and posterior modes -- on
the device in one pass (libtcw_hip, hand-written gfx950 kernels).  The batched
entry ``hip_compute_transient_fstat_maps`` replaces the per-template loop of
TransientGridSearch (pyfstat/grid_based_searches.py:1186-1189).

Differences from "pycuda" that are deliberate:
* the caller's windowRange is never modified (pycuda rewrites a
  TRANSIENT_NONE range in place, tcw_fstat_map_funcs.py:742-749; lalpulsar
  takes it by value);
* single-atom windows evaluate to the F = 2 fallback like pycuda; with
  ``lal_strict=True`` they raise like lalpulsar's XLAL_EDOM instead.
"""

import ctypes
import importlib
import logging
import os
from time import time

import numpy as np

from . import _hiplib
from .atoms import (
    TRANSIENT_EXPONENTIAL,
    TRANSIENT_LAST,
    TRANSIENT_NONE,
    TRANSIENT_RECTANGULAR,
    atoms_to_arrays,
    window_fields,
)

logger = logging.getLogger(__name__)


class pyTransientFstatMap:
    """F(t0, tau) map container (pyfstat/tcw_fstat_map_funcs.py:50-157).

    ``F_mn[m, n]`` is F (not 2F) at t0 = t0 + m*dt0, tau = tau + n*dtau.
    """

    def __init__(self, N_t0Range=None, N_tauRange=None, transientFstatMap_t=None,
                 from_file=None):
        if transientFstatMap_t is not None and from_file:
            raise ValueError(
                "Please choose either a transientFstatMap_t or file to init from."
            )
        if transientFstatMap_t is not None:
            self.F_mn = transientFstatMap_t.F_mn.data
            self.maxF = transientFstatMap_t.maxF
            self.t0_ML = transientFstatMap_t.t0_ML
            self.tau_ML = transientFstatMap_t.tau_ML
        elif from_file:
            self.read_from_file(from_file)
        else:
            if not N_t0Range or not N_tauRange:
                raise ValueError(
                    "Need either a transientFstatMap_t or a pair of (N_t0Range, N_tauRange)!."
                )
            self.F_mn = np.zeros((N_t0Range, N_tauRange), dtype=np.float32)
            self.maxF = float(-1.0)
            self.t0_ML = float(0.0)
            self.tau_ML = float(0.0)
        self.lnBtSG = np.nan
        self.t0_MP = np.nan
        self.tau_MP = np.nan

    def read_from_file(self, file):
        """Read a ``t0[s] tau[s] 2F`` text map (tcw_fstat_map_funcs.py:159-184)."""
        rows = np.atleast_2d(np.loadtxt(file, comments="#", dtype=np.float64))
        t0s = np.unique(rows[:, 0])
        taus = np.unique(rows[:, 1])
        self.F_mn = (0.5 * rows[:, 2]).reshape(t0s.size, taus.size)
        k = int(np.argmax(rows[:, 2]))
        self.t0_ML = rows[k, 0]
        self.tau_ML = rows[k, 1]
        self.maxF = 0.5 * rows[k, 2]

    def get_maxF_idx(self):
        """(m, n) of the first maximum in row-major order."""
        return np.unravel_index(np.argmax(self.F_mn), self.F_mn.shape)

    def _exp_rel(self):
        return np.exp(np.asarray(self.F_mn, dtype=np.float64) - float(self.maxF))

    def get_lnBtSG(self):
        """ln B_tSG = ln(70/(N_t0 N_tau)) + logsumexp(F_mn), rhohMax = 1
        (tcw_fstat_map_funcs.py:196-217, lalpulsar ComputeTransientBstat)."""
        logBhat = float(self.maxF) + np.log(np.sum(self._exp_rel()))
        self.lnBtSG = np.log(70.0 / np.prod(self.F_mn.shape)) + logBhat
        return self.lnBtSG

    def get_t0_max_posterior(self, windowRange):
        """Bin centre of the largest row sum of exp(F - maxF)
        (tcw_fstat_map_funcs.py:219-252)."""
        rows = np.sum(self._exp_rel(), axis=1)
        width = windowRange.t0Band / self.F_mn.shape[0]
        self.t0_MP = windowRange.t0 + (np.argmax(rows) + 0.5) * width
        return self.t0_MP

    def get_tau_max_posterior(self, windowRange):
        """Bin centre of the largest column sum of exp(F - maxF)
        (tcw_fstat_map_funcs.py:254-287)."""
        cols = np.sum(self._exp_rel(), axis=0)
        width = windowRange.tauBand / self.F_mn.shape[1]
        self.tau_MP = windowRange.tau + (np.argmax(cols) + 0.5) * width
        return self.tau_MP

    def write_F_mn_to_file(self, tCWfile, windowRange, header=[]):
        """Write ``t0 tau 2F`` rows in the lalpulsar.write_transientFstatMap_to_fp
        layout (tcw_fstat_map_funcs.py:289-317): header lines as ``# ...``, a
        ``# t0[s]     tau[s]     2F`` line, then ``  %10d %10d %- 11.8g`` rows."""
        N_t0, N_tau = self.F_mn.shape
        t0s = np.repeat(windowRange.t0 + np.arange(N_t0, dtype=np.int64) * windowRange.dt0,
                        N_tau)
        taus = np.tile(windowRange.tau + np.arange(N_tau, dtype=np.int64) * windowRange.dtau,
                       N_t0)
        twoF = 2.0 * np.asarray(self.F_mn, dtype=np.float64).ravel()
        with open(tCWfile, "w") as tfp:
            for hline in header:
                tfp.write("# {:s}\n".format(hline))
            tfp.write("# t0[s]     tau[s]     2F\n")
            tfp.writelines(
                "  %10d %10d %- 11.8g\n" % row for row in zip(t0s.tolist(), taus.tolist(),
                                                             twoF.tolist())
            )


# ----------------------------------------------------------------------------
# optional imports and features
# ----------------------------------------------------------------------------


def _optional_import(modulename):
    try:
        return importlib.import_module(modulename)
    except ImportError:
        logger.debug("Failed to import module {:s}.".format(modulename))
        return None


def _hip_available():
    try:
        lib = _hiplib.load()
    except OSError as e:
        logger.debug(f"libtcw_hip not loadable: {e}")
        return False
    return lib.tcw_device_count() > 0


def _get_transient_fstat_map_features():
    """Which map versions can run here (tcw_fstat_map_funcs.py:351-358)."""
    features = {}
    features["lal"] = (_optional_import("lal") is not None
                       and _optional_import("lalpulsar") is not None)
    features["pycuda"] = all(
        _optional_import(m) is not None
        for m in ("pycuda", "pycuda.driver", "pycuda.gpuarray", "pycuda.tools",
                  "pycuda.compiler", "pyfstat")
    )
    features["hip"] = _hip_available()
    return features


class _HipState:
    device = None
    handle = None


def _hip_device_names():
    lib = _hiplib.load()
    names = []
    for n in range(lib.tcw_device_count()):
        b = ctypes.create_string_buffer(256)
        _hiplib.check(lib.tcw_device_name(n, b, 256), "tcw_device_name")
        names.append(b.value.decode(errors="replace").replace(" ", "-").replace("_", "-"))
    return names


def _hip_select_device(devnum):
    lib = _hiplib.load()
    if _HipState.handle is not None and _HipState.device == devnum:
        return
    if _HipState.handle is not None:
        lib.tcw_destroy(_HipState.handle)
        _HipState.handle = None
    h = ctypes.c_void_p()
    _hiplib.check(lib.tcw_create(int(devnum), ctypes.byref(h)), "tcw_create")
    _HipState.handle = h
    _HipState.device = int(devnum)


def init_transient_fstat_map_features(feature="lal", cudaDeviceName=None):
    """Initialise a map version (tcw_fstat_map_funcs.py:361-492).

    For ``"hip"``: the device is chosen like the reference chooses a CUDA
    device -- a partial match of ``cudaDeviceName`` against the device names
    (spaces and underscores as dashes), else ``$CUDA_DEVICE``, else 0 -- and
    ``(features, None)`` is returned.
    """
    features = _get_transient_fstat_map_features()
    logger.debug("Got the following features for transient F-stat maps:")
    logger.debug(features)
    gpu_context = None
    if feature == "hip":
        if not features["hip"]:
            raise RuntimeError("hip use was requested, but libtcw_hip or a HIP device is unavailable.")
        devnames = _hip_device_names()
        num_gpus = len(devnames)
        logger.info("Found {} HIP device(s).".format(num_gpus))
        if cudaDeviceName:
            devmatches = [k for k, name in enumerate(devnames) if cudaDeviceName in name]
            if len(devmatches) == 0:
                raise RuntimeError(
                    'Requested HIP device "{}" not found. Available devices: [{}]'.format(
                        cudaDeviceName, ",".join(devnames))
                )
            devnum = devmatches[0]
            if len(devmatches) > 1:
                logger.warning(
                    'Found {} HIP devices matching name "{}". Choosing first one with index {}.'.format(
                        len(devmatches), cudaDeviceName, devnum)
                )
        elif "CUDA_DEVICE" in os.environ:
            devnum = int(os.environ["CUDA_DEVICE"])
        else:
            devnum = 0
        if devnum >= num_gpus:
            raise RuntimeError(
                "Requested HIP device number {} exceeds number of available devices!"
                " Please change through environment variable $CUDA_DEVICE.".format(devnum)
            )
        _hip_select_device(devnum)
        logger.info("Choosing HIP device {}, of {} devices present: {}".format(
            devnum, num_gpus, devnames[devnum]))
    elif feature == "pycuda":
        if not features["pycuda"]:
            raise RuntimeError("pycuda use was requested, but imports failed.")
        ref = importlib.import_module("pyfstat.tcw_fstat_map_funcs")
        return ref.init_transient_fstat_map_features("pycuda", cudaDeviceName)
    elif feature == "lal":
        gpu_context = None
    else:
        raise ValueError(
            "Unknown transient F-stat map computation feature" f"'{feature}' requested."
        )
    return features, gpu_context


def call_compute_transient_fstat_map(version, features, multiFstatAtoms=None,
                                     windowRange=None, BtSG=False):
    """Run map version ``version``; returns ``(FstatMap, timingFstatMap)``
    (tcw_fstat_map_funcs.py:495-544)."""
    if version in fstatmap_versions:
        if features[version]:
            time0 = time()
            FstatMap = fstatmap_versions[version](multiFstatAtoms, windowRange, BtSG)
            timingFstatMap = time() - time0
        else:
            raise Exception(
                "Required module(s) for transient F-stat map"
                ' method "{}" not available!'.format(version)
            )
    else:
        raise Exception(
            'Transient F-stat map method "{}"' " not implemented!".format(version)
        )
    return FstatMap, timingFstatMap


def lalpulsar_compute_transient_fstat_map(multiFstatAtoms, windowRange, BtSG=False):
    """The lalpulsar version, for environments that have lalsuite
    (tcw_fstat_map_funcs.py:547-587)."""
    lalpulsar = importlib.import_module("lalpulsar")
    FstatMap_lalpulsar = lalpulsar.ComputeTransientFstatMap(
        multiFstatAtoms=multiFstatAtoms, windowRange=windowRange, useFReg=False
    )
    FstatMap = pyTransientFstatMap(transientFstatMap_t=FstatMap_lalpulsar)
    if BtSG:
        FstatMap.lnBtSG = lalpulsar.ComputeTransientBstat(windowRange, FstatMap_lalpulsar)
        pdf_t0 = lalpulsar.ComputeTransientPosterior_t0(windowRange, FstatMap_lalpulsar)
        pdf_tau = lalpulsar.ComputeTransientPosterior_tau(windowRange, FstatMap_lalpulsar)
        FstatMap.t0_MP = lalpulsar.FindModeOfPDF1D(pdf_t0)
        FstatMap.tau_MP = lalpulsar.FindModeOfPDF1D(pdf_tau)
    return FstatMap


def pycuda_compute_transient_fstat_map(multiFstatAtoms, windowRange, BtSG=False):
    """The reference's pycuda version, delegated when pyfstat and pycuda exist."""
    ref = importlib.import_module("pyfstat.tcw_fstat_map_funcs")
    return ref.pycuda_compute_transient_fstat_map(multiFstatAtoms, windowRange, BtSG)


def reshape_FstatAtomsVector(atomsVector):
    """Dict of per-field arrays of one detector's atoms
    (tcw_fstat_map_funcs.py:590-639): ``timestamp`` uint32, the others float32
    with complex fields split into ``_re``/``_im``."""
    from .atoms import _single_to_arrays

    ts, cols = _single_to_arrays(atomsVector)
    return {
        "timestamp": ts,
        "a2_alpha": cols[0].copy(),
        "b2_alpha": cols[1].copy(),
        "ab_alpha": cols[2].copy(),
        "Fa_alpha_re": cols[3].copy(),
        "Fa_alpha_im": cols[4].copy(),
        "Fb_alpha_re": cols[5].copy(),
        "Fb_alpha_im": cols[6].copy(),
    }


# ----------------------------------------------------------------------------
# the "hip" version
# ----------------------------------------------------------------------------


def _torch():
    return importlib.import_module("torch")


def _ensure_handle():
    if _HipState.handle is None:
        devnum = int(os.environ.get("CUDA_DEVICE", "0"))
        _hip_select_device(devnum)
    return _HipState.handle


def _check_window_type(windowRange):
    if windowRange.type >= TRANSIENT_LAST:
        raise ValueError(
            "Unknown window-type ({}) passed as input. Allowed are [0,{}].".format(
                windowRange.type, TRANSIENT_LAST - 1)
        )


def _upload_packed(packed, dev):
    torch = _torch()
    ts = torch.from_numpy(packed.ts.view(np.int32)).to(dev, non_blocking=False)
    cols = torch.from_numpy(packed.cols).to(dev, non_blocking=False)
    return ts, cols


def _map_from_stats(F_mn, st, BtSG):
    FstatMap = pyTransientFstatMap.__new__(pyTransientFstatMap)
    FstatMap.F_mn = F_mn
    FstatMap.maxF = np.float32(st.maxF)
    FstatMap.t0_ML = int(st.t0_ML)
    FstatMap.tau_ML = int(st.tau_ML)
    FstatMap.lnBtSG = float(st.lnBtSG) if BtSG else np.nan
    FstatMap.t0_MP = float(st.t0_MP) if BtSG else np.nan
    FstatMap.tau_MP = float(st.tau_MP) if BtSG else np.nan
    return FstatMap


def hip_compute_transient_fstat_map(multiFstatAtoms, windowRange, BtSG=False,
                                    lal_strict=False, return_device=False,
                                    force_generic=False):
    """One transient F-statistic map on the HIP device.

    Counterpart of pycuda_compute_transient_fstat_map
    (tcw_fstat_map_funcs.py:656-834); same arguments and return.  Extra
    keywords: ``lal_strict`` raises on single-atom windows like lalpulsar;
    ``return_device`` leaves ``FstatMap.F_mn`` as a device torch tensor (no
    device-to-host copy); ``force_generic`` selects the generic kernels (tests).
    """
    _check_window_type(windowRange)
    torch = _torch()
    lib = _hiplib.load()
    handle = _ensure_handle()
    dev = torch.device("cuda", _HipState.device)
    packed = atoms_to_arrays(multiFstatAtoms)
    wfields = window_fields(windowRange)
    win = _hiplib.window_struct(wfields)
    nt0 = ctypes.c_int64()
    ntau = ctypes.c_int64()
    _hiplib.check(lib.tcw_map_shape(packed.num_binned, packed.TAtom, packed.t_min,
                                    ctypes.byref(win), ctypes.byref(nt0), ctypes.byref(ntau)),
                  "tcw_map_shape")
    ts_d, cols_d = _upload_packed(packed, dev)
    F_d = torch.empty((nt0.value, ntau.value), dtype=torch.float32, device=dev)
    at = _hiplib.Atoms()
    at.numDet = packed.numDet
    at.TAtom = packed.TAtom
    at.ts = ts_d.data_ptr()
    at.cols = cols_d.data_ptr()
    at.col_stride = packed.cols.shape[1]
    for X in range(packed.numDet + 1):
        at.det_off[X] = int(packed.det_off[X])
    at.t_min = packed.t_min
    at.t_max = packed.t_max
    flags = 0
    if BtSG:
        flags |= _hiplib.TCW_FLAG_BTSG
    if lal_strict:
        flags |= _hiplib.TCW_FLAG_LAL_STRICT
    if force_generic:
        flags |= _hiplib.TCW_FLAG_GENERIC
    st = _hiplib.Stats()
    stream = torch.cuda.current_stream(dev).cuda_stream
    _hiplib.check(lib.tcw_map(handle, ctypes.byref(at), ctypes.byref(win),
                              ctypes.c_void_p(F_d.data_ptr()), ctypes.byref(st), flags,
                              ctypes.c_void_p(stream)), "tcw_map")
    F_mn = F_d if return_device else F_d.cpu().numpy()
    FstatMap = _map_from_stats(F_mn, st, BtSG)
    logger.debug(
        "Done computing transient F-stat map. maxF={:.4f}, t0_ML={}, tau_ML={}".format(
            FstatMap.maxF, FstatMap.t0_ML, FstatMap.tau_ML)
    )
    return FstatMap


fstatmap_versions = {
    "lal": lambda multiFstatAtoms, windowRange, BtSG: lalpulsar_compute_transient_fstat_map(
        multiFstatAtoms, windowRange, BtSG
    ),
    "pycuda": lambda multiFstatAtoms, windowRange, BtSG: pycuda_compute_transient_fstat_map(
        multiFstatAtoms, windowRange, BtSG
    ),
    "hip": lambda multiFstatAtoms, windowRange, BtSG: hip_compute_transient_fstat_map(
        multiFstatAtoms, windowRange, BtSG
    ),
}
"""Map versions by name (tcw_fstat_map_funcs.py:320-327) plus "hip"."""


def register_into_pyfstat():
    """Add the "hip" version to an importable PyFstat, so its callers
    (pyfstat/core.py:1451-1457,1534-1540 via ComputeFstat(tCWFstatMapVersion="hip"))
    reach it unchanged.  Returns True if PyFstat was found."""
    ref = _optional_import("pyfstat.tcw_fstat_map_funcs")
    if ref is None:
        return False
    if getattr(ref, "_hip_registered", False):  # idempotent
        return True
    ref.fstatmap_versions["hip"] = fstatmap_versions["hip"]
    orig_features = ref._get_transient_fstat_map_features
    orig_init = ref.init_transient_fstat_map_features

    def _features():
        f = orig_features()
        f["hip"] = _hip_available()
        return f

    def _init(feature="lal", cudaDeviceName=None):
        if feature == "hip":
            return init_transient_fstat_map_features("hip", cudaDeviceName)
        return orig_init(feature, cudaDeviceName)

    ref._get_transient_fstat_map_features = _features
    ref.init_transient_fstat_map_features = _init
    ref._hip_registered = True
    return True


__all__ = [
    "TRANSIENT_NONE",
    "TRANSIENT_RECTANGULAR",
    "TRANSIENT_EXPONENTIAL",
    "TRANSIENT_LAST",
    "pyTransientFstatMap",
    "fstatmap_versions",
    "init_transient_fstat_map_features",
    "call_compute_transient_fstat_map",
    "lalpulsar_compute_transient_fstat_map",
    "hip_compute_transient_fstat_map",
    "reshape_FstatAtomsVector",
    "register_into_pyfstat",
]
