"""ctypes binding of libtcw_hip.so, the C ABI declared in include/tcw_hip.h.

The library is built in-tree by ``pyfstat_amd.build`` (hipcc, gfx950).  There
is no fallback: if the library is missing or cannot be loaded, ``load()``
raises, and the "hip" map version reports itself unavailable.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libtcw_hip.so")

TCW_NCOL = 7
TCW_MAX_DETECTORS = 10

TCW_OK = 0
TCW_ERR_INVALID = 1
TCW_ERR_DEVICE = 2
TCW_ERR_SINGLE_ATOM = 3
TCW_ERR_NOMEM = 4

TCW_FLAG_BTSG = 1
TCW_FLAG_LAL_STRICT = 2
TCW_FLAG_GENERIC = 4

# static device scratch of tcw_map_noalloc (TCW_NOALLOC_SCRATCH_BYTES)
TCW_NOALLOC_SCRATCH_BYTES = 32 << 20

MODE_RECT_FAST = 0
MODE_RECT_GENERIC = 1
MODE_EXP_DIRECT = 2
MODE_EXP_FAST = 3
MODE_COUNT = 4

EXPORTED_SYMBOLS = (
    "tcw_version",
    "tcw_last_error",
    "tcw_device_count",
    "tcw_device_name",
    "tcw_map_shape",
    "tcw_create",
    "tcw_destroy",
    "tcw_map",
    "tcw_map_noalloc",
    "tcw_batch_plan",
    "tcw_batch_plan_rows",
    "tcw_plan_scratch_bytes",
    "tcw_plan_items",
    "tcw_batch_run",
    "tcw_batch_run_timed",
    "tcw_plan_destroy",
    "tcw_merge_batch",
    "tcw_gather_bins",
    "tcw_refine_plan_create",
    "tcw_refine_plan_destroy",
    "tcw_refine_forward",
    "tcw_refine_backward",
    "tcw_refine_ascent",
    "tcw_demod_atoms",
    "tcw_demod_atoms_sky",
)

TCW_REFINE_SAVED = 10


class Window(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in
                ("type", "t0", "t0Band", "dt0", "tau", "tauBand", "dtau")]


class Atoms(ctypes.Structure):
    _fields_ = [
        ("numDet", ctypes.c_uint32),
        ("TAtom", ctypes.c_uint32),
        ("ts", ctypes.c_void_p),
        ("cols", ctypes.c_void_p),
        ("col_stride", ctypes.c_int64),
        ("det_off", ctypes.c_int64 * (TCW_MAX_DETECTORS + 1)),
        ("t_min", ctypes.c_uint32),
        ("t_max", ctypes.c_uint32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("maxF", ctypes.c_float),
        ("single_atom", ctypes.c_int32),
        ("m_ML", ctypes.c_int64),
        ("n_ML", ctypes.c_int64),
        ("N_t0", ctypes.c_int64),
        ("N_tau", ctypes.c_int64),
        ("t0_ML", ctypes.c_int64),
        ("tau_ML", ctypes.c_int64),
        ("lnBtSG", ctypes.c_double),
        ("m_MP", ctypes.c_int64),
        ("n_MP", ctypes.c_int64),
        ("t0_MP", ctypes.c_double),
        ("tau_MP", ctypes.c_double),
        ("numAtoms", ctypes.c_uint32),
        ("t0_data", ctypes.c_uint32),
    ]


STATS_BYTES = ctypes.sizeof(Stats)


class BatchDesc(ctypes.Structure):
    _fields_ = [
        ("T", ctypes.c_int64),
        ("TAtom", ctypes.c_uint32),
        ("seg_begin", ctypes.c_void_p),
        ("seg_off", ctypes.c_void_p),
        ("seg_tfirst", ctypes.c_void_p),
        ("seg_tlast", ctypes.c_void_p),
        ("windows", ctypes.c_void_p),
    ]


class RefineDesc(ctypes.Structure):
    _fields_ = [
        ("T", ctypes.c_int64),
        ("n_cand", ctypes.c_int64),
        ("TAtom", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("m_off", ctypes.c_void_p),
        ("t_data", ctypes.c_void_p),
        ("cand_off", ctypes.c_void_p),
        ("k_lo", ctypes.c_void_p),
        ("k_hi", ctypes.c_void_p),
    ]


class SftSet(ctypes.Structure):
    _fields_ = [
        ("n_sft", ctypes.c_int64),
        ("t_start", ctypes.c_void_p),
        ("bins", ctypes.c_void_p),
        ("n_bins", ctypes.c_int32),
        ("k_min", ctypes.c_int32),
        ("Tsft", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("delay", ctypes.c_void_p),
        ("delay_rate", ctypes.c_void_p),
        ("a", ctypes.c_void_p),
        ("b", ctypes.c_void_p),
    ]


class HipError(RuntimeError):
    """A HIP runtime failure reported by libtcw_hip."""


class SingleAtomError(ValueError):
    """lal-strict mode: a window spans a single atom (lalpulsar's XLAL_EDOM)."""


_lib = None


def load():
    """Load libtcw_hip.so (once).  Raises OSError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(
            f"{LIB_PATH} is not built; run pyfstat_amd.build.build() (hipcc, gfx950)."
        )
    # One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64.so.7,
    # and the library's NEEDED entry resolves to whichever copy is loaded
    # first.  Load torch first so that device memory, streams and our kernels
    # all live in torch's runtime.
    import torch  # noqa: F401

    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    lib.tcw_version.restype = ctypes.c_char_p
    lib.tcw_version.argtypes = []
    lib.tcw_last_error.restype = ctypes.c_char_p
    lib.tcw_last_error.argtypes = []
    lib.tcw_device_count.restype = ctypes.c_int
    lib.tcw_device_count.argtypes = []
    lib.tcw_device_name.restype = ctypes.c_int
    lib.tcw_device_name.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    lib.tcw_map_shape.restype = ctypes.c_int
    lib.tcw_map_shape.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                  P(Window), P(ctypes.c_int64), P(ctypes.c_int64)]
    lib.tcw_create.restype = ctypes.c_int
    lib.tcw_create.argtypes = [ctypes.c_int, P(vp)]
    lib.tcw_destroy.restype = ctypes.c_int
    lib.tcw_destroy.argtypes = [vp]
    lib.tcw_map.restype = ctypes.c_int
    lib.tcw_map.argtypes = [vp, P(Atoms), P(Window), vp, P(Stats), ctypes.c_uint32, vp]
    lib.tcw_map_noalloc.restype = ctypes.c_int
    lib.tcw_map_noalloc.argtypes = [vp, P(Atoms), P(Window), vp, vp, ctypes.c_uint32, vp]
    lib.tcw_batch_plan.restype = ctypes.c_int
    lib.tcw_batch_plan.argtypes = [vp, P(BatchDesc), ctypes.c_uint32, P(vp), vp]
    lib.tcw_batch_plan_rows.restype = ctypes.c_int
    lib.tcw_batch_plan_rows.argtypes = [vp, P(BatchDesc), ctypes.c_uint32, vp, P(vp), vp]
    lib.tcw_plan_scratch_bytes.restype = ctypes.c_int64
    lib.tcw_plan_scratch_bytes.argtypes = [vp]
    lib.tcw_plan_items.restype = ctypes.c_int
    lib.tcw_plan_items.argtypes = [vp, vp]
    lib.tcw_batch_run.restype = ctypes.c_int
    lib.tcw_batch_run.argtypes = [vp, vp, vp, vp, ctypes.c_int64, vp, vp, ctypes.c_int, vp]
    lib.tcw_batch_run_timed.restype = ctypes.c_int
    lib.tcw_batch_run_timed.argtypes = [vp, vp, vp, vp, ctypes.c_int64, vp, vp, vp, vp]
    lib.tcw_plan_destroy.restype = ctypes.c_int
    lib.tcw_plan_destroy.argtypes = [vp]
    i64 = ctypes.c_int64
    lib.tcw_merge_batch.restype = ctypes.c_int
    lib.tcw_merge_batch.argtypes = [vp, P(BatchDesc), vp, vp, i64, vp, i64, vp, vp, vp, vp]
    lib.tcw_gather_bins.restype = ctypes.c_int
    lib.tcw_gather_bins.argtypes = [vp, i64, vp, i64, vp, i64, vp]
    lib.tcw_refine_plan_create.restype = ctypes.c_int
    lib.tcw_refine_plan_create.argtypes = [vp, P(RefineDesc), P(vp)]
    lib.tcw_refine_plan_destroy.restype = ctypes.c_int
    lib.tcw_refine_plan_destroy.argtypes = [vp]
    lib.tcw_refine_forward.restype = ctypes.c_int
    lib.tcw_refine_forward.argtypes = [vp, vp, i64, vp, vp, vp, vp, vp, vp]
    lib.tcw_refine_backward.restype = ctypes.c_int
    lib.tcw_refine_backward.argtypes = [vp, vp, vp, vp, vp, vp, vp, i64, vp, vp, vp]
    lib.tcw_refine_ascent.restype = ctypes.c_int
    lib.tcw_refine_ascent.argtypes = [vp, vp, i64, vp, vp, vp, vp, vp, ctypes.c_int,
                                      ctypes.c_double, vp, vp, vp]
    lib.tcw_demod_atoms.restype = ctypes.c_int
    lib.tcw_demod_atoms.argtypes = [vp, P(SftSet), i64, vp, vp, ctypes.c_double, ctypes.c_int32,
                                    vp, vp, i64, vp]
    lib.tcw_demod_atoms_sky.restype = ctypes.c_int
    lib.tcw_demod_atoms_sky.argtypes = [vp, P(SftSet), i64, vp, vp, vp, vp, i64, vp, vp,
                                        ctypes.c_double, ctypes.c_int32, vp, vp, i64, vp]
    _lib = lib
    return lib


def check(rc, what="libtcw_hip"):
    """Map a TCW_* return code to the reference's exception classes."""
    if rc == TCW_OK:
        return
    msg = load().tcw_last_error().decode(errors="replace")
    if rc == TCW_ERR_INVALID:
        raise ValueError(msg)
    if rc == TCW_ERR_SINGLE_ATOM:
        raise SingleAtomError(msg)
    if rc == TCW_ERR_NOMEM:
        raise MemoryError(f"{what}: {msg}")
    raise HipError(f"{what}: {msg}")


def window_struct(fields):
    return Window(*[int(x) for x in fields])
