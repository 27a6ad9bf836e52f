"""tcw_map_noalloc (static device scratch, no allocation) against tcw_map and
the FP64 oracle, through the C ABI.

Two calls in a row on one device, the second with a smaller input, for a
rectangular and an exponential map (the exp path uses the largest scratch
tables): every F, and every stats field, bit-identical to tcw_map's result on
the same inputs, and the F map within the parity tolerance of the oracle.  A
map that needs more than the scratch holds is refused before anything runs.
"""

import ctypes

import numpy as np
import pytest

from pyfstat_amd import _hiplib, synth
from pyfstat_amd.atoms import TRANSIENT_EXPONENTIAL, atoms_to_arrays, window_fields
from tests.helpers import check_argmax, compare_maps, oracle_for

pytestmark = pytest.mark.gpu

STATS_FIELDS = [f for f, _ in _hiplib.Stats._fields_]


def _device_atoms(multi, dev):
    import torch

    packed = atoms_to_arrays(multi)
    ts = torch.from_numpy(packed.ts.view(np.int32)).to(dev)
    cols = torch.from_numpy(packed.cols).to(dev)
    at = _hiplib.Atoms()
    at.numDet = packed.numDet
    at.TAtom = packed.TAtom
    at.ts = ts.data_ptr()
    at.cols = cols.data_ptr()
    at.col_stride = packed.cols.shape[1]
    for X in range(packed.numDet + 1):
        at.det_off[X] = int(packed.det_off[X])
    at.t_min = packed.t_min
    at.t_max = packed.t_max
    return packed, at, (ts, cols)


def _shape(lib, packed, win):
    nt0, ntau = ctypes.c_int64(), ctypes.c_int64()
    _hiplib.check(lib.tcw_map_shape(packed.num_binned, packed.TAtom, packed.t_min,
                                    ctypes.byref(win), ctypes.byref(nt0), ctypes.byref(ntau)),
                  "tcw_map_shape")
    return nt0.value, ntau.value


def _both(multi, window, flags):
    """(F, stats) from tcw_map and from tcw_map_noalloc on the same inputs."""
    import torch

    from pyfstat_amd import tcw_fstat_map_funcs as tcw

    lib = _hiplib.load()
    handle = tcw._ensure_handle()
    dev = torch.device("cuda", tcw._HipState.device)
    packed, at, keep = _device_atoms(multi, dev)
    win = _hiplib.window_struct(window_fields(window))
    shape = _shape(lib, packed, win)
    stream = torch.cuda.current_stream(dev)
    sp = ctypes.c_void_p(stream.cuda_stream)
    F_ref = torch.empty(shape, dtype=torch.float32, device=dev)
    st_ref = _hiplib.Stats()
    _hiplib.check(lib.tcw_map(handle, ctypes.byref(at), ctypes.byref(win),
                              ctypes.c_void_p(F_ref.data_ptr()), ctypes.byref(st_ref), flags, sp),
                  "tcw_map")
    # poison the outputs so that a skipped write shows
    F_na = torch.full(shape, float("nan"), dtype=torch.float32, device=dev)
    st_d = torch.full((_hiplib.STATS_BYTES,), 0xAB, dtype=torch.uint8, device=dev)
    print(f"case: noalloc window={window} flags={flags} map={shape}", flush=True)
    _hiplib.check(lib.tcw_map_noalloc(handle, ctypes.byref(at), ctypes.byref(win),
                                      ctypes.c_void_p(F_na.data_ptr()),
                                      ctypes.c_void_p(st_d.data_ptr()), flags, sp),
                  "tcw_map_noalloc")
    stream.synchronize()
    st_na = _hiplib.Stats.from_buffer_copy(st_d.cpu().numpy().tobytes())
    del keep
    return F_ref.cpu().numpy(), st_ref, F_na.cpu().numpy(), st_na


def _check_case(multi, window, BtSG=True):
    flags = _hiplib.TCW_FLAG_BTSG if BtSG else 0
    F_ref, st_ref, F_na, st_na = _both(multi, window, flags)
    np.testing.assert_array_equal(F_na.view(np.uint32), F_ref.view(np.uint32))
    for f in STATS_FIELDS:
        a, b = getattr(st_na, f), getattr(st_ref, f)
        assert (a == b) or (a != a and b != b), (f, a, b)
    om = oracle_for(multi, window, BtSG=BtSG)
    compare_maps(F_na, om)
    check_argmax(F_na, int(st_na.m_ML), int(st_na.n_ML), om)
    if BtSG:
        assert abs(st_na.lnBtSG - om.lnBtSG) <= 1e-4 + 2e-7 * abs(om.maxF)
    return F_na.shape


@pytest.mark.parametrize("kind", ["rect", "exp"])
def test_noalloc_matches_tcw_map_twice_then_smaller(hip_ready, kind):
    if kind == "rect":
        big = synth.workload_rect_plain(n_sft=500, n_t0=400, n_tau=420, seed_index=71)
        small = synth.workload_rect_plain(n_sft=120, n_t0=50, n_tau=37, seed_index=72)
    else:
        big = synth.workload_exp_plain(n_sft=900, n_t0=300, n_tau=310, seed_index=73)
        small = synth.workload_exp_plain(n_sft=150, n_t0=40, n_tau=29, seed_index=74)
    s1 = _check_case(*big)
    s2 = _check_case(*small)
    assert s2[0] * s2[1] < s1[0] * s1[1]


def test_noalloc_refuses_a_map_beyond_the_scratch(hip_ready):
    import torch

    from pyfstat_amd import tcw_fstat_map_funcs as tcw

    lib = _hiplib.load()
    handle = tcw._ensure_handle()
    dev = torch.device("cuda", tcw._HipState.device)
    # an exp map of 6000 x 6000 cells on 12000 atoms needs ~0.1 GB of E tables
    multi, window = synth.workload_exp_plain(n_sft=12000, n_t0=6000, n_tau=6000,
                                             seed_index=75)
    assert window.type == TRANSIENT_EXPONENTIAL
    packed, at, keep = _device_atoms(multi, dev)
    win = _hiplib.window_struct(window_fields(window))
    F = torch.empty(16, dtype=torch.float32, device=dev)  # never written: refused first
    st_d = torch.empty(_hiplib.STATS_BYTES, dtype=torch.uint8, device=dev)
    rc = lib.tcw_map_noalloc(handle, ctypes.byref(at), ctypes.byref(win),
                             ctypes.c_void_p(F.data_ptr()), ctypes.c_void_p(st_d.data_ptr()),
                             _hiplib.TCW_FLAG_BTSG,
                             ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream))
    assert rc == _hiplib.TCW_ERR_INVALID
    assert b"static scratch holds" in lib.tcw_last_error()
    # the scratch still serves a map that fits
    _check_case(*synth.workload_rect_plain(n_sft=90, n_t0=20, n_tau=20, seed_index=76))
    del keep
