"""Atoms from SFTs on the device (tcw_demod_atoms) against the FP64 oracle
(oracle/demod_ref.py), and the chain SFTs -> device atoms -> batched
transient maps.

Tolerances: the kernel computes the phase in FP64 and the kernel weights,
the complex sum and the rotation in float32 (float32 atoms): atoms within
2e-6 of the template's largest |Fa|, |Fb| (a float32 emulation of the
kernel's operation order reaches ~2e-7 on these cases; phases of ~1e5 cycles
at these sizes, so the FP64 fractional cycle is exact to ~1e-11); a2, b2, ab
are the float32 products of the float32 antenna patterns, bit-exact.  Maps from
device atoms equal maps from the same atoms copied through the host
(bit-exact: the same kernels on the same numbers)."""

from types import SimpleNamespace

import numpy as np
import pytest

import oracle
from oracle import demod_ref as dr
from pyfstat_amd import synth
from pyfstat_amd.atoms import TRANSIENT_RECTANGULAR, PackedAtoms
from tests.helpers import compare_maps

pytestmark = pytest.mark.gpu

# |atom - oracle| / max over the template's SFTs of |Fa|, |Fb|
ATOL_REL = 2e-6


def _oracle_atoms(sft, freqs, fdots, dterms=8):
    return dr.demod_atoms(sft.t_start, sft.bins.astype(np.complex128), sft.k_min, sft.Tsft,
                          sft.delay, sft.delay_rate, sft.a, sft.b, freqs, fdots, sft.ref_time,
                          dterms=dterms)


def _check_atoms(got_ts, got_cols, sft, freqs, fdots, dterms=8):
    a2, b2, ab, Fa, Fb = _oracle_atoms(sft, freqs, fdots, dterms)
    n = sft.t_start.size
    T = len(freqs)
    assert np.array_equal(got_ts, np.tile(sft.t_start, T))
    cols = got_cols.reshape(7, T, n)
    a32 = sft.a.astype(np.float32)
    b32 = sft.b.astype(np.float32)
    assert np.array_equal(cols[0], np.broadcast_to(a32 * a32, (T, n)))
    assert np.array_equal(cols[1], np.broadcast_to(b32 * b32, (T, n)))
    assert np.array_equal(cols[2], np.broadcast_to(a32 * b32, (T, n)))
    for t in range(T):
        scale = max(np.abs(Fa[t]).max(), np.abs(Fb[t]).max(), 1e-30)
        for c, ref in ((3, Fa[t].real), (4, Fa[t].imag), (5, Fb[t].real), (6, Fb[t].imag)):
            err = np.abs(cols[c, t] - ref).max() / scale
            assert err < ATOL_REL, (t, c, err)


def test_atoms_vs_oracle(hip_ready):
    """Two detectors with gaps, noise plus a signal, templates around the
    injection and across the band, fdot != 0, dterms 8 and 3."""
    from pyfstat_amd.demod import demod_atoms

    inj = dict(f=100.0213, fdot=-3e-9, amp_a=0.4 + 0.2j, amp_b=0.1 - 0.3j)
    sft = synth.synthetic_sfts(n_det=2, n_sft=120, seed_index=70, inj=inj)
    sfts = sft.sft_set()
    freqs = np.array([inj["f"], inj["f"] + 2e-5, 100.0005, 100.049])
    fdots = np.array([inj["fdot"], 0.0, 1e-9, -5e-9])
    for dterms in (8, 3):
        print(f"case: atoms vs oracle, dterms {dterms}", flush=True)
        at = demod_atoms(sfts, freqs, fdots, sft.ref_time, dterms=dterms)
        ts, cols = at.host()
        _check_atoms(ts, cols, sft, freqs, fdots, dterms)


def test_band_edge_and_bin_centre(hip_ready):
    """Templates whose kernel window runs past the stored band (those bins
    count as zero), and, for SFTs with zero Doppler, templates whose observed
    frequency sits on a bin (the single-bin branch) or 1e-7 bins above or
    below one (delta near 0 and near 1: the weights' small denominators)."""
    from pyfstat_amd.demod import SFTSet, demod_atoms

    sft = synth.synthetic_sfts(n_det=1, n_sft=40, seed_index=71)
    sfts = sft.sft_set()
    f_lo = (sft.k_min + 2) / sft.Tsft
    f_hi = (sft.k_min + sft.bins.shape[1] - 3) / sft.Tsft
    at = demod_atoms(sfts, [f_lo, f_hi], [0.0, 0.0], sft.ref_time)
    ts, cols = at.host()
    _check_atoms(ts, cols, sft, [f_lo, f_hi], [0.0, 0.0])
    # zero delay rate: kappa = f Tsft exactly on bin k_min + 20
    zero = np.zeros(sft.t_start.size)
    flat = SFTSet(sft.t_start, sft.bins, sft.k_min, sft.Tsft, zero, zero, sft.a, sft.b,
                  sft.det_off)
    k0 = sft.k_min + 20
    fs = [k0 / sft.Tsft, (k0 + 1e-7) / sft.Tsft, (k0 - 1e-7) / sft.Tsft]
    at = demod_atoms(flat, fs, [0.0] * 3, sft.ref_time)
    ts, cols = at.host()
    a2, b2, ab, Fa, Fb = dr.demod_atoms(sft.t_start, sft.bins.astype(np.complex128), sft.k_min,
                                        sft.Tsft, zero, zero, sft.a, sft.b,
                                        fs, [0.0] * 3, sft.ref_time)
    n = sft.t_start.size
    for t in range(3):
        print(f"case: zero Doppler, kappa - k0 = {fs[t] * sft.Tsft - k0:.1e}", flush=True)
        scale = max(np.abs(Fa[t]).max(), np.abs(Fb[t]).max())
        for c, ref in ((3, Fa[t].real), (4, Fa[t].imag), (5, Fb[t].real), (6, Fb[t].imag)):
            err = np.abs(cols[c, t * n:(t + 1) * n] - ref).max() / scale
            assert err < ATOL_REL, (t, c, err)


@pytest.mark.parametrize("T", [2, 5, 40])
def test_template_tile_sizes(hip_ready, T):
    """Every tile shape the host picks (1, 8 and 64 template slots a
    workgroup for 2, 5 and 40 templates) with dterms 8 (unrolled) and 5
    (runtime run lengths), against the oracle; 40 templates leave a partial
    64-template tile and 260 SFTs a partial 64-SFT tile."""
    from pyfstat_amd.demod import demod_atoms

    sft = synth.synthetic_sfts(n_det=2, n_sft=130, gap_frac=0.0, seed_index=75)
    sfts = sft.sft_set()
    rng = np.random.default_rng(75 + T)
    freqs = rng.uniform(100.0, 100.05, T)
    fdots = rng.uniform(-5e-9, 5e-9, T)
    for dterms in (8, 5):
        print(f"case: {T} templates, dterms {dterms}", flush=True)
        at = demod_atoms(sfts, freqs, fdots, sft.ref_time, dterms=dterms)
        ts, cols = at.host()
        _check_atoms(ts, cols, sft, freqs, fdots, dterms)


def test_device_atoms_feed_batched_maps(hip_ready):
    """SFTs -> device atoms -> batched rect maps without a host copy: the maps
    equal those of the same atoms uploaded from the host, match the oracle's
    maps of those atoms, and the injected transient's template has the
    largest maxF at the injected start and duration."""
    from pyfstat_amd.batched import BatchedTransientMaps
    from pyfstat_amd.demod import batch_from_sfts, demod_atoms

    TS = synth.TSFT
    t_on, dur = synth.T_START + 40 * TS, 30 * TS
    inj = dict(f=100.02, fdot=0.0, amp_a=2.0, amp_b=0.8j, window=("rect", t_on, dur))
    sft = synth.synthetic_sfts(n_det=2, n_sft=160, seed_index=72, inj=inj)
    sfts = sft.sft_set()
    # the off-signal templates: 3 resolution widths of the transient away, and
    # far in the band
    freqs = np.array([inj["f"] - 3.0 / dur, inj["f"], 100.003])
    fdots = np.zeros(3)
    win = synth.grid_window(TRANSIENT_RECTANGULAR, synth.T_START, 100, 60)
    b = batch_from_sfts(sfts, freqs, fdots, sft.ref_time, win, BtSG=True)
    b.run()
    maps = b.maps()
    at = demod_atoms(sfts, freqs, fdots, sft.ref_time)
    ts, cols = at.host()
    h = BatchedTransientMaps(ts, cols, at.seg_off, at.seg_begin, [win] * 3, at.TAtom, BtSG=True)
    h.run()
    hmaps = h.maps()
    for t in range(3):
        print(f"case: device atoms, template {t}", flush=True)
        assert np.array_equal(maps[t].F_mn, hmaps[t].F_mn)
        assert maps[t].maxF == hmaps[t].maxF and maps[t].lnBtSG == hmaps[t].lnBtSG
        s0, s1 = at.seg_begin[t], at.seg_begin[t + 1]
        a, e = at.seg_off[s0], at.seg_off[s1]
        packed = PackedAtoms(at.TAtom, ts[a:e], np.ascontiguousarray(cols[:, a:e]),
                             at.seg_off[s0:s1 + 1] - a)
        om = oracle.compute_transient_fstat_map(packed, win, BtSG=True)
        compare_maps(maps[t].F_mn, om)
    assert maps[1].maxF > 2 * max(maps[0].maxF, maps[2].maxF)
    m, n = maps[1].get_maxF_idx()
    t0_ml = win.t0 + int(m) * win.dt0
    tau_ml = win.tau + int(n) * win.dtau
    # the start is sharp; the duration's likelihood is flat near the end of
    # the signal at this SNR
    assert abs(t0_ml - t_on) <= 2 * TS and abs(tau_ml - dur) <= 0.3 * dur
    b.close()
    h.close()


def test_validation(hip_ready):
    """Argument checks before anything is queued."""
    from pyfstat_amd.demod import SFTSet, demod_atoms

    sft = synth.synthetic_sfts(n_det=1, n_sft=10, seed_index=73)
    sfts = sft.sft_set()
    with pytest.raises(ValueError, match="dterms"):
        demod_atoms(sfts, [100.01], [0.0], sft.ref_time, dterms=0)
    with pytest.raises(ValueError, match="dterms"):
        demod_atoms(sfts, [100.01], [0.0], sft.ref_time, dterms=65)
    with pytest.raises(ValueError, match="same length"):
        demod_atoms(sfts, [100.01, 100.02], [0.0], sft.ref_time)
    with pytest.raises(ValueError, match="ascending"):
        SFTSet(sft.t_start[::-1], sft.bins, sft.k_min, sft.Tsft, sft.delay, sft.delay_rate,
               sft.a, sft.b, sft.det_off)
    with pytest.raises(ValueError, match="Tsft"):
        SFTSet(sft.t_start, sft.bins, sft.k_min, 1, sft.delay, sft.delay_rate, sft.a, sft.b,
               sft.det_off)
    with pytest.raises(ValueError, match="band"):
        SFTSet(sft.t_start, sft.bins, -1, sft.Tsft, sft.delay, sft.delay_rate, sft.a, sft.b,
               sft.det_off)
    # the C ABI checks the band itself (a descriptor filled by another caller)
    sfts.desc.k_min = 2**31 - 10
    with pytest.raises(ValueError, match="band"):
        demod_atoms(sfts, [100.01], [0.0], sft.ref_time)
    sfts.desc.k_min = sft.k_min


@pytest.mark.parametrize("T", [3, 40])
def test_sky_points_in_one_launch(hip_ready, T):
    """tcw_demod_atoms_sky: three sky points (their own per-SFT delay, rate
    and antenna patterns) in one launch equal, bit for bit, three single-sky
    launches with those arrays in the SFT set, and match the oracle."""
    from pyfstat_amd.demod import SFTSet, demod_atoms, demod_atoms_sky

    sft = synth.synthetic_sfts(n_det=2, n_sft=70, seed_index=76)
    sfts = sft.sft_set()
    n = sft.t_start.size
    rng = np.random.default_rng(76 + T)
    n_sky = 3
    delay = np.stack([sft.delay * (1 - 0.3 * s) + 0.01 * s for s in range(n_sky)])
    rate = np.stack([sft.delay_rate * (1 + 0.5 * s) for s in range(n_sky)])
    a = np.stack([np.cos(0.7 * s) * sft.a - np.sin(0.7 * s) * sft.b for s in range(n_sky)])
    b = np.stack([np.sin(0.7 * s) * sft.a + np.cos(0.7 * s) * sft.b for s in range(n_sky)])
    freqs = rng.uniform(100.0, 100.05, T)
    fdots = rng.uniform(-5e-9, 5e-9, T)
    at = demod_atoms_sky(sfts, delay, rate, a, b, freqs, fdots, sft.ref_time)
    assert at.T == n_sky * T
    ts, cols = at.host()
    for s in range(n_sky):
        print(f"case: {T} templates, sky point {s}", flush=True)
        one = SFTSet(sft.t_start, sft.bins, sft.k_min, sft.Tsft, delay[s], rate[s], a[s], b[s],
                     sft.det_off)
        ts1, cols1 = demod_atoms(one, freqs, fdots, sft.ref_time).host()
        blk = slice(s * T * n, (s + 1) * T * n)
        assert np.array_equal(ts[blk], ts1)
        assert np.array_equal(cols[:, blk], cols1)
        view = SimpleNamespace(t_start=sft.t_start, bins=sft.bins, k_min=sft.k_min,
                               Tsft=sft.Tsft, delay=delay[s], delay_rate=rate[s], a=a[s],
                               b=b[s], ref_time=sft.ref_time)
        _check_atoms(ts1, cols1, view, freqs, fdots)
    with pytest.raises(ValueError, match="n_sky, n_sft"):
        demod_atoms_sky(sfts, delay[:, :-1], rate, a, b, freqs, fdots, sft.ref_time)
    with pytest.raises(ValueError, match="same number of sky points"):
        demod_atoms_sky(sfts, delay, rate[:2], a, b, freqs, fdots, sft.ref_time)
    # the C ABI's own check: 1 .. 65535 sky points
    with pytest.raises(ValueError, match="sky points must number"):
        demod_atoms_sky(sfts, delay[:0], rate[:0], a[:0], b[:0], freqs, fdots, sft.ref_time)
