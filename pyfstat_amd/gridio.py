"""Grid-search output for a batch of transient templates: the results table of
TransientGridSearch (column names, order and text format) filled from one
batched device pass, the per-template F_mn file names, and a binary file that
holds every map of a batch.

Reference layout followed here:
  column order         TransientGridSearch._set_output_keys
                       (pyfstat/grid_based_searches.py:1023-1042)
  column values        TransientGridSearch._get_single_cand_results (:1080-1140):
                       t0_ML = windowRange.t0 + m*dt0, tau_ML = windowRange.tau +
                       n*dtau from the map's argmax; t0_MP/tau_MP from the map
  text file            GridSearch.save_array_to_disk (:471-501): np.savetxt of
                       np.nan_to_num(data) (a no-op on a structured array),
                       " " delimiter, header = header lines + the keys line
  formats              TransientGridSearch._get_savetxt_fmt_dict (:1232-1253),
                       fmt_doppler "%.16g" / fmt_detstat "%.9g" (core.py:123-126),
                       Doppler key list (utils/formatting.py:63-98)
  F_mn file names      TransientGridSearch.get_transient_fstat_map_filename
                       (:1201-1230)
The npz file has no counterpart in the reference (whose only map output is the
text writer, tcw_fstat_map_funcs.py:289-317, one file per template).
"""

import numpy as np

DEFAULT_SEARCH_KEYS = ["F0", "F1", "F2", "Alpha", "Delta"]
PULSAR_MAX_SPINS = 7  # lalpulsar's PULSAR_MAX_SPINS: F0 ... F6
DOPPLER_KEYS = [f"F{k}" for k in range(PULSAR_MAX_SPINS)] + [
    "Alpha", "Delta", "asini", "period", "ecc", "tp", "argp"]
FMT_DOPPLER = "%.16g"
FMT_DETSTAT = "%.9g"


def output_keys(search_keys=None, detector_names=(), singleFstats=False, transient=True,
                twoFXatMaxTwoF=False, detstat="maxTwoF", BtSG=False):
    """Column names in TransientGridSearch order: search keys, twoF, twoF{IFO}
    (singleFstats), maxTwoF (transient), twoF{IFO}atMaxTwoF, the detection
    statistic unless it is twoF/maxTwoF, then t0_ML, tau_ML (transient) and
    t0_MP, tau_MP (BtSG)."""
    keys = list(DEFAULT_SEARCH_KEYS if search_keys is None else search_keys)
    keys.append("twoF")
    if singleFstats:
        keys += [f"twoF{IFO}" for IFO in detector_names]
    if transient:
        keys.append("maxTwoF")
    if twoFXatMaxTwoF:
        keys += [f"twoF{IFO}atMaxTwoF" for IFO in detector_names]
    if detstat not in ("twoF", "maxTwoF"):
        keys.append(detstat)
    if transient:
        keys += ["t0_ML", "tau_ML"]
        if BtSG:
            keys += ["t0_MP", "tau_MP"]
    return keys


def savetxt_formats(keys, fmt_doppler=FMT_DOPPLER, fmt_detstat=FMT_DETSTAT):
    """One printf format per column, as _get_savetxt_fmt_dict assigns them:
    Doppler keys fmt_doppler, 2F-like statistics fmt_detstat, t0/tau "%d".
    Raises KeyError for a column it has no format for (the reference raises
    when the format list and the columns disagree)."""
    fmts = []
    for k in keys:
        if k in DOPPLER_KEYS:
            fmts.append(fmt_doppler)
        elif k in ("t0_ML", "tau_ML", "t0_MP", "tau_MP"):
            fmts.append("%d")
        elif k.startswith("twoF") or k in ("maxTwoF", "lnBtSG", "log10BSGL", "BSGL"):
            fmts.append(fmt_detstat)
        else:
            raise KeyError(f"no output format for column {k!r}")
    return fmts


def results_table(params, detstats, windows, twoF=None, twoFX=None, detector_names=(),
                  detstat="maxTwoF", BtSG=None, search_keys=None):
    """Structured float array with one row per template.

    params   : {search key: array [T]} (F0, F1, F2, Alpha, Delta, ...)
    detstats : the dict of pyfstat_amd.detstats.transient_detstats (maxTwoF,
               m_ML/n_ML or t0_ML/tau_ML, lnBtSG, t0_MP, tau_MP, twoFXatMaxTwoF)
    windows  : the caller's transientWindowRange_t per template (or one for
               all); t0_ML/tau_ML are recomputed from it and the argmax index,
               as the grid search does (grid_based_searches.py:1131-1133)
    twoF     : the fully coherent 2F per template (computed outside the map
               path); NaN if not given
    twoFX    : [T, X] per-detector coherent 2F for singleFstats columns
    """
    T = len(detstats["maxTwoF"])
    if not isinstance(windows, (list, tuple)):
        windows = [windows] * T
    if BtSG is None:
        BtSG = "t0_MP" in detstats
    fx_at_max = "twoFXatMaxTwoF" in detstats and len(detector_names) > 0
    keys = output_keys(search_keys if search_keys is not None else list(params),
                       detector_names, singleFstats=twoFX is not None, transient=True,
                       twoFXatMaxTwoF=fx_at_max, detstat=detstat, BtSG=BtSG)
    data = np.zeros(T, dtype={"names": keys, "formats": [float] * len(keys)})
    for k in params:
        data[k] = np.asarray(params[k], dtype=float)
    data["twoF"] = np.full(T, np.nan) if twoF is None else np.asarray(twoF, float)
    if twoFX is not None:
        for X, IFO in enumerate(detector_names):
            data[f"twoF{IFO}"] = np.asarray(twoFX, float)[:, X]
    data["maxTwoF"] = np.asarray(detstats["maxTwoF"], float)
    if fx_at_max:
        for X, IFO in enumerate(detector_names):
            data[f"twoF{IFO}atMaxTwoF"] = np.asarray(detstats["twoFXatMaxTwoF"], float)[:, X]
    if detstat not in ("twoF", "maxTwoF"):
        data[detstat] = np.asarray(detstats[detstat], float)
    m = np.asarray(detstats["m_ML"], dtype=np.int64)
    n = np.asarray(detstats["n_ML"], dtype=np.int64)
    data["t0_ML"] = [w.t0 + int(mm) * w.dt0 for w, mm in zip(windows, m)]
    data["tau_ML"] = [w.tau + int(nn) * w.dtau for w, nn in zip(windows, n)]
    if BtSG:
        data["t0_MP"] = np.asarray(detstats["t0_MP"], float)
        data["tau_MP"] = np.asarray(detstats["tau_MP"], float)
    return data


def write_results(path, data, header_lines=(), fmt_doppler=FMT_DOPPLER,
                  fmt_detstat=FMT_DETSTAT):
    """The grid search's results file: np.savetxt of np.nan_to_num(data) with a
    " " delimiter (the reference's call; on a structured array nan_to_num
    returns the data unchanged, so NaN is written as "nan"); the header is the
    header lines, then the column names."""
    keys = list(data.dtype.names)
    header = "\n".join(header_lines)
    header += "\n" + " ".join(keys)
    np.savetxt(path, np.nan_to_num(data), delimiter=" ", header=header,
               fmt=savetxt_formats(keys, fmt_doppler, fmt_detstat))


def read_results(path):
    """Structured array back from write_results' text (names from the last
    header line)."""
    names = None
    with open(path) as f:
        for line in f:
            if not line.startswith("#"):
                break
            names = line[1:].split()
    return np.genfromtxt(path, comments="#", names=names, dtype=float, ndmin=1)


def transient_fstat_map_filename(tCWfilebase, param_point, keys=None):
    """tCWfilebase + "F0_Alpha_Delta_F1_F2.dat" with each value as "{:.16g}";
    param_point is a dict, or a sequence / structured row ordered as keys."""
    fmt_keys = ["F0", "Alpha", "Delta", "F1", "F2"]
    if isinstance(param_point, np.void):
        param_point = {k: param_point[k] for k in param_point.dtype.names}
    if isinstance(param_point, dict):
        vals = [param_point[k] for k in fmt_keys]
    elif isinstance(param_point, (list, tuple, np.ndarray)):
        keys = list(DEFAULT_SEARCH_KEYS if keys is None else keys)
        vals = [list(param_point)[keys.index(k)] for k in fmt_keys]
    else:
        raise ValueError("param_point must be a dict, list, tuple or numpy array!")
    return tCWfilebase + "{:.16g}_{:.16g}_{:.16g}_{:.16g}_{:.16g}".format(*vals) + ".dat"


# ----------------------------------------------------------------------------
# all maps of a batch in one binary file
# ----------------------------------------------------------------------------

_WIN_FIELDS = ("type", "t0", "t0Band", "dt0", "tau", "tauBand", "dtau")


def write_maps_npz(path, batch, windows=None):
    """Every map of a BatchedTransientMaps (after run()) in one .npz: F (the
    concatenated float32 maps, F not 2F), F_off, shapes [T, 2], the windows
    [T, 7] and the per-template stats -- one device-to-host copy instead of a
    text writer per template."""
    wins = batch.windows if windows is None else windows
    st = batch.stats()
    np.savez(path, F=batch.F_host(), F_off=np.asarray(batch.F_off, np.int64),
             shapes=np.stack([np.asarray(st["N_t0"]), np.asarray(st["N_tau"])], axis=1),
             windows=np.array([[int(getattr(w, f)) for f in _WIN_FIELDS] for w in wins],
                              dtype=np.int64),
             **{f"stat_{k}": np.ascontiguousarray(st[k]) for k in st.dtype.names})


def read_maps_npz(path):
    """(maps, windows, stats) from write_maps_npz: maps is a list of
    [N_t0, N_tau] float32 arrays, windows a list of transientWindowRange_t."""
    from .atoms import transientWindowRange_t

    with np.load(path, allow_pickle=False) as z:
        F, off, shapes = z["F"], z["F_off"], z["shapes"]
        maps = [F[off[t]:off[t + 1]].reshape(int(shapes[t, 0]), int(shapes[t, 1]))
                for t in range(len(shapes))]
        windows = [transientWindowRange_t(*[int(x) for x in row]) for row in z["windows"]]
        stats = {k[5:]: z[k] for k in z.files if k.startswith("stat_")}
    return maps, windows, stats
