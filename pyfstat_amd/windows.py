"""Transient window ranges from ComputeFstat-style options.

The reference builds the ``transientWindowRange_t`` that feeds the map in two
places, and a caller that uses the "hip" map without PyFstat's ``core.py``
needs both:

* ``transient_window_range`` -- once per search, from the user options
  (pyfstat/core.py:821-899): window-type string -> lalpulsar enum, default
  spacings dt0 = dtau = Tsft, integer bands (non-integers truncated with a
  warning), tauMin defaulting to 2*Tsft, and the "none" type replaced by one
  rectangular window over [minStartTime, maxStartTime];
* ``set_window_times`` -- on every map call (pyfstat/core.py:1440-1449):
  t0 = tstart, and tau = tend - tstart when tauBand == 0.

Both write the same UINT4 fields as lalpulsar's struct (``atoms.
transientWindowRange_t``, or a real ``lalpulsar.transientWindowRange_t``
passed in as ``windowRange``).
"""

import logging

from .atoms import (
    TRANSIENT_EXPONENTIAL,
    TRANSIENT_NONE,
    TRANSIENT_RECTANGULAR,
    transientWindowRange_t,
)

logger = logging.getLogger(__name__)

# user strings accepted for transientWindowType (pyfstat/core.py:829-833)
TRANSIENT_WINDOW_TYPES = {
    "none": TRANSIENT_NONE,
    "rect": TRANSIENT_RECTANGULAR,
    "exp": TRANSIENT_EXPONENTIAL,
}


def _as_int(name, value):
    # core.py:858-865 (t0Band), 869-876 (tauBand), 884-891 (tauMin): anything
    # that is not a Python int is truncated by int() with a warning
    if isinstance(value, int):
        return value
    logger.warning("Casting non-integer {}={} to int...".format(name, value))
    return int(value)


def transient_window_range(transientWindowType, Tsft, minStartTime=None, maxStartTime=None,
                           t0Band=None, tauBand=None, tauMin=None, dt0=None, dtau=None,
                           windowRange=None):
    """The search-wide window range of ComputeFstat (pyfstat/core.py:821-899).

    Parameters follow ComputeFstat's attributes of the same names.  Returns
    ``windowRange`` (a new ``transientWindowRange_t`` unless one is passed
    in) with type, spacings and bands set; t0 is set here only for "none"
    (elsewhere by ``set_window_times`` at call time), tau for "none" and,
    when tauBand is given, from tauMin.

    Raises ValueError for an unknown type string, with the reference's text.
    """
    if transientWindowType not in TRANSIENT_WINDOW_TYPES:
        raise ValueError(
            "Unknown window-type ({}) passed as input, [{}] allows.".format(
                transientWindowType, ", ".join(TRANSIENT_WINDOW_TYPES)
            )
        )
    wr = transientWindowRange_t() if windowRange is None else windowRange
    wr.type = TRANSIENT_WINDOW_TYPES[transientWindowType]
    # default spacing (core.py:843-845)
    wr.dt0 = Tsft
    wr.dtau = Tsft
    if wr.type == TRANSIENT_NONE:
        # one rectangular window spanning all the data (core.py:847-853)
        if minStartTime is None or maxStartTime is None:
            raise ValueError("window type 'none' needs minStartTime and maxStartTime")
        wr.t0 = int(minStartTime)
        wr.t0Band = 0
        wr.tau = int(maxStartTime - minStartTime)
        wr.tauBand = 0
        return wr
    # user-set bands and spacings (core.py:854-892); a spacing is taken from
    # the options only together with its band, and tauMin only with tauBand
    if t0Band is None:
        wr.t0Band = 0
    else:
        wr.t0Band = _as_int("t0Band", t0Band)
        if dt0:
            wr.dt0 = dt0
    if tauBand is None:
        wr.tauBand = 0
    else:
        wr.tauBand = _as_int("tauBand", tauBand)
        if dtau:
            wr.dtau = dtau
        if tauMin is None:
            wr.tau = int(2 * Tsft)
        else:
            wr.tau = _as_int("tauMin", tauMin)
    return wr


def set_window_times(windowRange, tstart=None, tend=None, minStartTime=None,
                     maxStartTime=None):
    """Per-call window start and length (pyfstat/core.py:1440-1449).

    ``tstart``/``tend`` default to ``minStartTime``/``maxStartTime`` (a falsy
    value such as 0 falls back too, as ``tstart or self.minStartTime`` does);
    sets ``windowRange.t0 = int(tstart)`` and, when ``tauBand == 0``,
    ``windowRange.tau = int(tend - tstart)``.  Modifies and returns
    ``windowRange``, as ComputeFstat does with its own range.
    """
    tstart = tstart or minStartTime
    tend = tend or maxStartTime
    if tstart is None or tend is None:
        raise ValueError("Need tstart or self.minStartTime, and tend or self.maxStartTime!")
    windowRange.t0 = int(tstart)
    if windowRange.tauBand == 0:
        windowRange.tau = int(tend - tstart)
    return windowRange
