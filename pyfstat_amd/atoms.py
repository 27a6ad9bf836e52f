"""F-statistic atoms and transient window ranges: the inputs of the transient map.

These containers stand in for the lalpulsar SWIG types that the reference feeds
into ``call_compute_transient_fstat_map`` (pyfstat/tcw_fstat_map_funcs.py:495-544):

* ``FstatAtomVector`` / ``MultiFstatAtomVector`` mirror lalpulsar's
  ``FstatAtomVector`` / ``MultiFstatAtomVector`` (fields ``length``, ``TAtom``,
  ``data[X].data[i].{timestamp,a2_alpha,b2_alpha,ab_alpha,Fa_alpha,Fb_alpha}``,
  used at pyfstat/tcw_fstat_map_funcs.py:604-639,704-728 and filled atom by
  atom in tests/test_tcw_fstat_map_funcs.py:65-78), but keep every field in a
  numpy column so that packing for the device is a copy, not a per-atom loop.
* ``transientWindowRange_t`` mirrors lalpulsar's struct of UINT4 fields
  ``type, t0, t0Band, dt0, tau, tauBand, dtau`` (pyfstat/core.py:829-899).

Real lalpulsar objects are accepted too: ``atoms_to_arrays`` falls back to the
per-atom attribute walk of pyfstat/tcw_fstat_map_funcs.py:627-638 for them.
"""

import numpy as np

# lalpulsar transientWindowType_t values (pyfstat/core.py:829-833 maps the
# user strings onto these; tcw.py:691-697 checks type < TRANSIENT_LAST).
TRANSIENT_NONE = 0
TRANSIENT_RECTANGULAR = 1
TRANSIENT_EXPONENTIAL = 2
TRANSIENT_LAST = 3

# e-folding truncation of the exponential window, t1 = t0 + 3*tau
# (pyfstat/pyCUDAkernels/cudaTransientFstatExpWindow.cu:22,53).
TRANSIENT_EXP_EFOLDING = 3

# lalpulsar PULSAR_MAX_DETECTORS; also the device merge's detector limit.
PULSAR_MAX_DETECTORS = 10

ATOM_COLUMNS = (
    "a2_alpha",
    "b2_alpha",
    "ab_alpha",
    "Fa_alpha_re",
    "Fa_alpha_im",
    "Fb_alpha_re",
    "Fb_alpha_im",
)
"""Column order of the packed float32 atom matrix (tcw.py:711-721)."""

_UINT32_MAX = 0xFFFFFFFF


def _as_uint4(name, value):
    v = int(value)
    if v < 0 or v > _UINT32_MAX:
        raise ValueError(f"windowRange.{name}={value} does not fit a UINT4")
    return v


class transientWindowRange_t:
    """Transient window range: UINT4 fields as in lalpulsar.

    The map covers t0 in [t0, t0+t0Band] in steps of dt0 and tau in
    [tau, tau+tauBand] in steps of dtau (tcw.py:773-781).
    """

    _fields = ("type", "t0", "t0Band", "dt0", "tau", "tauBand", "dtau")

    def __init__(self, type=TRANSIENT_NONE, t0=0, t0Band=0, dt0=0, tau=0,
                 tauBand=0, dtau=0):
        self.type = type
        self.t0 = t0
        self.t0Band = t0Band
        self.dt0 = dt0
        self.tau = tau
        self.tauBand = tauBand
        self.dtau = dtau

    def __setattr__(self, name, value):
        if name in self._fields:
            value = _as_uint4(name, value)
        object.__setattr__(self, name, value)

    def copy(self):
        return transientWindowRange_t(**{f: getattr(self, f) for f in self._fields})

    def __repr__(self):
        inner = ", ".join(f"{f}={getattr(self, f)}" for f in self._fields)
        return f"transientWindowRange_t({inner})"


def window_fields(windowRange):
    """Read the 7 UINT4 fields from a window range (ours or lalpulsar's)."""
    return tuple(
        _as_uint4(f, getattr(windowRange, f)) for f in transientWindowRange_t._fields
    )


class _AtomView:
    """One atom of a FstatAtomVector; reads and writes go to the columns."""

    __slots__ = ("_vec", "_i")

    def __init__(self, vec, i):
        object.__setattr__(self, "_vec", vec)
        object.__setattr__(self, "_i", i)

    @property
    def timestamp(self):
        return int(self._vec.timestamp[self._i])

    @property
    def a2_alpha(self):
        return float(self._vec.a2_alpha[self._i])

    @property
    def b2_alpha(self):
        return float(self._vec.b2_alpha[self._i])

    @property
    def ab_alpha(self):
        return float(self._vec.ab_alpha[self._i])

    @property
    def Fa_alpha(self):
        return complex(self._vec.Fa_alpha[self._i])

    @property
    def Fb_alpha(self):
        return complex(self._vec.Fb_alpha[self._i])

    def __setattr__(self, name, value):
        if name not in ("timestamp", "a2_alpha", "b2_alpha", "ab_alpha",
                        "Fa_alpha", "Fb_alpha"):
            raise AttributeError(f"FstatAtom has no field {name}")
        if name == "timestamp":
            value = _as_uint4("timestamp", value)
        getattr(self._vec, name)[self._i] = value


class _AtomList:
    """``vec.data``: indexable/iterable atom views; assigning an atom copies it."""

    __slots__ = ("_vec",)

    def __init__(self, vec):
        self._vec = vec

    def __len__(self):
        return self._vec.length

    def __getitem__(self, i):
        n = self._vec.length
        if i < 0:
            i += n
        if not 0 <= i < n:
            raise IndexError(i)
        return _AtomView(self._vec, i)

    def __setitem__(self, i, atom):
        view = self[i]
        for f in ("timestamp", "a2_alpha", "b2_alpha", "ab_alpha", "Fa_alpha",
                  "Fb_alpha"):
            setattr(view, f, getattr(atom, f))

    def __iter__(self):
        for i in range(self._vec.length):
            yield _AtomView(self._vec, i)


class FstatAtomVector:
    """Per-detector atoms: uint32 timestamps, float32 a2/b2/ab, complex64 Fa/Fb."""

    def __init__(self, length=0, TAtom=0):
        self.TAtom = int(TAtom)
        self.timestamp = np.zeros(length, dtype=np.uint32)
        self.a2_alpha = np.zeros(length, dtype=np.float32)
        self.b2_alpha = np.zeros(length, dtype=np.float32)
        self.ab_alpha = np.zeros(length, dtype=np.float32)
        self.Fa_alpha = np.zeros(length, dtype=np.complex64)
        self.Fb_alpha = np.zeros(length, dtype=np.complex64)

    @classmethod
    def from_arrays(cls, TAtom, timestamp, a2, b2, ab, Fa, Fb):
        n = len(timestamp)
        v = cls(n, TAtom)
        v.timestamp[:] = np.asarray(timestamp, dtype=np.uint64).astype(np.uint32)
        v.a2_alpha[:] = a2
        v.b2_alpha[:] = b2
        v.ab_alpha[:] = ab
        v.Fa_alpha[:] = Fa
        v.Fb_alpha[:] = Fb
        return v

    @property
    def length(self):
        return int(self.timestamp.shape[0])

    @property
    def data(self):
        return _AtomList(self)

    def columns(self):
        """float32 [7, length] in ATOM_COLUMNS order."""
        out = np.empty((7, self.length), dtype=np.float32)
        out[0] = self.a2_alpha
        out[1] = self.b2_alpha
        out[2] = self.ab_alpha
        out[3] = self.Fa_alpha.real
        out[4] = self.Fa_alpha.imag
        out[5] = self.Fb_alpha.real
        out[6] = self.Fb_alpha.imag
        return out


class MultiFstatAtomVector:
    """Atoms of several detectors: ``length`` detectors, ``data[X]`` each."""

    def __init__(self, numDetectors=0):
        self.data = [None] * int(numDetectors)

    @property
    def length(self):
        return len(self.data)


def CreateFstatAtomVector(length):
    """Counterpart of lalpulsar.CreateFstatAtomVector (tests/test_tcw_fstat_map_funcs.py:66)."""
    return FstatAtomVector(length)


def CreateMultiFstatAtomVector(numDetectors):
    """Counterpart of lalpulsar.CreateMultiFstatAtomVector (tests/test_tcw_fstat_map_funcs.py:65)."""
    return MultiFstatAtomVector(numDetectors)


# One SWIG atom's fields as collected by _single_to_arrays (REAL4 fields land
# in float32 exactly; COMPLEX8 parts go through complex128 and back exactly).
_SWIG_ATOM = np.dtype([("t", np.uint32), ("a2", np.float32), ("b2", np.float32),
                       ("ab", np.float32), ("Fa", np.complex128), ("Fb", np.complex128)])


def _single_to_arrays(vec):
    """(timestamps uint32 [n], columns float32 [7, n]) of one detector."""
    if isinstance(vec, FstatAtomVector):
        return vec.timestamp.copy(), vec.columns()
    # lalpulsar SWIG object: the atoms' fields are only reachable one atom at a
    # time (tcw.py:627-638 walks them into a matrix element by element); one
    # pass collects each atom's six fields as a record, and the columns are
    # split off the record array as whole vectors.
    n = int(vec.length)
    rec = np.fromiter(
        ((a.timestamp, a.a2_alpha, a.b2_alpha, a.ab_alpha, a.Fa_alpha, a.Fb_alpha)
         for a in vec.data),
        dtype=_SWIG_ATOM, count=n)
    cols = np.empty((7, n), dtype=np.float32)
    cols[0] = rec["a2"]
    cols[1] = rec["b2"]
    cols[2] = rec["ab"]
    cols[3] = rec["Fa"].real
    cols[4] = rec["Fa"].imag
    cols[5] = rec["Fb"].real
    cols[6] = rec["Fb"].imag
    return np.ascontiguousarray(rec["t"]), cols


class PackedAtoms:
    """Detector-concatenated atoms ready for the device.

    ``ts`` uint32 [total], ``cols`` float32 [7, total], ``det_off`` int64
    [numDet+1]; TAtom is common to all detectors.
    """

    __slots__ = ("TAtom", "ts", "cols", "det_off")

    def __init__(self, TAtom, ts, cols, det_off):
        self.TAtom = int(TAtom)
        self.ts = ts
        self.cols = cols
        self.det_off = det_off

    @property
    def numDet(self):
        return len(self.det_off) - 1

    @property
    def t_min(self):
        return int(min(self.ts[self.det_off[X]] for X in range(self.numDet)))

    @property
    def t_max(self):
        return int(max(self.ts[self.det_off[X + 1] - 1] for X in range(self.numDet)))

    @property
    def num_binned(self):
        """numAtoms after the binned merge: floor((tMax-tMin)/TAtom)+1."""
        return (self.t_max - self.t_min) // self.TAtom + 1


def atoms_to_arrays(multiFstatAtoms):
    """Validate and pack a MultiFstatAtomVector (ours or lalpulsar's).

    Checks what lalpulsar's binned merge checks before summing: at least one
    detector, no empty detector, one common TAtom, ascending timestamps per
    detector.  Raises ValueError otherwise.
    """
    numDet = int(multiFstatAtoms.length)
    if numDet < 1:
        raise ValueError("multiFstatAtoms has no detectors.")
    if numDet > PULSAR_MAX_DETECTORS:
        raise ValueError(
            f"multiFstatAtoms has {numDet} detectors, more than {PULSAR_MAX_DETECTORS}."
        )
    TAtom = int(multiFstatAtoms.data[0].TAtom)
    if TAtom < 2:
        raise ValueError(f"TAtom={TAtom} must be at least 2 seconds.")
    ts_list, col_list, off = [], [], [0]
    for X in range(numDet):
        vec = multiFstatAtoms.data[X]
        if int(vec.TAtom) != TAtom:
            raise ValueError(
                f"Detector {X} has TAtom={vec.TAtom}, expected the common {TAtom}."
            )
        ts, cols = _single_to_arrays(vec)
        if ts.shape[0] == 0:
            raise ValueError(f"Detector {X} has no atoms.")
        if ts.shape[0] > 1 and np.any(ts[1:] < ts[:-1]):
            raise ValueError(f"Detector {X} timestamps are not in ascending order.")
        ts_list.append(ts)
        col_list.append(cols)
        off.append(off[-1] + ts.shape[0])
    ts = np.concatenate(ts_list)
    cols = np.ascontiguousarray(np.concatenate(col_list, axis=1))
    return PackedAtoms(TAtom, ts, cols, np.asarray(off, dtype=np.int64))


def write_atoms_to_file(fname, multiFstatAtoms, header=()):
    """Write atoms as text, one detector after the other.

    Column layout of lalpulsar.write_MultiFstatAtoms_to_fp as read back at
    tests/test_core.py:808-822: a ``%%``-commented header whose last line names
    the columns ``tGPS a2 b2 ab Fa_re Fa_im Fb_re Fb_im``.  Values are written
    with 9 significant digits so float32 atoms round-trip exactly.
    """
    packed = atoms_to_arrays(multiFstatAtoms)
    with open(fname, "w") as fp:
        for h in header:
            fp.write(f"%% {h}\n")
        fp.write("%% tGPS a2 b2 ab Fa_re Fa_im Fb_re Fb_im\n")
        for k in range(packed.ts.shape[0]):
            c = packed.cols[:, k]
            fp.write(
                "%d %.9g %.9g %.9g %.9g %.9g %.9g %.9g\n"
                % (packed.ts[k], c[0], c[1], c[2], c[3], c[4], c[5], c[6])
            )


def read_atoms_from_file(fname, TAtom, comments="%%"):
    """Read a single-detector atoms text file into a MultiFstatAtomVector.

    Same parsing as tests/test_tcw_fstat_map_funcs.py:60-78 (a ``%%`` header
    whose last line carries the column names), vectorised.
    """
    nhead = 0
    with open(fname) as f:
        for line in f:
            if not line.startswith(comments):
                break
            nhead += 1
    raw = np.atleast_2d(np.loadtxt(fname, comments=comments, dtype=np.float64))
    if raw.shape[1] != 8:
        raise ValueError(f"{fname}: expected 8 columns, found {raw.shape[1]}")
    multi = MultiFstatAtomVector(1)
    multi.data[0] = FstatAtomVector.from_arrays(
        TAtom,
        raw[:, 0].astype(np.uint64),
        raw[:, 1],
        raw[:, 2],
        raw[:, 3],
        raw[:, 4] + 1j * raw[:, 5],
        raw[:, 6] + 1j * raw[:, 7],
    )
    return multi


def extract_singleIFOmultiFatoms_from_multiAtoms(multiAtoms, X):
    """Length-1 MultiFstatAtomVector holding a deep copy of detector X.

    Counterpart of pyfstat/utils/atoms.py:8-41, used for per-detector maps
    (pyfstat/core.py:1527-1540).
    """
    if X >= multiAtoms.length:
        raise ValueError(
            f"Detector index {X} is out of range for multiAtoms of length {multiAtoms.length}."
        )
    ts, cols = _single_to_arrays(multiAtoms.data[X])
    single = MultiFstatAtomVector(1)
    single.data[0] = FstatAtomVector.from_arrays(
        multiAtoms.data[X].TAtom, ts, cols[0], cols[1], cols[2],
        cols[3] + 1j * cols[4], cols[5] + 1j * cols[6],
    )
    return single
