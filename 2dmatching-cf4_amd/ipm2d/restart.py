"""Particle-bundle restart files in the reference's formats.

The reference dumps one file per species and rank (ParticleBundle%Dump /
%Load, R/src/pic/particle/ParticleBundle.f90:86-300) either as HDF5
(DumpParticleBundleHDF5, lines 117-148) or as Fortran list-directed text
(DumpParticleBundleDat / LoadParticleBundleDat, lines 151-174 and 230-262).
This module reads and writes the text form; ``ipm2d.h5lite`` the HDF5 form.

A text dump is, record by record (each line below is one list-directed
WRITE, so one or more text lines):

    NPar size
    Charge Mass Weight
    XFactor VFactor
    LXScaled LVScaled
    PO%Z  (size values)   ... then R, Vz, Vr, Vt, Az, Ar, WQ

``size`` is the bundle's array size; entries past NPar are whatever the
array held.  LXScaled / LVScaled say whether positions / velocities are in
grid units (ParticleBundle.f90:606-645 switches them by multiplying with
XFactor or 1/XFactor, VFactor or 1/VFactor).  The device keeps grid units,
so ``to_grid_units`` applies those rescalings to a dump that is not scaled.
"""
import dataclasses
import re

import numpy as np

FIELDS = ("Z", "R", "Vz", "Vr", "Vt", "Az", "Ar", "WQ")


@dataclasses.dataclass
class BundleDump:
    npar: int
    size: int
    charge: float
    mass: float
    weight: float
    xfactor: float
    vfactor: float
    lxscaled: bool
    lvscaled: bool
    po: np.ndarray          # (8, size): Z R Vz Vr Vt Az Ar WQ

    def particles(self):
        """The NPar live particles, (8, NPar)."""
        return self.po[:, :self.npar]

    def to_grid_units(self):
        """A copy whose positions and velocities are in grid units, rescaled
        as PositionRescaleParticleBundle / VelocityRescaleParticleBundle do
        (ParticleBundle.f90:610-645; ParticleOne.f90:191-210)."""
        po = self.po.copy()
        n = self.npar
        if not self.lxscaled:
            f = 1.0 / self.xfactor
            po[0, :n] = po[0, :n] * f
            po[1, :n] = po[1, :n] * f
        if not self.lvscaled:
            f = 1.0 / self.vfactor
            po[2:5, :n] = po[2:5, :n] * f
        return dataclasses.replace(self, po=po, lxscaled=True, lvscaled=True)


# ------------------------------------------------------------ list-directed
_NUM = re.compile(r"^([+-]?(?:\d+\.?\d*|\.\d+))(?:[eEdDqQ]?([+-]?\d+))?$")


def _value(tok):
    """One list-directed numeric constant: Fortran's D/Q exponents and the
    exponent-letter-free form of three-digit exponents (1.0-300)."""
    t = tok.strip()
    m = _NUM.match(t)
    if not m:
        raise ValueError(f"not a Fortran number: {tok!r}")
    mant, exp = m.group(1), m.group(2)
    return float(mant + ("e" + exp if exp is not None else ""))


def _logical(tok):
    t = tok.strip().lstrip(".").upper()
    if not t or t[0] not in "TF":
        raise ValueError(f"not a Fortran logical: {tok!r}")
    return t[0] == "T"


def _items(line):
    """Values of one record: blanks and commas separate, r*c repeats."""
    out = []
    for tok in line.replace(",", " ").split():
        if tok.startswith("/"):
            raise ValueError("a '/' ends a list-directed read early; not expected in a bundle dump")
        if "*" in tok and not tok.startswith("'"):
            r, c = tok.split("*", 1)
            if not c:
                raise ValueError("null values (r*) are not expected in a bundle dump")
            out.extend([c] * int(r))
        else:
            out.append(tok)
    return out


class _Reader:
    """List-directed READ semantics: every read starts at a new record and
    takes values from as many records as it needs; what is left of its last
    record is skipped."""

    def __init__(self, lines):
        self.lines = lines
        self.k = 0

    def read(self, n):
        vals = []
        while len(vals) < n:
            if self.k >= len(self.lines):
                raise ValueError(f"end of file: wanted {n} values, got {len(vals)}")
            vals.extend(_items(self.lines[self.k]))
            self.k += 1
        return vals[:n]

    def read_array(self, n):
        """n reals, converted in bulk where the text allows it."""
        start = self.k
        count = 0
        fast = True
        while count < n:
            if self.k >= len(self.lines):
                raise ValueError(f"end of file: wanted {n} values, got {count}")
            line = self.lines[self.k]
            if "*" in line or "/" in line or "d" in line or "D" in line or "q" in line or "Q" in line:
                fast = False
            toks = line.replace(",", " ").split()
            if "*" in line:
                count += sum(int(t.split("*", 1)[0]) if "*" in t else 1 for t in toks)
            else:
                count += len(toks)
            self.k += 1
        chunk = self.lines[start:self.k]
        if fast and count == n:
            text = " ".join(chunk).replace(",", " ")
            try:
                return np.array(text.split(), dtype=np.float64)
            except ValueError:
                pass    # a three-digit exponent without its letter, e.g. 1.0-300
        vals = []
        for line in chunk:
            vals.extend(_items(line))
        return np.array([_value(v) for v in vals[:n]], dtype=np.float64)


def read_dat(path):
    """LoadParticleBundleDat, ParticleBundle.f90:230-262."""
    with open(path) as f:
        lines = f.read().splitlines()
    rd = _Reader(lines)
    npar, size = (int(_value(v)) for v in rd.read(2))
    charge, mass, weight = (_value(v) for v in rd.read(3))
    xfactor, vfactor = (_value(v) for v in rd.read(2))
    lx, lv = (_logical(v) for v in rd.read(2))
    if not 0 <= npar <= size:
        raise ValueError(f"NPar {npar} outside [0, size {size}]")
    po = np.empty((8, size), dtype=np.float64)
    for f in range(8):
        po[f] = rd.read_array(size)
    return BundleDump(npar, size, charge, mass, weight, xfactor, vfactor, lx, lv, po)


def _fmt(x):
    r = repr(float(x))
    if r in ("inf", "-inf", "nan"):
        raise ValueError("non-finite value in a bundle dump")
    return r


def write_dat(path, dump, per_line=4):
    """DumpParticleBundleDat, ParticleBundle.f90:151-174: the same records,
    each real written with the shortest text that reads back to the same
    double."""
    po = np.asarray(dump.po, dtype=np.float64)
    if po.shape != (8, dump.size):
        raise ValueError(f"po must be (8, size={dump.size})")
    with open(path, "w") as f:
        f.write(f" {int(dump.npar)} {int(dump.size)}\n")
        f.write(f" {_fmt(dump.charge)} {_fmt(dump.mass)} {_fmt(dump.weight)}\n")
        f.write(f" {_fmt(dump.xfactor)} {_fmt(dump.vfactor)}\n")
        f.write(f" {'T' if dump.lxscaled else 'F'} {'T' if dump.lvscaled else 'F'}\n")
        for k in range(8):
            row = po[k]
            for a in range(0, dump.size, per_line):
                f.write(" " + " ".join(_fmt(x) for x in row[a:a + per_line]) + "\n")
            if dump.size == 0:
                f.write("\n")


def dump_from_particles(P, charge, mass, weight, dx, dt):
    """A dump of live particles in grid units, as the reference's bundle
    holds them (XFactor = dx, VFactor = dx/dt: ParticleBundle.f90:343-346)."""
    P = np.ascontiguousarray(P, dtype=np.float64)
    n = P.shape[1]
    return BundleDump(n, n, float(charge), float(mass), float(weight), float(dx), float(dx) / float(dt),
                      True, True, P.copy())
