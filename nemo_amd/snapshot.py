"""NEMO structured-file (filestruct) snapshot codec, host side.

Reads and writes the binary item stream of NEMO snapshot files:
each item is a 2-byte magic (0x0992 singular, 0x0B92 plural), a NUL-terminated
type string, a NUL-terminated tag (absent for the end-of-set item ``)``),
a 0-terminated int dims vector for plural items, then the raw data
(src/kernel/io/filesecret.c:912-966 put, 1028-1137 get; inc/filestruct.h:15-49).
Byte-swapped files are detected from the magic and accepted
(filesecret.c CHKSWAP).  Output is written in native (little-endian) order.

Snapshot tags follow inc/snapshot/snapshot.h:22-56.  hackforce reads
``Mass`` and ``PhaseSpace`` with rigid double type (hackforce.c:150-151 via
get_data, filesecret.c:589-606): float-typed items are rejected there and here.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

SINGLE = 0x0992
PLURAL = 0x0B92
CS_CARTESIAN_3D_PHASE = 66306   # CSCode(Cartesian, 3, 2), inc/snapshot/snapshot.h:89-109

_SIZES = {"c": 1, "b": 1, "s": 2, "i": 4, "l": 8, "h": 2, "f": 4, "d": 8, "a": 1}
_NP = {"c": "S1", "b": "u1", "s": "i2", "i": "i4", "l": "i8", "h": "i2", "f": "f4", "d": "f8", "a": "u1"}


class SnapshotError(ValueError):
    """Malformed, truncated or unsupported structured file."""


@dataclass
class Item:
    tag: str
    type: str
    dims: list
    data: object = None                     # numpy array, bytes, or list[Item] for sets
    children: list = field(default_factory=list)

    def find(self, tag):
        for c in self.children:
            if c.tag == tag:
                return c
        return None


def parse(buf: bytes) -> list:
    """Parse a whole structured file into a list of top-level Items."""
    if len(buf) < 2:
        raise SnapshotError("file too short")
    head = buf[:2]
    if head in (b"\x09\x92", b"\x0b\x92"):
        end = ">"
    elif head in (b"\x92\x09", b"\x92\x0b"):
        end = "<"
    else:
        raise SnapshotError("bad magic %r" % head)
    pos = 0
    root: list = []
    stack = [root]
    sets: list = []
    n = len(buf)
    while pos < n:
        if pos + 2 > n:
            raise SnapshotError("truncated item header")
        magic = struct.unpack_from(end + "H", buf, pos)[0]
        pos += 2
        if magic not in (SINGLE, PLURAL):
            raise SnapshotError("bad magic 0x%04x at byte %d" % (magic, pos - 2))
        z = buf.find(b"\0", pos)
        if z < 0:
            raise SnapshotError("truncated type")
        typ = buf[pos:z].decode("latin-1")
        pos = z + 1
        if typ == ")":
            if not sets:
                raise SnapshotError("unbalanced end of set")
            sets.pop()
            stack.pop()
            continue
        z = buf.find(b"\0", pos)
        if z < 0:
            raise SnapshotError("truncated tag")
        tag = buf[pos:z].decode("latin-1")
        pos = z + 1
        dims = []
        if magic == PLURAL:
            while True:
                if pos + 4 > n:
                    raise SnapshotError("truncated dims")
                d = struct.unpack_from(end + "i", buf, pos)[0]
                pos += 4
                if d == 0:
                    break
                dims.append(d)
        if typ == "(":
            it = Item(tag, typ, [])
            stack[-1].append(it)
            sets.append(it)
            stack.append(it.children)
            continue
        if typ not in _SIZES:
            raise SnapshotError("unknown type %r for tag %s" % (typ, tag))
        count = int(np.prod(dims)) if dims else 1
        nbytes = count * _SIZES[typ]
        if pos + nbytes > n:
            raise SnapshotError("truncated data for tag %s" % tag)
        raw = buf[pos:pos + nbytes]
        pos += nbytes
        if typ in ("c",):
            data = raw
        else:
            data = np.frombuffer(raw, dtype=np.dtype(_NP[typ]).newbyteorder(end)).astype(_NP[typ])
            if dims:
                data = data.reshape(dims)
        stack[-1].append(Item(tag, typ, dims, data))
    if sets:
        raise SnapshotError("unterminated set %s" % sets[-1].tag)
    return root


@dataclass
class Snapshot:
    mass: np.ndarray          # [n] float64
    pos: np.ndarray           # [n,3] float64
    vel: np.ndarray           # [n,3] float64
    time: float = 0.0
    diagnostics: dict = field(default_factory=dict)
    extra: dict = field(default_factory=dict)   # Potential, Acceleration, ...


def _rigid_double(it: Item, tag: str):
    if it is None:
        raise SnapshotError("missing %s" % tag)
    if it.type != "d":
        raise SnapshotError("%s has type %r, expected double" % (tag, it.type))
    return np.asarray(it.data, dtype=np.float64)


def snapshots(items: list):
    """Yield Snapshot objects from parsed items, skipping SnapShot sets without
    Parameters or Particles (hackforce.c:121-143)."""
    for top in items:
        if top.tag != "SnapShot" or top.type != "(":
            continue
        par = top.find("Parameters")
        prt = top.find("Particles")
        if par is None or prt is None:
            continue
        nobj = par.find("Nobj")
        if nobj is None:
            raise SnapshotError("missing Nobj")
        n = int(np.asarray(nobj.data).reshape(-1)[0])
        if n < 1:
            raise SnapshotError("Nobj = %d is absurd" % n)
        t = par.find("Time")
        time = float(_rigid_double(t, "Time").reshape(-1)[0]) if t is not None else 0.0
        cs = prt.find("CoordSystem")
        if cs is None or int(np.asarray(cs.data).reshape(-1)[0]) != CS_CARTESIAN_3D_PHASE:
            raise SnapshotError("cannot handle CoordSystem")
        mass = _rigid_double(prt.find("Mass"), "Mass").reshape(n)
        ps = _rigid_double(prt.find("PhaseSpace"), "PhaseSpace").reshape(n, 2, 3)
        extra = {}
        for c in prt.children:
            if c.tag in ("Potential", "Acceleration") and c.type == "d":
                extra[c.tag] = np.asarray(c.data, dtype=np.float64)
        diag = {}
        dg = top.find("Diagnostics")
        if dg is not None:
            for c in dg.children:
                if c.type in ("d", "f", "i", "l"):
                    diag[c.tag] = np.asarray(c.data)
        yield Snapshot(mass.copy(), ps[:, 0, :].copy(), ps[:, 1, :].copy(), time, diag, extra)


def read(path: str) -> list:
    with open(path, "rb") as f:
        return list(snapshots(parse(f.read())))


# ---- writer ------------------------------------------------------------------

def _put(out: bytearray, tag: str, typ: str, data=None, dims=None):
    if typ in ("(", ")"):
        out += struct.pack("<H", SINGLE) + typ.encode() + b"\0"
        if typ == "(":
            out += tag.encode() + b"\0"
        return
    arr = np.ascontiguousarray(data, dtype=_NP[typ])
    plural = dims is not None and len(dims) > 0
    out += struct.pack("<H", PLURAL if plural else SINGLE) + typ.encode() + b"\0" + tag.encode() + b"\0"
    if plural:
        for d in dims:
            out += struct.pack("<i", int(d))
        out += struct.pack("<i", 0)
    out += arr.astype(np.dtype(_NP[typ]).newbyteorder("<")).tobytes()


def put_history(out: bytearray, text: str):
    b = text.encode() + b"\0"
    out += struct.pack("<H", PLURAL) + b"c\0History\0" + struct.pack("<ii", len(b), 0) + b


def put_force_snapshot(out: bytearray, mass, pos, vel, phi, acc, time, tol, eps,
                       n2btot, nbctot, cputree, cpufcal, options="mass,phase"):
    """Append one hackforce result snapshot (hackforce.c:231-272 layout)."""
    n = len(phi)
    opts = [o.strip() for o in options.split(",")]
    _put(out, "SnapShot", "(")
    _put(out, "Parameters", "(")
    _put(out, "Nobj", "i", np.int32(n))
    _put(out, "Time", "d", np.float64(time))
    _put(out, "tol", "d", np.float64(tol))
    _put(out, "eps", "d", np.float64(eps))
    _put(out, "", ")")
    _put(out, "Particles", "(")
    _put(out, "CoordSystem", "i", np.int32(CS_CARTESIAN_3D_PHASE))
    if "mass" in opts:
        _put(out, "Mass", "d", mass, [n])
    if "phase" in opts:
        ps = np.stack([np.asarray(pos), np.asarray(vel)], axis=1)
        _put(out, "PhaseSpace", "d", ps, [n, 2, 3])
    _put(out, "Potential", "d", phi, [n])
    _put(out, "Acceleration", "d", acc, [n, 3])
    _put(out, "", ")")
    _put(out, "Diagnostics", "(")
    _put(out, "n2btot", "i", np.int32(min(n2btot, 2**31 - 1)))
    _put(out, "nbctot", "i", np.int32(min(nbctot, 2**31 - 1)))
    _put(out, "cputree", "d", np.float64(cputree))
    _put(out, "cpufcal", "d", np.float64(cpufcal))
    _put(out, "", ")")
    _put(out, "", ")")


def put_snapshot(out: bytearray, mass, pos, vel, time, write_mass=True, write_phase=True, diag=None):
    """Append a hackcode1 put_snap frame (inc/snapshot/put_snap-old.c:437-451):
    Parameters{Nobj,Time}, Particles{CoordSystem,[Mass],[PhaseSpace]} if any
    particle bit, Diagnostics{Energy,KETensor,PETensor,AMTensor,CMPhaseSpace,cputime}."""
    n = len(mass)
    _put(out, "SnapShot", "(")
    _put(out, "Parameters", "(")
    _put(out, "Nobj", "i", np.int32(n))
    _put(out, "Time", "d", np.float64(time))
    _put(out, "", ")")
    if write_mass or write_phase:
        _put(out, "Particles", "(")
        _put(out, "CoordSystem", "i", np.int32(CS_CARTESIAN_3D_PHASE))
        if write_mass:
            _put(out, "Mass", "d", mass, [n])
        if write_phase:
            _put(out, "PhaseSpace", "d", np.stack([pos, vel], axis=1), [n, 2, 3])
        _put(out, "", ")")
    if diag is not None:
        d = np.asarray(diag, dtype=np.float64)
        _put(out, "Diagnostics", "(")
        _put(out, "Energy", "d", d[0:3], [3])
        _put(out, "KETensor", "d", d[3:12], [3, 3])
        _put(out, "PETensor", "d", d[12:21], [3, 3])
        _put(out, "AMTensor", "d", d[21:30], [3, 3])
        _put(out, "CMPhaseSpace", "d", d[30:36], [2, 3])
        _put(out, "cputime", "d", np.float64(0.0))
        _put(out, "", ")")
    _put(out, "", ")")


def bsf(buf: bytes, ignore: str = "cputime") -> str:
    """Five-number summary of every float/double value in a structured file
    (src/kernel/io/bsf.c:45-160): mean, sigma (population), min, max, count,
    each float printed with %g; items tagged ``ignore`` are skipped."""
    vals = []

    def walk(items):
        for it in items:
            if it.type == "(":
                walk(it.children)
            elif it.type in ("f", "d") and it.tag != ignore:
                vals.append(np.asarray(it.data, dtype=np.float64).reshape(-1))

    walk(parse(buf))
    v = np.concatenate(vals) if vals else np.zeros(0)
    s0 = float(len(v))
    s1 = float(np.sum(v))
    s2 = float(np.sum(v * v))
    mean = s1 / s0
    if v.min() == v.max():
        sig = 0.0
    else:
        t2 = s2 / s0 - mean * mean
        sig = float(np.sqrt(t2)) if t2 > 0 else 0.0
    return "%g %g %g %g %d" % (mean, sig, v.min(), v.max(), len(v))
