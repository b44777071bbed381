"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's cpu_baseline leg.  The product path
(nemo_amd/) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = C.POINTER(C.c_double)
_i64 = C.POINTER(C.c_int64)
_u64 = C.POINTER(C.c_uint64)
_i32 = C.POINTER(C.c_int)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        # NEMO_ORACLE_LIB: another build of the same sources (the sanitizer
        # build of oracle/san.mk)
        path = os.environ.get("NEMO_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_force.argtypes = [C.c_int, _d, _d, C.c_int, _d, C.c_double, C.c_double, C.c_int,
                                _d, _d, C.c_double, _d, _d, _i64, _i64]
        L.orc_force.restype = C.c_int
        L.orc_force_capped.argtypes = [C.c_int, _d, _d, C.c_int, _d, C.c_double, C.c_double, C.c_int,
                                       _d, _d, C.c_double, C.c_int, _d, _d, _i64, _i64, _i64]
        L.orc_force_capped.restype = C.c_int
        L.orc_direct.argtypes = [C.c_int, _d, _d, C.c_int, _i64, C.c_double, _d, _d]
        L.orc_direct.restype = None
        L.orc_step_force.argtypes = [C.c_int, _d, _d, _d, _d, _d, C.c_int, C.c_double, C.c_double,
                                     C.c_double, C.c_int, _d, _d, C.c_double, _i64, _i64]
        L.orc_step_force.restype = C.c_int
        L.orc_step_drift.argtypes = [C.c_int, _d, _d, _d, C.c_double]
        L.orc_step_drift.restype = None
        L.orc_force_sample.argtypes = [C.c_int, _d, _d, C.c_int, _i64, C.c_double, C.c_double, C.c_int,
                                       _d, _d, C.c_double, _d, _d, _i64, _i64, _d, _d]
        L.orc_force_sample.restype = C.c_int
        L.orc_force_sample_mt.argtypes = [C.c_int, _d, _d, C.c_int, _i64, C.c_double, C.c_double, C.c_int,
                                          _d, _d, C.c_double, C.c_int, _d, _d, _i64, _i64, _d, _d]
        L.orc_force_sample_mt.restype = C.c_int
        L.orc_diagnostics.argtypes = [C.c_int, _d, _d, _d, _d, _d, _d]
        L.orc_morton.argtypes = [C.c_int, _d, _d, C.c_double, _u64]
        L.orc_tree_build.argtypes = [C.POINTER(C.c_void_p), C.c_int, _d, _d, _d, _d, C.c_int, C.c_int]
        L.orc_tree_build.restype = C.c_int
        L.orc_tree_build_capped.argtypes = [C.POINTER(C.c_void_p), C.c_int, _d, _d, _d, _d, C.c_int, C.c_int,
                                            C.c_int]
        L.orc_tree_build_capped.restype = C.c_int
        L.orc_tree_nbucket.argtypes = [C.c_void_p]
        L.orc_tree_walk_idx.argtypes = [C.c_void_p, _d, C.c_int, _i64, C.c_double, C.c_double, C.c_int, _d, _d,
                                        _i64, _i64, _d]
        L.orc_tree_walk_idx.restype = C.c_int
        L.orc_tree_free.argtypes = [C.c_void_p]
        L.orc_tree_ncell.argtypes = [C.c_void_p]
        L.orc_tree_dump.argtypes = [C.c_void_p, _d, _d, _d, _i32, _i32]
        L.orc_tree_dump.restype = C.c_int
        L.tc_force.argtypes = [C.c_int, _d, _d, C.c_int, _d, _d, C.c_double, C.c_double, C.c_double, C.c_int,
                               C.c_int, _d, _d, _i64, _i64, _i64, C.c_int, _i32, _d, _d, _d]
        L.tc_force.restype = C.c_int
        L.lc_residual_field.argtypes = [C.c_int, _d, _d, C.c_double, _d, _d, C.c_double, _d, _d, _d]
        L.lc_residual_field.restype = None
        L.lc_features.argtypes = [_d, C.c_double, _d, _d, C.c_double, _d, _d, _d]
        L.lc_features.restype = C.c_double
        L.lc_apply.argtypes = [C.c_int, _d, _d, _d, _d, _d]
        L.lc_apply.restype = None
        L.le_coeffs.argtypes = [C.c_int, _d, _d, _d, C.c_double, C.c_double, _d]
        L.le_coeffs.restype = None
        L.le_eval.argtypes = [_d, C.c_double, _d, C.c_int, _d, _d, _d]
        L.le_eval.restype = None
        _LIB = L
    return _LIB


def _p(a, t=_d):
    return a.ctypes.data_as(t)


def force(mass, pos, tol=1.0, eps=0.05, usequad=False, rsize=4.0, rmin=None, fcells=2.0, tpos=None, depth_cap=0):
    """hackforce force_calc: returns (phi, acc, n2btot, nbctot, rmin, rsize).

    depth_cap=0 is the reference's tree (load.c:101-128 splits down to 2^30
    and fails on coincident bodies); depth_cap=21 stops cells at level 21, a
    colliding slot there becoming a bucket whose bodies are its children in
    input order: the device's 63-bit-key tree.  See force_capped."""
    out = force_capped(mass, pos, tol, eps, usequad, rsize, rmin, fcells, tpos, depth_cap)
    return out[:6]


def force_capped(mass, pos, tol=1.0, eps=0.05, usequad=False, rsize=4.0, rmin=None, fcells=2.0, tpos=None,
                 depth_cap=21):
    """As force(), with the tree's depth capped (default: the device's 21
    levels).  Returns (phi, acc, n2btot, nbctot, rmin, rsize, nbucket), where
    nbucket counts the bucket cells (cells at level depth_cap)."""
    mass = np.ascontiguousarray(mass, np.float64)
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    n = len(mass)
    if rmin is None:
        rmin = np.full(3, -rsize / 2.0)
    rmin = np.array(rmin, np.float64)
    rs = np.array([rsize], np.float64)
    nt = n if tpos is None else len(tpos)
    phi = np.zeros(nt)
    acc = np.zeros((nt, 3))
    a = np.zeros(1, np.int64)
    b = np.zeros(1, np.int64)
    nbk = np.zeros(1, np.int64)
    tp = None
    if tpos is not None:
        tpos = np.ascontiguousarray(tpos, np.float64).reshape(-1, 3)
        tp = _p(tpos)
    err = lib().orc_force_capped(n, _p(mass), _p(pos), nt, tp, tol, eps, int(usequad), _p(rmin), _p(rs),
                                 fcells, int(depth_cap), _p(phi), _p(acc), _p(a, _i64), _p(b, _i64), _p(nbk, _i64))
    if err:
        raise RuntimeError("oracle tree error %d" % err)
    return phi, acc, int(a[0]), int(b[0]), rmin, float(rs[0]), int(nbk[0])


def force_sample(mass, pos, idx, tol, eps, usequad, rsize=4.0, fcells=2.0, nthreads=1):
    """Full tree build plus the walk of the bodies idx (self skipped), timed
    in C; the walk runs on nthreads host threads over one shared tree.
    Returns (phi, acc, n2b, nbc, t_tree_s, t_walk_s)."""
    mass = np.ascontiguousarray(mass, np.float64)
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    idx = np.ascontiguousarray(idx, np.int64)
    rmin = np.full(3, -rsize / 2.0)
    rs = np.array([rsize])
    phi = np.zeros(len(idx))
    acc = np.zeros((len(idx), 3))
    a = np.zeros(1, np.int64)
    b = np.zeros(1, np.int64)
    tt = np.zeros(1)
    tw = np.zeros(1)
    err = lib().orc_force_sample_mt(len(mass), _p(mass), _p(pos), len(idx), _p(idx, _i64), tol, eps, int(usequad),
                                    _p(rmin), _p(rs), fcells, int(nthreads), _p(phi), _p(acc), _p(a, _i64),
                                    _p(b, _i64), _p(tt), _p(tw))
    if err:
        raise RuntimeError("oracle tree error %d" % err)
    return phi, acc, int(a[0]), int(b[0]), float(tt[0]), float(tw[0])


class Tree:
    """One reference tree (hackforce's box: rmin = -rsize/2) kept for several
    walks of a subset of its bodies, e.g. at different tol."""

    def __init__(self, mass, pos, usequad, rsize=4.0, fcells=2.0):
        self.mass = np.ascontiguousarray(mass, np.float64)
        self.pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        n = len(self.mass)
        self.rmin = np.full(3, -rsize / 2.0)
        rs = np.array([rsize])
        self._h = C.c_void_p()
        import time
        t0 = time.perf_counter()
        err = lib().orc_tree_build(C.byref(self._h), n, _p(self.mass), _p(self.pos), _p(self.rmin), _p(rs),
                                   int(fcells * n), int(bool(usequad)))
        self.t_build = time.perf_counter() - t0
        self.rsize = float(rs[0])
        if err:
            self.close()
            raise RuntimeError("oracle tree error %d" % err)

    def walk(self, idx, tol, eps, nthreads=1):
        """Field at the bodies idx (self skipped): (phi, acc, n2b, nbc, t_walk_s)."""
        idx = np.ascontiguousarray(idx, np.int64)
        phi = np.zeros(len(idx))
        acc = np.zeros((len(idx), 3))
        a = np.zeros(1, np.int64)
        b = np.zeros(1, np.int64)
        tw = np.zeros(1)
        lib().orc_tree_walk_idx(self._h, _p(self.pos), len(idx), _p(idx, _i64), float(tol), float(eps),
                                int(nthreads), _p(phi), _p(acc), _p(a, _i64), _p(b, _i64), _p(tw))
        return phi, acc, int(a[0]), int(b[0]), float(tw[0])

    def close(self):
        if self._h:
            lib().orc_tree_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()


def direct(mass, pos, eps, idx=None):
    mass = np.ascontiguousarray(mass, np.float64)
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    n = len(mass)
    if idx is None:
        nt, ip = n, None
    else:
        idx = np.ascontiguousarray(idx, np.int64)
        nt, ip = len(idx), _p(idx, _i64)
    phi = np.zeros(nt)
    acc = np.zeros((nt, 3))
    lib().orc_direct(n, _p(mass), _p(pos), nt, ip, eps, _p(phi), _p(acc))
    return phi, acc


def morton(pos, rmin, rsize):
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    keys = np.zeros(len(pos), np.uint64)
    rmin = np.ascontiguousarray(rmin, np.float64)
    lib().orc_morton(len(pos), _p(pos), _p(rmin), rsize, _p(keys, _u64))
    return keys


def tree_cells(mass, pos, rmin, rsize, usequad=True, maxcell=None, depth_cap=0):
    """Cells of the reference tree in DFS pre-order: dict of arrays
    (depth_cap as in force(); nbucket counts the bucket cells)."""
    mass = np.ascontiguousarray(mass, np.float64)
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    n = len(mass)
    rmin = np.array(rmin, np.float64)
    rs = np.array([rsize], np.float64)
    maxcell = maxcell or 2 * n + 64
    h = C.c_void_p()
    err = lib().orc_tree_build_capped(C.byref(h), n, _p(mass), _p(pos), _p(rmin), _p(rs), maxcell, int(usequad),
                                      int(depth_cap))
    try:
        if err:
            raise RuntimeError("oracle tree error %d" % err)
        nc = lib().orc_tree_ncell(h)
        m = np.zeros(nc)
        p = np.zeros((nc, 3))
        q = np.zeros((nc, 9))
        lv = np.zeros(nc, np.int32)
        nb = np.zeros(nc, np.int32)
        k = lib().orc_tree_dump(h, _p(m), _p(p), _p(q), _p(lv, _i32), _p(nb, _i32))
        return dict(mass=m[:k], pos=p[:k], quad=q[:k], level=lv[:k], nbody=nb[:k], rmin=rmin, rsize=float(rs[0]),
                    nbucket=int(lib().orc_tree_nbucket(h)))
    finally:
        lib().orc_tree_free(h)


class Leapfrog:
    """hackcode1 stepsystem() state machine (code.c:57-64,159-191)."""

    def __init__(self, mass, pos, vel, tol=1.0, eps=0.05, usequad=False, freq=32.0, fcells=2.0):
        self.mass = np.ascontiguousarray(mass, np.float64)
        self.pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3).copy()
        self.vel = np.ascontiguousarray(vel, np.float64).reshape(-1, 3).copy()
        n = len(self.mass)
        self.acc = np.zeros((n, 3))
        self.phi = np.zeros(n)
        self.nstep = np.zeros(1, np.int32)
        self.tol, self.eps, self.usequad, self.freq, self.fcells = tol, eps, usequad, freq, fcells
        self.rmin = np.full(3, -2.0)
        self.rsize = np.array([4.0])
        self.tnow = 0.0

    def force(self):
        """Force pass of stepsystem (with the second-order velocity correction
        when nstep > 0).  Returns (n2b, nbc); the state is now at output time."""
        n = len(self.mass)
        a = np.zeros(1, np.int64)
        b = np.zeros(1, np.int64)
        err = lib().orc_step_force(n, _p(self.mass), _p(self.pos), _p(self.vel), _p(self.acc),
                                   _p(self.phi), int(self.nstep[0]), 1.0 / self.freq, self.tol,
                                   self.eps, int(self.usequad), _p(self.rmin), _p(self.rsize),
                                   self.fcells, _p(a, _i64), _p(b, _i64))
        if err:
            raise RuntimeError("oracle tree error %d" % err)
        return int(a[0]), int(b[0])

    def diagnostics(self):
        d = np.zeros(36)
        lib().orc_diagnostics(len(self.mass), _p(self.mass), _p(self.pos), _p(self.vel), _p(self.acc),
                              _p(self.phi), _p(d))
        return d

    def drift(self):
        dt = 1.0 / self.freq
        lib().orc_step_drift(len(self.mass), _p(self.pos), _p(self.vel), _p(self.acc), dt)
        self.nstep[0] += 1
        self.tnow = self.tnow + dt

    def step(self):
        """One full step; returns (diag[36], n2b, nbc) with diag at output time."""
        n2b, nbc = self.force()
        d = self.diagnostics()
        self.drift()
        return d, n2b, nbc


TC_CRITERIA = {"rcrit": 0, "bh86": 1, "sw94": 2}


def tc_force(mass, pos, rmin, rsize, theta=1.0, eps=0.025, criterion="rcrit", usequad=False, tpos=None,
             cells=False):
    """treecode1 gravcalc over the root box (rmin, rsize): returns a dict with
    phi, acc, n2b, nbc per target (bodies, then test points), ncell and, with
    cells=True, the cells' level, mass, com and rcrit2 in DFS order."""
    mass = np.ascontiguousarray(mass, np.float64)
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    n = len(mass)
    tp = None if tpos is None else np.ascontiguousarray(tpos, np.float64).reshape(-1, 3)
    nt = 0 if tp is None else len(tp)
    phi = np.zeros(n + nt)
    acc = np.zeros((n + nt, 3))
    n2b = np.zeros(n + nt, np.int64)
    nbc = np.zeros(n + nt, np.int64)
    nc = np.zeros(1, np.int64)
    cap = 2 * n + 8 if cells else 0
    lev = np.zeros(max(cap, 1), np.int32)
    cm_ = np.zeros(max(cap, 1))
    ccm = np.zeros((max(cap, 1), 3))
    rc = np.zeros(max(cap, 1))
    rmin = np.ascontiguousarray(rmin, np.float64)
    err = lib().tc_force(n, _p(mass), _p(pos), nt, None if tp is None else _p(tp), _p(rmin), float(rsize),
                         float(theta), float(eps), TC_CRITERIA[criterion], int(bool(usequad)), _p(phi), _p(acc),
                         _p(n2b, _i64), _p(nbc, _i64), _p(nc, _i64), cap,
                         _p(lev, _i32) if cells else None, _p(cm_) if cells else None,
                         _p(ccm) if cells else None, _p(rc) if cells else None)
    if err:
        raise RuntimeError(f"tc_force failed with code {err}")
    out = dict(phi=phi, acc=acc, n2b=n2b, nbc=nbc, ncell=int(nc[0]))
    if cells:
        k = int(nc[0])
        out.update(level=lev[:k], mass=cm_[:k], com=ccm[:k], rcrit2=rc[:k])
    return out


# ---- learned correction data path (lc_oracle.c) -------------------------------

def residual_field(m, x, M, com, Q, eps, t):
    """Exact residual (softened pair sum over the cell's bodies m, x minus
    gravsub's multipole force of mass M at com, quadrupole Q[3][3] or None)
    at point t, and its Jacobian J[i][k] = d res_i / d t_k."""
    m = np.ascontiguousarray(m, np.float64)
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    com = np.ascontiguousarray(com, np.float64)
    t = np.ascontiguousarray(t, np.float64)
    q = None if Q is None else np.ascontiguousarray(Q, np.float64).reshape(9)
    res = np.zeros(3)
    J = np.zeros(9)
    lib().lc_residual_field(len(m), _p(m), _p(x), float(M), _p(com), None if q is None else _p(q), float(eps), _p(t),
                            _p(res), _p(J))
    return res, J.reshape(3, 3)


def features(com, M, Q, centre, size, t):
    """MLP input features of a cell seen from t: (f[12], M/r^2, 1/r)."""
    com = np.ascontiguousarray(com, np.float64)
    q = np.ascontiguousarray(Q, np.float64).reshape(9)
    cc = np.ascontiguousarray(centre, np.float64)
    t = np.ascontiguousarray(t, np.float64)
    f = np.zeros(12)
    ri = np.zeros(1)
    scale = lib().lc_features(_p(com), float(M), _p(q), _p(cc), float(size), _p(t), _p(f), _p(ri))
    return f, float(scale), float(ri[0])


def apply_first_order(x, centre, R, j6):
    """R + J (x - centre) for each point (J symmetric from xx xy xz yy yz zz)."""
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    c = np.ascontiguousarray(centre, np.float64)
    R = np.ascontiguousarray(R, np.float64)
    j6 = np.ascontiguousarray(j6, np.float64)
    out = np.zeros_like(x)
    lib().lc_apply(len(x), _p(x), _p(c), _p(R), _p(j6), _p(out))
    return out


# ---- group local expansion of far cells (le_oracle.c) -------------------------

def local_coeffs(m, s, centre, eps, ell):
    """Summed coefficients G[20] of the second-order local expansion about
    centre of the softened monopoles (masses m, c.o.m. s[k][3]), tensors
    scaled by powers of ell (le_oracle.c's order)."""
    m = np.ascontiguousarray(m, np.float64)
    s = np.ascontiguousarray(s, np.float64).reshape(-1, 3)
    c = np.ascontiguousarray(centre, np.float64)
    G = np.zeros(20)
    lib().le_coeffs(len(m), _p(m), _p(s), _p(c), float(eps), float(ell), _p(G))
    return G


def local_eval(G, ell, centre, x):
    """Potential (NEMO's sign) and field of the expansion G at points x."""
    G = np.ascontiguousarray(G, np.float64)
    c = np.ascontiguousarray(centre, np.float64)
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    phi = np.zeros(len(x))
    acc = np.zeros_like(x)
    lib().le_eval(_p(G), float(ell), _p(c), len(x), _p(x), _p(phi), _p(acc))
    return phi, acc
