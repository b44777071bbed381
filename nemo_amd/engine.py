"""Python host interface of the engine: ctypes over libnemo_bh.so (include/bh.h).

``hackforce(...)`` mirrors NEMO hackforce's keywords and semantics
(src/nbody/evolve/hackcode/hackcode1/hackforce.c:38-50,177-215): ``tol``,
``eps``, ``rsize``, ``rmin`` (default -rsize/2), quadrupole through
``usequad`` (hackforce_qp), test points through ``tpos`` (test=).  There is no
CPU fallback: if the shared library is missing or no device is present the
calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libnemo_bh.so")

MODE_FAST, MODE_STRICT, MODE_DIRECT = 0, 1, 2
DIRECT_NOALLOC_MAX = 32768     # BH_DIRECT_NOALLOC_MAX of include/bh.h

STATUS = {0: "OK", 1: "ERR_ARG", 2: "ERR_DEVICE", 3: "ERR_NOMEM", 4: "ERR_NONFINITE", 5: "ERR_CAPACITY",
          6: "ERR_STATE", 7: "ERR_INJECTED"}


class BHError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (STATUS.get(code, code), msg))
        self.code = code


class Params(C.Structure):
    _fields_ = [("tol", C.c_double), ("eps", C.c_double), ("usequad", C.c_int), ("mode", C.c_int),
                ("rsize", C.c_double), ("rmin", C.c_double * 3), ("rmin_given", C.c_int),
                ("group_size", C.c_int), ("learned", C.c_int), ("keep_box", C.c_int), ("nleaf", C.c_int),
                ("criterion", C.c_int), ("supergroup", C.c_int),
                ("local_alpha", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("n2btot", C.c_int64), ("nbctot", C.c_int64), ("ncell", C.c_int64), ("deep_buckets", C.c_int64),
                ("rmin", C.c_double * 3), ("rsize", C.c_double), ("t_tree_ms", C.c_double),
                ("t_walk_ms", C.c_double), ("n_mlp", C.c_int64), ("t_walk_kernel_ms", C.c_double),
                ("t_sort_ms", C.c_double), ("ngroups", C.c_int64), ("t_mlp_ms", C.c_double),
                ("n_local", C.c_int64)]

    def as_dict(self):
        return {k: (list(getattr(self, k)) if k == "rmin" else getattr(self, k)) for k, _ in self._fields_}


_LIB = None
_d = C.POINTER(C.c_double)
_i64 = C.POINTER(C.c_int64)
_u64 = C.POINTER(C.c_uint64)
_i32 = C.POINTER(C.c_int)
_vp = C.c_void_p
_f = C.POINTER(C.c_float)
MLP_PARAMS = 9737
MLP_IN = 12
MLP_OUT = 9      # residual (3) and its symmetric Jacobian (xx xy xz yy yz zz)

# every symbol include/bh.h declares, with its ctypes signature
SIGNATURES = {
    "bh_default_params": (None, [C.POINTER(Params)]),
    "bh_create": (C.c_int, [C.POINTER(_vp), C.c_int]),
    "bh_destroy": (None, [_vp]),
    "bh_last_error": (C.c_char_p, [_vp]),
    "bh_version": (C.c_char_p, []),
    "bh_force": (C.c_int, [_vp, C.POINTER(Params), C.c_int64, _d, _d, C.c_int64, _d, _d, _d, C.POINTER(Stats)]),
    "bh_upload": (C.c_int, [_vp, C.c_int64, _d, _d, C.c_int64, _d]),
    "bh_eval": (C.c_int, [_vp, C.POINTER(Params), C.POINTER(Stats)]),
    "bh_download": (C.c_int, [_vp, _d, _d]),
    "bh_direct": (C.c_int, [_vp, C.c_double, C.c_int64, _i64, _d, _d]),
    "bh_direct_noalloc": (C.c_int, [_vp, C.c_double, C.c_int64, _i64, _d, _d]),
    "bh_debug_samples": (C.c_int, [_vp, C.c_int64, _i32, _f, _i64]),
    "bh_debug_learned": (C.c_int, [_vp, C.c_int64, _i32, _i32, _f, _i64, C.c_int64, _f, _i64]),
    "bh_debug_cell_ranges": (C.c_int, [_vp, C.c_int64, _i32, _i32, _i64]),
    "bh_sync": (C.c_int, [_vp]),
    "bh_debug_keys": (C.c_int, [_vp, _u64]),
    "bh_debug_cells": (C.c_int, [_vp, C.c_int64, _i32, _i32, _d, _d, _d, _i64]),
    "bh_debug_inject": (None, [_vp, C.c_int]),
    "bh_debug_keep_learned": (None, [_vp, C.c_int]),
    "bh_debug_perm": (C.c_int, [_vp, C.POINTER(C.c_uint32)]),
    "bh_debug_groups": (C.c_int, [_vp, C.c_int64, _i32, _i32, _i64]),
    "bh_mkplummer": (None, [C.c_int64, C.c_uint64, _d, _d, _d]),
    "bh_mkdiskhalo": (None, [C.c_int64, C.c_int64, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_double,
                             C.c_double, C.c_double, _d, _d, _d]),
    "bh_lf_init": (C.c_int, [_vp, C.POINTER(Params), C.c_int64, _d, _d, _d, C.c_double]),
    "bh_lf_force": (C.c_int, [_vp, _d, _i64, _i64]),
    "bh_lf_drift": (C.c_int, [_vp]),
    "bh_lf_run": (C.c_int, [_vp, C.c_int64, C.c_int, _d, _i64, _i64]),
    "bh_lf_get": (C.c_int, [_vp, _d, _d, _d, _d, _i64]),
    "bh_lf_box": (C.c_int, [_vp, _d, _d]),
    "bh_lf_restore": (C.c_int, [_vp, _d, C.c_int64, _d, C.c_double]),
    "bh_mlp_init": (C.c_int, [_vp, C.c_uint64]),
    "bh_mlp_set": (C.c_int, [_vp, _f, _f]),
    "bh_mlp_get": (C.c_int, [_vp, _f, _f]),
    "bh_mlp_forward": (C.c_int, [_vp, C.c_int64, _f, _f]),
    "bh_mlp_grad": (C.c_int, [_vp, C.c_int64, _f, _f, _f, _f]),
    "bh_mlp_collect": (C.c_int, [_vp, C.POINTER(Params), C.c_int, C.c_double, C.c_double, _i64]),
    "bh_mlp_samples": (C.c_int, [_vp, C.c_int64, _f, _f, _i64]),
    "bh_mlp_clear_samples": (C.c_int, [_vp]),
    "bh_mlp_train": (C.c_int, [_vp, C.c_int, C.c_int, C.c_float, C.c_uint64, _f]),
    "bh_batch_upload": (C.c_int, [_vp, C.c_int, _i64, _d, _d, _d, _d]),
    "bh_batch_force": (C.c_int, [_vp, C.POINTER(Params), C.c_int, _i64, _d, _d, _d, _d, _d, _d, C.POINTER(Stats)]),
    "bh_batch_lf_init": (C.c_int, [_vp, C.POINTER(Params), C.c_int, _i64, _d, _d, _d, _d, _d, C.c_double]),
    "bh_autotune": (C.c_int, [_vp, C.POINTER(Params), C.c_int, C.POINTER(Params), _d]),
}

# bh_criterion: hackcode1's opening test, treecode1's default / bh86 / sw94
CRITERIA = {"hack": 0, "rcrit": 1, "bh86": 2, "sw94": 3}

AUTOTUNE_CANDIDATES = [(g, nl) for g in (64, 128, 192, 256) for nl in (0, 1, 2, 4, 8)]


def lib():
    """Load libnemo_bh.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise ImportError("libnemo_bh.so not built: run `make -C nemo_amd`")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _p(a, t=_d):
    return a.ctypes.data_as(t)


def default_params(**kw) -> Params:
    p = Params()
    lib().bh_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "rmin":
            if v is not None:
                p.rmin[:] = [float(x) for x in v]
                p.rmin_given = 1
        elif k == "mode" and isinstance(v, str):
            p.mode = {"fast": MODE_FAST, "strict": MODE_STRICT, "direct": MODE_DIRECT}[v]
        elif k == "criterion" and isinstance(v, str):
            p.criterion = CRITERIA[v]
        else:
            setattr(p, k, v)
    return p


def mkplummer(n, seed=123, with_vel=True):
    """Equal-mass Plummer sphere in virial units (mkplummer recipe, own RNG)."""
    mass = np.zeros(n)
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3)) if with_vel else None
    lib().bh_mkplummer(n, seed, _p(mass), _p(pos), _p(vel) if with_vel else None)
    return mass, pos, vel


def mkdiskhalo(ndisk=1 << 21, nhalo=1 << 22, seed=123, mdisk=0.1, rdisk=1.0, zdisk=0.1, q=1.5, mhalo=1.0, ahalo=4.0):
    """Exponential disk in a Hernquist halo (bh_mkdiskhalo); disk first."""
    n = ndisk + nhalo
    mass = np.zeros(n)
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    lib().bh_mkdiskhalo(ndisk, nhalo, seed, mdisk, rdisk, zdisk, q, mhalo, ahalo, _p(mass), _p(pos), _p(vel))
    return mass, pos, vel


class Engine:
    """One device context (bh_ctx)."""

    def __init__(self, device=0):
        self._h = _vp()
        st = lib().bh_create(C.byref(self._h), device)
        if st != 0:
            raise BHError(st, "bh_create(device=%d) failed: no usable HIP device" % device)
        self.ntest = 0

    def close(self):
        if self._h:
            lib().bh_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != 0:
            raise BHError(st, lib().bh_last_error(self._h).decode())

    def upload(self, mass, pos, tpos=None):
        self._mass = np.ascontiguousarray(mass, np.float64)
        self._pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        n = len(self._mass)
        if tpos is not None:
            self._tpos = np.ascontiguousarray(tpos, np.float64).reshape(-1, 3)
            nt = len(self._tpos)
            self._check(lib().bh_upload(self._h, n, _p(self._mass), _p(self._pos), nt, _p(self._tpos)))
        else:
            nt = n
            self._check(lib().bh_upload(self._h, n, _p(self._mass), _p(self._pos), 0, None))
        self.ntest = nt

    def eval(self, params: Params) -> Stats:
        st = Stats()
        self._check(lib().bh_eval(self._h, C.byref(params), C.byref(st)))
        return st

    def download(self):
        phi = np.zeros(self.ntest)
        acc = np.zeros((self.ntest, 3))
        self._check(lib().bh_download(self._h, _p(phi), _p(acc)))
        return phi, acc

    def direct(self, eps, idx):
        idx = np.ascontiguousarray(idx, np.int64)
        phi = np.zeros(len(idx))
        acc = np.zeros((len(idx), 3))
        self._check(lib().bh_direct(self._h, eps, len(idx), _p(idx, _i64), _p(phi), _p(acc)))
        return phi, acc

    def direct_noalloc(self, eps, idx):
        """bh_direct with its staging in the library's fixed device array
        (no device allocation; at most DIRECT_NOALLOC_MAX targets)."""
        idx = np.ascontiguousarray(idx, np.int64)
        phi = np.zeros(len(idx))
        acc = np.zeros((len(idx), 3))
        self._check(lib().bh_direct_noalloc(self._h, eps, len(idx), _p(idx, _i64), _p(phi), _p(acc)))
        return phi, acc

    def force(self, mass, pos, params: Params, tpos=None):
        mass = np.ascontiguousarray(mass, np.float64)
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        n = len(mass)
        if tpos is not None:
            tpos = np.ascontiguousarray(tpos, np.float64).reshape(-1, 3)
            nt, tp = len(tpos), _p(tpos)
        else:
            nt, tp = n, None
        phi = np.zeros(nt)
        acc = np.zeros((nt, 3))
        st = Stats()
        self._check(lib().bh_force(self._h, C.byref(params), n, _p(mass), _p(pos), nt if tpos is not None else 0, tp,
                                   _p(phi), _p(acc), C.byref(st)))
        self.ntest = nt
        return phi, acc, st

    def debug_keys(self):
        k = np.zeros(len(self._mass), np.uint64)
        self._check(lib().bh_debug_keys(self._h, _p(k, _u64)))
        return k

    def debug_cells(self):
        nc = np.zeros(1, np.int64)
        self._check(lib().bh_debug_cells(self._h, 0, None, None, None, None, None, _p(nc, _i64)))
        n = int(nc[0])
        lv = np.zeros(n, np.int32)
        nb = np.zeros(n, np.int32)
        m = np.zeros(n)
        com = np.zeros((n, 3))
        q = np.zeros((n, 6))
        self._check(lib().bh_debug_cells(self._h, n, _p(lv, _i32), _p(nb, _i32), _p(m), _p(com), _p(q), _p(nc, _i64)))
        return dict(level=lv, nbody=nb, mass=m, pos=com, quad6=q)

    def debug_perm(self):
        p = np.zeros(self.ntest, np.uint32)
        self._check(lib().bh_debug_perm(self._h, p.ctypes.data_as(C.POINTER(C.c_uint32))))
        return p

    def debug_groups(self):
        ng = np.zeros(1, np.int64)
        self._check(lib().bh_debug_groups(self._h, 0, None, None, _p(ng, _i64)))
        n = int(ng[0])
        f = np.zeros(n, np.int32)
        c = np.zeros(n, np.int32)
        self._check(lib().bh_debug_groups(self._h, n, _p(f, _i32), _p(c, _i32), _p(ng, _i64)))
        return f, c

    # -- learned correction -------------------------------------------------
    # ---- batched ensembles and the autotuner ----
    def batch_upload(self, off, mass, pos, eps, tol):
        """Systems s = 0..len(off)-2 own bodies [off[s], off[s+1])."""
        off = np.ascontiguousarray(off, np.int64)
        mass = np.ascontiguousarray(mass, np.float64)
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        eps = np.ascontiguousarray(eps, np.float64)
        tol = np.ascontiguousarray(tol, np.float64)
        nsys = len(off) - 1
        if len(eps) != nsys or len(tol) != nsys or off[-1] != len(mass):
            raise ValueError("inconsistent batch arrays")
        self._check(lib().bh_batch_upload(self._h, nsys, _p(off, _i64), _p(mass), _p(pos), _p(eps), _p(tol)))
        self.ntest = len(mass)

    def autotune(self, params: Params, reps=3):
        """Returns (best Params, ms per candidate in AUTOTUNE_CANDIDATES order)."""
        best = Params()
        ms = np.zeros(len(AUTOTUNE_CANDIDATES))
        self._check(lib().bh_autotune(self._h, C.byref(params), reps, C.byref(best), _p(ms)))
        return best, ms

    def mlp_init(self, seed=1):
        self._check(lib().bh_mlp_init(self._h, seed))

    def mlp_set(self, params, norm=None):
        p = np.ascontiguousarray(params, np.float32)
        nrm = None if norm is None else np.ascontiguousarray(norm, np.float32)
        self._check(lib().bh_mlp_set(self._h, _p(p, _f), None if nrm is None else _p(nrm, _f)))

    def mlp_get(self):
        p = np.zeros(MLP_PARAMS, np.float32)
        nrm = np.zeros(4, np.float32)
        self._check(lib().bh_mlp_get(self._h, _p(p, _f), _p(nrm, _f)))
        return p, nrm

    def mlp_forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((len(x), MLP_OUT), np.float32)
        self._check(lib().bh_mlp_forward(self._h, len(x), _p(x, _f), _p(y, _f)))
        return y

    def mlp_grad(self, x, y):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        self._check(lib().bh_mlp_grad(self._h, len(x), _p(x, _f), _p(y, _f), _p(g, _f), _p(loss, _f)))
        return g, float(loss[0])

    def mlp_collect(self, params: Params, stride=1, rlo=0.25, rhi=1.6):
        n = np.zeros(1, np.int64)
        self._check(lib().bh_mlp_collect(self._h, C.byref(params), stride, rlo, rhi, _p(n, _i64)))
        return int(n[0])

    def mlp_samples(self):
        n = np.zeros(1, np.int64)
        self._check(lib().bh_mlp_samples(self._h, 0, None, None, _p(n, _i64)))
        k = int(n[0])
        x = np.zeros((k, MLP_IN), np.float32)
        y = np.zeros((k, MLP_OUT), np.float32)
        self._check(lib().bh_mlp_samples(self._h, k, _p(x, _f), _p(y, _f), _p(n, _i64)))
        return x, y

    def debug_samples(self):
        """Cell id and group centre of each sample of the last collection."""
        n = np.zeros(1, np.int64)
        self._check(lib().bh_debug_samples(self._h, 0, None, None, _p(n, _i64)))
        k = int(n[0])
        cell = np.zeros(k, np.int32)
        ctr = np.zeros((k, 3), np.float32)
        self._check(lib().bh_debug_samples(self._h, k, _p(cell, _i32), _p(ctr, _f), _p(n, _i64)))
        return cell, ctr

    def debug_keep_learned(self, on=True):
        lib().bh_debug_keep_learned(self._h, int(bool(on)))

    def debug_learned(self, with_res=True):
        """Entries of the last learned evaluation: (cell, group, res[n][12]
        (None without with_res; needs debug_keep_learned before the
        evaluation), centre[ngroups][4])."""
        n = np.zeros(1, np.int64)
        ng = np.zeros(1, np.int64)
        self._check(lib().bh_debug_learned(self._h, 0, None, None, None, _p(n, _i64), 0, None, _p(ng, _i64)))
        k, g = int(n[0]), int(ng[0])
        cell = np.zeros(k, np.int32)
        grp = np.zeros(k, np.int32)
        res = np.zeros((k, 12), np.float32) if with_res else None
        ctr = np.zeros((g, 4), np.float32)
        self._check(lib().bh_debug_learned(self._h, k, _p(cell, _i32), _p(grp, _i32),
                                           _p(res, _f) if with_res else None, _p(n, _i64), g,
                                           _p(ctr, _f), _p(ng, _i64)))
        return cell, grp, res, ctr

    def debug_cell_ranges(self):
        n = np.zeros(1, np.int64)
        self._check(lib().bh_debug_cell_ranges(self._h, 0, None, None, _p(n, _i64)))
        k = int(n[0])
        first = np.zeros(k, np.int32)
        count = np.zeros(k, np.int32)
        self._check(lib().bh_debug_cell_ranges(self._h, k, _p(first, _i32), _p(count, _i32), _p(n, _i64)))
        return first, count

    def mlp_nsamples(self):
        n = np.zeros(1, np.int64)
        self._check(lib().bh_mlp_samples(self._h, 0, None, None, _p(n, _i64)))
        return int(n[0])

    def mlp_clear_samples(self):
        self._check(lib().bh_mlp_clear_samples(self._h))

    def mlp_train(self, epochs=1, batch=8192, lr=1e-3, seed=1):
        loss = np.zeros(epochs, np.float32)
        self._check(lib().bh_mlp_train(self._h, epochs, batch, lr, seed, _p(loss, _f)))
        return loss

    def inject(self, code):
        lib().bh_debug_inject(self._h, code)


class Leapfrog:
    """hackcode1 time integration on the device (bh_lf_*).  Implements the
    stepper interface of nemo_amd.hackcode1.run: force(), diagnostics(),
    drift(), and mass/pos/vel (downloaded on access)."""

    def __init__(self, engine: Engine, mass, pos, vel, dt, batch=None, **params):
        """batch = (off, eps, tol) integrates an ensemble of systems
        (bh_batch_lf_init); diagnostics then have one row of 36 per system."""
        params.setdefault("rsize", 4.0)          # hackcode1 box: rmin=-2, rsize=4 (code.c:119-120)
        self.engine = engine
        self.mass = np.ascontiguousarray(mass, np.float64)
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, np.float64).reshape(-1, 3)
        self.n = len(self.mass)
        self.dt = dt
        self.params = default_params(**params)
        if batch is None:
            self.nsys = 0
            engine._check(lib().bh_lf_init(engine._h, C.byref(self.params), self.n, _p(self.mass), _p(pos), _p(vel),
                                           dt))
        else:
            off, eps, tol = (np.ascontiguousarray(a, t) for a, t in zip(batch, (np.int64, np.float64, np.float64)))
            self.nsys = len(off) - 1
            engine._check(lib().bh_batch_lf_init(engine._h, C.byref(self.params), self.nsys, _p(off, _i64),
                                                 _p(self.mass), _p(pos), _p(vel), _p(eps), _p(tol), dt))
        self._diag = None

    def _dshape(self, nsteps=None):
        s = (36,) if self.nsys == 0 else (self.nsys, 36)
        return s if nsteps is None else (nsteps,) + s

    def force(self):
        d = np.zeros(self._dshape())
        a = np.zeros(1, np.int64)
        b = np.zeros(1, np.int64)
        self.engine._check(lib().bh_lf_force(self.engine._h, _p(d), _p(a, _i64), _p(b, _i64)))
        self._diag = d
        return int(a[0]), int(b[0])

    def diagnostics(self):
        return self._diag

    def drift(self):
        self.engine._check(lib().bh_lf_drift(self.engine._h))

    def run(self, nsteps, graph=True):
        """nsteps full steps; returns (diag[nsteps,36] (batches:
        [nsteps,nsys,36]), n2b[nsteps], nbc[nsteps])."""
        d = np.zeros(self._dshape(nsteps))
        a = np.zeros(nsteps, np.int64)
        b = np.zeros(nsteps, np.int64)
        self.engine._check(lib().bh_lf_run(self.engine._h, nsteps, int(bool(graph)), _p(d), _p(a, _i64), _p(b, _i64)))
        return d, a, b

    def state(self):
        pos = np.zeros((self.n, 3))
        vel = np.zeros((self.n, 3))
        acc = np.zeros((self.n, 3))
        phi = np.zeros(self.n)
        ns = np.zeros(1, np.int64)
        self.engine._check(lib().bh_lf_get(self.engine._h, _p(pos), _p(vel), _p(acc), _p(phi), _p(ns, _i64)))
        return pos, vel, acc, phi, int(ns[0])

    @property
    def pos(self):
        return self.state()[0]

    @property
    def vel(self):
        return self.state()[1]


@dataclass
class ForceResult:
    phi: np.ndarray
    acc: np.ndarray
    n2btot: int
    nbctot: int
    rmin: list
    rsize: float
    cputree: float   # minutes, as hackforce's Diagnostics
    cpufcal: float


def hackforce(mass, pos, tpos=None, tol=1.0, eps=0.05, rsize=4.0, rmin=None, usequad=False, mode="fast",
              engine: Engine | None = None, **kw) -> ForceResult:
    """force_calc of hackforce on the GPU engine (hackforce.c:177-215)."""
    eng = engine or Engine()
    p = default_params(tol=tol, eps=eps, rsize=rsize, rmin=rmin, usequad=int(bool(usequad)), mode=mode, **kw)
    phi, acc, st = eng.force(mass, pos, p, tpos=tpos)
    return ForceResult(phi, acc, st.n2btot, st.nbctot, list(st.rmin), st.rsize, st.t_tree_ms / 60000.0,
                       st.t_walk_ms / 60000.0)
