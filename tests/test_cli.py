"""The hackforce command (C, over the C ABI) and its snapshot codec.

CPU tests: the codec through `csf` (byte-swapped input, truncation, idempotent
copies) and hackforce's keyword and file checks, which fail before any GPU
call.  GPU tests: hackforce output against the oracle (bit-exact in strict
mode), the reference's known interaction total, test= and multi-snapshot
streams, options=, and the fast walk's error against the direct sum.
"""
import os
import subprocess

import numpy as np
import pytest

from nemo_amd import snapshot

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# NEMO_BIN: another build of the same programs (the sanitizer build of
# nemo_amd/san.mk)
BIN = os.environ.get("NEMO_BIN") or os.path.join(ROOT, "nemo_amd", "bin")


def run(prog, *args, cwd=None):
    return subprocess.run([os.path.join(BIN, prog)] + list(args), capture_output=True, cwd=cwd, timeout=300)


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(os.path.join(BIN, "csf")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "nemo_amd"), "bin/csf"], check=True)


def write_stream(path, snaps, history="test stream"):
    out = bytearray()
    snapshot.put_history(out, history)
    for m, x, v, t in snaps:
        snapshot.put_snapshot(out, m, x, v, t)
    with open(path, "wb") as f:
        f.write(bytes(out))


# ---- CPU: codec and argument handling ------------------------------------------------

@pytest.mark.parametrize("name", ["hack128.sun4", "plum10"])
def test_csf_copies_and_unswaps(golden, tmp_path, name):
    src = os.path.join(golden, name)
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    assert run("csf", src, a).returncode == 0
    assert run("csf", "in=" + a, "out=" + b).returncode == 0
    raw, cp = open(src, "rb").read(), open(a, "rb").read()
    assert len(raw) == len(cp)
    # big-endian input is written in host (little-endian) order, then copies are exact
    assert cp[:2] in (b"\x92\x09", b"\x92\x0b")
    assert open(b, "rb").read() == cp
    s0, s1 = snapshot.read(src)[0], snapshot.read(a)[0]
    assert np.array_equal(s0.mass, s1.mass) and np.array_equal(s0.pos, s1.pos) and np.array_equal(s0.vel, s1.vel)
    assert snapshot.bsf(raw) == snapshot.bsf(cp)


def test_csf_rejects_truncated_and_garbage(golden, tmp_path):
    raw = open(os.path.join(golden, "hack128.sun4"), "rb").read()
    for k, blob in enumerate((raw[:3000], raw[:-1], b"\x00\x01garbage")):
        p = tmp_path / ("t%d" % k)
        p.write_bytes(blob)
        r = run("csf", str(p), str(tmp_path / ("o%d" % k)))
        assert r.returncode == 255
        assert b"### Fatal error [csf]" in r.stderr


def test_hackforce_refuses_existing_output(golden, tmp_path):
    out = tmp_path / "exists"
    out.write_bytes(b"x")
    r = run("hackforce", os.path.join(golden, "hack128.sun4"), str(out))
    assert r.returncode == 255
    assert b"already exists" in r.stderr
    assert out.read_bytes() == b"x"


@pytest.mark.parametrize("args, msg", [
    (["out=/nonexistent/x"], b"must be given"),
    (["in=a", "out=b", "bogus=1"], b"unknown"),
    (["in=a", "out=b", "mode=fastest"], b"mode"),
    (["in=a", "out=b", "tol=abc"], b"not a number"),
    (["in=a", "out=b", "criterion=bh87"], b"criterion=bh87"),
    (["in=/nonexistent/a", "out=b"], b"cannot open"),
])
def test_hackforce_argument_errors(tmp_path, args, msg):
    r = run("hackforce", *args, cwd=str(tmp_path))
    assert r.returncode == 255, r
    assert b"### Fatal error [hackforce]" in r.stderr and msg in r.stderr, r.stderr


def test_hackforce_help_lists_reference_keywords():
    r = run("hackforce", "help")
    assert r.returncode == 0
    for k in (b"in=", b"out=", b"test=", b"tol=1.0", b"eps=0.05", b"rsize=4.0", b"rmin=", b"options=mass,phase",
              b"fcells=0.75"):
        assert k in r.stdout


# ---- GPU: results ---------------------------------------------------------------------

gpu = pytest.mark.gpu


def _items(path):
    return snapshot.parse(open(path, "rb").read())


@gpu
def test_hackforce_strict_matches_oracle_and_kat(golden, tmp_path):
    """Default keywords (tol=1, eps=0.05, monopole) on the reference's
    hack128 snapshot, strict mode: bit-identical to the oracle, and the
    reference's published interaction total 3864 at t=0."""
    from oracle import oracle

    src = os.path.join(golden, "hack128.sun4")
    out = str(tmp_path / "res")
    r = run("hackforce", "in=" + src, "out=" + out, "mode=strict")
    assert r.returncode == 0, r.stderr
    items = _items(out)
    hist = [it.data for it in items if it.tag == "History"]
    assert len(hist) == 2 and hist[-1].startswith(b"hackforce in=")
    snap = [it for it in items if it.tag == "SnapShot"]
    assert len(snap) == 1
    par, prt, dg = snap[0].find("Parameters"), snap[0].find("Particles"), snap[0].find("Diagnostics")
    assert [c.tag for c in par.children] == ["Nobj", "Time", "tol", "eps"]
    assert [c.tag for c in prt.children] == ["CoordSystem", "Mass", "PhaseSpace", "Potential", "Acceleration"]
    assert [c.tag for c in dg.children] == ["n2btot", "nbctot", "cputree", "cpufcal"]
    s = snapshot.read(src)[0]
    phi, acc, n2b, nbc, _, _ = oracle.force(s.mass, s.pos, tol=1.0, eps=0.05)
    assert np.array_equal(prt.find("Potential").data, phi)
    assert np.array_equal(prt.find("Acceleration").data, acc)
    assert int(dg.find("n2btot").data.reshape(-1)[0]) == n2b and int(dg.find("nbctot").data.reshape(-1)[0]) == nbc
    assert n2b + nbc == 3864
    assert np.array_equal(prt.find("PhaseSpace").data[:, 0, :], s.pos)


@gpu
@pytest.mark.parametrize("crit", ["rcrit", "bh86", "sw94"])
def test_hackforce_criterion_keyword_strict(tmp_path, crit):
    """criterion= in strict mode: the interaction totals written to
    Diagnostics equal treecode1's (restated) over the same root box, and the
    forces agree with it to FP64 rounding."""
    from nemo_amd import engine as E
    from oracle import oracle

    m, x, v = E.mkplummer(3000, seed=17)
    x = x * (1.9 / np.abs(x).max())      # inside the default box [-2, 2)^3
    write_stream(str(tmp_path / "in"), [(m, x, v, 0.0)])
    r = run("hackforce", "in", "out", "mode=strict", "tol=0.6", "eps=0.02", "usequad=t", "criterion=" + crit,
            cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    snap = [it for it in _items(str(tmp_path / "out")) if it.tag == "SnapShot"][0]
    prt, dg = snap.find("Particles"), snap.find("Diagnostics")
    t = oracle.tc_force(m, x, [-2.0, -2.0, -2.0], 4.0, theta=0.6, eps=0.02, criterion=crit, usequad=True)
    assert int(dg.find("n2btot").data.reshape(-1)[0]) == t["n2b"].sum()
    assert int(dg.find("nbctot").data.reshape(-1)[0]) == t["nbc"].sum()
    a = prt.find("Acceleration").data
    assert np.abs(a - t["acc"]).max() <= 1e-11 * np.abs(t["acc"]).max()


@gpu
def test_hackforce_qp_rmin_options(golden, tmp_path):
    """hackforce_qp (quadrupole cells), a given rmin= and options=phase."""
    from oracle import oracle

    src = os.path.join(golden, "hack128.sun4")
    out = str(tmp_path / "res")
    r = run("hackforce_qp", src, out, "tol=0.7", "eps=0.02", "rmin=-3,-3,-3", "rsize=6", "options=phase",
            "mode=strict")
    assert r.returncode == 0, r.stderr
    prt = [it for it in _items(out) if it.tag == "SnapShot"][0].find("Particles")
    assert prt.find("Mass") is None and prt.find("PhaseSpace") is not None
    s = snapshot.read(src)[0]
    phi, acc, _, _, _, _ = oracle.force(s.mass, s.pos, tol=0.7, eps=0.02, usequad=True, rsize=6.0,
                                        rmin=[-3.0, -3.0, -3.0])
    assert np.array_equal(prt.find("Potential").data, phi)
    assert np.array_equal(prt.find("Acceleration").data, acc)


@gpu
def test_hackforce_test_file_and_multisnapshot(tmp_path):
    """test=: masses from the first in= snapshot, one output per test
    snapshot; without test=, one output per in= snapshot."""
    from nemo_amd import engine as E
    from oracle import oracle

    m, x, v = E.mkplummer(3000, seed=5)
    tp1 = np.random.default_rng(1).normal(size=(500, 3))
    tp2 = np.random.default_rng(2).normal(size=(700, 3))
    write_stream(str(tmp_path / "mass"), [(m, x, v, 0.0), (m, x * 1.1, v, 1.0)])
    write_stream(str(tmp_path / "test"), [(np.zeros(500), tp1, tp1, 0.5), (np.zeros(700), tp2, tp2, 0.75)])
    r = run("hackforce", "in=mass", "out=o1", "test=test", "mode=strict", "tol=0.6", cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    outs = snapshot.read(str(tmp_path / "o1"))
    assert [len(o.mass) for o in outs] == [500, 700] and [o.time for o in outs] == [0.5, 0.75]
    for o, tp in zip(outs, (tp1, tp2)):
        phi, acc, _, _, _, _ = oracle.force(m, x, tol=0.6, eps=0.05, tpos=tp)
        assert np.array_equal(o.extra["Potential"], phi) and np.array_equal(o.extra["Acceleration"], acc)
    r = run("hackforce", "in=mass", "out=o2", "mode=strict", "tol=0.6", cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    outs = snapshot.read(str(tmp_path / "o2"))
    assert [o.time for o in outs] == [0.0, 1.0]
    phi, acc, _, _, _, _ = oracle.force(m, x * 1.1, tol=0.6, eps=0.05)
    assert np.array_equal(outs[1].extra["Acceleration"], acc)


@gpu
def test_hackforce_fast_and_direct_modes(tmp_path):
    """The default fast walk is within the tolerance band of the FP64
    direct sum (mode=direct), itself equal to the oracle's direct sum."""
    from nemo_amd import engine as E
    from oracle import oracle

    m, x, v = E.mkplummer(20000, seed=8)
    write_stream(str(tmp_path / "in"), [(m, x, v, 0.0)])
    assert run("hackforce_qp", "in", "fast", "tol=0.5", "eps=0.01", cwd=str(tmp_path)).returncode == 0
    r = run("hackforce", "in", "direct", "eps=0.01", "mode=direct", cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    fa = snapshot.read(str(tmp_path / "fast"))[0].extra["Acceleration"]
    da = snapshot.read(str(tmp_path / "direct"))[0].extra["Acceleration"]
    idx = np.arange(0, 20000, 97)
    _, oacc = oracle.direct(m, x, 0.01, idx)
    np.testing.assert_allclose(da[idx], oacc, rtol=1e-12, atol=1e-12)
    e = np.linalg.norm(fa - da, axis=1) / np.linalg.norm(da, axis=1)
    assert np.median(e) < 1e-3 and np.percentile(e, 99) < 1e-2


@gpu
def test_hackforce_local_keyword_equals_library(tmp_path):
    """local= sets bh_params.local_alpha: the program's output equals
    bh_force with the same parameters bit for bit; an out-of-range value is
    the library's argument error."""
    from nemo_amd import engine as E

    m, x, v = E.mkplummer(30000, seed=12)
    write_stream(str(tmp_path / "in"), [(m, x, v, 0.0)])
    r = run("hackforce", "in", "out", "tol=0.7", "eps=0.01", "group=192", "nleaf=1", "local=0.2", cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    got = snapshot.read(str(tmp_path / "out"))[0]
    e = E.Engine(0)
    try:
        phi, acc, st = e.force(m, x, E.default_params(tol=0.7, eps=0.01, rsize=4.0, group_size=192, nleaf=1,
                                                       local_alpha=0.2))
    finally:
        e.close()
    assert st.n_local > 0
    assert np.array_equal(got.extra["Potential"], phi) and np.array_equal(got.extra["Acceleration"], acc)
    r = run("hackforce", "in", "out2", "local=0.7", cwd=str(tmp_path))
    assert r.returncode == 255 and b"local_alpha" in r.stderr, r.stderr


# ---- hackcode1 -------------------------------------------------------------------------

def test_hackcode1_help_and_argument_errors(tmp_path):
    r = run("hackcode1", "help")
    assert r.returncode == 0
    for k in (b"in=", b"out=", b"restart=", b"continue=", b"save=", b"nbody=128", b"freq=32.0", b"eps=0.05",
              b"tol=1.0", b"options=mass,phase", b"tstop=2.0", b"freqout=4.0", b"minor_freqout=32.0", b"log=-"):
        assert k in r.stdout
    (tmp_path / "exists").write_bytes(b"x")
    for args, msg in ((["out=exists"], b"already exists"), (["bogus=1"], b"unknown"),
                      (["continue=/nonexistent/state"], b"cannot open"), (["in=/nonexistent/snap"], b"cannot open")):
        r = run("hackcode1", *args, cwd=str(tmp_path))
        assert r.returncode == 255 and b"### Fatal error [hackcode1]" in r.stderr and msg in r.stderr, (args, r.stderr)


def _frames(path):
    """(time, pos, vel) of the frames that hold particles."""
    out = []
    for it in snapshot.parse(open(path, "rb").read()):
        if it.tag != "SnapShot":
            continue
        t = float(it.find("Parameters").find("Time").data.reshape(-1)[0])
        prt = it.find("Particles")
        if prt is not None and prt.find("PhaseSpace") is not None:
            ps = prt.find("PhaseSpace").data
            out.append((t, ps[:, 0, :], ps[:, 1, :]))
    return out


@gpu
def test_hackcode1_default_run_reproduces_reference_kats(golden, tmp_path):
    """The 64-step default run from the reference's t=0 frame, strict mode:
    the output's bsf summary equals the reference Testfile KAT and the log's
    interaction totals equal the published 3864 (t=0) and 4899 (t=2)."""
    import json

    kats = json.load(open(os.path.join(golden, "kats.json")))
    src = os.path.join(golden, "hack128.sun4")
    r = run("hackcode1", "in=" + src, "out=run.out", "mode=strict", cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    assert snapshot.bsf(open(str(tmp_path / "run.out"), "rb").read()) == kats["hackcode1_default_bsf"]["value"]
    rows = [l.split() for l in r.stdout.decode().splitlines()]
    data = [x for x in rows if len(x) == 7 and x[0].replace(".", "").isdigit()]
    assert len(data) == 65
    assert int(data[0][3]) == kats["hackcode1_t0"]["nttot"] and int(data[-1][3]) == kats["hackcode1_t2"]["nttot"]
    assert b"particle data written" in r.stdout


@gpu
def test_hackcode1_fast_small_system(golden, tmp_path):
    """The default run in fast mode: 128 bodies form one group, so every
    body sees every other one; the t=0 energy equals the FP64 direct sum's
    (to the log's 4 decimals) and drifts less than the reference's own run
    (tree forces, -0.2675 -> -0.2647)."""
    from oracle import oracle

    src = os.path.join(golden, "hack128.sun4")
    r = run("hackcode1", "in=" + src, "out=fast", "mode=fast", cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    t = [l.split() for l in r.stdout.decode().splitlines()]
    rows = [x for x in t if len(x) == 7 and x[0].replace(".", "").isdigit()]
    assert len(rows) == 65
    s0 = snapshot.read(src)[0]
    phi, _ = oracle.direct(s0.mass, s0.pos, 0.05)
    e0 = 0.5 * (s0.mass * (s0.vel ** 2).sum(1)).sum() + 0.5 * (s0.mass * phi).sum()
    assert abs(float(rows[0][1]) - e0) <= 6e-5
    assert int(rows[0][3]) == 128 * 127
    assert max(abs(float(x[1]) - float(rows[0][1])) for x in rows) <= 0.0028


@gpu
def test_hackcode1_criterion_keyword(golden, tmp_path):
    """criterion= in hackcode1 (strict walk): the t=0 interaction total of the
    log equals treecode1's (restated) over hackcode1's box."""
    from oracle import oracle

    src = os.path.join(golden, "hack128.sun4")
    r = run("hackcode1", "in=" + src, "out=c", "mode=strict", "criterion=sw94", "tstop=0.25", cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    rows = [x for x in (l.split() for l in r.stdout.decode().splitlines())
            if len(x) == 7 and x[0].replace(".", "").isdigit()]
    assert len(rows) == 9
    s0 = snapshot.read(src)[0]
    assert np.abs(s0.pos).max() < 2.0
    t = oracle.tc_force(s0.mass, s0.pos, [-2.0, -2.0, -2.0], 4.0, theta=1.0, eps=0.05, criterion="sw94")
    assert int(rows[0][3]) == t["n2b"].sum() + t["nbc"].sum()


@gpu
def test_hackcode1_save_and_restart_continue_bitwise(golden, tmp_path):
    """A run saved at t=1 and restarted to t=2 ends bit-identical to a
    straight run to t=2 (strict and fast walk)."""
    src = os.path.join(golden, "hack128.sun4")
    for mode in ("strict", "fast"):
        d = tmp_path / mode
        d.mkdir()
        assert run("hackcode1", "in=" + src, "out=a", "mode=" + mode, "usequad=t", "tol=0.7",
                   cwd=str(d)).returncode == 0
        r = run("hackcode1", "in=" + src, "out=b1", "tstop=1", "save=state", "mode=" + mode, "usequad=t", "tol=0.7",
                cwd=str(d))
        assert r.returncode == 0, r.stderr
        r = run("hackcode1", "restart=state", "out=b2", "tstop=2", "mode=" + mode, "usequad=t", "tol=0.7",
                cwd=str(d))
        assert r.returncode == 0, r.stderr
        fa, fb = _frames(str(d / "a")), _frames(str(d / "b2"))
        assert fa[-1][0] == fb[-1][0] == 2.0
        assert np.array_equal(fa[-1][1], fb[-1][1]) and np.array_equal(fa[-1][2], fb[-1][2]), mode


# ---- acceleration(5) plugin ------------------------------------------------------------

def _plugin():
    import ctypes as C

    lib = C.CDLL(os.path.join(BIN, "bhtree.so"))
    ACC = C.CFUNCTYPE(None, C.c_int, C.c_double, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                      C.POINTER(C.c_int), C.c_void_p, C.c_void_p, C.c_int, C.c_char)
    lib.iniacceleration.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_char_p, C.POINTER(ACC),
                                    C.POINTER(C.c_bool), C.POINTER(C.c_bool)]
    lib.iniacceleration.restype = None
    return lib, ACC


def test_plugin_exports_iniacceleration():
    import ctypes as C

    lib, ACC = _plugin()
    pars = (C.c_double * 3)(0.7, 0.02, 1.0)
    f = ACC()
    nm, nv = C.c_bool(False), C.c_bool(True)
    lib.iniacceleration(pars, 3, None, C.byref(f), C.byref(nm), C.byref(nv))
    assert nm.value is True and nv.value is False and bool(f)


@gpu
def test_plugin_matches_bh_force():
    """acceleration(5) through bhtree.so: double and float arrays, active
    flags and the add/assign indicator, against bh_force."""
    import ctypes as C

    from nemo_amd import engine as E

    lib, ACC = _plugin()
    pars = (C.c_double * 3)(0.7, 0.02, 1.0)
    f = ACC()
    nm, nv = C.c_bool(), C.c_bool()
    lib.iniacceleration(pars, 3, None, C.byref(f), C.byref(nm), C.byref(nv))
    m, x, _ = E.mkplummer(5000, seed=4)
    eng = E.Engine(0)
    phi0, acc0, _ = eng.force(m, x, E.default_params(tol=0.7, eps=0.02, usequad=1))
    eng.close()
    pot = np.zeros(5000)
    acc = np.zeros((5000, 3))
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    f(3, 0.0, 5000, vp(m), vp(x), None, None, vp(pot), vp(acc), 0, b"d")
    assert np.array_equal(pot, phi0) and np.array_equal(acc, acc0)
    # add to existing values, only the active bodies
    flags = (np.arange(5000) % 3 == 0).astype(np.int32)
    pot2 = np.ones(5000)
    acc2 = np.ones((5000, 3))
    f(3, 0.0, 5000, vp(m), vp(x), None, flags.ctypes.data_as(C.POINTER(C.c_int)), vp(pot2), vp(acc2), 3, b"d")
    on = flags == 1
    assert np.array_equal(pot2[on], 1.0 + phi0[on]) and np.all(pot2[~on] == 1.0)
    assert np.array_equal(acc2[on], 1.0 + acc0[on]) and np.all(acc2[~on] == 1.0)
    # float arrays
    mf, xf = m.astype(np.float32), x.astype(np.float32)
    potf = np.zeros(5000, np.float32)
    accf = np.zeros((5000, 3), np.float32)
    f(3, 0.0, 5000, vp(mf), vp(xf), None, None, vp(potf), vp(accf), 0, b"f")
    np.testing.assert_allclose(accf, acc0, rtol=1e-3, atol=1e-3 * np.abs(acc0).max())
    # fourth parameter: treecode1's sw94 criterion (bh_criterion 3)
    pars4 = (C.c_double * 4)(0.7, 0.02, 1.0, 3.0)
    lib.iniacceleration(pars4, 4, None, C.byref(f), C.byref(nm), C.byref(nv))
    eng = E.Engine(0)
    phi3, acc3, _ = eng.force(m, x, E.default_params(tol=0.7, eps=0.02, usequad=1, criterion="sw94"))
    eng.close()
    f(3, 0.0, 5000, vp(m), vp(x), None, None, vp(pot), vp(acc), 0, b"d")
    assert np.array_equal(pot, phi3) and np.array_equal(acc, acc3)
    assert not np.array_equal(acc3, acc0)


def test_csf_convert_and_coerced_reads(golden, tmp_path):
    """csf convert=d2f / f2d (copy_item_cvt) and the codec's typed reads:
    get_data is rigid (a float PhaseSpace is refused, as hackforce refuses
    it), get_data_coerced converts float to double; shapes are checked."""
    src = os.path.join(golden, "plum10")
    f, d = str(tmp_path / "f"), str(tmp_path / "d")
    assert run("csf", "in=" + src, "out=" + f, "convert=d2f").returncode == 0
    r = run("csf", "in=" + f, "out=" + d, "convert=f2d,i2f")
    assert r.returncode == 0
    assert b"### Warning [csf]: Cannot convert i2f yet in Nobj" in r.stderr
    s0 = snapshot.read(src)[0]

    def particles(path):
        for it in snapshot.parse(open(path, "rb").read()):
            if it.tag == "SnapShot":
                return it.find("Particles")

    pf, pd = particles(f), particles(d)
    assert pf.find("PhaseSpace").type == "f" and pf.find("Mass").type == "f"
    assert pd.find("PhaseSpace").type == "d" and pd.find("CoordSystem").type == "i"
    x32 = np.asarray(pf.find("PhaseSpace").data, np.float32).reshape(-1, 2, 3)
    assert np.array_equal(x32[:, 0], s0.pos.astype(np.float32)) and np.array_equal(x32[:, 1], s0.vel.astype(np.float32))
    x64 = np.asarray(pd.find("PhaseSpace").data, np.float64).reshape(-1, 2, 3)
    assert np.array_equal(x64, x32.astype(np.float64))
    # float snapshots are refused by the rigid reader of hackforce
    r = run("hackforce", "in=" + f, "out=" + str(tmp_path / "o"))
    assert r.returncode == 255 and b"### Fatal error [hackforce]" in r.stderr and b"has type f, d expected" in r.stderr
    # typed reads through the C codec
    exe = str(tmp_path / "codec_check")
    cli = os.path.join(ROOT, "nemo_amd", "cli")
    san = os.environ.get("NEMO_SAN_CFLAGS", "").split()
    subprocess.run(["gcc", "-O1", "-std=gnu11"] + san + ["-I", cli, "-o", exe,
                    os.path.join(ROOT, "tests", "helpers", "codec_check.c"), os.path.join(cli, "filestruct.c")],
                   check=True)
    out = subprocess.run([exe, f], capture_output=True, text=True, check=True).stdout.splitlines()
    res = {ln.split()[0]: ln for ln in out}
    assert res["rigid"].startswith("rigid -1") and "types f, d don't match" in res["rigid"]
    assert res["coerced"] == "coerced 0 ok" and res["mass"] == "mass 0 ok"
    vals = np.array([float(v) for v in res["phase"].split()[1:]]).reshape(-1, 2, 3)
    assert np.array_equal(vals, x32.astype(np.float64))
    assert "dimensions don't match" in res["baddims"]
    assert "can't copy plural to scalar" in res["scalar"]
    assert "types f, i don't convert" in res["toint"]
    out = subprocess.run([exe, d], capture_output=True, text=True, check=True).stdout.splitlines()
    assert "rigid 0 ok" in out


def test_csf_item_select_headline_blocksize(tmp_path):
    """csf's other keywords (csf.c:27-34): item= keeps the named top-level
    items, select= the given positions among them, headline= writes a
    Headline item first, blocksize= copies raw bytes."""
    rng = np.random.default_rng(4)
    snaps = [(np.full(20, 0.05), rng.normal(size=(20, 3)), rng.normal(size=(20, 3)), float(t)) for t in range(3)]
    src = str(tmp_path / "s")
    write_stream(src, snaps)
    out = str(tmp_path / "o2")
    r = run("csf", "in=" + src, "out=" + out, "item=SnapShot", "select=2", "headline=picked")
    assert r.returncode == 0 and b"On top-level: 2 items  read, 1 items written" in r.stderr
    items = snapshot.parse(open(out, "rb").read())
    assert [it.tag for it in items] == ["Headline", "SnapShot"]
    assert bytes(items[0].data).rstrip(b"\x00") == b"picked"
    s = list(snapshot.snapshots(items))[0]
    assert s.time == 1.0 and np.array_equal(s.pos, snaps[1][1]) and np.array_equal(s.vel, snaps[1][2])
    r = run("csf", "in=" + src, "out=" + str(tmp_path / "h"), "item=History")
    assert r.returncode == 0 and [it.tag for it in snapshot.parse(open(tmp_path / "h", "rb").read())] == ["History"]
    r = run("csf", "in=" + src, "out=" + str(tmp_path / "all"), "select=1,3", "item=SnapShot")
    assert [s.time for s in snapshot.read(str(tmp_path / "all"))] == [0.0, 2.0]
    for bad in ("select=0", "select=3,2"):
        r = run("csf", "in=" + src, "out=" + str(tmp_path / ("b" + bad[-1])), bad)
        assert r.returncode == 255 and b"### Fatal error [csf]" in r.stderr
    r = run("csf", "in=" + src, "out=" + str(tmp_path / "raw"), "blocksize=7")
    assert r.returncode == 0 and open(tmp_path / "raw", "rb").read() == open(src, "rb").read()


@gpu
@pytest.mark.parametrize("cencon", ["false", "true"])
def test_hackcode1_test_model(tmp_path, cencon):
    """Without in=, hackcode1 draws the reference's test model (make_testdata,
    code.c:129-153): equal masses 1/nbody, positions uniform in the unit ball
    (M(r) = r^3) or with cencon on shells of radius uniform in [0, 1)
    (M(r) = r), velocities uniform in the unit ball, centre of mass at rest
    at the origin; the same seed gives the same model."""
    n = 4000
    outs = []
    for k in range(2):
        o = "m%d" % k
        r = run("hackcode1", "out=" + o, "nbody=%d" % n, "seed=7", "cencon=" + cencon, "tstop=0", cwd=str(tmp_path))
        assert r.returncode == 0, r.stderr
        outs.append(snapshot.read(str(tmp_path / o))[0])
    s = outs[0]
    assert np.array_equal(s.pos, outs[1].pos) and np.array_equal(s.vel, outs[1].vel)
    assert np.all(s.mass == 1.0 / n)
    assert np.abs(s.pos.mean(0)).max() < 1e-12 and np.abs(s.vel.mean(0)).max() < 1e-12
    r = np.linalg.norm(s.pos, axis=1)
    v = np.linalg.norm(s.vel, axis=1)
    assert r.max() < 1.1 and v.max() < 1.1
    frac = (r < 0.5).mean()
    assert abs(frac - (0.5 if cencon == "true" else 0.125)) < 0.03, frac
