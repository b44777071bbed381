"""The C-ABI library loads and exports every symbol include/bh.h declares
(no device needed); host-only entry points work on the CPU."""
import os
import re

import numpy as np

from nemo_amd import engine as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "bh.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bh_[a-z0-9_]+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    L = E.lib()
    names = _declared()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
    assert set(names) <= set(E.SIGNATURES), set(names) - set(E.SIGNATURES)


def test_default_params_match_hackforce_defaults():
    p = E.default_params()
    # hackforce.c:42-47: tol=1.0 eps=0.05 rsize=4.0 rmin=-rsize/2
    assert (p.tol, p.eps, p.rsize, p.rmin_given, p.usequad) == (1.0, 0.05, 4.0, 0, 0)


def test_mkplummer_virial_units():
    m, x, v = E.mkplummer(20000, seed=123)
    assert np.isclose(m.sum(), 1.0, rtol=1e-14)
    assert np.abs(x.mean(0)).max() < 1e-14 and np.abs(v.mean(0)).max() < 1e-14
    # virial units: T = 1/4 up to sampling noise; half-mass radius ~0.77
    T = 0.5 * (m * (v * v).sum(1)).sum()
    assert abs(T - 0.25) < 0.01
    r = np.sort(np.linalg.norm(x, axis=1))
    assert abs(r[len(r) // 2] - 0.769) < 0.02
    assert r[-1] < 22.81
    m2, x2, v2 = E.mkplummer(20000, seed=123)
    assert np.array_equal(x, x2) and np.array_equal(v, v2)


def test_create_without_device_reports_error():
    import ctypes as C
    if os.environ.get("HIP_VISIBLE_DEVICES", None) is None and os.path.exists("/dev/kfd"):
        return  # a device may be present; covered by the GPU tests
    h = C.c_void_p()
    st = E.lib().bh_create(C.byref(h), 0)
    assert st == 2


def test_every_hip_compile_disables_packed_fp32():
    """Every hipcc compile in nemo_amd/Makefile takes $(COMMON), which turns
    off the packed FP32 instructions; the MFMA register-form option is
    limited to the MLP kernels' file."""
    mk = open(os.path.join(ROOT, "nemo_amd", "Makefile")).read()
    common = re.search(r"^COMMON\s*=\s*(.*?)(?<!\\)\n", mk, flags=re.S | re.M).group(1)
    assert "-Xclang -target-feature -Xclang -packed-fp32-ops" in common
    compiles = [l.strip() for l in mk.splitlines() if l.startswith("\t$(HIPCC)") and " -c " in l]
    assert len(compiles) >= 6
    for line in compiles:
        assert "$(COMMON)" in line, line
    rules = re.findall(r"^(\S+\.o):.*\n\t(.*)$", mk, flags=re.M)
    with_vgpr = [obj for obj, cmd in rules if "$(MFMAVGPR)" in cmd]
    assert with_vgpr == ["$(OBJDIR)/bh_mlp.o"], with_vgpr
