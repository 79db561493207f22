"""
The oracle and the kernels' host arithmetic under AddressSanitizer and
UndefinedBehaviorSanitizer: sanitizer builds of oracle/hf_oracle.cpp (g++) and
of the harness's device derivative work items (clang, host code only; the s and
p quartet classes) run a
small RHF and UHF case in child processes with the matching sanitizer runtime
preloaded. Any sanitizer report fails the test.
"""

import os
import subprocess
import sys

import pytest

from tests.conftest import ROOT

ORACLE_CASE = r"""
import sys, numpy as np
sys.path.insert(0, {root!r})
from oracle import oracle as O
from tests.conftest import WATER, METHYL, bohr
O.set_threads(2)
for (s, x), basis, mult in ((bohr(WATER, 0.05, 1), "6-31g*", 1), (bohr(METHYL, 0.03, 1), "cc-pvdz", 2)):
    b = O.make_basis(s, x, basis)
    st = O.scf(b, mult=mult)
    g = O.scf_gradient(b, st)
    pr = O.Pairs(b)
    P = st.Ds[0] + st.Ds[-1]
    J, K = pr.jk_direct(P)
    O.one_electron_gradient_terms(b, P, P)
    print(s, basis, st.energy, float(np.abs(g).max()))
"""

HARNESS_CASE = r"""
import sys, numpy as np
sys.path.insert(0, {root!r}); sys.path.insert(0, {pkg!r})
from tests.harness.hostcheck import Harness
from tests.conftest import WATER, METHYL, bohr
from aimd_amd.qc.basis import build_basis
h = Harness({lib!r})
for (s, x), basis in ((bohr(WATER, 0.05, 1), "sto-3g"), (bohr(METHYL, 0.03, 1), "6-31g")):
    b = build_basis(s, x, basis)
    rng = np.random.default_rng(0)
    P = rng.standard_normal((b.nao, b.nao)) * 0.1
    P = P + P.T
    g = h.grad_two_variant(1, b, P, P[None], 0.5)
    print(s, basis, float(np.abs(g).max()))
"""


def _run(script, preload, env_extra):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1", LD_PRELOAD=preload,
               **env_extra)
    return subprocess.run([sys.executable, "-c", script], env=env, capture_output=True, text=True, timeout=600)


def _report(p):
    return p.returncode != 0 or "ERROR: AddressSanitizer" in p.stderr or "runtime error:" in p.stderr


def test_oracle_under_sanitizers():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "_build/liboracle_asan.so"], check=True,
                   stdout=subprocess.DEVNULL)
    gxx = lambda name: subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True,
                                      text=True).stdout.strip()
    p = _run(ORACLE_CASE.format(root=ROOT), f"{gxx('libasan.so')}:{gxx('libubsan.so')}",
             {"AIMD_ORACLE_LIB": os.path.join(ROOT, "oracle", "_build", "liboracle_asan.so")})
    assert not _report(p), p.stderr[-4000:]
    assert p.stdout.count("\n") == 2


def test_device_derivative_arithmetic_under_sanitizers():
    hdir = os.path.join(ROOT, "tests", "harness")
    subprocess.run(["make", "-C", hdir, "-j4", "_build/libhostcheck_asan.so"], check=True, stdout=subprocess.DEVNULL)
    rt = subprocess.run(["/opt/rocm/lib/llvm/bin/clang", "--print-file-name=libclang_rt.asan-x86_64.so"],
                        capture_output=True, text=True).stdout.strip()
    if not os.path.exists(rt):
        import glob
        found = glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
        if not found:
            pytest.skip("clang AddressSanitizer runtime not found")
        rt = found[0]
    p = _run(HARNESS_CASE.format(root=ROOT, pkg=os.path.join(ROOT, "ab-initio-md_amd"),
                                 lib=os.path.join(hdir, "_build", "libhostcheck_asan.so")), rt, {})
    assert not _report(p), p.stderr[-4000:]
    assert p.stdout.count("\n") == 2
