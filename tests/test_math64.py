"""The FP64 transcendentals of the converged reference (zillumgl_amd/csrc/zgl_math.h
sincosF64 / atan2F64 / logF64 / expF64 / powF64; the oracle's double mode restates them,
oracle/zgl_oracle.cpp sincosD ...).

They replace the libraries' double routines so that the GPU and the host give the same
bits.  Pinned here against numpy's double libm: within a few ulp of double (the methods
are Cody-Waite reduction plus series whose truncation error is far below an ulp; what is
left is the rounding of the Horner steps), the special cases (atan2(0, 0) = 0, log of 0
and of subnormals, exp overflow/underflow) and the product's host build against the
oracle's restatement bit for bit."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
OPS = {"sin": 0, "cos": 1, "atan2": 2, "log": 3, "exp": 4, "pow": 5}


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("math64") / "math64_probe.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC",
                    os.path.join(HERE, "native", "math64_probe.cpp"), "-o", out], check=True)
    L = C.CDLL(out)
    dp = C.POINTER(C.c_double)
    L.probe_dmath.argtypes = [C.c_int, dp, dp, dp, C.c_int64]
    L.probe_dmath.restype = None

    def run(op, a, b=None):
        x = np.ascontiguousarray(a, np.float64)
        y = np.ascontiguousarray(b if b is not None else np.zeros_like(x), np.float64)
        o = np.empty_like(x)
        L.probe_dmath(OPS[op], x.ctypes.data_as(dp), y.ctypes.data_as(dp), o.ctypes.data_as(dp), x.size)
        return o
    return run


def _ulps(got, ref, floor=None):
    sp = np.spacing(np.abs(ref))
    if floor is not None:
        sp = np.maximum(sp, floor)
    return np.abs(got - ref) / sp


def test_sin_cos_within_4ulp():
    x = np.random.default_rng(11).uniform(-20, 20, 200_000)
    s, c = orc.dmath("sin", x), orc.dmath("cos", x)
    # near the functions' zeros the error is absolute: measure it against an ulp of 1/4
    assert _ulps(s, np.sin(x), np.spacing(0.25)).max() <= 4.0
    assert _ulps(c, np.cos(x), np.spacing(0.25)).max() <= 4.0
    assert orc.dmath("sin", np.array([0.0]))[0] == 0.0 and orc.dmath("cos", np.array([0.0]))[0] == 1.0


def test_atan2_within_4ulp_and_special_cases():
    rng = np.random.default_rng(12)
    y, x = rng.normal(size=200_000), rng.normal(size=200_000)
    assert _ulps(orc.dmath("atan2", y, x), np.arctan2(y, x)).max() <= 4.0
    # the reduction's boundary and the axes
    t = np.tan(np.pi / 8)
    edge = np.array([t, np.nextafter(t, 0), np.nextafter(t, 1), 1.0, 1e-300, 1e300])
    assert _ulps(orc.dmath("atan2", edge, np.ones_like(edge)), np.arctan2(edge, 1.0)).max() <= 4.0
    assert orc.dmath("atan2", np.array([0.0]), np.array([0.0]))[0] == 0.0
    assert orc.dmath("atan2", np.array([1.0]), np.array([0.0]))[0] == np.pi / 2
    assert orc.dmath("atan2", np.array([0.0]), np.array([-1.0]))[0] == np.pi


def test_log_exp_pow_within_4ulp():
    rng = np.random.default_rng(13)
    x = np.exp(rng.uniform(-700, 700, 200_000))
    assert _ulps(orc.dmath("log", x), np.log(x), np.spacing(1.0)).max() <= 4.0
    sub = np.array([5e-324, 1e-310, 2.2e-308])
    assert _ulps(orc.dmath("log", sub), np.log(sub)).max() <= 4.0
    assert orc.dmath("log", np.array([0.0]))[0] == -np.inf and np.isnan(orc.dmath("log", np.array([-1.0]))[0])
    e = rng.uniform(-700, 700, 200_000)
    assert _ulps(orc.dmath("exp", e), np.exp(e)).max() <= 4.0
    assert orc.dmath("exp", np.array([800.0]))[0] == np.inf and orc.dmath("exp", np.array([-800.0]))[0] == 0.0
    a, b = rng.uniform(1e-4, 1.0, 20_000), rng.uniform(0.0, 1.0, 20_000)
    assert (np.abs(orc.dmath("pow", a, b) / np.power(a, b) - 1) < 1e-14).all()


@pytest.mark.parametrize("op", ["sin", "cos", "atan2", "log", "exp", "pow"])
def test_product_host_build_equals_oracle_bit_for_bit(probe, op):
    rng = np.random.default_rng(14 + OPS[op])
    if op in ("sin", "cos"):
        a, b = rng.uniform(-40, 40, 100_000), None
    elif op == "atan2":
        a, b = rng.normal(size=100_000), rng.normal(size=100_000)
    elif op == "log":
        a, b = np.exp(rng.uniform(-740, 700, 100_000)), None
    elif op == "exp":
        a, b = rng.uniform(-740, 700, 100_000), None
    else:
        a, b = rng.uniform(1e-6, 1.0, 100_000), rng.uniform(0.0, 1.0, 100_000)
    assert np.array_equal(probe(op, a, b).view(np.uint64), orc.dmath(op, a, b).view(np.uint64))
