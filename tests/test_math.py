"""CPU checks of the device math in arbor-cuda_amd/csrc/arbor_math.hpp, built
for the host from the same source (tests/mathcheck.cpp, no FMA contraction):

* sincosBounded against double-precision sin/cos: at most 1.6 ulp over the
  argument ranges the render path uses (and up to 2e4), including arguments
  next to multiples of pi/2 where the reduction cancels;
* divKnown(a, b, RN(1/b)) equal to IEEE a/b in every bit, for the divisors
  the leaf and bark code divide by (the seed-7 leaves' lengths and squared
  half-widths, 0.005, 2e = 0.0003, 2*pi, 0.19) and for random divisors.
"""
import ctypes as ct
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "arbor-cuda_amd", "csrc")


@pytest.fixture(scope="module")
def mc(tmp_path_factory):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("no host C++ compiler")
    out = str(tmp_path_factory.mktemp("mathcheck") / "libmathcheck.so")
    subprocess.check_call([gxx, "-O2", "-std=c++17", "-shared", "-fPIC", "-ffp-contract=off", "-I" + CSRC,
                           os.path.join(HERE, "mathcheck.cpp"), "-o", out])
    L = ct.CDLL(out)
    fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    L.mc_sincos.argtypes = [fp, fp, fp, ct.c_long]
    L.mc_div_known.argtypes = [fp, fp, fp, fp, ct.c_long]
    return L


def _ulps(approx, exact):
    ref = exact.astype(np.float32)
    return np.abs(approx.astype(np.float64) - exact) / np.spacing(np.abs(ref)).astype(np.float64)


@pytest.mark.parametrize("lo,hi", [(-4.0, 4.0), (-64.0, 64.0), (-1000.0, 1000.0), (-20000.0, 20000.0)])
def test_sincos_bounded_accuracy(mc, lo, hi):
    rng = np.random.default_rng(int(hi))
    x = rng.uniform(lo, hi, 1_000_000).astype(np.float32)
    k = rng.integers(int(lo / 1.5708), int(hi / 1.5708) + 1, 50_000)
    near = (k * (np.pi / 2)).astype(np.float32)
    x = np.concatenate([x, near, np.nextafter(near, np.float32(np.inf)), np.nextafter(near, np.float32(-np.inf)),
                        np.float32([0.0, -0.0, 1e-30, -1e-30, np.pi / 4, -np.pi / 4])]).astype(np.float32)
    s, c = np.empty_like(x), np.empty_like(x)
    mc.mc_sincos(x, s, c, len(x))
    xs = x.astype(np.float64)
    es, ec = _ulps(s, np.sin(xs)), _ulps(c, np.cos(xs))
    print("[%g, %g]: sin max %.3f ulp, cos max %.3f ulp" % (lo, hi, es.max(), ec.max()))
    assert es.max() <= 1.6 and ec.max() <= 1.6
    # a plausible slip (a dropped Taylor term, a wrong quadrant) shows up far above that
    assert np.abs(s * s + c * c - 1.0).max() < 1e-6


def test_div_known_is_ieee_division(mc, oracle):
    P = oracle.generate(7).P.reshape(-1, 24)
    leaves = P[P[:, 3] == 1]
    divisors = np.concatenate([leaves[:, 7], leaves[:, 11] * leaves[:, 11],
                               np.float32([0.005, 2.0 * np.float32(0.00015), 2.0 * np.float32(np.pi), 0.19, 7, 11, 23]),
                               np.random.default_rng(5).uniform(1e-6, 1e3, 2000).astype(np.float32)]).astype(np.float32)
    rng = np.random.default_rng(9)
    n_per = 400
    b = np.repeat(divisors, n_per).astype(np.float32)
    a = rng.uniform(-1.0, 1.0, len(b)).astype(np.float32) * np.float32(10.0) ** rng.integers(-6, 3, len(b)).astype(
        np.float32)
    a[::97] = 0.0
    a[1::97] = -0.0
    rb = (np.float32(1.0) / b).astype(np.float32)  # numpy float32 division: correctly rounded
    q = np.empty_like(a)
    mc.mc_div_known(a, b, rb, q, len(a))
    want = a / b
    assert np.array_equal(q.view(np.uint32), want.view(np.uint32))
    # and the check can fail: truncating the correction step leaves many quotients off by one ulp
    q_bad = (a * rb).astype(np.float32)
    assert not np.array_equal(q_bad.view(np.uint32), want.view(np.uint32))
