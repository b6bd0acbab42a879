"""The device powers of the physics phases (skewtcloudsim_amd/csrc/skt_device.h)
against the exact power, in ulps of the float result.

powNear1 raises numbers near 1 to powers of order 1: the pressure ratio
p / p_ground to Rd/cp (0.286) in every temperature conversion
(Game/kernelSky.cu:1445 and the other theta -> T sites), and the temperature
ratio T/273.15 to 1.94 and 1.5 in the vapour diffusivity and the air
viscosity (Game/meteoformulas.cuh). powPos takes every other fractional or
integer power of the microphysics rates (Game/microPhysics.cu: slopes to
0.25, powers of the slopes from -7 to 6.8, ...). The reference uses powf,
which is within an ulp of the exact power; the bounds here follow from the
hardware logarithm's and exponential's ulp (for powPos the logarithm's error
is multiplied by |y|), and each case prints what it measured.
"""
import ctypes
import os

import numpy as np
import pytest

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "libpowcheck.so")
KAPPA = np.float32(np.float32(287.0528) / np.float32(1003.5))


def _lib():
    if not os.path.exists(LIB):
        pytest.fail("tests/native/libpowcheck.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB)
    fp = ctypes.POINTER(ctypes.c_float)
    L.pow_eval.argtypes = [fp, fp, fp, fp, ctypes.c_long]
    L.pow_eval.restype = ctypes.c_int
    return L


def _eval(x, y):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    a, b = np.empty_like(x), np.empty_like(x)
    fp = ctypes.POINTER(ctypes.c_float)
    rc = _lib().pow_eval(x.ctypes.data_as(fp), y.ctypes.data_as(fp), a.ctypes.data_as(fp), b.ctypes.data_as(fp),
                         x.size)
    assert rc == 0
    return a, b


def _ulps(got, x, y):
    exact = np.power(x.astype(np.float64), y.astype(np.float64))
    ulp = np.spacing(np.abs(exact).astype(np.float32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - exact) / ulp


def test_pow_check_library_exports():
    assert hasattr(_lib(), "pow_eval")


@pytest.mark.gpu
@pytest.mark.parametrize("y,lo,hi", [(KAPPA, 0.12, 1.3), (np.float32(1.94), 0.7, 1.25),
                                     (np.float32(1.5), 0.7, 1.25)])
def test_pow_near1(y, lo, hi):
    rng = np.random.default_rng(11)
    x = rng.uniform(lo, hi, 1 << 20).astype(np.float32)
    x[:4] = [1.0, lo, hi, 0.5]
    yy = np.full_like(x, y)
    near1, _ = _eval(x, yy)
    e = _ulps(near1, x, yy)
    print(f"powNear1 y={float(y):.4f} x in [{lo}, {hi}]: max {e.max():.2f} ulp, mean {e.mean():.3f}")
    assert e.max() <= 3.0
    assert near1[0] == 1.0


EXPONENTS = [0.25, 0.8, 6.8, 3.8, 3.25, 3.5, -2.0, -2.75, -2.9, -2.625, -7.0, -0.5, 1.0 / 3.0, 2.0 / 3.0,
             -2.0 / 3.0, 1.0 / 7.0, 0.846, 2.0, 2.0, 4.0, 5.0, 6.0, 0.5, 0.137, 0.999]


@pytest.mark.gpu
@pytest.mark.parametrize("y", [np.float32(v) for v in EXPONENTS])
def test_pow_pos(y):
    rng = np.random.default_rng(int(abs(float(y)) * 1000) + 3)
    x = np.exp(rng.uniform(np.log(1e-30), np.log(1e30), 1 << 20)).astype(np.float32)
    # results in the normal float range
    r = np.power(x.astype(np.float64), float(y))
    x = x[(r > 1e-37) & (r < 1e37)]
    x[:3] = [1.0, 2.0, 0.5]
    yy = np.full_like(x, y)
    _, pos = _eval(x, yy)
    e = _ulps(pos, x, yy)
    bound = 2.5 + 0.4 * abs(float(y))
    print(f"powPos y={float(y):.4f}: max {e.max():.2f} ulp (bound {bound:.1f}), mean {e.mean():.3f}")
    assert e.max() <= bound
    assert pos[0] == 1.0


@pytest.mark.gpu
def test_pow_pos_special_bases():
    """0, subnormal, negative, infinite and NaN bases take the library powf."""
    x = np.array([0.0, 0.0, -8.0, -2.0, np.inf, np.nan, 1e-45, 3.4e38], dtype=np.float32)
    y = np.array([0.5, -1.0, 1.0 / 3.0, 2.0, 0.5, 0.5, 0.5, 0.5], dtype=np.float32)
    _, pos = _eval(x, y)
    assert pos[0] == 0.0 and np.isinf(pos[1]) and np.isnan(pos[2]) and pos[3] == 4.0
    assert np.isinf(pos[4]) and np.isnan(pos[5])
    e = _ulps(pos[6:], x[6:], y[6:])
    assert e.max() <= 2.5, e
