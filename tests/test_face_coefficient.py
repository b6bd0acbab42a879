"""The pressure operator's face coefficient with the short division
(skt::faceOfSum in skewtcloudsim_amd/csrc/skt_device.h) against the
reference's IEEE expression -max(1, 1.225 / (2 / s)) (Game/kernelSky.cu:1372-1377)
for every one of the 2^32 float inputs s: the two must agree bit for bit
(NaN against NaN). The SpMV uses this function, so this is what makes the
matrix-free operator equal to the reference's.
"""
import ctypes
import os

import pytest

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "libfacecheck.so")


def _lib():
    if not os.path.exists(LIB):
        pytest.fail("tests/native/libfacecheck.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB)
    L.face_check.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(ctypes.c_ulonglong),
                             ctypes.POINTER(ctypes.c_uint32)]
    L.face_check.restype = ctypes.c_int
    return L


def test_face_check_library_exports():
    assert hasattr(_lib(), "face_check")


@pytest.mark.gpu
def test_face_coefficient_all_inputs():
    L = _lib()
    bad, first = ctypes.c_ulonglong(0), ctypes.c_uint32(0)
    assert L.face_check(0, 1 << 32, ctypes.byref(bad), ctypes.byref(first)) == 0
    assert bad.value == 0, f"{bad.value} inputs differ, first bit pattern {first.value:#010x}"



def _rlib():
    L = _lib()
    L.recip_check.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_ulonglong),
                              ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_ulonglong)]
    L.recip_check.restype = ctypes.c_int
    return L


def test_recip_check_library_exports():
    assert hasattr(_lib(), "recip_check")


@pytest.mark.gpu
def test_fast_reciprocal_all_inputs():
    """skt::recipRNfast (hardware reciprocal, one FMA refinement), which the
    preconditioner, divergence and projection passes take for 1/density and
    the harmonic face densities, equals the IEEE 1.0f / s bit for bit for
    every float s with |s| in [2^-126, 2^126]."""
    L = _rlib()
    bad, first, skipped = ctypes.c_ulonglong(0), ctypes.c_uint32(0), ctypes.c_ulonglong(0)
    assert L.recip_check(0, 1 << 32, ctypes.byref(bad), ctypes.byref(first), ctypes.byref(skipped)) == 0
    assert bad.value == 0, f"{bad.value} inputs differ, first bit pattern {first.value:#010x}"
    # outside the range, per sign: zero and the subnormals (exponent field 0),
    # (2^126, 2^127) (field 253 with a nonzero mantissa), field 254, and the
    # infinities and NaNs (field 255)
    assert skipped.value == 2 * (4 * (1 << 23) - 1), skipped.value
