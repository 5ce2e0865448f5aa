"""The exact arithmetic shortcuts of the pair kernels against the IEEE operations, on the
GPU through the C ABI (fe_check_exact_arith): the corrected square root for every float in
[2^-96, 2^16), the boxed division for 2^26 pseudo-random operand pairs spanning
[2^-40, 2^40) in both signs, and r / h by the residual-corrected product for every float
r in [1e-6, 1.001 h] at the smoothing radii of the test and bench scenes. Every count of
mismatches must be zero. (The kernels also check r / h per scene before using it.)"""
import ctypes

import pytest

from fluid_engine import _ffi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("h", [0.04, 0.08, 0.0123, 1.0])
def test_shortcuts_equal_ieee(h):
    lib = _ffi.load_library()
    counts = (ctypes.c_uint64 * 6)()
    _ffi.check(lib.fe_check_exact_arith(0, ctypes.c_float(h), 1 << 26, counts), "fe_check_exact_arith")
    sq, sq_bad, dv, dv_bad, rh, rh_bad = list(counts)
    print(f"h={h}: sqrt {sq} checked, division {dv}, r/h {rh}")
    assert sq == 112 << 23 and dv == 1 << 26 and rh > 100_000_000
    assert (sq_bad, dv_bad, rh_bad) == (0, 0, 0)


def test_rejects_bad_arguments():
    lib = _ffi.load_library()
    counts = (ctypes.c_uint64 * 6)()
    assert lib.fe_check_exact_arith(0, ctypes.c_float(0.0), 0, counts) != 0
    assert lib.fe_check_exact_arith(999, ctypes.c_float(0.04), 0, counts) != 0
