"""The kernels' shortcuts for correctly rounded division and square root
against the compiler's correctly rounded sequences, on the GPU.

The lighting kernels (common.wgsl:67-83: brdf, direct, indirect) compute
1/sqrt(r2) with a shortened square root and a refined hardware reciprocal,
the diffuse term's division by PI with one Markstein remainder step, and the
half vector's normalisation, the specular and the inverse-square divisions
with the ranged quotient (an exact fallback takes operands outside its
domain).  The FP32 profile's parity with the FP32 restatement
(test_gpu_parity.py) needs their results bit for bit equal to the correctly
rounded ones: the device check compares every binary32 input of the
one-operand shortcuts' domains, and every numerator of the ranged quotient
over a divisor drawn across its domain."""
from __future__ import annotations

import ctypes

import pytest

from instantradiositytest_amd import _lib

pytestmark = pytest.mark.gpu


def test_shortcuts_equal_the_division_sequence_for_every_input(require_gpu):
    L = _lib.irt()
    out = (ctypes.c_uint64 * 4)()
    rc = L.irt_check_fast_division(out)
    assert rc == 0, L.irt_global_error()
    assert list(out) == [0, 0, 0, 0], (f"reciprocal mismatches {out[0]}, division-by-pi mismatches {out[1]}, "
                                       f"square root mismatches {out[2]}, ranged quotient mismatches {out[3]}")
