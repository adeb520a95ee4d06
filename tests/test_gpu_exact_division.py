"""The kernels' shortcuts for correctly rounded division against the division
sequence itself, on the GPU, for every binary32 input of their domains.

The lighting kernels compute 1/sqrt(r2) (common.wgsl:75-83, inverseSqrt) with
a refined hardware reciprocal and the BRDF's diffuse term
(common.wgsl:67-74, ... / PI) with one Markstein remainder step.  The FP32
profile's parity with the FP32 restatement (test_gpu_parity.py) needs their
results bit for bit equal to the correctly rounded quotient; this test
checks all 2^32 inputs, so it holds for every input the kernels can meet,
not only the scenes of the parity tests."""
from __future__ import annotations

import ctypes

import pytest

from instantradiositytest_amd import _lib

pytestmark = pytest.mark.gpu


def test_shortcuts_equal_the_division_sequence_for_every_input(require_gpu):
    L = _lib.irt()
    out = (ctypes.c_uint64 * 2)()
    rc = L.irt_check_fast_division(out)
    assert rc == 0, L.irt_global_error()
    assert list(out) == [0, 0], f"reciprocal mismatches {out[0]}, division-by-pi mismatches {out[1]}"
