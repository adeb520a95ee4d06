"""The reference's GPU integration invariants (src/gpu_tests.rs) on the CPU
restatement — they pin the oracle's end-to-end behaviour before it is used
as the checker of the HIP path."""
from __future__ import annotations

import pytest

from invariants import (OracleBackend, inverse_square_and_spot, lighting_history_and_motion, moving_occluder,
                        svgf_noise_edges_disocclusion)


def backend(scene, settings):
    return OracleBackend(scene, settings, bits=32)


def test_svgf_noise_edges_disocclusion_hdr():
    svgf_noise_edges_disocclusion(backend)


def test_lighting_history_and_motion():
    lighting_history_and_motion(backend)


def test_inverse_square_and_spot_cone():
    inverse_square_and_spot(backend)


def test_moving_occluder_blocks_light():
    moving_occluder(backend)


@pytest.mark.parametrize("test", [inverse_square_and_spot, moving_occluder])
def test_fp64_instance_agrees(test):
    test(lambda s, st: OracleBackend(s, st, bits=64))
