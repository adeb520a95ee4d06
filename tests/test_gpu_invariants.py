"""The reference's GPU integration invariants (src/gpu_tests.rs) on the HIP path."""
from __future__ import annotations

import pytest

from invariants import (GpuBackend, inverse_square_and_spot, lighting_history_and_motion, moving_occluder,
                        svgf_noise_edges_disocclusion)

pytestmark = pytest.mark.gpu

SCHEDULES = ["passes", "fused"]


@pytest.mark.parametrize("schedule", SCHEDULES)
def test_svgf_noise_edges_disocclusion_hdr(require_gpu, schedule):
    print(f"[invariants] svgf noise/edges/disocclusion/HDR schedule={schedule}", flush=True)
    svgf_noise_edges_disocclusion(lambda s, st: GpuBackend(s, st, schedule=schedule))


@pytest.mark.parametrize("schedule", SCHEDULES)
def test_lighting_history_and_motion(require_gpu, schedule):
    print(f"[invariants] lighting history schedule={schedule}", flush=True)
    lighting_history_and_motion(lambda s, st: GpuBackend(s, st, schedule=schedule))


@pytest.mark.parametrize("schedule", SCHEDULES)
def test_inverse_square_and_spot_cone(require_gpu, schedule):
    print(f"[invariants] inverse square / spot schedule={schedule}", flush=True)
    inverse_square_and_spot(lambda s, st: GpuBackend(s, st, schedule=schedule))


@pytest.mark.parametrize("schedule", SCHEDULES)
def test_moving_occluder_blocks_light(require_gpu, schedule):
    print(f"[invariants] moving occluder schedule={schedule}", flush=True)
    moving_occluder(lambda s, st: GpuBackend(s, st, schedule=schedule))
