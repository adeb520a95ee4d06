"""Writes pcg_vectors.json: the per-pixel random stream of the path.

The stream is fixed by integer arithmetic alone (common.wgsl:39-44):
  state = state * 747796405 + 2891336453            (mod 2^32)
  w     = ((state >> ((state >> 28) + 4)) ^ state) * 277803737   (mod 2^32)
  value = f32((w >> 22) ^ w) * 2^-32
  seed(i, salt) = i*1973 + frame*9277 + salt*26699 + 89173       (mod 2^32)
This script evaluates it with Python integers and exact rational rounding
to binary32, independently of the C/HIP code under test.
"""
import json
from fractions import Fraction
from pathlib import Path

M = 1 << 32


def f32_of_u32(v: int) -> float:
    """Round an integer < 2^32 to the nearest binary32 (ties to even), exactly."""
    if v == 0:
        return 0.0
    e = v.bit_length() - 24
    if e <= 0:
        return float(v)
    q, r = divmod(v, 1 << e)
    half = 1 << (e - 1)
    if r > half or (r == half and q & 1):
        q += 1
    return float(q << e)


def draw(state: int):
    state = (state * 747796405 + 2891336453) % M
    w = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) % M
    return state, f32_of_u32((w >> 22) ^ w) / 4294967296.0


def seed(i, salt, frame):
    return (i * 1973 + frame * 9277 + salt * 26699 + 89173) % M


def main():
    cases = []
    for (i, salt, frame) in [(0, 11, 0), (1, 13, 0), (12345, 17, 3), (2303999, 19, 127), (767, 23, 64),
                             (4294967295, 29, 4294967295)]:
        s = seed(i, salt, frame)
        out = []
        st = s
        for _ in range(16):
            st, v = draw(st)
            out.append(v)
        cases.append({"i": i, "salt": salt, "frame": frame, "seed": s, "values": out, "final_state": st})
    Path(__file__).with_name("pcg_vectors.json").write_text(json.dumps(cases, indent=1))


if __name__ == "__main__":
    main()
