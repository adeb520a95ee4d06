"""Captures and image comparison for parity reports.

``write_png`` is the PNG side of ``Renderer::capture`` (src/renderer.rs:2109-2163:
the RGBA8 display target saved as an 8-bit RGBA PNG).  The analysis follows
the reference's image_analysis_tools/render_debug, the checks that
tools/verify.py:61-80 runs on captures:

  - ``select_uv_rect``: a mask of the pixels whose centres ((x + 0.5) / W,
    (y + 0.5) / H) lie in [u_min, u_max] x [v_min, v_max], bounds included
    (the rule of selection.py:149-164);
  - ``occupancy``: how many selected pixels satisfy a predicate over the
    normalised channels R, G, B, A (0..1), written with comparisons, AND, OR
    and parentheses, as in ``R < 0.03 AND G < 0.03 AND B < 0.03``
    (analysis.py:302-310, predicate.py);
  - ``match``: the largest absolute difference of corresponding normalised
    channel samples against a tolerance (analysis.py:407-415);
  - ``psnr``: peak signal-to-noise ratio of the normalised RGB channels.

Command line (a subset of render_debug's)::

    python -m instantradiositytest_amd.images --select uv-rect 0.1 0.05 0.28 0.4 \\
        --evaluate occupancy --condition "R < 0.03 AND G < 0.03 AND B < 0.03" --images render.png
    python -m instantradiositytest_amd.images --select all --evaluate match --tolerance 0 \\
        --images repeat_a.png repeat_b.png
"""
from __future__ import annotations

import argparse
import json
import math
import struct
import sys
import zlib
from pathlib import Path

import numpy as np

_PNG_SIGNATURE = b"\x89PNG\r\n\x1a\n"


# ------------------------------------------------------------------- PNG I/O
def _chunk(kind: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def write_png(path, rgba: np.ndarray) -> None:
    """8-bit RGBA PNG of an [H, W, 4] uint8 image (filter type 0 on every row)."""
    rgba = np.ascontiguousarray(rgba, np.uint8)
    if rgba.ndim != 3 or rgba.shape[2] != 4:
        raise ValueError("expected an [H, W, 4] uint8 image")
    h, w, _ = rgba.shape
    raw = np.concatenate([np.zeros((h, 1), np.uint8), rgba.reshape(h, w * 4)], axis=1)
    png = _PNG_SIGNATURE + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0))
    png += _chunk(b"IDAT", zlib.compress(raw.tobytes(), 6)) + _chunk(b"IEND", b"")
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_bytes(png)


def _paeth(a: int, b: int, c: int) -> int:
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def read_png(path) -> np.ndarray:
    """An 8-bit grey, grey+alpha, RGB or RGBA non-interlaced PNG as [H, W, 4]
    uint8 RGBA (grey expands to RGB, a missing alpha is 255)."""
    data = Path(path).read_bytes()
    if data[:8] != _PNG_SIGNATURE:
        raise ValueError("not a PNG file")
    pos, idat, header = 8, [], None
    while pos < len(data):
        (length,) = struct.unpack(">I", data[pos:pos + 4])
        kind = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + length]
        pos += 12 + length
        if kind == b"IHDR":
            header = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat.append(body)
        elif kind == b"IEND":
            break
    if header is None:
        raise ValueError("PNG without IHDR")
    w, h, depth, ctype, _, _, interlace = header
    channels = {0: 1, 2: 3, 4: 2, 6: 4}.get(ctype)
    if depth != 8 or channels is None or interlace != 0:
        raise ValueError("only 8-bit, non-interlaced grey / grey+alpha / RGB / RGBA PNGs are supported")
    raw = zlib.decompress(b"".join(idat))
    stride = w * channels
    out = np.zeros((h, stride), np.uint8)
    prev = np.zeros(stride, np.int64)
    for y in range(h):
        ftype = raw[y * (stride + 1)]
        line = np.frombuffer(raw, np.uint8, stride, y * (stride + 1) + 1).astype(np.int64)
        if ftype == 0:
            cur = line
        elif ftype == 2:
            cur = (line + prev) & 0xFF
        else:
            cur = np.zeros(stride, np.int64)
            for i in range(stride):
                a = cur[i - channels] if i >= channels else 0
                b = prev[i]
                c = prev[i - channels] if i >= channels else 0
                if ftype == 1:
                    pred = a
                elif ftype == 3:
                    pred = (a + b) >> 1
                elif ftype == 4:
                    pred = _paeth(a, b, c)
                else:
                    raise ValueError(f"unknown PNG filter {ftype}")
                cur[i] = (line[i] + pred) & 0xFF
        out[y] = cur
        prev = cur
    px = out.reshape(h, w, channels)
    if channels == 4:
        return px.copy()
    rgba = np.full((h, w, 4), 255, np.uint8)
    if channels == 3:
        rgba[..., :3] = px
    elif channels == 1:
        rgba[..., :3] = px
    else:
        rgba[..., :3] = px[..., :1]
        rgba[..., 3] = px[..., 1]
    return rgba


# ----------------------------------------------------------------- selection
# A selection is a boolean mask over the image's pixels, [H, W].  The uv
# rectangle keeps the pixels whose centres ((x + 0.5) / W, (y + 0.5) / H)
# lie inside it, edges included; the mask is the outer product of a column
# test and a row test on those centres.

def select_all(width: int, height: int) -> np.ndarray:
    return np.ones((height, width), bool)


def select_uv_rect(width: int, height: int, u_min: float, v_min: float, u_max: float, v_max: float) -> np.ndarray:
    """Mask of the pixels whose centres lie in [u_min, u_max] x [v_min, v_max]."""
    box = np.array([u_min, v_min, u_max, v_max], np.float64)
    if not np.all(np.isfinite(box)) or box.min() < 0.0 or box.max() > 1.0:
        raise ValueError(f"uv-rect corners must be finite numbers in [0, 1], got {box.tolist()}")
    if box[2] < box[0] or box[3] < box[1]:
        raise ValueError(f"uv-rect is inverted: ({u_min}, {v_min}) lies beyond ({u_max}, {v_max})")
    cu = (np.arange(width) + 0.5) / width
    cv = (np.arange(height) + 0.5) / height
    cols = (cu >= box[0]) & (cu <= box[2])
    rows = (cv >= box[1]) & (cv <= box[3])
    return rows[:, None] & cols[None, :]


def selected(image: np.ndarray, mask: np.ndarray) -> np.ndarray:
    """Normalised float64 RGBA of the selected pixels in row-major order, [N, 4]."""
    if mask.shape != image.shape[:2]:
        raise ValueError(f"selection of shape {mask.shape} does not fit a {image.shape[1]}x{image.shape[0]} image")
    return image[mask].astype(np.float64) / 255.0


# ----------------------------------------------------------------- predicate
# Grammar: a condition is comparisons "<channel> <op> <number>" (or
# "<number> <op> <channel>") over the normalised channels R, G, B, A, joined
# by AND and OR, with AND binding tighter, and parentheses.  Channel names
# and the words AND / OR are case-insensitive.
#
# The scanner below classifies characters and emits typed tokens; the
# compiler turns them into reverse Polish notation with Dijkstra's
# shunting-yard method, where a comparison is one operand; evaluation runs
# the RPN with a stack of boolean arrays.

_CHANNELS = "RGBA"
_COMPARE = {"<": np.less, "<=": np.less_equal, ">": np.greater, ">=": np.greater_equal, "==": np.equal,
            "!=": np.not_equal}
_MIRROR = {"<": ">", "<=": ">=", ">": "<", ">=": "<=", "==": "==", "!=": "!="}
_PRECEDENCE = {"OR": 1, "AND": 2}


def _scan(text: str) -> list[tuple[str, object, int]]:
    """Tokens (kind, value, column): kinds 'chan', 'cmp', 'num', 'bool', '(' and ')'."""
    toks: list[tuple[str, object, int]] = []
    i, n = 0, len(text)
    while i < n:
        c = text[i]
        if c.isspace():
            i += 1
        elif c in "()":
            toks.append((c, c, i))
            i += 1
        elif c in "<>=!":
            two = text[i:i + 2]
            if two in _COMPARE:
                toks.append(("cmp", two, i))
                i += 2
            elif c in "<>":
                toks.append(("cmp", c, i))
                i += 1
            else:
                raise ValueError(f"column {i}: '{c}' must be followed by '=' in {text!r}")
        elif c.isdigit() or c in "+-.":
            j = i + 1 if c in "+-" else i
            while j < n and (text[j].isdigit() or text[j] == "."):
                j += 1
            if j < n and text[j] in "eE":
                k = j + 1
                if k < n and text[k] in "+-":
                    k += 1
                if k < n and text[k].isdigit():
                    j = k
                    while j < n and text[j].isdigit():
                        j += 1
            try:
                value = float(text[i:j])
            except ValueError:
                raise ValueError(f"column {i}: malformed number {text[i:j]!r} in {text!r}") from None
            toks.append(("num", value, i))
            i = j
        elif c.isalpha():
            j = i
            while j < n and (text[j].isalnum() or text[j] == "_"):
                j += 1
            word = text[i:j].upper()
            if word in _PRECEDENCE:
                toks.append(("bool", word, i))
            elif len(word) == 1 and word in _CHANNELS:
                toks.append(("chan", _CHANNELS.index(word), i))
            else:
                raise ValueError(f"column {i}: unknown word {text[i:j]!r} (channels are R, G, B, A) in {text!r}")
            i = j
        else:
            raise ValueError(f"column {i}: unexpected character {c!r} in {text!r}")
    return toks


def _compile(text: str) -> list[tuple]:
    """RPN program: ('test', channel, op, value) operands and ('AND',) / ('OR',) operators."""
    toks = _scan(text)
    if not toks:
        raise ValueError("empty condition")
    out: list[tuple] = []
    ops: list[tuple[str, int]] = []  # pending 'AND' / 'OR' / '(' with columns
    expect_operand = True
    i = 0
    while i < len(toks):
        kind, value, col = toks[i]
        if expect_operand:
            if kind == "(":
                ops.append(("(", col))
                i += 1
                continue
            atom = toks[i:i + 3]
            kinds = [t[0] for t in atom]
            if kinds == ["chan", "cmp", "num"]:
                out.append(("test", atom[0][1], atom[1][1], atom[2][1]))
            elif kinds == ["num", "cmp", "chan"]:
                out.append(("test", atom[2][1], _MIRROR[atom[1][1]], atom[0][1]))
            else:
                raise ValueError(f"column {col}: expected a comparison like 'R < 0.5' in {text!r}")
            i += 3
            expect_operand = False
        else:
            if kind == "bool":
                while ops and ops[-1][0] != "(" and _PRECEDENCE[ops[-1][0]] >= _PRECEDENCE[value]:
                    out.append((ops.pop()[0],))
                ops.append((value, col))
                expect_operand = True
            elif kind == ")":
                while ops and ops[-1][0] != "(":
                    out.append((ops.pop()[0],))
                if not ops:
                    raise ValueError(f"column {col}: ')' without a matching '(' in {text!r}")
                ops.pop()
            else:
                raise ValueError(f"column {col}: expected AND, OR or ')' in {text!r}")
            i += 1
    if expect_operand:
        raise ValueError(f"condition ends where a comparison is expected: {text!r}")
    while ops:
        op, col = ops.pop()
        if op == "(":
            raise ValueError(f"column {col}: '(' is never closed in {text!r}")
        out.append((op,))
    return out


def predicate_mask(condition: str, px: np.ndarray) -> np.ndarray:
    """Boolean mask over [N, 4] normalised pixels for the condition."""
    stack: list[np.ndarray] = []
    for instr in _compile(condition):
        if instr[0] == "test":
            _, ch, op, value = instr
            stack.append(_COMPARE[op](px[:, ch], value))
        else:
            b, a = stack.pop(), stack.pop()
            stack.append(a & b if instr[0] == "AND" else a | b)
    (mask,) = stack
    return np.asarray(mask, bool)


# ---------------------------------------------------------------- evaluations
def occupancy(image: np.ndarray, mask: np.ndarray, condition: str) -> dict:
    """How many selected pixels satisfy the condition."""
    px = selected(image, mask)
    hits = int(np.count_nonzero(predicate_mask(condition, px)))
    total = int(px.shape[0])
    share = hits / total if total else 0.0
    return {"condition": condition, "count": hits, "total": total, "fraction": share, "percentage": 100.0 * share}


def match(reference: np.ndarray, actual: np.ndarray, mask: np.ndarray, tolerance: float) -> dict:
    """Largest absolute difference of the selected normalised channel values against a tolerance."""
    if reference.shape != actual.shape:
        raise ValueError(f"image sizes differ: {reference.shape} and {actual.shape}")
    if not (0.0 <= tolerance <= 1.0):  # also refuses NaN
        raise ValueError(f"tolerance {tolerance} is not a number in [0, 1]")
    a, b = selected(reference, mask), selected(actual, mask)
    worst = float(np.max(np.abs(a - b))) if a.size else 0.0
    return {"match": worst <= tolerance, "tolerance": tolerance, "max_error": worst,
            "metric": "largest |reference - actual| over the selected RGBA values, each scaled to [0, 1]"}


def psnr(reference: np.ndarray, actual: np.ndarray, mask: np.ndarray | None = None) -> dict:
    """PSNR of the normalised RGB channels (peak 1); infinite for identical images."""
    if reference.shape != actual.shape:
        raise ValueError(f"image sizes differ: {reference.shape} and {actual.shape}")
    if mask is None:
        mask = select_all(reference.shape[1], reference.shape[0])
    err = selected(reference, mask)[:, :3] - selected(actual, mask)[:, :3]
    mse = float(np.mean(err * err))
    return {"psnr_db": math.inf if mse == 0 else -10.0 * math.log10(mse), "mse": mse}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--select", nargs="+", default=["all"], help="all | uv-rect u_min v_min u_max v_max")
    ap.add_argument("--evaluate", required=True, choices=["occupancy", "match", "psnr"])
    ap.add_argument("--condition")
    ap.add_argument("--tolerance", type=float)
    ap.add_argument("--images", nargs="+", required=True)
    args = ap.parse_args(argv)
    images = [read_png(p) for p in args.images]
    h, w, _ = images[0].shape
    if args.select[0] == "all":
        rect = select_all(w, h)
    elif args.select[0] == "uv-rect" and len(args.select) == 5:
        rect = select_uv_rect(w, h, *(float(v) for v in args.select[1:]))
    else:
        ap.error("--select all | uv-rect u_min v_min u_max v_max")
    if args.evaluate == "occupancy":
        if args.condition is None or len(images) != 1:
            ap.error("occupancy needs --condition and one image")
        result = occupancy(images[0], rect, args.condition)
    else:
        if len(images) != 2:
            ap.error(f"{args.evaluate} needs two images")
        if args.evaluate == "match":
            if args.tolerance is None:
                ap.error("--tolerance is required")
            result = match(images[0], images[1], rect, args.tolerance)
        else:
            result = psnr(images[0], images[1], rect)
    print(json.dumps({"result": result}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
