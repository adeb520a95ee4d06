"""Comparison helpers for GPU-vs-oracle parity tests."""
from __future__ import annotations

import numpy as np

import oracle
from instantradiositytest_amd.renderer import half_to_float

RES_INT = ("id", "source", "count", "age", "generation")
RES_FLT = ("sum", "selected_target", "pdf")


def to_oracle_reservoirs(r: np.ndarray) -> np.ndarray:
    """GPU reservoir records (RESERVOIR_DT) -> oracle records (FP64 fields)."""
    out = np.zeros(r.size, oracle.ORES_DT)
    flat = r.reshape(-1)
    for k in RES_INT:
        out[k] = flat[k]
    for k in RES_FLT:
        out[k] = flat[k].astype(np.float64)
    return out


def compare_reservoirs(gpu: np.ndarray, ref: np.ndarray) -> dict:
    """Fraction of pixels whose integer fields all match, and the worst
    relative error of the float fields over those pixels."""
    g = gpu.reshape(-1)
    same = np.ones(g.size, bool)
    for k in RES_INT:
        same &= g[k] == ref[k]
    rel, worst = 0.0, None
    idx = np.nonzero(same)[0]
    for k in RES_FLT:
        a = g[k][same].astype(np.float64)
        b = ref[k][same]
        fin = np.isfinite(a) & np.isfinite(b)
        d = np.abs(a[fin] - b[fin]) / np.maximum(np.abs(b[fin]), 1e-30)
        if d.size and float(d.max()) > rel:
            rel = float(d.max())
            j = int(idx[fin][np.argmax(d)])
            worst = {"pixel": j, "field": k, "gpu": {f: float(g[f][j]) for f in RES_FLT + RES_INT},
                     "ref": {f: float(ref[f][j]) for f in RES_FLT + RES_INT}}
    return {"match": float(same.mean()), "max_rel": rel, "mismatches": int((~same).sum()), "worst": worst}


def compare_images(gpu: np.ndarray, ref: np.ndarray, tol: float = 1e-3, floor: float = 1e-6) -> dict:
    """Per-pixel relative error (max over rgb, denominator max(|ref|, floor))
    and the image-level relative L2 error."""
    a = gpu[..., :3].astype(np.float64)
    b = ref[..., :3].astype(np.float64)
    per = np.max(np.abs(a - b) / np.maximum(np.abs(b), floor), axis=-1)
    ok = per <= tol
    l2 = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
    return {"within": float(ok.mean()), "rel_l2": l2, "max_rel": float(per.max()), "bad": int((~ok).sum())}


def psnr(a: np.ndarray, b: np.ndarray) -> float:
    a = a[..., :3].astype(np.float64)
    b = b[..., :3].astype(np.float64)
    peak = max(float(np.abs(b).max()), 1e-30)
    mse = float(np.mean((a - b) ** 2))
    return float("inf") if mse == 0 else 10.0 * np.log10(peak * peak / mse)


def rgba16_to_f64(bits: np.ndarray) -> np.ndarray:
    return half_to_float(bits).astype(np.float64)
