"""Host model of the block-scaled FP8 (OCP e4m3) operand planes of the FP8 trailing update,
used by the GPU tests to evaluate the same product in FP64.

Each row of A (each column of B) is cut into blocks of 32 consecutive K elements; a block with
largest magnitude m * 2^E (m in [0.5, 1)) gets the power-of-two scale 2^(E - 8); plane 0 holds
the scaled values rounded to nearest even e4m3 (3 mantissa bits, subnormals down to 2^-9), plane
p the same of the residual left by planes 0..p-1, each with its own block scales.
"""
import numpy as np


def e4m3_rne(v):
    """Round float64 values with |v| <= 448 to the nearest OCP e4m3 value (ties to even)."""
    v = np.asarray(v, dtype=np.float64)
    a = np.abs(v)
    out = np.zeros_like(a)
    nz = a > 0
    e = np.floor(np.log2(a[nz]))
    e = np.maximum(e, -6.0)  # below 2^-6 the spacing stays 2^-9 (subnormals)
    step = np.exp2(e - 3.0)
    out[nz] = np.round(a[nz] / step) * step
    return np.copysign(out, v)


def planes_along(x, planes, axis):
    """x (float32) split into `planes` dequantised e4m3 planes with one scale per 32 elements
    along `axis` (the K axis); returns float64 arrays of x's shape."""
    x = np.asarray(x, dtype=np.float32)
    xk = np.moveaxis(x, axis, -1)
    k = xk.shape[-1]
    kp = (k + 31) // 32 * 32
    r = np.zeros(xk.shape[:-1] + (kp,), dtype=np.float32)
    r[..., :k] = xk
    blocks = r.reshape(r.shape[:-1] + (kp // 32, 32))
    out = []
    for _ in range(planes):
        amax = np.abs(blocks).max(axis=-1, keepdims=True).astype(np.float64)
        _, E = np.frexp(amax)
        ex = np.where(amax > 0, E - 8, -127).astype(np.float64)
        q = e4m3_rne(blocks.astype(np.float64) * np.exp2(-ex))
        deq = q * np.exp2(ex)
        out.append(np.moveaxis(deq.reshape(r.shape)[..., :k], -1, axis))
        blocks = (blocks.astype(np.float64) - deq).astype(np.float32)
    return out


PAIRS = {1: [(0, 0)], 2: [(1, 0), (0, 1), (0, 0)], 3: [(2, 0), (0, 2), (1, 1), (1, 0), (0, 1), (0, 0)]}


def fp8_gemm_sub(a, b, c, planes):
    """C - sum over plane pairs (i, j), i + j < planes, of A_i B_j, in FP64, and the elementwise
    bound C_abs + sum |A_i| |B_j| used for the FP32-accumulation tolerance."""
    pa, pb = planes_along(a, planes, 1), planes_along(b, planes, 0)
    ref = np.asarray(c, dtype=np.float64) - sum(pa[i] @ pb[j] for i, j in PAIRS[planes])
    mag = np.abs(np.asarray(c, dtype=np.float64)) + sum(np.abs(pa[i]) @ np.abs(pb[j]) for i, j in PAIRS[planes])
    return ref, mag
