"""Adversarial (d2, r) float32 pairs for the square-root-free contact test
sqrt_lt (blobs_amd/csrc/common.hpp): radii across the exponent range (powers
of two, subnormals, the largest finite, infinity, zero, negative, NaN) and
squared distances within a few ulps of r*r and of the rounding midpoint
((pred(r) + r) / 2)^2, plus random pairs."""
import numpy as np


def cases(n_random: int = 200_000, seed: int = 7):
    rng = np.random.default_rng(seed)
    f32 = np.float32
    r = np.concatenate([
        (2.0 ** rng.uniform(-140, 60, 20_000)).astype(f32),
        (2.0 ** np.arange(-149, 64)).astype(f32),
        rng.uniform(0.05, 2.0, 20_000).astype(f32),
        np.array([np.finfo(f32).tiny, np.finfo(f32).max, 1.0, 3.0, 0.1, 0.5], f32),
        np.arange(1, 200, dtype=np.uint32).view(f32),  # subnormals
    ])
    d2, rr = [], []
    for base in (r.astype(np.float64) ** 2,
                 ((r.view(np.uint32) - 1).view(f32).astype(np.float64) + r.astype(np.float64)) ** 2 / 4.0):
        with np.errstate(over="ignore"):
            b = base.astype(f32)
        for k in range(-3, 4):
            bits = b.view(np.int32).astype(np.int64) + k
            ok = (bits >= 0) & (bits < 0x7F800000)
            d2.append(bits[ok].astype(np.int32).view(f32))
            rr.append(r[ok])
    d2.append(rng.uniform(0, 10, n_random).astype(f32))
    rr.append(rng.uniform(0, 4, n_random).astype(f32))
    special_d2 = np.array([0.0, -0.0, np.inf, np.nan, 1.0, 4.0, np.finfo(f32).max], f32)
    special_r = np.array([0.0, -1.0, np.inf, np.nan, 1.0, 2.0, np.finfo(f32).tiny], f32)
    d2.append(np.repeat(special_d2, special_r.size))
    rr.append(np.tile(special_r, special_d2.size))
    return np.concatenate(d2).astype(f32), np.concatenate(rr).astype(f32)


def reference(d2, r):
    """sqrtf(d2) < r with numpy's correctly rounded float32 square root."""
    with np.errstate(invalid="ignore"):
        return np.sqrt(d2.astype(np.float32)) < r
