"""The reference-mode collision divides by tau with IEEE float division;
the GPU computes q = x * RN64(1/tau) in double and rounds q to float (IEEE
division only below the float normal range).  This pins the claim that the
two agree bit for bit on normal results: random and adversarial x over
the whole exponent range, the relaxation times the tests and benchmarks
use and awkward ones."""
import numpy as np
import pytest

FLT_MIN = np.float32(1.1754943508222875e-38)


def gpu_rule(x: np.ndarray, tau: np.float32) -> np.ndarray:
    rinv = 1.0 / np.float64(tau)
    q = x.astype(np.float64) * rinv
    out = q.astype(np.float32)
    tiny = np.abs(q) < np.float64(FLT_MIN)
    out[tiny] = x[tiny] / tau
    return out


@pytest.mark.parametrize("tau", [0.6, 0.7, 1.0, 0.51, 0.8, 1.7, 0.5000001,
                                 1.9999999, 0.9, 0.55, 3.14159])
def test_double_product_rounds_like_float_division(tau):
    tau = np.float32(tau)
    rng = np.random.default_rng(int(tau * 1e6))
    # random significands over every exponent a population difference can take
    bits = rng.integers(0, 1 << 31, size=2_000_000, dtype=np.uint64)
    x = bits.astype(np.uint32).view(np.float32)
    x = x[np.isfinite(x)]
    # values near the quotients' rounding boundaries: x = RN32(m * tau) for
    # m halfway between floats
    m = rng.random(500_000).astype(np.float32)
    mid = (m.astype(np.float64) + np.spacing(m).astype(np.float64) / 2)
    near = (mid * np.float64(tau)).astype(np.float32)
    x = np.concatenate([x, near, -near, np.float32([0.0, -0.0, 1e-45, 3e38])])
    with np.errstate(over="ignore", under="ignore", invalid="ignore"):
        ref = x / tau
        got = gpu_rule(x, tau)
    same = (ref.view(np.uint32) == got.view(np.uint32)) | (np.isnan(ref) & np.isnan(got))
    assert same.all(), (x[~same][:5], ref[~same][:5], got[~same][:5])
