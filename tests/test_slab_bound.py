"""The traversal's ray-space slab test (csrc/tr_device.h, trace4_step) evaluates each quantized
plane as one fused multiply-add, q * (step * inv) + (origin - o) * inv, and widens the entry and exit
distances of each axis by e = 3e-7 ((255 step + |origin|) |inv| + |b|). This checks, on the CPU with
float32 arithmetic, that the widened interval always contains the exact distances of the planes as
the host rounds them (float32(origin + q * step)), so the box test stays conservative; and that the
bound is not loose by orders of magnitude (the widening stays a few float32 ulps of the slab's
distances scale)."""
import numpy as np

F = np.float32


def _fma32(a, b, c):
    # a * b is exact in float64 for float32 operands; the sum rounds once in float64, then to float32
    # (a double rounding: at most one float32 ulp away from a true fused multiply-add, which the
    # check below allows for)
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(F)


def _cases(n, rng):
    # node origins and power-of-two steps as the quantizer makes them, rays anywhere around
    origin = rng.uniform(-2.0, 2.0, n).astype(F)
    step = np.ldexp(F(1.0), rng.integers(-20, -2, n)).astype(F)
    q = rng.integers(0, 256, n).astype(F)
    o = (origin + rng.uniform(-1.0, 1.0, n) * rng.choice([1e-6, 1e-3, 1.0], n)).astype(F)
    d = rng.normal(size=n) * rng.choice([1.0, 1e-3, 1e-7, 0.0], n, p=[0.7, 0.1, 0.1, 0.1])
    d = d.astype(F)
    safe = np.where(np.abs(d) < F(1e-20), np.copysign(F(1e-20), d), d).astype(F)
    inv = (F(1.0) / safe).astype(F)
    return origin, step, q, o, inv


def test_widened_slab_contains_exact_plane_distance():
    rng = np.random.default_rng(5)
    origin, step, q, o, inv = _cases(400_000, rng)
    plane = (origin.astype(np.float64) + q.astype(np.float64) * step.astype(np.float64)).astype(F)  # host rounding
    exact = (plane.astype(np.float64) - o.astype(np.float64)) * inv.astype(np.float64)
    a = (step * inv).astype(F)
    b = ((origin - o).astype(F) * inv).astype(F)
    e = (F(3e-7) * _fma32(_fma32(np.full_like(step, 255.0), step, np.abs(origin)), np.abs(inv), np.abs(b))).astype(F)
    entry = _fma32(q, a, (b - e).astype(F))
    exit_ = _fma32(q, a, (b + e).astype(F))
    # one float32 ulp of slack for the emulated fused multiply-add's double rounding
    ulp = np.spacing(np.abs(entry)).astype(np.float64)
    assert (entry.astype(np.float64) - ulp <= exact).all()
    ulp = np.spacing(np.abs(exit_)).astype(np.float64)
    assert (exit_.astype(np.float64) + ulp >= exact).all()
    # the widening is a small multiple of the rounding scale of the slab's terms
    scale = (255.0 * step.astype(np.float64) + np.abs(origin)) * np.abs(inv) + np.abs(b)
    assert (e.astype(np.float64) <= 3.1e-7 * scale).all()


def test_a_dropped_widening_fails_the_check():
    """The same check on the unwidened distances finds plane distances outside the interval, so the
    check above can fail (the widening is needed)."""
    rng = np.random.default_rng(6)
    origin, step, q, o, inv = _cases(400_000, rng)
    plane = (origin.astype(np.float64) + q.astype(np.float64) * step.astype(np.float64)).astype(F)
    exact = (plane.astype(np.float64) - o.astype(np.float64)) * inv.astype(np.float64)
    a = (step * inv).astype(F)
    b = ((origin - o).astype(F) * inv).astype(F)
    t = _fma32(q, a, b).astype(np.float64)
    # unwidened, an entry distance can lie beyond the exact one (a box missed) and an exit before it
    assert (t > exact).sum() > 100 and (t < exact).sum() > 100
