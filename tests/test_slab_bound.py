"""The traversal's ray-space slab test (csrc/tr_device.h, trace4_step / trace4_packet, offsets and
margin from ray_pre) evaluates each quantized plane as two fused multiply-adds: b = fma(origin, inv, c)
with the ray's c = -(o * inv), then fma(q, step * inv, b -+ k |inv|), where b -+ k |inv| is itself a
fused multiply-add and k = 6e-7 (C + max |o|), C bounding |origin| + 255 step over all nodes
(qbvh_plane_bound, rounded up). This checks, on the CPU with float32 arithmetic, that the widened
interval always contains the exact distances of the planes as the host rounds them
(float32(origin + q * step)), so the box test stays conservative, with the tightest valid C (the
node's own |origin| + 255 step) and the ray's own |o| (the smallest margin the kernels can use); and
that the widening stays a small multiple of the slab's rounding scale."""
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


def _margin(origin, step, o):
    # the device's k = 6e-7f * (plane_bound + max|o|), with the smallest valid plane bound
    C = np.nextafter((np.abs(origin).astype(np.float64) + 255.0 * step.astype(np.float64)).astype(F), F(np.inf))
    return (F(6e-7) * (C + np.abs(o)).astype(F)).astype(F), C


def test_widened_slab_contains_exact_plane_distance():
    rng = np.random.default_rng(5)
    origin, step, q, o, inv = _cases(400_000, rng)
    plane = (origin.astype(np.float64) + q.astype(np.float64) * step.astype(np.float64)).astype(F)  # host rounding
    exact = (plane.astype(np.float64) - o.astype(np.float64)) * inv.astype(np.float64)
    a = (step * inv).astype(F)
    c = (-(o * inv).astype(F)).astype(F)
    k, C = _margin(origin, step, o)
    b = _fma32(origin, inv, c)
    ai = np.abs(inv).astype(F)
    entry = _fma32(q, a, _fma32(-k, ai, b))
    exit_ = _fma32(q, a, _fma32(k, ai, b))
    # one float32 ulp of slack for the emulated fused multiply-add's double rounding
    ulp = np.spacing(np.abs(entry)).astype(np.float64)
    assert (entry.astype(np.float64) - ulp <= exact).all()
    ulp = np.spacing(np.abs(exit_)).astype(np.float64)
    assert (exit_.astype(np.float64) + ulp >= exact).all()
    # the widening is a small multiple of the rounding scale of the slab's terms
    scale = (C.astype(np.float64) + np.abs(o)) * np.abs(inv)
    assert (k.astype(np.float64) * np.abs(inv) <= 6.1e-7 * scale).all()


def test_a_dropped_widening_fails_the_check():
    """The same check on the unwidened distances finds plane distances outside the interval, so the
    check above can fail (the widening is needed)."""
    rng = np.random.default_rng(6)
    origin, step, q, o, inv = _cases(400_000, rng)
    plane = (origin.astype(np.float64) + q.astype(np.float64) * step.astype(np.float64)).astype(F)
    exact = (plane.astype(np.float64) - o.astype(np.float64)) * inv.astype(np.float64)
    a = (step * inv).astype(F)
    c = (-(o * inv).astype(F)).astype(F)
    t = _fma32(q, a, _fma32(origin, inv, c)).astype(np.float64)
    # unwidened, an entry distance can lie beyond the exact one (a box missed) and an exit before it
    assert (t > exact).sum() > 100 and (t < exact).sum() > 100


def test_a_margin_too_small_fails_the_check():
    """A margin of a tenth of the kernels' (k = 6e-8 (C + |o|)) is not enough on these cases, so the
    margin's size is what the first check pins."""
    rng = np.random.default_rng(7)
    origin, step, q, o, inv = _cases(400_000, rng)
    plane = (origin.astype(np.float64) + q.astype(np.float64) * step.astype(np.float64)).astype(F)
    exact = (plane.astype(np.float64) - o.astype(np.float64)) * inv.astype(np.float64)
    a = (step * inv).astype(F)
    c = (-(o * inv).astype(F)).astype(F)
    k, _ = _margin(origin, step, o)
    k = (k * F(0.1)).astype(F)
    b = _fma32(origin, inv, c)
    ai = np.abs(inv).astype(F)
    entry = _fma32(q, a, _fma32(-k, ai, b)).astype(np.float64)
    exit_ = _fma32(q, a, _fma32(k, ai, b)).astype(np.float64)
    assert (entry > exact).sum() + (exit_ < exact).sum() > 0
