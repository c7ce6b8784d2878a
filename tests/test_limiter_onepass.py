"""The FP32 path evaluates both limiters from one pass over a cell's faces.

Per face f the reference (computeReconstruction.cpp:19-40, slopeLimiter.cpp:
20-25) takes 1 for a flat difference (|d_f| <= 1e-16 max(|mn|, |mx|)) and
phi(bound / d_f) otherwise; the limiter is min(1, min_f of those).  The kernel
keeps only max_f d_f and min_f d_f and forms min(1, phi(min ratio)), using that
phi rises to 1 at X1 (1 minmod, 2 smooth) and stays >= 1 beyond.  This checks
that identity in FP64 on random and adversarial faces: zero bounds, all-flat
sides, ratios near X1, and the slope limiter applied to the
reconstruction-limited differences.
"""
import numpy as np
import pytest

MINMOD, SMOOTH = 1, 2


def phi(mode, x):
    if mode == MINMOD:
        return np.minimum(1.0, x)
    e = 0.2 * 1.0e-16
    return (x * x + 2.0 * x + e) / (x * x + x + 2.0 + e)


def per_face(mode, d, mn, mx):
    """min(1, min_f limiter(d_f)) face by face, as the reference loops."""
    lm = 1.0
    tau = 1.0e-16 * max(abs(mx), abs(mn))
    for df in d:
        if abs(df) <= tau:
            v = 1.0
        else:
            v = phi(mode, (mn if df < 0 else mx) / df)
        lm = min(lm, v)
    return lm


def min_ratio(mn, mx, dmin, dmax, s):
    tau = 1.0e-16 * max(abs(mx), abs(mn))
    hi, lo = dmax * s, dmin * s
    a = (mx / hi if mx > 0 else 0.0) if hi > tau else np.inf
    b = (mn / lo if mn < 0 else 0.0) if lo < -tau else np.inf
    return min(a, b)


def lim_from_ratio(mode, x):
    return 1.0 if x >= (1.0 if mode == MINMOD else 2.0) else phi(mode, x)


def one_pass(mode, d, c, mn, mx):
    """The kernel's form: two running extremes of d_f; the slope limiter's
    y_f = d_f + c (c = Wt dt', the same on every face) takes its extremes from
    them, since rounding is monotone."""
    dmin, dmax = d.min(), d.max()
    dlo, dhi = min(0.0, dmin), max(0.0, dmax)
    ylo, yhi = min(0.0, dmin + c), max(0.0, dmax + c)
    lr = lim_from_ratio(mode, min_ratio(mn, mx, dlo, dhi, 1.0))
    return lr, lr * lim_from_ratio(mode, min_ratio(mn, mx, ylo, yhi, lr))


def reference(mode, d, c, mn, mx):
    lr = per_face(mode, d, mn, mx)
    return lr, lr * per_face(mode, lr * (d + c), mn, mx)


def _cases(rng):
    for _ in range(3000):
        nf = int(rng.integers(1, 30))
        mn = -abs(rng.normal()) * rng.choice([1.0, 0.0, 1e-3])
        mx = abs(rng.normal()) * rng.choice([1.0, 0.0, 1e-3])
        d = rng.normal(size=nf) * rng.choice([1.0, 0.1, 10.0])
        kind = rng.integers(0, 5)
        if kind == 1:          # some faces flat
            d[rng.random(nf) < 0.5] *= 1e-20
        elif kind == 2:        # every face flat
            d *= 1e-20
        elif kind == 3:        # one side only
            d = np.abs(d) * rng.choice([-1.0, 1.0])
        elif kind == 4 and mx > 0:  # ratio straddling X1 = 2 / 1
            d[0] = mx / rng.choice([1.999999, 2.0, 2.000001, 0.999999, 1.000001])
        c = rng.normal() * rng.choice([0.3, 0.0, 3.0])
        yield d, c, mn, mx


@pytest.mark.parametrize("mode", [MINMOD, SMOOTH])
def test_one_pass_limiters_equal_per_face_limiters(mode):
    rng = np.random.default_rng(7)
    n = 0
    for d, c, mn, mx in _cases(rng):
        a = np.array(one_pass(mode, d, c, mn, mx))
        b = np.array(reference(mode, d, c, mn, mx))
        assert np.allclose(a, b, rtol=1e-12, atol=1e-300), (mode, d, c, mn, mx, a, b)
        n += 1
    assert n == 3000


def test_phi_rises_to_one_at_x1_and_stays_above():
    x = np.linspace(0.0, 2.0, 20001)
    p = phi(SMOOTH, x)
    assert np.all(np.diff(p) > 0) and abs(p[-1] - 1.0) < 1e-15
    xs = np.concatenate([np.linspace(2.0, 1e3, 100001), np.geomspace(1e3, 1e15, 1000)])
    assert np.all(phi(SMOOTH, xs) >= 1.0 - 1e-15)
