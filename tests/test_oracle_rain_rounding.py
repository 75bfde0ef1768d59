"""Stochastic rounding of the rain source (an option of this build, off by
default; the reference rounds to nearest).

The oracle's restatement (``oracle/swe_oracle.py``: ``sr_hash``, ``sr_seed``,
``sr_add``) is the checker of the device's ``sr_*`` functions, so it is pinned
here against what defines it: the lowbias32 hash recomputed with Python
integers, the two float neighbours of the exact sum, the probability of each,
and the mass balance it is for.
"""

from __future__ import annotations

import numpy as np


def _lowbias32(x: int) -> int:
    x &= 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def test_hash_is_lowbias32():
    from oracle import swe_oracle as o

    rng = np.random.default_rng(11)
    xs = rng.integers(0, 2**32, size=2000, dtype=np.uint64)
    got = o.sr_hash(xs)
    assert got.dtype == np.uint32
    assert [int(v) for v in got] == [_lowbias32(int(x)) for x in xs]
    for step in (0, 1, 7, 123456, 2**31 - 1):
        assert int(o.sr_seed(step)) == _lowbias32((step * 0x9E3779B9 + 0x632BE5AB) & 0xFFFFFFFF)


def _neighbours(a: np.ndarray, r: np.ndarray):
    """Nearest float32 sum, and the other float32 bracketing the exact sum."""
    exact = a.astype(np.float64) + r.astype(np.float64)
    s = (a + r).astype(np.float32)
    up = np.nextafter(s, np.float32(np.inf))
    down = np.nextafter(s, np.float32(-np.inf))
    other = np.where(exact > s.astype(np.float64), up, np.where(exact < s, down, s))
    return exact, s, other


def test_result_is_one_of_the_two_neighbours():
    from oracle import swe_oracle as o

    rng = np.random.default_rng(3)
    n = 20000
    a = (rng.lognormal(-3, 2, n) * rng.choice([-1, 1], n, p=[0.1, 0.9])).astype(np.float32)
    r = (rng.lognormal(-14, 3, n)).astype(np.float32)
    r[:100] = 0.0
    a[100:200] = -r[100:200]                     # exact zero sums
    bits = o.sr_hash(np.arange(n, dtype=np.uint64))
    got = o.sr_add(a, r, bits)
    exact, s, other = _neighbours(a, r)
    assert np.all((got == s) | (got == other))
    exact_sum = exact == s.astype(np.float64)
    assert np.array_equal(got[exact_sum], s[exact_sum])   # representable sums are kept
    assert np.array_equal(got[:100], a[:100])             # adding no rain changes nothing
    moved = got != s
    assert 0.1 < moved.mean() < 0.9


def test_probability_makes_the_expected_sum_exact():
    """a + r with r a fixed fraction of a unit in the last place: the share of
    draws rounded up is that fraction, and the mean result is the exact sum."""
    from oracle import swe_oracle as o

    n = 1 << 16
    bits = o.sr_hash(np.arange(n, dtype=np.uint64) ^ np.uint64(o.sr_seed(5)))
    for a0, frac in ((1.0, 0.3), (0.0371, 0.77), (3.5, 0.05), (-0.25, 0.5)):
        a = np.full(n, a0, np.float32)
        ulp = float(np.spacing(np.float32(abs(a0))))
        r = np.full(n, frac * ulp if a0 > 0 else -frac * ulp, np.float32)
        got = o.sr_add(a, r, bits).astype(np.float64)
        exact = float(a[0]) + float(r[0])
        p = float(np.mean(got != float((a[0] + r[0]).astype(np.float32))))
        nearest_is_exact_side = abs(frac) < 0.5
        want = frac if nearest_is_exact_side else 1.0 - frac
        sigma = np.sqrt(want * (1 - want) / n)
        assert abs(p - want) < 5 * sigma + 1e-12, (a0, frac, p, want)
        assert abs(got.mean() - exact) < 5 * sigma * ulp + 1e-15


def test_stochastic_rain_closes_the_float32_balance():
    """The 128^2 urban workload (2400 s, dry start, rain, one open edge): with
    round-to-nearest the float32 oracle drifts to ~2.2e-7 (the rain addition's
    rounding, the same on every ponded cell); with stochastic rounding of the
    rain it closes ten times tighter, and its total depth is nearer the
    float64 oracle's."""
    from hybrid_flood_amd import synthetic as sy
    from oracle import swe_oracle as o

    u = sy.urban_case(128)
    z = np.zeros_like(u.h0)
    res = {}
    for dt, mode in ((np.float64, "nearest"), (np.float32, "nearest"), (np.float32, "stochastic")):
        op = o.OracleParams(bed=u.bed, manning_n=u.manning_n, domain_mask=u.domain_mask,
                            boundary_types=u.boundary_types, rainfall_times_s=u.rain_times_s,
                            rainfall_rates_m_s=u.rain_rates_m_s, dx=u.dx, cfl=u.cfl,
                            dry_tolerance=u.dry_tolerance, max_dt=u.max_dt, dtype=dt)
        res[(dt, mode)] = o.integrate_to_outputs(o.OracleState(u.h0, z, z), op, [0.0, 2400.0],
                                                 max_steps=100_000, rain_rounding=mode)
    r64, rn, sr = res[(np.float64, "nearest")], res[(np.float32, "nearest")], res[(np.float32, "stochastic")]
    assert rn.steps_used == sr.steps_used == r64.steps_used > 500
    assert abs(rn.mass_balance_rel) > 1e-7
    assert abs(sr.mass_balance_rel) < 0.1 * abs(rn.mass_balance_rel)
    v64 = r64.volume_final
    assert abs(sr.volume_final - v64) < 0.2 * abs(rn.volume_final - v64)
