"""Hybrid stochastic PT2 estimator (pt2_stoch_routines.irp.f) on synthetic
per-generator contributions: the comb geometry against brute force, the
estimator's unbiasedness and error bars over many seeds, and the exact sum
once every generator is computed."""
import numpy as np
import pytest

from quantum_package_amd.pt2_stoch import first_tooth, get_comb, make_comb, pt2_find, stochastic_pt2_state


def _system(n, seed):
    rng = np.random.default_rng(seed)
    c = np.sort(rng.pareto(1.5, n) + 0.01)[::-1] * np.where(rng.random(n) < 0.5, -1, 1)
    c[0] = 10.0 * np.abs(c).max()                    # an HF-like leading generator
    w = c ** 2
    # PT2 contributions correlated with the weight, with noise and both signs of deviation
    e = -w * (0.5 + rng.random(n)) - 1e-4 * rng.random(n)
    return c, e


def test_pt2_find_is_the_first_index_reaching_the_point():
    rng = np.random.default_rng(0)
    w = np.cumsum(rng.random(500))
    w /= w[-1]
    for _ in range(300):
        v = rng.random()
        lo = int(rng.integers(1, 400))
        hi = int(rng.integers(lo + 1, 502))
        got = pt2_find(v, w, lo, hi)
        want = next((i for i in range(lo, hi) if w[i - 1] >= v), hi)
        assert got == want


def test_comb_geometry():
    c, _ = _system(3000, 1)
    cb = make_comb(c)
    w = c ** 2 / np.sum(c ** 2)
    assert np.allclose(cb.weight, w, rtol=1e-14)
    assert np.allclose(cb.cweight, np.cumsum(w), atol=1e-12) and cb.cweight[-1] == 1.0
    f = cb.first_det_of_comb
    # the deterministic set: each of its generators weighs at least half a
    # comb step of what is left, the first one after it does not
    left = 1.0
    for i in range(f - 1):
        assert w[i] / left >= 0.5 / cb.teeth
        left -= w[i]
    assert w[f - 1] / left < 0.5 / cb.teeth
    # equal steps over the remaining weight
    assert cb.step == pytest.approx((1.0 - cb.cweight[f - 2]) / cb.teeth, rel=1e-14)
    fdt = cb.first_det_of_teeth
    assert fdt[1] == f and fdt[cb.teeth + 1] == cb.n + 1
    assert np.all(np.diff(fdt[1:]) >= 0)
    # tooth j starts with the generator whose cumulated weight first reaches
    # 1 - (teeth - j + 1) * step
    for j in range(2, cb.teeth + 1):
        start = 1.0 - (cb.teeth - j + 1) * cb.step
        want = int(np.searchsorted(cb.cweight, start, side="left")) + 1
        assert abs(fdt[j] - want) <= 1
        assert cb.cweight[fdt[j] - 1] >= start - 1e-12


def test_comb_picks_one_generator_per_tooth_by_weight():
    c, _ = _system(2000, 2)
    cb = make_comb(c)
    rng = np.random.default_rng(5)
    counts = np.zeros(cb.n)
    K = 20000
    for r in rng.random(K) * cb.step:
        d = get_comb(cb, r)
        for j in range(cb.teeth):
            lo, hi = cb.first_det_of_teeth[j + 1], cb.first_det_of_teeth[j + 2]
            assert lo <= d[j] <= max(lo, hi - 1) or d[j] == hi
        counts[d - 1] += 1
    # each generator past the first tooth's start is drawn with frequency
    # weight / step (its share of one tooth), within sampling error
    f = cb.first_det_of_teeth[2]
    sel = np.arange(f, cb.n)
    p = cb.weight[sel] / cb.step
    big = p > 0.05
    assert np.all(np.abs(counts[sel][big] / K - p[big]) < 5 * np.sqrt(p[big] / K) + 1e-3)


def test_exact_once_everything_is_computed():
    c, e = _system(700, 3)
    calls = []

    def compute(g):
        calls.append(g.size)
        return e[g - 1]

    est, err, info = stochastic_pt2_state(c, compute, relative_error=0.0, absolute_error=0.0, seed=1)
    assert info["generators_computed"] == 700
    assert err == 0.0
    assert est == pytest.approx(e.sum(), rel=1e-13)
    assert sum(calls) == 700          # every generator computed once


def test_unbiased_with_calibrated_errors():
    c, e = _system(5000, 4)
    exact = e.sum()
    ests, errs = [], []
    for seed in range(300):
        est, err, info = stochastic_pt2_state(c, lambda g: e[g - 1], relative_error=1.0, absolute_error=0.0,
                                              seed=seed, combs_per_batch=16)
        assert info["generators_computed"] < 5000
        ests.append(est)
        errs.append(err)
    ests, errs = np.array(ests), np.array(errs)
    # the mean of 300 estimates sits within 4 standard errors of the exact sum
    assert abs(ests.mean() - exact) < 4 * ests.std() / np.sqrt(len(ests))
    # the reported error is the spread of the estimates, and about 95 % of
    # them fall within two reported errors
    assert 0.6 < np.median(errs) / ests.std() < 1.6
    assert 0.85 < np.mean(np.abs(ests - exact) < 2 * errs) <= 1.0


def test_first_tooth_follows_the_computed_prefix():
    c, _ = _system(1500, 6)
    cb = make_comb(c)
    computed = np.zeros(cb.n, dtype=bool)
    computed[: cb.first_det_of_comb - 1] = True
    assert first_tooth(cb, computed) == 1
    for j in (2, 17, 60, cb.teeth):
        computed[:] = False
        computed[: cb.first_det_of_teeth[j]] = True     # through the first generator of tooth j
        assert first_tooth(cb, computed) == j
    computed[:] = True
    assert first_tooth(cb, computed) == cb.teeth + 1
