"""Pins the oracle of the search strategies (oracle/strategy_oracle.py): shcloc's choice against
a rejection sampler, the neighbourhood size against hand-computed values, the tie rules of the
greedy loops, and the limits where the stochastic strategies reduce to the greedy ones."""
import math

import numpy as np
import pytest

import strategy_oracle as S

f32 = np.float32


def _rejection_shcloc(corrs, prev, rng):
    """shcloc as the reference draws it: irnd_uni until an available index (simple_rnd.f90:273-295)."""
    avail = [(c - prev) >= 0 for c in corrs]
    if not any(avail):
        return -1
    while True:
        i = int(rng.integers(0, len(corrs)))
        if avail[i]:
            return i


def test_shcloc_is_the_rejection_samplers_distribution():
    rng = np.random.default_rng(0)
    corrs = rng.standard_normal(40).astype(f32)
    prev = f32(np.sort(corrs)[30])                        # 10 rotations at or above prev
    avail = [i for i in range(40) if corrs[i] >= prev]
    assert len(avail) == 10
    # the oracle's map: u in [j/10, (j+1)/10) -> the j-th available rotation, so u uniform gives
    # each available rotation probability 1/10
    for j in range(10):
        for u in (j / 10 + 1e-6, (j + 0.5) / 10, (j + 1) / 10 - 1e-6):
            assert S.shcloc(corrs, prev, u) == avail[j]
    # the reference's rejection sampling gives the same distribution
    draws = np.array([_rejection_shcloc(corrs, prev, rng) for _ in range(20000)])
    counts = np.array([(draws == a).sum() for a in avail])
    assert set(np.unique(draws)) == set(avail)
    assert np.all(np.abs(counts - 2000) < 5 * math.sqrt(2000 * 0.9))
    # nothing available; everything available
    assert S.shcloc(corrs, f32(corrs.max() + 1), 0.3) == -1
    assert S.shcloc(corrs, f32(corrs.min()), 0.999) == 39


@pytest.mark.parametrize("nrefs,sb,want", [(200, 0.0, 200), (200, 0.5, 100), (5, 0.5, 3), (7, 0.5, 4),
                                           (1, 0.0, 2), (100, 1.0, 2), (100, 0.95, 5), (3, 0.9, 2)])
def test_snhc_bound(nrefs, sb, want):
    # nint rounds halves away from zero: 5 x 0.5 = 2.5 -> 3, 7 x 0.5 = 3.5 -> 4
    assert S.snhc_bound(nrefs, sb) == want


def _corrs(rng, nrefs, nrots):
    return rng.standard_normal((nrefs, nrots)).astype(f32)


def test_greedy_2d_keeps_the_last_equal_maximum_and_skips_empty_classes():
    rng = np.random.default_rng(1)
    c = _corrs(rng, 6, 16)
    c[1, 5] = c[4, 9] = f32(10.0)
    ok = [True] * 6
    assert S.greedy_2d(c, ok)[:3] == (4, 9, f32(10.0))
    ok[4] = False
    assert S.greedy_2d(c, ok)[:3] == (1, 5, f32(10.0))
    # within a reference the first maximum (maxloc)
    c[2, 3] = c[2, 7] = f32(20.0)
    assert S.greedy_2d(c, [True] * 6)[:2] == (2, 3)
    # no class with members: the previous parameters stay
    assert S.greedy_2d(c, [False] * 6, prev_corr=0.25) == (-1, -1, 0.25, 6)


def test_snhc_limits():
    rng = np.random.default_rng(2)
    c = _corrs(rng, 12, 20)
    order = list(rng.permutation(12))
    ok = [True] * 12
    ok[order[0]] = False
    best = [float(c[r].max()) for r in range(12)]
    # prev above every correlation: no exit, the whole neighbourhood searched, the best of it in
    # search order (first of equal maxima), nrefs_eval = bound + 1 when the bound cuts the order
    r, t, cc, ne, found, searched = S.snhc_2d(c, order, ok, f32(100.0), 0.5, 0.0)
    assert not found and ne == 12 and sum(searched) == 11
    assert r == max((x for x in range(12) if ok[x]), key=lambda x: best[x]) and t == int(np.argmax(c[r]))
    r, t, cc, ne, found, searched = S.snhc_2d(c, order, ok, f32(100.0), 0.5, 0.5)
    assert not found and ne == 7 and [order[i] for i in range(12) if searched[order[i]]] == order[1:6]
    assert r == max(order[1:6], key=lambda x: best[x])
    # prev below every correlation: the first class with members exits with the u-th rotation
    r, t, cc, ne, found, searched = S.snhc_2d(c, order, ok, f32(-100.0), 0.0, 0.0)
    assert found and r == order[1] and t == 0 and ne == 2 and cc == c[r, 0]
    r, t, cc, ne, found, _ = S.snhc_2d(c, order, ok, f32(-100.0), 0.999, 0.0)
    assert t == 19
    # threshold between: the first position whose best reaches prev
    thr = f32(sorted(best)[8])
    first = next(i for i in range(12) if ok[order[i]] and best[order[i]] >= thr)
    r, t, cc, ne, found, _ = S.snhc_2d(c, order, ok, thr, 0.0, 0.0)
    assert found and r == order[first] and ne == first + 1 and c[r, t] >= thr
    assert t == min(j for j in range(20) if c[r, j] >= thr)


def test_shc_3d_reduces_to_greedy_and_stops_at_the_first_improvement():
    rng = np.random.default_rng(3)
    c = _corrs(rng, 15, 24)
    order = list(rng.permutation(15))
    ok = [i % 4 != 1 for i in range(15)]
    g = S.greedy_3d(c, ok)
    s = S.shc_3d(c, order, ok, f32(100.0))
    assert s[:4] == g[:4] and s[4] == sum(ok) and not s[5]
    best = [float(c[r].max()) for r in range(15)]
    thr = f32(np.median(best))
    first = next(i for i in range(15) if ok[order[i]] and best[order[i]] > thr)
    rc, rr, searched, srt, ne, found = S.shc_3d(c, order, ok, thr)
    assert found and ne == sum(ok[order[i]] for i in range(first + 1))
    assert [i for i in range(15) if searched[i]] == sorted(order[i] for i in range(first + 1) if ok[order[i]])
    assert srt[0] == max((r for r in range(15) if searched[r]), key=lambda r: rc[r])


def test_sort_corrs_against_argsort():
    rng = np.random.default_rng(4)
    vals = list(rng.standard_normal(50).astype(f32))
    searched = [i % 3 != 0 for i in range(50)]
    got = S.sort_corrs(vals, searched)
    v = np.array([vals[i] if searched[i] else -S.HUGE for i in range(50)], np.float64)
    assert got == list(np.argsort(-v, kind="stable"))
    # the unsearched trail in index order
    assert got[-17:] == [i for i in range(50) if not searched[i]]


def test_snhc_3d_prefix_and_full_neighbourhood():
    rng = np.random.default_rng(5)
    c = _corrs(rng, 14, 20)
    order = list(rng.permutation(14))
    full = S.snhc_3d(c, order, 14)
    g = S.greedy_3d(c, [True] * 14)
    assert full[:4] == g[:4] and full[4] == 14
    rc, rr, searched, srt, ne = S.snhc_3d(c, order, 5)
    assert ne == 5 and [i for i in range(14) if searched[i]] == sorted(order[:5])
    assert srt[:5] == sorted(order[:5], key=lambda r: (-float(c[r].max()), r))
