"""GPU parity of the batched search strategies (pftcc_strategy_srch) against the oracle loops
(oracle/strategy_oracle.py) run on the same correlation vectors.

The correlation vectors come from pftcc_gencorrs_batch, the same kernel and arguments the
strategy call runs (their values are checked against the FP64 oracle in test_gpu_parity.py), so
every output here is integer or selection work on identical FP32 values and must agree exactly:
references, rotations, correlations, counts, the searched set and the sorted order.
"""
import os

import numpy as np
import pytest

import strategy_oracle as S
from cases import make_case
from gpu_helpers import say, to_lib
from pftcc import SRCH_GREEDY2D, SRCH_GREEDY3D, SRCH_SHC3D, SRCH_SNHC2D, SRCH_SNHC3D, PftccError

pytestmark = pytest.mark.gpu

# ring length 308 runs the fused kernel of the benchmark; 56 has no fused instantiation (generic)
SHAPES = [dict(seed=31, n=308, kto=16, nrefs=40, nptcls=10, ldim=200),
          dict(seed=32, n=56, kto=10, nrefs=300, nptcls=6, ldim=64)]


def _setup(shape):
    pf, cps = make_case(seed=shape["seed"], n=shape["n"], kto=shape["kto"], ctf=True, nrefs=shape["nrefs"],
                        nptcls=shape["nptcls"], ldim=shape["ldim"], eo_mixed=True)
    lib = to_lib(pf, cps)
    nrefs, npt = shape["nrefs"], shape["nptcls"]
    ids = np.arange(1, npt + 1)
    cc = lib.gencorrs_batch(ids)[:, :, 0, :]                     # [np][nrefs][nrots]
    rng = np.random.default_rng(shape["seed"])
    ok = rng.random(nrefs) > 0.2
    order = np.stack([rng.permutation(nrefs) for _ in range(npt)])
    best = cc.max(axis=2)
    # thresholds that stop the stochastic scans early, late and never, and one below everything
    def thr(p):
        k = p % 5
        if k == 3:
            return best[p].max() + 1.0
        if k == 4:
            return best[p].min() - 1.0
        return np.quantile(best[p], (0.97, 0.8, 0.5)[k])
    prev = np.array([thr(p) for p in range(npt)], np.float32)
    rnd = rng.random(npt).astype(np.float32)
    return lib, ids, cc, ok, order, prev, rnd


@pytest.mark.parametrize("shape", SHAPES, ids=["n308_fused", "n56_generic"])
def test_strategies_match_oracle(shape):
    say(f"strategy parity n={shape['n']} nrefs={shape['nrefs']}")
    lib, ids, cc, ok, order, prev, rnd = _setup(shape)
    npt, nrefs = cc.shape[:2]

    out = lib.strategy_srch(SRCH_GREEDY2D, ids, ref_ok=ok, ref_corrs=True)
    for p in range(npt):
        r, t, c, ne = S.greedy_2d(cc[p], ok)
        assert (out["best_ref"][p], out["best_rot"][p], out["best_corr"][p], out["nrefs_eval"][p]) == \
               (r + 1, t + 1, c, ne), p
        np.testing.assert_array_equal(out["ref_rots"][p], np.where(ok, cc[p].argmax(axis=1) + 1, 0))

    for sb in (0.0, 0.6):
        say(f"  snhc stoch_bound {sb}")
        out = lib.strategy_srch(SRCH_SNHC2D, ids, srch_order=order + 1, ref_ok=ok, prev_corr=prev, rnd=rnd,
                                stoch_bound=sb, ref_corrs=True)
        for p in range(npt):
            r, t, c, ne, found, searched = S.snhc_2d(cc[p], list(order[p]), ok, prev[p], rnd[p], sb)
            got = (out["best_ref"][p], out["best_rot"][p], out["best_corr"][p], out["nrefs_eval"][p],
                   bool(out["found_better"][p]))
            assert got == (r + 1, t + 1, c, ne, found), (p, got, (r + 1, t + 1, c, ne, found))
            want = np.where(searched, cc[p].max(axis=1), -S.HUGE).astype(np.float32)
            np.testing.assert_array_equal(out["ref_corrs"][p], want)

    say("  greedy 3-D / shc 3-D")
    ntop = min(nrefs, 25)
    out = lib.strategy_srch(SRCH_GREEDY3D, ids, ref_ok=ok, ntop=ntop, ref_corrs=True)
    for p in range(npt):
        rc, rr, searched, srt, ne = S.greedy_3d(cc[p], ok)
        np.testing.assert_array_equal(out["sorted_refs"][p], np.array(srt[:ntop]) + 1)
        np.testing.assert_array_equal(out["ref_corrs"][p], np.array(rc, np.float32))
        assert (out["best_ref"][p], out["best_rot"][p], out["nrefs_eval"][p]) == (srt[0] + 1, rr[srt[0]] + 1, ne)
    out = lib.strategy_srch(SRCH_SHC3D, ids, srch_order=order + 1, ref_ok=ok, prev_corr=prev, ntop=nrefs,
                            ref_corrs=True)
    for p in range(npt):
        rc, rr, searched, srt, ne, found = S.shc_3d(cc[p], list(order[p]), ok, prev[p])
        np.testing.assert_array_equal(out["sorted_refs"][p], np.array(srt) + 1)
        np.testing.assert_array_equal(out["ref_rots"][p], np.array(rr) + 1)
        assert (out["nrefs_eval"][p], bool(out["found_better"][p])) == (ne, found), p
        assert (out["best_ref"][p], out["best_corr"][p]) == (srt[0] + 1, rc[srt[0]])
    szsn = max(1, (2 * nrefs) // 3)
    say(f"  snhc 3-D szsn {szsn}")
    out = lib.strategy_srch(SRCH_SNHC3D, ids, srch_order=order + 1, ref_ok=ok, ntop=nrefs, ref_corrs=True,
                            szsn=szsn)
    for p in range(npt):
        rc, rr, searched, srt, ne = S.snhc_3d(cc[p], list(order[p]), szsn)
        np.testing.assert_array_equal(out["sorted_refs"][p], np.array(srt) + 1)
        np.testing.assert_array_equal(out["ref_corrs"][p], np.array(rc, np.float32))
        assert (out["nrefs_eval"][p], out["found_better"][p]) == (ne, 0), p
        assert (out["best_ref"][p], out["best_rot"][p], out["best_corr"][p]) == (srt[0] + 1, rr[srt[0]] + 1, rc[srt[0]])
    lib.kill()


def test_greedy_3d_equals_device_search_and_errors():
    shape = SHAPES[0]
    say("strategy greedy 3-D vs pftcc_search; error paths")
    lib, ids, cc, ok, order, prev, rnd = _setup(shape)
    out = lib.strategy_srch(SRCH_GREEDY3D, ids, ntop=1)
    c0, ir0, _, rot0 = lib.search(ids)
    np.testing.assert_array_equal(out["best_ref"], ir0)
    np.testing.assert_array_equal(out["best_rot"], rot0)
    np.testing.assert_array_equal(out["best_corr"], c0)
    np.testing.assert_array_equal(out["sorted_refs"][:, 0], ir0)
    # no class with members / no state: nothing searched, the previous parameters stay
    none = np.zeros(cc.shape[1], bool)
    for kind, kw in ((SRCH_GREEDY2D, {}), (SRCH_SNHC2D, dict(srch_order=order + 1, rnd=rnd)),
                     (SRCH_GREEDY3D, dict(ntop=3)), (SRCH_SHC3D, dict(srch_order=order + 1, ntop=3))):
        out = lib.strategy_srch(kind, ids, ref_ok=none, prev_corr=prev, ref_corrs=True, **kw)
        np.testing.assert_array_equal(out["best_ref"], 0)
        np.testing.assert_array_equal(out["best_rot"], 0)
        np.testing.assert_array_equal(out["best_corr"], prev)
        np.testing.assert_array_equal(out["found_better"], 0)
        assert np.all(out["ref_corrs"] == -S.HUGE) and np.all(out["ref_rots"] == 0)
        if "sorted_refs" in out:
            np.testing.assert_array_equal(out["sorted_refs"], np.tile(np.arange(1, 4), (len(ids), 1)))
    bad = order.copy() + 1
    bad[1, 3] = bad[1, 4]
    with pytest.raises(PftccError, match="permutation"):
        lib.strategy_srch(SRCH_SHC3D, ids, srch_order=bad, prev_corr=prev)
    with pytest.raises(PftccError, match="uniform draw"):
        lib.strategy_srch(SRCH_SNHC2D, ids, srch_order=order + 1, prev_corr=prev)
    with pytest.raises(PftccError, match="outside"):
        lib.strategy_srch(SRCH_SNHC2D, ids, srch_order=order + 1, prev_corr=prev, rnd=np.ones(len(ids)))
    lib.kill()


def test_chunked_calls_equal_one_chunk():
    """A correlation budget of a few vectors splits the greedy kinds into many particle chunks and
    the rounds of the stochastic kinds into sub-chunks: every output is unchanged."""
    shape = SHAPES[1]
    say("strategy chunking: budget of 3 particles' vectors")
    lib, ids, cc, ok, order, prev, rnd = _setup(shape)
    nrots = cc.shape[2]
    runs = [dict(kind=SRCH_GREEDY2D, ref_ok=ok, ref_corrs=True),
            dict(kind=SRCH_SNHC2D, srch_order=order + 1, ref_ok=ok, prev_corr=prev, rnd=rnd, stoch_bound=0.3,
                 ref_corrs=True),
            dict(kind=SRCH_GREEDY3D, ref_ok=ok, ntop=7, ref_corrs=True),
            dict(kind=SRCH_SHC3D, srch_order=order + 1, ref_ok=ok, prev_corr=prev, ntop=7, ref_corrs=True),
            dict(kind=SRCH_SNHC3D, srch_order=order + 1, ntop=7, ref_corrs=True, szsn=150)]
    base = [lib.strategy_srch(r["kind"], ids, **{k: v for k, v in r.items() if k != "kind"}) for r in runs]
    os.environ["PFTCC_STRATEGY_CC_BYTES"] = str(3 * 40 * nrots * 4)
    try:
        small = [lib.strategy_srch(r["kind"], ids, **{k: v for k, v in r.items() if k != "kind"}) for r in runs]
    finally:
        del os.environ["PFTCC_STRATEGY_CC_BYTES"]
    for b, s_ in zip(base, small):
        assert b.keys() == s_.keys()
        for k in b:
            np.testing.assert_array_equal(b[k], s_[k], err_msg=k)
    lib.kill()
