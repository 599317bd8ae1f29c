"""Oracle (test infrastructure only): the reference-loop search strategies restated over given
correlation vectors, one particle at a time, in the reference's loop order.

Only tests may import this module; the product path never does.

Inputs per particle: corrs[nrefs][nrots] (what pftcc_glob%gencorrs(iref, iptcl, corrs) returns
for every reference), the 0-based search order, the per-reference availability (cls_pops > 0 /
state_exists), prev_corr and one uniform draw u for shcloc.

  greedy_2d   simple_strategy2D_greedy.f90:33-58
  snhc_2d     simple_strategy2D_snhc.f90:33-93 (+ shcloc, simple_rnd.f90:273-295)
  greedy_3d   simple_strategy3D_greedy.f90:40-77 (+ sort_corrs)
  shc_3d      simple_strategy3D_shc.f90:36-83 (+ sort_corrs)
  snhc_3d     simple_strategy3D_snhc.f90:37-60 (+ sort_corrs)
  sort_corrs  simple_strategy3D_utils.f90:100-124

shcloc draws irnd_uni until it hits an available rotation, so its choice is uniform over the
available set; the random stream of the reference cannot be reproduced, so the oracle takes the
draw as an argument and picks the floor(u count)-th available rotation, in rotation order (the
same uniform distribution; pinned against a rejection sampler in tests/test_oracle_strategy.py).
sort_corrs calls hpsort (a heapsort, not stable): equal correlations (in practice the -huge of
unsearched references) come out in an order the reference leaves unspecified; the oracle breaks
ties by the lower reference index.
"""
import math

import numpy as np

HUGE = float(np.finfo(np.float32).max)


def maxloc(v):
    """Fortran maxloc: the first index of the largest value (0-based here)."""
    best, bi = None, -1
    for i, x in enumerate(v):
        if bi < 0 or x > best:
            best, bi = x, i
    return bi


def shcloc(corrs, corr_prev, u):
    """simple_rnd.f90:273-295 with the draw u in [0, 1): 0-based rotation, or -1 if corrs <
    corr_prev everywhere (the reference returns 0)."""
    f32 = np.float32
    diffs = np.asarray(corrs, f32) - f32(corr_prev)
    avail = [i for i, d in enumerate(diffs) if not d < 0]
    if not avail:
        return -1
    k = min(max(int(math.floor(float(u) * len(avail))), 0), len(avail) - 1)
    return avail[k]


def snhc_bound(nrefs, stoch_bound):
    """nrefs_bound of srch_snhc (:45-46): max(2, min(nrefs, nint(real(nrefs) (1 - stoch_bound))))
    in default (single precision) real; nint rounds halves away from zero."""
    x = float(np.float32(nrefs) * (np.float32(1.0) - np.float32(stoch_bound)))
    n = int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))
    return max(2, min(nrefs, n))


def greedy_2d(corrs, ok, prev_corr=None):
    """srch_greedy (2D): returns (best_ref, best_rot, best_corr, nrefs_eval), refs/rots 0-based,
    best_ref -1 if no class has members."""
    corr = -HUGE
    best = (-1, -1, prev_corr if prev_corr is not None else -HUGE)
    nrefs = len(corrs)
    for iref in range(nrefs):
        if not ok[iref]:
            continue
        inpl = maxloc(corrs[iref])
        c = corrs[iref][inpl]
        if c >= corr:
            corr = c
            best = (iref, inpl, c)
    return best + (nrefs,)


def snhc_2d(corrs, order, ok, prev_corr, u, stoch_bound):
    """srch_snhc: (best_ref, best_rot, best_corr, nrefs_eval, found_better, searched[nrefs])."""
    nrefs = len(corrs)
    bound = snhc_bound(nrefs, stoch_bound)
    cc_glob, class_glob, inpl_glob = -HUGE, -1, -1
    found = False
    nrefs_eval = 0
    searched = [False] * nrefs
    best = None
    for isample in range(nrefs):
        iref = order[isample]
        nrefs_eval += 1
        if nrefs_eval > bound:
            break
        if not ok[iref]:
            continue
        searched[iref] = True
        inpl = shcloc(corrs[iref], prev_corr, u)
        if inpl < 0:
            inpl = maxloc(corrs[iref])
            c = corrs[iref][inpl]
        else:
            c = corrs[iref][inpl]
            best = (iref, inpl, c)
            found = True
        if c > cc_glob:
            cc_glob, class_glob, inpl_glob = c, iref, inpl
        if found:
            break
    if not found:
        best = (class_glob, inpl_glob, cc_glob if class_glob >= 0 else prev_corr)
    return best + (nrefs_eval, found, searched)


def sort_corrs(ref_corrs, searched):
    """Reference indices, best first (the reversed ascending order of sort_corrs); unsearched at
    -huge; ties to the lower index."""
    vals = [ref_corrs[i] if searched[i] else -HUGE for i in range(len(ref_corrs))]
    return sorted(range(len(vals)), key=lambda i: (-vals[i], i))


def greedy_3d(corrs, ok):
    """srch_greedy (3D): per-reference maxloc of every reference with a state, then sort_corrs.
    Returns (ref_corrs, ref_rots, searched, sorted_refs, nrefs_eval)."""
    nrefs = len(corrs)
    rc, rr, searched = [-HUGE] * nrefs, [-1] * nrefs, [False] * nrefs
    for iref in range(nrefs):
        if ok[iref]:
            inpl = maxloc(corrs[iref])
            rc[iref], rr[iref], searched[iref] = corrs[iref][inpl], inpl, True
    return rc, rr, searched, sort_corrs(rc, searched), nrefs


def shc_3d(corrs, order, ok, prev_corr):
    """srch_shc: stops after the first reference whose best exceeds prev_corr.
    Returns (ref_corrs, ref_rots, searched, sorted_refs, nrefs_eval, found)."""
    nrefs = len(corrs)
    rc, rr, searched = [-HUGE] * nrefs, [-1] * nrefs, [False] * nrefs
    nbetter = 0
    nrefs_eval = 0
    for isample in range(nrefs):
        iref = order[isample]
        if ok[iref]:
            inpl = maxloc(corrs[iref])
            rc[iref], rr[iref], searched[iref] = corrs[iref][inpl], inpl, True
            if corrs[iref][inpl] > prev_corr:
                nbetter += 1
            nrefs_eval += 1
        if nbetter > 0:
            break
    return rc, rr, searched, sort_corrs(rc, searched), nrefs_eval, nbetter > 0


def snhc_3d(corrs, order, szsn):
    """srch_snhc (3D): the first szsn references of the search order, every one searched (the loop
    has no state test), then sort_corrs.  Returns (ref_corrs, ref_rots, searched, sorted_refs,
    nrefs_eval)."""
    nrefs = len(corrs)
    rc, rr, searched = [-HUGE] * nrefs, [-1] * nrefs, [False] * nrefs
    for isample in range(szsn):
        iref = order[isample]
        inpl = maxloc(corrs[iref])
        rc[iref], rr[iref], searched[iref] = corrs[iref][inpl], inpl, True
    return rc, rr, searched, sort_corrs(rc, searched), szsn
