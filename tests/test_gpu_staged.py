"""Parity of the staged-reference plain kernel (the exhaustive shift-grid search) with the oracle.

The plain FP32 search switches to its staged-reference variant when a work-group's tile of
(reference, shift) pairs spans at most two references, i.e. when the number of shifts is at least
the tile: 9 at pftsz 308, 5 at 192 and 400, 32 at small sizes.  That variant stages the two
references' |CTF|-weighted rows per shell in LDS, sums their FP64 norms from the staged rows, and
orders work-groups tile-major over groups of 16 particles.  These tests run it (checked through
pftcc_get_corr_tile) with tiles that straddle two references (nr x ns not a multiple of the tile)
and particle counts that leave a partial group of 16, against:
  * the FP64 oracle of gencorrs_cc_2 (simple_polarft_corrcalc.f90:1502-1534) for every vector:
    within 1e-5 of the peak, best rotation identical where the top-2 gap exceeds 10x the error;
  * the device's own full vectors for pftcc_search (first maximum in (ref, shift, rot) order);
  * on a slice of the 2D-classification workload (box 200, kfromto [2,60], the +-3 px / 0.5 px grid
    of 169 shifts), the C++ restatement of the reference's FP32 FFT route and an FP64 brute force
    of the winning pair.
"""
import os

import numpy as np
import pytest

import cpu_port
import pftcc_oracle as O
from cases import make_case
from gpu_helpers import peak_rel_err, say, to_lib, top2_gap
import pftcc
from pftcc import Params, PolarFTCorrCalc, synth

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-5


def shift_grid(lim, step):
    g = np.arange(-lim, lim + 1e-6, step)
    return np.array([(x, y) for x in g for y in g], np.float32)


@pytest.mark.parametrize("n,kto,ldim,shifts,nptcls,nrefs", [
    (308, 60, 200, shift_grid(3.0, 0.5), 17, 3),     # 169 shifts, tile 9: 507 pairs, 3 left over
    (192, 40, 128, shift_grid(1.5, 1.5)[:7], 21, 3),  # 7 shifts, tile 5
    (400, 80, 256, shift_grid(1.0, 1.0), 17, 2),      # 9 shifts, tile 4
    (500, 40, 320, shift_grid(1.0, 1.0), 17, 2),      # 9 shifts, tile 4
    (16, 7, 32, shift_grid(2.0, 0.5)[:33], 18, 2),    # 33 shifts, tile 32
])
def test_staged_reference_kernel_matches_oracle(n, kto, ldim, shifts, nptcls, nrefs):
    ns = len(shifts)
    say(f"staged kernel: n={n} kto={kto} box={ldim} ns={ns} np={nptcls} nr={nrefs}")
    pf, cps = make_case(seed=500 + n, n=n, kfrom=2, kto=kto, ctf=True, astig=0.5, nrefs=nrefs, nptcls=nptcls,
                        ldim=ldim, eo_mixed=True)
    lib = to_lib(pf, cps)
    tile, staged = lib.get_corr_tile(ns)
    assert staged, (tile, ns)
    assert (nrefs * ns) % tile != 0 and tile <= ns
    ids = np.arange(1, nptcls + 1)
    full = lib.gencorrs_batch(ids, None, shifts)               # [np][nr][ns][nrots]
    worst = 0.0
    for i in range(nptcls):
        for r in range(nrefs):
            ref = pf.gencorrs_shifts(r, i, shifts)
            for s in range(ns):
                e = peak_rel_err(full[i, r, s], ref[s])
                worst = max(worst, e)
                assert e < FP32_TOL, (i, r, s, e)
                if top2_gap(ref[s]) > 10 * e * np.max(np.abs(ref[s])):
                    assert int(np.argmax(full[i, r, s])) == int(np.argmax(ref[s]))
    say(f"  max peak-relative error vs FP64 oracle {worst:.2e} (tile {tile})")
    cc, ir, ish, rot = lib.search(ids, None, shifts)
    for i in range(nptcls):
        r, s, t = np.unravel_index(int(np.argmax(full[i])), full[i].shape)
        assert (ir[i], ish[i], rot[i]) == (r + 1, s, t + 1)
        assert cc[i] == full[i, r, s, t]
    lib.kill()


def test_staged_kernel_on_2d_classification_slice():
    """16 particles of the 2D-classification workload against 40 of its 200 references (each
    particle's source reference among them) over the 169-shift grid: the device's best (reference,
    shift, rotation) equals the reference algorithm's (C++ FP32 FFT route) wherever that algorithm's
    own top-2 gap exceeds FP32 resolution, the best cc agrees within 1e-5, and the device's best cc
    equals the FP64 brute force of the same (reference, shift, rotation) within 1e-5."""
    box, smpd, kfromto = 200, 1.0, (2, 60)
    pftsz = pftcc.magic_pftsz(98)
    w = synth.Workload(box=box, smpd=smpd, kfromto=kfromto, pftsz=pftsz, nrefs=200, nptcls=16, snr=0.03,
                       seed=20251210, dfrange=(0.8, 3.5), astig=0.1, max_shift=3.0)
    synth.make_params(w)
    say("staged kernel: 2D-classification slice, 16 ptcls x 40 refs x 169 shifts, box 200")
    lib = PolarFTCorrCalc().new(w.nrefs, (1, w.nptcls), Params(boxmatch=box, kfromto=kfromto, pftsz=pftsz, ctf=True))
    synth.load_into(lib, w, chunk=16)
    shifts = shift_grid(3.0, 0.5)
    rng = np.random.default_rng(7)
    src = np.unique(w.prm["ref"] + 1)
    others = rng.permutation(np.setdiff1d(np.arange(1, 201), src))[:40 - len(src)]
    refs_sel = np.sort(np.concatenate([src, others])).astype(np.int32)
    nr = len(refs_sel)
    tile, staged = lib.get_corr_tile(len(shifts))
    assert staged and tile == 9
    ids = np.arange(1, w.nptcls + 1)
    gcc, gir, gish, grot = lib.search(ids, refs_sel, shifts)
    refs = np.stack([lib.get_ref_pft(int(r), True) for r in refs_sel])
    pt = np.stack([lib.get_ptcl_pft(i) for i in ids])
    ctf = np.stack([lib.get_ctfmat(i) for i in ids])
    sq = np.array([lib.get_sqsum_ptcl(i) for i in ids], np.float32)
    lib.kill()
    grid = O.make_grid(pftsz, kfromto[0], kfromto[1], box)
    shm = O.shmat_8_many(grid, shifts)
    nthr = min(16, os.cpu_count() or 1)
    pa, pb = cpu_port.memoize(pt, nthr)
    say(f"  CPU reference algorithm over {nr} refs x {len(shifts)} shifts")
    ccc, cir, cish, crot, allc = cpu_port.search_fp32(refs, ctf, sq, pa, pb, shm.astype(np.complex64), w.nk, nthr,
                                                      want_all=True)
    gpos = np.searchsorted(refs_sel, gir)          # device iref -> position in the subset
    same = (gpos + 1 == cir) & (gish == cish) & (grot == crot)
    for i in np.where(~same)[0]:
        v = allc[i]
        top = np.sort(v.ravel())
        assert top[-1] - top[-2] <= 1e-5 * top[-1], (i, top[-1] - top[-2])
        assert v[gpos[i], gish[i], grot[i] - 1] >= top[-1] - 1e-5 * top[-1]
    assert np.max(np.abs(gcc - ccc)) < 1e-5
    # FP64 brute force of each winner (gencorrs_cc_2: reference norm over all shells)
    for i in range(w.nptcls):
        d64 = cpu_port.gencorrs_direct64(refs[gpos[i]:gpos[i] + 1], pt[i:i + 1], ctf[i:i + 1],
                                         shm[gish[i]:gish[i] + 1], w.nk, True, 1)
        assert abs(float(d64[0, 0, grot[i] - 1]) - float(gcc[i])) < 1e-5
    say(f"  identical assignments {np.mean(same):.3f}; max |dcc| vs CPU {np.max(np.abs(gcc - ccc)):.2e}")
