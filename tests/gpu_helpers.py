"""Helpers that mirror an oracle case (tests/cases.py) into the gfx950 library."""
import sys

import numpy as np

import pftcc_oracle as O
from pftcc import PFTCC_OBJFUN_CC, PFTCC_OBJFUN_EUCLID, Params, PolarFTCorrCalc


def say(msg):
    """Names the case before each GPU step (flushed, so a fault still leaves the name)."""
    print(msg, flush=True)
    sys.stdout.flush()


def to_lib(pf: O.Pftcc, cps, pfrom=1, memoize=True, nthr=1):
    """Builds a library handle holding exactly the oracle's inputs.

    When the case has a CTF, the library computes its own |CTF| matrices from the parameters
    and the oracle then adopts them, so both sides correlate identical FP32 inputs (the CTF
    itself is checked separately against the FP64 CTF model).
    """
    nrefs = pf.refs_even.shape[0]
    nptcls = pf.ptcls.shape[0]
    prm = Params(boxmatch=pf.grid.ldim, kfromto=(pf.kfrom, pf.grid.kto), kstop=pf.kstop,
                 pftsz=pf.pftsz, ctf=cps is not None, match_filt=pf.match_filt, pssnr=pf.pssnr,
                 objfun=PFTCC_OBJFUN_EUCLID if pf.euclid else PFTCC_OBJFUN_CC, nthr=nthr)
    eo = np.where(pf.iseven, 0, 1).astype(np.int32)
    lib = PolarFTCorrCalc().new(nrefs, (pfrom, pfrom + nptcls - 1), prm, eoarr=eo)
    lib.set_ref_pfts(1, pf.refs_even, True)
    lib.set_ref_pfts(1, pf.refs_odd, False)
    lib.set_ptcl_pfts(pfrom, pf.ptcls)
    if pf.match_filt or pf.pssnr:
        for r in range(nrefs):
            lib.set_ref_optlp(r + 1, pf.ref_optlp[r])
    if pf.euclid:
        lib.assign_sigma2_noise(pf.sigma2)
    if cps is not None:
        lib.create_polar_absctfmats(cps)
        pf.ctfmats = np.stack([lib.get_ctfmat(pfrom + i) for i in range(nptcls)])
    if memoize:
        lib.memoize_ffts()
    return lib


def peak_rel_err(got, ref):
    """max |got - ref| / max |ref| (the correlation error relative to the peak)."""
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-30))


def top2_gap(v):
    s = np.sort(np.ravel(v))
    return float(s[-1] - s[-2])
