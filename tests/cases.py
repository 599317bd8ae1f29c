"""Seeded synthetic inputs for the oracle and parity tests (polar FTs with a decaying spectrum)."""
import numpy as np

import pftcc_oracle as O


def random_pfts(rng, count, nk, n, kfrom=2, decay=1.0):
    """Complex64 [count, nk, n] with per-shell amplitude ~ k^-decay (a cryo-EM-like spectrum)."""
    ks = np.arange(kfrom, kfrom + nk, dtype=np.float64)
    amp = (ks ** -decay)[None, :, None]
    z = rng.standard_normal((count, nk, n)) + 1j * rng.standard_normal((count, nk, n))
    return (z * amp).astype(np.complex64)


def random_ctfparams(rng, count, smpd=1.3, astig=0.1, dfrange=(1.0, 3.0), phaseplate=False):
    out = []
    for _ in range(count):
        df = rng.uniform(*dfrange)
        a = rng.uniform(0.0, astig)
        out.append(O.CtfParams(smpd=smpd, kv=300.0, cs=2.7, fraca=0.1, dfx=df + a / 2,
                               dfy=df - a / 2, angast=rng.uniform(0.0, 180.0),
                               phshift=rng.uniform(0, 1.0) if phaseplate else 0.0,
                               phaseplate=phaseplate))
    return out


def make_case(seed=0, n=16, kfrom=2, kto=7, kstop=None, nrefs=3, nptcls=4, ldim=None,
              ctf=True, astig=0.1, match_filt=False, euclid=False, eo_mixed=False):
    """A small Pftcc model with random refs/ptcls, optional CTF (from the FP64 CTF model)."""
    rng = np.random.default_rng(seed)
    kstop = kto if kstop is None else kstop
    ldim = ldim if ldim is not None else max(2 * kto + 8, 16)
    grid = O.make_grid(n, kfrom, kto, ldim)
    nk = kto - kfrom + 1
    refs_e = random_pfts(rng, nrefs, nk, n, kfrom)
    refs_o = random_pfts(rng, nrefs, nk, n, kfrom)
    ptcls = random_pfts(rng, nptcls, nk, n, kfrom)
    ctfmats = None
    cps = None
    if ctf:
        cps = random_ctfparams(rng, nptcls, smpd=1.3, astig=astig)
        ctfmats = np.stack([O.ctf_abs_mat(grid, p) for p in cps]).astype(np.float32)
    iseven = np.ones(nptcls, bool)
    if eo_mixed:
        iseven = rng.random(nptcls) < 0.5
    optlp = None
    if match_filt:
        optlp = rng.uniform(0.2, 1.0, (nrefs, kstop - kfrom + 1)).astype(np.float32)
    sigma2 = None
    if euclid:
        sigma2 = rng.uniform(0.5, 2.0, (nptcls, nk)).astype(np.float32)
    return O.Pftcc(grid=grid, kstop=kstop, refs_even=refs_e, refs_odd=refs_o, ptcls=ptcls,
                   ctfmats=ctfmats, iseven=iseven, match_filt=match_filt, ref_optlp=optlp,
                   sigma2=sigma2, euclid=euclid), cps
