"""Python mirror of SIMPLE's ``polarft_corrcalc`` over the gfx950 C ABI (``include/pftcc.h``).

Method names, argument meaning and index bases follow the reference type
(``src/main/simple_polarft_corrcalc.f90:68-198``): ``iref`` 1-based, ``iptcl`` a global particle
id in ``pfromto``, ``irot`` 1-based.  Polar FTs are complex64 arrays ``[nk, pftsz]`` (the memory
order of the reference's ``pft(pftsz, kfromto(1):kfromto(2))``).  Errors raise
``PftccError`` where the reference stops with THROW_HARD.

The library ``libpftcc_hip.so`` is built in-tree (``simple3.0_amd/lib``) by
``simple3.0_amd/build.py``; importing this module without it raises immediately -- there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(os.path.dirname(_HERE), "lib")
LIBPATH = os.path.join(LIBDIR, "libpftcc_hip.so")

PFTCC_OBJFUN_CC = 0
PFTCC_OBJFUN_EUCLID = 1
PFTCC_PREC_BF16 = 1
PFTCC_PREC_SPLIT_BF16 = 3
PFTCC_PREC_FP16 = 2
PFTCC_PREC_SPLIT_FP16 = 4
PFTCC_PREC_FP8 = 8


class PftccError(RuntimeError):
    pass


class Cfg(C.Structure):
    _fields_ = [("boxmatch", C.c_int32), ("ring2", C.c_int32), ("pftsz", C.c_int32),
                ("kfromto", C.c_int32 * 2), ("kstop", C.c_int32), ("objfun", C.c_int32),
                ("match_filt", C.c_int32), ("pssnr", C.c_int32), ("ctf", C.c_int32),
                ("clsfrcs", C.c_int32), ("nthr", C.c_int32)]


class CtfParamsC(C.Structure):
    _fields_ = [("smpd", C.c_float), ("kv", C.c_float), ("cs", C.c_float), ("fraca", C.c_float),
                ("dfx", C.c_float), ("dfy", C.c_float), ("angast", C.c_float),
                ("phshift", C.c_float), ("l_phaseplate", C.c_int32)]


class PrepCfgC(C.Structure):
    _fields_ = [("box", C.c_int32), ("boxmatch", C.c_int32), ("msk", C.c_float), ("inner", C.c_float),
                ("width", C.c_float), ("focusmsk", C.c_float), ("ctfflag", C.c_int32), ("griddev", C.c_int32)]


class StrategyCfgC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("stoch_bound", C.c_float), ("ntop", C.c_int32), ("szsn", C.c_int32)]


# search strategies of pftcc_strategy_srch
SRCH_GREEDY2D, SRCH_SNHC2D, SRCH_GREEDY3D, SRCH_SHC3D, SRCH_SNHC3D = 0, 1, 2, 3, 4


def _ctf_array(ctfparams):
    """ctypes array of pftcc_ctfparams from objects with smpd, kv, cs, fraca, dfx, dfy, angast[,
    phshift, phaseplate] attributes (an array built here passes through unchanged)."""
    if isinstance(ctfparams, C.Array):
        return ctfparams
    arr = (CtfParamsC * len(ctfparams))()
    for i, p in enumerate(ctfparams):
        arr[i] = CtfParamsC(p.smpd, p.kv, p.cs, p.fraca, p.dfx, p.dfy, p.angast,
                            getattr(p, "phshift", 0.0), int(getattr(p, "phaseplate", False)))
    return arr


ctf_params_array = _ctf_array


# (name, restype, argtypes) of every entry point declared in include/pftcc.h
_P = C.c_void_p
_I = C.c_int32
_F = C.c_float
_D = C.c_double
_PI = C.POINTER(C.c_int32)
_PF = C.POINTER(C.c_float)
_PD = C.POINTER(C.c_double)
SIGNATURES = [
    ("pftcc_new", _I, [C.POINTER(Cfg), _I, _I, _I, C.POINTER(C.c_int8), _PI, C.POINTER(_P)]),
    ("pftcc_kill", _I, [_P]),
    ("pftcc_last_error", C.c_char_p, [_P]),
    ("pftcc_set_stream", _I, [_P, _P]),
    ("pftcc_sync", _I, [_P]),
    ("pftcc_set_ref_pft", _I, [_P, _I, _PF, _I]),
    ("pftcc_set_ptcl_pft", _I, [_P, _I, _PF]),
    ("pftcc_set_ref_pfts", _I, [_P, _I, _I, _PF, _I]),
    ("pftcc_set_ptcl_pfts", _I, [_P, _I, _I, _PF]),
    ("pftcc_set_ref_pfts_dev", _I, [_P, _I, _I, _P, _I]),
    ("pftcc_set_ptcl_pfts_dev", _I, [_P, _I, _I, _P]),
    ("pftcc_set_ref_fcomp", _I, [_P, _I, _I, _I, _F, _F, _I]),
    ("pftcc_set_ptcl_fcomp", _I, [_P, _I, _I, _I, _F, _F]),
    ("pftcc_cp_even2odd_ref", _I, [_P, _I]),
    ("pftcc_cp_even_ref2ptcl", _I, [_P, _I, _I]),
    ("pftcc_swap_ptclsevenodd", _I, [_P]),
    ("pftcc_reallocate_ptcls", _I, [_P, _I, _PI]),
    ("pftcc_set_eo", _I, [_P, _I, _I]),
    ("pftcc_set_eos", _I, [_P, _PI]),
    ("pftcc_set_ref_optlp", _I, [_P, _I, _PF]),
    ("pftcc_assign_sigma2_noise", _I, [_P, _PF]),
    ("pftcc_zero_refs_beyond_kstop", _I, [_P]),
    ("pftcc_create_polar_absctfmats", _I, [_P, C.POINTER(CtfParamsC)]),
    ("pftcc_memoize_ffts", _I, [_P]),
    ("pftcc_get_nrots", _I, [_P, _PI]),
    ("pftcc_get_pdim", _I, [_P, _PI]),
    ("pftcc_get_pftsz", _I, [_P, _PI]),
    ("pftcc_get_box", _I, [_P, _PI]),
    ("pftcc_get_rot", _I, [_P, _I, _PF]),
    ("pftcc_get_roind", _I, [_P, _F, _PI]),
    ("pftcc_get_coord", _I, [_P, _I, _I, _PF]),
    ("pftcc_get_nrefs", _I, [_P, _PI]),
    ("pftcc_get_nptcls", _I, [_P, _PI]),
    ("pftcc_ptcl_iseven", _I, [_P, _I, _PI]),
    ("pftcc_get_ref_pft", _I, [_P, _I, _I, _PF]),
    ("pftcc_get_ptcl_pft", _I, [_P, _I, _PF]),
    ("pftcc_get_ctfmat", _I, [_P, _I, _PF]),
    ("pftcc_get_memo", _I, [_P, _I, _PF]),
    ("pftcc_get_sqsum_ptcl", _I, [_P, _I, _PF]),
    ("pftcc_gencorrs", _I, [_P, _I, _I, _PF, _PF]),
    ("pftcc_gencorr_for_rot", _I, [_P, _I, _I, _PF, _I, _PF]),
    ("pftcc_gencorr_for_rot_8", _I, [_P, _I, _I, _PD, _I, _PD]),
    ("pftcc_gencorr_grad_for_rot_8", _I, [_P, _I, _I, _PD, _I, _PD, _PD]),
    ("pftcc_gencorr_grad_only_for_rot_8", _I, [_P, _I, _I, _PD, _I, _PD]),
    ("pftcc_gencorr_cc_for_rot", _I, [_P, _I, _I, _PF, _I, _PF]),
    ("pftcc_gencorr_cc_for_rot_8", _I, [_P, _I, _I, _PD, _I, _PD]),
    ("pftcc_gencorr_cc_grad_for_rot_8", _I, [_P, _I, _I, _PD, _I, _PD, _PD]),
    ("pftcc_gencorr_cc_grad_only_for_rot_8", _I, [_P, _I, _I, _PD, _I, _PD]),
    ("pftcc_gencorr_euclid_for_rot", _I, [_P, _I, _I, _PF, _I, _PF]),
    ("pftcc_gencorr_euclid_for_rot_8", _I, [_P, _I, _I, _PD, _I, _PD]),
    ("pftcc_gencorr_euclid_grad_for_rot_8", _I, [_P, _I, _I, _PD, _I, _PD, _PD]),
    ("pftcc_gencorr_euclid_grad_only_for_rot_8", _I, [_P, _I, _I, _PD, _I, _PD]),
    ("pftcc_set_dref_fcomp", _I, [_P, _I, _I, _I, _PF, _I]),
    ("pftcc_set_dref_pft", _I, [_P, _I, _PF, _I]),
    ("pftcc_gencorr_cont_grad_cc_for_rot_8", _I, [_P, _I, _I, _PD, _I, _PD, _PD]),
    ("pftcc_gencorr_cont_shift_grad_cc_for_rot_8", _I, [_P, _I, _I, _PD, _I, _PD, _PD]),
    ("pftcc_gencorr_cont_grad_euclid_for_rot_8", _I, [_P, _I, _I, _PD, _I, _PD, _PD]),
    ("pftcc_gencorr_cont_shift_grad_euclid_for_rot_8", _I, [_P, _I, _I, _PD, _I, _PD, _PD]),
    ("pftcc_calc_ptcl_pspec", _I, [_P, _I, _PF]),
    ("pftcc_get_pfromto", _I, [_P, _PI]),
    ("pftcc_assign_pinds", _I, [_P, _PI]),
    ("pftcc_cp_refs", _I, [_P, _P]),
    ("pftcc_gencorr_sigma_contrib", _I, [_P, _I, _I, _PF, _I, _PF, _PF, _PF]),
    ("pftcc_specscore", _I, [_P, _I, _I, _I, _PF, _PF]),
    ("pftcc_calc_roinv_corrmat", _I, [_P, _PF]),
    ("pftcc_gencorrs_batch", _I, [_P, _I, _PI, _I, _PI, _I, _PF, _PF]),
    ("pftcc_search", _I, [_P, _I, _PI, _I, _PI, _I, _PF, _PF, _PI, _PI, _PI]),
    ("pftcc_strategy_srch", _I, [_P, C.POINTER(StrategyCfgC), _I, _PI, _PI, _P, _PF, _PF, _PF, _PI, _PI, _PI, _PI,
                                 _PF, _PI, _PI]),
    ("pftcc_search_dev", _I, [_P, _I, _P, _I, _P, _I, _P, _P, _P, _P, _P]),
    ("pftcc_search_dev_noalloc", _I, [_P, _I, _P, _I, _P, _I, _P, _P, _P, _P, _P]),
    ("pftcc_grad_batch", _I, [_P, _I, _PI, _PI, _PD, _PI, _PD, _PD]),
    ("pftcc_grad_batch_dev", _I, [_P, _I, _P, _P, _P, _P, _P, _P]),
    ("pftcc_prep_matchfilt", _I, [_P, _I, _I, _I]),
    ("pftcc_prep_matchfilt_batch", _I, [_P, _I, _PI, _PI, _PI]),
    ("pftcc_shsrch_grad_batch", _I, [_P, _I, _PI, _PI, _PF, _I, _I, _I, _PI, _PF]),
    ("pftcc_init_polarizer", _I, [_P, _I]),
    ("pftcc_polarize", _I, [_P, _I, _PI, _PF, _I, _I, C.POINTER(C.c_int8), _I]),
    ("pftcc_polarize_dev", _I, [_P, _I, _PI, _P, _I, _I, C.POINTER(C.c_int8), _I]),
    ("pftcc_prep_imgs", _I, [_P, _I, _PF, C.POINTER(PrepCfgC), _PF, C.POINTER(CtfParamsC), _PF]),
    ("pftcc_prep_imgs_dev", _I, [_P, _I, _P, C.POINTER(PrepCfgC), _PF, C.POINTER(CtfParamsC), _P]),
    ("pftcc_search_mfma", _I, [_P, _I, _PI, _I, _PI, _I, _PF, _I, _I, _PF, _PI, _PI, _PI]),
    ("pftcc_gencorrs_mfma_batch", _I, [_P, _I, _PI, _I, _PI, _I, _PF, _I, _PF]),
    ("pftcc_get_timing_channel", _I, [_P, _I, _PD, C.POINTER(C.c_int64)]),
    ("pftcc_set_timing", _I, [_P, _I]),
    ("pftcc_get_timing", _I, [_P, _PD, C.POINTER(C.c_int64)]),
    ("pftcc_get_corr_tile", _I, [_P, _I, _PI, _PI]),
    ("pftcc_magic_pftsz", _I, [_I]),
]

_lib = None


def load_library(path: str = LIBPATH):
    """Loads libpftcc_hip.so and declares every entry point; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise PftccError(f"{path} is missing: build it with python simple3.0_amd/build.py "
                         "(there is no CPU fallback)")
    lib = C.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _fp(a):
    return a.ctypes.data_as(_PF)


def _dp(a):
    return a.ctypes.data_as(_PD)


def _ip(a):
    return a.ctypes.data_as(_PI)


@dataclass
class Params:
    """The params_glob fields polarft_corrcalc reads (simple_parameters.f90:1182-1371)."""
    boxmatch: int
    kfromto: tuple
    kstop: int = 0
    ring2: int = 0
    pftsz: int = 0
    objfun: int = PFTCC_OBJFUN_CC
    match_filt: bool = False
    pssnr: bool = False
    ctf: bool = True
    clsfrcs: bool = False
    nthr: int = 1


class PolarFTCorrCalc:
    """type(polarft_corrcalc) backed by the gfx950 library (one instance per handle)."""

    def __init__(self):
        self._lib = load_library()
        self._h = C.c_void_p()
        self.nrots = self.pftsz = self.nk = 0

    # -- helpers ---------------------------------------------------------------------------
    def _chk(self, rc, what):
        if rc != 0:
            msg = self._lib.pftcc_last_error(self._h) if self._h else b""
            raise PftccError(f"{what}: {msg.decode() if msg else 'error'} (code {rc})")

    def _pft(self, pft):
        a = np.ascontiguousarray(pft, dtype=np.complex64)
        if a.shape != (self.nk, self.pftsz):
            raise PftccError(f"polar FT must have shape {(self.nk, self.pftsz)}, got {a.shape}")
        return a

    # -- constructor / destructor (new :209-422, kill :2234-2279) ----------------------------
    def new(self, nrefs, pfromto, params: Params, ptcl_mask=None, eoarr=None):
        self.kill()
        cfg = Cfg(boxmatch=params.boxmatch, ring2=params.ring2, pftsz=params.pftsz,
                  kfromto=(C.c_int32 * 2)(*params.kfromto),
                  kstop=params.kstop or params.kfromto[1], objfun=params.objfun,
                  match_filt=int(params.match_filt), pssnr=int(params.pssnr), ctf=int(params.ctf),
                  clsfrcs=int(params.clsfrcs), nthr=params.nthr)
        span = pfromto[1] - pfromto[0] + 1
        mask = None
        if ptcl_mask is not None:
            m = np.ascontiguousarray(ptcl_mask, dtype=np.int8)
            assert m.size == span
            mask = m.ctypes.data_as(C.POINTER(C.c_int8))
            self._mask = m
        eo = None
        if eoarr is not None:
            e = np.ascontiguousarray(eoarr, dtype=np.int32)
            assert e.size == span
            eo = _ip(e)
            self._eo = e
        rc = self._lib.pftcc_new(C.byref(cfg), nrefs, pfromto[0], pfromto[1], mask, eo,
                                 C.byref(self._h))
        if rc != 0:
            msg = self._lib.pftcc_last_error(self._h) if self._h else b""
            if self._h:
                self._lib.pftcc_kill(self._h)
                self._h = C.c_void_p()
            raise PftccError(f"new: {msg.decode() if msg else ''} (code {rc})")
        self.params = params
        self.pfromto = tuple(pfromto)
        self.nrefs = nrefs
        pd = np.zeros(3, np.int32)
        self._chk(self._lib.pftcc_get_pdim(self._h, _ip(pd)), "get_pdim")
        self.pftsz = int(pd[0])
        self.kfromto = (int(pd[1]), int(pd[2]))
        self.nk = self.kfromto[1] - self.kfromto[0] + 1
        self.nrots = 2 * self.pftsz
        return self

    def kill(self):
        if getattr(self, "_h", None):
            self._lib.pftcc_kill(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.kill()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # -- setters -----------------------------------------------------------------------------
    def set_ref_pft(self, iref, pft, iseven=True):
        a = self._pft(pft)
        self._chk(self._lib.pftcc_set_ref_pft(self._h, iref, _fp(a.view(np.float32)), int(iseven)),
                  "set_ref_pft")

    def set_ref_pfts(self, iref_first, pfts, iseven=True):
        a = np.ascontiguousarray(pfts, dtype=np.complex64)
        self._chk(self._lib.pftcc_set_ref_pfts(self._h, iref_first, a.shape[0],
                                               _fp(a.view(np.float32)), int(iseven)), "set_ref_pfts")

    def set_ptcl_pft(self, iptcl, pft):
        a = self._pft(pft)
        self._chk(self._lib.pftcc_set_ptcl_pft(self._h, iptcl, _fp(a.view(np.float32))),
                  "set_ptcl_pft")

    def set_ptcl_pfts(self, iptcl_first, pfts):
        a = np.ascontiguousarray(pfts, dtype=np.complex64)
        self._chk(self._lib.pftcc_set_ptcl_pfts(self._h, iptcl_first, a.shape[0],
                                                _fp(a.view(np.float32))), "set_ptcl_pfts")

    def set_ref_pfts_dev(self, iref_first, count, dptr, iseven=True):
        self._chk(self._lib.pftcc_set_ref_pfts_dev(self._h, iref_first, count, C.c_void_p(dptr),
                                                   int(iseven)), "set_ref_pfts_dev")

    def set_ptcl_pfts_dev(self, iptcl_first, count, dptr):
        self._chk(self._lib.pftcc_set_ptcl_pfts_dev(self._h, iptcl_first, count,
                                                    C.c_void_p(dptr)), "set_ptcl_pfts_dev")

    def set_ref_fcomp(self, iref, irot, k, comp, iseven=True):
        self._chk(self._lib.pftcc_set_ref_fcomp(self._h, iref, irot, k, float(np.real(comp)),
                                                float(np.imag(comp)), int(iseven)), "set_ref_fcomp")

    def set_ptcl_fcomp(self, iptcl, irot, k, comp):
        self._chk(self._lib.pftcc_set_ptcl_fcomp(self._h, iptcl, irot, k, float(np.real(comp)),
                                                 float(np.imag(comp))), "set_ptcl_fcomp")

    def cp_even2odd_ref(self, iref):
        self._chk(self._lib.pftcc_cp_even2odd_ref(self._h, iref), "cp_even2odd_ref")

    def cp_even_ref2ptcl(self, iref, iptcl):
        self._chk(self._lib.pftcc_cp_even_ref2ptcl(self._h, iref, iptcl), "cp_even_ref2ptcl")

    def swap_ptclsevenodd(self):
        self._chk(self._lib.pftcc_swap_ptclsevenodd(self._h), "swap_ptclsevenodd")

    def reallocate_ptcls(self, nptcls, pinds):
        p = np.ascontiguousarray(pinds, dtype=np.int32)
        self._chk(self._lib.pftcc_reallocate_ptcls(self._h, nptcls, _ip(p)), "reallocate_ptcls")
        self.pfromto = (int(p.min()), int(p.max()))

    def set_eo(self, iptcl, is_even):
        self._chk(self._lib.pftcc_set_eo(self._h, iptcl, int(is_even)), "set_eo")

    def set_eos(self, eoarr):
        e = np.ascontiguousarray(eoarr, dtype=np.int32)
        self._chk(self._lib.pftcc_set_eos(self._h, _ip(e)), "set_eos")

    def set_ref_optlp(self, iref, optlp):
        o = np.ascontiguousarray(optlp, dtype=np.float32)
        self._chk(self._lib.pftcc_set_ref_optlp(self._h, iref, _fp(o)), "set_ref_optlp")

    def assign_sigma2_noise(self, sigma2):
        """sigma2[iptcl - pfrom, k - kfrom] for the whole pfromto range."""
        s = np.ascontiguousarray(sigma2, dtype=np.float32)
        self._chk(self._lib.pftcc_assign_sigma2_noise(self._h, _fp(s)), "assign_sigma2_noise")

    def zero_refs_beyond_kstop(self):
        self._chk(self._lib.pftcc_zero_refs_beyond_kstop(self._h), "zero_refs_beyond_kstop")

    # -- memoizers ---------------------------------------------------------------------------
    def create_polar_absctfmats(self, ctfparams):
        """ctfparams: sequence (length pto-pfrom+1) of objects with smpd, kv, cs, fraca, dfx,
        dfy, angast[, phshift, phaseplate] attributes."""
        self._chk(self._lib.pftcc_create_polar_absctfmats(self._h, _ctf_array(ctfparams)), "create_polar_absctfmats")

    def memoize_ffts(self):
        self._chk(self._lib.pftcc_memoize_ffts(self._h), "memoize_ffts")

    # -- getters -------------------------------------------------------------------------------
    def _geti(self, fn, what):
        v = C.c_int32()
        self._chk(fn(self._h, C.byref(v)), what)
        return v.value

    def get_nrots(self):
        return self._geti(self._lib.pftcc_get_nrots, "get_nrots")

    def get_pftsz(self):
        return self._geti(self._lib.pftcc_get_pftsz, "get_pftsz")

    def get_box(self):
        return self._geti(self._lib.pftcc_get_box, "get_box")

    def get_nrefs(self):
        return self._geti(self._lib.pftcc_get_nrefs, "get_nrefs")

    def get_nptcls(self):
        return self._geti(self._lib.pftcc_get_nptcls, "get_nptcls")

    def get_pdim(self):
        pd = np.zeros(3, np.int32)
        self._chk(self._lib.pftcc_get_pdim(self._h, _ip(pd)), "get_pdim")
        return pd

    def get_rot(self, roind):
        v = C.c_float()
        self._chk(self._lib.pftcc_get_rot(self._h, roind, C.byref(v)), "get_rot")
        return v.value

    def get_roind(self, rot):
        v = C.c_int32()
        self._chk(self._lib.pftcc_get_roind(self._h, float(rot), C.byref(v)), "get_roind")
        return v.value

    def get_coord(self, irot, k):
        xy = np.zeros(2, np.float32)
        self._chk(self._lib.pftcc_get_coord(self._h, irot, k, _fp(xy)), "get_coord")
        return xy

    def ptcl_iseven(self, iptcl):
        v = C.c_int32()
        self._chk(self._lib.pftcc_ptcl_iseven(self._h, iptcl, C.byref(v)), "ptcl_iseven")
        return bool(v.value)

    def get_ref_pft(self, iref, iseven=True):
        out = np.zeros((self.nk, self.pftsz), np.complex64)
        self._chk(self._lib.pftcc_get_ref_pft(self._h, iref, int(iseven), _fp(out.view(np.float32))),
                  "get_ref_pft")
        return out

    def get_ptcl_pft(self, iptcl):
        out = np.zeros((self.nk, self.pftsz), np.complex64)
        self._chk(self._lib.pftcc_get_ptcl_pft(self._h, iptcl, _fp(out.view(np.float32))),
                  "get_ptcl_pft")
        return out

    def get_ctfmat(self, iptcl):
        out = np.zeros((self.nk, self.pftsz), np.float32)
        self._chk(self._lib.pftcc_get_ctfmat(self._h, iptcl, _fp(out)), "get_ctfmat")
        return out

    def get_memo(self, iptcl):
        out = np.zeros((self.nk, self.pftsz + 1), np.complex64)
        self._chk(self._lib.pftcc_get_memo(self._h, iptcl, _fp(out.view(np.float32))), "get_memo")
        return out

    def get_sqsum_ptcl(self, iptcl):
        v = C.c_float()
        self._chk(self._lib.pftcc_get_sqsum_ptcl(self._h, iptcl, C.byref(v)), "get_sqsum_ptcl")
        return v.value

    # -- calculators -------------------------------------------------------------------------
    def gencorrs(self, iref, iptcl, shvec=None):
        """gencorrs_1 (shvec None) / gencorrs_2: cc[nrots] float32."""
        cc = np.zeros(self.nrots, np.float32)
        sh = None
        if shvec is not None:
            s = np.asarray(shvec, np.float32)
            sh = _fp(s)
        self._chk(self._lib.pftcc_gencorrs(self._h, iref, iptcl, sh, _fp(cc)), "gencorrs")
        return cc

    def gencorr_for_rot(self, iref, iptcl, shvec, irot):
        s = np.asarray(shvec, np.float32)
        v = C.c_float()
        self._chk(self._lib.pftcc_gencorr_for_rot(self._h, iref, iptcl, _fp(s), irot, C.byref(v)),
                  "gencorr_for_rot")
        return v.value

    def gencorr_for_rot_8(self, iref, iptcl, shvec, irot):
        s = np.asarray(shvec, np.float64)
        v = C.c_double()
        self._chk(self._lib.pftcc_gencorr_for_rot_8(self._h, iref, iptcl, _dp(s), irot, C.byref(v)),
                  "gencorr_for_rot_8")
        return v.value

    def gencorr_grad_for_rot_8(self, iref, iptcl, shvec, irot):
        s = np.asarray(shvec, np.float64)
        f = C.c_double()
        g = np.zeros(2, np.float64)
        self._chk(self._lib.pftcc_gencorr_grad_for_rot_8(self._h, iref, iptcl, _dp(s), irot,
                                                         C.byref(f), _dp(g)), "gencorr_grad_for_rot_8")
        return f.value, g

    def gencorr_grad_only_for_rot_8(self, iref, iptcl, shvec, irot):
        s = np.asarray(shvec, np.float64)
        g = np.zeros(2, np.float64)
        self._chk(self._lib.pftcc_gencorr_grad_only_for_rot_8(self._h, iref, iptcl, _dp(s), irot,
                                                              _dp(g)), "gencorr_grad_only_for_rot_8")
        return g

    # objective-specific forms (gencorr_cc_* :1645-1733, :1863-2000; gencorr_euclid_* :2002-2148)
    def _for_rot_8_obj(self, fn, what, iref, iptcl, shvec, irot):
        s = np.asarray(shvec, np.float64)
        v = C.c_double()
        self._chk(fn(self._h, iref, iptcl, _dp(s), irot, C.byref(v)), what)
        return v.value

    def _for_rot_obj(self, fn, what, iref, iptcl, shvec, irot):
        s = np.asarray(shvec, np.float32)
        v = C.c_float()
        self._chk(fn(self._h, iref, iptcl, _fp(s), irot, C.byref(v)), what)
        return v.value

    def _grad_obj(self, fn, what, iref, iptcl, shvec, irot, with_f=True):
        s = np.asarray(shvec, np.float64)
        g = np.zeros(2, np.float64)
        if with_f:
            f = C.c_double()
            self._chk(fn(self._h, iref, iptcl, _dp(s), irot, C.byref(f), _dp(g)), what)
            return f.value, g
        self._chk(fn(self._h, iref, iptcl, _dp(s), irot, _dp(g)), what)
        return g

    def gencorr_cc_for_rot(self, iref, iptcl, shvec, irot):
        return self._for_rot_obj(self._lib.pftcc_gencorr_cc_for_rot, "gencorr_cc_for_rot", iref, iptcl, shvec, irot)

    def gencorr_cc_for_rot_8(self, iref, iptcl, shvec, irot):
        return self._for_rot_8_obj(self._lib.pftcc_gencorr_cc_for_rot_8, "gencorr_cc_for_rot_8", iref, iptcl,
                                   shvec, irot)

    def gencorr_euclid_for_rot(self, iref, iptcl, shvec, irot):
        return self._for_rot_obj(self._lib.pftcc_gencorr_euclid_for_rot, "gencorr_euclid_for_rot", iref, iptcl,
                                 shvec, irot)

    def gencorr_euclid_for_rot_8(self, iref, iptcl, shvec, irot):
        return self._for_rot_8_obj(self._lib.pftcc_gencorr_euclid_for_rot_8, "gencorr_euclid_for_rot_8", iref,
                                   iptcl, shvec, irot)

    def gencorr_cc_grad_for_rot_8(self, iref, iptcl, shvec, irot):
        return self._grad_obj(self._lib.pftcc_gencorr_cc_grad_for_rot_8, "gencorr_cc_grad_for_rot_8", iref, iptcl,
                              shvec, irot)

    def gencorr_cc_grad_only_for_rot_8(self, iref, iptcl, shvec, irot):
        return self._grad_obj(self._lib.pftcc_gencorr_cc_grad_only_for_rot_8, "gencorr_cc_grad_only_for_rot_8",
                              iref, iptcl, shvec, irot, with_f=False)

    def gencorr_euclid_grad_for_rot_8(self, iref, iptcl, shvec, irot):
        return self._grad_obj(self._lib.pftcc_gencorr_euclid_grad_for_rot_8, "gencorr_euclid_grad_for_rot_8", iref,
                              iptcl, shvec, irot)

    def gencorr_euclid_grad_only_for_rot_8(self, iref, iptcl, shvec, irot):
        return self._grad_obj(self._lib.pftcc_gencorr_euclid_grad_only_for_rot_8,
                              "gencorr_euclid_grad_only_for_rot_8", iref, iptcl, shvec, irot, with_f=False)

    # orientation derivatives (set_dref_fcomp :459, gencorr_cont_* :1735-1847, :2050-2069)
    def set_dref_fcomp(self, iref, irot, k, dcomp, iseven=True):
        d = np.ascontiguousarray(dcomp, np.complex64)
        if d.shape != (3,):
            raise PftccError("dcomp must hold 3 complex values")
        self._chk(self._lib.pftcc_set_dref_fcomp(self._h, iref, irot, k, _fp(d), int(bool(iseven))),
                  "set_dref_fcomp")

    def set_dref_pft(self, iref, dpft, iseven=True):
        d = np.ascontiguousarray(dpft, np.complex64)
        self._chk(self._lib.pftcc_set_dref_pft(self._h, iref, _fp(d), int(bool(iseven))), "set_dref_pft")

    def gencorr_cont_grad_cc_for_rot_8(self, iref, iptcl, shvec, irot):
        s = np.asarray(shvec, np.float64)
        cc = C.c_double()
        d = np.zeros(3, np.float64)
        self._chk(self._lib.pftcc_gencorr_cont_grad_cc_for_rot_8(self._h, iref, iptcl, _dp(s), irot, C.byref(cc),
                                                                 _dp(d)), "gencorr_cont_grad_cc_for_rot_8")
        return cc.value, d

    def gencorr_cont_shift_grad_cc_for_rot_8(self, iref, iptcl, shvec, irot):
        s = np.asarray(shvec, np.float64)
        f = C.c_double()
        g = np.zeros(5, np.float64)
        self._chk(self._lib.pftcc_gencorr_cont_shift_grad_cc_for_rot_8(self._h, iref, iptcl, _dp(s), irot,
                                                                       C.byref(f), _dp(g)),
                  "gencorr_cont_shift_grad_cc_for_rot_8")
        return f.value, g

    def gencorr_cont_grad_euclid_for_rot_8(self, iref, iptcl, shvec, irot):
        s = np.asarray(shvec, np.float64)
        cc = C.c_double()
        d = np.zeros(3, np.float64)
        self._chk(self._lib.pftcc_gencorr_cont_grad_euclid_for_rot_8(self._h, iref, iptcl, _dp(s), irot,
                                                                     C.byref(cc), _dp(d)),
                  "gencorr_cont_grad_euclid_for_rot_8")
        return cc.value, d

    def gencorr_cont_shift_grad_euclid_for_rot_8(self, iref, iptcl, shvec, irot):
        s = np.asarray(shvec, np.float64)
        f = C.c_double()
        g = np.zeros(5, np.float64)
        self._chk(self._lib.pftcc_gencorr_cont_shift_grad_euclid_for_rot_8(self._h, iref, iptcl, _dp(s), irot,
                                                                           C.byref(f), _dp(g)),
                  "gencorr_cont_shift_grad_euclid_for_rot_8")
        return f.value, g

    def calc_ptcl_pspec(self, iptcl):
        out = np.zeros(self.nk, np.float32)
        self._chk(self._lib.pftcc_calc_ptcl_pspec(self._h, iptcl, _fp(out)), "calc_ptcl_pspec")
        return out

    def assign_pinds(self):
        """pinds over pfromto: (pfrom, array of 1-based local indices, 0 = not held)."""
        pft = np.zeros(2, np.int32)
        self._chk(self._lib.pftcc_get_pfromto(self._h, _ip(pft)), "get_pfromto")
        out = np.zeros(int(pft[1] - pft[0] + 1), np.int32)
        self._chk(self._lib.pftcc_assign_pinds(self._h, _ip(out)), "assign_pinds")
        return int(pft[0]), out

    def cp_refs(self, other):
        self._chk(self._lib.pftcc_cp_refs(self._h, other.handle), "cp_refs")

    def gencorr_sigma_contrib(self, iref, iptcl, shvec, irot):
        s = np.asarray(shvec, np.float32)
        sc = np.zeros(self.nk, np.float32)
        ps = np.zeros(self.nk, np.float32)
        rs = np.zeros(self.nk, np.float32)
        self._chk(self._lib.pftcc_gencorr_sigma_contrib(self._h, iref, iptcl, _fp(s), irot, _fp(sc),
                                                        _fp(ps), _fp(rs)), "gencorr_sigma_contrib")
        return sc, ps, rs

    def specscore(self, iref, iptcl, irot, shvec=None):
        v = C.c_float()
        sh = None
        if shvec is not None:
            s = np.asarray(shvec, np.float32)
            sh = _fp(s)
        self._chk(self._lib.pftcc_specscore(self._h, iref, iptcl, irot, sh, C.byref(v)), "specscore")
        return v.value

    def calc_roinv_corrmat(self):
        n = self.get_nptcls()
        out = np.zeros((n, n), np.float32)
        self._chk(self._lib.pftcc_calc_roinv_corrmat(self._h, _fp(out)), "calc_roinv_corrmat")
        return out.T   # Fortran order -> corrmat[iref-1, iptcl-1]

    # -- batched fast paths ------------------------------------------------------------------
    def gencorrs_batch(self, iptcls, irefs=None, shifts=None):
        p = np.ascontiguousarray(iptcls, np.int32)
        r = None if irefs is None else np.ascontiguousarray(irefs, np.int32)
        nr = self.nrefs if r is None else r.size
        ns = 0 if shifts is None else len(shifts)
        s = None if shifts is None else np.ascontiguousarray(shifts, np.float32).reshape(-1, 2)
        out = np.zeros((p.size, nr, max(ns, 1), self.nrots), np.float32)
        self._chk(self._lib.pftcc_gencorrs_batch(self._h, p.size, _ip(p), nr,
                                                 None if r is None else _ip(r), ns,
                                                 None if s is None else _fp(s), _fp(out)),
                  "gencorrs_batch")
        return out

    def search(self, iptcls, irefs=None, shifts=None):
        p = np.ascontiguousarray(iptcls, np.int32)
        r = None if irefs is None else np.ascontiguousarray(irefs, np.int32)
        nr = self.nrefs if r is None else r.size
        ns = 0 if shifts is None else len(shifts)
        s = None if shifts is None else np.ascontiguousarray(shifts, np.float32).reshape(-1, 2)
        cc = np.zeros(p.size, np.float32)
        ir = np.zeros(p.size, np.int32)
        ish = np.zeros(p.size, np.int32)
        rot = np.zeros(p.size, np.int32)
        self._chk(self._lib.pftcc_search(self._h, p.size, _ip(p), nr, None if r is None else _ip(r),
                                         ns, None if s is None else _fp(s), _fp(cc), _ip(ir),
                                         _ip(ish), _ip(rot)), "search")
        return cc, ir, ish, rot

    def strategy_srch(self, kind, iptcls, srch_order=None, ref_ok=None, prev_corr=None, rnd=None,
                      stoch_bound=0.0, ntop=0, ref_corrs=False, szsn=0):
        """pftcc_strategy_srch: returns a dict of best_corr, best_ref, best_rot, nrefs_eval,
        found_better [np], plus ref_corrs/ref_rots [np][nrefs] (ref_corrs=True) and sorted_refs
        [np][ntop] (3-D kinds, ntop > 0)."""
        p = np.ascontiguousarray(iptcls, np.int32)
        n = p.size
        o = None if srch_order is None else np.ascontiguousarray(srch_order, np.int32).reshape(n, -1)
        ok = None if ref_ok is None else np.ascontiguousarray(np.asarray(ref_ok) != 0, np.int8)
        pc = None if prev_corr is None else np.ascontiguousarray(prev_corr, np.float32)
        u = None if rnd is None else np.ascontiguousarray(rnd, np.float32)
        out = {"best_corr": np.zeros(n, np.float32), "best_ref": np.zeros(n, np.int32),
               "best_rot": np.zeros(n, np.int32), "nrefs_eval": np.zeros(n, np.int32),
               "found_better": np.zeros(n, np.int32)}
        rc = rr = so = None
        if ref_corrs:
            rc = out["ref_corrs"] = np.zeros((n, self.nrefs), np.float32)
            rr = out["ref_rots"] = np.zeros((n, self.nrefs), np.int32)
        if kind in (SRCH_GREEDY3D, SRCH_SHC3D, SRCH_SNHC3D) and ntop > 0:
            so = out["sorted_refs"] = np.zeros((n, ntop), np.int32)
        cfg = StrategyCfgC(kind, stoch_bound, ntop, szsn)
        self._chk(self._lib.pftcc_strategy_srch(
            self._h, C.byref(cfg), n, _ip(p), None if o is None else _ip(o),
            None if ok is None else ok.ctypes.data_as(C.c_void_p), None if pc is None else _fp(pc),
            None if u is None else _fp(u), _fp(out["best_corr"]), _ip(out["best_ref"]), _ip(out["best_rot"]),
            _ip(out["nrefs_eval"]), _ip(out["found_better"]), None if rc is None else _fp(rc),
            None if rr is None else _ip(rr), None if so is None else _ip(so)), "strategy_srch")
        return out

    def search_dev(self, np_, d_iptcls, nr, d_irefs, ns, d_shifts, d_cc, d_iref, d_ishift, d_irot):
        self._chk(self._lib.pftcc_search_dev(self._h, np_, C.c_void_p(d_iptcls), nr,
                                             C.c_void_p(d_irefs) if d_irefs else None, ns,
                                             C.c_void_p(d_shifts) if d_shifts else None,
                                             C.c_void_p(d_cc), C.c_void_p(d_iref),
                                             C.c_void_p(d_ishift), C.c_void_p(d_irot)), "search_dev")

    def search_dev_noalloc(self, np_, d_iptcls, nr, d_irefs, ns, d_shifts, d_cc, d_iref, d_ishift, d_irot):
        """pftcc_search_dev without device allocation (static per-GPU scratch of 32 MiB)."""
        self._chk(self._lib.pftcc_search_dev_noalloc(self._h, np_, C.c_void_p(d_iptcls), nr,
                                                     C.c_void_p(d_irefs) if d_irefs else None, ns,
                                                     C.c_void_p(d_shifts) if d_shifts else None,
                                                     C.c_void_p(d_cc), C.c_void_p(d_iref),
                                                     C.c_void_p(d_ishift), C.c_void_p(d_irot)),
                  "search_dev_noalloc")

    def grad_batch(self, irefs, iptcls, shvecs, irots, with_grad=True):
        r = np.ascontiguousarray(irefs, np.int32)
        p = np.ascontiguousarray(iptcls, np.int32)
        s = np.ascontiguousarray(shvecs, np.float64).reshape(-1, 2)
        o = np.ascontiguousarray(irots, np.int32)
        f = np.zeros(r.size, np.float64)
        g = np.zeros((r.size, 3), np.float64) if with_grad else None
        self._chk(self._lib.pftcc_grad_batch(self._h, r.size, _ip(r), _ip(p), _dp(s), _ip(o), _dp(f),
                                             None if g is None else _dp(g)), "grad_batch")
        return (f, g) if with_grad else f

    def grad_batch_dev(self, npairs, d_irefs, d_iptcls, d_shvecs, d_irots, d_f, d_grad):
        self._chk(self._lib.pftcc_grad_batch_dev(self._h, npairs, C.c_void_p(d_irefs),
                                                 C.c_void_p(d_iptcls), C.c_void_p(d_shvecs),
                                                 C.c_void_p(d_irots), C.c_void_p(d_f),
                                                 C.c_void_p(d_grad) if d_grad else None),
                  "grad_batch_dev")

    def prep_matchfilt(self, iptcl, iref, irot):
        """prep_matchfilt (:987-1095): match-filter particle iptcl against iref at irot (iref 0:
        memoize only)."""
        self._chk(self._lib.pftcc_prep_matchfilt(self._h, int(iptcl), int(iref), int(irot)), "prep_matchfilt")

    def prep_matchfilt_batch(self, iptcls, irefs, irots):
        p = np.ascontiguousarray(iptcls, np.int32)
        r = np.ascontiguousarray(irefs, np.int32)
        o = np.ascontiguousarray(irots, np.int32)
        self._chk(self._lib.pftcc_prep_matchfilt_batch(self._h, p.size, _ip(p), _ip(r), _ip(o)),
                  "prep_matchfilt_batch")

    def shsrch_grad_batch(self, irefs, iptcls, trs=None, lims=None, maxits=60, opt_angle=True, irots=None,
                          coarse_init=False):
        """pftcc_shsrch_grad%minimize for pairs (irefs[i], iptcls[i]): returns (cxy [n][3], irot [n]),
        irot 0 where no better solution than the start was found.  lims: (xlo, ylo, xhi, yhi), or
        +-trs in both directions."""
        r = np.ascontiguousarray(np.atleast_1d(irefs), np.int32)
        p = np.ascontiguousarray(np.atleast_1d(iptcls), np.int32)
        if lims is None:
            lims = (-trs, -trs, trs, trs)
        lm = np.ascontiguousarray(lims, np.float32)
        rot = np.zeros(r.size, np.int32) if irots is None else np.array(np.atleast_1d(irots), np.int32)
        cxy = np.zeros((r.size, 3), np.float32)
        self._chk(self._lib.pftcc_shsrch_grad_batch(self._h, r.size, _ip(r), _ip(p), _fp(lm), int(maxits),
                                                    1 if opt_angle else 0, 1 if coarse_init else 0, _ip(rot),
                                                    _fp(cxy)), "shsrch_grad_batch")
        return cxy, rot

    def init_polarizer(self, griddev=False):
        """init_polarizer (simple_polarizer.f90:41-92): Kaiser-Bessel gridding tables."""
        self._chk(self._lib.pftcc_init_polarizer(self._h, 1 if griddev else 0), "init_polarizer")

    def polarize(self, img_inds, cmats, isptcl, iseven=True, mask=None):
        """polarize (simple_polarizer.f90:120-151) of a batch: cmats [count][box][box/2+1] complex64
        (the image's cmat, Fortran order), img_inds 1-based particle or reference ids, mask the
        reference's 1-based band mask (index k-1 for shell k) or None."""
        ids = np.ascontiguousarray(np.atleast_1d(img_inds), np.int32)
        c = np.ascontiguousarray(cmats, np.complex64)
        if c.ndim == 2:
            c = c[None]
        m = None if mask is None else np.ascontiguousarray(mask, np.int8)
        self._chk(self._lib.pftcc_polarize(self._h, ids.size, _ip(ids), c.ctypes.data_as(_PF), 1 if isptcl else 0,
                                           1 if iseven else 0,
                                           None if m is None else m.ctypes.data_as(C.POINTER(C.c_int8)),
                                           0 if m is None else m.size), "polarize")

    def polarize_dev(self, img_inds, d_cmats, isptcl, iseven=True, mask=None):
        ids = np.ascontiguousarray(np.atleast_1d(img_inds), np.int32)
        m = None if mask is None else np.ascontiguousarray(mask, np.int8)
        self._chk(self._lib.pftcc_polarize_dev(self._h, ids.size, _ip(ids), C.c_void_p(d_cmats), 1 if isptcl else 0,
                                               1 if iseven else 0,
                                               None if m is None else m.ctypes.data_as(C.POINTER(C.c_int8)),
                                               0 if m is None else m.size), "polarize_dev")

    @staticmethod
    def _prep_cfg(box, boxmatch, msk, inner, width, focusmsk, ctfflag, griddev):
        return PrepCfgC(int(box), int(boxmatch), float(msk), float(inner), float(width), float(focusmsk),
                        int(ctfflag), 1 if griddev else 0)

    def prep_imgs(self, imgs, shifts, box, msk, ctfflag=0, ctfparams=None, inner=0.0, width=10.0, focusmsk=0.0,
                  griddev=False, boxmatch=None):
        """prepimg4align (simple_strategy2D3D_common.f90:316-366) of a batch of real images
        imgs [count][box][box] (float32, Fortran rmat order: x fastest); shifts [count][2] the
        particles' (x, y); ctfflag 1 (CTFFLAG_YES) phase-flips with ctfparams (one per image).
        Returns the transforms [count][boxmatch][boxmatch/2+1] complex64 for polarize."""
        im = np.ascontiguousarray(imgs, np.float32)
        if im.ndim == 2:
            im = im[None]
        n = im.shape[0]
        sh = np.ascontiguousarray(np.asarray(shifts, np.float32).reshape(n, 2))
        M = self.params.boxmatch if boxmatch is None else int(boxmatch)
        out = np.zeros((n, M, M // 2 + 1), np.complex64)
        cfg = self._prep_cfg(box, M, msk, inner, width, focusmsk, ctfflag, griddev)
        arr = _ctf_array(ctfparams) if ctfparams is not None else None
        self._chk(self._lib.pftcc_prep_imgs(self._h, n, im.ctypes.data_as(_PF), C.byref(cfg), _fp(sh), arr,
                                            out.ctypes.data_as(_PF)), "prep_imgs")
        return out

    def prep_imgs_dev(self, count, d_imgs, shifts, box, msk, d_cmats, ctfflag=0, ctfparams=None, inner=0.0,
                      width=10.0, focusmsk=0.0, griddev=False):
        """prep_imgs on device-resident images and transforms (addresses d_imgs, d_cmats)."""
        sh = np.ascontiguousarray(np.asarray(shifts, np.float32).reshape(count, 2))
        cfg = self._prep_cfg(box, self.params.boxmatch, msk, inner, width, focusmsk, ctfflag, griddev)
        arr = _ctf_array(ctfparams) if ctfparams is not None else None
        self._chk(self._lib.pftcc_prep_imgs_dev(self._h, int(count), C.c_void_p(d_imgs), C.byref(cfg), _fp(sh), arr,
                                                C.c_void_p(d_cmats)), "prep_imgs_dev")

    def search_mfma(self, iptcls, irefs=None, prec=PFTCC_PREC_SPLIT_BF16, ntop=0, shifts=None):
        """Precision form of search(): (best_cc, best_iref, best_irot), plus best_ishift (0-based)
        as a fourth value when shifts [ns, 2] are given.

        ntop > 0 rescores the ntop best (reference, shift) rows per particle exactly in FP64."""
        p = np.ascontiguousarray(iptcls, np.int32)
        r = None if irefs is None else np.ascontiguousarray(irefs, np.int32)
        nr = self.nrefs if r is None else r.size
        sh = None if shifts is None else np.ascontiguousarray(shifts, np.float32).reshape(-1, 2)
        ns = 0 if sh is None else sh.shape[0]
        cc = np.zeros(p.size, np.float32)
        ir = np.zeros(p.size, np.int32)
        ish = np.zeros(p.size, np.int32)
        rot = np.zeros(p.size, np.int32)
        self._chk(self._lib.pftcc_search_mfma(self._h, p.size, _ip(p), nr, None if r is None else _ip(r), ns,
                                              None if sh is None else _fp(sh), prec, ntop, _fp(cc), _ip(ir),
                                              _ip(ish), _ip(rot)), "search_mfma")
        return (cc, ir, ish, rot) if sh is not None else (cc, ir, rot)

    def gencorrs_mfma_batch(self, iptcls, irefs=None, prec=PFTCC_PREC_SPLIT_BF16, shifts=None):
        """cc[np, nr, nrots] (or [np, nr, ns, nrots] with shifts) of the precision form."""
        p = np.ascontiguousarray(iptcls, np.int32)
        r = None if irefs is None else np.ascontiguousarray(irefs, np.int32)
        nr = self.nrefs if r is None else r.size
        sh = None if shifts is None else np.ascontiguousarray(shifts, np.float32).reshape(-1, 2)
        ns = 0 if sh is None else sh.shape[0]
        shape = (p.size, nr, ns, self.nrots) if sh is not None else (p.size, nr, self.nrots)
        out = np.zeros(shape, np.float32)
        self._chk(self._lib.pftcc_gencorrs_mfma_batch(self._h, p.size, _ip(p), nr, None if r is None else _ip(r), ns,
                                                      None if sh is None else _fp(sh), prec, _fp(out)),
                  "gencorrs_mfma_batch")
        return out

    def get_timing_channel(self, ch):
        ms = C.c_double()
        n = C.c_int64()
        self._chk(self._lib.pftcc_get_timing_channel(self._h, ch, C.byref(ms), C.byref(n)), "get_timing_channel")
        return ms.value, n.value

    def set_timing(self, enable=True):
        self._chk(self._lib.pftcc_set_timing(self._h, int(enable)), "set_timing")

    def get_timing(self):
        """(summed device ms of the correlation kernel launches, number of launches)."""
        ms = C.c_double()
        n = C.c_int64()
        self._chk(self._lib.pftcc_get_timing(self._h, C.byref(ms), C.byref(n)), "get_timing")
        return ms.value, n.value

    def get_corr_tile(self, nshifts=0):
        """(pairs per work-group of the plain kernel, staged-reference variant on) for an
        exhaustive search with nshifts shifts."""
        t = C.c_int32()
        st = C.c_int32()
        self._chk(self._lib.pftcc_get_corr_tile(self._h, int(nshifts), C.byref(t), C.byref(st)), "get_corr_tile")
        return t.value, bool(st.value)

    def set_stream(self, stream_ptr):
        self._chk(self._lib.pftcc_set_stream(self._h, C.c_void_p(stream_ptr)), "set_stream")

    def sync(self):
        self._chk(self._lib.pftcc_sync(self._h), "sync")


def magic_pftsz(ring2: int) -> int:
    return int(load_library().pftcc_magic_pftsz(ring2))
