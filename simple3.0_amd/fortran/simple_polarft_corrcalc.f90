! simple_polarft_corrcalc - drop-in Fortran module for SIMPLE's polar-FT correlation search,
! backed by the gfx950 HIP library (include/pftcc.h) through ISO_C_BINDING.
!
! Keeps the reference type and procedure names (src/main/simple_polarft_corrcalc.f90:68-198):
! new, set_ref_pft, set_ptcl_pft, memoize_ffts, create_polar_absctfmats, gencorrs (generic),
! gencorr_for_rot(_8), gencorr_grad_for_rot_8, gencorr_grad_only_for_rot_8, getters, kill, and
! the pftcc_glob pointer.  Indices keep the reference's meaning (iref 1..nrefs, iptcl global id,
! irot 1..nrots).  Errors stop the program with the library's message, as THROW_HARD does
! (src/fileio/simple_error.f90:8-31).
!
! Two interface differences, both because the module no longer reads global state:
!  * new() takes the params_glob fields it needs as a pftcc_params argument;
!  * create_polar_absctfmats() takes the per-particle ctfparams array the reference extracts
!    from sp_project%get_ctfparams (:924-953).
module simple_polarft_corrcalc
use, intrinsic :: iso_c_binding
implicit none

public :: polarft_corrcalc, pftcc_glob, pftcc_params, ctfparams, prep_cfg, OBJFUN_CC, OBJFUN_EUCLID
public :: strategy_cfg, SRCH_GREEDY2D, SRCH_SNHC2D, SRCH_GREEDY3D, SRCH_SHC3D, SRCH_SNHC3D
public :: PREC_BF16, PREC_SPLIT_BF16, PREC_FP16, PREC_SPLIT_FP16, PREC_FP8
private

integer, parameter :: sp = c_float, dp = c_double
integer, parameter :: OBJFUN_CC = 0, OBJFUN_EUCLID = 1
integer, parameter :: PREC_BF16 = 1, PREC_SPLIT_BF16 = 3, PREC_FP16 = 2, PREC_SPLIT_FP16 = 4, PREC_FP8 = 8   ! precision form

! params_glob fields read by polarft_corrcalc (simple_parameters.f90:1182-1371)
type :: pftcc_params
    integer :: boxmatch     = 0
    integer :: ring2        = 0
    integer :: pftsz        = 0          ! explicit ring length (0: magic_pftsz(ring2))
    integer :: kfromto(2)   = [2, 2]
    integer :: kstop        = 0
    integer :: cc_objfun    = OBJFUN_CC
    logical :: l_match_filt = .false.
    logical :: l_pssnr      = .false.
    logical :: with_ctf     = .true.      ! ctf /= 'no'
    logical :: l_clsfrcs    = .false.
    integer :: nthr         = 1
end type pftcc_params

type, bind(C) :: pftcc_cfg_c
    integer(c_int32_t) :: boxmatch, ring2, pftsz, kfromto(2), kstop, objfun, match_filt, pssnr, ctf, clsfrcs, nthr
end type pftcc_cfg_c

! pftcc_ctfparams
type, bind(C) :: ctfparams
    real(c_float)      :: smpd = 0., kv = 300., cs = 2.7, fraca = 0.1
    real(c_float)      :: dfx = 0., dfy = 0., angast = 0., phshift = 0.
    integer(c_int32_t) :: l_phaseplate = 0
end type ctfparams

! pftcc_prep_cfg (prepimg4align settings, params_glob fields of simple_strategy2D3D_common.f90:316-366)
type, bind(C) :: prep_cfg
    integer(c_int32_t) :: box = 0, boxmatch = 0
    real(c_float)      :: msk = 0., inner = 0., width = 10., focusmsk = 0.
    integer(c_int32_t) :: ctfflag = 0, griddev = 0
end type prep_cfg

! pftcc_strategy_cfg: which reference-loop strategy, snhc 2-D's spec%stoch_bound, the length of
! each sorted reference list (3-D kinds), snhc 3-D's spec%szsn
integer, parameter :: SRCH_GREEDY2D = 0, SRCH_SNHC2D = 1, SRCH_GREEDY3D = 2, SRCH_SHC3D = 3, SRCH_SNHC3D = 4
type, bind(C) :: strategy_cfg
    integer(c_int32_t) :: kind = SRCH_GREEDY2D
    real(c_float)      :: stoch_bound = 0.
    integer(c_int32_t) :: ntop = 0
    integer(c_int32_t) :: szsn = 0
end type strategy_cfg

type :: polarft_corrcalc
    private
    type(c_ptr) :: h = c_null_ptr
    integer     :: nrots = 0, pftsz = 0, kfromto(2) = 0, nrefs = 0, nptcls = 0
    logical     :: existence = .false.
  contains
    procedure          :: new
    procedure          :: set_ref_pft
    procedure          :: set_ptcl_pft
    procedure          :: set_ref_fcomp
    procedure          :: set_ptcl_fcomp
    procedure          :: cp_even2odd_ref
    procedure          :: cp_refs
    procedure          :: set_dref_fcomp
    procedure          :: cp_even_ref2ptcl
    procedure          :: swap_ptclsevenodd
    procedure          :: reallocate_ptcls
    procedure          :: set_eo
    procedure          :: set_eos
    procedure          :: set_ref_optlp
    procedure          :: assign_sigma2_noise
    procedure          :: zero_refs_beyond_kstop
    procedure          :: get_nrots
    procedure          :: get_pdim
    procedure          :: get_pftsz
    procedure          :: get_box
    procedure          :: get_rot
    procedure          :: get_roind
    procedure          :: get_coord
    procedure          :: get_ref_pft
    procedure          :: assign_pinds
    procedure          :: print
    procedure          :: get_ptcl_pft
    procedure          :: get_nrefs
    procedure          :: get_nptcls
    procedure          :: ptcl_iseven
    procedure          :: exists
    procedure          :: create_polar_absctfmats
    procedure          :: memoize_ffts
    procedure          :: calc_ptcl_pspec
    procedure, private :: gencorrs_1
    procedure, private :: gencorrs_2
    generic            :: gencorrs => gencorrs_1, gencorrs_2
    procedure          :: gencorr_for_rot
    procedure          :: gencorr_for_rot_8
    procedure          :: gencorr_grad_for_rot_8
    procedure          :: gencorr_grad_only_for_rot_8
    procedure          :: gencorr_sigma_contrib
    procedure          :: gencorr_cc_for_rot
    procedure          :: gencorr_cc_for_rot_8
    procedure          :: gencorr_cc_grad_for_rot_8
    procedure          :: gencorr_cc_grad_only_for_rot_8
    procedure          :: gencorr_euclid_for_rot
    procedure          :: gencorr_euclid_for_rot_8
    procedure          :: gencorr_euclid_grad_for_rot_8
    procedure          :: gencorr_euclid_grad_only_for_rot_8
    procedure          :: gencorr_cont_grad_cc_for_rot_8
    procedure          :: gencorr_cont_shift_grad_cc_for_rot_8
    procedure          :: gencorr_cont_grad_euclid_for_rot_8
    procedure          :: gencorr_cont_shift_grad_euclid_for_rot_8
    procedure, private :: specscore_1
    procedure, private :: specscore_2
    generic            :: specscore => specscore_1, specscore_2
    procedure          :: calc_roinv_corrmat
    procedure          :: prep_matchfilt
    procedure          :: search
    procedure          :: search_mfma
    procedure          :: strategy_srch
    procedure          :: grad_batch
    procedure          :: shsrch_grad_batch
    procedure          :: init_polarizer
    procedure          :: polarize
    procedure          :: prep_imgs
    procedure          :: kill
end type polarft_corrcalc

class(polarft_corrcalc), pointer :: pftcc_glob => null()

interface
    integer(c_int) function c_new(cfg, nrefs, pfrom, pto, mask, eo, h) bind(C, name='pftcc_new')
        import :: c_int, c_int32_t, c_ptr, pftcc_cfg_c
        type(pftcc_cfg_c),         intent(in)  :: cfg
        integer(c_int32_t), value              :: nrefs, pfrom, pto
        type(c_ptr),        value              :: mask, eo
        type(c_ptr),               intent(out) :: h
    end function
    integer(c_int) function c_kill(h) bind(C, name='pftcc_kill')
        import :: c_int, c_ptr
        type(c_ptr), value :: h
    end function
    type(c_ptr) function c_last_error(h) bind(C, name='pftcc_last_error')
        import :: c_ptr
        type(c_ptr), value :: h
    end function
    integer(c_int) function c_set_ref_pft(h, iref, pft, iseven) bind(C, name='pftcc_set_ref_pft')
        import :: c_int, c_int32_t, c_ptr, c_float_complex
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iseven
        complex(c_float_complex), intent(in) :: pft(*)
    end function
    integer(c_int) function c_set_ptcl_pft(h, iptcl, pft) bind(C, name='pftcc_set_ptcl_pft')
        import :: c_int, c_int32_t, c_ptr, c_float_complex
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iptcl
        complex(c_float_complex), intent(in) :: pft(*)
    end function
    integer(c_int) function c_set_ref_fcomp(h, iref, irot, k, re, im, iseven) bind(C, name='pftcc_set_ref_fcomp')
        import :: c_int, c_int32_t, c_ptr, c_float
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, irot, k, iseven
        real(c_float), value :: re, im
    end function
    integer(c_int) function c_set_ptcl_fcomp(h, iptcl, irot, k, re, im) bind(C, name='pftcc_set_ptcl_fcomp')
        import :: c_int, c_int32_t, c_ptr, c_float
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iptcl, irot, k
        real(c_float), value :: re, im
    end function
    integer(c_int) function c_cp_even2odd_ref(h, iref) bind(C, name='pftcc_cp_even2odd_ref')
        import :: c_int, c_int32_t, c_ptr
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref
    end function
    integer(c_int) function c_cp_even_ref2ptcl(h, iref, iptcl) bind(C, name='pftcc_cp_even_ref2ptcl')
        import :: c_int, c_int32_t, c_ptr
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iptcl
    end function
    integer(c_int) function c_swap_ptclsevenodd(h) bind(C, name='pftcc_swap_ptclsevenodd')
        import :: c_int, c_ptr
        type(c_ptr), value :: h
    end function
    integer(c_int) function c_reallocate_ptcls(h, nptcls, pinds) bind(C, name='pftcc_reallocate_ptcls')
        import :: c_int, c_int32_t, c_ptr
        type(c_ptr), value :: h
        integer(c_int32_t), value :: nptcls
        integer(c_int32_t), intent(in) :: pinds(*)
    end function
    integer(c_int) function c_set_eo(h, iptcl, is_even) bind(C, name='pftcc_set_eo')
        import :: c_int, c_int32_t, c_ptr
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iptcl, is_even
    end function
    integer(c_int) function c_set_eos(h, eoarr) bind(C, name='pftcc_set_eos')
        import :: c_int, c_int32_t, c_ptr
        type(c_ptr), value :: h
        integer(c_int32_t), intent(in) :: eoarr(*)
    end function
    integer(c_int) function c_set_ref_optlp(h, iref, optlp) bind(C, name='pftcc_set_ref_optlp')
        import :: c_int, c_int32_t, c_ptr, c_float
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref
        real(c_float), intent(in) :: optlp(*)
    end function
    integer(c_int) function c_assign_sigma2_noise(h, sigma2) bind(C, name='pftcc_assign_sigma2_noise')
        import :: c_int, c_ptr, c_float
        type(c_ptr), value :: h
        real(c_float), intent(in) :: sigma2(*)
    end function
    integer(c_int) function c_zero_refs_beyond_kstop(h) bind(C, name='pftcc_zero_refs_beyond_kstop')
        import :: c_int, c_ptr
        type(c_ptr), value :: h
    end function
    integer(c_int) function c_create_polar_absctfmats(h, params) bind(C, name='pftcc_create_polar_absctfmats')
        import :: c_int, c_ptr, ctfparams
        type(c_ptr), value :: h
        type(ctfparams), intent(in) :: params(*)
    end function
    integer(c_int) function c_memoize_ffts(h) bind(C, name='pftcc_memoize_ffts')
        import :: c_int, c_ptr
        type(c_ptr), value :: h
    end function
    integer(c_int) function c_get_nrots(h, v) bind(C, name='pftcc_get_nrots')
        import :: c_int, c_int32_t, c_ptr
        type(c_ptr), value :: h
        integer(c_int32_t), intent(out) :: v
    end function
    integer(c_int) function c_get_pftsz(h, v) bind(C, name='pftcc_get_pftsz')
        import :: c_int, c_int32_t, c_ptr
        type(c_ptr), value :: h
        integer(c_int32_t), intent(out) :: v
    end function
    integer(c_int) function c_get_box(h, v) bind(C, name='pftcc_get_box')
        import :: c_int, c_int32_t, c_ptr
        type(c_ptr), value :: h
        integer(c_int32_t), intent(out) :: v
    end function
    integer(c_int) function c_get_nrefs(h, v) bind(C, name='pftcc_get_nrefs')
        import :: c_int, c_int32_t, c_ptr
        type(c_ptr), value :: h
        integer(c_int32_t), intent(out) :: v
    end function
    integer(c_int) function c_get_nptcls(h, v) bind(C, name='pftcc_get_nptcls')
        import :: c_int, c_int32_t, c_ptr
        type(c_ptr), value :: h
        integer(c_int32_t), intent(out) :: v
    end function
    integer(c_int) function c_get_pdim(h, pdim) bind(C, name='pftcc_get_pdim')
        import :: c_int, c_int32_t, c_ptr
        type(c_ptr), value :: h
        integer(c_int32_t), intent(out) :: pdim(3)
    end function
    integer(c_int) function c_get_rot(h, roind, rot) bind(C, name='pftcc_get_rot')
        import :: c_int, c_int32_t, c_ptr, c_float
        type(c_ptr), value :: h
        integer(c_int32_t), value :: roind
        real(c_float), intent(out) :: rot
    end function
    integer(c_int) function c_get_roind(h, rot, ind) bind(C, name='pftcc_get_roind')
        import :: c_int, c_int32_t, c_ptr, c_float
        type(c_ptr), value :: h
        real(c_float), value :: rot
        integer(c_int32_t), intent(out) :: ind
    end function
    integer(c_int) function c_get_coord(h, irot, k, xy) bind(C, name='pftcc_get_coord')
        import :: c_int, c_int32_t, c_ptr, c_float
        type(c_ptr), value :: h
        integer(c_int32_t), value :: irot, k
        real(c_float), intent(out) :: xy(2)
    end function
    integer(c_int) function c_ptcl_iseven(h, iptcl, v) bind(C, name='pftcc_ptcl_iseven')
        import :: c_int, c_int32_t, c_ptr
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iptcl
        integer(c_int32_t), intent(out) :: v
    end function
    integer(c_int) function c_get_ref_pft(h, iref, iseven, pft) bind(C, name='pftcc_get_ref_pft')
        import :: c_int, c_int32_t, c_ptr, c_float_complex
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iseven
        complex(c_float_complex), intent(out) :: pft(*)
    end function
    integer(c_int) function c_get_ptcl_pft(h, iptcl, pft) bind(C, name='pftcc_get_ptcl_pft')
        import :: c_int, c_int32_t, c_ptr, c_float_complex
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iptcl
        complex(c_float_complex), intent(out) :: pft(*)
    end function
    integer(c_int) function c_gencorrs(h, iref, iptcl, shvec, cc) bind(C, name='pftcc_gencorrs')
        import :: c_int, c_int32_t, c_ptr, c_float
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iptcl
        type(c_ptr), value :: shvec
        real(c_float), intent(out) :: cc(*)
    end function
    integer(c_int) function c_gencorr_for_rot(h, iref, iptcl, shvec, irot, f) bind(C, name='pftcc_gencorr_for_rot')
        import :: c_int, c_int32_t, c_ptr, c_float
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iptcl, irot
        real(c_float), intent(in) :: shvec(2)
        real(c_float), intent(out) :: f
    end function
    integer(c_int) function c_gencorr_for_rot_8(h, iref, iptcl, shvec, irot, f) bind(C, name='pftcc_gencorr_for_rot_8')
        import :: c_int, c_int32_t, c_ptr, c_double
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iptcl, irot
        real(c_double), intent(in) :: shvec(2)
        real(c_double), intent(out) :: f
    end function
    integer(c_int) function c_gencorr_grad_for_rot_8(h, iref, iptcl, shvec, irot, f, grad) &
        &bind(C, name='pftcc_gencorr_grad_for_rot_8')
        import :: c_int, c_int32_t, c_ptr, c_double
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iptcl, irot
        real(c_double), intent(in) :: shvec(2)
        real(c_double), intent(out) :: f, grad(2)
    end function
    integer(c_int) function c_gencorr_grad_only_for_rot_8(h, iref, iptcl, shvec, irot, grad) &
        &bind(C, name='pftcc_gencorr_grad_only_for_rot_8')
        import :: c_int, c_int32_t, c_ptr, c_double
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iptcl, irot
        real(c_double), intent(in) :: shvec(2)
        real(c_double), intent(out) :: grad(2)
    end function
    integer(c_int) function c_gencorr_sigma_contrib(h, iref, iptcl, shvec, irot, sc, ps, rs) &
        &bind(C, name='pftcc_gencorr_sigma_contrib')
        import :: c_int, c_int32_t, c_ptr, c_float
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iptcl, irot
        real(c_float), intent(in) :: shvec(2)
        real(c_float), intent(out) :: sc(*)
        type(c_ptr), value :: ps, rs
    end function
    ! objective-specific single-rotation forms (gencorr_cc_* / gencorr_euclid_*)
    integer(c_int) function c_cc_for_rot(h, iref, iptcl, shvec, irot, f) bind(C, name='pftcc_gencorr_cc_for_rot')
        import :: c_int, c_int32_t, c_ptr, c_float
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iptcl, irot
        real(c_float), intent(in) :: shvec(2)
        real(c_float), intent(out) :: f
    end function
    integer(c_int) function c_euclid_for_rot(h, iref, iptcl, shvec, irot, f) &
        &bind(C, name='pftcc_gencorr_euclid_for_rot')
        import :: c_int, c_int32_t, c_ptr, c_float
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iptcl, irot
        real(c_float), intent(in) :: shvec(2)
        real(c_float), intent(out) :: f
    end function
    integer(c_int) function c_cc_for_rot_8(h, iref, iptcl, shvec, irot, f) bind(C, name='pftcc_gencorr_cc_for_rot_8')
        import :: c_int, c_int32_t, c_ptr, c_double
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iptcl, irot
        real(c_double), intent(in) :: shvec(2)
        real(c_double), intent(out) :: f
    end function
    integer(c_int) function c_euclid_for_rot_8(h, iref, iptcl, shvec, irot, f) &
        &bind(C, name='pftcc_gencorr_euclid_for_rot_8')
        import :: c_int, c_int32_t, c_ptr, c_double
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iptcl, irot
        real(c_double), intent(in) :: shvec(2)
        real(c_double), intent(out) :: f
    end function
    integer(c_int) function c_cc_grad_for_rot_8(h, iref, iptcl, shvec, irot, f, grad) &
        &bind(C, name='pftcc_gencorr_cc_grad_for_rot_8')
        import :: c_int, c_int32_t, c_ptr, c_double
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iptcl, irot
        real(c_double), intent(in) :: shvec(2)
        real(c_double), intent(out) :: f, grad(2)
    end function
    integer(c_int) function c_euclid_grad_for_rot_8(h, iref, iptcl, shvec, irot, f, grad) &
        &bind(C, name='pftcc_gencorr_euclid_grad_for_rot_8')
        import :: c_int, c_int32_t, c_ptr, c_double
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iptcl, irot
        real(c_double), intent(in) :: shvec(2)
        real(c_double), intent(out) :: f, grad(2)
    end function
    integer(c_int) function c_cc_grad_only_for_rot_8(h, iref, iptcl, shvec, irot, grad) &
        &bind(C, name='pftcc_gencorr_cc_grad_only_for_rot_8')
        import :: c_int, c_int32_t, c_ptr, c_double
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iptcl, irot
        real(c_double), intent(in) :: shvec(2)
        real(c_double), intent(out) :: grad(2)
    end function
    integer(c_int) function c_euclid_grad_only_for_rot_8(h, iref, iptcl, shvec, irot, grad) &
        &bind(C, name='pftcc_gencorr_euclid_grad_only_for_rot_8')
        import :: c_int, c_int32_t, c_ptr, c_double
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iptcl, irot
        real(c_double), intent(in) :: shvec(2)
        real(c_double), intent(out) :: grad(2)
    end function
    integer(c_int) function c_set_dref_fcomp(h, iref, irot, k, dcomp, iseven) bind(C, name='pftcc_set_dref_fcomp')
        import :: c_int, c_int32_t, c_ptr, c_float_complex
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, irot, k, iseven
        complex(c_float_complex), intent(in) :: dcomp(3)
    end function
    integer(c_int) function c_cont_grad_cc(h, iref, iptcl, shvec, irot, cc, dcc) &
        &bind(C, name='pftcc_gencorr_cont_grad_cc_for_rot_8')
        import :: c_int, c_int32_t, c_ptr, c_double
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iptcl, irot
        real(c_double), intent(in) :: shvec(2)
        real(c_double), intent(out) :: cc, dcc(3)
    end function
    integer(c_int) function c_cont_grad_euclid(h, iref, iptcl, shvec, irot, cc, dcc) &
        &bind(C, name='pftcc_gencorr_cont_grad_euclid_for_rot_8')
        import :: c_int, c_int32_t, c_ptr, c_double
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iptcl, irot
        real(c_double), intent(in) :: shvec(2)
        real(c_double), intent(out) :: cc, dcc(3)
    end function
    integer(c_int) function c_cont_shift_grad_cc(h, iref, iptcl, shvec, irot, f, grad) &
        &bind(C, name='pftcc_gencorr_cont_shift_grad_cc_for_rot_8')
        import :: c_int, c_int32_t, c_ptr, c_double
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iptcl, irot
        real(c_double), intent(in) :: shvec(2)
        real(c_double), intent(out) :: f, grad(5)
    end function
    integer(c_int) function c_cont_shift_grad_euclid(h, iref, iptcl, shvec, irot, f, grad) &
        &bind(C, name='pftcc_gencorr_cont_shift_grad_euclid_for_rot_8')
        import :: c_int, c_int32_t, c_ptr, c_double
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iptcl, irot
        real(c_double), intent(in) :: shvec(2)
        real(c_double), intent(out) :: f, grad(5)
    end function
    integer(c_int) function c_calc_ptcl_pspec(h, iptcl, pspec) bind(C, name='pftcc_calc_ptcl_pspec')
        import :: c_int, c_int32_t, c_ptr, c_float
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iptcl
        real(c_float), intent(out) :: pspec(*)
    end function
    integer(c_int) function c_get_pfromto(h, pfromto) bind(C, name='pftcc_get_pfromto')
        import :: c_int, c_int32_t, c_ptr
        type(c_ptr), value :: h
        integer(c_int32_t), intent(out) :: pfromto(2)
    end function
    integer(c_int) function c_assign_pinds(h, pinds) bind(C, name='pftcc_assign_pinds')
        import :: c_int, c_int32_t, c_ptr
        type(c_ptr), value :: h
        integer(c_int32_t), intent(out) :: pinds(*)
    end function
    integer(c_int) function c_cp_refs(h, src) bind(C, name='pftcc_cp_refs')
        import :: c_int, c_ptr
        type(c_ptr), value :: h, src
    end function
    integer(c_int) function c_specscore(h, iref, iptcl, irot, shvec, score) bind(C, name='pftcc_specscore')
        import :: c_int, c_int32_t, c_ptr, c_float
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iref, iptcl, irot
        type(c_ptr), value :: shvec
        real(c_float), intent(out) :: score
    end function
    integer(c_int) function c_calc_roinv_corrmat(h, corrmat) bind(C, name='pftcc_calc_roinv_corrmat')
        import :: c_int, c_ptr, c_float
        type(c_ptr), value :: h
        real(c_float), intent(out) :: corrmat(*)
    end function
    integer(c_int) function c_search(h, np, iptcls, nr, irefs, ns, shifts, best_cc, best_iref, best_ishift, best_irot) &
        &bind(C, name='pftcc_search')
        import :: c_int, c_int32_t, c_ptr, c_float
        type(c_ptr), value :: h
        integer(c_int32_t), value :: np, nr, ns
        integer(c_int32_t), intent(in) :: iptcls(*)
        type(c_ptr), value :: irefs, shifts
        real(c_float), intent(out) :: best_cc(*)
        integer(c_int32_t), intent(out) :: best_iref(*), best_ishift(*), best_irot(*)
    end function
    integer(c_int) function c_shsrch_grad_batch(h, n, irefs, iptcls, lims, maxits, opt_angle, coarse_init, irots, &
        &cxy) bind(C, name='pftcc_shsrch_grad_batch')
        import :: c_int, c_int32_t, c_ptr, c_float
        type(c_ptr), value :: h
        integer(c_int32_t), value :: n, maxits, opt_angle, coarse_init
        integer(c_int32_t), intent(in) :: irefs(*), iptcls(*)
        real(c_float), intent(in) :: lims(*)
        integer(c_int32_t), intent(inout) :: irots(*)
        real(c_float), intent(out) :: cxy(*)
    end function
    integer(c_int) function c_init_polarizer(h, griddev) bind(C, name='pftcc_init_polarizer')
        import :: c_int, c_int32_t, c_ptr
        type(c_ptr), value :: h
        integer(c_int32_t), value :: griddev
    end function
    integer(c_int) function c_polarize(h, count, img_inds, cmats, isptcl, iseven, mask, nmask) &
        &bind(C, name='pftcc_polarize')
        import :: c_int, c_int32_t, c_ptr
        type(c_ptr), value :: h, cmats, mask
        integer(c_int32_t), value :: count, isptcl, iseven, nmask
        integer(c_int32_t), intent(in) :: img_inds(*)
    end function
    integer(c_int) function c_prep_imgs(h, count, imgs, cfg, shifts, ctfp, cmats) bind(C, name='pftcc_prep_imgs')
        import :: c_int, c_int32_t, c_ptr, prep_cfg
        type(c_ptr), value :: h, imgs, shifts, ctfp, cmats
        integer(c_int32_t), value :: count
        type(prep_cfg), intent(in) :: cfg
    end function
    integer(c_int) function c_strategy_srch(h, cfg, np, iptcls, order, ok, prev, rnd, bc, br, bro, ne, fb, rc, rr, so) &
        &bind(C, name='pftcc_strategy_srch')
        import :: c_int, c_int32_t, c_ptr, strategy_cfg
        type(c_ptr), value :: h, order, ok, prev, rnd, rc, rr, so
        type(strategy_cfg), intent(in) :: cfg
        integer(c_int32_t), value :: np
        integer(c_int32_t), intent(in) :: iptcls(*)
        type(c_ptr), value :: bc, br, bro, ne, fb
    end function
    integer(c_int) function c_prep_matchfilt(h, iptcl, iref, irot) bind(C, name='pftcc_prep_matchfilt')
        import :: c_int, c_int32_t, c_ptr
        type(c_ptr), value :: h
        integer(c_int32_t), value :: iptcl, iref, irot
    end function
    integer(c_int) function c_search_mfma(h, np, iptcls, nr, irefs, ns, shifts, prec, ntop, best_cc, best_iref, &
        &best_ishift, best_irot) bind(C, name='pftcc_search_mfma')
        import :: c_int, c_int32_t, c_ptr, c_float
        type(c_ptr), value :: h
        integer(c_int32_t), value :: np, nr, ns, prec, ntop
        integer(c_int32_t), intent(in) :: iptcls(*)
        type(c_ptr), value :: irefs, shifts
        real(c_float), intent(out) :: best_cc(*)
        integer(c_int32_t), intent(out) :: best_iref(*), best_ishift(*), best_irot(*)
    end function
    integer(c_int) function c_grad_batch(h, npairs, irefs, iptcls, shvecs, irots, f, grad) bind(C, name='pftcc_grad_batch')
        import :: c_int, c_int32_t, c_ptr, c_double
        type(c_ptr), value :: h
        integer(c_int32_t), value :: npairs
        integer(c_int32_t), intent(in) :: irefs(*), iptcls(*), irots(*)
        real(c_double), intent(in) :: shvecs(*)
        real(c_double), intent(out) :: f(*)
        type(c_ptr), value :: grad
    end function
    integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
        import :: c_size_t, c_ptr
        type(c_ptr), value :: s
    end function
end interface

contains

    ! stops with the library's message: the reference's THROW_HARD behaviour
    subroutine chk( self, rc, where )
        class(polarft_corrcalc), intent(in) :: self
        integer(c_int),          intent(in) :: rc
        character(len=*),        intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        type(c_ptr) :: p
        integer :: n, i
        character(len=512) :: text
        if( rc == 0 ) return
        text = ''
        if( c_associated(self%h) )then
            p = c_last_error(self%h)
            n = int(c_strlen(p))
            call c_f_pointer(p, msg, [n])
            do i = 1, min(n, 512)
                text(i:i) = msg(i)
            end do
        endif
        write(*,'(a)') 'ERROR! '//trim(where)//': '//trim(text)
        error stop 1
    end subroutine chk

    ! CONSTRUCTOR (:209-422)
    subroutine new( self, nrefs, pfromto, params, ptcl_mask, eoarr )
        class(polarft_corrcalc), target, intent(inout) :: self
        integer,                         intent(in)    :: nrefs, pfromto(2)
        type(pftcc_params),              intent(in)    :: params
        logical, optional,               intent(in)    :: ptcl_mask(pfromto(1):pfromto(2))
        integer, optional,               intent(in)    :: eoarr(pfromto(1):pfromto(2))
        type(pftcc_cfg_c) :: cfg
        integer(c_int8_t),  allocatable, target :: mask(:)
        integer(c_int32_t), allocatable, target :: eo(:)
        type(c_ptr)   :: pm, pe
        integer(c_int32_t) :: pdim(3)
        integer(c_int) :: rc
        call self%kill
        cfg%boxmatch   = params%boxmatch
        cfg%ring2      = params%ring2
        cfg%pftsz      = params%pftsz
        cfg%kfromto    = params%kfromto
        cfg%kstop      = merge(params%kstop, params%kfromto(2), params%kstop > 0)
        cfg%objfun     = params%cc_objfun
        cfg%match_filt = merge(1, 0, params%l_match_filt)
        cfg%pssnr      = merge(1, 0, params%l_pssnr)
        cfg%ctf        = merge(1, 0, params%with_ctf)
        cfg%clsfrcs    = merge(1, 0, params%l_clsfrcs)
        cfg%nthr       = params%nthr
        pm = c_null_ptr
        pe = c_null_ptr
        if( present(ptcl_mask) )then
            allocate(mask(pfromto(2)-pfromto(1)+1))
            mask = merge(1_c_int8_t, 0_c_int8_t, ptcl_mask)
            pm = c_loc(mask)
        endif
        if( present(eoarr) )then
            allocate(eo(pfromto(2)-pfromto(1)+1))
            eo = eoarr
            pe = c_loc(eo)
        endif
        rc = c_new(cfg, int(nrefs,c_int32_t), int(pfromto(1),c_int32_t), int(pfromto(2),c_int32_t), pm, pe, self%h)
        call chk(self, rc, 'new')
        rc = c_get_pdim(self%h, pdim)
        self%pftsz      = pdim(1)
        self%kfromto    = pdim(2:3)
        self%nrots      = 2 * self%pftsz
        self%nrefs      = nrefs
        rc = c_get_nptcls(self%h, pdim(1))
        self%nptcls     = pdim(1)
        self%existence  = .true.
        pftcc_glob => self
    end subroutine new

    ! SETTERS
    subroutine set_ref_pft( self, iref, pft, iseven )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref
        complex(sp),             intent(in)    :: pft(:,:)
        logical,                 intent(in)    :: iseven
        call chk(self, c_set_ref_pft(self%h, int(iref,c_int32_t), pft, merge(1_c_int32_t, 0_c_int32_t, iseven)), &
            &'set_ref_pft')
    end subroutine set_ref_pft

    subroutine set_ptcl_pft( self, iptcl, pft )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iptcl
        complex(sp),             intent(in)    :: pft(:,:)
        call chk(self, c_set_ptcl_pft(self%h, int(iptcl,c_int32_t), pft), 'set_ptcl_pft')
    end subroutine set_ptcl_pft

    subroutine set_ref_fcomp( self, iref, irot, k, comp, iseven )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, irot, k
        complex(sp),             intent(in)    :: comp
        logical,                 intent(in)    :: iseven
        call chk(self, c_set_ref_fcomp(self%h, int(iref,c_int32_t), int(irot,c_int32_t), int(k,c_int32_t), &
            &real(comp), aimag(comp), merge(1_c_int32_t, 0_c_int32_t, iseven)), 'set_ref_fcomp')
    end subroutine set_ref_fcomp

    subroutine set_ptcl_fcomp( self, iptcl, irot, k, comp )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iptcl, irot, k
        complex(sp),             intent(in)    :: comp
        call chk(self, c_set_ptcl_fcomp(self%h, int(iptcl,c_int32_t), int(irot,c_int32_t), int(k,c_int32_t), &
            &real(comp), aimag(comp)), 'set_ptcl_fcomp')
    end subroutine set_ptcl_fcomp

    subroutine cp_even2odd_ref( self, iref )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref
        call chk(self, c_cp_even2odd_ref(self%h, int(iref,c_int32_t)), 'cp_even2odd_ref')
    end subroutine cp_even2odd_ref

    subroutine cp_even_ref2ptcl( self, iref, iptcl )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl
        call chk(self, c_cp_even_ref2ptcl(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t)), 'cp_even_ref2ptcl')
    end subroutine cp_even_ref2ptcl

    subroutine swap_ptclsevenodd( self )
        class(polarft_corrcalc), intent(inout) :: self
        call chk(self, c_swap_ptclsevenodd(self%h), 'swap_ptclsevenodd')
    end subroutine swap_ptclsevenodd

    subroutine reallocate_ptcls( self, nptcls, pinds )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: nptcls
        integer,                 intent(in)    :: pinds(nptcls)
        integer(c_int32_t) :: p(nptcls)
        p = pinds
        call chk(self, c_reallocate_ptcls(self%h, int(nptcls,c_int32_t), p), 'reallocate_ptcls')
        self%nptcls = nptcls
    end subroutine reallocate_ptcls

    subroutine set_eo( self, iptcl, is_even )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iptcl
        logical,                 intent(in)    :: is_even
        call chk(self, c_set_eo(self%h, int(iptcl,c_int32_t), merge(1_c_int32_t, 0_c_int32_t, is_even)), 'set_eo')
    end subroutine set_eo

    subroutine set_eos( self, eoarr )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: eoarr(self%nptcls)
        integer(c_int32_t) :: e(self%nptcls)
        e = eoarr
        call chk(self, c_set_eos(self%h, e), 'set_eos')
    end subroutine set_eos

    subroutine set_ref_optlp( self, iref, optlp )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref
        real(sp),                intent(in)    :: optlp(:)
        call chk(self, c_set_ref_optlp(self%h, int(iref,c_int32_t), optlp), 'set_ref_optlp')
    end subroutine set_ref_optlp

    ! sigma2_noise(kfromto(1):kfromto(2), pfromto(1):pfromto(2)), copied (the reference keeps a pointer)
    subroutine assign_sigma2_noise( self, sigma2_noise )
        class(polarft_corrcalc), intent(inout) :: self
        real(sp),                intent(in)    :: sigma2_noise(:,:)
        call chk(self, c_assign_sigma2_noise(self%h, sigma2_noise), 'assign_sigma2_noise')
    end subroutine assign_sigma2_noise

    subroutine zero_refs_beyond_kstop( self )
        class(polarft_corrcalc), intent(inout) :: self
        call chk(self, c_zero_refs_beyond_kstop(self%h), 'zero_refs_beyond_kstop')
    end subroutine zero_refs_beyond_kstop

    ! GETTERS
    integer function get_nrots( self )
        class(polarft_corrcalc), intent(in) :: self
        get_nrots = self%nrots
    end function get_nrots

    function get_pdim( self ) result( pdim )
        class(polarft_corrcalc), intent(in) :: self
        integer :: pdim(3)
        pdim = [self%pftsz, self%kfromto(1), self%kfromto(2)]
    end function get_pdim

    integer function get_pftsz( self )
        class(polarft_corrcalc), intent(in) :: self
        get_pftsz = self%pftsz
    end function get_pftsz

    integer function get_box( self )
        class(polarft_corrcalc), intent(in) :: self
        integer(c_int32_t) :: v
        call chk(self, c_get_box(self%h, v), 'get_box')
        get_box = v
    end function get_box

    real(sp) function get_rot( self, roind )
        class(polarft_corrcalc), intent(in) :: self
        integer,                 intent(in) :: roind
        real(c_float) :: r
        call chk(self, c_get_rot(self%h, int(roind,c_int32_t), r), 'get_rot')
        get_rot = r
    end function get_rot

    integer function get_roind( self, rot )
        class(polarft_corrcalc), intent(in) :: self
        real(sp),                intent(in) :: rot
        integer(c_int32_t) :: ind
        call chk(self, c_get_roind(self%h, rot, ind), 'get_roind')
        get_roind = ind
    end function get_roind

    function get_coord( self, rot, k ) result( xy )
        class(polarft_corrcalc), intent(in) :: self
        integer,                 intent(in) :: rot, k
        real(sp) :: xy(2)
        call chk(self, c_get_coord(self%h, int(rot,c_int32_t), int(k,c_int32_t), xy), 'get_coord')
    end function get_coord

    function get_ref_pft( self, iref, iseven ) result( pft )
        class(polarft_corrcalc), intent(in) :: self
        integer,                 intent(in) :: iref
        logical,                 intent(in) :: iseven
        complex(sp), allocatable :: pft(:,:)
        allocate(pft(self%pftsz, self%kfromto(1):self%kfromto(2)))
        call chk(self, c_get_ref_pft(self%h, int(iref,c_int32_t), merge(1_c_int32_t, 0_c_int32_t, iseven), pft), &
            &'get_ref_pft')
    end function get_ref_pft

    function get_ptcl_pft( self, iptcl ) result( pft )
        class(polarft_corrcalc), intent(in) :: self
        integer,                 intent(in) :: iptcl
        complex(sp), allocatable :: pft(:,:)
        allocate(pft(self%pftsz, self%kfromto(1):self%kfromto(2)))
        call chk(self, c_get_ptcl_pft(self%h, int(iptcl,c_int32_t), pft), 'get_ptcl_pft')
    end function get_ptcl_pft

    integer function get_nrefs( self )
        class(polarft_corrcalc), intent(in) :: self
        get_nrefs = self%nrefs
    end function get_nrefs

    integer function get_nptcls( self )
        class(polarft_corrcalc), intent(in) :: self
        get_nptcls = self%nptcls
    end function get_nptcls

    logical function ptcl_iseven( self, iptcl )
        class(polarft_corrcalc), intent(in) :: self
        integer,                 intent(in) :: iptcl
        integer(c_int32_t) :: v
        call chk(self, c_ptcl_iseven(self%h, int(iptcl,c_int32_t), v), 'ptcl_iseven')
        ptcl_iseven = v /= 0
    end function ptcl_iseven

    logical function exists( self )
        class(polarft_corrcalc), intent(in) :: self
        exists = self%existence
    end function exists

    ! MEMOIZERS
    ! params(pfromto(1):pfromto(2)): per-particle CTF parameters (from sp_project%get_ctfparams)
    subroutine create_polar_absctfmats( self, params )
        class(polarft_corrcalc), intent(inout) :: self
        type(ctfparams),         intent(in)    :: params(:)
        call chk(self, c_create_polar_absctfmats(self%h, params), 'create_polar_absctfmats')
    end subroutine create_polar_absctfmats

    subroutine memoize_ffts( self )
        class(polarft_corrcalc), intent(inout) :: self
        call chk(self, c_memoize_ffts(self%h), 'memoize_ffts')
    end subroutine memoize_ffts

    ! CALCULATORS
    subroutine gencorrs_1( self, iref, iptcl, cc )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl
        real(sp),                intent(out)   :: cc(self%nrots)
        call chk(self, c_gencorrs(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t), c_null_ptr, cc), 'gencorrs')
    end subroutine gencorrs_1

    subroutine gencorrs_2( self, iref, iptcl, shvec, cc )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl
        real(sp),       target,  intent(in)    :: shvec(2)
        real(sp),                intent(out)   :: cc(self%nrots)
        real(c_float), target :: sh(2)
        sh = shvec
        call chk(self, c_gencorrs(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t), c_loc(sh), cc), 'gencorrs')
    end subroutine gencorrs_2

    real(sp) function gencorr_for_rot( self, iref, iptcl, shvec, irot )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl, irot
        real(sp),                intent(in)    :: shvec(2)
        real(c_float) :: f
        call chk(self, c_gencorr_for_rot(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t), shvec, &
            &int(irot,c_int32_t), f), 'gencorr_for_rot')
        gencorr_for_rot = f
    end function gencorr_for_rot

    real(dp) function gencorr_for_rot_8( self, iref, iptcl, shvec, irot )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl, irot
        real(dp),                intent(in)    :: shvec(2)
        real(c_double) :: f
        call chk(self, c_gencorr_for_rot_8(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t), shvec, &
            &int(irot,c_int32_t), f), 'gencorr_for_rot_8')
        gencorr_for_rot_8 = f
    end function gencorr_for_rot_8

    subroutine gencorr_grad_for_rot_8( self, iref, iptcl, shvec, irot, f, grad )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl, irot
        real(dp),                intent(in)    :: shvec(2)
        real(dp),                intent(out)   :: f, grad(2)
        call chk(self, c_gencorr_grad_for_rot_8(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t), shvec, &
            &int(irot,c_int32_t), f, grad), 'gencorr_grad_for_rot_8')
    end subroutine gencorr_grad_for_rot_8

    subroutine gencorr_grad_only_for_rot_8( self, iref, iptcl, shvec, irot, grad )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl, irot
        real(dp),                intent(in)    :: shvec(2)
        real(dp),                intent(out)   :: grad(2)
        call chk(self, c_gencorr_grad_only_for_rot_8(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t), shvec, &
            &int(irot,c_int32_t), grad), 'gencorr_grad_only_for_rot_8')
    end subroutine gencorr_grad_only_for_rot_8

    subroutine gencorr_sigma_contrib( self, iref, iptcl, shvec, irot, sigma_contrib, ptcl_sumsq, ref_sumsq )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl, irot
        real(sp),                intent(in)    :: shvec(2)
        real(sp),                intent(out)   :: sigma_contrib(self%kfromto(1):self%kfromto(2))
        real(sp), optional,      intent(out)   :: ptcl_sumsq(self%kfromto(1):self%kfromto(2))
        real(sp), optional,      intent(out)   :: ref_sumsq(self%kfromto(1):self%kfromto(2))
        real(c_float), allocatable, target :: ps(:), rs(:)
        type(c_ptr) :: pps, prs
        pps = c_null_ptr
        prs = c_null_ptr
        if( present(ptcl_sumsq) )then
            allocate(ps(self%kfromto(2) - self%kfromto(1) + 1))
            pps = c_loc(ps)
        endif
        if( present(ref_sumsq) )then
            allocate(rs(self%kfromto(2) - self%kfromto(1) + 1))
            prs = c_loc(rs)
        endif
        call chk(self, c_gencorr_sigma_contrib(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t), shvec, &
            &int(irot,c_int32_t), sigma_contrib, pps, prs), 'gencorr_sigma_contrib')
        if( present(ptcl_sumsq) ) ptcl_sumsq = ps
        if( present(ref_sumsq)  ) ref_sumsq  = rs
    end subroutine gencorr_sigma_contrib

    ! the objective-specific forms that gencorr_for_rot(_8) dispatches to (:1645-2148)
    real(sp) function gencorr_cc_for_rot( self, iref, iptcl, shvec, irot )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl, irot
        real(sp),                intent(in)    :: shvec(2)
        real(c_float) :: f
        call chk(self, c_cc_for_rot(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t), shvec, &
            &int(irot,c_int32_t), f), 'gencorr_cc_for_rot')
        gencorr_cc_for_rot = f
    end function gencorr_cc_for_rot

    real(dp) function gencorr_cc_for_rot_8( self, iref, iptcl, shvec, irot )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl, irot
        real(dp),                intent(in)    :: shvec(2)
        real(c_double) :: f
        call chk(self, c_cc_for_rot_8(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t), shvec, &
            &int(irot,c_int32_t), f), 'gencorr_cc_for_rot_8')
        gencorr_cc_for_rot_8 = f
    end function gencorr_cc_for_rot_8

    subroutine gencorr_cc_grad_for_rot_8( self, iref, iptcl, shvec, irot, f, grad )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl, irot
        real(dp),                intent(in)    :: shvec(2)
        real(dp),                intent(out)   :: f, grad(2)
        call chk(self, c_cc_grad_for_rot_8(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t), shvec, &
            &int(irot,c_int32_t), f, grad), 'gencorr_cc_grad_for_rot_8')
    end subroutine gencorr_cc_grad_for_rot_8

    subroutine gencorr_cc_grad_only_for_rot_8( self, iref, iptcl, shvec, irot, grad )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl, irot
        real(dp),                intent(in)    :: shvec(2)
        real(dp),                intent(out)   :: grad(2)
        call chk(self, c_cc_grad_only_for_rot_8(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t), shvec, &
            &int(irot,c_int32_t), grad), 'gencorr_cc_grad_only_for_rot_8')
    end subroutine gencorr_cc_grad_only_for_rot_8

    real(sp) function gencorr_euclid_for_rot( self, iref, iptcl, shvec, irot )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl, irot
        real(sp),                intent(in)    :: shvec(2)
        real(c_float) :: f
        call chk(self, c_euclid_for_rot(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t), shvec, &
            &int(irot,c_int32_t), f), 'gencorr_euclid_for_rot')
        gencorr_euclid_for_rot = f
    end function gencorr_euclid_for_rot

    real(dp) function gencorr_euclid_for_rot_8( self, iref, iptcl, shvec, irot )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl, irot
        real(dp),                intent(in)    :: shvec(2)
        real(c_double) :: f
        call chk(self, c_euclid_for_rot_8(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t), shvec, &
            &int(irot,c_int32_t), f), 'gencorr_euclid_for_rot_8')
        gencorr_euclid_for_rot_8 = f
    end function gencorr_euclid_for_rot_8

    subroutine gencorr_euclid_grad_for_rot_8( self, iref, iptcl, shvec, irot, f, grad )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl, irot
        real(dp),                intent(in)    :: shvec(2)
        real(dp),                intent(out)   :: f, grad(2)
        call chk(self, c_euclid_grad_for_rot_8(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t), shvec, &
            &int(irot,c_int32_t), f, grad), 'gencorr_euclid_grad_for_rot_8')
    end subroutine gencorr_euclid_grad_for_rot_8

    subroutine gencorr_euclid_grad_only_for_rot_8( self, iref, iptcl, shvec, irot, grad )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl, irot
        real(dp),                intent(in)    :: shvec(2)
        real(dp),                intent(out)   :: grad(2)
        call chk(self, c_euclid_grad_only_for_rot_8(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t), shvec, &
            &int(irot,c_int32_t), grad), 'gencorr_euclid_grad_only_for_rot_8')
    end subroutine gencorr_euclid_grad_only_for_rot_8

    ! orientation derivatives of continuous refinement (:1735-1847, :2050-2069)
    function gencorr_cont_grad_cc_for_rot_8( self, iref, iptcl, shvec, irot, dcc ) result( cc )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl, irot
        real(dp),                intent(in)    :: shvec(2)
        real(dp),                intent(out)   :: dcc(3)
        real(dp) :: cc
        call chk(self, c_cont_grad_cc(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t), shvec, &
            &int(irot,c_int32_t), cc, dcc), 'gencorr_cont_grad_cc_for_rot_8')
    end function gencorr_cont_grad_cc_for_rot_8

    subroutine gencorr_cont_shift_grad_cc_for_rot_8( self, iref, iptcl, shvec, irot, f, grad )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl, irot
        real(dp),                intent(in)    :: shvec(2)
        real(dp),                intent(out)   :: f, grad(5)
        call chk(self, c_cont_shift_grad_cc(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t), shvec, &
            &int(irot,c_int32_t), f, grad), 'gencorr_cont_shift_grad_cc_for_rot_8')
    end subroutine gencorr_cont_shift_grad_cc_for_rot_8

    function gencorr_cont_grad_euclid_for_rot_8( self, iref, iptcl, shvec, irot, dcc ) result( cc )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl, irot
        real(dp),                intent(in)    :: shvec(2)
        real(dp),                intent(out)   :: dcc(3)
        real(dp) :: cc
        call chk(self, c_cont_grad_euclid(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t), shvec, &
            &int(irot,c_int32_t), cc, dcc), 'gencorr_cont_grad_euclid_for_rot_8')
    end function gencorr_cont_grad_euclid_for_rot_8

    subroutine gencorr_cont_shift_grad_euclid_for_rot_8( self, iref, iptcl, shvec, irot, f, grad )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl, irot
        real(dp),                intent(in)    :: shvec(2)
        real(dp),                intent(out)   :: f, grad(5)
        call chk(self, c_cont_shift_grad_euclid(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t), shvec, &
            &int(irot,c_int32_t), f, grad), 'gencorr_cont_shift_grad_euclid_for_rot_8')
    end subroutine gencorr_cont_shift_grad_euclid_for_rot_8

    subroutine set_dref_fcomp( self, iref, irot, k, dcomp, iseven )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, irot, k
        complex(sp),             intent(in)    :: dcomp(3)
        logical,                 intent(in)    :: iseven
        call chk(self, c_set_dref_fcomp(self%h, int(iref,c_int32_t), int(irot,c_int32_t), int(k,c_int32_t), &
            &dcomp, merge(1_c_int32_t, 0_c_int32_t, iseven)), 'set_dref_fcomp')
    end subroutine set_dref_fcomp

    subroutine cp_refs( self, self2 )
        class(polarft_corrcalc), intent(inout) :: self, self2
        call chk(self, c_cp_refs(self%h, self2%h), 'cp_refs')
    end subroutine cp_refs

    subroutine calc_ptcl_pspec( self, iptcl, pspec )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iptcl
        real,                    intent(inout) :: pspec(self%kfromto(1):self%kfromto(2))
        call chk(self, c_calc_ptcl_pspec(self%h, int(iptcl,c_int32_t), pspec), 'calc_ptcl_pspec')
    end subroutine calc_ptcl_pspec

    subroutine assign_pinds( self, pinds )
        class(polarft_corrcalc), intent(inout) :: self
        integer, allocatable,    intent(out)   :: pinds(:)
        integer(c_int32_t) :: pft(2)
        integer(c_int32_t), allocatable :: tmp(:)
        call chk(self, c_get_pfromto(self%h, pft), 'assign_pinds')
        allocate(tmp(pft(1):pft(2)))
        call chk(self, c_assign_pinds(self%h, tmp), 'assign_pinds')
        allocate(pinds(pft(1):pft(2)))
        pinds = tmp
    end subroutine assign_pinds

    subroutine print( self )
        class(polarft_corrcalc), intent(in) :: self
        integer(c_int32_t) :: pdim(3)
        integer(c_int) :: rc
        rc = c_get_pdim(self%h, pdim)
        write(*,*) "total n particles in partition         (self%nptcls): ", self%nptcls
        write(*,*) "number of references                    (self%nrefs): ", self%nrefs
        write(*,*) "number of rotations                     (self%nrots): ", self%nrots
        write(*,*) "size of pft                             (self%pftsz): ", self%pftsz
        write(*,*) "shells held                           (self%kfromto): ", pdim(2:3)
    end subroutine print

    real(sp) function specscore_1( self, iref, iptcl, irot )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl, irot
        real(c_float) :: s
        call chk(self, c_specscore(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t), int(irot,c_int32_t), &
            &c_null_ptr, s), 'specscore')
        specscore_1 = s
    end function specscore_1

    real(sp) function specscore_2( self, iref, iptcl, irot, shvec )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iref, iptcl, irot
        real(sp),                intent(in)    :: shvec(2)
        real(c_float), target :: sh(2), s
        sh = shvec
        call chk(self, c_specscore(self%h, int(iref,c_int32_t), int(iptcl,c_int32_t), int(irot,c_int32_t), &
            &c_loc(sh), s), 'specscore')
        specscore_2 = s
    end function specscore_2

    function calc_roinv_corrmat( self ) result( corrmat )
        class(polarft_corrcalc), intent(inout) :: self
        real(sp), allocatable :: corrmat(:,:)
        allocate(corrmat(self%nptcls, self%nptcls))
        call chk(self, c_calc_roinv_corrmat(self%h, corrmat), 'calc_roinv_corrmat')
    end function calc_roinv_corrmat

    ! MATCH FILTER (:987-1095): whiten/filter the particle against iref at irot, re-memoize
    subroutine prep_matchfilt( self, iptcl, iref, irot )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: iptcl, iref, irot
        call chk(self, c_prep_matchfilt(self%h, int(iptcl,c_int32_t), int(iref,c_int32_t), int(irot,c_int32_t)), &
            &'prep_matchfilt')
    end subroutine prep_matchfilt

    ! GRADIENT SHIFT SEARCH: pftcc_shsrch_grad%minimize (simple_pftcc_shsrch_grad.f90:254-338) for
    ! n (reference, particle) pairs; lims(2,2) as grad_shsrch_new; irots in (opt_angle = .false.)
    ! and out (0: no better solution); cxy(3, n) = (cc, shift in the reference frame)
    subroutine shsrch_grad_batch( self, irefs, iptcls, lims, maxits, opt_angle, irots, cxy, coarse_init )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: irefs(:), iptcls(:)
        real(sp),                intent(in)    :: lims(2,2)
        integer,                 intent(in)    :: maxits
        logical,                 intent(in)    :: opt_angle
        integer,                 intent(inout) :: irots(size(irefs))
        real(sp),                intent(out)   :: cxy(3, size(irefs))
        logical, optional,       intent(in)    :: coarse_init
        integer(c_int32_t) :: ir(size(irefs)), ip(size(irefs)), rt(size(irefs)), ci
        real(c_float) :: lm(4)
        ir = irefs
        ip = iptcls
        rt = irots
        lm = reshape(lims, [4])
        ci = 0
        if( present(coarse_init) )then
            if( coarse_init ) ci = 1
        endif
        call chk(self, c_shsrch_grad_batch(self%h, int(size(irefs),c_int32_t), ir, ip, lm, int(maxits,c_int32_t), &
            &merge(1_c_int32_t, 0_c_int32_t, opt_angle), ci, rt, cxy), 'shsrch_grad_batch')
        irots = rt
    end subroutine shsrch_grad_batch

    ! POLARIZER: init_polarizer / polarize (simple_polarizer.f90:41-151) for the handle's grid;
    ! cmats(box/2+1, box, n) are the images' Fourier transforms as the image object holds them
    subroutine init_polarizer( self, griddev )
        class(polarft_corrcalc), intent(inout) :: self
        logical, optional,       intent(in)    :: griddev
        integer(c_int32_t) :: gd
        gd = 0
        if( present(griddev) )then
            if( griddev ) gd = 1
        endif
        call chk(self, c_init_polarizer(self%h, gd), 'init_polarizer')
    end subroutine init_polarizer

    subroutine polarize( self, img_inds, cmats, isptcl, iseven, mask )
        class(polarft_corrcalc),   intent(inout) :: self
        integer,                   intent(in)    :: img_inds(:)
        complex(sp), target,       intent(in)    :: cmats(:,:,:)
        logical,                   intent(in)    :: isptcl, iseven
        logical, optional,         intent(in)    :: mask(:)
        integer(c_int32_t) :: ids(size(img_inds))
        integer(c_int8_t), allocatable, target :: mk(:)
        type(c_ptr) :: pm
        integer(c_int32_t) :: nm
        ids = img_inds
        pm  = c_null_ptr
        nm  = 0
        if( present(mask) )then
            allocate(mk(size(mask)))
            mk = merge(1_c_int8_t, 0_c_int8_t, mask)
            pm = c_loc(mk)
            nm = size(mask)
        endif
        call chk(self, c_polarize(self%h, int(size(img_inds),c_int32_t), ids, c_loc(cmats), &
            &merge(1_c_int32_t, 0_c_int32_t, isptcl), merge(1_c_int32_t, 0_c_int32_t, iseven), pm, nm), 'polarize')
    end subroutine polarize

    ! prepimg4align (simple_strategy2D3D_common.f90:316-366) for n particle images at once:
    ! imgs(box, box, n) real space, shifts(2, n) the particles' (x, y), ctfp(n) their CTF
    ! parameters (read when cfg%ctfflag == CTFFLAG_YES); cmats(boxmatch/2+1, boxmatch, n) the
    ! transforms, ready for polarize
    subroutine prep_imgs( self, imgs, shifts, cfg, cmats, ctfp )
        class(polarft_corrcalc),   intent(inout) :: self
        real(sp), target,          intent(in)    :: imgs(:,:,:)
        real(sp), target,          intent(in)    :: shifts(:,:)
        type(prep_cfg),            intent(in)    :: cfg
        complex(sp), target,       intent(inout) :: cmats(:,:,:)
        type(ctfparams), optional, target, intent(in) :: ctfp(:)
        type(c_ptr) :: pc
        pc = c_null_ptr
        if( present(ctfp) ) pc = c_loc(ctfp)
        call chk(self, c_prep_imgs(self%h, int(size(imgs,3),c_int32_t), c_loc(imgs), cfg, c_loc(shifts), pc, &
            &c_loc(cmats)), 'prep_imgs')
    end subroutine prep_imgs

    ! BATCHED FAST PATHS (no reference counterpart: the ref loop of strategy3D_greedy%srch,
    ! simple_strategy3D_greedy.f90:60-77, moved onto the device)
    subroutine search( self, iptcls, best_cc, best_iref, best_irot, shifts, best_ishift, irefs )
        class(polarft_corrcalc),    intent(inout) :: self
        integer,                    intent(in)    :: iptcls(:)
        real(sp),                   intent(out)   :: best_cc(size(iptcls))
        integer,                    intent(out)   :: best_iref(size(iptcls)), best_irot(size(iptcls))
        real(sp), optional, target, intent(in)    :: shifts(:,:)      ! (2, nshifts)
        integer,  optional,         intent(out)   :: best_ishift(size(iptcls))
        integer,  optional, target, intent(in)    :: irefs(:)
        integer(c_int32_t) :: ip(size(iptcls)), bir(size(iptcls)), bis(size(iptcls)), brt(size(iptcls))
        integer(c_int32_t), allocatable, target :: ir(:)
        real(c_float),      allocatable, target :: sh(:,:)
        type(c_ptr) :: pir, psh
        integer(c_int32_t) :: nr, ns
        ip  = iptcls
        pir = c_null_ptr
        nr  = self%nrefs
        if( present(irefs) )then
            allocate(ir(size(irefs)), source=int(irefs, c_int32_t))
            pir = c_loc(ir)
            nr  = size(irefs)
        endif
        psh = c_null_ptr
        ns  = 0
        if( present(shifts) )then
            allocate(sh(2, size(shifts,2)), source=real(shifts, c_float))
            psh = c_loc(sh)
            ns  = size(shifts,2)
        endif
        call chk(self, c_search(self%h, int(size(iptcls),c_int32_t), ip, nr, pir, ns, psh, best_cc, bir, bis, brt), &
            &'search')
        best_iref = bir
        best_irot = brt
        if( present(best_ishift) ) best_ishift = bis + 1
    end subroutine search

    ! The reference-loop strategies (strategy2D_greedy/snhc%srch, strategy3D_greedy/shc/snhc%srch
    ! with sort_corrs) for a batch of particles against all references.  srch_order(np, nrefs) as
    ! s2D%srch_order / s3D%srch_order; ref_ok(nrefs) cls_pops > 0 / state_exists; prev_corr(np) the
    ! prep4srch threshold; rnd(np) one uniform draw per particle for shcloc (snhc).  Out: the
    ! chosen reference and rotation, nrefs_eval and whether the stochastic exit was taken;
    ! optionally proj_space_corrs / inplinds (nrefs, np) and the sorted references (ntop, np).
    subroutine strategy_srch( self, cfg, iptcls, best_corr, best_ref, best_rot, nrefs_eval, found_better, &
            &srch_order, ref_ok, prev_corr, rnd, ref_corrs, ref_rots, sorted_refs )
        class(polarft_corrcalc),    intent(inout) :: self
        type(strategy_cfg),         intent(in)    :: cfg
        integer,                    intent(in)    :: iptcls(:)
        real(sp),           target, intent(out)   :: best_corr(size(iptcls))
        integer(c_int32_t), target, intent(out)   :: best_ref(size(iptcls)), best_rot(size(iptcls))
        integer(c_int32_t), target, intent(out)   :: nrefs_eval(size(iptcls)), found_better(size(iptcls))
        integer,  optional,         intent(in)    :: srch_order(:,:)
        logical,  optional,         intent(in)    :: ref_ok(:)
        real(sp), optional, target, intent(in)    :: prev_corr(:), rnd(:)
        real(sp), optional, target, intent(out)   :: ref_corrs(:,:)
        integer(c_int32_t), optional, target, intent(out) :: ref_rots(:,:), sorted_refs(:,:)
        integer(c_int32_t) :: ip(size(iptcls))
        integer(c_int32_t), allocatable, target :: ord(:,:)
        integer(c_int8_t),  allocatable, target :: ok(:)
        type(c_ptr) :: pord, pok, pprev, prnd, prc, prr, pso
        ip   = iptcls
        pord = c_null_ptr
        if( present(srch_order) )then
            allocate(ord(size(srch_order,2), size(srch_order,1)), source=int(transpose(srch_order), c_int32_t))
            pord = c_loc(ord)
        endif
        pok = c_null_ptr
        if( present(ref_ok) )then
            allocate(ok(size(ref_ok)), source=int(merge(1, 0, ref_ok), c_int8_t))
            pok = c_loc(ok)
        endif
        pprev = c_null_ptr
        if( present(prev_corr) ) pprev = c_loc(prev_corr)
        prnd = c_null_ptr
        if( present(rnd) ) prnd = c_loc(rnd)
        prc = c_null_ptr
        prr = c_null_ptr
        if( present(ref_corrs) .and. present(ref_rots) )then
            prc = c_loc(ref_corrs)
            prr = c_loc(ref_rots)
        endif
        pso = c_null_ptr
        if( present(sorted_refs) ) pso = c_loc(sorted_refs)
        call chk(self, c_strategy_srch(self%h, cfg, int(size(iptcls),c_int32_t), ip, pord, pok, pprev, prnd, &
            &c_loc(best_corr), c_loc(best_ref), c_loc(best_rot), c_loc(nrefs_eval), c_loc(found_better), &
            &prc, prr, pso), 'strategy_srch')
    end subroutine strategy_srch

    ! precision form of search without shifts: MFMA contraction per angular frequency, the ntop
    ! best references per particle rescored exactly in FP64 (ntop = 0: no rescoring)
    subroutine search_mfma( self, iptcls, best_cc, best_iref, best_irot, prec, ntop, irefs, shifts, best_ishift )
        class(polarft_corrcalc),    intent(inout) :: self
        integer,                    intent(in)    :: iptcls(:)
        real(sp),                   intent(out)   :: best_cc(size(iptcls))
        integer,                    intent(out)   :: best_iref(size(iptcls)), best_irot(size(iptcls))
        integer,                    intent(in)    :: prec, ntop
        integer,  optional, target, intent(in)    :: irefs(:)
        real(sp), optional, target, intent(in)    :: shifts(:,:)      ! (2, nshifts)
        integer,  optional,         intent(out)   :: best_ishift(size(iptcls))
        integer(c_int32_t) :: ip(size(iptcls)), bir(size(iptcls)), bis(size(iptcls)), brt(size(iptcls))
        integer(c_int32_t), allocatable, target :: ir(:)
        real(c_float),      allocatable, target :: sh(:,:)
        type(c_ptr) :: pir, psh
        integer(c_int32_t) :: nr, ns
        ip  = iptcls
        pir = c_null_ptr
        nr  = self%nrefs
        if( present(irefs) )then
            allocate(ir(size(irefs)), source=int(irefs, c_int32_t))
            pir = c_loc(ir)
            nr  = size(irefs)
        endif
        psh = c_null_ptr
        ns  = 0
        if( present(shifts) )then
            allocate(sh(2, size(shifts,2)), source=real(shifts, c_float))
            psh = c_loc(sh)
            ns  = size(shifts,2)
        endif
        call chk(self, c_search_mfma(self%h, int(size(iptcls),c_int32_t), ip, nr, pir, ns, psh, int(prec,c_int32_t), &
            &int(ntop,c_int32_t), best_cc, bir, bis, brt), 'search_mfma')
        best_iref = bir
        best_irot = brt
        if( present(best_ishift) ) best_ishift = bis + 1
    end subroutine search_mfma

    subroutine grad_batch( self, irefs, iptcls, shvecs, irots, f, grad )
        class(polarft_corrcalc), intent(inout) :: self
        integer,                 intent(in)    :: irefs(:), iptcls(:), irots(:)
        real(dp),                intent(in)    :: shvecs(:,:)      ! (2, npairs)
        real(dp),                intent(out)   :: f(size(irefs))
        real(dp), optional, target, contiguous, intent(out) :: grad(:,:)   ! (3, npairs): d/dx, d/dy, d/dpsi
        integer(c_int32_t) :: ir(size(irefs)), ip(size(irefs)), ro(size(irefs))
        type(c_ptr) :: pg
        ir = irefs
        ip = iptcls
        ro = irots
        pg = c_null_ptr
        if( present(grad) ) pg = c_loc(grad)
        call chk(self, c_grad_batch(self%h, int(size(irefs),c_int32_t), ir, ip, shvecs, ro, f, pg), 'grad_batch')
    end subroutine grad_batch

    ! DESTRUCTOR (:2234-2279)
    subroutine kill( self )
        class(polarft_corrcalc), intent(inout) :: self
        integer(c_int) :: rc
        if( c_associated(self%h) )then
            rc = c_kill(self%h)
            self%h = c_null_ptr
        endif
        self%existence = .false.
    end subroutine kill

end module simple_polarft_corrcalc
