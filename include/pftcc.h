/*
 * pftcc.h - C ABI of the MI355X (gfx950) polar-Fourier correlation search.
 *
 * Drop-in boundary for SIMPLE's type polarft_corrcalc
 * (src/main/simple_polarft_corrcalc.f90:68-198).  Every entry point below names the reference
 * procedure it replaces.  The Fortran module simple3.0_amd/fortran/simple_polarft_corrcalc.f90
 * binds these with ISO_C_BINDING and keeps the reference's type-bound procedure names.
 *
 * Conventions
 *  - Plain C types only: int32_t/int64_t/float/double and pointers.  Complex single precision
 *    values are interleaved (re, im) float pairs; a polar FT block is float[nk][pftsz][2],
 *    which is the memory order of the reference's complex(sp) pft(pftsz, kfromto(1):kfromto(2)).
 *  - Indices keep the reference's meaning and base: iref in [1, nrefs]; iptcl is a GLOBAL
 *    particle index in [pfrom, pto] mapped through the pinds table (:276-288, :509-554);
 *    irot in [1, nrots], irot = 1 <-> 0 degrees (:261-274).
 *  - Every function returns 0 on success and a non-zero PFTCC_E* code on failure; the message
 *    is available from pftcc_last_error(h).  The reference stops the program instead
 *    (THROW_HARD, src/fileio/simple_error.f90:8-31); the Fortran wrapper does the same.
 *  - Host pointers are borrowed for the duration of the call.  Entry points with the suffix
 *    _dev take device pointers (memory on the GPU the handle was created on).
 *  - Setters must not run concurrently with searches on the same handle.  Per-pair calls
 *    (gencorrs, gencorr_*_for_rot_8) may be issued from several host threads; they are
 *    serialised on the handle's stream.  Batched calls are the fast path.
 */
#ifndef PFTCC_H
#define PFTCC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFTCC_OK 0
#define PFTCC_E_ARG 1      /* invalid argument (index out of range, bad size, NULL pointer) */
#define PFTCC_E_STATE 2    /* call out of order (e.g. particle FFTs not memoized)           */
#define PFTCC_E_DEVICE 3   /* HIP runtime error or device out of memory                      */
#define PFTCC_E_UNSUPPORTED 4

#define PFTCC_OBJFUN_CC 0      /* OBJFUN_CC     (simple_defs)  */
#define PFTCC_OBJFUN_EUCLID 1  /* OBJFUN_EUCLID (simple_defs)  */

/* Explicit replacement of the params_glob fields the reference reads inside the module
 * (simple_parameters.f90:1182-1193 ring2, 1226-1229 boxmatch, 1262-1264 kfromto/kstop,
 * 1336-1371 objfun/match_filt/pssnr).  No hidden globals. */
typedef struct pftcc_cfg {
  int32_t boxmatch;    /* ldim(1:2) of the Cartesian image; must be even (new :236-243)      */
  int32_t ring2;       /* pftsz = magic_pftsz(ring2) when pftsz == 0 (new :251)              */
  int32_t pftsz;       /* explicit half-ring length (even) or 0                               */
  int32_t kfromto[2];  /* shell range held in memory                                         */
  int32_t kstop;       /* last shell used by the correlations (<= kfromto[1])                */
  int32_t objfun;      /* PFTCC_OBJFUN_CC or PFTCC_OBJFUN_EUCLID (cc_objfun)                  */
  int32_t match_filt;  /* l_match_filt                                                        */
  int32_t pssnr;       /* l_pssnr                                                             */
  int32_t ctf;         /* 0: ctf=no; 1: with_ctf (ctf /= 'no', :355-356)                      */
  int32_t clsfrcs;     /* l_clsfrcs (:415)                                                    */
  int32_t nthr;        /* host threads that may issue per-pair calls (informational)          */
} pftcc_cfg;

/* Per-particle CTF parameters (ctfparams from sp_project%get_ctfparams, used by
 * create_polar_absctfmats :924-953; units as in simple_ctf.f90:54-81). */
typedef struct pftcc_ctfparams {
  float smpd;      /* A/pixel                          */
  float kv;        /* acceleration voltage, kV         */
  float cs;        /* spherical aberration, mm         */
  float fraca;     /* amplitude contrast fraction      */
  float dfx, dfy;  /* defocus, microns                 */
  float angast;    /* astigmatism angle, degrees       */
  float phshift;   /* phase-plate shift, radians       */
  int32_t l_phaseplate;
} pftcc_ctfparams;

typedef struct pftcc pftcc_t;

/* --------------------------------------------------------------- construction / teardown -- */
/* new (:209-422).  ptcl_mask (int8[pto-pfrom+1], nonzero = keep) and eoarr (int32, 0 = even)
 * may be NULL.  The handle is created on the current HIP device. */
int pftcc_new(const pftcc_cfg* cfg, int32_t nrefs, int32_t pfrom, int32_t pto,
              const int8_t* ptcl_mask, const int32_t* eoarr, pftcc_t** out);
/* kill (:2234-2279) */
int pftcc_kill(pftcc_t* h);
const char* pftcc_last_error(const pftcc_t* h);
/* Asynchronous work is issued on this stream (hipStream_t; NULL = the null stream). */
int pftcc_set_stream(pftcc_t* h, void* hip_stream);
int pftcc_sync(pftcc_t* h);

/* --------------------------------------------------------------------------------- setters -- */
int pftcc_set_ref_pft(pftcc_t* h, int32_t iref, const float* pft, int32_t iseven);   /* :426 */
int pftcc_set_ptcl_pft(pftcc_t* h, int32_t iptcl, const float* pft);                /* :438 */
/* Bulk forms of the two setters (same semantics; count consecutive blocks). */
int pftcc_set_ref_pfts(pftcc_t* h, int32_t iref_first, int32_t count, const float* pfts,
                       int32_t iseven);
int pftcc_set_ptcl_pfts(pftcc_t* h, int32_t iptcl_first, int32_t count, const float* pfts);
int pftcc_set_ref_pfts_dev(pftcc_t* h, int32_t iref_first, int32_t count, const void* dpfts,
                           int32_t iseven);
int pftcc_set_ptcl_pfts_dev(pftcc_t* h, int32_t iptcl_first, int32_t count, const void* dpfts);
int pftcc_set_ref_fcomp(pftcc_t* h, int32_t iref, int32_t irot, int32_t k, float re, float im,
                        int32_t iseven);                                              /* :447 */
int pftcc_set_ptcl_fcomp(pftcc_t* h, int32_t iptcl, int32_t irot, int32_t k, float re,
                         float im);                                                   /* :471 */
int pftcc_cp_even2odd_ref(pftcc_t* h, int32_t iref);                                 /* :485 */
/* cp_refs (:491-495): both reference sets of src (same nrefs, pftsz, kfromto, same GPU) into h. */
int pftcc_cp_refs(pftcc_t* h, const pftcc_t* src);
int pftcc_cp_even_ref2ptcl(pftcc_t* h, int32_t iref, int32_t iptcl);                 /* :497 */
int pftcc_swap_ptclsevenodd(pftcc_t* h);                                             /* :504 */
int pftcc_reallocate_ptcls(pftcc_t* h, int32_t nptcls, const int32_t* pinds);        /* :509 */
int pftcc_set_eo(pftcc_t* h, int32_t iptcl, int32_t is_even);                        /* :557 */
int pftcc_set_eos(pftcc_t* h, const int32_t* eoarr /* [nptcls] */);                  /* :564 */
int pftcc_set_ref_optlp(pftcc_t* h, int32_t iref, const float* optlp /* [kfrom..kstop] */); /* :587 */
/* assign_sigma2_noise (:581-585): sigma2[(iptcl - pfrom) * nk + (k - kfrom)] for
 * iptcl in [pfrom, pto], k in [kfrom, kto].  The data is copied. */
int pftcc_assign_sigma2_noise(pftcc_t* h, const float* sigma2);
int pftcc_zero_refs_beyond_kstop(pftcc_t* h);                                        /* :747 */

/* ------------------------------------------------------------------------------ memoizers -- */
/* create_polar_absctfmats (:906-959).  params[i] belongs to global particle pfrom + i
 * (i in [0, pto - pfrom]); particles masked out are skipped. */
int pftcc_create_polar_absctfmats(pftcc_t* h, const pftcc_ctfparams* params);
int pftcc_memoize_ffts(pftcc_t* h);                                                  /* :848 */

/* --------------------------------------------------------------------------------- getters -- */
int pftcc_get_nrots(const pftcc_t* h, int32_t* nrots);                               /* :597 */
int pftcc_get_pdim(const pftcc_t* h, int32_t pdim[3]);                               /* :604 */
int pftcc_get_pftsz(const pftcc_t* h, int32_t* pftsz);                               /* :611 */
int pftcc_get_box(const pftcc_t* h, int32_t* box);                                   /* :616 */
int pftcc_get_rot(const pftcc_t* h, int32_t roind, float* rot);                      /* :623 */
int pftcc_get_roind(const pftcc_t* h, float rot, int32_t* roind);                    /* :635 */
int pftcc_get_coord(const pftcc_t* h, int32_t irot, int32_t k, float xy[2]);         /* :648 */
int pftcc_get_nrefs(const pftcc_t* h, int32_t* nrefs);                               /* :685 */
int pftcc_get_nptcls(const pftcc_t* h, int32_t* nptcls);                             /* :701 */
int pftcc_ptcl_iseven(const pftcc_t* h, int32_t iptcl, int32_t* iseven);            /* :695 */
int pftcc_get_ref_pft(pftcc_t* h, int32_t iref, int32_t iseven, float* pft);        /* :670 */
/* The particle range [pfrom, pto] of the index table (pfromto, :276; reallocate_ptcls :514-515). */
int pftcc_get_pfromto(const pftcc_t* h, int32_t pfromto[2]);
/* assign_pinds (:706-710): pinds[i] = the 1-based local index of global particle pfrom + i
 * (i in [0, pto - pfrom]), 0 when the particle is not held. */
int pftcc_assign_pinds(const pftcc_t* h, int32_t* pinds);
int pftcc_get_ptcl_pft(pftcc_t* h, int32_t iptcl, float* pft);                      /* :660 */
/* Device-resident memo and CTF matrices, for verification: ctfmat float[nk][pftsz];
 * memo float[nk][pftsz+1][2] = real FFT of length 2*pftsz of the particle's ring (the
 * reference's PA/PB memo re-packed: W[2t] = 2 PA[t], W[2t+1] = 2 PB[t]). */
int pftcc_get_ctfmat(pftcc_t* h, int32_t iptcl, float* ctfmat);
int pftcc_get_memo(pftcc_t* h, int32_t iptcl, float* memo);
int pftcc_get_sqsum_ptcl(pftcc_t* h, int32_t iptcl, float* sqsum);

/* ------------------------------------------------------------------------- calculators -- */
/* gencorrs (:173, 1594-1617): cc[nrots].  shvec == NULL selects gencorrs_1, else gencorrs_2
 * (shift applied to the reference).  objfun CC or EUCLID as configured. */
int pftcc_gencorrs(pftcc_t* h, int32_t iref, int32_t iptcl, const float* shvec, float* cc);
/* gencorr_for_rot (:1619) and gencorr_for_rot_8 (:1632) */
int pftcc_gencorr_for_rot(pftcc_t* h, int32_t iref, int32_t iptcl, const float shvec[2],
                          int32_t irot, float* f);
int pftcc_gencorr_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2],
                            int32_t irot, double* f);
/* gencorr_grad_for_rot_8 (:1849) and gencorr_grad_only_for_rot_8 (:1927) */
int pftcc_gencorr_grad_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2],
                                 int32_t irot, double* f, double grad[2]);
int pftcc_gencorr_grad_only_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl,
                                      const double shvec[2], int32_t irot, double grad[2]);
/* The objective-specific forms that gencorr_for_rot(_8) / gencorr_grad(_only)_for_rot_8 dispatch to
 * (:1632-1643): gencorr_cc_for_rot (:1645), gencorr_cc_for_rot_8 (:1690), gencorr_cc_grad_for_rot_8
 * (:1863), gencorr_cc_grad_only_for_rot_8 (:1927), gencorr_euclid_for_rot (:2002),
 * gencorr_euclid_for_rot_8 (:2026), gencorr_euclid_grad_for_rot_8 (:2071),
 * gencorr_euclid_grad_only_for_rot_8 (:2111).  The objective is fixed instead of read from
 * cc_objfun; the Euclid forms need assign_sigma2_noise. */
int pftcc_gencorr_cc_for_rot(pftcc_t* h, int32_t iref, int32_t iptcl, const float shvec[2], int32_t irot,
                             float* f);
int pftcc_gencorr_cc_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2], int32_t irot,
                               double* f);
int pftcc_gencorr_cc_grad_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2],
                                    int32_t irot, double* f, double grad[2]);
int pftcc_gencorr_cc_grad_only_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2],
                                         int32_t irot, double grad[2]);
int pftcc_gencorr_euclid_for_rot(pftcc_t* h, int32_t iref, int32_t iptcl, const float shvec[2], int32_t irot,
                                 float* f);
int pftcc_gencorr_euclid_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2],
                                   int32_t irot, double* f);
int pftcc_gencorr_euclid_grad_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2],
                                        int32_t irot, double* f, double grad[2]);
int pftcc_gencorr_euclid_grad_only_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2],
                                             int32_t irot, double grad[2]);

/* Orientation derivatives for continuous refinement (the callers are simple_projector.f90:264 and
 * simple_pftcc_orisrch_grad.f90:145-204).  set_dref_fcomp (:459-469): dcomp = 3 complex values
 * (re, im interleaved), the derivatives of reference component (irot in [1, pftsz], k in
 * [kfrom, kstop]) w.r.t. the three orientation angles.  As in the reference (pfts_drefs,
 * :321-322) the derivative slots are 1..max(nthr, 1) of the configuration, addressed by iref.
 * set_dref_pft: all three derivative blocks of slot iref at once, float[3][kstop-kfrom+1][pftsz][2]. */
int pftcc_set_dref_fcomp(pftcc_t* h, int32_t iref, int32_t irot, int32_t k, const float dcomp[6], int32_t iseven);
int pftcc_set_dref_pft(pftcc_t* h, int32_t iref, const float* dpft, int32_t iseven);
/* gencorr_cont_grad_cc_for_rot_8 (:1735-1786): cc and dcc[3], d cc / d(orientation angles), with
 * q = reference x shell filter x |CTF| x shift phase, dq = dref x |CTF| x shift phase,
 * cc = y / sqrt(S_q S_p), dcc = (T1 - y T2 / S_q) / sqrt(S_q S_p), S_p the memoized sqsum.  As in
 * the reference, T1 = Re sum dq conj(P) is taken against particle slot iref (its
 * calc_T1_T2_for_rot_8 call passes iref as the particle index, :1781), so iref <= nptcls. */
int pftcc_gencorr_cont_grad_cc_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2],
                                         int32_t irot, double* cc, double dcc[3]);
/* gencorr_cont_shift_grad_cc_for_rot_8 (:1788-1847): f and grad[5] = 3 orientation derivatives
 * (as above), then d/dshift_x and d/dshift_y (unweighted shells over the same denominator). */
int pftcc_gencorr_cont_shift_grad_cc_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl,
                                               const double shvec[2], int32_t irot, double* f,
                                               double grad[5]);
/* gencorr_cont_grad_euclid_for_rot_8 / gencorr_cont_shift_grad_euclid_for_rot_8 (:2050-2069):
 * unimplemented stubs in the reference (cc = 0, derivatives unset); here every output is 0. */
int pftcc_gencorr_cont_grad_euclid_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl, const double shvec[2],
                                             int32_t irot, double* cc, double dcc[3]);
int pftcc_gencorr_cont_shift_grad_euclid_for_rot_8(pftcc_t* h, int32_t iref, int32_t iptcl,
                                                   const double shvec[2], int32_t irot, double* f,
                                                   double grad[5]);
/* calc_ptcl_pspec (:897-904): pspec[k - kfrom] = sum_j |p_k[j]|^2 / pftsz for k in [kfrom, kto]. */
int pftcc_calc_ptcl_pspec(pftcc_t* h, int32_t iptcl, float* pspec);
/* gencorr_sigma_contrib (:2150): sigma_contrib[nk] (+ optional ptcl/ref sums of squares). */
int pftcc_gencorr_sigma_contrib(pftcc_t* h, int32_t iref, int32_t iptcl, const float shvec[2],
                                int32_t irot, float* sigma_contrib, float* ptcl_sumsq,
                                float* ref_sumsq);
/* specscore (:2190-2205); shvec may be NULL (specscore_1) */
int pftcc_specscore(pftcc_t* h, int32_t iref, int32_t iptcl, int32_t irot, const float* shvec,
                    float* score);
/* calc_roinv_corrmat (:2207-2230): corrmat[nptcls][nptcls] (requires nptcls == nrefs). */
int pftcc_calc_roinv_corrmat(pftcc_t* h, float* corrmat);

/* ----------------------------------------------------------------- batched fast paths -- */
/* Exhaustive correlation of particles x references x shifts, all rotations.
 * iptcls[np] (global ids), irefs[nr] (1-based; NULL = all refs 1..nr), shifts[ns][2]
 * (NULL with ns == 0 selects the unshifted gencorrs_1 semantics).
 * cc_out layout: [np][nr][max(ns,1)][nrots] (host memory).  This is what a loop of
 * gencorrs(iref, iptcl[, shvec], cc) over the same lists returns, bit-compatible with it. */
int pftcc_gencorrs_batch(pftcc_t* h, int32_t np, const int32_t* iptcls, int32_t nr,
                         const int32_t* irefs, int32_t ns, const float* shifts, float* cc_out);
/* The same search reduced on the device: for each particle the best (ref, shift, irot) over
 * all nr x max(ns,1) x nrots evaluations, ties to the lowest (ref, shift, irot) in that order.
 * Outputs (host memory, length np): best_cc, best_iref (1-based), best_ishift (0-based into
 * shifts, 0 if ns == 0), best_irot (1-based).  The loop it replaces is the greedy ref loop of
 * strategy3D_greedy%srch (simple_strategy3D_greedy.f90:60-77: gencorrs + maxloc). */
int pftcc_search(pftcc_t* h, int32_t np, const int32_t* iptcls, int32_t nr, const int32_t* irefs,
                 int32_t ns, const float* shifts, float* best_cc, int32_t* best_iref,
                 int32_t* best_ishift, int32_t* best_irot);
/* Device-pointer form of pftcc_search: all array arguments are device pointers; runs
 * asynchronously on the handle's stream. */
int pftcc_search_dev(pftcc_t* h, int32_t np, const int32_t* d_iptcls, int32_t nr,
                     const int32_t* d_irefs, int32_t ns, const float* d_shifts, float* d_best_cc,
                     int32_t* d_best_iref, int32_t* d_best_ishift, int32_t* d_best_irot);
/* pftcc_search_dev that allocates nothing.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and errors as pftcc_search_dev; it
 * calls no hipMalloc, hipFree or other allocating routine.  Its working memory (index lists, the
 * shift table, the per-particle best keys) is a static device array of 32 MiB per GPU, held
 * until the process ends; a call needing more (about np * 16 + nr * 4 + max(ns,1) *
 * (nk * pftsz * 8 + 16) bytes, so at most ~150 shifts at box 200) returns PFTCC_E_ARG.  Every
 * call on one GPU shares that array and leaves the previous call's contents in it: calls that
 * overlap on one GPU (several handles, several host threads, several streams) race and must be
 * serialised by the caller.  With match_filt or pssnr the shell-normalisation tables must be
 * current (PFTCC_E_STATE otherwise): one pftcc_search or pftcc_search_dev call after the last
 * reference or particle change builds them. */
int pftcc_search_dev_noalloc(pftcc_t* h, int32_t np, const int32_t* d_iptcls, int32_t nr,
                             const int32_t* d_irefs, int32_t ns, const float* d_shifts,
                             float* d_best_cc, int32_t* d_best_iref, int32_t* d_best_ishift,
                             int32_t* d_best_irot);
/* The reference-loop search strategies for a batch of particles, each particle against all nrefs
 * references with gencorrs (:1594, unshifted) and maxloc over the rotations, then:
 *   PFTCC_SRCH_GREEDY2D  strategy2D_greedy%srch (simple_strategy2D_greedy.f90:33-58): references
 *                        1..nrefs with ref_ok, "inpl_corr >= corr" (the last of equal maxima);
 *   PFTCC_SRCH_SNHC2D    strategy2D_snhc%srch (simple_strategy2D_snhc.f90:33-93): srch_order,
 *                        neighbourhood max(2, min(nrefs, nint(nrefs (1 - stoch_bound)))), empty
 *                        classes skipped but counted, exit at the first reference where shcloc
 *                        (simple_rnd.f90:273-295) finds a rotation with corrs >= prev_corr,
 *                        else the best of the neighbourhood in search order;
 *   PFTCC_SRCH_GREEDY3D  strategy3D_greedy%srch (simple_strategy3D_greedy.f90:40-77);
 *   PFTCC_SRCH_SHC3D     strategy3D_shc%srch (simple_strategy3D_shc.f90:36-83): srch_order, exit
 *                        after the first reference whose best exceeds prev_corr;
 *   PFTCC_SRCH_SNHC3D    strategy3D_snhc%srch (simple_strategy3D_snhc.f90:37-60): the first
 *                        cfg->szsn references of srch_order, every one searched (no state test);
 * the 3-D kinds then order the searched references by best correlation (sort_corrs,
 * simple_strategy3D_utils.f90:100-124; unsearched at -huge, ties to the lower index).
 * srch_order[np][nrefs]: 1-based permutations (the stochastic kinds; ignored by the greedy ones).
 * ref_ok[nrefs]: cls_pops > 0 (2-D) / state_exists (3-D), NULL = all.  prev_corr[np]: the
 * threshold prep4srch computes (NULL = -huge for the greedy kinds).  rnd[np]: one uniform draw in
 * [0, 1) per particle (snhc): shcloc's choice is the floor(rnd count)-th rotation, in rotation
 * order, of the count with corrs >= prev_corr, the same uniform distribution as its rejection
 * sampling.  Out [np]: best_corr, best_ref (1-based; 0 if no reference was searched, when
 * best_corr = prev_corr), best_rot (1-based), nrefs_eval (as the strategy counts it),
 * found_better (the stochastic exit was taken).  Optional: ref_corrs/ref_rots [np][nrefs]
 * (proj_space_corrs / inplinds: -FLT_MAX and 0 where unsearched); sorted_refs [np][cfg->ntop]
 * (3-D kinds: 1-based references, best first).  No counterpart entry point: it replaces the
 * per-particle srch procedures above. */
enum { PFTCC_SRCH_GREEDY2D = 0, PFTCC_SRCH_SNHC2D = 1, PFTCC_SRCH_GREEDY3D = 2, PFTCC_SRCH_SHC3D = 3,
       PFTCC_SRCH_SNHC3D = 4 };
typedef struct {
  int32_t kind;        /* PFTCC_SRCH_* */
  float stoch_bound;   /* snhc 2-D: spec%stoch_bound, in [0, 1] */
  int32_t ntop;        /* 3-D kinds: length of each sorted_refs row */
  int32_t szsn;        /* snhc 3-D: spec%szsn, the neighbourhood size, in [1, nrefs] */
} pftcc_strategy_cfg;
int pftcc_strategy_srch(pftcc_t* h, const pftcc_strategy_cfg* cfg, int32_t np, const int32_t* iptcls,
                        const int32_t* srch_order, const int8_t* ref_ok, const float* prev_corr,
                        const float* rnd, float* best_corr, int32_t* best_ref, int32_t* best_rot,
                        int32_t* nrefs_eval, int32_t* found_better, float* ref_corrs,
                        int32_t* ref_rots, int32_t* sorted_refs);
/* Batched gencorr_grad_for_rot_8 over npairs (iref, iptcl, shvec, irot) tuples (the callers are
 * pftcc_shsrch_grad's cost function, simple_pftcc_shsrch_grad.f90:114-131, per particle and
 * local reference).  f[np], grad[np][3]: d/dshift_x, d/dshift_y and d/dpsi, psi the continuous
 * in-plane angle in radians at psi = (irot-1)*pi/pftsz (the band-limited rotational
 * interpolant of the objective).  grad may be NULL for objective-only evaluation. */
int pftcc_grad_batch(pftcc_t* h, int32_t npairs, const int32_t* irefs, const int32_t* iptcls,
                     const double* shvecs, const int32_t* irots, double* f, double* grad);
int pftcc_grad_batch_dev(pftcc_t* h, int32_t npairs, const int32_t* d_irefs,
                         const int32_t* d_iptcls, const double* d_shvecs, const int32_t* d_irots,
                         double* d_f, double* d_grad);

/* prep_matchfilt (simple_polarft_corrcalc.f90:987-1095): with match_filt, whiten particle iptcl
 * against reference iref at rotation irot (1-based) -- shell powers normalised to optlp^2, the
 * reference-particle difference spectrum smoothed over shells (Savitzky-Golay, 7 points, order 4,
 * simple_math.f90:3511), particle scaled to power ssnr/(ssnr+1); with pssnr scaled by
 * sqrt(k/pw) sqrt(1 + ssnr |CTF|^2) -- store it and re-memoize its spectra.  iref = 0 or
 * match_filt off: memoize only.  The batch form takes distinct particles. */
int pftcc_prep_matchfilt(pftcc_t* h, int32_t iptcl, int32_t iref, int32_t irot);
int pftcc_prep_matchfilt_batch(pftcc_t* h, int32_t np, const int32_t* iptcls, const int32_t* irefs,
                               const int32_t* irots);

/* ----------------------------------------------------------------- polarizer (input side) -- */
/* init_polarizer (simple_polarizer.f90:41-92): Kaiser-Bessel gridding tables for the handle's
 * polar grid and box (window half-width KBWINSZ = 1.5, beta 7.4, 3 x 3 components per polar
 * point, weights divided by their sum unless griddev != 0, params_glob%griddev == 'yes').  The
 * oversampling alpha does not enter these weights; it only shapes the instrument function that the
 * caller divides real-space images by before their FFT (div_by_instrfun).  Requires
 * kfromto(2) <= boxmatch/2. */
int pftcc_init_polarizer(pftcc_t* h, int32_t griddev);
/* polarize (simple_polarizer.f90:120-151) for count images at once.  cmats: the images' Fourier
 * transforms as the image object holds them after fwd_ft (simple_image.f90:5773-5782), i.e.
 * Fortran cmat(boxmatch/2+1, boxmatch) of complex float per image, images consecutive.
 * img_inds: distinct 1-based particle ids (isptcl != 0; set_ptcl_pft, sqsums refreshed) or
 * reference ids (isptcl = 0, into the even or odd set by iseven; set_ref_pft).  mask: NULL, or the
 * reference's band mask(:) with nmask >= kfromto(2) entries; shell k is zeroed when mask[k-1] == 0.
 * The _dev form reads the transforms from device memory.  As after set_ptcl_pft, memoize_ffts
 * must follow before particles are searched. */
int pftcc_polarize(pftcc_t* h, int32_t count, const int32_t* img_inds, const float* cmats, int32_t isptcl,
                   int32_t iseven, const int8_t* mask, int32_t nmask);
int pftcc_polarize_dev(pftcc_t* h, int32_t count, const int32_t* img_inds, const void* d_cmats, int32_t isptcl,
                       int32_t iseven, const int8_t* mask, int32_t nmask);

/* Particle preparation ahead of the polarizer: prepimg4align (simple_strategy2D3D_common.f90:
 * 316-366) for count 2-D images at once, replacing the per-image host work before
 * polarize(isptcl = 1).  Per image: noise_norm over the background of the disc of radius msk
 * (build_glob%lmsk), fft, shift2Dserial(-x, -y) when |x| or |y| > 0.001, CTF phase flip
 * (ctfflag == CTFFLAG_YES, apply_serial 'flip' with ctfp[i]), ifft, real-space clip to
 * boxmatch, soft mask (zero_background + cosine edge of radius msk; inner/width with
 * l_innermsk; focusmsk instead of msk with l_focusmsk), div_by_instrfun when griddev, fft and
 * the factor boxmatch^2/box^2.  imgs: real images [count][box][box] (Fortran rmat(box, box)
 * order, x fastest); shifts: [count][2] particle (x, y); ctfp: [count] or NULL when ctfflag !=
 * CTFFLAG_YES.  cmats: [count] transforms in the layout polarize reads (Fortran
 * cmat(boxmatch/2+1, boxmatch)).  boxmatch must equal the handle's.  The _dev form takes device
 * imgs and cmats (shifts and ctfp stay on the host). */
typedef struct pftcc_prep_cfg {
  int32_t box;        /* input image size, even                                          */
  int32_t boxmatch;   /* output transform size, even, <= box; the handle's boxmatch         */
  float msk;          /* mask radius in pixels (params%msk)                                 */
  float inner;        /* l_innermsk: inner mask radius (0: off)                             */
  float width;        /* inner mask width (params%width)                                    */
  float focusmsk;     /* l_focusmsk: soft-mask radius used instead of msk (0: off)          */
  int32_t ctfflag;    /* CTFFLAG_NO 0, CTFFLAG_YES 1 (phase flip), CTFFLAG_FLIP 2 (simple_defs) */
  int32_t griddev;    /* params%griddev == 'yes'                                            */
} pftcc_prep_cfg;
int pftcc_prep_imgs(pftcc_t* h, int32_t count, const float* imgs, const pftcc_prep_cfg* cfg, const float* shifts,
                    const pftcc_ctfparams* ctfp, float* cmats);
int pftcc_prep_imgs_dev(pftcc_t* h, int32_t count, const void* d_imgs, const pftcc_prep_cfg* cfg,
                        const float* shifts, const pftcc_ctfparams* ctfp, void* d_cmats);

/* ------------------------------------------------------------ gradient shift search -- */
/* pftcc_shsrch_grad%minimize (simple_pftcc_shsrch_grad.f90:254-338) for n (reference,
 * particle) pairs at once (set_indices(irefs[i], iptcls[i]) then minimize(irots[i])).  Each pair
 * runs L-BFGS-B over the shift within lims = Fortran lims(2,2) = {xlo, ylo, xhi, yhi}
 * (grad_shsrch_new :46-82: m 15, factr 1e7, pgtol 1e-5, line search ftol 0.1 gtol 1e-3, maxits
 * on the objective evaluations) on gencorr_grad_for_rot_8.  opt_angle != 0: the in-plane rotation
 * starts at the best of gencorrs at shift 0 and is re-chosen (gencorrs + maxloc) at every new
 * iterate, for up to 5 shift-search / rotation cycles; opt_angle = 0: irots[i] is the fixed
 * rotation.  coarse_init != 0: the start comes from the 5 x 5 grid over lims first
 * (coarse_search_opt_angle / coarse_search, :340-395).  Out: irots[i] = the rotation, or 0 when nothing better than the start was found
 * (then cxy is not meaningful); cxy[i] = {cc, shift rotated to the reference frame}
 * (matmul(shift, rotmat2d(get_rot(irot)))).  No direct counterpart entry point: it replaces the
 * reference's per-particle calls in strategy2D_srch%inpl_srch / strategy3D_srch%inpl_srch
 * (simple_strategy2D_srch.f90:120-145, simple_strategy3D_srch.f90:190-205).  With opt_angle the
 * particles must be memoized (memoize_ffts), as for gencorrs. */
int pftcc_shsrch_grad_batch(pftcc_t* h, int32_t n, const int32_t* irefs, const int32_t* iptcls, const float* lims,
                            int32_t maxits, int32_t opt_angle, int32_t coarse_init, int32_t* irots, float* cxy);

/* -------------------------------------------------------- precision (MFMA) form -- */
/* Exhaustive search over references x rotations (gencorrs_1), or references x shifts x rotations
 * (gencorrs_2: shifts[nshifts][2], the reference norm over all shells), computed as one dense
 * complex contraction over shells per angular frequency m on MFMA, followed by the inverse FFT
 * over m; each (reference, shift) is one operand row.  Exact for ring-constant |CTF| (ctf = no or no astigmatism); with astigmatism the
 * particle's |CTF| is replaced by its ring mean per shell (the reference applies the particle's
 * angle-dependent |CTF| in the reference frame, :982, which no per-m contraction can express),
 * and ntop > 0 rescores the candidates exactly.
 * prec: PFTCC_PREC_SPLIT_BF16 (hi+lo split, FP32 accumulation), PFTCC_PREC_BF16, PFTCC_PREC_FP16,
 * PFTCC_PREC_SPLIT_FP16 or PFTCC_PREC_FP8.
 * Denominators sqrt(sum_k cbar_k^2 S_ref,k * S_ptcl) are formed in FP64.  No counterpart
 * entry point in the reference; it replaces the same greedy loop as pftcc_search
 * (simple_strategy3D_greedy.f90:60-77, gencorrs over all references + maxloc).
 * ntop = 0: the best (reference, rotation) of the contraction itself.  ntop in [1, 256]: the ntop
 * references with the highest contraction score per particle are rescored exactly in FP64 (the
 * reference's gencorrs formula with the particle's own |CTF|) at their best rotation and its two
 * neighbours, and the best rescored candidate is returned with its FP64 cc (rescoring needs
 * pftsz <= 1024, PFTCC_E_UNSUPPORTED otherwise). */
#define PFTCC_PREC_BF16 1
#define PFTCC_PREC_SPLIT_BF16 3
/* FP16 operands (one product) or split FP16 (hi+lo, three products), FP32 accumulation.  FP16 has
 * 5 exponent bits, so per-shell power-of-two scales balance the two operands and a global power of
 * two per operand maps its largest component to <= 2^15 (as for FP8); Z is scaled back exactly. */
#define PFTCC_PREC_FP16 2
#define PFTCC_PREC_SPLIT_FP16 4
/* FP8 e4m3 coarse pass: per-shell power-of-two scales balance the two operands, a global power of
 * two per operand maps its largest component to <= 448; meant for ntop > 0 (exact rescoring). */
#define PFTCC_PREC_FP8 8
int pftcc_search_mfma(pftcc_t* h, int32_t np, const int32_t* iptcls, int32_t nr, const int32_t* irefs,
                      int32_t nshifts, const float* shifts, int32_t prec, int32_t ntop, float* best_cc,
                      int32_t* best_iref, int32_t* best_ishift, int32_t* best_irot);
/* cc_out[np][nr][max(nshifts,1)][nrots] of the same computation (verification, small batches). */
int pftcc_gencorrs_mfma_batch(pftcc_t* h, int32_t np, const int32_t* iptcls, int32_t nr, const int32_t* irefs,
                              int32_t nshifts, const float* shifts, int32_t prec, float* cc_out);

/* ------------------------------------------------------------------- instrumentation -- */
/* No reference counterpart.  When enabled, each launch of the correlation kernels (the plain
 * FP32 search/gencorrs kernel) is bracketed by HIP events on the handle's stream;
 * pftcc_get_timing synchronises and returns the summed device time and launch count since the
 * last reset. */
int pftcc_set_timing(pftcc_t* h, int32_t enable);
int pftcc_get_timing(pftcc_t* h, double* total_ms, int64_t* launches);
/* Per kernel family: 0 plain correlation, 1 MFMA contraction, 2 inverse FFT of the MFMA form,
 * 3 operand packing, candidate selection and rescoring of the MFMA form, 4 the FP64 objective /
 * adjoint kernel (gencorr_for_rot_8, grad_batch). */
int pftcc_get_timing_channel(pftcc_t* h, int32_t channel, double* total_ms, int64_t* launches);

/* Tile of the plain FP32 correlation kernel for an exhaustive search with ns shifts (0 =
 * unshifted): *tile = (reference, shift) pairs per work-group of the compile-time kernel (0: the
 * runtime-size kernel runs), *staged = 1 when the staged-reference variant runs (a tile then spans
 * at most two references).  No reference counterpart (instrumentation). */
int pftcc_get_corr_tile(const pftcc_t* h, int32_t ns, int32_t* tile, int32_t* staged);

/* ------------------------------------------------------------------------- utilities -- */
/* magic_pftsz (simple_magic_boxes.f90:25-37). */
int32_t pftcc_magic_pftsz(int32_t ring2);

#ifdef __cplusplus
}
#endif
#endif /* PFTCC_H */
