! pftcc_fortran_driver - drives module simple_polarft_corrcalc the way SIMPLE's search code does.
!
!   pftcc_fortran_driver parity <in.bin> <out.bin>
!       reads references, particles, CTF parameters, shifts and single-rotation queries, runs
!       new / set_ref_pft / set_ptcl_pft / create_polar_absctfmats / memoize_ffts / gencorrs /
!       search / search_mfma / shsrch_grad_batch / strategy_srch / gencorr_for_rot_8 /
!       gencorr_grad_for_rot_8 / specscore and writes the results
!       (raw little-endian; layout in tests/test_gpu_fortran.py).
!   pftcc_fortran_driver omp <nthreads>
!       the per-pair calls issued from OpenMP threads as simple_strategy3D_matcher.f90:204 does
!       (!$omp parallel do over particles): gencorrs (with and without shift),
!       gencorr_grad_for_rot_8 and gencorr_cont_grad_cc_for_rot_8 for every (ref, ptcl) pair, once
!       serially and once in parallel; exits non-zero unless both give identical results.  Also
!       exercises the rest of the module's surface (cp_refs, assign_pinds, calc_ptcl_pspec, the
!       objective-specific and orientation-derivative forms).
!   pftcc_fortran_driver shiftsrch
!       the invariant of production/tests/test_shiftsrch (simple_test_shiftsrch.f90:40-118):
!       8 (ref, ptcl) pairs, the particle carrying the reference's phase for a +-1 px shift in
!       8 directions; a grid search over shifts in [-2,2]^2 with gencorrs must recover it.
!       Exits non-zero if any shift is not recovered.
program pftcc_fortran_driver
use, intrinsic :: iso_c_binding
use simple_polarft_corrcalc
implicit none
character(len=512) :: mode, fin, fout
call get_command_argument(1, mode)
select case(trim(mode))
    case('parity')
        call get_command_argument(2, fin)
        call get_command_argument(3, fout)
        call run_parity(trim(fin), trim(fout))
    case('shiftsrch')
        call run_shiftsrch
    case('omp')
        call get_command_argument(2, fin)
        call run_omp(trim(fin))
    case default
        write(*,*) 'usage: pftcc_fortran_driver parity <in> <out> | shiftsrch | omp <nthreads>'
        error stop 2
end select

contains

    subroutine run_parity( fin, fout )
        character(len=*), intent(in) :: fin, fout
        type(polarft_corrcalc) :: pftcc
        type(pftcc_params)     :: p
        integer(c_int32_t)     :: hdr(16)
        complex(c_float_complex), allocatable :: refse(:,:,:), refso(:,:,:), ptcls(:,:,:)
        type(ctfparams), allocatable :: cps(:)
        real(c_float),  allocatable :: shifts(:,:), cc(:), ccall(:,:,:,:), best_cc(:), scores(:), mcc(:)
        integer,        allocatable :: best_iref(:), best_irot(:), best_ishift(:), ids(:), miref(:), mirot(:)
        integer(c_int32_t), allocatable :: qref(:), qptcl(:), qrot(:)
        integer,        allocatable :: srefs(:), sptcls(:), srots(:)
        real(c_float),  allocatable :: scxy(:,:)
        real(c_float)               :: lims(2,2)
        real(c_double), allocatable :: qsh(:,:), f8(:), g8(:,:)
        type(strategy_cfg)          :: scfg
        real(c_float),      allocatable :: gbc(:), hbc(:), hprev(:)
        integer(c_int32_t), allocatable :: gbr(:), gbrot(:), gne(:), gfb(:), gsorted(:,:)
        integer(c_int32_t), allocatable :: hbr(:), hbrot(:), hne(:), hfb(:), hsorted(:,:)
        integer,            allocatable :: horder(:,:)
        integer :: nrefs, nptcls, n, kfrom, kto, kstop, box, withctf, nsh, nq, pfrom, nk, nrots
        integer :: u, i, r, s, q
        open(newunit=u, file=fin, access='stream', form='unformatted', status='old')
        read(u) hdr
        nrefs = hdr(1); nptcls = hdr(2); n = hdr(3); kfrom = hdr(4); kto = hdr(5); kstop = hdr(6)
        box = hdr(7); withctf = hdr(8); nsh = hdr(9); nq = hdr(10); pfrom = hdr(11)
        nk = kto - kfrom + 1
        nrots = 2 * n
        allocate(refse(n, kfrom:kto, nrefs), refso(n, kfrom:kto, nrefs), ptcls(n, kfrom:kto, nptcls))
        read(u) refse
        read(u) refso
        read(u) ptcls
        allocate(cps(nptcls))
        if( withctf /= 0 ) read(u) cps
        allocate(shifts(2, max(nsh,1)))
        if( nsh > 0 ) read(u) shifts(:, 1:nsh)
        allocate(qref(max(nq,1)), qptcl(max(nq,1)), qrot(max(nq,1)), qsh(2, max(nq,1)))
        if( nq > 0 )then
            read(u) qref(1:nq)
            read(u) qptcl(1:nq)
            read(u) qrot(1:nq)
            read(u) qsh(:, 1:nq)
        endif
        close(u)
        ! the calls SIMPLE makes (preppftcc4align / build_pftcc_batch_particles / srch)
        p%boxmatch = box
        p%pftsz    = n
        p%kfromto  = [kfrom, kto]
        p%kstop    = kstop
        p%with_ctf = withctf /= 0
        call pftcc%new(nrefs, [pfrom, pfrom + nptcls - 1], p)
        do r = 1, nrefs
            call pftcc%set_ref_pft(r, refse(:,:,r), .true.)
            call pftcc%set_ref_pft(r, refso(:,:,r), .false.)
        end do
        do i = 1, nptcls
            call pftcc%set_ptcl_pft(pfrom + i - 1, ptcls(:,:,i))
        end do
        if( withctf /= 0 ) call pftcc%create_polar_absctfmats(cps)
        call pftcc%memoize_ffts
        allocate(cc(nrots), ccall(nrots, 0:nsh, nrefs, nptcls))
        do i = 1, nptcls
            do r = 1, nrefs
                call pftcc_glob%gencorrs(r, pfrom + i - 1, cc)
                ccall(:, 0, r, i) = cc
                do s = 1, nsh
                    call pftcc_glob%gencorrs(r, pfrom + i - 1, shifts(:, s), cc)
                    ccall(:, s, r, i) = cc
                end do
            end do
        end do
        allocate(best_cc(nptcls), best_iref(nptcls), best_irot(nptcls), best_ishift(nptcls), ids(nptcls))
        ids = [(pfrom + i - 1, i = 1, nptcls)]
        if( nsh > 0 )then
            call pftcc%search(ids, best_cc, best_iref, best_irot, shifts=shifts(:, 1:nsh), best_ishift=best_ishift)
        else
            call pftcc%search(ids, best_cc, best_iref, best_irot)
            best_ishift = 0
        endif
        ! precision form with exact rescoring of the 3 best references
        allocate(mcc(nptcls), miref(nptcls), mirot(nptcls))
        call pftcc%search_mfma(ids, mcc, miref, mirot, PREC_SPLIT_BF16, 3)
        ! gradient shift search of every (reference, particle) pair, in-plane callback on
        allocate(srefs(nrefs*nptcls), sptcls(nrefs*nptcls), srots(nrefs*nptcls), scxy(3, nrefs*nptcls))
        q = 0
        do i = 1, nptcls
            do r = 1, nrefs
                q = q + 1
                srefs(q)  = r
                sptcls(q) = pfrom + i - 1
            end do
        end do
        srots = 0
        lims(:,1) = -3.
        lims(:,2) =  3.
        call pftcc%shsrch_grad_batch(srefs, sptcls, lims, 60, .true., srots, scxy)
        ! search strategies: greedy 3-D with every reference sorted, and shc 3-D over the
        ! references in reversed order with prev_corr 0
        allocate(gbc(nptcls), gbr(nptcls), gbrot(nptcls), gne(nptcls), gfb(nptcls), gsorted(nrefs, nptcls))
        allocate(hbc(nptcls), hbr(nptcls), hbrot(nptcls), hne(nptcls), hfb(nptcls), hsorted(nrefs, nptcls))
        allocate(horder(nptcls, nrefs), hprev(nptcls))
        scfg%kind = SRCH_GREEDY3D
        scfg%ntop = nrefs
        call pftcc%strategy_srch(scfg, ids, gbc, gbr, gbrot, gne, gfb, sorted_refs=gsorted)
        do i = 1, nptcls
            horder(i,:) = [(nrefs - r + 1, r = 1, nrefs)]
        end do
        hprev = 0.
        scfg%kind = SRCH_SHC3D
        call pftcc%strategy_srch(scfg, ids, hbc, hbr, hbrot, hne, hfb, srch_order=horder, prev_corr=hprev, &
            &sorted_refs=hsorted)
        allocate(f8(max(nq,1)), g8(2, max(nq,1)), scores(max(nq,1)))
        do q = 1, nq
            f8(q) = pftcc%gencorr_for_rot_8(qref(q), qptcl(q), qsh(:,q), qrot(q))
            call pftcc%gencorr_grad_for_rot_8(qref(q), qptcl(q), qsh(:,q), qrot(q), f8(q), g8(:,q))
            scores(q) = pftcc%specscore(qref(q), qptcl(q), qrot(q))
        end do
        call pftcc%kill
        open(newunit=u, file=fout, access='stream', form='unformatted', status='replace')
        write(u) ccall
        write(u) best_cc
        write(u) int(best_iref, c_int32_t)
        write(u) int(best_ishift, c_int32_t)
        write(u) int(best_irot, c_int32_t)
        if( nq > 0 )then
            write(u) f8(1:nq)
            write(u) g8(:, 1:nq)
            write(u) scores(1:nq)
        endif
        write(u) mcc
        write(u) int(miref, c_int32_t)
        write(u) int(mirot, c_int32_t)
        write(u) scxy
        write(u) int(srots, c_int32_t)
        write(u) gbc
        write(u) gbr
        write(u) gbrot
        write(u) gsorted
        write(u) hbc
        write(u) hbr
        write(u) hne
        write(u) hfb
        write(u) hsorted
        close(u)
        write(*,'(a,i0,a,i0,a,i0,a)') 'parity run done: ', nptcls, ' ptcls x ', nrefs, ' refs x ', nsh, ' shifts'
    end subroutine run_parity

    subroutine run_shiftsrch
        integer, parameter :: NPAIR = 8, BOX = 64, KFROM = 2, KTO = 20
        real,    parameter :: SHMAG = 1.0, PI = acos(-1.0)
        type(polarft_corrcalc) :: pftcc
        type(pftcc_params)     :: p
        complex(c_float_complex), allocatable :: ref(:,:), ptcl(:,:)
        real(c_float), allocatable :: corrs(:)
        real    :: dirs(2, NPAIR), corr, corrmax, xy(2), ph, rnd(2)
        integer :: i, j, k, xsh, ysh, xbest, ybest, nbad, n, seed(64)
        dirs = reshape([SHMAG,0., 0.,SHMAG, -SHMAG,0., 0.,-SHMAG, SHMAG,SHMAG, -SHMAG,-SHMAG, &
            &-SHMAG,SHMAG, SHMAG,-SHMAG], [2, NPAIR])
        p%boxmatch = BOX
        p%ring2    = 30
        p%kfromto  = [KFROM, KTO]
        p%with_ctf = .false.
        call pftcc%new(NPAIR, [1, NPAIR], p)
        n = pftcc%get_pftsz()
        allocate(ref(n, KFROM:KTO), ptcl(n, KFROM:KTO), corrs(pftcc%get_nrots()))
        seed = 42
        call random_seed(put=seed)
        do i = 1, NPAIR
            do k = KFROM, KTO
                do j = 1, n
                    call random_number(rnd)
                    ref(j,k) = cmplx(rnd(1) - 0.5, rnd(2) - 0.5) / real(k)
                    ! the particle carries the phase exp(i pi (x sx + y sy) / (box/2)) of the shift
                    xy = pftcc%get_coord(j, k)
                    ph = PI * (xy(1) * dirs(1,i) + xy(2) * dirs(2,i)) / real(BOX / 2)
                    ptcl(j,k) = ref(j,k) * cmplx(cos(ph), sin(ph))
                end do
            end do
            call pftcc%set_ref_pft(i, ref, .true.)
            call pftcc%set_ptcl_pft(i, ptcl)
        end do
        call pftcc%memoize_ffts
        nbad = 0
        do i = 1, NPAIR
            corrmax = -1.
            do xsh = -2, 2
                do ysh = -2, 2
                    call pftcc%gencorrs(i, i, real([xsh, ysh]), corrs)
                    corr = maxval(corrs)
                    if( corr > corrmax )then
                        corrmax = corr
                        xbest   = xsh
                        ybest   = ysh
                    endif
                end do
            end do
            write(*,'(a,i0,a,2i3,a,2f5.1,a,f9.6)') 'pair ', i, ' best shift', xbest, ybest, ' applied', &
                &dirs(:,i), ' corr', corrmax
            if( xbest /= nint(dirs(1,i)) .or. ybest /= nint(dirs(2,i)) .or. corrmax < 0.9999 ) nbad = nbad + 1
        end do
        call pftcc%kill
        if( nbad > 0 )then
            write(*,'(a,i0,a)') 'shiftsrch: ', nbad, ' shifts not recovered'
            error stop 1
        endif
        write(*,'(a)') 'shiftsrch: all shifts recovered'
    end subroutine run_shiftsrch

    subroutine run_omp( nthr_arg )
        use omp_lib
        character(len=*), intent(in) :: nthr_arg
        integer, parameter :: sp = c_float, dp = c_double
        integer, parameter :: NREFS = 6, NPTCLS = 12, BOX = 64, KFROM = 2, KTO = 20
        type(polarft_corrcalc) :: pftcc, pftcc2
        type(pftcc_params)     :: p
        type(ctfparams)        :: cps(NPTCLS)
        complex(c_float_complex), allocatable :: ref(:,:), ptcl(:,:)
        complex(sp) :: dcomp(3)
        real(c_float),  allocatable :: cc_s(:,:,:,:), cc_p(:,:,:,:), pspec(:)
        real(c_double), allocatable :: f_s(:,:), f_p(:,:), g_s(:,:,:), g_p(:,:,:)
        real(c_double), allocatable :: c_s(:,:), c_p(:,:), d_s(:,:,:), d_p(:,:,:)
        integer, allocatable :: pinds(:)
        real    :: rnd(2), shv(2)
        real(dp) :: sh8(2), f8, g8(2), g5(5), ce, de(3)
        integer :: i, r, j, k, n, nrots, nthr, irot, seed(64), nbad
        read(nthr_arg, *) nthr
        p%boxmatch = BOX
        p%ring2    = 30
        p%kfromto  = [KFROM, KTO]
        p%kstop    = KTO
        p%with_ctf = .true.
        p%nthr     = NREFS
        call pftcc%new(NREFS, [1, NPTCLS], p)
        n     = pftcc%get_pftsz()
        nrots = pftcc%get_nrots()
        allocate(ref(n, KFROM:KTO), ptcl(n, KFROM:KTO))
        seed = 7
        call random_seed(put=seed)
        do r = 1, NREFS
            do k = KFROM, KTO
                do j = 1, n
                    call random_number(rnd)
                    ref(j,k) = cmplx(rnd(1) - 0.5, rnd(2) - 0.5) / real(k)
                end do
            end do
            call pftcc%set_ref_pft(r, ref, .true.)
            call pftcc%cp_even2odd_ref(r)
            do k = KFROM, KTO
                do j = 1, n
                    call random_number(rnd)
                    dcomp = cmplx(rnd(1) - 0.5, rnd(2) - 0.5) / real(k)
                    call pftcc%set_dref_fcomp(r, j, k, dcomp, .true.)
                    call pftcc%set_dref_fcomp(r, j, k, dcomp, .false.)
                end do
            end do
        end do
        do i = 1, NPTCLS
            do k = KFROM, KTO
                do j = 1, n
                    call random_number(rnd)
                    ptcl(j,k) = cmplx(rnd(1) - 0.5, rnd(2) - 0.5) / real(k)
                end do
            end do
            call pftcc%set_ptcl_pft(i, ptcl)
            call random_number(rnd)
            cps(i)%smpd = 1.3
            cps(i)%dfx  = 1.0 + 2.0 * rnd(1)
            cps(i)%dfy  = cps(i)%dfx - 0.1 * rnd(2)
            cps(i)%angast = 30.
        end do
        call pftcc%set_eos([(mod(i, 2), i = 1, NPTCLS)])
        call pftcc%create_polar_absctfmats(cps)
        call pftcc%memoize_ffts
        allocate(cc_s(nrots, 2, NREFS, NPTCLS), cc_p(nrots, 2, NREFS, NPTCLS))
        allocate(f_s(NREFS, NPTCLS), f_p(NREFS, NPTCLS), g_s(2, NREFS, NPTCLS), g_p(2, NREFS, NPTCLS))
        allocate(c_s(NREFS, NPTCLS), c_p(NREFS, NPTCLS), d_s(3, NREFS, NPTCLS), d_p(3, NREFS, NPTCLS))
        call pairs(1, cc_s, f_s, g_s, c_s, d_s)
        call pairs(nthr, cc_p, f_p, g_p, c_p, d_p)
        nbad = count(cc_s /= cc_p) + count(f_s /= f_p) + count(g_s /= g_p) + count(c_s /= c_p) + count(d_s /= d_p)
        write(*,'(a,i0,a,i0,a,i0)') 'omp: ', NREFS * NPTCLS, ' pairs on ', nthr, ' threads, mismatches: ', nbad
        ! the rest of the surface: each call must succeed (errors stop the program)
        call pftcc%assign_pinds(pinds)
        if( lbound(pinds,1) /= 1 .or. ubound(pinds,1) /= NPTCLS .or. any(pinds /= [(i, i = 1, NPTCLS)]) ) nbad = nbad + 1
        allocate(pspec(KFROM:KTO))
        call pftcc%calc_ptcl_pspec(3, pspec)
        if( any(pspec <= 0.) ) nbad = nbad + 1
        sh8 = [0.5_dp, -0.25_dp]
        f8  = pftcc%gencorr_cc_for_rot_8(2, 3, sh8, 5)
        if( f8 /= pftcc%gencorr_for_rot_8(2, 3, sh8, 5) ) nbad = nbad + 1
        call pftcc%gencorr_cc_grad_for_rot_8(2, 3, sh8, 5, f8, g8)
        call pftcc%gencorr_cont_shift_grad_cc_for_rot_8(2, 3, sh8, 5, f8, g5)
        ce = pftcc%gencorr_cont_grad_euclid_for_rot_8(2, 3, sh8, 5, de)
        if( ce /= 0._dp .or. any(de /= 0._dp) ) nbad = nbad + 1
        call pftcc2%new(NREFS, [1, NPTCLS], p)
        call pftcc2%cp_refs(pftcc)
        call pftcc2%kill
        call pftcc%kill
        if( nbad > 0 )then
            write(*,'(a,i0)') 'omp: FAILED, mismatches ', nbad
            error stop 1
        endif
        write(*,'(a)') 'omp: parallel per-pair calls equal the serial ones'

    contains

        subroutine pairs( nt, cc, f, g, c, d )
            integer,        intent(in)  :: nt
            real(c_float),  intent(out) :: cc(nrots, 2, NREFS, NPTCLS)
            real(c_double), intent(out) :: f(NREFS, NPTCLS), g(2, NREFS, NPTCLS), c(NREFS, NPTCLS), d(3, NREFS, NPTCLS)
            integer  :: ip, ir, ir8
            real     :: s2(2)
            real(dp) :: s8(2)
            !$omp parallel do num_threads(nt) default(shared) private(ip,ir,ir8,s2,s8) schedule(static)
            do ip = 1, NPTCLS
                do ir = 1, NREFS
                    s2  = [0.25 * real(ir), -0.5 * real(ip) / real(NPTCLS)]
                    s8  = real(s2, dp)
                    ir8 = 1 + mod(7 * ip + 3 * ir, nrots)
                    call pftcc%gencorrs(ir, ip, cc(:, 1, ir, ip))
                    call pftcc%gencorrs(ir, ip, s2, cc(:, 2, ir, ip))
                    call pftcc%gencorr_grad_for_rot_8(ir, ip, s8, ir8, f(ir, ip), g(:, ir, ip))
                    c(ir, ip) = pftcc%gencorr_cont_grad_cc_for_rot_8(ir, ip, s8, ir8, d(:, ir, ip))
                end do
            end do
            !$omp end parallel do
        end subroutine pairs

    end subroutine run_omp

end program pftcc_fortran_driver
