! Fortran side of the drop-in boundary: modules with the reference's names and argument lists that forward,
! through iso_c_binding, to libsmcf (include/smcf.h).
!
!   module type_mat   mat.f90:1-7 (same derived type)
!   module smcf_c     C interfaces, the cached device context and problems, the keyed-RNG counters
!   module gfunc      gen_unit_slip, del_unit_slip, call_dc3d0, calc_responce, calc_responce_dist,
!                     calc_greens_func                                        gfunc.f90:6-348
!   module smc_slip   cdf_norm, pdf_norm, slip_calc_likelihood, slip_calc_prior, slip_calc_grad,
!                     slip_gen_init_particles, slip_calc_mean_std_vector, slip_find_next_gamma,
!                     slip_normalize_weights, slip_calc_mean_particles, slip_calc_cov_particles,
!                     slip_resample_particles, slip_leapfrog, slip_calc_gsd_gsg, slip_smc_exec
!                                                                             smc_slip.f90:40-1663
!   module smc_fault  fault_calc_neglog_q, fault_sample_init_particles, fault_calc_likelihood,
!                     work_eval_init_particles, calc_mean_std_vector, fault_find_next_gamma,
!                     fault_normalize_weights, fault_calc_mean_particles, fault_calc_cov_particles,
!                     fault_resample_particles, fault_reorder_to_send, work_mcmc_sampling, fault_smc_exec
!                                                                             smc_fault.f90:9-1075
!
! The computation runs in libsmcf (the device, and the host code of the outer stage).  Procedures whose reference
! argument lists carry the observation set and fault description (calc_greens_func, fault_calc_likelihood,
! work_eval_init_particles, work_mcmc_sampling, fault_smc_exec) (re)build the cached device problem from them;
! slip_smc_exec, slip_gen_init_particles, slip_leapfrog and slip_calc_gsd_gsg, whose lists lack the station
! coordinates, use the problem cached by such a call (as in the reference's own call order) and stop with a message
! when there is none.  Workspace arguments of the reference (slip_dist, response_dist, target_id_val, st_rand, ...)
! are accepted and left untouched: their contents are scratch in the reference too.
!
! Random numbers are keyed, not drawn from the intrinsic stream: smcf_fortran_seed(seed) sets the seed.  Inner SMC
! runs of fault_calc_likelihood / slip_smc_exec are keyed by (seed, call counter); the outer procedures by (seed,
! outer stage, particle), with the stage counted by fault_find_next_gamma and reset by fault_sample_init_particles,
! so a caller composing the reference's outer loop from these procedures gets what fault_smc_exec gets.
! The inner mover is the reference's (HMC, ntau 50, ESJD tuning) unless smcf_fortran_mover selects another.
module type_mat
    implicit none
    type mat
        double precision, allocatable :: body(:, :)
        integer :: nrow, ncol
    end type
end module type_mat

module smcf_c
    use iso_c_binding
    use type_mat
    implicit none

    type, bind(C) :: smcf_problem_desc
        integer(c_int) :: nplane
        type(c_ptr) :: nxi_ls, neta_ls
        real(c_double) :: xmin, xmax, zmin
        integer(c_int) :: fix_xbend
        type(c_ptr) :: xbend
        integer(c_int) :: nobs, nsar_total, ngnss, npath
        type(c_ptr) :: nsar_index, obs_points, obs_unitvec, obs_sigma, dvec, sar_sinv
        real(c_double) :: max_slip
        integer(c_int) :: nparticle_slip, mover, gibbs_sweeps, precision, hmc_ntau
        real(c_double) :: hmc_accept
    end type

    interface
        integer(c_int) function smcf_ctx_create(device, out) bind(C, name="smcf_ctx_create")
            import :: c_int, c_ptr
            integer(c_int), value :: device
            type(c_ptr) :: out
        end function
        integer(c_int) function smcf_problem_create(ctx, desc, max_batch, out) bind(C, name="smcf_problem_create")
            import :: c_int, c_ptr, smcf_problem_desc
            type(c_ptr), value :: ctx
            type(smcf_problem_desc) :: desc
            integer(c_int), value :: max_batch
            type(c_ptr) :: out
        end function
        integer(c_int) function smcf_problem_destroy(pb) bind(C, name="smcf_problem_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: pb
        end function
        integer(c_int) function smcf_calc_greens_func(pb, b, theta, gmat) bind(C, name="smcf_calc_greens_func")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: pb
            integer(c_int), value :: b
            real(c_double) :: theta(*), gmat(*)
        end function
        integer(c_int) function smcf_laplacian(pb, theta, li, lv, lti, ltv, ll) bind(C, name="smcf_laplacian")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: pb
            real(c_double) :: theta(*), lv(*), ltv(*), ll(*)
            integer(c_int) :: li(*), lti(*)
        end function
        integer(c_int) function smcf_discretize_fault(pb, theta, cny, coor, n2e, n2es, id_dof, luni, lmat) &
            bind(C, name="smcf_discretize_fault")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: pb
            real(c_double) :: theta(*), coor(*), luni(*), lmat(*)
            integer(c_int) :: cny(*), n2e(*), n2es(*), id_dof(*)
        end function
        integer(c_int) function smcf_call_dc3d0(ctx, n, in, uret) bind(C, name="smcf_call_dc3d0")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: n
            real(c_double) :: in(*), uret(*)
        end function
        integer(c_int) function smcf_calc_responce(ctx, npts, pts, npatch, cny, nnode, coor, slip_dist, xf, yf, &
                                                   zf, strike, dip, ntarget, target_id, uobs) &
            bind(C, name="smcf_calc_responce")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: npts, npatch, nnode, ntarget
            real(c_double) :: pts(*), coor(*), slip_dist(*), uobs(*)
            real(c_double), value :: xf, yf, zf, strike, dip
            integer(c_int) :: cny(*), target_id(*)
        end function
        integer(c_int) function smcf_slip_smc_exec(pb, theta, key, particles, likelihood, prior, mean, nstage, &
                                                   neglog, gamma_trace) bind(C, name="smcf_slip_smc_exec")
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: pb
            real(c_double) :: theta(*), particles(*), likelihood(*), prior(*), mean(*), neglog, gamma_trace(*)
            integer(c_int64_t), value :: key
            integer(c_int) :: nstage
        end function
        integer(c_int) function smcf_fault_calc_likelihood_batch(pb, b, theta, keys, neglog, nstage) &
            bind(C, name="smcf_fault_calc_likelihood_batch")
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: pb
            integer(c_int), value :: b
            real(c_double) :: theta(*), neglog(*)
            integer(c_int64_t) :: keys(*)
            integer(c_int) :: nstage(*)
        end function
        integer(c_int) function smcf_slip_likelihood_gram(ctx, ndim, n, particles, gsgmat, gsdvec, dsd, nsar, &
                                                          ngnss, lss2, lsg2, out) bind(C, name="smcf_slip_likelihood_gram")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: ndim, n, nsar, ngnss
            real(c_double) :: particles(*), gsgmat(*), gsdvec(*), out(*)
            real(c_double), value :: dsd, lss2, lsg2
        end function
        integer(c_int) function smcf_slip_prior_ell(ctx, nnode, ndim, n, particles, alpha2, li, lv, out) &
            bind(C, name="smcf_slip_prior_ell")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nnode, ndim, n
            real(c_double) :: particles(*), alpha2(*), lv(*), out(*)
            integer(c_int) :: li(*)
        end function
        integer(c_int) function smcf_slip_grad_ell(ctx, nnode, ndim, n, particles, gamma, gsgmat, gsdvec, alpha2, &
                                                   li, lv, lti, ltv, grad) bind(C, name="smcf_slip_grad_ell")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nnode, ndim, n
            real(c_double), value :: gamma
            real(c_double) :: particles(*), gsgmat(*), gsdvec(*), alpha2(*), lv(*), ltv(*), grad(*)
            integer(c_int) :: li(*), lti(*)
        end function
        integer(c_int) function smcf_slip_gen_init_particles(pb, theta, key, particles, likelihood, prior) &
            bind(C, name="smcf_slip_gen_init_particles")
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: pb
            integer(c_int64_t), value :: key
            real(c_double) :: theta(*), particles(*), likelihood(*), prior(*)
        end function
        integer(c_int) function smcf_slip_calc_moments(ctx, ndim, np, particles, weights, mean, cov) &
            bind(C, name="smcf_slip_calc_moments")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: ndim, np
            real(c_double) :: particles(*), weights(*), mean(*), cov(*)
        end function
        integer(c_int) function smcf_find_next_gamma(ctx, np, gprev, like, cv, tol, cap, gamma, w, nev) &
            bind(C, name="smcf_find_next_gamma")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: np, cap
            real(c_double), value :: gprev, cv, tol
            real(c_double) :: like(*), gamma, w(*), nev
        end function
        integer(c_int) function smcf_normalize_resample(ctx, np, w, uh, ul, wn, assigned) &
            bind(C, name="smcf_normalize_resample")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: np
            real(c_double), value :: uh, ul
            real(c_double) :: w(*), wn(*)
            integer(c_int) :: assigned(*)
        end function
        integer(c_int) function smcf_philox(ctx, n, ctr, key, out) bind(C, name="smcf_philox")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: n
            integer(c_int32_t) :: ctr(*), key(*), out(*)
        end function
        integer(c_int) function smcf_slip_leapfrog(pb, theta, gamma, n, cur, p0, cov, dtau, ntau, cand, pout, hc, &
                                                   hn, lk, pr) bind(C, name="smcf_slip_leapfrog")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: pb
            real(c_double), value :: gamma, dtau
            integer(c_int), value :: n, ntau
            real(c_double) :: theta(*), cur(*), p0(*), cov(*), cand(*), pout(*), hc(*), hn(*), lk(*), pr(*)
        end function
        integer(c_int) function smcf_slip_calc_gsd_gsg_gmat(pb, gmat, ss2, sg2, gsgmat, gsdvec, dsd) &
            bind(C, name="smcf_slip_calc_gsd_gsg_gmat")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: pb
            real(c_double), value :: ss2, sg2
            real(c_double) :: gmat(*), gsgmat(*), gsdvec(*), dsd
        end function
        integer(c_int) function smcf_fault_sample_init_particles(nd, n, seed, range, qmean, qchol, particles) &
            bind(C, name="smcf_fault_sample_init_particles")
            import :: c_int, c_ptr, c_double, c_int64_t
            integer(c_int), value :: nd, n
            integer(c_int64_t), value :: seed
            type(c_ptr), value :: range, qmean, qchol
            real(c_double) :: particles(*)
        end function
        integer(c_int) function smcf_fault_mean_std(n, vec, mean, std) bind(C, name="smcf_fault_mean_std")
            import :: c_int, c_double
            integer(c_int), value :: n
            real(c_double) :: vec(*), mean, std
        end function
        integer(c_int) function smcf_fault_find_next_gamma(n, gprev, like, q, w, gamma, nev) &
            bind(C, name="smcf_fault_find_next_gamma")
            import :: c_int, c_ptr, c_double
            integer(c_int), value :: n
            real(c_double), value :: gprev
            real(c_double) :: like(*), w(*), gamma, nev
            type(c_ptr), value :: q
        end function
        integer(c_int) function smcf_fault_normalize_weights(n, w) bind(C, name="smcf_fault_normalize_weights")
            import :: c_int, c_double
            integer(c_int), value :: n
            real(c_double) :: w(*)
        end function
        integer(c_int) function smcf_fault_calc_mean_particles(nd, n, x, w, mean) &
            bind(C, name="smcf_fault_calc_mean_particles")
            import :: c_int, c_double
            integer(c_int), value :: nd, n
            real(c_double) :: x(*), w(*), mean(*)
        end function
        integer(c_int) function smcf_fault_calc_cov_particles(nd, n, x, w, mean, cov, info) &
            bind(C, name="smcf_fault_calc_cov_particles")
            import :: c_int, c_double
            integer(c_int), value :: nd, n
            real(c_double) :: x(*), w(*), mean(*), cov(*)
            integer(c_int) :: info
        end function
        integer(c_int) function smcf_fault_resample_particles(n, w, seed, stage, assigned) &
            bind(C, name="smcf_fault_resample_particles")
            import :: c_int, c_double, c_int64_t
            integer(c_int), value :: n, stage
            integer(c_int64_t), value :: seed
            real(c_double) :: w(*)
            integer(c_int) :: assigned(*)
        end function
        integer(c_int) function smcf_fault_reorder_to_send(nd, n, assigned, x, like, q, sidx) &
            bind(C, name="smcf_fault_reorder_to_send")
            import :: c_int, c_ptr, c_double
            integer(c_int), value :: nd, n
            integer(c_int) :: assigned(*), sidx(*)
            real(c_double) :: x(*), like(*)
            type(c_ptr), value :: q
        end function
        integer(c_int) function smcf_work_mcmc_sampling(pb, seed, stage, gamma, range, chol, n, assigned, x, q, &
                                                        qmean, qinv, logdet, max_batch, xn, ln, qn, acc) &
            bind(C, name="smcf_work_mcmc_sampling")
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: pb, q, qmean, qinv
            integer(c_int64_t), value :: seed
            integer(c_int), value :: stage, n, max_batch
            real(c_double), value :: gamma, logdet
            real(c_double) :: range(*), chol(*), x(*), xn(*), ln(*), qn(*)
            integer(c_int) :: assigned(*), acc
        end function
        integer(c_int) function smcf_fault_smc_set_output_dir(pb, dir) bind(C, name="smcf_fault_smc_set_output_dir")
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: pb
            character(kind=c_char) :: dir(*)
        end function
        integer(c_int) function smcf_fault_smc_exec(pb, range, n, seed, max_stages, max_batch, particles, like, &
                                                    nstage, evidence, gtrace) bind(C, name="smcf_fault_smc_exec")
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: pb
            integer(c_int), value :: n, max_stages, max_batch
            integer(c_int64_t), value :: seed
            real(c_double) :: range(*), particles(*), like(*), evidence, gtrace(*)
            integer(c_int) :: nstage
        end function
        integer(c_int) function smcf_fault_smc_exec_gsmc(pb, range, n, seed, max_stages, max_batch, qmean, qchol, &
                                                         qinv, logdet, particles, like, nstage, evidence, gtrace) &
            bind(C, name="smcf_fault_smc_exec_gsmc")
            import :: c_int, c_ptr, c_double, c_int64_t
            type(c_ptr), value :: pb
            integer(c_int), value :: n, max_stages, max_batch
            integer(c_int64_t), value :: seed
            real(c_double), value :: logdet
            real(c_double) :: range(*), qmean(*), qchol(*), qinv(*), particles(*), like(*), evidence, gtrace(*)
            integer(c_int) :: nstage
        end function
        integer(c_int64_t) function smcf_inner_key(seed, stage, role, slot) bind(C, name="smcf_inner_key")
            import :: c_int64_t, c_int32_t
            integer(c_int64_t), value :: seed
            integer(c_int32_t), value :: stage, role, slot
        end function
        type(c_ptr) function smcf_last_error() bind(C, name="smcf_last_error")
            import :: c_ptr
        end function
    end interface

    ! cached context and problems: g_pb for evaluations (the caller's data and sampler settings), g_pbg for
    ! Green's matrices alone (unit sigma, zero data), so that the two kinds of call do not evict each other
    type(c_ptr), save :: g_ctx = c_null_ptr
    type(c_ptr), save :: g_pb = c_null_ptr, g_pbg = c_null_ptr
    integer, save :: g_sig(10) = -1, g_sigg(10) = -1
    double precision, save :: g_sum = -1d300, g_sumg = -1d300
    integer(c_int64_t), save :: g_seed = 0_c_int64_t
    integer(c_int32_t), save :: g_counter = 0, g_draws = 0
    integer, save :: g_stage = 0
    ! default: the reference mover (HMC, ntau 50, ESJD tuning, smc_slip.f90:591-1214)
    integer, save :: g_mover = 0, g_sweeps = 5, g_max_batch = 64, g_precision = 0, g_hmc_ntau = 10
    double precision, save :: g_hmc_accept = 0.65d0
    integer, save :: g_outer_max_stages = 1000
    ! arrays kept alive for the C side, per cached problem
    type keep_t
        integer(c_int), allocatable :: nxi(:), neta(:), nsi(:)
        real(c_double), allocatable :: xb(:), pts(:), uv(:), sig(:), d(:), sar(:)
    end type
    type(keep_t), target, save :: t_keep, t_keepg

contains
    subroutine smcf_fortran_seed(seed)
        integer(c_int64_t), intent(in) :: seed
        g_seed = seed
        g_counter = 0
        g_draws = 0
        g_stage = 0
    end subroutine

    ! inner mover: 0 = the reference HMC (default), 1 = truncated-MVN Gibbs (`sweeps` per stage), 2 = HMC with
    ! smcf_fortran_hmc's steps and target acceptance; max_batch = geometries per library call
    subroutine smcf_fortran_mover(mover, sweeps, max_batch)
        integer, intent(in) :: mover, sweeps, max_batch
        g_mover = mover
        g_sweeps = sweeps
        g_max_batch = max_batch
        g_sig = -1
    end subroutine

    subroutine smcf_fortran_hmc(ntau, accept)
        integer, intent(in) :: ntau
        double precision, intent(in) :: accept
        g_hmc_ntau = ntau
        g_hmc_accept = accept
        g_sig = -1
    end subroutine

    ! stage cap of fault_smc_exec (the reference runs to gamma = 1; default 1000)
    subroutine smcf_fortran_outer_stages(max_stages)
        integer, intent(in) :: max_stages
        g_outer_max_stages = max_stages
    end subroutine

    ! arithmetic of the Gibbs mover: 0 = FP64 (default), 1 = FP32 matrix products, 2 = split-BF16 products
    subroutine smcf_fortran_precision(precision)
        integer, intent(in) :: precision
        g_precision = precision
        g_sig = -1
    end subroutine

    subroutine smcf_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        integer :: n
        if (rc /= 0) then
            call c_f_pointer(smcf_last_error(), msg, [512])
            n = 0
            do while (n < 512)
                if (msg(n + 1) == c_null_char) exit
                n = n + 1
            end do
            print *, "libsmcf error in ", where, ": ", msg(1:n)
            stop 1
        end if
    end subroutine

    subroutine smcf_ensure_ctx()
        if (.not. c_associated(g_ctx)) call smcf_check(smcf_ctx_create(0_c_int, g_ctx), "smcf_ctx_create")
    end subroutine

    subroutine smcf_need_problem(where)
        character(*), intent(in) :: where
        if (.not. c_associated(g_pb)) then
            print *, where, ": no problem cached; call calc_greens_func or fault_calc_likelihood with the ", &
                "observation set first"
            stop 1
        end if
    end subroutine

    ! (re)create a cached device problem when the observation set, mesh or settings change
    ! sigma_sar_mat (optional): the inverse InSAR covariance per path (calc_sigma_sar_mat); without it the
    ! blocks are identities, which is enough for Green's matrices alone.  greens_only selects the Green's cache.
    subroutine smcf_ensure_problem(nplane, nxi_ls, neta_ls, xmin, xmax, zmin, fix_xbend, xbend, nobs, &
                                   nsar_total, ngnss, npath, nsar_index, obs_points, obs_unitvec, &
                                   obs_sigma, dvec, max_slip, nparticle_slip, sigma_sar_mat, greens_only)
        integer, intent(in) :: nplane, nxi_ls(:), neta_ls(:), nobs, nsar_total, ngnss, npath, nsar_index(:), &
                               nparticle_slip
        double precision, intent(in) :: xmin, xmax, zmin, xbend(:), obs_points(:, :), obs_unitvec(:, :), &
                                        obs_sigma(:), dvec(:), max_slip
        logical, intent(in) :: fix_xbend
        type(mat), intent(in), optional :: sigma_sar_mat(:)
        logical, intent(in), optional :: greens_only
        type(smcf_problem_desc) :: d
        integer :: sig(10), ip, n, off, k, nsinv
        double precision :: s
        logical :: gr
        gr = .false.
        if (present(greens_only)) gr = greens_only
        sig = [nplane, nobs, nsar_total, ngnss, npath, nparticle_slip, sum(nxi_ls) + 1000*sum(neta_ls), &
               merge(1, 0, fix_xbend), g_mover + 10*g_precision + 100*g_sweeps, g_hmc_ntau]
        s = sum(obs_points) + 3d0*sum(obs_unitvec) + 7d0*sum(obs_sigma) + 11d0*sum(dvec) + xmin + 2d0*xmax &
            + 5d0*zmin + max_slip + 17d0*g_hmc_accept
        nsinv = 0
        do ip = 1, npath
            n = nsar_index(ip + 1) - nsar_index(ip)
            nsinv = nsinv + n*n
            if (present(sigma_sar_mat) .and. n > 0) s = s + 13d0*sum(sigma_sar_mat(ip)%body(1:n, 1:n))
        end do
        if (gr) then
            if (c_associated(g_pbg) .and. all(sig == g_sigg) .and. s == g_sumg) return
        else
            if (c_associated(g_pb) .and. all(sig == g_sig) .and. s == g_sum) return
        end if
        call smcf_ensure_ctx()
        if (gr) then
            if (c_associated(g_pbg)) call smcf_check(smcf_problem_destroy(g_pbg), "smcf_problem_destroy")
            call fill_keep(t_keepg)
            call describe(t_keepg)
            call smcf_check(smcf_problem_create(g_ctx, d, 1_c_int, g_pbg), "smcf_problem_create")
            g_sigg = sig
            g_sumg = s
        else
            if (c_associated(g_pb)) call smcf_check(smcf_problem_destroy(g_pb), "smcf_problem_destroy")
            call fill_keep(t_keep)
            call describe(t_keep)
            call smcf_check(smcf_problem_create(g_ctx, d, int(g_max_batch, c_int), g_pb), "smcf_problem_create")
            g_sig = sig
            g_sum = s
        end if
    contains
        subroutine fill_keep(t)
            type(keep_t), intent(inout) :: t
            if (allocated(t%nxi)) deallocate (t%nxi, t%neta, t%nsi, t%xb, t%pts, t%uv, t%sig, t%d, t%sar)
            allocate (t%nxi(nplane), t%neta(nplane), t%nsi(npath + 1), t%xb(max(nplane - 1, 1)), &
                      t%pts(2*nobs), t%uv(3*nobs), t%sig(nobs), t%d(nobs), t%sar(max(nsinv, 1)))
            t%nxi = nxi_ls(1:nplane)
            t%neta = neta_ls(1:nplane)
            t%nsi = nsar_index(1:npath + 1)
            t%xb = 0d0
            if (nplane > 1) t%xb(1:nplane - 1) = xbend(1:nplane - 1)
            t%pts = reshape(obs_points(1:2, 1:nobs), [2*nobs])
            t%uv = reshape(obs_unitvec(1:3, 1:nobs), [3*nobs])
            t%sig = obs_sigma(1:nobs)
            t%d = dvec(1:nobs)
            t%sar = 0d0
            off = 0
            do ip = 1, npath
                n = nsar_index(ip + 1) - nsar_index(ip)
                if (n <= 0) cycle
                if (present(sigma_sar_mat)) then
                    t%sar(off + 1:off + n*n) = reshape(sigma_sar_mat(ip)%body(1:n, 1:n), [n*n])
                else
                    do k = 1, n
                        t%sar(off + (k - 1)*n + k) = 1d0
                    end do
                end if
                off = off + n*n
            end do
        end subroutine
        subroutine describe(t)
            type(keep_t), target, intent(in) :: t
            d%nplane = nplane
            d%nxi_ls = c_loc(t%nxi)
            d%neta_ls = c_loc(t%neta)
            d%xmin = xmin
            d%xmax = xmax
            d%zmin = zmin
            d%fix_xbend = merge(1, 0, fix_xbend)
            d%xbend = c_loc(t%xb)
            d%nobs = nobs
            d%nsar_total = nsar_total
            d%ngnss = ngnss
            d%npath = npath
            d%nsar_index = c_loc(t%nsi)
            d%obs_points = c_loc(t%pts)
            d%obs_unitvec = c_loc(t%uv)
            d%obs_sigma = c_loc(t%sig)
            d%dvec = c_loc(t%d)
            d%sar_sinv = c_loc(t%sar)
            d%max_slip = max_slip
            d%nparticle_slip = nparticle_slip
            d%mover = g_mover
            d%gibbs_sweeps = g_sweeps
            d%precision = g_precision
            d%hmc_ntau = g_hmc_ntau
            d%hmc_accept = g_hmc_accept
            if (gr) then
                d%mover = 1
                d%precision = 0
            end if
        end subroutine
    end subroutine

    integer(c_int64_t) function smcf_next_key()
        smcf_next_key = smcf_inner_key(g_seed, 0_c_int32_t, 2_c_int32_t, g_counter)
        g_counter = g_counter + 1
    end function

    ! one keyed uniform pair (hi + lo = a 106-bit uniform) for the shim's own draws (site 2, the resampling's)
    subroutine smcf_shim_uniform(uh, ul)
        double precision, intent(out) :: uh, ul
        integer(c_int32_t) :: ctr(4), key(2), out(4)
        integer(c_int64_t) :: w0, w1
        call smcf_ensure_ctx()
        ctr = [g_draws, 0_c_int32_t, 0_c_int32_t, int(z'02000000', c_int32_t)]
        key = [int(iand(g_seed, int(z'FFFFFFFF', c_int64_t)), c_int32_t), int(shiftr(g_seed, 32), c_int32_t)]
        call smcf_check(smcf_philox(g_ctx, 1_c_int, ctr, key, out), "smcf_philox")
        g_draws = g_draws + 1
        w0 = ior(shiftl(int(out(2), c_int64_t), 32), iand(int(out(1), c_int64_t), int(z'FFFFFFFF', c_int64_t)))
        w1 = ior(shiftl(int(out(4), c_int64_t), 32), iand(int(out(3), c_int64_t), int(z'FFFFFFFF', c_int64_t)))
        uh = dble(shiftr(w0, 11))*2d0**(-53)
        ul = dble(shiftr(w1, 11))*2d0**(-106)
    end subroutine

    ! the slip-particle file of flag_output = 1 (smc_slip.f90:1632-1660)
    subroutine smcf_write_slip_particles(output_path, nparticle, nnode, ndof, id_dof, particles, like, prior)
        character(*), intent(in) :: output_path
        integer, intent(in) :: nparticle, nnode, ndof, id_dof(:)
        double precision, intent(in) :: particles(:, :), like(:), prior(:)
        double precision, allocatable :: slip(:, :)
        integer :: iparticle, idof, inode, idirection
        allocate (slip(2, nnode))
        open (17, file=output_path, status='replace')
        do iparticle = 1, nparticle
            slip = 0d0
            do idof = 1, ndof
                inode = id_dof(idof)
                slip(1, inode) = particles(2*(idof - 1) + 1, iparticle)
                slip(2, inode) = particles(2*(idof - 1) + 2, iparticle)
            end do
            do inode = 1, nnode
                do idirection = 1, 2
                    write (17, "(f12.5)", advance="no") slip(idirection, inode)
                end do
            end do
            write (17, "(f12.5)", advance="no") like(iparticle)
            write (17, "(f12.5)") prior(iparticle)
        end do
        close (17)
        deallocate (slip)
    end subroutine
end module smcf_c

module gfunc
    use smcf_c
    use type_mat
    implicit none
contains
    ! gfunc.f90:6-19: unit slip at a node (strike component -1, dip component +1)
    subroutine gen_unit_slip(inode, idirection, slip_dist)
        integer, intent(in) :: inode, idirection
        double precision, intent(inout) :: slip_dist(:, :)
        if (idirection == 1) slip_dist(idirection, inode) = -1d0
        if (idirection == 2) slip_dist(idirection, inode) = 1d0
    end subroutine gen_unit_slip

    ! gfunc.f90:21-26
    subroutine del_unit_slip(inode, idirection, slip_dist)
        integer, intent(in) :: inode, idirection
        double precision, intent(inout) :: slip_dist(:, :)
        slip_dist(idirection, inode) = 0d0
    end subroutine del_unit_slip

    ! gfunc.f90:65-100, on the device
    subroutine call_dc3d0(xsource, ysource, zsource, xobs, yobs, uxi, ueta, dip, area, strike, uret)
        double precision, intent(in) :: xsource, ysource, zsource, xobs, yobs, uxi, ueta, area, strike
        double precision, intent(inout) :: uret(:)
        double precision, intent(in) :: dip
        real(c_double) :: a(10), u(3)
        a = [xsource, ysource, zsource, xobs, yobs, uxi, ueta, dip, area, strike]
        call smcf_ensure_ctx()
        call smcf_check(smcf_call_dc3d0(g_ctx, 1_c_int, a, u), "smcf_call_dc3d0")
        uret(1:3) = u
    end subroutine call_dc3d0

    ! gfunc.f90:102-194, on the device; workspace arguments untouched
    subroutine calc_responce(cny_fault, coor_fault, slip_dist, xobs, yobs, xf, yf, zf, strike, dip, target_id_val, &
                             target_id_size, node_id_in_patch, xinode, etanode, uxinode, uetanode, r1vec, r2vec, &
                             nvec, uobs, uret)
        double precision, intent(in) :: coor_fault(:, :), slip_dist(:, :)
        double precision, intent(in) :: xf, yf, zf, strike, dip, xobs, yobs
        double precision, intent(inout) :: xinode(:), etanode(:), uxinode(:), uetanode(:), r1vec(:), r2vec(:), &
                                           nvec(:), uobs(:), uret(:)
        integer, intent(in) :: cny_fault(:, :), target_id_val(:)
        integer, intent(in) :: target_id_size
        integer, intent(inout) :: node_id_in_patch(:)
        real(c_double) :: pts(2), u(3)
        integer(c_int), allocatable :: cny(:), tgt(:)
        pts = [xobs, yobs]
        allocate (cny(4*size(cny_fault, 2)), tgt(max(target_id_size, 1)))
        cny = reshape(cny_fault(1:4, :), [4*size(cny_fault, 2)])
        tgt = 1
        tgt(1:target_id_size) = target_id_val(1:target_id_size)
        call smcf_ensure_ctx()
        call smcf_check(smcf_calc_responce(g_ctx, 1_c_int, pts, int(size(cny_fault, 2), c_int), cny, &
                                           int(size(coor_fault, 2), c_int), reshape(coor_fault(1:2, :), &
                                           [2*size(coor_fault, 2)]), reshape(slip_dist(1:2, :), &
                                           [2*size(slip_dist, 2)]), xf, yf, zf, strike, dip, &
                                           int(target_id_size, c_int), tgt, u), "smcf_calc_responce")
        uobs(1:3) = u
    end subroutine calc_responce

    ! gfunc.f90:196-240: SAR points, then one point per GNSS station copied to its 3 rows
    subroutine calc_responce_dist(obs_points, cny_fault, coor_fault, slip_dist, xf, yf, zf, strike, dip, &
                                  target_id_val, target_id_size, ngnss, node_id_in_patch, xinode, etanode, uxinode, &
                                  uetanode, r1vec, r2vec, nvec, response_dist, uobs, uret, nsar_total)
        double precision, intent(in) :: obs_points(:, :), coor_fault(:, :), slip_dist(:, :)
        double precision, intent(in) :: xf, yf, zf, strike, dip
        double precision, intent(inout) :: xinode(:), etanode(:), uxinode(:), uetanode(:), r1vec(:), r2vec(:), &
                                           nvec(:), response_dist(:, :), uobs(:), uret(:)
        integer, intent(in) :: cny_fault(:, :), target_id_val(:)
        integer, intent(inout) :: node_id_in_patch(:)
        integer, intent(in) :: ngnss, target_id_size, nsar_total
        real(c_double), allocatable :: pts(:, :), u(:, :)
        integer(c_int), allocatable :: cny(:), tgt(:)
        integer :: iobs, np, k
        np = nsar_total + ngnss
        allocate (pts(2, max(np, 1)), u(3, max(np, 1)), cny(4*size(cny_fault, 2)), tgt(max(target_id_size, 1)))
        do iobs = 1, nsar_total
            pts(:, iobs) = obs_points(1:2, iobs)
        end do
        do iobs = 1, ngnss
            pts(:, nsar_total + iobs) = obs_points(1:2, nsar_total + 1 + 3*(iobs - 1))
        end do
        cny = reshape(cny_fault(1:4, :), [4*size(cny_fault, 2)])
        tgt = 1
        tgt(1:target_id_size) = target_id_val(1:target_id_size)
        call smcf_ensure_ctx()
        call smcf_check(smcf_calc_responce(g_ctx, int(np, c_int), pts, int(size(cny_fault, 2), c_int), cny, &
                                           int(size(coor_fault, 2), c_int), reshape(coor_fault(1:2, :), &
                                           [2*size(coor_fault, 2)]), reshape(slip_dist(1:2, :), &
                                           [2*size(slip_dist, 2)]), xf, yf, zf, strike, dip, &
                                           int(target_id_size, c_int), tgt, u), "smcf_calc_responce")
        do iobs = 1, nsar_total
            response_dist(1:3, iobs) = u(:, iobs)
        end do
        do iobs = 1, ngnss
            do k = 1, 3
                response_dist(1:3, nsar_total + 3*(iobs - 1) + k) = u(:, nsar_total + iobs)
            end do
        end do
    end subroutine calc_responce_dist

    ! gfunc.f90:242-348; workspace arguments are accepted and not used
    subroutine calc_greens_func(theta, nplane, nxi_ls, neta_ls, gmat, slip_dist, cny_fault, coor_fault, obs_points, &
                                obs_unitvec, node_to_elem_val, node_to_elem_size, &
                                id_dof, ngnss, nobs, nnode_total, ndof_total, ndof_index, target_id_val, &
                                node_id_in_patch, xinode, etanode, uxinode, uetanode, &
                                r1vec, r2vec, nvec, response_dist, uobs, uret, nsar_total, npath, &
                                nsar_index, gmat_arr, xmin, xmax, zmin, fix_xbend, xbend)
        double precision, intent(inout) :: gmat(:, :), slip_dist(:, :), &
            xinode(:), etanode(:), uxinode(:), uetanode(:), &
            r1vec(:), r2vec(:), nvec(:), response_dist(:, :), uobs(:), uret(:)
        integer, intent(inout) :: target_id_val(:), node_id_in_patch(:)
        integer, intent(in) :: cny_fault(:, :), node_to_elem_val(:, :), &
                               node_to_elem_size(:), id_dof(:), nsar_total
        double precision, intent(in) :: theta(:), coor_fault(:, :), obs_points(:, :), &
            obs_unitvec(:, :), xbend(:)
        integer, intent(in) :: nplane, nxi_ls(:), neta_ls(:), ngnss, nobs, &
                               nnode_total, ndof_total, ndof_index(:), npath, nsar_index(:)
        logical, intent(in) :: fix_xbend
        type(mat), intent(inout) :: gmat_arr(:)
        double precision, intent(in) :: xmin, xmax, zmin
        double precision, allocatable :: ones(:), zeros(:), th(:), g(:)
        integer :: ipath, k, iobs
        allocate (ones(nobs), zeros(nobs), th(size(theta)), g(nobs*2*ndof_total))
        ones = 1d0
        zeros = 0d0
        th = theta
        call smcf_ensure_problem(nplane, nxi_ls, neta_ls, xmin, xmax, zmin, fix_xbend, xbend, nobs, nsar_total, &
                                 ngnss, npath, nsar_index, obs_points, obs_unitvec, ones, zeros, 10d0, 16, &
                                 greens_only=.true.)
        call smcf_check(smcf_calc_greens_func(g_pbg, 1_c_int, th, g), "smcf_calc_greens_func")
        gmat(1:nobs, 1:2*ndof_total) = reshape(g, [nobs, 2*ndof_total])
        do ipath = 1, npath
            k = nsar_index(ipath)
            do iobs = nsar_index(ipath), nsar_index(ipath + 1) - 1
                gmat_arr(ipath)%body(iobs - k + 1, 1:2*ndof_total) = gmat(iobs, 1:2*ndof_total)
            end do
        end do
    end subroutine calc_greens_func
end module gfunc

module smc_slip
    use smcf_c
    use type_mat
    implicit none
contains
    ! smc_slip.f90:40-45
    double precision function cdf_norm(x, mu, sigma2)
        double precision, intent(in) :: x, mu, sigma2
        cdf_norm = (1d0 + erf((x - mu)/sqrt(2d0*sigma2)))/2d0
    end function cdf_norm

    ! smc_slip.f90:47-54
    double precision function pdf_norm(x, mu, sigma2)
        double precision, intent(in) :: x, mu, sigma2
        double precision :: pi
        pi = 2d0*asin(1d0)
        pdf_norm = exp(-((x - mu)**2d0)/(2d0*sigma2))/sqrt(2d0*pi*sigma2)
    end function pdf_norm

    ! smc_slip.f90:56-104 (Gram form), on the device
    double precision function slip_calc_likelihood(svec, dvec, sigma2_full, gmat, log_sigma_sar2, log_sigma_gnss2, &
                                                   ngnss, gsvec, nobs, ndof, gsgmat, gsdvec, dsd, nsar_total)
        double precision, intent(in) :: svec(:), dvec(:), sigma2_full(:), gmat(:, :), log_sigma_sar2, &
                                        log_sigma_gnss2, gsgmat(:, :), gsdvec(:), dsd
        integer, intent(in) :: ngnss, nobs, ndof, nsar_total
        double precision, intent(inout) :: gsvec(:)
        real(c_double) :: out(1)
        integer :: nd
        nd = 2*ndof
        call smcf_ensure_ctx()
        call smcf_check(smcf_slip_likelihood_gram(g_ctx, int(nd, c_int), 1_c_int, svec(1:nd), &
                                                  reshape(gsgmat(1:nd, 1:nd), [nd*nd]), gsdvec(1:nd), dsd, &
                                                  int(nsar_total, c_int), int(ngnss, c_int), log_sigma_sar2, &
                                                  log_sigma_gnss2, out), "smcf_slip_likelihood_gram")
        slip_calc_likelihood = out(1)
    end function slip_calc_likelihood

    ! smc_slip.f90:106-137, on the device
    double precision function slip_calc_prior(svec, alpha2_full, theta, nplane, lmat_index, lmat_val, lsvec, nnode)
        double precision, intent(in) :: svec(:), alpha2_full(:), lmat_val(:, :), theta(:)
        integer, intent(in) :: lmat_index(:, :), nnode, nplane
        double precision, intent(inout) :: lsvec(:)
        real(c_double) :: out(1)
        call smcf_ensure_ctx()
        call smcf_check(smcf_slip_prior_ell(g_ctx, int(nnode, c_int), int(size(svec), c_int), 1_c_int, svec, &
                                            alpha2_full(1:2*nnode), reshape(lmat_index(1:9, 1:2*nnode), &
                                            [18*nnode]), reshape(lmat_val(1:9, 1:2*nnode), [18*nnode]), out), &
                        "smcf_slip_prior_ell")
        slip_calc_prior = out(1)
    end function slip_calc_prior

    ! smc_slip.f90:139-182, on the device
    subroutine slip_calc_grad(grad, svec, gmat, lmat_index, lmat_val, ltmat_index, ltmat_val, dvec, gamma, gsvec, &
                              lsvec, sigma2_full, alpha2_full, nobs, ndof, nnode, ndim, gsdvec, gsgmat)
        double precision, intent(in) :: svec(:), gmat(:, :), dvec(:), gamma, sigma2_full(:), alpha2_full(:), &
                                        lmat_val(:, :), ltmat_val(:, :), gsdvec(:), gsgmat(:, :)
        integer, intent(in) :: nobs, ndof, nnode, ndim, lmat_index(:, :), ltmat_index(:, :)
        double precision, intent(inout) :: grad(:), gsvec(:), lsvec(:)
        real(c_double), allocatable :: g(:)
        allocate (g(ndim))
        call smcf_ensure_ctx()
        call smcf_check(smcf_slip_grad_ell(g_ctx, int(nnode, c_int), int(ndim, c_int), 1_c_int, svec(1:ndim), gamma, &
                                           reshape(gsgmat(1:ndim, 1:ndim), [ndim*ndim]), gsdvec(1:ndim), &
                                           alpha2_full(1:2*nnode), reshape(lmat_index(1:9, 1:2*nnode), [18*nnode]), &
                                           reshape(lmat_val(1:9, 1:2*nnode), [18*nnode]), &
                                           reshape(ltmat_index(1:9, 1:ndim), [9*ndim]), &
                                           reshape(ltmat_val(1:9, 1:ndim), [9*ndim]), g), "smcf_slip_grad_ell")
        grad(1:ndim) = g
    end subroutine slip_calc_grad

    ! smc_slip.f90:184-287: the sequential prior Gibbs chain of every particle, on the cached problem
    subroutine slip_gen_init_particles(particles, likelihood_ls, prior_ls, nparticle, ndim, llmat, max_slip, dvec, &
                                       sigma2_full, alpha2_full, theta, nplane, gmat, log_sigma_sar2, &
                                       log_sigma_gnss2, ngnss, nobs, ndof, lmat_index, lmat_val, nnode, particle_cur, &
                                       gsvec, lsvec, gsdvec, dsd, nsar_total, gsgmat)
        double precision, intent(in) :: max_slip, dvec(:), lmat_val(:, :), llmat(:, :), gmat(:, :), &
            sigma2_full(:), alpha2_full(:), log_sigma_sar2, log_sigma_gnss2, theta(:), gsdvec(:), dsd, gsgmat(:, :)
        integer, intent(in) :: nnode, ndof, ngnss, nobs, nparticle, ndim, lmat_index(:, :), nplane, nsar_total
        double precision, intent(inout) :: particles(:, :), likelihood_ls(:), prior_ls(:), particle_cur(:), &
                                           gsvec(:), lsvec(:)
        real(c_double), allocatable :: x(:), lk(:), pr(:), th(:)
        call smcf_need_problem("slip_gen_init_particles")
        if (nparticle /= g_sig(6)) then
            print *, "slip_gen_init_particles: nparticle differs from the cached problem's"
            stop 1
        end if
        allocate (x(ndim*nparticle), lk(nparticle), pr(nparticle), th(size(theta)))
        th = theta
        call smcf_check(smcf_slip_gen_init_particles(g_pb, th, smcf_next_key(), x, lk, pr), &
                        "smcf_slip_gen_init_particles")
        particles(1:ndim, 1:nparticle) = reshape(x, [ndim, nparticle])
        likelihood_ls(1:nparticle) = lk
        prior_ls(1:nparticle) = pr
    end subroutine slip_gen_init_particles

    ! smc_slip.f90:289-312
    subroutine slip_calc_mean_std_vector(vec, size, mean, std)
        double precision, intent(in) :: vec(:)
        integer, intent(in) :: size
        double precision, intent(inout) :: mean, std
        call smcf_check(smcf_fault_mean_std(int(size, c_int), vec(1:size), mean, std), "smcf_fault_mean_std")
    end subroutine slip_calc_mean_std_vector

    ! smc_slip.f90:314-387 (CV target 2, tolerance 1e-10), on the device
    double precision function slip_find_next_gamma(gamma_prev, likelihood_ls, weights, neglog_evidence, nparticle)
        double precision, intent(in) :: gamma_prev, likelihood_ls(:)
        integer, intent(in) :: nparticle
        double precision, intent(inout) :: neglog_evidence, weights(:)
        real(c_double) :: g, nev
        real(c_double), allocatable :: w(:)
        allocate (w(nparticle))
        call smcf_ensure_ctx()
        call smcf_check(smcf_find_next_gamma(g_ctx, int(nparticle, c_int), gamma_prev, likelihood_ls(1:nparticle), &
                                             2d0, 1d-10, 100000_c_int, g, w, nev), "smcf_find_next_gamma")
        weights(1:nparticle) = w
        neglog_evidence = nev
        slip_find_next_gamma = g
    end function slip_find_next_gamma

    ! smc_slip.f90:465-483 (real(16) sum; double-double on the device)
    subroutine slip_normalize_weights(weights, nparticle)
        double precision, intent(inout) :: weights(:)
        integer, intent(in) :: nparticle
        real(c_double), allocatable :: wn(:)
        integer(c_int), allocatable :: a(:)
        allocate (wn(nparticle), a(nparticle))
        call smcf_ensure_ctx()
        call smcf_check(smcf_normalize_resample(g_ctx, int(nparticle, c_int), weights(1:nparticle), 0.5d0, 0d0, wn, &
                                                a), "smcf_normalize_resample")
        weights(1:nparticle) = wn
    end subroutine slip_normalize_weights

    ! smc_slip.f90:485-505
    subroutine slip_calc_mean_particles(particles, weights, mean, nparticle, ndim)
        double precision, intent(in) :: particles(:, :), weights(:)
        double precision, intent(inout) :: mean(:)
        integer, intent(in) :: nparticle, ndim
        real(c_double), allocatable :: m(:), c(:)
        allocate (m(ndim), c(ndim))
        call smcf_ensure_ctx()
        call smcf_check(smcf_slip_calc_moments(g_ctx, int(ndim, c_int), int(nparticle, c_int), &
                                               reshape(particles(1:ndim, 1:nparticle), [ndim*nparticle]), &
                                               weights(1:nparticle), m, c), "smcf_slip_calc_moments")
        mean(1:ndim) = m
    end subroutine slip_calc_mean_particles

    ! smc_slip.f90:507-570: the variance diagonal (floored at 1e-3); the full covariance is not formed (the
    ! reference's is dead code)
    subroutine slip_calc_cov_particles(particles, weights, mean, cov, nparticle, ndim, cov_diag)
        double precision, intent(in) :: particles(:, :), weights(:), mean(:)
        integer, intent(in) :: nparticle, ndim
        double precision, intent(inout) :: cov(:, :), cov_diag(:)
        real(c_double), allocatable :: m(:), c(:)
        allocate (m(ndim), c(ndim))
        call smcf_ensure_ctx()
        call smcf_check(smcf_slip_calc_moments(g_ctx, int(ndim, c_int), int(nparticle, c_int), &
                                               reshape(particles(1:ndim, 1:nparticle), [ndim*nparticle]), &
                                               weights(1:nparticle), m, c), "smcf_slip_calc_moments")
        cov_diag(1:ndim) = c
    end subroutine slip_calc_cov_particles

    ! smc_slip.f90:572-589: systematic residual resampling with a real(16) running u, u from the keyed stream
    subroutine slip_resample_particles(nparticle, weights, assigned_num)
        double precision, intent(in) :: weights(:)
        integer, intent(in) :: nparticle
        integer, intent(inout) :: assigned_num(:)
        real(c_double), allocatable :: wn(:)
        integer(c_int), allocatable :: a(:)
        double precision :: uh, ul
        allocate (wn(nparticle), a(nparticle))
        call smcf_shim_uniform(uh, ul)
        call smcf_check(smcf_normalize_resample(g_ctx, int(nparticle, c_int), weights(1:nparticle), uh, ul, wn, a), &
                        "smcf_normalize_resample")
        assigned_num(1:nparticle) = a
    end subroutine slip_resample_particles

    ! smc_slip.f90:732-917: one trajectory on the device (the cached problem's geometry theta); on an escape
    ! (|x| > 1e3) ham_cand = 1e20 and particle_cand = particle_cur, as in the reference
    subroutine slip_leapfrog(ham_cur, post_cur, ndim, pvec, cov_diag, particle_cand, dtau, ntau, grad, gmat, &
                             lmat_index, lmat_val, ltmat_index, ltmat_val, dvec, gamma, gsvec, lsvec, sigma2_full, &
                             alpha2_full, nobs, ndof, nnode, max_slip, likelihood_cand, log_sigma_sar2, &
                             log_sigma_gnss2, ngnss, prior_cand, theta, nplane, post_cand, ham_cand, particle_cur, &
                             gsdvec, gsgmat, dsd, nsar_total)
        integer, intent(in) :: ndim, ntau, lmat_index(:, :), ltmat_index(:, :), nobs, ndof, nnode, ngnss, nplane, &
                               nsar_total
        double precision, intent(in) :: post_cur, dtau, gmat(:, :), lmat_val(:, :), ltmat_val(:, :), dvec(:), gamma, &
            sigma2_full(:), alpha2_full(:), max_slip, log_sigma_sar2, log_sigma_gnss2, theta(:), particle_cur(:), &
            gsdvec(:), gsgmat(:, :), dsd, cov_diag(:)
        double precision, intent(inout) :: ham_cur, pvec(:), particle_cand(:), grad(:), gsvec(:), lsvec(:), &
                                           likelihood_cand, prior_cand, post_cand, ham_cand
        real(c_double) :: hc(1), hn(1), lk(1), pr(1)
        real(c_double), allocatable :: cand(:), pout(:), th(:)
        call smcf_need_problem("slip_leapfrog")
        allocate (cand(ndim), pout(ndim), th(size(theta)))
        th = theta
        call smcf_check(smcf_slip_leapfrog(g_pb, th, gamma, 1_c_int, particle_cur(1:ndim), pvec(1:ndim), &
                                           cov_diag(1:ndim), dtau, int(ntau, c_int), cand, pout, hc, hn, lk, pr), &
                        "smcf_slip_leapfrog")
        ham_cur = hc(1)
        ham_cand = hn(1)
        particle_cand(1:ndim) = cand
        pvec(1:ndim) = pout
        if (hn(1) /= 1d20) then
            likelihood_cand = lk(1)
            prior_cand = pr(1)
            post_cand = gamma*lk(1) + pr(1)
        end if
    end subroutine slip_leapfrog

    ! smc_slip.f90:1347-1443 from the given Green's matrix, with the cached problem's observation set
    subroutine slip_calc_gsd_gsg(gmat, dvec, gsdvec, gsgmat, nobs, ndof, npath, nsar_index, nsar_total, &
                                 sigma_sar_mat, sigma_sar2, sigma_gnss2, obs_sigma, gmat_arr, dsd)
        integer, intent(in) :: nobs, ndof, npath, nsar_index(:), nsar_total
        type(mat), intent(in) :: sigma_sar_mat(:), gmat_arr(:)
        double precision, intent(in) :: gmat(:, :), dvec(:), sigma_sar2, sigma_gnss2, obs_sigma(:)
        double precision, intent(inout) :: gsdvec(:), gsgmat(:, :), dsd
        real(c_double), allocatable :: h(:), b(:)
        real(c_double) :: d
        integer :: nd
        call smcf_need_problem("slip_calc_gsd_gsg")
        if (any(dvec(1:nobs) /= t_keep%d) .or. any(obs_sigma(1:nobs) /= t_keep%sig)) then
            print *, "slip_calc_gsd_gsg: dvec / obs_sigma differ from the cached problem's observation set"
            stop 1
        end if
        nd = 2*ndof
        allocate (h(nd*nd), b(nd))
        call smcf_check(smcf_slip_calc_gsd_gsg_gmat(g_pb, reshape(gmat(1:nobs, 1:nd), [nobs*nd]), sigma_sar2, &
                                                    sigma_gnss2, h, b, d), "smcf_slip_calc_gsd_gsg_gmat")
        gsgmat(1:nd, 1:nd) = reshape(h, [nd, nd])
        gsdvec(1:nd) = b
        dsd = d
    end subroutine slip_calc_gsd_gsg

    ! smc_slip.f90:1445-1663: the whole inner SMC of geometry theta on the device (cached problem; the device
    ! forms G, L and the Gram quantities from theta, as fault_calc_likelihood does before calling it)
    subroutine slip_smc_exec(particles, particles_new, nparticle, ndim, flag_output, output_path, llmat, max_slip, &
                             dvec, sigma2_full, alpha2_full, theta, nplane, gmat, log_sigma_sar2, log_sigma_gnss2, &
                             ngnss, nobs, ndof, lmat_index, lmat_val, ltmat_index, ltmat_val, nnode, likelihood_ls, &
                             prior_ls, weights, mean, cov, likelihood_ls_new, prior_ls_new, assigned_num, id_start, &
                             id_dof, st_rand_ls, metropolis_ls, gsvec, lsvec, particle_cur, particle_cand, st_rand, &
                             neglog_ret, npath, nsar_index, nsar_total, sigma_sar_mat, sigma_sar2, sigma_gnss2, &
                             obs_sigma, gmat_arr)
        double precision, intent(in) :: max_slip, dvec(:), lmat_val(:, :), ltmat_val(:, :), llmat(:, :), &
            gmat(:, :), sigma2_full(:), alpha2_full(:), log_sigma_sar2, log_sigma_gnss2, theta(:), sigma_sar2, &
            sigma_gnss2, obs_sigma(:)
        integer, intent(in) :: nplane, nnode, ndof, ngnss, nobs, nparticle, ndim, flag_output, id_dof(:), &
                               lmat_index(:, :), ltmat_index(:, :), npath, nsar_total, nsar_index(:)
        character(*), intent(in) :: output_path
        type(mat), intent(in) :: sigma_sar_mat(:), gmat_arr(:)
        double precision, intent(inout) :: particles(:, :), particles_new(:, :), likelihood_ls(:), prior_ls(:), &
            weights(:), cov(:, :), mean(:), likelihood_ls_new(:), prior_ls_new(:), st_rand_ls(:, :), &
            metropolis_ls(:), gsvec(:), lsvec(:), particle_cur(:), particle_cand(:), st_rand(:), neglog_ret
        integer, intent(inout) :: assigned_num(:), id_start(:)
        real(c_double), allocatable :: x(:), lk(:), pr(:), mn(:), gt(:), th(:)
        real(c_double) :: neglog
        integer(c_int) :: nstage
        call smcf_need_problem("slip_smc_exec")
        if (nparticle /= g_sig(6) .or. any(dvec(1:nobs) /= t_keep%d)) then
            print *, "slip_smc_exec: nparticle or dvec differ from the cached problem's"
            stop 1
        end if
        allocate (x(ndim*nparticle), lk(nparticle), pr(nparticle), mn(ndim), gt(202), th(size(theta)))
        th = theta
        call smcf_check(smcf_slip_smc_exec(g_pb, th, smcf_next_key(), x, lk, pr, mn, nstage, neglog, gt), &
                        "smcf_slip_smc_exec")
        particles(1:ndim, 1:nparticle) = reshape(x, [ndim, nparticle])
        likelihood_ls(1:nparticle) = lk
        prior_ls(1:nparticle) = pr
        mean(1:ndim) = mn
        neglog_ret = neglog
        if (flag_output == 1) call smcf_write_slip_particles(output_path, nparticle, nnode, ndof, id_dof, &
                                                             particles, likelihood_ls, prior_ls)
    end subroutine slip_smc_exec
end module smc_slip

module smc_fault
    use smcf_c
    use type_mat
    implicit none
contains
    ! smc_fault.f90:9-27
    double precision function fault_calc_neglog_q(particle, qmean, qcov_inv, log_qcov_det, ndim)
        double precision, intent(in) :: particle(:), qmean(:), qcov_inv(:, :), log_qcov_det
        integer, intent(in) :: ndim
        double precision :: pi, tmp, delta(ndim)
        integer :: idim, jdim
        pi = 2d0*asin(1d0)
        delta = particle - qmean
        tmp = 0d0
        do jdim = 1, ndim
            do idim = 1, ndim
                tmp = tmp + delta(idim)*qcov_inv(idim, jdim)*delta(jdim)
            end do
        end do
        fault_calc_neglog_q = tmp/2d0 + log(2d0*pi)*ndim/2d0 + log_qcov_det/2d0
    end function fault_calc_neglog_q

    ! smc_fault.f90:29-64: uniform in range, or qmean + qcov z (qcov holds the lower Cholesky factor, as
    ! main.f90:332 leaves it); keyed by the seed; restarts the outer stage count
    subroutine fault_sample_init_particles(particles, nparticle, ndim, range, nplane, gsmc, qmean, qcov)
        double precision, intent(inout) :: particles(:, :)
        integer, intent(in) :: nparticle, ndim, nplane
        double precision, intent(in) :: range(:, :), qmean(:), qcov(:, :)
        logical, intent(in) :: gsmc
        real(c_double), allocatable, target :: r(:), qm(:), qc(:)
        real(c_double), allocatable :: x(:)
        allocate (r(2*ndim), qm(ndim), qc(ndim*ndim), x(ndim*nparticle))
        r = reshape(range(1:2, 1:ndim), [2*ndim])
        if (gsmc) then
            qm = qmean(1:ndim)
            qc = reshape(qcov(1:ndim, 1:ndim), [ndim*ndim])
            call smcf_check(smcf_fault_sample_init_particles(int(ndim, c_int), int(nparticle, c_int), g_seed, &
                                                             c_loc(r), c_loc(qm), c_loc(qc), x), &
                            "smcf_fault_sample_init_particles")
        else
            call smcf_check(smcf_fault_sample_init_particles(int(ndim, c_int), int(nparticle, c_int), g_seed, &
                                                             c_loc(r), c_null_ptr, c_null_ptr, x), &
                            "smcf_fault_sample_init_particles")
        end if
        particles(1:ndim, 1:nparticle) = reshape(x, [ndim, nparticle])
        g_stage = 0
    end subroutine fault_sample_init_particles

    ! smc_fault.f90:66-292: -log Z of the inner slip SMC for one geometry; the reference's workspace outputs of
    ! the mesh, Laplacian and Green's steps (cny_fault, coor_fault, node_to_elem_*, id_dof, luni, lmat, the ELL
    ! of L and L^T, llmat, gmat, gmat_arr, alpha2_full) are filled as the reference leaves them
    double precision function fault_calc_likelihood( &
        theta, nplane, nxi_ls, neta_ls, nnode_total, ndof_total, ndof_index, &
        ngnss, nobs, cny_fault, coor_fault, &
        node_to_elem_val, node_to_elem_size, id_dof, luni, &
        lmat, lmat_index, lmat_val, ltmat_index, ltmat_val, llmat, &
        gmat, slip_dist, obs_points, &
        obs_unitvec, obs_sigma, sigma2_full, alpha2_full, &
        target_id_val, node_id_in_patch, xinode, etanode, &
        uxinode, uetanode, r1vec, r2vec, &
        nvec, response_dist, uobs, uret, &
        slip_particles, slip_particles_new, &
        nparticle_slip, max_slip, dvec, &
        slip_likelihood_ls, slip_prior_ls, slip_weights, &
        slip_mean, slip_cov, slip_likelihood_ls_new, &
        slip_prior_ls_new, slip_assigned_num, slip_id_start, &
        slip_st_rand_ls, slip_metropolis_ls, gsvec, lsvec, &
        slip_particle_cur, slip_particle_cand, slip_st_rand, &
        flag_output, output_path, npath, nsar_index, &
        nsar_total, sigma_sar_mat, gmat_arr, xmin, xmax, zmin, &
        fix_xbend, xbend)
        double precision, intent(in) :: theta(:), obs_points(:, :), &
            obs_unitvec(:, :), obs_sigma(:), max_slip, dvec(:), &
            xmin, xmax, zmin, xbend(:)
        integer, intent(in) :: nxi_ls(:), neta_ls(:), nnode_total, ndof_total, ndof_index(:), ngnss, nobs, &
                               nparticle_slip, flag_output, nplane, npath, &
                               nsar_index(:), nsar_total
        character(*), intent(in) :: output_path
        type(mat), intent(in) :: sigma_sar_mat(:)
        logical, intent(in) :: fix_xbend
        double precision, intent(inout) :: coor_fault(:, :), luni(:, :), &
            lmat(:, :), lmat_val(:, :), ltmat_val(:, :), llmat(:, :), gmat(:, :), slip_dist(:, :), &
            sigma2_full(:), alpha2_full(:), xinode(:), etanode(:), uxinode(:), uetanode(:), &
            r1vec(:), r2vec(:), nvec(:), response_dist(:, :), uobs(:), uret(:), &
            slip_particles(:, :), slip_particles_new(:, :), slip_likelihood_ls(:), &
            slip_prior_ls(:), slip_weights(:), slip_mean(:), slip_cov(:, :), &
            slip_likelihood_ls_new(:), slip_prior_ls_new(:), &
            slip_st_rand_ls(:, :), slip_metropolis_ls(:), gsvec(:), lsvec(:), &
            slip_particle_cur(:), slip_particle_cand(:), slip_st_rand(:)
        integer, intent(inout) :: cny_fault(:, :), node_to_elem_val(:, :), &
                                  node_to_elem_size(:), id_dof(:), &
                                  lmat_index(:, :), ltmat_index(:, :), target_id_val(:), &
                                  node_id_in_patch(:), slip_assigned_num(:), &
                                  slip_id_start(:)
        type(mat), intent(inout) :: gmat_arr(:)
        real(c_double), allocatable :: th(:), x(:), lk(:), pr(:), mn(:), gt(:), g(:), cr(:), lu(:), lm(:), lv(:), &
                                       ltv(:), ll(:)
        integer(c_int), allocatable :: cn(:), n2e(:), n2es(:), idd(:), li(:), lti(:)
        real(c_double) :: neglog
        integer(c_int) :: nstage
        integer :: ndim, nn, npatch, ipath, k, iobs
        ndim = 2*ndof_total
        nn = nnode_total
        npatch = size(cny_fault, 2)
        call smcf_ensure_problem(nplane, nxi_ls, neta_ls, xmin, xmax, zmin, fix_xbend, xbend, nobs, nsar_total, &
                                 ngnss, npath, nsar_index, obs_points, obs_unitvec, obs_sigma, dvec, max_slip, &
                                 nparticle_slip, sigma_sar_mat)
        allocate (th(size(theta)), x(ndim*nparticle_slip), lk(nparticle_slip), pr(nparticle_slip), mn(ndim), &
                  gt(202), g(nobs*ndim), cr(2*nn), lu(nn*nn), lm(2*nn*ndim), lv(18*nn), ltv(9*ndim), ll(ndim*ndim), &
                  cn(4*npatch), n2e(4*nn), n2es(nn), idd(ndof_total), li(18*nn), lti(9*ndim))
        th = theta
        ! the reference's intermediate results (init.f90:198-541, gfunc.f90:242-348)
        call smcf_check(smcf_discretize_fault(g_pb, th, cn, cr, n2e, n2es, idd, lu, lm), "smcf_discretize_fault")
        cny_fault(1:4, 1:npatch) = reshape(cn, [4, npatch])
        coor_fault(1:2, 1:nn) = reshape(cr, [2, nn])
        node_to_elem_val(1:4, 1:nn) = reshape(n2e, [4, nn])
        node_to_elem_size(1:nn) = n2es
        id_dof(1:ndof_total) = idd
        luni(1:nn, 1:nn) = reshape(lu, [nn, nn])
        lmat(1:2*nn, 1:ndim) = reshape(lm, [2*nn, ndim])
        call smcf_check(smcf_laplacian(g_pb, th, li, lv, lti, ltv, ll), "smcf_laplacian")
        lmat_index(1:9, 1:2*nn) = reshape(li, [9, 2*nn])
        lmat_val(1:9, 1:2*nn) = reshape(lv, [9, 2*nn])
        ltmat_index(1:9, 1:ndim) = reshape(lti, [9, ndim])
        ltmat_val(1:9, 1:ndim) = reshape(ltv, [9, ndim])
        llmat(1:ndim, 1:ndim) = reshape(ll, [ndim, ndim])
        call smcf_check(smcf_calc_greens_func(g_pb, 1_c_int, th, g), "smcf_calc_greens_func")
        gmat(1:nobs, 1:ndim) = reshape(g, [nobs, ndim])
        do ipath = 1, npath
            k = nsar_index(ipath)
            do iobs = nsar_index(ipath), nsar_index(ipath + 1) - 1
                gmat_arr(ipath)%body(iobs - k + 1, 1:ndim) = gmat(iobs, 1:ndim)
            end do
        end do
        ! smc_fault.f90:170-173 (the log alpha^2 entry of theta)
        if (fix_xbend) then
            alpha2_full(1:2*nn) = exp(theta(2*nplane + 2))
        else
            alpha2_full(1:2*nn) = exp(theta(3*nplane + 1))
        end if
        call smcf_check(smcf_slip_smc_exec(g_pb, th, smcf_next_key(), x, lk, pr, mn, nstage, neglog, gt), &
                        "smcf_slip_smc_exec")
        slip_particles(1:ndim, 1:nparticle_slip) = reshape(x, [ndim, nparticle_slip])
        slip_likelihood_ls(1:nparticle_slip) = lk
        slip_prior_ls(1:nparticle_slip) = pr
        slip_mean(1:ndim) = mn
        if (flag_output == 1) call smcf_write_slip_particles(output_path, nparticle_slip, nn, ndof_total, id_dof, &
                                                             slip_particles, slip_likelihood_ls, slip_prior_ls)
        fault_calc_likelihood = neglog
    end function fault_calc_likelihood

    ! smc_fault.f90:294-355: all work_size particles as device batches (particle i keyed as the outer SMC's
    ! initial evaluation i)
    subroutine work_eval_init_particles(myid, work_size, nplane, ndim, particle_cur, &
                                        work_particles, work_likelihood_ls, nxi_ls, neta_ls, &
                                        nnode_total, ndof_total, ndof_index, ngnss, nobs, &
                                        cny_fault, coor_fault, node_to_elem_val, node_to_elem_size, &
                                        id_dof, luni, lmat, lmat_index, lmat_val, ltmat_index, ltmat_val, llmat, gmat, &
                                        slip_dist, obs_points, obs_unitvec, obs_sigma, sigma2_full, alpha2_full, &
                                        target_id_val, node_id_in_patch, xinode, etanode, uxinode, &
                                        uetanode, r1vec, r2vec, nvec, response_dist, uobs, uret, &
                                        slip_particles, slip_particles_new, nparticle_slip, max_slip, &
                                        dvec, slip_likelihood_ls, slip_prior_ls, slip_weights, slip_mean, &
                                        slip_cov, slip_likelihood_ls_new, slip_prior_ls_new, slip_assigned_num, &
                                        slip_id_start, slip_st_rand_ls, slip_metropolis_ls, gsvec, lsvec, &
                                        slip_particle_cur, slip_particle_cand, slip_st_rand, npath, nsar_index, &
                                        nsar_total, sigma_sar_mat, gmat_arr, xmin, xmax, zmin, &
                                        fix_xbend, xbend, gsmc, qmean, qcov_inv, log_qcov_det, work_q_ls)
        integer, intent(in) :: myid, work_size, nplane, ndim, nxi_ls(:), neta_ls(:), nnode_total, &
                               ndof_total, ndof_index(:), ngnss, nobs, nparticle_slip, npath, &
                               nsar_index(:), nsar_total
        double precision, intent(in) :: work_particles(:, :), obs_points(:, :), &
            obs_unitvec(:, :), obs_sigma(:), max_slip, dvec(:), xmin, xmax, zmin, xbend(:), &
            qmean(:), qcov_inv(:, :), log_qcov_det
        type(mat), intent(in) :: sigma_sar_mat(:)
        logical, intent(in) :: fix_xbend, gsmc
        integer, intent(inout) :: cny_fault(:, :), node_to_elem_val(:, :), node_to_elem_size(:), &
                                  id_dof(:), lmat_index(:, :), ltmat_index(:, :), target_id_val(:), node_id_in_patch(:), &
                                  slip_assigned_num(:), slip_id_start(:)
        double precision, intent(inout) :: work_likelihood_ls(:), particle_cur(:), coor_fault(:, :), luni(:, :), &
            lmat(:, :), lmat_val(:, :), ltmat_val(:, :), llmat(:, :), gmat(:, :), slip_dist(:, :), sigma2_full(:), &
            alpha2_full(:), xinode(:), &
            etanode(:), uxinode(:), uetanode(:), r1vec(:), r2vec(:), nvec(:), response_dist(:, :), &
            uobs(:), uret(:), slip_particles(:, :), slip_particles_new(:, :), &
            slip_likelihood_ls(:), slip_prior_ls(:), slip_weights(:), slip_mean(:), slip_cov(:, :), &
            slip_likelihood_ls_new(:), slip_prior_ls_new(:), slip_st_rand_ls(:, :), slip_metropolis_ls(:), &
            gsvec(:), lsvec(:), slip_particle_cur(:), slip_particle_cand(:), slip_st_rand(:), work_q_ls(:)
        type(mat), intent(inout) :: gmat_arr(:)
        double precision, allocatable :: th(:), out(:)
        integer(c_int64_t), allocatable :: keys(:)
        integer(c_int), allocatable :: ns(:)
        integer :: i0, nb, k
        call smcf_ensure_problem(nplane, nxi_ls, neta_ls, xmin, xmax, zmin, fix_xbend, xbend, nobs, nsar_total, &
                                 ngnss, npath, nsar_index, obs_points, obs_unitvec, obs_sigma, dvec, max_slip, &
                                 nparticle_slip, sigma_sar_mat)
        i0 = 1
        do while (i0 <= work_size)
            nb = min(g_max_batch, work_size - i0 + 1)
            allocate (th(ndim*nb), out(nb), keys(nb), ns(nb))
            th = reshape(work_particles(1:ndim, i0:i0 + nb - 1), [ndim*nb])
            do k = 1, nb
                keys(k) = smcf_inner_key(g_seed, 0_c_int32_t, 0_c_int32_t, int(i0 + k - 2, c_int32_t))
            end do
            call smcf_check(smcf_fault_calc_likelihood_batch(g_pb, int(nb, c_int), th, keys, out, ns), &
                            "smcf_fault_calc_likelihood_batch")
            work_likelihood_ls(i0:i0 + nb - 1) = out
            if (gsmc) then
                do k = i0, i0 + nb - 1
                    work_q_ls(k) = fault_calc_neglog_q(work_particles(:, k), qmean, qcov_inv, log_qcov_det, ndim)
                end do
            end if
            deallocate (th, out, keys, ns)
            i0 = i0 + nb
        end do
    end subroutine work_eval_init_particles

    ! smc_fault.f90:372-390
    subroutine calc_mean_std_vector(vec, size, mean, std)
        double precision, intent(in) :: vec(:)
        integer, intent(in) :: size
        double precision, intent(inout) :: mean, std
        call smcf_check(smcf_fault_mean_std(int(size, c_int), vec(1:size), mean, std), "smcf_fault_mean_std")
    end subroutine calc_mean_std_vector

    ! smc_fault.f90:392-485: gamma of the next outer stage; counts the stage for the keyed draws that follow
    subroutine fault_find_next_gamma(gamma, gamma_prev, likelihood_ls, weights, nparticle, fault_evidence, gsmc, q_ls)
        double precision, intent(inout) :: gamma, fault_evidence
        double precision, intent(in) :: gamma_prev
        double precision, intent(in) :: likelihood_ls(:), q_ls(:)
        double precision, intent(inout) :: weights(:)
        integer, intent(in) :: nparticle
        logical, intent(in) :: gsmc
        real(c_double), allocatable, target :: q(:)
        real(c_double), allocatable :: w(:)
        real(c_double) :: g, ev
        allocate (w(nparticle), q(nparticle))
        if (gsmc) then
            q = q_ls(1:nparticle)
            call smcf_check(smcf_fault_find_next_gamma(int(nparticle, c_int), gamma_prev, likelihood_ls(1:nparticle), &
                                                       c_loc(q), w, g, ev), "smcf_fault_find_next_gamma")
        else
            call smcf_check(smcf_fault_find_next_gamma(int(nparticle, c_int), gamma_prev, likelihood_ls(1:nparticle), &
                                                       c_null_ptr, w, g, ev), "smcf_fault_find_next_gamma")
        end if
        weights(1:nparticle) = w
        gamma = g
        fault_evidence = fault_evidence + ev
        g_stage = g_stage + 1
    end subroutine fault_find_next_gamma

    ! smc_fault.f90:487-501
    subroutine fault_normalize_weights(weights, nparticle)
        double precision, intent(inout) :: weights(:)
        integer, intent(in) :: nparticle
        real(c_double), allocatable :: w(:)
        allocate (w(nparticle))
        w = weights(1:nparticle)
        call smcf_check(smcf_fault_normalize_weights(int(nparticle, c_int), w), "smcf_fault_normalize_weights")
        weights(1:nparticle) = w
    end subroutine fault_normalize_weights

    ! smc_fault.f90:503-520
    subroutine fault_calc_mean_particles(particles, weights, mean, nparticle, ndim)
        double precision, intent(in) :: particles(:, :), weights(:)
        double precision, intent(inout) :: mean(:)
        integer, intent(in) :: nparticle, ndim
        real(c_double), allocatable :: m(:)
        allocate (m(ndim))
        call smcf_check(smcf_fault_calc_mean_particles(int(ndim, c_int), int(nparticle, c_int), &
                                                       reshape(particles(1:ndim, 1:nparticle), [ndim*nparticle]), &
                                                       weights(1:nparticle), m), "smcf_fault_calc_mean_particles")
        mean(1:ndim) = m
    end subroutine fault_calc_mean_particles

    ! smc_fault.f90:522-570: cov receives the lower Cholesky factor of 0.04 x the covariance (info ignored)
    subroutine fault_calc_cov_particles(particles, weights, mean, cov, nparticle, ndim)
        double precision, intent(in) :: particles(:, :), weights(:), mean(:)
        double precision, intent(inout) :: cov(:, :)
        integer, intent(in) :: nparticle, ndim
        real(c_double), allocatable :: c(:)
        integer(c_int) :: info
        allocate (c(ndim*ndim))
        call smcf_check(smcf_fault_calc_cov_particles(int(ndim, c_int), int(nparticle, c_int), &
                                                      reshape(particles(1:ndim, 1:nparticle), [ndim*nparticle]), &
                                                      weights(1:nparticle), mean(1:ndim), c, info), &
                        "smcf_fault_calc_cov_particles")
        cov(1:ndim, 1:ndim) = reshape(c, [ndim, ndim])
    end subroutine fault_calc_cov_particles

    ! smc_fault.f90:572-590, u keyed by (seed, outer stage)
    subroutine fault_resample_particles(nparticle, weights, assigned_num)
        integer, intent(in) :: nparticle
        double precision, intent(in) :: weights(:)
        integer, intent(inout) :: assigned_num(:)
        integer(c_int), allocatable :: a(:)
        allocate (a(nparticle))
        call smcf_check(smcf_fault_resample_particles(int(nparticle, c_int), weights(1:nparticle), g_seed, &
                                                      int(g_stage, c_int), a), "smcf_fault_resample_particles")
        assigned_num(1:nparticle) = a
    end subroutine fault_resample_particles

    ! smc_fault.f90:592-637 with numprocs = 1 (the tmp_ arrays keep the order before the sort, as in the reference)
    subroutine fault_reorder_to_send(assigned_num, particles, tmp_assigned_num, tmp_particles, sorted_idx, nparticle, &
                                     ndim, numprocs, work_size, tmp_likelihood_ls, likelihood_ls, gsmc, q_ls, tmp_q_ls)
        integer, intent(inout) :: assigned_num(:), tmp_assigned_num(:)
        double precision, intent(inout) :: particles(:, :), tmp_particles(:, :), tmp_likelihood_ls(:), &
                                           likelihood_ls(:), q_ls(:), tmp_q_ls(:)
        integer, intent(inout) :: sorted_idx(:)
        integer, intent(in) :: nparticle, ndim, numprocs, work_size
        logical, intent(in) :: gsmc
        integer(c_int), allocatable :: a(:), sidx(:)
        real(c_double), allocatable :: x(:), l(:)
        real(c_double), allocatable, target :: q(:)
        if (numprocs /= 1) then
            print *, "fault_reorder_to_send: this build runs one rank (numprocs = 1)"
            stop 1
        end if
        tmp_assigned_num(1:nparticle) = assigned_num(1:nparticle)
        tmp_particles(1:ndim, 1:nparticle) = particles(1:ndim, 1:nparticle)
        tmp_likelihood_ls(1:nparticle) = likelihood_ls(1:nparticle)
        if (gsmc) tmp_q_ls(1:nparticle) = q_ls(1:nparticle)
        allocate (a(nparticle), sidx(nparticle), x(ndim*nparticle), l(nparticle), q(nparticle))
        a = assigned_num(1:nparticle)
        x = reshape(particles(1:ndim, 1:nparticle), [ndim*nparticle])
        l = likelihood_ls(1:nparticle)
        if (gsmc) then
            q = q_ls(1:nparticle)
            call smcf_check(smcf_fault_reorder_to_send(int(ndim, c_int), int(nparticle, c_int), a, x, l, c_loc(q), &
                                                       sidx), "smcf_fault_reorder_to_send")
            q_ls(1:nparticle) = q
        else
            call smcf_check(smcf_fault_reorder_to_send(int(ndim, c_int), int(nparticle, c_int), a, x, l, c_null_ptr, &
                                                       sidx), "smcf_fault_reorder_to_send")
        end if
        assigned_num(1:nparticle) = a
        particles(1:ndim, 1:nparticle) = reshape(x, [ndim, nparticle])
        likelihood_ls(1:nparticle) = l
        sorted_idx(1:nparticle) = sidx
    end subroutine fault_reorder_to_send

    ! smc_fault.f90:639-783 with numprocs = 1: the parents' re-evaluation and every child as device batches
    subroutine work_mcmc_sampling(work_assigned_num, work_particles, work_particles_new, &
                                  work_likelihood_ls, work_likelihood_ls_new, nplane, id_start, &
                                  particle_cur, particle_cand, st_rand, work_size, ndim, &
                                  cov, gamma, myid, nxi_ls, neta_ls, nnode_total, ndof_total, ndof_index, &
                                  ngnss, nobs, cny_fault, coor_fault, &
                                  node_to_elem_val, node_to_elem_size, id_dof, &
                                  luni, lmat, lmat_index, lmat_val, ltmat_index, ltmat_val, llmat, gmat, &
                                  slip_dist, obs_points, obs_unitvec, obs_sigma, &
                                  sigma2_full, alpha2_full, target_id_val, node_id_in_patch, &
                                  xinode, etanode, uxinode, uetanode, r1vec, r2vec, &
                                  nvec, response_dist, uobs, uret, slip_particles, &
                                  slip_particles_new, nparticle_slip, max_slip, dvec, &
                                  slip_likelihood_ls, slip_prior_ls, slip_weights, &
                                  slip_mean, slip_cov, slip_likelihood_ls_new, slip_prior_ls_new, &
                                  slip_assigned_num, slip_id_start, slip_st_rand_ls, &
                                  slip_metropolis_ls, gsvec, lsvec, slip_particle_cur, &
                                  slip_particle_cand, slip_st_rand, work_acc_count, range, &
                                  npath, nsar_index, nsar_total, sigma_sar_mat, gmat_arr, &
                                  xmin, xmax, zmin, fix_xbend, xbend, gsmc, work_q_ls, work_q_ls_new, &
                                  qmean, qcov_inv, log_qcov_det)
        integer, intent(in) :: work_assigned_num(:), nplane, id_start(:), work_size, ndim, myid, &
                               nxi_ls(:), neta_ls(:), nnode_total, ndof_total, ndof_index(:), ngnss, nobs, &
                               nparticle_slip, npath, nsar_index(:), nsar_total
        double precision, intent(in) :: work_particles(:, :), work_likelihood_ls(:), cov(:, :), gamma, &
            obs_points(:, :), obs_unitvec(:, :), obs_sigma(:), max_slip, dvec(:), range(:, :), xmin, xmax, zmin, &
            xbend(:), qmean(:), qcov_inv(:, :), log_qcov_det
        type(mat), intent(in) :: sigma_sar_mat(:)
        logical, intent(in) :: fix_xbend, gsmc
        integer, intent(inout) :: cny_fault(:, :), node_to_elem_val(:, :), node_to_elem_size(:), &
                                  id_dof(:), lmat_index(:, :), ltmat_index(:, :), target_id_val(:), &
                                  node_id_in_patch(:), slip_assigned_num(:), slip_id_start(:), work_acc_count
        double precision, intent(inout) :: work_particles_new(:, :), work_likelihood_ls_new(:), &
            particle_cur(:), particle_cand(:), st_rand(:), coor_fault(:, :), luni(:, :), &
            lmat(:, :), lmat_val(:, :), ltmat_val(:, :), llmat(:, :), gmat(:, :), slip_dist(:, :), &
            sigma2_full(:), alpha2_full(:), xinode(:), etanode(:), uxinode(:), uetanode(:), r1vec(:), r2vec(:), &
            nvec(:), response_dist(:, :), uobs(:), uret(:), slip_particles(:, :), slip_particles_new(:, :), &
            slip_likelihood_ls(:), slip_prior_ls(:), slip_weights(:), slip_mean(:), slip_cov(:, :), &
            slip_likelihood_ls_new(:), slip_prior_ls_new(:), slip_st_rand_ls(:, :), slip_metropolis_ls(:), &
            gsvec(:), lsvec(:), slip_particle_cur(:), slip_particle_cand(:), slip_st_rand(:), work_q_ls(:), &
            work_q_ls_new(:)
        type(mat), intent(inout) :: gmat_arr(:)
        real(c_double), allocatable :: r(:), c(:), x(:), xn(:), ln(:), qn(:)
        real(c_double), allocatable, target :: q(:), qm(:), qi(:)
        integer(c_int), allocatable :: a(:)
        integer(c_int) :: acc
        integer :: n
        n = work_size
        call smcf_ensure_problem(nplane, nxi_ls, neta_ls, xmin, xmax, zmin, fix_xbend, xbend, nobs, nsar_total, &
                                 ngnss, npath, nsar_index, obs_points, obs_unitvec, obs_sigma, dvec, max_slip, &
                                 nparticle_slip, sigma_sar_mat)
        allocate (r(2*ndim), c(ndim*ndim), x(ndim*n), xn(ndim*n), ln(n), qn(n), q(n), qm(ndim), qi(ndim*ndim), a(n))
        r = reshape(range(1:2, 1:ndim), [2*ndim])
        c = reshape(cov(1:ndim, 1:ndim), [ndim*ndim])
        x = reshape(work_particles(1:ndim, 1:n), [ndim*n])
        a = work_assigned_num(1:n)
        if (gsmc) then
            q = work_q_ls(1:n)
            qm = qmean(1:ndim)
            qi = reshape(qcov_inv(1:ndim, 1:ndim), [ndim*ndim])
            call smcf_check(smcf_work_mcmc_sampling(g_pb, g_seed, int(g_stage, c_int), gamma, r, c, int(n, c_int), a, &
                                                    x, c_loc(q), c_loc(qm), c_loc(qi), log_qcov_det, &
                                                    int(g_max_batch, c_int), xn, ln, qn, acc), &
                            "smcf_work_mcmc_sampling")
            work_q_ls_new(1:n) = qn
        else
            call smcf_check(smcf_work_mcmc_sampling(g_pb, g_seed, int(g_stage, c_int), gamma, r, c, int(n, c_int), a, &
                                                    x, c_null_ptr, c_null_ptr, c_null_ptr, 0d0, &
                                                    int(g_max_batch, c_int), xn, ln, qn, acc), &
                            "smcf_work_mcmc_sampling")
        end if
        work_particles_new(1:ndim, 1:n) = reshape(xn, [ndim, n])
        work_likelihood_ls_new(1:n) = ln
        work_acc_count = acc
    end subroutine work_mcmc_sampling

    ! smc_fault.f90:785-1075 with numprocs = 1: the whole outer SMC (per-stage <output_dir><iter>.csv dumps)
    subroutine fault_smc_exec( &
        output_dir, range, nplane, nparticle, ndim, &
        myid, numprocs, nxi_ls, neta_ls, nnode_total, ndof_total, ndof_index, ngnss, nobs, cny_fault, &
        coor_fault, node_to_elem_val, node_to_elem_size, id_dof, luni, lmat, &
        lmat_index, lmat_val, ltmat_index, ltmat_val, llmat, gmat, slip_dist, obs_points, &
        obs_unitvec, obs_sigma, sigma2_full, alpha2_full, target_id_val, node_id_in_patch, &
        xinode, etanode, uxinode, uetanode, r1vec, r2vec, nvec, &
        response_dist, uobs, uret, slip_particles, &
        slip_particles_new, nparticle_slip, max_slip, dvec, gsvec, &
        lsvec, slip_likelihood_ls, slip_prior_ls, slip_weights, slip_mean, &
        slip_cov, slip_likelihood_ls_new, slip_prior_ls_new, &
        slip_st_rand, slip_particle_cur, slip_particle_cand, &
        slip_assigned_num, slip_id_start, slip_st_rand_ls, slip_metropolis_ls, &
        npath, nsar_index, nsar_total, sigma_sar_mat, gmat_arr, xmin, xmax, zmin, &
        fix_xbend, xbend, gsmc, qmean, qcov, qcov_inv, log_qcov_det)
        character(*), intent(in) :: output_dir
        double precision, intent(in) :: obs_points(:, :), &
            obs_unitvec(:, :), obs_sigma(:), max_slip, dvec(:), range(:, :), &
            xmin, xmax, zmin, xbend(:), qmean(:), qcov(:, :), qcov_inv(:, :), log_qcov_det
        integer, intent(in) :: nplane, nxi_ls(:), neta_ls(:), nnode_total, ndof_total, ndof_index(:), ngnss, nobs, &
                               nparticle_slip, nparticle, ndim, &
                               myid, numprocs, npath, nsar_index(:), nsar_total
        type(mat), intent(in) :: sigma_sar_mat(:)
        logical, intent(in) :: fix_xbend, gsmc
        double precision, intent(inout) :: coor_fault(:, :), luni(:, :), &
            lmat(:, :), lmat_val(:, :), ltmat_val(:, :), llmat(:, :), gmat(:, :), slip_dist(:, :), &
            sigma2_full(:), alpha2_full(:), xinode(:), etanode(:), uxinode(:), uetanode(:), &
            r1vec(:), r2vec(:), nvec(:), response_dist(:, :), uobs(:), uret(:), &
            slip_particles(:, :), slip_particles_new(:, :), gsvec(:), lsvec(:), &
            slip_likelihood_ls(:), slip_prior_ls(:), slip_weights(:), slip_mean(:), &
            slip_cov(:, :), slip_likelihood_ls_new(:), slip_prior_ls_new(:), &
            slip_st_rand(:), slip_particle_cur(:), slip_particle_cand(:), &
            slip_st_rand_ls(:, :), slip_metropolis_ls(:)
        integer, intent(inout) :: cny_fault(:, :), node_to_elem_val(:, :), node_to_elem_size(:), id_dof(:), &
                                  lmat_index(:, :), ltmat_index(:, :), target_id_val(:), node_id_in_patch(:), &
                                  slip_assigned_num(:), slip_id_start(:)
        type(mat), intent(inout) :: gmat_arr(:)
        real(c_double), allocatable :: r(:), x(:), l(:), gt(:), qm(:), qc(:), qi(:)
        real(c_double) :: ev
        integer(c_int) :: ns
        if (numprocs /= 1) then
            print *, "fault_smc_exec: this build runs one rank (numprocs = 1)"
            stop 1
        end if
        call smcf_ensure_problem(nplane, nxi_ls, neta_ls, xmin, xmax, zmin, fix_xbend, xbend, nobs, nsar_total, &
                                 ngnss, npath, nsar_index, obs_points, obs_unitvec, obs_sigma, dvec, max_slip, &
                                 nparticle_slip, sigma_sar_mat)
        allocate (r(2*ndim), x(ndim*nparticle), l(nparticle), gt(g_outer_max_stages + 1), qm(ndim), qc(ndim*ndim), &
                  qi(ndim*ndim))
        r = reshape(range(1:2, 1:ndim), [2*ndim])
        call smcf_check(smcf_fault_smc_set_output_dir(g_pb, trim(output_dir)//c_null_char), &
                        "smcf_fault_smc_set_output_dir")
        if (gsmc) then
            qm = qmean(1:ndim)
            qc = reshape(qcov(1:ndim, 1:ndim), [ndim*ndim])
            qi = reshape(qcov_inv(1:ndim, 1:ndim), [ndim*ndim])
            call smcf_check(smcf_fault_smc_exec_gsmc(g_pb, r, int(nparticle, c_int), g_seed, &
                                                     int(g_outer_max_stages, c_int), int(g_max_batch, c_int), qm, qc, &
                                                     qi, log_qcov_det, x, l, ns, ev, gt), "smcf_fault_smc_exec_gsmc")
        else
            call smcf_check(smcf_fault_smc_exec(g_pb, r, int(nparticle, c_int), g_seed, int(g_outer_max_stages, c_int), &
                                                int(g_max_batch, c_int), x, l, ns, ev, gt), "smcf_fault_smc_exec")
        end if
        call smcf_check(smcf_fault_smc_set_output_dir(g_pb, c_null_char), "smcf_fault_smc_set_output_dir")
    end subroutine fault_smc_exec
end module smc_fault
