! Fortran side of the drop-in boundary: modules with the reference's names and
! argument lists that forward, through iso_c_binding, to libsmcf (include/smcf.h).
!
!   module type_mat   mat.f90:1-7 (same derived type)
!   module smcf_c     C interfaces + the cached device context/problem
!   module gfunc      calc_greens_func            gfunc.f90:242-348
!   module smc_fault  fault_calc_likelihood       smc_fault.f90:66-292
!                     work_eval_init_particles    smc_fault.f90:294-355
!                     fault_calc_neglog_q         smc_fault.f90:9-27
!
! Workspace arguments of the reference (slip_dist, response_dist, target_id_val, ...)
! are accepted and left untouched: their contents are scratch in the reference too.
! Random numbers: each evaluation is keyed by (seed, call counter); set the seed with
! smcf_fortran_seed.  The inner mover is chosen with smcf_fortran_mover.
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
        integer(c_int64_t) function smcf_inner_key(seed, stage, role, slot) bind(C, name="smcf_inner_key")
            import :: c_int64_t, c_int32_t
            integer(c_int64_t), value :: seed
            integer(c_int32_t), value :: stage, role, slot
        end function
        type(c_ptr) function smcf_last_error() bind(C, name="smcf_last_error")
            import :: c_ptr
        end function
    end interface

    type(c_ptr), save :: g_ctx = c_null_ptr
    type(c_ptr), save :: g_pb = c_null_ptr
    integer, save :: g_sig(8) = -1
    double precision, save :: g_sum = -1d300
    integer(c_int64_t), save :: g_seed = 0_c_int64_t
    integer(c_int32_t), save :: g_counter = 0
    ! default: the reference mover (HMC, ntau 50, ESJD tuning, smc_slip.f90:591-1214)
    integer, save :: g_mover = 0, g_sweeps = 5, g_max_batch = 64, g_precision = 0, g_hmc_ntau = 10
    double precision, save :: g_hmc_accept = 0.65d0
    ! arrays kept alive for the C side
    integer(c_int), allocatable, target, save :: t_nxi(:), t_neta(:), t_nsi(:)
    real(c_double), allocatable, target, save :: t_xb(:), t_pts(:), t_uv(:), t_sig(:), t_d(:), t_sar(:)

contains
    subroutine smcf_fortran_seed(seed)
        integer(c_int64_t), intent(in) :: seed
        g_seed = seed
        g_counter = 0
    end subroutine

    subroutine smcf_fortran_mover(mover, sweeps, max_batch)
        integer, intent(in) :: mover, sweeps, max_batch
        g_mover = mover
        g_sweeps = sweeps
        g_max_batch = max_batch
        g_sig = -1
    end subroutine

    ! arithmetic of the Gibbs mover: 0 = FP64 (default), 1 = mixed (FP32 matrix products and draws)
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

    ! (re)create the device problem when the observation set or mesh changes
    ! sigma_sar_mat (optional): the inverse InSAR covariance per path (calc_sigma_sar_mat); without it the
    ! blocks are identities, which is enough for Green's matrices alone
    subroutine smcf_ensure_problem(nplane, nxi_ls, neta_ls, xmin, xmax, zmin, fix_xbend, xbend, nobs, &
                                   nsar_total, ngnss, npath, nsar_index, obs_points, obs_unitvec, &
                                   obs_sigma, dvec, max_slip, nparticle_slip, sigma_sar_mat)
        integer, intent(in) :: nplane, nxi_ls(:), neta_ls(:), nobs, nsar_total, ngnss, npath, nsar_index(:), &
                               nparticle_slip
        double precision, intent(in) :: xmin, xmax, zmin, xbend(:), obs_points(:, :), obs_unitvec(:, :), &
                                        obs_sigma(:), dvec(:), max_slip
        logical, intent(in) :: fix_xbend
        type(mat), intent(in), optional :: sigma_sar_mat(:)
        type(smcf_problem_desc) :: d
        integer :: sig(8), ip, n, off, k, nsinv
        double precision :: s
        sig = [nplane, nobs, nsar_total, ngnss, npath, nparticle_slip, sum(nxi_ls) + 1000*sum(neta_ls), &
               merge(1, 0, fix_xbend)]
        s = sum(obs_points) + 3d0*sum(obs_unitvec) + 7d0*sum(obs_sigma) + 11d0*sum(dvec) + xmin + 2d0*xmax &
            + 5d0*zmin + max_slip
        nsinv = 0
        do ip = 1, npath
            n = nsar_index(ip + 1) - nsar_index(ip)
            nsinv = nsinv + n*n
            if (present(sigma_sar_mat) .and. n > 0) s = s + 13d0*sum(sigma_sar_mat(ip)%body(1:n, 1:n))
        end do
        if (c_associated(g_pb) .and. all(sig == g_sig) .and. s == g_sum) return
        if (.not. c_associated(g_ctx)) call smcf_check(smcf_ctx_create(0_c_int, g_ctx), "smcf_ctx_create")
        if (c_associated(g_pb)) call smcf_check(smcf_problem_destroy(g_pb), "smcf_problem_destroy")
        if (allocated(t_nxi)) deallocate (t_nxi, t_neta, t_nsi, t_xb, t_pts, t_uv, t_sig, t_d, t_sar)
        allocate (t_nxi(nplane), t_neta(nplane), t_nsi(npath + 1), t_xb(max(nplane - 1, 1)), &
                  t_pts(2*nobs), t_uv(3*nobs), t_sig(nobs), t_d(nobs), t_sar(max(nsinv, 1)))
        t_nxi = nxi_ls(1:nplane)
        t_neta = neta_ls(1:nplane)
        t_nsi = nsar_index(1:npath + 1)
        t_xb = 0d0
        if (nplane > 1) t_xb(1:nplane - 1) = xbend(1:nplane - 1)
        t_pts = reshape(obs_points(1:2, 1:nobs), [2*nobs])
        t_uv = reshape(obs_unitvec(1:3, 1:nobs), [3*nobs])
        t_sig = obs_sigma(1:nobs)
        t_d = dvec(1:nobs)
        t_sar = 0d0
        off = 0
        do ip = 1, npath
            n = nsar_index(ip + 1) - nsar_index(ip)
            if (n <= 0) cycle
            if (present(sigma_sar_mat)) then
                t_sar(off + 1:off + n*n) = reshape(sigma_sar_mat(ip)%body(1:n, 1:n), [n*n])
            else
                do k = 1, n
                    t_sar(off + (k - 1)*n + k) = 1d0
                end do
            end if
            off = off + n*n
        end do
        d%nplane = nplane
        d%nxi_ls = c_loc(t_nxi)
        d%neta_ls = c_loc(t_neta)
        d%xmin = xmin
        d%xmax = xmax
        d%zmin = zmin
        d%fix_xbend = merge(1, 0, fix_xbend)
        d%xbend = c_loc(t_xb)
        d%nobs = nobs
        d%nsar_total = nsar_total
        d%ngnss = ngnss
        d%npath = npath
        d%nsar_index = c_loc(t_nsi)
        d%obs_points = c_loc(t_pts)
        d%obs_unitvec = c_loc(t_uv)
        d%obs_sigma = c_loc(t_sig)
        d%dvec = c_loc(t_d)
        d%sar_sinv = c_loc(t_sar)
        d%max_slip = max_slip
        d%nparticle_slip = nparticle_slip
        d%mover = g_mover
        d%gibbs_sweeps = g_sweeps
        d%precision = g_precision
        d%hmc_ntau = g_hmc_ntau
        d%hmc_accept = g_hmc_accept
        call smcf_check(smcf_problem_create(g_ctx, d, int(g_max_batch, c_int), g_pb), "smcf_problem_create")
        g_sig = sig
        g_sum = s
    end subroutine

    integer(c_int64_t) function smcf_next_key()
        smcf_next_key = smcf_inner_key(g_seed, 0_c_int32_t, 2_c_int32_t, g_counter)
        g_counter = g_counter + 1
    end function
end module smcf_c

module gfunc
    use smcf_c
    use type_mat
    implicit none
contains
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
                                 ngnss, npath, nsar_index, obs_points, obs_unitvec, ones, zeros, 10d0, 16)
        call smcf_check(smcf_calc_greens_func(g_pb, 1_c_int, th, g), "smcf_calc_greens_func")
        gmat(1:nobs, 1:2*ndof_total) = reshape(g, [nobs, 2*ndof_total])
        do ipath = 1, npath
            k = nsar_index(ipath)
            do iobs = nsar_index(ipath), nsar_index(ipath + 1) - 1
                gmat_arr(ipath)%body(iobs - k + 1, 1:2*ndof_total) = gmat(iobs, 1:2*ndof_total)
            end do
        end do
    end subroutine calc_greens_func
end module gfunc

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

    ! smc_fault.f90:66-292: -log Z of the inner slip SMC for one geometry
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
        double precision, allocatable :: th(:), x(:), lk(:), pr(:), mn(:), gt(:)
        double precision :: neglog
        integer(c_int) :: nstage
        integer :: ndim, iparticle, idof, inode, idirection
        double precision, allocatable :: slip(:, :)
        ndim = 2*ndof_total
        call smcf_ensure_problem(nplane, nxi_ls, neta_ls, xmin, xmax, zmin, fix_xbend, xbend, nobs, nsar_total, &
                                 ngnss, npath, nsar_index, obs_points, obs_unitvec, obs_sigma, dvec, max_slip, &
                                 nparticle_slip, sigma_sar_mat)
        allocate (th(size(theta)), x(ndim*nparticle_slip), lk(nparticle_slip), pr(nparticle_slip), mn(ndim), &
                  gt(202))
        th = theta
        call smcf_check(smcf_slip_smc_exec(g_pb, th, smcf_next_key(), x, lk, pr, mn, nstage, neglog, gt), &
                        "smcf_slip_smc_exec")
        slip_particles(1:ndim, 1:nparticle_slip) = reshape(x, [ndim, nparticle_slip])
        slip_likelihood_ls(1:nparticle_slip) = lk
        slip_prior_ls(1:nparticle_slip) = pr
        slip_mean(1:ndim) = mn
        if (flag_output == 1) then
            ! smc_slip.f90:1632-1660 output format
            allocate (slip(2, nnode_total))
            open (17, file=output_path, status='replace')
            do iparticle = 1, nparticle_slip
                slip = 0d0
                do idof = 1, ndof_total
                    inode = id_dof(idof)
                    slip(1, inode) = slip_particles(2*(idof - 1) + 1, iparticle)
                    slip(2, inode) = slip_particles(2*(idof - 1) + 2, iparticle)
                end do
                do inode = 1, nnode_total
                    do idirection = 1, 2
                        write (17, "(f12.5)", advance="no") slip(idirection, inode)
                    end do
                end do
                write (17, "(f12.5)", advance="no") slip_likelihood_ls(iparticle)
                write (17, "(f12.5)") slip_prior_ls(iparticle)
            end do
            close (17)
            deallocate (slip)
        end if
        fault_calc_likelihood = neglog
    end function fault_calc_likelihood

    ! smc_fault.f90:294-355: all work_size particles as device batches
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
                keys(k) = smcf_next_key()
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
end module smc_fault
