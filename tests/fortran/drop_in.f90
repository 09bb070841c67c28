! Drop-in check: a Fortran caller written against the reference's interfaces (gfunc / smc_slip / smc_fault module
! procedures, main.f90:349-390 workspace layout) linked against the libsmcf shim instead of the reference objects.
!
! usage: drop_in <in.bin> <out.bin> <outer_dump_dir/>
!   in : nplane fix nxi(nplane) neta(nplane) xmin xmax zmin [xbend] theta(ndim_fault)
!        nobs nsar_total ngnss npath nsar_index(npath+1) obs_points(2,nobs) obs_unitvec(3,nobs)
!        obs_sigma(nobs) dvec(nobs) nparticle_slip max_slip seed mover nouter nstage range(2,ndim_fault)
!   out: gmat(nobs, 2*ndof)  neglog
!        (GNSS-only data:) the workspace outputs of fault_calc_likelihood: cny coor node_to_elem_val
!        node_to_elem_size id_dof luni lmat lmat_index lmat_val ltmat_index ltmat_val llmat alpha2_full;
!        slip_smc_exec: particles(ndim, nparticle_slip) neglog;
!        an outer SMC of nstage stages composed from fault_sample_init_particles, work_eval_init_particles,
!        fault_find_next_gamma, fault_normalize_weights, fault_calc_mean_particles, fault_calc_cov_particles,
!        fault_resample_particles, fault_reorder_to_send and work_mcmc_sampling: particles(ndim_fault, nouter)
!        likelihood(nouter) evidence.  Then fault_smc_exec itself runs nstage stages and dumps its stages into
!        <outer_dump_dir>.
program drop_in
    use iso_c_binding
    use type_mat
    use gfunc
    use smc_slip
    use smc_fault
    use smcf_c, only: smcf_fortran_seed, smcf_fortran_mover, smcf_fortran_outer_stages
    implicit none
    character(len=512) :: fin, fout, fdir
    integer :: nplane, ifix, ndim_fault, nnode, ndof, npatch, ip, nobs, nsar_total, ngnss, npath, np, k, mover, &
               nouter, nstage, i, iter, acc
    integer(c_int64_t) :: seed
    integer, allocatable :: nxi_ls(:), neta_ls(:), ndof_index(:), nsar_index(:), cny(:, :), n2e(:, :), n2es(:), &
                            id_dof(:), target_id_val(:), node_id_in_patch(:), li(:, :), lti(:, :), &
                            assigned(:), id_start(:), o_assigned(:), o_tmp_assigned(:), o_sorted(:), o_id_start(:)
    double precision :: xmin, xmax, zmin, max_slip, neglog, neglog2, gamma, gamma_prev, evidence
    double precision, allocatable :: xbend(:), theta(:), obs_points(:, :), obs_unitvec(:, :), obs_sigma(:), dvec(:), &
                                     gmat(:, :), slip_dist(:, :), coor(:, :), response_dist(:, :), xinode(:), &
                                     etanode(:), uxinode(:), uetanode(:), r1vec(:), r2vec(:), nvec(:), uobs(:), uret(:), &
                                     luni(:, :), lmat(:, :), lv(:, :), ltv(:, :), llmat(:, :), sigma2_full(:), &
                                     alpha2_full(:), sp(:, :), spn(:, :), sl(:), spr(:), sw(:), smean(:), scov(:, :), &
                                     sln(:), sprn(:), strl(:, :), sml(:), gsvec(:), lsvec(:), scur(:), scand(:), sst(:), &
                                     range(:, :), o_x(:, :), o_xn(:, :), o_l(:), o_ln(:), o_w(:), o_mean(:), o_cov(:, :), &
                                     o_tmp_x(:, :), o_tmp_l(:), o_q(:), o_qn(:), o_tmp_q(:), o_cur(:), o_cand(:), &
                                     o_st(:), qdummy(:), qcdummy(:, :)
    type(mat), allocatable :: gmat_arr(:), sigma_sar_mat(:)
    logical :: fix_xbend
    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    call get_command_argument(3, fdir)
    open (21, file=trim(fin), access='stream', form='unformatted', status='old')
    read (21) nplane, ifix
    fix_xbend = ifix /= 0
    allocate (nxi_ls(nplane), neta_ls(nplane), ndof_index(nplane + 1), xbend(max(nplane - 1, 1)))
    read (21) nxi_ls, neta_ls
    read (21) xmin, xmax, zmin
    xbend = 0d0
    if (nplane > 1) read (21) xbend(1:nplane - 1)
    ndim_fault = merge(2*nplane + 4, 3*nplane + 3, fix_xbend)
    allocate (theta(ndim_fault), range(2, ndim_fault))
    read (21) theta
    read (21) nobs, nsar_total, ngnss, npath
    allocate (nsar_index(npath + 1), obs_points(2, nobs), obs_unitvec(3, nobs), obs_sigma(nobs), dvec(nobs))
    read (21) nsar_index, obs_points, obs_unitvec, obs_sigma, dvec
    read (21) np, max_slip, seed, mover, nouter, nstage
    read (21) range
    close (21)
    nnode = 0
    ndof = 0
    npatch = 0
    ndof_index(1) = 1
    do ip = 1, nplane
        nnode = nnode + (nxi_ls(ip) + 1)*(neta_ls(ip) + 1)
        npatch = npatch + nxi_ls(ip)*neta_ls(ip)
        k = (nxi_ls(ip) + 1)*(neta_ls(ip) + 1) - (nxi_ls(ip) + 1)
        if (ip == 1) k = k - neta_ls(ip)
        if (ip == nplane) k = k - neta_ls(ip)
        ndof = ndof + k
        ndof_index(ip + 1) = ndof_index(ip) + k
    end do
    ! workspaces as main.f90:349-390 allocates them
    allocate (cny(4, npatch), coor(2, nnode), n2e(4, nnode), n2es(nnode), id_dof(ndof), gmat(nobs, 2*ndof), &
              slip_dist(2, nnode), response_dist(3, nobs), luni(nnode, nnode), lmat(2*nnode, 2*ndof), &
              li(9, 2*nnode), lv(9, 2*nnode), lti(9, 2*ndof), ltv(9, 2*ndof), llmat(2*ndof, 2*ndof), &
              target_id_val(4), node_id_in_patch(4), xinode(4), etanode(4), uxinode(4), uetanode(4), r1vec(2), &
              r2vec(2), nvec(4), uobs(3), uret(3), sigma2_full(nobs), alpha2_full(2*nnode), &
              sp(2*ndof, np), spn(2*ndof, np), sl(np), spr(np), sw(np), smean(2*ndof), scov(2*ndof, 2*ndof), &
              sln(np), sprn(np), assigned(np), id_start(np), strl(2*ndof, np), sml(np), gsvec(nobs), &
              lsvec(2*nnode), scur(2*ndof), scand(2*ndof), sst(2*ndof), gmat_arr(npath), sigma_sar_mat(npath))
    do ip = 1, npath
        allocate (gmat_arr(ip)%body(max(nsar_index(ip + 1) - nsar_index(ip), 1), 2*ndof))
        allocate (sigma_sar_mat(ip)%body(1, 1))
    end do
    cny = 0
    coor = 0d0
    n2e = 0
    n2es = 0
    id_dof = 0
    call calc_greens_func(theta, nplane, nxi_ls, neta_ls, gmat, slip_dist, cny, coor, obs_points, &
                          obs_unitvec, n2e, n2es, id_dof, ngnss, nobs, nnode, ndof, ndof_index, target_id_val, &
                          node_id_in_patch, xinode, etanode, uxinode, uetanode, r1vec, r2vec, nvec, response_dist, &
                          uobs, uret, nsar_total, npath, nsar_index, gmat_arr, xmin, xmax, zmin, fix_xbend, xbend)
    open (22, file=trim(fout), access='stream', form='unformatted', status='replace')
    neglog = 0d0
    if (nsar_total /= 0) then
        write (22) gmat, neglog
        close (22)
        stop
    end if
    call smcf_fortran_seed(seed)
    call smcf_fortran_mover(mover, 3, 64)
    neglog = fault_calc_likelihood(theta, nplane, nxi_ls, neta_ls, nnode, ndof, ndof_index, ngnss, nobs, cny, &
                                   coor, n2e, n2es, id_dof, luni, lmat, li, lv, lti, ltv, llmat, gmat, slip_dist, &
                                   obs_points, obs_unitvec, obs_sigma, sigma2_full, alpha2_full, target_id_val, &
                                   node_id_in_patch, xinode, etanode, uxinode, uetanode, r1vec, r2vec, nvec, &
                                   response_dist, uobs, uret, sp, spn, np, max_slip, dvec, sl, spr, sw, smean, scov, &
                                   sln, sprn, assigned, id_start, strl, sml, gsvec, lsvec, scur, scand, sst, 0, "", &
                                   npath, nsar_index, nsar_total, sigma_sar_mat, gmat_arr, xmin, xmax, zmin, &
                                   fix_xbend, xbend)
    write (22) gmat, neglog
    write (22) cny, coor, n2e, n2es, id_dof, luni, lmat, li, lv, lti, ltv, llmat, alpha2_full
    ! the inner SMC alone, on the problem fault_calc_likelihood left cached (the second key of the seed)
    call slip_smc_exec(sp, spn, np, 2*ndof, 0, "", llmat, max_slip, dvec, sigma2_full, alpha2_full, theta, nplane, &
                       gmat, theta(ndim_fault - 1), theta(ndim_fault), ngnss, nobs, ndof, li, lv, lti, ltv, nnode, &
                       sl, spr, sw, smean, scov, sln, sprn, assigned, id_start, id_dof, strl, sml, gsvec, lsvec, scur, &
                       scand, sst, neglog2, npath, nsar_index, nsar_total, sigma_sar_mat, 1d0, 1d0, obs_sigma, &
                       gmat_arr)
    write (22) sp, neglog2
    ! an outer SMC composed from the module procedures, in the order of fault_smc_exec (smc_fault.f90:896-1069)
    allocate (o_x(ndim_fault, nouter), o_xn(ndim_fault, nouter), o_l(nouter), o_ln(nouter), o_w(nouter), &
              o_mean(ndim_fault), o_cov(ndim_fault, ndim_fault), o_tmp_x(ndim_fault, nouter), o_tmp_l(nouter), &
              o_q(nouter), o_qn(nouter), o_tmp_q(nouter), o_cur(ndim_fault), o_cand(ndim_fault), o_st(ndim_fault), &
              o_assigned(nouter), o_tmp_assigned(nouter), o_sorted(nouter), o_id_start(nouter), qdummy(ndim_fault), &
              qcdummy(ndim_fault, ndim_fault))
    qdummy = 0d0
    qcdummy = 0d0
    call smcf_fortran_seed(seed)
    call fault_sample_init_particles(o_x, nouter, ndim_fault, range, nplane, .false., qdummy, qcdummy)
    call work_eval_init_particles(0, nouter, nplane, ndim_fault, o_cur, o_x, o_l, nxi_ls, neta_ls, nnode, ndof, &
                                  ndof_index, ngnss, nobs, cny, coor, n2e, n2es, id_dof, luni, lmat, li, lv, lti, ltv, &
                                  llmat, gmat, slip_dist, obs_points, obs_unitvec, obs_sigma, sigma2_full, alpha2_full, &
                                  target_id_val, node_id_in_patch, xinode, etanode, uxinode, uetanode, r1vec, r2vec, &
                                  nvec, response_dist, uobs, uret, sp, spn, np, max_slip, dvec, sl, spr, sw, smean, &
                                  scov, sln, sprn, assigned, id_start, strl, sml, gsvec, lsvec, scur, scand, sst, &
                                  npath, nsar_index, nsar_total, sigma_sar_mat, gmat_arr, xmin, xmax, zmin, &
                                  fix_xbend, xbend, .false., qdummy, qcdummy, 0d0, o_q)
    gamma = 0d0
    evidence = 0d0
    do iter = 1, nstage
        gamma_prev = gamma
        call fault_find_next_gamma(gamma, gamma_prev, o_l, o_w, nouter, evidence, .false., o_q)
        call fault_normalize_weights(o_w, nouter)
        call fault_calc_mean_particles(o_x, o_w, o_mean, nouter, ndim_fault)
        call fault_calc_cov_particles(o_x, o_w, o_mean, o_cov, nouter, ndim_fault)
        call fault_resample_particles(nouter, o_w, o_assigned)
        call fault_reorder_to_send(o_assigned, o_x, o_tmp_assigned, o_tmp_x, o_sorted, nouter, ndim_fault, 1, &
                                   nouter, o_tmp_l, o_l, .false., o_q, o_tmp_q)
        o_id_start(1) = 1
        do i = 1, nouter - 1
            o_id_start(i + 1) = o_id_start(i) + o_assigned(i)
        end do
        call work_mcmc_sampling(o_assigned, o_x, o_xn, o_l, o_ln, nplane, o_id_start, o_cur, o_cand, o_st, nouter, &
                                ndim_fault, o_cov, gamma, 0, nxi_ls, neta_ls, nnode, ndof, ndof_index, ngnss, nobs, &
                                cny, coor, n2e, n2es, id_dof, luni, lmat, li, lv, lti, ltv, llmat, gmat, slip_dist, &
                                obs_points, obs_unitvec, obs_sigma, sigma2_full, alpha2_full, target_id_val, &
                                node_id_in_patch, xinode, etanode, uxinode, uetanode, r1vec, r2vec, nvec, &
                                response_dist, uobs, uret, sp, spn, np, max_slip, dvec, sl, spr, sw, smean, scov, &
                                sln, sprn, assigned, id_start, strl, sml, gsvec, lsvec, scur, scand, sst, acc, range, &
                                npath, nsar_index, nsar_total, sigma_sar_mat, gmat_arr, xmin, xmax, zmin, fix_xbend, &
                                xbend, .false., o_q, o_qn, qdummy, qcdummy, 0d0)
        o_x = o_xn
        o_l = o_ln
    end do
    write (22) o_x, o_l, evidence
    close (22)
    ! fault_smc_exec through the shim (per-stage dumps into fdir)
    call smcf_fortran_seed(seed)
    call smcf_fortran_outer_stages(nstage)
    call fault_smc_exec(trim(fdir), range, nplane, nouter, ndim_fault, 0, 1, nxi_ls, neta_ls, nnode, ndof, &
                        ndof_index, ngnss, nobs, cny, coor, n2e, n2es, id_dof, luni, lmat, li, lv, lti, ltv, llmat, &
                        gmat, slip_dist, obs_points, obs_unitvec, obs_sigma, sigma2_full, alpha2_full, target_id_val, &
                        node_id_in_patch, xinode, etanode, uxinode, uetanode, r1vec, r2vec, nvec, response_dist, uobs, &
                        uret, sp, spn, np, max_slip, dvec, gsvec, lsvec, sl, spr, sw, smean, scov, sln, sprn, sst, &
                        scur, scand, assigned, id_start, strl, sml, npath, nsar_index, nsar_total, sigma_sar_mat, &
                        gmat_arr, xmin, xmax, zmin, fix_xbend, xbend, .false., qdummy, qcdummy, qcdummy, 0d0)
end program drop_in
