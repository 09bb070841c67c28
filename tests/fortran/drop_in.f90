! Drop-in check: a Fortran caller written against the reference's interfaces
! (gfunc / smc_fault module procedures, main.f90:349-390 workspace layout) linked
! against the libsmcf shim instead of the reference objects.
!
! usage: drop_in <in.bin> <out.bin>
!   in : nplane fix nxi(nplane) neta(nplane) xmin xmax zmin [xbend] theta(ndim_fault)
!        nobs nsar_total ngnss npath nsar_index(npath+1) obs_points(2,nobs) obs_unitvec(3,nobs)
!        obs_sigma(nobs) dvec(nobs) nparticle_slip max_slip seed
!   out: gmat(nobs, 2*ndof)  neglog
program drop_in
    use iso_c_binding
    use type_mat
    use gfunc
    use smc_fault
    use smcf_c, only: smcf_fortran_seed, smcf_fortran_mover
    implicit none
    character(len=512) :: fin, fout
    integer :: nplane, ifix, ndim_fault, nnode, ndof, npatch, ip, nobs, nsar_total, ngnss, npath, np, k
    integer(c_int64_t) :: seed
    integer, allocatable :: nxi_ls(:), neta_ls(:), ndof_index(:), nsar_index(:), cny(:, :), n2e(:, :), n2es(:), &
                            id_dof(:), target_id_val(:), node_id_in_patch(:), li(:, :), lti(:, :), &
                            assigned(:), id_start(:)
    double precision :: xmin, xmax, zmin, max_slip, neglog
    double precision, allocatable :: xbend(:), theta(:), obs_points(:, :), obs_unitvec(:, :), obs_sigma(:), dvec(:), &
                                     gmat(:, :), slip_dist(:, :), coor(:, :), response_dist(:, :), xinode(:), &
                                     etanode(:), uxinode(:), uetanode(:), r1vec(:), r2vec(:), nvec(:), uobs(:), uret(:), &
                                     luni(:, :), lmat(:, :), lv(:, :), ltv(:, :), llmat(:, :), sigma2_full(:), &
                                     alpha2_full(:), sp(:, :), spn(:, :), sl(:), spr(:), sw(:), smean(:), scov(:, :), &
                                     sln(:), sprn(:), strl(:, :), sml(:), gsvec(:), lsvec(:), scur(:), scand(:), sst(:)
    type(mat), allocatable :: gmat_arr(:), sigma_sar_mat(:)
    logical :: fix_xbend
    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open (21, file=trim(fin), access='stream', form='unformatted', status='old')
    read (21) nplane, ifix
    fix_xbend = ifix /= 0
    allocate (nxi_ls(nplane), neta_ls(nplane), ndof_index(nplane + 1), xbend(max(nplane - 1, 1)))
    read (21) nxi_ls, neta_ls
    read (21) xmin, xmax, zmin
    xbend = 0d0
    if (nplane > 1) read (21) xbend(1:nplane - 1)
    ndim_fault = merge(2*nplane + 4, 3*nplane + 3, fix_xbend)
    allocate (theta(ndim_fault))
    read (21) theta
    read (21) nobs, nsar_total, ngnss, npath
    allocate (nsar_index(npath + 1), obs_points(2, nobs), obs_unitvec(3, nobs), obs_sigma(nobs), dvec(nobs))
    read (21) nsar_index, obs_points, obs_unitvec, obs_sigma, dvec
    read (21) np, max_slip, seed
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
    neglog = 0d0
    if (nsar_total == 0) then
        call smcf_fortran_seed(seed)
        neglog = fault_calc_likelihood(theta, nplane, nxi_ls, neta_ls, nnode, ndof, ndof_index, ngnss, nobs, cny, &
                                       coor, n2e, n2es, id_dof, luni, lmat, li, lv, lti, ltv, llmat, gmat, slip_dist, &
                                       obs_points, obs_unitvec, obs_sigma, sigma2_full, alpha2_full, target_id_val, &
                                       node_id_in_patch, xinode, etanode, uxinode, uetanode, r1vec, r2vec, nvec, &
                                       response_dist, uobs, uret, sp, spn, np, max_slip, dvec, sl, spr, sw, smean, scov, &
                                       sln, sprn, assigned, id_start, strl, sml, gsvec, lsvec, scur, scand, sst, 0, "", &
                                       npath, nsar_index, nsar_total, sigma_sar_mat, gmat_arr, xmin, xmax, zmin, &
                                       fix_xbend, xbend)
    end if
    open (22, file=trim(fout), access='stream', form='unformatted', status='replace')
    write (22) gmat, neglog
    close (22)
end program drop_in
