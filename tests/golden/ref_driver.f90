! Golden-vector driver: calls the reference's own module procedures (compiled from
! the reference sources by oracle/Makefile target `ref`) on inputs read from a
! binary file and writes their outputs.  Run only by tests/golden/make_golden.py
! in a container that holds the reference tree; the committed fixtures are the
! product of that run.
!
! usage: ref_driver <mode> <in.bin> <out.bin>
!   modes: dc3d0 mesh greens like gamma moments resample init
program ref_driver
    use init
    use gfunc
    use smc_slip
    use type_mat
    implicit none
    character(len=32) :: mode
    character(len=512) :: fin, fout
    call get_command_argument(1, mode)
    call get_command_argument(2, fin)
    call get_command_argument(3, fout)
    open (21, file=trim(fin), access='stream', form='unformatted', status='old')
    open (22, file=trim(fout), access='stream', form='unformatted', status='replace')
    select case (trim(mode))
    case ('dc3d0')
        call run_dc3d0()
    case ('mesh')
        call run_mesh(.false.)
    case ('greens')
        call run_mesh(.true.)
    case ('like')
        call run_like()
    case ('gamma')
        call run_gamma()
    case ('moments')
        call run_moments()
    case ('resample')
        call run_resample()
    case ('init')
        call run_init()
    case default
        stop 2
    end select
    close (21)
    close (22)
contains
    subroutine set_seed(s)
        integer, intent(in) :: s
        integer :: n
        integer, allocatable :: seed(:)
        call random_seed(size=n)
        allocate (seed(n))
        seed = s
        call random_seed(put=seed)
    end subroutine set_seed

    subroutine run_dc3d0()
        integer :: n, i, iret
        double precision :: a(10), u(12)
        read (21) n
        do i = 1, n
            read (21) a
            iret = -99
            u = 0d0
            call dc3d0(a(1), a(2), a(3), a(4), a(5), a(6), a(7), a(8), a(9), a(10), &
                       u(1), u(2), u(3), u(4), u(5), u(6), u(7), u(8), u(9), u(10), u(11), u(12), iret)
            write (22) u, iret
        end do
    end subroutine run_dc3d0

    subroutine run_mesh(do_greens)
        logical, intent(in) :: do_greens
        integer :: nplane, ifix, ndim_fault, nnode, ndof, npatch, ip, nobs, nsar_total, ngnss, npath
        integer, allocatable :: nxi_ls(:), neta_ls(:), ndof_index(:), cny(:, :), n2e(:, :), n2es(:), &
                                id_dof(:), li(:, :), lti(:, :), nsar_index(:), target_id_val(:), &
                                node_id_in_patch(:)
        double precision :: xmin, xmax, zmin
        double precision, allocatable :: xbend(:), theta(:), coor(:, :), luni(:, :), lmat(:, :), &
                                         lv(:, :), ltv(:, :), obs_points(:, :), obs_unitvec(:, :), &
                                         gmat(:, :), slip_dist(:, :), response_dist(:, :), &
                                         xinode(:), etanode(:), uxinode(:), uetanode(:), r1vec(:), &
                                         r2vec(:), nvec(:), uobs(:), uret(:)
        type(mat), allocatable :: gmat_arr(:)
        logical :: fix_xbend
        read (21) nplane, ifix
        fix_xbend = (ifix /= 0)
        allocate (nxi_ls(nplane), neta_ls(nplane), ndof_index(nplane + 1), xbend(max(nplane - 1, 1)))
        read (21) nxi_ls, neta_ls
        read (21) xmin, xmax, zmin
        xbend = 0d0
        if (nplane > 1) read (21) xbend(1:nplane - 1)
        if (fix_xbend) then
            ndim_fault = 2*nplane + 4
        else
            ndim_fault = 3*nplane + 3
        end if
        allocate (theta(ndim_fault))
        read (21) theta
        nnode = 0
        ndof = 0
        npatch = 0
        ndof_index(1) = 1
        do ip = 1, nplane
            nnode = nnode + (nxi_ls(ip) + 1)*(neta_ls(ip) + 1)
            npatch = npatch + nxi_ls(ip)*neta_ls(ip)
            ndof_index(ip + 1) = (nxi_ls(ip) + 1)*(neta_ls(ip) + 1) - (nxi_ls(ip) + 1)
            if (ip == 1) ndof_index(ip + 1) = ndof_index(ip + 1) - neta_ls(ip)
            if (ip == nplane) ndof_index(ip + 1) = ndof_index(ip + 1) - neta_ls(ip)
            ndof = ndof + ndof_index(ip + 1)
            ndof_index(ip + 1) = ndof_index(ip) + ndof_index(ip + 1)
        end do
        allocate (cny(4, npatch), coor(2, nnode), n2e(4, nnode), n2es(nnode), id_dof(ndof))
        call discretize_fault(theta, nplane, nxi_ls, neta_ls, cny, coor, n2e, n2es, id_dof, &
                              xmin, xmax, zmin, fix_xbend, xbend)
        if (.not. do_greens) then
            allocate (luni(nnode, nnode), lmat(2*nnode, 2*ndof), li(9, 2*nnode), lv(9, 2*nnode), &
                      lti(9, 2*ndof), ltv(9, 2*ndof))
            call gen_laplacian(theta, nplane, nnode, nxi_ls, neta_ls, id_dof, ndof, luni, lmat, &
                               xmin, xmax, zmin, fix_xbend, xbend)
            call gen_sparse_lmat(lmat, li, lv, lti, ltv, nnode, ndof)
            write (22) nnode, ndof, npatch
            write (22) coor, cny, n2e, n2es, id_dof
            write (22) luni, lmat, li, lv, lti, ltv
        else
            read (21) nobs, nsar_total, ngnss, npath
            allocate (nsar_index(npath + 1), obs_points(2, nobs), obs_unitvec(3, nobs))
            read (21) nsar_index, obs_points, obs_unitvec
            allocate (gmat(nobs, 2*ndof), slip_dist(2, nnode), response_dist(3, nobs), &
                      target_id_val(4), node_id_in_patch(4), xinode(4), etanode(4), uxinode(4), &
                      uetanode(4), r1vec(2), r2vec(2), nvec(4), uobs(3), uret(3), gmat_arr(npath))
            do ip = 1, npath
                allocate (gmat_arr(ip)%body(max(nsar_index(ip + 1) - nsar_index(ip), 1), 2*ndof))
                gmat_arr(ip)%nrow = nsar_index(ip + 1) - nsar_index(ip)
                gmat_arr(ip)%ncol = 2*ndof
            end do
            call calc_greens_func(theta, nplane, nxi_ls, neta_ls, gmat, slip_dist, cny, coor, &
                                  obs_points, obs_unitvec, n2e, n2es, id_dof, ngnss, nobs, nnode, &
                                  ndof, ndof_index, target_id_val, node_id_in_patch, xinode, etanode, &
                                  uxinode, uetanode, r1vec, r2vec, nvec, response_dist, uobs, uret, &
                                  nsar_total, npath, nsar_index, gmat_arr, xmin, xmax, zmin, &
                                  fix_xbend, xbend)
            write (22) nobs, ndof
            write (22) gmat
        end if
    end subroutine run_mesh

    subroutine run_like()
        integer :: ndim, nnode, np, nsar, ngnss, k
        double precision :: dsd, lss2, lsg2
        double precision, allocatable :: h(:, :), b(:), s(:, :), lv(:, :), a2(:), lsvec(:), &
                                         dummy(:), dummy2(:, :), gsvec(:), out(:, :), theta(:)
        integer, allocatable :: li(:, :)
        read (21) ndim, nnode, np, nsar, ngnss
        read (21) dsd, lss2, lsg2
        allocate (h(ndim, ndim), b(ndim), s(ndim, np), li(9, 2*nnode), lv(9, 2*nnode), a2(2*nnode), &
                  lsvec(2*nnode), dummy(1), dummy2(1, 1), gsvec(1), out(2, np), theta(6))
        read (21) h, b, s, li, lv, a2
        theta = 0d0
        do k = 1, np
            out(1, k) = slip_calc_likelihood(s(:, k), dummy, dummy, dummy2, lss2, lsg2, ngnss, gsvec, &
                                             1, ndim/2, h, b, dsd, nsar)
            out(2, k) = slip_calc_prior(s(:, k), a2, theta, 1, li, lv, lsvec, nnode)
        end do
        write (22) out
    end subroutine run_like

    subroutine run_gamma()
        integer :: np
        double precision :: gprev, gamma, ev
        double precision, allocatable :: l(:), w(:)
        read (21) np, gprev
        allocate (l(np), w(np))
        read (21) l
        gamma = slip_find_next_gamma(gprev, l, w, ev, np)
        write (22) gamma, ev, w
    end subroutine run_gamma

    subroutine run_moments()
        integer :: np, nd
        double precision, allocatable :: x(:, :), w(:), mean(:), cov(:, :), cd(:)
        read (21) np, nd
        allocate (x(nd, np), w(np), mean(nd), cov(1, 1), cd(nd))
        read (21) w, x
        call slip_normalize_weights(w, np)
        call slip_calc_mean_particles(x, w, mean, np, nd)
        call slip_calc_cov_particles(x, w, mean, cov, np, nd, cd)
        write (22) w, mean, cd
    end subroutine run_moments

    subroutine run_resample()
        integer :: np, seed
        double precision, allocatable :: w(:)
        integer, allocatable :: a(:)
        real(16) :: q
        double precision :: hi, lo
        read (21) np, seed
        allocate (w(np), a(np))
        read (21) w
        call set_seed(seed)
        call slip_resample_particles(np, w, a)
        call set_seed(seed)
        call random_number(q)
        hi = real(q, 8)
        lo = real(q - real(hi, 16), 8)
        write (22) hi, lo, a
    end subroutine run_resample

    subroutine run_init()
        integer :: nd, np, nnode, seed, k, i
        double precision :: max_slip, x
        double precision, allocatable :: ll(:, :), a2(:), parts(:, :), lik(:), pri(:), cur(:), &
                                         dummy(:), dummy2(:, :), gsvec(:), lsvec(:), lv(:, :), &
                                         theta(:), h(:, :), b(:), u(:)
        integer, allocatable :: li(:, :)
        read (21) nd, np, nnode, seed
        read (21) max_slip
        allocate (ll(nd, nd), a2(2*nnode), parts(nd, np), lik(np), pri(np), cur(nd), dummy(1), &
                  dummy2(1, 1), gsvec(1), lsvec(2*nnode), lv(9, 2*nnode), li(9, 2*nnode), theta(6), &
                  h(nd, nd), b(nd), u(2*nd*np))
        read (21) ll, a2, li, lv
        h = 0d0
        b = 0d0
        theta = 0d0
        call set_seed(seed)
        call slip_gen_init_particles(parts, lik, pri, np, nd, ll, max_slip, dummy, dummy, a2, theta, &
                                     1, dummy2, 0d0, 0d0, 0, 1, nd/2, li, lv, nnode, cur, gsvec, lsvec, &
                                     b, 0d0, 0, h)
        call set_seed(seed)
        do k = 1, np
            do i = 1, 2*nd
                call random_number(x)
                u((k - 1)*2*nd + i) = x
            end do
        end do
        write (22) u, parts, pri
    end subroutine run_init
end program ref_driver
