! A Fortran caller of the euler3d sweeps through mod_btdma_amd, in the order of the reference's
! ADI step (Fortran_Original/examples/euler3d/euler3d_mpi_gpu.f90:1260-1325: z, y, x sweeps, each
! solution the next sweep's right-hand side), on a 20 x 12 x 10 periodic grid:
!  * assembled: btdma_euler_assemble_amd -> btdma_many_cycl_amd -> btdma_sweep_scatter_amd, with the
!    periodic residual of every system checked on the host from the assembled blocks;
!  * matrix-free: btdma_euler_sweep_cyclic_amd in place over dU;
! and the two results compared (relative difference <= 1e-12).
program sample_euler_amd
    use iso_c_binding
    use mod_btdma_amd
    implicit none
    interface
        integer(c_int) function hipMalloc(ptr, bytes) bind(C, name='hipMalloc')
            import :: c_int, c_ptr, c_size_t
            type(c_ptr) :: ptr
            integer(c_size_t), value :: bytes
        end function
        integer(c_int) function hipFree(ptr) bind(C, name='hipFree')
            import :: c_int, c_ptr
            type(c_ptr), value :: ptr
        end function
        integer(c_int) function hipMemcpy(dst, src, bytes, kind) bind(C, name='hipMemcpy')
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: dst, src
            integer(c_size_t), value :: bytes
            integer(c_int), value :: kind
        end function
    end interface
    integer, parameter :: n1 = 20, n2 = 12, n3 = 10, m = 5
    integer(c_int), parameter :: H2D = 1, D2H = 2
    real(c_double), parameter :: dt = 0.02d0, eps = 0.01d0, pi = 3.14159265358979323846d0
    real(c_double), allocatable, target :: U(:, :, :, :), dU(:, :, :, :), Ra(:, :, :, :), Rb(:, :, :, :)
    real(c_double), allocatable, target :: A(:, :, :, :), B(:, :, :, :), C(:, :, :, :), D(:, :, :), X(:, :, :)
    type(c_ptr) :: dUf, dUa, dUb, pA, pB, pC, pD, pW, plan
    integer(c_size_t) :: ubytes, fbytes, mbytes, vbytes, wbytes
    integer :: i, j, k, v, axis, isw, nsys, nrow, s, q, qm, qp, failures
    integer, parameter :: order(3) = [2, 1, 0]
    real(c_double) :: xc, yc, zc, rho, uu, vv, ww, p, h, res, rmax, dmax, diff

    allocate (U(0:n1 + 1, 0:n2 + 1, 0:n3 + 1, m), dU(n1, n2, n3, m), Ra(n1, n2, n3, m), Rb(n1, n2, n3, m))
    do k = 1, n3
        do j = 1, n2
            do i = 1, n1
                xc = (i - 0.5d0) / n1; yc = (j - 0.5d0) / n2; zc = (k - 0.5d0) / n3
                rho = 1 + 0.2d0 * sin(2 * pi * xc) * cos(2 * pi * yc)
                uu = 0.3d0 * sin(2 * pi * yc); vv = 0.3d0 * cos(2 * pi * zc); ww = 0.2d0 * sin(2 * pi * xc)
                p = 1 + 0.1d0 * cos(2 * pi * (xc + yc + zc))
                U(i, j, k, 1) = rho
                U(i, j, k, 2) = rho * uu; U(i, j, k, 3) = rho * vv; U(i, j, k, 4) = rho * ww
                U(i, j, k, 5) = p / 0.4d0 + 0.5d0 * rho * (uu**2 + vv**2 + ww**2)
                do v = 1, m
                    dU(i, j, k, v) = sin(0.7d0 * i + 1.3d0 * j + 0.4d0 * k + v)
                end do
            end do
        end do
    end do
    ! periodic halo (one rank's halo exchange)
    U(0, :, :, :) = U(n1, :, :, :); U(n1 + 1, :, :, :) = U(1, :, :, :)
    U(:, 0, :, :) = U(:, n2, :, :); U(:, n2 + 1, :, :) = U(:, 1, :, :)
    U(:, :, 0, :) = U(:, :, n3, :); U(:, :, n3 + 1, :) = U(:, :, 1, :)

    ubytes = int(8, c_size_t) * size(U)
    fbytes = int(8, c_size_t) * size(dU)
    if (hipMalloc(dUf, ubytes) /= 0 .or. hipMalloc(dUa, fbytes) /= 0 .or. hipMalloc(dUb, fbytes) /= 0) &
        error stop 'hipMalloc'
    if (hipMemcpy(dUf, c_loc(U), ubytes, H2D) /= 0 .or. hipMemcpy(dUa, c_loc(dU), fbytes, H2D) /= 0 .or. &
        hipMemcpy(dUb, c_loc(dU), fbytes, H2D) /= 0) error stop 'hipMemcpy'

    failures = 0
    rmax = 0
    do isw = 1, 3
        axis = order(isw)
        select case (axis)
        case (0)
            nrow = n1; nsys = n2 * n3; h = 1.0d0 / n1
        case (1)
            nrow = n2; nsys = n1 * n3; h = 1.0d0 / n2
        case default
            nrow = n3; nsys = n1 * n2; h = 1.0d0 / n3
        end select
        mbytes = int(8, c_size_t) * nsys * nrow * m * m
        vbytes = int(8, c_size_t) * nsys * nrow * m
        wbytes = 2 * mbytes + vbytes + int(8, c_size_t) * nsys * (nrow + 2) * m
        allocate (A(nsys, nrow, m, m), B(nsys, nrow, m, m), C(nsys, nrow, m, m), D(nsys, nrow, m), X(nsys, nrow, m))
        if (hipMalloc(pA, mbytes) /= 0 .or. hipMalloc(pB, mbytes) /= 0 .or. hipMalloc(pC, mbytes) /= 0 .or. &
            hipMalloc(pD, vbytes) /= 0 .or. hipMalloc(pW, wbytes) /= 0) error stop 'hipMalloc'
        call btdma_makeplan_amd(plan, m, nsys, nrow, 1)
        ! assembled path, with the blocks kept on the host for the residual
        call btdma_euler_assemble_amd(plan, axis, dUf, dUa, n1, n2, n3, h, dt, eps, pA, pB, pC, pD)
        if (hipMemcpy(c_loc(A), pA, mbytes, D2H) /= 0 .or. hipMemcpy(c_loc(B), pB, mbytes, D2H) /= 0 .or. &
            hipMemcpy(c_loc(C), pC, mbytes, D2H) /= 0 .or. hipMemcpy(c_loc(D), pD, vbytes, D2H) /= 0) &
            error stop 'hipMemcpy'
        call btdma_many_cycl_amd(pA, pB, pC, pD, m, nsys, nrow, plan)
        if (hipMemcpy(c_loc(X), pD, vbytes, D2H) /= 0) error stop 'hipMemcpy'
        ! periodic residual A_q x_{q-1} + B_q x_q + C_q x_{q+1} - d_q of every system
        dmax = maxval(abs(D))
        do s = 1, nsys
            do q = 1, nrow
                qm = merge(nrow, q - 1, q == 1); qp = merge(1, q + 1, q == nrow)
                do i = 1, m
                    res = -D(s, q, i)
                    do j = 1, m
                        res = res + A(s, q, i, j) * X(s, qm, j) + B(s, q, i, j) * X(s, q, j) + C(s, q, i, j) * X(s, qp, j)
                    end do
                    rmax = max(rmax, abs(res) / dmax)
                end do
            end do
        end do
        call btdma_sweep_scatter_amd(plan, axis, pD, n1, n2, n3, dUa)
        ! matrix-free path, in place over its own dU
        call btdma_euler_sweep_cyclic_amd(plan, axis, dUf, dUb, n1, n2, n3, h, dt, eps, pW, dUb)
        call btdma_cleanplan_amd(plan)
        if (hipFree(pA) /= 0 .or. hipFree(pB) /= 0 .or. hipFree(pC) /= 0 .or. hipFree(pD) /= 0 .or. &
            hipFree(pW) /= 0) error stop 'hipFree'
        deallocate (A, B, C, D, X)
    end do
    if (hipMemcpy(c_loc(Ra), dUa, fbytes, D2H) /= 0 .or. hipMemcpy(c_loc(Rb), dUb, fbytes, D2H) /= 0) &
        error stop 'hipMemcpy'
    diff = maxval(abs(Ra - Rb)) / maxval(abs(Ra))
    write (*, '(a,es10.3,a,es10.3,a,es23.16)') 'max periodic residual ', rmax, '  assembled vs matrix-free ', &
        diff, '  sum(dU) ', sum(Ra)
    if (rmax > 1d-12) failures = failures + 1
    if (diff > 1d-12) failures = failures + 1
    if (.not. (maxval(abs(Ra - dU)) > 1d-6)) failures = failures + 1  ! the sweeps changed dU
    if (hipFree(dUf) /= 0 .or. hipFree(dUa) /= 0 .or. hipFree(dUb) /= 0) error stop 'hipFree'
    if (failures /= 0) then
        write (*, '(a)') 'FAILED'
        error stop 1
    end if
    write (*, '(a)') 'PASSED'
end program sample_euler_amd
