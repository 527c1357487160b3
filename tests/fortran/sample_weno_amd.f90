! A Fortran caller of the weno2d sweeps through mod_btdma_amd, in the order of the reference's
! SDIRK stage (Fortran_Original/examples/weno2d/weno_mpi_gpu.f90:1026-1070: x sweep, then y sweep
! on its solution) on a 24 x 18 periodic grid of a rotating scalar, each sweep
! btdma_weno_gather_amd -> btdma_weno_assemble_amd -> btdma_many_cycl_amd -> btdma_weno_scatter_amd.
! Checked on the host from the assembled blocks: every block row of I - gamma dt L sums to one
! (constants are preserved) and the periodic residual of every system is at rounding level; and
! the same two sweeps matrix-free (btdma_weno_sweep_cyclic_amd) give the same field bit for bit.
program sample_weno_amd
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
    integer, parameter :: n1 = 24, n2 = 18, m = 3, nh = 3
    integer(c_int), parameter :: H2D = 1, D2H = 2
    real(c_double), parameter :: pi = 3.14159265358979323846d0, omega = 2 * pi, gdt = 0.01d0
    real(c_double), allocatable, target :: U(:, :), F(:, :), R(:, :), G(:, :), vel(:)
    real(c_double), allocatable, target :: A(:, :, :, :), B(:, :, :, :), C(:, :, :, :), D(:, :, :), X(:, :, :)
    type(c_ptr) :: dU, dF, dG, pv, pA, pB, pC, pD, pW, plan
    integer(c_size_t) :: ubytes, fbytes, mbytes, vbytes
    integer :: i, j, axis, nsys, nrow, s, q, qm, qp, failures
    real(c_double) :: xc, yc, h, res, rmax, smax, dmax

    allocate (U(1 - nh:n1 + nh, 1 - nh:n2 + nh), F(n1, n2), R(n1, n2), G(n1, n2))
    do j = 1, n2
        do i = 1, n1
            xc = (i - 0.5d0) / n1; yc = (j - 0.5d0) / n2
            U(i, j) = merge(1.0d0, 0.0d0, (xc - 0.5d0)**2 + (yc - 0.75d0)**2 < 0.15d0**2)  ! a disc
            F(i, j) = U(i, j) + 0.1d0 * sin(2 * pi * xc) * cos(2 * pi * yc)
        end do
    end do
    ! periodic halo (one rank's halo exchange)
    U(1 - nh:0, 1:n2) = U(n1 - nh + 1:n1, 1:n2); U(n1 + 1:n1 + nh, 1:n2) = U(1:nh, 1:n2)
    U(:, 1 - nh:0) = U(:, n2 - nh + 1:n2); U(:, n2 + 1:n2 + nh) = U(:, 1:nh)

    ubytes = int(8, c_size_t) * size(U)
    fbytes = int(8, c_size_t) * size(F)
    if (hipMalloc(dU, ubytes) /= 0 .or. hipMalloc(dF, fbytes) /= 0 .or. hipMalloc(dG, fbytes) /= 0) &
        error stop 'hipMalloc'
    if (hipMemcpy(dU, c_loc(U), ubytes, H2D) /= 0 .or. hipMemcpy(dF, c_loc(F), fbytes, H2D) /= 0 .or. &
        hipMemcpy(dG, c_loc(F), fbytes, H2D) /= 0) error stop 'hipMemcpy'

    failures = 0
    rmax = 0
    smax = 0
    do axis = 0, 1
        if (axis == 0) then
            nsys = n2; nrow = n1 / m; h = 1.0d0 / n1
            allocate (vel(nsys))
            do j = 1, n2
                vel(j) = -omega * ((j - 0.5d0) / n2 - 0.5d0)  ! c_x(j) (weno_mpi_gpu.f90:438)
            end do
        else
            nsys = n1; nrow = n2 / m; h = 1.0d0 / n2
            allocate (vel(nsys))
            do i = 1, n1
                vel(i) = omega * ((i - 0.5d0) / n1 - 0.5d0)   ! c_y(i) (:467)
            end do
        end if
        mbytes = int(8, c_size_t) * nsys * nrow * m * m
        vbytes = int(8, c_size_t) * nsys * nrow * m
        allocate (A(nsys, nrow, m, m), B(nsys, nrow, m, m), C(nsys, nrow, m, m), D(nsys, nrow, m), X(nsys, nrow, m))
        if (hipMalloc(pA, mbytes) /= 0 .or. hipMalloc(pB, mbytes) /= 0 .or. hipMalloc(pC, mbytes) /= 0 .or. &
            hipMalloc(pD, vbytes) /= 0 .or. hipMalloc(pv, int(8, c_size_t) * nsys) /= 0 .or. &
            hipMalloc(pW, 2 * mbytes) /= 0) error stop 'hipMalloc'
        if (hipMemcpy(pv, c_loc(vel), int(8, c_size_t) * nsys, H2D) /= 0) error stop 'hipMemcpy'
        call btdma_makeplan_amd(plan, m, nsys, nrow, 1)
        call btdma_weno_gather_amd(plan, axis, dF, n1, n2, pD)
        call btdma_weno_assemble_amd(plan, axis, dU, pv, n1, n2, h, gdt, pA, pB, pC)
        if (hipMemcpy(c_loc(A), pA, mbytes, D2H) /= 0 .or. hipMemcpy(c_loc(B), pB, mbytes, D2H) /= 0 .or. &
            hipMemcpy(c_loc(C), pC, mbytes, D2H) /= 0 .or. hipMemcpy(c_loc(D), pD, vbytes, D2H) /= 0) &
            error stop 'hipMemcpy'
        call btdma_many_cycl_amd(pA, pB, pC, pD, m, nsys, nrow, plan)
        if (hipMemcpy(c_loc(X), pD, vbytes, D2H) /= 0) error stop 'hipMemcpy'
        call btdma_weno_scatter_amd(plan, axis, pD, n1, n2, dF)
        ! the same sweep matrix-free, on its own copy of the field
        call btdma_weno_gather_amd(plan, axis, dG, n1, n2, pD)
        call btdma_weno_sweep_cyclic_amd(plan, axis, dU, pv, n1, n2, h, gdt, pW, pD)
        call btdma_weno_scatter_amd(plan, axis, pD, n1, n2, dG)
        dmax = maxval(abs(D))
        do s = 1, nsys
            do q = 1, nrow
                qm = merge(nrow, q - 1, q == 1); qp = merge(1, q + 1, q == nrow)
                do i = 1, m
                    smax = max(smax, abs(sum(A(s, q, i, :)) + sum(B(s, q, i, :)) + sum(C(s, q, i, :)) - 1))
                    res = -D(s, q, i)
                    do j = 1, m
                        res = res + A(s, q, i, j) * X(s, qm, j) + B(s, q, i, j) * X(s, q, j) + C(s, q, i, j) * X(s, qp, j)
                    end do
                    rmax = max(rmax, abs(res) / dmax)
                end do
            end do
        end do
        call btdma_cleanplan_amd(plan)
        if (hipFree(pA) /= 0 .or. hipFree(pB) /= 0 .or. hipFree(pC) /= 0 .or. hipFree(pD) /= 0 .or. &
            hipFree(pv) /= 0 .or. hipFree(pW) /= 0) error stop 'hipFree'
        deallocate (A, B, C, D, X, vel)
    end do
    if (hipMemcpy(c_loc(R), dF, fbytes, D2H) /= 0 .or. hipMemcpy(c_loc(G), dG, fbytes, D2H) /= 0) &
        error stop 'hipMemcpy'
    write (*, '(a,es10.3,a,es10.3,a,es23.16,a,es10.3)') 'max periodic residual ', rmax, '  max |row sum - 1| ', &
        smax, '  sum(F) ', sum(R), '  matrix-free vs assembled ', maxval(abs(G - R))
    if (any(G /= R)) failures = failures + 1
    if (rmax > 1d-12) failures = failures + 1
    if (smax > 1d-13) failures = failures + 1
    if (abs(sum(R) - sum(F)) > 1d-9 * sum(abs(F))) failures = failures + 1  ! conservative sweeps keep the mass
    if (.not. (maxval(abs(R - F)) > 1d-6)) failures = failures + 1         ! the sweeps changed F
    if (hipFree(dU) /= 0 .or. hipFree(dF) /= 0 .or. hipFree(dG) /= 0) error stop 'hipFree'
    if (failures /= 0) then
        write (*, '(a)') 'FAILED'
        error stop 1
    end if
    write (*, '(a)') 'PASSED'
end program sample_weno_amd
