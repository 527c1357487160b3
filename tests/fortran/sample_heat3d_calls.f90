! A caller written against the reference's CUDA Fortran module, compiled against this library's
! call-compatible mod_btdma_gpu_v2: the module line and the plan, solve and clean-up calls are
! those of heat3d (Fortran_Original/examples/heat3d/heat3d_mpi_gpu.f90:700, 714, 825-827, 877,
! 915, 953, 1025-1027); only the makeplan's last argument is a partition count instead of a
! communicator. Three sweeps (x, y, z) of a 40 x 24 x 32 grid with m = 5 blocks, in the reference
! layout A(nsys, nblk, mb, mb). The blocks are diagonal with component- and system-dependent
! coefficients, so each solution component is a scalar tridiagonal system that the host solves
! independently (Thomas' algorithm) for the check; the profiled and periodic entry points and
! the timing reset are exercised too.
program sample_heat3d_calls
    use iso_c_binding
    use mod_btdma_gpu_v2
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
    integer(c_int), parameter :: H2D = 1, D2H = 2
    integer, parameter :: nx = 40, ny = 24, nz = 32, mb = 5
    integer, parameter :: nsys_x = ny * nz, nblk_x = nx, nsys_y = nx * nz, nblk_y = ny
    integer, parameter :: nsys_z = nx * ny, nblk_z = nz
    integer, parameter :: npart_x = 1, npart_y = 2, npart_z = 4
    type(BTDMA_PLAN_gpu_v2) :: plan_x, plan_y, plan_z
    type(BTDMA_TIMING_gpu_v2) :: timing
    real(c_double), pointer :: Ax_d(:, :, :, :), Bx_d(:, :, :, :), Cx_d(:, :, :, :), Dx_d(:, :, :)
    real(c_double), pointer :: Ay_d(:, :, :, :), By_d(:, :, :, :), Cy_d(:, :, :, :), Dy_d(:, :, :)
    real(c_double), pointer :: Az_d(:, :, :, :), Bz_d(:, :, :, :), Cz_d(:, :, :, :), Dz_d(:, :, :)
    real(c_double) :: err, worst
    integer :: failures

    failures = 0
    worst = 0

    call device_arrays(nsys_x, nblk_x, Ax_d, Bx_d, Cx_d, Dx_d)
    call device_arrays(nsys_y, nblk_y, Ay_d, By_d, Cy_d, Dy_d)
    call device_arrays(nsys_z, nblk_z, Az_d, Bz_d, Cz_d, Dz_d)

    !--- BTDMA plans (heat3d_mpi_gpu.f90:825-827, the communicator replaced by a partition count) ---
    call btdma_makeplan_gpu_v2(plan_x, mb, nsys_x, nblk_x, npart_x)
    call btdma_makeplan_gpu_v2(plan_y, mb, nsys_y, nblk_y, npart_y)
    call btdma_makeplan_gpu_v2(plan_z, mb, nsys_z, nblk_z, npart_z)

    ! x sweep (heat3d_mpi_gpu.f90:877)
    call fill(nsys_x, nblk_x, 1, Ax_d, Bx_d, Cx_d, Dx_d, .false.)
    call btdma_many_mpi_gpu_v2(Ax_d, Bx_d, Cx_d, Dx_d, mb, nsys_x, nblk_x, plan_x)
    call verify(nsys_x, nblk_x, 1, Dx_d, .false., 'x sweep', err)
    worst = max(worst, err)
    ! y sweep (:915)
    call fill(nsys_y, nblk_y, 2, Ay_d, By_d, Cy_d, Dy_d, .false.)
    call btdma_many_mpi_gpu_v2(Ay_d, By_d, Cy_d, Dy_d, mb, nsys_y, nblk_y, plan_y)
    call verify(nsys_y, nblk_y, 2, Dy_d, .false., 'y sweep', err)
    worst = max(worst, err)
    ! z sweep (:953)
    call fill(nsys_z, nblk_z, 3, Az_d, Bz_d, Cz_d, Dz_d, .false.)
    call btdma_many_mpi_gpu_v2(Az_d, Bz_d, Cz_d, Dz_d, mb, nsys_z, nblk_z, plan_z)
    call verify(nsys_z, nblk_z, 3, Dz_d, .false., 'z sweep', err)
    worst = max(worst, err)

    ! the profiled form (mod_btdma_gpu_v2.f90:169) and the timing reset (:58)
    call fill(nsys_z, nblk_z, 4, Az_d, Bz_d, Cz_d, Dz_d, .false.)
    call btdma_many_mpi_gpu_v2_profiled(Az_d, Bz_d, Cz_d, Dz_d, mb, nsys_z, nblk_z, plan_z, timing)
    call verify(nsys_z, nblk_z, 4, Dz_d, .false., 'z sweep, profiled', err)
    worst = max(worst, err)
    write (*, '(a,4es11.3)') 'timing total/local/reduced/update [s]: ', timing%total, timing%local_compute, &
        timing%reduced_compute, timing%update_compute
    if (.not. (timing%total > 0 .and. timing%local_compute > 0 .and. timing%reduced_compute > 0 .and. &
               timing%forward_exchange == 0 .and. timing%backward_exchange == 0)) then
        write (*, '(a)') 'unexpected phase times'
        failures = failures + 1
    end if
    call btdma_timing_reset_gpu_v2(timing)
    if (timing%total /= 0 .or. timing%local_compute /= 0) failures = failures + 1

    ! the periodic form (mod_btdma_gpu_v2.f90:200) on the y and x plans
    call fill(nsys_y, nblk_y, 5, Ay_d, By_d, Cy_d, Dy_d, .true.)
    call btdma_many_cycl_mpi_gpu_v2(Ay_d, By_d, Cy_d, Dy_d, mb, nsys_y, nblk_y, plan_y)
    call verify(nsys_y, nblk_y, 5, Dy_d, .true., 'y sweep, periodic', err)
    worst = max(worst, err)
    call fill(nsys_x, nblk_x, 6, Ax_d, Bx_d, Cx_d, Dx_d, .true.)
    call btdma_many_cycl_mpi_gpu_v2(Ax_d, Bx_d, Cx_d, Dx_d, mb, nsys_x, nblk_x, plan_x)
    call verify(nsys_x, nblk_x, 6, Dx_d, .true., 'x sweep, periodic', err)
    worst = max(worst, err)

    !--- clean-up (heat3d_mpi_gpu.f90:1025-1027) ---
    call btdma_cleanplan_gpu_v2(plan_x)
    call btdma_cleanplan_gpu_v2(plan_y)
    call btdma_cleanplan_gpu_v2(plan_z)

    write (*, '(a,es10.3)') 'max rel err vs host scalar solves: ', worst
    if (worst > 1.d-12) failures = failures + 1
    if (failures == 0) then
        write (*, '(a)') 'PASSED'
    else
        write (*, '(a,i0)') 'FAILED: ', failures
        error stop 1
    end if

contains

    ! device arrays in the reference's order, as Fortran pointers to hipMalloc'd memory
    subroutine device_arrays(nsys, nblk, A, B, C, D)
        integer, intent(in) :: nsys, nblk
        real(c_double), pointer :: A(:, :, :, :), B(:, :, :, :), C(:, :, :, :), D(:, :, :)
        type(c_ptr) :: p
        integer(c_size_t) :: mbytes, vbytes
        mbytes = int(8, c_size_t) * nsys * nblk * mb * mb
        vbytes = int(8, c_size_t) * nsys * nblk * mb
        if (hipMalloc(p, mbytes) /= 0) error stop 'hipMalloc'
        call c_f_pointer(p, A, [nsys, nblk, mb, mb])
        if (hipMalloc(p, mbytes) /= 0) error stop 'hipMalloc'
        call c_f_pointer(p, B, [nsys, nblk, mb, mb])
        if (hipMalloc(p, mbytes) /= 0) error stop 'hipMalloc'
        call c_f_pointer(p, C, [nsys, nblk, mb, mb])
        if (hipMalloc(p, vbytes) /= 0) error stop 'hipMalloc'
        call c_f_pointer(p, D, [nsys, nblk, mb])
    end subroutine

    ! coefficients of component p of system s, row q (diagonal blocks): a, b, c, rhs
    subroutine coef(seed, s, q, p, a, b, c, r)
        integer, intent(in) :: seed, s, q, p
        real(c_double), intent(out) :: a, b, c, r
        a = -0.25d0 - 0.01d0 * p - 0.001d0 * mod(s, 7)
        c = -0.20d0 - 0.02d0 * p + 0.001d0 * mod(s, 5)
        b = 1.d0 + 0.1d0 * p + 0.01d0 * mod(s + seed, 11)
        r = sin(0.1d0 * q + 0.3d0 * p + 0.01d0 * s + seed)
    end subroutine

    subroutine fill(nsys, nblk, seed, Ad, Bd, Cd, Dd, periodic)
        integer, intent(in) :: nsys, nblk, seed
        real(c_double), pointer :: Ad(:, :, :, :), Bd(:, :, :, :), Cd(:, :, :, :), Dd(:, :, :)
        logical, intent(in) :: periodic
        real(c_double), allocatable, target :: A(:, :, :, :), B(:, :, :, :), C(:, :, :, :), D(:, :, :)
        real(c_double) :: a_, b_, c_, r_
        integer :: s, q, p
        integer(c_size_t) :: mbytes, vbytes
        allocate (A(nsys, nblk, mb, mb), B(nsys, nblk, mb, mb), C(nsys, nblk, mb, mb), D(nsys, nblk, mb))
        A = 0; B = 0; C = 0
        do p = 1, mb
            do q = 1, nblk
                do s = 1, nsys
                    call coef(seed, s, q, p, a_, b_, c_, r_)
                    if (q > 1 .or. periodic) A(s, q, p, p) = a_
                    B(s, q, p, p) = b_
                    if (q < nblk .or. periodic) C(s, q, p, p) = c_
                    D(s, q, p) = r_
                end do
            end do
        end do
        mbytes = int(8, c_size_t) * nsys * nblk * mb * mb
        vbytes = int(8, c_size_t) * nsys * nblk * mb
        if (hipMemcpy(c_loc(Ad), c_loc(A), mbytes, H2D) /= 0 .or. hipMemcpy(c_loc(Bd), c_loc(B), mbytes, H2D) /= 0 &
            .or. hipMemcpy(c_loc(Cd), c_loc(C), mbytes, H2D) /= 0 .or. hipMemcpy(c_loc(Dd), c_loc(D), vbytes, H2D) /= 0) &
            error stop 'hipMemcpy'
    end subroutine

    ! x from the device against the host's scalar solve of every (system, component): Thomas'
    ! algorithm, and for the periodic systems the Sherman-Morrison correction of it
    subroutine verify(nsys, nblk, seed, Dd, periodic, what, err)
        integer, intent(in) :: nsys, nblk, seed
        real(c_double), pointer :: Dd(:, :, :)
        logical, intent(in) :: periodic
        character(*), intent(in) :: what
        real(c_double), intent(out) :: err
        real(c_double), allocatable, target :: X(:, :, :)
        real(c_double), allocatable :: a(:), b(:), c(:), r(:), y(:), z(:), u(:)
        real(c_double) :: gam, fac, num, den
        integer :: s, p, q
        integer(c_size_t) :: vbytes
        allocate (X(nsys, nblk, mb), a(nblk), b(nblk), c(nblk), r(nblk), y(nblk), z(nblk), u(nblk))
        vbytes = int(8, c_size_t) * nsys * nblk * mb
        if (hipMemcpy(c_loc(X), c_loc(Dd), vbytes, D2H) /= 0) error stop 'hipMemcpy'
        num = 0
        den = 0
        do s = 1, nsys
            do p = 1, mb
                do q = 1, nblk
                    call coef(seed, s, q, p, a(q), b(q), c(q), r(q))
                end do
                if (.not. periodic) then
                    call thomas(nblk, a, b, c, r, y)
                else
                    ! T = T0 + u v^T with T0 tridiagonal: u = (gam, 0.., c_n), v = (1, 0.., a_1/gam)
                    gam = -b(1)
                    b(1) = b(1) - gam
                    b(nblk) = b(nblk) - a(1) * c(nblk) / gam
                    call thomas(nblk, a, b, c, r, y)
                    u = 0
                    u(1) = gam
                    u(nblk) = c(nblk)
                    call thomas(nblk, a, b, c, u, z)
                    fac = (y(1) + a(1) * y(nblk) / gam) / (1.d0 + z(1) + a(1) * z(nblk) / gam)
                    y = y - fac * z
                end if
                num = max(num, maxval(abs(X(s, :, p) - y)))
                den = max(den, maxval(abs(y)))
            end do
        end do
        err = num / den
        write (*, '(a,a,es10.3)') what, ': max rel err ', err
    end subroutine

    subroutine thomas(n, a, b, c, r, y)
        integer, intent(in) :: n
        real(c_double), intent(in) :: a(n), b(n), c(n), r(n)
        real(c_double), intent(out) :: y(n)
        real(c_double) :: cp(n), dp(n), w
        integer :: i
        cp(1) = c(1) / b(1)
        dp(1) = r(1) / b(1)
        do i = 2, n
            w = b(i) - a(i) * cp(i - 1)
            cp(i) = c(i) / w
            dp(i) = (r(i) - a(i) * dp(i - 1)) / w
        end do
        y(n) = dp(n)
        do i = n - 1, 1, -1
            y(i) = dp(i) - cp(i) * y(i + 1)
        end do
    end subroutine

end program sample_heat3d_calls
