! A Fortran caller in the style of the reference's samples (Fortran_Original/examples),
! using mod_btdma_amd: the Study's diagonal workload (A = 0.25 I, B = 2 I, C = 0.25 I, D = 1),
! 32 x 32 systems of 128 block rows, m = 5, arrays in the reference layout
! A(nsys,nrow,m,m), solved with 1 and 4 row partitions and checked against the reference's
! published signature (Study/result/btdma_solution_signature_260702_193138.csv: sum
! 262605.6260190311, x at the first row 0.4508066615170332, at the middle row 0.4).
program sample_btdma_amd
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
    integer, parameter :: n1 = 32, n2 = 32, n3 = 128, m = 5, nsys = n1 * n2, nrow = n3
    integer(c_int), parameter :: H2D = 1, D2H = 2
    real(c_double), allocatable, target :: A(:, :, :, :), B(:, :, :, :), C(:, :, :, :), D(:, :, :), X(:, :, :)
    type(c_ptr) :: dA, dB, dC, dD, plan
    type(btdma_timing_amd) :: timing
    integer(c_size_t) :: mbytes, vbytes
    integer :: p, npart, ipart, failures
    real(c_double) :: s
    integer, parameter :: parts(2) = [1, 4]

    allocate (A(nsys, nrow, m, m), B(nsys, nrow, m, m), C(nsys, nrow, m, m), D(nsys, nrow, m), X(nsys, nrow, m))
    mbytes = int(8, c_size_t) * nsys * nrow * m * m
    vbytes = int(8, c_size_t) * nsys * nrow * m
    if (hipMalloc(dA, mbytes) /= 0 .or. hipMalloc(dB, mbytes) /= 0 .or. hipMalloc(dC, mbytes) /= 0 &
        .or. hipMalloc(dD, vbytes) /= 0) error stop 'hipMalloc'
    failures = 0
    do ipart = 1, 2
        npart = parts(ipart)
        A = 0; B = 0; C = 0; D = 1
        do p = 1, m
            A(:, :, p, p) = 0.25d0
            B(:, :, p, p) = 2.00d0
            C(:, :, p, p) = 0.25d0
        end do
        if (hipMemcpy(dA, c_loc(A), mbytes, H2D) /= 0 .or. hipMemcpy(dB, c_loc(B), mbytes, H2D) /= 0 .or. &
            hipMemcpy(dC, c_loc(C), mbytes, H2D) /= 0 .or. hipMemcpy(dD, c_loc(D), vbytes, H2D) /= 0) &
            error stop 'hipMemcpy'
        call btdma_makeplan_amd(plan, m, nsys, nrow, npart)
        call btdma_many_amd_profiled(dA, dB, dC, dD, m, nsys, nrow, plan, timing)
        call btdma_cleanplan_amd(plan)
        if (hipMemcpy(c_loc(X), dD, vbytes, D2H) /= 0) error stop 'hipMemcpy'
        s = sum(X)
        write (*, '(a,i0,a,es23.16,a,es23.16,a,es23.16,a,es10.3)') 'npart=', npart, ' sum=', s, &
            ' x(1,1,1)=', X(1, 1, 1), ' x(1,65,1)=', X(1, n3 / 2 + 1, 1), ' t=', timing%total
        if (abs(s / 262605.6260190311d0 - 1) > 1d-9) failures = failures + 1
        if (abs(X(1, 1, 1) - 0.4508066615170332d0) > 1d-14) failures = failures + 1
        if (abs(X(1, n3 / 2 + 1, 1) - 0.4d0) > 1d-14) failures = failures + 1
    end do
    if (hipFree(dA) /= 0 .or. hipFree(dB) /= 0 .or. hipFree(dC) /= 0 .or. hipFree(dD) /= 0) error stop 'hipFree'
    if (failures /= 0) then
        write (*, '(a)') 'FAILED'
        error stop 1
    end if
    write (*, '(a)') 'PASSED'
end program sample_btdma_amd
