! Call-compatible Fortran binding of libpascal_btdma_amd: a module of the reference's name
! (Fortran_Original/src/mod_btdma_gpu_v2.f90) with its public types and procedures, so a caller
! written against the reference keeps its `use mod_btdma_gpu_v2` line and its call sites, e.g.
! heat3d (Fortran_Original/examples/heat3d/heat3d_mpi_gpu.f90:825-827, 877, 915, 953, 1025-1027):
!     call btdma_makeplan_gpu_v2(plan_x, mb, nsys_x, nblk_x, npart_x)
!     call btdma_many_mpi_gpu_v2(Ax_d, Bx_d, Cx_d, Dx_d, mb, nsys_x, nblk_x, plan_x)
!     call btdma_cleanplan_gpu_v2(plan_x)
! The one change: the last argument of btdma_makeplan_gpu_v2 is the number of row partitions of
! every system on this GPU (npart; 1 = the single-partition solve, 0 = the library's choice)
! where the reference takes the MPI communicator whose ranks hold those partitions.
!
! Arrays: A, B, C (nsys, nrow_sub, m, m) and D (nsys, nrow_sub, m), in the reference's order, in
! device memory. Without CUDA Fortran's `device` attribute a caller holds device arrays as Fortran
! pointers associated with hipMalloc'd memory (c_f_pointer); they are passed to these procedures
! by address only and never read on the host. Every procedure also accepts the arrays as the
! type(c_ptr) device addresses themselves. The solve is in place, like the reference: x returns
! in D, A and C are overwritten, B is read. Calls run on the default stream. Errors (which the
! reference does not detect) stop the program with the library's message.
module mod_btdma_gpu_v2
    use iso_c_binding
    implicit none
    private

    integer(c_int), parameter :: BTD_OK = 0
    integer(c_int), parameter :: BTD_LAYOUT_REF_GPU = 0

    ! BTDMA_PLAN_gpu_v2 (mod_btdma_gpu_v2.f90:28-36): the plan's scratch lives in the library; the
    ! type holds its handle. nsys_sub is the number of systems whose reduced rows this GPU solves
    ! (all of them: there is one GPU).
    type, public :: BTDMA_PLAN_gpu_v2
        integer :: nsys_sub = 0
        type(c_ptr) :: handle = c_null_ptr
    end type BTDMA_PLAN_gpu_v2

    ! BTDMA_TIMING_gpu_v2 (mod_btdma_gpu_v2.f90:38-45): the six phase times in seconds; on one GPU
    ! the two exchange phases read 0.
    type, bind(C), public :: BTDMA_TIMING_gpu_v2
        real(c_double) :: total = 0.d0
        real(c_double) :: local_compute = 0.d0
        real(c_double) :: forward_exchange = 0.d0
        real(c_double) :: reduced_compute = 0.d0
        real(c_double) :: backward_exchange = 0.d0
        real(c_double) :: update_compute = 0.d0
    end type BTDMA_TIMING_gpu_v2

    interface
        integer(c_int) function btd_plan_create(plan, m, nsys, nrow, npart, layout) bind(C)
            import :: c_int, c_ptr
            type(c_ptr) :: plan
            integer(c_int), value :: m, nsys, nrow, npart, layout
        end function
        integer(c_int) function btd_plan_destroy(plan) bind(C)
            import :: c_int, c_ptr
            type(c_ptr), value :: plan
        end function
        integer(c_int) function btd_solve_noncyclic(plan, a, b, c, d, m, nsys, nrow, stream) bind(C)
            import :: c_int, c_ptr
            type(c_ptr), value :: plan, a, b, c, d, stream
            integer(c_int), value :: m, nsys, nrow
        end function
        integer(c_int) function btd_solve_noncyclic_profiled(plan, a, b, c, d, m, nsys, nrow, t, stream) bind(C)
            import :: c_int, c_ptr, BTDMA_TIMING_gpu_v2
            type(c_ptr), value :: plan, a, b, c, d, stream
            integer(c_int), value :: m, nsys, nrow
            type(BTDMA_TIMING_gpu_v2) :: t
        end function
        integer(c_int) function btd_solve_cyclic(plan, a, b, c, d, m, nsys, nrow, stream) bind(C)
            import :: c_int, c_ptr
            type(c_ptr), value :: plan, a, b, c, d, stream
            integer(c_int), value :: m, nsys, nrow
        end function
        integer(c_int) function btd_plan_npart(plan, npart) bind(C)
            import :: c_int, c_ptr
            type(c_ptr), value :: plan
            integer(c_int) :: npart
        end function
        type(c_ptr) function btd_last_error() bind(C)
            import :: c_ptr
        end function
        integer(c_size_t) function strlen(s) bind(C, name='strlen')
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function
    end interface

    ! mod_btdma_gpu_v2.f90:131 / :169 / :200: device arrays, or their addresses
    interface btdma_many_mpi_gpu_v2
        module procedure many_arrays, many_ptrs
    end interface
    interface btdma_many_mpi_gpu_v2_profiled
        module procedure many_profiled_arrays, many_profiled_ptrs
    end interface
    interface btdma_many_cycl_mpi_gpu_v2
        module procedure cycl_arrays, cycl_ptrs
    end interface

    public :: btdma_makeplan_gpu_v2, btdma_cleanplan_gpu_v2, btdma_timing_reset_gpu_v2
    public :: btdma_many_mpi_gpu_v2, btdma_many_mpi_gpu_v2_profiled, btdma_many_cycl_mpi_gpu_v2

contains

    subroutine check(st, what)
        integer(c_int), intent(in) :: st
        character(*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        type(c_ptr) :: p
        integer :: n
        if (st == BTD_OK) return
        p = btd_last_error()
        n = 0
        if (c_associated(p)) n = int(strlen(p))
        if (n > 0) then
            call c_f_pointer(p, msg, [n])
            write (*, '(a,a,i0,a,*(a))') what, ': pascal_btdma_amd status ', st, ': ', msg
        else
            write (*, '(a,a,i0)') what, ': pascal_btdma_amd status ', st
        end if
        error stop 1
    end subroutine

    ! btdma_timing_reset_gpu_v2 (mod_btdma_gpu_v2.f90:58-68)
    subroutine btdma_timing_reset_gpu_v2(timing)
        type(BTDMA_TIMING_gpu_v2), intent(out) :: timing
        timing = BTDMA_TIMING_gpu_v2()
    end subroutine

    ! btdma_makeplan_gpu_v2(plan, m, nsys, nrow_sub, comm) (mod_btdma_gpu_v2.f90:70-109); npart
    ! stands where the communicator stood
    subroutine btdma_makeplan_gpu_v2(plan, m, nsys, nrow_sub, npart)
        type(BTDMA_PLAN_gpu_v2) :: plan
        integer, intent(in) :: m, nsys, nrow_sub, npart
        call check(btd_plan_create(plan%handle, int(m, c_int), int(nsys, c_int), int(nrow_sub, c_int), &
                                   int(npart, c_int), BTD_LAYOUT_REF_GPU), 'btdma_makeplan_gpu_v2')
        plan%nsys_sub = nsys
    end subroutine

    ! btdma_cleanplan_gpu_v2(plan) (mod_btdma_gpu_v2.f90:111-129)
    subroutine btdma_cleanplan_gpu_v2(plan)
        type(BTDMA_PLAN_gpu_v2) :: plan
        call check(btd_plan_destroy(plan%handle), 'btdma_cleanplan_gpu_v2')
        plan%handle = c_null_ptr
        plan%nsys_sub = 0
    end subroutine

    subroutine many_ptrs(A, B, C, D, m, nsys, nrow_sub, plan)
        type(c_ptr), intent(in) :: A, B, C, D
        integer, intent(in) :: m, nsys, nrow_sub
        type(BTDMA_PLAN_gpu_v2) :: plan
        call check(btd_solve_noncyclic(plan%handle, A, B, C, D, int(m, c_int), int(nsys, c_int), &
                                       int(nrow_sub, c_int), c_null_ptr), 'btdma_many_mpi_gpu_v2')
    end subroutine

    subroutine many_arrays(A, B, C, D, m, nsys, nrow_sub, plan)
        integer, intent(in) :: m, nsys, nrow_sub
        real(c_double), target :: A(nsys, nrow_sub, m, m), B(nsys, nrow_sub, m, m)
        real(c_double), target :: C(nsys, nrow_sub, m, m), D(nsys, nrow_sub, m)
        type(BTDMA_PLAN_gpu_v2) :: plan
        call many_ptrs(c_loc(A), c_loc(B), c_loc(C), c_loc(D), m, nsys, nrow_sub, plan)
    end subroutine

    subroutine many_profiled_ptrs(A, B, C, D, m, nsys, nrow_sub, plan, timing)
        type(c_ptr), intent(in) :: A, B, C, D
        integer, intent(in) :: m, nsys, nrow_sub
        type(BTDMA_PLAN_gpu_v2) :: plan
        type(BTDMA_TIMING_gpu_v2), intent(out) :: timing
        call check(btd_solve_noncyclic_profiled(plan%handle, A, B, C, D, int(m, c_int), int(nsys, c_int), &
                                                int(nrow_sub, c_int), timing, c_null_ptr), &
                   'btdma_many_mpi_gpu_v2_profiled')
    end subroutine

    subroutine many_profiled_arrays(A, B, C, D, m, nsys, nrow_sub, plan, timing)
        integer, intent(in) :: m, nsys, nrow_sub
        real(c_double), target :: A(nsys, nrow_sub, m, m), B(nsys, nrow_sub, m, m)
        real(c_double), target :: C(nsys, nrow_sub, m, m), D(nsys, nrow_sub, m)
        type(BTDMA_PLAN_gpu_v2) :: plan
        type(BTDMA_TIMING_gpu_v2), intent(out) :: timing
        call many_profiled_ptrs(c_loc(A), c_loc(B), c_loc(C), c_loc(D), m, nsys, nrow_sub, plan, timing)
    end subroutine

    subroutine cycl_ptrs(A, B, C, D, m, nsys, nrow_sub, plan)
        type(c_ptr), intent(in) :: A, B, C, D
        integer, intent(in) :: m, nsys, nrow_sub
        type(BTDMA_PLAN_gpu_v2) :: plan
        call check(btd_solve_cyclic(plan%handle, A, B, C, D, int(m, c_int), int(nsys, c_int), &
                                    int(nrow_sub, c_int), c_null_ptr), 'btdma_many_cycl_mpi_gpu_v2')
    end subroutine

    subroutine cycl_arrays(A, B, C, D, m, nsys, nrow_sub, plan)
        integer, intent(in) :: m, nsys, nrow_sub
        real(c_double), target :: A(nsys, nrow_sub, m, m), B(nsys, nrow_sub, m, m)
        real(c_double), target :: C(nsys, nrow_sub, m, m), D(nsys, nrow_sub, m)
        type(BTDMA_PLAN_gpu_v2) :: plan
        call cycl_ptrs(c_loc(A), c_loc(B), c_loc(C), c_loc(D), m, nsys, nrow_sub, plan)
    end subroutine

end module mod_btdma_gpu_v2
