! Fortran binding of libpascal_btdma_amd (ISO_C_BINDING), the module a caller of the
! reference's CUDA Fortran solver (Fortran_Original/src/mod_btdma_gpu_v2.f90) uses instead.
! The procedures keep the reference's names with an _amd suffix, its argument order and
! meaning, and its array layout A(nsys,nrow_sub,m,m) / D(nsys,nrow_sub,m), which is
! BTD_LAYOUT_REF_GPU. Device arrays are passed by their C address (c_ptr from hipMalloc);
! the communicator argument becomes the number of row partitions npart of one GPU.
module mod_btdma_amd
    use iso_c_binding
    implicit none
    private

    integer(c_int), parameter, public :: BTD_OK = 0
    integer(c_int), parameter, public :: BTD_LAYOUT_REF_GPU = 0

    ! BTDMA_TIMING_gpu_v2 (mod_btdma_gpu_v2.f90:38-45), seconds
    type, bind(C), public :: btdma_timing_amd
        real(c_double) :: total = 0, local_compute = 0, forward_exchange = 0
        real(c_double) :: reduced_compute = 0, backward_exchange = 0, update_compute = 0
    end type

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
            import :: c_int, c_ptr, btdma_timing_amd
            type(c_ptr), value :: plan, a, b, c, d, stream
            integer(c_int), value :: m, nsys, nrow
            type(btdma_timing_amd) :: t
        end function
        integer(c_int) function btd_solve_cyclic(plan, a, b, c, d, m, nsys, nrow, stream) bind(C)
            import :: c_int, c_ptr
            type(c_ptr), value :: plan, a, b, c, d, stream
            integer(c_int), value :: m, nsys, nrow
        end function
        type(c_ptr) function btd_last_error() bind(C)
            import :: c_ptr
        end function
    end interface

    public :: btdma_makeplan_amd, btdma_many_amd, btdma_many_amd_profiled
    public :: btdma_many_cycl_amd, btdma_cleanplan_amd

contains

    subroutine check(st, what)
        integer(c_int), intent(in) :: st
        character(*), intent(in) :: what
        if (st /= BTD_OK) then
            write (*, '(a,a,i0)') what, ': pascal_btdma_amd status ', st
            error stop 1
        end if
    end subroutine

    ! btdma_makeplan_gpu_v2(plan, m, nsys, nrow_sub, comm) (mod_btdma_gpu_v2.f90:70)
    subroutine btdma_makeplan_amd(plan, m, nsys, nrow_sub, npart)
        type(c_ptr), intent(out) :: plan
        integer, intent(in) :: m, nsys, nrow_sub, npart
        call check(btd_plan_create(plan, m, nsys, nrow_sub, npart, BTD_LAYOUT_REF_GPU), 'btdma_makeplan_amd')
    end subroutine

    ! btdma_many_mpi_gpu_v2(A, B, C, D, m, nsys, nrow_sub, plan) (mod_btdma_gpu_v2.f90:131)
    subroutine btdma_many_amd(A, B, C, D, m, nsys, nrow_sub, plan)
        type(c_ptr), intent(in) :: A, B, C, D, plan
        integer, intent(in) :: m, nsys, nrow_sub
        call check(btd_solve_noncyclic(plan, A, B, C, D, m, nsys, nrow_sub, c_null_ptr), 'btdma_many_amd')
    end subroutine

    ! btdma_many_mpi_gpu_v2_profiled (mod_btdma_gpu_v2.f90:169)
    subroutine btdma_many_amd_profiled(A, B, C, D, m, nsys, nrow_sub, plan, timing)
        type(c_ptr), intent(in) :: A, B, C, D, plan
        integer, intent(in) :: m, nsys, nrow_sub
        type(btdma_timing_amd), intent(out) :: timing
        call check(btd_solve_noncyclic_profiled(plan, A, B, C, D, m, nsys, nrow_sub, timing, c_null_ptr), &
                   'btdma_many_amd_profiled')
    end subroutine

    ! btdma_many_cycl_mpi_gpu_v2 (mod_btdma_gpu_v2.f90:200)
    subroutine btdma_many_cycl_amd(A, B, C, D, m, nsys, nrow_sub, plan)
        type(c_ptr), intent(in) :: A, B, C, D, plan
        integer, intent(in) :: m, nsys, nrow_sub
        call check(btd_solve_cyclic(plan, A, B, C, D, m, nsys, nrow_sub, c_null_ptr), 'btdma_many_cycl_amd')
    end subroutine

    ! btdma_cleanplan_gpu_v2(plan) (mod_btdma_gpu_v2.f90:111)
    subroutine btdma_cleanplan_amd(plan)
        type(c_ptr), intent(inout) :: plan
        call check(btd_plan_destroy(plan), 'btdma_cleanplan_amd')
        plan = c_null_ptr
    end subroutine

end module mod_btdma_amd
