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
        integer(c_int) function btd_euler_assemble(plan, axis, U, rhs, n1, n2, n3, h, dt, eps, a, b, c, d, &
                                                   stream) bind(C)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: plan, U, rhs, a, b, c, d, stream
            integer(c_int), value :: axis, n1, n2, n3
            real(c_double), value :: h, dt, eps
        end function
        integer(c_int) function btd_sweep_scatter(plan, axis, d, n1, n2, n3, field, stream) bind(C)
            import :: c_int, c_ptr
            type(c_ptr), value :: plan, d, field, stream
            integer(c_int), value :: axis, n1, n2, n3
        end function
        integer(c_int) function btd_euler_sweep_cyclic(plan, axis, U, rhs, n1, n2, n3, h, dt, eps, work, out, &
                                                       stream) bind(C)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: plan, U, rhs, work, out, stream
            integer(c_int), value :: axis, n1, n2, n3
            real(c_double), value :: h, dt, eps
        end function
        integer(c_int) function btd_weno_assemble(plan, axis, U, vel, n1, n2, h, gdt, a, b, c, stream) bind(C)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: plan, U, vel, a, b, c, stream
            integer(c_int), value :: axis, n1, n2
            real(c_double), value :: h, gdt
        end function
        integer(c_int) function btd_weno_sweep_cyclic(plan, axis, U, vel, n1, n2, h, gdt, work, d, stream) &
                bind(C)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: plan, U, vel, work, d, stream
            integer(c_int), value :: axis, n1, n2
            real(c_double), value :: h, gdt
        end function
        integer(c_int) function btd_weno_gather(plan, axis, field, n1, n2, d, stream) bind(C)
            import :: c_int, c_ptr
            type(c_ptr), value :: plan, field, d, stream
            integer(c_int), value :: axis, n1, n2
        end function
        integer(c_int) function btd_weno_scatter(plan, axis, d, n1, n2, field, stream) bind(C)
            import :: c_int, c_ptr
            type(c_ptr), value :: plan, d, field, stream
            integer(c_int), value :: axis, n1, n2
        end function
    end interface

    public :: btdma_makeplan_amd, btdma_many_amd, btdma_many_amd_profiled
    public :: btdma_many_cycl_amd, btdma_cleanplan_amd
    public :: btdma_euler_assemble_amd, btdma_sweep_scatter_amd, btdma_euler_sweep_cyclic_amd
    public :: btdma_weno_assemble_amd, btdma_weno_gather_amd, btdma_weno_scatter_amd
    public :: btdma_weno_sweep_cyclic_amd

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

    ! euler3d sweeps (axis 0 = x, 1 = y, 2 = z; the plan is the sweep's: nsys = the product of the other
    ! two grid sizes, nrow = the axis' size). U(0:n1+1,0:n2+1,0:n3+1,5) with its halo filled and
    ! dU(n1,n2,n3,5) are device fields in the reference's order.
    ! build_matrix_{x,y,z}_kernel plus the transposes (euler3d_mpi_gpu.f90:803-1002): the sweep's
    ! A, B, C and D = dU reordered, in the plan's layout
    subroutine btdma_euler_assemble_amd(plan, axis, U, dU, n1, n2, n3, h, dt, eps, A, B, C, D)
        type(c_ptr), intent(in) :: plan, U, dU, A, B, C, D
        integer, intent(in) :: axis, n1, n2, n3
        real(c_double), intent(in) :: h, dt, eps
        call check(btd_euler_assemble(plan, axis, U, dU, n1, n2, n3, h, dt, eps, A, B, C, D, c_null_ptr), &
                   'btdma_euler_assemble_amd')
    end subroutine

    ! transpose_kj_kernel / transpose_ki_kernel (euler3d_mpi_gpu.f90:959-974, 1004-1017): the sweep's
    ! solution D back into the field dU
    subroutine btdma_sweep_scatter_amd(plan, axis, D, n1, n2, n3, dU)
        type(c_ptr), intent(in) :: plan, D, dU
        integer, intent(in) :: axis, n1, n2, n3
        call check(btd_sweep_scatter(plan, axis, D, n1, n2, n3, dU, c_null_ptr), 'btdma_sweep_scatter_amd')
    end subroutine

    ! the whole periodic sweep without storing A, B, C; the solution replaces dU when out = dU.
    ! work: device scratch of 2*nm + nv + nu doubles (nm = nsys*nrow*25, nv = nsys*nrow*5,
    ! nu = nsys*(nrow+2)*5 in the reference layout)
    subroutine btdma_euler_sweep_cyclic_amd(plan, axis, U, dU, n1, n2, n3, h, dt, eps, work, out)
        type(c_ptr), intent(in) :: plan, U, dU, work, out
        integer, intent(in) :: axis, n1, n2, n3
        real(c_double), intent(in) :: h, dt, eps
        call check(btd_euler_sweep_cyclic(plan, axis, U, dU, n1, n2, n3, h, dt, eps, work, out, c_null_ptr), &
                   'btdma_euler_sweep_cyclic_amd')
    end subroutine

    ! compute_weno_coeffs_{x,y}_kernel + build_matrix_{x,y}_kernel_weno5 (weno_mpi_gpu.f90:430-719):
    ! the weno2d sweep's m = 3 blocks A, B, C in the plan's layout from U(1-3:n1+3, 1-3:n2+3) and
    ! each line's speed vel (c_x(j) for axis 0, c_y(i) for axis 1); gdt = gamma*dt
    subroutine btdma_weno_assemble_amd(plan, axis, U, vel, n1, n2, h, gdt, A, B, C)
        type(c_ptr), intent(in) :: plan, U, vel, A, B, C
        integer, intent(in) :: axis, n1, n2
        real(c_double), intent(in) :: h, gdt
        call check(btd_weno_assemble(plan, axis, U, vel, n1, n2, h, gdt, A, B, C, c_null_ptr), &
                   'btdma_weno_assemble_amd')
    end subroutine

    ! the weno2d sweep's periodic solve with the blocks computed inside it (never stored): D, the
    ! gathered right-hand side, becomes the solution; work: 2*nsys*nrow*9 device doubles
    subroutine btdma_weno_sweep_cyclic_amd(plan, axis, U, vel, n1, n2, h, gdt, work, D)
        type(c_ptr), intent(in) :: plan, U, vel, work, D
        integer, intent(in) :: axis, n1, n2
        real(c_double), intent(in) :: h, gdt
        call check(btd_weno_sweep_cyclic(plan, axis, U, vel, n1, n2, h, gdt, work, D, c_null_ptr), &
                   'btdma_weno_sweep_cyclic_amd')
    end subroutine

    ! pack_scalar_to_xblock_kernel / the transpose before the y sweep (weno_mpi_gpu.f90:279-289,
    ! 332-343): the field F(n1, n2) as the sweep's vector D
    subroutine btdma_weno_gather_amd(plan, axis, F, n1, n2, D)
        type(c_ptr), intent(in) :: plan, F, D
        integer, intent(in) :: axis, n1, n2
        call check(btd_weno_gather(plan, axis, F, n1, n2, D, c_null_ptr), 'btdma_weno_gather_amd')
    end subroutine

    ! unpack_xblock_to_scalar_kernel / the transpose after the y sweep (:291-301, 345-356)
    subroutine btdma_weno_scatter_amd(plan, axis, D, n1, n2, F)
        type(c_ptr), intent(in) :: plan, D, F
        integer, intent(in) :: axis, n1, n2
        call check(btd_weno_scatter(plan, axis, D, n1, n2, F, c_null_ptr), 'btdma_weno_scatter_amd')
    end subroutine

    ! btdma_cleanplan_gpu_v2(plan) (mod_btdma_gpu_v2.f90:111)
    subroutine btdma_cleanplan_amd(plan)
        type(c_ptr), intent(inout) :: plan
        call check(btd_plan_destroy(plan), 'btdma_cleanplan_amd')
        plan = c_null_ptr
    end subroutine

end module mod_btdma_amd
