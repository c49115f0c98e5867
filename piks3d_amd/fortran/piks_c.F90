! piks_c -- ISO_C_BINDING interface to libpiks3d.so (include/piks3d.h).
!
! The derived types are bind(C), so their layout is the C structs' layout;
! arrays are passed with c_loc and stay in Fortran (column-major) order, which
! is the order the library returns cell arrays and f in.
module piks_c
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: PIKS_PREC_FP64 = 0, PIKS_PREC_FP32_REFINED = 1

    type, bind(C) :: piks_params
        integer(c_int32_t) :: nx, ny, nz
        integer(c_int32_t) :: wall_ext_order
        integer(c_int32_t) :: fluid_layer
        real(c_double)     :: ref_l
        integer(c_int32_t) :: nc_fundamental
        integer(c_int32_t) :: half_range
        integer(c_int32_t) :: chk_converge_step
        real(c_double)     :: kn
        real(c_double)     :: press_drop
        real(c_double)     :: accom
        integer(c_int32_t) :: precision
        integer(c_int32_t) :: device
        integer(c_int32_t) :: exact_quadrature
        integer(c_int32_t) :: exact_order
        integer(c_int32_t) :: sweep, tile
        integer(c_int32_t) :: reserved(2)
    end type piks_params

    type, bind(C) :: piks_info
        integer(c_int32_t) :: nx, ny, nz
        integer(c_int32_t) :: nxt, nyt, nzt
        integer(c_int64_t) :: ntotal
        integer(c_int32_t) :: nc8
        integer(c_int32_t) :: nwall, ncorner
        integer(c_int64_t) :: nfluid
        integer(c_int64_t) :: updates_per_iter
        real(c_double)     :: ds, mu, diff_flux
        real(c_double)     :: mass
        integer(c_int64_t) :: device_bytes
        integer(c_int32_t) :: coupled_walls
        integer(c_int32_t) :: precision
        integer(c_int64_t) :: steps_fp32
        integer(c_int32_t) :: dc_block
        integer(c_int32_t) :: sweep_path
    end type piks_info

    interface
        integer(c_int) function piks_create(p, raw_image, image_bytes, ctx) bind(C, name='piks_create')
            import :: c_int, c_int64_t, c_ptr, piks_params
            type(piks_params), intent(in) :: p
            type(c_ptr), value :: raw_image
            integer(c_int64_t), value :: image_bytes
            type(c_ptr), intent(out) :: ctx
        end function piks_create

        integer(c_int) function piks_iterate(ctx, nsteps) bind(C, name='piks_iterate')
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nsteps
        end function piks_iterate

        integer(c_int) function piks_set_precision(ctx, precision) bind(C, name='piks_set_precision')
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: precision
        end function piks_set_precision
        integer(c_int) function piks_set_correction_block(ctx, m) bind(C, name='piks_set_correction_block')
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: m
        end function piks_set_correction_block

        integer(c_int) function piks_chk_converge(ctx, out) bind(C, name='piks_chk_converge')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: out(3)
        end function piks_chk_converge

        integer(c_int) function piks_get_macro(ctx, rho, ux, uy, uz) bind(C, name='piks_get_macro')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, rho, ux, uy, uz
        end function piks_get_macro

        integer(c_int) function piks_create_multi(p, raw_image, image_bytes, nrhs, kn, ctx) &
                bind(C, name='piks_create_multi')
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr, piks_params
            type(piks_params), intent(in) :: p
            type(c_ptr), value :: raw_image
            integer(c_int64_t), value :: image_bytes
            integer(c_int32_t), value :: nrhs
            real(c_double), intent(in) :: kn(*)
            type(c_ptr), intent(out) :: ctx
        end function piks_create_multi

        integer(c_int) function piks_chk_converge_multi(ctx, out) bind(C, name='piks_chk_converge_multi')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: out(3, *)
        end function piks_chk_converge_multi

        integer(c_int) function piks_get_macro_rhs(ctx, rhs, rho, ux, uy, uz) bind(C, name='piks_get_macro_rhs')
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: rhs
            type(c_ptr), value :: rho, ux, uy, uz
        end function piks_get_macro_rhs

        integer(c_int) function piks_get_labels(ctx, labels) bind(C, name='piks_get_labels')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, labels
        end function piks_get_labels

        integer(c_int) function piks_get_info(ctx, info) bind(C, name='piks_get_info')
            import :: c_int, c_ptr, piks_info
            type(c_ptr), value :: ctx
            type(piks_info), intent(out) :: info
        end function piks_get_info

        integer(c_int) function piks_destroy(ctx) bind(C, name='piks_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function piks_destroy

        type(c_ptr) function piks_last_error() bind(C, name='piks_last_error')
            import :: c_ptr
        end function piks_last_error
    end interface

contains

    ! The last error message of the library as a Fortran string.
    function piks_error_message() result(msg)
        character(len=512) :: msg
        character(kind=c_char), pointer :: chars(:)
        type(c_ptr) :: p
        integer :: i
        msg = ' '
        p = piks_last_error()
        if (.not. c_associated(p)) return
        call c_f_pointer(p, chars, [512])
        do i = 1, 512
            if (chars(i) == c_null_char) exit
            msg(i:i) = chars(i)
        enddo
    end function piks_error_message

end module piks_c
