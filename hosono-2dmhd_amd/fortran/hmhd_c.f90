! hmhd_c.f90 -- iso_c_binding interface to libhmhd.so (include/hmhd.h).
!
! A Fortran driver uses this module in place of the reference's calls of
! cfl_computedt (src/mhd_driver.f90:66,78) and timeStep_advanceSolution
! (src/mhd_driver.f90:81); the state arrays keep the reference's layout
! U(8, minIdx:maxIdx, minIdx:maxIdx), V(10, ...), scheme(...) and are passed
! without copies (contiguous, column-major).
module hmhd_c
    use, intrinsic :: iso_c_binding
    implicit none
    private

    integer(c_int32_t), parameter, public :: HMHD_OK = 0, HMHD_ERR_ARG = 1, HMHD_ERR_DEVICE = 2, &
                                             HMHD_ERR_GP = 3, HMHD_ERR_TIMEOUT = 4
    integer(c_int32_t), parameter, public :: HMHD_BC_PERIODIC = 0, HMHD_BC_OUTFLOW = 1
    integer(c_int32_t), parameter, public :: HMHD_RIEMANN_HLLC = 1

    ! mirrors struct hmhd_config field by field
    type, bind(C), public :: hmhd_config
        integer(c_int32_t) :: nx, ny
        real(c_double) :: xbeg, xend
        real(c_double) :: ybeg, yend
        integer(c_int32_t) :: R
        real(c_double) :: ell_over_delta
        real(c_double) :: gamma
        real(c_double) :: small_p, small_rho
        real(c_double) :: small_B
        integer(c_int32_t) :: force_hydro
        integer(c_int32_t) :: bc
        real(c_double) :: cfl
        integer(c_int32_t) :: riemann
    end type hmhd_config

    public :: hmhd_create, hmhd_destroy, hmhd_last_error, hmhd_dims, hmhd_upload_state, hmhd_cfl_dt, &
              hmhd_advance, hmhd_download, hmhd_stats, hmhd_get_gp, hmhd_set_gp, hmhd_synchronize, &
              hmhd_error_message

    interface
        integer(c_int32_t) function hmhd_create(cfg, h) bind(C, name="hmhd_create")
            import :: c_int32_t, c_ptr, hmhd_config
            type(hmhd_config), intent(in) :: cfg
            type(c_ptr), intent(out) :: h
        end function hmhd_create

        subroutine hmhd_destroy(h) bind(C, name="hmhd_destroy")
            import :: c_ptr
            type(c_ptr), value :: h
        end subroutine hmhd_destroy

        type(c_ptr) function hmhd_last_error(h) bind(C, name="hmhd_last_error")
            import :: c_ptr
            type(c_ptr), value :: h
        end function hmhd_last_error

        integer(c_int32_t) function hmhd_dims(h, ngc, mx, my) bind(C, name="hmhd_dims")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), intent(out) :: ngc, mx, my
        end function hmhd_dims

        ! U(8,mx,my) required; V(10,mx,my) may be c_null_ptr
        integer(c_int32_t) function hmhd_upload_state(h, U, V) bind(C, name="hmhd_upload_state")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h, U, V
        end function hmhd_upload_state

        integer(c_int32_t) function hmhd_cfl_dt(h, dt) bind(C, name="hmhd_cfl_dt")
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: dt
        end function hmhd_cfl_dt

        integer(c_int32_t) function hmhd_advance(h, dt) bind(C, name="hmhd_advance")
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), value :: dt
        end function hmhd_advance

        ! any pointer may be c_null_ptr
        integer(c_int32_t) function hmhd_download(h, U, V, scheme) bind(C, name="hmhd_download")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h, U, V, scheme
        end function hmhd_download

        integer(c_int32_t) function hmhd_stats(h, passes, inner) bind(C, name="hmhd_stats")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), intent(out) :: passes(3), inner(3)
        end function hmhd_stats

        ! predvec(nsMax,nqMax,4,R), quadw(nqMax,R), stenc(2,nsMax,R) as in src/GP.f90:12-21
        integer(c_int32_t) function hmhd_get_gp(h, predvec, quadw, stenc) bind(C, name="hmhd_get_gp")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h, predvec, quadw, stenc
        end function hmhd_get_gp

        integer(c_int32_t) function hmhd_set_gp(h, predvec, quadw) bind(C, name="hmhd_set_gp")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h, predvec, quadw
        end function hmhd_set_gp

        integer(c_int32_t) function hmhd_synchronize(h) bind(C, name="hmhd_synchronize")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
        end function hmhd_synchronize

        integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function c_strlen
    end interface

contains

    ! the message of the last failed call on h (h = c_null_ptr: of the
    ! last failed hmhd_create) as a Fortran string
    function hmhd_error_message(h) result(msg)
        type(c_ptr), intent(in) :: h
        character(len=:), allocatable :: msg
        type(c_ptr) :: p
        character(kind=c_char), pointer :: chars(:)
        integer :: n, k
        p = hmhd_last_error(h)
        n = int(c_strlen(p))
        call c_f_pointer(p, chars, [n])
        allocate(character(len=n) :: msg)
        do k = 1, n
            msg(k:k) = chars(k)
        end do
    end function hmhd_error_message

end module hmhd_c
