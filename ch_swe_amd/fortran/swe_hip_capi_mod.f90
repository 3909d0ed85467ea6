! iso_c_binding interface to libchswe.so (include/ch_swe.h).
module swe_hip_capi_mod

    use iso_c_binding, only : c_int, c_double, c_float, c_ptr, c_char, c_null_char, c_associated, c_f_pointer

    implicit none

    private

    integer(c_int), parameter, public :: SWE_FP64 = 0, SWE_FP32 = 1

    ! swe_config (include/ch_swe.h)
    type, bind(C), public :: swe_config_t
        integer(c_int) :: N
        integer(c_int) :: order
        integer(c_int) :: precision
        integer(c_int) :: coriolis_const
        real(c_double) :: radius
        real(c_double) :: grav
        real(c_double) :: h_mean
        real(c_double) :: omega
        real(c_double) :: axis(3)
        integer(c_int) :: flags
        integer(c_int) :: block_threads
        integer(c_int) :: tile_x0
        integer(c_int) :: tile_y0
    end type swe_config_t

    public :: swe_plan_create, swe_plan_destroy, swe_state_create, swe_state_destroy, &
              swe_state_upload, swe_state_download, swe_state_upload_halo, swe_state_download_halo, &
              swe_apply, swe_assign, swe_update,  &
              swe_rk4_steps, swe_integrals, swe_integrals_diag, swe_integral_tends, swe_latlon_h, swe_latlon_uv, swe_cfl_diag, &
              swe_solve, swe_implicit_euler_steps, swe_crank_nicolson_steps, &
              swe_synchronize, swe_error_message

    interface
        integer(c_int) function swe_plan_create(cfg, plan) bind(C, name="swe_plan_create")
            import :: c_int, c_ptr, swe_config_t
            type(swe_config_t), intent(in) :: cfg
            type(c_ptr), intent(out) :: plan
        end function
        integer(c_int) function swe_plan_destroy(plan) bind(C, name="swe_plan_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: plan
        end function
        integer(c_int) function swe_synchronize(plan) bind(C, name="swe_synchronize")
            import :: c_int, c_ptr
            type(c_ptr), value :: plan
        end function
        integer(c_int) function swe_state_create(plan, st) bind(C, name="swe_state_create")
            import :: c_int, c_ptr
            type(c_ptr), value :: plan
            type(c_ptr), intent(out) :: st
        end function
        integer(c_int) function swe_state_destroy(st) bind(C, name="swe_state_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: st
        end function
        integer(c_int) function swe_state_upload(st, u, v, h) bind(C, name="swe_state_upload")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: st
            real(c_double), intent(in) :: u(*), v(*), h(*)
        end function
        integer(c_int) function swe_state_download(st, u, v, h) bind(C, name="swe_state_download")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: st
            real(c_double), intent(out) :: u(*), v(*), h(*)
        end function
        ! the reference's halo'd tile arrays p(1-hw:nx+hw, 1-hw:ny+hw), one tile per panel
        integer(c_int) function swe_state_upload_halo(st, hw, u, v, h) bind(C, name="swe_state_upload_halo")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: st
            integer(c_int), value :: hw
            real(c_double), intent(in) :: u(*), v(*), h(*)
        end function
        integer(c_int) function swe_state_download_halo(st, hw, u, v, h) bind(C, name="swe_state_download_halo")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: st
            integer(c_int), value :: hw
            real(c_double), intent(inout) :: u(*), v(*), h(*)
        end function
        integer(c_int) function swe_apply(plan, vout, vin) bind(C, name="swe_apply")
            import :: c_int, c_ptr
            type(c_ptr), value :: plan, vout, vin
        end function
        integer(c_int) function swe_assign(plan, out, s1, v1, s2, v2) bind(C, name="swe_assign")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: plan, out, v1, v2
            real(c_double), value :: s1, s2
        end function
        integer(c_int) function swe_update(plan, out, s1, v1, s2, v2) bind(C, name="swe_update")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: plan, out, v1, v2
            real(c_double), value :: s1, s2
        end function
        integer(c_int) function swe_rk4_steps(plan, v0, dt, nsteps) bind(C, name="swe_rk4_steps")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: plan, v0
            real(c_double), value :: dt
            integer(c_int), value :: nsteps
        end function
        integer(c_int) function swe_integrals(plan, st, mass, energy) bind(C, name="swe_integrals")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: plan, st
            real(c_double), intent(out) :: mass, energy
        end function
        integer(c_int) function swe_integrals_diag(plan, st, grav, out) bind(C, name="swe_integrals_diag")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: plan, st
            real(c_double), value :: grav
            real(c_double), intent(out) :: out(7)
        end function
        integer(c_int) function swe_solve(plan, vout, rhs, dt, max_iter, rel_tol, iters, relres) &
                bind(C, name="swe_solve")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: plan, vout, rhs
            real(c_double), value :: dt, rel_tol
            integer(c_int), value :: max_iter
            integer(c_int), intent(out) :: iters
            real(c_double), intent(out) :: relres
        end function
        integer(c_int) function swe_implicit_euler_steps(plan, v0, dt, nsteps, max_iter, rel_tol, iters, relres) &
                bind(C, name="swe_implicit_euler_steps")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: plan, v0
            real(c_double), value :: dt, rel_tol
            integer(c_int), value :: nsteps, max_iter
            integer(c_int), intent(out) :: iters
            real(c_double), intent(out) :: relres
        end function
        integer(c_int) function swe_crank_nicolson_steps(plan, v0, dt, nsteps, max_iter, rel_tol, iters, relres) &
                bind(C, name="swe_crank_nicolson_steps")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: plan, v0
            real(c_double), value :: dt, rel_tol
            integer(c_int), value :: nsteps, max_iter
            integer(c_int), intent(out) :: iters
            real(c_double), intent(out) :: relres
        end function
        integer(c_int) function swe_cfl_diag(plan, st, dt, grav, out) bind(C, name="swe_cfl_diag")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: plan, st
            real(c_double), value :: dt, grav
            real(c_double), intent(out) :: out(3)
        end function
        integer(c_int) function swe_latlon_uv(plan, st, nlat, nlon, uo, vo) bind(C, name="swe_latlon_uv")
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: plan, st
            integer(c_int), value :: nlat, nlon
            real(c_float), intent(out) :: uo(*), vo(*)
        end function
        integer(c_int) function swe_latlon_h(plan, st, nlat, nlon, out) bind(C, name="swe_latlon_h")
            import :: c_int, c_ptr, c_float
            type(c_ptr), value :: plan, st
            integer(c_int), value :: nlat, nlon
            real(c_float), intent(out) :: out(*)
        end function
        integer(c_int) function swe_integral_tends(plan, st, grav, out) bind(C, name="swe_integral_tends")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: plan, st
            real(c_double), value :: grav
            real(c_double), intent(out) :: out(2)
        end function
        type(c_ptr) function swe_last_error() bind(C, name="swe_last_error")
            import :: c_ptr
        end function
    end interface

contains

    ! message of the last failure (swe_last_error) as a Fortran string
    function swe_error_message() result(msg)
        character(len=:), allocatable :: msg
        character(kind=c_char), pointer :: s(:)
        type(c_ptr) :: p
        integer :: n
        p = swe_last_error()
        msg = ""
        if (.not. c_associated(p)) return
        call c_f_pointer(p, s, [4096])
        n = 0
        do while (n < 4096)
            if (s(n + 1) == c_null_char) exit
            n = n + 1
        end do
        allocate(character(len=n) :: msg)
        msg = transfer(s(1:n), msg)
    end function swe_error_message

end module swe_hip_capi_mod
