! Fortran host layer over libchswe.so, shaped like the reference's interface
! for this path so that it drops in where operator_swm_lin_t, stvec_swm_t and
! rk4_t sit:
!
!   reference                                   here
!   operator_t%apply(vout, vin, domain)         operator_swm_lin_hip_t%apply
!     (src/operator_mod.f90 L11-41)
!   stvec_t%assign / %update / %create_similar  stvec_hip_t%assign / %update / %create_similar
!     (src/stvec/stvec_mod.f90 L10-30)
!   timescheme_t%step(v0, operator, domain, dt) rk4_hip_t%step
!     (src/time_schemes/timescheme_mod.f90 L9-24, rk4_mod.f90 L20-42)
!   parcomm_global%abort(message)               swe_hip_abort (error stop)
!
! The abstract types below restate the reference's deferred bindings that this
! path uses; a build inside the reference extends the reference's own
! operator_t / stvec_t / timescheme_t instead (same procedures, same calls).
module swe_hip_mod

    use iso_c_binding, only : c_ptr, c_null_ptr, c_associated, c_int, c_double
    use swe_hip_capi_mod

    implicit none

    private

    public :: SWE_FP64, SWE_FP32
    public :: domain_hip_t, stvec_t, operator_t, timescheme_t
    public :: stvec_hip_t, operator_swm_lin_hip_t, rk4_hip_t, implicit_hip_t
    public :: create_domain_hip, create_linear_swm_operator_hip, create_stvec_hip, create_rk4_hip, create_implicit_hip
    public :: swe_hip_abort

    ! the "domain": geometry, edge maps and stream live in the device plan
    type :: domain_hip_t
        type(c_ptr)     :: plan = c_null_ptr
        integer(c_int)  :: N = 0, order = 0, precision = SWE_FP64
    contains
        procedure :: destroy => destroy_domain
    end type domain_hip_t

    type, abstract :: stvec_t
    contains
        procedure(assign_s1v1s2v2_i), deferred :: assign_s1v1s2v2
        procedure(assign_s1v1_i),     deferred :: assign_s1v1
        generic :: assign => assign_s1v1s2v2, assign_s1v1
        procedure(assign_s1v1s2v2_i), deferred :: update_s1v1s2v2
        procedure(assign_s1v1_i),     deferred :: update_s1v1
        generic :: update => update_s1v1s2v2, update_s1v1
    end type stvec_t

    type, abstract :: operator_t
    contains
        procedure(apply_i), deferred :: apply   ! vout = A * vin
    end type operator_t

    type, abstract :: timescheme_t
    contains
        procedure(step_i), deferred :: step
    end type timescheme_t

    abstract interface
        subroutine assign_s1v1s2v2_i(this, scalar1, v1, scalar2, v2, domain)
            import :: stvec_t, domain_hip_t
            class(stvec_t),     intent(inout) :: this
            real(kind=8),       intent(in)    :: scalar1, scalar2
            class(stvec_t),     intent(in)    :: v1, v2
            type(domain_hip_t), intent(in)    :: domain
        end subroutine
        subroutine assign_s1v1_i(this, scalar1, v1, domain)
            import :: stvec_t, domain_hip_t
            class(stvec_t),     intent(inout) :: this
            real(kind=8),       intent(in)    :: scalar1
            class(stvec_t),     intent(in)    :: v1
            type(domain_hip_t), intent(in)    :: domain
        end subroutine
        subroutine apply_i(this, vout, vin, domain)
            import :: operator_t, stvec_t, domain_hip_t
            class(operator_t),  intent(inout) :: this
            class(stvec_t),     intent(inout) :: vout   ! inout to enable preallocated vectors
            class(stvec_t),     intent(inout) :: vin
            type(domain_hip_t), intent(in)    :: domain
        end subroutine
        subroutine step_i(this, v0, operator, domain, dt)
            import :: timescheme_t, stvec_t, operator_t, domain_hip_t
            class(timescheme_t), intent(inout) :: this
            class(stvec_t),      intent(inout) :: v0
            class(operator_t),   intent(inout) :: operator
            type(domain_hip_t),  intent(in)    :: domain
            real(kind=8),        intent(in)    :: dt
        end subroutine
    end interface

    ! device-resident (u, v, h): stvec_swm_t counterpart
    type, extends(stvec_t) :: stvec_hip_t
        type(c_ptr) :: st = c_null_ptr
    contains
        procedure :: assign_s1v1s2v2 => stvec_assign_s1v1s2v2
        procedure :: assign_s1v1     => stvec_assign_s1v1
        procedure :: update_s1v1s2v2 => stvec_update_s1v1s2v2
        procedure :: update_s1v1     => stvec_update_s1v1
        procedure :: create_similar  => stvec_create_similar
        procedure :: upload, download, destroy => destroy_stvec
    end type stvec_hip_t

    type, extends(operator_t) :: operator_swm_lin_hip_t
    contains
        procedure :: apply => operator_apply
        procedure :: integrals => operator_integrals
        procedure :: integrals_diag => operator_integrals_diag
        procedure :: integral_tends => operator_integral_tends
        procedure :: latlon_h => operator_latlon_h
        procedure :: latlon_uv => operator_latlon_uv
        procedure :: cfl_diag => operator_cfl_diag
        procedure :: solve_h => operator_solve
    end type operator_swm_lin_hip_t

    type, extends(timescheme_t) :: rk4_hip_t
    contains
        procedure :: step => rk4_step
        procedure :: steps => rk4_steps
    end type rk4_hip_t

    ! implicit_Eul_t / crank_nikolson_t on the operator's solve, with the
    ! helm_solver_cfg settings; iters / relres: the largest over the last
    ! call's steps
    type, extends(timescheme_t) :: implicit_hip_t
        logical :: crank_nicolson = .false.
        integer :: max_iter = 100
        real(kind=8) :: rel_tol = 1.0e-10_8
        integer :: iters = 0
        real(kind=8) :: relres = 0.0_8
    contains
        procedure :: step => implicit_step
        procedure :: steps => implicit_steps
    end type implicit_hip_t

contains

    subroutine swe_hip_abort(message)
        character(len=*), intent(in) :: message
        write(*, '(A)') "ch_swe_amd failure: "//message
        error stop 1
    end subroutine swe_hip_abort

    subroutine check(rc, what)
        integer(c_int),   intent(in) :: rc
        character(len=*), intent(in) :: what
        if (rc /= 0) call swe_hip_abort(what//": "//swe_error_message())
    end subroutine check

    ! create_domain + the linear operator's constants (shallow_water_mod.f90 L70-93)
    ! axis: rotation axis (default the z axis); flags: SWE_FLAG_* bits of
    ! include/ch_swe.h (default 0)
    function create_domain_hip(N, order, precision, radius, grav, h_mean, omega, coriolis_const, axis, flags) &
            result(domain)
        integer,      intent(in) :: N, order, precision
        real(kind=8), intent(in) :: radius, grav, h_mean, omega
        logical,      intent(in) :: coriolis_const
        real(kind=8), intent(in), optional :: axis(3)
        integer,      intent(in), optional :: flags
        type(domain_hip_t)       :: domain
        type(swe_config_t)       :: cfg
        cfg%N = N
        cfg%order = order
        cfg%precision = precision
        cfg%coriolis_const = merge(1, 0, coriolis_const)
        cfg%radius = radius
        cfg%grav = grav
        cfg%h_mean = h_mean
        cfg%omega = omega
        cfg%axis = [0.0_8, 0.0_8, 1.0_8]
        if (present(axis)) cfg%axis = axis
        cfg%flags = 0
        if (present(flags)) cfg%flags = flags
        cfg%block_threads = 0
        cfg%tile_x0 = 0
        cfg%tile_y0 = 0
        call check(swe_plan_create(cfg, domain%plan), "create_domain_hip")
        domain%N = N
        domain%order = order
        domain%precision = precision
    end function create_domain_hip

    subroutine destroy_domain(this)
        class(domain_hip_t), intent(inout) :: this
        integer(c_int) :: rc
        if (c_associated(this%plan)) rc = swe_plan_destroy(this%plan)
        this%plan = c_null_ptr
    end subroutine destroy_domain

    function create_linear_swm_operator_hip() result(op)
        type(operator_swm_lin_hip_t) :: op
    end function create_linear_swm_operator_hip

    function create_rk4_hip() result(ts)
        type(rk4_hip_t) :: ts
    end function create_rk4_hip

    ! timescheme_factory_mod.f90 L31-34: 'implicit_Eul' or 'crank_nikolson'
    function create_implicit_hip(crank_nicolson, max_iter, rel_tol) result(ts)
        logical,      intent(in) :: crank_nicolson
        integer,      intent(in) :: max_iter
        real(kind=8), intent(in) :: rel_tol
        type(implicit_hip_t) :: ts
        ts%crank_nicolson = crank_nicolson
        ts%max_iter = max_iter
        ts%rel_tol = rel_tol
    end function create_implicit_hip

    subroutine create_stvec_hip(v, domain)
        type(stvec_hip_t),  intent(out) :: v
        type(domain_hip_t), intent(in)  :: domain
        call check(swe_state_create(domain%plan, v%st), "create_stvec_hip")
    end subroutine create_stvec_hip

    subroutine stvec_create_similar(this, destination, domain)
        class(stvec_hip_t), intent(in)  :: this
        type(stvec_hip_t),  intent(out) :: destination
        type(domain_hip_t), intent(in)  :: domain
        call create_stvec_hip(destination, domain)
    end subroutine stvec_create_similar

    subroutine destroy_stvec(this)
        class(stvec_hip_t), intent(inout) :: this
        integer(c_int) :: rc
        if (c_associated(this%st)) rc = swe_state_destroy(this%st)
        this%st = c_null_ptr
    end subroutine destroy_stvec

    ! packed interior layout: u(N, N+1, 6), v(N+1, N, 6), h(N+1, N+1, 6)
    subroutine upload(this, u, v, h)
        class(stvec_hip_t), intent(inout) :: this
        real(kind=8),       intent(in)    :: u(:,:,:), v(:,:,:), h(:,:,:)
        call check(swe_state_upload(this%st, u, v, h), "stvec upload")
    end subroutine upload

    subroutine download(this, u, v, h)
        class(stvec_hip_t), intent(inout) :: this
        real(kind=8),       intent(out)   :: u(:,:,:), v(:,:,:), h(:,:,:)
        call check(swe_state_download(this%st, u, v, h), "stvec download")
    end subroutine download

    function hip_state(v) result(p)
        class(stvec_t), intent(in) :: v
        type(c_ptr) :: p
        select type (v)
        class is (stvec_hip_t)
            p = v%st
        class default
            call swe_hip_abort("stvec of wrong type (expected stvec_hip_t)")
        end select
    end function hip_state

    subroutine stvec_assign_s1v1s2v2(this, scalar1, v1, scalar2, v2, domain)
        class(stvec_hip_t), intent(inout) :: this
        real(kind=8),       intent(in)    :: scalar1, scalar2
        class(stvec_t),     intent(in)    :: v1, v2
        type(domain_hip_t), intent(in)    :: domain
        call check(swe_assign(domain%plan, this%st, scalar1, hip_state(v1), scalar2, hip_state(v2)), "assign")
    end subroutine stvec_assign_s1v1s2v2

    subroutine stvec_assign_s1v1(this, scalar1, v1, domain)
        class(stvec_hip_t), intent(inout) :: this
        real(kind=8),       intent(in)    :: scalar1
        class(stvec_t),     intent(in)    :: v1
        type(domain_hip_t), intent(in)    :: domain
        call check(swe_assign(domain%plan, this%st, scalar1, hip_state(v1), 0.0_8, c_null_ptr), "assign")
    end subroutine stvec_assign_s1v1

    subroutine stvec_update_s1v1s2v2(this, scalar1, v1, scalar2, v2, domain)
        class(stvec_hip_t), intent(inout) :: this
        real(kind=8),       intent(in)    :: scalar1, scalar2
        class(stvec_t),     intent(in)    :: v1, v2
        type(domain_hip_t), intent(in)    :: domain
        call check(swe_update(domain%plan, this%st, scalar1, hip_state(v1), scalar2, hip_state(v2)), "update")
    end subroutine stvec_update_s1v1s2v2

    subroutine stvec_update_s1v1(this, scalar1, v1, domain)
        class(stvec_hip_t), intent(inout) :: this
        real(kind=8),       intent(in)    :: scalar1
        class(stvec_t),     intent(in)    :: v1
        type(domain_hip_t), intent(in)    :: domain
        call check(swe_update(domain%plan, this%st, scalar1, hip_state(v1), 0.0_8, c_null_ptr), "update")
    end subroutine stvec_update_s1v1

    ! operator_swm_lin_mod.f90 apply L66-108: vout = A(vin)
    subroutine operator_apply(this, vout, vin, domain)
        class(operator_swm_lin_hip_t), intent(inout) :: this
        class(stvec_t),                intent(inout) :: vout
        class(stvec_t),                intent(inout) :: vin
        type(domain_hip_t),            intent(in)    :: domain
        call check(swe_apply(domain%plan, hip_state(vout), hip_state(vin)), "swm operator apply")
    end subroutine operator_apply

    subroutine operator_integrals(this, v, domain, mass, energy)
        class(operator_swm_lin_hip_t), intent(inout) :: this
        class(stvec_t),                intent(in)    :: v
        type(domain_hip_t),            intent(in)    :: domain
        real(kind=8),                  intent(out)   :: mass, energy
        call check(swe_integrals(domain%plan, hip_state(v), mass, energy), "integrals")
    end subroutine operator_integrals

    ! swm_output_diag_mod.f90 print_integrals_diag L156-184: diag = hmin, hmax,
    ! mass, TE, KE, PE, enstrophy (grav: the PE's gravity, Earth_grav there)
    subroutine operator_integrals_diag(this, v, domain, grav, diag)
        class(operator_swm_lin_hip_t), intent(inout) :: this
        class(stvec_t),                intent(in)    :: v
        type(domain_hip_t),            intent(in)    :: domain
        real(kind=8),                  intent(in)    :: grav
        real(kind=8),                  intent(out)   :: diag(7)
        call check(swe_integrals_diag(domain%plan, hip_state(v), grav, diag), "integrals_diag")
    end subroutine operator_integrals_diag

    ! swm_output_diag_mod.f90 print_integral_tends L236-342 of this operator:
    ! tends = TE_tend, Enstrophy_tend
    subroutine operator_integral_tends(this, v, domain, grav, tends)
        class(operator_swm_lin_hip_t), intent(inout) :: this
        class(stvec_t),                intent(in)    :: v
        type(domain_hip_t),            intent(in)    :: domain
        real(kind=8),                  intent(in)    :: grav
        real(kind=8),                  intent(out)   :: tends(2)
        call check(swe_integral_tends(domain%plan, hip_state(v), grav, tends), "integral_tends")
    end subroutine operator_integral_tends

    ! swm_output_diag_mod.f90 write_fields L52-73 (h_surf = 0) through
    ! latlon_outputer_mod.f90 latlon_write L33-63: h on the lat-lon grid,
    ! hll(nlon, nlat) as one record of h.dat
    subroutine operator_latlon_h(this, v, domain, hll)
        class(operator_swm_lin_hip_t), intent(inout) :: this
        class(stvec_t),                intent(in)    :: v
        type(domain_hip_t),            intent(in)    :: domain
        real(kind=4),                  intent(out)   :: hll(:, :)
        call check(swe_latlon_h(domain%plan, hip_state(v), int(size(hll, 2), c_int), int(size(hll, 1), c_int), hll), &
                   "latlon_h")
    end subroutine operator_latlon_h

    ! swm_output_diag_mod.f90 write_fields L75-76 through latlon_outputer_mod.f90
    ! latlon_vec_write L65-107: zonal ull and meridional vll winds, (nlon, nlat)
    ! each, one record of u.dat / v.dat
    subroutine operator_latlon_uv(this, v, domain, ull, vll)
        class(operator_swm_lin_hip_t), intent(inout) :: this
        class(stvec_t),                intent(in)    :: v
        type(domain_hip_t),            intent(in)    :: domain
        real(kind=4),                  intent(out)   :: ull(:, :), vll(:, :)
        call check(swe_latlon_uv(domain%plan, hip_state(v), int(size(ull, 2), c_int), int(size(ull, 1), c_int), &
                                 ull, vll), "latlon_uv")
    end subroutine operator_latlon_uv

    ! swm_output_diag_mod.f90 print_CFL_diag L110-154: cfl = gravity-wave
    ! CFL, advective CFL along x and y
    subroutine operator_cfl_diag(this, v, domain, dt, grav, cfl)
        class(operator_swm_lin_hip_t), intent(inout) :: this
        class(stvec_t),                intent(in)    :: v
        type(domain_hip_t),            intent(in)    :: domain
        real(kind=8),                  intent(in)    :: dt, grav
        real(kind=8),                  intent(out)   :: cfl(3)
        call check(swe_cfl_diag(domain%plan, hip_state(v), dt, grav, cfl), "cfl_diag")
    end subroutine operator_cfl_diag

    ! operator_swm_lin_mod.f90 solve L110-146: vout = (I - dt L0)^-1 rhs (no
    ! Coriolis), CG on the Helmholtz problem for h
    subroutine operator_solve(this, vout, rhs, domain, dt, max_iter, rel_tol, iters, relres)
        class(operator_swm_lin_hip_t), intent(inout) :: this
        class(stvec_t),                intent(inout) :: vout
        class(stvec_t),                intent(in)    :: rhs
        type(domain_hip_t),            intent(in)    :: domain
        real(kind=8),                  intent(in)    :: dt, rel_tol
        integer,                       intent(in)    :: max_iter
        integer,                       intent(out)   :: iters
        real(kind=8),                  intent(out)   :: relres
        integer(c_int) :: it
        call check(swe_solve(domain%plan, hip_state(vout), hip_state(rhs), dt, int(max_iter, c_int), rel_tol, it, &
                             relres), "solve")
        iters = it
    end subroutine operator_solve

    subroutine implicit_step(this, v0, operator, domain, dt)
        class(implicit_hip_t), intent(inout) :: this
        class(stvec_t),        intent(inout) :: v0
        class(operator_t),     intent(inout) :: operator
        type(domain_hip_t),    intent(in)    :: domain
        real(kind=8),          intent(in)    :: dt
        call this%steps(v0, domain, dt, 1)
    end subroutine implicit_step

    ! implicit_Eul_mod.f90 L20-30 / crank_nikolson_mod.f90 L20-32, nsteps times
    subroutine implicit_steps(this, v0, domain, dt, nsteps)
        class(implicit_hip_t), intent(inout) :: this
        class(stvec_t),        intent(inout) :: v0
        type(domain_hip_t),    intent(in)    :: domain
        real(kind=8),          intent(in)    :: dt
        integer,               intent(in)    :: nsteps
        integer(c_int) :: it
        if (this%crank_nicolson) then
            call check(swe_crank_nicolson_steps(domain%plan, hip_state(v0), dt, int(nsteps, c_int), &
                                                int(this%max_iter, c_int), this%rel_tol, it, this%relres), &
                       "crank_nikolson steps")
        else
            call check(swe_implicit_euler_steps(domain%plan, hip_state(v0), dt, int(nsteps, c_int), &
                                                int(this%max_iter, c_int), this%rel_tol, it, this%relres), &
                       "implicit_Eul steps")
        end if
        this%iters = it
    end subroutine implicit_steps

    ! rk4_mod.f90 step_rk4 L20-42 (fused stages on the device)
    subroutine rk4_step(this, v0, operator, domain, dt)
        class(rk4_hip_t),   intent(inout) :: this
        class(stvec_t),     intent(inout) :: v0
        class(operator_t),  intent(inout) :: operator
        type(domain_hip_t), intent(in)    :: domain
        real(kind=8),       intent(in)    :: dt
        call check(swe_rk4_steps(domain%plan, hip_state(v0), dt, 1_c_int), "rk4 step")
    end subroutine rk4_step

    subroutine rk4_steps(this, v0, domain, dt, nsteps)
        class(rk4_hip_t),   intent(inout) :: this
        class(stvec_t),     intent(inout) :: v0
        type(domain_hip_t), intent(in)    :: domain
        real(kind=8),       intent(in)    :: dt
        integer,            intent(in)    :: nsteps
        call check(swe_rk4_steps(domain%plan, hip_state(v0), dt, int(nsteps, c_int)), "rk4 steps")
    end subroutine rk4_steps

end module swe_hip_mod
