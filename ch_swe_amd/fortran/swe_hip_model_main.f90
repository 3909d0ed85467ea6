! The linear shallow-water model driven by the reference's own configuration
! files, on the GPU: SHALLOW_WATER_MAIN (src/models/shallow_water/
! shallow_water_main.f90) -> run_shallow_water_model (shallow_water_mod.f90
! L29-145) with the Ch-grid linear operator of libchswe.so.
!
! usage: swe_hip_model config.cfg [out.bin] [fp64|fp32|fp64-compensated|fp32-compensated]
!
! Configurations accepted (the others stop with a message naming the key):
!   domain:  topology_type 'cube', horizontal_staggering 'Ch', Nz 1,
!            metric_2d_type 'ecs', planet_radius, omega, rotation_axis,
!            coriolis_constant, rotation_matrix only the identity
!            (metric_factory_mod.f90 L28-60, L176-200 with its defaults:
!            radius 1, omega 1, axis [0,0,1]);
!   operator (swm_operator_section): swm_operator_type 'linear_swm_operator',
!            v_components_type 'covariant', h_mean, div_op_name
!            'div_ch_sbp_sat_XX_proj', grad_op_name 'grad_ch_sbp_sat_XX',
!            co2contra_op_name 'co2contra_ch_sbpXX', coriolis_op_name
!            'coriolis_Ch_sbpXX', quadrature_name 'SBP_CXX_quadrature' with
!            one XX in {21, 42, 63} (operator_swm_factory_mod.f90 L586-685;
!            sbp63 plans run on the multi-kernel path);
!   timescheme_name 'rk4', 'implicit_Eul' or 'crank_nikolson' (the last two
!            with the operator section's helm_solver_cfg: solver_name 'cg',
!            max_iter, rel_tol, verbose; iterative_solver_factory_mod.f90
!            L27-36);
!   testcase (testcase_section -> testcase_name): 'Gaussian_hill_linear'
!            (phi0, lam0, d0, refsol_path; gaussian_hill_linear_test_mod.f90
!            L30-67) or 'Williamson_test2' / 'TS2' / 'test2' with linear_mode
!            .true. (u0, h_mean; Williamson_test2_mod.f90 L32-112);
!   diffusion_section absent or 'None'.
! g is Earth_grav (shallow_water_mod.f90 L93, const_mod.f90 L6).
!
! Output, as the reference prints it: print_CFL_diag's line
! (swm_output_diag_mod.f90 L110-154, swe_cfl_diag) first, the scores at t = 0 and at every
! diagnostics step (Gaussian hill: against the reference solution file
! refsol_path when it exists, gaussian_hill_linear_test_mod.f90 L71-129;
! TS2: relative l2 / linf error of h, Williamson_test2_mod.f90 L115-143),
! 'rec n' at every write step, and at every diagnostics step the line of
! print_integrals_diag (swm_output_diag_mod.f90 L156-184): hmin, hmax, mass,
! TE, KE, PE, enstrophy, computed on the device (swe_integrals_diag) with the
! Ch operators of the paper's output_diag section, followed by the line of
! print_integral_tends (L236-342, "dynamics: TE_tend = ..., Enstrophy_tend =
! ...", swe_integral_tends).  h.dat, u.dat and v.dat in the working directory
! get the lat-lon h and zonal / meridional winds (2N+1 latitudes, 4N
! longitudes, float32; swe_latlon_h, swe_latlon_uv) at t = 0 and at every
! write step, as write_fields writes them.
! The final state is written to out.bin (default swe_hip_state.bin) in the
! packed (u, v, h) layout of include/ch_swe.h.
program swe_hip_model

    use swe_hip_mod
    use swe_hip_cfg_mod

    implicit none

    real(kind=8), parameter :: pi = acos(-1.0_8)
    real(kind=8), parameter :: Earth_grav = 9.80616_8, Earth_radii = 0.6371229e7_8
    real(kind=8), parameter :: Earth_sidereal_T = 23 * 60 * 60 + 56 * 60 + 4
    real(kind=8), parameter :: Earth_omega = 2.0_8 * pi / Earth_sidereal_T
    integer, parameter :: FLAG_COMPENSATED = 8
    ! panel frames: topology/cubed_sphere_topology_mod.f90 L29-55
    integer, parameter :: ex(3, 6) = reshape([-1, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, 1, 1, 0, 0, 1, 0, 0], [3, 6])
    integer, parameter :: ey(3, 6) = reshape([0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 0, 1, 0, 0, -1], [3, 6])
    integer, parameter :: nn(3, 6) = reshape([0, 0, 1, -1, 0, 0, 0, 0, -1, 1, 0, 0, 0, 1, 0, 0, -1, 0], [3, 6])

    type(config_hip_t)           :: config
    type(domain_hip_t)           :: domain
    type(operator_swm_lin_hip_t) :: operator
    type(rk4_hip_t)              :: timescheme
    type(implicit_hip_t)         :: implicit_ts
    logical                      :: implicit, cg_verbose
    type(stvec_hip_t)            :: state
    real(kind=8), allocatable    :: u(:,:,:), v(:,:,:), h(:,:,:), h0(:,:,:), jac(:,:)
    character(len=:), allocatable :: opsec, tcsec, tcname, name, refsol_path
    character(len=512) :: cfg_file, out_file, arg
    integer :: N, order, precision, flags, nstep_write, nstep_diagnostics, nsteps, it, nrun, p, i, j, k
    real(kind=8) :: dt, tau_write, tau_diagnostics, simulation_time, radius, omega, h_mean, axis(3), rm(9)
    real(kind=8) :: lam0, phi0, d0, u0, tc_h_mean, an(3), c(3), r(3), al, be, X, Y, s, vel(3), a1(3), a2(3)
    real(kind=8) :: diag(7), tends(2), cfl(3), l2_h0, linf_h0
    real(kind=4), allocatable :: hll(:, :), ull(:, :), vll(:, :)
    logical :: coriolis_const, linear_mode, gauss, have_refsol

    call get_command_argument(1, cfg_file)
    if (len_trim(cfg_file) == 0) &
        call swe_hip_abort("usage: swe_hip_model config.cfg [out.bin] [fp64|fp32|fp64-compensated|fp32-compensated]")
    out_file = "swe_hip_state.bin"
    if (command_argument_count() >= 2) call get_command_argument(2, out_file)
    precision = 0
    flags = 0
    if (command_argument_count() >= 3) then
        call get_command_argument(3, arg)
        select case (trim(arg))
        case ("fp64")
        case ("fp32")
            precision = 1
        case ("fp64-compensated")
            flags = FLAG_COMPENSATED
        case ("fp32-compensated")
            precision = 1
            flags = FLAG_COMPENSATED
        case default
            call swe_hip_abort("unknown precision "//trim(arg)//" (fp64, fp32, fp64-compensated, fp32-compensated)")
        end select
    end if

    config = parse_config_file(trim(cfg_file))

    ! time parameters: shallow_water_mod.f90 L60-66
    dt = config%get_real("dt")
    tau_write = config%get_real("tau_write")
    tau_diagnostics = config%get_real("tau_diagnostics")
    simulation_time = config%get_real("simulation_time")
    nstep_write = nint(tau_write / dt)
    nstep_diagnostics = nint(tau_diagnostics / dt)
    if (nstep_write < 1 .or. nstep_diagnostics < 1) &
        call swe_hip_abort("tau_write and tau_diagnostics must be at least dt")
    nsteps = int(simulation_time / dt)

    ! domain: domain_factory_mod.f90 L70-130, metric_factory_mod.f90 L28-200
    N = config%get_int("domain%N")
    call expect_str("domain%topology_type", "cube", "cube")
    call expect_str("domain%horizontal_staggering", "", "Ch")
    if (config%get_int("domain%Nz", 1) /= 1) call swe_hip_abort("domain%Nz: only Nz = 1 is supported")
    if (config%has("domain%metric%metric_2d_type")) call expect_str("domain%metric%metric_2d_type", "", "ecs")
    radius = config%get_real("domain%metric%planet_radius", 1.0_8)
    omega = config%get_real("domain%metric%omega", 1.0_8)
    coriolis_const = config%get_logical("domain%metric%coriolis_constant", .false.)
    axis = [0.0_8, 0.0_8, 1.0_8]
    if (config%has("domain%metric%rotation_axis%length")) then
        if (config%get_int("domain%metric%rotation_axis%length") /= 3) &
            call swe_hip_abort("domain%metric%rotation_axis must have 3 elements")
        do k = 1, 3
            axis(k) = config%get_real("domain%metric%rotation_axis%"//achar(48 + k))
        end do
    end if
    if (config%has("domain%metric%rotation_matrix%length")) then
        if (config%get_int("domain%metric%rotation_matrix%length") /= 9) &
            call swe_hip_abort("domain%metric%rotation_matrix must have 9 elements")
        do k = 1, 9
            rm(k) = config%get_real("domain%metric%rotation_matrix%"//achar(48 + k))
        end do
        if (any(rm /= [1.0_8, 0.0_8, 0.0_8, 0.0_8, 1.0_8, 0.0_8, 0.0_8, 0.0_8, 1.0_8])) &
            call swe_hip_abort("domain%metric%rotation_matrix: only the identity is supported")
    end if

    ! operator: operator_swm_factory_mod.f90 L586-685
    opsec = config%get_str("swm_operator_section")
    call expect_str(opsec//"%swm_operator_type", "", "linear_swm_operator")
    call expect_str(opsec//"%v_components_type", "", "covariant")
    h_mean = config%get_real(opsec//"%h_mean")
    order = sbp_order(opsec//"%div_op_name", "div_ch_sbp_sat_", "_proj")
    call same_order(sbp_order(opsec//"%grad_op_name", "grad_ch_sbp_sat_", ""), opsec//"%grad_op_name")
    call same_order(sbp_order(opsec//"%co2contra_op_name", "co2contra_ch_sbp", ""), opsec//"%co2contra_op_name")
    call same_order(sbp_order(opsec//"%coriolis_op_name", "coriolis_Ch_sbp", ""), opsec//"%coriolis_op_name")
    call same_order(sbp_order(opsec//"%quadrature_name", "SBP_C", "_quadrature"), opsec//"%quadrature_name")

    name = config%get_str("timescheme_name")
    implicit = name == "implicit_Eul" .or. name == "crank_nikolson"
    if (.not. implicit .and. name /= "rk4") &
        call swe_hip_abort("timescheme_name = '"//name//"': only rk4, implicit_Eul and crank_nikolson run here")
    if (implicit) then
        call expect_str(opsec//"%helm_solver_cfg%solver_name", "", "cg")
        implicit_ts = create_implicit_hip(name == "crank_nikolson", config%get_int(opsec//"%helm_solver_cfg%max_iter"), &
                                          config%get_real(opsec//"%helm_solver_cfg%rel_tol"))
        cg_verbose = config%get_logical(opsec//"%helm_solver_cfg%verbose", .false.)
    end if
    name = config%get_str("diffusion_section", "None")
    if (name /= "None") call swe_hip_abort("diffusion_section: horizontal diffusion is not supported")

    ! initial conditions: shallow_water_mod.f90 setup_testcase L147-192
    tcsec = config%get_str("testcase_section")
    tcname = config%get_str(tcsec//"%testcase_name")
    allocate(u(N, N + 1, 6), v(N + 1, N, 6), h(N + 1, N + 1, 6), h0(N + 1, N + 1, 6), jac(N + 1, N + 1))
    u = 0.0_8
    v = 0.0_8
    gauss = .false.
    have_refsol = .false.
    select case (tcname)
    case ("Gaussian_hill_linear")
        ! gaussian_hill_linear_test_mod.f90 L30-67, test_fields_mod.f90 L491-514
        gauss = .true.
        phi0 = config%get_real(tcsec//"%phi0", 0.0_8)
        lam0 = config%get_real(tcsec//"%lam0", pi)
        d0 = config%get_real(tcsec//"%d0", 0.25_8)
        refsol_path = config%get_str(tcsec//"%refsol_path", "ref_solutions/gauss_linear/gauss_h_cub128_xy.dat")
        inquire(file=refsol_path, exist=have_refsol)
        if (.not. have_refsol) write(*, '(A)') "reference solution "//refsol_path//" not found: scores not printed"
        if (have_refsol .and. mod(192, N) /= 0) then
            write(*, '(A)') "N does not divide 192: reference-solution scores not printed"
            have_refsol = .false.
        end if
        c = [cos(phi0) * cos(lam0), cos(phi0) * sin(lam0), sin(phi0)]
        do p = 1, 6
            do j = 1, N + 1
                do i = 1, N + 1
                    call point(p, i, j, 0.0_8, 0.0_8, r)
                    h(i, j, p) = exp(-(acos(min(1.0_8, max(-1.0_8, sum(r * c)))) / d0)**2)
                end do
            end do
        end do
    case ("Williamson_test2", "TS2", "test2")
        ! Williamson_test2_mod.f90 L32-112; test_fields_mod.f90 L586-612 and L857-894
        linear_mode = config%get_logical(tcsec//"%linear_mode", .false.)
        if (.not. linear_mode) call swe_hip_abort(tcsec//"%linear_mode: the linear operator needs linear_mode = .true.")
        u0 = config%get_real(tcsec//"%u0", pi * Earth_radii / 6.0_8 / Earth_sidereal_T)
        tc_h_mean = config%get_real(tcsec//"%h_mean", 29400 / Earth_grav)
        if (u0 /= pi * Earth_radii / 6.0_8 / Earth_sidereal_T) &
            print *, "WARNING: Williamson test 2 using non-standard u0 = ", u0
        if (tc_h_mean /= 29400 / Earth_grav) &
            print *, "WARNING: Williamson test 2 using non-standard h_mean = ", tc_h_mean
        if (omega /= Earth_omega) print *, "WARNING: Williamson test 2 using non-standard omega = ", omega
        if (radius /= Earth_radii) print *, "WARNING: Williamson test 2 using non-standard planet radius = ", radius
        an = axis / sqrt(sum(axis**2))
        do p = 1, 6
            do j = 1, N + 1
                do i = 1, N + 1
                    call point(p, i, j, 0.0_8, 0.0_8, r)
                    h(i, j, p) = tc_h_mean - radius * omega * u0 / Earth_grav * sin2_lat(sum(r * an))
                end do
            end do
            ! covariant components V . a1 on the u points, V . a2 on the v points,
            ! V = u0 axis x r
            do j = 1, N + 1
                do i = 1, N
                    call point(p, i, j, 0.5_8, 0.0_8, r, a1, a2)
                    vel = u0 * [an(2) * r(3) - an(3) * r(2), an(3) * r(1) - an(1) * r(3), an(1) * r(2) - an(2) * r(1)]
                    u(i, j, p) = sum(vel * a1)
                end do
            end do
            do j = 1, N
                do i = 1, N + 1
                    call point(p, i, j, 0.0_8, 0.5_8, r, a1, a2)
                    vel = u0 * [an(2) * r(3) - an(3) * r(2), an(3) * r(1) - an(1) * r(3), an(1) * r(2) - an(2) * r(1)]
                    v(i, j, p) = sum(vel * a2)
                end do
            end do
        end do
    case default
        call swe_hip_abort("testcase_name "//tcname//": only Gaussian_hill_linear and Williamson_test2 (linear_mode)"// &
                           " run on the linear operator")
    end select
    h0 = h
    ! cell Jacobian of the xy points, for the TS2 scorer's l2 norm (vec_math_mod.f90 L89-130)
    do j = 1, N + 1
        do i = 1, N + 1
            X = tan(-0.25_8 * pi + (i - 1) * 0.5_8 * pi / N)
            Y = tan(-0.25_8 * pi + (j - 1) * 0.5_8 * pi / N)
            jac(i, j) = (1.0_8 + X * X) * (1.0_8 + Y * Y) / sqrt(1.0_8 + X * X + Y * Y)**3
        end do
    end do
    l2_h0 = sqrt(sum(h0**2 * spread(jac, 3, 6)))
    linf_h0 = maxval(abs(h0))
    call config%destroy()

    domain = create_domain_hip(N, order, precision, radius, Earth_grav, h_mean, omega, coriolis_const, axis, flags)
    operator = create_linear_swm_operator_hip()
    timescheme = create_rk4_hip()
    call create_stvec_hip(state, domain)
    call state%upload(u, v, h)

    ! print_CFL_diag (swm_output_diag_mod.f90 L110-154), then write_fields
    ! (L52-73): h.dat record 1 holds the initial state, record k + 1 the
    ! state after k write intervals
    call operator%cfl_diag(state, domain, dt, Earth_grav, cfl)
    print '(A,F15.7,A,2F15.7)', "Gravity waves CFL-max = ", cfl(1), ", Advective CFL-max (x,y) = ", cfl(2), cfl(3)
    allocate(hll(4 * N, 2 * N + 1), ull(4 * N, 2 * N + 1), vll(4 * N, 2 * N + 1))
    call write_h(1)
    call print_scores(0)
    it = 0
    do while (it < nsteps)
        ! run up to the next diagnostics / write step in one call
        nrun = min(nsteps, (it / nstep_diagnostics + 1) * nstep_diagnostics, (it / nstep_write + 1) * nstep_write) - it
        if (implicit) then
            call implicit_ts%steps(state, domain, dt, nrun)
            ! cg_solver_mod.f90 L82-88, once per chunk of steps (the largest
            ! iteration count over them)
            if (implicit_ts%iters > implicit_ts%max_iter) then
                print *, "CG failed to converge in ", implicit_ts%max_iter, " iterations"
            else if (cg_verbose) then
                print *, "CG converged in ", implicit_ts%iters, " iterations"
            end if
        else
            call timescheme%steps(state, domain, dt, nrun)
        end if
        it = it + nrun
        if (mod(it, nstep_diagnostics) == 0) then
            ! print_integrals_diag (swm_output_diag_mod.f90 L156-184) and
            ! print_integral_tends (L236-342, shallow_water_mod.f90 L126-127),
            ! then the scores
            call operator%integrals_diag(state, domain, Earth_grav, diag)
            print '(7(A,E24.16))', "hmin = ", diag(1), " hmax = ", diag(2), " mass = ", diag(3), &
                " TE = ", diag(4), " KE = ", diag(5), " PE = ", diag(6), " Enstrophy = ", diag(7)
            call operator%integral_tends(state, domain, Earth_grav, tends)
            print '(2(A,E20.12))', "dynamics: TE_tend = ", tends(1), ", Enstrophy_tend = ", tends(2)
            if (.not. gauss .or. have_refsol) then
                call state%download(u, v, h)
                call print_scores(it)
            end if
        end if
        if (mod(it, nstep_write) == 0) then
            call write_h(it / nstep_write + 1)
            print *, "rec", it / nstep_write + 1
        end if
    end do
    call state%download(u, v, h)

    open(unit=11, file=trim(out_file), access="stream", form="unformatted", status="replace")
    write(11) u, v, h
    close(11)

    call state%destroy()
    call domain%destroy()

contains

    subroutine expect_str(key, default, expected)
        character(len=*), intent(in) :: key, default, expected
        character(len=:), allocatable :: val
        if (len(default) > 0) then
            val = config%get_str(key, default)
        else
            val = config%get_str(key)
        end if
        if (val /= expected) call swe_hip_abort(key//" = '"//val//"': only '"//expected//"' is supported")
    end subroutine expect_str

    ! the SBP order XX of an operator name prefixXXsuffix (21, 42 or 63)
    integer function sbp_order(key, prefix, suffix)
        character(len=*), intent(in) :: key, prefix, suffix
        character(len=:), allocatable :: val
        val = config%get_str(key)
        sbp_order = 0
        if (val == prefix//"21"//suffix) sbp_order = 21
        if (val == prefix//"42"//suffix) sbp_order = 42
        if (val == prefix//"63"//suffix) sbp_order = 63
        if (sbp_order == 0) call swe_hip_abort(key//" = '"//val//"': supported are "// &
                                               prefix//"21"//suffix//", "//prefix//"42"//suffix//" and "// &
                                               prefix//"63"//suffix)
    end function sbp_order

    subroutine same_order(o, key)
        integer,          intent(in) :: o
        character(len=*), intent(in) :: key
        if (o /= order) call swe_hip_abort(key//": SBP order differs from div_op_name's")
    end subroutine same_order

    ! unit-sphere point of mesh node (i, j) of panel p, shifted by (sx, sy)
    ! cells, and the covariant basis a1 = dr/dalpha, a2 = dr/dbeta there
    ! (metric_2d_ecs_mod.f90 L138-160, topology frames as above)
    subroutine point(p, i, j, sx, sy, r, a1, a2)
        integer,      intent(in)            :: p, i, j
        real(kind=8), intent(in)            :: sx, sy
        real(kind=8), intent(out)           :: r(3)
        real(kind=8), intent(out), optional :: a1(3), a2(3)
        real(kind=8) :: X, Y, s, s3, cx, cy
        X = tan(-0.25_8 * pi + (i - 1 + sx) * 0.5_8 * pi / N)
        Y = tan(-0.25_8 * pi + (j - 1 + sy) * 0.5_8 * pi / N)
        s = sqrt(1.0_8 + X * X + Y * Y)
        r = (ex(:, p) * X + ey(:, p) * Y - nn(:, p)) / s
        if (present(a1)) then
            s3 = s * s * s
            cx = 1.0_8 + X * X
            cy = 1.0_8 + Y * Y
            a1 = (ex(:, p) * (cx * cy) - ey(:, p) * (X * Y * cx) + nn(:, p) * (X * cx)) / s3
            a2 = (-ex(:, p) * (X * Y * cy) + ey(:, p) * (cx * cy) + nn(:, p) * (Y * cy)) / s3
        end if
    end subroutine point

    ! sin^2 of the latitude about the rotation axis, from sin(phi) = r . axis
    real(kind=8) function sin2_lat(sp)
        real(kind=8), intent(in) :: sp
        sin2_lat = sin(asin(min(1.0_8, max(-1.0_8, sp))))**2
    end function sin2_lat

    subroutine print_scores(step)
        integer, intent(in) :: step
        integer, parameter :: NREF = 192
        real(kind=8), allocatable :: buffer(:,:,:)
        character(len=256) :: str_buff
        real(kind=8) :: l2, linf, e, time
        integer(kind=8) :: pos, fsize
        integer :: inc, rec, q, ii, jj
        time = step * dt
        if (gauss) then
            if (.not. have_refsol) return
            ! gaussian_hill_linear_test_mod.f90 L71-129: records of 193 x 193 x 6
            ! doubles, one per hour of model time
            rec = nint(time / 3600.0_8)
            inc = NREF / N
            allocate(buffer(NREF + 1, NREF + 1, 6))
            pos = int(rec, 8) * 8_8 * 6_8 * int(NREF + 1, 8)**2 + 1_8
            inquire(file=refsol_path, size=fsize)
            if (fsize < pos - 1 + 8_8 * 6_8 * int(NREF + 1, 8)**2) then
                write(*, '(A,I0,A)') "reference solution has no record ", rec + 1, ": scores not printed"
                return
            end if
            open(unit=12, file=refsol_path, access="stream", form="unformatted", status="old", action="read")
            read(12, pos=pos) buffer
            close(12)
            l2 = 0.0_8
            linf = 0.0_8
            do q = 1, 6
                do jj = 1, N + 1
                    do ii = 1, N + 1
                        e = h(ii, jj, q) - buffer((ii - 1) * inc + 1, (jj - 1) * inc + 1, q)
                        linf = max(abs(e), linf)
                        l2 = l2 + e**2
                    end do
                end do
            end do
            l2 = sqrt(l2 / (6 * (N + 1)**2))
            write(str_buff, *) "Errors, t=", time / 3600.0_8, " hours, l2 = ", l2, ", linf=", linf
        else
            ! Williamson_test2_mod.f90 L115-143
            l2 = sqrt(sum((h - h0)**2 * spread(jac, 3, 6))) / l2_h0
            linf = maxval(abs(h - h0)) / linf_h0
            write(str_buff, *) "Errors, t=", time / 3600.0_8, " hours, l2_h = ", l2, ", linf_h =", linf
        end if
        write(*, '(A)') trim(str_buff)
    end subroutine print_scores

    ! write_fields (swm_output_diag_mod.f90 L52-76) through
    ! latlon_outputer_mod.f90 latlon_write / latlon_vec_write L33-107: record
    ! irec of the direct-access float32 files h.dat, u.dat, v.dat (record
    ! length nlat * nlon values, longitude fastest); the files are started
    ! afresh with record 1
    subroutine write_h(irec)
        integer, intent(in) :: irec
        call operator%latlon_h(state, domain, hll)
        call operator%latlon_uv(state, domain, ull, vll)
        call write_record("h.dat", hll, irec)
        call write_record("u.dat", ull, irec)
        call write_record("v.dat", vll, irec)
    end subroutine write_h

    subroutine write_record(file, f, irec)
        character(len=*), intent(in) :: file
        real(kind=4),     intent(in) :: f(:, :)
        integer,          intent(in) :: irec
        integer :: unit, rl
        inquire(iolength=rl) f
        if (irec == 1) then
            open(newunit=unit, file=file, access="direct", form="unformatted", recl=rl, status="replace")
        else
            open(newunit=unit, file=file, access="direct", form="unformatted", recl=rl, status="old")
        end if
        write(unit, rec=irec) f
        close(unit)
    end subroutine write_record

end program swe_hip_model
