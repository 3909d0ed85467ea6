! Gaussian-hill linear SWE run on the GPU through the Fortran host layer: the
! counterpart of SHALLOW_WATER_MAIN with testcase 'Gaussian_hill_linear' and
! timescheme 'rk4' (models/shallow_water/shallow_water_mod.f90 L29-145,
! testcases/gaussian_hill_linear_test_mod.f90 L30-67).
!
! usage: swe_hip_gaussian N order precision(0 fp64|1 fp32) nsteps dt centre(0|1|2) out.bin
!   centre 0: panel interior (lam0 = pi, phi0 = 0); 1: cube corner
!   (lam0 = pi/4, phi0 = asin(1/sqrt(3))); 2: 1 cell off the corner
!   (lam0 = 0.27 pi, phi0 = asin(1/sqrt(3)) - 0.02 pi) -- linear_gauss_run.sh L41-52.
! Physical constants: a = 6.371e6 m, g = 9.81, H0 = 1000 m, Omega = 7.292e-5,
! f = 2 Omega sin(phi); h0 = 10 m exp(-(d / 500 km)^2), u = v = 0.
! Writes the final (u, v, h) in the packed layout of include/ch_swe.h.
program swe_hip_gaussian

    use swe_hip_mod

    implicit none

    real(kind=8), parameter :: pi = acos(-1.0_8)
    real(kind=8), parameter :: radius = 6.371e6_8, grav = 9.81_8, h_mean = 1000.0_8, omega = 7.292e-5_8
    ! panel frames: topology/cubed_sphere_topology_mod.f90 L29-55
    integer, parameter :: ex(3, 6) = reshape([-1, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, 1, 1, 0, 0, 1, 0, 0], [3, 6])
    integer, parameter :: ey(3, 6) = reshape([0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 0, 1, 0, 0, -1], [3, 6])
    integer, parameter :: nn(3, 6) = reshape([0, 0, 1, -1, 0, 0, 0, 0, -1, 1, 0, 0, 0, 1, 0, 0, -1, 0], [3, 6])

    type(domain_hip_t)           :: domain
    type(operator_swm_lin_hip_t) :: operator
    type(rk4_hip_t)              :: timescheme
    type(stvec_hip_t)            :: state
    real(kind=8), allocatable    :: u(:,:,:), v(:,:,:), h(:,:,:)
    character(len=256)           :: arg, out_file
    integer :: N, order, precision, nsteps, centre, it, p, i, j, every
    real(kind=8) :: dt, lam0, phi0, d0, c(3), r(3), al, be, X, Y, s, dist, mass0, energy0, mass, energy, hx

    call get_command_argument(1, arg); read(arg, *) N
    call get_command_argument(2, arg); read(arg, *) order
    call get_command_argument(3, arg); read(arg, *) precision
    call get_command_argument(4, arg); read(arg, *) nsteps
    call get_command_argument(5, arg); read(arg, *) dt
    call get_command_argument(6, arg); read(arg, *) centre
    call get_command_argument(7, out_file)

    select case (centre)
    case (0)
        lam0 = pi; phi0 = 0.0_8
    case (1)
        lam0 = 0.25_8 * pi; phi0 = asin(1.0_8 / sqrt(3.0_8))
    case default
        lam0 = 0.27_8 * pi; phi0 = asin(1.0_8 / sqrt(3.0_8)) - 0.02_8 * pi
    end select
    d0 = 5.0e5_8 / radius
    c = [cos(phi0) * cos(lam0), cos(phi0) * sin(lam0), sin(phi0)]

    allocate(u(N, N + 1, 6), v(N + 1, N, 6), h(N + 1, N + 1, 6))
    u = 0.0_8
    v = 0.0_8
    hx = 0.5_8 * pi / N
    do p = 1, 6
        do j = 1, N + 1
            be = -0.25_8 * pi + (j - 1) * hx
            do i = 1, N + 1
                al = -0.25_8 * pi + (i - 1) * hx
                X = tan(al); Y = tan(be); s = sqrt(1.0_8 + X * X + Y * Y)
                r = (ex(:, p) * X + ey(:, p) * Y - nn(:, p)) / s
                dist = acos(min(1.0_8, max(-1.0_8, sum(r * c))))
                h(i, j, p) = 10.0_8 * exp(-(dist / d0)**2)
            end do
        end do
    end do

    domain = create_domain_hip(N, order, precision, radius, grav, h_mean, omega, .false.)
    operator = create_linear_swm_operator_hip()
    timescheme = create_rk4_hip()
    call create_stvec_hip(state, domain)
    call state%upload(u, v, h)

    call operator%integrals(state, domain, mass0, energy0)
    write(*, '(A,ES24.16,A,ES24.16)') "step 0 mass = ", mass0, " energy = ", energy0
    every = max(1, nsteps / 4)
    do it = 1, nsteps
        call timescheme%step(state, operator, domain, dt)
        if (mod(it, every) == 0 .or. it == nsteps) then
            call operator%integrals(state, domain, mass, energy)
            write(*, '(A,I6,A,ES24.16,A,ES24.16)') "step ", it, " mass = ", mass, " energy = ", energy
        end if
    end do
    call state%download(u, v, h)

    open(unit=11, file=trim(out_file), access="stream", form="unformatted", status="replace")
    write(11) u, v, h
    close(11)

    call state%destroy()
    call domain%destroy()
    write(*, '(A)') "done"

end program swe_hip_gaussian
