! hmhd_ic.f90 -- initial conditions of the Fortran driver.
!
! The reference's three (src/initialCondition.f90:86-226): explosion2d,
! sedov2d and OrszagTang2d, read from the same .par parameters and set as
! point values at cell centres; and three the reference cannot express,
! with the workloads' parameters:
!   OrszagTang2dWorkload  rho = gamma^2, p = gamma, v = (-sin 2 pi y,
!                         sin 2 pi x), B = (-sin 2 pi y, sin 4 pi x)/sqrt(4 pi)
!   mhdBlast2d            rho = 1, p = IC_presIns for r < IC_shockRad else
!                         IC_presOut, v = 0, B = (IC_magx, IC_magy, IC_magz)
!   sedov2dCells          rho = 1, v = 0, p = IC_presOut, B = (IC_magx,
!                         IC_magy, IC_magz), and IC_energy of thermal energy
!                         deposited in the 3 x 3 interior cells centred on
!                         cell (Nx/2, Ny/2) (0-based)
! With IC_noiseSeed = s (a non-negative integer) in the .par file, rho, v,
! p and B of the interior are then multiplied by (1 + IC_noiseAmp u), with
! IC_noiseAmp 0.01 unless given and u in [-1, 1) from a counter-based hash
! (noise_u below), the same values hosono_2dmhd_amd.ics.perturb uses.
! Then, as initialCondition_set does (src/initialCondition.f90:62-78): eint
! and gamma on the interior, U = prim2cons(V) (src/convert.f90:159-196) and
! the guard-cell fills of U and V (src/boundaryConditions.f90:118-187).
! Arrays: U(8, mx, my), V(10, mx, my), 1-based, NGC guard cells per side.
module hmhd_ic
    use hmhd_par
    implicit none
    private

    real(8), parameter :: pi = acos(-1.0d0)
    integer, parameter :: dens_v = 1, velx_v = 2, vely_v = 3, velz_v = 4, pres_v = 5, &
                          magx_v = 6, magy_v = 7, magz_v = 8, eint_v = 9, gamm_v = 10
    integer(8), parameter :: mask32 = 4294967295_8

    public :: ic_set, prim2cons, bc_fill, hash32

contains

    subroutine ic_set(parfile, nx, ny, ngc, xbeg, xend, ybeg, yend, gamma, force_hydro, periodic, U, V)
        character(len=*), intent(in) :: parfile
        integer, intent(in) :: nx, ny, ngc
        real(8), intent(in) :: xbeg, xend, ybeg, yend, gamma
        logical, intent(in) :: force_hydro, periodic
        real(8), intent(out) :: U(:, :, :), V(:, :, :)
        character(len=par_len) :: ic_type
        real(8) :: x(nx + 2 * ngc), y(ny + 2 * ngc), dx, dy
        integer :: i, j, s, ex, ey

        dx = (xend - xbeg) / nx
        dy = (yend - ybeg) / ny
        do i = 1, nx + 2 * ngc
            x(i) = (i - ngc - 0.5d0) * dx + xbeg
        end do
        do j = 1, ny + 2 * ngc
            y(j) = (j - ngc - 0.5d0) * dy + ybeg
        end do
        s = ngc + 1
        ex = ngc + nx
        ey = ngc + ny
        V = 0.0d0
        U = 0.0d0

        ic_type = par_char(parfile, "IC_type")
        select case (trim(ic_type))
        case ("explosion2d")
            call explosion2d(parfile, x, y, s, ex, ey, V)
        case ("sedov2d")
            call sedov2d(parfile, x, y, s, ex, ey, min(dx, dy), gamma, V)
        case ("OrszagTang2d")
            call orszag_tang2d(parfile, x, y, s, ex, ey, gamma, V)
        case ("OrszagTang2dWorkload")
            call orszag_tang_workload(x, y, s, ex, ey, gamma, V)
        case ("mhdBlast2d")
            call mhd_blast(parfile, x, y, s, ex, ey, V)
        case ("sedov2dCells")
            call sedov_cells(parfile, nx, ny, ngc, dx, dy, gamma, V)
        case default
            write(*,*) "=========================================================================="
            write(*,*) "Unrecognized choice of initial condition: ", trim(ic_type)
            write(*,*) "Please check that, in your parameter file, the value of"
            write(*,*) "IC_type is set to a valid string, e.g., 'OrszagTang2D'"
            write(*,*) "=========================================================================="
            stop
        end select
        if (par_has(parfile, "IC_noiseSeed")) call add_noise(parfile, nx, s, ex, ey, V)

        do j = s, ey
            do i = s, ex
                V(eint_v, i, j) = V(pres_v, i, j) / (gamma - 1.0d0) / V(dens_v, i, j)
                V(gamm_v, i, j) = gamma
                call prim2cons(V(:, i, j), gamma, force_hydro, U(:, i, j))
            end do
        end do
        call bc_fill(U, nx, ny, ngc, periodic)
        call bc_fill(V, nx, ny, ngc, periodic)
    end subroutine ic_set

    ! rho, v, p and B of the interior times (1 + amp u), u = noise_u(seed, k, c)
    ! for the k-th of those eight variables (k = 0..7) and interior cell
    ! c = jj nx + ii (0-based)
    subroutine add_noise(parfile, nx, s, ex, ey, V)
        character(len=*), intent(in) :: parfile
        integer, intent(in) :: nx, s, ex, ey
        real(8), intent(inout) :: V(:, :, :)
        integer :: seed, k, i, j
        integer(8) :: base, c
        real(8) :: amp, u

        seed = par_int(parfile, "IC_noiseSeed")
        if (seed < 0) then
            write(*,*) "IC_noiseSeed must be a non-negative integer"
            stop
        end if
        amp = 0.01d0
        if (par_has(parfile, "IC_noiseAmp")) amp = par_real(parfile, "IC_noiseAmp")
        do k = 0, 7
            base = hash32(iand(8_8 * seed + k, mask32))
            do j = s, ey
                do i = s, ex
                    c = int(j - s, 8) * nx + (i - s)
                    u = real(hash32(iand(c + base, mask32)), 8) * 2.0d0**(-31) - 1.0d0
                    V(dens_v + k, i, j) = V(dens_v + k, i, j) * (1.0d0 + amp * u)
                end do
            end do
        end do
    end subroutine add_noise

    ! lowbias32, C. Wellons' 32-bit integer hash ("Prospecting for hash
    ! functions", 2018): x ^= x >> 16; x *= 0x7feb352d; x ^= x >> 15;
    ! x *= 0x846ca68b; x ^= x >> 16, modulo 2^32, in 64-bit integers
    integer(8) function hash32(x0)
        integer(8), intent(in) :: x0
        integer(8) :: x

        x = iand(x0, mask32)
        x = ieor(x, shiftr(x, 16))
        x = mulmod32(x, 2146121005_8)
        x = ieor(x, shiftr(x, 15))
        x = mulmod32(x, 2221713035_8)
        hash32 = ieor(x, shiftr(x, 16))
    end function hash32

    ! x c modulo 2^32 for x, c in [0, 2^32), from x c = x c_hi 2^16 + x c_lo:
    ! every product stays below 2^48
    integer(8) function mulmod32(x, c)
        integer(8), intent(in) :: x, c
        integer(8) :: t

        t = iand(x * shiftr(c, 16), 65535_8)
        mulmod32 = iand(shiftl(t, 16) + x * iand(c, 65535_8), mask32)
    end function mulmod32

    ! src/convert.f90:159-196
    subroutine prim2cons(Vc, gamma, force_hydro, Uc)
        real(8), intent(in) :: Vc(:), gamma
        logical, intent(in) :: force_hydro
        real(8), intent(out) :: Uc(:)
        real(8) :: Bp, rho_eint, E

        if (force_hydro) then
            Bp = 0.0d0
        else
            Bp = 0.5d0 * (Vc(magx_v) * Vc(magx_v) + Vc(magy_v) * Vc(magy_v) + Vc(magz_v) * Vc(magz_v))
        end if
        rho_eint = Vc(pres_v) / (gamma - 1.0d0)
        E = rho_eint + Vc(dens_v) * (0.5d0 * (Vc(velx_v) * Vc(velx_v) + Vc(vely_v) * Vc(vely_v) &
                                              + Vc(velz_v) * Vc(velz_v))) + Bp
        Uc(1) = Vc(dens_v)
        Uc(2) = Vc(dens_v) * Vc(velx_v)
        Uc(3) = Vc(dens_v) * Vc(vely_v)
        Uc(4) = Vc(dens_v) * Vc(velz_v)
        if (force_hydro) then
            Uc(6:8) = 0.0d0
        else
            Uc(6:8) = Vc(magx_v:magz_v)
        end if
        Uc(5) = E
    end subroutine prim2cons

    ! guard cells: outflow copies the nearest interior cell, periodic wraps
    ! by N; x bands over all rows, then y bands over all columns
    subroutine bc_fill(A, nx, ny, ngc, periodic)
        real(8), intent(inout) :: A(:, :, :)
        integer, intent(in) :: nx, ny, ngc
        logical, intent(in) :: periodic
        integer :: i, j, src

        do j = 1, ny + 2 * ngc
            do i = 1, ngc
                src = merge(i + nx, ngc + 1, periodic)
                A(:, i, j) = A(:, src, j)
                src = merge(ngc + nx + i - nx, ngc + nx, periodic)
                A(:, ngc + nx + i, j) = A(:, src, j)
            end do
        end do
        do j = 1, ngc
            do i = 1, nx + 2 * ngc
                src = merge(j + ny, ngc + 1, periodic)
                A(:, i, j) = A(:, i, src)
                src = merge(ngc + ny + j - ny, ngc + ny, periodic)
                A(:, i, ngc + ny + j) = A(:, i, src)
            end do
        end do
    end subroutine bc_fill

    ! src/initialCondition.f90:86-136
    subroutine explosion2d(parfile, x, y, s, ex, ey, V)
        character(len=*), intent(in) :: parfile
        real(8), intent(in) :: x(:), y(:)
        integer, intent(in) :: s, ex, ey
        real(8), intent(inout) :: V(:, :, :)
        real(8) :: cx, cy, rad, ins(5), out(5)
        integer :: i, j

        cx = par_real(parfile, "IC_shockCenter_x")
        cy = par_real(parfile, "IC_shockCenter_y")
        rad = par_real(parfile, "IC_shockRad")
        ins(1) = par_real(parfile, "IC_densIns")
        out(1) = par_real(parfile, "IC_densOut")
        ins(2) = par_real(parfile, "IC_velxIns")
        ins(3) = par_real(parfile, "IC_velyIns")
        ins(4) = par_real(parfile, "IC_velzIns")
        out(2) = par_real(parfile, "IC_velxOut")
        out(3) = par_real(parfile, "IC_velyOut")
        out(4) = par_real(parfile, "IC_velzOut")
        ins(5) = par_real(parfile, "IC_presIns")
        out(5) = par_real(parfile, "IC_presOut")
        do j = s, ey
            do i = s, ex
                if (sqrt((x(i) - cx) * (x(i) - cx) + (y(j) - cy) * (y(j) - cy)) <= rad) then
                    V(1:5, i, j) = ins
                else
                    V(1:5, i, j) = out
                end if
            end do
        end do
    end subroutine explosion2d

    ! src/initialCondition.f90:138-187 (disc of radius 3.5 min(dx, dy))
    subroutine sedov2d(parfile, x, y, s, ex, ey, mindl, gamma, V)
        character(len=*), intent(in) :: parfile
        real(8), intent(in) :: x(:), y(:), mindl, gamma
        integer, intent(in) :: s, ex, ey
        real(8), intent(inout) :: V(:, :, :)
        real(8) :: cx, cy, rad, p_in
        integer :: i, j

        cx = par_real(parfile, "IC_shockCenter_x")
        cy = par_real(parfile, "IC_shockCenter_y")
        rad = 3.5d0 * mindl
        p_in = 3 * (gamma - 1) * 1.0d0 / ((2 + 1) * pi * (rad * rad))
        do j = s, ey
            do i = s, ex
                V(dens_v, i, j) = 1.0d0
                if (sqrt((x(i) - cx) * (x(i) - cx) + (y(j) - cy) * (y(j) - cy)) <= rad) then
                    V(pres_v, i, j) = p_in
                else
                    V(pres_v, i, j) = 1d-5
                end if
            end do
        end do
    end subroutine sedov2d

    ! src/initialCondition.f90:189-226 (p = B0 = 1/gamma)
    subroutine orszag_tang2d(parfile, x, y, s, ex, ey, gamma, V)
        character(len=*), intent(in) :: parfile
        real(8), intent(in) :: x(:), y(:), gamma
        integer, intent(in) :: s, ex, ey
        real(8), intent(inout) :: V(:, :, :)
        real(8) :: u0, dens, b0, pres
        integer :: i, j

        u0 = par_real(parfile, "IC_U0")
        dens = par_real(parfile, "IC_dens")
        b0 = 1 / gamma
        pres = 1 / gamma
        do j = s, ey
            do i = s, ex
                V(dens_v, i, j) = dens
                V(velx_v, i, j) = -u0 * sin(pi * y(j) * 2.0d0)
                V(vely_v, i, j) = u0 * sin(pi * x(i) * 2.0d0)
                V(magx_v, i, j) = -b0 * sin(pi * y(j) * 2.0d0)
                V(magy_v, i, j) = b0 * sin(pi * x(i) * 4.0d0)
                V(pres_v, i, j) = pres
            end do
        end do
    end subroutine orszag_tang2d

    subroutine orszag_tang_workload(x, y, s, ex, ey, gamma, V)
        real(8), intent(in) :: x(:), y(:), gamma
        integer, intent(in) :: s, ex, ey
        real(8), intent(inout) :: V(:, :, :)
        real(8) :: b0
        integer :: i, j

        b0 = 1.0d0 / sqrt(4.0d0 * pi)
        do j = s, ey
            do i = s, ex
                V(dens_v, i, j) = gamma * gamma
                V(velx_v, i, j) = -sin(pi * y(j) * 2.0d0)
                V(vely_v, i, j) = sin(pi * x(i) * 2.0d0)
                V(magx_v, i, j) = -b0 * sin(pi * y(j) * 2.0d0)
                V(magy_v, i, j) = b0 * sin(pi * x(i) * 4.0d0)
                V(pres_v, i, j) = gamma
            end do
        end do
    end subroutine orszag_tang_workload

    subroutine mhd_blast(parfile, x, y, s, ex, ey, V)
        character(len=*), intent(in) :: parfile
        real(8), intent(in) :: x(:), y(:)
        integer, intent(in) :: s, ex, ey
        real(8), intent(inout) :: V(:, :, :)
        real(8) :: rad, p_in, p_out, b(3)
        integer :: i, j

        rad = par_real(parfile, "IC_shockRad")
        p_in = par_real(parfile, "IC_presIns")
        p_out = par_real(parfile, "IC_presOut")
        b(1) = par_real(parfile, "IC_magx")
        b(2) = par_real(parfile, "IC_magy")
        b(3) = par_real(parfile, "IC_magz")
        do j = s, ey
            do i = s, ex
                V(dens_v, i, j) = 1.0d0
                if (sqrt(x(i) * x(i) + y(j) * y(j)) < rad) then
                    V(pres_v, i, j) = p_in
                else
                    V(pres_v, i, j) = p_out
                end if
                V(magx_v:magz_v, i, j) = b
            end do
        end do
    end subroutine mhd_blast

    subroutine sedov_cells(parfile, nx, ny, ngc, dx, dy, gamma, V)
        character(len=*), intent(in) :: parfile
        integer, intent(in) :: nx, ny, ngc
        real(8), intent(in) :: dx, dy, gamma
        real(8), intent(inout) :: V(:, :, :)
        real(8) :: energy, p_bg, b(3), e_dens
        integer :: i, j, ci, cj

        energy = par_real(parfile, "IC_energy")
        p_bg = par_real(parfile, "IC_presOut")
        b(1) = par_real(parfile, "IC_magx")
        b(2) = par_real(parfile, "IC_magy")
        b(3) = par_real(parfile, "IC_magz")
        do j = ngc + 1, ngc + ny
            do i = ngc + 1, ngc + nx
                V(dens_v, i, j) = 1.0d0
                V(pres_v, i, j) = p_bg
                V(magx_v:magz_v, i, j) = b
            end do
        end do
        ! 1-based index of interior cell (Nx/2, Ny/2), 0-based
        ci = ngc + nx / 2 + 1
        cj = ngc + ny / 2 + 1
        e_dens = energy / (9.0d0 * dx * dy)
        V(pres_v, ci - 1:ci + 1, cj - 1:cj + 1) = (gamma - 1.0d0) * e_dens
    end subroutine sedov_cells

end module hmhd_ic
