! hmhd_driver.f90 -- the Fortran driver: runs a reference .par file on the
! GPU through libhmhd.so.
!
! The program flow is that of the reference's main program
! (src/mhd_driver.f90:19-112): read the parameters, set the grid and the
! GP tables (hmhd_create), set the initial condition on the host and
! upload it, take the first CFL dt and check it, write the initial output,
! then cfl -> validTimeStep -> advance until sim_tmax or sim_nstepmax,
! printing one line per step and writing outputs at the requested
! frequency (the state comes back from the device only for those), and a
! final output.  Outputs use the build's own file format (hmhd_out.f90)
! in place of HDF5.
!
! Usage: hmhd_driver file.par [--ic-only]
! With --ic-only the driver writes the initial condition as output 0 and
! stops before touching the device.
program hmhd_driver
    use, intrinsic :: iso_c_binding
    use hmhd_c
    use hmhd_par
    use hmhd_ic
    use hmhd_out
    implicit none

    character(len=par_len) :: parfile, base, bcname, riemann, opt
    type(hmhd_config) :: cfg
    type(c_ptr) :: h
    integer(c_int32_t) :: rc, ngc, mx, my
    real(8) :: gamma, tmax, cfl, freq_time, t, last_time
    real(c_double) :: dtc
    real(8) :: dt
    integer :: nstepmax, freq_step, nstep, last_step, counter
    logical :: force_hydro, periodic, out_hdf5, out_ascii
    real(8), allocatable, target :: U(:, :, :), V(:, :, :)
    integer(c_int32_t), allocatable, target :: scheme(:, :)

    call get_command_argument(1, parfile)
    if (len_trim(parfile) == 0) then
        write(*,*) "=============================================================="
        write(*,*) "No parameter file specified."
        write(*,*) "Please specify a parameter file, e.g."
        write(*,*) "./hmhd_driver myfile.par"
        write(*,*) "=============================================================="
        stop
    end if
    write(*,*) "=============================================================="
    write(*,*) "Reading input parameters from ", trim(parfile)
    write(*,*) "=============================================================="

    ! simulation parameters (src/simulation.f90:43-72)
    cfg%small_B = par_real(parfile, 'sim_smallB')
    cfg%small_p = par_real(parfile, 'sim_smallPressure')
    cfg%small_rho = par_real(parfile, 'sim_smallDensity')
    gamma = par_real(parfile, 'sim_gamma')
    riemann = par_char(parfile, 'sim_riemannSolver')
    cfl = par_real(parfile, 'sim_cfl')
    tmax = par_real(parfile, 'sim_tmax')
    nstepmax = par_int(parfile, 'sim_nstepmax')
    bcname = par_char(parfile, 'sim_BC')
    base = par_char(parfile, 'sim_dataFileBaseName')
    freq_time = par_real(parfile, 'sim_outputFreqTime')
    freq_step = par_int(parfile, 'sim_outputFreqStep')
    out_hdf5 = par_logical(parfile, 'sim_outputHdf5')
    out_ascii = par_logical(parfile, 'sim_outputAscii')
    force_hydro = par_logical(parfile, 'sim_forceHydro')
    select case (trim(riemann))
    case ('hllc', 'hllcmhd')
        cfg%riemann = HMHD_RIEMANN_HLLC
    case default
        write(*,*) "==================================================================="
        write(*,*) "Unrecognized choice of riemann solver: ", trim(riemann)
        write(*,*) "Please check that, in your parameter file, the value of"
        write(*,*) "sim_riemannSolver is set to one of hllc, hllcmhd"
        write(*,*) "==================================================================="
        stop
    end select
    select case (trim(bcname))
    case ('periodic')
        periodic = .true.
        cfg%bc = HMHD_BC_PERIODIC
    case ('outflow')
        periodic = .false.
        cfg%bc = HMHD_BC_OUTFLOW
    case default
        write(*,*) "=========================================================================="
        write(*,*) "Unrecognized choice of boundary condition: ", trim(bcname)
        write(*,*) "Please check that, in your parameter file, the value of"
        write(*,*) "sim_BC is set to a valid string, e.g., 'outflow' or 'periodic'"
        write(*,*) "=========================================================================="
        stop
    end select
    if (out_ascii) then
        write(*,*) 'Ascii output not supported right now'
        stop
    end if
    cfg%gamma = gamma
    cfg%cfl = cfl
    cfg%force_hydro = merge(1, 0, force_hydro)

    ! GP and grid parameters (src/GP.f90:58-69, src/grid.f90:58-63)
    cfg%R = par_int(parfile, "GP_maxRadius")
    cfg%ell_over_delta = par_real(parfile, "GP_ellOverDelta")
    cfg%nx = par_int(parfile, "grid_Nx")
    cfg%ny = par_int(parfile, "grid_Ny")
    cfg%xbeg = par_real(parfile, "grid_xBeg")
    cfg%xend = par_real(parfile, "grid_xEnd")
    cfg%ybeg = par_real(parfile, "grid_yBeg")
    cfg%yend = par_real(parfile, "grid_yEnd")

    ! NGC = GP_maxRadius + 2 guard cells per side (src/grid.f90:71)
    ngc = cfg%R + 2
    mx = cfg%nx + 2 * ngc
    my = cfg%ny + 2 * ngc
    call get_command_argument(2, opt)
    allocate(U(8, mx, my), V(10, mx, my), scheme(mx, my))
    scheme = 0
    if (trim(opt) == '--ic-only') then
        call ic_set(parfile, cfg%nx, cfg%ny, ngc, cfg%xbeg, cfg%xend, cfg%ybeg, cfg%yend, gamma, force_hydro, &
                    periodic, U, V)
        call output_file(trim(base)//'_00000.hmhd', 0, 0.0d0, 0.0d0, 0, int(cfg%nx), int(cfg%ny), int(ngc), &
                         cfg%xbeg, cfg%xend, cfg%ybeg, cfg%yend, U, V, scheme)
    else
        call simulate()
    end if

contains

    subroutine simulate()
        ! the run itself: handle, upload, the main loop, outputs
        rc = hmhd_create(cfg, h)
        if (rc /= HMHD_OK) then
            write(*,*) trim(hmhd_error_message(c_null_ptr))
            ! configuration errors stop as the reference does; the rest are internal
            if (rc == HMHD_ERR_ARG) stop
            error stop 2
        end if

        ! initial condition (src/initialCondition.f90:23-84)
        call ic_set(parfile, cfg%nx, cfg%ny, ngc, cfg%xbeg, cfg%xend, cfg%ybeg, cfg%yend, gamma, force_hydro, &
                    periodic, U, V)
        call check(hmhd_upload_state(h, c_loc(U), c_loc(V)))

        ! the main loop (src/mhd_driver.f90:64-101)
        nstep = 0
        t = 0.0d0
        call check(hmhd_cfl_dt(h, dtc))
        dt = dtc
        call valid_time_step(dt, t, tmax)
        last_step = 0
        last_time = 0.0d0
        counter = 0
        call write_output(.true.)
        do while (t < tmax .and. nstep < nstepmax)
            call check(hmhd_cfl_dt(h, dtc))
            dt = dtc
            call valid_time_step(dt, t, tmax)
            call check(hmhd_advance(h, dt))
            t = t + dt
            nstep = nstep + 1
            write(*, '(1x,i5,f16.8,1x,f16.8)') nstep, t, dt
            call write_output(.false.)
        end do
        if (last_step < nstep) call write_output(.true.)

        call hmhd_destroy(h)
        write(*,*) "=============================================================="
        write(*,*) "Simulation has ended."
        write(*,*) "=============================================================="
    end subroutine simulate

    subroutine check(code)
        integer(c_int32_t), intent(in) :: code

        if (code /= HMHD_OK) then
            write(*,*) trim(hmhd_error_message(h))
            error stop 2
        end if
    end subroutine check

    ! src/mhd_driver.f90:117-144
    subroutine valid_time_step(delta_t, current_time, t_end)
        real(8), intent(inout) :: delta_t
        real(8), intent(in) :: current_time, t_end

        if (current_time + delta_t > t_end) then
            delta_t = t_end - current_time
        else if (delta_t < 0.0d0) then
            write(*,*) "=========================================================================="
            write(*,*) "dt is negative. Aborting simulation."
            write(*,*) "=========================================================================="
            stop
        else if (delta_t /= delta_t) then
            write(*,*) "=========================================================================="
            write(*,*) "dt is NAN. Aborting simulation."
            write(*,*) "=========================================================================="
            stop
        else if (delta_t - delta_t /= 0.0d0) then
            write(*,*) "=========================================================================="
            write(*,*) "dt is infinite. Aborting simulation."
            write(*,*) "=========================================================================="
            stop
        end if
    end subroutine valid_time_step

    ! src/output.f90:28-112
    subroutine write_output(force)
        logical, intent(in) :: force
        character(len=5) :: tag

        if (.not. output_due(nstep, t, last_step, last_time, freq_step, freq_time, force)) return
        write(tag, '(i5.5)') counter
        if (out_hdf5) then
            if (nstep > 0) call check(hmhd_download(h, c_loc(U), c_loc(V), c_loc(scheme)))
            call output_file(trim(base)//'_'//tag//'.hmhd', counter, t, dt, nstep, int(cfg%nx), int(cfg%ny), &
                             int(ngc), cfg%xbeg, cfg%xend, cfg%ybeg, cfg%yend, U, V, scheme)
        end if
        write(*,*) ''
        write(*,*) ' Output no.', counter, 'has been written      '
        write(*,*) '================================================='
        write(*,*) '   Steps      Time              dt               '
        write(*,*) '================================================='
        write(*,*) ''
        last_step = nstep
        last_time = counter * freq_time
        counter = counter + 1
    end subroutine write_output

end program hmhd_driver
