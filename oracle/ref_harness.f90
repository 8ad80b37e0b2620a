! ref_harness.f90 -- C-callable harness around the reference's own Fortran
! modules, compiled from the reference sources by oracle/Makefile into
! oracle/_ref/.  TEST INFRASTRUCTURE ONLY: it lets tests/ and bench.py's
! cpu_baseline leg run the reference's hot path (src/cfl.f90:14-52 and
! src/timeStep.f90:22-171) on inputs the tests choose, and set the
! reference's initial conditions (src/initialCondition.f90:23-84).  Nothing in the
! product library links or calls it.
!
! Layouts match the reference: U(8, 1:mx, 1:my), V(10, 1:mx, 1:my),
! scheme(1:mx, 1:my), faceVals(8, nq, 4, 1:mx, 1:my), with mx = Nx + 2*NGC.
module ref_harness

    use, intrinsic :: iso_c_binding
    use definitions, only: max_string_length, nConsVars, nPrimVars, xdir, ydir
    use simulation, only: simulation_init, sim_BC
    use GP, only: GP_init, GP_finalize, GP_maxRadius, GP_nQuadratureMax, GP_nStencMax, &
        GP_predictionVectors, GP_quadratureWeights, GP_stencIdxs
    use grid, only: grid_init, grid_finalize, grid_block, grid_state
    use cfl, only: cfl_computedt
    use timeStep, only: timeStep_advanceSolution
    use reconstruct, only: reconstruct_faceValsWithGP
    use getFluxes, only: getFluxes_
    use FLICDdivB, only: FLICDdivB_fluxCorrections
    use detect, only: detect_ALL
    use boundaryConditions, only: boundaryConditions_applyBlk, boundaryConditions_applyMOOD
    use initialCondition, only: initialCondition_set

    implicit none

    logical, save :: ref_ready = .false.

contains

    subroutine c2f_string(cstr, fstr)
        character(kind=c_char), intent(in) :: cstr(*)
        character(len=*), intent(out) :: fstr
        integer :: k
        fstr = ' '
        k = 1
        do while (cstr(k) /= c_null_char .and. k <= len(fstr))
            fstr(k:k) = cstr(k)
            k = k + 1
        end do
    end subroutine c2f_string

    ! Read a .par file exactly as mhd_driver does (src/mhd_driver.f90:40-52).
    subroutine ref_init(parfile) bind(C, name='ref_init')
        character(kind=c_char), intent(in) :: parfile(*)
        character(len=max_string_length) :: pf
        call c2f_string(parfile, pf)
        if (ref_ready) call ref_finalize()
        call simulation_init(pf)
        call GP_init(pf)
        call grid_init(pf, GP_maxRadius, GP_nQuadratureMax)
        ref_ready = .true.
    end subroutine ref_init

    ! src/initialCondition.f90:23-84 on the grid ref_init set up: IC_type
    ! and its IC_* parameters read from the same .par file
    subroutine ref_initial_condition(parfile) bind(C, name='ref_initial_condition')
        character(kind=c_char), intent(in) :: parfile(*)
        character(len=max_string_length) :: pf
        call c2f_string(parfile, pf)
        call initialCondition_set(pf, grid_block, grid_state)
    end subroutine ref_initial_condition

    subroutine ref_finalize() bind(C, name='ref_finalize')
        if (ref_ready) then
            call grid_finalize()
            call GP_finalize()
        end if
        ref_ready = .false.
    end subroutine ref_finalize

    ! flush the Fortran runtime's buffered stdout (the reference's banners)
    subroutine ref_flush() bind(C, name='ref_flush')
        flush(6)
    end subroutine ref_flush

    subroutine ref_dims(ngc, mx, my, nsmax, nqmax, rmax) bind(C, name='ref_dims')
        integer(c_int), intent(out) :: ngc, mx, my, nsmax, nqmax, rmax
        ngc = grid_block%NGC
        mx = grid_block%maxIdx(xdir)
        my = grid_block%maxIdx(ydir)
        nsmax = GP_nStencMax
        nqmax = GP_nQuadratureMax
        rmax = GP_maxRadius
    end subroutine ref_dims

    subroutine ref_gp_tables(pv, qw, st) bind(C, name='ref_gp_tables')
        real(c_double), intent(out) :: pv(GP_nStencMax, GP_nQuadratureMax, 4, GP_maxRadius)
        real(c_double), intent(out) :: qw(GP_nQuadratureMax, GP_maxRadius)
        integer(c_int), intent(out) :: st(2, GP_nStencMax, GP_maxRadius)
        pv = GP_predictionVectors
        qw = GP_quadratureWeights
        st = GP_stencIdxs
    end subroutine ref_gp_tables

    subroutine ref_set_state(U, V) bind(C, name='ref_set_state')
        real(c_double), intent(in) :: U(nConsVars, grid_block%maxIdx(xdir), grid_block%maxIdx(ydir))
        real(c_double), intent(in) :: V(nPrimVars, grid_block%maxIdx(xdir), grid_block%maxIdx(ydir))
        grid_state%U = U
        grid_state%V = V
    end subroutine ref_set_state

    subroutine ref_get_state(U, V, scheme) bind(C, name='ref_get_state')
        real(c_double), intent(out) :: U(nConsVars, grid_block%maxIdx(xdir), grid_block%maxIdx(ydir))
        real(c_double), intent(out) :: V(nPrimVars, grid_block%maxIdx(xdir), grid_block%maxIdx(ydir))
        integer(c_int), intent(out) :: scheme(grid_block%maxIdx(xdir), grid_block%maxIdx(ydir))
        U = grid_state%U
        V = grid_state%V
        scheme = grid_block%scheme
    end subroutine ref_get_state

    ! src/cfl.f90:14-52 on the current state
    function ref_cfl_dt() bind(C, name='ref_cfl_dt') result(dt)
        real(c_double) :: dt
        dt = cfl_computedt(grid_block, grid_state)
    end function ref_cfl_dt

    ! src/timeStep.f90:22-43 on the current state
    subroutine ref_advance(dt) bind(C, name='ref_advance')
        real(c_double), value :: dt
        call timeStep_advanceSolution(dt, grid_block, grid_state)
    end subroutine ref_advance

    ! Advance nsteps full steps with the driver's dt rule
    ! (src/mhd_driver.f90:76-95 without output); returns the final time.
    function ref_run(nsteps, tmax) bind(C, name='ref_run') result(t)
        integer(c_int), value :: nsteps
        real(c_double), value :: tmax
        real(c_double) :: t, dt
        integer :: n
        t = 0.0
        n = 0
        do while ((t < tmax) .and. (n < nsteps))
            dt = cfl_computedt(grid_block, grid_state)
            if (t + dt > tmax) dt = tmax - t
            call timeStep_advanceSolution(dt, grid_block, grid_state)
            t = t + dt
            n = n + 1
        end do
    end function ref_run

    ! reconstruct_faceValsWithGP with every cell at `radius`, as in
    ! test/test_reconstruct_faceValsWithGP.f90:24-34
    subroutine ref_reconstruct_all(radius, fv) bind(C, name='ref_reconstruct_all')
        integer(c_int), value :: radius
        real(c_double), intent(out) :: fv(nConsVars, GP_nQuadratureMax, 4, &
                                          grid_block%maxIdx(xdir), grid_block%maxIdx(ydir))
        grid_block%scheme = radius
        grid_block%recompute = .true.
        grid_block%recomputeFLICD = .true.
        grid_block%faceVals = 0.0
        call reconstruct_faceValsWithGP(grid_block, grid_state%U)
        fv = grid_block%faceVals
    end subroutine ref_reconstruct_all

    ! One SSP-RK3 step re-assembled from the reference's public routines
    ! (same sequence as src/timeStep.f90:45-171), recording the final
    ! scheme of each stage and the MOOD pass count.  Leaves the state
    ! untouched; returns the stepped U in Uout so callers can check it
    ! against ref_advance bit for bit.
    subroutine ref_step_masks(dt, Uout, schemes, passes) bind(C, name='ref_step_masks')
        real(c_double), value :: dt
        real(c_double), intent(out) :: Uout(nConsVars, grid_block%maxIdx(xdir), grid_block%maxIdx(ydir))
        integer(c_int), intent(out) :: schemes(grid_block%maxIdx(xdir), grid_block%maxIdx(ydir), 3)
        integer(c_int), intent(out) :: passes(3)
        real(c_double), allocatable :: U(:,:,:), U1(:,:,:), U2(:,:,:), U3(:,:,:)
        allocate(U, source=grid_state%U)
        allocate(U1, U2, U3, mold=grid_state%U)
        U1 = 0.0; U2 = 0.0; U3 = 0.0
        call stage(dt, U, U1, 1)
        call boundaryConditions_applyBlk(grid_block, U1)
        call stage(dt, U1, U2, 2)
        call boundaryConditions_applyBlk(grid_block, U2)
        U2 = (3*U + U2)/4.0
        call boundaryConditions_applyBlk(grid_block, U2)
        call stage(dt, U2, U3, 3)
        call boundaryConditions_applyBlk(grid_block, U3)
        Uout = (U + 2*U3)/3.0
        call boundaryConditions_applyBlk(grid_block, Uout)
    contains
        subroutine stage(dt_, Uin, Unew, s)
            real(c_double), intent(in) :: dt_
            real(c_double), intent(in) :: Uin(:,:,:)
            real(c_double), intent(inout) :: Unew(:,:,:)
            integer, intent(in) :: s
            logical :: done
            integer :: i, j, io, jo
            real(c_double) :: dx, dy
            dx = grid_block%dl(xdir)
            dy = grid_block%dl(ydir)
            io = 1 - grid_block%minIdx(xdir)
            jo = 1 - grid_block%minIdx(ydir)
            grid_block%scheme = GP_maxRadius
            grid_block%recompute = .true.
            grid_block%recomputeFLICD = .true.
            passes(s) = 0
            done = .false.
            do while (.not. done)
                passes(s) = passes(s) + 1
                call reconstruct_faceValsWithGP(grid_block, Uin)
                call getFluxes_(grid_block)
                call FLICDdivB_fluxCorrections(grid_block)
                do j = grid_block%strtIdx(ydir), grid_block%stopIdx(ydir)
                    do i = grid_block%strtIdx(xdir), grid_block%stopIdx(xdir)
                        if (grid_block%recomputeFLICD(i,j)) then
                            Unew(:,i+io,j+jo) = Uin(:,i+io,j+jo) &
                                - (dt_/dx)*(grid_block%flux(:,xdir,i+1,j) - grid_block%flux(:,xdir,i,j)) &
                                - (dt_/dy)*(grid_block%flux(:,ydir,i,j+1) - grid_block%flux(:,ydir,i,j))
                        end if
                    end do
                end do
                call detect_ALL(grid_block, Uin, Unew, done)
                call boundaryConditions_applyMOOD(grid_block)
            end do
            schemes(:,:,s) = grid_block%scheme
        end subroutine stage
    end subroutine ref_step_masks

end module ref_harness
