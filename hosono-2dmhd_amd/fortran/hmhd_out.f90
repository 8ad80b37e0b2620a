! hmhd_out.f90 -- output files of the Fortran driver.
!
! The datasets of the reference's output_writeHdf5 (src/output.f90:135-330)
! -- output_number, time, dt, step_number, dens, velx, vely, velz, magx,
! magy, magz, pres, eint, gama, ener, magp, divb, scheme, ordr, xmin, xmax,
! ymin, ymax, nx, ny -- with the same derived fields (ener, magp, divb from
! V with its guard cells, ordr = 2 scheme + 1), plus `cons`, the interior
! U(8, nx, ny), so a run can be resumed or checked bit for bit.  The
! format is the build's own (no HDF5 library for this compiler): a stream
! file of little-endian records
!     magic 'HMHDOUT1', int32 record count, then per record:
!     name (16 bytes, blank padded), int32 type (1 = int32, 2 = float64),
!     int32 rank, int32 dims(3) (unused dims 1), the data in Fortran order.
! hosono_2dmhd_amd/outfile.py reads it.  The decision of when to write is
! output_write's (src/output.f90:28-112).
module hmhd_out
    implicit none
    private

    public :: output_due, output_file

contains

    ! src/output.f90:62-79: whether step nstep at time t writes a file
    logical function output_due(nstep, t, last_step, last_time, freq_step, freq_time, force)
        integer, intent(in) :: nstep, last_step, freq_step
        real(8), intent(in) :: t, last_time, freq_time
        logical, intent(in) :: force

        if (freq_time <= 0 .and. freq_step <= 0) then
            write(*,*) "=========================================================================="
            write(*,*) "No ouptputs will be written because both sim_outputFreqStep and"
            write(*,*) "sim_outputFreqTime are not positive. Please set at least one of them"
            write(*,*) "to be greater than zero."
            write(*,*) "=========================================================================="
            stop
        end if
        output_due = force
        if (nstep - last_step == freq_step .and. freq_step > 0) then
            output_due = .true.
        else if (t - last_time >= freq_time .and. freq_time > 0) then
            output_due = .true.
        end if
    end function output_due

    subroutine put_header(u, name, typ, rank, d1, d2, d3)
        integer, intent(in) :: u, typ, rank, d1, d2, d3
        character(len=*), intent(in) :: name
        character(len=16) :: nm

        nm = name
        write(u) nm, int(typ, 4), int(rank, 4), int(d1, 4), int(d2, 4), int(d3, 4)
    end subroutine put_header

    subroutine put_r0(u, name, x)
        integer, intent(in) :: u
        character(len=*), intent(in) :: name
        real(8), intent(in) :: x

        call put_header(u, name, 2, 1, 1, 1, 1)
        write(u) x
    end subroutine put_r0

    subroutine put_i0(u, name, k)
        integer, intent(in) :: u, k
        character(len=*), intent(in) :: name

        call put_header(u, name, 1, 1, 1, 1, 1)
        write(u) int(k, 4)
    end subroutine put_i0

    subroutine put_r2(u, name, a)
        integer, intent(in) :: u
        character(len=*), intent(in) :: name
        real(8), intent(in) :: a(:, :)

        call put_header(u, name, 2, 2, size(a, 1), size(a, 2), 1)
        write(u) a
    end subroutine put_r2

    subroutine put_i2(u, name, a)
        integer, intent(in) :: u
        character(len=*), intent(in) :: name
        integer(4), intent(in) :: a(:, :)

        call put_header(u, name, 1, 2, size(a, 1), size(a, 2), 1)
        write(u) a
    end subroutine put_i2

    ! one output file: V(10, mx, my) with guard cells, U(8, mx, my),
    ! scheme(mx, my); interior [ngc+1, ngc+nx] x [ngc+1, ngc+ny]
    subroutine output_file(path, counter, t, dt, nstep, nx, ny, ngc, xbeg, xend, ybeg, yend, U, V, scheme)
        character(len=*), intent(in) :: path
        integer, intent(in) :: counter, nstep, nx, ny, ngc
        real(8), intent(in) :: t, dt, xbeg, xend, ybeg, yend
        real(8), intent(in) :: U(:, :, :), V(:, :, :)
        integer(4), intent(in) :: scheme(:, :)
        real(8), allocatable :: f(:, :), ener(:, :), magp(:, :), divb(:, :)
        real(8) :: dl(2)
        integer :: u_, ios, s, ex, ey, k, d, ii, jj
        character(len=16), parameter :: names(10) = [character(len=16) :: "dens", "velx", "vely", "velz", "pres", &
                                                     "magx", "magy", "magz", "eint", "gama"]

        s = ngc + 1
        ex = ngc + nx
        ey = ngc + ny
        dl(1) = (xend - xbeg) / nx
        dl(2) = (yend - ybeg) / ny
        open(newunit=u_, file=path, access='stream', form='unformatted', status='replace', iostat=ios)
        if (ios /= 0) then
            write(*,*) "=========================================================================="
            write(*,*) "Could not create the output file: ", path
            write(*,*) "Check that the directory it lives in exists and is writable."
            write(*,*) "(the directory comes from sim_dataFileBaseName in the parameter file)"
            write(*,*) "=========================================================================="
            stop
        end if
        write(u_) 'HMHDOUT1', int(26, 4)
        call put_i0(u_, "output_number", counter)
        call put_r0(u_, "time", t)
        call put_r0(u_, "dt", dt)
        call put_i0(u_, "step_number", nstep)
        allocate(f(nx, ny))
        do k = 1, 10
            f = V(k, s:ex, s:ey)
            call put_r2(u_, trim(names(k)), f)
        end do
        ! src/output.f90:248-274, in its order of operations
        allocate(ener(nx, ny), magp(nx, ny), divb(nx, ny))
        ener = V(1, s:ex, s:ey) * (V(2, s:ex, s:ey)**2 + V(3, s:ex, s:ey)**2 + V(4, s:ex, s:ey)**2) / 2 &
               + V(1, s:ex, s:ey) * V(9, s:ex, s:ey)
        magp = (V(6, s:ex, s:ey)**2 + V(7, s:ex, s:ey)**2 + V(8, s:ex, s:ey)**2) / 2
        ener = ener + magp
        divb = 0.0d0
        do d = 1, 2
            ii = merge(1, 0, d == 1)
            jj = merge(1, 0, d == 2)
            divb = divb + 1 / (2 * dl(d)) * (V(5 + d, s + ii:ex + ii, s + jj:ey + jj) &
                                             - V(5 + d, s - ii:ex - ii, s - jj:ey - jj))
        end do
        call put_r2(u_, "ener", ener)
        call put_r2(u_, "magp", magp)
        call put_r2(u_, "divb", divb)
        call put_i2(u_, "scheme", scheme(s:ex, s:ey))
        call put_i2(u_, "ordr", 2 * scheme(s:ex, s:ey) + 1)
        call put_r0(u_, "xmin", xbeg)
        call put_r0(u_, "xmax", xend)
        call put_r0(u_, "ymin", ybeg)
        call put_r0(u_, "ymax", yend)
        call put_i0(u_, "nx", nx)
        call put_i0(u_, "ny", ny)
        ! the interior conserved state, cons(8, nx, ny)
        call put_header(u_, "cons", 2, 3, 8, nx, ny)
        write(u_) U(:, s:ex, s:ey)
        close(u_)
    end subroutine output_file

end module hmhd_out
