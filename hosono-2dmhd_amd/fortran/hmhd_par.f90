! hmhd_par.f90 -- reader for the reference's .par files.
!
! Same syntax and lookup as src/readParamFile.f90: a line that starts with
! '#' is a comment, text after a '#' elsewhere is dropped, the name is the
! text before the first '=' (compared without trailing blanks), the first
! matching line wins, and the value is read list-directed (numbers,
! 'quoted strings', .true./.false.).  A missing parameter stops the run
! with the reference's message.
module hmhd_par
    implicit none
    private

    integer, parameter, public :: par_len = 800

    public :: par_int, par_real, par_char, par_logical, par_has

contains

    subroutine find_value(file, key, val, found)
        character(len=*), intent(in) :: file, key
        character(len=par_len), intent(out) :: val
        logical, intent(out) :: found
        character(len=par_len) :: line
        integer :: u, ios, k

        found = .false.
        val = ''
        open(newunit=u, file=file, status='old', action='read', form='formatted', iostat=ios)
        if (ios /= 0) then
            write(*,*) "Could not open the parameter file ", trim(file)
            stop
        end if
        do
            read(u, '(A)', iostat=ios) line
            if (ios /= 0) exit
            if (line(1:1) == '#') cycle
            k = index(line, '#')
            if (k > 0) line = line(1:k-1)
            k = index(line, '=')
            if (k < 2) cycle
            if (trim(line(1:k-1)) == key) then
                val = line(k+1:)
                found = .true.
                exit
            end if
        end do
        close(u)
    end subroutine find_value

    subroutine value_of(file, key, val)
        character(len=*), intent(in) :: file, key
        character(len=par_len), intent(out) :: val
        logical :: found

        call find_value(file, key, val, found)
        if (.not. found) then
            write(*,*) "====================================================================================="
            write(*,*) "UNSUCCESFUL INPUT FILE READ."
            write(*,*) "The parameter            '", trim(key), "'"
            write(*,*) "was not found inside of  '", trim(file), "'"
            write(*,*) "---------------------------------------------------------------------------------"
            write(*,*) "Make sure the parameter is present in the input"
            write(*,*) "file and is spelled correctly."
            write(*,*) "====================================================================================="
            stop
        end if
        write(*,*) trim(key)//" set to "//trim(val)
    end subroutine value_of

    logical function par_has(file, key)
        character(len=*), intent(in) :: file, key
        character(len=par_len) :: val

        call find_value(file, key, val, par_has)
    end function par_has

    integer function par_int(file, key)
        character(len=*), intent(in) :: file, key
        character(len=par_len) :: val

        call value_of(file, key, val)
        read(val, *) par_int
    end function par_int

    real(8) function par_real(file, key)
        character(len=*), intent(in) :: file, key
        character(len=par_len) :: val

        call value_of(file, key, val)
        read(val, *) par_real
    end function par_real

    function par_char(file, key) result(s)
        character(len=*), intent(in) :: file, key
        character(len=par_len) :: s
        character(len=par_len) :: val

        call value_of(file, key, val)
        read(val, *) s
    end function par_char

    logical function par_logical(file, key)
        character(len=*), intent(in) :: file, key
        character(len=par_len) :: val

        call value_of(file, key, val)
        read(val, *) par_logical
    end function par_logical

end module hmhd_par
