! screen_output.f90 -- module ScreenOutput: the two banners the reference's
! main program prints around the solve (input_output/screen_output.f90:120-144),
! shown from verbosity 2 on (the reference's pretty_messages_verbosity).
module ScreenOutput
  implicit none
  private
  public :: solve_message, finalize_message, set_screen_verbosity
  integer :: verbosity = 2
  integer, parameter :: pretty_messages_verbosity = 2
  character(len=*), parameter :: rule = &
       '==========================================================================================='

contains

  subroutine set_screen_verbosity(v)
    integer, intent(in) :: v
    verbosity = v
  end subroutine set_screen_verbosity

  subroutine solve_message()
    if (verbosity < pretty_messages_verbosity) return
    write(*, '(a)') rule
    write(*, '(a)') '                                    Beginning Solve'
    write(*, '(a)') rule
  end subroutine solve_message

  subroutine finalize_message()
    if (verbosity < pretty_messages_verbosity) return
    write(*, '(a)') rule
    write(*, '(a)') '                                   Finishing Solve'
    write(*, '(a)') rule
  end subroutine finalize_message

end module ScreenOutput
