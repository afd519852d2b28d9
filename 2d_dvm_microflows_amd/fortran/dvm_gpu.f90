! dvm_gpu.f90 -- runs a trunk input deck on the GPU: the main program of the
! reference (main.f90) over the drop-in DVMSolver and ScreenOutput modules.
!
!   dvm_gpu <deck> [<velocity-grid file>] [--steps N] [--check] [--dump-init F]
!           [--dump F] [--props F]
program dvm_gpu
  use DVMSolver
  use ScreenOutput
  implicit none
  call initialize_solver()
  call solve_message()
  call solve()
  call finalize_message()
  call destroy_solver()
end program dvm_gpu
