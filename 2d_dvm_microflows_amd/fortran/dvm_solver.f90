! dvm_solver.f90 -- module DVMSolver: the reference's solver module
! (solver/dvm_solver.f90:86-88) with the same public procedures,
! initialize_solver, solve and destroy_solver, so that a main program written
! against the reference (main.f90: use DVMSolver; use ScreenOutput; call
! initialize_solver(); call solve_message(); call solve(); call
! finalize_message(); call destroy_solver()) builds unchanged against this
! module and runs its deck on the GPU through the DVMHip C-ABI.
!
! Command line (command_line.f90:45-86): <deck> [<velocity-grid file>], plus
!   --check      parse the deck, write reference_values.txt and print the
!                derived configuration ("key value" lines); no device work
!   --steps N    last step N instead of the deck's ntstep
!   --dump-init  raw FP64 initial distribution [cell][node]
!   --dump       raw FP64 distribution after the last step
!   --props      final moments, one line per cell (density u v w T_tr T p <|v|>)
! With read_rst = 1 the run starts from the deck's restart file (its step
! count and distribution) and steps ntime_init+1 .. N, as the solve() loop
! does (dvm_solver.f90:265-299); with write_rst = 1 it writes the restart
! file every write_freq steps and at step N (restart.f90:92-145).
module DVMSolver
  use iso_c_binding
  use DVMHip
  use DVMDeck
  use ScreenOutput, only: set_screen_verbosity
  implicit none
  private
  public :: initialize_solver, solve, destroy_solver

  ! the run's state (the reference's module globals phi, properties, ...)
  type(deck_t) :: d
  type(run_t), target :: r
  type(c_ptr) :: ctx = c_null_ptr
  character(len=512) :: deck_file, grid_file, dump_init, dump_final, props_file
  logical :: check = .false.
  integer :: steps = -1, rc, n, ncell, loc, ntime_init = 0
  integer(c_int32_t) :: ntime
  real(c_double), allocatable, target :: f(:,:), props(:,:), ext(:,:), heat(:,:)
  integer(c_int32_t) :: heat_nt
  real(c_double), target :: phase(DVM_NPHASE)
  integer(c_int64_t) :: calls
  real(c_double) :: t0, t1

contains

  ! initialize_solver (solver/dvm_solver.f90:92-229): command line, deck,
  ! scaling (reference_values.txt), velocity grid, initial state or restart,
  ! output file headers.
  subroutine initialize_solver()
    character(len=512) :: arg, msg
    logical :: ok
    integer :: i, nargs, npos
    deck_file = ''
    grid_file = ''
    dump_init = ''
    dump_final = ''
    props_file = ''
    npos = 0
    nargs = command_argument_count()
    i = 1
    do while (i <= nargs)
       call get_command_argument(i, arg)
       select case (trim(arg))
       case ('--check')
          check = .true.
       case ('--steps')
          i = i + 1
          call get_command_argument(i, arg)
          read(arg, *) steps
       case ('--dump-init')
          i = i + 1
          call get_command_argument(i, dump_init)
       case ('--dump')
          i = i + 1
          call get_command_argument(i, dump_final)
       case ('--props')
          i = i + 1
          call get_command_argument(i, props_file)
       case default
          npos = npos + 1
          if (npos == 1) deck_file = arg
          if (npos == 2) grid_file = arg
       end select
       i = i + 1
    end do
    if (len_trim(deck_file) == 0) then
       write(*, '(a)') 'usage: <program> <deck> [<velocity-grid file>] [--steps N] [--check] [--dump-init F] ' // &
            '[--dump F] [--props F]'
       stop 1
    end if
    call deck_read(trim(deck_file), d, ok, msg)
    if (.not. ok) call fail(msg)
    call run_from_deck(d, r, ok, msg, trim(grid_file))
    if (.not. ok) call fail(msg)
    ! --check prints only the configuration
    call set_screen_verbosity(merge(0, r%verbosity, check))
    if (steps < 0) steps = r%ntstep
    r%cfg%vx = c_loc(r%vx)
    r%cfg%vy = c_loc(r%vy)
    r%cfg%vz = c_loc(r%vz)
    call write_reference_values(r, 'reference_values.txt')
    n = size(r%f0, 1)
    ncell = size(r%f0, 2)
    ntime_init = 0
    if (r%read_rst) then
       call read_restart(r%read_rst_file, ntime_init, r%f0, ok, msg)
       if (.not. ok) call fail(msg)
    end if
    if (len_trim(dump_init) > 0) call dump(dump_init, r%f0)
    if (r%props_vis) then
       call props_header(r, steps)
       call species_header(r, steps)
    end if
    call df_headers(r)
    call wall_headers(r)
  end subroutine initialize_solver

  ! The time loop of solve() (solver/dvm_solver.f90:231-369) on the GPU: one
  ! fused dvm_step per step, with the screen, property, distribution-function,
  ! wall and restart output of the deck at their intervals.
  subroutine solve()
    integer :: u, c
    if (check) then
       call print_config()
       return
    end if
    rc = dvm_ctx_create(r%cfg, ctx)
    if (rc /= DVM_OK) call fail(dvm_error_message(c_null_ptr))
    allocate(f(n, ncell), props(DVM_NPROP, ncell), ext(DVM_NEXT, ncell))
    allocate(heat(max(r%cfg%nx, r%cfg%ny), 4))
    if (r%heat_freq > 0 .and. r%cfg%nx > 1 .and. r%cfg%ny > 1) then
       rc = dvm_set_wall_heat_flux_freq(ctx, int(r%heat_freq, c_int32_t))
       if (rc /= DVM_OK) call fail(dvm_error_message(ctx))
    end if
    rc = dvm_upload_f(ctx, 0_c_int32_t, r%f0)
    if (rc == DVM_OK) rc = dvm_moments(ctx, props)
    if (rc /= DVM_OK) call fail(dvm_error_message(ctx))
    rc = dvm_timing_reset(ctx)
    loc = r%out_x + r%cfg%nx * (r%out_y - 1)
    if (r%verbosity >= 2) call screen(ntime_init)
    if (r%props_vis) call write_props(ntime_init)
    if (r%df_vis) call write_df(ntime_init)
    call cpu_time(t0)
    do ntime = ntime_init + 1, steps
       rc = dvm_step(ctx, ntime)
       if (rc /= DVM_OK) call fail(dvm_error_message(ctx))
       if (r%heat_freq > 0 .and. r%cfg%nx > 1 .and. r%cfg%ny > 1) then
          if (mod(int(ntime), r%heat_freq) == 0) then
             rc = dvm_wall_heat_flux(ctx, heat, heat_nt)
             if (rc /= DVM_OK) call fail(dvm_error_message(ctx))
             call wall_records(r, int(ntime), heat)
          end if
       end if
       if (restart_due(r, int(ntime), steps)) then
          rc = dvm_download_f(ctx, 0_c_int32_t, f)
          if (rc /= DVM_OK) call fail(dvm_error_message(ctx))
          call write_restart(r%write_rst_file, int(ntime), f)
       end if
       if (r%verbosity >= 2 .and. mod(int(ntime), r%display_freq) == 0) then
          rc = dvm_moments(ctx, props)
          if (rc /= DVM_OK) call fail(dvm_error_message(ctx))
          call screen(int(ntime))
       end if
       if (r%props_vis) then
          if (mod(int(ntime), steps) == 0 .or. mod(int(ntime), r%props_freq) == 0) call write_props(int(ntime))
       end if
       if (r%df_vis) then
          if (mod(int(ntime), steps) == 0 .or. mod(int(ntime), r%df_freq) == 0) call write_df(int(ntime))
       end if
    end do
    rc = dvm_moments(ctx, props)
    if (rc /= DVM_OK) call fail(dvm_error_message(ctx))
    rc = dvm_timing_read(ctx, phase, calls)
    call cpu_time(t1)
    if (r%verbosity >= 1) then
       write(*, '(a,i0,a,f10.3,a)') 'steps: ', max(0, steps - ntime_init), '   wall (host cpu_time): ', t1 - t0, ' s'
       write(*, '(a,7f10.4)') 'device ms (cdf prep collide conserve walls sweep moments): ', phase
    end if
    if (len_trim(dump_final) > 0) then
       rc = dvm_download_f(ctx, 0_c_int32_t, f)
       if (rc /= DVM_OK) call fail(dvm_error_message(ctx))
       call dump(dump_final, f)
    end if
    if (len_trim(props_file) > 0) then
       open(newunit=u, file=trim(props_file), status='replace', action='write')
       do c = 1, ncell
          write(u, '(8es25.16e3)') props(:, c)
       end do
       close(u)
    end if
  end subroutine solve

  ! destroy_solver (solver/dvm_solver.f90:371-414): frees the context and the
  ! host arrays.
  subroutine destroy_solver()
    if (c_associated(ctx)) call dvm_ctx_destroy(ctx)
    ctx = c_null_ptr
    if (allocated(f)) deallocate(f)
    if (allocated(props)) deallocate(props)
    if (allocated(ext)) deallocate(ext)
    if (allocated(heat)) deallocate(heat)
  end subroutine destroy_solver

  ! Moments and output moments of the resident f, then one record of the total
  ! and the per-species property files.
  subroutine write_props(nt)
    integer, intent(in) :: nt
    rc = dvm_moments_ext(ctx, ext)
    if (rc == DVM_OK) rc = dvm_moments(ctx, props)
    if (rc == DVM_OK) rc = dvm_download_f(ctx, 0_c_int32_t, f)
    if (rc /= DVM_OK) call fail(dvm_error_message(ctx))
    call props_record(r, dble(nt) * r%cfg%dt, props, ext)
    call species_record(r, dble(nt) * r%cfg%dt, f, props, ext)
  end subroutine write_props

  ! Distribution-function records of the resident f and its moments.
  subroutine write_df(nt)
    integer, intent(in) :: nt
    rc = dvm_moments(ctx, props)
    if (rc == DVM_OK) rc = dvm_download_f(ctx, 0_c_int32_t, f)
    if (rc /= DVM_OK) call fail(dvm_error_message(ctx))
    call df_records(r, dble(nt) * r%cfg%dt, f, props)
  end subroutine write_df

  subroutine fail(m)
    character(len=*), intent(in) :: m
    write(*, '(a,a)') 'Error: ', trim(m)
    stop 2
  end subroutine fail

  subroutine dump(file, a)
    character(len=*), intent(in) :: file
    real(c_double), intent(in) :: a(:,:)
    integer :: u
    open(newunit=u, file=trim(file), access='stream', form='unformatted', status='replace', action='write')
    write(u) a
    close(u)
  end subroutine dump

  subroutine screen(nt)
    integer, intent(in) :: nt
    if (nt == 0 .or. mod(nt, 20 * r%display_freq) == 0) &
         write(*, '(8x,6a20)') 'density  ', 'u-velocity ', 'trans. temp', 'pressure ', 'time step ', ''
    write(*, '(a11,4e20.10,i8,a6,i7)') 'total:', props(1 + 0, loc), props(1 + 1, loc), props(1 + 4, loc), &
         props(1 + 6, loc), nt, ' of ', steps
  end subroutine screen

  subroutine print_config()
    integer :: k
    write(*, '(a,1x,i0)') 'dim', r%cfg%dim
    write(*, '(a,1x,i0)') 'nx', r%cfg%nx
    write(*, '(a,1x,i0)') 'ny', r%cfg%ny
    write(*, '(a,1x,es25.16e3)') 'dx', r%cfg%dx
    write(*, '(a,1x,es25.16e3)') 'dy', r%cfg%dy
    write(*, '(a,1x,i0)') 'nvx', r%cfg%nvx
    write(*, '(a,1x,i0)') 'nvy', r%cfg%nvy
    write(*, '(a,1x,i0)') 'nvz', r%cfg%nvz
    write(*, '(a,1x,es25.16e3)') 'vx0', r%vx(1)
    write(*, '(a,1x,es25.16e3)') 'vx1', r%vx(size(r%vx))
    write(*, '(a,1x,es25.16e3)') 'vy0', r%vy(1)
    write(*, '(a,1x,es25.16e3)') 'vy1', r%vy(size(r%vy))
    write(*, '(a,1x,es25.16e3)') 'vz0', r%vz(1)
    write(*, '(a,1x,es25.16e3)') 'vz1', r%vz(size(r%vz))
    write(*, '(a,1x,es25.16e3)') 'dt', r%cfg%dt
    write(*, '(a,1x,es25.16e3)') 'kn', r%cfg%kn
    write(*, '(a,1x,es25.16e3)') 'mass', r%cfg%mass
    write(*, '(a,1x,es25.16e3)') 'diam', r%cfg%diam
    write(*, '(a,1x,es25.16e3)') 'omega', r%cfg%omega
    write(*, '(a,1x,es25.16e3)') 'alpha', r%cfg%alpha
    write(*, '(a,1x,es25.16e3)') 'viscosity', r%cfg%viscosity
    write(*, '(a,1x,es25.16e3)') 't_viscosity', r%cfg%t_viscosity
    write(*, '(a,1x,i0)') 'csection_coeff', r%cfg%csection_coeff
    write(*, '(a,1x,i0)') 'coll_solver', r%cfg%coll_solver
    write(*, '(a,1x,es25.16e3)') 'cdf_tolerance', r%cfg%cdf_tolerance
    write(*, '(a,1x,i0)') 'num_repl', r%cfg%num_repl
    write(*, '(a,1x,es25.16e3)') 'coln_rms', r%cfg%coln_rms
    write(*, '(a,1x,i0)') 'sample_mode', r%cfg%sample_mode
    write(*, '(a,1x,i0)') 'samples_per_cell', r%cfg%samples_per_cell
    write(*, '(a,1x,i0)') 'seed', r%cfg%seed
    write(*, '(a,1x,i0)') 'precision', r%cfg%precision
    write(*, '(a,1x,i0)') 'enable_convection', r%cfg%enable_convection
    write(*, '(a,1x,i0)') 'bc_left', r%cfg%bc_left
    write(*, '(a,1x,i0)') 'bc_right', r%cfg%bc_right
    write(*, '(a,1x,es25.16e3)') 'acc_coeff', r%cfg%acc_coeff
    write(*, '(a,1x,i0)') 'conv_order', r%cfg%conv_order
    do k = 1, 4
       write(*, '(a,i0,5(1x,es25.16e3))') 'wall', k - 1, r%cfg%wall(:, k)
    end do
    write(*, '(a,1x,i0)') 'ntstep', r%ntstep
    write(*, '(a,1x,i0)') 'ntime_init', ntime_init
    write(*, '(a,1x,a)') 'species', trim(r%species)
  end subroutine print_config

end module DVMSolver
