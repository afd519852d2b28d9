! dvm_deck.f90 -- reads a trunk input deck (software/trunk/input_files/*.inp,
! the keys of input_output/read_input_2D.f90:75-777) and turns it into the
! dvm_config, velocity axes and initial distribution of the GPU library.
!
! Deck format: one "key = value" per line; '#' starts a comment; section
! headers ("[phy_grid]") and lines without '=' are skipped, as by the
! reference's file_read (input_output/input.f90). The reference keeps the
! quotes of quoted values ('1.0d0') and its numeric reads then fail (Q16);
! here quotes are stripped. Vector values (one entry per species) take their
! first entry: the library runs one monatomic species.
!
! Supported: num_species = 1 of He, Ne, Ar, Xe (species_data.f90:245-524),
! csection_method 1-4, csection_coeff 1-3, uniform velocity axes
! (x/y/z_vel_grid_type = 1) and axes read from the velocity-grid file of the
! command line (type 0, read_input_2D.f90:779-829), iColl 0, 2 or 3, iConv 0
! or 1, InitC 1-5 (4: the Rankine-Hugoniot shock profile) with InitDF 1-4
! (Maxwellian, BKW, delta function, exact Gaussian), LWallBC / RWallBC 1-4.
! Anything else stops with a message, as the reference stops on an invalid
! deck value.
!
! Extensions (keys the reference does not read):
!   dvm_precision    64 (default) or 32: FP32 collision arithmetic
!   dvm_dim          1: zone-grid 1-D semantics for nx_space > 1, ny_space = 1
!   dvm_conv_order   2: limited second-order 1-D sweep
!   dvm_samples      > 0: that many samples per cell per step instead of the
!                    ColnRMS count rule
module DVMDeck
  use iso_c_binding
  use DVMHip
  implicit none
  private

  integer, parameter :: max_keys = 1000
  real(c_double), parameter :: pi_ref = 3.1459265358979323d0   ! constants.f90:30
  real(c_double), parameter :: kb = 1.3805d-23                 ! constants.f90:31

  type, public :: deck_t
     integer :: n = 0
     character(len=64) :: key(max_keys)
     character(len=256) :: val(max_keys)
  end type deck_t

  ! Everything a run needs, derived from the deck.
  type, public :: run_t
     type(dvm_config) :: cfg
     real(c_double), allocatable :: vx(:), vy(:), vz(:)
     real(c_double), allocatable :: f0(:,:)            ! (nnode, ncell)
     integer :: ntstep = 0, display_freq = 1, verbosity = 2
     integer :: out_x = 1, out_y = 1                   ! screen_output_loc
     real(c_double) :: c_ref, sigma_ref, l_scale, t_scale, mass_ref, diam_ref, dens_ref, temp_ref
     character(len=16) :: species = ''
     ! macroscopic-properties output (visualization.f90:270-340, 615-660, 1591-1716)
     logical :: props_vis = .false.
     character(len=128) :: props_file = 'macro_properties'
     integer :: props_freq = 1
     logical :: to_output(27) = .false.
     integer :: n_stress = 0, n_moment = 0
     integer :: stress_dirs(6) = 0, moment_nums(16) = 0
     real(c_double) :: gamma = 5.0d0 / 3.0d0, mol_mass = 1.0d0
     ! distribution-function output (visualization.f90:497-560, 920-1215)
     logical :: df_vis = .false.
     character(len=128) :: df_file = 'dist_func'
     integer :: df_freq = 1, df_dim = 3, n_loc = 0
     integer :: loc_x(64) = 1, loc_y(64) = 1
     integer :: i_zero = 0, j_zero = 0, k_zero = 0
     ! wall heat-flux files (boundary_conditions.f90:108-116, 550-575, 872-923)
     integer :: heat_freq = 0
     ! restart (restart.f90:66-80, read_input_2D.f90:745-760)
     logical :: read_rst = .false., write_rst = .false.
     integer :: write_freq = 0
     character(len=128) :: read_rst_file = 'restart.rst', write_rst_file = 'restart.rst'
  end type run_t

  ! deck switches of the 27 output properties, in the order of
  ! visualization.f90:276-311
  character(len=20), parameter :: prop_keys(27) = [character(len=20) :: &
       'write_dens', 'write_x_vel', 'write_y_vel', 'write_z_vel', 'write_speed', 'write_mach', &
       'write_tr_temp', 'write_rot_temp', 'write_vib_temp', 'write_tot_temp', 'write_temp_x', 'write_temp_y', &
       'write_temp_z', 'write_tr_energy', 'write_rot_energy', 'write_vib_energy', 'write_tot_energy', &
       'write_entropy', 'write_heat_flux_x', 'write_heat_flux_y', 'write_heat_flux_z', 'write_rot_dof', &
       'write_vib_dof', 'write_rot_temp_relax', 'write_vib_temp_relax', 'write_kin_temp_relax', 'write_pressure']
  ! property names of visualization.f90:314-340
  character(len=24), parameter, public :: prop_names(27) = [character(len=24) :: &
       '"Density"', '"X-Velocity"', '"Y-Velocity"', '"Z-Velocity"', '"Speed"', '"Mach Number"', &
       '"Trans. Temp."', '"Rot. Temp."', '"Vib. Temp."', '"Total Temp."', '"Temp. X"', '"Temp. Y"', &
       '"Temp. Z"', '"Trans. Energy"', '"Rot. Energy"', '"Vib. Energy"', '"Total Energy"', '"Entropy"', &
       '"Heat Flux X"', '"Heat Flux Y"', '"Heat Flux Z"', '"Rot DOF"', '"Vib DOF"', '"Rot. Temp Analytic"', &
       '"Vib. Temp Analytic"', '"Trans. Temp Analytic"', '"Pressure"']

  public :: deck_read, deck_has, deck_real, deck_int, deck_str, run_from_deck, write_reference_values
  public :: props_header, props_record, write_restart, read_restart, restart_due, df_headers, df_records
  public :: wall_headers, wall_records, species_header, species_record

contains

  subroutine deck_read(filename, d, ok, msg)
    character(len=*), intent(in) :: filename
    type(deck_t), intent(out) :: d
    logical, intent(out) :: ok
    character(len=*), intent(out) :: msg
    character(len=512) :: line
    integer :: u, ios, p, q
    ok = .false.
    msg = ''
    open(newunit=u, file=filename, status='old', action='read', iostat=ios)
    if (ios /= 0) then
       msg = 'cannot open deck ' // trim(filename)
       return
    end if
    do
       read(u, '(a)', iostat=ios) line
       if (ios /= 0) exit
       p = index(line, '#')
       if (p > 0) line = line(1:p-1)
       p = index(line, '=')
       if (p <= 1) cycle
       if (d%n >= max_keys) exit
       d%n = d%n + 1
       d%key(d%n) = adjustl(line(1:p-1))
       d%val(d%n) = adjustl(line(p+1:))
       ! strip quotes
       do q = 1, len_trim(d%val(d%n))
          if (d%val(d%n)(q:q) == "'" .or. d%val(d%n)(q:q) == '"') d%val(d%n)(q:q) = ' '
       end do
       d%val(d%n) = adjustl(d%val(d%n))
    end do
    close(u)
    ok = .true.
  end subroutine deck_read

  integer function find_key(d, key)
    type(deck_t), intent(in) :: d
    character(len=*), intent(in) :: key
    integer :: i
    find_key = 0
    do i = 1, d%n
       if (trim(d%key(i)) == key) then
          find_key = i
          return
       end if
    end do
  end function find_key

  logical function deck_has(d, key)
    type(deck_t), intent(in) :: d
    character(len=*), intent(in) :: key
    deck_has = find_key(d, key) > 0
  end function deck_has

  ! First whitespace- or comma-separated token of a value (one species).
  function first_token(v) result(t)
    character(len=*), intent(in) :: v
    character(len=256) :: t
    integer :: i
    t = adjustl(v)
    do i = 1, len_trim(t)
       if (t(i:i) == ' ' .or. t(i:i) == ',') then
          t = t(1:i-1)
          return
       end if
    end do
  end function first_token

  real(c_double) function deck_real(d, key, default, token)
    type(deck_t), intent(in) :: d
    character(len=*), intent(in) :: key
    real(c_double), intent(in) :: default
    integer, intent(in), optional :: token
    integer :: i, ios, k, n
    character(len=256) :: rest
    deck_real = default
    i = find_key(d, key)
    if (i == 0) return
    rest = d%val(i)
    n = 1
    if (present(token)) n = token
    do k = 1, n - 1
       rest = adjustl(rest(len_trim(first_token(rest)) + 2:))
    end do
    read(first_token(rest), *, iostat=ios) deck_real
    if (ios /= 0) deck_real = default
  end function deck_real

  integer function deck_int(d, key, default, token)
    type(deck_t), intent(in) :: d
    character(len=*), intent(in) :: key
    integer, intent(in) :: default
    integer, intent(in), optional :: token
    real(c_double) :: r
    if (present(token)) then
       r = deck_real(d, key, real(default, c_double), token)
    else
       r = deck_real(d, key, real(default, c_double))
    end if
    deck_int = nint(r)
  end function deck_int

  function deck_str(d, key, default) result(s)
    type(deck_t), intent(in) :: d
    character(len=*), intent(in) :: key, default
    character(len=256) :: s
    integer :: i
    s = default
    i = find_key(d, key)
    if (i > 0) s = first_token(d%val(i))
  end function deck_str

  ! Monatomic species constants of species_data.f90:245-524, kept as the
  ! reference writes them (default-real literals for diameter and mass).
  subroutine species_constants(code, diam, mass, omega, alpha, visc, tvisc, ok)
    character(len=*), intent(in) :: code
    real(c_double), intent(out) :: diam, mass, omega, alpha, visc, tvisc
    logical, intent(out) :: ok
    ok = .true.
    visc = 0.0d0
    tvisc = 0.0d0
    select case (trim(code))
    case ('He')
       diam = 2.33e-10; mass = 6.65e-27; omega = 0.66d0; alpha = 1.31d0
    case ('Ne')
       diam = 2.77e-10; mass = 33.5e-27; omega = 0.66d0; alpha = 1.31d0
    case ('Ar')
       diam = 4.17e-10; mass = 66.3e-27; omega = 0.81d0; alpha = 1.40d0
       visc = 2.75e-5; tvisc = 380.0d0
    case ('Xe')
       diam = 5.74e-10; mass = 218.e-27; omega = 0.85d0; alpha = 1.44d0
    case default
       ok = .false.
    end select
  end subroutine species_constants

  ! The deck's run: scaling (scaling.f90:96-130), species constants scaled by
  ! the reference species (species_data.f90:155-229), grids, boundary states
  ! and the initial distribution (initial_conditions.f90:212-650).
  subroutine run_from_deck(d, r, ok, msg, grid_file)
    type(deck_t), intent(in) :: d
    type(run_t), intent(inout) :: r
    logical, intent(out) :: ok
    character(len=*), intent(out) :: msg
    ! the velocity-grid file of the command line (command_line.f90:45-86), read
    ! for the axes whose x/y/z_vel_grid_type is 0 (read_input_2D.f90:779-829)
    character(len=*), intent(in), optional :: grid_file
    character(len=16) :: ref_species
    real(c_double) :: dref, mref, d_s, m_s, om, al, vis, tvis, dd, mm, oo, aa, vv, tt
    real(c_double) :: vmin(3), vmax(3), dens, u, v, w, temp, opt, coeff, s, x, y, z, b
    integer :: npt(3), method, icoll, iconv, initc, initdf, mid, nx, ny, ncell, n, i, j, k, c, ix, iy, gt, gts(3)
    real(c_double), allocatable :: ga(:), gb(:), gc(:), sdens(:), su(:), stemp(:)
    real(c_double) :: xk, bnorm, c2, x1, x2, y1, y2, z1, z2, dl(3)
    logical :: sok
    type(dvm_config) :: cfg
    character(len=2), parameter :: side(4) = ['LW', 'RW', 'BW', 'TW']
    ok = .false.
    msg = ''
    if (deck_int(d, 'num_species', 1) /= 1) then
       msg = 'num_species must be 1 (one monatomic species)'
       return
    end if
    r%species = deck_str(d, 'species_codes', 'Ar')
    ref_species = deck_str(d, 'species_ref', r%species)
    call species_constants(r%species, d_s, m_s, om, al, vis, tvis, sok)
    if (.not. sok) then
       msg = 'species ' // trim(r%species) // ' is not a supported monatomic species (He, Ne, Ar, Xe)'
       return
    end if
    call species_constants(ref_species, dref, mref, oo, aa, vv, tt, sok)
    if (.not. sok) then
       msg = 'species_ref ' // trim(ref_species) // ' is not a supported monatomic species'
       return
    end if
    ! scaling.f90:96-108 set_scaling_parameters
    r%dens_ref = deck_real(d, 'n_ref', 1.0d20)
    r%temp_ref = deck_real(d, 'T_ref', 273.0d0)
    r%mass_ref = mref
    r%diam_ref = dref
    cfg%kn = deck_real(d, 'Kn', 1.0d0)
    r%c_ref = sqrt(2.0d0 * kb * r%temp_ref / mref)
    r%sigma_ref = pi_ref * dref * dref
    r%l_scale = 1.0d0 / (cfg%kn * r%dens_ref * r%sigma_ref)
    r%t_scale = r%l_scale / r%c_ref
    ! species_data.f90:155-229 and the collision-model overrides :503-520
    method = deck_int(d, 'csection_method', 1)
    select case (method)
    case (1); om = 0.5d0; al = 1.0d0
    case (2); om = 1.0d0; al = 1.0d0
    case (3); al = 1.0d0
    case (4)
    case default
       msg = 'csection_method must be 1-4'
       return
    end select
    cfg%mass = m_s / mref
    cfg%diam = d_s / dref
    cfg%omega = om
    cfg%alpha = al
    cfg%viscosity = vis * r%sigma_ref / (mref * r%c_ref)
    cfg%t_viscosity = tvis / r%temp_ref
    cfg%csection_coeff = deck_int(d, 'csection_coeff', 1)
    if (cfg%csection_coeff == DVM_CS_VKI .and. vis == 0.0d0) then
       msg = 'csection_coeff 3 (vki) needs the species viscosity (Ar only)'
       return
    end if
    ! physical grid
    nx = deck_int(d, 'nx_space', 1)
    ny = deck_int(d, 'ny_space', 1)
    cfg%nx = nx
    cfg%ny = ny
    cfg%dx = deck_real(d, 'delta_x', 1.0d0)
    cfg%dy = deck_real(d, 'delta_y', 1.0d0)
    if (nx == 1 .and. ny == 1) then
       cfg%dim = DVM_DIM_0D
    else if (deck_int(d, 'dvm_dim', 2) == 1) then
       cfg%dim = DVM_DIM_1D
    else
       cfg%dim = DVM_DIM_2D
    end if
    ! time stepping
    cfg%dt = deck_real(d, 'deltat', 0.01d0)
    r%ntstep = deck_int(d, 'ntstep', 0)
    ! velocity grid (velocity_grid.f90:636-648 generate_uniform_grid)
    npt = [deck_int(d, 'num_points_x', 11), deck_int(d, 'num_points_y', 11), deck_int(d, 'num_points_z', 11)]
    vmin = [deck_real(d, 'xv_min', -3.0d0), deck_real(d, 'yv_min', -3.0d0), deck_real(d, 'zv_min', -3.0d0)]
    vmax = [deck_real(d, 'xv_max', 3.0d0), deck_real(d, 'yv_max', 3.0d0), deck_real(d, 'zv_max', 3.0d0)]
    gts = [deck_int(d, 'x_vel_grid_type', 1), deck_int(d, 'y_vel_grid_type', 1), deck_int(d, 'z_vel_grid_type', 1)]
    do k = 1, 3
       if (gts(k) /= 0 .and. gts(k) /= 1) then
          msg = 'x/y/z_vel_grid_type must be 0 (read from the velocity-grid file) or 1 (uniform)'
          return
       end if
    end do
    if (any(gts == 0)) then
       if (.not. present(grid_file)) then
          msg = 'a velocity grid type 0 needs the velocity-grid file (second command-line argument)'
          return
       end if
       if (len_trim(grid_file) == 0) then
          msg = 'a velocity grid type 0 needs the velocity-grid file (second command-line argument)'
          return
       end if
       call read_grid_file(grid_file, ga, gb, gc, sok, msg)
       if (.not. sok) return
    end if
    ! generate_uniform_grid (velocity_grid.f90:636-648) or the file's axis
    ! (generate_file_read_grid)
    if (gts(1) == 0) then
       npt(1) = size(ga)
       allocate(r%vx(npt(1)))
       r%vx = ga
    else
       allocate(r%vx(npt(1)))
       do i = 0, npt(1) - 1
          r%vx(i + 1) = vmin(1) + dble(i) * (vmax(1) - vmin(1)) / dble(npt(1) - 1)
       end do
    end if
    if (gts(2) == 0) then
       npt(2) = size(gb)
       allocate(r%vy(npt(2)))
       r%vy = gb
    else
       allocate(r%vy(npt(2)))
       do i = 0, npt(2) - 1
          r%vy(i + 1) = vmin(2) + dble(i) * (vmax(2) - vmin(2)) / dble(npt(2) - 1)
       end do
    end if
    if (gts(3) == 0) then
       npt(3) = size(gc)
       allocate(r%vz(npt(3)))
       r%vz = gc
    else
       allocate(r%vz(npt(3)))
       do i = 0, npt(3) - 1
          r%vz(i + 1) = vmin(3) + dble(i) * (vmax(3) - vmin(3)) / dble(npt(3) - 1)
       end do
    end if
    cfg%nvx = npt(1)
    cfg%nvy = npt(2)
    cfg%nvz = npt(3)
    ! collisions
    icoll = deck_int(d, 'iColl', 3)
    select case (icoll)
    case (0); cfg%coll_solver = DVM_COLL_NONE
    case (2); cfg%coll_solver = DVM_COLL_FMC
    case (3); cfg%coll_solver = DVM_COLL_VR
    case default
       msg = 'iColl must be 0 (none), 2 (full Monte Carlo) or 3 (variance reduction)'
       return
    end select
    cfg%cdf_tolerance = deck_real(d, 'cdf_tolerance', 0.01d0)
    cfg%num_repl = deck_int(d, 'NRepl', 1)
    cfg%coln_rms = deck_real(d, 'ColnRMS', 0.001d0)
    cfg%sample_mode = DVM_SAMPLES_REFERENCE
    cfg%samples_per_cell = 0
    if (deck_int(d, 'dvm_samples', 0) > 0) then
       cfg%sample_mode = DVM_SAMPLES_EXPLICIT
       cfg%samples_per_cell = deck_int(d, 'dvm_samples', 0)
    end if
    cfg%seed = int(deck_int(d, 'seed', 0), c_int64_t)
    cfg%precision = DVM_PREC_FP64
    if (deck_int(d, 'dvm_precision', 64) == 32) cfg%precision = DVM_PREC_FP32
    ! convection and boundaries
    iconv = deck_int(d, 'iConv', 1)
    if (iconv /= 0 .and. iconv /= 1) then
       msg = 'iConv must be 0 or 1 (first-order upwind)'
       return
    end if
    cfg%enable_convection = iconv
    cfg%conv_order = deck_int(d, 'dvm_conv_order', 1)
    cfg%bc_left = deck_int(d, 'LWallBC', DVM_BC_WALL)
    cfg%bc_right = deck_int(d, 'RWallBC', DVM_BC_WALL)
    cfg%acc_coeff = deck_real(d, 'acc_coeff', 1.0d0)
    do k = 1, 4
       cfg%wall(1, k) = deck_real(d, 'dens_' // side(k), 1.0d0)
       cfg%wall(2, k) = deck_real(d, 'x_vel_' // side(k), 0.0d0)
       cfg%wall(3, k) = deck_real(d, 'y_vel_' // side(k), 0.0d0)
       cfg%wall(4, k) = deck_real(d, 'z_vel_' // side(k), 0.0d0)
       cfg%wall(5, k) = deck_real(d, 'tr_temp_' // side(k), 1.0d0)
    end do
    ! properties output
    r%props_vis = deck_int(d, 'enable_properties_vis', 0) /= 0
    r%props_file = deck_str(d, 'properties_vis_filename', 'macro_properties')
    r%props_freq = max(1, deck_int(d, 'properties_vis_dump_freq', 1))
    do k = 1, 27
       r%to_output(k) = deck_int(d, prop_keys(k), 0) /= 0
    end do
    r%n_stress = min(6, max(0, deck_int(d, 'write_stress', 0)))
    do k = 1, r%n_stress
       r%stress_dirs(k) = deck_int(d, 'stress_directions', 11, k)
    end do
    r%n_moment = min(16, max(0, deck_int(d, 'write_moment', 0)))
    do k = 1, r%n_moment
       r%moment_nums(k) = deck_int(d, 'moment_numbers', 4, k)
    end do
    ! calculate_specific_heat_ratio / calculate_molecular_mass
    ! (shock_conditions.f90:222-314) for one monatomic species (3 dof),
    ! weighted by the right-wall density
    opt = cfg%mass * deck_real(d, 'dens_RW', 1.0d0)
    coeff = opt * (3.0d0 * 0.5d0 + 1.0d0) / opt
    r%gamma = coeff / (coeff - 1.0d0)
    r%mol_mass = opt / deck_real(d, 'dens_RW', 1.0d0)
    ! distribution-function output of a single monatomic species (no rotational
    ! or vibrational levels to write)
    r%df_vis = deck_int(d, 'enable_vel_df_vis', 0) /= 0
    if (deck_int(d, 'enable_rot_df_vis', 0) /= 0 .or. deck_int(d, 'enable_vib_df_vis', 0) /= 0) then
       msg = 'enable_rot_df_vis / enable_vib_df_vis need internal energy levels (monatomic species only)'
       return
    end if
    if (r%df_vis) then
       r%df_file = deck_str(d, 'df_vis_filename', 'dist_func')
       r%df_freq = max(1, deck_int(d, 'df_vis_dump_freq', 1))
       r%df_dim = deck_int(d, 'df_output_dimension', 3)
       r%n_loc = deck_int(d, 'num_output_locations', 1)
       if (r%df_dim < 1 .or. r%df_dim > 3) then
          msg = 'df_output_dimension must be 1, 2 or 3'
          return
       end if
       if (r%n_loc < 0 .or. r%n_loc > size(r%loc_x)) then
          msg = 'num_output_locations must be 0..64'
          return
       end if
       ! locations outside the physical grid (out-of-bounds cells in the
       ! reference) are skipped with a warning
       mid = 0
       do k = 1, r%n_loc
          i = deck_int(d, 'x_output_locations', 1, k)
          j = deck_int(d, 'y_output_locations', 1, k)
          if (i < 1 .or. i > nx .or. j < 1 .or. j > ny) then
             write(*, '(a,i0,a,i0,a)') 'Warning: distribution-function output location (', i, ',', j, &
                  ') is outside the physical grid and is skipped'
             cycle
          end if
          mid = mid + 1
          r%loc_x(mid) = i
          r%loc_y(mid) = j
       end do
       r%n_loc = mid
       ! the planes of the 1-D and 2-D outputs: first x (y) node with
       ! |x| < double_tol (visualization.f90:998-1012) and the z zero point
       ! (find_zero_point, velocity_grid.f90:653-689)
       r%i_zero = zero_node(r%vx)
       r%j_zero = zero_node(r%vy)
       r%k_zero = 0
       do k = 1, size(r%vz)
          if (r%vz(k) <= 0.0d0) r%k_zero = k
       end do
       if (r%df_dim == 1 .and. r%j_zero == 0 .or. r%df_dim < 3 .and. r%k_zero == 0) then
          msg = 'df_output_dimension 1 needs a y node at zero, 1 and 2 a z node at or below zero'
          return
       end if
    end if
    r%heat_freq = max(0, deck_int(d, 'wall_heat_flux_dump_freq', 0))
    ! restart files
    r%read_rst = deck_int(d, 'read_rst', 0) /= 0
    r%write_rst = deck_int(d, 'write_rst', 0) /= 0
    r%write_freq = deck_int(d, 'write_freq', 0)
    r%read_rst_file = deck_str(d, 'read_restart_filename', 'restart.rst')
    r%write_rst_file = deck_str(d, 'write_restart_filename', 'restart.rst')
    ! screen output
    r%verbosity = deck_int(d, 'verbosity', 2)
    r%display_freq = max(1, deck_int(d, 'display_freq', 1))
    r%out_x = min(max(deck_int(d, 'screen_output_loc', 1, 1), 1), nx)
    r%out_y = min(max(deck_int(d, 'screen_output_loc', 1, 2), 1), ny)
    ! initial distribution (initialize_dist_func, initial_conditions.f90:212-650):
    ! per cell a state from InitC, the InitDF shape times beta^3 at every node,
    ! renormalised to the cell's density
    initdf = deck_int(d, 'InitDF', 1)
    if (initdf < 1 .or. initdf > 4) then
       msg = 'InitDF must be 1 (Maxwellian), 2 (BKW), 3 (delta function) or 4 (exact Gaussian)'
       return
    end if
    initc = deck_int(d, 'InitC', 5)
    mid = deck_int(d, 'spatial_middle', nx / 2)
    if (initc < 1 .or. initc > 5) then
       msg = 'InitC must be 1 (left wall), 2 (right wall), 3 (left | right), 4 (Rankine-Hugoniot shock) or 5 (user input)'
       return
    end if
    ! the state of each column ix (the profile depends on x only)
    allocate(sdens(nx), su(nx), stemp(nx))
    if (initc == 4) then
       call shock_profile(cfg, deck_int(d, 'shock_thickness', 1), mid, sdens, su, stemp, sok, msg)
       if (.not. sok) return
    end if
    dl = [deck_real(d, 'x_delta_loc', 0.0d0), deck_real(d, 'y_delta_loc', 0.0d0), deck_real(d, 'z_delta_loc', 0.0d0)]
    ncell = nx * ny
    n = npt(1) * npt(2) * npt(3)
    allocate(r%f0(n, ncell))
    do iy = 1, ny
       do ix = 1, nx
          c = ix + nx * (iy - 1)
          select case (initc)
          case (1)
             dens = cfg%wall(1, 1); u = cfg%wall(2, 1); v = cfg%wall(3, 1); w = cfg%wall(4, 1); temp = cfg%wall(5, 1)
          case (2)
             dens = cfg%wall(1, 2); u = cfg%wall(2, 2); v = cfg%wall(3, 2); w = cfg%wall(4, 2); temp = cfg%wall(5, 2)
          case (3)
             if (ix > mid) then
                dens = cfg%wall(1, 2); u = cfg%wall(2, 2); v = cfg%wall(3, 2); w = cfg%wall(4, 2); temp = cfg%wall(5, 2)
             else
                dens = cfg%wall(1, 1); u = cfg%wall(2, 1); v = cfg%wall(3, 1); w = cfg%wall(4, 1); temp = cfg%wall(5, 1)
             end if
          case (4)
             ! v and w stay the left wall's (initial_conditions.f90:338-347)
             dens = sdens(ix); u = su(ix); v = cfg%wall(3, 1); w = cfg%wall(4, 1); temp = stemp(ix)
          case default
             dens = deck_real(d, 'dens_init', 1.0d0)
             u = deck_real(d, 'x_vel_init', 0.0d0)
             v = deck_real(d, 'y_vel_init', 0.0d0)
             w = deck_real(d, 'z_vel_init', 0.0d0)
             temp = deck_real(d, 'tr_temp_init', 1.0d0)
          end select
          if (dens == 0.0d0) then
             r%f0(:, c) = 0.0d0
             cycle
          end if
          opt = cfg%mass / (pi_ref * temp)
          coeff = dens * sqrt(opt * opt * opt)
          ! compute_bkw at time 0 (initial_conditions.f90:483-485)
          xk = 1.0d0 - 0.4d0 * exp(-0.0d0 / 6.0d0)
          bnorm = 0.5d0 * dens * sqrt(opt * opt * opt) * xk**(-2.5d0)
          c2 = cfg%mass / temp
          s = 0.0d0
          i = 0
          do k = 1, npt(3)
             z = r%vz(k)
             z1 = z - beta(r%vz, k)
             if (k > 1) z1 = r%vz(k - 1)
             z2 = z + beta(r%vz, k)
             if (k < npt(3)) z2 = r%vz(k + 1)
             do j = 1, npt(2)
                y = r%vy(j)
                y1 = y - beta(r%vy, j)
                if (j > 1) y1 = r%vy(j - 1)
                y2 = y + beta(r%vy, j)
                if (j < npt(2)) y2 = r%vy(j + 1)
                do gt = 1, npt(1)
                   x = r%vx(gt)
                   i = i + 1
                   b = beta(r%vx, gt) * beta(r%vy, j) * beta(r%vz, k)
                   select case (initdf)
                   case (1)
                      ! compute_maxwellian (dist_func.f90:317-344)
                      opt = (x - u)**2 + (y - v)**2 + (z - w)**2
                      r%f0(i, c) = coeff * exp(-opt * cfg%mass / temp)
                   case (2)
                      ! compute_bkw (dist_func.f90:374-401)
                      opt = (x - u)**2 + (y - v)**2 + (z - w)**2
                      r%f0(i, c) = bnorm * (5.0d0 * xk - 3.0d0 + 2.0d0 * (1.0d0 - xk) * opt * cfg%mass / (xk * temp)) * &
                           exp(-opt * cfg%mass / (xk * temp))
                   case (3)
                      ! a delta function at the node (x/y/z_delta_loc)
                      if (abs(dl(1) - x) < 1.0d-13 .and. abs(dl(2) - y) < 1.0d-13 .and. abs(dl(3) - z) < 1.0d-13) then
                         r%f0(i, c) = dens / b
                      else
                         r%f0(i, c) = 0.0d0
                      end if
                   case default
                      ! compute_exact_gaussian (dist_func.f90:346-372): the
                      ! Maxwellian integrated over the node's cell
                      x1 = x - beta(r%vx, gt)
                      if (gt > 1) x1 = r%vx(gt - 1)
                      x2 = x + beta(r%vx, gt)
                      if (gt < npt(1)) x2 = r%vx(gt + 1)
                      r%f0(i, c) = 0.125d0 * coeff * (pi_ref / c2)**(1.5d0) * &
                           (erf(sqrt(c2) * (x + 0.5d0 * (x2 - x) - u)) - erf(sqrt(c2) * (x1 + 0.5d0 * (x - x1) - u))) * &
                           (erf(sqrt(c2) * (y1 + 0.5d0 * (y - y1) - v)) - erf(sqrt(c2) * (y + 0.5d0 * (y2 - y) - v))) * &
                           (erf(sqrt(c2) * (z1 + 0.5d0 * (z - z1) - w)) - erf(sqrt(c2) * (z + 0.5d0 * (z2 - z) - w)))
                   end select
                   r%f0(i, c) = r%f0(i, c) * b
                   s = s + r%f0(i, c)
                end do
             end do
          end do
          if (s == 0.0d0) then
             msg = 'the initial distribution of a cell is zero everywhere (a delta location that is not a node?)'
             return
          end if
          r%f0(:, c) = (dens / s) * r%f0(:, c)
       end do
    end do
    r%cfg = cfg
    ok = .true.
  end subroutine run_from_deck

  ! First node with |a| < double_tol (1e-13), 0 if none.
  ! read_vel_grid (read_input_2D.f90:779-829) for one species: the point
  ! counts of x, y and z, then the x, y and z values, one per record
  subroutine read_grid_file(fname, ax, ay, az, ok, msg)
    character(len=*), intent(in) :: fname
    real(c_double), allocatable, intent(out) :: ax(:), ay(:), az(:)
    logical, intent(out) :: ok
    character(len=*), intent(out) :: msg
    integer :: u, ios, n(3), i
    ok = .false.
    open(newunit=u, file=trim(fname), status='old', action='read', iostat=ios)
    if (ios /= 0) then
       msg = 'cannot open the velocity-grid file ' // trim(fname)
       return
    end if
    read(u, *, iostat=ios) n(1)
    if (ios == 0) read(u, *, iostat=ios) n(2)
    if (ios == 0) read(u, *, iostat=ios) n(3)
    if (ios /= 0 .or. any(n < 2)) then
       msg = 'the velocity-grid file must start with three point counts >= 2'
       close(u)
       return
    end if
    allocate(ax(n(1)), ay(n(2)), az(n(3)))
    do i = 1, n(1)
       if (ios == 0) read(u, *, iostat=ios) ax(i)
    end do
    do i = 1, n(2)
       if (ios == 0) read(u, *, iostat=ios) ay(i)
    end do
    do i = 1, n(3)
       if (ios == 0) read(u, *, iostat=ios) az(i)
    end do
    close(u)
    if (ios /= 0) then
       msg = 'the velocity-grid file ends before its point values'
       return
    end if
    ok = .true.
  end subroutine read_grid_file

  ! InitC = 4: the shock jump of initialize_normal_shock_props
  ! (shock_conditions.f90:54-171) for one monatomic species (3 d.o.f., so
  ! gamma = cp / (cp - 1) = 5/3, shock_conditions.f90:255-314), applied to the
  ! left wall (initial_conditions.f90:309-336), and the per-column profile
  ! between the two walls (:394-454, initial_shock_profile :652-671).
  subroutine shock_profile(cfg, thick, space_mid, dens_x, u_x, temp_x, ok, msg)
    type(dvm_config), intent(inout) :: cfg
    integer, intent(in) :: thick, space_mid
    real(c_double), intent(out) :: dens_x(:), u_x(:), temp_x(:)
    logical, intent(out) :: ok
    character(len=*), intent(out) :: msg
    real(c_double) :: dens_up, t_up, u_up, gamma, a_up, a, b, c, u_s, m_inf, t_ratio, dens_ratio, m_down
    real(c_double) :: a_down, u_down, dens_lw, u_lw, temp_lw, dens_rw, u_rw, temp_rw, dl, ul, tl, ratio, cp
    integer :: nx, half, dmid, tmid
    logical :: reset
    ok = .false.
    nx = size(dens_x)
    ! calculate_upstream_conditions: the right wall, or the left when the
    ! right wall's flow is positive
    dens_up = cfg%wall(1, 2); u_up = cfg%wall(2, 2); t_up = cfg%wall(5, 2)
    if (u_up > 0.0d0) then
       dens_up = cfg%wall(1, 1); u_up = cfg%wall(2, 1); t_up = cfg%wall(5, 1)
    end if
    cp = 3.0d0 * 0.5d0 + 1.0d0
    gamma = cp / (cp - 1.0d0)
    a_up = sqrt(gamma * t_up / (2.0d0 * cfg%mass))
    if (cfg%bc_left == DVM_BC_WALL .or. cfg%bc_right == DVM_BC_WALL) then
       ! a shock moving away from a reflecting wall
       a = 1.0d0
       b = 0.5d0 * (3.0d0 - gamma) * u_up
       c = -0.5d0 * (gamma + (gamma - 1.0d0) * u_up * u_up)
       u_s = abs((-b - sqrt(b * b - 4.0d0 * a * c)) / (2.0d0 * a))
       m_inf = (abs(u_up) + u_s) / a_up
    else
       if (u_up > 0.0d0) then
          msg = 'InitC 4: the flow of a stationary shock must be negative (towards -x)'
          return
       end if
       u_s = 0.0d0
       m_inf = abs(u_up) / a_up
    end if
    if (m_inf < 1.0d-13) then
       t_ratio = 1.0d0; dens_ratio = 1.0d0; u_down = 0.0d0
    else
       t_ratio = (1.0d0 + 2.0d0 * gamma * (m_inf * m_inf - 1.0d0) / (gamma + 1.0d0)) * &
            (2.0d0 + (gamma - 1.0d0) * m_inf * m_inf) / ((gamma + 1.0d0) * m_inf * m_inf)
       dens_ratio = (gamma + 1.0d0) * m_inf * m_inf / (2.0d0 + (gamma - 1.0d0) * m_inf * m_inf)
       m_down = sqrt((1.0d0 + 0.5d0 * (gamma - 1.0d0) * m_inf * m_inf) / &
            (gamma * m_inf * m_inf - 0.5d0 * (gamma - 1.0d0)))
       a_down = sqrt(gamma * t_ratio * t_up / (2.0d0 * cfg%mass))
       u_down = sign(1.0d0, u_up) * a_down * m_down + u_s
    end if
    dens_rw = cfg%wall(1, 2); u_rw = cfg%wall(2, 2); temp_rw = cfg%wall(5, 2)
    dens_lw = dens_ratio * dens_rw
    u_lw = u_down
    temp_lw = t_ratio * temp_rw
    cfg%wall(1, 1) = dens_lw
    cfg%wall(2, 1) = u_lw
    cfg%wall(5, 1) = temp_lw
    half = ceiling(0.5d0 * dble(thick))
    if (half < 1) then
       msg = 'InitC 4 needs shock_thickness >= 1'
       return
    end if
    dmid = space_mid
    tmid = ceiling(dble(space_mid) + 0.5d0 * dble(half))
    dl = dens_lw; ul = u_lw; tl = temp_lw
    reset = .false.
    do nx = 1, size(dens_x)
       if (.not. reset) then
          if (nx > dmid + half) then
             dl = dens_rw; ul = u_rw
          else if (nx > dmid) then
             dl = profile(dens_lw, dens_rw, half, dmid, nx)
             ratio = (dl - dens_rw) / (dens_lw - dens_rw)
             ul = u_rw + (u_lw - u_rw) * ratio
          else if (nx >= dmid - half) then
             dl = profile(dens_rw, dens_lw, half, nx, dmid)
             ratio = (dl - dens_rw) / (dens_lw - dens_rw)
             ul = u_rw + (u_lw - u_rw) * ratio
          else
             dl = dens_lw; ul = u_lw
          end if
          if (nx > tmid + half) then
             tl = temp_rw
             reset = .true.
          else if (nx > tmid) then
             tl = profile(temp_lw, temp_rw, half, tmid, nx)
          else if (nx >= tmid - half) then
             tl = profile(temp_rw, temp_lw, half, nx, tmid)
          else
             tl = temp_lw
          end if
       end if
       dens_x(nx) = dl; u_x(nx) = ul; temp_x(nx) = tl
    end do
    ok = .true.
  contains
    ! initial_shock_profile (initial_conditions.f90:652-671)
    real(c_double) function profile(lw, rw, th, midpoint, n)
      real(c_double), intent(in) :: lw, rw
      integer, intent(in) :: th, midpoint, n
      real(c_double) :: expo, lx
      expo = 0.5d0 * (lw - rw) / dble(th * th)
      lx = dble(midpoint - n + th)
      profile = rw + expo * lx * lx
    end function profile
  end subroutine shock_profile

  integer function zero_node(a)
    real(c_double), intent(in) :: a(:)
    integer :: i
    zero_node = 0
    do i = 1, size(a)
       if (abs(a(i)) < 1.0d-13) then
          zero_node = i
          return
       end if
    end do
  end function zero_node

  ! velocity_grid.f90:486-540 compute_beta
  real(c_double) function beta(a, i)
    real(c_double), intent(in) :: a(:)
    integer, intent(in) :: i
    integer :: m
    m = size(a)
    if (i == 1) then
       beta = a(2) - a(1)
    else if (i == m) then
       beta = a(m) - a(m - 1)
    else
       beta = (a(i + 1) - a(i)) / 2.0d0 + (a(i) - a(i - 1)) / 2.0d0
    end if
  end function beta

  ! reference_values.txt in the format of scaling.f90:203-237.
  subroutine write_reference_values(r, filename)
    type(run_t), intent(in) :: r
    character(len=*), intent(in) :: filename
    integer :: u
    real(c_double) :: press, hf, e
    press = 2.0d0 * kb * r%dens_ref * r%temp_ref
    hf = r%mass_ref * r%c_ref * r%c_ref * r%c_ref
    e = r%mass_ref * r%c_ref * r%c_ref
    open(newunit=u, file=filename, status='replace', action='write')
    write(u, '(a)') 'Reference values:'
    write(u, '(a7,es16.8)') 'Mass = ', r%mass_ref
    write(u, '(a11,es16.8)') 'Diameter = ', r%diam_ref
    write(u, '(a16,es16.8)') 'Cross-section = ', r%sigma_ref
    write(u, '(a10,es16.8)') 'Density = ', r%dens_ref
    write(u, '(a14,es16.8)') 'Temperature = ', r%temp_ref
    write(u, '(a8,es16.8)') 'Speed = ', r%c_ref
    write(u, '(a11,es16.8)') 'Pressure = ', press
    write(u, '(a12,es16.8)') 'Heat Flux = ', hf
    write(u, '(a9,es16.8)') 'Energy = ', e
    write(u, '(a22,es16.8)') 'Flow Knudsen number = ', r%cfg%kn
    write(u, '(a30,es16.8)') 'Characteristic length scale = ', r%l_scale
    write(u, '(a13,es16.8)') 'Time scale = ', r%t_scale
    close(u)
  end subroutine write_reference_values

  ! Header of <properties_vis_filename>.dat (visualization.f90:615-645).
  subroutine props_header(r, end_time)
    type(run_t), intent(in) :: r
    integer, intent(in) :: end_time
    integer :: u, i, print_times
    character(len=24) :: name
    open(newunit=u, file=trim(adjustl(r%props_file)) // '.dat', status='replace', action='write')
    write(u, advance='yes', fmt='(a)') 'title = "Macroscopic Properties"'
    write(u, advance='no', fmt='(a)') 'variables = "Time", "x-Location", "y-Location"'
    do i = 1, 27
       name = prop_names(i)
       ! the names are blank-padded variables, as in the reference: a14 keeps
       ! their leftmost 14 characters
       if (r%to_output(i)) write(u, advance='no', fmt='(a3,a14)') ', ', name
    end do
    do i = 1, r%n_stress
       write(u, advance='no', fmt='(a7,i2,a1)') ', "Tau ', r%stress_dirs(i), '"'
    end do
    do i = 1, r%n_moment
       write(u, advance='no', fmt='(a10,i2,a1)') ', "Moment ', r%moment_nums(i), '"'
    end do
    write(u, advance='yes', fmt='(1x)')
    if (r%cfg%nx == 1) then
       if (mod(end_time, r%props_freq) == 0) then
          print_times = end_time / r%props_freq + 1
       else
          print_times = end_time / r%props_freq + 2
       end if
       write(u, advance='yes', fmt='(a7,i4)') 'zone i=', print_times
    end if
    close(u)
  end subroutine props_header

  ! One record of the total-mixture file (visualization.f90:1591-1646 with the
  ! values of get_data_to_output :2296-2633 for one monatomic species: no
  ! internal energy, and the directional temperatures, DOF, analytic
  ! relaxation temperatures and moments written as zero in that file).
  ! props(DVM_NPROP, ncell), ext(DVM_NEXT, ncell).
  subroutine props_record(r, time, props, ext)
    type(run_t), intent(in) :: r
    real(c_double), intent(in) :: time, props(:,:), ext(:,:)
    integer :: u, nx, ny, c, i, prop
    real(c_double) :: xl, yl, val, dx, dy
    open(newunit=u, file=trim(adjustl(r%props_file)) // '.dat', status='old', position='append', action='write')
    if (r%cfg%nx > 1 .or. r%cfg%ny > 1) write(u, advance='yes', fmt='(a7,i5)') 'zone i=', r%cfg%nx * r%cfg%ny
    dx = r%cfg%dx
    dy = r%cfg%dy
    xl = 0.5 * dx
    do nx = 1, r%cfg%nx
       yl = 0.5 * dy
       do ny = 1, r%cfg%ny
          c = nx + r%cfg%nx * (ny - 1)
          write(u, advance='no', fmt='(3e24.15)') time, xl, yl
          do prop = 1, 27
             if (.not. r%to_output(prop)) cycle
             select case (prop)
             case (1); val = props(1, c)
             case (2); val = props(2, c)
             case (3); val = props(3, c)
             case (4); val = props(4, c)
             case (5); val = sqrt(props(2, c) * props(2, c) + props(3, c) * props(3, c) + props(4, c) * props(4, c))
             case (6); val = props(2, c) / sqrt(0.5d0 * r%gamma * props(5, c) / r%mol_mass)
             case (7); val = props(5, c)
             case (10); val = props(6, c)
             case (14, 17); val = 0.75d0 * props(1, c) * props(5, c)
             case (18); val = ext(10, c)
             case (19); val = ext(1, c)
             case (20); val = ext(2, c)
             case (21); val = ext(3, c)
             case (27); val = 0.5d0 * props(1, c) * props(5, c)
             case default; val = 0.0d0
             end select
             write(u, advance='no', fmt='(e24.16)') val
          end do
          do i = 1, r%n_stress
             select case (r%stress_dirs(i))
             case (11); val = ext(4, c)
             case (22); val = ext(5, c)
             case (33); val = ext(6, c)
             case (12); val = ext(7, c)
             case (13); val = ext(8, c)
             case (23); val = ext(9, c)
             case default; val = 0.0d0
             end select
             write(u, advance='no', fmt='(e24.16)') val
          end do
          do i = 1, r%n_moment
             write(u, advance='no', fmt='(e24.16)') 0.0d0
          end do
          write(u, advance='yes', fmt='(1x)')
          yl = yl + dy
       end do
       xl = xl + dx
    end do
    close(u)
  end subroutine props_record

  ! write_restart (restart.f90:92-145): unformatted sequential records,
  ! ntime first, then the distribution of every cell (one record of
  ! num_points values each, nodes in (i, j, k) order with i fastest). The
  ! reference indexes phi(species, nx) (1-D); cells here run x fastest, then
  ! y, which is the same order for ny = 1.
  subroutine write_restart(filename, ntime, f)
    character(len=*), intent(in) :: filename
    integer, intent(in) :: ntime
    real(c_double), intent(in) :: f(:,:)
    integer :: u, c
    open(newunit=u, file=trim(filename), status='replace', form='unformatted', action='write')
    write(u) ntime
    do c = 1, size(f, 2)
       write(u) f(:, c)
    end do
    close(u)
  end subroutine write_restart

  ! read_restart (restart.f90:147-250) for one monatomic species: ntime, then
  ! one record per cell into f (the caller's shape fixes the sizes).
  subroutine read_restart(filename, ntime, f, ok, msg)
    character(len=*), intent(in) :: filename
    integer, intent(out) :: ntime
    real(c_double), intent(inout) :: f(:,:)
    logical, intent(out) :: ok
    character(len=*), intent(out) :: msg
    integer :: u, c, ios
    ok = .false.
    ntime = 0
    open(newunit=u, file=trim(filename), status='old', form='unformatted', action='read', iostat=ios)
    if (ios /= 0) then
       msg = 'cannot open restart file ' // trim(filename)
       return
    end if
    read(u, iostat=ios) ntime
    do c = 1, size(f, 2)
       if (ios /= 0) exit
       read(u, iostat=ios) f(:, c)
    end do
    close(u)
    if (ios /= 0) then
       msg = 'restart file ' // trim(filename) // ' is shorter than the deck''s grid'
       return
    end if
    ok = .true.
  end subroutine read_restart

  ! write_restart's schedule: every write_freq steps and at the last step
  ! (write_freq = 0: the last step only).
  logical function restart_due(r, ntime, last)
    type(run_t), intent(in) :: r
    integer, intent(in) :: ntime, last
    restart_due = .false.
    if (.not. r%write_rst) return
    if (ntime == last) then
       restart_due = .true.
    else if (r%write_freq > 0) then
       restart_due = mod(ntime, r%write_freq) == 0
    end if
  end function restart_due

  ! File of output location q: <df_vis_filename>_<species>_<x>_<y>.dat
  ! (visualization.f90:505-508, int2char = I31 left-adjusted).
  function df_name(r, q) result(fn)
    type(run_t), intent(in) :: r
    integer, intent(in) :: q
    character(len=256) :: fn
    character(len=32) :: a, b
    write(a, '(i0)') r%loc_x(q)
    write(b, '(i0)') r%loc_y(q)
    fn = trim(adjustl(r%df_file)) // '_1_' // trim(a) // '_' // trim(b) // '.dat'
  end function df_name

  ! write_headers, distribution-function part (visualization.f90:497-560).
  subroutine df_headers(r)
    type(run_t), intent(in) :: r
    integer :: u, q
    if (.not. r%df_vis) return
    do q = 1, r%n_loc
       open(newunit=u, file=trim(df_name(r, q)), status='replace', action='write')
       write(u, advance='yes', fmt='(a)') 'title = "Distribution Function"'
       select case (r%df_dim)
       case (1)
          write(u, advance='no', fmt='(a)') 'variables = "X-Velocity", "Beta3", "Time"'
       case (2)
          write(u, advance='no', fmt='(a)') 'variables = "X-Velocity", "Y-Velocity", "Beta3", "Time"'
       case default
          write(u, advance='no', fmt='(a)') 'variables = "X-Velocity", "Y-Velocity", "Z-Velocity", "Beta3", "Time"'
       end select
       write(u, advance='no', fmt='(a)') ', "Vel DistFunc", "Deviation"'
       write(u, advance='yes', fmt='(1x)')
       close(u)
    end do
  end subroutine df_headers

  ! write_df (visualization.f90:920-1215) for one species: per output cell a
  ! zone of f and f - f_eq on the x line (1), the z = 0 plane (2) or the whole
  ! grid (3); f_eq is the Maxwellian of the cell's density, velocity and
  ! mixture temperature with the reference's pi, times beta^3, evaluated at
  ! (x, 0, 0) or (x, y, 0) for the line and plane. props: (DVM_NPROP, ncell).
  subroutine df_records(r, time, f, props)
    type(run_t), intent(in) :: r
    real(c_double), intent(in) :: time, f(:,:), props(:,:)
    integer :: u, q, c, i, j, k, nvx, nvy, nvz, l
    real(c_double) :: mass, dens, uu, vv, ww, temp, coeff, x, y, z, b3, feq
    if (.not. r%df_vis) return
    nvx = size(r%vx)
    nvy = size(r%vy)
    nvz = size(r%vz)
    mass = r%cfg%mass
    do q = 1, r%n_loc
       c = r%loc_x(q) + r%cfg%nx * (r%loc_y(q) - 1)
       dens = props(DVM_P_DENS, c)
       uu = props(DVM_P_U, c)
       vv = props(DVM_P_V, c)
       ww = props(DVM_P_W, c)
       temp = props(DVM_P_TEMP, c)
       if (dens > 0.0d0) then
          coeff = dens * sqrt(mass * mass * mass) / sqrt(pi_ref * pi_ref * pi_ref * temp * temp * temp)
       else
          coeff = 0.0d0
       end if
       open(newunit=u, file=trim(df_name(r, q)), status='old', position='append', action='write')
       select case (r%df_dim)
       case (1)
          write(u, advance='yes', fmt='(a7,i4)') 'zone i=', nvx
          j = r%j_zero
          k = r%k_zero
          do i = 1, nvx
             x = r%vx(i)
             b3 = beta(r%vx, i) * beta(r%vy, j) * beta(r%vz, k)
             feq = 0.0d0
             if (temp > 0.0d0) feq = maxwell(x, 0.0d0, 0.0d0, coeff, mass, uu, vv, ww, temp)
             feq = feq * b3
             l = i + nvx * ((j - 1) + nvy * (k - 1))
             write(u, advance='no', fmt='(3e24.16)') x, b3, time
             write(u, advance='no', fmt='(2e24.16)') f(l, c), f(l, c) - feq
             write(u, advance='yes', fmt='(1x)')
          end do
       case (2)
          write(u, advance='yes', fmt='(a7,i3,a4,i3)') 'zone i=', nvx, ', j=', nvy
          k = r%k_zero
          do j = 1, nvy
             y = r%vy(j)
             do i = 1, nvx
                x = r%vx(i)
                b3 = beta(r%vx, i) * beta(r%vy, j) * beta(r%vz, k)
                feq = 0.0d0
                if (temp > 0.0d0) feq = maxwell(x, y, 0.0d0, coeff, mass, uu, vv, ww, temp)
                feq = feq * b3
                l = i + nvx * ((j - 1) + nvy * (k - 1))
                write(u, advance='no', fmt='(4e24.16)') x, y, b3, time
                write(u, advance='no', fmt='(2e24.16)') f(l, c), f(l, c) - feq
                write(u, advance='yes', fmt='(1x)')
             end do
          end do
       case default
          write(u, advance='yes', fmt='(a7,i4,a9,i4,a9,i4)') 'zone i=', nvx, ', j=', nvy, ', k=', nvz
          do k = 1, nvz
             z = r%vz(k)
             do j = 1, nvy
                y = r%vy(j)
                do i = 1, nvx
                   x = r%vx(i)
                   b3 = beta(r%vx, i) * beta(r%vy, j) * beta(r%vz, k)
                   feq = 0.0d0
                   if (temp > 0.0d0) feq = maxwell(x, y, z, coeff, mass, uu, vv, ww, temp)
                   feq = feq * b3
                   l = i + nvx * ((j - 1) + nvy * (k - 1))
                   write(u, advance='no', fmt='(5e24.16)') x, y, z, b3, time
                   write(u, advance='no', fmt='(2e24.16)') f(l, c), f(l, c) - feq
                   write(u, advance='yes', fmt='(1x)')
                end do
             end do
          end do
       end select
       close(u)
    end do
  end subroutine df_records

  ! compute_maxwellian (dist_func.f90:317-344)
  real(c_double) function maxwell(x, y, z, coeff, mass, u, v, w, temp)
    real(c_double), intent(in) :: x, y, z, coeff, mass, u, v, w, temp
    real(c_double) :: cx, cy, cz
    cx = (x - u)
    cx = cx * cx
    cy = (y - v)
    cy = cy * cy
    cz = (z - w)
    cz = cz * cz
    maxwell = coeff * exp(-(cx + cy + cz) * mass / temp)
  end function maxwell

  ! The four wall files, created with their headers when the physical grid is
  ! 2-D (initialize_solver, dvm_solver.f90:208-226).
  subroutine wall_headers(r)
    type(run_t), intent(in) :: r
    character(len=6), parameter :: wname(4) = [character(len=6) :: 'left', 'right', 'bottom', 'top']
    integer :: u, s
    if (.not. (r%cfg%nx > 1 .and. r%cfg%ny > 1)) return
    do s = 1, 4
       open(newunit=u, file=trim(wname(s)) // '_wall.dat', status='replace', action='write')
       write(u, fmt='(a)') 'title = "' // trim(wname(s)) // ' wall fluxes"'
       write(u, fmt='(a)') 'variables = "Time", "Location", "Heat Flux"'
       close(u)
    end do
  end subroutine wall_headers

  ! One dump step: a zone line in every wall file (apply_boundary_conditions,
  ! boundary_conditions.f90:563-575), then a record per cell along each
  ! diffuse / specular wall (time, cell-centre location, heat flux;
  ! :920-922 and analogues). q(pos, side) as returned by dvm_wall_heat_flux.
  subroutine wall_records(r, ntime, q)
    type(run_t), intent(in) :: r
    integer, intent(in) :: ntime
    real(c_double), intent(in) :: q(:,:)
    character(len=6), parameter :: wname(4) = [character(len=6) :: 'left', 'right', 'bottom', 'top']
    integer :: u, s, n, k
    logical :: wall
    if (.not. (r%cfg%nx > 1 .and. r%cfg%ny > 1)) return
    do s = 1, 4
       n = r%cfg%ny
       if (s > 2) n = r%cfg%nx
       open(newunit=u, file=trim(wname(s)) // '_wall.dat', status='old', position='append', action='write')
       write(u, advance='yes', fmt='(a7,i4)') 'zone i=', n
       wall = .true.
       if (s == 1) wall = r%cfg%bc_left == DVM_BC_WALL
       if (s == 2) wall = r%cfg%bc_right == DVM_BC_WALL
       if (wall) then
          do k = 1, n
             if (s <= 2) then
                write(u, fmt='(5e24.14)') ntime * r%cfg%dt, (k - 1) * r%cfg%dy + r%cfg%dy * 0.5d0, q(k, s)
             else
                write(u, fmt='(5e24.14)') ntime * r%cfg%dt, (k - 1) * r%cfg%dx + r%cfg%dx * 0.5d0, q(k, s)
             end if
          end do
       end if
       close(u)
    end do
  end subroutine wall_records

  ! Header of the per-species file <properties_vis_filename>_1.dat
  ! (visualization.f90:660-700): names trimmed, not padded as in the total
  ! file; the same 1-D zone line.
  subroutine species_header(r, end_time)
    type(run_t), intent(in) :: r
    integer, intent(in) :: end_time
    integer :: u, i, print_times
    open(newunit=u, file=trim(adjustl(r%props_file)) // '_1.dat', status='replace', action='write')
    write(u, advance='yes', fmt='(a)') 'title = "Macroscopic Properties"'
    write(u, advance='no', fmt='(a)') 'variables = "Time", "x-Location", "y-Location"'
    do i = 1, 27
       if (r%to_output(i)) write(u, advance='no', fmt='(a2,a)') ', ', trim(adjustl(prop_names(i)))
    end do
    do i = 1, r%n_stress
       write(u, advance='no', fmt='(a7,i2,a1)') ', "Tau ', r%stress_dirs(i), '"'
    end do
    do i = 1, r%n_moment
       write(u, advance='no', fmt='(a10,i2,a1)') ', "Moment ', r%moment_nums(i), '"'
    end do
    write(u, advance='yes', fmt='(1x)')
    if (r%cfg%nx == 1) then
       if (mod(end_time, r%props_freq) == 0) then
          print_times = end_time / r%props_freq + 1
       else
          print_times = end_time / r%props_freq + 2
       end if
       write(u, advance='yes', fmt='(a7,i4)') 'zone i=', print_times
    end if
    close(u)
  end subroutine species_header

  ! One record of the per-species file (visualization.f90:1666-1713) with the
  ! species values of get_data_to_output (:2296-2633) for one monatomic
  ! species: as the total file, plus the directional temperatures
  ! (compute_directional_temperature, physical_props.f90:1811-1869), the
  ! analytic relaxation temperatures and compute_moment (:2073-2150), summed
  ! from f in the reference's loop order. Monatomic species leave the rotational
  ! and vibrational collision numbers cZr / cZv unset; columns 24 and 25 take
  ! the limit of a zero collision number, T_eq. Locations start at 0 (:1677-1680)
  ! and the zone line is written only when nx > 1. f(N, ncell).
  subroutine species_record(r, time, f, props, ext)
    type(run_t), intent(in) :: r
    real(c_double), intent(in) :: time, f(:,:), props(:,:), ext(:,:)
    integer :: u, nx, ny, c, i, prop
    real(c_double) :: xl, yl, val, dens, uu, vv, ww, ttr, tr_e, teq, sd24, sd25, td(3)
    open(newunit=u, file=trim(adjustl(r%props_file)) // '_1.dat', status='old', position='append', action='write')
    if (r%cfg%nx > 1) write(u, advance='yes', fmt='(a7,i5)') 'zone i=', r%cfg%nx * r%cfg%ny
    xl = 0.0d0
    do nx = 1, r%cfg%nx
       yl = 0.0d0
       do ny = 1, r%cfg%ny
          c = nx + r%cfg%nx * (ny - 1)
          dens = props(DVM_P_DENS, c)
          uu = props(DVM_P_U, c)
          vv = props(DVM_P_V, c)
          ww = props(DVM_P_W, c)
          ttr = props(DVM_P_TRTEMP, c)
          tr_e = 0.75d0 * dens * ttr
          teq = 4.0d0 * (tr_e + 0.0d0) / (dens * (3.0d0 + 0.0d0))
          sd24 = 0.0d0
          sd25 = 0.0d0
          if (abs(time) >= 1.0d-13) then
             sd24 = teq
             sd25 = teq
          end if
          if (r%to_output(11) .or. r%to_output(12) .or. r%to_output(13)) call dir_temps(r, f(:, c), dens, uu, vv, ww, td)
          write(u, advance='no', fmt='(3e24.15)') time, xl, yl
          do prop = 1, 27
             if (.not. r%to_output(prop)) cycle
             select case (prop)
             case (1); val = dens
             case (2); val = uu
             case (3); val = vv
             case (4); val = ww
             case (5); val = sqrt(uu * uu + vv * vv + ww * ww)
             case (6); val = uu / sqrt(0.5d0 * r%gamma * ttr / r%mol_mass)
             case (7); val = ttr
             case (10); val = props(DVM_P_TEMP, c)
             case (11); val = td(1)
             case (12); val = td(2)
             case (13); val = td(3)
             case (14, 17); val = tr_e
             case (18); val = ext(10, c)
             case (19); val = ext(1, c)
             case (20); val = ext(2, c)
             case (21); val = ext(3, c)
             case (24); val = sd24
             case (25); val = sd25
             case (26)
                if (abs(time) < 1.0d-13) then
                   val = ttr
                else
                   val = (4.0d0 / 3.0d0) * (tr_e - (0.0d0 / 4.0d0) * sd24 - (0.0d0 / 4.0d0) * sd25)
                end if
             case (27); val = 0.5d0 * dens * ttr
             case default; val = 0.0d0
             end select
             write(u, advance='no', fmt='(e24.16)') val
          end do
          do i = 1, r%n_stress
             select case (r%stress_dirs(i))
             case (11); val = ext(4, c)
             case (22); val = ext(5, c)
             case (33); val = ext(6, c)
             case (12); val = ext(7, c)
             case (13); val = ext(8, c)
             case (23); val = ext(9, c)
             case default; val = 0.0d0
             end select
             write(u, advance='no', fmt='(e24.16)') val
          end do
          do i = 1, r%n_moment
             write(u, advance='no', fmt='(e24.16)') moment(r, f(:, c), r%moment_nums(i), dens, uu, vv, ww, ttr)
          end do
          write(u, advance='yes', fmt='(1x)')
          yl = yl + r%cfg%dy
       end do
       xl = xl + r%cfg%dx
    end do
    close(u)
  end subroutine species_record

  ! compute_directional_temperature (physical_props.f90:1811-1869) for x, y, z.
  subroutine dir_temps(r, fc, dens, u, v, w, td)
    type(run_t), intent(in) :: r
    real(c_double), intent(in) :: fc(:), dens, u, v, w
    real(c_double), intent(out) :: td(3)
    integer :: i, j, k, l
    real(c_double) :: cs
    td = 0.0d0
    l = 0
    do k = 1, size(r%vz)
       do j = 1, size(r%vy)
          do i = 1, size(r%vx)
             l = l + 1
             cs = (r%vx(i) - u) * (r%vx(i) - u)
             td(1) = td(1) + fc(l) * cs
             cs = (r%vy(j) - v) * (r%vy(j) - v)
             td(2) = td(2) + fc(l) * cs
             cs = (r%vz(k) - w) * (r%vz(k) - w)
             td(3) = td(3) + fc(l) * cs
          end do
       end do
    end do
    td = (2.0d0 / 3.0d0) * r%cfg%mass * td / dens
  end subroutine dir_temps

  ! compute_moment (physical_props.f90:2073-2150): normalised n-th moment of
  ! the peculiar speed against its Maxwellian value.
  real(c_double) function moment(r, fc, n, dens, u, v, w, temp)
    type(run_t), intent(in) :: r
    real(c_double), intent(in) :: fc(:), dens, u, v, w, temp
    integer, intent(in) :: n
    integer :: i, j, k, l, in
    real(c_double) :: d, nfact, mom_mb, czsq, cysq, cxsq, cc
    d = dble(n)
    nfact = 1.0d0
    in = 1
    do
       if (n + in >= 1) then
          nfact = nfact * dble(n + in)
          in = in - 2
       else
          exit
       end if
    end do
    mom_mb = nfact * temp**(1.5d0 + d) * (2.0d0**(-0.5d0 * d))
    moment = 0.0d0
    l = 0
    do k = 1, size(r%vz)
       czsq = (r%vz(k) - w)
       czsq = czsq * czsq
       do j = 1, size(r%vy)
          cysq = (r%vy(j) - v)
          cysq = cysq * cysq
          do i = 1, size(r%vx)
             l = l + 1
             cxsq = (r%vx(i) - u)
             cxsq = cxsq * cxsq
             cc = sqrt(cxsq + cysq + czsq)
             moment = moment + cc**d * fc(l)
          end do
       end do
    end do
    moment = r%cfg%mass**(0.5d0 * d) * moment / (dens * mom_mb)
  end function moment

end module DVMDeck
