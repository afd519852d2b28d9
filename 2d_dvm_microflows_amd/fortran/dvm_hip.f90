! dvm_hip.f90 -- ISO_C_BINDING view of include/dvm_hip.h for Fortran hosts.
!
! The reference's own C interop style (random_number_generation.f90:41-80,
! physical_props.f90:95-117): bind(c) interfaces, opaque handles passed as
! type(c_ptr), value; scalars by value; arrays by reference.
module DVMHip
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: DVM_OK = 0
  integer(c_int32_t), parameter, public :: DVM_COLL_NONE = 0, DVM_COLL_FMC = 2, DVM_COLL_VR = 3
  integer(c_int32_t), parameter, public :: DVM_BC_WALL = 1, DVM_BC_INOUT = 2, DVM_BC_ZEROGRAD = 3, &
       DVM_BC_COUETTE = 4
  integer(c_int32_t), parameter, public :: DVM_CS_CLARKE = 1, DVM_CS_BIRD = 2, DVM_CS_VKI = 3
  integer(c_int32_t), parameter, public :: DVM_DIM_0D = 0, DVM_DIM_1D = 1, DVM_DIM_2D = 2
  integer(c_int32_t), parameter, public :: DVM_PREC_FP64 = 0, DVM_PREC_FP32 = 1, DVM_PREC_FP32_REMAP32 = 2
  integer(c_int32_t), parameter, public :: DVM_SAMPLES_REFERENCE = 0, DVM_SAMPLES_EXPLICIT = 1
  integer(c_int32_t), parameter, public :: DVM_CONV_UPWIND1 = 1, DVM_CONV_MINMOD2 = 2
  integer, parameter, public :: DVM_NPROP = 8, DVM_NPHASE = 7, DVM_NEXT = 10
  integer, parameter, public :: DVM_P_DENS = 1, DVM_P_U = 2, DVM_P_V = 3, DVM_P_W = 4, DVM_P_TRTEMP = 5, &
       DVM_P_TEMP = 6, DVM_P_PRESSURE = 7, DVM_P_AVGSPEED = 8   ! 1-based positions in a props row

  ! struct dvm_config (field order and types as in the C header)
  type, bind(c), public :: dvm_config
     integer(c_int32_t) :: dim
     integer(c_int32_t) :: nx, ny
     real(c_double)     :: dx, dy
     integer(c_int32_t) :: nvx, nvy, nvz
     type(c_ptr)        :: vx, vy, vz
     real(c_double)     :: dt, kn
     real(c_double)     :: mass, diam, omega, alpha, viscosity, t_viscosity
     integer(c_int32_t) :: csection_coeff
     integer(c_int32_t) :: coll_solver
     real(c_double)     :: cdf_tolerance
     integer(c_int32_t) :: num_repl
     real(c_double)     :: coln_rms
     integer(c_int32_t) :: sample_mode
     integer(c_int64_t) :: samples_per_cell
     integer(c_int64_t) :: seed
     integer(c_int32_t) :: precision
     integer(c_int32_t) :: enable_convection
     integer(c_int32_t) :: bc_left, bc_right
     real(c_double)     :: acc_coeff
     real(c_double)     :: wall(5,4)   ! wall(q, side): C wall[side][q]
     integer(c_int32_t) :: conv_order = 1
  end type dvm_config

  public :: dvm_ctx_create, dvm_ctx_destroy, dvm_last_error, dvm_upload_f, dvm_download_f, &
       dvm_moments, dvm_moments_ext, dvm_collide, dvm_convect, dvm_step, dvm_timing_reset, dvm_timing_read, &
       dvm_set_wall_heat_flux_freq, dvm_wall_heat_flux
  public :: dvm_batch_create, dvm_batch_destroy, dvm_batch_last_error, dvm_batch_admit, dvm_batch_retire, &
       dvm_batch_step, dvm_batch_synchronize, dvm_batch_download_f, dvm_batch_moments, dvm_batch_l1, &
       dvm_batch_tune
  public :: dvm_error_message

  interface
     integer(c_int) function dvm_ctx_create(cfg, ctx) bind(c, name="dvm_ctx_create")
       import :: c_int, c_ptr, dvm_config
       type(dvm_config), intent(in) :: cfg
       type(c_ptr), intent(out) :: ctx
     end function dvm_ctx_create
     subroutine dvm_ctx_destroy(ctx) bind(c, name="dvm_ctx_destroy")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine dvm_ctx_destroy
     type(c_ptr) function dvm_last_error(ctx) bind(c, name="dvm_last_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function dvm_last_error
     integer(c_int) function dvm_upload_f(ctx, species, f) bind(c, name="dvm_upload_f")
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: species
       real(c_double), intent(in) :: f(*)
     end function dvm_upload_f
     integer(c_int) function dvm_download_f(ctx, species, f) bind(c, name="dvm_download_f")
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: species
       real(c_double), intent(out) :: f(*)
     end function dvm_download_f
     integer(c_int) function dvm_moments(ctx, props) bind(c, name="dvm_moments")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: props(*)
     end function dvm_moments
     integer(c_int) function dvm_moments_ext(ctx, ext) bind(c, name="dvm_moments_ext")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: ext(*)
     end function dvm_moments_ext
     integer(c_int) function dvm_set_wall_heat_flux_freq(ctx, every) bind(c, name="dvm_set_wall_heat_flux_freq")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: every
     end function dvm_set_wall_heat_flux_freq
     integer(c_int) function dvm_wall_heat_flux(ctx, q, ntime) bind(c, name="dvm_wall_heat_flux")
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: q(*)
       integer(c_int32_t), intent(out) :: ntime
     end function dvm_wall_heat_flux
     integer(c_int) function dvm_collide(ctx, ntime, samples) bind(c, name="dvm_collide")
       import :: c_int, c_int32_t, c_int64_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ntime
       integer(c_int64_t), intent(out) :: samples(*)
     end function dvm_collide
     integer(c_int) function dvm_convect(ctx, species, ntime) bind(c, name="dvm_convect")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: species, ntime
     end function dvm_convect
     integer(c_int) function dvm_step(ctx, ntime) bind(c, name="dvm_step")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ntime
     end function dvm_step
     ! distance of each cell's f from its equilibrium Maxwellian (no reference counterpart)
     integer(c_int) function dvm_l1(ctx, l1) bind(c, name="dvm_l1")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: l1(*)
     end function dvm_l1
     ! parity mode: uniforms from an explicit list (replaces get_rand,
     ! utils/random_number_generation.f90:140-162, for parity runs)
     integer(c_int) function dvm_set_uniform_stream(ctx, u, n) bind(c, name="dvm_set_uniform_stream")
       import :: c_int, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: u(*)
       integer(c_int64_t), value :: n
     end function dvm_set_uniform_stream
     integer(c_int) function dvm_uniform_stream_position(ctx, consumed) bind(c, name="dvm_uniform_stream_position")
       import :: c_int, c_int64_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(out) :: consumed
     end function dvm_uniform_stream_position
     integer(c_int) function dvm_timing_reset(ctx) bind(c, name="dvm_timing_reset")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function dvm_timing_reset
     integer(c_int) function dvm_timing_read(ctx, phase_ms, calls) bind(c, name="dvm_timing_read")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: phase_ms(DVM_NPHASE)
       integer(c_int64_t), intent(out) :: calls
     end function dvm_timing_read

     ! batched solves (dvm_batch_* of include/dvm_hip.h)
     integer(c_int) function dvm_batch_create(species, max_slots, node_capacity, cell_capacity, b) &
          bind(c, name="dvm_batch_create")
       import :: c_int, c_int32_t, c_int64_t, c_ptr, dvm_config
       type(dvm_config), intent(in) :: species
       integer(c_int32_t), value :: max_slots, cell_capacity
       integer(c_int64_t), value :: node_capacity
       type(c_ptr), intent(out) :: b
     end function dvm_batch_create
     subroutine dvm_batch_destroy(b) bind(c, name="dvm_batch_destroy")
       import :: c_ptr
       type(c_ptr), value :: b
     end subroutine dvm_batch_destroy
     type(c_ptr) function dvm_batch_last_error(b) bind(c, name="dvm_batch_last_error")
       import :: c_ptr
       type(c_ptr), value :: b
     end function dvm_batch_last_error
     integer(c_int) function dvm_batch_admit(b, inst, f0, slot) bind(c, name="dvm_batch_admit")
       import :: c_int, c_int32_t, c_ptr, c_double, dvm_config
       type(c_ptr), value :: b
       type(dvm_config), intent(in) :: inst
       real(c_double), intent(in) :: f0(*)
       integer(c_int32_t), intent(out) :: slot
     end function dvm_batch_admit
     integer(c_int) function dvm_batch_retire(b, slot) bind(c, name="dvm_batch_retire")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: b
       integer(c_int32_t), value :: slot
     end function dvm_batch_retire
     integer(c_int) function dvm_batch_step(b, nsteps) bind(c, name="dvm_batch_step")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: b
       integer(c_int32_t), value :: nsteps
     end function dvm_batch_step
     integer(c_int) function dvm_batch_synchronize(b) bind(c, name="dvm_batch_synchronize")
       import :: c_int, c_ptr
       type(c_ptr), value :: b
     end function dvm_batch_synchronize
     integer(c_int) function dvm_batch_download_f(b, slot, f) bind(c, name="dvm_batch_download_f")
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: b
       integer(c_int32_t), value :: slot
       real(c_double), intent(out) :: f(*)
     end function dvm_batch_download_f
     integer(c_int) function dvm_batch_moments(b, slot, props) bind(c, name="dvm_batch_moments")
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: b
       integer(c_int32_t), value :: slot
       real(c_double), intent(out) :: props(*)
     end function dvm_batch_moments
     integer(c_int) function dvm_batch_l1(b, l1) bind(c, name="dvm_batch_l1")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: b
       real(c_double), intent(out) :: l1(*)
     end function dvm_batch_l1
     integer(c_int) function dvm_batch_tune(b, steps, tiles, ms) bind(c, name="dvm_batch_tune")
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: b
       integer(c_int32_t), value :: steps
       integer(c_int32_t), intent(out) :: tiles(3)
       real(c_double), intent(out) :: ms
     end function dvm_batch_tune
  end interface

contains

  ! Message of the last failing call on ctx (c_null_ptr: the last failing create).
  function dvm_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=512) :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: i
    msg = ''
    p = dvm_last_error(ctx)
    if( .not. c_associated(p) ) return
    call c_f_pointer( p, s, [512] )
    do i = 1, 512
       if( s(i) == c_null_char ) exit
       msg(i:i) = s(i)
    end do
  end function dvm_error_message

end module DVMHip
