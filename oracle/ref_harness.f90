! ref_harness.f90 -- drives the reference's own Fortran modules (compiled from
! the reference sources by oracle/Makefile target `ref`) to produce golden
! vectors. TEST INFRASTRUCTURE ONLY: nothing here is part of the product.
!
! Reads commands from standard input, writes results to standard output with
! doubles printed as their IEEE-754 bit patterns (integer(8)) so that the
! vectors are exact. Each command word is on a line of its own and its
! arguments start on the next line. Commands:
!   grid  type nvx nvy nvz            (type 1 uniform: then 6 bounds; type 0: then the axes)
!   remap n f_off                     then n lines "u v w" (bit patterns)
!   search n_points n_u               then cumul(0:n_points), ref(1000,2) as ints, u values
!   maxw  n                           then n lines of x y z coeff mass u v w temp
!   csec  model coeff m_red omega d1 d2 visc1 visc2 tv1 tv2
!   g2l   l
!   end
program ref_harness
  use iso_fortran_env, only : int64
  use Constants
  use VelocityGrid
  use Remapping
  use SortAndSearch
  use DistFunc, only : compute_maxwellian
  use Scaling
  implicit none

  type(VelocityGridType) :: grid
  type(MappingResultType) :: mapping
  character(len=16) :: cmd
  integer :: gtype, nvx, nvy, nvz, n, q, npts, nu, l, i, j, k
  integer :: model, coeff_kind
  integer(int64) :: b(10)
  double precision :: lo(3), hi(3), f_off, u, v, w, val, sigma
  double precision, allocatable :: ax(:), ay(:), az(:), cumul(:)
  integer, allocatable :: refa(:,:)
  double precision :: d2(2), visc(2), tv(2)
  logical :: have_grid

  have_grid = .false.
  do
     read(*,*) cmd
     select case( trim(cmd) )
     case( 'end' )
        exit
     case( 'grid' )
        read(*,*) gtype, nvx, nvy, nvz
        call initialize_grid_input_arrays( 1 )
        call set_vel_grid_type( gtype, gtype, gtype, 1 )
        call set_vel_grid_num_points( nvx, nvy, nvz, 1 )
        if( gtype .eq. 1 )then
           read(*,*) b(1:6)
           lo = transfer( b(1:3), lo )
           hi = transfer( b(4:6), hi )
           call set_vel_grid_bounds( lo(1), hi(1), lo(2), hi(2), lo(3), hi(3), 1 )
        else
           allocate( ax(nvx), ay(nvy), az(nvz) )
           do q = 1, nvx
              read(*,*) b(1)
              ax(q) = transfer( b(1), val )
           end do
           do q = 1, nvy
              read(*,*) b(1)
              ay(q) = transfer( b(1), val )
           end do
           do q = 1, nvz
              read(*,*) b(1)
              az(q) = transfer( b(1), val )
           end do
           call initialize_file_read_arrays( 1 )
           call set_velocity_vectors( ax, ay, az, 1 )
           deallocate( ax, ay, az )
        end if
        call create_velocity_grid( grid, 1 )
        have_grid = .true.
        write(*,'(a,9i8)') 'GRIDI ', grid%num_points, grid%i_zero, grid%i_neg, grid%i_pos, &
             grid%j_zero, grid%j_neg, grid%j_pos, grid%k_zero, grid%k_neg
        write(*,'(a,2i8)') 'GRIDK ', grid%k_pos, grid%num_points_x
        write(*,'(a,i22)') 'B3MIN ', transfer( grid%beta3_min, 0_int64 )
        do q = 0, grid%num_points_x - 1
           write(*,'(a,2i22)') 'AX ', transfer( grid%x(q), 0_int64 ), transfer( grid%beta_x(q), 0_int64 )
        end do
        do q = 0, grid%num_points_y - 1
           write(*,'(a,2i22)') 'AY ', transfer( grid%y(q), 0_int64 ), transfer( grid%beta_y(q), 0_int64 )
        end do
        do q = 0, grid%num_points_z - 1
           write(*,'(a,2i22)') 'AZ ', transfer( grid%z(q), 0_int64 ), transfer( grid%beta_z(q), 0_int64 )
        end do
     case( 'remap' )
        read(*,*) n, b(1)
        f_off = transfer( b(1), f_off )
        do q = 1, n
           read(*,*) b(1:3)
           u = transfer( b(1), u )
           v = transfer( b(2), v )
           w = transfer( b(3), w )
           call perform_remapping( mapping, u, v, w, f_off, grid )
           write(*,'(a,7i6,7i22)') 'RM ', mapping%stencil(1:7), &
                ( transfer( mapping%mass(i), 0_int64 ), i = 1, 7 )
        end do
     case( 'search' )
        read(*,*) npts, nu
        allocate( cumul(0:npts), refa(1000,2) )
        do q = 0, npts
           read(*,*) b(1)
           cumul(q) = transfer( b(1), val )
        end do
        do q = 1, 1000
           read(*,*) refa(q,1), refa(q,2)
        end do
        do q = 1, nu
           read(*,*) b(1)
           u = transfer( b(1), u )
           call referenced_search( cumul, refa, npts, u, l )
           call binary_search( cumul, npts, u*cumul(npts), i )
           write(*,'(a,2i10)') 'SR ', l, i
        end do
        deallocate( cumul, refa )
     case( 'maxw' )
        read(*,*) n
        do q = 1, n
           read(*,*) b(1:9)
           call compute_maxwellian( transfer(b(1),val), transfer(b(2),val), transfer(b(3),val), &
                transfer(b(4),val), transfer(b(5),val), transfer(b(6),val), transfer(b(7),val), &
                transfer(b(8),val), transfer(b(9),val), val )
           write(*,'(a,i22)') 'MX ', transfer( val, 0_int64 )
        end do
     case( 'csec' )
        read(*,*) model, coeff_kind, b(1:8)
        call set_cross_section( model, coeff_kind )
        d2 = transfer( b(3:4), d2 )
        visc = transfer( b(5:6), visc )
        tv = transfer( b(7:8), tv )
        call calculate_cross_section( sigma, transfer(b(1),val), transfer(b(2),val), d2, visc, tv )
        write(*,'(a,i22)') 'CS ', transfer( sigma, 0_int64 )
     case( 'g2l' )
        read(*,*) l
        call global2local_map( l, grid%i_min, grid%j_min, grid%k_min, grid%num_points_x, &
             grid%num_points_y, grid%div_Nxy, grid%div_Nx, i, j, k )
        write(*,'(a,3i8)') 'GL ', i, j, k
     case default
        write(*,*) 'unknown command ', trim(cmd)
        stop 1
     end select
  end do
  if( have_grid ) call destroy_velocity_grid( grid )
end program ref_harness
