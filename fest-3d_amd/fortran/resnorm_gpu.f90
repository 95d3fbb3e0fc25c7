!< resnorm_gpu.f90 -- the reference's `resnorm` module (src/resnorm.f90)
!< beside the device `update` module (update_gpu.f90).  Same public surface:
!<
!<   setup_resnorm(files, control, scheme, flow)      src/resnorm.f90:40-53
!<   find_resnorm(file_handler, residue, F, G, H, control, scheme, dims)
!<                                                    src/resnorm.f90:62-91
!<   Res_abs, Res_rel
!<
!< When the update module keeps the state on the device (gpu_resident), the
!< absolute norms -- sum(R_l^2)/scale_l^2 and the boundary mass imbalance,
!< assembled as one process -- come from the device (fest_resnorm) and the
!< host arrays passed in are not read; otherwise they are summed from those
!< arrays.  The relative norms (against the norms of the first Res_itr = 3
!< iterations, or the restart's previous_Res) and the resnorm file's columns
!< follow src/resnorm.f90:227-377.  One process: the MPI gather of the
!< reference's multi-process runs is not part of this module.
module resnorm
  use iso_c_binding
  use vartypes
  use fest_gpu, only : fest_resnorm, fest_check
  use update, only : engine, gpu_resident
  implicit none
  private

  integer, parameter :: Res_itr = 3
  real(wp), dimension(:), allocatable, target :: Res_abs, Res_rel, Res_save, Res_scale

  public :: setup_resnorm, find_resnorm, Res_abs, Res_rel

contains

  subroutine setup_resnorm(files, control, scheme, flow)
    type(filetype), intent(in) :: files
    type(controltype), intent(in) :: control
    type(schemetype), intent(in) :: scheme
    type(flowtype), intent(in) :: flow
    integer :: l, n
    n = control%n_var
    if (allocated(Res_abs)) deallocate(Res_abs, Res_rel, Res_save, Res_scale)
    allocate(Res_abs(0:n), Res_rel(0:n), Res_save(0:n), Res_scale(0:n))
    Res_abs = 0.; Res_rel = 0.; Res_save = 1.; Res_scale = 1.
    ! scales, src/resnorm.f90:128-169
    Res_scale(1) = flow%density_inf*flow%vel_mag
    Res_scale(2:4) = flow%density_inf*flow%vel_mag*flow%vel_mag
    Res_scale(5) = 0.5*flow%density_inf*flow%vel_mag**3 + (flow%gm/(flow%gm - 1.))*flow%pressure_inf
    select case (trim(scheme%turbulence))
      case ('sst', 'sst2003', 'kw', 'des')
        Res_scale(6) = flow%density_inf*flow%vel_mag*flow%tk_inf
        Res_scale(7) = flow%density_inf*flow%vel_mag*flow%tw_inf
      case ('kkl')
        Res_scale(6) = flow%density_inf*flow%vel_mag*flow%tk_inf
        Res_scale(7) = flow%density_inf*flow%vel_mag*flow%tkl_inf
      case ('sa', 'saBC')
        Res_scale(6) = flow%density_inf*flow%vel_mag*flow%tv_inf
      case default
        continue
    end select
    ! lctm2015: the reference leaves Res_scale(8) unset; the device scales
    ! the intermittency's norm by rho V gamma_inf, and so does this table
    if (trim(scheme%transition) == 'lctm2015' .and. n >= 8) &
      Res_scale(8) = flow%density_inf*flow%vel_mag*flow%tgm_inf
    if (process_id == 0) then
      if (control%start_from == 0) then
        open(files%RESNORM_FILE_UNIT, file=files%resnorm_file)
      else
        open(files%RESNORM_FILE_UNIT, file=files%resnorm_file, status='old', position='append', action='write')
      end if
      write(files%RESNORM_FILE_UNIT, '(A,2x)', advance='no') "Iteration"
      do l = 1, control%Res_count
        write(files%RESNORM_FILE_UNIT, '(A,2x)', advance='no') trim(control%Res_list(l))
      end do
      write(files%RESNORM_FILE_UNIT, *)
    end if
  end subroutine setup_resnorm

  subroutine find_resnorm(file_handler, residue, F, G, H, control, scheme, dims)
    integer, intent(in) :: file_handler
    type(controltype), intent(inout) :: control
    type(schemetype), intent(in) :: scheme
    type(extent), intent(in) :: dims
    real(wp), dimension(:, :, :, :), intent(in) :: residue, F, G, H
    integer :: l
    integer(c_int) :: rc
    if (gpu_resident .and. c_associated(engine)) then
      rc = fest_resnorm(engine, Res_abs)
      call fest_check(rc, engine, 'fest_resnorm')
    else
      do l = 1, control%n_var
        Res_abs(l) = sqrt(sum(residue(:, :, :, l)**2)/Res_scale(l)**2)
      end do
      Res_abs(0) = abs(sum(F(1, 1:dims%jmx-1, 1:dims%kmx-1, 1)) - sum(F(dims%imx, 1:dims%jmx-1, 1:dims%kmx-1, 1)) &
                     + sum(G(1:dims%imx-1, 1, 1:dims%kmx-1, 1)) - sum(G(1:dims%imx-1, dims%jmx, 1:dims%kmx-1, 1)) &
                     + sum(H(1:dims%imx-1, 1:dims%jmx-1, 1, 1)) - sum(H(1:dims%imx-1, 1:dims%jmx-1, dims%kmx, 1)))
    end if
    ! relative to the first iterations' norms, or the restart's
    if (control%current_iter <= Res_itr) Res_save = Res_abs
    if (control%start_from /= 0) then
      Res_save = control%previous_Res
    else
      control%previous_Res = Res_save
    end if
    Res_rel = Res_abs/Res_save
    if ((mod(control%current_iter, control%res_write_interval) == 0 .or. control%current_iter == Res_itr .or. &
         control%current_iter == 1) .and. process_id == 0) call write_norms(file_handler, control, scheme)
  end subroutine find_resnorm

  !< one line of the resnorm file: the iteration, then the columns the
  !< output control lists (src/resnorm.f90:241-377)
  subroutine write_norms(u, control, scheme)
    integer, intent(in) :: u
    type(controltype), intent(in) :: control
    type(schemetype), intent(in) :: scheme
    character(len=20) :: frm
    character(len=16) :: tm
    logical :: two_eq, sa
    integer :: n, l
    n = control%write_percision
    write(frm, '(A,I0,A,I0,A)') "(e", n + 8, ".", n, "E2, 4x)"
    tm = trim(scheme%turbulence)
    two_eq = tm == 'sst' .or. tm == 'kkl' .or. tm == 'sst2003'
    sa = tm == 'sa' .or. tm == 'saBC'
    write(u, '(I0,4x)', advance='no') control%current_iter + control%last_iter
    do l = 1, control%Res_count
      select case (trim(control%Res_list(l)))
        case ('Mass_abs');       call col(Res_abs(0))
        case ('Resnorm_abs');    call col(sqrt(sum(Res_abs(1:)**2)))
        case ('Viscous_abs');    call col(sqrt(sum(Res_abs(1:5)**2)))
        case ('Turbulent_abs');  if (tm /= 'none') call col(sqrt(sum(Res_abs(6:)**2)))
        case ('Continuity_abs'); call col(Res_abs(1))
        case ('X_mom_abs');      call col(Res_abs(2))
        case ('Y_mom_abs');      call col(Res_abs(3))
        case ('Z_mom_abs');      call col(Res_abs(4))
        case ('Energy_abs');     call col(Res_abs(5))
        case ('Mass_rel');       call col(Res_rel(0))
        case ('Resnorm_rel');    call col(sqrt(sum(Res_rel(1:)**2)))
        case ('Viscous_rel');    call col(sqrt(sum(Res_rel(1:5)**2)))
        case ('Turbulent_rel');  if (tm /= 'none') call col(sqrt(sum(Res_rel(6:)**2)))
        case ('Continuity_rel'); call col(Res_rel(1))
        case ('X-mom_rel');      call col(Res_rel(2))
        case ('Y-mom_rel');      call col(Res_rel(3))
        case ('Z-mom_rel');      call col(Res_rel(4))
        case ('Energy_rel');     call col(Res_rel(5))
        case ('TKE_abs');        if (two_eq) call col(Res_abs(6))
        case ('Tv_abs');         if (sa) call col(Res_abs(6))
        case ('Dissipation_abs'); if (tm == 'ke') call col(Res_abs(7))
        case ('Omega_abs');      if (tm == 'sst' .or. tm == 'sst2003') call col(Res_abs(7))
        case ('Kl_abs');         if (tm == 'kkl') call col(Res_abs(7))
        case ('TKE_rel');        if (two_eq) call col(Res_rel(6))
        case ('Tv_rel');         if (sa) call col(Res_rel(6))
        case ('Dissipation_rel'); if (tm == 'ke') call col(Res_rel(7))
        case ('Omega_rel');      if (tm == 'sst' .or. tm == 'sst2003') call col(Res_rel(7))
        case ('Kl_rel');         if (tm == 'kkl') call col(Res_rel(7))
        case default;            call col(sqrt(sum(Res_abs(1:)**2)))
      end select
    end do
    write(u, *)
  contains
    subroutine col(v)
      real(wp), intent(in) :: v
      write(u, frm, advance='no') v
    end subroutine col
  end subroutine write_norms

end module resnorm
