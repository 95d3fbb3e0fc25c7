!< fest_gpu.f90 -- ISO_C_BINDING interface to libfest_hip.so, the MI355X
!< explicit-residual engine (C declarations in include/fest_hip.h).
!<
!< The derived types mirror the C structs field for field; the interfaces
!< bind the C entry points by name.  Arrays are passed as the reference holds
!< them: column-major with the reference's lower bounds, one contiguous
!< plane per variable or component.  fest_check turns a nonzero return code
!< into the print + STOP of the reference's Fatal_error (src/error.h:17).
module fest_gpu
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: FEST_OK = 0
  integer(c_int), parameter, public :: FEST_E_NONPHYSICAL = -1
  integer(c_int), parameter, public :: FEST_E_NAN = -2
  integer(c_int), parameter, public :: FEST_E_CONFIG = -3
  integer(c_int), parameter, public :: FEST_E_HIP = -4
  integer(c_int), parameter, public :: FEST_E_ARG = -5

  integer(c_int32_t), parameter, public :: FEST_FLUX_AUSM = 0, FEST_FLUX_AUSMP = 1, &
      FEST_FLUX_AUSMUP = 2, FEST_FLUX_SLAU = 3, FEST_FLUX_LDFSS0 = 4, FEST_FLUX_VANLEER = 5
  integer(c_int32_t), parameter, public :: FEST_RECON_NONE = 0, FEST_RECON_MUSCL = 1, &
      FEST_RECON_PPM = 2, FEST_RECON_WENO = 3, FEST_RECON_WENO_NM = 4
  integer(c_int32_t), parameter, public :: FEST_LIM_KOREN = 0, FEST_LIM_VAN_ALBADA = 1
  integer(c_int32_t), parameter, public :: FEST_TRANS_NONE = 0, FEST_TRANS_BC = 1, FEST_TRANS_LCTM2015 = 2
  integer(c_int32_t), parameter, public :: FEST_PREC_FP64 = 0, FEST_PREC_FP32_FLUX = 1
  integer(c_int32_t), parameter, public :: FEST_TURB_NONE = 0, FEST_TURB_SST = 1, FEST_TURB_SST2003 = 2, &
                                           FEST_TURB_SA = 3, FEST_TURB_KKL = 4
  integer(c_int32_t), parameter, public :: FEST_TA_NONE = 0, FEST_TA_RK2 = 2, FEST_TA_RK4 = 4, &
      FEST_TA_TVDRK2 = 12, FEST_TA_TVDRK3 = 13, FEST_TA_IMPLICIT = 20

  integer(c_int32_t), parameter, public :: FEST_DL_QP = 1, FEST_DL_TEMP = 2, &
      FEST_DL_RESIDUE = 4, FEST_DL_F = 8, FEST_DL_G = 16, FEST_DL_H = 32, FEST_DL_DT = 64, &
      FEST_DL_MU = 128, FEST_DL_MUT = 256, FEST_DL_GRADS = 512, FEST_DL_F1 = 1024, &
      FEST_DL_DIST = 2048, FEST_DL_CELLS = 4096, FEST_DL_FACES = 8192

  !< type(extent), src/vartypes.f90:21-26
  type, bind(c), public :: fest_extent
    integer(c_int32_t) :: imx, jmx, kmx, n_var
  end type fest_extent

  !< hot-path fields of type(schemetype), src/vartypes.f90:195-236
  type, bind(c), public :: fest_scheme
    integer(c_int32_t) :: flux, recon, turb, time_local, time_accuracy, accur
    integer(c_int32_t) :: lim(3), tlim(3)
    real(c_double) :: global_time_step
    integer(c_int32_t) :: transition
    integer(c_int32_t) :: pbs(3)
    integer(c_int32_t) :: muscl_limiter
  end type fest_scheme

  !< hot-path fields of type(flowtype), src/vartypes.f90:239-290
  type, bind(c), public :: fest_flow
    real(c_double) :: density_inf, x_speed_inf, y_speed_inf, z_speed_inf, pressure_inf
    real(c_double) :: tk_inf, tw_inf, vel_mag, MInf
    real(c_double) :: gm, R_gas, mu_ref, T_ref, Sutherland_temp, Pr, tPr
    integer(c_int32_t) :: mu_sutherland
    real(c_double) :: tv_inf, tu_inf
    real(c_double) :: tkl_inf
    real(c_double) :: tgm_inf
  end type fest_flow

  !< hot-path fields of type(boundarytype), src/vartypes.f90:293-365
  type, bind(c), public :: fest_bc
    integer(c_int32_t) :: id(6)
    real(c_double) :: fixed_density(6), fixed_x_speed(6), fixed_y_speed(6), fixed_z_speed(6)
    real(c_double) :: fixed_pressure(6), fixed_tk(6), fixed_tw(6), fixed_wall_temperature(6)
    real(c_double) :: fixed_Tpressure(6)
    real(c_double) :: fixed_tv(6)
    real(c_double) :: fixed_tkl(6)
    real(c_double) :: fixed_tgm(6)
  end type fest_bc

  !< one interface link of a multiblock grid (the reference's map line,
  !< src/mapping.f90:64-258): the ghost layers of `face` of `block` take the
  !< interior layers next to `from_face` of `from_block`, walked through the
  !< receiving side's node ranges lo1..hi1, lo2..hi2 (and dir_switch)
  type, bind(c), public :: fest_link
    integer(c_int32_t) :: block, face, from_block, from_face
    integer(c_int32_t) :: lo1, hi1, lo2, hi2
    integer(c_int32_t) :: dir_switch
  end type fest_link

  public :: fest_create, fest_destroy, fest_last_error, fest_upload_geometry, fest_set_grid, &
            fest_upload_state, fest_get_next_solution, fest_residual, fest_vjp, fest_time_step, &
            fest_download, fest_resnorm, fest_sync, fest_check, &
            fest_group_create, fest_group_destroy, fest_group_exchange, fest_group_vjp, fest_group_wall_dist, &
            fest_group_resnorm, fest_group_get_next_solution

  interface
    integer(c_int) function fest_create(h, dims, scheme, flow, bc, precision) bind(c, name='fest_create')
      import :: c_int, c_ptr, fest_extent, fest_scheme, fest_flow, fest_bc
      type(c_ptr), intent(out) :: h
      type(fest_extent), intent(in) :: dims
      type(fest_scheme), intent(in) :: scheme
      type(fest_flow), intent(in) :: flow
      type(fest_bc), intent(in) :: bc
      integer(c_int), value :: precision
    end function fest_create

    subroutine fest_destroy(h) bind(c, name='fest_destroy')
      import :: c_ptr
      type(c_ptr), value :: h
    end subroutine fest_destroy

    type(c_ptr) function fest_last_error_c(h) bind(c, name='fest_last_error')
      import :: c_ptr
      type(c_ptr), value :: h
    end function fest_last_error_c

    integer(c_int) function fest_upload_geometry(h, cells4, if4, jf4, kf4, dist) &
        bind(c, name='fest_upload_geometry')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: cells4(*), if4(*), jf4(*), kf4(*)
      type(c_ptr), value :: dist
    end function fest_upload_geometry

    integer(c_int) function fest_set_grid(h, nodes, with_wall_dist) bind(c, name='fest_set_grid')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: nodes(*)
      integer(c_int), value :: with_wall_dist
    end function fest_set_grid

    integer(c_int) function fest_upload_state(h, qp, temp) bind(c, name='fest_upload_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: qp(*)
      type(c_ptr), value :: temp
    end function fest_upload_state

    integer(c_int) function fest_get_next_solution(h, current_iter, cfl) bind(c, name='fest_get_next_solution')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: current_iter
      real(c_double), value :: cfl
    end function fest_get_next_solution

    integer(c_int) function fest_residual(h, current_iter) bind(c, name='fest_residual')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: current_iter
    end function fest_residual

    ! the adjoint of the inviscid residual (no counterpart in the reference):
    ! bar_qp(-2:imx+2,-2:jmx+2,-2:kmx+2,1:5) = (dR/dqp)^T bar_residue; pass
    ! c_null_ptr for bar_nodes
    integer(c_int) function fest_vjp(h, bar_residue, bar_qp, bar_nodes) bind(c, name='fest_vjp')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: bar_residue(*)
      real(c_double), intent(out) :: bar_qp(*)
      type(c_ptr), value :: bar_nodes
    end function fest_vjp

    integer(c_int) function fest_time_step(h, cfl) bind(c, name='fest_time_step')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: cfl
    end function fest_time_step

    !< every array argument is a c_ptr (c_loc of the Fortran array, or
    !< c_null_ptr when its bit is not set in `what`); fi, gj, hk are the
    !< I/J/K face fluxes F, G, H
    integer(c_int) function fest_download(handle, what, qp, temp, residue, fi, gj, hk, delta_t, mu, &
        mu_t, grads, sst_f1, dist, cells4, faces) bind(c, name='fest_download')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: handle
      integer(c_int32_t), value :: what
      type(c_ptr), value :: qp, temp, residue, fi, gj, hk, delta_t, mu, mu_t, grads, sst_f1, dist, &
                            cells4, faces
    end function fest_download

    integer(c_int) function fest_resnorm(h, res_abs) bind(c, name='fest_resnorm')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: res_abs(*)
    end function fest_resnorm

    integer(c_int) function fest_sync(h) bind(c, name='fest_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function fest_sync

    !< blocks: the handles of the group's blocks (type(c_ptr) array)
    integer(c_int) function fest_group_create(g, blocks, nblocks, links, nlinks) bind(c, name='fest_group_create')
      import :: c_int, c_ptr, fest_link
      type(c_ptr), intent(out) :: g
      type(c_ptr), intent(in) :: blocks(*)
      integer(c_int), value :: nblocks
      type(fest_link), intent(in) :: links(*)
      integer(c_int), value :: nlinks
    end function fest_group_create

    subroutine fest_group_destroy(g) bind(c, name='fest_group_destroy')
      import :: c_ptr
      type(c_ptr), value :: g
    end subroutine fest_group_destroy

    integer(c_int) function fest_group_exchange(g) bind(c, name='fest_group_exchange')
      import :: c_int, c_ptr
      type(c_ptr), value :: g
    end function fest_group_exchange

    ! fest_vjp over the group's blocks (no counterpart in the reference):
    ! arrays of c_loc pointers, one residue adjoint and one cell adjoint per
    ! block, and optionally one node adjoint per block
    integer(c_int) function fest_group_vjp(g, bar_residues, bar_qps, bar_nodes) bind(c, name='fest_group_vjp')
      import :: c_int, c_ptr
      type(c_ptr), value :: g
      type(c_ptr), intent(in) :: bar_residues(*)
      type(c_ptr), intent(in) :: bar_qps(*)
      type(c_ptr), value :: bar_nodes  ! c_null_ptr, or c_loc of an array of per-block pointers
    end function fest_group_vjp

    ! wall distance of every block to the no-slip walls of all blocks
    integer(c_int) function fest_group_wall_dist(g) bind(c, name='fest_group_wall_dist')
      import :: c_int, c_ptr
      type(c_ptr), value :: g
    end function fest_group_wall_dist

    integer(c_int) function fest_group_resnorm(g, res_abs) bind(c, name='fest_group_resnorm')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: g
      real(c_double), intent(out) :: res_abs(*)
    end function fest_group_resnorm

    integer(c_int) function fest_group_get_next_solution(g, current_iter, cfl) &
        bind(c, name='fest_group_get_next_solution')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: g
      integer(c_int), value :: current_iter
      real(c_double), value :: cfl
    end function fest_group_get_next_solution
  end interface

contains

  !< message of the last failed call on handle h
  function fest_last_error(h) result(msg)
    type(c_ptr), intent(in) :: h
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: n
    p = fest_last_error_c(h)
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
      if (s(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function fest_last_error

  !< print the engine's message and stop, as Fatal_error does (src/error.h:17)
  subroutine fest_check(rc, h, where)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: h
    character(len=*), intent(in) :: where
    if (rc == FEST_OK) return
    print *, 'fest_gpu: ', where, ' failed (', rc, '): ', fest_last_error(h)
    stop 1
  end subroutine fest_check

end module fest_gpu
