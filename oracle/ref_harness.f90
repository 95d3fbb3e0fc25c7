! ref_harness.f90 -- test-infrastructure driver around the reference's own
! Fortran modules (compiled in place from /root/reference/src by
! oracle/Makefile into oracle/_ref/libfestref.so).  It exposes bind(c)
! entry points so the tests can run the reference's residual passes on the
! same inputs as the oracle and the GPU path.  Nothing here is product code.
!
! The call order in ref_residual is that of get_total_conservative_Residue
! (src/update.f90:534-545).  apply_interface (src/interface1.f90) is not
! called: it needs MPI and is a no-op for a single block without interfaces.
! The boundary setup mirrors setup_bc (src/boundary/bc.f90:18-68) without its
! file reader.
module ref_harness
  use iso_c_binding
  use vartypes
  use geometry,                      only: setup_geometry
  use face_interpolant,              only: x_qp_left, x_qp_right, y_qp_left, y_qp_right, &
                                           z_qp_left, z_qp_right, pdif
  use scheme,                        only: setup_scheme, compute_fluxes, compute_residue
  use bc_primitive,                  only: populate_ghost_primitive
  use face_interpolant,              only: compute_face_interpolant
  use boundary_state_reconstruction, only: reconstruct_boundary_state
  use gradients,                     only: setup_gradients, evaluate_all_gradients, &
                                           gradqp_x, gradqp_y, gradqp_z
  use viscosity,                     only: setup_viscosity, calculate_viscosity, mu, mu_t
  use viscous,                       only: compute_viscous_fluxes
  use source,                        only: add_source_term_residue
  use global_sst,                    only: sst_F1
  use wall_dist,                     only: setup_wall_dist, find_wall_dist, dist
  use lusgs,                         only: setup_lusgs, update_with_lusgs
  use CC,                            only: setupCC, CCnormalX
  implicit none
  private

  type, bind(c) :: hcase
    integer(c_int32_t) :: imx, jmx, kmx, n_var
    integer(c_int32_t) :: flux, recon, turb, time_local, time_accuracy, accur
    integer(c_int32_t) :: lim(3), tlim(3)
    integer(c_int32_t) :: mu_sutherland
    integer(c_int32_t) :: bc_id(6)
    real(c_double) :: global_time_step, CFL
    real(c_double) :: rho_inf, u_inf, v_inf, w_inf, p_inf, tk_inf, tw_inf
    real(c_double) :: vel_mag, MInf
    real(c_double) :: gm, R_gas, mu_ref, T_ref, S_ref, Pr, tPr
    real(c_double) :: fix_rho(6), fix_u(6), fix_v(6), fix_w(6), fix_p(6)
    real(c_double) :: fix_tk(6), fix_tw(6), fix_twall(6)
    real(c_double) :: fix_tp(6)
    real(c_double) :: tv_inf, fix_tv(6)
    integer(c_int32_t) :: transition
    real(c_double) :: tu_inf
    real(c_double) :: tkl_inf, fix_tkl(6)
    integer(c_int32_t) :: pbs(3), muscl_limiter
    real(c_double) :: tgm_inf, fix_tgm(6)
  end type hcase

  type(extent) :: dims
  type(controltype) :: control
  type(schemetype) :: schm
  type(flowtype) :: flow
  type(boundarytype) :: bc
  type(filetype) :: files
  real(wp), allocatable, target :: qp(:,:,:,:), Temp(:,:,:)
  real(wp), allocatable :: F(:,:,:,:), G(:,:,:,:), H(:,:,:,:), residue(:,:,:,:)
  type(celltype), allocatable :: cells(:,:,:)
  type(facetype), allocatable :: Ifaces(:,:,:), Jfaces(:,:,:), Kfaces(:,:,:)
  type(nodetype), allocatable :: nodes(:,:,:)
  integer :: imx, jmx, kmx, nv

  public :: ref_init, ref_set_nodes, ref_set_state, ref_set_dist, ref_residual, ref_get, &
            ref_wall_dist, ref_stage, ref_lusgs

contains

  subroutine ref_init(hc) bind(c, name='ref_init')
    type(hcase), intent(in) :: hc
    character(len=16), parameter :: fluxes(0:5) = [character(len=16) :: 'ausm', 'ausmP', 'ausmUP', &
                                                   'slau', 'ldfss0', 'van_leer']
    character(len=16), parameter :: recons(0:4) = [character(len=16) :: 'none', 'muscl', 'ppm', 'weno', &
                                                   'weno_NM']
    imx = hc%imx; jmx = hc%jmx; kmx = hc%kmx; nv = hc%n_var
    dims%imx = imx; dims%jmx = jmx; dims%kmx = kmx; dims%n_var = nv
    control%n_var = nv
    control%CFL = hc%CFL
    control%current_iter = 1
    schm%scheme_name = trim(fluxes(hc%flux))
    schm%interpolant = trim(recons(hc%recon))
    if (hc%turb == 1) then
      schm%turbulence = 'sst'
    else if (hc%turb == 2) then
      schm%turbulence = 'sst2003'
    else if (hc%turb == 3) then
      schm%turbulence = 'sa'
    else if (hc%turb == 4) then
      schm%turbulence = 'kkl'
    else
      schm%turbulence = 'none'
    end if
    schm%transition = 'none'
    if (hc%transition == 1) schm%transition = 'bc'
    if (hc%transition == 2) schm%transition = 'lctm2015'
    schm%accur = hc%accur
    schm%ilimiter_switch = hc%lim(1); schm%jlimiter_switch = hc%lim(2); schm%klimiter_switch = hc%lim(3)
    schm%itlimiter_switch = hc%tlim(1); schm%jtlimiter_switch = hc%tlim(2); schm%ktlimiter_switch = hc%tlim(3)
    schm%iPB_switch = hc%pbs(1); schm%jPB_switch = hc%pbs(2); schm%kPB_switch = hc%pbs(3)
    if (hc%time_local == 1) then
      schm%time_stepping_method = 'l'
    else
      schm%time_stepping_method = 'g'
    end if
    schm%time_step_accuracy = 'RK4'
    schm%global_time_step = hc%global_time_step
    flow%density_inf = hc%rho_inf; flow%x_speed_inf = hc%u_inf
    flow%y_speed_inf = hc%v_inf; flow%z_speed_inf = hc%w_inf
    flow%pressure_inf = hc%p_inf; flow%tk_inf = hc%tk_inf; flow%tw_inf = hc%tw_inf
    flow%tv_inf = hc%tv_inf
    flow%tkl_inf = hc%tkl_inf
    flow%tu_inf = hc%tu_inf
    flow%tgm_inf = hc%tgm_inf
    flow%vel_mag = hc%vel_mag; flow%MInf = hc%MInf
    flow%gm = hc%gm; flow%R_gas = hc%R_gas; flow%mu_ref = hc%mu_ref
    ! set as state.f90:89 sets it
    if (hc%mu_ref /= 0.0) flow%Reynolds_number = flow%density_inf*flow%vel_mag*1.0/flow%mu_ref
    flow%T_ref = hc%T_ref; flow%Sutherland_temp = hc%S_ref; flow%Pr = hc%Pr; flow%tPr = hc%tPr
    if (hc%mu_sutherland == 1) then
      flow%mu_variation = 'sutherland_law'
    else
      flow%mu_variation = 'constant'
    end if
    ! boundary setup as src/boundary/bc.f90:18-68, values from the case
    bc%imin_id = hc%bc_id(1); bc%imax_id = hc%bc_id(2); bc%jmin_id = hc%bc_id(3)
    bc%jmax_id = hc%bc_id(4); bc%kmin_id = hc%bc_id(5); bc%kmax_id = hc%bc_id(6)
    bc%face_names = [character(len=4) :: 'imin', 'imax', 'jmin', 'jmax', 'kmin', 'kmax']
    bc%id = hc%bc_id
    bc%c2 = 1 + schm%accur
    bc%c3 = 0.5*schm%accur
    bc%c1 = bc%c2 - bc%c3
    bc%fixed_density = hc%fix_rho; bc%fixed_x_speed = hc%fix_u; bc%fixed_y_speed = hc%fix_v
    bc%fixed_z_speed = hc%fix_w; bc%fixed_pressure = hc%fix_p
    bc%fixed_tk = hc%fix_tk; bc%fixed_tw = hc%fix_tw; bc%fixed_wall_temperature = hc%fix_twall
    bc%fixed_Tpressure = hc%fix_tp
    bc%fixed_tv = hc%fix_tv
    bc%fixed_tkl = hc%fix_tkl
    bc%fixed_tgm = hc%fix_tgm
    if (allocated(bc%make_F_flux_zero)) deallocate(bc%make_F_flux_zero, bc%make_G_flux_zero, bc%make_H_flux_zero)
    allocate(bc%make_F_flux_zero(1:imx), bc%make_G_flux_zero(1:jmx), bc%make_H_flux_zero(1:kmx))
    bc%make_F_flux_zero = 1; bc%make_G_flux_zero = 1; bc%make_H_flux_zero = 1
    if (any(bc%imin_id == [-5, -6, -7])) bc%make_F_flux_zero(1) = 0
    if (any(bc%jmin_id == [-5, -6, -7])) bc%make_G_flux_zero(1) = 0
    if (any(bc%kmin_id == [-5, -6, -7])) bc%make_H_flux_zero(1) = 0
    if (any(bc%imax_id == [-5, -6, -7])) bc%make_F_flux_zero(imx) = 0
    if (any(bc%jmax_id == [-5, -6, -7])) bc%make_G_flux_zero(jmx) = 0
    if (any(bc%kmax_id == [-5, -6, -7])) bc%make_H_flux_zero(kmx) = 0

    ! the reference's allocators stop on an already-allocated array
    if (allocated(x_qp_left)) deallocate(x_qp_left, x_qp_right, y_qp_left, y_qp_right, &
                                         z_qp_left, z_qp_right, pdif)
    if (allocated(gradqp_x)) deallocate(gradqp_x, gradqp_y, gradqp_z)
    if (allocated(mu)) deallocate(mu)
    if (allocated(mu_t)) deallocate(mu_t)
    if (allocated(sst_F1)) deallocate(sst_F1)
    if (allocated(qp)) deallocate(qp, Temp, nodes)
    allocate(qp(-2:imx+2, -2:jmx+2, -2:kmx+2, 1:nv), Temp(-2:imx+2, -2:jmx+2, -2:kmx+2))
    allocate(nodes(-2:imx+3, -2:jmx+3, -2:kmx+3))
    qp = 0.0; Temp = 0.0
    call setup_scheme(residue, F, G, H, control, dims)
    call setup_gradients(control, schm, flow, dims)
    call setup_viscosity(schm, flow, dims)
    if (allocated(dist)) deallocate(dist)
    allocate(dist(-2:imx+2, -2:jmx+2, -2:kmx+2))
    dist = 1.0
  end subroutine ref_init

  ! nodes: three planes (x, y, z) over (-2:imx+3,-2:jmx+3,-2:kmx+3), ghost
  ! nodes already filled by the caller
  subroutine ref_set_nodes(x) bind(c, name='ref_set_nodes')
    real(c_double), intent(in) :: x(-2:imx+3, -2:jmx+3, -2:kmx+3, 3)
    nodes(:,:,:)%x = x(:,:,:,1)
    nodes(:,:,:)%y = x(:,:,:,2)
    nodes(:,:,:)%z = x(:,:,:,3)
    call setup_geometry(cells, Ifaces, Jfaces, Kfaces, nodes, bc, dims)
  end subroutine ref_set_nodes

  ! writes the no-slip surface nodes in the reference's ES18.10E3 format
  ! (src/wall/wall.f90:101) and runs setup_wall_dist/find_wall_dist
  subroutine ref_wall_dist(path, plen) bind(c, name='ref_wall_dist')
    character(kind=c_char), intent(in) :: path(*)
    integer(c_int), value :: plen
    integer :: i, j, k, f, n, u
    integer :: lo(3), hi(3)
    character(len=128) :: fname
    fname = ''
    do i = 1, plen
      fname(i:i) = path(i)
    end do
    n = 0
    do f = 1, 6
      if (bc%id(f) /= -5) cycle
      call face_box(f, lo, hi)
      n = n + (hi(1)-lo(1)+1)*(hi(2)-lo(2)+1)*(hi(3)-lo(3)+1)
    end do
    open(newunit=u, file=trim(fname), status='replace')
    write(u, '(I0)') n
    do f = 1, 6
      if (bc%id(f) /= -5) cycle
      call face_box(f, lo, hi)
      do k = lo(3), hi(3)
        do j = lo(2), hi(2)
          do i = lo(1), hi(1)
            write(u, '(3(ES18.10E3,4x))') nodes(i,j,k)%x, nodes(i,j,k)%y, nodes(i,j,k)%z
          end do
        end do
      end do
    end do
    close(u)
    files%surface_node_points = trim(fname)
    if (allocated(dist)) deallocate(dist)
    call setup_wall_dist(files, dims)
    call find_wall_dist(nodes, dims)
    call cc_setup()
  end subroutine ref_wall_dist

  ! the cell-centre normals of the lctm2015 sensor, from the wall distance,
  ! as solver.f90:131-135 sets them up after find_wall_dist
  subroutine cc_setup()
    if (allocated(CCnormalX)) return
    call setupCC(schm, cells, Ifaces, Jfaces, Kfaces, dims)
  end subroutine cc_setup

  subroutine face_box(f, lo, hi)
    integer, intent(in) :: f
    integer, intent(out) :: lo(3), hi(3)
    lo = 1; hi = [imx, jmx, kmx]
    select case (f)
      case (1); hi(1) = 1
      case (2); lo(1) = imx
      case (3); hi(2) = 1
      case (4); lo(2) = jmx
      case (5); hi(3) = 1
      case (6); lo(3) = kmx
    end select
  end subroutine face_box

  subroutine ref_set_state(q, t) bind(c, name='ref_set_state')
    real(c_double), intent(in) :: q(-2:imx+2, -2:jmx+2, -2:kmx+2, nv)
    real(c_double), intent(in) :: t(-2:imx+2, -2:jmx+2, -2:kmx+2)
    qp = q
    Temp = t
  end subroutine ref_set_state

  subroutine ref_set_dist(d) bind(c, name='ref_set_dist')
    real(c_double), intent(in) :: d(-2:imx+2, -2:jmx+2, -2:kmx+2)
    dist = d
    call cc_setup()
  end subroutine ref_set_dist

  ! stage selects how far to go: 1 ghosts, 2 +face states, 3 +fluxes,
  ! 4 +gradients/viscosity/viscous fluxes, 5 +residue and source
  subroutine ref_stage(current_iter, stage) bind(c, name='ref_stage')
    integer(c_int), value :: current_iter, stage
    control%current_iter = current_iter
    call populate_ghost_primitive(qp, Ifaces, Jfaces, Kfaces, control, schm, flow, bc, dims)
    if (stage < 2) return
    call compute_face_interpolant(qp, cells, schm, flow, dims)
    call reconstruct_boundary_state(qp, control, schm, bc, dims)
    if (stage < 3) return
    call compute_fluxes(F, G, H, Ifaces, Jfaces, Kfaces, schm, flow, bc, dims)
    if (stage < 4) return
    if (flow%mu_ref /= 0.0) then
      call evaluate_all_gradients(qp, Temp, cells, Ifaces, Jfaces, Kfaces, schm, bc, dims)
      call calculate_viscosity(qp, schm, flow, bc, dims)
      call compute_viscous_fluxes(F, G, H, qp, cells, Ifaces, Jfaces, Kfaces, schm, flow, dims)
    end if
    if (stage < 5) return
    call compute_residue(residue, F, G, H, dims)
    call add_source_term_residue(qp, residue, cells, Ifaces, Jfaces, Kfaces, schm, flow, dims)
  end subroutine ref_stage

  subroutine ref_residual(current_iter) bind(c, name='ref_residual')
    integer(c_int), value :: current_iter
    call ref_stage(current_iter, 5)
  end subroutine ref_residual

  ! one LU-SGS update (src/lusgs.f90:134-183) of the current state with the
  ! residue of the last ref_stage(.., 5) and the time step dt; the module
  ! allocates its sweep arrays once, so one call per process
  subroutine ref_lusgs(dt) bind(c, name='ref_lusgs')
    real(c_double), intent(in) :: dt(1:imx-1, 1:jmx-1, 1:kmx-1)
    real(wp), allocatable :: delta_t(:,:,:)
    schm%time_step_accuracy = 'implicit'
    call setup_lusgs(control, schm, flow, dims)
    allocate(delta_t(1:imx-1, 1:jmx-1, 1:kmx-1))
    delta_t = dt
    call update_with_lusgs(qp, residue, delta_t, cells, Ifaces, Jfaces, Kfaces, schm, dims)
  end subroutine ref_lusgs

  ! copies one array out in the oracle's layout; ids follow orc_array()
  subroutine ref_get(id, out) bind(c, name='ref_get')
    integer(c_int), value :: id
    real(c_double), intent(out) :: out(*)
    integer :: n
    select case (id)
      case (0); n = size(qp); out(1:n) = reshape(qp, [n])
      case (1); n = size(Temp); out(1:n) = reshape(Temp, [n])
      case (2); n = size(mu); out(1:n) = reshape(mu, [n])
      case (3); n = size(mu_t); out(1:n) = reshape(mu_t, [n])
      case (4); n = size(sst_F1); out(1:n) = reshape(sst_F1, [n])
      case (5); n = size(dist); out(1:n) = reshape(dist, [n])
      case (6); n = size(cells); out(1:n) = reshape(cells(:,:,:)%volume, [n])
      case (7); n = size(cells); out(1:n) = reshape(cells(:,:,:)%centerx, [n])
      case (8); n = size(cells); out(1:n) = reshape(cells(:,:,:)%centery, [n])
      case (9); n = size(cells); out(1:n) = reshape(cells(:,:,:)%centerz, [n])
      case (10); n = size(Ifaces); out(1:n) = reshape(Ifaces(:,:,:)%A, [n])
      case (11); n = size(Ifaces); out(1:n) = reshape(Ifaces(:,:,:)%nx, [n])
      case (12); n = size(Ifaces); out(1:n) = reshape(Ifaces(:,:,:)%ny, [n])
      case (13); n = size(Ifaces); out(1:n) = reshape(Ifaces(:,:,:)%nz, [n])
      case (14); n = size(Jfaces); out(1:n) = reshape(Jfaces(:,:,:)%A, [n])
      case (15); n = size(Jfaces); out(1:n) = reshape(Jfaces(:,:,:)%nx, [n])
      case (16); n = size(Jfaces); out(1:n) = reshape(Jfaces(:,:,:)%ny, [n])
      case (17); n = size(Jfaces); out(1:n) = reshape(Jfaces(:,:,:)%nz, [n])
      case (18); n = size(Kfaces); out(1:n) = reshape(Kfaces(:,:,:)%A, [n])
      case (19); n = size(Kfaces); out(1:n) = reshape(Kfaces(:,:,:)%nx, [n])
      case (20); n = size(Kfaces); out(1:n) = reshape(Kfaces(:,:,:)%ny, [n])
      case (21); n = size(Kfaces); out(1:n) = reshape(Kfaces(:,:,:)%nz, [n])
      case (22); n = size(x_qp_left); out(1:n) = reshape(x_qp_left, [n])
      case (23); n = size(x_qp_right); out(1:n) = reshape(x_qp_right, [n])
      case (24); n = size(y_qp_left); out(1:n) = reshape(y_qp_left, [n])
      case (25); n = size(y_qp_right); out(1:n) = reshape(y_qp_right, [n])
      case (26); n = size(z_qp_left); out(1:n) = reshape(z_qp_left, [n])
      case (27); n = size(z_qp_right); out(1:n) = reshape(z_qp_right, [n])
      case (28); n = size(F); out(1:n) = reshape(F, [n])
      case (29); n = size(G); out(1:n) = reshape(G, [n])
      case (30); n = size(H); out(1:n) = reshape(H, [n])
      case (31); n = size(residue); out(1:n) = reshape(residue, [n])
      case (35); n = size(gradqp_x); out(1:n) = reshape(gradqp_x, [n])
      case (36); n = size(gradqp_y); out(1:n) = reshape(gradqp_y, [n])
      case (37); n = size(gradqp_z); out(1:n) = reshape(gradqp_z, [n])
      case default
        out(1) = 0.0
    end select
  end subroutine ref_get

end module ref_harness
