!< fest_gpu_drive.f90 -- a Fortran host for the engine through the fest_gpu
!< ISO_C_BINDING module: reads a block description, its nodes and its state
!< from a stream file, runs `nsteps` calls of fest_get_next_solution (the
!< device's get_next_solution, src/update.f90:129-226) and writes the state,
!< the residue and the residual norms back.
!<
!<   fest_gpu_drive IN OUT
!<
!< IN (stream, native endianness), in this order:
!<   int32  imx, jmx, kmx, n_var, flux, recon, turb, time_local,
!<          time_accuracy, accur, lim(3), tlim(3), mu_sutherland, bc_id(6),
!<          nsteps, with_wall_dist, transition, pbs(3), muscl_limiter
!<   real64 global_time_step, CFL, density_inf, x_speed_inf, y_speed_inf,
!<          z_speed_inf, pressure_inf, tk_inf, tw_inf, vel_mag, MInf, gm,
!<          R_gas, mu_ref, T_ref, Sutherland_temp, Pr, tPr, tv_inf, tu_inf,
!<          tkl_inf, tgm_inf,
!<          fixed_{density, x_speed, y_speed, z_speed, pressure, tk, tw,
!<          wall_temperature, Tpressure, tv, tkl, tgm}(6)
!<   real64 nodes(-2:imx+3,-2:jmx+3,-2:kmx+3,3), interior 1..imx filled
!<   real64 qp(-2:imx+2,-2:jmx+2,-2:kmx+2,n_var)
!< OUT: real64 qp (same extent), residue(1:imx-1,1:jmx-1,1:kmx-1,n_var),
!<      Res_abs(0:n_var)
program fest_gpu_drive
  use iso_c_binding
  use fest_gpu
  implicit none
  type(fest_extent) :: e
  type(fest_scheme) :: s
  type(fest_flow) :: f
  type(fest_bc) :: b
  type(c_ptr) :: h
  integer(c_int32_t) :: hdr(30)
  real(c_double) :: cfl
  real(c_double), allocatable, target :: nodes(:,:,:,:), qp(:,:,:,:), res(:,:,:,:)
  real(c_double) :: norms(0:7)
  character(len=1024) :: fin, fout
  integer :: u, it, nsteps, wall
  integer(c_int) :: rc

  if (command_argument_count() /= 2) then
    print *, 'usage: fest_gpu_drive IN OUT'
    stop 2
  end if
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)

  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read(u) hdr
  e%imx = hdr(1); e%jmx = hdr(2); e%kmx = hdr(3); e%n_var = hdr(4)
  s%flux = hdr(5); s%recon = hdr(6); s%turb = hdr(7); s%time_local = hdr(8)
  s%time_accuracy = hdr(9); s%accur = hdr(10)
  s%lim = hdr(11:13); s%tlim = hdr(14:16)
  f%mu_sutherland = hdr(17)
  b%id = hdr(18:23)
  nsteps = hdr(24)
  wall = hdr(25)
  s%transition = hdr(26)
  s%pbs = hdr(27:29)
  s%muscl_limiter = hdr(30)
  read(u) s%global_time_step, cfl
  read(u) f%density_inf, f%x_speed_inf, f%y_speed_inf, f%z_speed_inf, f%pressure_inf, &
          f%tk_inf, f%tw_inf, f%vel_mag, f%MInf, f%gm, f%R_gas, f%mu_ref, f%T_ref, &
          f%Sutherland_temp, f%Pr, f%tPr, f%tv_inf, f%tu_inf, f%tkl_inf, f%tgm_inf
  read(u) b%fixed_density, b%fixed_x_speed, b%fixed_y_speed, b%fixed_z_speed, &
          b%fixed_pressure, b%fixed_tk, b%fixed_tw, b%fixed_wall_temperature, b%fixed_Tpressure, &
          b%fixed_tv, b%fixed_tkl, b%fixed_tgm
  allocate(nodes(-2:e%imx+3, -2:e%jmx+3, -2:e%kmx+3, 3))
  allocate(qp(-2:e%imx+2, -2:e%jmx+2, -2:e%kmx+2, e%n_var))
  allocate(res(1:e%imx-1, 1:e%jmx-1, 1:e%kmx-1, e%n_var))
  read(u) nodes
  read(u) qp
  close(u)

  rc = fest_create(h, e, s, f, b, 0_c_int)
  call fest_check(rc, h, 'fest_create')
  rc = fest_set_grid(h, nodes, int(wall, c_int))
  call fest_check(rc, h, 'fest_set_grid')
  rc = fest_upload_state(h, qp, c_null_ptr)
  call fest_check(rc, h, 'fest_upload_state')
  do it = 1, nsteps
    rc = fest_get_next_solution(h, int(it, c_int), cfl)
    call fest_check(rc, h, 'fest_get_next_solution')
  end do
  rc = fest_download(h, ior(FEST_DL_QP, FEST_DL_RESIDUE), c_loc(qp), c_null_ptr, c_loc(res), &
                     c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                     c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr)
  call fest_check(rc, h, 'fest_download')
  norms = 0
  rc = fest_resnorm(h, norms)
  call fest_check(rc, h, 'fest_resnorm')
  call fest_destroy(h)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  write(u) qp
  write(u) res
  write(u) norms(0:e%n_var)
  close(u)
  print '(a,i0,a)', 'fest_gpu_drive: ', nsteps, ' step(s) done'
end program fest_gpu_drive
