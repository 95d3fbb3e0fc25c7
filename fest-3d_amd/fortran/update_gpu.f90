!< update_gpu.f90 -- the reference's `update` module (src/update.f90) with the
!< explicit time step moved to the MI355X engine.  It keeps the public
!< surface the solver calls (src/solver.f90:126 setup, :184 per iteration):
!<
!<   setup_update(control, scheme, flow, dims)            src/update.f90:85-126
!<   get_next_solution(qp, Temp, residue, delta_t, cells, F, G, H,
!<                     Ifaces, Jfaces, Kfaces, control, scheme, flow, bc, dims)
!<                                                        src/update.f90:129-226
!<
!< and is compiled into the reference build in place of src/update.f90.  The
!< first call creates the device block from the solver's derived types and
!< uploads the geometry (cells, faces, wall distance).  By default every call
!< uploads qp, runs one whole time step on the device (Temp refresh, all RK
!< stages or the LU-SGS update, the stage-1 time step) and downloads qp,
!< Temp, residue, delta_t and the face fluxes F, G, H that find_resnorm and
!< the writers read next.  With gpu_resident = .true. the state stays on the
!< device: qp is uploaded by the first call only, and qp and Temp come back
!< only at the solver's checkpoint iterations (src/read_write/write/
!< dump_solution.f90:42-44) or through gpu_state_to_host; the residual norms
!< then come from the device (resnorm_gpu.f90, the drop-in resnorm module).
!< The parallel LU-SGS variant 'plusgs' stays on the host path.
module update
  use iso_c_binding
  use vartypes
  use wall_dist, only : dist
  use fest_gpu
  implicit none
  private

  !< the device block; read by the drop-in resnorm module
  type(c_ptr), public, protected, save :: engine = c_null_ptr
  !< .true.: the state stays on the device between calls (see above)
  logical, public, save :: gpu_resident = .false.

  public :: setup_update
  public :: get_next_solution
  public :: destroy_update
  public :: gpu_state_to_host

contains

  subroutine setup_update(control, scheme, flow, dims)
    type(controltype), intent(in) :: control
    type(schemetype), intent(in) :: scheme
    type(flowtype), intent(in) :: flow
    type(extent), intent(in) :: dims
    select case (trim(scheme%time_step_accuracy))
      case ("none", "RK2", "RK4", "TVDRK2", "TVDRK3", "implicit")
        continue
      case default
        print *, "update: time_step_accuracy '", trim(scheme%time_step_accuracy), &
                 "' is not on the device path"
        stop 1
    end select
    if (control%n_var /= dims%n_var .or. flow%gm <= 0.) then
      print *, "update: inconsistent n_var or gamma"
      stop 1
    end if
  end subroutine setup_update

  subroutine destroy_update()
    if (c_associated(engine)) call fest_destroy(engine)
    engine = c_null_ptr
  end subroutine destroy_update

  !< schemetype / flowtype / boundarytype strings and values -> the C structs
  subroutine describe(scheme, flow, bc, dims, e, s, f, b)
    type(schemetype), intent(in) :: scheme
    type(flowtype), intent(in) :: flow
    type(boundarytype), intent(in) :: bc
    type(extent), intent(in) :: dims
    type(fest_extent), intent(out) :: e
    type(fest_scheme), intent(out) :: s
    type(fest_flow), intent(out) :: f
    type(fest_bc), intent(out) :: b

    e = fest_extent(dims%imx, dims%jmx, dims%kmx, dims%n_var)
    select case (trim(scheme%scheme_name))   ! src/face/flux/convective/scheme.f90:68-109
      case ("ausm");     s%flux = FEST_FLUX_AUSM
      case ("ausmP");    s%flux = FEST_FLUX_AUSMP
      case ("ausmUP");   s%flux = FEST_FLUX_AUSMUP
      case ("slau");     s%flux = FEST_FLUX_SLAU
      case ("ldfss0");   s%flux = FEST_FLUX_LDFSS0
      case ("van_leer"); s%flux = FEST_FLUX_VANLEER
      case default
        print *, "update: flux scheme '", trim(scheme%scheme_name), "' is not on the device path"
        stop 1
    end select
    select case (trim(scheme%interpolant))   ! src/face/state/face_interpolant.f90:91-104
      case ("none");  s%recon = FEST_RECON_NONE
      case ("muscl"); s%recon = FEST_RECON_MUSCL
      case ("ppm");   s%recon = FEST_RECON_PPM
      case ("weno");  s%recon = FEST_RECON_WENO
      case ("weno_NM");  s%recon = FEST_RECON_WENO_NM
      case default
        print *, "update: interpolant '", trim(scheme%interpolant), "' is not on the device path"
        stop 1
    end select
    select case (trim(scheme%turbulence))
      case ("none"); s%turb = FEST_TURB_NONE
      case ("sst");  s%turb = FEST_TURB_SST
      case ("sst2003");  s%turb = FEST_TURB_SST2003
      case ("sa");  s%turb = FEST_TURB_SA
      case ("kkl");  s%turb = FEST_TURB_KKL
      case default
        print *, "update: turbulence model '", trim(scheme%turbulence), "' is not on the device path"
        stop 1
    end select
    s%time_local = merge(1, 0, scheme%time_stepping_method == 'l')
    select case (trim(scheme%time_step_accuracy))
      case ("RK4"); s%time_accuracy = FEST_TA_RK4
      case ("RK2"); s%time_accuracy = FEST_TA_RK2
      case ("TVDRK2"); s%time_accuracy = FEST_TA_TVDRK2
      case ("TVDRK3"); s%time_accuracy = FEST_TA_TVDRK3
      case ("implicit"); s%time_accuracy = FEST_TA_IMPLICIT
      case default; s%time_accuracy = FEST_TA_NONE
    end select
    s%accur = scheme%accur
    s%lim = [scheme%ilimiter_switch, scheme%jlimiter_switch, scheme%klimiter_switch]
    s%tlim = [scheme%itlimiter_switch, scheme%jtlimiter_switch, scheme%ktlimiter_switch]
    s%pbs = [scheme%iPB_switch, scheme%jPB_switch, scheme%kPB_switch]   ! src/face/state/muscl.f90:231-241
    s%muscl_limiter = FEST_LIM_KOREN
    s%global_time_step = scheme%global_time_step
    select case (trim(scheme%transition))
      case ("none"); s%transition = FEST_TRANS_NONE
      case ("bc"); s%transition = FEST_TRANS_BC
      case ("lctm2015"); s%transition = FEST_TRANS_LCTM2015
      case default
        print *, "update: transition model '", trim(scheme%transition), "' is not on the device path"
        stop 1
    end select

    f%density_inf = flow%density_inf
    f%x_speed_inf = flow%x_speed_inf
    f%y_speed_inf = flow%y_speed_inf
    f%z_speed_inf = flow%z_speed_inf
    f%pressure_inf = flow%pressure_inf
    f%tk_inf = flow%tk_inf
    f%tw_inf = flow%tw_inf
    f%vel_mag = flow%vel_mag
    f%MInf = flow%MInf
    f%gm = flow%gm
    f%R_gas = flow%R_gas
    f%mu_ref = flow%mu_ref
    f%T_ref = flow%T_ref
    f%Sutherland_temp = flow%Sutherland_temp
    f%Pr = flow%Pr
    f%tPr = flow%tPr
    f%mu_sutherland = merge(1, 0, trim(flow%mu_variation) == 'sutherland_law')
    f%tv_inf = flow%tv_inf
    f%tu_inf = flow%tu_inf
    f%tkl_inf = flow%tkl_inf
    f%tgm_inf = flow%tgm_inf

    b%id = bc%id
    b%fixed_density = bc%fixed_density
    b%fixed_x_speed = bc%fixed_x_speed
    b%fixed_y_speed = bc%fixed_y_speed
    b%fixed_z_speed = bc%fixed_z_speed
    b%fixed_pressure = bc%fixed_pressure
    b%fixed_tk = bc%fixed_tk
    b%fixed_tw = bc%fixed_tw
    b%fixed_wall_temperature = bc%fixed_wall_temperature
    b%fixed_Tpressure = bc%fixed_Tpressure
    b%fixed_tv = bc%fixed_tv
    b%fixed_tkl = bc%fixed_tkl
    b%fixed_tgm = bc%fixed_tgm
  end subroutine describe

  !< first call: the device block and its geometry, copied out of the
  !< solver's array-of-structures types into component planes
  subroutine create_engine(cells, Ifaces, Jfaces, Kfaces, scheme, flow, bc, dims)
    type(extent), intent(in) :: dims
    type(celltype), dimension(-2:dims%imx+2,-2:dims%jmx+2,-2:dims%kmx+2), intent(in) :: cells
    type(facetype), dimension(-2:dims%imx+3,-2:dims%jmx+2,-2:dims%kmx+2), intent(in) :: Ifaces
    type(facetype), dimension(-2:dims%imx+2,-2:dims%jmx+3,-2:dims%kmx+2), intent(in) :: Jfaces
    type(facetype), dimension(-2:dims%imx+2,-2:dims%jmx+2,-2:dims%kmx+3), intent(in) :: Kfaces
    type(schemetype), intent(in) :: scheme
    type(flowtype), intent(in) :: flow
    type(boundarytype), intent(in) :: bc
    type(fest_extent) :: e
    type(fest_scheme) :: s
    type(fest_flow) :: f
    type(fest_bc) :: b
    real(c_double), allocatable, target :: c4(:,:,:,:), i4(:,:,:,:), j4(:,:,:,:), k4(:,:,:,:)
    real(c_double), allocatable, target :: d(:,:,:)
    type(c_ptr) :: pdist
    integer(c_int) :: rc

    call describe(scheme, flow, bc, dims, e, s, f, b)
    rc = fest_create(engine, e, s, f, b, 0_c_int)
    call fest_check(rc, engine, 'fest_create')

    allocate(c4(-2:dims%imx+2,-2:dims%jmx+2,-2:dims%kmx+2,4))
    c4(:,:,:,1) = cells%volume
    c4(:,:,:,2) = cells%centerx
    c4(:,:,:,3) = cells%centery
    c4(:,:,:,4) = cells%centerz
    allocate(i4(-2:dims%imx+3,-2:dims%jmx+2,-2:dims%kmx+2,4))
    allocate(j4(-2:dims%imx+2,-2:dims%jmx+3,-2:dims%kmx+2,4))
    allocate(k4(-2:dims%imx+2,-2:dims%jmx+2,-2:dims%kmx+3,4))
    i4(:,:,:,1) = Ifaces%A;  i4(:,:,:,2) = Ifaces%nx; i4(:,:,:,3) = Ifaces%ny; i4(:,:,:,4) = Ifaces%nz
    j4(:,:,:,1) = Jfaces%A;  j4(:,:,:,2) = Jfaces%nx; j4(:,:,:,3) = Jfaces%ny; j4(:,:,:,4) = Jfaces%nz
    k4(:,:,:,1) = Kfaces%A;  k4(:,:,:,2) = Kfaces%nx; k4(:,:,:,3) = Kfaces%ny; k4(:,:,:,4) = Kfaces%nz
    pdist = c_null_ptr
    if (allocated(dist)) then  ! src/wall/wall_dist.f90:22, (-2:imx+2,-2:jmx+2,-2:kmx+2)
      allocate(d(-2:dims%imx+2,-2:dims%jmx+2,-2:dims%kmx+2))
      d = dist
      pdist = c_loc(d)
    end if
    rc = fest_upload_geometry(engine, c4, i4, j4, k4, pdist)
    call fest_check(rc, engine, 'fest_upload_geometry')
    deallocate(c4, i4, j4, k4)
    if (allocated(d)) deallocate(d)
  end subroutine create_engine

  subroutine get_next_solution(qp, Temp, residue, delta_t, cells, F, G, H, Ifaces, Jfaces, Kfaces, &
                               control, scheme, flow, bc, dims)
    type(controltype), intent(in) :: control
    type(schemetype), intent(in) :: scheme
    type(flowtype), intent(in) :: flow
    type(boundarytype), intent(in) :: bc
    type(extent), intent(in) :: dims
    real(wp), dimension(-2:dims%imx+2,-2:dims%jmx+2,-2:dims%kmx+2,1:dims%n_var), intent(inout), target :: qp
    real(wp), dimension(-2:dims%imx+2,-2:dims%jmx+2,-2:dims%kmx+2), intent(inout), target :: Temp
    real(wp), dimension(:, :, :, :), intent(inout), target, contiguous :: residue
    real(wp), dimension(1:dims%imx-1, 1:dims%jmx-1, 1:dims%kmx-1), intent(inout), target :: delta_t
    real(wp), dimension(:, :, :, :), intent(inout), target, contiguous :: F
    real(wp), dimension(:, :, :, :), intent(inout), target, contiguous :: G
    real(wp), dimension(:, :, :, :), intent(inout), target, contiguous :: H
    type(celltype), dimension(-2:dims%imx+2,-2:dims%jmx+2,-2:dims%kmx+2), intent(in) :: cells
    type(facetype), dimension(-2:dims%imx+3,-2:dims%jmx+2,-2:dims%kmx+2), intent(in) :: Ifaces
    type(facetype), dimension(-2:dims%imx+2,-2:dims%jmx+3,-2:dims%kmx+2), intent(in) :: Jfaces
    type(facetype), dimension(-2:dims%imx+2,-2:dims%jmx+2,-2:dims%kmx+3), intent(in) :: Kfaces
    integer(c_int) :: rc
    integer(c_int32_t) :: what

    logical :: first

    first = .not. c_associated(engine)
    if (first) call create_engine(cells, Ifaces, Jfaces, Kfaces, scheme, flow, bc, dims)
    if (first .or. .not. gpu_resident) then
      rc = fest_upload_state(engine, qp, c_null_ptr)
      call fest_check(rc, engine, 'fest_upload_state')
    end if
    rc = fest_get_next_solution(engine, int(control%current_iter, c_int), real(control%CFL, c_double))
    call fest_check(rc, engine, 'fest_get_next_solution')
    if (gpu_resident) then
      ! the state the checkpoint is about to dump (dump_solution.f90:42-44)
      if (control%checkpoint_iter /= 0) then
        if (mod(control%current_iter, control%checkpoint_iter) == 0 .or. &
            control%current_iter == control%max_iters) call gpu_state_to_host(qp, Temp, dims)
      end if
      return
    end if
    what = ior(ior(ior(FEST_DL_QP, FEST_DL_TEMP), ior(FEST_DL_RESIDUE, FEST_DL_DT)), &
               ior(ior(FEST_DL_F, FEST_DL_G), FEST_DL_H))
    rc = fest_download(engine, what, c_loc(qp), c_loc(Temp), c_loc(residue), c_loc(F), c_loc(G), &
                       c_loc(H), c_loc(delta_t), c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                       c_null_ptr, c_null_ptr, c_null_ptr)
    call fest_check(rc, engine, 'fest_download')
  end subroutine get_next_solution

  !< qp and Temp of the device block into the host arrays
  subroutine gpu_state_to_host(qp, Temp, dims)
    type(extent), intent(in) :: dims
    real(wp), dimension(-2:dims%imx+2,-2:dims%jmx+2,-2:dims%kmx+2,1:dims%n_var), intent(inout), target :: qp
    real(wp), dimension(-2:dims%imx+2,-2:dims%jmx+2,-2:dims%kmx+2), intent(inout), target :: Temp
    integer(c_int) :: rc
    if (.not. c_associated(engine)) return
    rc = fest_download(engine, ior(FEST_DL_QP, FEST_DL_TEMP), c_loc(qp), c_loc(Temp), c_null_ptr, c_null_ptr, &
                       c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                       c_null_ptr, c_null_ptr, c_null_ptr)
    call fest_check(rc, engine, 'fest_download')
  end subroutine gpu_state_to_host

end module update
