! Golden-vector harness for the WAVETRISK reference (test infrastructure only).
!
! Drives the reference's own modules (compiled from /root/reference/src by
! oracle/Makefile.ref) through the initialisation sequence of main.f90
! (init_basic main.f90:23-34, init_structures main.f90:482-524) without the
! driver's checkpoint/restart, sets a uniform single-level grid with every
! node and edge active, applies an analytic initial state, and dumps
!   * every per-domain array (nodes incl. ghost cells) with coordinates,
!   * geometry (len, pedlen, areas, triarea, coriolis, topography),
!   * the state after the first ghost fill and trend_ml's output
!     (multi_level.f90:5-47) with its work arrays,
!   * the state after each of nsteps RK4 steps (time_integr.f90:78-103),
!     or after every `every`-th step (tags numbered step / every).
!
! Namelist &run (stdin): case ('tc5' | 'ocean'), zlev, nsteps, dt, out, ntimed, every.
! vremap ('0', '1', '2PPM', 'PPR', ...; default none): after the steps, the
! reference's remap_vertical_coordinates (remap.f90:18-75) onto the hybrid
! coordinates of coords ('uniform' | 'chebyshev': b_vert as the test cases'
! initialize_a_b_vert, a_vert = 1 - b_vert), dumped with the tag 'r0_'.
! uniform_first (logical): before that remap, a piecewise-constant remap onto
! uniform coordinates (so the old layers are evenly spaced), dumped as 'p0_'.
! vdiff (1 analytic eddy coefficients, 2 TKE closure): after that, two calls of
! the reference's vertical_diffusion (vert_diffusion.f90:66-86, istep 1 then 2,
! dt = step) with position-dependent forcing hooks (vd_forcing, dumped as
! 'vf_*'), each followed by update_bdry and dumps 'v<n>_' of the state and of
! Kt, Kv (interfaces 0..zlevels) and tke ('v0_': the coefficients before).
! vdinline (logical, with vdiff): instead, vertical_diffusion after every RK4
! step (istep = step number, as time_step orders them, main.f90:384-420); the
! step dumps hold the state after both, 'v9_' the coefficients at the end.
! ntimed > 0: after the dumps, ntimed more RK4 steps timed by the wall clock
! (system_clock) and printed as "RK4 <ms> ms/step" (the CPU reference's cost).
! lsclr, ldivu, lrotu (0, 1, 2) and vsclr(2), vdivu, vrotu: Laplacian
! diffusion (order 1) or iterated hyperdiffusion (order 2) through the
! physics hooks, in the form the reference's test cases give them (a scalar
! flux (-1)^Laplace_sclr visc_sclr grad(q or Laplacian_scalar) pedlen, a
! velocity source (-1)^(Laplace_rotu-1) (visc_divu grad divu - visc_rotu
! curl rotu)); the reference computes divu, the Laplacians and rot rot of
! the vorticity itself (multi_level.f90:26-30, 99, 190-309, 579-617).
! Output: one stream file of tagged records (tag*16, kind int32 (1=i32,
! 2=f64), count int64, data), read by tests/golden/make_golden.py.
module harness_hooks
  use domain_mod
  use geom_mod
  implicit none
  real(8), target :: vd_friction = 4d-4  ! bottom_friction of the vertical diffusion
contains
  function no_scalar_flux (q, dom, id, idE, idNE, idN, v, zlev, type)
    real(8), dimension(1:EDGE)                           :: no_scalar_flux
    type(Float_Field), dimension(1:N_VARIABLE,1:zlevels) :: q
    type(domain)                                         :: dom
    integer                                              :: d, id, idE, idNE, idN, v, zlev
    logical, optional                                    :: type
    no_scalar_flux = 0d0
  end function no_scalar_flux

  function no_velo_source (dom, i, j, zlev, offs, dims)
    real(8), dimension(1:EDGE)     :: no_velo_source
    type(domain)                   :: dom
    integer                        :: i, j, zlev
    integer, dimension(N_BDRY+1)   :: offs
    integer, dimension(2,N_BDRY+1) :: dims
    no_velo_source = 0d0
  end function no_velo_source

  ! edge fluxes of diffusion of scalar v (step1's physics_scalar_flux):
  ! edges E, NE, N of id, or (sw) the W, SW, S edges with the neighbours
  ! passed as idE, idNE, idN
  function diff_scalar_flux (q, dom, id, idE, idNE, idN, v, zlev, type)
    real(8), dimension(1:EDGE)                           :: diff_scalar_flux
    type(Float_Field), dimension(1:N_VARIABLE,1:zlevels) :: q
    type(domain)                                         :: dom
    integer                                              :: d, id, idE, idNE, idN, v, zlev
    logical, optional                                    :: type
    real(8), dimension(1:EDGE) :: dl, pl, g
    real(8) :: f0, fE, fNE, fN
    logical :: sw
    diff_scalar_flux = 0d0
    if (Laplace_sclr == 0) return
    sw = .false.
    if (present(type)) sw = type
    d = dom%id + 1
    if (sw) then
       pl = (/ dom%pedlen%elts(EDGE*idE+RT+1), dom%pedlen%elts(EDGE*idNE+DG+1), dom%pedlen%elts(EDGE*idN+UP+1) /)
       dl = (/ -dom%len%elts(EDGE*idE+RT+1), -dom%len%elts(EDGE*idNE+DG+1), -dom%len%elts(EDGE*idN+UP+1) /)
    else
       pl = dom%pedlen%elts(EDGE*id+1:EDGE*id+EDGE)
       dl = dom%len%elts(EDGE*id+1:EDGE*id+EDGE)
    end if
    if (Laplace_sclr == 1) then
       f0 = q(v,zlev)%data(d)%elts(id+1);  fE = q(v,zlev)%data(d)%elts(idE+1)
       fNE = q(v,zlev)%data(d)%elts(idNE+1); fN = q(v,zlev)%data(d)%elts(idN+1)
    else
       f0 = Laplacian_scalar(v)%data(d)%elts(id+1);  fE = Laplacian_scalar(v)%data(d)%elts(idE+1)
       fNE = Laplacian_scalar(v)%data(d)%elts(idNE+1); fN = Laplacian_scalar(v)%data(d)%elts(idN+1)
    end if
    g(RT+1) = (fE - f0) / dl(RT+1)
    g(DG+1) = (f0 - fNE) / dl(DG+1)
    g(UP+1) = (fN - f0) / dl(UP+1)
    diff_scalar_flux = (-1)**Laplace_sclr * visc_sclr(v) * g * pl
  end function diff_scalar_flux

  ! diffusion of the velocity (du_source's physics_velo_source) at the
  ! edges of (i, j): divu is grid%divu (order 1) or its Laplacian (order 2),
  ! vort the vorticity (order 1) or rot rot of it (order 2)
  function diff_velo_source (dom, i, j, zlev, offs, dims)
    real(8), dimension(1:EDGE)     :: diff_velo_source
    type(domain)                   :: dom
    integer                        :: i, j, zlev
    integer, dimension(N_BDRY+1)   :: offs
    integer, dimension(2,N_BDRY+1) :: dims
    integer :: id, idE, idN, idNE, idS, idW
    real(8), dimension(1:EDGE) :: gd, cr
    diff_velo_source = 0d0
    if (Laplace_rotu == 0) return
    id   = idx (i,   j,   offs, dims)
    idE  = idx (i+1, j,   offs, dims)
    idN  = idx (i,   j+1, offs, dims)
    idNE = idx (i+1, j+1, offs, dims)
    idS  = idx (i,   j-1, offs, dims)
    idW  = idx (i-1, j,   offs, dims)
    gd = 0d0
    if (Laplace_divu /= 0) then
       gd(RT+1) = (divu(idE+1) - divu(id+1))   / dom%len%elts(EDGE*id+RT+1)
       gd(DG+1) = (divu(id+1)  - divu(idNE+1)) / dom%len%elts(EDGE*id+DG+1)
       gd(UP+1) = (divu(idN+1) - divu(id+1))   / dom%len%elts(EDGE*id+UP+1)
    end if
    cr(RT+1) = (vort(TRIAG*id +LORT+1) - vort(TRIAG*idS+UPLT+1)) / dom%pedlen%elts(EDGE*id+RT+1)
    cr(DG+1) = (vort(TRIAG*id +LORT+1) - vort(TRIAG*id +UPLT+1)) / dom%pedlen%elts(EDGE*id+DG+1)
    cr(UP+1) = (vort(TRIAG*idW+LORT+1) - vort(TRIAG*id +UPLT+1)) / dom%pedlen%elts(EDGE*id+UP+1)
    diff_velo_source = (-1)**(Laplace_rotu-1) * (visc_divu * gd - visc_rotu * cr)
  end function diff_velo_source

  ! forcing of the vertical diffusion (the test cases' bottom_buoy_flux,
  ! top_buoy_flux, wind_flux, tau_mag and bottom_friction): smooth functions
  ! of the node position, also evaluated by vd_forcing for the dump
  subroutine vd_forcing (p, bot, top, tau, wind)
    type(Coord) :: p
    real(8)     :: bot, top, tau, wind(1:EDGE), lon, lat
    integer     :: e
    call cart2sph (p, lon, lat)
    bot = 2d-7 * sin (lat) * cos (2d0*lon)
    top = 1d-6 * cos (lon) * cos (lat) + 3d-7
    tau = 0.1d0 * (1d0 + 0.5d0 * sin (lat) * cos (lon))
    do e = 1, EDGE
       wind(e) = 1d-4 * (1d0 + 0.3d0*e) * sin (2d0*lat) * cos (lon - 0.5d0*e)
    end do
  end subroutine vd_forcing

  real(8) function vd_bottom_buoy (dom, i, j, z_lev, offs, dims)
    type(Domain)                   :: dom
    integer                        :: i, j, z_lev
    integer, dimension(N_BDRY+1)   :: offs
    integer, dimension(2,N_BDRY+1) :: dims
    real(8) :: top, tau, wind(1:EDGE)
    call vd_forcing (dom%node%elts(idx (i, j, offs, dims)+1), vd_bottom_buoy, top, tau, wind)
  end function vd_bottom_buoy

  real(8) function vd_top_buoy (dom, i, j, z_lev, offs, dims)
    type(Domain)                   :: dom
    integer                        :: i, j, z_lev
    integer, dimension(N_BDRY+1)   :: offs
    integer, dimension(2,N_BDRY+1) :: dims
    real(8) :: bot, tau, wind(1:EDGE)
    call vd_forcing (dom%node%elts(idx (i, j, offs, dims)+1), bot, vd_top_buoy, tau, wind)
  end function vd_top_buoy

  function vd_wind (dom, i, j, z_lev, offs, dims)
    real(8), dimension(1:EDGE)     :: vd_wind
    type(Domain)                   :: dom
    integer                        :: i, j, z_lev
    integer, dimension(N_BDRY+1)   :: offs
    integer, dimension(2,N_BDRY+1) :: dims
    real(8) :: bot, top, tau
    call vd_forcing (dom%node%elts(idx (i, j, offs, dims)+1), bot, top, tau, vd_wind)
  end function vd_wind

  real(8) function vd_tau (p)
    type(Coord) :: p
    real(8) :: bot, top, wind(1:EDGE)
    call vd_forcing (p, bot, top, vd_tau, wind)
  end function vd_tau
end module harness_hooks

module harness_ic
  use domain_mod
  use geom_mod
  implicit none
  character(16) :: ic_case = 'tc5'
  real(8)       :: u0, h0, b0, rmount, lonc, latc
contains
  ! Williamson TC5-like zonal flow: u = u0 cos(lat)
  subroutine zonal_vel (lon, lat, u, v)
    real(8) :: lon, lat, u, v
    u = u0 * cos (lat)
    if (ic_case == 'ident') then
       ! non-symmetric field: every edge and node gets a distinct value
       ! single-valued at the poles (cos(lat) factor): the reference keeps one
       ! copy of each pole per lozenge and never reconciles them
       u = u0 * cos (lat) * (sin (7d0*lon + 3d0*lat) + 0.37d0*cos (11d0*lat - 2d0*lon))
       v = u0 * cos (lat) * (cos (5d0*lon - 9d0*lat) + 0.29d0*sin (13d0*lon))
       return
    end if
    if (ic_case == 'ocean') u = u0 * cos (lat) * (1d0 + 0.3d0*sin (3d0*lon)*cos (lat))
    v = 0d0
    if (ic_case == 'ocean') v = 0.2d0 * u0 * sin (2d0*lon) * cos (lat)**2
  end subroutine zonal_vel

  real(8) function bottom (p)
    type(Coord) :: p
    real(8)     :: lon, lat, r
    call cart2sph (p, lon, lat)
    if (ic_case == 'tc5') then
       r = min (rmount, sqrt ((lon - lonc)**2 + (lat - latc)**2))
       bottom = b0 * (1d0 - r/rmount)
    elseif (ic_case == 'ident') then
       bottom = cos (lat) * (100d0 * sin (3d0*lon + 5d0*lat) + 37d0 * cos (7d0*lon - 4d0*lat))
    else
       bottom = -4000d0 + 500d0 * sin (2d0*lon) * cos (lat)**2 + 300d0 * sin (3d0*lat)
    end if
  end function bottom

  real(8) function surface (p)
    type(Coord) :: p
    real(8)     :: lon, lat
    call cart2sph (p, lon, lat)
    if (ic_case == 'tc5') then
       surface = h0 - (radius*omega*u0 + 0.5d0*u0**2) * sin (lat)**2 / grav_accel
    elseif (ic_case == 'ident') then
       surface = h0 + cos (lat) * (50d0 * sin (9d0*lon - 2d0*lat) + 20d0 * cos (4d0*lon + 6d0*lat))
    else
       surface = 0.5d0 * exp (-((lon - 0.5d0)**2 + (lat - 0.3d0)**2) / 0.04d0) &
            - (radius*omega*u0 + 0.5d0*u0**2) * sin (lat)**2 / grav_accel
    end if
  end function surface

  ! layer density (k=1 bottom), stably stratified for the ocean case
  real(8) function layer_density (k)
    integer :: k
    if (zlevels == 1) then
       layer_density = ref_density
    else
       layer_density = 1028d0 - 4d0 * dble(k-1) / dble(zlevels-1)
    end if
  end function layer_density

  subroutine init_sol (dom, i, j, zlev, offs, dims)
    type(Domain)                   :: dom
    integer                        :: i, j, zlev
    integer, dimension(N_BDRY+1)   :: offs
    integer, dimension(2,N_BDRY+1) :: dims

    integer :: d, id, idN, idE, idNE, k
    real(8) :: b, s, dz, rho

    d    = dom%id + 1
    id   = idx (i,   j,   offs, dims)
    idN  = idx (i,   j+1, offs, dims)
    idE  = idx (i+1, j,   offs, dims)
    idNE = idx (i+1, j+1, offs, dims)

    b = bottom (dom%node%elts(id+1))
    s = surface (dom%node%elts(id+1))
    topography%data(d)%elts(id+1) = b
    do k = 1, zlevels
       dz  = (s - b) / dble(zlevels) * (1d0 + 0.1d0*dble(k-1)/dble(max(zlevels-1,1)))
       rho = layer_density (k)
       sol(S_MASS,k)%data(d)%elts(id+1) = ref_density * dz
       sol(S_TEMP,k)%data(d)%elts(id+1) = ref_density * dz * (1d0 - rho/ref_density)
       sol_mean(S_MASS,k)%data(d)%elts(id+1) = 0d0
       sol_mean(S_TEMP,k)%data(d)%elts(id+1) = 0d0
       sol(S_VELO,k)%data(d)%elts(EDGE*id+RT+1) = &
            proj_vel (zonal_vel, dom%node%elts(id+1),   dom%node%elts(idE+1)) * (1d0 - 0.05d0*dble(k-1))
       sol(S_VELO,k)%data(d)%elts(EDGE*id+DG+1) = &
            proj_vel (zonal_vel, dom%node%elts(idNE+1), dom%node%elts(id+1))  * (1d0 - 0.05d0*dble(k-1))
       sol(S_VELO,k)%data(d)%elts(EDGE*id+UP+1) = &
            proj_vel (zonal_vel, dom%node%elts(id+1),   dom%node%elts(idN+1)) * (1d0 - 0.05d0*dble(k-1))
       sol_mean(S_VELO,k)%data(d)%elts(EDGE*id+RT+1:EDGE*id+UP+1) = 0d0
    end do
  end subroutine init_sol
end module harness_ic

program dump_trisk
  use time_integr_mod
  use remap_mod
  use vert_diffusion_mod
  use harness_hooks
  use harness_ic
  implicit none

  character(16)  :: case
  character(255) :: out
  integer        :: zlev, nsteps, d, k, l, n, u, v, ntimed, lsclr, ldivu, lrotu, every
  integer(8)     :: t_start, t_end, t_rate
  real(8)        :: step, vsclr(2), vdivu, vrotu
  logical        :: nanfill, uniform_first
  integer        :: vdiff, id
  logical        :: vdinline
  real(8)        :: lon, lat
  character(16)  :: vremap, coords
  namelist /run/ case, zlev, nsteps, step, out, nanfill, ntimed, lsclr, ldivu, lrotu, vsclr, vdivu, vrotu, every, &
                 vremap, coords, uniform_first, vdiff, vdinline

  case = 'tc5'; zlev = 1; nsteps = 0; step = 0d0; out = 'dump.bin'; nanfill = .false.; ntimed = 0; every = 1
  lsclr = 0; ldivu = 0; lrotu = 0; vsclr = 0d0; vdivu = 0d0; vrotu = 0d0; vremap = ''; coords = 'uniform'
  uniform_first = .false.; vdiff = 0; vdinline = .false.
  read (5, nml=run)
  ic_case = case

  ! --- parameters (shared.f90:259-442 defaults, overridden as a test case would) ---
  call init_shared_mod
  zlevels       = zlev
  max_level     = MIN_LEVEL
  level_start   = MIN_LEVEL
  level_end     = MIN_LEVEL
  compressible  = .false.
  mode_split    = .false.
  optimize_grid = NO_OPTIM
  vert_diffuse  = vdiff > 0
  tke_closure   = vdiff == 2
  Laplace_sclr  = lsclr; Laplace_divu = ldivu; Laplace_rotu = lrotu
  visc_sclr(S_MASS) = vsclr(1); visc_sclr(S_TEMP) = vsclr(2)
  visc_divu     = vdivu
  visc_rotu     = vrotu
  radius        = 6.371d6
  omega         = 7.292d-5
  grav_accel    = 9.80616d0
  ref_density   = 1028d0
  tol           = 0d0
  u0 = 20d0; h0 = 5960d0; b0 = 2000d0; rmount = MATH_PI/9d0
  lonc = -MATH_PI/2d0; latc = MATH_PI/6d0
  if (ic_case == 'ocean') u0 = 0.5d0

  allocate (a_vert(0:zlevels), b_vert(0:zlevels), a_vert_mass(1:zlevels), b_vert_mass(1:zlevels))
  do k = 0, zlevels
     a_vert(k) = 0d0; b_vert(k) = dble(k)/dble(zlevels)
  end do
  do k = 1, zlevels
     a_vert_mass(k) = 0d0; b_vert_mass(k) = b_vert(k) - b_vert(k-1)
  end do

  bottom_friction  => vd_friction
  bottom_buoy_flux => vd_bottom_buoy
  top_buoy_flux    => vd_top_buoy
  wind_flux        => vd_wind
  tau_mag          => vd_tau
  physics_scalar_flux => no_scalar_flux
  physics_velo_source => no_velo_source
  if (lsclr /= 0 .or. lrotu /= 0) then
     physics_scalar_flux => diff_scalar_flux
     physics_velo_source => diff_velo_source
  end if

  ! --- init_basic (main.f90:23-34), minus io ---
  call init_comm_mod
  call init_init_mod
  call init_refine_patch_mod
  call init_time_integr_mod
  call init_wavelet_mod
  call init_mask_mod
  call init_adapt_mod

  ! --- init_structures (main.f90:482-524) ---
  level_start = min_level
  level_end   = level_start
  call distribute_grid (cp_idx, 'harness')
  call init_grid
  call init_comm_mpi
  call init_geometry
  call comm_nodes3_mpi (get_coord, set_coord, NONE)
  call precompute_geometry
  call init_wavelets
  call init_masks
  call add_second_level
  call apply_onescale2 (set_level, level_start, z_null, -BDRY_THICKNESS, +BDRY_THICKNESS)
  call apply_interscale (mask_adj_child, level_start-1, z_null, 0, 1)
  call init_RK_mem

  ! uniform non-adaptive grid: every node and edge active
  do d = 1, size(grid)
     grid(d)%mask_n%elts = TOLRNZ
     grid(d)%mask_e%elts = TOLRNZ
     do k = 1, zmax
        penal_node(k)%data(d)%elts = 0d0
        penal_edge(k)%data(d)%elts = 0d0
     end do
  end do

  do l = level_start, level_end
     call apply_onescale (init_sol, l, z_null, -1, 1)
  end do

  ! optional: poison every cell that is neither inside a patch of level_start
  ! nor a pole, so that ghost values the ghost fill does not provide show up
  ! as NaN wherever a stencil reads them
  if (nanfill) call poison_ghosts

  open (newunit=u, file=trim(out), access='stream', form='unformatted', status='replace')
  call wr_i ('min_level', (/ MIN_LEVEL, DOMAIN_LEVEL, PATCH_SIZE, zlevels, size(grid) /))
  call wr_f ('params', (/ radius, omega, grav_accel, ref_density, step /))
  call dump_grid
  call update_bdry (sol(:,1:zlevels), NONE, 1)
  call dump_state ('s0_')
  call trend_ml (sol(:,1:zlevels), trend(:,1:zlevels))
  call dump_trend ('t0_')
  call dump_work
  if (vdiff > 0 .and. vdinline) then
     call init_vdiff
     dt = step
  end if
  do n = 1, nsteps
     call RK4 (sol(:,1:zlevels), wav_coeff(:,1:zlevels), trend_ml, step)
     call update_bdry (sol(:,1:zlevels), NONE, 1)
     if (vdiff > 0 .and. vdinline) then  ! time_step's order: the dynamics, then the vertical diffusion
        istep = n
        call vertical_diffusion
        call update_bdry (sol(:,1:zlevels), NONE, 1)
     end if
     if (mod(n, every) == 0) &
          call dump_state ('s'//achar(48+(n/every)/10)//achar(48+mod(n/every,10))//'_')
  end do
  if (len_trim(vremap) > 0 .and. uniform_first) then
     do k = 0, zlevels
        b_vert(k) = 1d0 - dble(k)/dble(zlevels)
     end do
     a_vert = 1d0 - b_vert
     remap_type = '0'
     sigma_z = .false.
     call remap_vertical_coordinates
     call update_bdry (sol(:,1:zlevels), NONE, 1)
     call dump_state ('p0_')
  end if
  if (len_trim(vremap) > 0) then
     do k = 0, zlevels
        if (trim(coords) == 'chebyshev') then
           b_vert(k) = (1d0 + cos (dble(k)/dble(zlevels) * MATH_PI)) / 2d0
        else
           b_vert(k) = 1d0 - dble(k)/dble(zlevels)
        end if
     end do
     a_vert = 1d0 - b_vert
     remap_type = trim(vremap)
     sigma_z = .false.
     call remap_vertical_coordinates
     call update_bdry (sol(:,1:zlevels), NONE, 1)
     call dump_state ('r0_')
  end if
  if (vdiff > 0 .and. vdinline) call dump_coefficients ('v9_')
  if (vdiff > 0 .and. .not. vdinline) then
     call init_vdiff
     dt = step
     do n = 1, 2
        istep = n
        call vertical_diffusion
        call update_bdry (sol(:,1:zlevels), NONE, 1)
        call dump_state ('v'//achar(48+n)//'_')
        call dump_coefficients ('v'//achar(48+n)//'_')
     end do
  end if
  close (u)
  if (ntimed > 0) then
     call system_clock (t_start, t_rate)
     do n = 1, ntimed
        call RK4 (sol(:,1:zlevels), wav_coeff(:,1:zlevels), trend_ml, step)
     end do
     call system_clock (t_end)
     write (6, '(a,f12.3,a)') 'RK4 ', 1d3 * dble(t_end - t_start) / dble(t_rate) / dble(ntimed), ' ms/step'
  end if
contains
  subroutine wr_i (tag, a)
    character(*)          :: tag
    integer, dimension(:) :: a
    character(16)         :: t
    t = tag
    write (u) t, 1_4, int(size(a),8), a
  end subroutine wr_i

  subroutine wr_f (tag, a)
    character(*)          :: tag
    real(8), dimension(:) :: a
    character(16)         :: t
    t = tag
    write (u) t, 2_4, int(size(a),8), a
  end subroutine wr_f

  subroutine dump_grid
    integer :: dd, j, p
    integer, allocatable :: starts(:)
    call wr_i ('nelts', (/ (grid(dd)%node%length, dd = 1, size(grid)) /))
    do dd = 1, size(grid)
       allocate (starts(grid(dd)%lev(level_start)%length))
       do j = 1, grid(dd)%lev(level_start)%length
          p = grid(dd)%lev(level_start)%elts(j)
          starts(j) = grid(dd)%patch%elts(p+1)%elts_start
       end do
       call wr_i ('patch_starts', starts)
       deallocate (starts)
       call dump_patch_tables (dd)
       call wr_f ('node_x', grid(dd)%node%elts(1:grid(dd)%node%length)%x)
       call wr_f ('node_y', grid(dd)%node%elts(1:grid(dd)%node%length)%y)
       call wr_f ('node_z', grid(dd)%node%elts(1:grid(dd)%node%length)%z)
       call wr_f ('len', grid(dd)%len%elts(1:EDGE*grid(dd)%node%length))
       call wr_f ('pedlen', grid(dd)%pedlen%elts(1:EDGE*grid(dd)%node%length))
       do j = 1, 6
          call wr_f ('area_part', grid(dd)%areas%elts(1:grid(dd)%node%length)%part(j))
       end do
       call wr_f ('hex_inv', grid(dd)%areas%elts(1:grid(dd)%node%length)%hex_inv)
       call wr_f ('triarea', grid(dd)%triarea%elts(1:TRIAG*grid(dd)%node%length))
       call wr_f ('coriolis', grid(dd)%coriolis%elts(1:TRIAG*grid(dd)%node%length))
       call wr_f ('topo', topography%data(dd)%elts(1:grid(dd)%node%length))
       call wr_i ('pole_master', (/ merge(1,0,grid(dd)%pole_master(1)), merge(1,0,grid(dd)%pole_master(2)) /))
    end do
  end subroutine dump_grid

  subroutine poison_ghosts
    integer :: dd, j, p, e, kk, s0
    logical, allocatable :: own(:)
    real(8) :: nan
    nan = 1d20
    do dd = 1, size(grid)
       allocate (own(grid(dd)%node%length)); own = .false.
       do j = 1, grid(dd)%lev(level_start)%length
          p = grid(dd)%lev(level_start)%elts(j)
          s0 = grid(dd)%patch%elts(p+1)%elts_start
          own(s0+1:s0+PATCH_SIZE**2) = .true.
       end do
       do e = 1, grid(dd)%node%length
          if (abs(abs(grid(dd)%node%elts(e)%z) - radius) < 1d-6*radius) own(e) = .true.
       end do
       do kk = 1, zlevels
          do e = 1, grid(dd)%node%length
             if (own(e)) cycle
             sol(S_MASS,kk)%data(dd)%elts(e) = nan
             sol(S_TEMP,kk)%data(dd)%elts(e) = nan
             sol(S_VELO,kk)%data(dd)%elts(EDGE*(e-1)+1:EDGE*e) = nan
          end do
       end do
       deallocate (own)
    end do
  end subroutine poison_ghosts

  subroutine dump_patch_tables (dd)
    ! For each child c of the root patch (main.f90 add_second_level): the
    ! child index and idx() over (i,j) in [-BDRY_THICKNESS, PATCH_SIZE+BDRY_THICKNESS)
    integer :: dd, c, p, i, j, m
    integer, dimension(N_BDRY+1)   :: offs
    integer, dimension(2,N_BDRY+1) :: dims
    integer, allocatable :: tab(:)
    m = PATCH_SIZE + 2*BDRY_THICKNESS
    allocate (tab(m*m + 2))
    do c = 1, N_CHDRN
       p = grid(dd)%patch%elts(2)%children(c)
       call get_offs_Domain (grid(dd), p, offs, dims)
       tab(1) = c; tab(2) = p
       do j = -BDRY_THICKNESS, PATCH_SIZE+BDRY_THICKNESS-1
          do i = -BDRY_THICKNESS, PATCH_SIZE+BDRY_THICKNESS-1
             tab(3 + (j+BDRY_THICKNESS)*m + (i+BDRY_THICKNESS)) = idx (i, j, offs, dims)
          end do
       end do
       call wr_i ('patch_idx', tab)
    end do
    deallocate (tab)
  end subroutine dump_patch_tables

  subroutine dump_state (pre)
    character(*) :: pre
    integer :: dd, kk
    do kk = 1, zlevels
       do dd = 1, size(grid)
          call wr_f (pre//'mass', sol(S_MASS,kk)%data(dd)%elts)
          call wr_f (pre//'temp', sol(S_TEMP,kk)%data(dd)%elts)
          call wr_f (pre//'velo', sol(S_VELO,kk)%data(dd)%elts)
       end do
    end do
  end subroutine dump_state

  subroutine init_vdiff
    ! initial TKE (Kt and Kv are recomputed from it at istep 1) and the forcing
    integer :: dd, ii, kk
    do dd = 1, size(grid)
       do ii = 1, grid(dd)%node%length
          call cart2sph (grid(dd)%node%elts(ii), lon, lat)
          do kk = 1, zlevels
             tke(kk)%data(dd)%elts(ii) = 1d-5 * (1d0 + 0.5d0 * sin (3d0*lon + dble(kk)) * cos (lat))
          end do
          do kk = 0, zlevels
             Kt(kk)%data(dd)%elts(ii) = Kt_min
             Kv(kk)%data(dd)%elts(ii) = Kv_min
          end do
       end do
       call dump_forcing (dd)
    end do
    call dump_coefficients ('v0_')
  end subroutine init_vdiff

  subroutine dump_coefficients (pre)
    character(*) :: pre
    integer :: dd, kk
    do kk = 0, zlevels
       do dd = 1, size(grid)
          call wr_f (pre//'kt', Kt(kk)%data(dd)%elts)
          call wr_f (pre//'kv', Kv(kk)%data(dd)%elts)
       end do
    end do
    do kk = 1, zlevels
       do dd = 1, size(grid)
          call wr_f (pre//'tke', tke(kk)%data(dd)%elts)
       end do
    end do
  end subroutine dump_coefficients

  subroutine dump_forcing (dd)
    integer :: dd, ii
    real(8), dimension(:), allocatable :: bot, top, tau, wind
    allocate (bot(grid(dd)%node%length), top(grid(dd)%node%length), tau(grid(dd)%node%length), &
              wind(EDGE*grid(dd)%node%length))
    do ii = 1, grid(dd)%node%length
       call vd_forcing (grid(dd)%node%elts(ii), bot(ii), top(ii), tau(ii), wind(EDGE*(ii-1)+1:EDGE*ii))
    end do
    call wr_f ('vf_bot', bot)
    call wr_f ('vf_top', top)
    call wr_f ('vf_tau', tau)
    call wr_f ('vf_wind', wind)
  end subroutine dump_forcing

  subroutine dump_trend (pre)
    character(*) :: pre
    integer :: dd, kk
    do kk = 1, zlevels
       do dd = 1, size(grid)
          call wr_f (pre//'dmass', trend(S_MASS,kk)%data(dd)%elts)
          call wr_f (pre//'dtemp', trend(S_TEMP,kk)%data(dd)%elts)
          call wr_f (pre//'dvelo', trend(S_VELO,kk)%data(dd)%elts)
       end do
    end do
  end subroutine dump_trend

  subroutine dump_work
    ! work arrays of the last layer evaluated by trend_ml
    integer :: dd
    do dd = 1, size(grid)
       call wr_f ('w_qe', grid(dd)%qe%elts)
       call wr_f ('w_ke', grid(dd)%ke%elts)
       call wr_f ('w_bern', grid(dd)%bernoulli%elts)
       call wr_f ('w_exner', exner_fun(zlevels)%data(dd)%elts)
       call wr_f ('w_flux_m', horiz_flux(S_MASS)%data(dd)%elts)
       call wr_f ('w_flux_t', horiz_flux(S_TEMP)%data(dd)%elts)
       call wr_f ('w_press', grid(dd)%press%elts)
       call wr_f ('w_geopot', grid(dd)%geopot%elts)
    end do
  end subroutine dump_work
end program dump_trisk
