! Golden-vector harness for the multi-level explicit step of the WAVETRISK
! reference (test infrastructure only).
!
! Initialises like dump_trisk.f90 (incompressible ocean, hooks zero, no
! diffusion), refines uniformly from MIN_LEVEL to maxlev with the reference's
! own adapt at zero thresholds (as the driver's initial refinement does,
! main.f90:100-128), applies the initial state on every level and its
! forward wavelet transform (count_active, main.f90:137-143), then dumps
!   * per level and domain: the element ids of the level's patch interiors,
!     their coordinates, masks (mask_n, mask_e), overlap areas a(1..4) and
!     split(1..2), velocity interpolation weights I_u_wgt (9) and flux
!     restriction weights R_F_wgt (3);
!   * whole per-domain arrays: s0 (state after the transform), w0 (its
!     wavelet coefficients), i0 (inverse_wavelet_transform of (w0, s0)),
!     t0 (trend_ml of s0, every level), a0 (WT_after_step of s0 + dt/4 t0),
!     and the state after each of nsteps RK4 steps (time_integr.f90:78-103).
!
! With split = 1 the model runs in mode-split form (mode_split, zmax =
! zlevels+1: the free surface eta is the extra layer's mass) and the steps
! are RK4_split (time_integr.f90:360-382) with the FMG free-surface solve to
! tol_solve (lin_solve.f90:33-140); eta is dumped per level as 'eta'.
!
! Namelist &run (stdin): maxlev, zlev, nsteps, step, split, th1, th2,
! tol_solve, out.
module ml_hooks
  use domain_mod
  implicit none
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

  subroutine no_update
  end subroutine no_update
end module ml_hooks

module ml_ic
  use domain_mod
  use geom_mod
  implicit none
  real(8) :: u0
contains
  subroutine flow (lon, lat, u, v)
    real(8) :: lon, lat, u, v
    u = u0 * cos (lat) * (1d0 + 0.3d0*sin (3d0*lon)*cos (lat))
    v = 0.2d0 * u0 * sin (2d0*lon) * cos (lat)**2
  end subroutine flow

  real(8) function bottom (p)
    type(Coord) :: p
    real(8)     :: lon, lat
    call cart2sph (p, lon, lat)
    bottom = -4000d0 + 500d0 * sin (2d0*lon) * cos (lat)**2 + 300d0 * sin (3d0*lat)
  end function bottom

  real(8) function surface (p)
    type(Coord) :: p
    real(8)     :: lon, lat
    call cart2sph (p, lon, lat)
    surface = 0.5d0 * exp (-((lon - 0.5d0)**2 + (lat - 0.3d0)**2) / 0.04d0) &
         - (radius*omega*u0 + 0.5d0*u0**2) * sin (lat)**2 / grav_accel
  end function surface

  real(8) function eta_init (p)
    type(Coord) :: p
    real(8)     :: lon, lat
    call cart2sph (p, lon, lat)
    eta_init = 0.5d0 * exp (-((lon - 0.5d0)**2 + (lat - 0.3d0)**2) / 0.04d0) + 0.05d0 * cos (lat)**2 * sin (3d0*lon)
  end function eta_init

  real(8) function layer_density (k)
    ! a single layer is lighter than the reference density, so that its
    ! temperature variable (and its flux restriction) is not zero
    integer :: k
    if (zlevels == 1) then
       layer_density = 1026d0
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
            proj_vel (flow, dom%node%elts(id+1),   dom%node%elts(idE+1)) * (1d0 - 0.05d0*dble(k-1))
       sol(S_VELO,k)%data(d)%elts(EDGE*id+DG+1) = &
            proj_vel (flow, dom%node%elts(idNE+1), dom%node%elts(id+1))  * (1d0 - 0.05d0*dble(k-1))
       sol(S_VELO,k)%data(d)%elts(EDGE*id+UP+1) = &
            proj_vel (flow, dom%node%elts(id+1),   dom%node%elts(idN+1)) * (1d0 - 0.05d0*dble(k-1))
       sol_mean(S_VELO,k)%data(d)%elts(EDGE*id+RT+1:EDGE*id+UP+1) = 0d0
    end do
    if (mode_split) then
       sol(S_MASS,zlevels+1)%data(d)%elts(id+1) = eta_init (dom%node%elts(id+1))
       sol(S_TEMP,zlevels+1)%data(d)%elts(id+1) = 0d0
       sol(S_VELO,zlevels+1)%data(d)%elts(EDGE*id+RT+1:EDGE*id+UP+1) = 0d0
       sol_mean(S_MASS,zlevels+1)%data(d)%elts(id+1) = 0d0
       sol_mean(S_TEMP,zlevels+1)%data(d)%elts(id+1) = 0d0
       sol_mean(S_VELO,zlevels+1)%data(d)%elts(EDGE*id+RT+1:EDGE*id+UP+1) = 0d0
    end if
  end subroutine init_sol
end module ml_ic

subroutine set_thresholds_zero
  use shared_mod
  implicit none
  threshold = 0d0
end subroutine set_thresholds_zero

program dump_mlstep
  use time_integr_mod
  use ml_hooks
  use ml_ic
  implicit none
  external :: set_thresholds_zero

  character(255) :: out
  integer        :: maxlev, zlev, nsteps, d, k, l, n, u, split
  real(8)        :: step, th1, th2, tol_solve
  type(Float_Field), dimension(:,:), allocatable :: cp
  namelist /run/ maxlev, zlev, nsteps, step, split, th1, th2, tol_solve, out

  maxlev = MIN_LEVEL + 1; zlev = 2; nsteps = 2; step = 600d0; out = 'ml.bin'
  split = 0; th1 = 1d0; th2 = 1d0; tol_solve = 1d-9
  read (5, nml=run)

  call init_shared_mod
  zlevels       = zlev
  max_level     = maxlev
  level_start   = MIN_LEVEL
  level_end     = MIN_LEVEL
  compressible  = .false.
  mode_split    = split /= 0
  theta1        = th1
  theta2        = th2
  dt            = step
  fine_tol      = tol_solve
  coarse_tol    = tol_solve
  coarse_iter   = 400
  max_vcycle    = 50
  optimize_grid = NO_OPTIM
  Laplace_sclr  = 0; Laplace_divu = 0; Laplace_rotu = 0
  radius        = 6.371d6
  omega         = 7.292d-5
  grav_accel    = 9.80616d0
  ref_density   = 1028d0
  tol           = 0d0
  u0            = 0.5d0

  allocate (a_vert(0:zlevels), b_vert(0:zlevels), a_vert_mass(1:zlevels), b_vert_mass(1:zlevels))
  do k = 0, zlevels
     a_vert(k) = 0d0; b_vert(k) = dble(k)/dble(zlevels)
  end do
  do k = 1, zlevels
     a_vert_mass(k) = 0d0; b_vert_mass(k) = b_vert(k) - b_vert(k-1)
  end do

  physics_scalar_flux => no_scalar_flux
  physics_velo_source => no_velo_source
  update => no_update
  elliptic_solver => FMG

  call init_comm_mod
  call init_init_mod
  call init_refine_patch_mod
  call init_time_integr_mod
  call init_wavelet_mod
  call init_mask_mod
  call init_adapt_mod

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
  allocate (n_patch_old(size(grid)), n_node_old(size(grid))); n_patch_old = 2
  call init_RK_mem
  allocate (threshold(1:N_VARIABLE,zmin:zmax), threshold_def(1:N_VARIABLE,zmin:zmax))
  threshold = 0d0; threshold_def = 0d0
  do d = 1, size(grid)
     do k = 1, zmax
        penal_node(k)%data(d)%elts = 0d0
        penal_edge(k)%data(d)%elts = 0d0
     end do
  end do

  ! initial refinement (main.f90:100-128) with zero thresholds
  call set_all
  do while (level_end < max_level)
     call adapt (set_thresholds_zero)
     call set_all
  end do
  call adapt (set_thresholds_zero)
  call set_all

  open (newunit=u, file=trim(out), access='stream', form='unformatted', status='replace')
  call wr_i ('min_level', (/ MIN_LEVEL, DOMAIN_LEVEL, PATCH_SIZE, zlevels, size(grid), level_start, level_end /))
  call wr_f ('params', (/ radius, omega, grav_accel, ref_density, step /))
  call wr_i ('nelts', (/ (grid(d)%node%length, d = 1, size(grid)) /))
  do l = level_start, level_end
     call dump_level (l)
  end do
  call dump_whole
  call dump_state ('s0_', sol)
  call dump_state ('w0_', wav_coeff)
  if (mode_split) then
     call update_bdry (sol(S_MASS,zlevels+1), NONE, 1)
     call dump_eta ('e00_')
     do n = 1, nsteps
        call RK4_split (dt)
        call update_bdry (sol(:,1:zmax), NONE, 1)
        call dump_state ('s'//achar(48+n/10)//achar(48+mod(n,10))//'_', sol)
        call dump_eta ('e'//achar(48+n/10)//achar(48+mod(n,10))//'_')
     end do
     close (u)
     stop
  end if

  ! inverse transform of (w0, s0) on a copy
  allocate (cp(1:N_VARIABLE,1:zlevels))
  cp = sol(:,1:zlevels)
  call inverse_wavelet_transform (wav_coeff(:,1:zlevels), cp)
  call update_bdry (cp, NONE, 1)
  call dump_state ('i0_', cp)

  ! trend on every level
  call update_bdry (sol(:,1:zlevels), NONE, 1)
  call trend_ml (sol(:,1:zlevels), trend(:,1:zlevels))
  call dump_state ('t0_', trend)

  ! one stage: q1 = s0 + dt/4 t0, then WT_after_step
  cp = sol(:,1:zlevels)
  call stage_axpy (cp, trend(:,1:zlevels), step/4d0)
  call WT_after_step (cp, wav_coeff(:,1:zlevels))
  call update_bdry (cp, NONE, 1)
  call dump_state ('a0_', cp)

  do n = 1, nsteps
     call RK4 (sol(:,1:zlevels), wav_coeff(:,1:zlevels), trend_ml, step)
     call update_bdry (sol(:,1:zlevels), NONE, 1)
     call dump_state ('s'//achar(48+n/10)//achar(48+mod(n,10))//'_', sol)
  end do
  close (u)
contains
  subroutine set_all
    integer :: ll
    do ll = level_start, level_end
       call apply_onescale (init_sol, ll, z_null, -1, 1)
    end do
    call forward_wavelet_transform (sol, wav_coeff)
  end subroutine set_all

  subroutine stage_axpy (q, dq, h)
    ! RK_sub_step's update q := q + h dq over whole arrays from patch 2 on (time_integr.f90:206-225)
    type(Float_Field), dimension(:,:) :: q, dq
    real(8)                           :: h
    integer :: dd, kk, v, ibeg, iend
    do v = 1, N_VARIABLE
       do dd = 1, size(grid)
          ibeg = (1+2*(POSIT(v)-1)) * grid(dd)%patch%elts(2+1)%elts_start + 1
          iend = q(v,1)%data(dd)%length
          do kk = 1, size(q,2)
             q(v,kk)%data(dd)%elts(ibeg:iend) = q(v,kk)%data(dd)%elts(ibeg:iend) + h * dq(v,kk)%data(dd)%elts(ibeg:iend)
          end do
       end do
    end do
    q%bdry_uptodate = .false.
  end subroutine stage_axpy

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

  subroutine dump_whole
    ! per-domain arrays of inter-level data over every element
    integer :: dd, j, nn
    do dd = 1, size(grid)
       nn = grid(dd)%node%length
       call wr_i ('mask_n', grid(dd)%mask_n%elts(1:nn))
       call wr_i ('mask_e', grid(dd)%mask_e%elts(1:EDGE*nn))
       call wr_f ('node_x', grid(dd)%node%elts(1:nn)%x)
       call wr_f ('node_y', grid(dd)%node%elts(1:nn)%y)
       call wr_f ('node_z', grid(dd)%node%elts(1:nn)%z)
       do j = 1, 4
          call wr_f ('ovl_a', grid(dd)%overl_areas%elts(1:nn)%a(j))
       end do
       do j = 1, 2
          call wr_f ('ovl_split', grid(dd)%overl_areas%elts(1:nn)%split(j))
       end do
       do j = 1, 9
          call wr_f ('iu_wgt', grid(dd)%I_u_wgt%elts(1:nn)%enc(j))
       end do
       do j = 1, 3
          call wr_f ('rf_wgt', grid(dd)%R_F_wgt%elts(1:nn)%enc(j))
       end do
       call wr_f ('topo', topography%data(dd)%elts(1:nn))
    end do
  end subroutine dump_whole

  subroutine dump_state (pre, q)
    character(*)                      :: pre
    type(Float_Field), dimension(:,:) :: q
    integer :: dd, kk
    do kk = 1, zlevels
       do dd = 1, size(grid)
          call wr_f (pre//'mass', q(S_MASS,kk)%data(dd)%elts(1:grid(dd)%node%length))
          call wr_f (pre//'temp', q(S_TEMP,kk)%data(dd)%elts(1:grid(dd)%node%length))
          call wr_f (pre//'velo', q(S_VELO,kk)%data(dd)%elts(1:EDGE*grid(dd)%node%length))
       end do
    end do
  end subroutine dump_state

  subroutine dump_eta (pre)
    character(*) :: pre
    integer :: dd
    do dd = 1, size(grid)
       call wr_f (pre//'eta', sol(S_MASS,zlevels+1)%data(dd)%elts(1:grid(dd)%node%length))
    end do
  end subroutine dump_eta

  function pole_ids (dd, l)
    ! elements of the level-l patches of domain dd (with one boundary row) that are poles
    integer :: dd, l
    integer, dimension(2) :: pole_ids
    integer :: j, p, i, jj, id
    integer, dimension(N_BDRY+1)   :: offs
    integer, dimension(2,N_BDRY+1) :: dims
    pole_ids = 0
    do j = 1, grid(dd)%lev(l)%length
       p = grid(dd)%lev(l)%elts(j)
       call get_offs_Domain (grid(dd), p, offs, dims)
       do jj = -1, PATCH_SIZE
          do i = -1, PATCH_SIZE
             id = idx (i, jj, offs, dims) + 1
             if (id < 1 .or. id > grid(dd)%node%length) cycle
             if (abs (grid(dd)%node%elts(id)%z) > radius * (1d0 - 1d-12)) then
                if (grid(dd)%node%elts(id)%z < 0d0) then
                   pole_ids(1) = id
                else
                   pole_ids(2) = id
                end if
             end if
          end do
       end do
    end do
  end function pole_ids

  subroutine dump_level (l)
    ! element ids of the level's patch interiors, and per patch idx() over
    ! (i,j) in [-BDRY_THICKNESS, PATCH_SIZE+BDRY_THICKNESS)
    integer :: l, dd, j, p, i, jj, m, nn
    integer, dimension(N_BDRY+1)   :: offs
    integer, dimension(2,N_BDRY+1) :: dims
    integer, allocatable :: ids(:), tab(:)
    m = PATCH_SIZE + 2*BDRY_THICKNESS
    allocate (tab(m*m + 2))
    do dd = 1, size(grid)
       do j = 1, grid(dd)%lev(l)%length
          p = grid(dd)%lev(l)%elts(j)
          call get_offs_Domain (grid(dd), p, offs, dims)
          tab(1) = l; tab(2) = dd
          do jj = -BDRY_THICKNESS, PATCH_SIZE+BDRY_THICKNESS-1
             do i = -BDRY_THICKNESS, PATCH_SIZE+BDRY_THICKNESS-1
                tab(3 + (jj+BDRY_THICKNESS)*m + (i+BDRY_THICKNESS)) = idx (i, jj, offs, dims)
             end do
          end do
          call wr_i ('ptab', tab)
       end do
    end do
    deallocate (tab)
    do dd = 1, size(grid)
       nn = grid(dd)%lev(l)%length * PATCH_SIZE**2
       allocate (ids(nn))
       m = 0
       do j = 1, grid(dd)%lev(l)%length
          p = grid(dd)%lev(l)%elts(j)
          call get_offs_Domain (grid(dd), p, offs, dims)
          do jj = 0, PATCH_SIZE-1
             do i = 0, PATCH_SIZE-1
                m = m + 1
                ids(m) = idx (i, jj, offs, dims) + 1
             end do
          end do
       end do
       call wr_i ('lev_ids', (/ l, ids /))
       call wr_i ('lev_poles', pole_ids (dd, l))
       deallocate (ids)
    end do
  end subroutine dump_level
end program dump_mlstep
