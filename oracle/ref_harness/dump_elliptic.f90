! Golden-vector harness for the free-surface elliptic path of the WAVETRISK
! reference (test infrastructure only).
!
! Same initialisation as dump_trisk.f90 (init_basic main.f90:23-34,
! init_structures main.f90:482-524) with mode_split on (zmax = zlevels+1,
! init.f90:145-150), a uniform single-level grid, every node active, and an
! ocean state.  It then evaluates the reference's own elliptic routines and
! dumps, per domain array:
!   e_q, e_etan         operand q and old free surface eta^n (Laplacian_scalar(S_TEMP))
!   e_Lq                elliptic_lo (q, l)          (barotropic_2d.f90:187-253)
!   e_hflux             its edge fluxes (horiz_flux(S_MASS) after the ghost fill)
!   e_diag              elliptic_lo_diag (q, l)     (barotropic_2d.f90:255-324)
!   e_wsum              visit count of the (0,1) sweep used by l2/dp/lc (lin_solve.f90:433-518)
!   e_red               l2 (q), dp (q, eta^n), l2 (eta^n)
!   e_jac               one Jacobi_iteration of u = eta^n with f, weight wjac (lin_solve.f90:271-340)
!   e_divF, e_rhs       flux_divergence of sol and the RHS of rhs_elliptic (barotropic_2d.f90:151-185, 326-366)
!   e_sol, e_iters      bicgstab (u = eta^n, f = e_rhs) to tol_solve (lin_solve.f90:361-431)
! and with split = 1, each from the initial state again:
!   bc_mass, bc_temp, bc_velo   barotropic_correction of the state
!   ge_flux                     grad_eta: g grad(eta) on the edges (horiz_flux(S_TEMP))
!   sp_mass, sp_temp, sp_velo, sp_eta   the state after one RK4_split step
!                                       (nstage = 3, 2, 1: RK3_split, RK2_split,
!                                       Euler_split, time_integr.f90:384-451)
!
! Namelist &run (stdin): zlev, step (dt), theta_1, theta_2, wjac, tol_solve, maxit, split, nstage, out.
module ell_hooks
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
end module ell_hooks

module ell_ic
  use domain_mod
  use geom_mod
  implicit none
  real(8) :: u0
  type(Float_Field), target :: qf, etan, ff
  integer, allocatable :: visits(:,:)
contains
  real(8) function bottom (p)
    type(Coord) :: p
    real(8)     :: lon, lat
    call cart2sph (p, lon, lat)
    bottom = -4000d0 + cos (lat)**2 * (500d0 * sin (2d0*lon) + 300d0 * sin (3d0*lat))
  end function bottom

  real(8) function eta_old (p)
    type(Coord) :: p
    real(8)     :: lon, lat
    call cart2sph (p, lon, lat)
    eta_old = cos (lat)**2 * (0.5d0 * exp (-((lon - 0.5d0)**2 + (lat - 0.3d0)**2) / 0.04d0) + 0.05d0 * sin (3d0*lon))
  end function eta_old

  real(8) function q_field (p)
    ! distinct values everywhere; the cos(lat)**2 factor makes every pole
    ! copy (one per lozenge, each from its own coordinates) the same value
    type(Coord) :: p
    real(8)     :: lon, lat
    call cart2sph (p, lon, lat)
    q_field = 0.3d0 + cos (lat)**2 * (0.7d0 * sin (7d0*lon + 3d0*lat) + 0.21d0 * cos (11d0*lat - 2d0*lon))
  end function q_field

  subroutine zonal_vel (lon, lat, u, v)
    real(8) :: lon, lat, u, v
    u = u0 * cos (lat) * (1d0 + 0.3d0*sin (3d0*lon)*cos (lat))
    v = 0.2d0 * u0 * sin (2d0*lon) * cos (lat)**2
  end subroutine zonal_vel

  subroutine init_sol (dom, i, j, zlev, offs, dims)
    type(Domain)                   :: dom
    integer                        :: i, j, zlev
    integer, dimension(N_BDRY+1)   :: offs
    integer, dimension(2,N_BDRY+1) :: dims

    integer :: d, id, idN, idE, idNE, k
    real(8) :: b, dz, rho

    d    = dom%id + 1
    id   = idx (i,   j,   offs, dims)
    idN  = idx (i,   j+1, offs, dims)
    idE  = idx (i+1, j,   offs, dims)
    idNE = idx (i+1, j+1, offs, dims)

    b = bottom (dom%node%elts(id+1))
    topography%data(d)%elts(id+1) = b
    do k = 1, zlevels
       dz  = -b / dble(zlevels) * (1d0 + 0.1d0*dble(k-1)/dble(max(zlevels-1,1)))
       rho = 1028d0 - 4d0 * dble(k-1) / dble(max(zlevels-1,1))
       sol(S_MASS,k)%data(d)%elts(id+1) = ref_density * dz * (1d0 + 0.01d0 * q_field (dom%node%elts(id+1)))
       sol(S_TEMP,k)%data(d)%elts(id+1) = sol(S_MASS,k)%data(d)%elts(id+1) * (1d0 - rho/ref_density)
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
    sol(S_MASS,zlevels+1)%data(d)%elts(id+1) = eta_old (dom%node%elts(id+1))
    sol(S_TEMP,zlevels+1)%data(d)%elts(id+1) = 0d0
    sol_mean(S_MASS,zlevels+1)%data(d)%elts(id+1) = 0d0
    sol_mean(S_TEMP,zlevels+1)%data(d)%elts(id+1) = 0d0
    etan%data(d)%elts(id+1) = eta_old (dom%node%elts(id+1))
    qf%data(d)%elts(id+1)   = q_field (dom%node%elts(id+1))
  end subroutine init_sol

  subroutine count_visit (dom, i, j, zlev, offs, dims)
    type(Domain)                   :: dom
    integer                        :: i, j, zlev
    integer, dimension(N_BDRY+1)   :: offs
    integer, dimension(2,N_BDRY+1) :: dims
    integer :: id
    id = idx (i, j, offs, dims) + 1
    visits(id, dom%id+1) = visits(id, dom%id+1) + 1
  end subroutine count_visit
end module ell_ic

program dump_elliptic
  use time_integr_mod
  use ell_hooks
  use ell_ic
  implicit none

  character(255) :: out
  integer        :: zlev, maxit, d, k, l, u, iters, j, nmax, split, nstage
  real(8)        :: step, theta_1, theta_2, wjac, tol_solve, err, nrm
  type(Float_Field) :: Lq, dgl, uu
  namelist /run/ zlev, step, theta_1, theta_2, wjac, tol_solve, maxit, split, nstage, out

  zlev = 2; step = 600d0; theta_1 = 1d0; theta_2 = 1d0; wjac = 0.7d0; tol_solve = 1d-9; maxit = 200; split = 0
  nstage = 4
  out = 'ell.bin'
  read (5, nml=run)

  call init_shared_mod
  zlevels       = zlev
  max_level     = MIN_LEVEL
  level_start   = MIN_LEVEL
  level_end     = MIN_LEVEL
  compressible  = .false.
  mode_split    = .true.
  optimize_grid = NO_OPTIM
  Laplace_sclr  = 0; Laplace_divu = 0; Laplace_rotu = 0
  radius        = 6.371d6
  omega         = 7.292d-5
  grav_accel    = 9.80616d0
  ref_density   = 1028d0
  tol           = 0d0
  theta1        = theta_1
  theta2        = theta_2
  dt            = step
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
  call init_RK_mem

  do d = 1, size(grid)
     grid(d)%mask_n%elts = TOLRNZ
     grid(d)%mask_e%elts = TOLRNZ
     do k = 1, zmax
        penal_node(k)%data(d)%elts = 0d0
        penal_edge(k)%data(d)%elts = 0d0
     end do
  end do

  qf = sol(S_MASS,1); etan = sol(S_MASS,1); ff = sol(S_MASS,1)
  l = level_start
  call apply_onescale (init_sol, l, z_null, -1, 1)
  call update_bdry (sol(:,1:zmax), NONE, 1)
  call update_bdry (qf, NONE, 1)
  call update_bdry (etan, NONE, 1)

  open (newunit=u, file=trim(out), access='stream', form='unformatted', status='replace')
  call wr_i ('min_level', (/ MIN_LEVEL, DOMAIN_LEVEL, PATCH_SIZE, zlevels, size(grid) /))
  call wr_f ('params', (/ radius, omega, grav_accel, ref_density, dt, theta1, theta2, wjac, tol_solve /))
  call dump_grid
  call dump_field ('e_q', qf)
  call dump_field ('e_etan', etan)
  do k = 1, zlevels
     do d = 1, size(grid)
        call wr_f ('s0_mass', sol(S_MASS,k)%data(d)%elts)
        call wr_f ('s0_temp', sol(S_TEMP,k)%data(d)%elts)
        call wr_f ('s0_velo', sol(S_VELO,k)%data(d)%elts)
     end do
  end do

  ! old free surface used by the operator and its diagonal
  call equals_float_field (Laplacian_scalar(S_TEMP), etan, AT_NODE)

  Lq = elliptic_lo (qf, l)
  call dump_field ('e_Lq', Lq)
  do d = 1, size(grid)
     call wr_f ('e_hflux', horiz_flux(S_MASS)%data(d)%elts(1:EDGE*grid(d)%node%length))
  end do
  dgl = elliptic_lo_diag (qf, l)
  call dump_field ('e_diag', dgl)

  ! visit counts of the (0,1) sweep
  nmax = 0
  do d = 1, size(grid)
     nmax = max (nmax, grid(d)%node%length)
  end do
  allocate (visits(nmax, size(grid))); visits = 0
  do d = 1, size(grid)
     do j = 1, grid(d)%lev(l)%length
        call apply_onescale_to_patch (count_visit, grid(d), grid(d)%lev(l)%elts(j), z_null, 0, 1)
     end do
     call wr_i ('e_wsum', visits(1:grid(d)%node%length, d))
  end do
  call wr_f ('e_red', (/ l2 (qf, l), dp (qf, etan, l), l2 (etan, l) /))

  ! RHS of the elliptic equation: flux divergence of the layer-integrated
  ! velocity (stored in trend(S_MASS,zlevels+1)), previous divergence = 0
  do d = 1, size(grid)
     trend(S_TEMP,zlevels+1)%data(d)%elts = 0d0
  end do
  call rhs_elliptic
  call dump_field ('e_divF', trend(S_MASS,zlevels+1))
  call dump_field ('e_rhs', sol(S_TEMP,zlevels+1))
  call equals_float_field (Laplacian_scalar(S_TEMP), sol(S_MASS,zlevels+1), AT_NODE)

  ! one weighted Jacobi sweep from u = eta^n on L u = rhs
  uu = etan
  call Jacobi_iteration (uu, sol(S_TEMP,zlevels+1), wjac, elliptic_lo, elliptic_lo_diag, l)
  call update_bdry (uu, NONE, 1)
  call dump_field ('e_jac', uu)

  ! BiCGSTAB from u = eta^n
  uu = etan
  call update_bdry (sol(S_TEMP,zlevels+1), NONE, 1)
  nrm = l2 (sol(S_TEMP,zlevels+1), l)
  call bicgstab (uu, sol(S_TEMP,zlevels+1), nrm, elliptic_lo, l, tol_solve, maxit, err, iters)
  call update_bdry (uu, NONE, 1)
  call dump_field ('e_sol', uu)
  call wr_i ('e_iters', (/ iters /))
  call wr_f ('e_err', (/ err, nrm /))

  if (split > 0) then
     ! barotropic_correction of the initial state (barotropic_2d.f90:52-116)
     call reset_state
     call barotropic_correction (sol(:,1:zmax))
     call dump_layers ('bc')
     ! external pressure gradient of the initial free surface (grad_eta, step1 itype 3)
     call reset_state
     call grad_eta
     do d = 1, size(grid)
        call wr_f ('ge_flux', horiz_flux(S_TEMP)%data(d)%elts(1:EDGE*grid(d)%node%length))
     end do
     ! one RK4_split step (time_integr.f90:360-382) with the SRJ solver
     ! (main.f90:50), which on one level is BiCGSTAB to coarse_tol
     call reset_state
     elliptic_solver => SRJ
     coarse_tol = tol_solve
     fine_tol   = tol_solve
     coarse_iter = maxit
     log_iter = .false.
     select case (nstage)
     case (3)
        call RK3_split (dt)
     case (2)
        call RK2_split (dt)
     case (1)
        call Euler_split (dt)
     case default
        call RK4_split (dt)
     end select
     call update_bdry (sol(:,1:zmax), NONE, 1)
     call dump_layers ('sp')
     call dump_field ('sp_eta', sol(S_MASS,zlevels+1))
  end if
  close (u)
contains
  subroutine reset_state
    call apply_onescale (init_sol, level_start, z_null, -1, 1)
    call update_bdry (sol(:,1:zmax), NONE, 1)
    do d = 1, size(grid)
       trend(S_MASS,zlevels+1)%data(d)%elts = 0d0
       trend(S_TEMP,zlevels+1)%data(d)%elts = 0d0
    end do
  end subroutine reset_state

  subroutine dump_layers (pre)
    character(*) :: pre
    integer :: kk, dd
    do kk = 1, zlevels
       do dd = 1, size(grid)
          call wr_f (pre // '_mass', sol(S_MASS,kk)%data(dd)%elts)
          call wr_f (pre // '_temp', sol(S_TEMP,kk)%data(dd)%elts)
          call wr_f (pre // '_velo', sol(S_VELO,kk)%data(dd)%elts)
       end do
    end do
  end subroutine dump_layers

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

  subroutine dump_field (tag, f)
    character(*)      :: tag
    type(Float_Field) :: f
    integer :: dd
    do dd = 1, size(grid)
       call wr_f (tag, f%data(dd)%elts(1:grid(dd)%node%length))
    end do
  end subroutine dump_field

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
       call wr_f ('topo', topography%data(dd)%elts(1:grid(dd)%node%length))
    end do
  end subroutine dump_grid

  subroutine dump_patch_tables (dd)
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
end program dump_elliptic
