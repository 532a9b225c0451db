! Golden-vector harness for the multi-level (wavelet) part of the WAVETRISK
! reference (test infrastructure only).
!
! Initialises like dump_elliptic.f90 (mode split, every node active), then
! refines the grid uniformly from MIN_LEVEL to max_level with the
! reference's own adapt (adapt.f90:14-58) at zero thresholds, as the
! driver's initial refinement does (main.f90:103-116).  Dumps, per level l
! and per domain, the node elements of the level's patches with their
! coordinates, hex_inv and overlap areas (overl_areas a(1..4),
! wavelet.f90:1408-1475 via init_wavelets / refine_patch), then
!   x_*     a scalar field given at every level
!   fst_*   forward_scalar_transform (wavelet.f90:104-150) of x from
!           level_end down to level_start: the restricted scalars
!   fsw_*   its wavelet coefficients
!   ist_*   inverse_scalar_transform (wavelet.f90:298-346) of (fsw, fst)
!   fmg_*   FMG (lin_solve.f90:33-140) of elliptic_lo u = f from u = 0,
!           with eta^n and f given at every level
!   srj_*   SRJ (lin_solve.f90:142-231) of the same problem from u = 0
! The solvers' logs (residuals, iterations per level) go to stdout, SRJ's
! after a line 'SRJ'.
!
! Namelist &run (stdin): maxlev, zlev, step, tol_solve, out.
module lev_ic
  use domain_mod
  use geom_mod
  implicit none
  type(Float_Field), target :: xf, etan, ff
  integer :: hu, cur_dom, cur_l
contains
  real(8) function bottom (p)
    type(Coord) :: p
    real(8)     :: lon, lat
    call cart2sph (p, lon, lat)
    bottom = -4000d0 + cos (lat)**2 * (500d0 * sin (2d0*lon) + 300d0 * sin (3d0*lat))
  end function bottom

  real(8) function x_field (p)
    type(Coord) :: p
    real(8)     :: lon, lat
    call cart2sph (p, lon, lat)
    x_field = 0.3d0 + cos (lat)**2 * (0.7d0 * sin (7d0*lon + 3d0*lat) + 0.21d0 * cos (11d0*lat - 2d0*lon))
  end function x_field

  real(8) function eta_old (p)
    type(Coord) :: p
    real(8)     :: lon, lat
    call cart2sph (p, lon, lat)
    eta_old = cos (lat)**2 * (0.5d0 * exp (-((lon - 0.5d0)**2 + (lat - 0.3d0)**2) / 0.04d0) + 0.05d0 * sin (3d0*lon))
  end function eta_old

  real(8) function f_field (p)
    type(Coord) :: p
    real(8)     :: lon, lat
    call cart2sph (p, lon, lat)
    f_field = -eta_old (p) + 0.01d0 * cos (lat)**2 * sin (5d0*lon - 2d0*lat)
  end function f_field

  subroutine init_fields (dom, i, j, zlev, offs, dims)
    type(Domain)                   :: dom
    integer                        :: i, j, zlev
    integer, dimension(N_BDRY+1)   :: offs
    integer, dimension(2,N_BDRY+1) :: dims
    integer :: d, id, k
    d  = dom%id + 1
    id = idx (i, j, offs, dims) + 1
    topography%data(d)%elts(id) = bottom (dom%node%elts(id))
    do k = 1, zmax
       sol(S_MASS,k)%data(d)%elts(id) = 0d0
       sol(S_TEMP,k)%data(d)%elts(id) = 0d0
       sol_mean(S_MASS,k)%data(d)%elts(id) = 0d0
       sol_mean(S_TEMP,k)%data(d)%elts(id) = 0d0
    end do
    do k = 1, zlevels
       sol(S_MASS,k)%data(d)%elts(id) = -ref_density * bottom (dom%node%elts(id)) / dble(zlevels)
    end do
  end subroutine init_fields

  subroutine init_test_fields (dom, i, j, zlev, offs, dims)
    type(Domain)                   :: dom
    integer                        :: i, j, zlev
    integer, dimension(N_BDRY+1)   :: offs
    integer, dimension(2,N_BDRY+1) :: dims
    integer :: d, id
    d  = dom%id + 1
    id = idx (i, j, offs, dims) + 1
    xf%data(d)%elts(id)   = x_field (dom%node%elts(id))
    etan%data(d)%elts(id) = eta_old (dom%node%elts(id))
    ff%data(d)%elts(id)   = f_field (dom%node%elts(id))
  end subroutine init_test_fields

  subroutine pole_stencil (dom, i_par, j_par, i_chd, j_chd, zlev, offs_par, dims_par, offs_chd, dims_chd)
    ! records, for a parent that is a pole, the child and the twelve lifting
    ! neighbours in the order of Restrict_scalar (wavelet.f90:542-553)
    type(Domain)                   :: dom
    integer                        :: i_par, j_par, i_chd, j_chd, zlev
    integer, dimension(N_BDRY+1)   :: offs_par, offs_chd
    integer, dimension(2,N_BDRY+1) :: dims_par, dims_chd
    integer :: id_par, k
    integer, dimension(13) :: ids
    real(8), dimension(13) :: buf
    id_par = idx (i_par, j_par, offs_par, dims_par)
    if (abs (dom%node%elts(id_par+1)%z) < radius * (1d0 - 1d-12)) return
    ids = (/ idx (i_chd, j_chd, offs_chd, dims_chd), &
         idx (i_chd+1, j_chd,   offs_chd, dims_chd), idx (i_chd+1, j_chd+1, offs_chd, dims_chd), &
         idx (i_chd+2, j_chd+1, offs_chd, dims_chd), idx (i_chd+1, j_chd+2, offs_chd, dims_chd), &
         idx (i_chd,   j_chd+1, offs_chd, dims_chd), idx (i_chd-1, j_chd,   offs_chd, dims_chd), &
         idx (i_chd-1, j_chd+1, offs_chd, dims_chd), idx (i_chd-2, j_chd-1, offs_chd, dims_chd), &
         idx (i_chd-1, j_chd-1, offs_chd, dims_chd), idx (i_chd,   j_chd-1, offs_chd, dims_chd), &
         idx (i_chd-1, j_chd-2, offs_chd, dims_chd), idx (i_chd+1, j_chd-1, offs_chd, dims_chd) /) + 1
    call mwr_i ('pst_ids', (/ cur_dom, cur_l, id_par + 1, i_par, j_par, i_chd, j_chd, &
         dom%mask_n%elts(ids(1)), ids /))
    do k = 1, 4
       buf = dom%overl_areas%elts(ids)%a(k); call mwr_f ('pst_ovl', buf)
    end do
    buf = dom%node%elts(ids)%x; call mwr_f ('pst_x', buf)
    buf = dom%node%elts(ids)%y; call mwr_f ('pst_y', buf)
    buf = dom%node%elts(ids)%z; call mwr_f ('pst_z', buf)
    buf(1) = dom%areas%elts(id_par+1)%hex_inv; call mwr_f ('pst_hinv', buf(1:1))
  end subroutine pole_stencil

  subroutine mwr_i (tag, a)
    character(*)          :: tag
    integer, dimension(:) :: a
    character(16)         :: t
    t = tag
    write (hu) t, 1_4, int(size(a),8), a
  end subroutine mwr_i

  subroutine mwr_f (tag, a)
    character(*)          :: tag
    real(8), dimension(:) :: a
    character(16)         :: t
    t = tag
    write (hu) t, 2_4, int(size(a),8), a
  end subroutine mwr_f

  subroutine no_update
    ! the test case's "update" hook (init.f90:120), called by adapt: nothing to re-evaluate
  end subroutine no_update
end module lev_ic

! threshold hook passed to adapt (a test case module provides it upstream)
subroutine set_thresholds_zero
  use shared_mod
  implicit none
  threshold = 0d0
end subroutine set_thresholds_zero

program dump_levels
  use time_integr_mod
  use lev_ic
  implicit none
  external :: set_thresholds_zero

  character(255) :: out
  integer        :: maxlev, zlev, d, k, l, u
  real(8)        :: step, tol_solve
  type(Float_Field) :: wav, scl, uu
  namelist /run/ maxlev, zlev, step, tol_solve, out

  maxlev = MIN_LEVEL + 2; zlev = 2; step = 600d0; tol_solve = 1d-9; out = 'lev.bin'
  read (5, nml=run)

  call init_shared_mod
  zlevels       = zlev
  max_level     = maxlev
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
  dt            = step
  fine_tol      = tol_solve
  coarse_tol    = tol_solve
  coarse_iter   = 400
  max_vcycle    = 50

  allocate (a_vert(0:zlevels), b_vert(0:zlevels), a_vert_mass(1:zlevels), b_vert_mass(1:zlevels))
  do k = 0, zlevels
     a_vert(k) = 0d0; b_vert(k) = dble(k)/dble(zlevels)
  end do
  do k = 1, zlevels
     a_vert_mass(k) = 0d0; b_vert_mass(k) = b_vert(k) - b_vert(k-1)
  end do

  update => no_update

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
  allocate (n_patch_old(size(grid)), n_node_old(size(grid))); n_patch_old = 2   ! init_structures, main.f90:517
  call init_RK_mem
  allocate (threshold(1:N_VARIABLE,zmin:zmax), threshold_def(1:N_VARIABLE,zmin:zmax))
  threshold = 0d0; threshold_def = 0d0

  do d = 1, size(grid)
     do k = 1, zmax
        penal_node(k)%data(d)%elts = 0d0
        penal_edge(k)%data(d)%elts = 0d0
     end do
  end do
  call set_all
  ! uniform refinement to max_level (main.f90:103-116 with zero thresholds)
  do while (level_end < max_level)
     call adapt (set_thresholds_zero)
     call set_all
  end do
  ! test fields on the final grid (copies of a field sized for every level)
  xf = sol(S_MASS,1); etan = sol(S_MASS,1); ff = sol(S_MASS,1)
  do l = level_start, level_end
     call apply_onescale (init_test_fields, l, z_null, -1, 1)
  end do

  open (newunit=u, file=trim(out), access='stream', form='unformatted', status='replace')
  call wr_i ('min_level', (/ MIN_LEVEL, DOMAIN_LEVEL, PATCH_SIZE, zlevels, size(grid), level_start, level_end /))
  call wr_f ('params', (/ radius, omega, grav_accel, ref_density, dt, theta1, theta2, tol_solve /))
  do l = level_start, level_end
     call dump_level (l)
  end do
  ! the fine elements that Restrict_scalar reads for each pole parent
  do l = level_start, level_end - 1
     do d = 1, size(grid)
        cur_dom = d; cur_l = l; hu = u
        call apply_interscale_d (pole_stencil, grid(d), l, z_null, 0, 1)
     end do
  end do
  call dump_field ('x', xf)
  call dump_field ('etan', etan)
  call dump_field ('f', ff)
  call dump_field ('topo', topography)

  ! forward and inverse scalar transforms over the whole hierarchy
  wav = xf; scl = xf
  call forward_scalar_transform (scl, wav, level_start, level_end)
  call update_bdry (scl, NONE, 1)
  call dump_field ('fst', scl)
  call dump_field ('fsw', wav)
  call inverse_scalar_transform (wav, scl, level_start, level_end)
  call update_bdry (scl, NONE, 1)
  call dump_field ('ist', scl)

  ! FMG on the hierarchy: elliptic_lo u = f, eta^n in Laplacian_scalar(S_TEMP)
  call equals_float_field (Laplacian_scalar(S_TEMP), etan, AT_NODE)
  uu = etan
  do d = 1, size(grid)
     uu%data(d)%elts = 0d0
  end do
  log_iter = .true.
  call FMG (uu, ff, elliptic_lo, elliptic_lo_diag)
  call update_bdry (uu, NONE, 1)
  call dump_field ('fmg', uu)

  ! SRJ (lin_solve.f90:142-231, the driver's default elliptic solver) from u = 0
  write (6,'(a)') 'SRJ'
  do d = 1, size(grid)
     uu%data(d)%elts = 0d0
  end do
  call SRJ (uu, ff, elliptic_lo, elliptic_lo_diag)
  call update_bdry (uu, NONE, 1)
  call dump_field ('srj', uu)
  close (u)
contains
  subroutine set_all
    integer :: ll
    do ll = level_start, level_end
       call apply_onescale (init_fields, ll, z_null, -1, 1)
    end do
    ! wavelets of the new state, as count_active does (main.f90:137-141)
    call forward_wavelet_transform (sol, wav_coeff)
  end subroutine set_all

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
    ! element ids of the level's patch interiors, their coordinates, hex_inv, overlap areas
    integer :: l, dd, j, p, i, jj, m, n
    integer, dimension(N_BDRY+1)   :: offs
    integer, dimension(2,N_BDRY+1) :: dims
    integer, allocatable :: ids(:)
    real(8), allocatable :: buf(:)
    do dd = 1, size(grid)
       n = grid(dd)%lev(l)%length * PATCH_SIZE**2
       allocate (ids(n), buf(n))
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
       buf = grid(dd)%node%elts(ids)%x; call wr_f ('lev_x', buf)
       buf = grid(dd)%node%elts(ids)%y; call wr_f ('lev_y', buf)
       buf = grid(dd)%node%elts(ids)%z; call wr_f ('lev_z', buf)
       buf = grid(dd)%areas%elts(ids)%hex_inv; call wr_f ('lev_hexinv', buf)
       do i = 1, 4
          buf = grid(dd)%overl_areas%elts(ids)%a(i); call wr_f ('lev_ovl', buf)
       end do
       deallocate (ids, buf)
       ! nodes (-2..1, -2..1) around a pentagon at the south-west corner of a
       ! level-l patch: element ids, overlap areas, coordinates
       do j = 1, grid(dd)%lev(l)%length
          p = grid(dd)%lev(l)%elts(j)
          if (.not. is_penta (grid(dd), p, IJMINUS - 1)) cycle
          call get_offs_Domain (grid(dd), p, offs, dims)
          allocate (ids(16), buf(16))
          m = 0
          do jj = -2, 1
             do i = -2, 1
                m = m + 1
                ids(m) = idx (i, jj, offs, dims) + 1
             end do
          end do
          call wr_i ('pen_ids', (/ l, ids /))
          do i = 1, 4
             buf = grid(dd)%overl_areas%elts(ids)%a(i); call wr_f ('pen_ovl', buf)
          end do
          buf = grid(dd)%node%elts(ids)%x; call wr_f ('pen_x', buf)
          buf = grid(dd)%node%elts(ids)%y; call wr_f ('pen_y', buf)
          buf = grid(dd)%node%elts(ids)%z; call wr_f ('pen_z', buf)
          deallocate (ids, buf)
       end do
    end do
  end subroutine dump_level
end program dump_levels
