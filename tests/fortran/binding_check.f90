! Calls libwavetrisk_gpu.so through the ISO_C_BINDING module
! (wavetrisk_hydrostatic_amd/fortran/wavetrisk_gpu.f90), as a WAVETRISK
! caller would, for tests/test_fortran_binding.py.
!
!   binding_check grid <in> <out>   grid only (no GPU): info and hex_inv
!   binding_check rk4  <in> <out>   one level: wtg_rk4 steps on the device
!   binding_check rk4d <in> <out>   the same with hyperdiffusion (wtg_set_diffusion:
!                                   orders 2, 2, 2, viscosities 4e16, 2e16, 6e16, 4e16)
!   binding_check ml   <in> <out>   a hierarchy: wtg_ml_rk4_split steps with
!                                   the FMG free surface (wtg_mg_*)
!   binding_check vdr  <in> <out>   one level: wtg_vertical_diffusion (TKE
!                                   closure, Q_sr 20, bottom friction 4e-4,
!                                   uniform forcing and coefficients, istep 1,
!                                   time step dt) then wtg_remap (PPR) onto
!                                   b_vert(k) = 1 - (k/nlayers)^2
!
! <in>: int32 level_start, level_end, min_level, domain_level, nlayers,
! nsteps; float64 dt, theta1, theta2; then per level coarse to fine: topo,
! mass, temp, velo, eta (stream access, the layout's cell order).
! <out>: per level mass, temp, velo, eta; then the int32 iteration count.
program binding_check
  use iso_c_binding
  use wavetrisk_gpu
  implicit none
  character(len=256) :: mode, fin, fout
  integer(c_int) :: hdr(6), nl, l, nlay, nsteps, iters
  real(c_double) :: par(3)
  type(wtg_grid_desc) :: desc
  type(wtg_grid_info) :: info
  type(wtg_params) :: prm
  type(wtg_fs_params) :: fsp
  type(wtg_mg_params) :: mgp
  type(wtg_diffusion) :: dif
  type(wtg_vdiff) :: vdp
  real(c_double), allocatable :: fb(:), ftp(:), fta(:), fw(:), kt(:), kv(:), tk(:), av(:), bv(:)
  integer :: k
  type(c_ptr), target :: grids(16), ctxs(16)
  type(c_ptr) :: ml, mg
  type arr
     real(c_double), allocatable :: topo(:), mass(:), temp(:), velo(:), eta(:)
  end type
  type(arr) :: s(16)
  integer(c_int64_t) :: nc
  real(c_double), allocatable, target :: hexinv(:)

  call get_command_argument(1, mode)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)
  open (10, file=trim(fin), access='stream', form='unformatted', status='old')
  read (10) hdr, par
  nl = hdr(2) - hdr(1) + 1
  nlay = hdr(5)
  nsteps = hdr(6)
  do l = 1, nl
     desc = wtg_grid_desc(hdr(1) + l - 1, hdr(3), hdr(4), 6.371d6, 7.292d-5)
     call wtg_check(wtg_grid_create(desc, grids(l)))
  end do
  open (11, file=trim(fout), access='stream', form='unformatted', status='replace')

  if (trim(mode) == 'grid') then
     do l = 1, nl
        call wtg_check(wtg_grid_info_get(grids(l), info))
        write (11) info%n, info%S, info%halo, info%ncells, info%south_pole, info%north_pole
        allocate (hexinv(info%ncells))
        call wtg_check(wtg_grid_get(grids(l), 'hex_inv' // c_null_char, c_loc(hexinv), &
                                    int(8 * info%ncells, c_size_t)))
        write (11) hexinv
        deallocate (hexinv)
     end do
     close (11)
     stop
  end if

  prm = wtg_params(nlay, 9.80616d0, 1028d0, 0d0)
  fsp = wtg_fs_params(par(1), par(2), par(3))
  do l = 1, nl
     call wtg_check(wtg_grid_info_get(grids(l), info))
     nc = info%ncells
     allocate (s(l)%topo(nc), s(l)%mass(nlay * nc), s(l)%temp(nlay * nc), s(l)%velo(3 * nlay * nc), s(l)%eta(nc))
     read (10) s(l)%topo, s(l)%mass, s(l)%temp, s(l)%velo, s(l)%eta
     call wtg_check(wtg_create(grids(l), prm, 2, ctxs(l)))
     call wtg_check(wtg_set_topography(ctxs(l), s(l)%topo), ctxs(l))
     call wtg_check(wtg_upload_state(ctxs(l), 0, s(l)%mass, s(l)%temp, s(l)%velo), ctxs(l))
  end do
  iters = 0
  if (trim(mode) == 'rk4' .or. trim(mode) == 'rk4d') then
     if (trim(mode) == 'rk4d') then
        dif = wtg_diffusion(2, 2, 2, 4d16, 2d16, 6d16, 4d16)
        call wtg_check(wtg_set_diffusion(ctxs(1), dif), ctxs(1))
     end if
     call wtg_check(wtg_rk4(ctxs(1), 0, par(1), nsteps), ctxs(1))
  else if (trim(mode) == 'vdr') then
     call wtg_vdiff_defaults(vdp)
     vdp%tke_closure = 1
     vdp%Q_sr = 20d0
     vdp%bottom_friction = 4d-4
     call wtg_check(wtg_set_vdiff(ctxs(1), vdp), ctxs(1))
     call wtg_check(wtg_grid_info_get(grids(1), info))
     nc = info%ncells
     allocate (fb(nc), ftp(nc), fta(nc), fw(3 * nc), kt((nlay + 1) * nc), kv((nlay + 1) * nc), tk(nlay * nc))
     fb = 1d-7; ftp = 2d-7; fta = 0.1d0; fw = 1d-4; kt = 1.2d-5; kv = 1.2d-4; tk = 2d-5
     call wtg_check(wtg_vdiff_set_coefficients(ctxs(1), kt, kv, tk), ctxs(1))
     call wtg_check(wtg_vdiff_set_forcing(ctxs(1), fb, ftp, fta, fw), ctxs(1))
     call wtg_check(wtg_vertical_diffusion(ctxs(1), 0, par(1), 1), ctxs(1))
     allocate (av(0:nlay), bv(0:nlay))
     do k = 0, nlay
        bv(k) = 1d0 - (dble(k) / dble(nlay)) * (dble(k) / dble(nlay))
     end do
     av = 1d0 - bv
     call wtg_check(wtg_remap(ctxs(1), 0, WTG_REMAP_PPR, av, bv), ctxs(1))
  else if (trim(mode) == 'ml') then
     do l = 1, nl
        call wtg_check(wtg_fs_set_params(ctxs(l), fsp), ctxs(l))
     end do
     call wtg_check(wtg_mg_create(grids, ctxs, nl, mg))
     call wtg_check(wtg_ml_create(grids, ctxs, nl, ml))
     do l = 1, nl
        call wtg_check(wtg_fs_upload(ctxs(l), WTG_FS_ETA, s(l)%eta), ctxs(l))
     end do
     call wtg_mg_default_params(mgp)
     mgp%coarse_tol = 1d-9
     mgp%fine_tol = 1d-9
     call wtg_check(wtg_ml_rk4_split(ml, mg, 0, nsteps, mgp, iters), ctxs(nl))
     call wtg_ml_destroy(ml)
     call wtg_mg_destroy(mg)
  else
     write (0, *) 'unknown mode ', trim(mode)
     error stop 2
  end if
  do l = 1, nl
     call wtg_check(wtg_download_state(ctxs(l), 0, s(l)%mass, s(l)%temp, s(l)%velo), ctxs(l))
     if (trim(mode) == 'ml') call wtg_check(wtg_fs_download(ctxs(l), WTG_FS_ETA, s(l)%eta), ctxs(l))
     write (11) s(l)%mass, s(l)%temp, s(l)%velo, s(l)%eta
  end do
  write (11) iters
  close (11)
  do l = 1, nl
     call wtg_destroy(ctxs(l))
     call wtg_grid_destroy(grids(l))
  end do
end program binding_check
