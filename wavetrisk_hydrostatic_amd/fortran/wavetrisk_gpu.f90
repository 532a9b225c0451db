! ISO_C_BINDING interface to libwavetrisk_gpu.so (include/wavetrisk_gpu.h).
!
! A WAVETRISK maintainer adds this module to src/ and links
!   -L<repo>/wavetrisk_hydrostatic_amd -lwavetrisk_gpu
! Every function returns 0 or a negative WTG_E_* code; wtg_check stops the
! program on a non-zero code (the reference's own convention is
! arch_mod's abort, which a caller inside WAVETRISK may call instead).
! The derived types mirror the C structs field by field (bind(C)).
module wavetrisk_gpu
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: WTG_FS_ETA = 0, WTG_FS_RHS = 1, WTG_FS_ETA_OLD = 2, WTG_FS_DIVF = 3, &
                               WTG_FS_DIVF_OLD = 4
  integer(c_int), parameter :: WTG_SOLVE_BICGSTAB = 0, WTG_SOLVE_JACOBI = 1, WTG_SOLVE_FMG = 2, &
                               WTG_SOLVE_SRJ = 3, WTG_SOLVE_VCYCLE = 4, WTG_SOLVE_LEARNED = 5
  integer(c_int), parameter :: WTG_PRECISION_FP64 = 0, WTG_PRECISION_FP32_BULK = 1
  ! remap_type "0", "1", "2PPM", "PPR", "2S", "2W" and "4" of remap_vertical_coordinates
  integer(c_int), parameter :: WTG_REMAP_PIECEWISE_CONSTANT = 0, WTG_REMAP_PIECEWISE_LINEAR = 1, &
                               WTG_REMAP_PARABOLIC_PPM = 2, WTG_REMAP_PPR = 3, WTG_REMAP_PARABOLIC_SPLINE = 4, &
                               WTG_REMAP_PARABOLIC_WENO = 5, WTG_REMAP_QUARTIC_WENO = 6

  type, bind(C) :: wtg_grid_desc
     integer(c_int) :: level, min_level, domain_level
     real(c_double) :: radius, omega
  end type
  type, bind(C) :: wtg_grid_info
     integer(c_int)     :: n, S, halo
     integer(c_int64_t) :: ncells, south_pole, north_pole
  end type
  type, bind(C) :: wtg_params
     integer(c_int) :: nlayers
     real(c_double) :: grav, ref_density, p_top
  end type
  type, bind(C) :: wtg_fs_params
     real(c_double) :: dt, theta1, theta2
  end type
  ! solver parameters (lin_solve.f90:11-26); fill with wtg_mg_default_params
  ! first so that learned_cycles, which has no reference counterpart, gets
  ! its default
  ! Laplace_sclr, Laplace_divu, Laplace_rotu (0, 1 or 2) and visc_sclr(S_MASS),
  ! visc_sclr(S_TEMP), visc_divu, visc_rotu (shared.f90:226, 238)
  type, bind(C) :: wtg_diffusion
     integer(c_int) :: laplace_sclr, laplace_divu, laplace_rotu
     real(c_double) :: visc_mass, visc_temp, visc_divu, visc_rotu
  end type
  ! vert_diffusion_mod's module variables, tke_closure, c_p and the test case's bottom_friction
  type, bind(C) :: wtg_vdiff
     integer(c_int) :: tke_closure, patankar, enhance_diff, mode_split
     real(c_double) :: C_e, C_eps, C_l, C_k, C_srf, e_0, e_min, e_min_srf, eps_s, kappa_VK, Kt_enh, Kt_max, Kt_min, &
                       Kv_max, Kv_min, mixed_layer, l_0, l_min, Neps_sq, Nsq_min, Q_sr, c_p, bottom_friction
  end type
  type, bind(C) :: wtg_mg_params
     real(c_double) :: coarse_tol, fine_tol
     integer(c_int) :: coarse_iter, max_vcycle, pre_iter, down_iter, up_iter, post_iter, max_srj_iter, &
                       learned_cycles
  end type

  interface
     ! ---- grid (init_grid, precompute_geometry) ----
     integer(c_int) function wtg_grid_create (desc, grid) bind(C)
       import; type(wtg_grid_desc), intent(in) :: desc; type(c_ptr) :: grid
     end function
     integer(c_int) function wtg_grid_info_get (grid, info) bind(C)
       import; type(c_ptr), value :: grid; type(wtg_grid_info) :: info
     end function
     ! name: a NUL-terminated array name ("hex_inv", "x", "node_src", ...)
     integer(c_int) function wtg_grid_get (grid, name, out, bytes) bind(C)
       import; type(c_ptr), value :: grid; character(kind=c_char), intent(in) :: name(*)
       type(c_ptr), value :: out; integer(c_size_t), value :: bytes
     end function
     subroutine wtg_grid_destroy (grid) bind(C)
       import; type(c_ptr), value :: grid
     end subroutine
     ! ---- one level's state (sol, trend_ml, RK4) ----
     integer(c_int) function wtg_create (grid, params, nslots, ctx) bind(C)
       import; type(c_ptr), value :: grid; type(wtg_params), intent(in) :: params
       integer(c_int), value :: nslots; type(c_ptr) :: ctx
     end function
     subroutine wtg_destroy (ctx) bind(C)
       import; type(c_ptr), value :: ctx
     end subroutine
     integer(c_int) function wtg_last_error (ctx, buf, n) bind(C)
       import; type(c_ptr), value :: ctx; character(kind=c_char) :: buf(*); integer(c_size_t), value :: n
     end function
     integer(c_int) function wtg_set_topography (ctx, topo) bind(C)
       import; type(c_ptr), value :: ctx; real(c_double), intent(in) :: topo(*)
     end function
     integer(c_int) function wtg_upload_state (ctx, slot, mass, temp, velo) bind(C)
       import; type(c_ptr), value :: ctx; integer(c_int), value :: slot
       real(c_double), intent(in) :: mass(*), temp(*), velo(*)
     end function
     integer(c_int) function wtg_download_state (ctx, slot, mass, temp, velo) bind(C)
       import; type(c_ptr), value :: ctx; integer(c_int), value :: slot
       real(c_double) :: mass(*), temp(*), velo(*)
     end function
     integer(c_int) function wtg_trend (ctx, q_slot, dq_slot) bind(C)
       import; type(c_ptr), value :: ctx; integer(c_int), value :: q_slot, dq_slot
     end function
     integer(c_int) function wtg_rk4 (ctx, slot, dt, nsteps) bind(C)
       import; type(c_ptr), value :: ctx; integer(c_int), value :: slot, nsteps
       real(c_double), value :: dt
     end function
     integer(c_int) function wtg_ghost_fill (ctx, slot) bind(C)
       import; type(c_ptr), value :: ctx; integer(c_int), value :: slot
     end function
     ! precision of wtg_rk4 and the split steps: WTG_PRECISION_FP64 (bit-exact) or _FP32_BULK
     integer(c_int) function wtg_set_precision (ctx, precision) bind(C)
       import; type(c_ptr), value :: ctx; integer(c_int), value :: precision
     end function
     ! FP32-bulk launch shape: 0 = auto (timed), 1 = layer pairs, 2 = edge tasks, 3 = one layer
     integer(c_int) function wtg_set_stage_shape (ctx, shape) bind(C)
       import; type(c_ptr), value :: ctx; integer(c_int), value :: shape
     end function
     ! 1: wtg_rk4 replays its steps from a captured HIP graph
     integer(c_int) function wtg_set_graph (ctx, enable) bind(C)
       import; type(c_ptr), value :: ctx; integer(c_int), value :: enable
     end function
     ! Laplacian diffusion / hyperdiffusion in the trend (FP64, single level)
     integer(c_int) function wtg_set_diffusion (ctx, d) bind(C)
       import; type(c_ptr), value :: ctx; type(wtg_diffusion), intent(in) :: d
     end function
     ! remap_vertical_coordinates onto z = a_vert eta + b_vert z_s (a_vert, b_vert(0:zlevels))
     integer(c_int) function wtg_remap (ctx, slot, scheme, a_vert, b_vert) bind(C)
       import; type(c_ptr), value :: ctx; integer(c_int), value :: slot, scheme
       real(c_double), intent(in) :: a_vert(*), b_vert(*)
     end function
     ! ---- vertical_diffusion (Kt, Kv(0:zlevels), tke(1:zlevels) and the forcing per node) ----
     subroutine wtg_vdiff_defaults (p) bind(C)
       import; type(wtg_vdiff), intent(out) :: p
     end subroutine
     integer(c_int) function wtg_set_vdiff (ctx, p) bind(C)
       import; type(c_ptr), value :: ctx; type(wtg_vdiff), intent(in) :: p
     end function
     integer(c_int) function wtg_vdiff_set_coefficients (ctx, Kt, Kv, tke) bind(C)
       import; type(c_ptr), value :: ctx; real(c_double), intent(in) :: Kt(*), Kv(*), tke(*)
     end function
     integer(c_int) function wtg_vdiff_get_coefficients (ctx, Kt, Kv, tke) bind(C)
       import; type(c_ptr), value :: ctx; real(c_double), intent(out) :: Kt(*), Kv(*), tke(*)
     end function
     integer(c_int) function wtg_vdiff_set_forcing (ctx, bottom_buoy_flux, top_buoy_flux, tau_mag, wind_flux) bind(C)
       import; type(c_ptr), value :: ctx
       real(c_double), intent(in) :: bottom_buoy_flux(*), top_buoy_flux(*), tau_mag(*), wind_flux(*)
     end function
     integer(c_int) function wtg_vertical_diffusion (ctx, slot, dt, istep) bind(C)
       import; type(c_ptr), value :: ctx; integer(c_int), value :: slot, istep; real(c_double), value :: dt
     end function
     ! ---- free surface (mode_split) ----
     integer(c_int) function wtg_fs_set_params (ctx, p) bind(C)
       import; type(c_ptr), value :: ctx; type(wtg_fs_params), intent(in) :: p
     end function
     integer(c_int) function wtg_fs_upload (ctx, field, host) bind(C)
       import; type(c_ptr), value :: ctx; integer(c_int), value :: field; real(c_double), intent(in) :: host(*)
     end function
     integer(c_int) function wtg_fs_download (ctx, field, host) bind(C)
       import; type(c_ptr), value :: ctx; integer(c_int), value :: field; real(c_double) :: host(*)
     end function
     integer(c_int) function wtg_rk4_split (ctx, slot, nsteps, rel_tol, max_iter, iters) bind(C)
       import; type(c_ptr), value :: ctx; integer(c_int), value :: slot, nsteps, max_iter
       real(c_double), value :: rel_tol; integer(c_int) :: iters
     end function
     integer(c_int) function wtg_elliptic_solve (ctx, method, rel_tol, max_iter, iters, rel_resid) bind(C)
       import; type(c_ptr), value :: ctx; integer(c_int), value :: method, max_iter
       real(c_double), value :: rel_tol; integer(c_int) :: iters; real(c_double) :: rel_resid
     end function
     ! ---- hierarchy of uniform levels (grids(l), ctxs(l) coarse to fine) ----
     subroutine wtg_mg_default_params (p) bind(C)
       import; type(wtg_mg_params) :: p
     end subroutine
     integer(c_int) function wtg_mg_create (grids, ctxs, nlevels, mg) bind(C)
       import; type(c_ptr), intent(in) :: grids(*), ctxs(*); integer(c_int), value :: nlevels; type(c_ptr) :: mg
     end function
     subroutine wtg_mg_destroy (mg) bind(C)
       import; type(c_ptr), value :: mg
     end subroutine
     integer(c_int) function wtg_mg_solve (mg, method, p, iters, rel_resid) bind(C)
       import; type(c_ptr), value :: mg; integer(c_int), value :: method
       type(wtg_mg_params), intent(in) :: p; integer(c_int) :: iters(*); real(c_double) :: rel_resid(*)
     end function
     ! V-cycles: WTG_PRECISION_FP64 or FP32 smoothing with FP64 residuals
     integer(c_int) function wtg_mg_set_precision (mg, precision) bind(C)
       import; type(c_ptr), value :: mg; integer(c_int), value :: precision
     end function
     ! multi-level step: trend_ml with compute-or-restrict, WT_after_step, RK4, RK4_split
     integer(c_int) function wtg_ml_create (grids, ctxs, nlevels, ml) bind(C)
       import; type(c_ptr), intent(in) :: grids(*), ctxs(*); integer(c_int), value :: nlevels; type(c_ptr) :: ml
     end function
     subroutine wtg_ml_destroy (ml) bind(C)
       import; type(c_ptr), value :: ml
     end subroutine
     integer(c_int) function wtg_ml_forward (ml, slot) bind(C)
       import; type(c_ptr), value :: ml; integer(c_int), value :: slot
     end function
     integer(c_int) function wtg_ml_rk4 (ml, slot, dt, nsteps) bind(C)
       import; type(c_ptr), value :: ml; integer(c_int), value :: slot, nsteps; real(c_double), value :: dt
     end function
     integer(c_int) function wtg_ml_rk4_split (ml, mg, slot, nsteps, p, iters) bind(C)
       import; type(c_ptr), value :: ml, mg; integer(c_int), value :: slot, nsteps
       type(wtg_mg_params), intent(in) :: p; integer(c_int) :: iters
     end function
  end interface

contains

  subroutine wtg_check (rc, ctx)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in), optional :: ctx
    character(kind=c_char) :: buf(512)
    integer :: i, ierr
    if (rc == 0) return
    if (present(ctx)) then
       ierr = wtg_last_error(ctx, buf, int(size(buf), c_size_t))
       write (0, '(a,i0,a)', advance='no') 'wavetrisk_gpu: error ', rc, ': '
       do i = 1, size(buf)
          if (buf(i) == c_null_char) exit
          write (0, '(a)', advance='no') buf(i)
       end do
       write (0, *)
    else
       write (0, '(a,i0)') 'wavetrisk_gpu: error ', rc
    end if
    error stop 1
  end subroutine
end module wavetrisk_gpu
