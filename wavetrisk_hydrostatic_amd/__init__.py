"""MI355X-native TRiSK dynamical-core hot path (host-side Python mirror).

Thin ctypes binding of the C ABI in include/wavetrisk_gpu.h.  The compute
path is the in-tree HIP library libwavetrisk_gpu.so; there is no CPU
fallback: every entry point raises if the library is missing or fails.

    grid = Grid(level=5, min_level=5, domain_level=2)       # host-side grid
    ctx = Context(grid, nlayers=1)                           # needs a GPU
    ctx.set_topography(topo); ctx.upload_state(0, mass, temp, velo)
    ctx.rk4(0, dt, nsteps)                                   # dt_step => RK4
    mass, temp, velo = ctx.download_state(0)

Array layouts follow the header: scalars [nlayers, ncells], velocity
[nlayers, 3, ncells] (RT, DG, UP), float64, C-contiguous.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .layout import Layout

__all__ = ["Grid", "Context", "Hierarchy", "MgParams", "MLP", "Layout", "WtgError", "library_path", "load_library",
           "FS_ETA", "FS_RHS", "FS_ETA_OLD", "FS_DIVF", "FS_DIVF_OLD", "SOLVE_BICGSTAB", "SOLVE_JACOBI", "SOLVE_FMG",
           "SOLVE_SRJ", "SOLVE_VCYCLE", "SOLVE_LEARNED", "PRECISION_FP64", "PRECISION_FP32_BULK",
           "SHAPE_AUTO", "SHAPE_LAYER_PAIRS", "SHAPE_EDGE_TASKS", "SHAPE_ONE_LAYER"]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_NAME = "libwavetrisk_gpu.so"
_lib = None


class WtgError(RuntimeError):
    pass


def library_path() -> str:
    return os.path.join(_HERE, _LIB_NAME)


class _GridDesc(C.Structure):
    _fields_ = [("level", C.c_int), ("min_level", C.c_int), ("domain_level", C.c_int),
                ("radius", C.c_double), ("omega", C.c_double)]


class _GridInfo(C.Structure):
    _fields_ = [("n", C.c_int), ("S", C.c_int), ("halo", C.c_int), ("ncells", C.c_int64),
                ("south_pole", C.c_int64), ("north_pole", C.c_int64)]


class _Params(C.Structure):
    _fields_ = [("nlayers", C.c_int), ("grav", C.c_double), ("ref_density", C.c_double), ("p_top", C.c_double)]


class _VDiff(C.Structure):
    _fields_ = ([(k, C.c_int) for k in ("tke_closure", "patankar", "enhance_diff", "mode_split")] +
                [(k, C.c_double) for k in ("C_e", "C_eps", "C_l", "C_k", "C_srf", "e_0", "e_min", "e_min_srf", "eps_s",
                                           "kappa_VK", "Kt_enh", "Kt_max", "Kt_min", "Kv_max", "Kv_min",
                                           "mixed_layer", "l_0", "l_min", "Neps_sq", "Nsq_min", "Q_sr", "c_p",
                                           "bottom_friction")])


class _Diffusion(C.Structure):
    _fields_ = [("laplace_sclr", C.c_int), ("laplace_divu", C.c_int), ("laplace_rotu", C.c_int),
                ("visc_mass", C.c_double), ("visc_temp", C.c_double), ("visc_divu", C.c_double),
                ("visc_rotu", C.c_double)]


_P = C.c_void_p
_D = C.POINTER(C.c_double)

# name -> (restype, argtypes); every function of include/wavetrisk_gpu.h
SIGNATURES = {
    "wtg_grid_create": (C.c_int, [C.POINTER(_GridDesc), C.POINTER(_P)]),
    "wtg_grid_info_get": (C.c_int, [_P, C.POINTER(_GridInfo)]),
    "wtg_grid_get": (C.c_int, [_P, C.c_char_p, _P, C.c_size_t]),
    "wtg_grid_destroy": (None, [_P]),
    "wtg_grid_overlap_areas": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "wtg_grid_transfer_weights": (C.c_int, [_P, _P, C.c_char_p, _P, C.c_size_t]),
    "wtg_create": (C.c_int, [_P, C.POINTER(_Params), C.c_int, C.POINTER(_P)]),
    "wtg_create_ensemble": (C.c_int, [_P, C.POINTER(_Params), C.c_int, C.c_int, C.POINTER(_P)]),
    "wtg_destroy": (None, [_P]),
    "wtg_set_topography": (C.c_int, [_P, _P]),
    "wtg_upload_state": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "wtg_download_state": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "wtg_ghost_fill": (C.c_int, [_P, C.c_int]),
    "wtg_trend": (C.c_int, [_P, C.c_int, C.c_int]),
    "wtg_rk4": (C.c_int, [_P, C.c_int, C.c_double, C.c_int]),
    "wtg_rk4_timed": (C.c_int, [_P, C.c_int, C.c_double, C.c_int, _D, _D]),
    "wtg_set_precision": (C.c_int, [_P, C.c_int]),
    "wtg_set_stage_shape": (C.c_int, [_P, C.c_int]),
    "wtg_set_graph": (C.c_int, [_P, C.c_int]),
    "wtg_set_diffusion": (C.c_int, [_P, C.POINTER(_Diffusion)]),
    "wtg_remap": (C.c_int, [_P, C.c_int, C.c_int, _P, _P]),
    "wtg_vdiff_defaults": (None, [C.POINTER(_VDiff)]),
    "wtg_set_vdiff": (C.c_int, [_P, C.POINTER(_VDiff)]),
    "wtg_vdiff_set_coefficients": (C.c_int, [_P, _P, _P, _P]),
    "wtg_vdiff_get_coefficients": (C.c_int, [_P, _P, _P, _P]),
    "wtg_vdiff_set_forcing": (C.c_int, [_P, _P, _P, _P, _P]),
    "wtg_vertical_diffusion": (C.c_int, [_P, C.c_int, C.c_double, C.c_int]),
    "wtg_grid_shard_halo": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int64)]),
    "wtg_ghost_fill_split_timed": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _D, _D]),
    "wtg_fs_set_graph": (C.c_int, [_P, C.c_int]),
    "wtg_get_stage_shape": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_float)]),
    "wtg_state_device_ptrs": (C.c_int, [_P, C.c_int, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]),
    "wtg_synchronize": (C.c_int, [_P]),
    "wtg_fs_dots_noalloc": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "wtg_last_error": (C.c_int, [_P, C.c_char_p, C.c_size_t]),
    "wtg_fs_set_params": (C.c_int, [_P, _P]),
    "wtg_fs_upload": (C.c_int, [_P, C.c_int, _P]),
    "wtg_fs_download": (C.c_int, [_P, C.c_int, _P]),
    "wtg_elliptic_apply": (C.c_int, [_P, C.c_int, C.c_int]),
    "wtg_elliptic_diag": (C.c_int, [_P, C.c_int]),
    "wtg_jacobi": (C.c_int, [_P, C.c_int, C.c_int, C.c_double, C.c_int]),
    "wtg_fs_dot": (C.c_int, [_P, C.c_int, C.c_int, _D]),
    "wtg_flux_divergence": (C.c_int, [_P, C.c_int]),
    "wtg_rhs_elliptic": (C.c_int, [_P]),
    "wtg_elliptic_solve": (C.c_int, [_P, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_int), _D]),
    "wtg_rk4_split": (C.c_int, [_P, C.c_int, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_int)]),
    "wtg_rk_split": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_int)]),
    "wtg_mg_default_params": (None, [_P]),
    "wtg_mg_create": (C.c_int, [_P, _P, C.c_int, C.POINTER(_P)]),
    "wtg_mg_member_results": (C.c_int, [_P, _P, _P]),
    "wtg_ml_create": (C.c_int, [_P, _P, C.c_int, C.POINTER(_P)]),
    "wtg_ml_destroy": (None, [_P]),
    "wtg_ml_forward": (C.c_int, [_P, C.c_int]),
    "wtg_ml_inverse": (C.c_int, [_P, C.c_int]),
    "wtg_ml_wavelets": (C.c_int, [_P, C.c_int, C.c_int, _P, _P, _P]),
    "wtg_ml_trend": (C.c_int, [_P, C.c_int, C.c_int]),
    "wtg_ml_wt_after_step": (C.c_int, [_P, C.c_int]),
    "wtg_ml_rk4": (C.c_int, [_P, C.c_int, C.c_double, C.c_int]),
    "wtg_ml_rk4_split": (C.c_int, [_P, _P, C.c_int, C.c_int, _P, C.POINTER(C.c_int)]),
    "wtg_ml_rk_split": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, _P, C.POINTER(C.c_int)]),
    "wtg_mg_destroy": (None, [_P]),
    "wtg_mg_solve": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "wtg_mg_forward": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int]),
    "wtg_mg_inverse": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int]),
    "wtg_mg_rk4_split": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, C.POINTER(C.c_int)]),
    "wtg_mg_set_precision": (C.c_int, [_P, C.c_int]),
    "wtg_restrict_velo": (C.c_int, [_P, C.c_int, _P, C.c_int]),
    "wtg_mg_restrict_ss": (C.c_int, [_P, C.c_int, C.c_int, C.c_int]),
    "wtg_mg_npatches": (C.c_int, [_P]),
    "wtg_mg_set_learned": (C.c_int, [_P, _P]),
    "wtg_mg_features": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "wtg_mlp_create": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(_P)]),
    "wtg_mlp_destroy": (None, [_P]),
    "wtg_mlp_last_error": (C.c_int, [_P, C.c_char_p, C.c_size_t]),
    "wtg_mlp_get_params": (C.c_int, [_P, _P, C.c_size_t]),
    "wtg_mlp_set_params": (C.c_int, [_P, _P, C.c_size_t]),
    "wtg_mlp_forward": (C.c_int, [_P, _P, C.c_int, _P]),
    "wtg_mlp_forward_timed": (C.c_int, [_P, C.c_int, C.c_int, _D]),
    "wtg_mlp_gradients": (C.c_int, [_P, _P, _P, C.c_int, C.POINTER(C.c_float), _P]),
    "wtg_mlp_set_dataset": (C.c_int, [_P, _P, _P, C.c_int64]),
    "wtg_mlp_train": (C.c_int, [_P, C.c_int, C.c_int, C.c_float, C.c_uint64, _P]),
}

# free-surface fields and solver methods (include/wavetrisk_gpu.h)
FS_ETA, FS_RHS, FS_ETA_OLD, FS_DIVF, FS_DIVF_OLD = 0, 1, 2, 3, 4
SOLVE_BICGSTAB, SOLVE_JACOBI, SOLVE_FMG, SOLVE_SRJ, SOLVE_VCYCLE, SOLVE_LEARNED = 0, 1, 2, 3, 4, 5
# precision of Context.rk4 / rk4_timed: FP64 (bit-exact with the reference's
# RK4 of trend_ml) or FP32 bulk stencils with FP64 column scan and updates
PRECISION_FP64, PRECISION_FP32_BULK = 0, 1
REMAP_PIECEWISE_CONSTANT, REMAP_PIECEWISE_LINEAR, REMAP_PARABOLIC_PPM, REMAP_PPR = 0, 1, 2, 3
REMAP_PARABOLIC_SPLINE, REMAP_PARABOLIC_WENO, REMAP_QUARTIC_WENO, REMAP_MAX_LAYERS = 4, 5, 6, 64
# launch shape of the FP32-bulk stage kernel (AUTO: timed on the first step)
SHAPE_AUTO, SHAPE_LAYER_PAIRS, SHAPE_EDGE_TASKS, SHAPE_ONE_LAYER = 0, 1, 2, 3


class _FsParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("theta1", C.c_double), ("theta2", C.c_double)]


class MgParams(C.Structure):
    """wtg_mg_params: tolerances and iteration counts of FMG / SRJ
    (lin_solve.f90:11-26); MgParams.defaults() gives the reference's."""
    _fields_ = [("coarse_tol", C.c_double), ("fine_tol", C.c_double), ("coarse_iter", C.c_int),
                ("max_vcycle", C.c_int), ("pre_iter", C.c_int), ("down_iter", C.c_int), ("up_iter", C.c_int),
                ("post_iter", C.c_int), ("max_srj_iter", C.c_int), ("learned_cycles", C.c_int)]

    @classmethod
    def defaults(cls, **overrides):
        p = cls()
        load_library().wtg_mg_default_params(C.byref(p))
        for k, v in overrides.items():
            setattr(p, k, v)
        return p


def load_library():
    """Loads the in-tree HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise WtgError(f"{path} is missing: build it with `python -m wavetrisk_hydrostatic_amd.build`")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _err(ctx_handle) -> str:
    buf = C.create_string_buffer(1024)
    load_library().wtg_last_error(ctx_handle, buf, len(buf))
    return buf.value.decode(errors="replace")


def _check(rc: int, ctx_handle=None, what: str = ""):
    if rc != 0:
        raise WtgError(f"{what} failed ({rc}): {_err(ctx_handle)}")


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


class Grid:
    """Host-side icosahedral grid of one uniform level (no GPU needed)."""

    def __init__(self, level: int, min_level: int | None = None, domain_level: int = 2,
                 radius: float = 6.371e6, omega: float = 7.292e-5):
        lib = load_library()
        if min_level is None:
            min_level = level
        self.desc = _GridDesc(level, min_level, domain_level, radius, omega)
        h = _P()
        _check(lib.wtg_grid_create(C.byref(self.desc), C.byref(h)), None, "wtg_grid_create")
        self._h = h
        info = _GridInfo()
        _check(lib.wtg_grid_info_get(h, C.byref(info)), None, "wtg_grid_info_get")
        self.layout = Layout(level=level, n=info.n, S=info.S, halo=info.halo)
        assert self.layout.ncells == info.ncells
        assert self.layout.south_pole == info.south_pole and self.layout.north_pole == info.north_pole
        self.radius, self.omega = radius, omega

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.wtg_grid_destroy(self._h)
            self._h = None

    _SHAPES = {"x": (3, np.float64), "len": (3, np.float64), "pedlen": (3, np.float64), "area": (6, np.float64),
               "hex_inv": (1, np.float64), "triarea": (2, np.float64), "coriolis": (2, np.float64),
               "node_src": (1, np.int64), "edge_src": (3, np.int64), "edge_sign": (3, np.int8),
               "red_weight": (1, np.float64)}

    def overlap_areas(self, coarse: "Grid") -> np.ndarray:
        """a(1..4) of the scalar wavelet transform between this grid and the
        next coarser one, [4, ncells]."""
        out = np.empty(4 * self.layout.ncells)
        _check(load_library().wtg_grid_overlap_areas(self._h, coarse._h, _ptr(out), out.nbytes), None,
               "wtg_grid_overlap_areas")
        return out.reshape(4, self.layout.ncells)

    WIDE_HALO = 5

    def wide_src(self) -> np.ndarray:
        """Owned cell of every lattice point within WIDE_HALO cells of each
        lozenge's owned square, [10, n+2W, n+2W] flattened (-1: no lattice)."""
        ws = self.layout.n + 2 * self.WIDE_HALO
        out = np.empty(10 * ws * ws, dtype=np.int64)
        _check(load_library().wtg_grid_get(self._h, b"wide_src", _ptr(out), out.nbytes), None, "wtg_grid_get(wide_src)")
        return out

    def transfer_weights(self, coarse: "Grid", name: str) -> np.ndarray:
        """Velocity/flux transfer weights to the next coarser grid: "split"
        [2, ncells], "iu" [9, ncells] (fine cells), "rf" [4, 6, coarse ncells]."""
        shape = {"split": (2, self.layout.ncells), "iu": (9, self.layout.ncells),
                 "rf": (4, 6, coarse.layout.ncells)}[name]
        out = np.empty(int(np.prod(shape)))
        _check(load_library().wtg_grid_transfer_weights(self._h, coarse._h, name.encode(), _ptr(out), out.nbytes),
               None, "wtg_grid_transfer_weights")
        return out.reshape(shape)

    def array(self, name: str) -> np.ndarray:
        comp, dt = self._SHAPES[name]
        nc = self.layout.ncells
        out = np.empty(comp * nc, dtype=dt)
        _check(load_library().wtg_grid_get(self._h, name.encode(), _ptr(out), out.nbytes), None, f"wtg_grid_get({name})")
        if name in ("edge_src", "edge_sign"):
            return out  # slot = 3*cell + e
        return out.reshape(comp, nc) if comp > 1 else out

    def shard_halo(self, nshards: int):
        """Ghost entries of one layer for whole-lozenge shards: (node local,
        node remote, edge local, edge remote)."""
        counts = (C.c_int64 * 4)()
        _check(load_library().wtg_grid_shard_halo(self._h, nshards, counts), None, "wtg_grid_shard_halo")
        return tuple(int(v) for v in counts)

    @property
    def handle(self):
        return self._h


class Context:
    """Device state of one grid level on the current HIP device."""

    def __init__(self, grid: Grid, nlayers: int, grav: float = 9.80616, ref_density: float = 1028.0,
                 p_top: float = 0.0, nslots: int = 2, members: int = 1):
        """members > 1: an ensemble of independent states (wtg_create_ensemble);
        state arrays then have members * nlayers layers, member-major."""
        lib = load_library()
        self.grid = grid
        self.layout = grid.layout
        self.nlayers = nlayers
        self.members = members
        self.params = _Params(nlayers, grav, ref_density, p_top)
        h = _P()
        _check(lib.wtg_create_ensemble(grid.handle, C.byref(self.params), members, nslots, C.byref(h)), None,
               "wtg_create_ensemble")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.wtg_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _shape(self):
        L, nc = self.nlayers * self.members, self.layout.ncells
        return (L, nc), (L, 3, nc)

    def set_topography(self, topo: np.ndarray):
        """[ncells], or [members, ncells] for an ensemble (one array is repeated)"""
        topo = np.ascontiguousarray(topo, dtype=np.float64)
        if self.members > 1 and topo.shape == (self.layout.ncells,):
            topo = np.ascontiguousarray(np.broadcast_to(topo, (self.members, self.layout.ncells)))
        assert topo.shape in ((self.layout.ncells,), (self.members, self.layout.ncells))
        _check(_lib.wtg_set_topography(self._h, _ptr(topo)), self._h, "wtg_set_topography")

    def upload_state(self, slot: int, mass, temp, velo):
        s_shape, e_shape = self._shape()
        mass = np.ascontiguousarray(mass, dtype=np.float64).reshape(s_shape)
        temp = np.ascontiguousarray(temp, dtype=np.float64).reshape(s_shape)
        velo = np.ascontiguousarray(velo, dtype=np.float64).reshape(e_shape)
        _check(_lib.wtg_upload_state(self._h, slot, _ptr(mass), _ptr(temp), _ptr(velo)), self._h, "wtg_upload_state")

    def download_state(self, slot: int):
        s_shape, e_shape = self._shape()
        mass, temp, velo = np.empty(s_shape), np.empty(s_shape), np.empty(e_shape)
        _check(_lib.wtg_download_state(self._h, slot, _ptr(mass), _ptr(temp), _ptr(velo)), self._h,
               "wtg_download_state")
        return mass, temp, velo

    def ghost_fill(self, slot: int):
        _check(_lib.wtg_ghost_fill(self._h, slot), self._h, "wtg_ghost_fill")

    def ghost_fill_split_timed(self, slot: int, nshards: int, reps: int = 20):
        """The ghost fill as local and would-be-remote launches: (ms, ms)."""
        a, b = C.c_double(), C.c_double()
        _check(_lib.wtg_ghost_fill_split_timed(self._h, slot, nshards, reps, C.byref(a), C.byref(b)), self._h,
               "wtg_ghost_fill_split_timed")
        return a.value, b.value

    def trend(self, q_slot: int, dq_slot: int):
        _check(_lib.wtg_trend(self._h, q_slot, dq_slot), self._h, "wtg_trend")

    def rk4(self, slot: int, dt: float, nsteps: int = 1):
        _check(_lib.wtg_rk4(self._h, slot, dt, nsteps), self._h, "wtg_rk4")

    def set_precision(self, precision: int):
        """PRECISION_FP64 or PRECISION_FP32_BULK for rk4 / rk4_timed."""
        _check(_lib.wtg_set_precision(self._h, precision), self._h, "wtg_set_precision")

    def fs_set_graph(self, enable: bool):
        """BiCGSTAB chunks replayed from a captured HIP graph (default off)."""
        _check(_lib.wtg_fs_set_graph(self._h, int(bool(enable))), self._h, "wtg_fs_set_graph")

    def set_graph(self, enable: bool):
        """rk4 replays its steps from a captured HIP graph (identical results)."""
        _check(_lib.wtg_set_graph(self._h, int(bool(enable))), self._h, "wtg_set_graph")

    def remap(self, slot: int, scheme: int, a_vert, b_vert):
        """remap_vertical_coordinates of slot onto z = a_vert eta + b_vert z_s
        (nlayers + 1 interfaces, bottom up): REMAP_PIECEWISE_CONSTANT,
        REMAP_PIECEWISE_LINEAR, REMAP_PARABOLIC_PPM, REMAP_PPR,
        REMAP_PARABOLIC_SPLINE, REMAP_PARABOLIC_WENO or REMAP_QUARTIC_WENO
        (the reference's remap_type "0", "1", "2PPM", its default "PPR",
        "2S", "2W" and "4")"""
        a = np.ascontiguousarray(a_vert, dtype=np.float64)
        b = np.ascontiguousarray(b_vert, dtype=np.float64)
        L = self.params.nlayers
        if a.shape != (L + 1,) or b.shape != (L + 1,):
            raise ValueError(f"a_vert and b_vert need {L + 1} interfaces")
        _check(_lib.wtg_remap(self._h, slot, scheme, _ptr(a), _ptr(b)), self._h, "wtg_remap")

    def set_vdiff(self, **params):
        """vertical_diffusion's switches and constants (wtg_vdiff: tke_closure,
        patankar, enhance_diff, mode_split, C_e, ..., bottom_friction); those
        not given keep the reference's defaults"""
        p = _VDiff()
        _lib.wtg_vdiff_defaults(C.byref(p))
        for k, v in params.items():
            if k not in dict(_VDiff._fields_):
                raise ValueError(f"unknown vertical-diffusion parameter {k}")
            setattr(p, k, v)
        _check(_lib.wtg_set_vdiff(self._h, C.byref(p)), self._h, "wtg_set_vdiff")

    def vdiff_set_coefficients(self, Kt=None, Kv=None, tke=None):
        """Kt, Kv [members * (nlayers + 1), ncells], tke [members * nlayers, ncells]"""
        a = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in (Kt, Kv, tke)]
        _check(_lib.wtg_vdiff_set_coefficients(self._h, *[None if x is None else _ptr(x) for x in a]), self._h,
               "wtg_vdiff_set_coefficients")

    def vdiff_coefficients(self):
        """(Kt, Kv, tke) as set by the last call"""
        L, M, NC = self.params.nlayers, self.members, self.layout.ncells
        Kt, Kv, tke = np.empty((M * (L + 1), NC)), np.empty((M * (L + 1), NC)), np.empty((M * L, NC))
        _check(_lib.wtg_vdiff_get_coefficients(self._h, _ptr(Kt), _ptr(Kv), _ptr(tke)), self._h,
               "wtg_vdiff_get_coefficients")
        return Kt, Kv, tke

    def vdiff_set_forcing(self, bottom_buoy_flux=None, top_buoy_flux=None, tau_mag=None, wind_flux=None):
        """the test case's hooks per node: [members * ncells] fluxes and wind
        stress, [members * 3, ncells] wind flux of the RT, DG, UP edges"""
        a = [None if x is None else np.ascontiguousarray(x, dtype=np.float64)
             for x in (bottom_buoy_flux, top_buoy_flux, tau_mag, wind_flux)]
        _check(_lib.wtg_vdiff_set_forcing(self._h, *[None if x is None else _ptr(x) for x in a]), self._h,
               "wtg_vdiff_set_forcing")

    def vertical_diffusion(self, slot: int, dt: float, istep: int):
        """vertical_diffusion (vert_diffusion.f90:66-86) of slot; istep 1
        recomputes Kt, Kv from the TKE first"""
        _check(_lib.wtg_vertical_diffusion(self._h, slot, float(dt), int(istep)), self._h, "wtg_vertical_diffusion")

    def set_diffusion(self, laplace=(0, 0, 0), visc=(0.0, 0.0, 0.0, 0.0)):
        """Laplacian diffusion (order 1) or hyperdiffusion (order 2) in the
        trend: laplace = (Laplace_sclr, Laplace_divu, Laplace_rotu), visc =
        (visc of mass, of temperature, visc_divu, visc_rotu); all orders 0
        switch it off.  FP64 precision on a single level only."""
        d = _Diffusion(*[int(x) for x in laplace], *[float(x) for x in visc])
        _check(_lib.wtg_set_diffusion(self._h, C.byref(d)), self._h, "wtg_set_diffusion")

    def set_stage_shape(self, shape: int):
        """SHAPE_AUTO, SHAPE_LAYER_PAIRS, SHAPE_EDGE_TASKS or SHAPE_ONE_LAYER
        for the FP32-bulk stage kernel."""
        _check(_lib.wtg_set_stage_shape(self._h, shape), self._h, "wtg_set_stage_shape")

    def stage_shape(self):
        """(shape in use, [tuner ms of shapes 1..3])."""
        sh, ms = C.c_int(), (C.c_float * 3)()
        _check(_lib.wtg_get_stage_shape(self._h, C.byref(sh), ms), self._h, "wtg_get_stage_shape")
        return sh.value, list(ms)

    def rk4_timed(self, slot: int, dt: float, nsteps: int):
        tot, ker = C.c_double(), C.c_double()
        _check(_lib.wtg_rk4_timed(self._h, slot, dt, nsteps, C.byref(tot), C.byref(ker)), self._h, "wtg_rk4_timed")
        return tot.value, ker.value

    def synchronize(self):
        _check(_lib.wtg_synchronize(self._h), self._h, "wtg_synchronize")

    def fs_dots_noalloc(self, a_ptrs, b_ptrs, out_ptr: int):
        """wtg_fs_dots_noalloc: len(a_ptrs) weighted dot products of device
        arrays (addresses) into the device array at out_ptr, queued on the
        context's stream (call synchronize before reading out)."""
        nd = len(a_ptrs)
        A = (_P * max(nd, 1))(*[_P(int(p)) for p in a_ptrs])
        B = (_P * max(nd, 1))(*[_P(int(p)) for p in b_ptrs])
        _check(_lib.wtg_fs_dots_noalloc(self._h, nd, A, B, _P(int(out_ptr))), self._h, "wtg_fs_dots_noalloc")

    # --- free-surface elliptic problem (node fields [ncells]) ---------------
    def fs_set_params(self, dt: float, theta1: float = 1.0, theta2: float = 1.0):
        p = _FsParams(dt, theta1, theta2)
        _check(_lib.wtg_fs_set_params(self._h, C.byref(p)), self._h, "wtg_fs_set_params")

    def restrict_velo(self, slot: int, coarse: "Context", coarse_slot: int):
        """velocity of `slot` restricted onto the next coarser level's slot
        (Restrict_velo, wavelet.f90:1154-1174), ghost filled there"""
        _check(_lib.wtg_restrict_velo(self._h, slot, coarse._h, coarse_slot), self._h, "wtg_restrict_velo")

    def fs_upload(self, field: int, values):
        """[ncells], or [members, ncells] for an ensemble"""
        a = np.ascontiguousarray(values, dtype=np.float64)
        shape = (self.layout.ncells,) if self.members == 1 else (self.members, self.layout.ncells)
        assert a.shape == shape, (a.shape, shape)
        _check(_lib.wtg_fs_upload(self._h, field, _ptr(a)), self._h, "wtg_fs_upload")

    def fs_download(self, field: int) -> np.ndarray:
        a = np.empty(self.layout.ncells if self.members == 1 else (self.members, self.layout.ncells))
        _check(_lib.wtg_fs_download(self._h, field, _ptr(a)), self._h, "wtg_fs_download")
        return a

    def elliptic_apply(self, in_field: int, out_field: int):
        _check(_lib.wtg_elliptic_apply(self._h, in_field, out_field), self._h, "wtg_elliptic_apply")

    def elliptic_diag(self, out_field: int):
        _check(_lib.wtg_elliptic_diag(self._h, out_field), self._h, "wtg_elliptic_diag")

    def jacobi(self, u_field: int, f_field: int, w: float, nsweeps: int = 1):
        _check(_lib.wtg_jacobi(self._h, u_field, f_field, w, nsweeps), self._h, "wtg_jacobi")

    def fs_dot(self, a_field: int, b_field: int) -> float:
        out = C.c_double()
        _check(_lib.wtg_fs_dot(self._h, a_field, b_field, C.byref(out)), self._h, "wtg_fs_dot")
        return out.value

    def flux_divergence(self, slot: int):
        _check(_lib.wtg_flux_divergence(self._h, slot), self._h, "wtg_flux_divergence")

    def rhs_elliptic(self):
        _check(_lib.wtg_rhs_elliptic(self._h), self._h, "wtg_rhs_elliptic")

    def rk4_split(self, slot: int, nsteps: int = 1, rel_tol: float = 1e-9, max_iter: int = 400,
                  nstages: int = 4) -> int:
        """split-explicit RK4 steps with implicit free surface (FS_ETA); dt
        from fs_set_params; returns the last step's BiCGSTAB iterations.
        nstages 3, 2, 1: RK3_split, RK2_split, Euler_split (wtg_rk_split)"""
        it = C.c_int()
        if nstages == 4:
            _check(_lib.wtg_rk4_split(self._h, slot, nsteps, rel_tol, max_iter, C.byref(it)), self._h,
                   "wtg_rk4_split")
        else:
            _check(_lib.wtg_rk_split(self._h, slot, nstages, nsteps, rel_tol, max_iter, C.byref(it)), self._h,
                   "wtg_rk_split")
        return it.value

    def elliptic_solve(self, method: int = SOLVE_BICGSTAB, rel_tol: float = 1e-9, max_iter: int = 200):
        it, res = C.c_int(), C.c_double()
        _check(_lib.wtg_elliptic_solve(self._h, method, rel_tol, max_iter, C.byref(it), C.byref(res)), self._h,
               "wtg_elliptic_solve")
        return it.value, res.value


class Hierarchy:
    """Free-surface contexts on consecutive uniform levels, coarse to fine
    (wtg_mg_*): one level of the scalar wavelet transform each way, and the
    FMG / SRJ elliptic solvers.  While it exists the contexts run on the
    finest one's stream; close it before the contexts."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        n = len(self.contexts)
        grids = (_P * n)(*[c.grid.handle.value for c in self.contexts])
        ctxs = (_P * n)(*[c._h.value for c in self.contexts])
        h = _P()
        _check(load_library().wtg_mg_create(grids, ctxs, n, C.byref(h)), None, "wtg_mg_create")
        self._h = h

    @property
    def _top(self):
        return self.contexts[-1]._h

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.wtg_mg_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def solve(self, method: int = SOLVE_FMG, params: MgParams | None = None):
        """u on every level's FS_ETA; returns (iterations, relative residuals) per level"""
        p = params if params is not None else MgParams.defaults()
        n = len(self.contexts)
        it = np.zeros(n, dtype=np.int32)
        res = np.zeros(n)
        _check(_lib.wtg_mg_solve(self._h, method, C.byref(p), _ptr(it), _ptr(res)), self._top, "wtg_mg_solve")
        return it, res

    def rk4_split(self, slot: int, nsteps: int = 1, method: int = SOLVE_VCYCLE, params: MgParams | None = None) -> int:
        """split-explicit steps on the finest context with the free surface
        solved on the hierarchy; returns the last step's cycle count"""
        p = params if params is not None else MgParams.defaults()
        it = C.c_int()
        _check(_lib.wtg_mg_rk4_split(self._h, slot, nsteps, method, C.byref(p), C.byref(it)), self._top,
               "wtg_mg_rk4_split")
        return it.value

    def member_results(self):
        """per member of an ensemble hierarchy: (cycles, relative residual) of
        the last member-batched solve"""
        M = self.contexts[-1].members
        cyc = np.zeros(M, dtype=np.int32)
        res = np.zeros(M)
        _check(_lib.wtg_mg_member_results(self._h, _ptr(cyc), _ptr(res)), self._top, "wtg_mg_member_results")
        return cyc, res

    @property
    def npatches(self) -> int:
        return _lib.wtg_mg_npatches(self._h)

    def set_learned(self, mlp: "MLP | None"):
        self._mlp = mlp
        _check(_lib.wtg_mg_set_learned(self._h, mlp._h if mlp is not None else None), self._top, "wtg_mg_set_learned")

    def restrict_ss(self, fine: int, fine_field: int, coarse_field: int):
        """coarse_field of level fine - 1 := fine_field of level fine at the
        coincident nodes (Restrict_ss, wavelet.f90:823-837)"""
        _check(_lib.wtg_mg_restrict_ss(self._h, fine, fine_field, coarse_field), self._top, "wtg_mg_restrict_ss")

    def set_precision(self, precision: int):
        """PRECISION_FP64 or PRECISION_FP32_BULK (FP32 smoothing and transfers
        of each V-cycle's correction, FP64 residuals and solution update)"""
        _check(_lib.wtg_mg_set_precision(self._h, precision), self._top, "wtg_mg_set_precision")

    def features(self, field: int):
        """network input rows [npatch, 192], patch RMS sigma [npatch] and patch
        node cells [npatch, 64] of the finest level's field"""
        n = self.npatches
        f = np.empty((n, MLP_IN), dtype=np.float32)
        sg = np.empty(n)
        cells = np.empty((n, 64), dtype=np.int64)
        _check(_lib.wtg_mg_features(self._h, field, _ptr(f), _ptr(sg), _ptr(cells)), self._top, "wtg_mg_features")
        return f, sg, cells

    def forward(self, fine: int, s_field: int, wc_field: int, coarse_field: int):
        _check(_lib.wtg_mg_forward(self._h, fine, s_field, wc_field, coarse_field), self._top, "wtg_mg_forward")

    def inverse(self, fine: int, coarse_field: int, wc_field: int, fine_field: int):
        _check(_lib.wtg_mg_inverse(self._h, fine, coarse_field, wc_field, fine_field), self._top, "wtg_mg_inverse")


MLP_IN, MLP_HIDDEN, MLP_OUT, MLP_NPARAMS = 192, 128, 64, 49472


class MultiLevel:
    """The multi-level explicit step on a full hierarchy of uniform levels
    (wtg_ml_*): contexts coarse to fine with the same layers and slots, each
    holding its level's state.  While it exists the contexts run on the
    finest one's stream; close it before the contexts."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        n = len(self.contexts)
        grids = (_P * n)(*[c.grid.handle.value for c in self.contexts])
        ctxs = (_P * n)(*[c._h.value for c in self.contexts])
        h = _P()
        _check(load_library().wtg_ml_create(grids, ctxs, n, C.byref(h)), None, "wtg_ml_create")
        self._h = h

    @property
    def _top(self):
        return self.contexts[-1]._h

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.wtg_ml_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def upload(self, slot: int, mass, temp, velo):
        """per-level lists of [L, nc] / [L, 3, nc] arrays"""
        for c, m, t, u in zip(self.contexts, mass, temp, velo):
            c.upload_state(slot, m, t, u)

    def download(self, slot: int):
        out = [c.download_state(slot) for c in self.contexts]
        return [o[0] for o in out], [o[1] for o in out], [o[2] for o in out]

    def forward(self, slot: int):
        _check(_lib.wtg_ml_forward(self._h, slot), self._top, "wtg_ml_forward")

    def inverse(self, slot: int):
        _check(_lib.wtg_ml_inverse(self._h, slot), self._top, "wtg_ml_inverse")

    def wavelets(self, level: int):
        c = self.contexts[level]
        (L, nc), (_, _, _) = c._shape()
        m, t, u = np.empty((L, nc)), np.empty((L, nc)), np.empty((L, 3, nc))
        _check(_lib.wtg_ml_wavelets(self._h, level, 0, _ptr(m), _ptr(t), _ptr(u)), self._top, "wtg_ml_wavelets")
        return m, t, u

    def set_wavelets(self, level: int, mass, temp, velo):
        c = self.contexts[level]
        (L, nc), _ = c._shape()
        m = np.ascontiguousarray(mass, dtype=np.float64).reshape(L, nc)
        t = np.ascontiguousarray(temp, dtype=np.float64).reshape(L, nc)
        u = np.ascontiguousarray(velo, dtype=np.float64).reshape(L, 3, nc)
        _check(_lib.wtg_ml_wavelets(self._h, level, 1, _ptr(m), _ptr(t), _ptr(u)), self._top, "wtg_ml_wavelets")

    def trend(self, q_slot: int, dq_slot: int):
        _check(_lib.wtg_ml_trend(self._h, q_slot, dq_slot), self._top, "wtg_ml_trend")

    def wt_after_step(self, slot: int):
        _check(_lib.wtg_ml_wt_after_step(self._h, slot), self._top, "wtg_ml_wt_after_step")

    def rk4(self, slot: int, dt: float, nsteps: int = 1):
        _check(_lib.wtg_ml_rk4(self._h, slot, dt, nsteps), self._top, "wtg_ml_rk4")

    def rk4_split(self, hierarchy: "Hierarchy", slot: int, nsteps: int = 1, params: MgParams | None = None,
                  nstages: int = 4) -> int:
        """RK4_split steps on every level with the FMG free surface of
        `hierarchy` (a Hierarchy of the same contexts; dt, theta1, theta2
        from each context's fs_set_params, the free surface in FS_ETA of
        every level); returns the FMG V-cycles summed over levels and steps"""
        p = params if params is not None else MgParams.defaults()
        it = C.c_int()
        if nstages == 4:
            _check(_lib.wtg_ml_rk4_split(self._h, hierarchy._h, slot, nsteps, C.byref(p), C.byref(it)), self._top,
                   "wtg_ml_rk4_split")
        else:
            _check(_lib.wtg_ml_rk_split(self._h, hierarchy._h, slot, nstages, nsteps, C.byref(p), C.byref(it)),
                   self._top, "wtg_ml_rk_split")
        return it.value


class MLP:
    """The learned correction's network on the current HIP device
    (wtg_mlp_*): BF16 weights and activations, FP32 accumulation on the
    matrix cores, Adam training on device-resident pairs."""

    def __init__(self, max_batch: int = 4096, seed: int = 0):
        h = _P()
        _check(load_library().wtg_mlp_create(max_batch, seed, C.byref(h)), None, "wtg_mlp_create")
        self._h = h
        self.max_batch = max_batch

    def _check(self, rc, what):
        if rc != 0:
            buf = C.create_string_buffer(1024)
            _lib.wtg_mlp_last_error(self._h, buf, len(buf))
            raise WtgError(f"{what} failed ({rc}): {buf.value.decode(errors='replace')}")

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.wtg_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def params(self) -> np.ndarray:
        out = np.empty(MLP_NPARAMS, dtype=np.float32)
        self._check(_lib.wtg_mlp_get_params(self._h, _ptr(out), out.size), "wtg_mlp_get_params")
        return out

    @params.setter
    def params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        self._check(_lib.wtg_mlp_set_params(self._h, _ptr(a), a.size), "wtg_mlp_set_params")

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty((x.shape[0], MLP_OUT), dtype=np.float32)
        self._check(_lib.wtg_mlp_forward(self._h, _ptr(x), x.shape[0], _ptr(y)), "wtg_mlp_forward")
        return y

    def forward_timed(self, batch: int, nrep: int = 20) -> float:
        """mean ms of a forward pass of the first `batch` rows on the device"""
        ms = C.c_double()
        self._check(_lib.wtg_mlp_forward_timed(self._h, batch, nrep, C.byref(ms)), "wtg_mlp_forward_timed")
        return ms.value

    def gradients(self, x, y_true):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(y_true, dtype=np.float32)
        g = np.empty(MLP_NPARAMS, dtype=np.float32)
        loss = C.c_float()
        self._check(_lib.wtg_mlp_gradients(self._h, _ptr(x), _ptr(t), x.shape[0], C.byref(loss), _ptr(g)),
                    "wtg_mlp_gradients")
        return loss.value, g

    def set_dataset(self, x, y_true):
        self._x = np.ascontiguousarray(x, dtype=np.float32)
        self._t = np.ascontiguousarray(y_true, dtype=np.float32)
        self._check(_lib.wtg_mlp_set_dataset(self._h, _ptr(self._x), _ptr(self._t), self._x.shape[0]),
                    "wtg_mlp_set_dataset")

    def train(self, nsteps: int, batch: int, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        losses = np.empty(nsteps, dtype=np.float32)
        self._check(_lib.wtg_mlp_train(self._h, nsteps, batch, lr, seed, _ptr(losses)), "wtg_mlp_train")
        return losses

