"""ctypes binding of libfsg.so (include/fsg.h).

``Domain`` mirrors the reference's FluidDomain + FluidSolver pair
(include/FluidDomain.h:37-76, include/FluidSolver.h:45-54) with the state held on
the GPU.  There is no CPU fallback: if libfsg.so is missing or cannot be loaded,
``load()`` raises.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import sys
import threading
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfsg.so")
_LIB = None
_LOCK = threading.Lock()
# Handles still open.  They are destroyed by an atexit hook (registered by load()),
# which runs while the HIP runtime is intact; a handle's __del__ that runs later,
# during interpreter finalisation, makes no device call.
_LIVE: "weakref.WeakSet" = weakref.WeakSet()


_finalizing = sys.is_finalizing


def _close_live():
    for d in list(_LIVE):
        try:
            d.close()
        except Exception:
            pass

OK = 0
LIQUID, AIR, SOLID = 0, 1, 2
SEMI_LAGRANGIAN, PIC, FLIP, PICFLIP, GRID_ADVECT = 0, 1, 2, 3, 4
RK3, EULER, RK2, RK4 = 0, 1, 2, 3
SOLVER_PLAIN, SOLVER_PERSISTENT, SOLVER_REFINED, SOLVER_MG = 0, 1, 2, 3
VEL_FRONT, VEL_BACK, VEL_PREV, VEL_DIFF = 0, 1, 2, 3
PH = dict(CLASSIFY=1, P2G=2, SAVE_PREV=3, GRAVITY=4, DIRICHLET=5, EXTEND=6,
          PRESSURE=7, DIFF=8, G2P_PIC=9, G2P_FLIP=10, G2P_PICFLIP=11, ADVECT=12,
          ADVECT_REJECT=13, SL_SCATTER=14, SL_GATHER=15, ADVECT_GRID=16,
          CLASSIFY_LEVELSET=17, EXTEND_AVERAGE=18)

# every entry point declared in include/fsg.h
EXPORTS = [
    "fsg_version", "fsg_params_default", "fsg_create", "fsg_destroy", "fsg_last_error",
    "fsg_set_params", "fsg_get_params", "fsg_set_pic_ratio", "fsg_get_geometry",
    "fsg_synchronize", "fsg_set_particles", "fsg_add_particles", "fsg_num_particles",
    "fsg_get_particles", "fsg_set_velocity", "fsg_get_velocity", "fsg_swap_velocity",
    "fsg_set_labels", "fsg_get_labels", "fsg_set_solid_mask", "fsg_set_level_set",
    "fsg_step", "fsg_run_phase", "fsg_validate_pool", "fsg_get_pressure", "fsg_get_rhs",
    "fsg_get_fluid_indices", "fsg_get_valid_masks", "fsg_get_ext_info",
    "fsg_get_reject_flags", "fsg_get_stats", "fsg_reset_stats", "fsg_set_timing",
    "fsg_get_timing", "fsg_reset_timing", "fsg_set_timing_prefix", "fsg_emit_lattice",
    "fsg_advect_level_set", "fsg_get_level_set", "fsg_batch_create", "fsg_batch_num_scenes",
    "fsg_batch_scene_rect", "fsg_batch_set_particles", "fsg_batch_set_solid_mask",
    "fsg_batch_num_particles", "fsg_batch_get_particles", "fsg_batch_get_velocity",
    "fsg_batch_get_labels", "fsg_checkpoint_save", "fsg_checkpoint_load",
    "fsg_checkpoint_clear", "fsg_adjoint_set", "fsg_adjoint_get", "fsg_adjoint_step",
    "fsg_adjoint_last_iterations", "fsg_particle_density", "fsg_adjoint_seed_density",
    "fsg_adjoint_set_front", "fsg_adjoint_get_front",
    "fsg_set_kernel_clock", "fsg_get_kernel_clock", "fsg_get_step_decisions",
    "fsg_p2g_noalloc", "fsg_autotune",
]


class FsgError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [
        ("gravity_x", C.c_float), ("gravity_y", C.c_float), ("cg_tol", C.c_double),
        ("cg_max_iter", C.c_int32), ("ext_layers", C.c_int32), ("integrator", C.c_int32),
        ("solver", C.c_int32), ("project_mean", C.c_int32), ("check_every", C.c_int32),
        ("refine_inner_tol", C.c_double), ("refine_max_outer", C.c_int32),
        ("record_ext_info", C.c_int32), ("pcg_blocks", C.c_int32), ("pcg_kpt", C.c_int32),
        ("warm_start", C.c_int32), ("particle_pressure_k", C.c_float),
        ("closed_walls", C.c_int32), ("graph", C.c_int32), ("deterministic", C.c_int32),
        ("record_decisions", C.c_int32), ("mg_smoother", C.c_int32),
        ("mg_coarse_sweeps", C.c_int32), ("graph_iters", C.c_int32),
        ("mg_fuse_update", C.c_int32), ("splat_order", C.c_int32),
        ("mg_tile_threads", C.c_int32), ("mg_level_sweeps", C.c_int32),
        ("mg_level0_sweeps", C.c_int32), ("mg_spread_tiles", C.c_int32),
    ]


class Tune(C.Structure):
    _fields_ = [
        ("mg_smoother", C.c_int32), ("mg_coarse_sweeps", C.c_int32),
        ("mg_fuse_update", C.c_int32), ("graph_iters", C.c_int32),
        ("splat_order", C.c_int32), ("candidates", C.c_int32), ("cg_iterations", C.c_int32),
        ("mg_tile_threads", C.c_int32),
        ("ms_per_step", C.c_double), ("ms_per_step_before", C.c_double),
        ("mg_level_sweeps", C.c_int32), ("mg_level0_sweeps", C.c_int32),
        ("reserved", C.c_int64 * 2),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("n_fluid", C.c_int32), ("cg_iterations", C.c_int32), ("cg_error", C.c_double),
        ("refine_sweeps", C.c_int32), ("mg_live_tiles", C.c_int32),
        ("total_unknown_iterations", C.c_int64), ("total_solves", C.c_int64),
        ("reserved", C.c_int64 * 4),
    ]


def load(path: str | None = None):
    """Load libfsg.so (built by __graft_entry__.build()); raise if unavailable."""
    global _LIB
    with _LOCK:
        if _LIB is not None:
            return _LIB
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise FsgError(f"libfsg.so not found at {p}: build it first "
                           "(python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(p)
        H = C.c_void_p
        P = C.POINTER
        sig = {
            "fsg_version": ([], C.c_char_p),
            "fsg_params_default": ([P(Params)], None),
            "fsg_create": ([P(H), C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_float,
                            C.c_float], C.c_int),
            "fsg_destroy": ([H], C.c_int),
            "fsg_last_error": ([H], C.c_char_p),
            "fsg_set_params": ([H, P(Params)], C.c_int),
            "fsg_get_params": ([H, P(Params)], C.c_int),
            "fsg_set_pic_ratio": ([H, C.c_float], C.c_int),
            "fsg_get_geometry": ([H, P(C.c_int32), P(C.c_int32), P(C.c_float), P(C.c_float),
                                  P(C.c_float), P(C.c_float)], C.c_int),
            "fsg_synchronize": ([H], C.c_int),
            "fsg_set_particles": ([H, C.c_void_p, C.c_int64], C.c_int),
            "fsg_add_particles": ([H, C.c_void_p, C.c_int64], C.c_int),
            "fsg_num_particles": ([H, P(C.c_int64)], C.c_int),
            "fsg_get_particles": ([H, C.c_void_p, C.c_int64], C.c_int),
            "fsg_set_velocity": ([H, C.c_int32, C.c_void_p, C.c_void_p], C.c_int),
            "fsg_get_velocity": ([H, C.c_int32, C.c_void_p, C.c_void_p], C.c_int),
            "fsg_swap_velocity": ([H], C.c_int),
            "fsg_set_labels": ([H, C.c_void_p], C.c_int),
            "fsg_get_labels": ([H, C.c_void_p], C.c_int),
            "fsg_set_solid_mask": ([H, C.c_void_p], C.c_int),
            "fsg_set_level_set": ([H, C.c_void_p], C.c_int),
            "fsg_step": ([H, C.c_int32, C.c_float], C.c_int),
            "fsg_run_phase": ([H, C.c_int32, C.c_float], C.c_int),
            "fsg_validate_pool": ([H, C.c_int32, C.c_int32, C.c_float, C.c_float], C.c_int),
            "fsg_get_pressure": ([H, C.c_void_p], C.c_int),
            "fsg_get_rhs": ([H, C.c_void_p], C.c_int),
            "fsg_get_fluid_indices": ([H, C.c_void_p], C.c_int),
            "fsg_get_valid_masks": ([H, C.c_void_p, C.c_void_p], C.c_int),
            "fsg_get_ext_info": ([H, C.c_void_p, C.c_void_p], C.c_int),
            "fsg_get_reject_flags": ([H, C.c_void_p], C.c_int),
            "fsg_get_stats": ([H, P(Stats)], C.c_int),
            "fsg_reset_stats": ([H], C.c_int),
            "fsg_set_timing": ([H, C.c_int32], C.c_int),
            "fsg_get_timing": ([H, C.c_int32, P(C.c_char_p), P(C.c_double), P(C.c_int64)],
                               C.c_int),
            "fsg_reset_timing": ([H], C.c_int),
            "fsg_set_timing_prefix": ([H, C.c_char_p], C.c_int),
            "fsg_emit_lattice": ([H, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_float,
                                  C.c_float], C.c_int),
            "fsg_advect_level_set": ([H, C.c_float], C.c_int),
            "fsg_get_level_set": ([H, C.c_void_p], C.c_int),
            "fsg_batch_create": ([P(H), C.c_int32, C.c_void_p, C.c_void_p, C.c_float, C.c_float,
                                  C.c_float], C.c_int),
            "fsg_batch_num_scenes": ([H, P(C.c_int32)], C.c_int),
            "fsg_batch_scene_rect": ([H, C.c_int32, C.c_void_p], C.c_int),
            "fsg_batch_set_particles": ([H, C.c_void_p, C.c_void_p], C.c_int),
            "fsg_batch_set_solid_mask": ([H, C.c_int32, C.c_void_p], C.c_int),
            "fsg_batch_num_particles": ([H, C.c_int32, P(C.c_int64)], C.c_int),
            "fsg_batch_get_particles": ([H, C.c_int32, C.c_void_p, C.c_int64], C.c_int),
            "fsg_batch_get_velocity": ([H, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p], C.c_int),
            "fsg_batch_get_labels": ([H, C.c_int32, C.c_void_p], C.c_int),
            "fsg_checkpoint_save": ([H, C.c_int32], C.c_int),
            "fsg_checkpoint_load": ([H, C.c_int32], C.c_int),
            "fsg_checkpoint_clear": ([H], C.c_int),
            "fsg_adjoint_set": ([H, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
            "fsg_adjoint_get": ([H, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
            "fsg_adjoint_step": ([H, C.c_int32, C.c_float], C.c_int),
            "fsg_adjoint_set_front": ([H, C.c_void_p, C.c_void_p], C.c_int),
            "fsg_adjoint_get_front": ([H, C.c_void_p, C.c_void_p], C.c_int),
            "fsg_adjoint_last_iterations": ([H, P(C.c_int32)], C.c_int),
            "fsg_particle_density": ([H, C.c_void_p], C.c_int),
            "fsg_adjoint_seed_density": ([H, C.c_void_p], C.c_int),
            "fsg_set_kernel_clock": ([H, C.c_int32], C.c_int),
            "fsg_get_kernel_clock": ([H, P(C.c_double), P(C.c_int64)], C.c_int),
            "fsg_get_step_decisions": ([H, C.c_void_p], C.c_int),
            "fsg_autotune": ([H, C.c_int32, C.c_float, C.c_int32, P(Tune)], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _LIB = L
        atexit.register(_close_live)
        return L


def default_params() -> Params:
    p = Params()
    load().fsg_params_default(C.byref(p))
    return p


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != shape:
        a = a.reshape(shape)
    return a


class Domain:
    """FluidDomain + FluidSolver on the GPU (one HIP stream per domain)."""

    def __init__(self, nx: int, ny: int, length_x: float, length_y: float, density: float,
                 pic_ratio: float):
        self._L = load()
        self._h = C.c_void_p()
        rc = self._L.fsg_create(C.byref(self._h), nx, ny, length_x, length_y, density, pic_ratio)
        if rc != OK:
            raise FsgError(f"fsg_create failed ({rc})")
        self.nx, self.ny = nx, ny
        _LIVE.add(self)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.fsg_destroy(self._h)
            self._h = C.c_void_p()
        _LIVE.discard(self)

    def __del__(self):
        # at interpreter exit the atexit hook has already closed every handle; one
        # collected later (finalisation) must not call into HIP, whose runtime may be
        # torn down by then (nor can it rely on module globals such as sys, which
        # finalisation may already have cleared: any failure here is ignored)
        try:
            if _finalizing():
                return
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != OK:
            msg = self._L.fsg_last_error(self._h)
            raise FsgError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    # -- params ---------------------------------------------------------------
    @property
    def params(self) -> Params:
        p = Params()
        self._chk(self._L.fsg_get_params(self._h, C.byref(p)), "fsg_get_params")
        return p

    def set_params(self, **kw):
        p = self.params
        for k, v in kw.items():
            setattr(p, k, v)
        self._chk(self._L.fsg_set_params(self._h, C.byref(p)), "fsg_set_params")

    def geometry(self):
        nx, ny = C.c_int32(), C.c_int32()
        dx, dy, rho, r = C.c_float(), C.c_float(), C.c_float(), C.c_float()
        self._chk(self._L.fsg_get_geometry(self._h, C.byref(nx), C.byref(ny), C.byref(dx),
                                           C.byref(dy), C.byref(rho), C.byref(r)),
                  "fsg_get_geometry")
        return dict(nx=nx.value, ny=ny.value, dx=dx.value, dy=dy.value, density=rho.value,
                    pic_ratio=r.value)

    def set_pic_ratio(self, r):
        self._chk(self._L.fsg_set_pic_ratio(self._h, r), "fsg_set_pic_ratio")

    # -- particles ------------------------------------------------------------
    def set_particles(self, parts):
        a = _f32(parts).reshape(-1, 4)
        self._chk(self._L.fsg_set_particles(self._h, a.ctypes.data, len(a)), "fsg_set_particles")

    def add_particles(self, parts):
        a = _f32(parts).reshape(-1, 4)
        self._chk(self._L.fsg_add_particles(self._h, a.ctypes.data, len(a)), "fsg_add_particles")

    def num_particles(self) -> int:
        n = C.c_int64()
        self._chk(self._L.fsg_num_particles(self._h, C.byref(n)), "fsg_num_particles")
        return n.value

    def particles(self) -> np.ndarray:
        n = self.num_particles()
        out = np.zeros((n, 4), dtype=np.float32)
        self._chk(self._L.fsg_get_particles(self._h, out.ctypes.data, n), "fsg_get_particles")
        return out

    # -- grids ----------------------------------------------------------------
    def set_velocity(self, u, v, which=VEL_FRONT):
        sh = (self.ny, self.nx)
        uu = _f32(u, sh) if u is not None else None
        vv = _f32(v, sh) if v is not None else None
        self._chk(self._L.fsg_set_velocity(self._h, which, uu.ctypes.data if uu is not None else None,
                                           vv.ctypes.data if vv is not None else None),
                  "fsg_set_velocity")

    def velocity(self, which=VEL_FRONT):
        u = np.zeros((self.ny, self.nx), np.float32)
        v = np.zeros((self.ny, self.nx), np.float32)
        self._chk(self._L.fsg_get_velocity(self._h, which, u.ctypes.data, v.ctypes.data),
                  "fsg_get_velocity")
        return u, v

    def swap_velocity(self):
        self._chk(self._L.fsg_swap_velocity(self._h), "fsg_swap_velocity")

    def set_labels(self, lab):
        a = np.ascontiguousarray(lab, dtype=np.uint8).reshape(self.ny, self.nx)
        self._chk(self._L.fsg_set_labels(self._h, a.ctypes.data), "fsg_set_labels")

    def labels(self) -> np.ndarray:
        a = np.zeros((self.ny, self.nx), np.uint8)
        self._chk(self._L.fsg_get_labels(self._h, a.ctypes.data), "fsg_get_labels")
        return a

    def set_solid_mask(self, mask):
        if mask is None:
            self._chk(self._L.fsg_set_solid_mask(self._h, None), "fsg_set_solid_mask")
            return
        a = np.ascontiguousarray(mask, dtype=np.uint8).reshape(self.ny, self.nx)
        self._chk(self._L.fsg_set_solid_mask(self._h, a.ctypes.data), "fsg_set_solid_mask")

    def set_level_set(self, phi):
        a = _f32(phi, (self.ny, self.nx))
        self._chk(self._L.fsg_set_level_set(self._h, a.ctypes.data), "fsg_set_level_set")

    # -- stepping -------------------------------------------------------------
    def step(self, scheme: int, dt: float):
        self._chk(self._L.fsg_step(self._h, scheme, dt), "fsg_step")

    def run_phase(self, phase, dt: float = 0.0):
        ph = PH[phase] if isinstance(phase, str) else phase
        self._chk(self._L.fsg_run_phase(self._h, ph, dt), f"fsg_run_phase({phase})")

    def validate_pool(self, nx, ny, dx, dy):
        return self._L.fsg_validate_pool(self._h, nx, ny, dx, dy)

    def synchronize(self):
        self._chk(self._L.fsg_synchronize(self._h), "fsg_synchronize")

    # -- inspection -----------------------------------------------------------
    def _grid(self, fn, dtype):
        a = np.zeros((self.ny, self.nx), dtype)
        self._chk(getattr(self._L, fn)(self._h, a.ctypes.data), fn)
        return a

    def pressure(self):
        return self._grid("fsg_get_pressure", np.float32)

    def rhs(self):
        return self._grid("fsg_get_rhs", np.float32)

    def fluid_indices(self):
        return self._grid("fsg_get_fluid_indices", np.int32)

    def valid_masks(self):
        mu = np.zeros((self.ny, self.nx), np.uint8)
        mv = np.zeros((self.ny, self.nx), np.uint8)
        self._chk(self._L.fsg_get_valid_masks(self._h, mu.ctypes.data, mv.ctypes.data),
                  "fsg_get_valid_masks")
        return mu, mv

    def ext_info(self):
        iu = np.zeros((self.ny, self.nx), np.uint8)
        iv = np.zeros((self.ny, self.nx), np.uint8)
        self._chk(self._L.fsg_get_ext_info(self._h, iu.ctypes.data, iv.ctypes.data),
                  "fsg_get_ext_info")
        return iu, iv

    def reject_flags(self):
        n = self.num_particles()
        a = np.zeros(n, np.uint8)
        self._chk(self._L.fsg_get_reject_flags(self._h, a.ctypes.data), "fsg_get_reject_flags")
        return a

    def autotune(self, scheme: int, dt: float, steps: int = 3) -> dict:
        """fsg_autotune: time every candidate multigrid configuration over ``steps``
        steps from the current state, keep the fastest, restore the state."""
        t = Tune()
        self._chk(self._L.fsg_autotune(self._h, scheme, dt, steps, C.byref(t)), "fsg_autotune")
        return {k: getattr(t, k) for k, _ in Tune._fields_ if not k.startswith("reserved")}

    def step_decisions(self):
        """The last step's discrete decisions per cell (fsg_params.record_decisions):
        bit 0 / 1 u / v normalised by P2G, bit 2 / 3 u / v zeroed by the first
        Dirichlet pass, bit 4 / 5 by the second."""
        return self._grid("fsg_get_step_decisions", np.uint8)

    def stats(self) -> Stats:
        s = Stats()
        self._chk(self._L.fsg_get_stats(self._h, C.byref(s)), "fsg_get_stats")
        return s

    def reset_stats(self):
        self._chk(self._L.fsg_reset_stats(self._h), "fsg_reset_stats")

    def emit_lattice(self, xs, ys, vx=0.0, vy=0.0):
        """FluidSource::update's lattice (xs[i], ys[j]) appended on the device."""
        xs = _f32(np.asarray(xs, np.float32).ravel())
        ys = _f32(np.asarray(ys, np.float32).ravel())
        self._chk(self._L.fsg_emit_lattice(self._h, xs.ctypes.data, len(xs), ys.ctypes.data,
                                           len(ys), float(vx), float(vy)), "fsg_emit_lattice")

    def advect_level_set(self, dt):
        self._chk(self._L.fsg_advect_level_set(self._h, float(dt)), "fsg_advect_level_set")

    def level_set(self):
        return self._grid("fsg_get_level_set", np.float32)

    def set_timing(self, mode):
        """0 off, 1 (or True) every kernel, 2 only the pressure-solver kernels."""
        self._chk(self._L.fsg_set_timing(self._h, int(mode)), "fsg_set_timing")

    def set_timing_prefix(self, prefix: str):
        """Timing mode 2 times only kernels whose name starts with `prefix`."""
        self._chk(self._L.fsg_set_timing_prefix(self._h, prefix.encode()),
                  "fsg_set_timing_prefix")

    def timing(self) -> dict:
        n = self._L.fsg_get_timing(self._h, 0, None, None, None)
        if n < 0:
            self._chk(n, "fsg_get_timing")
        names = (C.c_char_p * max(n, 1))()
        ms = (C.c_double * max(n, 1))()
        cnt = (C.c_int64 * max(n, 1))()
        self._L.fsg_get_timing(self._h, n, names, ms, cnt)
        return {names[q].decode(): (ms[q], cnt[q]) for q in range(n)}

    def reset_timing(self):
        self._chk(self._L.fsg_reset_timing(self._h), "fsg_reset_timing")

    # -- checkpoint ---------------------------------------------------------------
    def save_state(self) -> dict:
        """The state carried from one step to the next: particles, front and back
        velocity buffers, prev and diff grids, labels (the buffers the reference
        keeps alive across steps, include/MacGrid.h:162-176)."""
        st = {"particles": self.particles(), "labels": self.labels()}
        for name, which in (("front", VEL_FRONT), ("back", VEL_BACK), ("prev", VEL_PREV),
                            ("diff", VEL_DIFF)):
            st[name] = self.velocity(which)
        return st

    def load_state(self, st: dict):
        self.set_particles(st["particles"])
        for name, which in (("front", VEL_FRONT), ("back", VEL_BACK), ("prev", VEL_PREV),
                            ("diff", VEL_DIFF)):
            u, v = st[name]
            self.set_velocity(u, v, which)
        self.set_labels(st["labels"])

    # -- construction from a synthetic scene -----------------------------------
    @classmethod
    def from_scene(cls, scene, **params):
        d = cls(scene.nx, scene.ny, scene.length_x, scene.length_y, scene.density,
                scene.pic_ratio)
        if params:
            d.set_params(**params)
        if len(scene.particles):
            d.set_particles(scene.particles)
        if scene.u is not None:
            d.set_velocity(scene.u, scene.v)
        if scene.solid_mask is not None:
            d.set_solid_mask(scene.solid_mask)
        if scene.phi is not None:
            d.set_level_set(scene.phi)
        return d


    def set_kernel_clock(self, enable: bool, kernel: str = "coarse"):
        """Device-clock timing of k_mg_coarse (``kernel="coarse"``) or of level 0's
        down kernel (``"down0"``: earliest block start to latest block end)."""
        code = {"coarse": 1, "down0": 2}[kernel] if enable else 0
        self._chk(self._L.fsg_set_kernel_clock(self._h, code), "fsg_set_kernel_clock")

    def kernel_clock(self):
        """(microseconds, launches) of the clocked kernel by the device clock."""
        us, n = C.c_double(), C.c_int64()
        self._chk(self._L.fsg_get_kernel_clock(self._h, C.byref(us), C.byref(n)),
                  "fsg_get_kernel_clock")
        return us.value, n.value

    # -- checkpoints and reverse mode ---------------------------------------------
    def checkpoint_save(self, slot: int):
        self._chk(self._L.fsg_checkpoint_save(self._h, slot), "fsg_checkpoint_save")

    def checkpoint_load(self, slot: int):
        self._chk(self._L.fsg_checkpoint_load(self._h, slot), "fsg_checkpoint_load")

    def checkpoint_clear(self):
        self._chk(self._L.fsg_checkpoint_clear(self._h), "fsg_checkpoint_clear")

    def adjoint_set(self, pbar=None, bbar_u=None, bbar_v=None):
        """Adjoint state: pbar (n, 4) per particle, the back buffer's (ny, nx) u and v
        (None = zeros)."""
        args = []
        keep = []
        for a, shape in ((pbar, (self.num_particles(), 4)), (bbar_u, (self.ny, self.nx)),
                         (bbar_v, (self.ny, self.nx))):
            if a is None:
                args.append(None)
            else:
                arr = np.ascontiguousarray(a, np.float32).reshape(shape)
                keep.append(arr)
                args.append(arr.ctypes.data)
        self._chk(self._L.fsg_adjoint_set(self._h, *args), "fsg_adjoint_set")

    def adjoint_get(self):
        n = self.num_particles()
        pbar = np.zeros((n, 4), np.float32)
        bu = np.zeros((self.ny, self.nx), np.float32)
        bv = np.zeros((self.ny, self.nx), np.float32)
        self._chk(self._L.fsg_adjoint_get(self._h, pbar.ctypes.data if n else None,
                                          bu.ctypes.data, bv.ctypes.data), "fsg_adjoint_get")
        return pbar, bu, bv

    def adjoint_set_front(self, fbar_u=None, fbar_v=None):
        """The front buffer's adjoint, (ny, nx) u and v (None = zeros)."""
        keep = [None if a is None else
                np.ascontiguousarray(a, np.float32).reshape(self.ny, self.nx)
                for a in (fbar_u, fbar_v)]
        self._chk(self._L.fsg_adjoint_set_front(
            self._h, *(None if a is None else a.ctypes.data for a in keep)),
            "fsg_adjoint_set_front")

    def adjoint_get_front(self):
        fu = np.zeros((self.ny, self.nx), np.float32)
        fv = np.zeros((self.ny, self.nx), np.float32)
        self._chk(self._L.fsg_adjoint_get_front(self._h, fu.ctypes.data, fv.ctypes.data),
                  "fsg_adjoint_get_front")
        return fu, fv

    def adjoint_step(self, scheme: int, dt: float):
        self._chk(self._L.fsg_adjoint_step(self._h, scheme, dt), "fsg_adjoint_step")

    def adjoint_last_iterations(self) -> int:
        it = C.c_int32()
        self._chk(self._L.fsg_adjoint_last_iterations(self._h, C.byref(it)),
                  "fsg_adjoint_last_iterations")
        return it.value

    def density(self) -> np.ndarray:
        rho = np.zeros((self.ny, self.nx), np.float32)
        self._chk(self._L.fsg_particle_density(self._h, rho.ctypes.data), "fsg_particle_density")
        return rho

    def adjoint_seed_density(self, rho_bar):
        rb = np.ascontiguousarray(rho_bar, np.float32).reshape(self.ny, self.nx)
        self._chk(self._L.fsg_adjoint_seed_density(self._h, rb.ctypes.data),
                  "fsg_adjoint_seed_density")


class Batch(Domain):
    """Many scenes packed into one grid and stepped together (fsg_batch_create):
    one kernel launch per phase for all scenes; particles in scene-local
    coordinates; the pressure solve bounds every scene's own relative residual."""

    def __init__(self, sizes, delta: float, density: float, pic_ratio: float):
        self._L = load()
        self._h = C.c_void_p()
        nx = np.ascontiguousarray([s[0] for s in sizes], np.int32)
        ny = np.ascontiguousarray([s[1] for s in sizes], np.int32)
        rc = self._L.fsg_batch_create(C.byref(self._h), len(sizes), nx.ctypes.data,
                                      ny.ctypes.data, delta, density, pic_ratio)
        if rc != OK:
            raise FsgError(f"fsg_batch_create failed ({rc})")
        _LIVE.add(self)
        self.sizes = [(int(a), int(b)) for a, b in zip(nx, ny)]
        self.n_scenes = len(self.sizes)
        geo = self.geometry()
        self.nx, self.ny = geo["nx"], geo["ny"]  # the packed grid

    @classmethod
    def from_scenes(cls, scenes, **params):
        """One batch holding ``scenes`` (all with the same cell width length / n,
        density and pic_ratio); scene s keeps index s."""
        if not scenes:
            raise FsgError("a batch needs at least one scene")
        sc0 = scenes[0]
        delta = sc0.length_x / sc0.nx
        for sc in scenes:
            if (sc.length_x / sc.nx != delta or sc.length_y / sc.ny != delta
                    or sc.density != sc0.density or sc.pic_ratio != sc0.pic_ratio):
                raise FsgError("batched scenes must share spacing, density and pic_ratio")
        b = cls([(sc.nx, sc.ny) for sc in scenes], delta, sc0.density, sc0.pic_ratio)
        if params:
            b.set_params(**params)
        b.set_scene_particles([sc.particles for sc in scenes])
        for s, sc in enumerate(scenes):
            if sc.solid_mask is not None:
                b.set_scene_solid_mask(s, sc.solid_mask)
        return b

    def scene_rect(self, s):
        r = (C.c_int32 * 4)()
        self._chk(self._L.fsg_batch_scene_rect(self._h, s, r), "fsg_batch_scene_rect")
        return tuple(int(v) for v in r)

    def set_scene_particles(self, parts_list):
        counts = np.ascontiguousarray([len(p) for p in parts_list], np.int64)
        allp = _f32(np.concatenate([np.asarray(p, np.float32).reshape(-1, 4) for p in parts_list])
                    if counts.sum() else np.zeros((0, 4), np.float32))
        self._chk(self._L.fsg_batch_set_particles(self._h, counts.ctypes.data,
                                                  allp.ctypes.data if len(allp) else None),
                  "fsg_batch_set_particles")

    def set_scene_solid_mask(self, s, mask):
        m = np.ascontiguousarray(mask, np.uint8)
        self._chk(self._L.fsg_batch_set_solid_mask(self._h, s, m.ctypes.data),
                  "fsg_batch_set_solid_mask")

    def scene_particles(self, s):
        n = C.c_int64()
        self._chk(self._L.fsg_batch_num_particles(self._h, s, C.byref(n)), "num_particles")
        out = np.zeros((n.value, 4), np.float32)
        if n.value:
            self._chk(self._L.fsg_batch_get_particles(self._h, s, out.ctypes.data, n.value),
                      "fsg_batch_get_particles")
        return out

    def scene_velocity(self, s, which=VEL_FRONT):
        nx, ny = self.sizes[s]
        u = np.zeros((ny, nx), np.float32)
        v = np.zeros((ny, nx), np.float32)
        self._chk(self._L.fsg_batch_get_velocity(self._h, s, which, u.ctypes.data, v.ctypes.data),
                  "fsg_batch_get_velocity")
        return u, v

    def scene_labels(self, s):
        nx, ny = self.sizes[s]
        lab = np.zeros((ny, nx), np.uint8)
        self._chk(self._L.fsg_batch_get_labels(self._h, s, lab.ctypes.data),
                  "fsg_batch_get_labels")
        return lab

