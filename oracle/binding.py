"""ctypes binding of liboracle.so (test infrastructure only; see oracle.h)."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_LOCK = threading.Lock()

LIQUID, AIR, SOLID = 0, 1, 2
RK3, EULER, RK2, RK4 = 0, 1, 2, 3
SCHEME_SEMI_LAGRANGIAN, SCHEME_PIC, SCHEME_FLIP, SCHEME_PICFLIP, SCHEME_GRID = range(5)
PH = dict(CLASSIFY=1, P2G=2, SAVE_PREV=3, GRAVITY=4, DIRICHLET=5, EXTEND=6,
          PRESSURE=7, DIFF=8, G2P_PIC=9, G2P_FLIP=10, G2P_PICFLIP=11, ADVECT=12,
          ADVECT_REJECT=13, SL_SCATTER=14, SL_GATHER=15, ADVECT_GRID=16,
          CLASSIFY_LEVELSET=17, EXTEND_AVERAGE=18)


def build() -> str:
    """Compile liboracle.so with the committed Makefile; returns its path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def _state_type(R):
    P = C.POINTER(R)
    U8 = C.POINTER(C.c_uint8)

    class State(C.Structure):
        _fields_ = [
            ("nx", C.c_int32), ("ny", C.c_int32), ("dx", R), ("dy", R),
            ("density", R), ("pic_ratio", R), ("gravity_x", R), ("gravity_y", R),
            ("ext_layers", C.c_int32), ("cg_max_iter", C.c_int32),
            ("cg_tol", C.c_double), ("integrator", C.c_int32), ("front", C.c_int32),
            ("project_mean", C.c_int32), ("mask_front", C.c_int32),
            ("u0", P), ("u1", P), ("v0", P), ("v1", P),
            ("u_prev", P), ("v_prev", P), ("u_diff", P), ("v_diff", P),
            ("label", U8), ("solid_mask", U8), ("phi", P),
            ("n_particles", C.c_int64), ("particles", P), ("reject_flag", U8),
            ("acc_u", P), ("acc_wu", P), ("acc_v", P), ("acc_wv", P),
            ("mask_u0", U8), ("mask_u1", U8), ("mask_v0", U8), ("mask_v1", U8),
            ("ext_info_u", U8), ("ext_info_v", U8),
            ("fluid_index", C.POINTER(C.c_int32)), ("pressure", P), ("rhs", P),
            ("n_fluid", C.c_int32), ("cg_iterations", C.c_int32),
            ("cg_error", C.c_double), ("particle_pressure_k", R),
            ("closed_walls", C.c_int32),
            ("replay_label", U8), ("replay_face", U8), ("replay_reject", U8),
            ("record_face", U8), ("dirichlet_pass", C.c_int32),
        ]
    return State


State32 = _state_type(C.c_float)
State64 = _state_type(C.c_double)


def lib():
    global _LIB
    with _LOCK:
        if _LIB is None:
            path = os.path.join(_HERE, "liboracle.so")
            if not os.path.exists(path):
                build()
            L = C.CDLL(path)
            for sfx, S, R in (("_f32", State32, C.c_float), ("_f64", State64, C.c_double)):
                getattr(L, "orc_defaults" + sfx).argtypes = [C.POINTER(S), C.c_int32,
                                                             C.c_int32, R, R, R, R]
                getattr(L, "orc_defaults" + sfx).restype = None
                getattr(L, "orc_run_phase" + sfx).argtypes = [C.POINTER(S), C.c_int32, R]
                getattr(L, "orc_step" + sfx).argtypes = [C.POINTER(S), C.c_int32, R]
                getattr(L, "orc_validate" + sfx).argtypes = [C.POINTER(S), C.c_int32,
                                                             C.c_int32, R, R]
                getattr(L, "orc_interp" + sfx).argtypes = [C.POINTER(R), C.c_int32,
                                                           C.c_int32, R, R, R, R]
                getattr(L, "orc_interp" + sfx).restype = R
                getattr(L, "orc_splat" + sfx).argtypes = [C.POINTER(R), C.c_int32,
                                                          C.c_int32, R, R, R, R, R]
                getattr(L, "orc_splat" + sfx).restype = None
                getattr(L, "orc_rk_displacement" + sfx).argtypes = [
                    C.POINTER(S), R, R, R, C.POINTER(R), C.POINTER(R)]
                getattr(L, "orc_rk_displacement" + sfx).restype = None
                getattr(L, "orc_advect_level_set" + sfx).argtypes = [C.POINTER(S), C.POINTER(R), R]
                getattr(L, "orc_advect_level_set" + sfx).restype = None
            _LIB = L
    return _LIB


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class OracleSim:
    """A FluidDomain + FluidSolver pair on the CPU, in float32 or float64.

    Grids are numpy arrays of shape (ny, nx) indexed [j, i]."""

    def __init__(self, nx, ny, length_x, length_y, density, pic_ratio,
                 precision: str = "f32"):
        assert precision in ("f32", "f64")
        self.sfx = "_" + precision
        self.R = np.float32 if precision == "f32" else np.float64
        self.cR = C.c_float if precision == "f32" else C.c_double
        self.s = (State32 if precision == "f32" else State64)()
        L = lib()
        getattr(L, "orc_defaults" + self.sfx)(C.byref(self.s), nx, ny, length_x, length_y,
                                              density, pic_ratio)
        sh = (ny, nx)
        z = lambda dt=None: np.zeros(sh, dtype=dt or self.R)  # noqa: E731
        self.u0, self.u1, self.v0, self.v1 = z(), z(), z(), z()
        self.u_prev, self.v_prev, self.u_diff, self.v_diff = z(), z(), z(), z()
        # MacGrid's constructor calls clearCellTypeBuffer (src/MacGrid.cpp:24,32-50)
        self.label = np.full(sh, AIR, dtype=np.uint8)
        self.label[0, :] = self.label[-1, :] = SOLID
        self.label[:, 0] = self.label[:, -1] = SOLID
        self.acc = [z() for _ in range(4)]
        self.masks = [z(np.uint8) for _ in range(4)]
        self.ext_info_u, self.ext_info_v = z(np.uint8), z(np.uint8)
        self.fluid_index = z(np.int32)
        self.pressure, self.rhs = z(), z()
        self.particles = np.zeros((0, 4), dtype=self.R)
        self.reject_flag = np.zeros(0, dtype=np.uint8)
        self.solid_mask = None
        self.phi = None
        s, R = self.s, self.cR
        for name in ("u0", "u1", "v0", "v1", "u_prev", "v_prev", "u_diff", "v_diff",
                     "pressure", "rhs"):
            setattr(s, name, _ptr(getattr(self, name), R))
        s.label = _ptr(self.label, C.c_uint8)
        s.acc_u, s.acc_wu, s.acc_v, s.acc_wv = (_ptr(a, R) for a in self.acc)
        s.mask_u0, s.mask_u1, s.mask_v0, s.mask_v1 = (_ptr(a, C.c_uint8) for a in self.masks)
        s.ext_info_u = _ptr(self.ext_info_u, C.c_uint8)
        s.ext_info_v = _ptr(self.ext_info_v, C.c_uint8)
        s.fluid_index = _ptr(self.fluid_index, C.c_int32)
        self.set_particles(self.particles)

    # -- state ----------------------------------------------------------------
    @property
    def nx(self):
        return self.s.nx

    @property
    def ny(self):
        return self.s.ny

    def set_particles(self, parts):
        self.particles = np.array(parts, dtype=self.R, copy=True, order="C").reshape(-1, 4)
        self.reject_flag = np.zeros(len(self.particles), dtype=np.uint8)
        self.s.particles = _ptr(self.particles, self.cR)
        self.s.reject_flag = _ptr(self.reject_flag, C.c_uint8)
        self.s.n_particles = len(self.particles)

    def set_solid_mask(self, mask):
        if mask is None:
            self.solid_mask = None
            self.s.solid_mask = C.POINTER(C.c_uint8)()
            return
        self.solid_mask = np.array(mask, dtype=np.uint8, copy=True, order="C")
        self.s.solid_mask = _ptr(self.solid_mask, C.c_uint8)

    def set_phi(self, phi):
        self.phi = np.array(phi, dtype=self.R, copy=True, order="C")
        self.s.phi = _ptr(self.phi, self.cR)

    def advect_level_set(self, dt):
        """FluidDomain::advectLevelSet on this sim's level set (set_phi)."""
        getattr(lib(), "orc_advect_level_set" + self.sfx)(C.byref(self.s),
                                                          _ptr(self.phi, self.cR), dt)

    def vel(self, which="front"):
        f = self.s.front
        if which == "front":
            return (self.u1, self.v1) if f else (self.u0, self.v0)
        return (self.u0, self.v0) if f else (self.u1, self.v1)

    def set_vel(self, u, v, which="front"):
        uu, vv = self.vel(which)
        uu[...] = u
        vv[...] = v

    def valid_masks(self):
        """Front valid masks (u, v) after the last extension."""
        mf = self.s.mask_front
        return (self.masks[1], self.masks[3]) if mf else (self.masks[0], self.masks[2])

    def set_params(self, **kw):
        for k, v in kw.items():
            setattr(self.s, k, v)

    # -- decision replay --------------------------------------------------------
    def record_faces(self, on=True):
        """Record each step's face decisions in ``self.face_record`` (one byte per cell:
        bit 0 / 1 u / v normalised by P2G, bit 2 / 3 zeroed by the first Dirichlet
        pass, bit 4 / 5 by the second)."""
        if on:
            self.face_record = np.zeros((self.ny, self.nx), np.uint8)
            self.s.record_face = _ptr(self.face_record, C.c_uint8)
        else:
            self.face_record = None
            self.s.record_face = C.POINTER(C.c_uint8)()

    def set_replay(self, label=None, face=None, reject=None):
        """The next steps take these decisions instead of computing them: labels
        (ny, nx), face decisions (ny, nx; bits as in record_faces), reject flags (one
        per particle).  None leaves that decision to the step."""
        sh = (self.ny, self.nx)
        self._replay = []
        for name, a, n in (("replay_label", label, sh), ("replay_face", face, sh),
                           ("replay_reject", reject, None)):
            if a is None:
                setattr(self.s, name, C.POINTER(C.c_uint8)())
                continue
            a = np.ascontiguousarray(a, dtype=np.uint8)
            if n is not None:
                a = a.reshape(n)
            elif a.size != len(self.particles):
                raise ValueError("replayed reject flags: one per particle")
            self._replay.append(a)
            setattr(self.s, name, _ptr(a, C.c_uint8))

    # -- running --------------------------------------------------------------
    def run_phase(self, phase, dt=0.0):
        ph = PH[phase] if isinstance(phase, str) else phase
        rc = getattr(lib(), "orc_run_phase" + self.sfx)(C.byref(self.s), ph, dt)
        if rc != 0:
            raise RuntimeError(f"oracle phase {phase} failed ({rc})")

    def step(self, scheme, dt):
        rc = getattr(lib(), "orc_step" + self.sfx)(C.byref(self.s), scheme, dt)
        if rc != 0:
            raise RuntimeError(f"oracle step {scheme} failed ({rc})")

    def rk_displacement(self, x, y, h):
        ox, oy = self.cR(), self.cR()
        getattr(lib(), "orc_rk_displacement" + self.sfx)(C.byref(self.s), x, y, h,
                                                         C.byref(ox), C.byref(oy))
        return ox.value, oy.value

    def interp(self, grid, x, y):
        g = np.ascontiguousarray(grid, dtype=self.R)
        return getattr(lib(), "orc_interp" + self.sfx)(_ptr(g, self.cR), self.nx, self.ny,
                                                       self.s.dx, self.s.dy, x, y)

    def splat(self, grid, x, y, val):
        assert grid.dtype == self.R and grid.flags.c_contiguous
        getattr(lib(), "orc_splat" + self.sfx)(_ptr(grid, self.cR), self.nx, self.ny,
                                               self.s.dx, self.s.dy, x, y, val)

    @classmethod
    def from_scene(cls, scene, precision="f32"):
        sim = cls(scene.nx, scene.ny, scene.length_x, scene.length_y, scene.density,
                  scene.pic_ratio, precision)
        sim.set_particles(scene.particles)
        if scene.u is not None:
            sim.set_vel(scene.u, scene.v)
        if scene.solid_mask is not None:
            sim.set_solid_mask(scene.solid_mask)
        if scene.phi is not None:
            sim.set_phi(scene.phi)
        return sim
