"""fsg_p2g_noalloc, the entry point that allocates nothing, through the C ABI: its P2G
grid equals the handle's deterministic P2G phase (fsg_run_phase(FSG_PH_P2G) with
fsg_params.deterministic = 1, transferVelocityToGridSpread, src/FluidSolver.cpp:873-919)
bit for bit, on two calls in a row, the second on a smaller grid with fewer particles
(the static scratch keeps the first call's contents and must be cleared)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from fluid_simulation_amd import fsg, scenes

pytestmark = pytest.mark.gpu


def _noalloc():
    lib = fsg.load()
    f = lib.fsg_p2g_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                  ctypes.c_float, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_void_p]
    f.restype = ctypes.c_int
    return f


def _handle_p2g(n, steps):
    """A developed dam break: particles, the back buffer before P2G, and the handle's
    deterministic P2G result (its new front)."""
    sc = scenes.dam_break(n)
    d = fsg.Domain.from_scene(sc, solver=fsg.SOLVER_MG, cg_tol=1e-6, cg_max_iter=200,
                              deterministic=1)
    for _ in range(steps):
        d.step(fsg.PICFLIP, sc.dt)
    parts = d.particles()
    bu, bv = d.velocity(fsg.VEL_BACK)
    d.run_phase("P2G")
    ru, rv = d.velocity(fsg.VEL_FRONT)
    d.close()
    return sc, parts, bu, bv, ru, rv


def test_noalloc_p2g_equals_the_handle_twice():
    import torch
    f = _noalloc()
    stream = torch.cuda.current_stream().cuda_stream
    for n, steps, keep in ((256, 40, None), (96, 20, 5000)):
        sc, parts, bu, bv, ru, rv = _handle_p2g(n, steps)
        if keep is not None:  # the second call: fewer particles than the first
            parts = parts[:keep]
            d = fsg.Domain.from_scene(sc, deterministic=1)
            d.set_particles(parts)
            d.set_velocity(bu, bv, fsg.VEL_BACK)
            d.run_phase("P2G")
            ru, rv = d.velocity(fsg.VEL_FRONT)
            d.close()
        print(f"fsg_p2g_noalloc: {n} x {n}, {len(parts)} particles", flush=True)
        tp = torch.from_numpy(np.ascontiguousarray(parts)).cuda()
        tu = torch.from_numpy(np.ascontiguousarray(bu)).cuda()
        tv = torch.from_numpy(np.ascontiguousarray(bv)).cuda()
        rc = f(tp.data_ptr(), len(parts), sc.nx, sc.ny, sc.length_x, sc.length_y,
               tu.data_ptr(), tv.data_ptr(), stream)
        torch.cuda.synchronize()
        assert rc == fsg.OK
        gu, gv = tu.cpu().numpy(), tv.cpu().numpy()
        assert np.array_equal(gu.view(np.uint32), ru.view(np.uint32))
        assert np.array_equal(gv.view(np.uint32), rv.view(np.uint32))
