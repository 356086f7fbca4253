"""The C ABI library loads and exports every entry point include/fsg.h declares
(no device calls: this runs without a GPU)."""
import ctypes
import os
import re

from fluid_simulation_amd import fsg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "fsg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fsg_[a-z_0-9]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared_functions() == sorted(fsg.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(fsg.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_params_struct_layout_and_defaults():
    # layout: 2 floats, double, 6 int32, double, 5 int32, 1 float, 13 int32, padded to
    # the doubles' 8-byte alignment
    assert ctypes.sizeof(fsg.Params) == 4 * 2 + 8 + 4 * 6 + 8 + 4 * 5 + 4 + 4 * 13 + 4
    assert fsg.Params.mg_spread_tiles.offset == 120
    assert fsg.Params.mg_level_sweeps.offset == 112
    assert fsg.Params.particle_pressure_k.offset == 68
    assert fsg.Params.closed_walls.offset == 72
    assert fsg.Params.graph.offset == 76
    assert fsg.Params.deterministic.offset == 80
    assert fsg.Params.record_decisions.offset == 84
    assert fsg.Params.mg_smoother.offset == 88
    assert fsg.Params.graph_iters.offset == 96
    assert fsg.Params.mg_tile_threads.offset == 108
    p = fsg.default_params()
    assert p.gravity_y == ctypes.c_float(-9.82).value      # src/FluidSolver.cpp:232
    assert p.cg_max_iter == 100                              # src/FluidSolver.cpp:81
    assert p.ext_layers == 2                                 # src/FluidSolver.cpp:234
    assert abs(p.cg_tol - 2.0 ** -23) == 0                   # Eigen default for float
    assert p.integrator == fsg.RK3                           # include/FluidSolver.h:144
    assert p.particle_pressure_k == 0.0                      # src/FluidSolver.cpp:443
    assert p.closed_walls == 0                               # src/FluidSolver.cpp:306-319
    assert p.deterministic == 0 and p.record_decisions == 0 and p.mg_smoother == 1
    assert p.mg_coarse_sweeps == 0 and p.graph_iters == 0 and p.mg_fuse_update == 1
    assert p.splat_order == 1
    assert p.mg_tile_threads == 512 and p.mg_level_sweeps == 5 and p.mg_level0_sweeps == 3
    assert p.mg_spread_tiles == 1


def test_noalloc_rejects_bad_sizes_before_any_device_call():
    """fsg_p2g_noalloc checks its sizes and buffers before its first HIP call (so this
    runs without a GPU): a grid above FSG_NOALLOC_MAX_CELLS, a negative count and
    null buffers return FSG_E_INVALID."""
    f = fsg.load().fsg_p2g_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                  ctypes.c_float, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_void_p]
    f.restype = ctypes.c_int
    dummy = ctypes.c_void_p(16)
    assert f(dummy, 1, 1025, 1024, 1025.0, 1024.0, dummy, dummy, None) == -1
    assert f(dummy, -1, 64, 64, 64.0, 64.0, dummy, dummy, None) == -1
    assert f(None, 10, 64, 64, 64.0, 64.0, dummy, dummy, None) == -1
    assert f(dummy, 10, 64, 64, 64.0, 64.0, None, dummy, None) == -1
    assert f(dummy, 10, 64, 64, 0.0, 64.0, dummy, dummy, None) == -1


def test_version_string():
    assert b"gfx950" in fsg.load().fsg_version()


def test_cpp_layer_loads_and_exports_reference_api():
    """libfluidsim_amd.so (the drop-in C++ classes) loads on top of libfsg.so and
    exports the reference's public C++ entry points (include/FluidSolver.h:51-54,
    include/FluidDomain.h:40-46, include/MacGrid.h)."""
    path = os.path.join(ROOT, "fluid_simulation_amd", "libfluidsim_amd.so")
    assert os.path.exists(path), "libfluidsim_amd.so is not built"
    lib = ctypes.CDLL(path)
    for sym in ("_ZN11FluidSolver18stepSemiLagrangianER11FluidDomainf",
                "_ZN11FluidSolver7stepPICER11FluidDomainf",
                "_ZN11FluidSolver8stepFLIPER11FluidDomainf",
                "_ZN11FluidSolver11stepPICFLIPER11FluidDomainf",
                "_ZN11FluidDomainC1Eiiffff",
                "_ZN7MacGridC1Eiiff",
                "_ZN11FluidDomain13classifyCellsER17MarkerParticleSet",
                "_ZN17MarkerParticleSet6advectEf"):
        assert hasattr(lib, sym), sym
