"""The drop-in C++ classes (FluidDomain / FluidSolver / MacGrid / MarkerParticleSet)
against the C ABI path and the oracle.

The example driver fsg_dam_break runs the reference's simple example
(examples/simple.cpp:21-62) through the C++ API: FluidSource emission on the host,
FluidSolver::step* on the GPU.  Its initial particles seed an fsg.Domain (the same
kernels through ctypes) and the FP32 oracle, which take the same steps.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from fluid_simulation_amd import fsg

from helpers import rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fluid_simulation_amd", "fsg_dam_break")


def _run_example(tmp_path, n, steps, scheme, pic=0.05, solver=0, tol=1e-7, max_iter=20000,
                 ppm=None):
    if not os.path.exists(EXE):
        pytest.fail("fsg_dam_break is not built (run __graft_entry__.build())")
    out = os.path.join(str(tmp_path), f"dam_{n}_{scheme}.bin")
    print(f"fsg_dam_break n={n} steps={steps} scheme={scheme} solver={solver}", flush=True)
    cmd = [EXE, str(n), str(steps), str(scheme), out, repr(pic), str(solver), repr(tol),
           str(max_iter)] + ([ppm] if ppm else [])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(out, "rb").read()
    hdr = np.frombuffer(raw[:20], np.int32)
    n_, steps_, scheme_, n0, n1 = (int(x) for x in hdr)
    assert (n_, steps_, scheme_) == (n, steps, scheme)
    off = 20
    init = np.frombuffer(raw[off:off + 16 * n0], np.float32).reshape(n0, 4)
    off += 16 * n0
    fin = np.frombuffer(raw[off:off + 16 * n1], np.float32).reshape(n1, 4)
    off += 16 * n1
    u = np.frombuffer(raw[off:off + 4 * n * n], np.float32).reshape(n, n)
    off += 4 * n * n
    v = np.frombuffer(raw[off:off + 4 * n * n], np.float32).reshape(n, n)
    off += 4 * n * n
    lab = np.frombuffer(raw[off:off + n * n], np.uint8).reshape(n, n)
    assert off + n * n == len(raw)
    return init, fin, u, v, lab


def _emitter_lattice(n):
    """FluidSource::update's lattice for the example's area, accumulated in float32
    (src/FluidDomain.cpp:36-46): y rows outer, x columns inner, step delta / 2.5."""
    f32 = np.float32
    delta = 1.0 / n
    step = f32(np.float64(f32(delta)) / 2.5)   # MyFloat / 2.5 is a double divide
    x0, x1 = f32(2.0 / n), f32(0.35)
    y0, y1 = f32(2.0 / n), f32(1 - 2.0 / n)
    xs, x = [], x0
    while x < x1:
        xs.append(x)
        x = f32(x + step)
    ys, y = [], y0
    while y < y1:
        ys.append(y)
        y = f32(y + step)
    out = np.zeros((len(ys) * len(xs), 4), np.float32)
    out[:, 0] = np.tile(np.array(xs, np.float32), len(ys))
    out[:, 1] = np.repeat(np.array(ys, np.float32), len(xs))
    return out


@pytest.mark.parametrize("scheme", [fsg.SEMI_LAGRANGIAN, fsg.PIC, fsg.FLIP, fsg.PICFLIP])
def test_cpp_dam_break(tmp_path, scheme):
    from oracle import OracleSim
    n, steps, dt = 64, 4, 0.01
    init, fin, u, v, lab = _run_example(tmp_path, n, steps, scheme)
    # host emission: the lattice of the example's fluid source
    assert np.array_equal(init, _emitter_lattice(n))

    d = fsg.Domain(n, n, 1.0, 1.0, 0.01, 0.05)
    d.set_params(cg_tol=1e-7, cg_max_iter=20000, solver=fsg.SOLVER_PLAIN)
    d.set_particles(init)
    o = OracleSim(n, n, 1.0, 1.0, 0.01, 0.05, "f32")
    o.set_params(cg_tol=1e-7, cg_max_iter=20000)
    o.set_particles(init)
    for _ in range(steps):
        d.step(scheme, dt)
        o.step(scheme, dt)
    ua, va = d.velocity(fsg.VEL_FRONT)
    uo, vo = o.vel("front")
    assert np.array_equal(lab, d.labels())
    assert np.array_equal(lab, o.label)
    e_abi = max(rel_err(fin, d.particles()), rel_err(u, ua), rel_err(v, va))
    e_orc = max(rel_err(fin, o.particles), rel_err(u, uo), rel_err(v, vo))
    print(f"scheme {scheme}: C++ vs ABI {e_abi:.2e}, C++ vs oracle {e_orc:.2e}")
    assert e_abi <= 1e-5
    assert e_orc <= 1e-4


def test_cpp_reference_defaults_run(tmp_path):
    """The reference's own solver settings (CG at most 100 iterations, tolerance
    FLT_EPSILON) through the C++ API: the example runs and particles stay finite."""
    init, fin, u, v, lab = _run_example(tmp_path, 64, 10, fsg.PICFLIP, tol=1.1920929e-07,
                                        max_iter=100)
    assert len(fin) == len(init) > 0
    assert np.isfinite(fin).all() and np.isfinite(u).all() and np.isfinite(v).all()


def _render_reference_frame(n, parts, lab, w=512):
    """The example's frame (Renderer::clearCanvas, renderGridCellsToCanvas,
    renderParticlesToCanvas, writeCanvasToPpm; src/Renderer.cpp:14-56,141-162,
    217-248) for a 1 x 1 world drawn on a w x w canvas, computed independently with
    numpy: cells are rectangles [int(i c), int((i+1) c)] (c = w / n pixels) with a
    white outline, particles 3 x 3 squares at int(w x)."""
    f32 = np.float32
    img = np.ones((w, w, 3), np.float32)
    c = w // n
    fill = {0: (0.7, 0.7, 1.0), 1: (1.0, 1.0, 1.0), 2: (0.5, 0.5, 0.5)}
    cols = np.array([fill[int(k)] for k in range(3)], np.float32)
    yy, xx = np.mgrid[0:w, 0:w]
    cell_lab = lab[np.minimum(yy // c, n - 1), np.minimum(xx // c, n - 1)]
    img[...] = cols[cell_lab]
    edge = (xx % c == 0) | (yy % c == 0) | (xx == w - 1) | (yy == w - 1)
    img[edge] = 1.0
    px = (f32(w) * parts[:, 0].astype(np.float32)).astype(np.int64)
    py = (f32(w) * parts[:, 1].astype(np.float32)).astype(np.int64)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            x = np.clip(px + dx, 0, w - 1)
            y = np.clip(py + dy, 0, w - 1)
            img[y, x] = (f32(0.3), f32(0.6), f32(0.9))
    return (np.clip(img, 0, 1) * f32(255)).astype(np.uint8)


def test_cpp_render_frame(tmp_path):
    """The C++ Renderer draws the host mirrors after GPU steps; the PPM matches an
    independent numpy rendering of the same particles and labels pixel for pixel."""
    ppm = os.path.join(str(tmp_path), "frame.ppm")
    init, fin, u, v, lab = _run_example(tmp_path, 64, 3, fsg.PICFLIP, ppm=ppm)
    raw = open(ppm, "rb").read()
    head = b"P6\n512 512\n255\n"
    assert raw.startswith(head)
    img = np.frombuffer(raw[len(head):], np.uint8).reshape(512, 512, 3)
    ref = _render_reference_frame(64, fin, lab)
    assert np.array_equal(img, ref), int((img != ref).any(-1).sum())
    assert (img == (76, 153, 229)).all(-1).sum() > len(fin)   # particle colour present
