"""Device-clock timing of the multigrid kernels (fsg_set_kernel_clock): every launch is
counted, and a launch's device-clock duration is positive and no longer than the HIP
events around it, which also count the gap before the kernel starts."""
from __future__ import annotations

import pytest

from fluid_simulation_amd import fsg, scenes

pytestmark = pytest.mark.gpu


# at 64^2 the one-workgroup stage is its own kernel (one tile level); from two tile
# levels on the last tile level's down kernel runs it
@pytest.mark.parametrize("kernel,timer,n", [("coarse", "cg_mg_coarse", 64),
                                            ("down0", "cg_mg_down0", 256)])
def test_kernel_clock_counts_every_launch(kernel, timer, n):
    sc = scenes.dam_break(n)
    d = fsg.Domain.from_scene(sc, cg_tol=1e-6, cg_max_iter=200, solver=fsg.SOLVER_MG)
    for _ in range(2):
        d.step(fsg.PICFLIP, sc.dt)
    d.synchronize()
    d.reset_timing()
    d.set_timing_prefix(timer)
    d.set_timing(2)
    d.set_kernel_clock(True, kernel)
    for _ in range(3):
        d.step(fsg.PICFLIP, sc.dt)
    d.synchronize()
    us, n = d.kernel_clock()
    ev = d.timing()
    d.set_kernel_clock(False)
    d.set_timing(0)
    ev_ms, ev_n = ev[timer]
    assert n == ev_n > 0
    assert 0.0 < us / n < 1e4
    # per launch the device clock sees only the kernel; the events add the launch gap
    assert us * 1e-3 <= ev_ms * 1.02
    d.close()


def test_kernel_clock_rejects_unknown_mode():
    sc = scenes.dam_break(64)
    d = fsg.Domain.from_scene(sc, solver=fsg.SOLVER_MG)
    with pytest.raises(fsg.FsgError):
        d._chk(d._L.fsg_set_kernel_clock(d._h, 3), "fsg_set_kernel_clock")
    d.close()
