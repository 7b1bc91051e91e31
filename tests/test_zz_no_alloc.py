"""The allocation-free single-pass step (hp_grid_step_fused_noalloc) against
hp_grid_step_fused with a caller workspace, through the C ABI: bitwise equal
state, h_max and time control on two calls in a row, the second on a smaller
domain (the library-owned partial records keep the first call's contents),
and the size limit rejected before any launch."""
import ctypes
import sys

import pytest
import torch

from hipims_amd import _lib
from hipims_amd import generators as gen
from hipims_amd._lib import HpGridDesc, lib
from hipims_amd.grid import GridSolver

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def solver(n, dtype, order):
    c = gen.rainfall_runoff(n, 5.0, rain_mmh=90.0)
    s = GridSolver(c.mask, c.h, c.qx, c.qy, c.z, c.dx, cfl=c.cfl, order=order, dtype=dtype, first_dt=c.first_dt,
                   manning=c.manning, landuse=c.landuse)
    s.set_station_rain(c.rain_table)
    s.ensure_out()
    return s


def noalloc_steps(s, n):
    d = HpGridDesc.from_buffer_copy(s.desc)
    d.work = None  # ignored by the allocation-free entry point
    d.work_bytes = 0
    _lib.check(lib().hp_grid_step_fused_noalloc(ctypes.byref(d), int(n), _lib.stream_ptr(s.device)),
               "hp_grid_step_fused_noalloc")


def snapshot(s):
    return [s.h.clone(), s.qx.clone(), s.qy.clone(), s.h_max.clone(), s.ctrl[: len(_lib.CTRL_FIELDS)].clone()]


@pytest.mark.parametrize("dtype,order", [(torch.float32, 2), (torch.float64, 2), (torch.float32, 1)])
def test_noalloc_matches_workspace_step(cuda, dtype, order):
    for n, steps in ((160, 7), (72, 4)):  # the second call on a smaller domain
        say(f"[noalloc {dtype} order {order}: {n}^2, {steps} steps]")
        ref, mine = solver(n, dtype, order), solver(n, dtype, order)
        ref.step_fused(steps)
        noalloc_steps(mine, steps)
        torch.cuda.synchronize()
        for name, x, y in zip(("h", "qx", "qy", "h_max", "ctrl"), snapshot(ref), snapshot(mine)):
            assert torch.equal(x, y), f"{n}^2: {name} differs from the workspace step"
        assert ref.ctrl_dict()["steps"] == steps


def test_noalloc_rejects_domains_beyond_its_records(cuda):
    s = solver(40, torch.float32, 2)
    d = HpGridDesc.from_buffer_copy(s.desc)
    # 20000 x 20000 cells need about 31 MB of tile records: rejected before
    # any launch (the small planes behind the pointers are never touched)
    d.nx, d.ny, d.pitch = 20000, 20000, 20032
    rc = lib().hp_grid_step_fused_noalloc(ctypes.byref(d), 2, _lib.stream_ptr(s.device))
    assert rc == _lib.HP_ESPACE
    assert b"allocation-free" in lib().hp_last_error()
    assert lib().hp_grid_noalloc_max_bytes() == 8 << 20
