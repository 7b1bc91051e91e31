"""Ensemble launches and graph replay (GPU).  One launch set serving many
domains must give every domain exactly (bitwise) what a standalone run of that
domain gives: the per-cell arithmetic and the per-chunk reductions are the
same, only the block-to-domain mapping differs."""
import sys

import numpy as np
import pytest
import torch

from hipims_amd import _lib
from hipims_amd import generators as gen
from hipims_amd.ensemble import Ensemble
from hipims_amd.grid import GridSolver

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def scenario(n, seed, peak_mmh, dur, dtype, order=2):
    c = gen.ensemble_case(n, seed, gen.mmh(peak_mmh), dur)
    s = GridSolver(c.mask, c.h, c.qx, c.qy, c.z, c.dx, cfl=c.cfl, order=order, dtype=dtype, first_dt=c.first_dt,
                   manning=c.manning, landuse=c.landuse)
    s.set_station_rain(c.rain_table)
    return s


SPECS = [(40, 2001, 80.0, 1800.0), (72, 2002, 120.0, 900.0), (100, 2003, 30.0, 3600.0), (33, 2004, 60.0, 600.0)]


def state(s):
    """State planes, h_max and the control fields a step defines (not
    stopped_synced, which records the copy state of a stopped domain's planes)."""
    return [s.h.clone(), s.qx.clone(), s.qy.clone(), s.h_max.clone(), s.ctrl[: len(_lib.CTRL_FIELDS)].clone()]


@pytest.mark.parametrize("single_pass", [True, False])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_ensemble_matches_standalone_bitwise(cuda, dtype, single_pass):
    """The single-pass ensemble against standalone single-pass steps, the
    two-pass ensemble against standalone two-pass steps."""
    say(f"[ensemble vs standalone {dtype}, single_pass {single_pass}]")
    ens_s = [scenario(*sp, dtype) for sp in SPECS]
    alone = [scenario(*sp, dtype) for sp in SPECS]
    e = Ensemble(ens_s, single_pass=single_pass)
    e.step(40)
    for s in alone:
        if single_pass:
            s.step_fused(40)
        else:
            s.step(40)
    torch.cuda.synchronize()
    for k, (a, b) in enumerate(zip(ens_s, alone)):
        for x, y in zip(state(a), state(b)):
            assert torch.equal(x, y), f"domain {k} differs"
        assert a.ctrl_dict()["steps"] == 40


def test_single_pass_ensemble_order1_matches_standalone_bitwise(cuda):
    say("[single-pass ensemble vs standalone, order 1]")
    ens_s = [scenario(*sp, torch.float32, order=1) for sp in SPECS]
    alone = [scenario(*sp, torch.float32, order=1) for sp in SPECS]
    Ensemble(ens_s).step(24)
    for s in alone:
        s.step_fused(24)
    torch.cuda.synchronize()
    for k, (a, b) in enumerate(zip(ens_s, alone)):
        for x, y in zip(state(a), state(b)):
            assert torch.equal(x, y), f"domain {k} differs"


@pytest.mark.parametrize("single_pass", [True, False])
def test_graph_replay_matches_eager(cuda, single_pass):
    say(f"[graph replay, single_pass {single_pass}]")
    eager = [scenario(*sp, torch.float32) for sp in SPECS]
    graphed = [scenario(*sp, torch.float32) for sp in SPECS]
    Ensemble(eager, single_pass=single_pass).step(30)
    e = Ensemble(graphed, single_pass=single_pass)
    e.capture(10)
    for r in range(3):
        e.replay()
        # eager work and a synchronisation between replays
        _ = [s.ctrl.sum().item() for s in graphed]
    torch.cuda.synchronize()
    for k, (a, b) in enumerate(zip(eager, graphed)):
        for x, y in zip(state(a), state(b)):
            assert torch.equal(x, y), f"domain {k} differs after replays"


@pytest.mark.parametrize("single_pass", [True, False])
def test_stop_time_freezes_each_domain(cuda, single_pass):
    say(f"[stop time, single_pass {single_pass}]")
    ss = [scenario(*sp, torch.float32) for sp in SPECS]
    for s in ss:
        s.ctrl[_lib.CTRL_EXPORT_T] = 120.0  # the export clamp lands t exactly on 120 s
    e = Ensemble(ss, t_stop=120.0, single_pass=single_pass)
    replays = e.run(steps_per_replay=20)
    torch.cuda.synchronize()
    snap = [state(s) for s in ss]
    for s in ss:
        assert s.t == 120.0
    e.replay()
    torch.cuda.synchronize()
    for a, b in zip(ss, snap):
        for x, y in zip(state(a), b):
            assert torch.equal(x, y)
    say(f"  replays {replays}, steps {[s.ctrl_dict()['steps'] for s in ss]}")


def test_compaction_with_retuned_unit_height_matches_standalone(cuda):
    """Staggered stop times make run() rebuild its launch maps (4 -> 3 -> 2
    or fewer active), choosing the unit height again each time; every domain's
    state, h_max, t, dt and step count equal a standalone single-pass run
    to the same stop time, bitwise (the per-cell arithmetic does not depend
    on the unit height)."""
    say("[compaction with retuned unit height]")
    # early steps of the dry start are tens of seconds long; 300 s apart, the
    # stops fall in different replays of two steps
    stops = [300.0, 600.0, 900.0, 900.0]
    ens_s = [scenario(*sp, torch.float32) for sp in SPECS]
    alone = [scenario(*sp, torch.float32) for sp in SPECS]
    for s, a, t in zip(ens_s, alone, stops):
        s.set_t_stop(t)
        a.set_t_stop(t)
    for s in ens_s:
        s.set_tile_rows(16)
    e = Ensemble(ens_s)
    heights = []

    def progress(k, n_act):
        heights.append((n_act, len(e.members), ens_s[-1].tile_rows))

    # retune_below=1: the height is timed again at every rebuild
    e.run(steps_per_replay=2, progress=progress, retune_below=1.0)
    torch.cuda.synchronize()
    # rebuilt at least twice: over 3 domains, then over 2 or fewer
    assert {4, 3} <= {m for _, m, _ in heights} and len(e.members) <= 2, heights
    assert len(e.retunes) >= 2, e.retunes
    for a in alone:
        while float(a.ctrl[_lib.CTRL_T]) < float(a.ctrl[_lib.CTRL_T_STOP]):
            a.step_fused(2)
    torch.cuda.synchronize()
    keys = ("t", "dt", "steps")
    for k, (a, b) in enumerate(zip(ens_s, alone)):
        for x, y in zip(state(a)[:4], state(b)[:4]):
            assert torch.equal(x, y), f"domain {k} differs"
        da, db = a.ctrl_dict(), b.ctrl_dict()
        assert all(da[f] == db[f] for f in keys), (k, da, db)
        assert da["t"] >= stops[k]
    say(f"  (active, members, unit height) per check: {sorted(set(heights))}; retunes {e.retunes}")


def test_single_pass_ensemble_rejects_odd_step_counts(cuda):
    ss = [scenario(*sp, torch.float32) for sp in SPECS[:2]]
    with pytest.raises(_lib.HipimsError, match="even number of steps"):
        Ensemble(ss).step(3)


def test_host_write_to_a_stopped_domain_survives_the_next_step(cuda):
    """A stopped domain whose state the host edits between calls keeps the
    edit through single-pass steps (odd and even counts)."""
    say("[stopped domain, host edit]")
    s = scenario(*SPECS[0], torch.float32)
    s.set_t_stop(60.0)
    s.ctrl[_lib.CTRL_EXPORT_T] = 60.0
    while s.t < 60.0:
        s.step_fused(2)
    s.step_fused(2)  # both sets of planes now hold the stopped state
    torch.cuda.synchronize()
    assert s.ctrl[_lib.CTRL_STOPPED_SYNCED].item() == 1.0
    for n in (1, 2, 3):
        edit = s.h.clone()
        edit[s.layout.ny // 2, : s.layout.nx] += 0.125
        s.h.copy_(edit)
        s.step_fused(n)
        torch.cuda.synchronize()
        assert torch.equal(s.h, edit), f"host edit lost after step_fused({n})"

