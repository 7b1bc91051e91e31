"""Host-side layout: the structured planes must describe exactly the
reference's compact cells and neighbour table (SWE_CUDA.py:115-227)."""
import numpy as np
import pytest

from hipims_amd import generators as gen
from hipims_amd.grid import make_layout


def _irregular_mask(seed=3, ny=23, nx=37):
    rng = np.random.Generator(np.random.PCG64(seed))
    m = np.where(rng.uniform(0, 1, (ny, nx)) < 0.8, 1, -9999).astype(np.int32)
    m[0, :] = m[-1, :] = m[:, 0] = m[:, -1] = -9999
    codes = np.array([30, 40, 50, 61, 612, 7, 80, 9])
    edge = (m > 0) & ((np.roll(m, 1, 0) <= 0) | (np.roll(m, -1, 0) <= 0) | (np.roll(m, 1, 1) <= 0)
                      | (np.roll(m, -1, 1) <= 0))
    m[edge] = codes[rng.integers(0, codes.size, edge.sum())]
    return m


def test_offsets_and_neighbours_match_reference_index(oracle_lib):
    mask = _irregular_mask()
    L = make_layout(mask)
    idx = oracle_lib.build_index(mask).reshape(5, -1)
    n = idx.shape[1]
    assert L.n_valid == n
    assert (L.pitch % 64 == 0) and L.pitch >= mask.shape[1]
    flat_mask = np.full(L.ny * L.pitch, -9999, dtype=np.int64)
    flat_mask.reshape(L.ny, L.pitch)[:, : L.nx] = mask
    # compact order, codes
    assert (flat_mask[L.offsets] == idx[0]).all()
    # neighbours E, N, W, S by address arithmetic
    inv = {int(o): i for i, o in enumerate(L.offsets)}
    for k, step in enumerate((1, -L.pitch, -1, L.pitch)):
        nb = L.offsets + step
        ours = np.array([inv.get(int(o), -1) for o in nb])
        assert (ours == idx[k + 1]).all()
    # slot table reproduces every code
    slots = L.type_slot.reshape(-1)[L.offsets]
    assert (L.slot_code[slots] == idx[0]).all()
    assert (L.type_slot.reshape(L.ny, L.pitch)[:, L.nx:] == 0).all()
    ys, xs = np.nonzero(mask > 0)
    assert L.cell0 == (xs[0], ys[0])


def test_layout_rejects_valid_cells_on_the_outer_ring():
    m = gen.ring_mask(4, 4)
    m[0, 2] = 1
    with pytest.raises(ValueError, match="outer ring"):
        make_layout(m)


def test_generators_are_deterministic_and_shaped():
    a = gen.rainfall_runoff(64, 5.0)
    b = gen.rainfall_runoff(64, 5.0)
    assert np.array_equal(a.z, b.z) and np.array_equal(a.landuse, b.landuse)
    assert a.mask.shape == (66, 66) and a.n_valid == 64 * 64
    zin = a.z[1:-1, 1:-1]
    assert zin.max() <= 50.0 + 1e-9 and zin.min() >= -2.0 - 1e-9
    assert set(np.unique(a.landuse[1:-1, 1:-1])) <= set(range(6))
    assert (a.mask[1:-1, -2] == 50).all() and (a.mask[1, 1:-2] == 40).all()
    c = gen.rainfall_runoff(64, 5.0, seed=7)
    assert not np.array_equal(a.z, c.z)
    f = gen.gaussian_random_field(128, 96, 2.0, 100.0, 1.0, 0.3, seed=1)
    assert f.shape == (128, 96) and f.min() >= 0.0 and abs(f.mean() - 1.0) < 0.05
    specs = gen.ensemble_specs(64)
    assert len(specs) == 64 and {s[0] for s in specs} <= set(gen.ENSEMBLE_SIZES)


def test_manning_table_must_cover_every_landuse_class():
    """A valid cell's class indexes the Manning table on the GPU: a table
    shorter than the classes in use is refused, in either order of setting."""
    import torch
    from hipims_amd.grid import GridSolver
    c = gen.rainfall_runoff(24, 5.0)
    s = GridSolver(c.mask, c.h, c.qx, c.qy, c.z, c.dx, manning=c.manning, landuse=c.landuse,
                   device=torch.device("cpu"))
    with pytest.raises(ValueError, match="manning table"):
        s.set_manning([0.03])
    s.set_manning(list(c.manning) + [0.1])
    with pytest.raises(ValueError, match="manning table"):
        s.set_landuse(np.full_like(c.landuse, 9))
