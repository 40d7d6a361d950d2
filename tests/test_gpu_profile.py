"""Per-kernel event timing (mfg_profile_enable): every kernel of the step, or the neighbour and
pair passes only; the neighbour pass's opt-in pair counts (mfg_pair_stats_enable)."""
import numpy as np
import pytest

import helpers
from iwf_mfree_gpu_3d_amd import _lib, api, scenes
from iwf_mfree_gpu_3d_amd._lib import MfgError

pytestmark = pytest.mark.gpu


def recorded(level, steps=3):
    sc = scenes.Scene(1, 0.5, 3)
    parts, grid = scenes.setup_cut(sc)
    api.global_time_current = sc.params.dt
    lf = api.leap_frog(sc.n, grid.num_cell())
    api.profile_reset()
    api.profile_enable(level)
    try:
        scenes.run(lf, parts, grid, steps)
        api.synchronize()
    finally:
        api.profile_enable(False)
    return api.profile_read()


def test_every_kernel_level(hip_device):
    prof = recorded(_lib.MFG_PROFILE_ALL)
    for k in ("blank_hash_count", "permute_predict", "neighbours", "pair_fp64", "particle_update"):
        assert prof[k][0] == 3, (k, prof)
        assert prof[k][1] > 0.0


def test_passes_level_records_the_neighbour_and_pair_passes_only(hip_device):
    prof = recorded(_lib.MFG_PROFILE_PASSES)
    assert set(prof) == {"neighbours", "pair_fp64"}, prof
    assert all(v[0] == 3 and v[1] > 0.0 for v in prof.values()), prof


def test_unknown_level_is_rejected(hip_device):
    with pytest.raises(MfgError) as e:
        api.profile_enable(3)
    assert e.value.code == _lib.MFG_ERR_ARG


def stencil_candidates(start, end, nx, ny, nz):
    """Sum over particles of the particles in their cell's 5x5x5 stencil (clipped at the grid's
    faces), from the Green cell tables (start = end = -1 for an empty cell)."""
    n = np.where(start >= 0, end - start, 0).reshape(nx, ny, nz).astype(np.int64)
    p = np.pad(n, 2)
    t = sum(p[2 + a:2 + a + nx, 2 + b:2 + b + ny, 2 + c:2 + c + nz]
            for a in range(-2, 3) for b in range(-2, 3) for c in range(-2, 3))
    return int((n * t).sum())


def test_pair_stats_are_counted_on_request(hip_device):
    """Without counting the pair statistics are an error; with it, the interacting pairs equal the
    oracle's list lengths and the candidates the stencil sum over the cell tables."""
    sc = scenes.Scene(1, 0.5, 3)
    sim = helpers.oracle_from_scene(sc)
    assert sim.sort_only() == 0
    parts, grid = scenes.setup_cut(sc)
    grid.assign_hashes(parts)
    grid.adapt_particle_number(parts)
    grid.sort(parts)
    api.pair_stats_enable(False)
    grid.neighbours(parts)
    with pytest.raises(MfgError) as e:
        parts.pair_stats()
    assert e.value.code == _lib.MFG_ERR_STATE
    api.pair_stats_enable(True)
    try:
        c_g, _ = grid.neighbours(parts)
        cand, pairs = parts.pair_stats()
    finally:
        api.pair_stats_enable(False)
    c_o, _ = sim.neighbours()
    s_o, e_o = sim.cells()
    g = sim.grid()
    assert pairs == int(c_o.sum()) == int(c_g.sum()) > 0
    assert cand == stencil_candidates(s_o, e_o, g["nx"], g["ny"], g["nz"]) > pairs
