"""The reference's run-time self-checks on the device: the NaN/Inf rate guard that blanks a
particle (src/actions_gpu.cu:302-331) against the oracle, the neighbour-list capacity check, and
the all-blanked stop (src/grid_gpu_base.cu:159-163)."""
import ctypes as C

import numpy as np
import pytest

import helpers
import oracle as O
from iwf_mfree_gpu_3d_amd import _lib, api, scenes
from iwf_mfree_gpu_3d_amd._lib import MfgError
from test_gpu_parity import perturbed_scene, rel_err

pytestmark = pytest.mark.gpu


def test_nonfinite_velocity_blanks_like_the_oracle(hip_device):
    """A NaN velocity makes the particle's rates and, through the pair terms, its neighbours'
    rates non-finite: all of them are blanked in the same step, and the rest of the block steps
    on as in the oracle."""
    sc = perturbed_scene(1, 0.5, 7)
    k = int(np.argmin(np.linalg.norm(sc.pos - sc.pos.mean(axis=0), axis=1)))
    sc.vel[k] = np.nan
    sim = helpers.oracle_from_scene(sc)
    dt = sc.params.dt
    sim.set_time(dt, dt)
    parts, grid = scenes.setup_cut(sc)
    api.global_time_current = dt
    lf = api.leap_frog(sc.n, grid.num_cell())
    assert sim.run(2) == 0
    scenes.run(lf, parts, grid, 2)
    og = np.argsort(parts.download(_lib.F_ID).astype(np.int64))
    oo = np.argsort(sim.field(O.F_ID).astype(np.int64))
    bl_g = parts.download(_lib.F_BLANKED)[og]
    bl_o = sim.field(O.F_BLANKED)[oo]
    np.testing.assert_array_equal(bl_g, bl_o)
    assert bl_o[k] == 1.0 and bl_o.sum() > 10, "the particle and its neighbours are blanked"
    assert parts.N == sim.counts()[0] == sc.n - int(bl_o.sum())
    keep = bl_o != 1.0
    for f, of in [(_lib.F_POS, O.F_POS), (_lib.F_VEL, O.F_VEL), (_lib.F_S, O.F_S), (_lib.F_RHO, O.F_RHO)]:
        g = parts.download(f)[og][keep]
        o = sim.field(of)[oo][keep]
        if f == _lib.F_S:
            o = 0.5 * (o + o.reshape(-1, 3, 3).transpose(0, 2, 1).reshape(-1, 9))
        assert np.isfinite(g).all()
        assert rel_err(g, o) < 1e-12, (f, rel_err(g, o))


def test_neighbour_capacity_overflow_is_reported(hip_device):
    """Lists of 32 entries on a lattice with about 100 neighbours per particle: the step keeps
    every list within its capacity and the next counter read reports MFG_ERR_CAPACITY."""
    sc = scenes.Scene(1, 0.5, 3)
    parts, grid = scenes.setup_cut(sc, max_neighbours=32)
    api.global_time_current = sc.params.dt
    lf = api.leap_frog(sc.n, grid.num_cell())
    with pytest.raises(MfgError) as e:
        scenes.run(lf, parts, grid, 1)
        parts.counts()
    assert e.value.code == _lib.MFG_ERR_CAPACITY


def test_all_blanked_is_reported(hip_device):
    """A blanking box away from the workpiece blanks every particle; adapt_particle_number then
    stops with MFG_ERR_ALL_BLANKED."""
    sc = scenes.Scene(1, 0.5, 3)
    parts, grid = scenes.setup_cut(sc)
    far = sc.bbmax() + 1.0
    api.global_blanking = api.blanking(far, far + 0.1, max_vel=sc.params.max_vel)
    api.global_time_current = sc.params.dt
    lf = api.leap_frog(sc.n, grid.num_cell())
    scenes.run(lf, parts, grid, 1)
    with pytest.raises(MfgError) as e:
        grid.adapt_particle_number(parts)
    assert e.value.code == _lib.MFG_ERR_ALL_BLANKED
    n, ni = C.c_uint(99), C.c_uint(0)
    assert _lib.load().mfg_particle_gpu_counts(parts._h, C.byref(n), C.byref(ni)) == _lib.MFG_ERR_ALL_BLANKED
    assert n.value == 0
