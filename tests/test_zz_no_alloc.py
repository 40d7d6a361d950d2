"""The allocation-free step (mfg_leap_frog_step_noalloc) against the step (mfg_leap_frog_step).

Both entry points run the same wedge cut (exhaustive contact, so the contact search's slots, list
and results live in the step's work memory; visiting-order strips of 8 columns, so the pair
pass's segment offsets do too) from the same inputs.  Every field, the particle count and the
recorded tool forces must be bit-identical.  Two cuts in a row, the second smaller than the
first: the library's fixed work array keeps the first cut's contents when the second starts.
"""
import numpy as np
import pytest

from iwf_mfree_gpu_3d_amd import _lib, api, scenes

pytestmark = pytest.mark.gpu

FIELDS = [_lib.F_ID, _lib.F_POS, _lib.F_VEL, _lib.F_S, _lib.F_T, _lib.F_RHO, _lib.F_EPS_PL]


def cut(sc, steps, allocation_free):
    parts, grid = scenes.setup_cut(sc, pair_strip=8)
    api.global_time_current = sc.params.dt
    lf = api.leap_frog(sc.n, grid.num_cell())
    for _ in range(steps):
        lf.step(parts, grid, record_forces=True, allocation_free=allocation_free)
        api.global_time_step += 1
        api.global_time_current += api.global_time_dt
    n = parts.counts()
    fields = {f: parts.download(f) for f in FIELDS}
    forces = api.global_tool_forces.get_forces_host().copy()
    return n, fields, forces


@pytest.mark.parametrize("scale,steps", [(0.6, 40), (0.45, 40)], ids=["first", "second_smaller"])
def test_allocation_free_step_matches_the_step(hip_device, scale, steps):
    sc = scenes.Scene(1, scale, 3)
    print(f"allocation-free step: wedge cut at scale {scale}, {sc.n} particles, {steps} steps", flush=True)
    n_ref, ref, f_ref = cut(sc, steps, False)
    n_got, got, f_got = cut(sc, steps, True)
    assert n_got == n_ref
    assert np.abs(f_ref).max() > 0.0, "the tool must be in contact"
    np.testing.assert_array_equal(f_got, f_ref)
    for f in FIELDS:
        np.testing.assert_array_equal(got[f], ref[f], err_msg=f"field {f}")
