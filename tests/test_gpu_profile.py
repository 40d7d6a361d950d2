"""Per-kernel event timing (mfg_profile_enable): every kernel of the step, or the neighbour and
pair passes only."""
import pytest

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
