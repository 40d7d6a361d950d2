"""Forms of the pair pass against each other.

- The cell-tiled pass (MFG_PAIR_TILED) against the list-gather pass (MFG_PAIR_LIST): both form
  the same neighbour sums in the same order (every particle walks its own list from first to
  last entry), so FP64 results are bit-identical; in the FP32 modes only the FMA contraction the
  compiler picks in each kernel can differ, a rounding-level difference.  The 1M-particle case
  puts more than 64 particles into every grid column of the workpiece, so the tiled pass splits
  columns into several chunks there.
- The list pass's visiting order (strips of grid columns) against the sorted order: the order
  only decides which thread takes which particle, so every precision is bit-identical.
- The neighbour pass's cell tile (cells per wavefront) against one cell per wavefront: each cell
  is still one wavefront's work in the same order, so lists, PSE sums and every field are
  bit-identical in every heat mode (PSE in FP64 and FP32, Brookshaw).
"""
import numpy as np
import pytest

from iwf_mfree_gpu_3d_amd import _lib, api, scenes

pytestmark = pytest.mark.gpu

FIELDS = [_lib.F_POS, _lib.F_VEL, _lib.F_RHO, _lib.F_S, _lib.F_T, _lib.F_EPS_PL, _lib.F_POS_T,
          _lib.F_VEL_T, _lib.F_RHO_T, _lib.F_S_T, _lib.F_T_T, _lib.F_FC]


def run_case(preset, scale, seed, precision, pair_pass, steps, pair_strip=0, cell_tile=0, pair_block=0):
    sc = scenes.Scene(preset, scale, seed)
    parts, grid = scenes.setup_cut(sc, precision=precision, pair_pass=pair_pass, pair_strip=pair_strip,
                                   cell_tile=cell_tile, pair_block=pair_block)
    api.global_time_current = sc.params.dt
    lf = api.leap_frog(sc.n, grid.num_cell())
    scenes.run(lf, parts, grid, steps)
    order = np.argsort(parts.download(_lib.F_ID).astype(np.int64))
    out = {f: parts.download(f)[order] for f in FIELDS}
    out["N"] = parts.N
    return out


def rel_err(a, b):
    n = np.linalg.norm(b.ravel())
    d = np.linalg.norm((a - b).ravel())
    return d / n if n > 0 else d


@pytest.mark.parametrize("precision", [_lib.MFG_FP64, _lib.MFG_FP32_KAHAN, _lib.MFG_FP32_WIDE],
                         ids=["fp64", "kahan", "wide"])
def test_tiled_matches_list_thermal_contact(hip_device, precision):
    """Single grain, PSE heat, thermal contact and plasticity, 20 steps."""
    a = run_case(2, 0.2, 3, precision, api.MFG_PAIR_LIST, 20)
    b = run_case(2, 0.2, 3, precision, api.MFG_PAIR_TILED, 20)
    assert a["N"] == b["N"]
    for f in FIELDS:
        if precision == _lib.MFG_FP64:
            np.testing.assert_array_equal(b[f], a[f], err_msg=f"field {f}")
        else:
            assert rel_err(b[f], a[f]) < 1e-5, (f, rel_err(b[f], a[f]))


def test_tiled_matches_list_multichunk_columns(hip_device):
    """1M-particle cut (about 110 particles per grid column), one FP64 step: bit-identical."""
    a = run_case(2, 1.0, 1, _lib.MFG_FP64, api.MFG_PAIR_LIST, 1)
    b = run_case(2, 1.0, 1, _lib.MFG_FP64, api.MFG_PAIR_TILED, 1)
    assert a["N"] == b["N"]
    for f in FIELDS:
        np.testing.assert_array_equal(b[f], a[f], err_msg=f"field {f}")


@pytest.mark.parametrize("precision", [_lib.MFG_FP64, _lib.MFG_FP32_KAHAN, _lib.MFG_FP32_WIDE],
                         ids=["fp64", "kahan", "wide"])
@pytest.mark.parametrize("strip", [3, 16])
def test_visiting_order_does_not_change_results(hip_device, precision, strip):
    """Strips of 3 and 16 grid columns against the sorted order, thermal-contact grain, 10 steps."""
    a = run_case(2, 0.2, 3, precision, api.MFG_PAIR_LIST, 10, pair_strip=-1)
    b = run_case(2, 0.2, 3, precision, api.MFG_PAIR_LIST, 10, pair_strip=strip)
    assert a["N"] == b["N"]
    for f in FIELDS:
        np.testing.assert_array_equal(b[f], a[f], err_msg=f"field {f}")


@pytest.mark.parametrize("case", [(2, 0.2, 3, _lib.MFG_FP64), (2, 0.2, 3, _lib.MFG_FP32_KAHAN),
                                  (1, 1.0, 1, _lib.MFG_FP64)], ids=["grain_fp64_pse", "grain_fp32_pse", "wedge_fp64"])
@pytest.mark.parametrize("tile", [2, 4, 8])
def test_cell_tile_does_not_change_results(hip_device, case, tile):
    """Tiles of 2, 4 and 8 cells per wavefront against one, 30 steps: bit-identical."""
    preset, scale, seed, precision = case
    a = run_case(preset, scale, seed, precision, api.MFG_PAIR_LIST, 30, cell_tile=1)
    b = run_case(preset, scale, seed, precision, api.MFG_PAIR_LIST, 30, cell_tile=tile)
    assert a["N"] == b["N"]
    for f in FIELDS:
        np.testing.assert_array_equal(b[f], a[f], err_msg=f"field {f}")


@pytest.mark.parametrize("block", [64, 128])
@pytest.mark.parametrize("precision", [_lib.MFG_FP64, _lib.MFG_FP32_KAHAN, _lib.MFG_FP32_WIDE],
                         ids=["fp64", "kahan", "wide"])
def test_pair_block_does_not_change_results(hip_device, precision, block):
    """The list pair pass with 64- or 128-thread workgroups: every field bit-identical to the
    default 256 after 30 steps of the single-grain thermal-contact cut (each wavefront works
    alone)."""
    a = run_case(2, 0.2, 3, precision, api.MFG_PAIR_LIST, 30)
    b = run_case(2, 0.2, 3, precision, api.MFG_PAIR_LIST, 30, pair_block=block)
    assert a["N"] == b["N"]
    for f in FIELDS:
        np.testing.assert_array_equal(a[f], b[f], err_msg=f"field {f}")
