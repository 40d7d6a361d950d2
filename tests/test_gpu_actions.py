"""The reference's per-function entry points (actions, pair passes, one tool's contact, the
leap-frog phases) on the GPU.

leap_frog::step composed call by call from them (api.leap_frog.step_by_actions, the order of
src/leap_frog.cu:121-195) runs the same arithmetic as the fused step, so every field, N and the
recorded tool forces must be bit-identical to mfg_leap_frog_step's: the wedge cut (Brookshaw
heat, CSPM, exhaustive contact) in FP64 and FP32 + Kahan, and the single-grain thermal-contact
cut (PSE heat, tool particles, friction heat).  The fields only the reference keeps (pressure,
artificial stress, velocity gradient, stress divergence terms, contact normal) are compared with
the oracle's after its step.
"""
import numpy as np
import pytest

import helpers
import oracle as O
from iwf_mfree_gpu_3d_amd import _lib, api, scenes
from test_gpu_parity import rel_err

pytestmark = pytest.mark.gpu

FIELDS = [_lib.F_ID, _lib.F_POS, _lib.F_VEL, _lib.F_S, _lib.F_T, _lib.F_RHO, _lib.F_EPS_PL, _lib.F_POS_T,
          _lib.F_VEL_T, _lib.F_RHO_T, _lib.F_S_T, _lib.F_T_T, _lib.F_FC, _lib.F_FT, _lib.F_BLANKED]


def run(preset, scale, seed, precision, steps, composed):
    sc = scenes.Scene(preset, scale, seed)
    parts, grid = scenes.setup_cut(sc, precision=precision)
    api.global_time_current = sc.params.dt
    lf = api.leap_frog(sc.n, grid.num_cell())
    for _ in range(steps):
        if composed:
            lf.step_by_actions(parts, grid, record_forces=True)
        else:
            lf.step(parts, grid, record_forces=True)
        api.global_time_current += api.global_time_dt
    out = {f: parts.download(f) for f in FIELDS}
    out["N"] = parts.counts()
    out["forces"] = api.global_tool_forces.get_forces_host().copy()
    return sc, parts, out


@pytest.mark.parametrize("case", [(1, 1.0, 1, _lib.MFG_FP64, 40), (1, 1.0, 1, _lib.MFG_FP32_KAHAN, 40),
                                  (2, 0.2, 3, _lib.MFG_FP64, 30)],
                         ids=["wedge_fp64", "wedge_fp32_kahan", "thermal_grain_fp64"])
def test_step_by_actions_matches_the_fused_step(hip_device, case):
    preset, scale, seed, precision, steps = case
    _, _, fused = run(preset, scale, seed, precision, steps, composed=False)
    _, _, comp = run(preset, scale, seed, precision, steps, composed=True)
    assert comp["N"] == fused["N"]
    assert np.abs(fused[_lib.F_VEL]).max() > 0.0
    for f in FIELDS:
        np.testing.assert_array_equal(comp[f], fused[f], err_msg=f"field {f}")
    np.testing.assert_array_equal(comp["forces"], fused["forces"])


def test_reference_arrays_match_the_oracle(hip_device):
    """After 30 composed steps of the wedge cut: pressure, artificial stress (of the last half
    step), velocity gradient and stress divergence terms of the last pair pass, and the contact
    normals, against the oracle's arrays after the same steps."""
    sc = scenes.Scene(1, 1.0, 1)
    sim = helpers.oracle_from_scene(sc)
    dt = sc.params.dt
    sim.set_time(dt, dt)
    assert sim.run(30, record=True) == 0
    _, parts, _ = run(1, 1.0, 1, _lib.MFG_FP64, 30, composed=True)
    assert np.array_equal(parts.download(_lib.F_ID).astype(np.int64), sim.field(O.F_ID).astype(np.int64))
    N = parts.N
    for f, of, sym in [(_lib.F_P, O.F_P, False), (_lib.F_R, O.F_R, True), (_lib.F_V_DER, O.F_V_DER, False),
                       (_lib.F_S_DER, O.F_S_DER, False), (_lib.F_N, O.F_N, False)]:
        g = parts.download(f)[:N]
        o = sim.field(of)[:N]
        if sym:
            o = 0.5 * (o + o.reshape(-1, 3, 3).transpose(0, 2, 1).reshape(-1, 9))
        assert np.linalg.norm(o) > 0.0, f"field {f} must be active"
        e = rel_err(g, o)
        print(f"field {f}: relative L2 error {e:.2e}")
        assert e < 1e-10, (f, e)
