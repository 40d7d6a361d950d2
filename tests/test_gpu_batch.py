"""Batched steps (mfg_batch_*): every kernel of the step launched once over many instances.

An instance of a batch must compute exactly what leap_frog.step computes for it alone (the
single-instance path is pinned to the oracle by the parity tests): every field, the particle
count and the recorded tool forces bit-identical.  Covered: wedge cuts of different sizes, speeds,
Johnson-Cook constants and seeds in FP64 and both FP32 modes; the single-grain thermal-contact
cut (PSE heat, tool particles, exhaustive contact with the lane-parallel search) in FP64 and
FP32 + Kahan; steps replayed from a captured graph, checked after every replay against eager
steps with downloads and synchronisations between the replays; an instance removed from the
batch and stepped on alone; the checks at mfg_batch_add.
"""
import numpy as np
import pytest

from iwf_mfree_gpu_3d_amd import _lib, api, scenes

pytestmark = pytest.mark.gpu

FIELDS = [_lib.F_ID, _lib.F_POS, _lib.F_VEL, _lib.F_S, _lib.F_T, _lib.F_RHO, _lib.F_EPS_PL]

# (preset, scale, seed, overrides): wedge cuts that differ in size, speed and material
WEDGES = [(1, 1.0, 1, {}), (1, 0.6, 2, {"speed": 3e-3, "jc_A": 0.0121}),
          (1, 0.8, 3, {"speed": 2e-3, "rake_deg": -10.0, "depth_of_cut": 2e-3})]
GRAINS = [(2, 0.2, 1, {}), (2, 0.16, 4, {})]


class Inst:
    """One instance: its scene, device objects and the constants installed for it."""

    def __init__(self, spec, precision):
        preset, scale, seed, kw = spec
        self.sc = scenes.Scene(preset, scale, seed, **kw)
        self.parts, self.grid = scenes.setup_cut(self.sc, precision=precision)
        self.tools = list(api.global_tool)
        self.bl = api.global_blanking
        self.forces = api.global_tool_forces
        self.dt = self.sc.params.dt

    def fields(self):
        out = {f: self.parts.download(f) for f in FIELDS}
        out["N"] = self.parts.counts()
        out["forces"] = self.forces.get_forces_host().copy() if self.forces is not None else None
        return out


def alone(spec, precision, steps):
    inst = Inst(spec, precision)
    api.global_time_current = inst.dt
    lf = api.leap_frog(inst.sc.n, inst.grid.num_cell())
    scenes.run(lf, inst.parts, inst.grid, steps, record_forces=True)
    return inst.fields()


def batched(specs, precision):
    b = api.batch()
    insts = []
    for spec in specs:
        inst = Inst(spec, precision)
        b.add(inst.parts, inst.grid, inst.tools, inst.bl, inst.forces, dt=inst.dt, t_current=inst.dt)
        insts.append(inst)
    return b, insts


def assert_same(got, ref, what):
    assert got["N"] == ref["N"], what
    for f in FIELDS:
        np.testing.assert_array_equal(got[f], ref[f], err_msg=f"{what}: field {f}")
    if ref["forces"] is not None:
        np.testing.assert_array_equal(got["forces"], ref["forces"], err_msg=f"{what}: forces")


@pytest.mark.parametrize("precision", [_lib.MFG_FP64, _lib.MFG_FP32_KAHAN, _lib.MFG_FP32_WIDE],
                         ids=["fp64", "kahan", "wide"])
def test_batch_matches_alone_wedges(hip_device, precision):
    steps = 40
    refs = [alone(spec, precision, steps) for spec in WEDGES]
    b, insts = batched(WEDGES, precision)
    b.step(steps, record_forces=True)
    for k, (inst, ref) in enumerate(zip(insts, refs)):
        assert abs(b.time(k) - (inst.dt + steps * inst.dt)) <= 1e-12 * inst.dt * steps
        got = inst.fields()
        # penalty contact of these small blocks is intermittent (the force record holds the last
        # step only): engagement shows as particles set in motion
        assert np.abs(ref[_lib.F_VEL]).max() > 0.0, "the tool must have engaged"
        assert_same(got, ref, f"wedge {k}")
    print(f"batch of {len(insts)} wedge cuts ({[i.sc.n for i in insts]} particles), {steps} steps: "
          "bit-identical to each alone")


@pytest.mark.parametrize("precision", [_lib.MFG_FP64, _lib.MFG_FP32_KAHAN], ids=["fp64", "kahan"])
def test_batch_matches_alone_thermal_grains(hip_device, precision):
    steps = 30
    refs = [alone(spec, precision, steps) for spec in GRAINS]
    b, insts = batched(GRAINS, precision)
    b.step(steps, record_forces=True)
    for k, (inst, ref) in enumerate(zip(insts, refs)):
        assert_same(inst.fields(), ref, f"grain {k}")
    print(f"batch of {len(insts)} thermal-contact grains ({[i.sc.n for i in insts]} particles): bit-identical")


def test_batch_graph_replays_match_eager_steps(hip_device):
    """Three replays of 4 captured steps each, with downloads between them, against the eager
    batch at the same points."""
    precision = _lib.MFG_FP32_KAHAN
    checkpoints = 3
    eager, eager_insts = batched(WEDGES[:2], precision)
    want = []
    for _ in range(checkpoints):
        eager.step(4, record_forces=True)
        want.append([i.fields() for i in eager_insts])
    g, g_insts = batched(WEDGES[:2], precision)
    for c in range(checkpoints):
        g.step(4, record_forces=True, use_graph=True)
        api.synchronize()
        for k, inst in enumerate(g_insts):
            assert_same(inst.fields(), want[c][k], f"replay {c}, instance {k}")


def test_batch_remove_then_step_alone(hip_device):
    """Instance 1 leaves the batch after 15 steps and takes 15 more alone; the others take 15
    more in the batch.  Everyone matches 30 steps alone."""
    precision = _lib.MFG_FP64
    refs = [alone(spec, precision, 30) for spec in WEDGES]
    b, insts = batched(WEDGES, precision)
    b.step(15, record_forces=True)
    t1 = b.time(1)
    b.remove(1)
    b.step(15, record_forces=True)
    leaver = insts[1]
    scenes.setup_cut(leaver.sc, precision=precision)  # its constants back in the library
    api.global_tool = leaver.tools
    api.global_blanking = leaver.bl
    api.global_tool_forces = leaver.forces
    api.global_time_current = t1
    lf = api.leap_frog(leaver.sc.n, leaver.grid.num_cell())
    scenes.run(lf, leaver.parts, leaver.grid, 15, record_forces=True)
    for k, inst in enumerate(insts):
        assert_same(inst.fields(), refs[k], f"instance {k}")


def test_batch_add_checks(hip_device):
    """Instances of one batch share precision and heat model; a grid must be locked and still;
    a particle set or tool joins once."""
    b, insts = batched(WEDGES[:1], _lib.MFG_FP64)
    other = Inst(WEDGES[1], _lib.MFG_FP32_KAHAN)
    with pytest.raises(_lib.MfgError):
        b.add(other.parts, other.grid, other.tools, other.bl, other.forces, dt=other.dt, t_current=other.dt)
    i0 = insts[0]
    with pytest.raises(_lib.MfgError):
        b.add(i0.parts, i0.grid, i0.tools, i0.bl, i0.forces, dt=i0.dt, t_current=i0.dt)
    moving = Inst(WEDGES[2], _lib.MFG_FP64)
    moving.grid.set_bbox_vel(np.array([1e-3, 0.0, 0.0]))
    with pytest.raises(_lib.MfgError):
        b.add(moving.parts, moving.grid, moving.tools, moving.bl, moving.forces, dt=moving.dt, t_current=moving.dt)
    assert len(b) == 1


def test_snapshot_restore_continues_bit_identically(hip_device):
    """An instance snapshotted after 12 steps (particle state and tool motion), set up again from
    scratch, restored and stepped 12 more times equals the same instance stepped 24 times; a
    snapshot of another size is refused."""
    spec, precision = WEDGES[0], _lib.MFG_FP32_KAHAN
    ref = alone(spec, precision, 24)
    a = Inst(spec, precision)
    api.global_time_current = a.dt
    lf = api.leap_frog(a.sc.n, a.grid.num_cell())
    scenes.run(lf, a.parts, a.grid, 12, record_forces=True)
    snap = a.parts.snapshot()
    motion = [t.get_motion() for t in a.tools]
    t_now = api.global_time_current
    del a, lf
    b = Inst(spec, precision)  # fresh objects, the scene's initial state
    b.parts.restore(snap)
    for t, m in zip(b.tools, motion):
        t.set_motion(m)
    api.global_time_current = t_now
    lf = api.leap_frog(b.sc.n, b.grid.num_cell())
    scenes.run(lf, b.parts, b.grid, 12, record_forces=True)
    assert_same(b.fields(), ref, "restored")
    other = Inst(WEDGES[1], precision)
    with pytest.raises(_lib.MfgError):
        other.parts.restore(snap)
