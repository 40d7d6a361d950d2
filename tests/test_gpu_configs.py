"""The workload configurations on the GPU against the oracle, and the reference's own runs.

- 1M single grain (PSE heat, thermal contact, 1,001,367 particles) at full size: 3 steps.
- three grains (precision form) at reduced size: FP64 and both FP32 modes against the oracle;
  at full size (8M particles) the FP32 modes against the FP64 run.
- the batched form at full size (256 instances, 13.7M particles, 300 steps) and a short served
  stream with an FP64 request checked against the oracle.
- the reference-shaped C++ driver over the C ABI, and the reference's single-grain runs with its
  shipped tool meshes (setup_single_grain_1tool and _realscale).
"""
import os
import re
import subprocess

import numpy as np
import pytest

import helpers
import oracle as O
from iwf_mfree_gpu_3d_amd import _lib, api, scenes, serve
from test_gpu_parity import compare_fields, rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def run_both(sc, n, precision=_lib.MFG_FP64, record=True):
    sim = helpers.oracle_from_scene(sc)
    dt = sc.params.dt
    sim.set_time(dt, dt)
    parts, grid = scenes.setup_cut(sc, precision=precision)
    api.global_time_current = dt
    lf = api.leap_frog(sc.n, grid.num_cell())
    assert sim.run(n, record=record) == 0
    scenes.run(lf, parts, grid, n, record_forces=record)
    return sim, parts


def errors_by_id(parts, sim, T0, fields=(O.F_S, O.F_EPS_PL, O.F_T, O.F_POS, O.F_VEL)):
    """Relative L2 and max-abs error per field, particles matched by persistent id; T is the rise
    T - T_init."""
    og = np.argsort(parts.download(_lib.F_ID).astype(np.int64))
    oo = np.argsort(sim.field(O.F_ID).astype(np.int64))
    out = {}
    for f in fields:
        g = parts.download(f)[og]
        o = sim.field(f)[oo]
        if f == O.F_S:
            o = 0.5 * (o + o.reshape(-1, 3, 3).transpose(0, 2, 1).reshape(-1, 9))
        if f == O.F_T:
            g, o = g - T0, o - T0
        out[f] = (rel_err(g, o), float(np.abs(g - o).max()), float(np.linalg.norm(o)))
    return out


def test_single_grain_full_size_three_steps(hip_device):
    """The 1M-particle single-grain cut at full size (1,001,367 particles): 3 steps, every field
    within 1e-11 of the oracle."""
    sc = scenes.Scene(2, 1.0, 1)
    print(f"single grain full size: {sc.n} particles", flush=True)
    sim, parts = run_both(sc, 3)
    assert parts.N == sim.counts()[0]
    errs = compare_fields(parts, sim, 1e-11)
    print("single grain 1M, 3 steps, relative errors", errs)


def three_grain_scene():
    """The three-grain form at 15 % of the lattice (60 x 30 x 15) with 12 um grains 1 um apart
    along the cut and 20 um across it, tool speed 400 m/s, so that all three grains engage within
    100 steps."""
    return scenes.Scene(3, 0.15, 1, grain_edge=12e-4, grain_offset_x=1e-4, grain_offset_y=2e-3, speed=4e-2)


def test_three_grains_reduced_fp64(hip_device):
    """Three grains, three tools with tool particles, PSE heat, thermal contact and friction heat:
    100 steps, every field within 1e-9 of the oracle, forces of every tool non-zero."""
    sc = three_grain_scene()
    sim, parts = run_both(sc, 100)
    f_o = sim.forces()
    assert np.all(np.abs(f_o).max(axis=1) > 0), "every grain must be in contact"
    errs = compare_fields(parts, sim, 1e-9)
    print("three grains fp64 relative errors", errs)
    f_g = api.global_tool_forces.get_forces_host()
    np.testing.assert_allclose(f_g, f_o, rtol=1e-4, atol=1e-3 * np.abs(f_o).max())


@pytest.mark.parametrize("precision", [_lib.MFG_FP32_KAHAN, _lib.MFG_FP32_WIDE], ids=["kahan", "wide"])
def test_three_grains_reduced_fp32(hip_device, precision):
    """The same run with FP32 pair arithmetic: stress, plastic strain, temperature rise,
    position and velocity within 1e-4 relative L2 of the FP64 oracle (particles matched by id)."""
    sc = three_grain_scene()
    sim, parts = run_both(sc, 100, precision=precision)
    errs = errors_by_id(parts, sim, sc.params.trml_wp.T_init)
    print(f"three grains precision {precision}: (rel L2, max abs, |ref|)", errs)
    for f, (e, _, ref) in errs.items():
        assert ref > 0, f"field {f} must be active"
        assert e < 1e-4, (f, e)


def test_three_grains_full_size_precision_report(hip_device):
    """The precision form at full size (8M particles, three grains): 20 steps in FP64 and in both
    FP32 modes; stress, temperature rise, plastic strain (where it has started) and position of
    the FP32 runs within 1e-4 relative L2 of the FP64 run, matched by persistent id."""
    sc = scenes.Scene(3, 1.0, 1)
    print(f"three grains full size: {sc.n} particles", flush=True)
    out = {}
    for precision in (_lib.MFG_FP64, _lib.MFG_FP32_KAHAN, _lib.MFG_FP32_WIDE):
        parts, grid = scenes.setup_cut(sc, precision=precision)
        api.global_time_current = sc.params.dt
        lf = api.leap_frog(sc.n, grid.num_cell())
        scenes.run(lf, parts, grid, 20)
        order = np.argsort(parts.download(_lib.F_ID).astype(np.int64))
        out[precision] = {f: parts.download(f)[order] for f in (_lib.F_S, _lib.F_T, _lib.F_EPS_PL, _lib.F_POS)}
        del parts, grid, lf
    T0 = sc.params.trml_wp.T_init
    ref = out[_lib.MFG_FP64]
    for precision in (_lib.MFG_FP32_KAHAN, _lib.MFG_FP32_WIDE):
        for f, a in out[precision].items():
            g, o = a, ref[f]
            if f == _lib.F_T:
                g, o = g - T0, o - T0
            if np.linalg.norm(o) == 0.0:
                assert np.linalg.norm(g) == 0.0, f
                continue
            e = rel_err(g, o)
            print(f"8M three grains, precision {precision}, field {f}: rel L2 {e:.3e}, max abs {np.abs(g - o).max():.3e}")
            assert e < 1e-4, (precision, f, e)


def test_batched_full_configuration(hip_device):
    """The batched form at full size: 256 instances of 5k-200k particles, FP32 + Kahan, 300 steps
    each, one launch per kernel over all of them.  Every instance ends with finite fields and its
    particle count; three sampled instances are bit-identical when they run alone (their own
    single-instance steps)."""
    reqs = serve.sweep_requests(count=256, batched=True, n_min=5e3, n_max=2e5, steps=(300, 300), seed=1,
                                fp32_share=1.0)
    sample = {0, 100, 255}
    FIELDS = [_lib.F_POS, _lib.F_VEL, _lib.F_S, _lib.F_T, _lib.F_EPS_PL]
    seen, mix = {}, {}

    def on_done(inst):
        n, n_init = inst.particles.counts()
        seen[inst.req.rid] = (n, n_init)
        if inst.req.rid in sample:
            order = np.argsort(inst.particles.download(_lib.F_ID).astype(np.int64))
            mix[inst.req.rid] = {f: inst.particles.download(f)[order] for f in FIELDS}
        else:
            for f in (_lib.F_POS, _lib.F_VEL):
                assert np.isfinite(inst.particles.download(f)).all(), (inst.req.rid, f)

    res = serve.serve(reqs, on_done=on_done)
    print(f"batched 256: {res['particle_steps']:.3e} particle-steps in {res['makespan_s']:.2f} s, "
          f"{res['particle_steps_per_s']:.3e}/s, p50 {res['p50_s']:.2f} s, p99 {res['p99_s']:.2f} s")
    assert len(seen) == 256 and all("error" not in r for r in res["requests"])
    assert all(0 < n <= n_init for n, n_init in seen.values())
    for rid in sorted(sample):
        alone = {}

        def grab(inst):
            order = np.argsort(inst.particles.download(_lib.F_ID).astype(np.int64))
            alone[inst.req.rid] = {f: inst.particles.download(f)[order] for f in FIELDS}

        serve.serve([reqs[rid]], on_done=grab, mode="streams")
        for f in FIELDS:
            assert np.isfinite(mix[rid][f]).all()
            np.testing.assert_array_equal(mix[rid][f], alone[rid][f], err_msg=f"request {rid} field {f}")


def test_served_stream_fp64_request_matches_the_oracle(hip_device):
    """A short served stream (Poisson arrivals at 4/s for 2 s, 20k-30k particles, 100-150 steps,
    70 % FP32 + Kahan): every request finishes, and the FP64 request's result after its last
    step matches the oracle run of the same cut within 1e-9."""
    reqs = serve.sweep_requests(rate=4.0, duration=2.0, seed=5, n_min=2e4, n_max=3e4, steps=(100, 150))
    fp64 = [r for r in reqs if r.precision == _lib.MFG_FP64]
    if not fp64:
        reqs[0].precision = _lib.MFG_FP64
        fp64 = [reqs[0]]
    target = min(fp64, key=lambda r: r.n_target * r.steps).rid
    got = {}

    def on_done(inst):
        if inst.req.rid == target:
            got["scene"] = inst.scene
            got["fields"] = {f: inst.particles.download(f) for f in (_lib.F_ID, _lib.F_POS, _lib.F_VEL, _lib.F_S,
                                                                     _lib.F_T, _lib.F_RHO, _lib.F_EPS_PL)}
            got["steps"] = inst.req.steps

    res = serve.serve(reqs, on_done=on_done)
    print(f"served {len(reqs)} requests: p50 {res['p50_s']:.2f} s, p99 {res['p99_s']:.2f} s, "
          f"{res['particle_steps_per_s']:.3e} particle-steps/s")
    assert all("time_to_result" in r for r in res["requests"])
    sc = got["scene"]
    sim = helpers.oracle_from_scene(sc)
    dt = sc.params.dt
    sim.set_time(dt, dt)
    assert sim.run(got["steps"]) == 0
    assert np.array_equal(got["fields"][_lib.F_ID].astype(np.int64), sim.field(O.F_ID).astype(np.int64))
    for f, of in [(_lib.F_POS, O.F_POS), (_lib.F_VEL, O.F_VEL), (_lib.F_S, O.F_S), (_lib.F_T, O.F_T),
                  (_lib.F_RHO, O.F_RHO), (_lib.F_EPS_PL, O.F_EPS_PL)]:
        o = sim.field(of)
        if f == _lib.F_S:
            o = 0.5 * (o + o.reshape(-1, 3, 3).transpose(0, 2, 1).reshape(-1, 9))
        e = rel_err(got["fields"][f], o)
        assert e < 1e-9, (f, e)


def test_cpp_driver_matches_the_python_path(hip_device):
    """The reference-shaped C++ driver (mfree_gpu_3d.cpp's main over the C++ shim) runs the wedge
    cut at half size for 20 steps in its own process; its particle count and its tool-force
    report of step 20 equal the same run through the Python mirror of the call surface, and the
    driver's run with every step composed call by call from the shim's actions, pair passes and
    contact calls reports the same."""
    exe = os.path.join(ROOT, "iwf_mfree_gpu_3d_amd", "mfree_gpu_3d")
    assert os.path.exists(exe), "the driver is built with the library"
    r = subprocess.run([exe, "1", "0.5", "20", "0", "1"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr
    line = [ln for ln in r.stdout.splitlines() if ln.startswith("000020 of 000020")][0]
    n_drv = int(re.search(r"N (\d+)", line).group(1))
    f_drv = np.array([float(x) for x in re.search(r"F0 (\S+) (\S+) (\S+)", line).groups()], dtype=np.float32)
    sc = scenes.Scene(1, 0.5, 1)
    parts, grid = scenes.setup_cut(sc)
    api.global_time_current = sc.params.dt
    lf = api.leap_frog(sc.n, grid.num_cell())
    scenes.run(lf, parts, grid, 19)
    scenes.run(lf, parts, grid, 1, record_forces=True)
    assert parts.N == n_drv
    np.testing.assert_array_equal(api.global_tool_forces.get_forces_host()[0], f_drv)
    # the same run with every step composed from the per-function calls through the C++ shim
    r2 = subprocess.run([exe, "1", "0.5", "20", "0", "1", "1"], capture_output=True, text=True, timeout=120)
    assert r2.returncode == 0, r2.stderr
    line2 = [ln for ln in r2.stdout.splitlines() if ln.startswith("000020 of 000020")][0]
    assert re.search(r"N (\d+)", line2).group(1) == str(n_drv)
    assert re.search(r"F0 (\S+) (\S+) (\S+)", line2).groups() == re.search(r"F0 (\S+) (\S+) (\S+)", line).groups()


@pytest.mark.parametrize("mesh,realscale,feed,steps", [("pin25_reduziert_20062018", False, 4, 300),
                                                        ("diamant_mansur_04072018", True, 3, 150)],
                         ids=["single_grain_1tool_pin", "realscale_diamond"])
def test_reference_single_grain_runs(hip_device, mesh, realscale, feed, steps):
    """The reference's single-grain set-ups with its shipped meshes (read into the fixtures under
    tests/golden by make_mesh_fixtures.py): setup_single_grain_1tool with the pin mesh and
    setup_single_grain_1tool_realscale with the diamond, at 4 and 3 particle layers per feed, the
    tool started where its lower front meets the block.  Contact, plasticity and heating engage;
    every field within 1e-9 of the oracle, tool forces within float rounding."""
    d = np.load(os.path.join(GOLDEN, mesh + ".npz"))
    sc0 = scenes.single_grain_1tool(d["verts"], d["tris"], feed_particles=feed, realscale=realscale)
    sc = scenes.single_grain_1tool(d["verts"], d["tris"], feed_particles=feed, realscale=realscale,
                                   advance=scenes.engagement_advance(sc0))
    sim, parts = run_both(sc, steps)
    assert parts.N == sim.counts()[0]
    assert sim.field(O.F_EPS_PL).max() > 0.0 and sim.field(O.F_T).max() > 300.0
    errs = compare_fields(parts, sim, 1e-9)
    print(f"{mesh} relative errors", errs)
    f_g = api.global_tool_forces.get_forces_host()
    f_o = sim.forces()
    assert np.abs(f_o).max() > 0.0
    np.testing.assert_allclose(f_g, f_o, rtol=1e-4, atol=1e-3 * np.abs(f_o).max())
