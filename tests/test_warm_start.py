"""Warm starts: particle positions from a VTK file, steady-state temperature samples, and the SPH
interpolation that carries a steady temperature field onto new tool particles.

CPU: the readers (round trip through the VTK writer, the reference's line rules) and the oracle's
interpolation pinned by closed forms.  GPU: the HIP interpolation against the oracle, bit-exact.
"""
import numpy as np
import pytest

import oracle as O
from iwf_mfree_gpu_3d_amd import api, scenes, vtk_writer
from iwf_mfree_gpu_3d_amd._lib import MfgError


def test_vtk_positions_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    n = 50
    pos = rng.uniform(-0.01, 0.02, (n, 3))
    blanked = np.zeros(n)
    blanked[[3, 17]] = 1.0
    z = np.zeros(n)
    path = str(tmp_path / "p.vtk")
    written = vtk_writer.write_particle_arrays(path, pos, z + 4.43, np.zeros((n, 3)), z + 300.0, z, z, z,
                                               np.zeros((n, 9)), z, blanked)
    got = api.read_vtk_positions(path)
    keep = blanked != 1.0
    want = np.array([[float("%f" % v) for v in row] for row in pos[keep]])
    assert written == got.shape[0] == keep.sum()
    np.testing.assert_array_equal(got, want)


def test_vtk_reader_line_rules(tmp_path):
    """vtk_read_init_pos (src/vtk_reader.cpp:29-69): after the POINTS line, lines of exactly three
    single-space-separated fields, up to the CELLS line; a line of one space is skipped."""
    text = "\n".join([
        "# vtk DataFile Version 2.0", "x", "ASCII", "", "DATASET UNSTRUCTURED_GRID",
        "POINTS 5 float",
        "1 2 3",
        "1  2 3",          # two spaces: four fields, skipped
        " ",
        "4 5 6 7",         # four fields, skipped
        "7.5 8 -9e-3",
        "CELLS 2 4",
        "10 11 12",        # after CELLS: not read
    ]) + "\n"
    path = tmp_path / "odd.vtk"
    path.write_text(text)
    got = api.read_vtk_positions(str(path))
    np.testing.assert_array_equal(got, [[1, 2, 3], [7.5, 8, -9e-3]])


def test_steady_reader_line_rules(tmp_path):
    text = "0 0 0 300\n1 2 3\n1  0 0 400\n1 0 0 500 \n-1.5e-3 2 3 301.25\n"
    path = tmp_path / "steady.txt"
    path.write_text(text)
    got = api.read_steady_temperatures(str(path))
    np.testing.assert_array_equal(got, [[0, 0, 0, 300], [1, 0, 0, 500], [-1.5e-3, 2, 3, 301.25]])


def test_reader_errors(tmp_path):
    with pytest.raises(MfgError):
        api.read_vtk_positions(str(tmp_path / "missing.vtk"))
    bad = tmp_path / "bad.vtk"
    bad.write_text("POINTS 1 float\na b c\nCELLS 0 0\n")
    with pytest.raises(MfgError):
        api.read_vtk_positions(str(bad))


def spline_value(r, h):
    """M4 cubic spline value in 3-D, 1/(pi h^3) normalisation, r rounded through float32 as the
    reference's sqrtf site does."""
    r = float(np.sqrt(np.float32(r * r)))
    q = r / h
    if r * r < 1e-8 or q >= 2.0:
        return 0.0
    if q > 1.0:
        return 0.25 * (2 - q) ** 3 / (np.pi * h ** 3)
    return (1 - 1.5 * q * q + 0.75 * q ** 3) / (np.pi * h ** 3)


@pytest.mark.parametrize("r", [0.0, 5e-5, 0.3, 1.0, 1.2, 1.99, 2.0, 2.5])
def test_oracle_interp_single_sample_closed_form(r):
    h, v, wgt = 1.0, 2.5, 0.7
    got = O.interp_temps(np.array([[0.1, -0.2, 0.3, v]]), wgt, h, np.array([[0.1 + r, -0.2, 0.3]]))
    want = wgt * spline_value(r, h) * v
    assert got[0] == pytest.approx(want, rel=1e-12, abs=1e-300)


def test_oracle_interp_lattice_partition_of_unity():
    """A constant field on a cubic lattice (spacing dx, h = 1.5 dx, weights dx^3) interpolates to
    itself at interior points off the lattice, to the quadrature error of the normalised kernel.
    On a lattice point the coincident sample drops out (r^2 < 1e-8 gives W = 0), which leaves
    1 - W(0) dx^3 of the field."""
    dx = 1.0
    h = 1.5 * dx
    g = np.arange(-4, 5) * dx
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    samples = np.stack([X.ravel(), Y.ravel(), Z.ravel(), np.full(X.size, 7.0)], axis=1)
    pts = np.array([[0, 0, 0], [0.3, 0.1, -0.2], [0.5, 0.5, 0.5]]) * dx
    got = O.interp_temps(samples, dx ** 3, h, pts)
    w0 = 1.0 / (np.pi * h ** 3)
    np.testing.assert_allclose(got, [7.0 * (1 - w0 * dx ** 3), 7.0, 7.0], rtol=2e-2)
    # linear in the sample values and the weights
    s2 = samples.copy()
    s2[:, 3] *= 3.0
    np.testing.assert_allclose(O.interp_temps(s2, 2 * dx ** 3, 1.5 * dx, pts), 6.0 * got, rtol=1e-14)


@pytest.mark.gpu
def test_gpu_interp_matches_oracle(hip_device):
    rng = np.random.default_rng(9)
    samples = np.concatenate([rng.uniform(0, 0.02, (1500, 3)), rng.uniform(300, 900, (1500, 1))], axis=1)
    points = rng.uniform(-0.001, 0.021, (2100, 3))  # more than 8 tiles of points, partial last tile
    h = 1.5e-3
    got = api.interp_temps(samples, 1e-9, h, points)
    want = O.interp_temps(samples, 1e-9, h, points)
    assert np.count_nonzero(want) > 1000
    np.testing.assert_array_equal(got, want)


@pytest.mark.gpu
def test_gpu_steady_tool_temperatures(hip_device):
    """The single-grain warm start: a steady field sampled on a lattice, shifted onto the tool
    samples' box, interpolated and clamped at T_ref."""
    dx_s = 2e-4
    g = np.arange(10) * dx_s
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    T_s = 300.0 + 5e4 * X.ravel()  # linear steady field along x
    steady = np.stack([X.ravel() + 1.0, Y.ravel() - 2.0, Z.ravel() + 0.5, T_s], axis=1)
    rng = np.random.default_rng(5)
    samples = rng.uniform(0.0, 9 * dx_s, (400, 3)) + np.array([0.3, 0.2, 0.1])
    T = scenes.steady_tool_temperatures(samples, steady, dz=1.5e-4, hdx=1.7, T_ref=300.0)
    dx_s = np.max(np.abs(steady[0, :3] - steady[1, :3]))  # the reference's spacing estimate
    shifted = steady.copy()  # x - bbmin_steady + bbmin_tool, in the reference's order
    shifted[:, :3] = steady[:, :3] - steady[:, :3].min(axis=0) + samples.min(axis=0)
    want = np.maximum(O.interp_temps(shifted, dx_s ** 3, max(1.5 * dx_s, 1.7 * 1.5e-4), samples), 300.0)
    np.testing.assert_array_equal(T, want)
    assert (T >= 300.0).all() and T.max() > 320.0
