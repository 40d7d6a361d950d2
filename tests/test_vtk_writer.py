"""Legacy-VTK writer (src/vtk_writer.cpp:78-427): structure and values on host arrays."""
import numpy as np

from iwf_mfree_gpu_3d_amd import vtk_writer


def _parse(path):
    lines = open(path).read().split("\n")
    out, i = {}, 0
    while i < len(lines):
        ln = lines[i]
        if ln.startswith("POINTS"):
            n = int(ln.split()[1])
            out["points"] = np.array([[float(v) for v in lines[i + 1 + k].split()] for k in range(n)])
            i += n
        elif ln.startswith("CELLS"):
            n = int(ln.split()[1])
            out["cells"] = [list(map(int, lines[i + 1 + k].split())) for k in range(n)]
            i += n
        elif ln.startswith("SCALARS"):
            name = ln.split()[1]
            n = len(out["points"])
            out[name] = np.array([float(lines[i + 2 + k]) for k in range(n)])
            i += n + 1
        i += 1
    return out


def test_particle_file_skips_blanked_and_writes_fields(tmp_path):
    rng = np.random.default_rng(1)
    n = 12
    pos = rng.uniform(0, 1e-2, (n, 3))
    rho = rng.uniform(4.4, 4.5, n)
    fc = np.zeros((n, 3))
    fc[3] = [0.0, 1e-7, 0.0]
    T = rng.uniform(300, 400, n)
    fixed = (np.arange(n) % 5 == 0).astype(float)
    tool = (np.arange(n) == 7).astype(float)
    eps = rng.uniform(0, 0.5, n)
    S = np.zeros((n, 3, 3))
    S[:, 0, 0] = 0.01  # uniaxial deviatoric part, p = 0: von Mises = |sxx|
    S[2] = np.diag([0.0, 0.0, 0.0])
    S[2, 0, 1] = S[2, 1, 0] = 0.004  # pure shear: von Mises = sqrt(3) |sxy|
    p = np.zeros(n)
    blanked = np.zeros(n)
    blanked[[4, 9]] = 1.0
    path = str(tmp_path / "p.vtk")
    assert vtk_writer.write_particle_arrays(path, pos, rho, fc, T, fixed, tool, eps, S.reshape(n, 9), p, blanked) == 10
    d = _parse(path)
    keep = blanked != 1.0
    np.testing.assert_allclose(d["points"], pos[keep], atol=1e-6)
    assert [c[1] for c in d["cells"]] == list(np.nonzero(keep)[0])
    np.testing.assert_allclose(d["density"], rho[keep], atol=1e-6)
    np.testing.assert_array_equal(d["contact"], (np.arange(n) == 3)[keep].astype(float))
    np.testing.assert_allclose(d["Temperature"], T[keep], atol=1e-6)
    np.testing.assert_array_equal(d["Fixed"], fixed[keep])
    np.testing.assert_array_equal(d["Tool"], tool[keep])
    svm = np.where(np.arange(n) == 2, np.sqrt(3) * 0.004, 0.01)
    np.testing.assert_allclose(d["Svm"], svm[keep], atol=1e-6)


def test_tool_file_concatenates_meshes(tmp_path):
    v = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    t = np.array([[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]])
    path = str(tmp_path / "t.vtk")
    vtk_writer.write_tools(path, [(v, t), (v, t)], offsets=[np.zeros(3), np.array([2.0, 0, 0])])
    d = _parse(path)
    assert d["points"].shape == (8, 3) and len(d["cells"]) == 8
    np.testing.assert_allclose(d["points"][4:], v + [2.0, 0, 0])
    assert d["cells"][4] == [3, 4, 5, 6]
