"""Tool mesh readers (src/ldynak_reader.cpp:308-434, src/json_reader.cpp:108-117).

The reference ships its own diamond and test meshes (Release/*.k, Release/*.json); they are read in
place when the reference tree is present (this checks the reader on real inputs) and the tests
are skipped otherwise.  Small synthetic files check the same rules everywhere.
"""
import collections
import os

import numpy as np
import pytest

from iwf_mfree_gpu_3d_amd import api

REF = "/root/reference/Release"


def closed_oriented(tris):
    """Every undirected edge in exactly two triangles, every directed edge exactly once."""
    und = collections.Counter()
    dire = collections.Counter()
    for a, b, c in tris:
        for u, v in ((a, b), (b, c), (c, a)):
            und[(min(u, v), max(u, v))] += 1
            dire[(u, v)] += 1
    return all(n == 2 for n in und.values()) and all(n == 1 for n in dire.values())


def signed_volume(verts, tris):
    a, b, c = verts[tris[:, 0]], verts[tris[:, 1]], verts[tris[:, 2]]
    return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)
    return str(path)


def test_ldynak_two_tets_synthetic(tmp_path):
    """Two tetrahedra sharing a face: the shared face is interior, 6 boundary faces remain, the
    surface is closed and consistently oriented, vertices ordered by node id."""
    text = """*KEYWORD
*NODE
$#   nid x y z
        7 0.0 0.0 0.0 0 0
        3 1.0 0.0 0.0 0 0
        5 0.0 1.0 0.0 0 0
        9 0.0 0.0 1.0 0 0
        2 0.3 0.3 -1.0 0 0
*ELEMENT_SOLID
       1       1       7       3       5       9       9       9       9       9
       2       1       3       7       5       2       2       2       2       2
*END
"""
    v, t = api.read_mesh_ldynak(_write(tmp_path / "two.k", text))
    assert v.shape == (5, 3) and t.shape == (6, 3)
    np.testing.assert_array_equal(v[:, 0], [0.3, 1.0, 0.0, 0.0, 0.0])  # node ids 2, 3, 5, 7, 9
    assert closed_oriented(t)
    assert abs(abs(signed_volume(v, t)) - (1 / 6 + 1 / 6)) < 1e-12


def test_ldynak_hex_synthetic(tmp_path):
    """One hexahedron: 6 boundary quads, each split into two triangles."""
    nodes = [(1, 0, 0, 0), (2, 1, 0, 0), (3, 1, 1, 0), (4, 0, 1, 0), (5, 0, 0, 1), (6, 1, 0, 1),
             (7, 1, 1, 1), (8, 0, 1, 1)]
    text = "*KEYWORD\n*NODE\n" + "".join(f"{n} {x} {y} {z} 0 0\n" for n, x, y, z in nodes)
    text += "*ELEMENT_SOLID\n1 1 1 2 3 4 5 6 7 8\n*END\n"
    v, t = api.read_mesh_ldynak(_write(tmp_path / "hex.k", text), quads=True)
    assert v.shape == (8, 3) and t.shape == (12, 3)
    assert closed_oriented(t)
    assert abs(abs(signed_volume(v, t)) - 1.0) < 1e-12


def test_json_synthetic(tmp_path):
    text = """{
  "vertices" :
  [
    {
      "values"     :
      [
        0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1
      ]
    }
  ],
  "connectivity" :
  [
    {
      "indices"     :
      [
        0, 2, 1, 0, 1, 3, 1, 2, 3, 0, 3, 2
      ]
    }
  ]
}
"""
    v, t = api.read_mesh_json(_write(tmp_path / "tet.json", text))
    assert v.shape == (4, 3) and t.shape == (4, 3)
    assert closed_oriented(t)
    assert signed_volume(v, t) == pytest.approx(1 / 6)


def test_reader_errors(tmp_path):
    with pytest.raises(api.MfgError):
        api.read_mesh_ldynak(str(tmp_path / "missing.k"))
    with pytest.raises(api.MfgError):
        api.read_mesh_json(_write(tmp_path / "empty.json", "{}\n"))


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")
@pytest.mark.parametrize("name", ["pin25_reduziert_20062018.k", "Diamant_von_Mansur_04072018_SDB1125-2025-D851.k"])
def test_reference_tet_meshes(name):
    """The shipped diamond tet meshes: the surface found equals the faces used by exactly one
    tetrahedron (counted independently here), and it is closed and consistently oriented."""
    path = os.path.join(REF, name)
    v, t = api.read_mesh_ldynak(path)
    # independent count of the boundary faces from the raw element lines
    elems, sec = [], None
    for line in open(path, errors="replace"):
        if "$" in line:
            continue
        if "*" in line:
            sec = "E" if "*ELEMENT" in line else None
            continue
        tk = line.split()
        if sec == "E" and len(tk) == 10:
            elems.append([int(x) for x in tk[2:6]])
    faces = collections.Counter()
    for e in elems:
        for f in ((0, 1, 3), (2, 1, 0), (1, 2, 3), (3, 2, 0)):
            faces[tuple(sorted(e[k] for k in f))] += 1
    assert t.shape[0] == sum(1 for n in faces.values() if n == 1) > 0
    assert closed_oriented(t)
    assert abs(signed_volume(v, t)) > 0


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")
@pytest.mark.parametrize("name", ["cube_rot.json", "sphere_coarse.json"])
def test_reference_json_meshes(name):
    v, t = api.read_mesh_json(os.path.join(REF, name))
    assert t.max() < v.shape[0] and t.shape[0] > 0
    assert closed_oriented(t)
    if name == "cube_rot.json":
        assert v.shape == (8, 3) and t.shape == (12, 3)
        d = np.linalg.norm(v[:, None, :] - v[None, :, :], axis=2)
        edge = d[d > 0].min()                     # a rotated cube: volume = edge^3
        assert abs(signed_volume(v, t)) == pytest.approx(edge ** 3, rel=1e-4)
