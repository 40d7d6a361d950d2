"""Writes the tool meshes the reference ships (LS-DYNA keyword tetrahedral meshes under
Release/) as data fixtures: the boundary triangles and vertices as the library's reader
mfg_mesh_read_ldynak returns them (ldynak_read_triangles_from_tetmesh,
src/ldynak_reader.cpp:371-434), in the file's units.  Run once in a tree that has the reference
at /root/reference; the GPU tests read only the .npz files."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from iwf_mfree_gpu_3d_amd import api  # noqa: E402

REF = "/root/reference/Release"
MESHES = {"pin25_reduziert_20062018": "pin25_reduziert_20062018.k",
          "diamant_mansur_04072018": "Diamant_von_Mansur_04072018_SDB1125-2025-D851.k"}

if __name__ == "__main__":
    for name, fn in MESHES.items():
        verts, tris = api.read_mesh_ldynak(os.path.join(REF, fn))
        out = os.path.join(HERE, name + ".npz")
        np.savez(out, verts=verts, tris=tris.astype(np.int32))
        print(f"{out}: {verts.shape[0]} vertices, {tris.shape[0]} triangles")
