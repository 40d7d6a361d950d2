"""Legacy-VTK output of particles and tool meshes, in the layout of the reference's writer
(src/vtk_writer.cpp:78-427): ASCII unstructured grid, one vertex cell per unblanked particle, and
point data density, contact, Temperature, Fixed, Tool, EquivAccumPlasticStrain and Svm (von Mises
stress of S - pI); tools as triangle cells.  Host-side I/O only: field values are downloaded from
the device through the C ABI."""
from __future__ import annotations

import numpy as np

from . import _lib


def _scalars(fp, name: str, values) -> None:
    fp.write(f"SCALARS {name} float 1\n")
    fp.write("LOOKUP_TABLE default\n")
    for v in values:
        fp.write("%f\n" % v)
    fp.write("\n")


def von_mises(S9: np.ndarray, p: np.ndarray) -> np.ndarray:
    """sqrt of the von Mises invariant of S - p I per particle (src/vtk_writer.cpp:241-254); S9 is
    9 doubles per particle in GLM column-major order (symmetric, so the order does not matter)."""
    S = S9.reshape(-1, 3, 3)
    sxx, syy, szz = S[:, 0, 0] - p, S[:, 1, 1] - p, S[:, 2, 2] - p
    sxy, syz, sxz = S[:, 0, 1], S[:, 1, 2], S[:, 0, 2]
    svm2 = (sxx * sxx + syy * syy + szz * szz) - sxx * syy - sxx * szz - syy * szz \
        + 3.0 * (sxy * sxy + syz * syz + sxz * sxz)
    return np.where(svm2 > 0, np.sqrt(np.maximum(svm2, 0.0)), 0.0)


def write_particle_arrays(path: str, pos, rho, fc, T, fixed, tool, eps_pl, S9, p, blanked) -> int:
    """Write the particles whose blanked flag is not 1; returns the number written."""
    keep = np.asarray(blanked) != 1.0
    idx = np.nonzero(keep)[0]
    n = idx.size
    with open(path, "w") as fp:
        fp.write("# vtk DataFile Version 2.0\nmfree iwf\nASCII\n\n")
        fp.write("DATASET UNSTRUCTURED_GRID\n")
        fp.write(f"POINTS {n} float\n")
        for i in idx:
            fp.write("%f %f %f\n" % (pos[i, 0], pos[i, 1], pos[i, 2]))
        fp.write("\n")
        fp.write(f"CELLS {n} {2 * n}\n")
        for i in idx:  # the reference writes the particle's index in the current order
            fp.write(f"1 {i}\n")
        fp.write("\n")
        fp.write(f"CELL_TYPES {n}\n")
        fp.write("1\n" * n)
        fp.write("\n")
        fp.write(f"POINT_DATA {n}\n")
        _scalars(fp, "density", np.asarray(rho)[idx])
        contact = (np.linalg.norm(np.asarray(fc)[idx], axis=1) > 0.0).astype(float)
        _scalars(fp, "contact", contact)
        _scalars(fp, "Temperature", np.asarray(T)[idx])
        _scalars(fp, "Fixed", np.asarray(fixed)[idx])
        _scalars(fp, "Tool", np.asarray(tool)[idx])
        _scalars(fp, "EquivAccumPlasticStrain", np.asarray(eps_pl)[idx])
        _scalars(fp, "Svm", von_mises(np.asarray(S9)[idx], np.asarray(p)[idx]))
    return n


def write_particles(path: str, particles) -> int:
    """vtk_writer_write(particle_gpu*): the first N particles (those in cells) of the current
    sorted order, blanked ones skipped."""
    N = particles.N
    d = lambda f: particles.download(f)[:N]
    return write_particle_arrays(path, d(_lib.F_POS), d(_lib.F_RHO), d(_lib.F_FC), d(_lib.F_T),
                                 d(_lib.F_FIXED), d(_lib.F_TOOL), d(_lib.F_EPS_PL), d(_lib.F_S),
                                 d(_lib.F_P), d(_lib.F_BLANKED))


def write_tools(path: str, meshes, offsets=None) -> None:
    """vtk_writer_write_unified: all tool meshes in one file.  meshes = [(verts (n, 3), tris (m, 3))];
    offsets = per-tool translation (the tool's velocity times the current time) or None."""
    pts, cells = [], []
    base = 0
    for k, (verts, tris) in enumerate(meshes):
        v = np.asarray(verts, dtype=np.float64)
        if offsets is not None:
            v = v + np.asarray(offsets[k], dtype=np.float64)
        pts.append(v)
        cells.append(np.asarray(tris, dtype=np.int64) + base)
        base += v.shape[0]
    P = np.concatenate(pts) if pts else np.zeros((0, 3))
    T = np.concatenate(cells) if cells else np.zeros((0, 3), dtype=np.int64)
    with open(path, "w") as fp:
        fp.write("# vtk DataFile Version 2.0\nmfree iwf\nASCII\n\n")
        fp.write("DATASET UNSTRUCTURED_GRID\n")
        fp.write(f"POINTS {P.shape[0]} float\n")
        for q in P:
            fp.write("%f %f %f\n" % (q[0], q[1], q[2]))
        fp.write("\n")
        fp.write(f"CELLS {T.shape[0]} {4 * T.shape[0]}\n")
        for t in T:
            fp.write(f"3 {t[0]} {t[1]} {t[2]}\n")
        fp.write("\n")
        fp.write(f"CELL_TYPES {T.shape[0]}\n")
        fp.write("5\n" * T.shape[0])
