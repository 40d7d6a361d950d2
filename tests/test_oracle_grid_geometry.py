"""Pins of the oracle's grid, neighbour search, contact geometry and step invariants against
brute force and textbook results.

- hashes, the stable sort and the Green cell tables against a numpy brute force on a small
  lattice with particles on cell faces, just below bbmin (truncation toward zero) and blanked
  (src/grid_gpu_base.cu:41-63,141-164, src/grid_gpu_green.cu:265-306)
- neighbour lists against an O(N^2) search with the same predicate (src/interactions_gpu.cu:
  298-336, src/kernels.cuh:41)
- closest point, Moller-Trumbore and the triangle/box overlap against textbook constructions
  (src/tool_3d_gpu.cu:101-280, src/surface_tri_mesh.cpp:151-278)
- linear momentum: without CSPM, contact or boundary conditions the pair terms are antisymmetric,
  so one step leaves the total momentum unchanged to round-off
"""
import numpy as np
import pytest

import oracle as O

INT_MAX = 2 ** 31 - 1


def _lattice(nx, ny, nz, dx, jitter, seed):
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    return g * dx + rng.uniform(-jitter, jitter, size=g.shape) * dx


def _sim_with_grid(pos, h, bbmin, bbmax, blank_box=None):
    n = pos.shape[0]
    sim = O.OracleSim(pos, h=np.full(n, h), rho=np.full(n, 4.43))
    sim.set_grid_locked(np.asarray(bbmin, float), np.asarray(bbmax, float), h)
    if blank_box is not None:
        sim.set_blanking(np.asarray(blank_box[0], float), np.asarray(blank_box[1], float))
    return sim


def test_hash_sort_cells_brute_force():
    dx = 1e-3
    h = 1.5 * dx
    pos = _lattice(9, 7, 5, dx, 0.3, 7)
    bbmin = np.array([0.0, 0.0, 0.0])
    bbmax = pos.max(0) + 2 * dx
    sim = _sim_with_grid(pos, h, bbmin, bbmax)
    g = sim.grid()
    gmin, gdx = g["bbmin"], g["dx"]
    # particles exactly on cell faces and slightly below bbmin (-> cell 0 by truncation)
    pos[0] = gmin + np.array([2, 3, 1]) * gdx
    pos[1] = gmin + np.array([0.0, 1.0, 2.0]) * gdx
    pos[2] = gmin - np.array([0.4, 0.1, 0.9]) * gdx
    pos[3] = gmin + np.array([-0.999, 2.5, -0.5]) * gdx
    # a blanking box that leaves the top x-slab of particles out
    xcut = np.sort(pos[:, 0])[-40]
    sim = _sim_with_grid(pos, h, bbmin, bbmax, blank_box=(bbmin - 1.0, [xcut, 1.0, 1.0]))
    assert sim.sort_only() == 0
    g = sim.grid()
    ny, nz = g["ny"], g["nz"]
    # brute force: same expression, truncation toward zero, INT_MAX for blanked particles
    blanked = pos[:, 0] > xcut
    q = (pos - g["bbmin"]) / g["dx"]
    ijk = np.trunc(q).astype(np.int64)
    key = ijk[:, 0] * ny * nz + ijk[:, 1] * nz + ijk[:, 2]
    key[blanked] = INT_MAX
    order = np.argsort(key, kind="stable")
    n_active = int((~blanked).sum())
    N, N_init = sim.counts()
    assert (N, N_init) == (n_active, pos.shape[0])
    np.testing.assert_array_equal(sim.field(O.F_ID).astype(np.int64), order)
    np.testing.assert_array_equal(sim.field(O.F_HASH).astype(np.int64), key[order])
    np.testing.assert_array_equal(sim.field(O.F_POS), pos[order])
    start, end = sim.cells()
    ks = key[order][:n_active]
    exp_s = np.full(g["num_cell"], -1)
    exp_e = np.full(g["num_cell"], -1)
    for c in np.unique(ks):
        idx = np.nonzero(ks == c)[0]
        exp_s[c], exp_e[c] = idx[0], idx[-1] + 1
    np.testing.assert_array_equal(start, exp_s)
    np.testing.assert_array_equal(end, exp_e)
    # the face / below-bbmin particles landed in the cells the truncation rule gives
    assert tuple(ijk[2]) == (0, 0, 0) and tuple(ijk[3]) == (0, 2, 0) and tuple(ijk[0]) == (2, 3, 1)


def test_neighbour_lists_brute_force():
    dx = 1e-3
    h = 1.5 * dx
    pos = _lattice(8, 6, 6, dx, 0.2, 3)
    bbmin = np.array([-dx, -dx, -dx])
    bbmax = pos.max(0) + dx
    xcut = np.sort(pos[:, 0])[-30]
    sim = _sim_with_grid(pos, h, bbmin, bbmax, blank_box=(bbmin - 1.0, [xcut, 1.0, 1.0]))
    assert sim.sort_only() == 0
    N, _ = sim.counts()
    counts, lists = sim.neighbours(cap=256)
    P = sim.field(O.F_POS)[:N]
    Hs = sim.field(O.F_H)[:N]
    B = sim.field(O.F_BLANKED)[:N]
    assert N < pos.shape[0] and np.all(B == 0)
    for i in range(N):
        d = P[i] - P
        rr2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        ok = (rr2 < 4 * Hs[i] * Hs[i]) & (rr2 >= 1e-8) & (B == 0)
        # the stencil visits cells in ascending (ii, jj, kk), i.e. ascending hash, i.e. ascending
        # sorted index: the reference order is the ascending list of qualifying indices
        expect = np.nonzero(ok)[0]
        assert counts[i] == expect.size
        np.testing.assert_array_equal(lists[i, :counts[i]], expect)


# ---- contact geometry ------------------------------------------------------------------------
def _tri(p1, p2, p3):
    n = np.cross(np.subtract(p2, p1), np.subtract(p3, p1))
    n = n / np.linalg.norm(n)
    ne = [np.array([0.0, 0.0, 1.0]) * (k + 2) for k in range(3)]  # distinct markers
    npv = [np.array([1.0, 0.0, 0.0]) * (k + 2) for k in range(3)]
    return O.tri30(p1, p2, p3, n, npv[0], npv[1], npv[2], ne[0], ne[1], ne[2]), n


def _closest_exact(p1, p2, p3, q):
    """Closest point on an acute triangle: the projection if it lies inside, else the closest of
    the three edge projections (Ericson, Real-Time Collision Detection, 5.1.5)."""
    a, b, c = map(np.asarray, (p1, p2, p3))
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n)
    pr = q - np.dot(q - a, n) * n
    M = np.stack([b - a, c - a], 1)
    uv = np.linalg.lstsq(M, pr - a, rcond=None)[0]
    if uv[0] >= 0 and uv[1] >= 0 and uv.sum() <= 1:
        return pr
    best = None
    for s, e in ((a, b), (b, c), (c, a)):
        t = np.clip(np.dot(q - s, e - s) / np.dot(e - s, e - s), 0, 1)
        cpt = s + t * (e - s)
        if best is None or np.sum((cpt - q) ** 2) < np.sum((best - q) ** 2):
            best = cpt
    return best


def test_closest_point_textbook():
    rng = np.random.default_rng(5)
    p1, p2, p3 = [0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.4, 0.9, 0.1]  # all angles acute
    t30, n = _tri(p1, p2, p3)
    for _ in range(400):
        q = rng.uniform(-0.8, 1.8, 3)
        cp, nrm = O.closest_point(t30, q)
        np.testing.assert_allclose(cp, _closest_exact(p1, p2, p3, q), atol=1e-12)
    # interior projection returns the face normal; vertex regions the vertex normal; edges the
    # edge normal (p1p2 -> ne1, p2p3 -> ne2, p3p1 -> ne3)
    cp, nrm = O.closest_point(t30, np.array([0.45, 0.3, 0.5]))
    np.testing.assert_allclose(nrm, n)
    cp, nrm = O.closest_point(t30, np.array([-0.5, -0.5, 0.0]))
    np.testing.assert_array_equal(cp, p1)
    np.testing.assert_array_equal(nrm, [2.0, 0.0, 0.0])
    cp, nrm = O.closest_point(t30, np.array([0.5, -0.5, 0.0]))
    np.testing.assert_array_equal(nrm, [0.0, 0.0, 2.0])
    assert abs(cp[2]) < 1e-15 and 0 < cp[0] < 1


def test_moller_trumbore_textbook():
    t30, _ = _tri([0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0])
    hit, t = O.ray_intersect(t30, np.array([0.2, 0.3, -2.0]), np.array([0.0, 0.0, 1.0]))
    assert hit and t == pytest.approx(2.0, rel=1e-15)
    hit, t = O.ray_intersect(t30, np.array([0.2, 0.3, -2.0]), np.array([0.0, 0.0, 0.5]))
    assert hit and t == pytest.approx(4.0, rel=1e-15)               # t in units of |dir|
    assert not O.ray_intersect(t30, np.array([0.8, 0.8, -1.0]), np.array([0.0, 0.0, 1.0]))[0]
    assert not O.ray_intersect(t30, np.array([0.2, 0.3, 1.0]), np.array([0.0, 0.0, 1.0]))[0]  # behind
    assert not O.ray_intersect(t30, np.array([0.2, 0.3, 0.0]), np.array([1.0, 0.0, 0.0]))[0]  # parallel
    # the absolute determinant threshold of 1e-7 (cm^2) rejects a small but valid triangle
    tiny, _ = _tri([0.0, 0.0, 0.0], [2e-4, 0.0, 0.0], [0.0, 2e-4, 0.0])
    assert not O.ray_intersect(tiny, np.array([5e-5, 5e-5, -1.0]), np.array([0.0, 0.0, 1.0]))[0]


def test_tri_aabb_overlap_textbook():
    t30, _ = _tri([0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0])
    assert O.tri_aabb(t30, np.array([0.1, 0.1, -0.1]), np.array([0.2, 0.2, 0.1]))      # box on the face
    assert not O.tri_aabb(t30, np.array([0.1, 0.1, 0.1]), np.array([0.2, 0.2, 0.2]))   # above the plane
    assert not O.tri_aabb(t30, np.array([0.6, 0.6, -0.1]), np.array([0.9, 0.9, 0.1]))  # past the hypotenuse
    assert O.tri_aabb(t30, np.array([0.45, 0.45, -0.1]), np.array([0.9, 0.9, 0.1]))   # straddles it
    assert O.tri_aabb(t30, np.array([-1.0, -1.0, -1.0]), np.array([2.0, 2.0, 2.0]))   # contains it
    # consistency with sampling on random boxes: a sampled triangle point inside the box implies
    # overlap
    rng = np.random.default_rng(9)
    uv = rng.uniform(0, 1, (4000, 2))
    uv = uv[uv.sum(1) <= 1]
    pts = np.stack([uv[:, 0], uv[:, 1], np.zeros(len(uv))], 1)
    for _ in range(200):
        lo = rng.uniform(-0.5, 1.0, 3)
        hi = lo + rng.uniform(0.05, 0.6, 3)
        inside = np.all((pts >= lo) & (pts <= hi), 1).any()
        if inside:
            assert O.tri_aabb(t30, lo, hi)


# ---- whole-step invariant ----------------------------------------------------------------------
def test_step_conserves_linear_momentum():
    """No CSPM, no contact, no fixed particles, equal h: every pair term of the momentum equation
    (stress, artificial stress, viscosity) is antisymmetric in i <-> j, so sum_i v_t = 0 and one
    step conserves the total momentum to round-off."""
    dx = 1e-3
    pos = _lattice(8, 7, 6, dx, 0.05, 11)
    n = pos.shape[0]
    rng = np.random.default_rng(2)
    vel = rng.normal(size=(n, 3)) * 1e-4
    sim = O.OracleSim(pos, vel3=vel, rho=np.full(n, 4.43), T=np.full(n, 300.0), h=np.full(n, 1.5 * dx))
    sim.set_options(cspm=False, heat=O.HEAT_NONE)
    E, nu, rho0 = 1.1, 0.35, 4.43
    K, G = E / (3 * (1 - 2 * nu)), E / (2 * (1 + nu))
    phys = O.Phys(E, nu, rho0, K, G, rho0 * dx ** 3)
    wdp = O.cubic_spline(np.zeros(3), np.array([dx, 0.0, 0.0]), 1.5 * dx, False)[0]
    corr = O.Corr(wdp, 0.3, 0.5, 1.0, 1.0, 0.1)
    sim.setup_constants(phys, corr, None, None)
    sim.set_grid_locked(pos.min(0) - 5 * dx, pos.max(0) + 5 * dx, 1.5 * dx)
    sim.set_time(1e-4, 1e-4)
    p0 = sim.field(O.F_VEL).sum(0)
    assert sim.step() == 0
    vt = sim.field(O.F_VEL_T)
    scale = np.abs(vt).sum(0)
    assert np.all(np.abs(vt.sum(0)) <= 1e-12 * scale)
    p1 = sim.field(O.F_VEL).sum(0)
    np.testing.assert_allclose(p1, p0, rtol=0, atol=1e-12 * np.abs(vel).sum())
    assert np.abs(vt).max() > 0   # the step did produce forces
