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


# ---- radial return, leap-frog, contact law, mesh normals ---------------------------------------
def _material(dx=1e-3):
    E, nu, rho0 = 1.1, 0.35, 4.43
    K, G = E / (3 * (1 - 2 * nu)), E / (2 * (1 + nu))
    phys = O.Phys(E, nu, rho0, K, G, rho0 * dx ** 3)
    wdp = O.cubic_spline(np.zeros(3), np.array([dx, 0.0, 0.0]), 1.5 * dx, False)[0]
    corr = O.Corr(wdp, 0.3, 0.5, 1.0, 1.0, 0.1)
    joco = O.Joco(0.01098, 0.01092, 0.014, 0.93, 1.1, 1878.0, 300.0, 1e-6, 1.0)
    trml = O.Trml(553e-8, 0.9, 7.1e-13, 7.1e-13 / (4.43 * 553e-8), 300.0, 0.0)
    return phys, corr, joco, trml


def test_radial_return_consistency():
    """After the return the deviatoric stress lies on the yield surface,
    |S_new| = sqrt(2/3) sigma_Y(eps_new, eps_dot_new, T) (up to the secant tolerance), it is parallel
    to the trial stress, eps_pl grows by sqrt(2/3) dlambda and T by tq/(cp rho) d_eps sigma_Y;
    below the yield surface nothing changes but eps_pl_dot is reset (src/actions_gpu.cu:371-449)."""
    rng = np.random.default_rng(4)
    n = 40
    pos = np.stack([np.arange(n) * 1.0, np.zeros(n), np.zeros(n)], 1)
    sim = O.OracleSim(pos, rho=np.full(n, 4.43), T=np.linspace(300.0, 900.0, n), h=np.full(n, 1.5e-3))
    sim.set_options(cspm=False, heat=O.HEAT_NONE, single_sqrt=False)
    phys, corr, joco, trml = _material()
    sim.setup_constants(phys, corr, joco, trml)
    dt = 1e-5
    sim.set_time(dt, dt)
    S = np.zeros((n, 3, 3))
    for i in range(n):
        a = rng.normal(size=(3, 3))
        a = a + a.T
        a -= np.trace(a) / 3 * np.eye(3)
        S[i] = a / np.linalg.norm(a) * (0.004 if i % 4 == 0 else 0.03)  # every 4th below yield
    eps0 = rng.uniform(0.05, 0.3, n)
    epsd0 = rng.uniform(0.0, 1e-2, n)
    sim.set_field(O.F_S, np.stack([O.glm_from_rowmajor(m) for m in S]))
    sim.set_field(O.F_EPS_PL, eps0)
    sim.set_field(O.F_EPS_PL_DOT, epsd0)
    T0 = sim.field(O.F_T).copy()
    sim.plasticity_only()
    Sn = np.stack([O.rowmajor_from_glm(m) for m in sim.field(O.F_S)])
    eps1, epsd1, T1 = sim.field(O.F_EPS_PL), sim.field(O.F_EPS_PL_DOT), sim.field(O.F_T)
    mu = phys.G
    plastic = 0
    for i in range(n):
        norm = np.sqrt(np.sum(S[i] * S[i]))
        sy = O.sigma_yield(joco, eps0[i], epsd0[i], min(T0[i], joco.Tmelt - 1))
        if np.sqrt(1.5) * norm < sy:
            np.testing.assert_array_equal(Sn[i], S[i])
            assert eps1[i] == eps0[i] and epsd1[i] == 0.0 and T1[i] == T0[i]
            continue
        plastic += 1
        dl = (eps1[i] - eps0[i]) / np.sqrt(2 / 3)
        assert dl > 0
        assert epsd1[i] == pytest.approx(np.sqrt(2 / 3) * dl / dt, rel=1e-12)
        np.testing.assert_allclose(Sn[i], S[i] * (1 - 2 * mu * dl / norm), rtol=1e-12, atol=1e-18)
        sy_new = O.sigma_yield(joco, eps1[i], epsd1[i], T0[i])
        assert np.sqrt(np.sum(Sn[i] * Sn[i])) == pytest.approx(np.sqrt(2 / 3) * sy_new, rel=1e-4)
        assert T1[i] - T0[i] == pytest.approx(trml.tq / (trml.cp * 4.43) * (eps1[i] - eps0[i]) * sy_new,
                                              rel=1e-12)
    assert plastic >= n // 2


def test_leap_frog_free_particles_move_linearly():
    """Particles farther apart than the kernel support feel no force: the midpoint predictor /
    corrector then moves them exactly along x0 + t v (src/leap_frog.cu:30-100)."""
    rng = np.random.default_rng(8)
    n = 27
    g = np.stack(np.meshgrid(np.arange(3), np.arange(3), np.arange(3), indexing="ij"), -1).reshape(-1, 3)
    pos = g * 1e-2 + 5e-2
    vel = rng.normal(size=(n, 3)) * 1e-4
    sim = O.OracleSim(pos, vel3=vel, rho=np.full(n, 4.43), T=np.full(n, 300.0), h=np.full(n, 1.5e-3))
    sim.set_options(cspm=True, heat=O.HEAT_BROOKSHAW)
    phys, corr, joco, trml = _material()
    sim.setup_constants(phys, corr, joco, trml)
    sim.set_grid_locked(np.zeros(3), np.full(3, 0.2), 1.5e-3)
    dt, steps = 2e-4, 25
    sim.set_time(dt, dt)
    assert sim.run(steps) == 0
    ids = sim.field(O.F_ID).astype(np.int64)
    np.testing.assert_allclose(sim.field(O.F_POS), (pos + steps * dt * vel)[ids], rtol=0, atol=1e-15)
    np.testing.assert_array_equal(sim.field(O.F_VEL), vel[ids])
    assert np.all(sim.field(O.F_S) == 0.0)


def _box_mesh(lo, hi):
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], float)
    # counter-clockwise seen from outside: outward face normals
    t = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                  [1, 2, 6], [1, 6, 5], [2, 3, 7], [2, 7, 6], [3, 0, 4], [3, 4, 7]], np.int32)
    return v, t


def test_mesh_normals_of_a_box():
    """Face normals are the outward axes; angle-weighted vertex normals at a box corner are the
    normalised sum of its three faces (each face contributes 90 degrees whatever the
    triangulation); edge normals are the normalised sum of the two adjacent triangles' normals
    (src/surface_tri_mesh.cpp:410-568)."""
    v, t = _box_mesh((0.0, 0.0, 0.0), (1.0, 2.0, 0.5))
    out = O.mesh_normals(v, t)
    c = v.mean(0)
    for k in range(t.shape[0]):
        P = out[k, 0:9].reshape(3, 3)
        n = out[k, 9:12]
        np.testing.assert_array_equal(P, v[t[k]])
        axis = np.zeros(3)
        j = int(np.argmax(np.abs(n)))
        axis[j] = np.sign(P.mean(0)[j] - c[j])
        np.testing.assert_allclose(n, axis, atol=1e-15)
        for q in range(3):
            np_q = out[k, 12 + 3 * q:15 + 3 * q]
            np.testing.assert_allclose(np_q, np.sign(P[q] - c) / np.sqrt(3), atol=1e-15)
        for q, (a, b) in enumerate(((0, 1), (1, 2), (2, 0))):
            ne = out[k, 21 + 3 * q:24 + 3 * q]
            mid = 0.5 * (P[a] + P[b])
            # an edge of the box: two faces; a face diagonal: one face
            on = [ax for ax in range(3) if np.isclose(mid[ax], v[:, ax].min()) or np.isclose(mid[ax], v[:, ax].max())]
            expect = np.zeros(3)
            for ax in on:
                expect[ax] = np.sign(mid[ax] - c[ax])
            expect /= np.linalg.norm(expect)
            np.testing.assert_allclose(ne, expect, atol=1e-15)


def test_contact_force_law_on_a_plane():
    """One particle a depth d below the top face of a resting box tool: the exhaustive search finds
    that face, and the penalty force is m d alpha / dt^2 along the outward normal, recorded in
    fc and in the tool force sum (src/tool_3d_gpu.cu:455-583); no friction with mu = 0."""
    phys, corr, joco, trml = _material()
    v, t = _box_mesh((0.0, 0.0, -1.0), (1.0, 1.0, 0.0))
    d = 2e-4
    pos = np.array([[0.31, 0.47, -d]])
    sim = O.OracleSim(pos, rho=np.full(1, 4.43), T=np.full(1, 300.0), h=np.full(1, 1.5e-3))
    sim.set_options(cspm=True, heat=O.HEAT_NONE)
    sim.setup_constants(phys, corr, joco, trml)
    gmin, gmax = np.array([0.25, 0.4, -0.05]), np.array([0.4, 0.55, 0.05])
    sim.set_grid_locked(gmin, gmax, 1.5e-3)
    alpha = 0.1
    sim.add_tool(v, t, vel=np.zeros(3), mass=phys.mass, alpha=alpha, algo=0, mu=0.0)
    dt = 1e-5
    sim.set_time(dt, dt)
    assert sim.step(record=True) == 0
    fc = sim.field(O.F_FC)[0]
    expect = phys.mass * d * alpha / dt ** 2
    np.testing.assert_allclose(fc, [0.0, 0.0, expect], rtol=1e-12, atol=1e-12 * expect)
    np.testing.assert_allclose(sim.forces()[0], [0.0, 0.0, expect], rtol=1e-6)
    # a particle outside the box feels nothing
    sim2 = O.OracleSim(np.array([[0.31, 0.47, +d]]), rho=np.full(1, 4.43), T=np.full(1, 300.0),
                       h=np.full(1, 1.5e-3))
    sim2.set_options(cspm=True, heat=O.HEAT_NONE)
    sim2.setup_constants(phys, corr, joco, trml)
    sim2.set_grid_locked(gmin, gmax, 1.5e-3)
    sim2.add_tool(v, t, vel=np.zeros(3), mass=phys.mass, alpha=alpha, algo=0, mu=0.0)
    sim2.set_time(dt, dt)
    assert sim2.step(record=True) == 0
    assert np.all(sim2.field(O.F_FC) == 0.0)


@pytest.mark.parametrize("vel", [(10.0, 5.0, 3.0), (2.0, 1.0, 0.5)], ids=["slip", "stick"])
def test_friction_heat_closed_form(vel):
    """One particle a depth d under the top face of a resting box tool, moving at v, mu = 0.3.
    The tool's normals point into the tool (src/tool_3d.cpp:294-360), so the contact normal is
    n = (0, 0, -1).
    - Coulomb friction (src/tool_3d_gpu.cu:455-583; the trial force is reset each step):
      v_r = v - v * n with GLM's component-wise product, = (vx, vy, 2 vz) here; the tangential
      force is -alpha m v_r / dt (stick) or mu |f_N| along it (slip, |v_r| > mu d / dt).
    - Friction heat (src/actions_gpu.cu:192-220) raises T by eta dt |f_T| |v_h| / (cp m), where
      v_h = v - (v . n) (1, 1, 1) subtracts the scalar v . n from every component:
      (vx + vz, vy + vz, 2 vz) here.
    Both differ from the tangential velocity (vx, vy, 0); the test fails if either is replaced
    by it."""
    phys, corr, joco, trml = _material()
    trml = O.Trml(trml.cp, trml.tq, trml.k, trml.alpha, trml.T_init, 0.9)   # eta = 0.9
    v, t = _box_mesh((0.0, 0.0, -1.0), (1.0, 1.0, 0.0))
    d, dt, alpha, mu = 2e-4, 1e-5, 0.1, 0.3
    vx, vy, vz = vel
    sim = O.OracleSim(np.array([[0.31, 0.47, -d]]), vel3=np.array([vel]), rho=np.full(1, 4.43),
                      T=np.full(1, 300.0), h=np.full(1, 1.5e-3))
    sim.set_options(cspm=True, heat=O.HEAT_NONE)
    sim.setup_constants(phys, corr, joco, trml)
    sim.set_grid_locked(np.array([0.25, 0.4, -0.05]), np.array([0.4, 0.55, 0.05]), 1.5e-3)
    sim.add_tool(v, t, vel=np.zeros(3), mass=phys.mass, alpha=alpha, algo=0, mu=mu)
    sim.set_time(dt, dt)
    assert sim.step(record=True) == 0
    np.testing.assert_array_equal(sim.field(O.F_N)[0], [0.0, 0.0, -1.0])
    m = phys.mass
    depth = d - vz * dt                        # the particle moved at v during the step
    f_n = m * depth * alpha / dt ** 2
    np.testing.assert_allclose(sim.field(O.F_FC)[0], [0.0, 0.0, f_n], rtol=1e-9)
    v_r = np.array([vx, vy, 2.0 * vz])
    slip = np.linalg.norm(v_r) > mu * depth / dt
    assert slip == (vx == 10.0)
    f_t = mu * f_n if slip else alpha * m * np.linalg.norm(v_r) / dt
    np.testing.assert_allclose(sim.field(O.F_FT)[0], -f_t * v_r / np.linalg.norm(v_r), rtol=1e-6)
    v_h = np.linalg.norm([vx + vz, vy + vz, 2.0 * vz])
    dT = 0.9 * dt * f_t * v_h / (trml.cp * m)
    np.testing.assert_allclose(sim.field(O.F_T)[0] - 300.0, dT, rtol=1e-6)
    # the tangential speed would give a different rise and force
    assert abs(np.hypot(vx, vy) / v_h - 1.0) > 0.1


def test_boundary_conditions_fixed_particles():
    """Fixed particles end the step with the prescribed velocity, zero contact force and zero
    position / velocity rates; with the thermal boundary condition their temperature is reset to
    the workpiece T_init (src/actions_gpu.cu:451-487)."""
    dx = 1e-3
    pos = _lattice(6, 6, 4, dx, 0.0, 1)
    n = pos.shape[0]
    rng = np.random.default_rng(6)
    vel = rng.normal(size=(n, 3)) * 1e-4
    fixed = (pos[:, 2] < 0.5 * dx).astype(float)     # bottom layer
    T = 300.0 + rng.uniform(0, 50, n)
    sim = O.OracleSim(pos, vel3=vel, rho=np.full(n, 4.43), T=T, h=np.full(n, 1.5 * dx), fixed=fixed)
    sim.set_options(cspm=True, heat=O.HEAT_BROOKSHAW)
    phys, corr, joco, trml = _material(dx)
    sim.setup_constants(phys, corr, joco, trml)
    sim.set_grid_locked(pos.min(0) - 3 * dx, pos.max(0) + 3 * dx, 1.5 * dx)
    sim.set_time(1e-5, 1e-5)
    assert sim.step() == 0
    f = sim.field(O.F_FIXED) == 1.0
    assert f.sum() == (fixed == 1).sum() > 0
    np.testing.assert_array_equal(sim.field(O.F_VEL)[f], sim.field(O.F_VEL_BC)[f])
    assert np.all(sim.field(O.F_FC)[f] == 0.0)
    assert np.all(sim.field(O.F_POS_T)[f] == 0.0) and np.all(sim.field(O.F_VEL_T)[f] == 0.0)
    np.testing.assert_array_equal(sim.field(O.F_T)[f], trml.T_init)
    # free particles moved and kept their own temperature evolution
    assert np.any(sim.field(O.F_VEL)[~f] != vel[sim.field(O.F_ID).astype(int)][~f])
    assert np.all(sim.field(O.F_T)[~f] != trml.T_init)
