"""Pins of the oracle's pair-loop terms, heat sums, hashed contact and grid variants against
mathematics that does not share their code.

- CSPM (src/interactions_gpu.cu:227-294): with the Randles-Libersky correction the SPH gradient
  of a linear velocity field is exact (the corrected kernel gradient reproduces linear functions
  by construction), on a jittered lattice with uneven volumes, for every particle.
- Monaghan viscosity and XSPH (src/interactions_gpu.cu:427-451): two particles, approaching and
  receding, against the textbook forms (Monaghan 1992; Monaghan 1989) evaluated with the closed
  M4 spline gradient.
- Brookshaw heat (src/interactions_gpu.cu:453-468): zero rate for uniform T, the two-particle
  closed form (Brookshaw 1985) with heat flowing from hot to cold, and the lattice Laplacian of a
  quadratic field.
- PSE heat (src/interactions_gpu.cu:89-150): the two-particle closed form, no support cutoff and
  blanked neighbours included (Q15), and the lattice Laplacian of a quadratic field (second
  moment of the Gaussian kernel = 2).
- Hashed contact (src/tool_3d_gpu.cu:282-444): ray parity and shell search on a subdivided,
  rotated box against the geometric inside test and the closest point over all triangles.
- Triangle grid (src/grid_cpu_tri.cpp:48-175): every (cell, triangle) pair against a separating-
  axis test over the whole grid, written here from the 13 axes.
- Roethlin tables (src/grid_gpu_rothlin.cu:36-144) and the unlocked geometry
  (src/grid_gpu_base.cu:185-233) against numpy.
"""
import math

import numpy as np
import pytest

import oracle as O

DX = 1e-3
H = 1.5 * DX
RHO0 = 4.43


def m4_grad(r, h):
    """dW/dr of the M4 spline with the 3-D normalisation 1/(pi h^3), from the textbook form."""
    q = r / h
    s = 1.0 / (math.pi * h ** 3)
    if q < 1.0:
        return s / h * (-3.0 * q + 2.25 * q * q)
    if q < 2.0:
        return -s / h * 0.75 * (2.0 - q) ** 2
    return 0.0


def m4(r, h):
    q = r / h
    s = 1.0 / (math.pi * h ** 3)
    if q < 1.0:
        return s * (1.0 - 1.5 * q * q + 0.75 * q ** 3)
    if q < 2.0:
        return s * 0.25 * (2.0 - q) ** 3
    return 0.0


def material(dx=DX, alpha_visc=1.0, beta_visc=1.0, eta=0.1, xspheps=0.5, wdeltap=None):
    E, nu = 1.1, 0.35
    K, G = E / (3 * (1 - 2 * nu)), E / (2 * (1 + nu))
    phys = O.Phys(E, nu, RHO0, K, G, RHO0 * dx ** 3)
    wdp = m4(dx, 1.5 * dx) if wdeltap is None else wdeltap
    corr = O.Corr(wdp, 0.3, xspheps, alpha_visc, beta_visc, eta)
    return phys, corr


def lattice(nx, ny, nz, jitter, seed, dx=DX):
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    return (g + rng.uniform(-jitter, jitter, size=g.shape)) * dx


def make_sim(pos, vel=None, rho=None, T=None, cspm=False, heat=O.HEAT_NONE, trml=None, phys_corr=None,
             margin=4.0, blanked=None, hard_blank=True):
    n = pos.shape[0]
    sim = O.OracleSim(pos, vel3=vel, rho=np.full(n, RHO0) if rho is None else rho,
                      T=np.full(n, 300.0) if T is None else T, h=np.full(n, H))
    sim.set_options(cspm=cspm, heat=heat, single_sqrt=False)
    phys, corr = phys_corr if phys_corr is not None else material()
    sim.setup_constants(phys, corr, None, trml)
    sim.set_grid_locked(pos.min(0) - margin * DX, pos.max(0) + margin * DX, H)
    if not hard_blank:
        sim.set_grid_params(hard_blank=False)
    if blanked is not None:
        sim.set_field(O.F_BLANKED, blanked)
    sim.set_time(1e-6, 1e-6)
    return sim, phys, corr


def by_id(sim, f):
    """Field in creation order (the oracle's arrays are in the sorted order)."""
    ids = sim.field(O.F_ID).astype(np.int64)
    v = sim.field(f)
    out = np.empty_like(v)
    out[ids] = v
    return out


# ---- CSPM ---------------------------------------------------------------------------------------
def test_cspm_gradient_of_a_linear_field_is_exact():
    """v = G x on a jittered lattice with densities varied by +-5 %: the corrected velocity
    gradient equals G to round-off for every particle, corners and faces included (where the
    uncorrected sum is off by tens of per cent).  Using B instead of inv(B), dropping the minus
    sign of the B sum, or weighting by the particle's own volume fails this."""
    pos = lattice(7, 6, 5, 0.15, 3)
    n = pos.shape[0]
    rng = np.random.default_rng(5)
    G = rng.normal(size=(3, 3)) * 1e-3
    vel = pos @ G.T
    rho = RHO0 * (1 + rng.uniform(-0.05, 0.05, n))
    sim, _, _ = make_sim(pos, vel=vel, rho=rho, cspm=True)
    assert sim.step() == 0
    vd = by_id(sim, O.F_V_DER).reshape(n, 3, 3)   # [c][r] = d v_c / d x_r (GLM column c, row r)
    err = np.abs(vd - G[None]).max()
    assert err < 1e-10 * np.abs(G).max(), err
    # the same sums without the correction are only approximate (boundary particles)
    sim2, _, _ = make_sim(pos, vel=vel, rho=rho, cspm=False)
    assert sim2.step() == 0
    vd2 = by_id(sim2, O.F_V_DER).reshape(n, 3, 3)
    assert np.abs(vd2 - G[None]).max() > 0.1 * np.abs(G).max()
    # continuity follows: rho_t = -rho tr(grad v)
    np.testing.assert_allclose(by_id(sim, O.F_RHO_T), -by_id(sim, O.F_RHO) * 0 - rho * np.trace(G),
                               rtol=1e-9)


# ---- viscosity and XSPH ---------------------------------------------------------------------------
@pytest.mark.parametrize("r_over_dx", [1.0, 1.8])
@pytest.mark.parametrize("approach", [True, False])
def test_viscosity_and_xsph_two_particles(approach, r_over_dx):
    """Two particles at distance r on the x axis moving towards (or away from) each other at u.
    Stress-free at the reference density, so the acceleration is the viscosity alone:
    a_i = -m Pi_ij dW/dr e_ij with Pi_ij = (-alpha c mu + beta mu^2) / rho for v_ij . x_ij < 0,
    mu = h v_ij . x_ij / (r^2 + eta^2 h^2), c = sqrt(K / rho); and Pi_ij = 0 when they recede.
    XSPH: dx_i/dt = v_i + eps m / rho W(r) (v_j - v_i)."""
    r = r_over_dx * DX
    u = 2e-3
    pos = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    s = 1.0 if approach else -1.0
    vel = np.array([[s * u, 0.0, 0.0], [-s * u, 0.0, 0.0]])
    sim, phys, corr = make_sim(pos, vel=vel)
    assert sim.step() == 0
    at = by_id(sim, O.F_VEL_T)
    xt = by_id(sim, O.F_POS_T)
    m, K = phys.mass, phys.K
    # pair i = 0, j = 1: x_ij = (-r, 0, 0), v_ij = (2 s u, 0, 0)
    vx = -2 * s * u * r
    grad_i = m4_grad(r, H) * np.array([-1.0, 0.0, 0.0])   # grad_i W = W'(r) x_ij / r
    if vx < 0:
        mu = H * vx / (r * r + corr.eta ** 2 * H * H)
        c = math.sqrt(K / RHO0)
        Pi = (-corr.alpha * c * mu + corr.beta * mu * mu) / RHO0
        a0 = -m * Pi * grad_i
    else:
        a0 = np.zeros(3)
    np.testing.assert_allclose(at[0], a0, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(at[1], -a0, rtol=1e-10, atol=1e-14)
    if approach:
        assert at[0, 0] < 0 < at[1, 0], "approaching particles repel"
    w = m4(r, H)
    for i, j in ((0, 1), (1, 0)):
        expect = vel[i] + corr.xspheps * m / RHO0 * w * (vel[j] - vel[i])
        np.testing.assert_allclose(xt[i], expect, rtol=1e-12)


# ---- Brookshaw ----------------------------------------------------------------------------------
def brook_trml(alpha=1e-3):
    return O.Trml(553e-8, 0.9, 7.1e-13, alpha, 300.0, 0.0)


def test_brookshaw_two_particles_closed_form():
    """T_t,i = alpha 2 (m / rho_j) (T_i - T_j) W'(r) / r (Brookshaw 1985): heat flows from the
    hot particle to the cold one, equal and opposite for equal masses."""
    r = 1.3 * DX
    pos = np.array([[0.0, 0.0, 0.0], [0.0, r, 0.0]])
    T = np.array([400.0, 300.0])
    trml = brook_trml()
    sim, phys, _ = make_sim(pos, T=T, heat=O.HEAT_BROOKSHAW, trml=trml)
    assert sim.step() == 0
    tt = by_id(sim, O.F_T_T)
    expect0 = trml.alpha * 2.0 * (phys.mass / RHO0) * (T[0] - T[1]) * m4_grad(r, H) / r
    np.testing.assert_allclose(tt[0], expect0, rtol=1e-12)
    np.testing.assert_allclose(tt[1], -expect0, rtol=1e-12)
    assert tt[0] < 0 < tt[1]


@pytest.mark.parametrize("heat", [O.HEAT_BROOKSHAW, O.HEAT_PSE])
def test_heat_rates_of_uniform_and_quadratic_fields(heat):
    """On a regular lattice the heat rate of a uniform temperature is exactly zero, and that of
    T = T0 + a x^2 approximates alpha * 2a in the bulk (Brookshaw: the lattice quadrature of the
    Laplacian; PSE: the Gaussian kernel's second moment is 2).  A dropped factor 2 or a sign
    error is off by 100 %."""
    pos = lattice(13, 13, 13, 0.0, 0)
    n = pos.shape[0]
    trml = brook_trml()
    sim, _, _ = make_sim(pos, T=np.full(n, 300.0), heat=heat, trml=trml)
    assert sim.step() == 0
    assert np.all(by_id(sim, O.F_T_T) == 0.0)
    a = 1e5
    x = pos[:, 0]
    T = 300.0 + a * (x - x.mean()) ** 2
    sim, _, _ = make_sim(pos, T=T, heat=heat, trml=trml)
    assert sim.step() == 0
    tt = by_id(sim, O.F_T_T)
    g = np.rint(pos / DX).astype(int)
    bulk = np.all((g >= 4) & (g <= 8), axis=1)
    ratio = tt[bulk] / (trml.alpha * 2 * a)
    assert np.all(np.abs(ratio - 1.0) < 0.05), (ratio.min(), ratio.max())


def test_pse_two_particles_no_cutoff_and_blanked_neighbour():
    """PSE: T_t,i = alpha sum_j (T_j - T_i) m / rho_j 4 / (h^5 pi^1.5) exp(-r^2 / h^2).  The sum
    has no support cutoff: a neighbour at 2.2 h (outside the spline's support, two cells away,
    inside the 5x5x5 cell stencil) still contributes; and a blanked neighbour still contributes
    when blanked particles stay in the cells (soft blanking), Q15."""
    r = 2.2 * H
    margin = 2.1 * H / DX   # particle 0 sits 0.1 h into cell 2, particle 1 in cell 4
    pos = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, r]])
    T = np.array([350.0, 320.0])
    trml = brook_trml()

    def closed(i, j):
        eta = 4.0 / (H ** 5 * math.pi ** 1.5) * math.exp(-r * r / (H * H))
        return trml.alpha * (T[j] - T[i]) * RHO0 * DX ** 3 / RHO0 * eta

    sim, _, _ = make_sim(pos, T=T, heat=O.HEAT_PSE, trml=trml, margin=margin)
    assert sim.step() == 0
    tt = by_id(sim, O.F_T_T)
    np.testing.assert_allclose(tt, [closed(0, 1), closed(1, 0)], rtol=1e-12)
    assert tt[0] < 0 < tt[1]
    # particle 1 blanked, soft blanking: particle 0 still receives its heat
    sim, _, _ = make_sim(pos, T=T, heat=O.HEAT_PSE, trml=trml, blanked=np.array([0.0, 1.0]),
                         hard_blank=False, margin=margin)
    assert sim.step() == 0
    tt = by_id(sim, O.F_T_T)
    np.testing.assert_allclose(tt[0], closed(0, 1), rtol=1e-12)


# ---- hashed contact -------------------------------------------------------------------------------
def subdivided_box(lo, hi, k, rot):
    """Closed box mesh, each face split into k x k quads (2 triangles each), rotated by `rot`
    about the box centre; outward counter-clockwise triangles."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    verts, index, tris = [], {}, []

    def vid(p):
        key = tuple(np.round(p, 12))
        if key not in index:
            index[key] = len(verts)
            verts.append(p)
        return index[key]

    for axis in range(3):
        for side in (0, 1):
            u, v = [a for a in range(3) if a != axis]
            for a in range(k):
                for b in range(k):
                    quad = []
                    for da, db in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        p = np.empty(3)
                        p[axis] = hi[axis] if side else lo[axis]
                        p[u] = lo[u] + (hi[u] - lo[u]) * (a + da) / k
                        p[v] = lo[v] + (hi[v] - lo[v]) * (b + db) / k
                        quad.append(vid(p))
                    q = quad if (side == 1) == ((axis == 1)) else quad[::-1]
                    tris.append([q[0], q[1], q[2]])
                    tris.append([q[0], q[2], q[3]])
    V = np.array(verts)
    c = 0.5 * (lo + hi)
    V = (V - c) @ rot.T + c
    T = np.array(tris, dtype=np.int32)
    # orient outward: the face normal must point away from the centre
    for t in range(T.shape[0]):
        a, b, cc = V[T[t]]
        nrm = np.cross(b - a, cc - a)
        if np.dot(nrm, (a + b + cc) / 3 - c) < 0:
            T[t] = T[t][[0, 2, 1]]
    return V, T


def rotation(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def test_hashed_contact_against_geometry():
    """Points within half a triangle-grid cell of the surface of a rotated, subdivided box: the
    ray-parity test (+z ray, intersections counted in the cell that holds them) and the normal
    test both agree with the geometric inside test, and the shell search returns the closest
    point over all triangles (the nearest surface is within the first shell)."""
    lo, hi = np.array([0.0, 0.0, 0.0]), np.array([0.012, 0.008, 0.006])
    R = rotation(7)
    V, T = subdivided_box(lo, hi, 4, R)
    sim = O.OracleSim(np.zeros((1, 3)), h=np.full(1, H))
    k = sim.add_tool(V, T, vel=np.zeros(3), mass=1.0, algo=1)
    grid = sim.tool_grid(k)
    gdx = grid["dx"]
    rng = np.random.default_rng(3)
    # sample points near the surface: a random surface point pushed along the face normal
    c = 0.5 * (lo + hi)
    pts, inside = [], []
    while len(pts) < 600:
        local = rng.uniform(lo, hi)
        axis = rng.integers(3)
        side = rng.integers(2)
        local[axis] = hi[axis] if side else lo[axis]
        d = rng.uniform(-0.45, 0.45) * gdx
        if abs(d) < 1e-6 * gdx:
            continue
        nrm = np.zeros(3)
        nrm[axis] = 1.0 if side else -1.0
        q_local = local + d * nrm
        pin = np.all((q_local > lo) & (q_local < hi))
        pts.append((q_local - c) @ R.T + c)
        inside.append(pin)
    pts, inside = np.array(pts), np.array(inside)
    assert 150 < inside.sum() < 450
    by_ray, by_normal = sim.tool_inside(k, pts)
    np.testing.assert_array_equal(by_ray, inside)
    np.testing.assert_array_equal(by_normal, inside)
    # closest point: the shell search against the minimum over every triangle
    tris30 = grid["tris"]
    cps, _ = sim.tool_closest(k, pts)
    for q, cp in zip(pts, cps):
        best = min(np.sum((O.closest_point(t, q)[0] - q) ** 2) for t in tris30)
        assert abs(np.sum((cp - q) ** 2) - best) <= 1e-12 * max(best, 1e-30) + 1e-30


def sat_overlap(tri, lo, hi):
    """Triangle / box overlap by the separating-axis theorem: the three box axes, the triangle
    normal and the nine cross products of box axes with triangle edges."""
    c = 0.5 * (lo + hi)
    e = 0.5 * (hi - lo)
    v = tri - c
    axes = [np.eye(3)[a] for a in range(3)]
    edges = [v[1] - v[0], v[2] - v[1], v[0] - v[2]]
    axes.append(np.cross(edges[0], edges[1]))
    for a in range(3):
        for ed in edges:
            axes.append(np.cross(np.eye(3)[a], ed))
    for ax in axes:
        if np.dot(ax, ax) < 1e-30:
            continue
        p = v @ ax
        rad = np.abs(ax) @ e
        if p.min() > rad or p.max() < -rad:
            return False
    return True


def test_triangle_grid_against_separating_axes():
    """Every (cell, triangle) pair of the tool grid (cells of the mesh's mean edge length over its
    bounding box padded by 1e-4) is in the CSR table exactly when the separating-axis test over
    the whole grid says they overlap; triangles of a cell in ascending index order."""
    R = rotation(11)
    V, T = subdivided_box([0.0, 0.0, 0.0], [0.01, 0.007, 0.005], 3, R)
    sim = O.OracleSim(np.zeros((1, 3)), h=np.full(1, H))
    k = sim.add_tool(V, T, vel=np.zeros(3), mass=1.0, algo=1)
    g = sim.tool_grid(k)
    nx, ny, nz = (int(a) for a in g["dims"])
    lo_all = V.min(0) - 1e-4
    np.testing.assert_array_equal(g["bbmin"], lo_all)
    edges = np.concatenate([np.linalg.norm(V[T[:, a]] - V[T[:, b]], axis=1) for a, b in ((0, 1), (1, 2), (2, 0))])
    assert g["dx"] == pytest.approx(0.5 * (edges.min() + edges.max()), rel=1e-12)
    assert (nx, ny, nz) == tuple(int(math.ceil((V.max(0)[a] + 1e-4 - lo_all[a]) / g["dx"])) for a in range(3))
    tris = V[T]
    expect = []
    for cell in range(nx * ny * nz):
        a, b, cc = cell // (ny * nz), (cell // nz) % ny, cell % nz
        lo = lo_all + np.array([a, b, cc]) * g["dx"]
        hi = lo_all + np.array([a + 1, b + 1, cc + 1]) * g["dx"]
        expect.append([t for t in range(T.shape[0]) if sat_overlap(tris[t], lo, hi)])
    cells = g["cells"].astype(np.int64)
    assert cells[-1] == sum(len(e) for e in expect) == g["n_boxed"]
    boxed = g["boxed"][:, :9].reshape(-1, 3, 3)
    for cell, want in enumerate(expect):
        got = boxed[cells[cell]:cells[cell + 1]]
        assert got.shape[0] == len(want), cell
        # boxed copies carry the vertices of the listed triangles, in ascending index order
        np.testing.assert_array_equal(got, tris[want] if want else got)


# ---- grid variants -----------------------------------------------------------------------------
def test_rothlin_tables_against_numpy():
    """Roethlin layout: start[c] = number of sorted keys below c, end[c] = number at most c, over
    the N unblanked particles (empty cells start == end)."""
    pos = lattice(9, 8, 6, 0.3, 4)
    n = pos.shape[0]
    sim = O.OracleSim(pos, h=np.full(n, H), rho=np.full(n, RHO0))
    sim.set_options(rothlin=True)
    sim.set_grid_locked(pos.min(0) - 2 * DX, pos.max(0) + 2 * DX, H)
    xcut = np.sort(pos[:, 0])[-30]
    sim.set_blanking(pos.min(0) - 1.0, np.array([xcut, 1.0, 1.0]))
    assert sim.sort_only() == 0
    N, _ = sim.counts()
    keys = sim.field(O.F_HASH).astype(np.int64)[:N]
    assert N == n - 29 and np.all(np.diff(keys) >= 0)
    s, e = sim.cells()
    cells = np.arange(sim.grid()["num_cell"])
    np.testing.assert_array_equal(s, np.searchsorted(keys, cells, side="left"))
    np.testing.assert_array_equal(e, np.searchsorted(keys, cells, side="right"))


def test_unlocked_geometry_against_numpy():
    """Geometry that follows the particles: bbmin / bbmax = min / max over the first N positions
    -/+ 1e-6, dx = max h * kernel_width / 2 (kernel width 2), n = ceil(l / dx); the hashes use
    that geometry.  After a step that blanks the top layer, the next geometry covers only the
    N particles left."""
    pos = lattice(9, 7, 5, 0.2, 6)
    n = pos.shape[0]
    rng = np.random.default_rng(2)
    h = H * (1 + rng.uniform(0, 0.1, n))
    sim = O.OracleSim(pos, h=h, rho=np.full(n, RHO0), T=np.full(n, 300.0))
    sim.set_options(cspm=False, heat=O.HEAT_NONE)
    phys, corr = material()
    sim.setup_constants(phys, corr, None, None)
    sim.set_grid_unlocked(100000)
    zcut = np.sort(pos[:, 2])[-40]
    sim.set_blanking(pos.min(0) - 1.0, np.array([1.0, 1.0, zcut]))
    sim.set_time(1e-9, 1e-9)
    assert sim.step() == 0

    def expect(p, hh):
        mn, mx = p.min(0) - 1e-6, p.max(0) + 1e-6
        dx = hh.max() * 2.0 / 2
        nn = np.ceil((mx - mn) / dx).astype(int)
        return mn, mx, dx, nn

    mn, mx, dx, nn = expect(pos, h)   # step 1: N = N_init, before the sort
    g = sim.grid()
    np.testing.assert_array_equal(g["bbmin"], mn)
    np.testing.assert_array_equal(g["bbmax"], mx)
    assert g["dx"] == dx and (g["nx"], g["ny"], g["nz"]) == tuple(nn)
    N, _ = sim.counts()
    assert N == n - 39
    q = sim.field(O.F_POS)[:N]
    hq = sim.field(O.F_H)[:N]
    ids_q = sim.field(O.F_ID).astype(np.int64)[:N]
    assert sim.step() == 0
    mn, mx, dx, nn = expect(q, hq)
    g = sim.grid()
    np.testing.assert_array_equal(g["bbmin"], mn)
    np.testing.assert_array_equal(g["bbmax"], mx)
    assert g["dx"] == dx and (g["nx"], g["ny"], g["nz"]) == tuple(nn)
    # the hashes of this step come from the positions at its start, on the new geometry, and
    # the stable sort orders the particles by them
    ijk = np.trunc((q - mn) / dx).astype(np.int64)
    key_q = (ijk[:, 0] * nn[1] + ijk[:, 1]) * nn[2] + ijk[:, 2]
    order = np.argsort(key_q, kind="stable")
    np.testing.assert_array_equal(sim.field(O.F_HASH).astype(np.int64)[:N], key_q[order])
    np.testing.assert_array_equal(sim.field(O.F_ID).astype(np.int64)[:N], ids_q[order])
