// scene.cpp -- set-up time builders for the orthogonal-cut workloads (host code).
//
// Follows the recipe of the reference's cutting set-ups (src/Benchmarks/benchmarks_single_grain.cu:
// 674-876 for a single tool, 29-471 for tool particles and multi-grain): cubic lattice
// workpiece with a fixed bottom layer, h = hdx*dx, mass = rho0*dx^3, wdeltap = W(dx, h),
// dt = 0.5*CFL*h/(sqrt(v_tool) + c0), locked Green grid and blanking box padded by 15 dx.
// All quantities are in the reference's cm-g-us unit system (stress in Mbar).
//
// Tools are synthetic: a 12-triangle hexahedral wedge (rake / clearance angles) or a rounded
// truncated-octahedron grain tessellated from a cube sphere; tool particles for thermal contact
// are sampled on a dx lattice by ray parity (src/tool_3d.cpp:253-290).

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

#include "../../include/mfree_gpu.h"

namespace {

struct vec {
  double x, y, z;
};
inline vec operator+(vec a, vec b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec operator-(vec a, vec b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec operator*(double s, vec a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dotp(vec a, vec b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline vec crossp(vec a, vec b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(vec a) { return std::sqrt(dotp(a, a)); }

// SplitMix64: counter-based, identical on every platform
struct splitmix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  double sym(double a) { return a * (2.0 * uniform() - 1.0); }                       // [-a,a)
};

struct mesh {
  std::vector<vec> v;
  std::vector<int> t;  // 3 per triangle
};

// hexahedral wedge: profile in the xz-plane, extruded over [y0, y1]; outward-oriented faces
mesh make_wedge(double x_edge, double z_edge, double height, double length, double rake_deg,
                double clear_deg, double y0, double y1) {
  const double d2r = M_PI / 180.0;
  double tr = std::tan(rake_deg * d2r), tc = std::tan(clear_deg * d2r);
  // E: cutting edge, R: top of the rake face, B: top back, C: back end of the clearance face
  double px[4] = {x_edge, x_edge - height * tr, x_edge - length, x_edge - length};
  double pz[4] = {z_edge, z_edge + height, z_edge + height, z_edge + length * tc};
  mesh m;
  for (int s = 0; s < 2; s++)
    for (int k = 0; k < 4; k++) m.v.push_back({px[k], s == 0 ? y0 : y1, pz[k]});
  // profile E,R,B,C is clockwise when seen from +y (x right, z up): E(bottom right) -> R(top
  // right) -> B(top left) -> C(left); caps and sides built so that normals point outward.
  auto quad = [&](int a, int b, int c, int d) {
    m.t.insert(m.t.end(), {a, b, c, a, c, d});
  };
  // y0 cap (normal -y) and y1 cap (normal +y)
  quad(0, 1, 2, 3);
  quad(7, 6, 5, 4);
  // side faces between consecutive profile points k -> k+1 (normal outward)
  for (int k = 0; k < 4; k++) {
    int a = k, b = (k + 1) % 4;
    quad(a, a + 4, b + 4, b);
  }
  // fix orientation so that every face normal points away from the centroid
  vec c{0, 0, 0};
  for (auto &p : m.v) c = c + p;
  c = (1.0 / m.v.size()) * c;
  for (size_t i = 0; i < m.t.size(); i += 3) {
    vec a = m.v[m.t[i]], b = m.v[m.t[i + 1]], d = m.v[m.t[i + 2]];
    vec n = crossp(b - a, d - a);
    vec ctr = (1.0 / 3.0) * (a + b + d);
    if (dotp(n, ctr - c) < 0) std::swap(m.t[i + 1], m.t[i + 2]);
  }
  return m;
}

// rounded truncated octahedron: smooth maximum of the 14 face planes (log-sum-exp with width
// `radius`), tessellated by radial projection of a cube sphere with `res` x `res` quads per face.
mesh make_grain(double edge, double radius, double perturb, int res, splitmix &rng) {
  std::vector<vec> nrm;
  std::vector<double> dist;
  const double dsq = std::sqrt(2.0) * edge;        // square faces
  const double dhex = std::sqrt(6.0) / 2.0 * edge;  // hexagonal faces
  for (int a = 0; a < 3; a++)
    for (int s = -1; s <= 1; s += 2) {
      vec n{0, 0, 0};
      (a == 0 ? n.x : (a == 1 ? n.y : n.z)) = s;
      nrm.push_back(n);
      dist.push_back(dsq + rng.sym(perturb));
    }
  const double r3 = 1.0 / std::sqrt(3.0);
  for (int sx = -1; sx <= 1; sx += 2)
    for (int sy = -1; sy <= 1; sy += 2)
      for (int sz = -1; sz <= 1; sz += 2) {
        nrm.push_back({sx * r3, sy * r3, sz * r3});
        dist.push_back(dhex + rng.sym(perturb));
      }
  auto sdf = [&](vec p) {
    double m = -DBL_MAX;
    std::vector<double> a(nrm.size());
    for (size_t k = 0; k < nrm.size(); k++) {
      a[k] = dotp(nrm[k], p) - dist[k];
      m = std::max(m, a[k]);
    }
    double sum = 0;
    for (size_t k = 0; k < nrm.size(); k++) sum += std::exp((a[k] - m) / radius);
    return m + radius * std::log(sum);
  };
  mesh out;
  std::map<std::tuple<int, int, int>, int> vid;
  auto vertex = [&](int ix, int iy, int iz) {
    auto key = std::make_tuple(ix, iy, iz);
    auto it = vid.find(key);
    if (it != vid.end()) return it->second;
    // cube point in [-1,1]^3 with an equal-angle warp, then radial projection
    auto warp = [&](int i) { return std::tan(M_PI / 4.0 * (2.0 * i / res - 1.0)); };
    vec d{warp(ix), warp(iy), warp(iz)};
    d = (1.0 / norm(d)) * d;
    double lo = 0.0, hi = 4.0 * edge;
    for (int it2 = 0; it2 < 80; it2++) {
      double mid = 0.5 * (lo + hi);
      if (sdf(mid * d) < 0) lo = mid;
      else hi = mid;
    }
    int id = (int)out.v.size();
    out.v.push_back((0.5 * (lo + hi)) * d);
    vid[key] = id;
    return id;
  };
  for (int a = 0; a < 3; a++)
    for (int s = 0; s <= res; s += res) {
      for (int i = 0; i < res; i++)
        for (int j = 0; j < res; j++) {
          auto lat = [&](int u, int v) {
            int c[3];
            c[a] = s;
            c[(a + 1) % 3] = u;
            c[(a + 2) % 3] = v;
            return vertex(c[0], c[1], c[2]);
          };
          int v00 = lat(i, j), v10 = lat(i + 1, j), v11 = lat(i + 1, j + 1), v01 = lat(i, j + 1);
          out.t.insert(out.t.end(), {v00, v10, v11, v00, v11, v01});
        }
    }
  // outward orientation (convex, star-shaped about the origin)
  for (size_t i = 0; i < out.t.size(); i += 3) {
    vec p0 = out.v[out.t[i]], p1 = out.v[out.t[i + 1]], p2 = out.v[out.t[i + 2]];
    vec n = crossp(p1 - p0, p2 - p0);
    if (dotp(n, p0 + p1 + p2) < 0) std::swap(out.t[i + 1], out.t[i + 2]);
  }
  return out;
}

void rotate_y(mesh &m, double phi) {
  double c = std::cos(phi), s = std::sin(phi);
  for (auto &p : m.v) {
    double x = p.x * c + p.z * s, z = -p.x * s + p.z * c;
    p.x = x;
    p.z = z;
  }
}

void bbox(const mesh &m, vec &mn, vec &mx) {
  mn = {DBL_MAX, DBL_MAX, DBL_MAX};
  mx = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (auto &p : m.v) {
    mn = {std::min(mn.x, p.x), std::min(mn.y, p.y), std::min(mn.z, p.z)};
    mx = {std::max(mx.x, p.x), std::max(mx.y, p.y), std::max(mx.z, p.z)};
  }
}

// Moller-Trumbore, +z ray parity: inside test for tool sampling
bool inside_by_parity(const mesh &m, vec o) {
  const vec dir{0, 0, 1};
  int hits = 0;
  for (size_t i = 0; i < m.t.size(); i += 3) {
    vec v0 = m.v[m.t[i]], v1 = m.v[m.t[i + 1]], v2 = m.v[m.t[i + 2]];
    vec e1 = v1 - v0, e2 = v2 - v0;
    vec h = crossp(dir, e2);
    double a = dotp(e1, h);
    if (a > -1e-14 && a < 1e-14) continue;
    double f = 1.0 / a;
    vec s = o - v0;
    double u = f * dotp(s, h);
    if (u < 0.0 || u > 1.0) continue;
    vec q = crossp(s, e1);
    double v = f * dotp(dir, q);
    if (v < 0.0 || u + v > 1.0) continue;
    double t = f * dotp(e2, q);
    if (t > 1e-12) hits++;
  }
  return (hits % 2) != 0;
}

}  // namespace

struct mfg_scene {
  mfg_cut_config cfg;
  std::vector<double> pos, vel, rho, T, h, fixed, tool_p;
  std::vector<mesh> tools;
  unsigned n_workpiece = 0;
  mfg_scene_params params;
};

extern "C" {

int mfg_scene_preset(int which, double scale, unsigned long long seed, mfg_cut_config *c) {
  if (!c) return MFG_ERR_ARG;
  std::memset(c, 0, sizeof(*c));
  c->seed = seed;
  c->jitter = 0.01;
  c->hdx = 1.5;
  c->cfl = 0.3;
  // Ti6Al4V-like: E=110 GPa, nu=0.35, rho=4430 kg/m^3, cp=553 J/kgK, k=7.1 W/mK
  c->E = 1.1; c->nu = 0.35; c->rho0 = 4.43;
  c->cp = 553e-8; c->k = 7.1e-13; c->tq = 0.9; c->T_init = 300.;
  // Johnson-Cook A=1098 MPa, B=1092 MPa, n=0.93, C=0.014, m=1.1, eps0=1/s, Tref=300, Tmelt=1878
  c->jc_A = 0.01098; c->jc_B = 0.01092; c->jc_n = 0.93; c->jc_C = 0.014; c->jc_m = 1.1;
  c->jc_eps_dot_ref = 1e-6; c->jc_Tref = 300.; c->jc_Tmelt = 1878.; c->jc_clamp_temp = 1.;
  c->visc_alpha = 1.0; c->visc_beta = 1.0; c->visc_eta = 0.1; c->xspheps = 0.5; c->stresseps = 0.3;
  c->mu = 0.35; c->contact_alpha = 0.1;
  c->cspm = 1;
  c->max_vel = 4.0;
  c->clearance_deg = 7.0;
  c->tool_rho = 3.5; c->tool_cp = 520e-8; c->tool_k = 0.;
  c->grain_edge = 25e-4; c->edge_radius = 2e-4; c->facet_perturb = 5e-5; c->grain_res = 20;
  c->n_grains = 1;
  if (scale <= 0.) scale = 1.;
  auto sc = [&](int n) { return std::max(4, (int)std::lround(n * scale)); };
  switch (which) {
    case 1:  // 8,000-particle orthogonal cut, wedge tool, Brookshaw heat
      c->nx = sc(40); c->ny = sc(20); c->nz = sc(10); c->dx = 10e-4;
      c->tool_kind = MFG_TOOL_WEDGE; c->rake_deg = 0.; c->speed = 1e-3; c->depth_of_cut = 30e-4;
      c->heat = MFG_HEAT_BROOKSHAW; c->steps = 1000; c->eta_fric = 0.;
      c->algo = MFG_CONTACT_EXHAUSTIVE;
      break;
    case 2:  // 1M single grain, PSE heat + thermal contact
      c->nx = sc(200); c->ny = sc(100); c->nz = sc(50); c->dx = 5e-4;
      c->tool_kind = MFG_TOOL_GRAIN; c->rake_deg = -20.; c->speed = 2e-3; c->depth_of_cut = 15e-4;
      c->heat = MFG_HEAT_PSE; c->steps = 500; c->eta_fric = 0.9;
      c->tool_particles = 1; c->tool_k = 2e-10;
      // exhaustive search: the hashed algorithm's ray-parity test rejects facets whose doubled
      // projected area is below an absolute 1e-7 cm^2 (src/tool_3d_gpu.cu:218,229), which the
      // ~3 um facets of a 4,800-triangle grain do, and then misclassifies particles
      c->algo = MFG_CONTACT_EXHAUSTIVE;
      break;
    case 3:  // 8M, three grains
      c->nx = sc(400); c->ny = sc(200); c->nz = sc(100); c->dx = 2.5e-4;
      c->tool_kind = MFG_TOOL_GRAIN; c->rake_deg = -20.; c->speed = 2e-3; c->depth_of_cut = 10e-4;
      c->heat = MFG_HEAT_PSE; c->steps = 200; c->eta_fric = 0.9;
      c->tool_particles = 1; c->tool_k = 2e-10;
      c->n_grains = 3; c->grain_offset_x = 1e-2; c->grain_offset_y = 2e-3;
      // exhaustive search: the hashed algorithm's ray-parity test rejects facets whose doubled
      // projected area is below an absolute 1e-7 cm^2 (src/tool_3d_gpu.cu:218,229), which the
      // ~3 um facets of a 4,800-triangle grain do, and then misclassifies particles
      c->algo = MFG_CONTACT_EXHAUSTIVE;
      break;
    default:
      return MFG_ERR_ARG;
  }
  return MFG_OK;
}

mfg_scene *mfg_scene_build(const mfg_cut_config *cfg) {
  if (!cfg || cfg->nx <= 0 || cfg->ny <= 0 || cfg->nz <= 0 || cfg->dx <= 0.) return nullptr;
  mfg_scene *sc = new mfg_scene();
  sc->cfg = *cfg;
  const mfg_cut_config &c = *cfg;
  splitmix rng{c.seed * 0x2545F4914F6CDD1Dull + 1};
  const double dx = c.dx;
  const double lx = (c.nx - 1) * dx, ly = (c.ny - 1) * dx, lz = (c.nz - 1) * dx;
  const double hh = c.hdx * dx;
  // workpiece lattice: x in [0, lx], y centred, top surface at z = 0
  for (int i = 0; i < c.nx; i++)
    for (int j = 0; j < c.ny; j++)
      for (int k = 0; k < c.nz; k++) {
        double px = i * dx + rng.sym(c.jitter * dx);
        double py = j * dx - 0.5 * ly + rng.sym(c.jitter * dx);
        double pz = -k * dx + rng.sym(c.jitter * dx);
        sc->pos.insert(sc->pos.end(), {px, py, pz});
        sc->fixed.push_back(pz < -lz + 0.5 * dx ? 1. : 0.);
      }
  sc->n_workpiece = (unsigned)(sc->pos.size() / 3);
  // tools
  vec tmin{DBL_MAX, DBL_MAX, DBL_MAX}, tmax{-DBL_MAX, -DBL_MAX, -DBL_MAX};
  int ntools = c.tool_kind == MFG_TOOL_WEDGE ? 1 : std::max(1, c.n_grains);
  for (int g = 0; g < ntools; g++) {
    mesh m;
    if (c.tool_kind == MFG_TOOL_WEDGE) {
      double margin = 5 * dx;
      m = make_wedge(-2.0 * c.jitter * dx, -c.depth_of_cut, c.depth_of_cut + 10 * dx, 12 * dx,
                     c.rake_deg, c.clearance_deg, -0.5 * ly - margin, 0.5 * ly + margin);
    } else {
      m = make_grain(c.grain_edge, c.edge_radius, c.facet_perturb, c.grain_res, rng);
      rotate_y(m, -c.rake_deg * M_PI / 180.0);
      vec mn, mx;
      bbox(m, mn, mx);
      // lowest point at -depth_of_cut; grain g shifted back along the cut and sideways
      double sy = (g == 0) ? 0. : ((g % 2) ? c.grain_offset_y : -c.grain_offset_y);
      vec shift{0., sy - 0.5 * (mn.y + mx.y), -c.depth_of_cut - mn.z};
      for (auto &p : m.v) p = p + shift;
      // leading point of the part below the workpiece surface (z <= 0) just touching x = 0
      double front = -DBL_MAX;
      for (auto &p : m.v)
        if (p.z <= 0.) front = std::max(front, p.x);
      if (front == -DBL_MAX) front = mx.x;
      double sx = -2.0 * c.jitter * dx - front - g * c.grain_offset_x;
      for (auto &p : m.v) p.x += sx;
    }
    vec mn, mx;
    bbox(m, mn, mx);
    tmin = {std::min(tmin.x, mn.x), std::min(tmin.y, mn.y), std::min(tmin.z, mn.z)};
    tmax = {std::max(tmax.x, mx.x), std::max(tmax.y, mx.y), std::max(tmax.z, mx.z)};
    sc->tools.push_back(m);
  }
  unsigned nwp = sc->n_workpiece;
  sc->vel.assign(3 * nwp, 0.);
  sc->rho.assign(nwp, c.rho0);
  sc->T.assign(nwp, c.T_init);
  sc->h.assign(nwp, hh);
  sc->tool_p.assign(nwp, 0.);
  // tool particles: dx lattice over each tool's box, kept if inside by ray parity
  if (c.tool_particles) {
    for (auto &m : sc->tools) {
      vec mn, mx;
      bbox(m, mn, mx);
      mn = mn - vec{1e-4, 1e-4, 1e-4};
      mx = mx + vec{1e-4, 1e-4, 1e-4};
      int nx = (int)std::ceil((mx.x - mn.x) / dx), ny = (int)std::ceil((mx.y - mn.y) / dx),
          nz = (int)std::ceil((mx.z - mn.z) / dx);
      for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
          for (int k = 0; k < nz; k++) {
            vec p{mn.x + i * dx, mn.y + j * dx, mn.z + k * dx};
            if (!inside_by_parity(m, p)) continue;
            sc->pos.insert(sc->pos.end(), {p.x, p.y, p.z});
            sc->vel.insert(sc->vel.end(), {0., 0., 0.});
            sc->rho.push_back(c.tool_rho);
            sc->T.push_back(c.T_init);
            sc->h.push_back(hh);
            sc->tool_p.push_back(1.);
            // the top layer of the grain is held (velocity and temperature boundary condition)
            sc->fixed.push_back(p.z > mx.z - 1.5 * dx ? 1. : 0.);
          }
    }
  }
  // constants (reference recipe)
  mfg_scene_params &P = sc->params;
  std::memset(&P, 0, sizeof(P));
  double G = c.E / (2. * (1. + c.nu));
  double K = 2.0 * G * (1 + c.nu) / (3 * (1 - 2 * c.nu));
  P.phys = {c.E, c.nu, c.rho0, K, G, dx * dx * dx * c.rho0};
  {
    double h1 = 1. / (c.hdx * dx);
    double q = dx * h1;
    double fac = (M_1_PI)*h1 * h1 * h1;
    double wdp = q > 1 ? fac * 0.25 * (2 - q) * (2 - q) * (2 - q) : fac * (1 - 1.5 * q * q * (1 - 0.5 * q));
    P.corr = {wdp, c.stresseps, c.xspheps, c.visc_alpha, c.visc_beta, c.visc_eta};
  }
  P.joco = {c.jc_A, c.jc_B, c.jc_C, c.jc_n, c.jc_m, c.jc_Tmelt, c.jc_Tref, c.jc_eps_dot_ref, c.jc_clamp_temp};
  P.trml_wp = {c.cp, c.tq, c.k, c.k / (c.rho0 * c.cp), c.T_init, c.eta_fric};
  if (c.tool_particles && c.tool_k > 0.)
    P.trml_tool = {c.tool_cp, 0., c.tool_k, c.tool_k / (c.tool_rho * c.tool_cp), c.T_init, 0.};
  double c0 = std::sqrt(K / c.rho0);
  double dt_max = c.cfl * c.hdx * dx / (std::sqrt(c.speed) + c0);
  P.dt = 0.5 * dt_max;
  P.steps = c.steps;
  P.tool_vel[0] = c.speed;
  P.tool_vel[1] = 0.;
  P.tool_vel[2] = 0.;
  P.h = hh;
  // locked grid and blanking box, padded by 15 dx around workpiece and tools
  double pad = 15.0 * dx;
  P.bbmin[0] = std::min(-pad, tmin.x - pad);
  P.bbmin[1] = std::min(-0.5 * ly - pad, tmin.y - pad);
  P.bbmin[2] = std::min(-1.5 * lz - pad, tmin.z - pad);
  P.bbmax[0] = std::max(lx + pad, tmax.x + pad);
  P.bbmax[1] = std::max(0.5 * ly + pad, tmax.y + pad);
  P.bbmax[2] = std::max(pad, tmax.z + pad);
  P.max_vel = c.max_vel;
  P.cspm = c.cspm;
  P.heat = c.heat;
  P.algo = c.algo;
  P.mu = c.mu;
  P.contact_alpha = c.contact_alpha;
  P.n_tools = (int)sc->tools.size();
  P.n_particles = (unsigned)sc->rho.size();
  P.n_workpiece = nwp;
  return sc;
}

void mfg_scene_free(mfg_scene *s) { delete s; }

int mfg_scene_get_params(const mfg_scene *s, mfg_scene_params *out) {
  if (!s || !out) return MFG_ERR_ARG;
  *out = s->params;
  return MFG_OK;
}

int mfg_scene_get_particles(const mfg_scene *s, double *pos3, double *vel3, double *rho, double *T,
                            double *h, double *fixed, double *tool_p) {
  if (!s) return MFG_ERR_ARG;
  size_t n = s->rho.size();
  if (pos3) std::memcpy(pos3, s->pos.data(), sizeof(double) * 3 * n);
  if (vel3) std::memcpy(vel3, s->vel.data(), sizeof(double) * 3 * n);
  if (rho) std::memcpy(rho, s->rho.data(), sizeof(double) * n);
  if (T) std::memcpy(T, s->T.data(), sizeof(double) * n);
  if (h) std::memcpy(h, s->h.data(), sizeof(double) * n);
  if (fixed) std::memcpy(fixed, s->fixed.data(), sizeof(double) * n);
  if (tool_p) std::memcpy(tool_p, s->tool_p.data(), sizeof(double) * n);
  return MFG_OK;
}

int mfg_scene_tool_size(const mfg_scene *s, int k, int *n_verts, int *n_tris) {
  if (!s || k < 0 || k >= (int)s->tools.size()) return MFG_ERR_ARG;
  *n_verts = (int)s->tools[k].v.size();
  *n_tris = (int)s->tools[k].t.size() / 3;
  return MFG_OK;
}

int mfg_scene_tool_mesh(const mfg_scene *s, int k, double *verts3, int *tris3) {
  if (!s || k < 0 || k >= (int)s->tools.size()) return MFG_ERR_ARG;
  const mesh &m = s->tools[k];
  for (size_t i = 0; i < m.v.size(); i++) {
    verts3[3 * i] = m.v[i].x;
    verts3[3 * i + 1] = m.v[i].y;
    verts3[3 * i + 2] = m.v[i].z;
  }
  std::memcpy(tris3, m.t.data(), sizeof(int) * m.t.size());
  return MFG_OK;
}

}  // extern "C"
