// mfree_oracle.cpp -- CPU FP64 restatement of the iwf_mfree_gpu_3d SPH cutting step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product path: only the
// tests, the smoke check and the CPU-baseline leg of the benchmark load it.  It is never
// linked into, called by, or used as a fallback for the product library.
//
// It restates, in plain host C++ with double precision ("USE_DOUBLE" build of the
// reference), the per-particle formulas and the phase order of the reference:
//   src/kernels.cuh:26-92            cubic_spline, lapl_pse
//   src/plasticity.cuh:28-93         sigma_yield, opt_fun, solve_secant
//   src/eigen_solver.cu:19-302       closed-form symmetric 3x3 eigen decomposition
//   src/interactions_gpu.cu:79-471   hash/unhash, PSE heat, Monaghan pair loop (+CSPM, Brookshaw)
//   src/actions_gpu.cu:40-531        per-particle actions
//   src/leap_frog.cu:30-195          init/predict/correct, reset contact forces, step()
//   src/grid_gpu_base.cu:41-273      hashing, particle-number adaption, geometry
//   src/grid_gpu_green.cu:24-306     stable sort, reorder, cell start/end tables
//   src/grid_gpu_rothlin.cu:36-144   offset-style cell tables
//   src/tool_3d_gpu.cu:74-681,873-948  tool motion and contact
//   src/surface_tri_mesh.cpp:42-584  closest point, ray test, tri/AABB test, mesh normals
//   src/grid_cpu_tri.cpp:48-175      triangle binning
//   src/tool_3d.cpp:28-360           hashed inside test, normal orientation, sampling
//   src/blanking.cpp:20-67           blanking box
// GLM semantics (column-major mat3, m[col][row]) are restated in the m3 type below; the
// operation order of glm::determinant / glm::inverse is the cofactor expansion recalled from
// GLM 0.9.9 (GLM is not vendored by the reference, so that order is unpinned).
//
// Compile with -ffp-contract=off so that every expression is evaluated exactly as written.

#include <algorithm>
#include <parallel/algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <unordered_map>
#include <vector>

#include <omp.h>

namespace orc {

// ------------------------------------------------------------------------------------------
// vector / matrix types with GLM semantics
// ------------------------------------------------------------------------------------------
struct v3 {
  double x = 0., y = 0., z = 0.;
  v3() {}
  v3(double a, double b, double c) : x(a), y(b), z(c) {}
  double &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
static inline v3 operator+(v3 a, v3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 operator-(v3 a, v3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 operator-(v3 a) { return v3(-a.x, -a.y, -a.z); }
static inline v3 operator*(double s, v3 a) { return v3(s * a.x, s * a.y, s * a.z); }
static inline v3 operator*(v3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline v3 operator*(v3 a, v3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 operator/(v3 a, double s) { return v3(a.x / s, a.y / s, a.z / s); }
static inline bool operator==(v3 a, v3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
// glm::dot for vec3: tmp = a*b; tmp.x + tmp.y + tmp.z
static inline double dot(v3 a, v3 b) {
  v3 t = a * b;
  return t.x + t.y + t.z;
}
static inline v3 cross(v3 a, v3 b) {
  return v3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
static inline double length(v3 a) { return std::sqrt(dot(a, a)); }
// glm::normalize: x * inversesqrt(dot(x, x))
static inline v3 normalize(v3 a) { return a * (1.0 / std::sqrt(dot(a, a))); }

struct m3 {
  double m[3][3];  // m[col][row]
  m3() { std::memset(m, 0, sizeof(m)); }
  explicit m3(double s) {
    std::memset(m, 0, sizeof(m));
    m[0][0] = m[1][1] = m[2][2] = s;
  }
  // glm column-major constructor: first three values are column 0
  m3(double x0, double y0, double z0, double x1, double y1, double z1, double x2, double y2,
     double z2) {
    m[0][0] = x0; m[0][1] = y0; m[0][2] = z0;
    m[1][0] = x1; m[1][1] = y1; m[1][2] = z1;
    m[2][0] = x2; m[2][1] = y2; m[2][2] = z2;
  }
  double *operator[](int c) { return m[c]; }
  const double *operator[](int c) const { return m[c]; }
};
static inline m3 operator+(const m3 &a, const m3 &b) {
  m3 r;
  for (int c = 0; c < 3; c++)
    for (int k = 0; k < 3; k++) r.m[c][k] = a.m[c][k] + b.m[c][k];
  return r;
}
static inline m3 operator-(const m3 &a, const m3 &b) {
  m3 r;
  for (int c = 0; c < 3; c++)
    for (int k = 0; k < 3; k++) r.m[c][k] = a.m[c][k] - b.m[c][k];
  return r;
}
static inline m3 operator*(double s, const m3 &a) {
  m3 r;
  for (int c = 0; c < 3; c++)
    for (int k = 0; k < 3; k++) r.m[c][k] = s * a.m[c][k];
  return r;
}
// glm mat3 product: (A*B)[j][i] = A[0][i]*B[j][0] + A[1][i]*B[j][1] + A[2][i]*B[j][2]
static inline m3 operator*(const m3 &A, const m3 &B) {
  m3 r;
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++)
      r.m[j][i] = A.m[0][i] * B.m[j][0] + A.m[1][i] * B.m[j][1] + A.m[2][i] * B.m[j][2];
  return r;
}
static inline m3 transpose(const m3 &a) {
  m3 r;
  for (int c = 0; c < 3; c++)
    for (int k = 0; k < 3; k++) r.m[c][k] = a.m[k][c];
  return r;
}
static inline double determinant(const m3 &m) {
  return +m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
         m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
         m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
}
static inline m3 inverse(const m3 &m) {
  double od = 1.0 / (+m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
                     m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
                     m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]));
  m3 r;
  r[0][0] = +(m[1][1] * m[2][2] - m[2][1] * m[1][2]) * od;
  r[1][0] = -(m[1][0] * m[2][2] - m[2][0] * m[1][2]) * od;
  r[2][0] = +(m[1][0] * m[2][1] - m[2][0] * m[1][1]) * od;
  r[0][1] = -(m[0][1] * m[2][2] - m[2][1] * m[0][2]) * od;
  r[1][1] = +(m[0][0] * m[2][2] - m[2][0] * m[0][2]) * od;
  r[2][1] = -(m[0][0] * m[2][1] - m[2][0] * m[0][1]) * od;
  r[0][2] = +(m[0][1] * m[1][2] - m[1][1] * m[0][2]) * od;
  r[1][2] = -(m[0][0] * m[1][2] - m[1][0] * m[0][2]) * od;
  r[2][2] = +(m[0][0] * m[1][1] - m[1][0] * m[0][1]) * od;
  return r;
}

// The reference's FP64 build still calls single-precision sqrtf / fmaxf at a few sites
// (src/kernels.cuh:48, src/actions_gpu.cu:50,203,217,420-421, src/interactions_gpu.cu:414-415).
// With `single_sqrt` on, those sites round through float exactly like that build.
static inline double sqrt_site(double a, bool single) {
  return single ? (double)sqrtf((float)a) : std::sqrt(a);
}

// ------------------------------------------------------------------------------------------
// SPH kernel functions  (src/kernels.cuh:26-92)
// ------------------------------------------------------------------------------------------
struct w4 {
  double w = 0., wx = 0., wy = 0., wz = 0.;
};

static inline w4 cubic_spline(v3 pi, v3 pj, double hi, bool single) {
  w4 w;
  double xij = pi.x - pj.x;
  double yij = pi.y - pj.y;
  double zij = pi.z - pj.z;
  double rr2 = xij * xij + yij * yij + zij * zij;
  double h1 = 1 / hi;
  double fourh2 = 4 * hi * hi;
  if (rr2 >= fourh2 || rr2 < 1e-8) return w;
  double der, val;
  double rad = sqrt_site(rr2, single);
  double q = rad * h1;
  double fac = (M_1_PI)*h1 * h1 * h1;
  if (q > 1) {
    double a = 2 - q;
    double a2 = a * a;
    val = 0.25 * a2 * a;
    der = -0.75 * a2 * h1 / rad;
  } else {
    val = 1 - 1.5 * q * q * (1 - 0.5 * q);
    der = -3.0 * q * (1 - 0.75 * q) * h1 / rad;
  }
  w.w = val * fac;
  w.wx = der * xij * fac;
  w.wy = der * yij * fac;
  w.wz = der * zij * fac;
  return w;
}

static inline double lapl_pse(v3 pi, v3 pj, double hi) {
  double xij = pi.x - pj.x;
  double yij = pi.y - pj.y;
  double zij = pi.z - pj.z;
  double xx = std::sqrt(xij * xij + yij * yij + zij * zij);
  double h2 = hi * hi;
  double h4 = h2 * h2;
  double h5 = hi * h4;
  return +4. / (h5 * M_PI * std::sqrt(M_PI)) * std::exp(-xx * xx / (h2));
}

// ------------------------------------------------------------------------------------------
// constants  (src/constants_structs.h:25-92)
// ------------------------------------------------------------------------------------------
struct phys_c { double E, nu, rho0, K, G, mass; };
struct trml_c { double cp, tq, k, alpha, T_init, eta; };
struct corr_c { double wdeltap, stresseps, xspheps, alpha, beta, eta; };
struct joco_c { double A, B, C, n, m, Tmelt, Tref, eps_dot_ref, clamp_temp; };

// ------------------------------------------------------------------------------------------
// Johnson-Cook plasticity  (src/plasticity.cuh:28-93)
// ------------------------------------------------------------------------------------------
static inline double sigma_yield(const joco_c &jc, double eps_pl, double eps_pl_dot, double t) {
  double theta = (t - jc.Tref) / (jc.Tmelt - jc.Tref);
  double term_a = jc.A + jc.B * std::pow(eps_pl, jc.n);
  double term_b = 1.0;
  double eps_dot = eps_pl_dot / jc.eps_dot_ref;
  if (eps_dot > 1.0) {
    term_b = 1.0 + jc.C * std::log(eps_dot);
  } else {
    term_b = std::pow((1.0 + eps_dot), jc.C);
  }
  double term_c = 1.0 - std::pow(theta, jc.m);
  return term_a * term_b * term_c;
}

static inline double opt_fun(const joco_c &jc, double dl, double norm_strial, double eps_pl_init,
                             double t, double delta_t, double mu) {
  double eps_eq = eps_pl_init + std::sqrt(2.0 / 3.0) * std::fmax(dl, 0.);
  double eps_eq_dot = std::sqrt(2.0 / 3.0) * std::fmax(dl, 0.) / delta_t;
  double sy = sigma_yield(jc, eps_eq, eps_eq_dot, t);
  return norm_strial - 2 * mu * dl - std::sqrt(2.0 / 3.0) * sy;
}

static inline double solve_secant(const joco_c &jc, double init, double tol, double norm_strial,
                                  double eps_pl_init, double t, double delta_t, double mu,
                                  bool &failed, int *iters = nullptr) {
  double dl = init;
  double dl_old = dl;
  int iter = 0;
  const int max_iter = 100;
  while (true) {
    double g = opt_fun(jc, dl, norm_strial, eps_pl_init, t, delta_t, mu);
    double dl1 = dl * 1.01;
    double g1 = opt_fun(jc, dl1, norm_strial, eps_pl_init, t, delta_t, mu);
    double slope = (g - g1) / (dl - dl1);
    dl_old = dl;
    dl = dl - g / slope;
    if (dl < 0.) dl = -0.1 * dl;
    if (std::fabs(dl - dl_old) < tol) {
      if (iters) *iters = iter;
      return dl;
    }
    if (iter > max_iter) {
      failed = true;
      if (iters) *iters = iter;
      return dl;
    }
    iter++;
  }
}

// ------------------------------------------------------------------------------------------
// closed-form symmetric 3x3 eigen decomposition (Eberly)  (src/eigen_solver.cu:19-302)
// ------------------------------------------------------------------------------------------
static inline v3 e_div(v3 u, double s) {
  double inv = 1.0 / s;
  return v3(u.x * inv, u.y * inv, u.z * inv);
}
static inline v3 e_cross(v3 u, v3 v) {
  return v3(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
}
static inline double e_dot(v3 u, v3 v) { return u.x * v.x + u.y * v.y + u.z * v.z; }

static inline void orthogonal_complement(v3 w, v3 &u, v3 &v) {
  double inv_len;
  if (std::fabs(w.x) > std::fabs(w.y)) {
    inv_len = 1.0 / std::sqrt(w.x * w.x + w.z * w.z);
    u = v3(-w.z * inv_len, 0., +w.x * inv_len);
  } else {
    inv_len = 1.0 / std::sqrt(w.y * w.y + w.z * w.z);
    u = v3(0., +w.z * inv_len, -w.y * inv_len);
  }
  v = e_cross(w, u);
}

static inline void eigenvector0(double a00, double a01, double a02, double a11, double a12,
                                double a22, double ev0, v3 &out) {
  v3 r0(a00 - ev0, a01, a02);
  v3 r1(a01, a11 - ev0, a12);
  v3 r2(a02, a12, a22 - ev0);
  v3 c01 = e_cross(r0, r1), c02 = e_cross(r0, r2), c12 = e_cross(r1, r2);
  double d0 = e_dot(c01, c01), d1 = e_dot(c02, c02), d2 = e_dot(c12, c12);
  double dmax = d0;
  int imax = 0;
  if (d1 > dmax) { dmax = d1; imax = 1; }
  if (d2 > dmax) { imax = 2; }
  if (imax == 0) out = e_div(c01, std::sqrt(d0));
  else if (imax == 1) out = e_div(c02, std::sqrt(d1));
  else out = e_div(c12, std::sqrt(d2));
}

static inline void eigenvector1(double a00, double a01, double a02, double a11, double a12,
                                double a22, v3 ev0, double e1, v3 &out) {
  v3 U, V;
  orthogonal_complement(ev0, U, V);
  v3 AU(a00 * U.x + a01 * U.y + a02 * U.z, a01 * U.x + a11 * U.y + a12 * U.z,
        a02 * U.x + a12 * U.y + a22 * U.z);
  v3 AV(a00 * V.x + a01 * V.y + a02 * V.z, a01 * V.x + a11 * V.y + a12 * V.z,
        a02 * V.x + a12 * V.y + a22 * V.z);
  double m00 = U.x * AU.x + U.y * AU.y + U.z * AU.z - e1;
  double m01 = U.x * AV.x + U.y * AV.y + U.z * AV.z;
  double m11 = V.x * AV.x + V.y * AV.y + V.z * AV.z - e1;
  double am00 = std::fabs(m00), am01 = std::fabs(m01), am11 = std::fabs(m11);
  if (am00 >= am11) {
    double mx = std::fmax(am00, am01);
    if (mx > 0.) {
      if (am00 >= am01) {
        m01 /= m00;
        m00 = 1.0 / std::sqrt(1.0 + m01 * m01);
        m01 *= m00;
      } else {
        m00 /= m01;
        m01 = 1.0 / std::sqrt(1.0 + m00 * m00);
        m00 *= m01;
      }
      out = v3(m01 * U.x - m00 * V.x, m01 * U.y - m00 * V.y, m01 * U.z - m00 * V.z);
    } else {
      out = U;
    }
  } else {
    double mx = std::fmax(am11, am01);
    if (mx > 0.) {
      if (am11 >= am01) {
        m01 /= m11;
        m11 = 1.0 / std::sqrt(1.0 + m01 * m01);
        m01 *= m11;
      } else {
        m11 /= m01;
        m01 = 1.0 / std::sqrt(1.0 + m11 * m11);
        m11 *= m01;
      }
      out = v3(m11 * U.x - m01 * V.x, m11 * U.y - m01 * V.y, m11 * U.z - m01 * V.z);
    } else {
      out = U;
    }
  }
}

static inline void solve_eigen(double a00, double a01, double a02, double a11, double a12,
                               double a22, v3 &eval, v3 &e1, v3 &e2, v3 &e3) {
  double mx0 = std::fmax(std::fabs(a00), std::fabs(a01));
  double mx1 = std::fmax(std::fabs(a02), std::fabs(a11));
  double mx2 = std::fmax(std::fabs(a12), std::fabs(a22));
  double mabs = std::fmax(std::fmax(mx0, mx1), mx2);
  if (mabs < 1e-8) {
    eval = v3(0., 0., 0.);
    e1 = v3(1., 0., 0.);
    e2 = v3(0., 1., 0.);
    e3 = v3(0., 0., 1.);
    return;
  }
  double inv = 1.0 / mabs;
  a00 *= inv; a01 *= inv; a02 *= inv; a11 *= inv; a12 *= inv; a22 *= inv;
  double norm = a01 * a01 + a02 * a02 + a12 * a12;
  if (norm > 0.) {
    double tr3 = (a00 + a11 + a22) / 3.0;
    double b00 = a00 - tr3, b11 = a11 - tr3, b22 = a22 - tr3;
    double denom = std::sqrt((b00 * b00 + b11 * b11 + b22 * b22 + norm * 2.0) / 6.0);
    double c00 = b11 * b22 - a12 * a12;
    double c01 = a01 * b22 - a12 * a02;
    double c02 = a01 * a12 - b11 * a02;
    double det = (b00 * c00 - a01 * c01 + a02 * c02) / (denom * denom * denom);
    double half_det = det * 0.5;
    half_det = std::fmin(std::fmax(half_det, -1.0), 1.0);
    double angle = std::acos(half_det) / 3.0;
    const double two_thirds_pi = 2.09439510239319549;
    double beta2 = std::cos(angle) * 2.0;
    double beta0 = std::cos(angle + two_thirds_pi) * 2.0;
    double beta1 = -(beta0 + beta2);
    eval.x = tr3 + denom * beta0;
    eval.y = tr3 + denom * beta1;
    eval.z = tr3 + denom * beta2;
    if (half_det >= 0.) {
      eigenvector0(a00, a01, a02, a11, a12, a22, eval.z, e3);
      eigenvector1(a00, a01, a02, a11, a12, a22, e3, eval.y, e2);
      e1 = e_cross(e2, e3);
    } else {
      eigenvector0(a00, a01, a02, a11, a12, a22, eval.x, e1);
      eigenvector1(a00, a01, a02, a11, a12, a22, e1, eval.y, e2);
      e3 = e_cross(e1, e2);
    }
  } else {
    eval.x = a00;
    eval.y = a11;
    eval.z = a22;
    e1 = v3(1., 0., 0.);
    e2 = v3(0., 1., 0.);
    e3 = v3(0., 0., 1.);
  }
  eval.x *= mabs;
  eval.y *= mabs;
  eval.z *= mabs;
}

// Gray artificial stress of one particle (src/actions_gpu.cu:70-132)
static inline m3 art_stress(double eps, double rhoi, double pi, const m3 &Si) {
  double rhoi21 = 1. / (rhoi * rhoi);
  double sxx = Si[0][0] - pi, syy = Si[1][1] - pi, szz = Si[2][2] - pi;
  double sxy = Si[0][1], syz = Si[1][2], sxz = Si[0][2];
  v3 ev, e1, e2, e3;
  solve_eigen(sxx, sxy, sxz, syy, syz, szz, ev, e1, e2, e3);
  m3 rot(e1.x, e2.x, e3.x, e1.y, e2.y, e3.y, e1.z, e2.z, e3.z);
  m3 srot(ev.x, 0., 0., 0., ev.y, 0., 0., 0., ev.z);
  for (int k = 0; k < 3; k++) {
    if (srot[k][k] > 0) srot[k][k] = -eps * srot[k][k] * rhoi21;
    else srot[k][k] = 0.;
  }
  return rot * srot * transpose(rot);
}

// Jaumann stress rate (src/actions_gpu.cu:134-190)
static inline m3 jaumann(double G, const m3 &vd, const m3 &Si) {
  double vx_x = vd[0][0], vx_y = vd[0][1], vx_z = vd[0][2];
  double vy_x = vd[1][0], vy_y = vd[1][1], vy_z = vd[1][2];
  double vz_x = vd[2][0], vz_y = vd[2][1], vz_z = vd[2][2];
  double Sxx = Si[0][0], Sxy = Si[0][1], Sxz = Si[0][2];
  double Syy = Si[1][1], Syz = Si[1][2], Szz = Si[2][2];
  const m3 epsdot(vx_x, 0.5 * (vx_y + vy_x), 0.5 * (vx_z + vz_x), 0.5 * (vx_y + vy_x), vy_y,
                  0.5 * (vy_z + vz_y), 0.5 * (vx_z + vz_x), 0.5 * (vy_z + vz_y), vz_z);
  const m3 omega(0., 0.5 * (vy_x - vx_y), 0.5 * (vz_x - vx_z), 0.5 * (vx_y - vy_x), 0.,
                 0.5 * (vz_y - vy_z), 0.5 * (vx_z - vz_x), 0.5 * (vy_z - vz_y), 0.);
  const m3 S(Sxx, Sxy, Sxz, Sxy, Syy, Syz, Sxz, Syz, Szz);
  const m3 I(1.);
  double tr = epsdot[0][0] + epsdot[1][1] + epsdot[2][2];
  return (2.0 * G) * (epsdot - ((1. / 3.) * tr) * I) + omega * S + S * transpose(omega);
}

// ------------------------------------------------------------------------------------------
// triangle geometry  (src/surface_tri_mesh.cpp, src/tool_3d_gpu.cu:101-280)
// ------------------------------------------------------------------------------------------
struct tri {
  int i1 = 0, i2 = 0, i3 = 0;
  v3 p1, p2, p3, p1_init, p2_init, p3_init;
  v3 normal, np1, np2, np3, ne1, ne2, ne3;
};

static inline v3 closest_on_segment(v3 p1, v3 p2, v3 p) {
  double dx = p1.x - p2.x, dy = p1.y - p2.y, dz = p1.z - p2.z;
  double ns2 = dx * dx + dy * dy + dz * dz;
  if (ns2 == 0.) return p1;
  double t = ((p2.x - p1.x) * (p.x - p1.x) + (p2.y - p1.y) * (p.y - p1.y) +
              (p2.z - p1.z) * (p.z - p1.z)) /
             ns2;
  if (t > 1.) return p2;
  if (t < 0.) return p1;
  return (1 - t) * p1 + t * p2;
}

static inline double distsquared(v3 a, v3 b) {
  double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  return dx * dx + dy * dy + dz * dz;
}

static inline void closest_point(const tri &t, v3 qp, v3 &cp, v3 &np) {
  v3 p1 = t.p1, p2 = t.p2, p3 = t.p3;
  v3 p1p2 = p2 - p1, p1p3 = p3 - p1, pp1 = p1 - qp;
  double B = dot(p1p3, p1p2), E = dot(p1p2, p1p2), C = dot(p1p3, p1p3);
  double det = B * B - E * C;
  double A = dot(p1p3, pp1), D = dot(p1p2, pp1);
  double t1 = (D * C - A * B) / det;
  double t2 = (A * E - D * B) / det;
  if (t1 < 0.) {
    cp = closest_on_segment(p1, p3, qp);
    np = (cp == p1) ? t.np1 : ((cp == p3) ? t.np3 : t.ne3);
  } else if (t2 < 0.) {
    cp = closest_on_segment(p1, p2, qp);
    np = (cp == p1) ? t.np1 : ((cp == p2) ? t.np2 : t.ne1);
  } else if (t1 + t2 > 1.) {
    cp = closest_on_segment(p2, p3, qp);
    np = (cp == p2) ? t.np2 : ((cp == p3) ? t.np3 : t.ne2);
  } else {
    cp = p1 + t1 * p1p2 + t2 * p1p3;
    np = t.normal;
  }
}

static inline bool ray_intersect(const tri &tr, v3 o, v3 dir, double &t) {
  const double EPS = 0.0000001;
  v3 e1 = tr.p2 - tr.p1, e2 = tr.p3 - tr.p1;
  v3 h = cross(dir, e2);
  double a = dot(e1, h);
  if (a > -EPS && a < EPS) return false;
  double f = 1. / a;
  v3 s = o - tr.p1;
  double u = f * (dot(s, h));
  if (u < 0.0 || u > 1.0) return false;
  v3 q = cross(s, e1);
  double v = f * dot(dir, q);
  if (v < 0.0 || u + v > 1.0) return false;
  t = f * dot(e2, q);
  return (t > EPS);
}

// Akenine-Moller triangle/box overlap (src/surface_tri_mesh.cpp:151-278): separating axis
// test over the 9 edge cross axes, the 3 box axes and the triangle plane.
static bool plane_box_overlap(v3 normal, v3 vert, v3 maxbox) {
  v3 vmin, vmax;
  for (int q = 0; q < 3; q++) {
    double v = vert[q];
    if (normal[q] > 0.0) {
      vmin[q] = -maxbox[q] - v;
      vmax[q] = maxbox[q] - v;
    } else {
      vmin[q] = maxbox[q] - v;
      vmax[q] = -maxbox[q] - v;
    }
  }
  if (dot(normal, vmin) > 0.0) return false;
  if (dot(normal, vmax) >= 0.0) return true;
  return false;
}

static bool axis_reject(double pa, double pb, double rad) {
  double mn = pa < pb ? pa : pb, mx = pa < pb ? pb : pa;
  return (mn > rad || mx < -rad);
}

static bool cover_aabb(const tri &t, v3 bbmin, v3 bbmax) {
  v3 c = 0.5 * (bbmin + bbmax);
  v3 hs = 0.5 * (bbmax - bbmin);
  v3 v0 = t.p1 - c, v1 = t.p2 - c, v2 = t.p3 - c;
  v3 e0 = t.p2 - t.p1, e1 = t.p3 - t.p2, e2 = t.p1 - t.p3;
  double fex, fey, fez;
  // edge e0: X01, Y02, Z12
  fex = std::fabs(e0.x); fey = std::fabs(e0.y); fez = std::fabs(e0.z);
  if (axis_reject(e0.z * v0.y - e0.y * v0.z, e0.z * v2.y - e0.y * v2.z, fez * hs.y + fey * hs.z)) return false;
  if (axis_reject(-e0.z * v0.x + e0.x * v0.z, -e0.z * v2.x + e0.x * v2.z, fez * hs.x + fex * hs.z)) return false;
  if (axis_reject(e0.y * v1.x - e0.x * v1.y, e0.y * v2.x - e0.x * v2.y, fey * hs.x + fex * hs.y)) return false;
  // edge e1: X01, Y02, Z0   (fabsf in the reference: the magnitudes round through float)
  fex = std::fabs((float)e1.x); fey = std::fabs((float)e1.y); fez = std::fabs((float)e1.z);
  if (axis_reject(e1.z * v0.y - e1.y * v0.z, e1.z * v2.y - e1.y * v2.z, fez * hs.y + fey * hs.z)) return false;
  if (axis_reject(-e1.z * v0.x + e1.x * v0.z, -e1.z * v2.x + e1.x * v2.z, fez * hs.x + fex * hs.z)) return false;
  if (axis_reject(e1.y * v0.x - e1.x * v0.y, e1.y * v1.x - e1.x * v1.y, fey * hs.x + fex * hs.y)) return false;
  // edge e2: X2, Y1, Z12
  fex = std::fabs((float)e2.x); fey = std::fabs((float)e2.y); fez = std::fabs((float)e2.z);
  if (axis_reject(e2.z * v0.y - e2.y * v0.z, e2.z * v1.y - e2.y * v1.z, fez * hs.y + fey * hs.z)) return false;
  if (axis_reject(-e2.z * v0.x + e2.x * v0.z, -e2.z * v1.x + e2.x * v1.z, fez * hs.x + fex * hs.z)) return false;
  if (axis_reject(e2.y * v1.x - e2.x * v1.y, e2.y * v2.x - e2.x * v2.y, fey * hs.x + fex * hs.y)) return false;
  // box axes
  double mn, mx;
  mn = std::min(v0.x, std::min(v1.x, v2.x)); mx = std::max(v0.x, std::max(v1.x, v2.x));
  if (mn > hs.x || mx < -hs.x) return false;
  mn = std::min(v0.y, std::min(v1.y, v2.y)); mx = std::max(v0.y, std::max(v1.y, v2.y));
  if (mn > hs.y || mx < -hs.y) return false;
  mn = std::min(v0.z, std::min(v1.z, v2.z)); mx = std::max(v0.z, std::max(v1.z, v2.z));
  if (mn > hs.z || mx < -hs.z) return false;
  v3 nrm = cross(e0, e1);
  if (!plane_box_overlap(nrm, v0, hs)) return false;
  return true;
}

static void tri_bbox(const tri &t, v3 &mn, v3 &mx) {
  mn = v3(std::fmin(std::fmin(t.p1.x, t.p2.x), t.p3.x), std::fmin(std::fmin(t.p1.y, t.p2.y), t.p3.y),
          std::fmin(std::fmin(t.p1.z, t.p2.z), t.p3.z));
  mx = v3(std::fmax(std::fmax(t.p1.x, t.p2.x), t.p3.x), std::fmax(std::fmax(t.p1.y, t.p2.y), t.p3.y),
          std::fmax(std::fmax(t.p1.z, t.p2.z), t.p3.z));
}

// angle weighted vertex / edge normals (src/surface_tri_mesh.cpp:369-568)
static void mesh_normals(std::vector<tri> &tris, const std::vector<v3> &pts) {
  std::vector<std::vector<int>> vtx_tris(pts.size());
  std::unordered_map<uint64_t, std::vector<int>> edges;
  auto ekey = [](int a, int b) -> uint64_t {
    uint32_t lo = (uint32_t)std::min(a, b), hi = (uint32_t)std::max(a, b);
    return ((uint64_t)hi << 32) | lo;
  };
  for (int i = 0; i < (int)tris.size(); i++) {
    vtx_tris[tris[i].i1].push_back(i);
    vtx_tris[tris[i].i2].push_back(i);
    vtx_tris[tris[i].i3].push_back(i);
    edges[ekey(tris[i].i1, tris[i].i2)].push_back(i);
    edges[ekey(tris[i].i2, tris[i].i3)].push_back(i);
    edges[ekey(tris[i].i3, tris[i].i1)].push_back(i);
  }
  for (auto &t : tris) {
    t.p1 = pts[t.i1]; t.p2 = pts[t.i2]; t.p3 = pts[t.i3];
    t.p1_init = t.p1; t.p2_init = t.p2; t.p3_init = t.p3;
    t.normal = normalize(cross(t.p2 - t.p1, t.p3 - t.p1));
  }
  for (auto &t : tris) {
    const std::vector<int> &a = edges[ekey(t.i1, t.i2)];
    const std::vector<int> &b = edges[ekey(t.i2, t.i3)];
    const std::vector<int> &c = edges[ekey(t.i3, t.i1)];
    t.ne1 = normalize(tris[a[0]].normal + tris[a[1]].normal);
    t.ne2 = normalize(tris[b[0]].normal + tris[b[1]].normal);
    t.ne3 = normalize(tris[c[0]].normal + tris[c[1]].normal);
  }
  auto vnormal = [&](int vid) {
    v3 pos = pts[vid];
    v3 np(0., 0., 0.);
    for (int ti : vtx_tris[vid]) {
      const tri &ct = tris[ti];
      v3 f1, f2;
      if (length(ct.p1 - pos) < 1e-12) { f1 = ct.p2; f2 = ct.p3; }
      if (length(ct.p2 - pos) < 1e-12) { f1 = ct.p1; f2 = ct.p3; }
      if (length(ct.p3 - pos) < 1e-12) { f1 = ct.p1; f2 = ct.p2; }
      v3 f1p = f1 - pos, f2p = f2 - pos;
      double ang = std::acos(std::fabs(dot(f1p, f2p) / (length(f1p) * length(f2p))));
      np = np + ang * ct.normal;
    }
    return normalize(np);
  };
  for (auto &t : tris) {
    t.np1 = vnormal(t.i1);
    t.np2 = vnormal(t.i2);
    t.np3 = vnormal(t.i3);
  }
}

static double mesh_average_edge_length(const std::vector<tri> &tris) {
  double lmin = DBL_MAX, lmax = 0;
  for (const auto &t : tris) {
    double l1 = length(t.p1 - t.p2), l2 = length(t.p2 - t.p3), l3 = length(t.p3 - t.p1);
    lmin = std::fmin(std::fmin(std::fmin(lmin, l1), l2), l3);
    lmax = std::fmax(std::fmax(std::fmax(lmax, l1), l2), l3);
  }
  return 0.5 * (lmax + lmin);
}

// ------------------------------------------------------------------------------------------
// tool: triangle grid + contact  (src/grid_cpu_tri.cpp, src/tool_3d.cpp, src/tool_3d_gpu.cu)
// ------------------------------------------------------------------------------------------
struct tool {
  // host tool grid (grid_cpu_tri)
  double gdx = 0.;
  unsigned gnx = 0, gny = 0, gnz = 0;
  double gbbmin[3], gbbmax[3];
  std::vector<unsigned> cells;   // num_cell + 1
  std::vector<tri> boxed;        // per-cell triangle copies
  std::vector<tri> tris;         // original order
  // device-side state of tool_3d_gpu
  int algo = 0;                  // 0 exhaustive, 1 spatial hashing
  double vel[3] = {0, 0, 0};     // m_vel
  double tool_vel_const[3] = {0, 0, 0};  // __constant__ tool_vel, set only by set_vel (Q4)
  double mu = 0., alpha = 0.1, slave_mass = 0.;
  double birth = 0., death = FLT_MAX;
  // geometry constant (moving bbox + grid) as uploaded to the device
  double bbmin[3], bbmax[3];
};

static unsigned tool_unhash_host(const tool &T, v3 p, unsigned &ix, unsigned &iy, unsigned &iz) {
  ix = (unsigned)((p.x - T.gbbmin[0]) / T.gdx);
  iy = (unsigned)((p.y - T.gbbmin[1]) / T.gdx);
  iz = (unsigned)((p.z - T.gbbmin[2]) / T.gdx);
  return ix * (T.gny) * (T.gnz) + iy * (T.gnz) + iz;
}

static void tool_build_grid(tool &T, const std::vector<tri> &tris, double h_avg) {
  double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (const auto &t : tris) {
    v3 a, b;
    tri_bbox(t, a, b);
    for (int k = 0; k < 3; k++) {
      mn[k] = std::fmin(a[k], mn[k]);
      mx[k] = std::fmax(b[k], mx[k]);
    }
  }
  for (int k = 0; k < 3; k++) {
    T.gbbmin[k] = mn[k] - 1e-4;
    T.gbbmax[k] = mx[k] + 1e-4;
  }
  T.gdx = h_avg;
  T.gnx = (unsigned)std::ceil((T.gbbmax[0] - T.gbbmin[0]) / T.gdx);
  T.gny = (unsigned)std::ceil((T.gbbmax[1] - T.gbbmin[1]) / T.gdx);
  T.gnz = (unsigned)std::ceil((T.gbbmax[2] - T.gbbmin[2]) / T.gdx);
  size_t nc = (size_t)T.gnx * T.gny * T.gnz;
  std::vector<std::vector<int>> grid(nc);
  for (int ti = 0; ti < (int)tris.size(); ti++) {
    v3 a, b;
    tri_bbox(tris[ti], a, b);
    unsigned xl, yl, zl, xh, yh, zh;
    tool_unhash_host(T, a, xl, yl, zl);
    tool_unhash_host(T, b, xh, yh, zh);
    for (unsigned ii = xl; ii <= xh; ii++)
      for (unsigned jj = yl; jj <= yh; jj++)
        for (unsigned kk = zl; kk <= zh; kk++) {
          unsigned lin = ii * (T.gny) * (T.gnz) + jj * (T.gnz) + kk;
          v3 lo(T.gbbmin[0] + ii * T.gdx, T.gbbmin[1] + jj * T.gdx, T.gbbmin[2] + kk * T.gdx);
          v3 hi(T.gbbmin[0] + (ii + 1) * T.gdx, T.gbbmin[1] + (jj + 1) * T.gdx,
                T.gbbmin[2] + (kk + 1) * T.gdx);
          if (cover_aabb(tris[ti], lo, hi)) grid[lin].push_back(ti);
        }
  }
  T.cells.assign(nc + 1, 0);
  unsigned off = 0;
  for (size_t c = 0; c < nc; c++) {
    T.cells[c] = off;
    off += (unsigned)grid[c].size();
  }
  T.cells[nc] = off;
  T.boxed.clear();
  for (size_t c = 0; c < nc; c++)
    for (int ti : grid[c]) T.boxed.push_back(tris[ti]);
}

// hashed closest-point search with expanding shells (src/tool_3d.cpp:44-148 on the host,
// src/tool_3d_gpu.cu:282-380 on the device); returns true if "inside" by the normal test.
static bool contact_by_normal(const tool &T, const double gbbmin[3], v3 qp, v3 &cp, v3 &nrm,
                              bool device_semantics) {
  int ix, iy, iz;
  if (device_semantics) {
    ix = (int)(unsigned)((qp.x - gbbmin[0]) / T.gdx);
    iy = (int)(unsigned)((qp.y - gbbmin[1]) / T.gdx);
    iz = (int)(unsigned)((qp.z - gbbmin[2]) / T.gdx);
  } else {
    unsigned a, b, c;
    tool_unhash_host(T, qp, a, b, c);
    ix = (int)a; iy = (int)b; iz = (int)c;
  }
  int nx = (int)T.gnx, ny = (int)T.gny, nz = (int)T.gnz;
  // a point on the grid's upper faces (the constructor's outside point bbmax + 1e-6, or a
  // particle at the edge of the moving box) can index one cell past the grid, where the
  // reference reads past its cell table; the index is clamped to the last cell, as the device does
  ix = std::min(ix, nx - 1);
  iy = std::min(iy, ny - 1);
  iz = std::min(iz, nz - 1);
  int box = ix * ny * nz + iy * nz + iz;
  double min_dist = DBL_MAX;
  v3 best_cp, best_n;
  bool first_empty = (T.cells[box] == T.cells[box + 1]);
  for (unsigned p = T.cells[box]; p < T.cells[box + 1]; p++) {
    v3 c, n;
    closest_point(T.boxed[p], qp, c, n);
    double d = distsquared(c, qp);
    if (d < min_dist) { min_dist = d; best_cp = c; best_n = n; }
  }
  int offset = 1;
  bool all_done = false, all_empty = first_empty;
  while (true) {
    int lo_i = ix - offset < 0 ? 0 : ix - offset;
    int lo_j = iy - offset < 0 ? 0 : iy - offset;
    int lo_k = iz - offset < 0 ? 0 : iz - offset;
    int hi_i = ix + offset + 1 > nx ? nx : ix + offset + 1;
    int hi_j = iy + offset + 1 > ny ? ny : iy + offset + 1;
    int hi_k = iz + offset + 1 > nz ? nz : iz + offset + 1;
    for (int a = lo_i; a < hi_i; a++)
      for (int b = lo_j; b < hi_j; b++)
        for (int c = lo_k; c < hi_k; c++) {
          bool on = (a == lo_i || a == hi_i - 1) || (b == lo_j || b == hi_j - 1) ||
                    (c == lo_k || c == hi_k - 1);
          if (!on) continue;
          int bi = a * ny * nz + b * nz + c;
          for (unsigned p = T.cells[bi]; p < T.cells[bi + 1]; p++) {
            all_empty = false;
            v3 cc, nn;
            closest_point(T.boxed[p], qp, cc, nn);
            double d = distsquared(cc, qp);
            if (d < min_dist) { min_dist = d; best_cp = cc; best_n = nn; }
          }
        }
    if (all_done) break;
    if (!all_empty) {
      offset += 1;
      all_done = true;
      continue;
    }
    offset += 1;
  }
  v3 diff = best_cp - qp;
  double sign = (dot(diff, best_n) > 0.) ? 1. : -1.;
  cp = best_cp;
  nrm = best_n;
  return sign < 0.;
}

// ray-parity inside test in the particle's own box column (src/tool_3d_gpu.cu:382-444)
static bool contact_by_ray(const tool &T, const double gbbmin[3], v3 qp) {
  int ix = (int)(unsigned)((qp.x - gbbmin[0]) / T.gdx);
  int iy = (int)(unsigned)((qp.y - gbbmin[1]) / T.gdx);
  int iz = (int)(unsigned)((qp.z - gbbmin[2]) / T.gdx);
  int ny = (int)T.gny, nz = (int)T.gnz;
  ix = std::min(ix, (int)T.gnx - 1);  // clamped as in contact_by_normal
  iy = std::min(iy, ny - 1);
  iz = std::min(iz, nz - 1);
  v3 dir(0., 0., 1.);
  unsigned n_inter = 0;
  for (; iz < nz; ++iz) {
    unsigned bi = ix * ny * nz + iy * nz + iz;
    for (unsigned p = T.cells[bi]; p < T.cells[bi + 1]; p++) {
      double t;
      if (ray_intersect(T.boxed[p], qp, dir, t)) {
        v3 ip = qp + t * dir;
        double lo_x = gbbmin[0] + ix * T.gdx, lo_y = gbbmin[1] + iy * T.gdx, lo_z = gbbmin[2] + iz * T.gdx;
        double hi_x = gbbmin[0] + (ix + 1) * T.gdx, hi_y = gbbmin[1] + (iy + 1) * T.gdx,
               hi_z = gbbmin[2] + (iz + 1) * T.gdx;
        bool in = lo_x <= ip.x && ip.x < hi_x && lo_y <= ip.y && ip.y < hi_y && lo_z <= ip.z &&
                  ip.z < hi_z;
        if (in) n_inter++;
      }
    }
  }
  return (n_inter % 2) != 0;
}

// tool_3d ctor (src/tool_3d.cpp:294-360): build grid, orient normals to point inward
static void tool_construct(tool &T, std::vector<tri> tris, double h_avg) {
  tool_build_grid(T, tris, h_avg);
  T.tris = tris;
  v3 outside(T.gbbmax[0] + 1e-6, T.gbbmax[1] + 1e-6, T.gbbmax[2] + 1e-6);
  v3 cp, n;
  bool inside = contact_by_normal(T, T.gbbmin, outside, cp, n, false);
  if (inside) {
    auto flip = [](tri &t) {
      t.ne1 = -t.ne1; t.ne2 = -t.ne2; t.ne3 = -t.ne3;
      t.np1 = -t.np1; t.np2 = -t.np2; t.np3 = -t.np3;
      t.normal = -t.normal;
    };
    for (auto &t : T.tris) flip(t);
    for (auto &t : T.boxed) flip(t);
  }
  for (int k = 0; k < 3; k++) {
    T.bbmin[k] = T.gbbmin[k];
    T.bbmax[k] = T.gbbmax[k];
  }
}

// ------------------------------------------------------------------------------------------
// simulation state
// ------------------------------------------------------------------------------------------
enum heat_model { HEAT_NONE = 0, HEAT_BROOKSHAW = 1, HEAT_PSE = 2 };

struct sim {
  // options (compile-time macros of the reference, src/types.h:32-108)
  bool cspm = false;
  int heat = HEAT_BROOKSHAW;
  bool single_sqrt = true;
  bool rothlin = false;
  // constants; the actions and interactions translation units hold separate copies
  phys_c a_phys{}, i_phys{};
  corr_c a_corr{}, i_corr{};
  joco_c jc{};
  bool plastic = false;
  trml_c a_twp{}, a_ttool{}, i_twp{}, i_ttool{};
  bool a_thermal = false, a_fric = false, i_thermal_wp = false;
  // particles (src/particle_gpu.h:36-99)
  unsigned N = 0, N_init = 0;
  std::vector<v3> pos, vel, vel_bc, fc, ft, nrm, pos_t, vel_t;
  std::vector<double> h, rho, p, fixed, blanked, tool_p, eps_pl, eps_pl_dot, T, rho_t, T_t;
  std::vector<m3> S, R, S_t, S_der, v_der;
  std::vector<int> idx, hash;
  std::vector<uint32_t> id;  // persistent particle id (not in the reference; for matching)
  // leap_frog init copies
  std::vector<v3> pos0, vel0;
  std::vector<m3> S0;
  std::vector<double> rho0, T0;
  // grid (src/grid_gpu_base.h)
  double g_dx = 0, g_bbmin[3] = {0, 0, 0}, g_bbmax[3] = {0, 0, 0};
  double g_bbmin_init[3] = {0, 0, 0}, g_bbmax_init[3] = {0, 0, 0};
  long g_nx = 0, g_ny = 0, g_nz = 0, g_num_cell = 0, g_max_cell = 0;
  bool g_locked = false, g_hard_blank = true, g_min_fixed = false, g_max_fixed = false;
  double g_vel_bbox[3] = {0, 0, 0};
  std::vector<int> cell_start, cell_end;
  // interactions geometry constants (uploaded each step)
  int i_nx = 0, i_ny = 0, i_nz = 0;
  // tools
  std::vector<tool> tools;
  std::vector<float> forces;
  // blanking (src/blanking.h)
  bool has_blanking = false;
  double b_bbmin[3] = {0, 0, 0}, b_bbmax[3] = {0, 0, 0}, b_bbmin_init[3] = {0, 0, 0},
         b_bbmax_init[3] = {0, 0, 0}, b_vel[3] = {0, 0, 0};
  double b_vmax2 = FLT_MAX;
  bool b_min_fixed = false, b_max_fixed = false;
  // time (globals of src/mfree_gpu_3d.cpp:45-48)
  double dt = 0., t = 0.;
  int status = 0;  // 0 ok, 1 all particles blanked, 2 hash out of grid
  long secant_failures = 0;
  int nthreads = 0;
};

static bool isnaninf(double v) { return std::isnan(v) || std::isinf(v); }

// -- blanking (src/blanking.cpp:43-51, src/actions_gpu.cu:489-531) -------------------------
static void blanking_update(sim &s) {
  for (int k = 0; k < 3; k++) {
    if (!s.b_min_fixed) s.b_bbmin[k] = s.b_bbmin_init[k] + s.t * s.b_vel[k];
    if (!s.b_max_fixed) s.b_bbmax[k] = s.b_bbmax_init[k] + s.t * s.b_vel[k];
  }
}

static bool rates_invalid(const sim &s, unsigned i) {
  return isnaninf(s.pos_t[i].x) || isnaninf(s.pos_t[i].y) || isnaninf(s.pos_t[i].z) ||
         isnaninf(s.vel_t[i].x) || isnaninf(s.vel_t[i].y) || isnaninf(s.vel_t[i].z) ||
         isnaninf(s.S_t[i][0][0]) || isnaninf(s.S_t[i][1][1]) || isnaninf(s.S_t[i][2][2]) ||
         isnaninf(s.S_t[i][0][1]) || isnaninf(s.S_t[i][1][2]) || isnaninf(s.S_t[i][2][0]) ||
         isnaninf(s.rho_t[i]) || isnaninf(s.T_t[i]);
}

static void perform_blanking(sim &s) {
#pragma omp parallel for
  for (long i = 0; i < (long)s.N_init; i++) {
    v3 p = s.pos[i], v = s.vel[i];
    bool too_fast = v.x * v.x + v.y * v.y + v.z * v.z >= s.b_vmax2;
    bool in_x = p.x >= s.b_bbmin[0] && p.x <= s.b_bbmax[0];
    bool in_y = p.y >= s.b_bbmin[1] && p.y <= s.b_bbmax[1];
    bool in_z = p.z >= s.b_bbmin[2] && p.z <= s.b_bbmax[2];
    if (too_fast || !in_x || !in_y || !in_z) {
      s.blanked[i] = 1.;
    } else if (!rates_invalid(s, (unsigned)i)) {
      s.blanked[i] = 0.;
    }
  }
}

// -- grid (src/grid_gpu_base.cu) ------------------------------------------------------------
static void grid_set_geometry(sim &s, const double bbmin[3], const double bbmax[3], double h) {
  for (int k = 0; k < 3; k++) {
    s.g_bbmin[k] = bbmin[k] - 1e-6;
    s.g_bbmax[k] = bbmax[k] + 1e-6;
    s.g_bbmin_init[k] = bbmin[k] - 1e-6;
    s.g_bbmax_init[k] = bbmax[k] + 1e-6;
  }
  double kernel_width = 2.;
  s.g_dx = h * kernel_width / 2.;
  s.g_nx = (long)std::ceil((s.g_bbmax[0] - s.g_bbmin[0]) / s.g_dx);
  s.g_ny = (long)std::ceil((s.g_bbmax[1] - s.g_bbmin[1]) / s.g_dx);
  s.g_nz = (long)std::ceil((s.g_bbmax[2] - s.g_bbmin[2]) / s.g_dx);
  s.g_num_cell = s.g_nx * s.g_ny * s.g_nz;
  s.g_max_cell = s.g_num_cell;
  s.g_locked = true;
  s.cell_start.assign(s.g_max_cell, -1);
  s.cell_end.assign(s.g_max_cell, -1);
}

static void grid_update_geometry(sim &s, double kernel_width) {
  if (s.g_locked) {
    if (std::fabs(s.g_vel_bbox[0]) > 0 || std::fabs(s.g_vel_bbox[1]) > 0 || std::fabs(s.g_vel_bbox[2]) > 0) {
      for (int k = 0; k < 3; k++) {
        if (!s.g_min_fixed) s.g_bbmin[k] = s.g_bbmin_init[k] + s.t * s.g_vel_bbox[k];
        if (!s.g_max_fixed) s.g_bbmax[k] = s.g_bbmax_init[k] + s.t * s.g_vel_bbox[k];
      }
    }
    return;
  }
  double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX}, hmax = -DBL_MAX;
  for (unsigned i = 0; i < s.N; i++) {
    for (int k = 0; k < 3; k++) {
      if (s.pos[i][k] < mn[k]) mn[k] = s.pos[i][k];
      if (s.pos[i][k] > mx[k]) mx[k] = s.pos[i][k];
    }
    if (s.h[i] > hmax) hmax = s.h[i];
  }
  for (int k = 0; k < 3; k++) {
    s.g_bbmin[k] = mn[k] - 1e-6;
    s.g_bbmax[k] = mx[k] + 1e-6;
  }
  s.g_dx = hmax * kernel_width / 2;
  s.g_nx = (long)std::ceil((s.g_bbmax[0] - s.g_bbmin[0]) / s.g_dx);
  s.g_ny = (long)std::ceil((s.g_bbmax[1] - s.g_bbmin[1]) / s.g_dx);
  s.g_nz = (long)std::ceil((s.g_bbmax[2] - s.g_bbmin[2]) / s.g_dx);
  s.g_num_cell = s.g_nx * s.g_ny * s.g_nz;
  if (s.g_num_cell >= s.g_max_cell) s.status = 3;
}

static void assign_hashes(sim &s) {
#pragma omp parallel for
  for (long i = 0; i < (long)s.N_init; i++) {
    if (s.blanked[i] == 1. && s.g_hard_blank) {
      s.hash[i] = INT_MAX;
      continue;
    }
    int ix = (int)((s.pos[i].x - s.g_bbmin[0]) / s.g_dx);
    int iy = (int)((s.pos[i].y - s.g_bbmin[1]) / s.g_dx);
    int iz = (int)((s.pos[i].z - s.g_bbmin[2]) / s.g_dx);
    s.hash[i] = (int)(ix * s.g_ny * s.g_nz + iy * s.g_nz + iz);
  }
}

static void adapt_particle_number(sim &s) {
  if (!s.g_hard_blank) return;
  long num_out = 0;
#pragma omp parallel for reduction(+ : num_out)
  for (long i = 0; i < (long)s.N_init; i++)
    if (s.hash[i] == INT_MAX) num_out++;
  s.N = s.N_init - num_out;
  if (s.N == 0) s.status = 1;
}

template <class T>
static void permute(std::vector<T> &v, const std::vector<int> &idx, unsigned n) {
  std::vector<T> tmp(v);
#pragma omp parallel for
  for (long i = 0; i < (long)n; i++) v[i] = tmp[idx[i]];
}

// Green sort (src/grid_gpu_green.cu:265-301): stable sort_by_key over N_init, reorder the 19
// persistent fields, cell start/end tables from the sorted keys.
static void grid_sort(sim &s) {
  unsigned n = s.N_init;
  std::vector<int> order(n);
  for (unsigned i = 0; i < n; i++) order[i] = (int)i;
  // ordered by (key, original index): a total order, so any sort gives the stable sort by key
  __gnu_parallel::sort(order.begin(), order.end(), [&](int a, int b) {
    return s.hash[a] < s.hash[b] || (s.hash[a] == s.hash[b] && a < b);
  });
  std::vector<int> keys(n);
#pragma omp parallel for
  for (long i = 0; i < (long)n; i++) keys[i] = s.hash[order[i]];
  s.hash = keys;
  s.idx = order;
  std::fill(s.cell_start.begin(), s.cell_start.end(), -1);
  std::fill(s.cell_end.begin(), s.cell_end.end(), -1);
  permute(s.pos, order, n); permute(s.vel, order, n); permute(s.vel_bc, order, n);
  permute(s.fc, order, n); permute(s.ft, order, n); permute(s.h, order, n);
  permute(s.rho, order, n); permute(s.S, order, n); permute(s.eps_pl, order, n);
  permute(s.eps_pl_dot, order, n); permute(s.T, order, n); permute(s.fixed, order, n);
  permute(s.blanked, order, n); permute(s.tool_p, order, n);
  permute(s.pos_t, order, n); permute(s.vel_t, order, n); permute(s.rho_t, order, n);
  permute(s.S_t, order, n); permute(s.T_t, order, n); permute(s.id, order, n);
  unsigned ncell_part = s.g_hard_blank ? s.N : s.N_init;
  for (unsigned i = 0; i < ncell_part; i++) {
    int hh = s.hash[i];
    if (!s.g_hard_blank && hh == INT_MAX) continue;
    if (hh < 0 || hh >= s.g_num_cell) {
      s.status = 2;
      continue;
    }
    if (i == 0 || hh != s.hash[i - 1]) {
      s.cell_start[hh] = (int)i;
      if (i > 0 && s.hash[i - 1] >= 0 && s.hash[i - 1] < s.g_num_cell) s.cell_end[s.hash[i - 1]] = (int)i;
    }
    if (i == ncell_part - 1) s.cell_end[hh] = (int)i + 1;
  }
  if (s.rothlin) {
    // offset-style tables (src/grid_gpu_rothlin.cu:36-144): start = #keys < c, end = #keys < c+1
    std::vector<int> off(s.g_num_cell + 1, 0);
    unsigned cnt = s.N;
    for (unsigned i = 0; i < cnt; i++) off[s.hash[i] + 1]++;
    for (long c = 0; c < s.g_num_cell; c++) off[c + 1] += off[c];
    for (long c = 0; c < s.g_num_cell; c++) {
      s.cell_start[c] = off[c];
      s.cell_end[c] = off[c + 1];
    }
  }
}

// -- leap frog (src/leap_frog.cu:30-100) ------------------------------------------------------
static void lf_init(sim &s) {
#pragma omp parallel for
  for (long i = 0; i < (long)s.N_init; i++) {
    if (s.blanked[i] == 1.) continue;
    s.pos0[i] = s.pos[i]; s.vel0[i] = s.vel[i]; s.S0[i] = s.S[i];
    s.rho0[i] = s.rho[i]; s.T0[i] = s.T[i];
  }
}

static void lf_advance(sim &s, double f) {
#pragma omp parallel for
  for (long i = 0; i < (long)s.N_init; i++) {
    if (s.blanked[i] == 1.) continue;
    s.pos[i].x = s.pos0[i].x + f * s.pos_t[i].x;
    s.pos[i].y = s.pos0[i].y + f * s.pos_t[i].y;
    s.pos[i].z = s.pos0[i].z + f * s.pos_t[i].z;
    s.vel[i].x = s.vel0[i].x + f * s.vel_t[i].x;
    s.vel[i].y = s.vel0[i].y + f * s.vel_t[i].y;
    s.vel[i].z = s.vel0[i].z + f * s.vel_t[i].z;
    for (int c = 0; c < 3; c++)
      for (int r = 0; r < 3; r++) s.S[i][c][r] = s.S0[i][c][r] + f * s.S_t[i][c][r];
    s.rho[i] = s.rho0[i] + f * s.rho_t[i];
    s.T[i] = s.T0[i] + f * s.T_t[i];
  }
}

// -- actions before the pair loop -------------------------------------------------------------
static void material_eos(sim &s) {
  double rho0 = s.a_phys.rho0;
  double c0 = sqrt_site(s.a_phys.K / rho0, s.single_sqrt);
#pragma omp parallel for
  for (long i = 0; i < (long)s.N; i++) {
    if (s.blanked[i] == 1. || s.tool_p[i] == 1.) continue;
    s.p[i] = c0 * c0 * (s.rho[i] - rho0);
  }
}

static void corrector_artificial_stress(sim &s) {
  double eps = s.a_corr.stresseps;
  if (eps == 0.) return;
#pragma omp parallel for
  for (long i = 0; i < (long)s.N; i++) {
    if (s.blanked[i] == 1. || s.tool_p[i] == 1.) continue;
    s.R[i] = art_stress(eps, s.rho[i], s.p[i], s.S[i]);
  }
}

// -- pair loops (src/interactions_gpu.cu:89-471) ----------------------------------------------
static inline void unhash(const sim &s, int &i, int &j, int &k, int idx) {
  i = idx / (s.i_nz * s.i_ny);
  j = (idx - i * s.i_ny * s.i_nz) / s.i_nz;
  k = idx % s.i_nz;
}

static void interactions_monaghan(sim &s) {
  const phys_c &ph = s.i_phys;
  const corr_c &cr = s.i_corr;
  double mass = ph.mass, K = ph.K;
  double wdeltap = cr.wdeltap, alpha = cr.alpha, beta = cr.beta, eta = cr.eta, eps = cr.xspheps;
  bool brook = s.heat == HEAT_BROOKSHAW;
  bool single = s.single_sqrt;
#pragma omp parallel for schedule(dynamic, 64)
  for (long pidx = 0; pidx < (long)s.N; pidx++) {
    if (s.blanked[pidx] == 1.) continue;
    bool tool_i = s.tool_p[pidx] == 1.;
    double thermal_alpha = (s.tool_p[pidx] == 0.) ? s.i_twp.alpha : s.i_ttool.alpha;
    v3 pi = s.pos[pidx], vi = s.vel[pidx];
    const m3 &Si = s.S[pidx];
    const m3 &Ri = s.R[pidx];
    double hi = s.h[pidx], rhoi = s.rho[pidx], prsi = s.p[pidx];
    double Ti = 0.;
    if (brook && thermal_alpha != 0.) Ti = s.T[pidx];
    double rhoi21 = 1. / (rhoi * rhoi);
    int gi, gj, gk;
    unhash(s, gi, gj, gk, s.hash[pidx]);
    int lo_i = gi - 2 < 0 ? 0 : gi - 2, lo_j = gj - 2 < 0 ? 0 : gj - 2, lo_k = gk - 2 < 0 ? 0 : gk - 2;
    int hi_i = gi + 3 > s.i_nx ? s.i_nx : gi + 3, hi_j = gj + 3 > s.i_ny ? s.i_ny : gj + 3,
        hi_k = gk + 3 > s.i_nz ? s.i_nz : gk + 3;
    m3 vd(0.), sd(0.);
    v3 vt(0., 0., 0.), xt(0., 0., 0.);
    double T_lapl = 0.;
    m3 invB(1.);
    if (s.cspm) {
      m3 B(0.);
      if (!tool_i) {
        for (int ii = lo_i; ii < hi_i; ii++)
          for (int jj = lo_j; jj < hi_j; jj++)
            for (int kk = lo_k; kk < hi_k; kk++) {
              int cid = ii * s.i_ny * s.i_nz + jj * s.i_nz + kk;
              int cs = s.cell_start[cid], ce = s.cell_end[cid];
              if (cs == -1) continue;
              for (int it = cs; it < ce; it++) {
                if (s.blanked[it] == 1.) continue;
                if (s.tool_p[it] == 1.) continue;
                v3 pj = s.pos[it];
                double volj = mass / s.rho[it];
                w4 ww = cubic_spline(pi, pj, hi, single);
                double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                B[0][0] -= volj * dx * ww.wx; B[1][0] -= volj * dx * ww.wy; B[2][0] -= volj * dx * ww.wz;
                B[0][1] -= volj * dy * ww.wx; B[1][1] -= volj * dy * ww.wy; B[2][1] -= volj * dy * ww.wz;
                B[0][2] -= volj * dz * ww.wx; B[1][2] -= volj * dz * ww.wy; B[2][2] -= volj * dz * ww.wz;
              }
            }
      }
      double detB = determinant(B);
      if (detB > 1e-8) invB = inverse(B);
    }
    for (int ii = lo_i; ii < hi_i; ii++)
      for (int jj = lo_j; jj < hi_j; jj++)
        for (int kk = lo_k; kk < hi_k; kk++) {
          int cid = ii * s.i_ny * s.i_nz + jj * s.i_nz + kk;
          int cs = s.cell_start[cid], ce = s.cell_end[cid];
          if (cs == -1) continue;
          for (int it = cs; it < ce; it++) {
            if (s.blanked[it] == 1.) continue;
            v3 pj = s.pos[it], vj = s.vel[it];
            const m3 &Sj = s.S[it];
            const m3 &Rj = s.R[it];
            double hj = s.h[it], rhoj = s.rho[it], prsj = s.p[it];
            double Tj = 0.;
            if (brook && thermal_alpha != 0.) Tj = s.T[it];
            bool tool_j = s.tool_p[it] == 1.;
            double volj = mass / rhoj;
            double rhoj21 = 1. / (rhoj * rhoj);
            w4 ww = cubic_spline(pi, pj, hi, single);
            double w = ww.w, w_x, w_y, w_z;
            if (!s.cspm) {
              w_x = ww.wx; w_y = ww.wy; w_z = ww.wz;
            } else {
              w_x = (ww.wx * invB[0][0] + ww.wy * invB[1][0] + ww.wz * invB[2][0]);
              w_y = (ww.wx * invB[0][1] + ww.wy * invB[1][1] + ww.wz * invB[2][1]);
              w_z = (ww.wx * invB[0][2] + ww.wy * invB[1][2] + ww.wz * invB[2][2]);
            }
            if (!tool_i && !tool_j) {
              vd[0][0] += (vj.x - vi.x) * w_x * volj;
              vd[0][1] += (vj.x - vi.x) * w_y * volj;
              vd[0][2] += (vj.x - vi.x) * w_z * volj;
              vd[1][0] += (vj.y - vi.y) * w_x * volj;
              vd[1][1] += (vj.y - vi.y) * w_y * volj;
              vd[1][2] += (vj.y - vi.y) * w_z * volj;
              vd[2][0] += (vj.z - vi.z) * w_x * volj;
              vd[2][1] += (vj.z - vi.z) * w_y * volj;
              vd[2][2] += (vj.z - vi.z) * w_z * volj;
              double Rxx = 0., Ryy = 0., Rzz = 0., Rxy = 0., Rxz = 0., Ryz = 0.;
              if (wdeltap > 0) {
                double fab = w / wdeltap;
                fab *= fab;
                fab *= fab;
                Rxx = fab * (Ri[0][0] + Rj[0][0]);
                Rxy = fab * (Ri[0][1] + Rj[0][1]);
                Ryy = fab * (Ri[1][1] + Rj[1][1]);
                Rxz = fab * (Ri[0][2] + Rj[0][2]);
                Ryz = fab * (Ri[1][2] + Rj[1][2]);
                Rzz = fab * (Ri[2][2] + Rj[2][2]);
              }
              sd[0][0] += mass * ((Si[0][0] - prsi) * rhoi21 + (Sj[0][0] - prsj) * rhoj21 + Rxx) * w_x;
              sd[0][1] += mass * (Si[0][1] * rhoi21 + Sj[0][1] * rhoj21 + Rxy) * w_y;
              sd[0][2] += mass * (Si[0][2] * rhoi21 + Sj[0][2] * rhoj21 + Rxz) * w_z;
              sd[1][0] += mass * (Si[1][0] * rhoi21 + Sj[1][0] * rhoj21 + Rxy) * w_x;
              sd[1][1] += mass * ((Si[1][1] - prsi) * rhoi21 + (Sj[1][1] - prsj) * rhoj21 + Ryy) * w_y;
              sd[1][2] += mass * (Si[1][2] * rhoi21 + Sj[1][2] * rhoj21 + Ryz) * w_z;
              sd[2][0] += mass * (Si[2][0] * rhoi21 + Sj[2][0] * rhoj21 + Rxz) * w_x;
              sd[2][1] += mass * (Si[2][1] * rhoi21 + Sj[2][1] * rhoj21 + Ryz) * w_y;
              sd[2][2] += mass * ((Si[2][2] - prsi) * rhoi21 + (Sj[2][2] - prsj) * rhoj21 + Rzz) * w_z;
              double xij = pi.x - pj.x, yij = pi.y - pj.y, zij = pi.z - pj.z;
              double vijx = vi.x - vj.x, vijy = vi.y - vj.y, vijz = vi.z - vj.z;
              double vijposij = vijx * xij + vijy * yij + vijz * zij;
              double rhoij = 0.5 * (rhoi + rhoj);
              if (vijposij < 0.) {
                double ci = sqrt_site(K / rhoi, single);
                double cj = sqrt_site(K / rhoj, single);
                double cij = 0.5 * (ci + cj);
                double hij = 0.5 * (hi + hj);
                double r2ij = xij * xij + yij * yij + zij * zij;
                double muij = (hij * vijposij) / (r2ij + eta * eta * hij * hij);
                double piij = (-alpha * cij * muij + beta * muij * muij) / rhoij;
                vt.x += -mass * piij * w_x;
                vt.y += -mass * piij * w_y;
                vt.z += -mass * piij * w_z;
              }
              xt.x += -eps * w * mass / rhoij * vijx;
              xt.y += -eps * w * mass / rhoij * vijy;
              xt.z += -eps * w * mass / rhoij * vijz;
            }
            if (brook && thermal_alpha != 0.) {
              double xij = pi.x - pj.x, yij = pi.y - pj.y, zij = pi.z - pj.z;
              double rij = std::sqrt(xij * xij + yij * yij + zij * zij);
              if (rij > 1e-8) {
                double ex = xij / rij, ey = yij / rij, ez = zij / rij;
                double rij1 = 1. / rij;
                T_lapl += 2.0 * (mass / rhoj) * (Ti - Tj) * rij1 * (ex * w_x + ey * w_y + ez * w_z);
              }
            }
          }
        }
    s.S_der[pidx] = sd;
    s.v_der[pidx] = vd;
    s.pos_t[pidx] = xt;
    s.vel_t[pidx] = vt;
    if (brook && thermal_alpha != 0.) s.T_t[pidx] = thermal_alpha * T_lapl;
  }
}

static void interactions_heat_pse(sim &s) {
  if (!s.i_thermal_wp) return;
  double mass = s.i_phys.mass;
  double a_wp = s.i_twp.alpha, a_tool = s.i_ttool.alpha;
#pragma omp parallel for schedule(dynamic, 64)
  for (long pidx = 0; pidx < (long)s.N; pidx++) {
    v3 pi = s.pos[pidx];
    double hi = s.h[pidx], Ti = s.T[pidx];
    double alpha = (s.tool_p[pidx] == 1.) ? a_tool : a_wp;
    int gi, gj, gk;
    unhash(s, gi, gj, gk, s.hash[pidx]);
    int lo_i = gi - 2 < 0 ? 0 : gi - 2, lo_j = gj - 2 < 0 ? 0 : gj - 2, lo_k = gk - 2 < 0 ? 0 : gk - 2;
    int hi_i = gi + 3 > s.i_nx ? s.i_nx : gi + 3, hi_j = gj + 3 > s.i_ny ? s.i_ny : gj + 3,
        hi_k = gk + 3 > s.i_nz ? s.i_nz : gk + 3;
    double T_ti = 0.;
    for (int ii = lo_i; ii < hi_i; ii++)
      for (int jj = lo_j; jj < hi_j; jj++)
        for (int kk = lo_k; kk < hi_k; kk++) {
          int cid = ii * s.i_ny * s.i_nz + jj * s.i_nz + kk;
          int cs = s.cell_start[cid], ce = s.cell_end[cid];
          if (cs == -1) continue;
          for (int it = cs; it < ce; it++) {
            double w2 = lapl_pse(pi, s.pos[it], hi);
            T_ti += (s.T[it] - Ti) * w2 * mass / s.rho[it];
          }
        }
    s.T_t[pidx] = alpha * T_ti;
  }
}

// -- actions after the pair loop --------------------------------------------------------------
static void rate_actions(sim &s) {
  double G = s.a_phys.G, mass = s.a_phys.mass;
#pragma omp parallel for
  for (long i = 0; i < (long)s.N; i++) {
    if (s.blanked[i] == 1. || s.tool_p[i] == 1.) continue;
    // jaumann (src/actions_gpu.cu:134-190)
    s.S_t[i] = jaumann(G, s.v_der[i], s.S[i]);
    // continuity (src/actions_gpu.cu:222-238)
    const m3 &vd = s.v_der[i];
    s.rho_t[i] = -s.rho[i] * (vd[0][0] + vd[1][1] + vd[2][2]);
    // momentum (src/actions_gpu.cu:240-281)
    const m3 &sd = s.S_der[i];
    s.vel_t[i].x += sd[0][0] + sd[0][1] + sd[0][2] + s.fc[i].x / mass + s.ft[i].x / mass;
    s.vel_t[i].y += sd[1][0] + sd[1][1] + sd[1][2] + s.fc[i].y / mass + s.ft[i].y / mass;
    s.vel_t[i].z += sd[2][0] + sd[2][1] + sd[2][2] + s.fc[i].z / mass + s.ft[i].z / mass;
    // advection (src/actions_gpu.cu:283-300)
    s.pos_t[i].x = s.pos_t[i].x + s.vel[i].x;
    s.pos_t[i].y = s.pos_t[i].y + s.vel[i].y;
    s.pos_t[i].z = s.pos_t[i].z + s.vel[i].z;
  }
}

static void invalidate_rates(sim &s) {
#pragma omp parallel for
  for (long i = 0; i < (long)s.N; i++)
    if (rates_invalid(s, (unsigned)i)) s.blanked[i] = 1.;
}

// radial return (src/actions_gpu.cu:371-449)
static void plasticity_johnson_cook(sim &s) {
  if (!s.plastic) return;
  const joco_c &jc = s.jc;
  double mu = s.a_phys.G, dt = s.dt;
  long failures = 0;
#pragma omp parallel for reduction(+ : failures)
  for (long i = 0; i < (long)s.N; i++) {
    if (s.blanked[i] == 1. || s.tool_p[i] == 1.) continue;
    const m3 &S = s.S[i];
    double sxx = S[0][0], syy = S[1][1], szz = S[2][2], sxy = S[0][1], syz = S[1][2], szx = S[2][0];
    double norm = std::sqrt(sxx * sxx + syy * syy + szz * szz + 2 * (sxy * sxy + syz * syz + szx * szx));
    double eps_pl = s.eps_pl[i], eps_pl_dot = s.eps_pl_dot[i], T = s.T[i];
    if (jc.clamp_temp != 0.) T = (T > jc.Tmelt) ? jc.Tmelt - 1 : T;
    double svm = std::sqrt(3.0 / 2.0) * norm;
    double sy = sigma_yield(jc, eps_pl, eps_pl_dot, T);
    if (svm < sy) {
      s.eps_pl_dot[i] = 0.;
      continue;
    }
    bool failed = false;
    double dl = solve_secant(jc, std::fmax(eps_pl_dot * dt * std::sqrt(2. / 3.), 1e-8), 1e-6, norm, eps_pl, T,
                             dt, mu, failed);
    if (failed) failures++;
    double eps_pl_new, eps_pl_dot_new;
    if (s.single_sqrt) {
      float f = sqrtf((float)(2.0 / 3.0)) * fmaxf((float)dl, 0.f);
      eps_pl_new = eps_pl + (double)f;
      eps_pl_dot_new = (double)f / dt;
    } else {
      eps_pl_new = eps_pl + std::sqrt(2.0 / 3.0) * std::fmax(dl, 0.);
      eps_pl_dot_new = std::sqrt(2.0 / 3.0) * std::fmax(dl, 0.) / dt;
    }
    s.eps_pl[i] = eps_pl_new;
    s.eps_pl_dot[i] = eps_pl_dot_new;
    m3 Sn;
    Sn[0][0] = sxx - sxx / norm * dl * 2. * mu;
    Sn[1][1] = syy - syy / norm * dl * 2. * mu;
    Sn[2][2] = szz - szz / norm * dl * 2. * mu;
    Sn[0][1] = sxy - sxy / norm * dl * 2. * mu;
    Sn[1][0] = sxy - sxy / norm * dl * 2. * mu;
    Sn[1][2] = syz - syz / norm * dl * 2. * mu;
    Sn[2][1] = syz - syz / norm * dl * 2. * mu;
    Sn[2][0] = szx - szx / norm * dl * 2. * mu;
    Sn[0][2] = szx - szx / norm * dl * 2. * mu;
    s.S[i] = Sn;
    if (s.a_twp.tq != 0.) {
      double d_eps = eps_pl_new - eps_pl;
      double sy2 = sigma_yield(jc, eps_pl_new, eps_pl_dot_new, T);
      double rho = s.rho[i];
      s.T[i] += s.a_twp.tq / (s.a_twp.cp * rho) * d_eps * sy2;
    }
  }
  s.secant_failures += failures;
}

// -- contact (src/tool_3d_gpu.cu:455-681,873-948) ---------------------------------------------
static void update_tool(sim &s, tool &T) {
  if (T.boxed.empty()) return;
  std::vector<tri> &tr = (T.algo == 1) ? T.boxed : T.tris;
  for (auto &t : tr) {
    t.p1.x = t.p1_init.x + T.vel[0] * s.t; t.p1.y = t.p1_init.y + T.vel[1] * s.t; t.p1.z = t.p1_init.z + T.vel[2] * s.t;
    t.p2.x = t.p2_init.x + T.vel[0] * s.t; t.p2.y = t.p2_init.y + T.vel[1] * s.t; t.p2.z = t.p2_init.z + T.vel[2] * s.t;
    t.p3.x = t.p3_init.x + T.vel[0] * s.t; t.p3.y = t.p3_init.y + T.vel[1] * s.t; t.p3.z = t.p3_init.z + T.vel[2] * s.t;
  }
  for (int k = 0; k < 3; k++) {
    T.bbmin[k] += s.dt * T.vel[k];
    T.bbmax[k] += s.dt * T.vel[k];
  }
}

static void contact_force(sim &s, tool &T, int tool_idx, bool record) {
  if (T.boxed.empty()) return;
  if (s.t < T.birth || s.t > T.death) return;
  bool hashing = T.algo == 1;
  bool safe = hashing;  // Q22
  const std::vector<tri> &tr = hashing ? T.boxed : T.tris;
  double dt = s.dt;
  std::vector<float> fx(s.N, 0.f), fy(s.N, 0.f), fz(s.N, 0.f);
  std::vector<char> hit(s.N, 0);
#pragma omp parallel for schedule(dynamic, 256)
  for (long pidx = 0; pidx < (long)s.N; pidx++) {
    if (s.blanked[pidx] == 1. || s.tool_p[pidx] == 1.) continue;
    v3 qp = s.pos[pidx];
    bool in_x = qp.x > T.bbmin[0] && qp.x < T.bbmax[0];
    bool in_y = qp.y > T.bbmin[1] && qp.y < T.bbmax[1];
    bool in_z = qp.z > T.bbmin[2] && qp.z < T.bbmax[2];
    if (!(in_x && in_y && in_z)) continue;
    v3 cp, nrm;
    if (hashing) {
      if (!contact_by_ray(T, T.bbmin, qp)) continue;
      contact_by_normal(T, T.bbmin, qp, cp, nrm, true);
    } else {
      double min_dist = FLT_MAX;
      for (size_t i = 0; i < tr.size(); i++) {
        v3 c, n;
        closest_point(tr[i], qp, c, n);
        double d = distsquared(c, qp);
        if (d < min_dist) { min_dist = d; cp = c; nrm = n; }
      }
      v3 diff = cp - qp;
      double sign = (dot(diff, nrm) > 0.) ? 1. : -1.;
      if (sign > 0.) continue;
    }
    double dt2 = dt * dt;
    double gN = dot((qp - cp), nrm);
    v3 fN = (-T.slave_mass * gN * nrm) / dt2 * T.alpha;
    v3 fT(0., 0., 0.);
    v3 vm(T.tool_vel_const[0], T.tool_vel_const[1], T.tool_vel_const[2]);
    if (T.mu != 0.) {
      v3 vs = s.vel[pidx];
      v3 fric = s.ft[pidx];
      v3 v = vs - vm;
      v3 vr = v - v * nrm;
      v3 kdeltae = T.alpha * T.slave_mass * vr / dt;
      double fyv = T.mu * length(fN);
      v3 fstar = fric - kdeltae;
      if (length(fstar) > fyv) fT = fyv * fstar / length(fstar);
      else fT = fstar;
    }
    s.fc[pidx] = fN;
    s.ft[pidx] = fT;
    s.nrm[pidx] = nrm;
    if (record) {
      float a = (float)(fN.x + fT.x), b = (float)(fN.y + fT.y), c = (float)(fN.z + fT.z);
      if (!hashing) {
        a = std::isnan(a) ? 0.f : a;
        b = std::isnan(b) ? 0.f : b;
        c = std::isnan(c) ? 0.f : c;
      }
      fx[pidx] = a; fy[pidx] = b; fz[pidx] = c;
      hit[pidx] = 1;
    }
  }
  if (record) {
    // the reference sums float atomics in an unspecified order; the oracle sums in index order
    for (unsigned i = 0; i < s.N; i++)
      if (hit[i]) {
        s.forces[3 * tool_idx + 0] += fx[i];
        s.forces[3 * tool_idx + 1] += fy[i];
        s.forces[3 * tool_idx + 2] += fz[i];
      }
  }
}

static void fric_heat_gen(sim &s, const double vt[3]) {
  if (!s.a_fric) return;
  double eta = s.a_twp.eta;
#pragma omp parallel for
  for (long i = 0; i < (long)s.N; i++) {
    if (s.blanked[i] == 1. || s.tool_p[i] == 1.) continue;
    v3 f = s.ft[i];
    double fmag = sqrt_site(f.x * f.x + f.y * f.y + f.z * f.z, s.single_sqrt);
    if (fmag == 0.) continue;
    v3 n = s.nrm[i], v = s.vel[i];
    v3 vd(v.x - vt[0], v.y - vt[1], v.z - vt[2]);
    double vdn = vd.x * n.x + vd.y * n.y + vd.z * n.z;
    v3 vr(vd.x - vdn, vd.y - vdn, vd.z - vdn);
    double vmag = sqrt_site(vr.x * vr.x + vr.y * vr.y + vr.z * vr.z, s.single_sqrt);
    s.T[i] += eta * s.dt * fmag * vmag / (s.a_twp.cp * s.a_phys.mass);
  }
}

static void boundary_conditions(sim &s) {
#pragma omp parallel for
  for (long i = 0; i < (long)s.N; i++) {
    if (s.blanked[i] == 1. || s.tool_p[i] == 1.) continue;
    if (s.fixed[i] != 0.) {
      s.vel[i] = s.vel_bc[i];
      s.fc[i] = v3(0., 0., 0.);
      s.pos_t[i] = v3(0., 0., 0.);
      s.vel_t[i] = v3(0., 0., 0.);
    }
  }
}

static void boundary_conditions_thermal(sim &s) {
  if (!s.a_thermal) return;
#pragma omp parallel for
  for (long i = 0; i < (long)s.N; i++)
    if (s.fixed[i] == 1.) s.T[i] = s.a_twp.T_init;
}

static void move_tool_particles(sim &s, const double v[3]) {
#pragma omp parallel for
  for (long i = 0; i < (long)s.N_init; i++) {
    if (s.tool_p[i] != 1.) continue;
    s.pos[i].x = s.pos[i].x + v[0] * s.dt;
    s.pos[i].y = s.pos[i].y + v[1] * s.dt;
    s.pos[i].z = s.pos[i].z + v[2] * s.dt;
  }
}

// leap_frog::step (src/leap_frog.cu:121-195)
static void step(sim &s, bool record) {
  if (s.has_blanking) {
    blanking_update(s);
    perform_blanking(s);
  }
  grid_update_geometry(s, 2.0);
  assign_hashes(s);
  if (s.has_blanking) adapt_particle_number(s);
  if (s.status == 1) return;
  grid_sort(s);
  lf_init(s);
  lf_advance(s, 0.5 * s.dt);
  material_eos(s);
  corrector_artificial_stress(s);
  s.i_nx = (int)s.g_nx; s.i_ny = (int)s.g_ny; s.i_nz = (int)s.g_nz;
  interactions_monaghan(s);
  if (s.heat == HEAT_PSE) interactions_heat_pse(s);
  rate_actions(s);
  invalidate_rates(s);
  lf_advance(s, s.dt);
  plasticity_johnson_cook(s);
#pragma omp parallel for
  for (long i = 0; i < (long)s.N; i++) {
    s.fc[i] = v3(0., 0., 0.);
    s.ft[i] = v3(0., 0., 0.);
    s.nrm[i] = v3(0., 0., 0.);
  }
  if (record) std::fill(s.forces.begin(), s.forces.end(), 0.f);
  for (size_t k = 0; k < s.tools.size(); k++) {
    update_tool(s, s.tools[k]);
    contact_force(s, s.tools[k], (int)k, record);
  }
  double tv[3] = {0, 0, 0};
  if (!s.tools.empty())
    for (int k = 0; k < 3; k++) tv[k] = s.tools[0].vel[k];
  fric_heat_gen(s, tv);
  boundary_conditions(s);
  boundary_conditions_thermal(s);
  move_tool_particles(s, tv);
}

// neighbour list of particle i in reference order: stencil (ii, jj, kk ascending), then sorted
// index; filtered by the support predicate of cubic_spline and by blanking of j.
static int neighbour_list(const sim &s, unsigned pidx, int *out, int cap) {
  v3 pi = s.pos[pidx];
  double hi = s.h[pidx];
  double fourh2 = 4 * hi * hi;
  int gi, gj, gk;
  unhash(s, gi, gj, gk, s.hash[pidx]);
  int lo_i = gi - 2 < 0 ? 0 : gi - 2, lo_j = gj - 2 < 0 ? 0 : gj - 2, lo_k = gk - 2 < 0 ? 0 : gk - 2;
  int hi_i = gi + 3 > s.i_nx ? s.i_nx : gi + 3, hi_j = gj + 3 > s.i_ny ? s.i_ny : gj + 3,
      hi_k = gk + 3 > s.i_nz ? s.i_nz : gk + 3;
  int n = 0;
  for (int ii = lo_i; ii < hi_i; ii++)
    for (int jj = lo_j; jj < hi_j; jj++)
      for (int kk = lo_k; kk < hi_k; kk++) {
        int cid = ii * s.i_ny * s.i_nz + jj * s.i_nz + kk;
        int cs = s.cell_start[cid], ce = s.cell_end[cid];
        if (cs == -1) continue;
        for (int it = cs; it < ce; it++) {
          if (s.blanked[it] == 1.) continue;
          v3 pj = s.pos[it];
          double xij = pi.x - pj.x, yij = pi.y - pj.y, zij = pi.z - pj.z;
          double rr2 = xij * xij + yij * yij + zij * zij;
          if (rr2 >= fourh2 || rr2 < 1e-8) continue;
          if (n < cap) out[n] = it;
          n++;
        }
      }
  return n;
}

}  // namespace orc

// ==========================================================================================
// C interface for the Python test harness (ctypes)
// ==========================================================================================
using namespace orc;

extern "C" {

typedef struct { double E, nu, rho0, K, G, mass; } o_phys;
typedef struct { double cp, tq, k, alpha, T_init, eta; } o_trml;
typedef struct { double wdeltap, stresseps, xspheps, alpha, beta, eta; } o_corr;
typedef struct { double A, B, C, n, m, Tmelt, Tref, eps_dot_ref, clamp_temp; } o_joco;

void *o_create(unsigned N, const double *pos3, const double *vel3, const double *rho,
               const double *T, const double *h, const double *fixed, const double *tool_p,
               const double *S9) {
  sim *s = new sim();
  s->N = s->N_init = N;
  s->pos.resize(N); s->vel.resize(N); s->vel_bc.resize(N); s->fc.resize(N); s->ft.resize(N);
  s->nrm.resize(N); s->pos_t.resize(N); s->vel_t.resize(N);
  s->h.assign(N, 0.); s->rho.assign(N, 0.); s->p.assign(N, 0.); s->fixed.assign(N, 0.);
  s->blanked.assign(N, 0.); s->tool_p.assign(N, 0.); s->eps_pl.assign(N, 0.);
  s->eps_pl_dot.assign(N, 0.); s->T.assign(N, 0.); s->rho_t.assign(N, 0.); s->T_t.assign(N, 0.);
  s->S.assign(N, m3()); s->R.assign(N, m3()); s->S_t.assign(N, m3()); s->S_der.assign(N, m3());
  s->v_der.assign(N, m3()); s->idx.assign(N, 0); s->hash.assign(N, 0); s->id.resize(N);
  s->pos0.resize(N); s->vel0.resize(N); s->S0.assign(N, m3()); s->rho0.assign(N, 0.);
  s->T0.assign(N, 0.);
  for (unsigned i = 0; i < N; i++) {
    s->pos[i] = v3(pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2]);
    if (vel3) s->vel[i] = v3(vel3[3 * i], vel3[3 * i + 1], vel3[3 * i + 2]);
    if (rho) s->rho[i] = rho[i];
    if (T) s->T[i] = T[i];
    if (h) s->h[i] = h[i];
    if (fixed) s->fixed[i] = fixed[i];
    if (tool_p) s->tool_p[i] = tool_p[i];
    if (S9)
      for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) s->S[i][c][r] = S9[9 * i + 3 * c + r];
    s->idx[i] = (int)i;
    s->id[i] = i;
  }
  return s;
}

void o_destroy(void *p) { delete (sim *)p; }

void o_set_options(void *p, int cspm, int heat, int single_sqrt, int rothlin) {
  sim *s = (sim *)p;
  s->cspm = cspm != 0;
  s->heat = heat;
  s->single_sqrt = single_sqrt != 0;
  s->rothlin = rothlin != 0;
}

// mirrors actions_setup_* / interactions_setup_* (src/actions_gpu.cu:659-710,
// src/interactions_gpu.cu:552-608); `which`: 0 both consumers, 1 actions only, 2 interactions only
void o_setup_physical_constants(void *p, const o_phys *c, int which) {
  sim *s = (sim *)p;
  phys_c v{c->E, c->nu, c->rho0, c->K, c->G, c->mass};
  if (which != 2) s->a_phys = v;
  if (which != 1) s->i_phys = v;
}
void o_setup_corrector_constants(void *p, const o_corr *c, int which) {
  sim *s = (sim *)p;
  corr_c v{c->wdeltap, c->stresseps, c->xspheps, c->alpha, c->beta, c->eta};
  if (which != 2) s->a_corr = v;
  if (which != 1) s->i_corr = v;
}
void o_setup_johnson_cook_constants(void *p, const o_joco *c) {
  sim *s = (sim *)p;
  s->jc = joco_c{c->A, c->B, c->C, c->n, c->m, c->Tmelt, c->Tref, c->eps_dot_ref, c->clamp_temp};
  s->plastic = true;
}
void o_setup_thermal_constants_wp(void *p, const o_trml *c, int which) {
  sim *s = (sim *)p;
  trml_c v{c->cp, c->tq, c->k, c->alpha, c->T_init, c->eta};
  if (which != 2) {
    s->a_twp = v;
    if (v.eta != 0.) s->a_fric = true;
    if (v.alpha != 0.) s->a_thermal = true;
  }
  if (which != 1) {
    s->i_twp = v;
    s->i_thermal_wp = v.alpha != 0.;
  }
}
void o_setup_thermal_constants_tool(void *p, const o_trml *c, int which) {
  sim *s = (sim *)p;
  trml_c v{c->cp, c->tq, c->k, c->alpha, c->T_init, c->eta};
  if (which != 2) {
    s->a_ttool = v;
    if (v.alpha != 0.) s->a_thermal = true;
  }
  if (which != 1) s->i_ttool = v;
}

void o_set_grid_locked(void *p, const double *bbmin, const double *bbmax, double h) {
  grid_set_geometry(*(sim *)p, bbmin, bbmax, h);
}
// unlocked grid (src/grid_gpu_base.cu:185-233,275-278): geometry follows the particles
void o_set_grid_unlocked(void *p, long max_cell) {
  sim *s = (sim *)p;
  s->g_locked = false;
  s->g_max_cell = max_cell;
  s->g_num_cell = max_cell;
  s->cell_start.assign(max_cell, -1);
  s->cell_end.assign(max_cell, -1);
}
void o_set_grid_params(void *p, const double *vel_bbox, int min_fixed, int max_fixed, int hard_blank) {
  sim *s = (sim *)p;
  for (int k = 0; k < 3; k++) s->g_vel_bbox[k] = vel_bbox[k];
  s->g_min_fixed = min_fixed != 0;
  s->g_max_fixed = max_fixed != 0;
  s->g_hard_blank = hard_blank != 0;
}
void o_get_grid(void *p, long *n_xyz_cells, double *bbmin_bbmax_dx) {
  sim *s = (sim *)p;
  n_xyz_cells[0] = s->g_nx; n_xyz_cells[1] = s->g_ny; n_xyz_cells[2] = s->g_nz;
  n_xyz_cells[3] = s->g_num_cell;
  for (int k = 0; k < 3; k++) {
    bbmin_bbmax_dx[k] = s->g_bbmin[k];
    bbmin_bbmax_dx[3 + k] = s->g_bbmax[k];
  }
  bbmin_bbmax_dx[6] = s->g_dx;
}

void o_set_blanking(void *p, int enabled, const double *bbmin, const double *bbmax,
                    const double *vel, double max_vel_squared, int min_fixed, int max_fixed) {
  sim *s = (sim *)p;
  s->has_blanking = enabled != 0;
  for (int k = 0; k < 3; k++) {
    s->b_bbmin[k] = s->b_bbmin_init[k] = bbmin[k];
    s->b_bbmax[k] = s->b_bbmax_init[k] = bbmax[k];
    s->b_vel[k] = vel[k];
  }
  s->b_vmax2 = max_vel_squared;
  s->b_min_fixed = min_fixed != 0;
  s->b_max_fixed = max_fixed != 0;
}

// tool from an indexed mesh; normals computed as mesh_compute_vertex_normals does
int o_add_tool(void *p, const double *verts, int nverts, const int *tri_idx, int ntris,
               double h_avg, const double *vel, double mass, double alpha, int algo, double mu,
               double birth, double death) {
  sim *s = (sim *)p;
  std::vector<v3> pts(nverts);
  for (int i = 0; i < nverts; i++) pts[i] = v3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
  std::vector<tri> tris(ntris);
  for (int i = 0; i < ntris; i++) {
    tris[i].i1 = tri_idx[3 * i];
    tris[i].i2 = tri_idx[3 * i + 1];
    tris[i].i3 = tri_idx[3 * i + 2];
  }
  mesh_normals(tris, pts);
  if (h_avg <= 0.) h_avg = mesh_average_edge_length(tris);
  tool T;
  tool_construct(T, tris, h_avg);
  T.algo = algo;
  for (int k = 0; k < 3; k++) T.vel[k] = vel[k];
  T.slave_mass = mass;
  T.alpha = alpha;
  T.mu = mu;
  T.birth = birth;
  T.death = death;
  s->tools.push_back(T);
  s->forces.assign(3 * s->tools.size(), 0.f);
  return (int)s->tools.size() - 1;
}

void o_tool_set_vel_constant(void *p, int tool_idx, const double *v) {
  sim *s = (sim *)p;
  for (int k = 0; k < 3; k++) s->tools[tool_idx].tool_vel_const[k] = v[k];
}

// grid dims, boxed triangle count, cells (CSR), per boxed triangle vertices and normals
void o_tool_info(void *p, int tool_idx, long *dims, double *bb_dx) {
  sim *s = (sim *)p;
  const tool &T = s->tools[tool_idx];
  dims[0] = T.gnx; dims[1] = T.gny; dims[2] = T.gnz;
  dims[3] = (long)T.boxed.size();
  dims[4] = (long)T.tris.size();
  for (int k = 0; k < 3; k++) {
    bb_dx[k] = T.gbbmin[k];
    bb_dx[3 + k] = T.gbbmax[k];
  }
  bb_dx[6] = T.gdx;
}
void o_tool_cells(void *p, int tool_idx, unsigned *cells) {
  sim *s = (sim *)p;
  const tool &T = s->tools[tool_idx];
  std::memcpy(cells, T.cells.data(), sizeof(unsigned) * T.cells.size());
}
// which: 0 = original order, 1 = boxed; out: 21 doubles per triangle
// (p1, p2, p3, normal, np1, np2, np3 ... ne1, ne2, ne3 -> 3*10 = 30 doubles)
void o_tool_tris(void *p, int tool_idx, int which, double *out) {
  sim *s = (sim *)p;
  const tool &T = s->tools[tool_idx];
  const std::vector<tri> &tr = which ? T.boxed : T.tris;
  for (size_t i = 0; i < tr.size(); i++) {
    const tri &t = tr[i];
    const v3 vs[10] = {t.p1, t.p2, t.p3, t.normal, t.np1, t.np2, t.np3, t.ne1, t.ne2, t.ne3};
    for (int k = 0; k < 10; k++) {
      out[30 * i + 3 * k + 0] = vs[k].x;
      out[30 * i + 3 * k + 1] = vs[k].y;
      out[30 * i + 3 * k + 2] = vs[k].z;
    }
  }
}
void o_tool_inside(void *p, int tool_idx, int n, const double *pts, int *by_ray, int *by_normal) {
  sim *s = (sim *)p;
  const tool &T = s->tools[tool_idx];
  for (int i = 0; i < n; i++) {
    v3 q(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    v3 cp, nn;
    by_ray[i] = contact_by_ray(T, T.bbmin, q) ? 1 : 0;
    by_normal[i] = contact_by_normal(T, T.bbmin, q, cp, nn, true) ? 1 : 0;
  }
}

// closest point and normal that the hashed algorithm's shell search returns for each point
void o_tool_closest(void *p, int tool_idx, int n, const double *pts, double *cp_out, double *n_out) {
  sim *s = (sim *)p;
  const tool &T = s->tools[tool_idx];
  for (int i = 0; i < n; i++) {
    v3 q(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    v3 cp, nn;
    contact_by_normal(T, T.bbmin, q, cp, nn, true);
    cp_out[3 * i] = cp.x; cp_out[3 * i + 1] = cp.y; cp_out[3 * i + 2] = cp.z;
    n_out[3 * i] = nn.x; n_out[3 * i + 1] = nn.y; n_out[3 * i + 2] = nn.z;
  }
}

void o_set_time(void *p, double dt, double t) {
  sim *s = (sim *)p;
  s->dt = dt;
  s->t = t;
}
void o_set_threads(void *p, int n) {
  ((sim *)p)->nthreads = n;
  if (n > 0) omp_set_num_threads(n);
}

int o_step(void *p, int record) {
  sim *s = (sim *)p;
  step(*s, record != 0);
  return s->status;
}

void o_advance_time(void *p) {
  sim *s = (sim *)p;
  s->t += s->dt;
}

// run `n` steps of the driver loop: step(), then t += dt (src/mfree_gpu_3d.cpp:155-205)
int o_run(void *p, int n, int record) {
  sim *s = (sim *)p;
  for (int k = 0; k < n; k++) {
    step(*s, record != 0);
    if (s->status) return s->status;
    s->t += s->dt;
  }
  return s->status;
}

// partial step for tests: only the spatial sorting part of leap_frog::step
int o_sort_only(void *p) {
  sim *s = (sim *)p;
  if (s->has_blanking) {
    blanking_update(*s);
    perform_blanking(*s);
  }
  grid_update_geometry(*s, 2.0);
  assign_hashes(*s);
  if (s->has_blanking) adapt_particle_number(*s);
  if (s->status == 1) return 1;
  grid_sort(*s);
  s->i_nx = (int)s->g_nx; s->i_ny = (int)s->g_ny; s->i_nz = (int)s->g_nz;
  return s->status;
}

// the Johnson-Cook radial return alone on the current state (src/actions_gpu.cu:371-449); for
// the consistency checks of the return mapping
void o_plasticity_only(void *p) { plasticity_johnson_cook(*(sim *)p); }

void o_counts(void *p, unsigned *N, unsigned *N_init) {
  sim *s = (sim *)p;
  *N = s->N;
  *N_init = s->N_init;
}
long o_secant_failures(void *p) { return ((sim *)p)->secant_failures; }

// field ids (per particle, current sorted order)
enum {
  F_POS = 0, F_VEL = 1, F_VEL_BC = 2, F_FC = 3, F_FT = 4, F_H = 5, F_RHO = 6, F_P = 7, F_S = 8,
  F_R = 9, F_EPS_PL = 10, F_EPS_PL_DOT = 11, F_T = 12, F_FIXED = 13, F_BLANKED = 14, F_TOOL = 15,
  F_POS_T = 16, F_VEL_T = 17, F_RHO_T = 18, F_S_T = 19, F_T_T = 20, F_S_DER = 21, F_V_DER = 22,
  F_HASH = 23, F_IDX = 24, F_ID = 25, F_N = 26
};

int o_field_width(int f) {
  switch (f) {
    case F_POS: case F_VEL: case F_VEL_BC: case F_FC: case F_FT: case F_POS_T: case F_VEL_T: case F_N:
      return 3;
    case F_S: case F_R: case F_S_T: case F_S_DER: case F_V_DER:
      return 9;
    default:
      return 1;
  }
}

void o_get_field(void *p, int f, double *out) {
  sim *s = (sim *)p;
  unsigned n = s->N_init;
  auto put3 = [&](const std::vector<v3> &v) {
    for (unsigned i = 0; i < n; i++) { out[3 * i] = v[i].x; out[3 * i + 1] = v[i].y; out[3 * i + 2] = v[i].z; }
  };
  auto put9 = [&](const std::vector<m3> &v) {
    for (unsigned i = 0; i < n; i++)
      for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) out[9 * i + 3 * c + r] = v[i][c][r];
  };
  auto put1 = [&](const std::vector<double> &v) {
    for (unsigned i = 0; i < n; i++) out[i] = v[i];
  };
  switch (f) {
    case F_POS: put3(s->pos); break;
    case F_VEL: put3(s->vel); break;
    case F_VEL_BC: put3(s->vel_bc); break;
    case F_FC: put3(s->fc); break;
    case F_FT: put3(s->ft); break;
    case F_N: put3(s->nrm); break;
    case F_POS_T: put3(s->pos_t); break;
    case F_VEL_T: put3(s->vel_t); break;
    case F_S: put9(s->S); break;
    case F_R: put9(s->R); break;
    case F_S_T: put9(s->S_t); break;
    case F_S_DER: put9(s->S_der); break;
    case F_V_DER: put9(s->v_der); break;
    case F_H: put1(s->h); break;
    case F_RHO: put1(s->rho); break;
    case F_P: put1(s->p); break;
    case F_EPS_PL: put1(s->eps_pl); break;
    case F_EPS_PL_DOT: put1(s->eps_pl_dot); break;
    case F_T: put1(s->T); break;
    case F_FIXED: put1(s->fixed); break;
    case F_BLANKED: put1(s->blanked); break;
    case F_TOOL: put1(s->tool_p); break;
    case F_RHO_T: put1(s->rho_t); break;
    case F_T_T: put1(s->T_t); break;
    case F_HASH: for (unsigned i = 0; i < n; i++) out[i] = s->hash[i]; break;
    case F_IDX: for (unsigned i = 0; i < n; i++) out[i] = s->idx[i]; break;
    case F_ID: for (unsigned i = 0; i < n; i++) out[i] = s->id[i]; break;
  }
}

// write a per-particle field (tests that plant NaN rates or move particles)
void o_set_field(void *p, int f, const double *in) {
  sim *s = (sim *)p;
  unsigned n = s->N_init;
  auto get3 = [&](std::vector<v3> &v) {
    for (unsigned i = 0; i < n; i++) v[i] = v3(in[3 * i], in[3 * i + 1], in[3 * i + 2]);
  };
  auto get9 = [&](std::vector<m3> &v) {
    for (unsigned i = 0; i < n; i++)
      for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) v[i][c][r] = in[9 * i + 3 * c + r];
  };
  auto get1 = [&](std::vector<double> &v) {
    for (unsigned i = 0; i < n; i++) v[i] = in[i];
  };
  switch (f) {
    case F_POS: get3(s->pos); break;
    case F_VEL: get3(s->vel); break;
    case F_VEL_BC: get3(s->vel_bc); break;
    case F_FC: get3(s->fc); break;
    case F_FT: get3(s->ft); break;
    case F_POS_T: get3(s->pos_t); break;
    case F_VEL_T: get3(s->vel_t); break;
    case F_S: get9(s->S); break;
    case F_S_T: get9(s->S_t); break;
    case F_RHO: get1(s->rho); break;
    case F_T: get1(s->T); break;
    case F_EPS_PL: get1(s->eps_pl); break;
    case F_EPS_PL_DOT: get1(s->eps_pl_dot); break;
    case F_FIXED: get1(s->fixed); break;
    case F_BLANKED: get1(s->blanked); break;
    case F_TOOL: get1(s->tool_p); break;
    case F_RHO_T: get1(s->rho_t); break;
    case F_T_T: get1(s->T_t); break;
  }
}

void o_get_cells(void *p, int *start, int *end) {
  sim *s = (sim *)p;
  std::memcpy(start, s->cell_start.data(), sizeof(int) * s->g_num_cell);
  std::memcpy(end, s->cell_end.data(), sizeof(int) * s->g_num_cell);
}

void o_get_forces(void *p, float *out) {
  sim *s = (sim *)p;
  std::memcpy(out, s->forces.data(), sizeof(float) * s->forces.size());
}

// neighbour lists after a sort (o_sort_only or o_step); counts[i] plus, if `lists` is given,
// `cap` entries per particle in reference order
void o_neighbours(void *p, int *counts, int *lists, int cap) {
  sim *s = (sim *)p;
  std::vector<int> tmp(4096);
  for (unsigned i = 0; i < s->N; i++) {
    if (s->blanked[i] == 1.) {
      counts[i] = 0;
      continue;
    }
    int n = neighbour_list(*s, i, tmp.data(), (int)tmp.size());
    counts[i] = n;
    if (lists)
      for (int k = 0; k < std::min(n, cap); k++) lists[(size_t)i * cap + k] = tmp[k];
  }
}

// ---- warm start ----
// interp_temps (src/Benchmarks/interp_utils.cpp:65-99): for each point, the sum over the samples
// in order of weights * W * v, W the spline of interp_utils.cpp:19-63 (the kernels.cuh form with
// its sqrtf site).  The reference's debug dumps of the inputs and outputs are not restated.
void o_interp_temps(const double *in4, int n_in, double weights, double h, const double *pts3,
                    int n_out, double *out) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n_out; i++) {
    const v3 p(pts3[3 * i], pts3[3 * i + 1], pts3[3 * i + 2]);
    double T = 0.;
    for (int j = 0; j < n_in; j++) {
      const w4 w = cubic_spline(p, v3(in4[4 * j], in4[4 * j + 1], in4[4 * j + 2]), h, true);
      T += weights * w.w * in4[4 * j + 3];
    }
    out[i] = T;
  }
}

// ---- stand-alone device functions for closed-form tests ----
void o_cubic_spline(const double *pi, const double *pj, double h, int single, double *out) {
  w4 w = cubic_spline(v3(pi[0], pi[1], pi[2]), v3(pj[0], pj[1], pj[2]), h, single != 0);
  out[0] = w.w; out[1] = w.wx; out[2] = w.wy; out[3] = w.wz;
}
double o_lapl_pse(const double *pi, const double *pj, double h) {
  return lapl_pse(v3(pi[0], pi[1], pi[2]), v3(pj[0], pj[1], pj[2]), h);
}
double o_sigma_yield(const o_joco *c, double eps_pl, double eps_pl_dot, double t) {
  joco_c jc{c->A, c->B, c->C, c->n, c->m, c->Tmelt, c->Tref, c->eps_dot_ref, c->clamp_temp};
  return sigma_yield(jc, eps_pl, eps_pl_dot, t);
}
double o_opt_fun(const o_joco *c, double dl, double norm, double eps_pl, double t, double dt, double mu) {
  joco_c jc{c->A, c->B, c->C, c->n, c->m, c->Tmelt, c->Tref, c->eps_dot_ref, c->clamp_temp};
  return opt_fun(jc, dl, norm, eps_pl, t, dt, mu);
}
double o_solve_secant(const o_joco *c, double init, double tol, double norm, double eps_pl, double t,
                      double dt, double mu, int *failed, int *iters) {
  joco_c jc{c->A, c->B, c->C, c->n, c->m, c->Tmelt, c->Tref, c->eps_dot_ref, c->clamp_temp};
  bool f = false;
  double r = solve_secant(jc, init, tol, norm, eps_pl, t, dt, mu, f, iters);
  *failed = f ? 1 : 0;
  return r;
}
void o_solve_eigen(const double *a6, double *eval, double *e1, double *e2, double *e3) {
  v3 ev, a, b, c;
  solve_eigen(a6[0], a6[1], a6[2], a6[3], a6[4], a6[5], ev, a, b, c);
  eval[0] = ev.x; eval[1] = ev.y; eval[2] = ev.z;
  e1[0] = a.x; e1[1] = a.y; e1[2] = a.z;
  e2[0] = b.x; e2[1] = b.y; e2[2] = b.z;
  e3[0] = c.x; e3[1] = c.y; e3[2] = c.z;
}
// m9 / out9 in glm storage m[col][row] flattened as 3*col+row
void o_mat3_det_inv(const double *m9, double *det, double *inv9) {
  m3 m;
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) m[c][r] = m9[3 * c + r];
  *det = determinant(m);
  m3 iv = inverse(m);
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) inv9[3 * c + r] = iv[c][r];
}
void o_mat3_mul(const double *a9, const double *b9, double *out9) {
  m3 a, b;
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) { a[c][r] = a9[3 * c + r]; b[c][r] = b9[3 * c + r]; }
  m3 o = a * b;
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) out9[3 * c + r] = o[c][r];
}
void o_art_stress(double eps, double rho, double p, const double *S9, double *out9) {
  m3 S;
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) S[c][r] = S9[3 * c + r];
  m3 R = art_stress(eps, rho, p, S);
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) out9[3 * c + r] = R[c][r];
}
void o_jaumann(double G, const double *vd9, const double *S9, double *out9) {
  m3 vd, S;
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) { vd[c][r] = vd9[3 * c + r]; S[c][r] = S9[3 * c + r]; }
  m3 o = jaumann(G, vd, S);
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++) out9[3 * c + r] = o[c][r];
}
// triangle as 30 doubles (p1,p2,p3,normal,np1,np2,np3,ne1,ne2,ne3)
static tri tri_from(const double *t30) {
  tri t;
  v3 *dst[10] = {&t.p1, &t.p2, &t.p3, &t.normal, &t.np1, &t.np2, &t.np3, &t.ne1, &t.ne2, &t.ne3};
  for (int k = 0; k < 10; k++) *dst[k] = v3(t30[3 * k], t30[3 * k + 1], t30[3 * k + 2]);
  return t;
}
void o_closest_point(const double *t30, const double *qp, double *cp, double *np) {
  tri t = tri_from(t30);
  v3 c, n;
  closest_point(t, v3(qp[0], qp[1], qp[2]), c, n);
  cp[0] = c.x; cp[1] = c.y; cp[2] = c.z;
  np[0] = n.x; np[1] = n.y; np[2] = n.z;
}
int o_ray_intersect(const double *t30, const double *o, const double *dir, double *t) {
  tri tr = tri_from(t30);
  return ray_intersect(tr, v3(o[0], o[1], o[2]), v3(dir[0], dir[1], dir[2]), *t) ? 1 : 0;
}
int o_tri_aabb(const double *t30, const double *bbmin, const double *bbmax) {
  tri tr = tri_from(t30);
  return cover_aabb(tr, v3(bbmin[0], bbmin[1], bbmin[2]), v3(bbmax[0], bbmax[1], bbmax[2])) ? 1 : 0;
}
// mesh normals for an indexed mesh; out: 30 doubles per triangle as in o_tool_tris
void o_mesh_normals(const double *verts, int nverts, const int *tri_idx, int ntris, double *out) {
  std::vector<v3> pts(nverts);
  for (int i = 0; i < nverts; i++) pts[i] = v3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
  std::vector<tri> tris(ntris);
  for (int i = 0; i < ntris; i++) {
    tris[i].i1 = tri_idx[3 * i];
    tris[i].i2 = tri_idx[3 * i + 1];
    tris[i].i3 = tri_idx[3 * i + 2];
  }
  mesh_normals(tris, pts);
  for (int i = 0; i < ntris; i++) {
    const tri &t = tris[i];
    const v3 vs[10] = {t.p1, t.p2, t.p3, t.normal, t.np1, t.np2, t.np3, t.ne1, t.ne2, t.ne3};
    for (int k = 0; k < 10; k++) {
      out[30 * i + 3 * k + 0] = vs[k].x;
      out[30 * i + 3 * k + 1] = vs[k].y;
      out[30 * i + 3 * k + 2] = vs[k].z;
    }
  }
}
double o_mesh_average_edge_length(const double *verts, const int *tri_idx, int ntris) {
  std::vector<tri> tris(ntris);
  for (int i = 0; i < ntris; i++) {
    tri &t = tris[i];
    const int a = tri_idx[3 * i], b = tri_idx[3 * i + 1], c = tri_idx[3 * i + 2];
    t.p1 = v3(verts[3 * a], verts[3 * a + 1], verts[3 * a + 2]);
    t.p2 = v3(verts[3 * b], verts[3 * b + 1], verts[3 * b + 2]);
    t.p3 = v3(verts[3 * c], verts[3 * c + 1], verts[3 * c + 2]);
  }
  return mesh_average_edge_length(tris);
}

}  // extern "C"
