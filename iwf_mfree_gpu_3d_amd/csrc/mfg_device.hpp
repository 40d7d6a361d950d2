// mfg_device.hpp -- per-particle device math of the SPH cutting step (FP64 unless noted).
//
// The files are compiled with -ffp-contract=off: every expression rounds exactly as written, so
// the FP64 path evaluates the reference formulas in the reference's operation order; fused
// multiply-adds are only used where they are spelled out (the FP32 pair arithmetic).
#pragma once

#include <hip/hip_runtime.h>

#include "mfg_internal.hpp"

namespace mfg {

// Workgroups of one launch are handed to the 8 XCDs of the chip round robin (workgroup b runs on
// XCD b % 8), and every XCD has its own L2.  Renumbering them so that each XCD sweeps one
// contiguous range of the sorted particles keeps the neighbour data of spatially close particles
// in the same L2.  The map is a bijection of [0, gridDim.x).
#ifndef MFG_XCD_REMAP
#define MFG_XCD_REMAP 1
#endif
// The same map for workgroup b of nb (a batched launch gives each instance a range of a multiple
// of 8 workgroups, so b % 8 is still the XCD).
__device__ __forceinline__ unsigned xcd_map(unsigned b, unsigned nb) {
  if (!MFG_XCD_REMAP) return b;
  const unsigned x = b & 7u, q = nb >> 3, r = nb & 7u;
  return x * q + (x < r ? x : r) + (b >> 3);
}
__device__ __forceinline__ unsigned xcd_block() { return xcd_map(blockIdx.x, gridDim.x); }

// Instance of this workgroup in a batched launch, its workgroup index b inside the instance's
// range and the range's size nb: the last k with off[k] <= blockIdx.x (empty ranges are skipped).
__device__ __forceinline__ int batch_locate(const BatchMap &m, unsigned &b, unsigned &nb) {
  const unsigned g = blockIdx.x;
  int lo = 0, hi = m.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (m.off[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  b = g - m.off[lo];
  nb = m.off[lo + 1] - m.off[lo];
  return lo;
}


constexpr double PI_D = 3.14159265358979323846;
constexpr double INV_PI_D = 0.31830988618379067154;  // M_1_PI

// Correctly rounded square roots.  HIP's __fsqrt_rn maps to the approximate hardware sqrt
// unless OCML_BASIC_ROUNDED_OPERATIONS is defined; the builtins are IEEE-rounded under the
// default -fhip-fp32-correctly-rounded-divide-sqrt.
__device__ __forceinline__ float sqrt_rn(float a) { return __builtin_sqrtf(a); }
__device__ __forceinline__ double sqrt_rn(double a) { return __builtin_sqrt(a); }

// sqrt site of the reference's FP64 build that calls sqrtf (src/kernels.cuh:48,
// src/actions_gpu.cu:50,203,217, src/interactions_gpu.cu:414-415)
__device__ __forceinline__ double sqrt_site(double a, int single) {
  return single ? (double)sqrt_rn((float)a) : sqrt_rn(a);
}

// ---- small symmetric / general 3x3 helpers in GLM storage m[col][row] ------------------------
struct M3 {
  double m[3][3];
};

// (A*B)[j][i] = A[0][i]*B[j][0] + A[1][i]*B[j][1] + A[2][i]*B[j][2]
__device__ __forceinline__ void mat_mul(const M3 &A, const M3 &B, M3 &R) {
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int i = 0; i < 3; i++)
      R.m[j][i] = A.m[0][i] * B.m[j][0] + A.m[1][i] * B.m[j][1] + A.m[2][i] * B.m[j][2];
}

__device__ __forceinline__ double det3(const M3 &a) {
  const double(*m)[3] = a.m;
  return +m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
         m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
         m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
}

// adjugate / determinant, cofactor order of glm::inverse (GLM 0.9.9)
__device__ __forceinline__ void inv3(const M3 &a, M3 &r) {
  const double(*m)[3] = a.m;
  double od = 1.0 / (+m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
                     m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
                     m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]));
  r.m[0][0] = +(m[1][1] * m[2][2] - m[2][1] * m[1][2]) * od;
  r.m[1][0] = -(m[1][0] * m[2][2] - m[2][0] * m[1][2]) * od;
  r.m[2][0] = +(m[1][0] * m[2][1] - m[2][0] * m[1][1]) * od;
  r.m[0][1] = -(m[0][1] * m[2][2] - m[2][1] * m[0][2]) * od;
  r.m[1][1] = +(m[0][0] * m[2][2] - m[2][0] * m[0][2]) * od;
  r.m[2][1] = -(m[0][0] * m[2][1] - m[2][0] * m[0][1]) * od;
  r.m[0][2] = +(m[0][1] * m[1][2] - m[1][1] * m[0][2]) * od;
  r.m[1][2] = -(m[0][0] * m[1][2] - m[1][0] * m[0][2]) * od;
  r.m[2][2] = +(m[0][0] * m[1][1] - m[1][0] * m[0][1]) * od;
}

// ---- Eberly closed-form eigen decomposition of a symmetric 3x3 (src/eigen_solver.cu:19-302) --
struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 d3(double a, double b, double c) { return D3{a, b, c}; }
__device__ __forceinline__ D3 cross_e(D3 u, D3 v) {
  return d3(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
}
__device__ __forceinline__ double dot_e(D3 u, D3 v) { return u.x * v.x + u.y * v.y + u.z * v.z; }
__device__ __forceinline__ D3 div_e(D3 u, double s) {
  double inv = 1.0 / s;
  return d3(u.x * inv, u.y * inv, u.z * inv);
}

__device__ __forceinline__ void eig_vec0(double a00, double a01, double a02, double a11, double a12,
                                         double a22, double ev, D3 &out) {
  D3 r0 = d3(a00 - ev, a01, a02), r1 = d3(a01, a11 - ev, a12), r2 = d3(a02, a12, a22 - ev);
  D3 c01 = cross_e(r0, r1), c02 = cross_e(r0, r2), c12 = cross_e(r1, r2);
  double d0 = dot_e(c01, c01), d1 = dot_e(c02, c02), d2 = dot_e(c12, c12);
  double dmax = d0;
  int imax = 0;
  if (d1 > dmax) {
    dmax = d1;
    imax = 1;
  }
  if (d2 > dmax) imax = 2;
  if (imax == 0) out = div_e(c01, sqrt_rn(d0));
  else if (imax == 1) out = div_e(c02, sqrt_rn(d1));
  else out = div_e(c12, sqrt_rn(d2));
}

__device__ __forceinline__ void eig_vec1(double a00, double a01, double a02, double a11, double a12,
                                         double a22, D3 w, double ev, D3 &out) {
  D3 U, V;
  if (fabs(w.x) > fabs(w.y)) {
    double il = 1.0 / sqrt_rn(w.x * w.x + w.z * w.z);
    U = d3(-w.z * il, 0., +w.x * il);
  } else {
    double il = 1.0 / sqrt_rn(w.y * w.y + w.z * w.z);
    U = d3(0., +w.z * il, -w.y * il);
  }
  V = cross_e(w, U);
  D3 AU = d3(a00 * U.x + a01 * U.y + a02 * U.z, a01 * U.x + a11 * U.y + a12 * U.z,
             a02 * U.x + a12 * U.y + a22 * U.z);
  D3 AV = d3(a00 * V.x + a01 * V.y + a02 * V.z, a01 * V.x + a11 * V.y + a12 * V.z,
             a02 * V.x + a12 * V.y + a22 * V.z);
  double m00 = U.x * AU.x + U.y * AU.y + U.z * AU.z - ev;
  double m01 = U.x * AV.x + U.y * AV.y + U.z * AV.z;
  double m11 = V.x * AV.x + V.y * AV.y + V.z * AV.z - ev;
  double b00 = fabs(m00), b01 = fabs(m01), b11 = fabs(m11);
  if (b00 >= b11) {
    if (fmax(b00, b01) > 0.) {
      if (b00 >= b01) {
        m01 /= m00;
        m00 = 1.0 / sqrt_rn(1.0 + m01 * m01);
        m01 *= m00;
      } else {
        m00 /= m01;
        m01 = 1.0 / sqrt_rn(1.0 + m00 * m00);
        m00 *= m01;
      }
      out = d3(m01 * U.x - m00 * V.x, m01 * U.y - m00 * V.y, m01 * U.z - m00 * V.z);
    } else {
      out = U;
    }
  } else {
    if (fmax(b11, b01) > 0.) {
      if (b11 >= b01) {
        m01 /= m11;
        m11 = 1.0 / sqrt_rn(1.0 + m01 * m01);
        m01 *= m11;
      } else {
        m11 /= m01;
        m01 = 1.0 / sqrt_rn(1.0 + m11 * m11);
        m11 *= m01;
      }
      out = d3(m11 * U.x - m01 * V.x, m11 * U.y - m01 * V.y, m11 * U.z - m01 * V.z);
    } else {
      out = U;
    }
  }
}

__device__ __forceinline__ void solve_eigen(double a00, double a01, double a02, double a11,
                                            double a12, double a22, D3 &eval, D3 &e1, D3 &e2,
                                            D3 &e3) {
  double mabs = fmax(fmax(fmax(fabs(a00), fabs(a01)), fmax(fabs(a02), fabs(a11))),
                     fmax(fabs(a12), fabs(a22)));
  if (mabs < 1e-8) {
    eval = d3(0., 0., 0.);
    e1 = d3(1., 0., 0.);
    e2 = d3(0., 1., 0.);
    e3 = d3(0., 0., 1.);
    return;
  }
  double inv = 1.0 / mabs;
  a00 *= inv; a01 *= inv; a02 *= inv; a11 *= inv; a12 *= inv; a22 *= inv;
  double nrm = a01 * a01 + a02 * a02 + a12 * a12;
  if (nrm > 0.) {
    double tr3 = (a00 + a11 + a22) / 3.0;
    double b00 = a00 - tr3, b11 = a11 - tr3, b22 = a22 - tr3;
    double den = sqrt_rn((b00 * b00 + b11 * b11 + b22 * b22 + nrm * 2.0) / 6.0);
    double c00 = b11 * b22 - a12 * a12;
    double c01 = a01 * b22 - a12 * a02;
    double c02 = a01 * a12 - b11 * a02;
    double hd = ((b00 * c00 - a01 * c01 + a02 * c02) / (den * den * den)) * 0.5;
    hd = fmin(fmax(hd, -1.0), 1.0);
    double ang = acos(hd) / 3.0;
    const double two_thirds_pi = 2.09439510239319549;
    double beta2 = cos(ang) * 2.0;
    double beta0 = cos(ang + two_thirds_pi) * 2.0;
    double beta1 = -(beta0 + beta2);
    eval.x = tr3 + den * beta0;
    eval.y = tr3 + den * beta1;
    eval.z = tr3 + den * beta2;
    if (hd >= 0.) {
      eig_vec0(a00, a01, a02, a11, a12, a22, eval.z, e3);
      eig_vec1(a00, a01, a02, a11, a12, a22, e3, eval.y, e2);
      e1 = cross_e(e2, e3);
    } else {
      eig_vec0(a00, a01, a02, a11, a12, a22, eval.x, e1);
      eig_vec1(a00, a01, a02, a11, a12, a22, e1, eval.y, e2);
      e3 = cross_e(e1, e2);
    }
  } else {
    eval = d3(a00, a11, a22);
    e1 = d3(1., 0., 0.);
    e2 = d3(0., 1., 0.);
    e3 = d3(0., 0., 1.);
  }
  eval.x *= mabs;
  eval.y *= mabs;
  eval.z *= mabs;
}

// Gray artificial stress (src/actions_gpu.cu:70-132): R = Rot * Srot * transpose(Rot) in GLM
// storage, Rot built column-major from the eigenvectors; out = xx, yy, zz, xy, yz, xz (R[0][1],
// R[1][2], R[0][2] -- the entries the pair loop reads, src/interactions_gpu.cu:380-385)
__device__ __forceinline__ void art_stress(double eps, double rho, double p, const double S[6],
                                           double out[6]) {
  double rho21 = 1. / (rho * rho);
  D3 ev, e1, e2, e3;
  solve_eigen(S[0] - p, S[3], S[5], S[1] - p, S[4], S[2] - p, ev, e1, e2, e3);
  M3 rot, srot, tmp, rt, r;
  rot.m[0][0] = e1.x; rot.m[0][1] = e2.x; rot.m[0][2] = e3.x;
  rot.m[1][0] = e1.y; rot.m[1][1] = e2.y; rot.m[1][2] = e3.y;
  rot.m[2][0] = e1.z; rot.m[2][1] = e2.z; rot.m[2][2] = e3.z;
  double d[3] = {ev.x, ev.y, ev.z};
#pragma unroll
  for (int k = 0; k < 3; k++) d[k] = (d[k] > 0) ? -eps * d[k] * rho21 : 0.;
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int q = 0; q < 3; q++) {
      srot.m[c][q] = (c == q) ? d[c] : 0.;
      rt.m[c][q] = rot.m[q][c];
    }
  mat_mul(rot, srot, tmp);
  mat_mul(tmp, rt, r);
  out[0] = r.m[0][0];
  out[1] = r.m[1][1];
  out[2] = r.m[2][2];
  out[3] = r.m[0][1];
  out[4] = r.m[1][2];
  out[5] = r.m[0][2];
}

// Jaumann rate 2G(eps_dot - tr/3 I) + Omega S + S Omega^T (src/actions_gpu.cu:134-190) in GLM
// storage; vd[c][r] = v_der[c][r] = dv_c/dx_r; S as xx yy zz xy yz zx; out the same 6 entries
// (out xy = [0][1], yz = [1][2], zx = [2][0])
__device__ __forceinline__ void jaumann(double G, const M3 &vd, const double S6[6], double out[6]) {
  double vx_x = vd.m[0][0], vx_y = vd.m[0][1], vx_z = vd.m[0][2];
  double vy_x = vd.m[1][0], vy_y = vd.m[1][1], vy_z = vd.m[1][2];
  double vz_x = vd.m[2][0], vz_y = vd.m[2][1], vz_z = vd.m[2][2];
  M3 ed, om, S, t1, os, so, omt;
  ed.m[0][0] = vx_x; ed.m[0][1] = 0.5 * (vx_y + vy_x); ed.m[0][2] = 0.5 * (vx_z + vz_x);
  ed.m[1][0] = 0.5 * (vx_y + vy_x); ed.m[1][1] = vy_y; ed.m[1][2] = 0.5 * (vy_z + vz_y);
  ed.m[2][0] = 0.5 * (vx_z + vz_x); ed.m[2][1] = 0.5 * (vy_z + vz_y); ed.m[2][2] = vz_z;
  om.m[0][0] = 0.; om.m[0][1] = 0.5 * (vy_x - vx_y); om.m[0][2] = 0.5 * (vz_x - vx_z);
  om.m[1][0] = 0.5 * (vx_y - vy_x); om.m[1][1] = 0.; om.m[1][2] = 0.5 * (vz_y - vy_z);
  om.m[2][0] = 0.5 * (vx_z - vz_x); om.m[2][1] = 0.5 * (vy_z - vz_y); om.m[2][2] = 0.;
  S.m[0][0] = S6[0]; S.m[0][1] = S6[3]; S.m[0][2] = S6[5];
  S.m[1][0] = S6[3]; S.m[1][1] = S6[1]; S.m[1][2] = S6[4];
  S.m[2][0] = S6[5]; S.m[2][1] = S6[4]; S.m[2][2] = S6[2];
  double tr = ed.m[0][0] + ed.m[1][1] + ed.m[2][2];
  double s = (1. / 3.) * tr;
  double g2 = 2.0 * G;
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int r = 0; r < 3; r++) {
      t1.m[c][r] = g2 * (ed.m[c][r] - s * ((c == r) ? 1. : 0.));
      omt.m[c][r] = om.m[r][c];
    }
  mat_mul(om, S, os);
  mat_mul(S, omt, so);
  out[0] = t1.m[0][0] + os.m[0][0] + so.m[0][0];
  out[1] = t1.m[1][1] + os.m[1][1] + so.m[1][1];
  out[2] = t1.m[2][2] + os.m[2][2] + so.m[2][2];
  out[3] = t1.m[0][1] + os.m[0][1] + so.m[0][1];
  out[4] = t1.m[1][2] + os.m[1][2] + so.m[1][2];
  out[5] = t1.m[2][0] + os.m[2][0] + so.m[2][0];
}

// ---- Johnson-Cook (src/plasticity.cuh:28-93) -------------------------------------------------
__device__ __forceinline__ double sigma_yield(const Consts &c, double eps_pl, double eps_pl_dot,
                                              double t) {
  double theta = (t - c.jc_Tref) / (c.jc_Tmelt - c.jc_Tref);
  double ta = c.jc_A + c.jc_B * pow(eps_pl, c.jc_n);
  double tb;
  double ed = eps_pl_dot / c.jc_eps_dot_ref;
  if (ed > 1.0) tb = 1.0 + c.jc_C * log(ed);
  else tb = pow((1.0 + ed), c.jc_C);
  double tc = 1.0 - pow(theta, c.jc_m);
  return ta * tb * tc;
}

__device__ __forceinline__ double opt_fun(const Consts &c, double dl, double norm, double e0,
                                          double t, double dt, double mu) {
  const double s23 = sqrt_rn(2.0 / 3.0);
  double eq = e0 + s23 * fmax(dl, 0.);
  double eqd = s23 * fmax(dl, 0.) / dt;
  double sy = sigma_yield(c, eq, eqd, t);
  return norm - 2 * mu * dl - s23 * sy;
}

__device__ __forceinline__ double solve_secant(const Consts &c, double init, double tol,
                                               double norm, double e0, double t, double dt,
                                               double mu, bool &failed) {
  double dl = init, dl_old;
  int iter = 0;
  while (true) {
    double g = opt_fun(c, dl, norm, e0, t, dt, mu);
    double dl1 = dl * 1.01;
    double g1 = opt_fun(c, dl1, norm, e0, t, dt, mu);
    double slope = (g - g1) / (dl - dl1);
    dl_old = dl;
    dl = dl - g / slope;
    if (dl < 0.) dl = -0.1 * dl;
    if (fabs(dl - dl_old) < tol) return dl;
    if (iter > 100) {
      failed = true;
      return dl;
    }
    iter++;
  }
}

// Johnson-Cook radial return of one particle (src/actions_gpu.cu:371-449,
// src/plasticity.cuh:28-93): S (xx yy zz xy yz zx) after the step is returned to the yield surface
// when its von Mises stress reaches the yield stress at (eps_pl, eps_pl_dot, T); eps_pl,
// eps_pl_dot and, with plastic heating, T follow.  rho: the density after the step.
__device__ __forceinline__ void jc_radial_return(const Consts &c, double dt, const State &st,
                                                 unsigned i, double S[6], double &T, double rho) {
  double sxx = S[0], syy = S[1], szz = S[2], sxy = S[3], syz = S[4], szx = S[5];
  double nrm = sqrt_rn(sxx * sxx + syy * syy + szz * szz + 2 * (sxy * sxy + syz * syz + szx * szx));
  double ep = st.eps_pl[i], epd = st.eps_pl_dot[i];
  double Tc = T;
  if (c.jc_clamp != 0.) Tc = (Tc > c.jc_Tmelt) ? c.jc_Tmelt - 1 : Tc;
  double svm = sqrt_rn(3.0 / 2.0) * nrm;
  double sy = sigma_yield(c, ep, epd, Tc);
  if (svm < sy) {
    st.eps_pl_dot[i] = 0.;
    return;
  }
  bool failed = false;
  double dl = solve_secant(c, fmax(epd * dt * sqrt_rn(2. / 3.), 1e-8), 1e-6, nrm, ep, Tc, dt, c.a_G,
                           failed);
  double ep_new, epd_new;
  if (c.single_sqrt) {
    float f = sqrt_rn((float)(2.0 / 3.0)) * fmaxf((float)dl, 0.f);
    ep_new = ep + (double)f;
    epd_new = (double)f / dt;
  } else {
    double f = sqrt_rn(2.0 / 3.0) * fmax(dl, 0.);
    ep_new = ep + f;
    epd_new = sqrt_rn(2.0 / 3.0) * fmax(dl, 0.) / dt;
  }
  st.eps_pl[i] = ep_new;
  st.eps_pl_dot[i] = epd_new;
  const double mu = c.a_G;
  S[0] = sxx - sxx / nrm * dl * 2. * mu;
  S[1] = syy - syy / nrm * dl * 2. * mu;
  S[2] = szz - szz / nrm * dl * 2. * mu;
  S[3] = sxy - sxy / nrm * dl * 2. * mu;
  S[4] = syz - syz / nrm * dl * 2. * mu;
  S[5] = szx - szx / nrm * dl * 2. * mu;
  if (c.twp_tq != 0.) {
    double sy2 = sigma_yield(c, ep_new, epd_new, Tc);
    T += c.twp_tq / (c.twp_cp * rho) * (ep_new - ep) * sy2;
  }
}

// ---- cubic spline (src/kernels.cuh:26-69), returns W and grad W ------------------------------
__device__ __forceinline__ void cubic_spline64(double xij, double yij, double zij, double rr2,
                                               double h1, double fac, int single, double &w,
                                               double &wx, double &wy, double &wz) {
  double rad = sqrt_site(rr2, single);
  double q = rad * h1;
  double val, der;
  if (q > 1) {
    double a = 2 - q;
    double a2 = a * a;
    val = 0.25 * a2 * a;
    der = -0.75 * a2 * h1 / rad;
  } else {
    val = 1 - 1.5 * q * q * (1 - 0.5 * q);
    der = -3.0 * q * (1 - 0.75 * q) * h1 / rad;
  }
  w = val * fac;
  wx = der * xij * fac;
  wy = der * yij * fac;
  wz = der * zij * fac;
}

// ---- triangle geometry (src/tool_3d_gpu.cu:101-280) ------------------------------------------
struct V3 {
  double x, y, z;
};
__device__ __forceinline__ V3 v3(double a, double b, double c) { return V3{a, b, c}; }
__device__ __forceinline__ V3 vload(const double *p) { return V3{p[0], p[1], p[2]}; }
__device__ __forceinline__ V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 vadd(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 vscale(double s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
// glm::dot (a*b then x + y + z)
__device__ __forceinline__ double vdot(V3 a, V3 b) {
  double x = a.x * b.x, y = a.y * b.y, z = a.z * b.z;
  return x + y + z;
}
__device__ __forceinline__ V3 vcross(V3 a, V3 b) {
  return v3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
__device__ __forceinline__ bool veq(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
__device__ __forceinline__ double vlen(V3 a) { return sqrt_rn(vdot(a, a)); }

__device__ __forceinline__ V3 closest_on_segment(V3 p1, V3 p2, V3 p) {
  double dx = p1.x - p2.x, dy = p1.y - p2.y, dz = p1.z - p2.z;
  double ns2 = dx * dx + dy * dy + dz * dz;
  if (ns2 == 0.) return p1;
  double t = ((p2.x - p1.x) * (p.x - p1.x) + (p2.y - p1.y) * (p.y - p1.y) +
              (p2.z - p1.z) * (p.z - p1.z)) /
             ns2;
  if (t > 1.) return p2;
  if (t < 0.) return p1;
  return vadd(vscale(1 - t, p1), vscale(t, p2));
}

__device__ __forceinline__ double dist2(V3 a, V3 b) {
  double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  return dx * dx + dy * dy + dz * dz;
}

// closest point on a triangle and the normal at it (face, edge or vertex normal), gts algorithm
__device__ __forceinline__ void closest_point(const Tri &t, V3 qp, V3 &cp, V3 &np) {
  V3 p1 = vload(t.p[0]), p2 = vload(t.p[1]), p3 = vload(t.p[2]);
  V3 p1p2 = vsub(p2, p1), p1p3 = vsub(p3, p1), pp1 = vsub(p1, qp);
  double B = vdot(p1p3, p1p2), E = vdot(p1p2, p1p2), C = vdot(p1p3, p1p3);
  double det = B * B - E * C;
  double A = vdot(p1p3, pp1), D = vdot(p1p2, pp1);
  double t1 = (D * C - A * B) / det;
  double t2 = (A * E - D * B) / det;
  if (t1 < 0.) {
    cp = closest_on_segment(p1, p3, qp);
    np = veq(cp, p1) ? vload(t.np[0]) : (veq(cp, p3) ? vload(t.np[2]) : vload(t.ne[2]));
  } else if (t2 < 0.) {
    cp = closest_on_segment(p1, p2, qp);
    np = veq(cp, p1) ? vload(t.np[0]) : (veq(cp, p2) ? vload(t.np[1]) : vload(t.ne[0]));
  } else if (t1 + t2 > 1.) {
    cp = closest_on_segment(p2, p3, qp);
    np = veq(cp, p2) ? vload(t.np[1]) : (veq(cp, p3) ? vload(t.np[2]) : vload(t.ne[1]));
  } else {
    cp = vadd(vadd(p1, vscale(t1, p1p2)), vscale(t2, p1p3));
    np = vload(t.n);
  }
}

// Moller-Trumbore with the +z ray used by the inside tests
__device__ __forceinline__ bool ray_hit(const Tri &tr, V3 o, V3 dir, double &t) {
  const double EPS = 0.0000001;
  V3 v0 = vload(tr.p[0]);
  V3 e1 = vsub(vload(tr.p[1]), v0), e2 = vsub(vload(tr.p[2]), v0);
  V3 h = vcross(dir, e2);
  double a = vdot(e1, h);
  if (a > -EPS && a < EPS) return false;
  double f = 1. / a;
  V3 s = vsub(o, v0);
  double u = f * (vdot(s, h));
  if (u < 0.0 || u > 1.0) return false;
  V3 q = vcross(s, e1);
  double v = f * vdot(dir, q);
  if (v < 0.0 || u + v > 1.0) return false;
  t = f * vdot(e2, q);
  return (t > EPS);
}

}  // namespace mfg
