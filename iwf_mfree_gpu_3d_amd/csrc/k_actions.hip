// k_actions.hip -- the reference's per-function entry points, one kernel each over a particle
// set's current state: the actions (src/actions_gpu.h:41-80), the pair passes on the current
// state (interactions_monaghan / interactions_heat_pse, src/interactions_gpu.h:51-52), one tool's
// contact force (tool_3d_gpu::compute_contact_force, src/tool_3d_gpu.h:120) and the leap-frog
// phases (init, predict / correct, src/leap_frog.cu:30-100).  The fused step (k_grid.hip,
// k_pair.hip, k_tool.hip) does the same work in fewer passes; these let a caller compose the
// reference's step, or an integrator of its own, call by call.  The arithmetic is the same
// device code: art_stress, jaumann, jc_radial_return, emit_record, tool_contact.
#include "mfg_device.hpp"
#include "mfg_records.hpp"

namespace mfg {

constexpr int BA = 256;

__device__ __forceinline__ bool skip_blanked_tool(uint32_t fl) {
  return (fl & FL_BLANKED) || (fl & FL_TOOL);
}

// material_eos (src/actions_gpu.cu:40-53)
__global__ void __launch_bounds__(BA) k_act_eos(State st, const Counters *cnt, Consts c, Aux a) {
  const unsigned i = blockIdx.x * BA + threadIdx.x;
  if (i >= cnt->N || skip_blanked_tool(st.flags[i])) return;
  const double c0 = sqrt_site(c.a_K / c.a_rho0, c.single_sqrt);
  a.p[i] = c0 * c0 * (st.rho[i] - c.a_rho0);
}

// corrector_artificial_stress (src/actions_gpu.cu:70-132)
__global__ void __launch_bounds__(BA) k_act_art_stress(State st, const Counters *cnt, Consts c,
                                                       Aux a) {
  const unsigned i = blockIdx.x * BA + threadIdx.x;
  if (i >= cnt->N || skip_blanked_tool(st.flags[i])) return;
  double S[6], R[6];
#pragma unroll
  for (int k = 0; k < 6; k++) S[k] = st.S[k][i];
  art_stress(c.a_stresseps, st.rho[i], a.p[i], S, R);
#pragma unroll
  for (int k = 0; k < 6; k++) a.R[6 * (size_t)i + k] = R[k];
}

__device__ __forceinline__ M3 load_m3(const double *m, size_t n, unsigned i) {
  M3 r;
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) r.m[a][b] = m[(3 * a + b) * n + i];
  return r;
}

// material_stress_rate_jaumann, contmech_continuity, contmech_momentum, contmech_advection
// (src/actions_gpu.cu:134-190,222-300), selected by `which`
__global__ void __launch_bounds__(BA) k_act_rates(State st, const Counters *cnt, Consts c, Aux a,
                                                  int which) {
  const unsigned i = blockIdx.x * BA + threadIdx.x;
  if (i >= cnt->N || skip_blanked_tool(st.flags[i])) return;
  if (which == 0) {
    double S[6], St[6];
#pragma unroll
    for (int k = 0; k < 6; k++) S[k] = st.S[k][i];
    jaumann(c.a_G, load_m3(a.vder, a.n, i), S, St);
#pragma unroll
    for (int k = 0; k < 6; k++) st.St[k][i] = St[k];
  } else if (which == 1) {
    const M3 vd = load_m3(a.vder, a.n, i);
    st.rho_t[i] = -st.rho[i] * (vd.m[0][0] + vd.m[1][1] + vd.m[2][2]);
  } else if (which == 2) {
    const M3 sd = load_m3(a.sder, a.n, i);
    const double m = c.a_mass;
    st.vxt[i] = st.vxt[i] + (sd.m[0][0] + sd.m[0][1] + sd.m[0][2] + st.fcx[i] / m + st.ftx[i] / m);
    st.vyt[i] = st.vyt[i] + (sd.m[1][0] + sd.m[1][1] + sd.m[1][2] + st.fcy[i] / m + st.fty[i] / m);
    st.vzt[i] = st.vzt[i] + (sd.m[2][0] + sd.m[2][1] + sd.m[2][2] + st.fcz[i] / m + st.ftz[i] / m);
  } else {
    st.xt[i] = st.xt[i] + st.vx[i];
    st.yt[i] = st.yt[i] + st.vy[i];
    st.zt[i] = st.zt[i] + st.vz[i];
  }
}

// debug_invalidate (src/actions_gpu.cu:302-331): non-finite rates blank the particle
__global__ void __launch_bounds__(BA) k_act_invalidate(State st, const Counters *cnt) {
  const unsigned i = blockIdx.x * BA + threadIdx.x;
  if (i >= cnt->N) return;
  auto bad = [](double v) { return isnan(v) || isinf(v); };
  bool inv = bad(st.xt[i]) || bad(st.yt[i]) || bad(st.zt[i]) || bad(st.vxt[i]) || bad(st.vyt[i]) ||
             bad(st.vzt[i]) || bad(st.rho_t[i]) || bad(st.T_t[i]);
#pragma unroll
  for (int k = 0; k < 6; k++) inv = inv || bad(st.St[k][i]);
  if (inv) st.flags[i] |= FL_BLANKED;
}

// plasticity_johnson_cook (src/actions_gpu.cu:371-449)
__global__ void __launch_bounds__(BA) k_act_jc(State st, const Counters *cnt, Consts c) {
  const unsigned i = blockIdx.x * BA + threadIdx.x;
  if (i >= cnt->N || skip_blanked_tool(st.flags[i])) return;
  double S[6];
#pragma unroll
  for (int k = 0; k < 6; k++) S[k] = st.S[k][i];
  double T = st.T[i];
  jc_radial_return(c, c.dt, st, i, S, T, st.rho[i]);
#pragma unroll
  for (int k = 0; k < 6; k++) st.S[k][i] = S[k];
  st.T[i] = T;
}

// material_fric_heat_gen (src/actions_gpu.cu:192-220,712-723): vt the first tool's velocity
__global__ void __launch_bounds__(BA) k_act_fric_heat(State st, const Counters *cnt, Consts c,
                                                      Aux a, double v0, double v1, double v2) {
  const unsigned i = blockIdx.x * BA + threadIdx.x;
  if (i >= cnt->N || skip_blanked_tool(st.flags[i])) return;
  const double fx = st.ftx[i], fy = st.fty[i], fz = st.ftz[i];
  const double fmag = sqrt_site(fx * fx + fy * fy + fz * fz, c.single_sqrt);
  if (fmag == 0.) return;
  const double nx = a.nrm[i], ny = a.nrm[a.n + i], nz = a.nrm[2 * a.n + i];
  const double dx = st.vx[i] - v0, dy = st.vy[i] - v1, dz = st.vz[i] - v2;
  const double vdn = dx * nx + dy * ny + dz * nz;
  const double rx = dx - vdn, ry = dy - vdn, rz = dz - vdn;
  const double vmag = sqrt_site(rx * rx + ry * ry + rz * rz, c.single_sqrt);
  st.T[i] += c.twp_eta * c.dt * fmag * vmag / (c.twp_cp * c.a_mass);
}

// perform_boundary_conditions / _thermal (src/actions_gpu.cu:451-487)
__global__ void __launch_bounds__(BA) k_act_bc(State st, const Counters *cnt, Consts c, int thermal) {
  const unsigned i = blockIdx.x * BA + threadIdx.x;
  if (i >= cnt->N) return;
  const uint32_t fl = st.flags[i];
  if (thermal) {
    if (fl & FL_FIXED_ONE) st.T[i] = c.twp_T_init;
    return;
  }
  if (skip_blanked_tool(fl) || !(fl & FL_FIXED)) return;
  st.vx[i] = st.vbx[i]; st.vy[i] = st.vby[i]; st.vz[i] = st.vbz[i];
  st.fcx[i] = 0.; st.fcy[i] = 0.; st.fcz[i] = 0.;
  st.xt[i] = 0.; st.yt[i] = 0.; st.zt[i] = 0.;
  st.vxt[i] = 0.; st.vyt[i] = 0.; st.vzt[i] = 0.;
}

// actions_move_tool_particles (src/actions_gpu.cu:55-68,620-629), over N_init
__global__ void __launch_bounds__(BA) k_act_move_tool(State st, unsigned n_init, double dt,
                                                      double v0, double v1, double v2) {
  const unsigned i = blockIdx.x * BA + threadIdx.x;
  if (i >= n_init || !(st.flags[i] & FL_TOOL)) return;
  st.x[i] = st.x[i] + v0 * dt;
  st.y[i] = st.y[i] + v1 * dt;
  st.z[i] = st.z[i] + v2 * dt;
}

// reset of the contact forces and normals (src/leap_frog.cu:102-119)
__global__ void __launch_bounds__(BA) k_act_reset_contact(State st, const Counters *cnt, Aux a) {
  const unsigned i = blockIdx.x * BA + threadIdx.x;
  if (i >= cnt->N) return;
  st.fcx[i] = 0.; st.fcy[i] = 0.; st.fcz[i] = 0.;
  st.ftx[i] = 0.; st.fty[i] = 0.; st.ftz[i] = 0.;
  a.nrm[i] = 0.; a.nrm[a.n + i] = 0.; a.nrm[2 * a.n + i] = 0.;
}

// leap-frog init (start-of-step copies) and advance x = x0 + f x_t (predict f = dt / 2, correct
// f = dt) of every unblanked particle (src/leap_frog.cu:30-100)
__global__ void __launch_bounds__(BA) k_lf_init(State st, unsigned n_init, Aux a) {
  const unsigned i = blockIdx.x * BA + threadIdx.x;
  if (i >= n_init || (st.flags[i] & FL_BLANKED)) return;
  const size_t n = a.n;
  double *x0 = a.x0;
  x0[i] = st.x[i]; x0[n + i] = st.y[i]; x0[2 * n + i] = st.z[i];
  x0[3 * n + i] = st.vx[i]; x0[4 * n + i] = st.vy[i]; x0[5 * n + i] = st.vz[i];
#pragma unroll
  for (int k = 0; k < 6; k++) x0[(6 + k) * n + i] = st.S[k][i];
  x0[12 * n + i] = st.rho[i];
  x0[13 * n + i] = st.T[i];
}

__global__ void __launch_bounds__(BA) k_lf_advance(State st, unsigned n_init, Aux a, double f) {
  const unsigned i = blockIdx.x * BA + threadIdx.x;
  if (i >= n_init || (st.flags[i] & FL_BLANKED)) return;
  const size_t n = a.n;
  const double *x0 = a.x0;
  st.x[i] = x0[i] + f * st.xt[i];
  st.y[i] = x0[n + i] + f * st.yt[i];
  st.z[i] = x0[2 * n + i] + f * st.zt[i];
  st.vx[i] = x0[3 * n + i] + f * st.vxt[i];
  st.vy[i] = x0[4 * n + i] + f * st.vyt[i];
  st.vz[i] = x0[5 * n + i] + f * st.vzt[i];
#pragma unroll
  for (int k = 0; k < 6; k++) st.S[k][i] = x0[(6 + k) * n + i] + f * st.St[k][i];
  st.rho[i] = x0[12 * n + i] + f * st.rho_t[i];
  st.T[i] = x0[13 * n + i] + f * st.T_t[i];
}

// the pair records of the current state, with the pressure and artificial stress the actions
// computed (zero for blanked and tool particles, as the reference's arrays hold for them)
template <class RecT>
__global__ void __launch_bounds__(BA) k_act_records(State st, unsigned n_init, const Counters *cnt,
                                                    Consts c, Aux a, RecT *rec, double4 *nl4,
                                                    double2 *pse2, double4 *cs4) {
  const unsigned i = blockIdx.x * BA + threadIdx.x;
  if (i >= n_init) return;
  const uint32_t fl = st.flags[i];
  const bool blanked = (fl & FL_BLANKED) != 0, tool = (fl & FL_TOOL) != 0;
  const double x[3] = {st.x[i], st.y[i], st.z[i]};
  const double v[3] = {st.vx[i], st.vy[i], st.vz[i]};
  double S[6], R[6] = {0., 0., 0., 0., 0., 0.};
#pragma unroll
  for (int k = 0; k < 6; k++) S[k] = st.S[k][i];
  double p = 0.;
  if (i < cnt->N && !blanked && !tool) {
    p = a.p[i];
#pragma unroll
    for (int k = 0; k < 6; k++) R[k] = a.R[6 * (size_t)i + k];
  }
  emit_record<RecT>(i, x, st.h[i], v, st.rho[i], S, p, R, st.T[i], tool, blanked, c, rec, nl4, pse2,
                    cs4);
}

// the reference's fields of interactions_monaghan from the pair sums: v_der, S_der, pos_t
// (XSPH), vel_t and, with Brookshaw conduction, T_t (src/interactions_gpu.cu:152-471)
__global__ void __launch_bounds__(BA) k_act_sums(State st, const Counters *cnt, Consts c, Aux a,
                                                 const double *sums, size_t stride,
                                                 const int *poison) {
  const unsigned i = blockIdx.x * BA + threadIdx.x;
  if (i >= cnt->N) return;
  const uint32_t fl = st.flags[i];
  if (fl & FL_BLANKED) return;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    a.vder[k * a.n + i] = sums[k * stride + i];
    a.sder[k * a.n + i] = sums[(9 + k) * stride + i];
  }
  // a non-finite record in the stencil carries into the sums (0 * NaN in the reference)
  if (poison[i]) a.vder[i] = __builtin_nan("");
  st.vxt[i] = sums[18 * stride + i];
  st.vyt[i] = sums[19 * stride + i];
  st.vzt[i] = sums[20 * stride + i];
  st.xt[i] = sums[21 * stride + i];
  st.yt[i] = sums[22 * stride + i];
  st.zt[i] = sums[23 * stride + i];
  const double alpha = (fl & FL_TOOL_NZ) ? c.alpha_tool : c.alpha_wp;
  if (c.heat == MFG_HEAT_BROOKSHAW && alpha != 0.) st.T_t[i] = alpha * sums[24 * stride + i];
}

// ---------------------------------------------------------------------------------------------
static inline unsigned nba(long n) { return (unsigned)((n + BA - 1) / BA); }

void launch_action(hipStream_t s, int which, const State &st, unsigned n_init, const Counters *cnt,
                   const Consts &c, const Aux &a, const double *vel) {
  const dim3 g(nba(n_init)), b(BA);
  switch (which) {
    case ACT_EOS: hipLaunchKernelGGL(k_act_eos, g, b, 0, s, st, cnt, c, a); break;
    case ACT_ART_STRESS: hipLaunchKernelGGL(k_act_art_stress, g, b, 0, s, st, cnt, c, a); break;
    case ACT_JAUMANN: hipLaunchKernelGGL(k_act_rates, g, b, 0, s, st, cnt, c, a, 0); break;
    case ACT_CONTINUITY: hipLaunchKernelGGL(k_act_rates, g, b, 0, s, st, cnt, c, a, 1); break;
    case ACT_MOMENTUM: hipLaunchKernelGGL(k_act_rates, g, b, 0, s, st, cnt, c, a, 2); break;
    case ACT_ADVECTION: hipLaunchKernelGGL(k_act_rates, g, b, 0, s, st, cnt, c, a, 3); break;
    case ACT_INVALIDATE: hipLaunchKernelGGL(k_act_invalidate, g, b, 0, s, st, cnt); break;
    case ACT_JOHNSON_COOK: hipLaunchKernelGGL(k_act_jc, g, b, 0, s, st, cnt, c); break;
    case ACT_FRIC_HEAT:
      hipLaunchKernelGGL(k_act_fric_heat, g, b, 0, s, st, cnt, c, a, vel[0], vel[1], vel[2]);
      break;
    case ACT_BC: hipLaunchKernelGGL(k_act_bc, g, b, 0, s, st, cnt, c, 0); break;
    case ACT_BC_THERMAL: hipLaunchKernelGGL(k_act_bc, g, b, 0, s, st, cnt, c, 1); break;
    case ACT_MOVE_TOOL:
      hipLaunchKernelGGL(k_act_move_tool, g, b, 0, s, st, n_init, c.dt, vel[0], vel[1], vel[2]);
      break;
    case ACT_RESET_CONTACT: hipLaunchKernelGGL(k_act_reset_contact, g, b, 0, s, st, cnt, a); break;
    case ACT_LF_INIT: hipLaunchKernelGGL(k_lf_init, g, b, 0, s, st, n_init, a); break;
    default: break;
  }
}

void launch_lf_advance(hipStream_t s, const State &st, unsigned n_init, const Aux &a, double f) {
  hipLaunchKernelGGL(k_lf_advance, dim3(nba(n_init)), dim3(BA), 0, s, st, n_init, a, f);
}

void launch_act_records(hipStream_t s, const State &st, unsigned n_init, const Counters *cnt,
                        const Consts &c, const Aux &a, void *rec, double4 *nl4, double2 *pse2,
                        double4 *cs4) {
  const dim3 g(nba(n_init)), b(BA);
  if (c.precision == MFG_FP64)
    hipLaunchKernelGGL(k_act_records<Rec64>, g, b, 0, s, st, n_init, cnt, c, a, (Rec64 *)rec, nl4,
                       pse2, cs4);
  else
    hipLaunchKernelGGL(k_act_records<Rec32>, g, b, 0, s, st, n_init, cnt, c, a, (Rec32 *)rec, nl4,
                       pse2, cs4);
}

void launch_act_sums(hipStream_t s, const State &st, unsigned n_init, const Counters *cnt,
                     const Consts &c, const Aux &a, const double *sums, size_t stride,
                     const int *poison) {
  hipLaunchKernelGGL(k_act_sums, dim3(nba(n_init)), dim3(BA), 0, s, st, cnt, c, a, sums, stride,
                     poison);
}

}  // namespace mfg
