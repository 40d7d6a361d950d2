// mfg_records.hpp -- the per-particle records a step's pair passes read (built by the permute +
// predict kernel of the fused step, and from the state arrays by the per-action entry points).
#pragma once

#include "mfg_device.hpp"

namespace mfg {

// what the CSPM sum reads of a neighbour, 32 bytes: FP64 (x, y, z, vol), the volume of a tool
// particle replaced by the NaN pattern CSPM_TOOL_BITS (no volume m / rho has that payload, so a
// workpiece particle of negative or non-finite density still counts as the reference counts it);
// FP32 the record's first 32 bytes (x, y, z, {tool, vol})
template <class RecT>
__device__ __forceinline__ double4 cspm_entry(const double *xh, double vol, bool tool);
template <>
__device__ __forceinline__ double4 cspm_entry<Rec64>(const double *xh, double vol, bool tool) {
  return make_double4(xh[0], xh[1], xh[2], tool ? __longlong_as_double(CSPM_TOOL_BITS) : vol);
}
template <>
__device__ __forceinline__ double4 cspm_entry<Rec32>(const double *xh, double vol, bool tool) {
  struct {
    float tool, vol;
  } tv = {tool ? 1.f : 0.f, (float)vol};
  return make_double4(xh[0], xh[1], xh[2], __builtin_bit_cast(double, tv));
}

template <class RecT>
__device__ __forceinline__ void store_rec(RecT *rec, unsigned i, const double *xh, double h,
                                          const double *vh, double rho, const double *A,
                                          const double *R, double T, double vol, double c,
                                          double tool);

template <>
__device__ __forceinline__ void store_rec<Rec64>(Rec64 *rec, unsigned i, const double *xh,
                                                 double h, const double *vh, double rho,
                                                 const double *A, const double *R, double T,
                                                 double vol, double c, double tool) {
  Rec64 r;
  r.x = xh[0]; r.y = xh[1]; r.z = xh[2]; r.h = h;
  r.vx = vh[0]; r.vy = vh[1]; r.vz = vh[2]; r.rho = rho;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    r.A[k] = A[k];
    r.R[k] = R[k];
  }
  r.T = T; r.vol = vol; r.c = c; r.tool = tool;
  rec[i] = r;
}

template <>
__device__ __forceinline__ void store_rec<Rec32>(Rec32 *rec, unsigned i, const double *xh,
                                                 double h, const double *vh, double rho,
                                                 const double *A, const double *R, double T,
                                                 double vol, double c, double tool) {
  Rec32 r;
  r.x = xh[0]; r.y = xh[1]; r.z = xh[2];
  r.h = (float)h; r.vx = (float)vh[0]; r.vy = (float)vh[1]; r.vz = (float)vh[2];
  r.rho = (float)rho;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    r.A[k] = (float)A[k];
    r.R[k] = (float)R[k];
  }
  r.T = (float)T; r.vol = (float)vol; r.c = (float)c; r.tool = (float)tool; r.pad = 0.f;
  rec[i] = r;
}

// The records of particle i from its half-step (or current) state: the pair record (FP64 or FP32
// layout), what the CSPM sum reads, the neighbour pass's candidate entry and the PSE entry.
// p and R: pressure and artificial stress (zero for blanked and tool particles).
template <class RecT>
__device__ __forceinline__ void emit_record(unsigned i, const double xh[3], double h,
                                            const double vh[3], double rhoh, const double Sh[6],
                                            double p, const double R[6], double Th, bool tool,
                                            bool blanked, const Consts &c, RecT *rec, double4 *nl4,
                                            double2 *pse2, double4 *cs4) {
  double rho21 = 1. / (rhoh * rhoh);
  double A[6] = {(Sh[0] - p) * rho21, (Sh[1] - p) * rho21, (Sh[2] - p) * rho21,
                 Sh[3] * rho21,       Sh[4] * rho21,       Sh[5] * rho21};
  double vol = c.i_mass / rhoh;
  double cs = sqrt_site(c.i_K / rhoh, c.single_sqrt);
  store_rec<RecT>(rec, i, xh, h, vh, rhoh, A, R, Th, vol, cs, tool ? 1.0 : 0.0);
  cs4[i] = cspm_entry<RecT>(xh, vol, tool);
  // candidate flag of the neighbour pass: 1 blanked, 2 a non-finite value among the fields a
  // neighbour contributes (the reference's loops visit every candidate of the stencil, so
  // 0 * NaN carries such a value into the sums of every particle around it)
  bool finite = isfinite(vol) && isfinite(Th);
#pragma unroll
  for (int k = 0; k < 3; k++) finite = finite && isfinite(xh[k]) && isfinite(vh[k]);
#pragma unroll
  for (int k = 0; k < 6; k++) finite = finite && isfinite(A[k]) && isfinite(R[k]);
  nl4[i] = make_double4(xh[0], xh[1], xh[2], blanked ? 1.0 : (finite ? 0.0 : 2.0));
  if (pse2) pse2[i] = make_double2(Th, c.i_mass / rhoh);  // (T, m / rho) of the PSE sum
}

}  // namespace mfg
