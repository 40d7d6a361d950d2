// k_tool.hip -- tool mesh on the GPU: triangle binning, motion, contact, boundary conditions.
//
//   triangle binning          grid_cpu_tri::get_cells + mesh_triangle::cover_aabb
//                             (src/grid_cpu_tri.cpp:96-175, src/surface_tri_mesh.cpp:151-278),
//                             done here with a counting sort over cells on the device
//   normal orientation        tool_3d ctor (src/tool_3d.cpp:318-358)
//   do_update_tool            src/tool_3d_gpu.cu:74-99
//   contact + boundary pass   do_reset_contact_forces, do_compute_contact_forces[_n2],
//                             do_material_fric_heat_gen, do_boundary_conditions[_thermal],
//                             do_move_tool_particles  (src/leap_frog.cu:102-119,
//                             src/tool_3d_gpu.cu:282-681, src/actions_gpu.cu:55-68,192-220,451-487)
//   tool force recording      per-block partial sums reduced in a fixed order (the reference
//                             uses float atomics, src/tool_3d_gpu.cu:579-581)
#include <cfloat>

#include "mfg_device.hpp"

namespace mfg {

constexpr int BST = 256;

// ---- triangle / box overlap ------------------------------------------------------------------
__device__ __forceinline__ bool axis_reject(double pa, double pb, double rad) {
  double mn = pa < pb ? pa : pb, mx = pa < pb ? pb : pa;
  return (mn > rad || mx < -rad);
}

__device__ bool plane_box_overlap(V3 nrm, V3 vert, V3 mb) {
  double n[3] = {nrm.x, nrm.y, nrm.z}, v[3] = {vert.x, vert.y, vert.z}, m[3] = {mb.x, mb.y, mb.z};
  double vmin[3], vmax[3];
  for (int q = 0; q < 3; q++) {
    if (n[q] > 0.0) {
      vmin[q] = -m[q] - v[q];
      vmax[q] = m[q] - v[q];
    } else {
      vmin[q] = m[q] - v[q];
      vmax[q] = -m[q] - v[q];
    }
  }
  if (vdot(nrm, v3(vmin[0], vmin[1], vmin[2])) > 0.0) return false;
  if (vdot(nrm, v3(vmax[0], vmax[1], vmax[2])) >= 0.0) return true;
  return false;
}

__device__ bool tri_box(const Tri &t, V3 bmin, V3 bmax) {
  V3 c = vscale(0.5, vadd(bmin, bmax));
  V3 hs = vscale(0.5, vsub(bmax, bmin));
  V3 p1 = vload(t.p[0]), p2 = vload(t.p[1]), p3 = vload(t.p[2]);
  V3 v0 = vsub(p1, c), v1 = vsub(p2, c), v2 = vsub(p3, c);
  V3 e0 = vsub(p2, p1), e1 = vsub(p3, p2), e2 = vsub(p1, p3);
  double fx = fabs(e0.x), fy = fabs(e0.y), fz = fabs(e0.z);
  if (axis_reject(e0.z * v0.y - e0.y * v0.z, e0.z * v2.y - e0.y * v2.z, fz * hs.y + fy * hs.z)) return false;
  if (axis_reject(-e0.z * v0.x + e0.x * v0.z, -e0.z * v2.x + e0.x * v2.z, fz * hs.x + fx * hs.z)) return false;
  if (axis_reject(e0.y * v1.x - e0.x * v1.y, e0.y * v2.x - e0.x * v2.y, fy * hs.x + fx * hs.y)) return false;
  // the second and third edge magnitudes are taken in single precision as the reference does
  fx = fabsf((float)e1.x); fy = fabsf((float)e1.y); fz = fabsf((float)e1.z);
  if (axis_reject(e1.z * v0.y - e1.y * v0.z, e1.z * v2.y - e1.y * v2.z, fz * hs.y + fy * hs.z)) return false;
  if (axis_reject(-e1.z * v0.x + e1.x * v0.z, -e1.z * v2.x + e1.x * v2.z, fz * hs.x + fx * hs.z)) return false;
  if (axis_reject(e1.y * v0.x - e1.x * v0.y, e1.y * v1.x - e1.x * v1.y, fy * hs.x + fx * hs.y)) return false;
  fx = fabsf((float)e2.x); fy = fabsf((float)e2.y); fz = fabsf((float)e2.z);
  if (axis_reject(e2.z * v0.y - e2.y * v0.z, e2.z * v1.y - e2.y * v1.z, fz * hs.y + fy * hs.z)) return false;
  if (axis_reject(-e2.z * v0.x + e2.x * v0.z, -e2.z * v1.x + e2.x * v1.z, fz * hs.x + fx * hs.z)) return false;
  if (axis_reject(e2.y * v1.x - e2.x * v1.y, e2.y * v2.x - e2.x * v2.y, fy * hs.x + fx * hs.y)) return false;
  double mn, mx;
  mn = fmin(v0.x, fmin(v1.x, v2.x)); mx = fmax(v0.x, fmax(v1.x, v2.x));
  if (mn > hs.x || mx < -hs.x) return false;
  mn = fmin(v0.y, fmin(v1.y, v2.y)); mx = fmax(v0.y, fmax(v1.y, v2.y));
  if (mn > hs.y || mx < -hs.y) return false;
  mn = fmin(v0.z, fmin(v1.z, v2.z)); mx = fmax(v0.z, fmax(v1.z, v2.z));
  if (mn > hs.z || mx < -hs.z) return false;
  return plane_box_overlap(vcross(e0, e1), v0, hs);
}

struct TriGrid {
  double bbmin[3];
  double dx;
  int nx, ny, nz;
};

__device__ __forceinline__ void tri_cell_range(const Tri &t, const TriGrid &g, unsigned lo[3],
                                               unsigned hi[3]) {
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double mn = fmin(fmin(t.p[0][k], t.p[1][k]), t.p[2][k]);
    double mx = fmax(fmax(t.p[0][k], t.p[1][k]), t.p[2][k]);
    lo[k] = (unsigned)((mn - g.bbmin[k]) / g.dx);
    hi[k] = (unsigned)((mx - g.bbmin[k]) / g.dx);
  }
  // the grid is padded around every triangle; the clamp only guards rounding at the last cell
  const unsigned n[3] = {(unsigned)g.nx, (unsigned)g.ny, (unsigned)g.nz};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if (hi[k] >= n[k]) hi[k] = n[k] - 1;
    if (lo[k] > hi[k]) lo[k] = hi[k];
  }
}

__global__ void k_tri_bin(const Tri *tris, int n, TriGrid g, int *count, int *fill, int *out) {
  int t = blockIdx.x * BST + threadIdx.x;
  if (t >= n) return;
  const Tri tr = tris[t];
  unsigned lo[3], hi[3];
  tri_cell_range(tr, g, lo, hi);
  for (unsigned a = lo[0]; a <= hi[0]; a++)
    for (unsigned b = lo[1]; b <= hi[1]; b++)
      for (unsigned cc = lo[2]; cc <= hi[2]; cc++) {
        unsigned lin = a * (g.ny) * (g.nz) + b * (g.nz) + cc;
        V3 bl = v3(g.bbmin[0] + a * g.dx, g.bbmin[1] + b * g.dx, g.bbmin[2] + cc * g.dx);
        V3 bh = v3(g.bbmin[0] + (a + 1) * g.dx, g.bbmin[1] + (b + 1) * g.dx,
                   g.bbmin[2] + (cc + 1) * g.dx);
        if (!tri_box(tr, bl, bh)) continue;
        if (!out) atomicAdd(&count[lin], 1);
        else out[count[lin] + atomicAdd(&fill[lin], 1)] = t;
      }
}

// exclusive scan of the per-cell triangle counts (tool grids are small: one block)
__global__ void __launch_bounds__(1024) k_tri_scan(int *count, int n) {
  __shared__ int part[1024];
  int per = (n + 1023) / 1024;
  int beg = threadIdx.x * per, end = min(beg + per, n);
  int s = 0;
  for (int k = beg; k < end; k++) s += count[k];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int k = 0; k < 1024; k++) {
      int v = part[k];
      part[k] = acc;
      acc += v;
    }
    count[n] = acc;
  }
  __syncthreads();
  int off = part[threadIdx.x];
  for (int k = beg; k < end; k++) {
    int v = count[k];
    count[k] = off;
    off += v;
  }
}

__global__ void k_tri_cell_sort(const int *off, int n, int *idx) {
  int c = blockIdx.x * BST + threadIdx.x;
  if (c >= n) return;
  int s = off[c], e = off[c + 1];
  for (int a = s + 1; a < e; a++) {
    int v = idx[a], b = a - 1;
    while (b >= s && idx[b] > v) {
      idx[b + 1] = idx[b];
      b--;
    }
    idx[b + 1] = v;
  }
}

__global__ void k_gather_tris(const Tri *src, const int *idx, int n, Tri *dst) {
  int k = blockIdx.x * BST + threadIdx.x;
  if (k < n) dst[k] = src[idx[k]];
}

// ---- inside tests -------------------------------------------------------------------------
// closest point by expanding shells of the triangle grid (src/tool_3d_gpu.cu:282-380);
// returns true if inside by the normal test
__device__ bool contact_by_normal(const Tri *tris, const int *cells, int nx, int ny, int nz,
                                  int ix, int iy, int iz, V3 qp, V3 &cp, V3 &nrm) {
  int box = ix * ny * nz + iy * nz + iz;
  double min_dist = DBL_MAX;
  V3 bc = v3(0, 0, 0), bn = v3(0, 0, 0);
  bool all_empty = (cells[box] == cells[box + 1]);
  for (int p = cells[box]; p < cells[box + 1]; p++) {
    V3 c, n;
    closest_point(tris[p], qp, c, n);
    double d = dist2(c, qp);
    if (d < min_dist) {
      min_dist = d;
      bc = c;
      bn = n;
    }
  }
  int offset = 1;
  bool all_done = false;
  while (true) {
    int lo_i = ix - offset < 0 ? 0 : ix - offset, hi_i = ix + offset + 1 > nx ? nx : ix + offset + 1;
    int lo_j = iy - offset < 0 ? 0 : iy - offset, hi_j = iy + offset + 1 > ny ? ny : iy + offset + 1;
    int lo_k = iz - offset < 0 ? 0 : iz - offset, hi_k = iz + offset + 1 > nz ? nz : iz + offset + 1;
    for (int a = lo_i; a < hi_i; a++)
      for (int b = lo_j; b < hi_j; b++)
        for (int cc = lo_k; cc < hi_k; cc++) {
          bool on = (a == lo_i || a == hi_i - 1) || (b == lo_j || b == hi_j - 1) ||
                    (cc == lo_k || cc == hi_k - 1);
          if (!on) continue;
          int bi = a * ny * nz + b * nz + cc;
          for (int p = cells[bi]; p < cells[bi + 1]; p++) {
            all_empty = false;
            V3 c, n;
            closest_point(tris[p], qp, c, n);
            double d = dist2(c, qp);
            if (d < min_dist) {
              min_dist = d;
              bc = c;
              bn = n;
            }
          }
        }
    if (all_done) break;
    if (!all_empty) {
      offset += 1;
      all_done = true;
      continue;
    }
    offset += 1;
    // every shell beyond the grid is empty: the search cannot find a triangle
    if (offset > nx + ny + nz) break;
  }
  cp = bc;
  nrm = bn;
  V3 diff = vsub(bc, qp);
  return !(vdot(diff, bn) > 0.);
}

// +z ray parity inside the particle's own box column (src/tool_3d_gpu.cu:382-444)
__device__ bool contact_by_ray(const Tri *tris, const int *cells, int ny, int nz, int ix, int iy,
                               int iz, V3 qp, const double bbmin[3], double dx) {
  const V3 dir = v3(0., 0., 1.);
  unsigned n_inter = 0;
  for (; iz < nz; ++iz) {
    unsigned bi = ix * ny * nz + iy * nz + iz;
    for (int p = cells[bi]; p < cells[bi + 1]; p++) {
      double t;
      if (ray_hit(tris[p], qp, dir, t)) {
        V3 ip = vadd(qp, vscale(t, dir));
        double lo_x = bbmin[0] + ix * dx, lo_y = bbmin[1] + iy * dx, lo_z = bbmin[2] + iz * dx;
        double hi_x = bbmin[0] + (ix + 1) * dx, hi_y = bbmin[1] + (iy + 1) * dx,
               hi_z = bbmin[2] + (iz + 1) * dx;
        if (lo_x <= ip.x && ip.x < hi_x && lo_y <= ip.y && ip.y < hi_y && lo_z <= ip.z && ip.z < hi_z)
          n_inter++;
      }
    }
  }
  return (n_inter % 2) != 0;
}

// ---- closest-triangle candidate lists ----------------------------------------------------------
// The exhaustive contact search of do_compute_contact_forces_n2 (src/tool_3d_gpu.cu:501-513) takes
// the triangle of smallest squared distance over ALL triangles, the first index on ties.  Here
// every cell of the triangle grid keeps the triangles that can be that argmin for a query point
// in the cell, ascending by index, so scanning the list with a strict '<' returns the same
// triangle as the full loop.  With the cell box widened by `slack` on every side, corners c_k,
// centre m and half diagonal r:
//   ub = min_T max_k |c_k - T| bounds the closest distance of every point of the box from above
//        (the distance to a convex set is convex, so its maximum over a box is at a corner);
//   |m - T| - r <= |p - T| for every point p of the box;
// so a triangle with |m - T| - r > ub is never the closest one.  A relative margin of 1e-6 on
// the cut covers rounding.  The grid moves rigidly with the mesh: the lists hold for all time.
__device__ void cell_geometry(const TriGrid &g, int cell, double slack, V3 c[8], V3 &m,
                              double &r) {
  const int a = cell / (g.ny * g.nz), b = (cell / g.nz) % g.ny, cc = cell % g.nz;
  const double lo[3] = {g.bbmin[0] + a * g.dx - slack, g.bbmin[1] + b * g.dx - slack,
                        g.bbmin[2] + cc * g.dx - slack};
  const double w = g.dx + 2. * slack;
#pragma unroll
  for (int k = 0; k < 8; k++)
    c[k] = v3(lo[0] + ((k & 1) ? w : 0.), lo[1] + ((k & 2) ? w : 0.), lo[2] + ((k & 4) ? w : 0.));
  m = v3(lo[0] + 0.5 * w, lo[1] + 0.5 * w, lo[2] + 0.5 * w);
  r = 0.5 * w * sqrt_rn(3.);
}

__device__ __forceinline__ double tri_dist2(const Tri &t, V3 q) {
  V3 cp, n;
  closest_point(t, q, cp, n);
  return dist2(cp, q);
}

static_assert(BST == 256, "the block reductions below assume four wavefronts");

// block of 256 per cell: the cut radius^2 of the cell and the number of candidates
__global__ void __launch_bounds__(BST)
    k_cand_count(const Tri *tris, int n, TriGrid g, double slack, double *cut2, int *count) {
  __shared__ double red[BST / 64];
  __shared__ int redi[BST / 64];
  const int cell = blockIdx.x;
  V3 c[8], m;
  double r;
  cell_geometry(g, cell, slack, c, m, r);
  double best = DBL_MAX;
  for (int t = threadIdx.x; t < n; t += BST) {
    double mx = 0.;
#pragma unroll
    for (int k = 0; k < 8; k++) mx = fmax(mx, tri_dist2(tris[t], c[k]));
    best = fmin(best, mx);
  }
  for (int o = 32; o > 0; o >>= 1) best = fmin(best, __shfl_xor(best, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
  __syncthreads();
  best = fmin(fmin(red[0], red[1]), fmin(red[2], red[3]));
  const double ub = sqrt_rn(best);
  const double lim = (ub + r) * (1. + 1e-6);
  const double lim2 = lim * lim;
  int cnt = 0;
  for (int t = threadIdx.x; t < n; t += BST) cnt += tri_dist2(tris[t], m) <= lim2 ? 1 : 0;
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
  if ((threadIdx.x & 63) == 0) redi[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    count[cell] = redi[0] + redi[1] + redi[2] + redi[3];
    cut2[cell] = lim2;
  }
}

// block of 256 per cell: write the candidates in ascending index order
__global__ void __launch_bounds__(BST)
    k_cand_fill(const Tri *tris, int n, TriGrid g, double slack, const double *cut2,
                const int *off, int *idx) {
  __shared__ int wsum[BST / 64];
  const int cell = blockIdx.x;
  V3 c[8], m;
  double r;
  cell_geometry(g, cell, slack, c, m, r);
  const double lim2 = cut2[cell];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int base = off[cell];
  const int end = off[cell + 1];  // the count pass evaluated the same predicate
  for (int t0 = 0; t0 < n; t0 += BST) {
    const int t = t0 + threadIdx.x;
    const bool take = t < n && tri_dist2(tris[t], m) <= lim2;
    const unsigned long long mask = __ballot(take);
    const int before = __popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[w] = __popcll(mask);
    __syncthreads();
    int wbase = 0;
    for (int k = 0; k < w; k++) wbase += wsum[k];
    if (take && base + wbase + before < end) idx[base + wbase + before] = t;
    base += wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
  }
}

// closest triangle of the whole mesh for a point in cell `cell` (see above)
__device__ __forceinline__ void closest_by_candidates(const ToolView &tv, V3 qp, int cell, V3 &cp,
                                                      V3 &nrm) {
  double best = FLT_MAX;
  for (int p = tv.cand_off[cell]; p < tv.cand_off[cell + 1]; p++) {
    V3 c2, n2;
    closest_point(tv.tris[tv.cand_idx[p]], qp, c2, n2);
    const double d2 = dist2(c2, qp);
    if (d2 < best) {
      best = d2;
      cp = c2;
      nrm = n2;
    }
  }
}

// ---- exhaustive contact, lane-parallel search ----------------------------------------------------
__device__ __forceinline__ bool in_tool_box(const ToolView &tv, V3 qp) {
  return qp.x > tv.bbmin[0] && qp.x < tv.bbmax[0] && qp.y > tv.bbmin[1] && qp.y < tv.bbmax[1] &&
         qp.z > tv.bbmin[2] && qp.z < tv.bbmax[2];
}

__global__ void k_reset_u32(unsigned *p) {
  if (threadIdx.x == 0) *p = 0u;
}

// list every eligible particle inside the box of an active exhaustive tool (the same test as
// k_contact_bc) and remember its entry
__device__ __forceinline__ void contact_list(unsigned b, const State &st, unsigned n_init,
                                             const ToolSet &ts, const Counters *cnt,
                                             const ContactSearch &cs) {
  const unsigned i = b * BST + threadIdx.x;
  if (i >= n_init || i >= cnt->N) return;
  const uint32_t fl = st.flags[i];
  if ((fl & FL_BLANKED) || (fl & FL_TOOL)) return;
  const V3 qp = v3(st.x[i], st.y[i], st.z[i]);
  for (int k = 0; k < ts.n; k++) {
    const ToolView &tv = ts.t[k];
    if (!tv.active || tv.algo == MFG_CONTACT_HASHED || !tv.slot) continue;
    if (!in_tool_box(tv, qp)) continue;
    const unsigned pos = atomicAdd(cs.list_n, 1u);
    int sl = -1;
    if (pos < cs.cap) {
      cs.list[pos] = make_int2((int)i, k);
      sl = (int)pos;
    }
    cs.slot[(size_t)k * n_init + i] = sl;  // -1: searched per thread by k_contact_bc
  }
}
__global__ void __launch_bounds__(BST)
    k_contact_list(State st, unsigned n_init, ToolSet ts, const Counters *cnt, ContactSearch cs) {
  contact_list(blockIdx.x, st, n_init, ts, cnt, cs);
}

// one wavefront per listed pair: lanes take the candidate triangles (the cell's list, or all
// triangles outside the grid) in strides of 64; the smallest squared distance wins, the
// smaller triangle index on ties -- the first minimum of the reference's ascending loop
__device__ __forceinline__ void contact_search(unsigned b, unsigned nb, const State &st,
                                               const ToolSet &ts, const ContactSearch &cs) {
  const int lane = threadIdx.x & 63;
  const unsigned nw = nb * (BST / 64);
  const unsigned w0 = b * (BST / 64) + (threadIdx.x >> 6);
  const unsigned n = min(*cs.list_n, cs.cap);
  for (unsigned e = w0; e < n; e += nw) {
    const int2 ent = cs.list[e];
    const int k = __builtin_amdgcn_readfirstlane(ent.y);  // one entry per wavefront
    const unsigned i = (unsigned)__builtin_amdgcn_readfirstlane(ent.x);
    const ToolView &tv = ts.t[k];
    const V3 qp = v3(st.x[i], st.y[i], st.z[i]);
    const double fx = (qp.x - tv.corg[0]) / tv.dx, fy = (qp.y - tv.corg[1]) / tv.dx,
                 fz = (qp.z - tv.corg[2]) / tv.dx;
    const bool in_grid = tv.cand_off && fx >= 0. && fy >= 0. && fz >= 0. && fx < tv.nx &&
                         fy < tv.ny && fz < tv.nz;
    int beg = 0, end = tv.n_tris;
    if (in_grid) {
      const int cell = ((int)fx * tv.ny + (int)fy) * tv.nz + (int)fz;
      beg = tv.cand_off[cell];
      end = tv.cand_off[cell + 1];
    }
    double best = FLT_MAX;
    int best_t = 0x7fffffff;
    V3 cp = v3(0., 0., 0.), nrm = v3(0., 0., 0.);
    for (int q = beg + lane; q < end; q += 64) {
      const int t = in_grid ? tv.cand_idx[q] : q;
      V3 c2, n2;
      closest_point(tv.tris[t], qp, c2, n2);
      const double d2 = dist2(c2, qp);
      if (d2 < best) {  // a lane's candidates ascend, so its first minimum has its smallest index
        best = d2;
        best_t = t;
        cp = c2;
        nrm = n2;
      }
    }
    double gb = best;
    int gt = best_t;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ob = __shfl_xor(gb, o, 64);
      const int ot = __shfl_xor(gt, o, 64);
      if (ob < gb || (ob == gb && ot < gt)) {
        gb = ob;
        gt = ot;
      }
    }
    double *r = cs.res + 6 * (size_t)e;
    if (gt == 0x7fffffff) {  // no distance below FLT_MAX: the per-thread loop leaves zeros
      if (lane == 0)
        for (int k = 0; k < 6; k++) r[k] = 0.;
    } else if (best_t == gt && best == gb) {
      r[0] = cp.x; r[1] = cp.y; r[2] = cp.z;
      r[3] = nrm.x; r[4] = nrm.y; r[5] = nrm.z;
    }
  }
}
__global__ void __launch_bounds__(BST)
    k_contact_search(State st, ToolSet ts, ContactSearch cs) {
  contact_search(blockIdx.x, gridDim.x, st, ts, cs);
}

// normal orientation test of the tool_3d constructor (host unhash: unsigned casts)
__global__ void k_orient(const Tri *boxed, const int *cells, int nx, int ny, int nz, double b0,
                         double b1, double b2, double dx, double q0, double q1, double q2,
                         int *inside) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  unsigned ix = (unsigned)((q0 - b0) / dx), iy = (unsigned)((q1 - b1) / dx),
           iz = (unsigned)((q2 - b2) / dx);
  // the constructed outside point can fall one cell past the grid; clamp to the last cell
  if (ix >= (unsigned)nx) ix = nx - 1;
  if (iy >= (unsigned)ny) iy = ny - 1;
  if (iz >= (unsigned)nz) iz = nz - 1;
  V3 cp, n;
  *inside = contact_by_normal(boxed, cells, nx, ny, nz, (int)ix, (int)iy, (int)iz, v3(q0, q1, q2),
                              cp, n)
                ? 1
                : 0;
}

__global__ void k_flip(Tri *tris, int n) {
  int t = blockIdx.x * BST + threadIdx.x;
  if (t >= n) return;
  Tri &r = tris[t];
  for (int k = 0; k < 3; k++) {
    r.n[k] = -r.n[k];
    for (int q = 0; q < 3; q++) {
      r.np[q][k] = -r.np[q][k];
      r.ne[q][k] = -r.ne[q][k];
    }
  }
}

// triangle k at time t: its vertices at time 0 plus v t (update_tool, src/tool_3d_gpu.cu:74-99)
__device__ __forceinline__ void move_tri(Tri *tris, const double *p_init, int k, double t,
                                         const double v[3]) {
#pragma unroll
  for (int q = 0; q < 3; q++)
#pragma unroll
    for (int a = 0; a < 3; a++) tris[k].p[q][a] = p_init[9 * k + 3 * q + a] + v[a] * t;
}

// two triangle arrays in one launch: threads [0, n) move the first, [n, n + n2) the second
__global__ void k_tool_update(Tri *tris, const double *p_init, int n, Tri *tris2,
                              const double *p_init2, int n2, double t, double v0, double v1,
                              double v2) {
  int k = blockIdx.x * BST + threadIdx.x;
  if (k >= n) {
    k -= n;
    if (k >= n2) return;
    tris = tris2;
    p_init = p_init2;
  }
  const double v[3] = {v0, v1, v2};
  move_tri(tris, p_init, k, t, v);
}

// Contact of the particle at qp (velocity vel, tangential force of before the contact ft) with
// one tool (src/tool_3d_gpu.cu:455-583 exhaustive, :585-681 hashed): false outside the tool's
// moving box, outside the mesh or for an inactive tool; else the normal penalty force fN, the
// Coulomb-limited tangential force fT and the contact normal nrm.
__device__ __forceinline__ bool tool_contact(const ToolView &tv, unsigned i, V3 qp, V3 vel, V3 ft,
                                             double dt, V3 &fN, V3 &fT, V3 &nrm) {
  if (!tv.active) return false;
  if (!in_tool_box(tv, qp)) return false;
  V3 cp = v3(0, 0, 0);
  nrm = v3(0, 0, 0);
  if (tv.algo == MFG_CONTACT_HASHED) {
    int ix = (int)(unsigned)((qp.x - tv.bbmin[0]) / tv.dx);
    int iy = (int)(unsigned)((qp.y - tv.bbmin[1]) / tv.dx);
    int iz = (int)(unsigned)((qp.z - tv.bbmin[2]) / tv.dx);
    // inside the (moving) box the cell index is < n; clamp against rounding at the edge
    ix = min(ix, tv.nx - 1);
    iy = min(iy, tv.ny - 1);
    iz = min(iz, tv.nz - 1);
    if (!contact_by_ray(tv.tris, tv.cells, tv.ny, tv.nz, ix, iy, iz, qp, tv.bbmin, tv.dx))
      return false;
    contact_by_normal(tv.tris, tv.cells, tv.nx, tv.ny, tv.nz, ix, iy, iz, qp, cp, nrm);
  } else {
    // cell of the point in the mesh's own frame; outside the grid, the full loop
    const double fx = (qp.x - tv.corg[0]) / tv.dx, fy = (qp.y - tv.corg[1]) / tv.dx,
                 fz = (qp.z - tv.corg[2]) / tv.dx;
    const bool in_grid = tv.cand_off && fx >= 0. && fy >= 0. && fz >= 0. &&
                         fx < tv.nx && fy < tv.ny && fz < tv.nz;
    const int sl = tv.slot ? tv.slot[i] : -1;
    if (sl >= 0) {
      const double *r = tv.res + 6 * (size_t)sl;
      cp = v3(r[0], r[1], r[2]);
      nrm = v3(r[3], r[4], r[5]);
    } else if (in_grid) {
      const int cell = ((int)fx * tv.ny + (int)fy) * tv.nz + (int)fz;
      closest_by_candidates(tv, qp, cell, cp, nrm);
    } else {
      double min_dist = FLT_MAX;
      for (int q = 0; q < tv.n_tris; q++) {
        V3 cc, n2;
        closest_point(tv.tris[q], qp, cc, n2);
        double d = dist2(cc, qp);
        if (d < min_dist) {
          min_dist = d;
          cp = cc;
          nrm = n2;
        }
      }
    }
    if (vdot(vsub(cp, qp), nrm) > 0.) return false;
  }
  const double dt2 = dt * dt;
  double gN = vdot(vsub(qp, cp), nrm);
  double s = -tv.mass * gN;
  fN = v3(s * nrm.x / dt2 * tv.alpha, s * nrm.y / dt2 * tv.alpha, s * nrm.z / dt2 * tv.alpha);
  fT = v3(0., 0., 0.);
  if (tv.mu != 0.) {
    V3 vv = vsub(vel, v3(tv.vel_const[0], tv.vel_const[1], tv.vel_const[2]));
    V3 vr = v3(vv.x - vv.x * nrm.x, vv.y - vv.y * nrm.y, vv.z - vv.z * nrm.z);
    double am = tv.alpha * tv.mass;
    V3 kd = v3(am * vr.x / dt, am * vr.y / dt, am * vr.z / dt);
    double fy = tv.mu * vlen(fN);
    V3 fstar = vsub(ft, kd);
    double lf = vlen(fstar);
    if (lf > fy) fT = v3(fy * fstar.x / lf, fy * fstar.y / lf, fy * fstar.z / lf);
    else fT = fstar;
  }
  return true;
}

// ---- contact, friction heat, boundary conditions, tool particle motion ------------------------
__device__ __forceinline__ void contact_bc(unsigned b, const State &st, unsigned n_init,
                                           const ToolSet &ts, const Consts &c,
                                           const Counters *cnt, float *partials) {
  __shared__ float red[MAX_TOOLS * 3][BST / 64];
  const unsigned i = b * BST + threadIdx.x;
  const unsigned N = cnt->N;
  float frec[MAX_TOOLS * 3];
  if (ts.record)
    for (int q = 0; q < ts.n * 3; q++) frec[q] = 0.f;
  if (i < n_init) {
    const uint32_t fl = st.flags[i];
    if (i < N) {
      // reset contact forces (src/leap_frog.cu:102-119)
      V3 fc = v3(0., 0., 0.), ft = v3(0., 0., 0.), nn = v3(0., 0., 0.);
      const bool eligible = !(fl & FL_BLANKED) && !(fl & FL_TOOL);
      V3 vel = v3(st.vx[i], st.vy[i], st.vz[i]);
      double T = st.T[i];
      bool T_dirty = false;
      if (eligible) {
        const V3 qp = v3(st.x[i], st.y[i], st.z[i]);
        for (int k = 0; k < ts.n; k++) {
          const ToolView &tv = ts.t[k];
          V3 fN, fT, nrm;
          if (!tool_contact(tv, i, qp, vel, ft, c.dt, fN, fT, nrm)) continue;
          fc = fN;
          ft = fT;
          nn = nrm;
          if (ts.record) {
            float a = (float)(fN.x + fT.x), b = (float)(fN.y + fT.y), d2 = (float)(fN.z + fT.z);
            if (tv.algo != MFG_CONTACT_HASHED) {
              a = isnan(a) ? 0.f : a;
              b = isnan(b) ? 0.f : b;
              d2 = isnan(d2) ? 0.f : d2;
            }
            frec[3 * k] += a;
            frec[3 * k + 1] += b;
            frec[3 * k + 2] += d2;
          }
        }
        // friction heat (src/actions_gpu.cu:192-220)
        if (c.a_fric) {
          double fmag = sqrt_site(ft.x * ft.x + ft.y * ft.y + ft.z * ft.z, c.single_sqrt);
          if (fmag != 0.) {
            V3 vd = v3(vel.x - ts.fric_vel[0], vel.y - ts.fric_vel[1], vel.z - ts.fric_vel[2]);
            double vdn = vd.x * nn.x + vd.y * nn.y + vd.z * nn.z;
            V3 vr = v3(vd.x - vdn, vd.y - vdn, vd.z - vdn);
            double vmag = sqrt_site(vr.x * vr.x + vr.y * vr.y + vr.z * vr.z, c.single_sqrt);
            T += c.twp_eta * c.dt * fmag * vmag / (c.twp_cp * c.a_mass);
            T_dirty = true;
          }
        }
        // velocity boundary condition (src/actions_gpu.cu:464-487)
        if (fl & FL_FIXED) {
          st.vx[i] = st.vbx[i];
          st.vy[i] = st.vby[i];
          st.vz[i] = st.vbz[i];
          fc = v3(0., 0., 0.);
          st.xt[i] = 0.; st.yt[i] = 0.; st.zt[i] = 0.;
          st.vxt[i] = 0.; st.vyt[i] = 0.; st.vzt[i] = 0.;
        }
      }
      // thermal boundary condition (src/actions_gpu.cu:451-462)
      if (c.a_thermal && (fl & FL_FIXED_ONE)) {
        T = c.twp_T_init;
        T_dirty = true;
      }
      if (T_dirty) st.T[i] = T;
      st.fcx[i] = fc.x; st.fcy[i] = fc.y; st.fcz[i] = fc.z;
      st.ftx[i] = ft.x; st.fty[i] = ft.y; st.ftz[i] = ft.z;
    }
    // tool particles move with the first tool (src/actions_gpu.cu:55-68)
    if (fl & FL_TOOL) {
      st.x[i] = st.x[i] + ts.fric_vel[0] * c.dt;
      st.y[i] = st.y[i] + ts.fric_vel[1] * c.dt;
      st.z[i] = st.z[i] + ts.fric_vel[2] * c.dt;
    }
  }
  if (ts.record) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int q = 0; q < ts.n * 3; q++) {
      float v = frec[q];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
      if (lane == 0) red[q][wid] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int q = 0; q < ts.n * 3; q++) {
        float v = 0.f;
        for (int w = 0; w < BST / 64; w++) v += red[q][w];
        partials[(size_t)b * MAX_TOOLS * 3 + q] = v;
      }
  }
}
__global__ void __launch_bounds__(BST)
    k_contact_bc(State st, unsigned n_init, ToolSet ts, Consts c, const Counters *cnt,
                 float *partials) {
  contact_bc(blockIdx.x, st, n_init, ts, c, cnt, partials);
}

__device__ __forceinline__ void force_reduce(const float *partials, int nblocks, int n_tools,
                                             float *forces) {
  int q = threadIdx.x;
  if (q >= n_tools * 3) return;
  float s = 0.f;
  for (int b = 0; b < nblocks; b++) s += partials[(size_t)b * MAX_TOOLS * 3 + q];
  forces[q] += s;
}
__global__ void k_force_reduce(const float *partials, int nblocks, int n_tools, float *forces) {
  force_reduce(partials, nblocks, n_tools, forces);
}

// ---- NaN scan (debug_check_valid_full) ---------------------------------------------------------
__global__ void k_check_valid(State st, unsigned n, int *flag) {
  unsigned i = blockIdx.x * BST + threadIdx.x;
  if (i >= n) return;
  if (st.flags[i] & FL_BLANKED) return;
  auto bad = [](double v) { return isnan(v) || isinf(v); };
  bool b = bad(st.x[i]) || bad(st.y[i]) || bad(st.z[i]) || bad(st.vx[i]) || bad(st.vy[i]) ||
           bad(st.vz[i]) || bad(st.rho[i]) || bad(st.T[i]) || bad(st.eps_pl[i]) ||
           bad(st.eps_pl_dot[i]);
  for (int k = 0; k < 6; k++) b = b || bad(st.S[k][i]);
  if (b) atomicOr(flag, 1);
}

// ---------------------------------------------------------------------------------------------
static inline unsigned nbk(long n) { return (unsigned)((n + BST - 1) / BST); }

int tool_bin_triangles(hipStream_t s, const Tri *tris, int n_tris, const double bbmin[3],
                       double dx, int nx, int ny, int nz, int **cells_out, int **tri_idx_out,
                       int *n_boxed_out) {
  long nc = (long)nx * ny * nz;
  int *count = nullptr, *fill = nullptr, *idx = nullptr;
  if (hipMalloc(&count, sizeof(int) * (nc + 1)) != hipSuccess) return MFG_ERR_NOMEM;
  if (hipMalloc(&fill, sizeof(int) * (nc + 1)) != hipSuccess) {
    (void)hipFree(count);
    return MFG_ERR_NOMEM;
  }
  (void)hipMemsetAsync(count, 0, sizeof(int) * (nc + 1), s);
  (void)hipMemsetAsync(fill, 0, sizeof(int) * (nc + 1), s);
  TriGrid g;
  for (int k = 0; k < 3; k++) g.bbmin[k] = bbmin[k];
  g.dx = dx;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  hipLaunchKernelGGL(k_tri_bin, dim3(nbk(n_tris)), dim3(BST), 0, s, tris, n_tris, g, count, fill,
                     (int *)nullptr);
  hipLaunchKernelGGL(k_tri_scan, dim3(1), dim3(1024), 0, s, count, (int)nc);
  int total = 0;
  if (hipMemcpyAsync(&total, count + nc, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    (void)hipFree(count);
    (void)hipFree(fill);
    return MFG_ERR_HIP;
  }
  if (hipMalloc(&idx, sizeof(int) * (total > 0 ? total : 1)) != hipSuccess) {
    (void)hipFree(count);
    (void)hipFree(fill);
    return MFG_ERR_NOMEM;
  }
  hipLaunchKernelGGL(k_tri_bin, dim3(nbk(n_tris)), dim3(BST), 0, s, tris, n_tris, g, count, fill,
                     idx);
  hipLaunchKernelGGL(k_tri_cell_sort, dim3(nbk(nc)), dim3(BST), 0, s, count, (int)nc, idx);
  (void)hipStreamSynchronize(s);
  (void)hipFree(fill);
  *cells_out = count;
  *tri_idx_out = idx;
  *n_boxed_out = total;
  return MFG_OK;
}

int tool_candidate_lists(hipStream_t s, const Tri *tris, int n_tris, const double bbmin[3],
                         double dx, int nx, int ny, int nz, int **off_out, int **idx_out,
                         long *total_out) {
  const long nc = (long)nx * ny * nz;
  if (nc <= 0 || nc > 0x7fffffffL) return MFG_ERR_ARG;
  TriGrid g;
  for (int k = 0; k < 3; k++) g.bbmin[k] = bbmin[k];
  g.dx = dx;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  const double slack = 1e-3 * dx;
  int *off = nullptr, *idx = nullptr;
  double *cut2 = nullptr;
  if (hipMalloc(&off, sizeof(int) * (nc + 1)) != hipSuccess) return MFG_ERR_NOMEM;
  if (hipMalloc(&cut2, sizeof(double) * nc) != hipSuccess) {
    (void)hipFree(off);
    return MFG_ERR_NOMEM;
  }
  hipLaunchKernelGGL(k_cand_count, dim3((unsigned)nc), dim3(BST), 0, s, tris, n_tris, g, slack,
                     cut2, off);
  hipLaunchKernelGGL(k_tri_scan, dim3(1), dim3(1024), 0, s, off, (int)nc);
  int total = 0;
  if (hipMemcpyAsync(&total, off + nc, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess || total < 0) {
    (void)hipFree(off);
    (void)hipFree(cut2);
    return MFG_ERR_HIP;
  }
  if (hipMalloc(&idx, sizeof(int) * (total > 0 ? total : 1)) != hipSuccess) {
    (void)hipFree(off);
    (void)hipFree(cut2);
    return MFG_ERR_NOMEM;
  }
  hipLaunchKernelGGL(k_cand_fill, dim3((unsigned)nc), dim3(BST), 0, s, tris, n_tris, g, slack,
                     cut2, off, idx);
  const hipError_t e = hipStreamSynchronize(s);
  (void)hipFree(cut2);
  if (e != hipSuccess) {
    (void)hipFree(off);
    (void)hipFree(idx);
    return MFG_ERR_HIP;
  }
  *off_out = off;
  *idx_out = idx;
  *total_out = total;
  return MFG_OK;
}

void launch_gather_tris(hipStream_t s, const Tri *src, const int *idx, int n, Tri *dst) {
  if (n > 0) hipLaunchKernelGGL(k_gather_tris, dim3(nbk(n)), dim3(BST), 0, s, src, idx, n, dst);
}

void launch_orient_test(hipStream_t s, const Tri *boxed, const int *cells, int nx, int ny, int nz,
                        const double bbmin[3], double dx, const double q[3], int *inside) {
  hipLaunchKernelGGL(k_orient, dim3(1), dim3(64), 0, s, boxed, cells, nx, ny, nz, bbmin[0],
                     bbmin[1], bbmin[2], dx, q[0], q[1], q[2], inside);
}

void launch_flip_normals(hipStream_t s, Tri *tris, int n) {
  if (n > 0) hipLaunchKernelGGL(k_flip, dim3(nbk(n)), dim3(BST), 0, s, tris, n);
}

void launch_tool_update(hipStream_t s, Tri *tris, const double *p_init, int n, double t,
                        const double vel[3], Tri *tris2, const double *p_init2, int n2) {
  if (n2 < 0 || !tris2) n2 = 0;
  if (n + n2 > 0)
    hipLaunchKernelGGL(k_tool_update, dim3(nbk(n + n2)), dim3(BST), 0, s, tris, p_init, n, tris2,
                       p_init2, n2, t, vel[0], vel[1], vel[2]);
}

void launch_contact_search(hipStream_t s, const State &st, unsigned n_init, const ToolSet &ts,
                           const Counters *cnt, const ContactSearch &cs) {
  hipLaunchKernelGGL(k_reset_u32, dim3(1), dim3(64), 0, s, cs.list_n);
  hipLaunchKernelGGL(k_contact_list, dim3(nbk(n_init)), dim3(BST), 0, s, st, n_init, ts, cnt, cs);
  hipLaunchKernelGGL(k_contact_search, dim3(1024), dim3(BST), 0, s, st, ts, cs);
}

void launch_contact_bc(hipStream_t s, const State &st, unsigned n_init, const ToolSet &ts,
                       const Consts &c, const Counters *cnt, float *partials, int nblocks) {
  hipLaunchKernelGGL(k_contact_bc, dim3(nblocks), dim3(BST), 0, s, st, n_init, ts, c, cnt,
                     partials);
}

void launch_force_reduce(hipStream_t s, const float *partials, int nblocks, int n_tools,
                         float *forces) {
  hipLaunchKernelGGL(k_force_reduce, dim3(1), dim3(64), 0, s, partials, nblocks, n_tools, forces);
}

// one tool's contact (tool_3d_gpu::compute_contact_force, src/tool_3d_gpu.cu:455-681): fc, ft
// and the normal of every particle in contact (the tangential force of before the call is the
// friction law's previous force), and the force record's per-workgroup partials
__global__ void __launch_bounds__(BST)
    k_act_contact(State st, unsigned n_init, ToolView tv, double dt, const Counters *cnt, Aux a,
                  int record, float *partials) {
  __shared__ float red[3][BST / 64];
  const unsigned i = blockIdx.x * BST + threadIdx.x;
  float fr[3] = {0.f, 0.f, 0.f};
  if (i < n_init && i < cnt->N) {
    const uint32_t fl = st.flags[i];
    if (!(fl & FL_BLANKED) && !(fl & FL_TOOL)) {
      const V3 qp = v3(st.x[i], st.y[i], st.z[i]);
      const V3 vel = v3(st.vx[i], st.vy[i], st.vz[i]);
      const V3 ft = v3(st.ftx[i], st.fty[i], st.ftz[i]);
      V3 fN, fT, nrm;
      if (tool_contact(tv, i, qp, vel, ft, dt, fN, fT, nrm)) {
        st.fcx[i] = fN.x; st.fcy[i] = fN.y; st.fcz[i] = fN.z;
        st.ftx[i] = fT.x; st.fty[i] = fT.y; st.ftz[i] = fT.z;
        a.nrm[i] = nrm.x; a.nrm[a.n + i] = nrm.y; a.nrm[2 * a.n + i] = nrm.z;
        fr[0] = (float)(fN.x + fT.x);
        fr[1] = (float)(fN.y + fT.y);
        fr[2] = (float)(fN.z + fT.z);
        if (tv.algo != MFG_CONTACT_HASHED)
          for (int q = 0; q < 3; q++) fr[q] = isnan(fr[q]) ? 0.f : fr[q];
      }
    }
  }
  if (!record) return;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int q = 0; q < 3; q++) {
    float v = fr[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) red[q][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 0; q < 3; q++) {
      float v = 0.f;
      for (int w = 0; w < BST / 64; w++) v += red[q][w];
      partials[(size_t)blockIdx.x * MAX_TOOLS * 3 + q] = v;
    }
}

void launch_act_contact(hipStream_t s, const State &st, unsigned n_init, const ToolView &tv,
                        double dt, const Counters *cnt, const Aux &a, int record, float *partials,
                        int nblocks) {
  hipLaunchKernelGGL(k_act_contact, dim3(nblocks), dim3(BST), 0, s, st, n_init, tv, dt, cnt, a,
                     record, partials);
}

// clears n floats with a kernel (a memset node in a captured graph is what this avoids)
__global__ void k_zero_f32(float *a, unsigned n) {
  const unsigned i = blockIdx.x * BST + threadIdx.x;
  if (i < n) a[i] = 0.f;
}

void launch_zero_f32(hipStream_t s, float *a, unsigned n) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_zero_f32, dim3(nbk(n)), dim3(BST), 0, s, a, n);
}

// ---------------------------------------------------------------------------------------------
// batched launches: tool motion on the device clock, contact, and the end of the step
// ---------------------------------------------------------------------------------------------
// Workgroup k of an instance moves its tool k to the instance's time t and advances the tool's
// moving box as mfg_tool_update does on the host (a tool without boxed triangles is left as it
// is and inactive); workgroup 0 also clears the force record and the search list counter.
__global__ void __launch_bounds__(BST) kb_tool_advance(Inst *in, BatchMap m, int record) {
  unsigned b, nb;
  Inst &d = in[batch_locate(m, b, nb)];
  const double t = d.t, dt = d.c.dt;
  if (b == 0) {
    if (record && d.forces)
      for (int q = threadIdx.x; q < 3 * d.n_forces; q += BST) d.forces[q] = 0.f;
    if (d.search && threadIdx.x == 0) *d.cs.list_n = 0u;
  }
  if ((int)b >= d.ts.n) return;
  ToolMotion &tm = d.tm[b];
  ToolView &tv = d.ts.t[b];
  if (tm.n_boxed == 0) {
    if (threadIdx.x == 0) tv.active = 0;
    return;
  }
  if (tm.algo != MFG_CONTACT_HASHED)
    for (int k = threadIdx.x; k < tm.n_tris; k += BST) move_tri(tm.tris, tm.p_init, k, t, tm.vel);
  for (int k = threadIdx.x; k < tm.n_boxed; k += BST)
    move_tri(tm.boxed, tm.p_init_boxed, k, t, tm.vel);
  if (threadIdx.x == 0) {
    tm.mesh_tc = t;
    for (int a = 0; a < 3; a++) {
      tv.bbmin[a] += dt * tm.vel[a];
      tv.bbmax[a] += dt * tm.vel[a];
      tv.corg[a] = tm.gbbmin[a] + tm.vel[a] * t;
    }
    tv.active = (t >= tm.birth && t <= tm.death) ? 1 : 0;
  }
}

__global__ void __launch_bounds__(BST) kb_contact_list(const Inst *in, BatchMap m) {
  unsigned b, nb;
  const Inst &d = in[batch_locate(m, b, nb)];
  if (!d.search) return;
  contact_list(b, d.buf[d.cur ^ 1], d.n_init, d.ts, d.cnt, d.cs);
}

__global__ void __launch_bounds__(BST) kb_contact_search(const Inst *in, BatchMap m) {
  unsigned b, nb;
  const Inst &d = in[batch_locate(m, b, nb)];
  if (!d.search) return;
  contact_search(b, nb, d.buf[d.cur ^ 1], d.ts, d.cs);
}

__global__ void __launch_bounds__(BST) kb_contact_bc(const Inst *in, BatchMap m) {
  unsigned b, nb;
  const Inst &d = in[batch_locate(m, b, nb)];
  contact_bc(b, d.buf[d.cur ^ 1], d.n_init, d.ts, d.c, d.cnt, d.partials);
}

// one workgroup per instance: the force record's fixed-order sum, then the clock advances by
// dt (the driver's global_time_current += dt) and the state buffers swap roles
__global__ void __launch_bounds__(64) kb_step_end(Inst *in, BatchMap m, int record) {
  unsigned b, nb;
  Inst &d = in[batch_locate(m, b, nb)];
  if (record && d.forces && d.ts.n > 0) force_reduce(d.partials, d.contact_blocks, d.ts.n, d.forces);
  __syncthreads();
  if (threadIdx.x == 0) {
    d.t = d.t + d.c.dt;
    d.cur ^= 1;
  }
}

void launch_batch_contact(hipStream_t s, const BatchLaunch &L, int record) {
  Inst *in = const_cast<Inst *>(L.inst);
  hipLaunchKernelGGL(kb_tool_advance, dim3(L.tools.blocks), dim3(BST), 0, s, in, L.tools.m, record);
  if (L.search.blocks > 0) {
    hipLaunchKernelGGL(kb_contact_list, dim3(L.parts.blocks), dim3(BST), 0, s, L.inst, L.parts.m);
    hipLaunchKernelGGL(kb_contact_search, dim3(L.search.blocks), dim3(BST), 0, s, L.inst,
                       L.search.m);
  }
  hipLaunchKernelGGL(kb_contact_bc, dim3(L.parts.blocks), dim3(BST), 0, s, L.inst, L.parts.m);
  hipLaunchKernelGGL(kb_step_end, dim3(L.one.blocks), dim3(64), 0, s, in, L.one.m, record);
}

// ---------------------------------------------------------------------------------------------
// Warm start: uniformly weighted SPH interpolation of a scalar carried by samples (x, y, z, v)
// onto points (src/Benchmarks/interp_utils.cpp:65-99): v(p) = sum_j weights W(|p - x_j|, h) v_j
// over all samples in sample order, W the cubic spline with the reference's sqrtf site
// (interp_utils.cpp:19-63).  One thread per point, the samples staged through LDS in tiles; every
// sample's term is added (zero outside the support), as the reference does.
constexpr int INTERP_TILE = 256;

__global__ void __launch_bounds__(INTERP_TILE)
    k_interp_temps(const double4 *__restrict__ in, int n_in, double weights, double h,
                   const double *__restrict__ pts3, int n_out, double *__restrict__ out) {
  __shared__ double4 s_in[INTERP_TILE];
  const int i = blockIdx.x * INTERP_TILE + threadIdx.x;
  const bool act = i < n_out;
  const double px = act ? pts3[3 * i] : 0., py = act ? pts3[3 * i + 1] : 0.,
               pz = act ? pts3[3 * i + 2] : 0.;
  const double h1 = 1 / h;
  const double fourh2 = 4 * h * h;
  const double fac = (INV_PI_D)*h1 * h1 * h1;
  double v = 0.;
  for (int t0 = 0; t0 < n_in; t0 += INTERP_TILE) {
    const int j = t0 + (int)threadIdx.x;
    if (j < n_in) s_in[threadIdx.x] = in[j];
    __syncthreads();
    const int m = min(INTERP_TILE, n_in - t0);
    if (act) {
      for (int k = 0; k < m; k++) {
        const double4 q = s_in[k];
        const double xij = px - q.x, yij = py - q.y, zij = pz - q.z;
        const double rr2 = xij * xij + yij * yij + zij * zij;
        double w = 0., wx, wy, wz;
        if (!(rr2 >= fourh2 || rr2 < 1e-8)) cubic_spline64(xij, yij, zij, rr2, h1, fac, 1, w, wx, wy, wz);
        v += weights * w * q.w;
      }
    }
    __syncthreads();
  }
  if (act) out[i] = v;
}

void launch_interp_temps(hipStream_t s, const double4 *in, int n_in, double weights, double h,
                         const double *pts3, int n_out, double *out) {
  hipLaunchKernelGGL(k_interp_temps, dim3((n_out + INTERP_TILE - 1) / INTERP_TILE),
                     dim3(INTERP_TILE), 0, s, in, n_in, weights, h, pts3, n_out, out);
}

int valid_check(hipStream_t s, const State &st, unsigned n, int *d_flag) {
  (void)hipMemsetAsync(d_flag, 0, sizeof(int), s);
  hipLaunchKernelGGL(k_check_valid, dim3(nbk(n)), dim3(BST), 0, s, st, n, d_flag);
  int h = 0;
  if (hipMemcpyAsync(&h, d_flag, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) return MFG_ERR_HIP;
  if (hipStreamSynchronize(s) != hipSuccess) return MFG_ERR_HIP;
  return h ? MFG_ERR_NONFINITE : MFG_OK;
}

}  // namespace mfg
