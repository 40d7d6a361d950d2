// k_grid.hip -- blanking, cell hashing, stable counting sort, cell tables, permute + predict.
//
// Replaces (one pass each, all on one stream, no host round trip):
//   do_blanking + compute_hashes + adapt_particle_number   src/actions_gpu.cu:489-531,
//                                                          src/grid_gpu_base.cu:41-63,141-164
//   thrust::sort_by_key (stable) + find_cell_start         src/grid_gpu_green.cu:24-70,265-301
//   reorder_data + init + predict + material_eos +
//   do_corrector_artificial_stress                         src/grid_gpu_green.cu:72-125,
//                                                          src/leap_frog.cu:30-70,
//                                                          src/actions_gpu.cu:40-53,70-132
//
// Sort: keys are cell hashes < num_cell or INT_MAX.  A counting sort by cell gives the cell
// offsets directly (exclusive scan of the per-cell counts); particles are placed with an atomic
// slot inside their cell and each cell is then re-ordered by previous index, which makes the
// result identical to a stable sort.  INT_MAX keys go to the tail in previous-index order by a
// block-wise compaction.  Cell start/end of the reference are off[c], off[c+1] for non-empty
// cells.
#include "mfg_device.hpp"
#include "mfg_records.hpp"

namespace mfg {

constexpr int BS = 256;          // threads per block of the particle passes
constexpr int SCAN_ITEMS = 16;   // items per thread of the cell scan
constexpr int SCAN_TILE = BS * SCAN_ITEMS;

__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// exclusive scan across a block of NT threads; lds must hold NT/64 + 1 ints
template <int NT>
__device__ __forceinline__ int block_excl_scan(int v, int *lds, int &total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int incl = wave_incl_scan(v);
  if (lane == 63) lds[wid] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int w = 0; w < NT / 64; w++) {
      int t = lds[w];
      lds[w] = acc;
      acc += t;
    }
    lds[NT / 64] = acc;
  }
  __syncthreads();
  int r = incl - v + lds[wid];
  total = lds[NT / 64];
  __syncthreads();
  return r;
}

__device__ __forceinline__ bool rates_invalid(const State &st, unsigned i) {
  auto bad = [](double v) { return isnan(v) || isinf(v); };
  return bad(st.xt[i]) || bad(st.yt[i]) || bad(st.zt[i]) || bad(st.vxt[i]) || bad(st.vyt[i]) ||
         bad(st.vzt[i]) || bad(st.St[0][i]) || bad(st.St[1][i]) || bad(st.St[2][i]) ||
         bad(st.St[3][i]) || bad(st.St[4][i]) || bad(st.St[5][i]) || bad(st.rho_t[i]) ||
         bad(st.T_t[i]);
}

// Each kernel below is a body run by workgroup `b` (blockIdx.x of a single-instance launch, the
// workgroup inside its instance's range of a batched launch) and a __global__ wrapper.
__device__ __forceinline__ void zero_cells(unsigned b, int *count, int *fill, long n) {
  long c = (long)b * BS + threadIdx.x;
  if (c < n) {
    count[c] = 0;
    fill[c] = 0;
  }
  if (c == n) count[c] = 0;
}
__global__ void __launch_bounds__(BS) k_zero_cells(int *count, int *fill, long n) {
  zero_cells(blockIdx.x, count, fill, n);
}

__device__ __forceinline__ void blank_hash_count(unsigned b, const State &st, unsigned n_init,
                                                 const GridInfo &g, int blanking, double b0,
                                                 double b1, double b2, double b3, double b4,
                                                 double b5, double vmax2, int *key, int *count,
                                                 int *blk, Counters *cnt) {
  __shared__ int lds[BS / 64 + 1];
  const unsigned i = b * BS + threadIdx.x;
  int blank_key = 0;
  if (i < n_init) {
    uint32_t fl = st.flags[i];
    if (blanking) {
      double px = st.x[i], py = st.y[i], pz = st.z[i];
      double vx = st.vx[i], vy = st.vy[i], vz = st.vz[i];
      bool too_fast = vx * vx + vy * vy + vz * vz >= vmax2;
      bool in = px >= b0 && px <= b3 && py >= b1 && py <= b4 && pz >= b2 && pz <= b5;
      uint32_t nf = fl;
      if (too_fast || !in) nf |= FL_BLANKED;
      // inside and slow: finite rates unblank (src/actions_gpu.cu:506-529), non-finite ones keep
      // the flag as it is, so the rates matter only for a blanked particle
      else if ((fl & FL_BLANKED) && !rates_invalid(st, i)) nf &= ~FL_BLANKED;
      if (nf != fl) st.flags[i] = nf;
      fl = nf;
    }
    int k;
    if ((fl & FL_BLANKED) && g.hard_blank) {
      k = KEY_BLANK;
    } else {
      int ix = (int)((st.x[i] - g.bbmin[0]) / g.dx);
      int iy = (int)((st.y[i] - g.bbmin[1]) / g.dx);
      int iz = (int)((st.z[i] - g.bbmin[2]) / g.dx);
      long hh = (long)ix * g.ny * g.nz + (long)iy * g.nz + iz;
      k = (int)hh;
      if (k < 0 || (long)k >= g.num_cell) {
        atomicOr(&cnt->err_grid, 1);
        k = KEY_BLANK;
      }
    }
    key[i] = k;
    if (k != KEY_BLANK) atomicAdd(&count[k], 1);
    blank_key = (k == KEY_BLANK) ? 1 : 0;
  }
  int total;
  block_excl_scan<BS>(blank_key, lds, total);
  if (threadIdx.x == 0) blk[b] = total;
}
__global__ void __launch_bounds__(BS)
    k_blank_hash_count(State st, unsigned n_init, GridInfo g, int blanking, double b0, double b1,
                       double b2, double b3, double b4, double b5, double vmax2, int *key,
                       int *count, int *blk, Counters *cnt) {
  blank_hash_count(blockIdx.x, st, n_init, g, blanking, b0, b1, b2, b3, b4, b5, vmax2, key, count,
                   blk, cnt);
}

// tile reduction of the cell counts
__device__ __forceinline__ void scan_reduce(unsigned b, const int *count, long n, int *part) {
  __shared__ int lds[BS / 64 + 1];
  long base = (long)b * SCAN_TILE + (long)threadIdx.x * SCAN_ITEMS;
  int s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++)
    if (base + k < n) s += count[base + k];
  int total;
  block_excl_scan<BS>(s, lds, total);
  if (threadIdx.x == 0) part[b] = total;
}
__global__ void __launch_bounds__(BS) k_scan_reduce(const int *count, long n, int *part) {
  scan_reduce(blockIdx.x, count, n, part);
}

// single-block exclusive scan of two small arrays (tile sums of the cells, block blank counts)
constexpr int SB = 1024;
__device__ void block_scan_array(int *a, int n, int *lds, int &total) {
  int per = (n + SB - 1) / SB;
  int beg = threadIdx.x * per, end = min(beg + per, n);
  int s = 0;
  for (int k = beg; k < end; k++) s += a[k];
  int off = block_excl_scan<SB>(s, lds, total);
  for (int k = beg; k < end; k++) {
    int t = a[k];
    a[k] = off;
    off += t;
  }
  __syncthreads();
}

__device__ __forceinline__ void scan_top(int *part, int nparts, int *blk, int nblk, int *count,
                                         long num_cell, Counters *cnt) {
  __shared__ int lds[SB / 64 + 1];
  int tot_valid, tot_blank;
  block_scan_array(part, nparts, lds, tot_valid);
  block_scan_array(blk, nblk, lds, tot_blank);
  if (threadIdx.x == 0) {
    count[num_cell] = tot_valid;
    cnt->N = (unsigned)tot_valid;
    cnt->n_blank_key = (unsigned)tot_blank;
    cnt->candidates = 0ull;
    cnt->pairs = 0ull;
  }
}
__global__ void __launch_bounds__(SB) k_scan_top(int *part, int nparts, int *blk, int nblk,
                                                 int *count, long num_cell, Counters *cnt) {
  scan_top(part, nparts, blk, nblk, count, num_cell, cnt);
}

__device__ __forceinline__ void scan_down(unsigned b, int *count, long n, const int *part) {
  __shared__ int lds[BS / 64 + 1];
  long base = (long)b * SCAN_TILE + (long)threadIdx.x * SCAN_ITEMS;
  int v[SCAN_ITEMS];
  int s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    v[k] = (base + k < n) ? count[base + k] : 0;
    s += v[k];
  }
  int total;
  int off = block_excl_scan<BS>(s, lds, total) + part[b];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++)
    if (base + k < n) {
      count[base + k] = off;
      off += v[k];
    }
}
__global__ void __launch_bounds__(BS) k_scan_down(int *count, long n, const int *part) {
  scan_down(blockIdx.x, count, n, part);
}

__device__ __forceinline__ void scatter(unsigned b, unsigned n_init, const int *key,
                                        const int *off, int *fill, const int *blk,
                                        const Counters *cnt, int *order) {
  __shared__ int lds[BS / 64 + 1];
  const unsigned i = b * BS + threadIdx.x;
  int k = (i < n_init) ? key[i] : 0;
  int is_blank = (i < n_init && k == KEY_BLANK) ? 1 : 0;
  int total;
  int rank = block_excl_scan<BS>(is_blank, lds, total);
  if (i >= n_init) return;
  int pos;
  if (is_blank) pos = (int)cnt->N + blk[b] + rank;
  else pos = off[k] + atomicAdd(&fill[k], 1);
  order[pos] = (int)i;
}
__global__ void __launch_bounds__(BS) k_scatter(unsigned n_init, const int *key, const int *off,
                                                int *fill, const int *blk, const Counters *cnt,
                                                int *order) {
  scatter(blockIdx.x, n_init, key, off, fill, blk, cnt, order);
}

// restore previous-index order inside every cell (stable sort)
__device__ __forceinline__ void fixup(unsigned wg, const int *off, long num_cell, int *order) {
  long c = (long)wg * BS + threadIdx.x;
  if (c >= num_cell) return;
  int s = off[c], e = off[c + 1];
  if (e - s < 2) return;
  for (int a = s + 1; a < e; a++) {
    int v = order[a];
    int b = a - 1;
    while (b >= s && order[b] > v) {
      order[b + 1] = order[b];
      b--;
    }
    order[b + 1] = v;
  }
}
__global__ void __launch_bounds__(BS) k_fixup(const int *off, long num_cell, int *order) {
  fixup(blockIdx.x, off, num_cell, order);
}

// gather the persistent state by the sort permutation, then predict the half step and build the
// pair records (EOS and artificial stress included)
template <class RecT>
__device__ __forceinline__ void permute_predict(unsigned b, const State &src, const State &dst,
                                                unsigned n_init, const int *order,
                                                const int *key, int *key_sorted,
                                                const Counters *cnt, const Consts &c, int predict,
                                                RecT *rec, double4 *nl4, double2 *pse2,
                                                double4 *cs4, double *p_dbg, double *R_dbg) {
  const unsigned i = b * BS + threadIdx.x;
  if (i >= n_init) return;
  const unsigned N = cnt->N;
  // a null order: the particles keep their places (a step between sorts)
  const int s = order ? order[i] : (int)i;
  if (order) key_sorted[i] = key[s];
  double x[3] = {src.x[s], src.y[s], src.z[s]};
  double v[3] = {src.vx[s], src.vy[s], src.vz[s]};
  double S[6], St[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    S[k] = src.S[k][s];
    St[k] = src.St[k][s];
  }
  double xt[3] = {src.xt[s], src.yt[s], src.zt[s]};
  double vt[3] = {src.vxt[s], src.vyt[s], src.vzt[s]};
  double rho = src.rho[s], T = src.T[s], rho_t = src.rho_t[s], T_t = src.T_t[s], h = src.h[s];
  uint32_t fl = src.flags[s];
  dst.x[i] = x[0]; dst.y[i] = x[1]; dst.z[i] = x[2];
  dst.vx[i] = v[0]; dst.vy[i] = v[1]; dst.vz[i] = v[2];
  dst.vbx[i] = src.vbx[s]; dst.vby[i] = src.vby[s]; dst.vbz[i] = src.vbz[s];
  dst.fcx[i] = src.fcx[s]; dst.fcy[i] = src.fcy[s]; dst.fcz[i] = src.fcz[s];
  dst.ftx[i] = src.ftx[s]; dst.fty[i] = src.fty[s]; dst.ftz[i] = src.ftz[s];
  dst.h[i] = h; dst.rho[i] = rho;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    dst.S[k][i] = S[k];
    dst.St[k][i] = St[k];
  }
  dst.eps_pl[i] = src.eps_pl[s]; dst.eps_pl_dot[i] = src.eps_pl_dot[s]; dst.T[i] = T;
  dst.xt[i] = xt[0]; dst.yt[i] = xt[1]; dst.zt[i] = xt[2];
  dst.vxt[i] = vt[0]; dst.vyt[i] = vt[1]; dst.vzt[i] = vt[2];
  dst.rho_t[i] = rho_t; dst.T_t[i] = T_t;
  dst.flags[i] = fl;
  dst.id[i] = src.id[s];
  if (!predict) return;

  // predict (src/leap_frog.cu:42-70): half step from the rates of the previous step
  const bool blanked = (fl & FL_BLANKED) != 0;
  const bool tool = (fl & FL_TOOL) != 0;
  double xh[3], vh[3], Sh[6], rhoh = rho, Th = T;
  if (!blanked) {
    const double hdt = 0.5 * c.dt;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      xh[k] = x[k] + hdt * xt[k];
      vh[k] = v[k] + hdt * vt[k];
    }
#pragma unroll
    for (int k = 0; k < 6; k++) Sh[k] = S[k] + hdt * St[k];
    rhoh = rho + hdt * rho_t;
    Th = T + hdt * T_t;
  } else {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      xh[k] = x[k];
      vh[k] = v[k];
    }
#pragma unroll
    for (int k = 0; k < 6; k++) Sh[k] = S[k];
  }
  // equation of state and artificial stress (src/actions_gpu.cu:40-53,70-132)
  double p = 0., R[6] = {0., 0., 0., 0., 0., 0.};
  if (i < N && !blanked && !tool) {
    double c0 = sqrt_site(c.a_K / c.a_rho0, c.single_sqrt);
    p = c0 * c0 * (rhoh - c.a_rho0);
    if (c.a_stresseps != 0.) art_stress(c.a_stresseps, rhoh, p, Sh, R);
  }
  emit_record<RecT>(i, xh, h, vh, rhoh, Sh, p, R, Th, tool, blanked, c, rec, nl4, pse2, cs4);
  if (p_dbg) {
    p_dbg[i] = p;
#pragma unroll
    for (int k = 0; k < 6; k++) R_dbg[6 * (size_t)i + k] = R[k];
  }
}
template <class RecT>
__global__ void __launch_bounds__(BS)
    k_permute_predict(State src, State dst, unsigned n_init, const int *order, const int *key,
                      int *key_sorted, const Counters *cnt, Consts c, int predict, RecT *rec,
                      double4 *nl4, double2 *pse2, double4 *cs4, double *p_dbg, double *R_dbg) {
  permute_predict<RecT>(blockIdx.x, src, dst, n_init, order, key, key_sorted, cnt, c, predict, rec,
                        nl4, pse2, cs4, p_dbg, R_dbg);
}

// Green (0xffffffff for empty cells) or Rothlin (start == end) tables from the offsets
__global__ void __launch_bounds__(BS) k_cells_to_tables(const int *off, long n, int rothlin,
                                                        int *start, int *end) {
  long c = (long)blockIdx.x * BS + threadIdx.x;
  if (c >= n) return;
  int s = off[c], e = off[c + 1];
  if (!rothlin && s == e) {
    start[c] = -1;
    end[c] = -1;
  } else {
    start[c] = s;
    end[c] = e;
  }
}

// bounding box and largest h over the first N particles (unlocked grids)
__global__ void __launch_bounds__(BS) k_bbox(State st, const Counters *cnt, double *part) {
  __shared__ double red[7][BS];
  const unsigned N = cnt->N;
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300}, hm = -1e300;
  for (unsigned i = blockIdx.x * BS + threadIdx.x; i < N; i += gridDim.x * BS) {
    double p[3] = {st.x[i], st.y[i], st.z[i]};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      mn[k] = fmin(mn[k], p[k]);
      mx[k] = fmax(mx[k], p[k]);
    }
    hm = fmax(hm, st.h[i]);
  }
  for (int k = 0; k < 3; k++) {
    red[k][threadIdx.x] = mn[k];
    red[3 + k][threadIdx.x] = mx[k];
  }
  red[6][threadIdx.x] = hm;
  __syncthreads();
  for (int o = BS / 2; o > 0; o >>= 1) {
    if (threadIdx.x < (unsigned)o) {
      for (int k = 0; k < 3; k++) red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + o]);
      for (int k = 3; k < 7; k++) red[k][threadIdx.x] = fmax(red[k][threadIdx.x], red[k][threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 7; k++) part[7 * blockIdx.x + k] = red[k][0];
}

// ---------------------------------------------------------------------------------------------
static inline unsigned nblocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

void launch_blank_hash_count(hipStream_t s, const State &st, unsigned n_init, const GridInfo &g,
                             int blanking, const double *bb, double vmax2, GridBuffers &gb,
                             Counters *cnt) {
  hipLaunchKernelGGL(k_zero_cells, dim3(nblocks(g.num_cell + 1, BS)), dim3(BS), 0, s, gb.count,
                     gb.fill, g.num_cell);
  double z[6] = {0, 0, 0, 0, 0, 0};
  const double *b = bb ? bb : z;
  hipLaunchKernelGGL(k_blank_hash_count, dim3(gb.nblk), dim3(BS), 0, s, st, n_init, g, blanking,
                     b[0], b[1], b[2], b[3], b[4], b[5], vmax2, gb.key, gb.count, gb.blk, cnt);
}

void launch_scan_cells(hipStream_t s, const GridInfo &g, GridBuffers &gb, unsigned n_init,
                       Counters *cnt) {
  (void)n_init;
  int nparts = (int)nblocks(g.num_cell, SCAN_TILE);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nparts), dim3(BS), 0, s, gb.count, g.num_cell, gb.blk2);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SB), 0, s, gb.blk2, nparts, gb.blk, gb.nblk,
                     gb.count, g.num_cell, cnt);
  hipLaunchKernelGGL(k_scan_down, dim3(nparts), dim3(BS), 0, s, gb.count, g.num_cell, gb.blk2);
}

void launch_scatter(hipStream_t s, unsigned n_init, const GridInfo &g, GridBuffers &gb,
                    Counters *cnt) {
  (void)g;
  hipLaunchKernelGGL(k_scatter, dim3(gb.nblk), dim3(BS), 0, s, n_init, gb.key, gb.count, gb.fill,
                     gb.blk, cnt, gb.order);
}

void launch_fixup(hipStream_t s, const GridInfo &g, GridBuffers &gb) {
  hipLaunchKernelGGL(k_fixup, dim3(nblocks(g.num_cell, BS)), dim3(BS), 0, s, gb.count, g.num_cell,
                     gb.order);
}

void launch_permute_predict_ex(hipStream_t s, const State &src, const State &dst, unsigned n_init,
                               const GridBuffers &gb, const Counters *cnt, const Consts &c,
                               int predict, void *rec, double4 *nl4, double2 *pse2, double4 *cs4,
                               double *p_dbg, double *R_dbg) {
  if (c.precision == MFG_FP64)
    hipLaunchKernelGGL(k_permute_predict<Rec64>, dim3(nblocks(n_init, BS)), dim3(BS), 0, s, src,
                       dst, n_init, gb.order, gb.key, gb.key_sorted, cnt, c, predict, (Rec64 *)rec,
                       nl4, pse2, cs4, p_dbg, R_dbg);
  else
    hipLaunchKernelGGL(k_permute_predict<Rec32>, dim3(nblocks(n_init, BS)), dim3(BS), 0, s, src,
                       dst, n_init, gb.order, gb.key, gb.key_sorted, cnt, c, predict, (Rec32 *)rec,
                       nl4, pse2, cs4, p_dbg, R_dbg);
}

void launch_permute_predict(hipStream_t s, const State &src, const State &dst, unsigned n_init,
                            const GridBuffers &gb, const Counters *cnt, const Consts &c,
                            int predict, void *rec, double4 *nl4, double2 *pse2, double4 *cs4) {
  launch_permute_predict_ex(s, src, dst, n_init, gb, cnt, c, predict, rec, nl4, pse2, cs4, nullptr,
                            nullptr);
}

void launch_cells_to_tables(hipStream_t s, const GridInfo &g, const int *offsets, int rothlin,
                            int *start, int *end) {
  hipLaunchKernelGGL(k_cells_to_tables, dim3(nblocks(g.num_cell, BS)), dim3(BS), 0, s, offsets,
                     g.num_cell, rothlin, start, end);
}

void launch_bbox_reduce(hipStream_t s, const State &st, const Counters *cnt, unsigned n_init,
                        double *partials, int nb) {
  (void)n_init;
  hipLaunchKernelGGL(k_bbox, dim3(nb), dim3(BS), 0, s, st, cnt, partials);
}

// ---------------------------------------------------------------------------------------------
// batched launches: the same bodies, arguments from each instance's descriptor.  Before the
// permutation an instance's state is buf[cur], after it buf[cur ^ 1].
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(BS) kb_zero_cells(const Inst *in, BatchMap m) {
  unsigned b, nb;
  const Inst &d = in[batch_locate(m, b, nb)];
  zero_cells(b, d.gb.count, d.gb.fill, d.g.num_cell);
}

__global__ void __launch_bounds__(BS) kb_blank_hash_count(const Inst *in, BatchMap m) {
  unsigned b, nb;
  const Inst &d = in[batch_locate(m, b, nb)];
  blank_hash_count(b, d.buf[d.cur], d.n_init, d.g, d.blanking, d.bb[0], d.bb[1], d.bb[2], d.bb[3],
                   d.bb[4], d.bb[5], d.vmax2, d.gb.key, d.gb.count, d.gb.blk, d.cnt);
}

__global__ void __launch_bounds__(BS) kb_scan_reduce(const Inst *in, BatchMap m) {
  unsigned b, nb;
  const Inst &d = in[batch_locate(m, b, nb)];
  scan_reduce(b, d.gb.count, d.g.num_cell, d.gb.blk2);
}

__global__ void __launch_bounds__(SB) kb_scan_top(const Inst *in, BatchMap m) {
  unsigned b, nb;
  const Inst &d = in[batch_locate(m, b, nb)];
  scan_top(d.gb.blk2, d.nparts, d.gb.blk, d.gb.nblk, d.gb.count, d.g.num_cell, d.cnt);
}

__global__ void __launch_bounds__(BS) kb_scan_down(const Inst *in, BatchMap m) {
  unsigned b, nb;
  const Inst &d = in[batch_locate(m, b, nb)];
  scan_down(b, d.gb.count, d.g.num_cell, d.gb.blk2);
}

__global__ void __launch_bounds__(BS) kb_scatter(const Inst *in, BatchMap m) {
  unsigned b, nb;
  const Inst &d = in[batch_locate(m, b, nb)];
  scatter(b, d.n_init, d.gb.key, d.gb.count, d.gb.fill, d.gb.blk, d.cnt, d.gb.order);
}

__global__ void __launch_bounds__(BS) kb_fixup(const Inst *in, BatchMap m) {
  unsigned b, nb;
  const Inst &d = in[batch_locate(m, b, nb)];
  fixup(b, d.gb.count, d.g.num_cell, d.gb.order);
}

template <class RecT>
__global__ void __launch_bounds__(BS) kb_permute_predict(const Inst *in, BatchMap m) {
  unsigned b, nb;
  const Inst &d = in[batch_locate(m, b, nb)];
  permute_predict<RecT>(b, d.buf[d.cur], d.buf[d.cur ^ 1], d.n_init, d.gb.order, d.gb.key,
                        d.gb.key_sorted, d.cnt, d.c, 1, (RecT *)d.rec, d.nl4, d.pse2, d.cs4,
                        nullptr, nullptr);
}

// hash + blanking, scan, scatter and fix-up of every instance (the sort permutation and cell
// offsets), as sort_keys does for one
void launch_batch_sort(hipStream_t s, const BatchLaunch &L) {
  hipLaunchKernelGGL(kb_zero_cells, dim3(L.cells1.blocks), dim3(BS), 0, s, L.inst, L.cells1.m);
  hipLaunchKernelGGL(kb_blank_hash_count, dim3(L.parts.blocks), dim3(BS), 0, s, L.inst, L.parts.m);
  hipLaunchKernelGGL(kb_scan_reduce, dim3(L.tiles.blocks), dim3(BS), 0, s, L.inst, L.tiles.m);
  hipLaunchKernelGGL(kb_scan_top, dim3(L.one.blocks), dim3(SB), 0, s, L.inst, L.one.m);
  hipLaunchKernelGGL(kb_scan_down, dim3(L.tiles.blocks), dim3(BS), 0, s, L.inst, L.tiles.m);
  hipLaunchKernelGGL(kb_scatter, dim3(L.parts.blocks), dim3(BS), 0, s, L.inst, L.parts.m);
  hipLaunchKernelGGL(kb_fixup, dim3(L.cells.blocks), dim3(BS), 0, s, L.inst, L.cells.m);
}

void launch_batch_permute(hipStream_t s, const BatchLaunch &L) {
  if (L.precision == MFG_FP64)
    hipLaunchKernelGGL(kb_permute_predict<Rec64>, dim3(L.parts.blocks), dim3(BS), 0, s, L.inst,
                       L.parts.m);
  else
    hipLaunchKernelGGL(kb_permute_predict<Rec32>, dim3(L.parts.blocks), dim3(BS), 0, s, L.inst,
                       L.parts.m);
}

}  // namespace mfg
