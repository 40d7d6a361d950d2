// abi.hip -- host side of the C ABI (include/mfree_gpu.h): objects, memory, step orchestration.
//
// leap_frog::step (src/leap_frog.cu:121-195) becomes one stream of kernels with no host round
// trip for a locked grid:
//   zero cells | blank + hash + count | scan (N on the device) | scatter | cell fix-up |
//   permute + predict + EOS + artificial stress | neighbour lists (+ PSE) |
//   pair sums + rates + correct + Johnson-Cook | tool motion | contact + BCs | force reduction
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cstdarg>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "mfg_internal.hpp"

using namespace mfg;

namespace mfg {
void launch_permute_predict_ex(hipStream_t s, const State &src, const State &dst, unsigned n_init,
                               const GridBuffers &gb, const Counters *cnt, const Consts &c,
                               int predict, void *rec, double4 *nl4, double2 *pse2, double4 *cs4,
                               double *p_dbg, double *R_dbg);
}

// ---------------------------------------------------------------------------------------------
// library state
// ---------------------------------------------------------------------------------------------
namespace {

// pool streams created at mfg_init (the library stream included)
constexpr int MFG_EARLY_STREAMS = 4;

struct Lib {
  bool ready = false;
  int device = -1;
  hipStream_t stream = nullptr;  // the stream calls queue on: streams[k] of mfg_select_stream
  hipStream_t streams[MFG_MAX_STREAMS] = {};
  int n_streams = 0;  // streams of the pool (mfg_stream_pool)
  int n_created = 0;  // streams created, >= n_streams
  mfg_options opt{0, MFG_HEAT_BROOKSHAW, 1, 256, MFG_PAIR_LIST, 0, 0, 0, 0};
  // exhaustive contact: closest triangle found by one wavefront per (particle, tool) pair;
  // MFG_LANE_CONTACT=0 in the environment selects the one-thread-per-particle search instead
  bool opt_contact_search = true;
  // constants as the two reference translation units hold them
  mfg_phys_constants a_phys{}, i_phys{};
  mfg_corr_constants a_corr{}, i_corr{};
  mfg_joco_constants joco{};
  mfg_trml_constants a_twp{}, a_ttool{}, i_twp{}, i_ttool{};
  bool plastic = false, a_thermal = false, a_fric = false, i_thermal_wp = false;
  // profiling: 0 off, MFG_PROFILE_ALL every kernel, MFG_PROFILE_PASSES the neighbour and pair
  // passes only
  int prof = 0;
  // whether the neighbour passes count their candidate and interacting pairs
  bool count_pairs = false;
  std::vector<hipEvent_t> pool;
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<std::string> names;
  std::map<std::string, std::pair<long, double>> acc;
};

Lib g_lib;
thread_local std::string g_err;

// Work memory of the allocation-free step (mfg_leap_frog_step_noalloc): one array per GPU, held
// until the process ends.  Host code never passes `scratch` itself to a kernel or copy (in host
// code the name is a stand-in, not a device address); it uses the device address that mfg_init
// resolves with the library's device current (g_scratch).
static __device__ __align__(256) unsigned char scratch[MFG_STEP_SCRATCH_BYTES];
unsigned char *g_scratch = nullptr;

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(MFG_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                  __LINE__);                                                             \
  } while (0)

#define NEED_LIB()                                                         \
  do {                                                                     \
    if (!g_lib.ready) return fail(MFG_ERR_STATE, "mfg_init was not called"); \
  } while (0)

hipEvent_t prof_event() {
  if (!g_lib.pool.empty()) {
    hipEvent_t e = g_lib.pool.back();
    g_lib.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

struct ProfScope {
  hipEvent_t b = nullptr;
  std::string name;
  ProfScope(const char *n) {
    if (!g_lib.prof) return;
    if (g_lib.prof == MFG_PROFILE_PASSES && std::strcmp(n, "neighbours") != 0 &&
        std::strncmp(n, "pair_fp", 7) != 0)
      return;
    name = n;
    hipEvent_t a = prof_event();
    b = prof_event();
    (void)hipEventRecord(a, g_lib.stream);
    g_lib.pending.push_back({name, a, b});
  }
  ~ProfScope() {
    if (b) (void)hipEventRecord(b, g_lib.stream);
  }
};

void prof_collect() {
  if (g_lib.pending.empty()) return;
  (void)hipStreamSynchronize(g_lib.stream);
  for (auto &p : g_lib.pending) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, p.a, p.b);
    auto it = g_lib.acc.find(p.name);
    if (it == g_lib.acc.end()) {
      g_lib.names.push_back(p.name);
      g_lib.acc[p.name] = {1, (double)ms};
    } else {
      it->second.first += 1;
      it->second.second += ms;
    }
    g_lib.pool.push_back(p.a);
    g_lib.pool.push_back(p.b);
  }
  g_lib.pending.clear();
}

Consts make_consts(double dt, int precision) {
  Consts c;
  std::memset(&c, 0, sizeof(c));
  const Lib &L = g_lib;
  c.a_rho0 = L.a_phys.rho0; c.a_K = L.a_phys.K; c.a_G = L.a_phys.G; c.a_mass = L.a_phys.mass;
  c.a_stresseps = L.a_corr.stresseps;
  c.jc_A = L.joco.A; c.jc_B = L.joco.B; c.jc_C = L.joco.C; c.jc_n = L.joco.n; c.jc_m = L.joco.m;
  c.jc_Tmelt = L.joco.Tmelt; c.jc_Tref = L.joco.Tref; c.jc_eps_dot_ref = L.joco.eps_dot_ref;
  c.jc_clamp = L.joco.clamp_temp;
  c.twp_cp = L.a_twp.cp; c.twp_tq = L.a_twp.tq; c.twp_T_init = L.a_twp.T_init; c.twp_eta = L.a_twp.eta;
  c.plastic = L.plastic; c.a_thermal = L.a_thermal; c.a_fric = L.a_fric;
  c.i_mass = L.i_phys.mass; c.i_K = L.i_phys.K;
  c.wdeltap = L.i_corr.wdeltap; c.v_alpha = L.i_corr.alpha; c.v_beta = L.i_corr.beta;
  c.v_eta = L.i_corr.eta; c.xspheps = L.i_corr.xspheps;
  c.alpha_wp = L.i_twp.alpha; c.alpha_tool = L.i_ttool.alpha;
  c.pse_on = L.i_thermal_wp;
  c.cspm = L.opt.cspm; c.heat = L.opt.heat; c.single_sqrt = L.opt.single_sqrt;
  c.precision = precision;
  c.dt = dt;
  return c;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// objects
// ---------------------------------------------------------------------------------------------
struct mfg_particles {
  unsigned N_init = 0;
  unsigned N_host = 0;
  int precision = MFG_FP64;
  double *arena[2] = {nullptr, nullptr};
  uint32_t *uarena[2] = {nullptr, nullptr};
  State buf[2];
  int cur = 0;
  void *rec = nullptr;
  double4 *nl4 = nullptr;
  double2 *pse2 = nullptr;
  double4 *cs4 = nullptr;  // what the CSPM sum reads of each record, 32 bytes
  double *sums = nullptr;  // neighbour sums of the pair pass, PAIR_SUM_COUNT x sum_stride (SoA)
  size_t sum_stride = 0;
  int *trav = nullptr;     // visiting order of the list pair pass (N_init)
  int *poison = nullptr;   // per particle: a non-finite record in its stencil (N_init)
  int *seg_pos = nullptr;  // its segment offsets (lazily sized)
  int seg_cap = 0;
  int *nbr = nullptr;
  int *count = nullptr;
  int nl_stride = 0;
  int cap = 0;
  int *key = nullptr, *key_sorted = nullptr, *order = nullptr, *blk = nullptr;
  int nblk = 0;
  Counters *cnt = nullptr;     // device
  Counters *cnt_h = nullptr;   // pinned host mirror
  float *partials = nullptr;   // contact-force partial sums
  int contact_blocks = 0;
  double *p_dbg = nullptr, *R_dbg = nullptr;  // pressure / artificial stress (per-action p and R)
  // per-action entry points: velocity gradient, stress divergence terms, contact normal and the
  // leap-frog start-of-step copies (lazily allocated)
  double *vder = nullptr, *sder = nullptr, *nrm = nullptr, *x0 = nullptr;
  int *flag_d = nullptr;
  unsigned long long *stat_part = nullptr;  // per-workgroup (candidates, pairs) of the last list build
  int stat_cap = 0, stat_n = 0;
  // lane-parallel exhaustive contact search (lazily sized)
  int2 *cs_list = nullptr;
  unsigned *cs_n = nullptr;
  int *cs_slot = nullptr;
  double *cs_res = nullptr;
  unsigned cs_cap = 0;
  int cs_tools = 0;
  // the last sort (re-sort interval): the grid and its geometry then, and the steps since
  const mfg_grid *sorted_grid = nullptr;
  double sorted_geom[7] = {0, 0, 0, 0, 0, 0, 0};
  int since_sort = 0;
  const State &st() const { return buf[cur]; }
};

struct mfg_grid {
  int kind = MFG_GRID_GREEN;
  bool locked = false, hard_blank = true, min_fixed = false, max_fixed = false;
  double dx = 0.;
  double bbmin[3] = {0, 0, 0}, bbmax[3] = {0, 0, 0};
  double bbmin_init[3] = {0, 0, 0}, bbmax_init[3] = {0, 0, 0};
  double vel_bbox[3] = {0, 0, 0};
  long nx = 0, ny = 0, nz = 0, num_cell = 0, max_cell = 0;
  int *count = nullptr;  // max_cell + 1: counts, then offsets
  int *fill = nullptr;
  int *parts = nullptr;
  int nparts_cap = 0;
  double *bbox_part = nullptr;
  std::vector<double> bbox_host;  // host copy of the bbox partials (sized at creation)
  // keys, offsets and sort permutation made by a hash (mfg_grid_assign_hashes) for sorted_for
  // and not yet applied; a step or a sort applies them
  bool scanned = false;
  const mfg_particles *sorted_for = nullptr;  // the particle set the last sort was made for
  GridInfo info() const {
    GridInfo g;
    g.nx = (int)nx; g.ny = (int)ny; g.nz = (int)nz;
    g.hard_blank = hard_blank ? 1 : 0;
    for (int k = 0; k < 3; k++) g.bbmin[k] = bbmin[k];
    g.dx = dx;
    g.num_cell = num_cell;
    return g;
  }
};

struct mfg_tool {
  int n_tris = 0, n_boxed = 0;
  Tri *tris = nullptr;       // all triangles (original order)
  Tri *boxed = nullptr;      // per-cell copies
  double *p_init_all = nullptr, *p_init_boxed = nullptr;
  int *cells = nullptr;      // CSR offsets, num_cell + 1
  int *boxed_idx = nullptr;  // original triangle index of every boxed copy
  int *cand_off = nullptr;   // closest-triangle candidate lists (exhaustive algorithm only)
  int *cand_idx = nullptr;
  long n_cand = 0;
  double mesh_tc = 0.;       // time of the last translation of the triangles
  int gnx = 0, gny = 0, gnz = 0;
  double gbbmin[3], gbbmax[3], gdx = 0.;
  double bbmin[3], bbmax[3];  // moving copy (geometry constant of the reference)
  int algo = MFG_CONTACT_HASHED;
  double vel[3] = {0, 0, 0};
  double vel_const[3] = {0, 0, 0};
  double mu = 0., alpha = 0.1, mass = 0.;
  double birth = 0., death = FLT_MAX;
  bool thermal = false;
};

struct mfg_tool_forces {
  unsigned n = 0;
  float *d = nullptr;
};

struct mfg_leap_frog {
  unsigned num_part = 0, num_cell = 0;
};

// ---------------------------------------------------------------------------------------------
// allocation helpers
// ---------------------------------------------------------------------------------------------
namespace {

void bind_state(State &s, double *a, uint32_t *u, size_t n) {
  double **slots[] = {&s.x, &s.y, &s.z, &s.vx, &s.vy, &s.vz, &s.vbx, &s.vby, &s.vbz,
                      &s.fcx, &s.fcy, &s.fcz, &s.ftx, &s.fty, &s.ftz, &s.h, &s.rho,
                      &s.S[0], &s.S[1], &s.S[2], &s.S[3], &s.S[4], &s.S[5],
                      &s.eps_pl, &s.eps_pl_dot, &s.T, &s.xt, &s.yt, &s.zt,
                      &s.vxt, &s.vyt, &s.vzt, &s.rho_t, &s.T_t,
                      &s.St[0], &s.St[1], &s.St[2], &s.St[3], &s.St[4], &s.St[5]};
  static_assert(sizeof(slots) / sizeof(slots[0]) == STATE_DOUBLES, "state arrays");
  for (int k = 0; k < STATE_DOUBLES; k++) *slots[k] = a + (size_t)k * n;
  s.flags = u;
  s.id = u + n;
}

size_t padded(size_t n) { return (n + 63) / 64 * 64; }

void free_particles(mfg_particles *p) {
  void *ptrs[] = {p->arena[0], p->arena[1], p->uarena[0], p->uarena[1], p->rec, p->nl4, p->pse2,
                  p->cs4, p->sums, p->trav, p->seg_pos, p->poison, p->nbr, p->count, p->key, p->key_sorted, p->order, p->blk, p->cnt, p->partials,
                  p->p_dbg, p->R_dbg, p->vder, p->sder, p->nrm, p->x0, p->flag_d, p->stat_part,
                  p->cs_list, p->cs_n,
                  p->cs_slot, p->cs_res};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  if (p->cnt_h) (void)hipHostFree(p->cnt_h);
}

int grid_alloc(mfg_grid *g) {
  HIP_TRY(hipMalloc(&g->count, sizeof(int) * (g->max_cell + 1)));
  HIP_TRY(hipMalloc(&g->fill, sizeof(int) * (g->max_cell + 1)));
  g->nparts_cap = (int)((g->max_cell + 4095) / 4096) + 1;
  HIP_TRY(hipMalloc(&g->parts, sizeof(int) * g->nparts_cap));
  HIP_TRY(hipMalloc(&g->bbox_part, sizeof(double) * 7 * 1024));
  g->bbox_host.assign(7 * 1024, 0.);
  return MFG_OK;
}

GridBuffers grid_buffers(mfg_grid *g, mfg_particles *p) {
  GridBuffers gb;
  gb.key = p->key;
  gb.key_sorted = p->key_sorted;
  gb.order = p->order;
  gb.count = g->count;
  gb.fill = g->fill;
  gb.blk = p->blk;
  gb.blk2 = g->parts;
  gb.nblk = p->nblk;
  return gb;
}

// host reduce of the unlocked-grid bounding box (src/grid_gpu_base.cu:185-233)
int grid_update_unlocked(mfg_grid *g, mfg_particles *p, double kernel_width) {
  const int nb = 1024;
  launch_bbox_reduce(g_lib.stream, p->st(), p->cnt, p->N_init, g->bbox_part, nb);
  std::vector<double> &part = g->bbox_host;
  HIP_TRY(hipMemcpyAsync(part.data(), g->bbox_part, sizeof(double) * 7 * nb, hipMemcpyDeviceToHost,
                         g_lib.stream));
  HIP_TRY(hipStreamSynchronize(g_lib.stream));
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300}, hm = -1e300;
  for (int b = 0; b < nb; b++) {
    for (int k = 0; k < 3; k++) {
      mn[k] = std::fmin(mn[k], part[7 * b + k]);
      mx[k] = std::fmax(mx[k], part[7 * b + 3 + k]);
    }
    hm = std::fmax(hm, part[7 * b + 6]);
  }
  for (int k = 0; k < 3; k++) {
    g->bbmin[k] = mn[k] - 1e-6;
    g->bbmax[k] = mx[k] + 1e-6;
  }
  g->dx = hm * kernel_width / 2;
  g->nx = (long)std::ceil((g->bbmax[0] - g->bbmin[0]) / g->dx);
  g->ny = (long)std::ceil((g->bbmax[1] - g->bbmin[1]) / g->dx);
  g->nz = (long)std::ceil((g->bbmax[2] - g->bbmin[2]) / g->dx);
  g->num_cell = g->nx * g->ny * g->nz;
  if (g->num_cell >= g->max_cell)
    return fail(MFG_ERR_GRID, "num cell %ld exceeds max cell %ld", g->num_cell, g->max_cell);
  return MFG_OK;
}

int grid_update_geometry(mfg_grid *g, mfg_particles *p, double kernel_width, double t) {
  if (g->locked) {
    if (std::fabs(g->vel_bbox[0]) > 0 || std::fabs(g->vel_bbox[1]) > 0 || std::fabs(g->vel_bbox[2]) > 0) {
      for (int k = 0; k < 3; k++) {
        if (!g->min_fixed) g->bbmin[k] = g->bbmin_init[k] + t * g->vel_bbox[k];
        if (!g->max_fixed) g->bbmax[k] = g->bbmax_init[k] + t * g->vel_bbox[k];
      }
    }
    return MFG_OK;
  }
  return grid_update_unlocked(g, p, kernel_width);
}

// hash (+ blanking), scan, scatter, per-cell fix-up: the sort permutation and cell offsets
// what decides whether the cell tables of a sort still describe a grid: origin, cell size and
// cells per axis
void grid_geometry(const mfg_grid *g, double geom[7]) {
  for (int k = 0; k < 3; k++) geom[k] = g->bbmin[k];
  geom[3] = g->dx;
  geom[4] = (double)g->nx;
  geom[5] = (double)g->ny;
  geom[6] = (double)g->nz;
}

// Whether a step may keep the particle order and cell tables of the last sort (the opt-in
// re-sort interval): the interval is not over, and the set was last sorted on this grid, which
// has sorted no other set since, is locked, does not move and has the geometry it had then.
bool keep_last_sort(const mfg_grid *g, const mfg_particles *p) {
  const int k = g_lib.opt.sort_interval;
  if (k <= 1 || p->sorted_grid != g || g->sorted_for != p || p->since_sort + 1 >= k || !g->locked)
    return false;
  if (g->vel_bbox[0] != 0. || g->vel_bbox[1] != 0. || g->vel_bbox[2] != 0.) return false;
  double geom[7];
  grid_geometry(g, geom);
  for (int a = 0; a < 7; a++)
    if (geom[a] != p->sorted_geom[a]) return false;
  return true;
}

void sort_keys(mfg_grid *g, mfg_particles *p, const mfg_blanking_state *b) {
  GridBuffers gb = grid_buffers(g, p);
  GridInfo gi = g->info();
  double bb[6];
  double vmax2 = FLT_MAX;
  if (b) {
    for (int k = 0; k < 3; k++) {
      bb[k] = b->bbmin[k];
      bb[3 + k] = b->bbmax[k];
    }
    vmax2 = b->max_vel_squared;
  }
  {
    ProfScope ps("blank_hash_count");
    launch_blank_hash_count(g_lib.stream, p->st(), p->N_init, gi, b ? 1 : 0, b ? bb : nullptr,
                            vmax2, gb, p->cnt);
  }
  {
    ProfScope ps("scan_cells");
    launch_scan_cells(g_lib.stream, gi, gb, p->N_init, p->cnt);
  }
  {
    ProfScope ps("scatter");
    launch_scatter(g_lib.stream, p->N_init, gi, gb, p->cnt);
  }
  {
    ProfScope ps("cell_fixup");
    launch_fixup(g_lib.stream, gi, gb);
  }
  g->scanned = true;
  g->sorted_for = p;
  p->sorted_grid = g;
  grid_geometry(g, p->sorted_geom);
  p->since_sort = 0;
}

// identity: the particle order stays (a step between sorts): no gather, keys kept
void permute(mfg_grid *g, mfg_particles *p, const Consts &c, int predict, bool identity = false) {
  GridBuffers gb = grid_buffers(g, p);
  if (identity) gb.order = nullptr;
  ProfScope ps(predict ? "permute_predict" : "permute");
  int nxt = 1 - p->cur;
  launch_permute_predict_ex(g_lib.stream, p->buf[p->cur], p->buf[nxt], p->N_init, gb, p->cnt, c,
                            predict, p->rec, p->nl4, p->pse2, p->cs4, p->p_dbg, p->R_dbg);
  p->cur = nxt;
}

// the per-action arrays of a particle set (allocated and zeroed on first use)
int ensure_aux(mfg_particles *p) {
  if (p->vder) return MFG_OK;
  const size_t n = padded(p->N_init);
  hipStream_t s = g_lib.stream;
  if (!p->p_dbg) {
    if (hipMalloc(&p->p_dbg, sizeof(double) * n) != hipSuccess ||
        hipMalloc(&p->R_dbg, sizeof(double) * 6 * n) != hipSuccess)
      return fail(MFG_ERR_NOMEM, "pressure / artificial stress arrays");
    HIP_TRY(hipMemsetAsync(p->p_dbg, 0, sizeof(double) * n, s));
    HIP_TRY(hipMemsetAsync(p->R_dbg, 0, sizeof(double) * 6 * n, s));
  }
  if (hipMalloc(&p->vder, sizeof(double) * 9 * n) != hipSuccess ||
      hipMalloc(&p->sder, sizeof(double) * 9 * n) != hipSuccess ||
      hipMalloc(&p->nrm, sizeof(double) * 3 * n) != hipSuccess ||
      hipMalloc(&p->x0, sizeof(double) * 14 * n) != hipSuccess)
    return fail(MFG_ERR_NOMEM, "per-action arrays (%u particles)", p->N_init);
  HIP_TRY(hipMemsetAsync(p->vder, 0, sizeof(double) * 9 * n, s));
  HIP_TRY(hipMemsetAsync(p->sder, 0, sizeof(double) * 9 * n, s));
  HIP_TRY(hipMemsetAsync(p->nrm, 0, sizeof(double) * 3 * n, s));
  HIP_TRY(hipMemsetAsync(p->x0, 0, sizeof(double) * 14 * n, s));
  return MFG_OK;
}

Aux aux_of(const mfg_particles *p) {
  Aux a;
  a.p = p->p_dbg;
  a.R = p->R_dbg;
  a.vder = p->vder;
  a.sder = p->sder;
  a.nrm = p->nrm;
  a.x0 = p->x0;
  a.n = padded(p->N_init);
  return a;
}

// buffers of the lane-parallel exhaustive contact search for n_tools tools
int contact_search_buffers(mfg_particles *p, int n_tools, ContactSearch &cs) {
  const unsigned cap = (unsigned)std::min<size_t>((size_t)p->N_init * (size_t)n_tools, (size_t)1 << 22);
  if (n_tools > p->cs_tools || cap > p->cs_cap) {
    void *old[] = {p->cs_list, p->cs_n, p->cs_slot, p->cs_res};
    for (void *q : old)
      if (q) (void)hipFree(q);
    p->cs_list = nullptr; p->cs_n = nullptr; p->cs_slot = nullptr; p->cs_res = nullptr;
    p->cs_cap = 0; p->cs_tools = 0;
    if (hipMalloc(&p->cs_list, sizeof(int2) * cap) != hipSuccess ||
        hipMalloc(&p->cs_n, sizeof(unsigned)) != hipSuccess ||
        hipMalloc(&p->cs_slot, sizeof(int) * (size_t)p->N_init * n_tools) != hipSuccess ||
        hipMalloc(&p->cs_res, sizeof(double) * 6 * (size_t)cap) != hipSuccess)
      return fail(MFG_ERR_NOMEM, "contact search buffers (%u entries)", cap);
    p->cs_cap = cap;
    p->cs_tools = n_tools;
  }
  cs.list = p->cs_list;
  cs.list_n = p->cs_n;
  cs.cap = p->cs_cap;
  cs.slot = p->cs_slot;
  cs.res = p->cs_res;
  return MFG_OK;
}

// neighbour lists (+ PSE) of the current records on grid g; stats: keep the per-workgroup
// candidate / pair totals (a buffer sized on first use) for mfg_particle_gpu_pair_stats
// stats: the pass may count its pairs (when counting is on, mfg_pair_stats_enable)
int run_neighbours(mfg_grid *g, mfg_particles *p, const Consts &c, int want_pse, bool stats = true) {
  const GridInfo gi = g->info();
  const int tile = resolve_cell_tile(g_lib.opt.cell_tile, gi.num_cell, p->N_init);
  if (!stats || !g_lib.count_pairs) {
    p->stat_n = 0;
    launch_neighbours(g_lib.stream, gi, g->count, p->nl4, p->pse2, p->st().h, p->st().flags, c,
                      p->nbr, p->nl_stride, p->count, p->st().T_t, p->cnt, want_pse, nullptr,
                      p->poison, tile);
    HIP_TRY(hipGetLastError());
    return MFG_OK;
  }
  const int nb = neighbour_blocks(gi, tile);
  if (nb > p->stat_cap) {
    if (p->stat_part) (void)hipFree(p->stat_part);
    p->stat_part = nullptr;
    p->stat_cap = 0;
    if (hipMalloc(&p->stat_part, sizeof(unsigned long long) * 2 * (size_t)nb) != hipSuccess)
      return fail(MFG_ERR_NOMEM, "neighbour statistics (%d workgroups)", nb);
    p->stat_cap = nb;
  }
  p->stat_n = nb;
  launch_neighbours(g_lib.stream, gi, g->count, p->nl4, p->pse2, p->st().h, p->st().flags, c,
                    p->nbr, p->nl_stride, p->count, p->st().T_t, p->cnt, want_pse, p->stat_part,
                    p->poison, tile);
  HIP_TRY(hipGetLastError());
  return MFG_OK;
}

int check_counters(mfg_particles *p) {
  HIP_TRY(hipMemcpyAsync(p->cnt_h, p->cnt, sizeof(Counters), hipMemcpyDeviceToHost, g_lib.stream));
  HIP_TRY(hipStreamSynchronize(g_lib.stream));
  p->N_host = p->cnt_h->N;
  if (p->cnt_h->err_capacity)
    return fail(MFG_ERR_CAPACITY, "a particle has %d neighbours, list capacity is %d",
                p->cnt_h->max_count, p->cap);
  if (p->cnt_h->err_grid) return fail(MFG_ERR_GRID, "unblanked particle outside the grid");
  if (p->cnt_h->err_list)
    return fail(MFG_ERR_STATE, "a neighbour list entry lies outside the particle's cell stencil");
  // the reference stops the run here (src/grid_gpu_base.cu:159-163); N is 0 only when every
  // particle was hard-blanked
  if (p->N_host == 0 && p->N_init > 0)
    return fail(MFG_ERR_ALL_BLANKED, "all particles left bbox due to blanking");
  return MFG_OK;
}

// ---- mesh normals (set-up time, host): face, angle-weighted edge and vertex normals ---------
struct HV {
  double x, y, z;
};
inline HV hsub(HV a, HV b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline HV hadd(HV a, HV b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline HV hmul(double s, HV a) { return {s * a.x, s * a.y, s * a.z}; }
inline double hdot(HV a, HV b) {
  double x = a.x * b.x, y = a.y * b.y, z = a.z * b.z;
  return x + y + z;
}
inline HV hcross(HV a, HV b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
inline double hlen(HV a) { return std::sqrt(hdot(a, a)); }
inline HV hnorm(HV a) { return hmul(1.0 / std::sqrt(hdot(a, a)), a); }

int mesh_to_tris(const double *verts, int nv, const int *t3, int nt, std::vector<Tri> &out,
                 double &h_avg_out) {
  std::vector<HV> P(nv);
  for (int i = 0; i < nv; i++) P[i] = {verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
  std::vector<std::vector<int>> vt(nv);
  std::unordered_map<uint64_t, std::vector<int>> edges;
  auto key = [](int a, int b) {
    uint64_t lo = (uint32_t)std::min(a, b), hi = (uint32_t)std::max(a, b);
    return (hi << 32) | lo;
  };
  for (int t = 0; t < nt; t++) {
    int a = t3[3 * t], b = t3[3 * t + 1], c = t3[3 * t + 2];
    if (a < 0 || b < 0 || c < 0 || a >= nv || b >= nv || c >= nv)
      return fail(MFG_ERR_ARG, "triangle %d references a missing vertex", t);
    vt[a].push_back(t);
    vt[b].push_back(t);
    vt[c].push_back(t);
    edges[key(a, b)].push_back(t);
    edges[key(b, c)].push_back(t);
    edges[key(c, a)].push_back(t);
  }
  for (auto &e : edges)
    if (e.second.size() != 2)
      return fail(MFG_ERR_ARG, "mesh is not closed: an edge is shared by %zu triangles",
                  e.second.size());
  std::vector<HV> fn(nt);
  for (int t = 0; t < nt; t++) {
    HV p1 = P[t3[3 * t]], p2 = P[t3[3 * t + 1]], p3 = P[t3[3 * t + 2]];
    fn[t] = hnorm(hcross(hsub(p2, p1), hsub(p3, p1)));
  }
  auto vnormal = [&](int v) {
    HV pos = P[v], np{0., 0., 0.};
    for (int t : vt[v]) {
      HV q[3] = {P[t3[3 * t]], P[t3[3 * t + 1]], P[t3[3 * t + 2]]};
      HV f1{0, 0, 0}, f2{0, 0, 0};
      if (hlen(hsub(q[0], pos)) < 1e-12) { f1 = q[1]; f2 = q[2]; }
      if (hlen(hsub(q[1], pos)) < 1e-12) { f1 = q[0]; f2 = q[2]; }
      if (hlen(hsub(q[2], pos)) < 1e-12) { f1 = q[0]; f2 = q[1]; }
      HV a = hsub(f1, pos), b = hsub(f2, pos);
      double ang = std::acos(std::fabs(hdot(a, b) / (hlen(a) * hlen(b))));
      np = hadd(np, hmul(ang, fn[t]));
    }
    return hnorm(np);
  };
  std::vector<HV> vn(nv);
  std::vector<char> have(nv, 0);
  out.resize(nt);
  double lmin = DBL_MAX, lmax = 0.;
  for (int t = 0; t < nt; t++) {
    int id[3] = {t3[3 * t], t3[3 * t + 1], t3[3 * t + 2]};
    Tri &r = out[t];
    for (int q = 0; q < 3; q++) {
      r.p[q][0] = P[id[q]].x;
      r.p[q][1] = P[id[q]].y;
      r.p[q][2] = P[id[q]].z;
      if (!have[id[q]]) {
        vn[id[q]] = vnormal(id[q]);
        have[id[q]] = 1;
      }
      r.np[q][0] = vn[id[q]].x;
      r.np[q][1] = vn[id[q]].y;
      r.np[q][2] = vn[id[q]].z;
      const std::vector<int> &e = edges[key(id[q], id[(q + 1) % 3])];
      HV ne = hnorm(hadd(fn[e[0]], fn[e[1]]));
      r.ne[q][0] = ne.x;
      r.ne[q][1] = ne.y;
      r.ne[q][2] = ne.z;
    }
    r.n[0] = fn[t].x;
    r.n[1] = fn[t].y;
    r.n[2] = fn[t].z;
    HV p1 = P[id[0]], p2 = P[id[1]], p3 = P[id[2]];
    double l1 = hlen(hsub(p1, p2)), l2 = hlen(hsub(p2, p3)), l3 = hlen(hsub(p3, p1));
    lmin = std::fmin(std::fmin(std::fmin(lmin, l1), l2), l3);
    lmax = std::fmax(std::fmax(std::fmax(lmax, l1), l2), l3);
  }
  h_avg_out = 0.5 * (lmax + lmin);
  return MFG_OK;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// kernels local to the ABI: blanking only, field conversion
// ---------------------------------------------------------------------------------------------
namespace {

__global__ void k_blanking_only(State st, unsigned n, double b0, double b1, double b2, double b3,
                                double b4, double b5, double vmax2) {
  unsigned i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double px = st.x[i], py = st.y[i], pz = st.z[i];
  double vx = st.vx[i], vy = st.vy[i], vz = st.vz[i];
  bool too_fast = vx * vx + vy * vy + vz * vz >= vmax2;
  bool in = px >= b0 && px <= b3 && py >= b1 && py <= b4 && pz >= b2 && pz <= b5;
  uint32_t fl = st.flags[i];
  if (too_fast || !in) {
    fl |= FL_BLANKED;
  } else {
    auto bad = [](double v) { return isnan(v) || isinf(v); };
    bool inv = bad(st.xt[i]) || bad(st.yt[i]) || bad(st.zt[i]) || bad(st.vxt[i]) ||
               bad(st.vyt[i]) || bad(st.vzt[i]) || bad(st.St[0][i]) || bad(st.St[1][i]) ||
               bad(st.St[2][i]) || bad(st.St[3][i]) || bad(st.St[4][i]) || bad(st.St[5][i]) ||
               bad(st.rho_t[i]) || bad(st.T_t[i]);
    if (!inv) fl &= ~FL_BLANKED;
  }
  st.flags[i] = fl;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------------
extern "C" {

int mfg_version(void) { return 1; }

const char *mfg_last_error(void) { return g_err.c_str(); }

int mfg_init(int device) {
  if (g_lib.ready) {
    if (device == g_lib.device) return MFG_OK;
    return fail(MFG_ERR_STATE, "library already bound to device %d", g_lib.device);
  }
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(MFG_ERR_ARG, "device %d of %d", device, n);
  HIP_TRY(hipSetDevice(device));
  HIP_TRY(hipStreamCreateWithFlags(&g_lib.stream, hipStreamNonBlocking));
  g_lib.streams[0] = g_lib.stream;
  g_lib.n_streams = 1;
  // The first pool streams are created here, before any blocking copy brings up the runtime's
  // own stream: streams take the process's hardware queues (4 by default) in creation order, and
  // a pool created later shared a queue between two of its streams (the batched form ran 23 %
  // slower when an instance had been set up before the pool).  Default priority, like the
  // library stream.
  for (int k = 1; k < MFG_EARLY_STREAMS; k++)
    HIP_TRY(hipStreamCreateWithFlags(&g_lib.streams[k], hipStreamNonBlocking));
  g_lib.n_created = MFG_EARLY_STREAMS;
  {
    void *ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(MFG_ERR_HIP, "no device address for the step's work memory");
    g_scratch = (unsigned char *)ptr;
  }
  const char *lane_contact = std::getenv("MFG_LANE_CONTACT");
  g_lib.opt_contact_search = !(lane_contact && lane_contact[0] == '0');
  g_lib.device = device;
  g_lib.ready = true;
  return MFG_OK;
}

void *mfg_get_stream(void) { return (void *)g_lib.stream; }

int mfg_synchronize(void) {
  NEED_LIB();
  HIP_TRY(hipStreamSynchronize(g_lib.stream));
  return MFG_OK;
}

int mfg_stream_pool(int n) {
  NEED_LIB();
  if (n < 1 || n > MFG_MAX_STREAMS) return fail(MFG_ERR_ARG, "%d streams (1 .. %d)", n, MFG_MAX_STREAMS);
  // default priority, like the library stream
  for (int k = g_lib.n_created; k < n; k++)
    HIP_TRY(hipStreamCreateWithFlags(&g_lib.streams[k], hipStreamNonBlocking));
  if (n > g_lib.n_created) g_lib.n_created = n;
  if (n > g_lib.n_streams) g_lib.n_streams = n;
  return MFG_OK;
}

int mfg_select_stream(int k) {
  NEED_LIB();
  if (k < 0 || k >= g_lib.n_streams) return fail(MFG_ERR_ARG, "stream %d of %d", k, g_lib.n_streams);
  g_lib.stream = g_lib.streams[k];
  return MFG_OK;
}

int mfg_device_synchronize(void) {
  NEED_LIB();
  for (int k = 0; k < g_lib.n_streams; k++) HIP_TRY(hipStreamSynchronize(g_lib.streams[k]));
  return MFG_OK;
}

int mfg_constants_reset(void) {
  g_lib.a_phys = g_lib.i_phys = mfg_phys_constants{};
  g_lib.a_corr = g_lib.i_corr = mfg_corr_constants{};
  g_lib.joco = mfg_joco_constants{};
  g_lib.a_twp = g_lib.a_ttool = g_lib.i_twp = g_lib.i_ttool = mfg_trml_constants{};
  g_lib.plastic = g_lib.a_thermal = g_lib.a_fric = g_lib.i_thermal_wp = false;
  return MFG_OK;
}

int mfg_set_options(const mfg_options *o) {
  if (!o) return fail(MFG_ERR_ARG, "null options");
  if (o->heat < MFG_HEAT_NONE || o->heat > MFG_HEAT_PSE) return fail(MFG_ERR_ARG, "heat model %d", o->heat);
  if (o->pair_pass != MFG_PAIR_LIST && o->pair_pass != MFG_PAIR_TILED)
    return fail(MFG_ERR_ARG, "pair pass %d", o->pair_pass);
  if (o->cell_tile != 0 && o->cell_tile != 1 && o->cell_tile != 2 && o->cell_tile != 4 &&
      o->cell_tile != 8 && o->cell_tile != 16)
    return fail(MFG_ERR_ARG, "cell tile %d (0, 1, 2, 4, 8 or 16)", o->cell_tile);
  if (o->sort_interval < 0 || o->sort_interval > 1000)
    return fail(MFG_ERR_ARG, "sort interval %d (0 to 1000)", o->sort_interval);
  if (o->pair_block != 0 && o->pair_block != 64 && o->pair_block != 128 && o->pair_block != 256)
    return fail(MFG_ERR_ARG, "pair-pass workgroup size %d (0, 64, 128 or 256)", o->pair_block);
  g_lib.opt = *o;
  if (g_lib.opt.max_neighbours <= 0) g_lib.opt.max_neighbours = 256;
  // list rows are read as 16-byte vectors: a multiple of 4 entries, at least 8
  g_lib.opt.max_neighbours = std::max(8, (g_lib.opt.max_neighbours + 3) / 4 * 4);
  return MFG_OK;
}

int mfg_get_options(mfg_options *o) {
  if (!o) return fail(MFG_ERR_ARG, "null options");
  *o = g_lib.opt;
  return MFG_OK;
}

// ---- constants ----
int mfg_actions_setup_physical_constants(const mfg_phys_constants *c) {
  if (!c) return fail(MFG_ERR_ARG, "null constants");
  g_lib.a_phys = *c;
  return MFG_OK;
}
int mfg_actions_setup_corrector_constants(const mfg_corr_constants *c) {
  if (!c) return fail(MFG_ERR_ARG, "null constants");
  g_lib.a_corr = *c;
  return MFG_OK;
}
int mfg_actions_setup_johnson_cook_constants(const mfg_joco_constants *c) {
  if (!c) return fail(MFG_ERR_ARG, "null constants");
  g_lib.joco = *c;
  g_lib.plastic = true;
  return MFG_OK;
}
int mfg_actions_setup_thermal_constants_wp(const mfg_trml_constants *c) {
  if (!c) return fail(MFG_ERR_ARG, "null constants");
  g_lib.a_twp = *c;
  if (c->eta != 0.) g_lib.a_fric = true;
  if (c->alpha != 0.) g_lib.a_thermal = true;
  return MFG_OK;
}
int mfg_actions_setup_thermal_constants_tool(const mfg_trml_constants *c) {
  if (!c) return fail(MFG_ERR_ARG, "null constants");
  g_lib.a_ttool = *c;
  if (c->alpha != 0.) g_lib.a_thermal = true;
  return MFG_OK;
}
int mfg_interactions_setup_physical_constants(const mfg_phys_constants *c) {
  if (!c) return fail(MFG_ERR_ARG, "null constants");
  g_lib.i_phys = *c;
  return MFG_OK;
}
int mfg_interactions_setup_corrector_constants(const mfg_corr_constants *c) {
  if (!c) return fail(MFG_ERR_ARG, "null constants");
  g_lib.i_corr = *c;
  return MFG_OK;
}
int mfg_interactions_setup_thermal_constants_workpiece(const mfg_trml_constants *c) {
  if (!c) return fail(MFG_ERR_ARG, "null constants");
  g_lib.i_twp = *c;
  g_lib.i_thermal_wp = c->alpha != 0.;
  return MFG_OK;
}
int mfg_interactions_setup_thermal_constants_tool(const mfg_trml_constants *c, mfg_tool *tool) {
  if (!c) return fail(MFG_ERR_ARG, "null constants");
  g_lib.i_ttool = *c;
  if (tool && c->alpha != 0.) tool->thermal = true;
  return MFG_OK;
}
int mfg_interactions_setup_geometry_constants(const mfg_grid *g) {
  // the grid geometry is passed to every kernel launch; nothing to upload
  return g ? MFG_OK : fail(MFG_ERR_ARG, "null grid");
}

// ---- particles ----
int mfg_field_width(int f) {
  switch (f) {
    case MFG_F_POS: case MFG_F_VEL: case MFG_F_VEL_BC: case MFG_F_FC: case MFG_F_FT:
    case MFG_F_POS_T: case MFG_F_VEL_T:
      return 3;
    case MFG_F_S: case MFG_F_R: case MFG_F_S_T: case MFG_F_S_DER: case MFG_F_V_DER:
      return 9;
    case MFG_F_N:
      return 3;
    case MFG_F_H: case MFG_F_RHO: case MFG_F_P: case MFG_F_EPS_PL: case MFG_F_EPS_PL_DOT:
    case MFG_F_T: case MFG_F_FIXED: case MFG_F_BLANKED: case MFG_F_TOOL: case MFG_F_RHO_T:
    case MFG_F_T_T: case MFG_F_HASH: case MFG_F_IDX: case MFG_F_ID:
      return 1;
    default:
      return 0;
  }
}

int mfg_particle_gpu_create(unsigned N, const double *pos4, const double *vel4, const double *rho,
                            const double *T_init, const double *h, const double *fixed,
                            const double *tool_p, const double *S9, int precision,
                            mfg_particles **out) {
  NEED_LIB();
  if (!out || !pos4 || N == 0) return fail(MFG_ERR_ARG, "particle_gpu: need N > 0 and positions");
  if (precision != MFG_FP64 && precision != MFG_FP32_KAHAN && precision != MFG_FP32_WIDE)
    return fail(MFG_ERR_ARG, "precision %d", precision);
  mfg_particles *p = new mfg_particles();
  p->N_init = N;
  p->N_host = N;
  p->precision = precision;
  const size_t n = padded(N);
  auto cleanup = [&](int code) {
    free_particles(p);
    delete p;
    return code;
  };
  for (int b = 0; b < 2; b++) {
    if (hipMalloc(&p->arena[b], sizeof(double) * STATE_DOUBLES * n) != hipSuccess ||
        hipMalloc(&p->uarena[b], sizeof(uint32_t) * 2 * n) != hipSuccess)
      return cleanup(fail(MFG_ERR_NOMEM, "particle state (%u particles)", N));
    bind_state(p->buf[b], p->arena[b], p->uarena[b], n);
  }
  size_t rec_bytes = precision == MFG_FP64 ? sizeof(Rec64) : sizeof(Rec32);
  p->cap = g_lib.opt.max_neighbours > 0 ? g_lib.opt.max_neighbours : 256;
  p->nl_stride = p->cap;  // row-major lists: nbr[i * cap + k]
  p->sum_stride = n;
  p->nblk = (int)((N + 255) / 256);
  p->contact_blocks = p->nblk;
  if (hipMalloc(&p->rec, rec_bytes * n) != hipSuccess || hipMalloc(&p->nl4, sizeof(double4) * n) != hipSuccess ||
      hipMalloc(&p->pse2, sizeof(double2) * n) != hipSuccess ||
      hipMalloc(&p->cs4, sizeof(double4) * n) != hipSuccess ||
      hipMalloc(&p->sums, sizeof(double) * PAIR_SUM_COUNT * n) != hipSuccess ||
      hipMalloc(&p->trav, sizeof(int) * n) != hipSuccess ||
      hipMalloc(&p->poison, sizeof(int) * n) != hipSuccess ||
      hipMalloc(&p->nbr, sizeof(int) * (size_t)p->cap * n) != hipSuccess ||
      hipMalloc(&p->count, sizeof(int) * n) != hipSuccess || hipMalloc(&p->key, sizeof(int) * n) != hipSuccess ||
      hipMalloc(&p->key_sorted, sizeof(int) * n) != hipSuccess ||
      hipMalloc(&p->order, sizeof(int) * n) != hipSuccess ||
      hipMalloc(&p->blk, sizeof(int) * (p->nblk + 1)) != hipSuccess ||
      hipMalloc(&p->cnt, sizeof(Counters)) != hipSuccess ||
      hipMalloc(&p->partials, sizeof(float) * MAX_TOOLS * 3 * p->contact_blocks) != hipSuccess ||
      hipMalloc(&p->flag_d, sizeof(int)) != hipSuccess ||
      hipHostMalloc(&p->cnt_h, sizeof(Counters)) != hipSuccess)
    return cleanup(fail(MFG_ERR_NOMEM, "particle buffers (%u particles)", N));
  hipStream_t s = g_lib.stream;
  State &st = p->buf[0];
  (void)hipMemsetAsync(p->arena[0], 0, sizeof(double) * STATE_DOUBLES * n, s);
  (void)hipMemsetAsync(p->arena[1], 0, sizeof(double) * STATE_DOUBLES * n, s);
  (void)hipMemsetAsync(p->uarena[1], 0, sizeof(uint32_t) * 2 * n, s);
  (void)hipMemsetAsync(p->count, 0, sizeof(int) * n, s);
  (void)hipMemsetAsync(p->poison, 0, sizeof(int) * n, s);
  (void)hipMemsetAsync(p->key_sorted, 0, sizeof(int) * n, s);
  (void)hipMemsetAsync(p->cnt, 0, sizeof(Counters), s);
  // host staging in the device layout
  std::vector<double> tmp(N);
  auto up = [&](double *dst, const double *src, int stride, int comp) -> bool {
    for (unsigned i = 0; i < N; i++) tmp[i] = src[(size_t)stride * i + comp];
    return hipMemcpyAsync(dst, tmp.data(), sizeof(double) * N, hipMemcpyHostToDevice, s) == hipSuccess &&
           hipStreamSynchronize(s) == hipSuccess;
  };
  bool ok = up(st.x, pos4, 4, 0) && up(st.y, pos4, 4, 1) && up(st.z, pos4, 4, 2);
  if (vel4) ok = ok && up(st.vx, vel4, 4, 0) && up(st.vy, vel4, 4, 1) && up(st.vz, vel4, 4, 2);
  if (rho) ok = ok && up(st.rho, rho, 1, 0);
  if (T_init) ok = ok && up(st.T, T_init, 1, 0);
  if (h) ok = ok && up(st.h, h, 1, 0);
  if (S9) {
    // xx yy zz xy yz zx from GLM storage m[col][row] = S9[3*col+row]
    const int src_idx[6] = {0, 4, 8, 1, 5, 2};
    for (int k = 0; k < 6; k++) ok = ok && up(st.S[k], S9, 9, src_idx[k]);
  }
  std::vector<uint32_t> fu(2 * n, 0);
  for (unsigned i = 0; i < N; i++) {
    uint32_t f = 0;
    double fx = fixed ? fixed[i] : 0., tp = tool_p ? tool_p[i] : 0.;
    if (fx != 0.) f |= FL_FIXED;
    if (fx == 1.) f |= FL_FIXED_ONE;
    if (tp == 1.) f |= FL_TOOL;
    if (tp != 0.) f |= FL_TOOL_NZ;
    fu[i] = f;
    fu[n + i] = i;
  }
  // the permutation starts as the identity (the ids): a set is never gathered through
  // unwritten indices
  ok = ok && hipMemcpyAsync(p->uarena[0], fu.data(), sizeof(uint32_t) * 2 * n, hipMemcpyHostToDevice, s) == hipSuccess &&
       hipMemcpyAsync(p->order, fu.data() + n, sizeof(int) * n, hipMemcpyHostToDevice, s) == hipSuccess &&
       hipStreamSynchronize(s) == hipSuccess;
  if (!ok) return cleanup(fail(MFG_ERR_HIP, "particle upload failed"));
  // counters: N = N_init until the first sort
  Counters c0;
  std::memset(&c0, 0, sizeof(c0));
  c0.N = N;
  if (hipMemcpy(p->cnt, &c0, sizeof(c0), hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(MFG_ERR_HIP, "counter upload failed"));
  *out = p;
  return MFG_OK;
}

int mfg_particle_gpu_destroy(mfg_particles *p) {
  if (!p) return MFG_OK;
  if (g_lib.ready) (void)hipStreamSynchronize(g_lib.stream);
  free_particles(p);
  delete p;
  return MFG_OK;
}

int mfg_particle_gpu_pair_stats(mfg_particles *p, unsigned long long *cand,
                                unsigned long long *pairs) {
  NEED_LIB();
  if (!p) return fail(MFG_ERR_ARG, "null particles");
  int rc = check_counters(p);
  if (rc == MFG_OK && p->stat_n == 0)
    rc = fail(MFG_ERR_STATE, "the last neighbour pass did not count its pairs (mfg_pair_stats_enable)");
  unsigned long long a = 0, b = 0;
  if (p->stat_n > 0) {
    std::vector<unsigned long long> h(2 * (size_t)p->stat_n);
    HIP_TRY(hipMemcpy(h.data(), p->stat_part, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < p->stat_n; k++) {
      a += h[2 * k];
      b += h[2 * k + 1];
    }
  }
  if (cand) *cand = a;
  if (pairs) *pairs = b;
  return rc;
}

int mfg_particle_gpu_precision(const mfg_particles *p) { return p ? p->precision : MFG_ERR_ARG; }

int mfg_particle_gpu_counts(mfg_particles *p, unsigned *N, unsigned *N_init) {
  NEED_LIB();
  if (!p) return fail(MFG_ERR_ARG, "null particles");
  prof_collect();
  int rc = check_counters(p);
  if (N) *N = p->N_host;
  if (N_init) *N_init = p->N_init;
  return rc;
}

// snapshot layout: header, then the current state buffer as the device holds it (the double
// arena, STATE_DOUBLES arrays of padded(N_init), and the flag / id arena, 2 x padded(N_init) u32)
struct SnapshotHeader {
  unsigned long long magic;
  unsigned N_init, N;
  int precision, n_doubles;
  unsigned long long padded_n;
};
constexpr unsigned long long SNAPSHOT_MAGIC = 0x6d66672d736e6170ull;  // "mfg-snap"

int mfg_particle_gpu_snapshot_bytes(const mfg_particles *p, unsigned long long *bytes) {
  if (!p || !bytes) return fail(MFG_ERR_ARG, "null argument");
  const size_t n = padded(p->N_init);
  *bytes = sizeof(SnapshotHeader) + sizeof(double) * STATE_DOUBLES * n + sizeof(uint32_t) * 2 * n;
  return MFG_OK;
}

int mfg_particle_gpu_snapshot(mfg_particles *p, void *host) {
  NEED_LIB();
  if (!p || !host) return fail(MFG_ERR_ARG, "null argument");
  int rc = check_counters(p);  // waits for the queued steps; N of the last one
  if (rc != MFG_OK && rc != MFG_ERR_ALL_BLANKED) return rc;
  const size_t n = padded(p->N_init);
  SnapshotHeader h{SNAPSHOT_MAGIC, p->N_init, p->N_host, p->precision, STATE_DOUBLES, n};
  unsigned char *dst = (unsigned char *)host;
  std::memcpy(dst, &h, sizeof(h));
  dst += sizeof(h);
  HIP_TRY(hipMemcpy(dst, p->arena[p->cur], sizeof(double) * STATE_DOUBLES * n, hipMemcpyDeviceToHost));
  dst += sizeof(double) * STATE_DOUBLES * n;
  HIP_TRY(hipMemcpy(dst, p->uarena[p->cur], sizeof(uint32_t) * 2 * n, hipMemcpyDeviceToHost));
  return MFG_OK;
}

int mfg_particle_gpu_restore(mfg_particles *p, const void *host) {
  NEED_LIB();
  if (!p || !host) return fail(MFG_ERR_ARG, "null argument");
  p->sorted_grid = nullptr;  // another order: the next step sorts
  SnapshotHeader h;
  std::memcpy(&h, host, sizeof(h));
  const size_t n = padded(p->N_init);
  if (h.magic != SNAPSHOT_MAGIC || h.n_doubles != STATE_DOUBLES || h.padded_n != n)
    return fail(MFG_ERR_ARG, "not a snapshot of this library's particle layout");
  if (h.N_init != p->N_init || h.precision != p->precision || h.N > h.N_init)
    return fail(MFG_ERR_ARG, "snapshot of %u particles (precision %d) into a set of %u (precision %d)",
                h.N_init, h.precision, p->N_init, p->precision);
  hipStream_t s = g_lib.stream;
  HIP_TRY(hipStreamSynchronize(s));  // nothing queued may still use the buffers
  const unsigned char *src = (const unsigned char *)host + sizeof(h);
  HIP_TRY(hipMemcpy(p->arena[p->cur], src, sizeof(double) * STATE_DOUBLES * n, hipMemcpyHostToDevice));
  src += sizeof(double) * STATE_DOUBLES * n;
  HIP_TRY(hipMemcpy(p->uarena[p->cur], src, sizeof(uint32_t) * 2 * n, hipMemcpyHostToDevice));
  // N as the snapshot's last step left it; errors of the set's earlier life are cleared
  Counters c0;
  std::memset(&c0, 0, sizeof(c0));
  c0.N = h.N;
  HIP_TRY(hipMemcpy(p->cnt, &c0, sizeof(c0), hipMemcpyHostToDevice));
  p->N_host = h.N;
  p->stat_n = 0;
  return MFG_OK;
}

int mfg_particle_gpu_download(mfg_particles *p, int f, double *host) {
  NEED_LIB();
  if (!p || !host) return fail(MFG_ERR_ARG, "null argument");
  // errors of the queued steps surface here too (the copies below synchronise anyway)
  const int rc0 = check_counters(p);
  if (rc0 != MFG_OK) return rc0;
  const unsigned N = p->N_init;
  const State &st = p->st();
  hipStream_t s = g_lib.stream;
  std::vector<double> a(N);
  auto get = [&](const double *src) -> bool {
    return hipMemcpyAsync(a.data(), src, sizeof(double) * N, hipMemcpyDeviceToHost, s) == hipSuccess &&
           hipStreamSynchronize(s) == hipSuccess;
  };
  auto put3 = [&](const double *x, const double *y, const double *z) -> int {
    const double *src[3] = {x, y, z};
    for (int k = 0; k < 3; k++) {
      if (!get(src[k])) return fail(MFG_ERR_HIP, "download failed");
      for (unsigned i = 0; i < N; i++) host[3 * (size_t)i + k] = a[i];
    }
    return MFG_OK;
  };
  auto put_sym = [&](double *const *S6) -> int {
    // GLM storage m[col][row]; symmetric entries mirrored
    const int dst_of[6][2] = {{0, 0}, {4, 4}, {8, 8}, {1, 3}, {5, 7}, {2, 6}};
    for (int k = 0; k < 6; k++) {
      if (!get(S6[k])) return fail(MFG_ERR_HIP, "download failed");
      for (unsigned i = 0; i < N; i++) {
        host[9 * (size_t)i + dst_of[k][0]] = a[i];
        host[9 * (size_t)i + dst_of[k][1]] = a[i];
      }
    }
    return MFG_OK;
  };
  auto put1 = [&](const double *x) -> int {
    if (!get(x)) return fail(MFG_ERR_HIP, "download failed");
    std::memcpy(host, a.data(), sizeof(double) * N);
    return MFG_OK;
  };
  auto put_u = [&](const void *src, int kind) -> int {
    std::vector<uint32_t> u(N);
    if (hipMemcpyAsync(u.data(), src, sizeof(uint32_t) * N, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail(MFG_ERR_HIP, "download failed");
    for (unsigned i = 0; i < N; i++) {
      uint32_t v = u[i];
      switch (kind) {
        case 0: host[i] = (v & FL_FIXED_ONE) ? 1. : ((v & FL_FIXED) ? 2. : 0.); break;
        case 1: host[i] = (v & FL_BLANKED) ? 1. : 0.; break;
        case 2: host[i] = (v & FL_TOOL) ? 1. : ((v & FL_TOOL_NZ) ? 2. : 0.); break;
        default: host[i] = (double)(int)v; break;
      }
    }
    return MFG_OK;
  };
  switch (f) {
    case MFG_F_POS: return put3(st.x, st.y, st.z);
    case MFG_F_VEL: return put3(st.vx, st.vy, st.vz);
    case MFG_F_VEL_BC: return put3(st.vbx, st.vby, st.vbz);
    case MFG_F_FC: return put3(st.fcx, st.fcy, st.fcz);
    case MFG_F_FT: return put3(st.ftx, st.fty, st.ftz);
    case MFG_F_POS_T: return put3(st.xt, st.yt, st.zt);
    case MFG_F_VEL_T: return put3(st.vxt, st.vyt, st.vzt);
    case MFG_F_S: return put_sym(st.S);
    case MFG_F_S_T: return put_sym(st.St);
    case MFG_F_H: return put1(st.h);
    case MFG_F_RHO: return put1(st.rho);
    case MFG_F_EPS_PL: return put1(st.eps_pl);
    case MFG_F_EPS_PL_DOT: return put1(st.eps_pl_dot);
    case MFG_F_T: return put1(st.T);
    case MFG_F_RHO_T: return put1(st.rho_t);
    case MFG_F_T_T: return put1(st.T_t);
    case MFG_F_FIXED: return put_u(st.flags, 0);
    case MFG_F_BLANKED: return put_u(st.flags, 1);
    case MFG_F_TOOL: return put_u(st.flags, 2);
    case MFG_F_ID: return put_u(st.id, 3);
    case MFG_F_HASH: return put_u(p->key_sorted, 3);
    case MFG_F_IDX: return put_u(p->order, 3);
    case MFG_F_S_DER:
    case MFG_F_V_DER: {
      const double *m = f == MFG_F_V_DER ? p->vder : p->sder;
      if (!m) return fail(MFG_ERR_STATE, "no velocity gradient / stress divergence yet (interactions_monaghan computes them)");
      const size_t n = padded(N);
      for (int k = 0; k < 9; k++) {
        if (!get(m + k * n)) return fail(MFG_ERR_HIP, "download failed");
        for (unsigned i = 0; i < N; i++) host[9 * (size_t)i + k] = a[i];
      }
      return MFG_OK;
    }
    case MFG_F_N: {
      if (!p->nrm) return fail(MFG_ERR_STATE, "no contact normals yet (tool compute_contact_force writes them)");
      const size_t n = padded(N);
      return put3(p->nrm, p->nrm + n, p->nrm + 2 * n);
    }
    case MFG_F_P:
    case MFG_F_R: {
      if (!p->p_dbg) {
        // pressure and artificial stress live in the pair records; keep copies from now on
        if (hipMalloc(&p->p_dbg, sizeof(double) * padded(N)) != hipSuccess ||
            hipMalloc(&p->R_dbg, sizeof(double) * 6 * padded(N)) != hipSuccess)
          return fail(MFG_ERR_NOMEM, "debug copies");
        (void)hipMemsetAsync(p->p_dbg, 0, sizeof(double) * padded(N), s);
        (void)hipMemsetAsync(p->R_dbg, 0, sizeof(double) * 6 * padded(N), s);
        return fail(MFG_ERR_STATE, "pressure / artificial stress are recorded from the next step on");
      }
      if (f == MFG_F_P) return put1(p->p_dbg);
      std::vector<double> r(6 * (size_t)N);
      if (hipMemcpyAsync(r.data(), p->R_dbg, sizeof(double) * 6 * N, hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess)
        return fail(MFG_ERR_HIP, "download failed");
      const int dst_of[6][2] = {{0, 0}, {4, 4}, {8, 8}, {1, 3}, {5, 7}, {2, 6}};
      for (unsigned i = 0; i < N; i++)
        for (int k = 0; k < 6; k++) {
          host[9 * (size_t)i + dst_of[k][0]] = r[6 * (size_t)i + k];
          host[9 * (size_t)i + dst_of[k][1]] = r[6 * (size_t)i + k];
        }
      return MFG_OK;
    }
    default:
      return fail(MFG_ERR_ARG, "unknown field %d", f);
  }
}

int mfg_particle_gpu_upload(mfg_particles *p, int f, const double *host) {
  NEED_LIB();
  if (!p || !host) return fail(MFG_ERR_ARG, "null argument");
  p->sorted_grid = nullptr;  // new state: the next step sorts
  const unsigned N = p->N_init;
  const State &st = p->st();
  hipStream_t s = g_lib.stream;
  std::vector<double> a(N);
  auto set = [&](double *dst) -> bool {
    return hipMemcpyAsync(dst, a.data(), sizeof(double) * N, hipMemcpyHostToDevice, s) == hipSuccess &&
           hipStreamSynchronize(s) == hipSuccess;
  };
  auto set3 = [&](double *x, double *y, double *z) -> int {
    double *dst[3] = {x, y, z};
    for (int k = 0; k < 3; k++) {
      for (unsigned i = 0; i < N; i++) a[i] = host[3 * (size_t)i + k];
      if (!set(dst[k])) return fail(MFG_ERR_HIP, "upload failed");
    }
    return MFG_OK;
  };
  auto set1 = [&](double *x) -> int {
    std::memcpy(a.data(), host, sizeof(double) * N);
    return set(x) ? MFG_OK : fail(MFG_ERR_HIP, "upload failed");
  };
  auto set_sym = [&](double *const *S6) -> int {
    const int src_of[6] = {0, 4, 8, 1, 5, 2};
    for (int k = 0; k < 6; k++) {
      for (unsigned i = 0; i < N; i++) a[i] = host[9 * (size_t)i + src_of[k]];
      if (!set(S6[k])) return fail(MFG_ERR_HIP, "upload failed");
    }
    return MFG_OK;
  };
  auto set_flag = [&](uint32_t bit_eq1, uint32_t bit_nz) -> int {
    std::vector<uint32_t> u(N);
    if (hipMemcpyAsync(u.data(), st.flags, sizeof(uint32_t) * N, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail(MFG_ERR_HIP, "flag download failed");
    for (unsigned i = 0; i < N; i++) {
      u[i] &= ~(bit_eq1 | bit_nz);
      if (host[i] == 1.) u[i] |= bit_eq1;
      if (host[i] != 0.) u[i] |= bit_nz;
    }
    if (hipMemcpyAsync(st.flags, u.data(), sizeof(uint32_t) * N, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail(MFG_ERR_HIP, "flag upload failed");
    return MFG_OK;
  };
  switch (f) {
    case MFG_F_POS: return set3(st.x, st.y, st.z);
    case MFG_F_VEL: return set3(st.vx, st.vy, st.vz);
    case MFG_F_VEL_BC: return set3(st.vbx, st.vby, st.vbz);
    case MFG_F_FC: return set3(st.fcx, st.fcy, st.fcz);
    case MFG_F_FT: return set3(st.ftx, st.fty, st.ftz);
    case MFG_F_POS_T: return set3(st.xt, st.yt, st.zt);
    case MFG_F_VEL_T: return set3(st.vxt, st.vyt, st.vzt);
    case MFG_F_S: return set_sym(st.S);
    case MFG_F_S_T: return set_sym(st.St);
    case MFG_F_H: return set1(st.h);
    case MFG_F_RHO: return set1(st.rho);
    case MFG_F_EPS_PL: return set1(st.eps_pl);
    case MFG_F_EPS_PL_DOT: return set1(st.eps_pl_dot);
    case MFG_F_T: return set1(st.T);
    case MFG_F_RHO_T: return set1(st.rho_t);
    case MFG_F_T_T: return set1(st.T_t);
    case MFG_F_FIXED: return set_flag(FL_FIXED_ONE, FL_FIXED);
    case MFG_F_BLANKED: return set_flag(FL_BLANKED, 0);
    case MFG_F_TOOL: return set_flag(FL_TOOL, FL_TOOL_NZ);
    default: return fail(MFG_ERR_ARG, "field %d cannot be uploaded", f);
  }
}

// ---- grid ----
int mfg_grid_create_locked(int kind, int num_part, const double bbmin[3], const double bbmax[3],
                           double h, mfg_grid **out) {
  NEED_LIB();
  if (!out || !bbmin || !bbmax || !(h > 0.) || num_part <= 0) return fail(MFG_ERR_ARG, "grid arguments");
  if (kind != MFG_GRID_GREEN && kind != MFG_GRID_ROTHLIN) return fail(MFG_ERR_ARG, "grid kind %d", kind);
  mfg_grid *g = new mfg_grid();
  g->kind = kind;
  // set_geometry (src/grid_gpu_base.cu:237-273)
  for (int k = 0; k < 3; k++) {
    g->bbmin[k] = g->bbmin_init[k] = bbmin[k] - 1e-6;
    g->bbmax[k] = g->bbmax_init[k] = bbmax[k] + 1e-6;
  }
  double kernel_width = 2.;
  g->dx = h * kernel_width / 2.;
  g->nx = (long)std::ceil((g->bbmax[0] - g->bbmin[0]) / g->dx);
  g->ny = (long)std::ceil((g->bbmax[1] - g->bbmin[1]) / g->dx);
  g->nz = (long)std::ceil((g->bbmax[2] - g->bbmin[2]) / g->dx);
  g->num_cell = g->nx * g->ny * g->nz;
  g->max_cell = g->num_cell;
  g->locked = true;
  if (g->num_cell <= 0 || g->num_cell >= INT_MAX) {
    delete g;
    return fail(MFG_ERR_ARG, "grid of %ld cells", g->num_cell);
  }
  int rc = grid_alloc(g);
  if (rc != MFG_OK) {
    delete g;
    return rc;
  }
  *out = g;
  return MFG_OK;
}

int mfg_grid_create_unlocked(int kind, long max_cell, int num_part, mfg_grid **out) {
  NEED_LIB();
  if (!out || max_cell <= 0 || max_cell >= INT_MAX || num_part <= 0) return fail(MFG_ERR_ARG, "grid arguments");
  mfg_grid *g = new mfg_grid();
  g->kind = kind;
  g->max_cell = max_cell;
  g->num_cell = max_cell;
  g->locked = false;
  int rc = grid_alloc(g);
  if (rc != MFG_OK) {
    delete g;
    return rc;
  }
  *out = g;
  return MFG_OK;
}

int mfg_grid_destroy(mfg_grid *g) {
  if (!g) return MFG_OK;
  if (g_lib.ready) (void)hipStreamSynchronize(g_lib.stream);
  void *ptrs[] = {g->count, g->fill, g->parts, g->bbox_part};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  delete g;
  return MFG_OK;
}

int mfg_grid_get_geometry(const mfg_grid *g, mfg_geom_constants *geom, long *num_cell, int *locked) {
  if (!g) return fail(MFG_ERR_ARG, "null grid");
  if (geom) {
    geom->nx = (int)g->nx; geom->ny = (int)g->ny; geom->nz = (int)g->nz;
    geom->bbmin_x = g->bbmin[0]; geom->bbmin_y = g->bbmin[1]; geom->bbmin_z = g->bbmin[2];
    geom->bbmax_x = g->bbmax[0]; geom->bbmax_y = g->bbmax[1]; geom->bbmax_z = g->bbmax[2];
    geom->dx = g->dx;
  }
  if (num_cell) *num_cell = g->num_cell;
  if (locked) *locked = g->locked ? 1 : 0;
  return MFG_OK;
}

int mfg_grid_set_bbox_vel(mfg_grid *g, const double v[3]) {
  if (!g || !v) return fail(MFG_ERR_ARG, "null argument");
  for (int k = 0; k < 3; k++) g->vel_bbox[k] = v[k];
  return MFG_OK;
}
int mfg_grid_set_max_fixed(mfg_grid *g, int f) {
  if (!g) return fail(MFG_ERR_ARG, "null grid");
  g->max_fixed = f != 0;
  return MFG_OK;
}
int mfg_grid_set_min_fixed(mfg_grid *g, int f) {
  if (!g) return fail(MFG_ERR_ARG, "null grid");
  g->min_fixed = f != 0;
  return MFG_OK;
}
int mfg_grid_set_hard_blank(mfg_grid *g, int hb) {
  if (!g) return fail(MFG_ERR_ARG, "null grid");
  g->hard_blank = hb != 0;
  return MFG_OK;
}

int mfg_grid_update_geometry(mfg_grid *g, mfg_particles *p, double kernel_width, double t) {
  NEED_LIB();
  if (!g || !p) return fail(MFG_ERR_ARG, "null argument");
  return grid_update_geometry(g, p, kernel_width, t);
}

int mfg_grid_assign_hashes(const mfg_grid *gc, mfg_particles *p) {
  NEED_LIB();
  mfg_grid *g = const_cast<mfg_grid *>(gc);
  if (!g || !p) return fail(MFG_ERR_ARG, "null argument");
  sort_keys(g, p, nullptr);
  return MFG_OK;
}

int mfg_grid_adapt_particle_number(const mfg_grid *g, mfg_particles *p) {
  NEED_LIB();
  if (!g || !p) return fail(MFG_ERR_ARG, "null argument");
  return check_counters(p);
}

int mfg_grid_sort(mfg_grid *g, mfg_particles *p) {
  NEED_LIB();
  if (!g || !p) return fail(MFG_ERR_ARG, "null argument");
  // the hash of mfg_grid_assign_hashes is used only if it was made for this set (its
  // permutation indexes this set's fields) and not yet applied
  if (!g->scanned || g->sorted_for != p) sort_keys(g, p, nullptr);
  const int rc = check_counters(p);
  if (rc != MFG_OK) return rc;
  Consts c = make_consts(0., p->precision);
  permute(g, p, c, 0);
  // the keys now describe the sorted order; a further sort hashes again
  g->scanned = false;
  return MFG_OK;
}

int mfg_grid_get_cells(mfg_grid *g, mfg_particles *p, int *d_start, int *d_end) {
  NEED_LIB();
  if (!g || !p || !d_start || !d_end) return fail(MFG_ERR_ARG, "null argument");
  launch_cells_to_tables(g_lib.stream, g->info(), g->count, g->kind == MFG_GRID_ROTHLIN, d_start, d_end);
  HIP_TRY(hipGetLastError());
  return MFG_OK;
}

int mfg_grid_download_cells(mfg_grid *g, int *h_start, int *h_end) {
  NEED_LIB();
  if (!g || !h_start || !h_end) return fail(MFG_ERR_ARG, "null argument");
  int *ds = nullptr, *de = nullptr;
  HIP_TRY(hipMalloc(&ds, sizeof(int) * g->num_cell));
  HIP_TRY(hipMalloc(&de, sizeof(int) * g->num_cell));
  launch_cells_to_tables(g_lib.stream, g->info(), g->count, g->kind == MFG_GRID_ROTHLIN, ds, de);
  hipError_t e1 = hipMemcpyAsync(h_start, ds, sizeof(int) * g->num_cell, hipMemcpyDeviceToHost, g_lib.stream);
  hipError_t e2 = hipMemcpyAsync(h_end, de, sizeof(int) * g->num_cell, hipMemcpyDeviceToHost, g_lib.stream);
  hipError_t e3 = hipStreamSynchronize(g_lib.stream);
  (void)hipFree(ds);
  (void)hipFree(de);
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) return fail(MFG_ERR_HIP, "cell download failed");
  return MFG_OK;
}

int mfg_grid_neighbours(mfg_grid *g, mfg_particles *p, int *h_counts, int *h_lists, int cap) {
  NEED_LIB();
  if (!g || !p || !h_counts) return fail(MFG_ERR_ARG, "null argument");
  // positions of the current (sorted) state as neighbour-pass records
  const State &st = p->st();
  Consts c = make_consts(0., p->precision);
  c.heat = MFG_HEAT_NONE;
  // build nl4 from the current state: a permute with the identity order would do; use the
  // predict path with dt = 0 on a copy of the order (half step == current state)
  GridBuffers gb = grid_buffers(g, p);
  std::vector<int> ident(p->N_init);
  for (unsigned i = 0; i < p->N_init; i++) ident[i] = (int)i;
  int *saved_order = nullptr;
  HIP_TRY(hipMalloc(&saved_order, sizeof(int) * p->N_init));
  HIP_TRY(hipMemcpyAsync(saved_order, p->order, sizeof(int) * p->N_init, hipMemcpyDeviceToDevice, g_lib.stream));
  int *saved_keys = nullptr;
  HIP_TRY(hipMalloc(&saved_keys, sizeof(int) * p->N_init));
  HIP_TRY(hipMemcpyAsync(saved_keys, p->key, sizeof(int) * p->N_init, hipMemcpyDeviceToDevice, g_lib.stream));
  HIP_TRY(hipMemcpyAsync(p->order, ident.data(), sizeof(int) * p->N_init, hipMemcpyHostToDevice, g_lib.stream));
  HIP_TRY(hipMemcpyAsync(p->key, p->key_sorted, sizeof(int) * p->N_init, hipMemcpyDeviceToDevice, g_lib.stream));
  HIP_TRY(hipStreamSynchronize(g_lib.stream));
  launch_permute_predict_ex(g_lib.stream, st, p->buf[1 - p->cur], p->N_init, gb, p->cnt, c, 1,
                            p->rec, p->nl4, p->pse2, p->cs4, nullptr, nullptr);
  p->cur = 1 - p->cur;
  HIP_TRY(hipMemcpyAsync(p->order, saved_order, sizeof(int) * p->N_init, hipMemcpyDeviceToDevice, g_lib.stream));
  HIP_TRY(hipMemcpyAsync(p->key, saved_keys, sizeof(int) * p->N_init, hipMemcpyDeviceToDevice, g_lib.stream));
  int rn = run_neighbours(g, p, c, 0);
  if (rn != MFG_OK) return rn;
  HIP_TRY(hipStreamSynchronize(g_lib.stream));
  (void)hipFree(saved_order);
  (void)hipFree(saved_keys);
  int rc = check_counters(p);
  unsigned N = p->N_host;
  std::vector<int> cnts(N);
  HIP_TRY(hipMemcpy(cnts.data(), p->count, sizeof(int) * N, hipMemcpyDeviceToHost));
  for (unsigned i = 0; i < N; i++) h_counts[i] = cnts[i];
  if (h_lists && cap > 0) {
    std::vector<int> nb((size_t)p->nl_stride * N);
    HIP_TRY(hipMemcpy(nb.data(), p->nbr, sizeof(int) * nb.size(), hipMemcpyDeviceToHost));
    for (unsigned i = 0; i < N; i++)
      for (int k = 0; k < std::min(cnts[i], cap); k++)
        h_lists[(size_t)i * cap + k] = nb[(size_t)i * p->nl_stride + k];
  }
  return rc;
}

// ---- tool ----
int mfg_tool_3d_gpu_create(const double *verts3, int n_verts, const int *tris3, int n_tris,
                           double h_avg, const double *vel, const mfg_phys_constants *phys,
                           double alpha, mfg_tool **out) {
  NEED_LIB();
  if (!out || !verts3 || !tris3 || n_verts < 3 || n_tris < 1) return fail(MFG_ERR_ARG, "tool mesh arguments");
  std::vector<Tri> tris;
  double havg = 0.;
  int rc = mesh_to_tris(verts3, n_verts, tris3, n_tris, tris, havg);
  if (rc != MFG_OK) return rc;
  if (h_avg <= 0.) h_avg = havg;
  mfg_tool *t = new mfg_tool();
  t->n_tris = n_tris;
  t->alpha = alpha;
  if (vel)
    for (int k = 0; k < 3; k++) t->vel[k] = vel[k];
  if (phys) t->mass = phys->mass;
  // triangle grid geometry (src/grid_cpu_tri.cpp:48-86)
  double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (const Tri &r : tris)
    for (int k = 0; k < 3; k++) {
      mn[k] = std::fmin(std::fmin(std::fmin(r.p[0][k], r.p[1][k]), r.p[2][k]), mn[k]);
      mx[k] = std::fmax(std::fmax(std::fmax(r.p[0][k], r.p[1][k]), r.p[2][k]), mx[k]);
    }
  for (int k = 0; k < 3; k++) {
    t->gbbmin[k] = mn[k] - 1e-4;
    t->gbbmax[k] = mx[k] + 1e-4;
  }
  t->gdx = h_avg;
  t->gnx = (int)std::ceil((t->gbbmax[0] - t->gbbmin[0]) / t->gdx);
  t->gny = (int)std::ceil((t->gbbmax[1] - t->gbbmin[1]) / t->gdx);
  t->gnz = (int)std::ceil((t->gbbmax[2] - t->gbbmin[2]) / t->gdx);
  hipStream_t s = g_lib.stream;
  auto bail = [&](int code) {
    void *ptrs[] = {t->tris, t->boxed, t->p_init_all, t->p_init_boxed, t->cells, t->boxed_idx,
                    t->cand_off, t->cand_idx};
    for (void *q : ptrs)
      if (q) (void)hipFree(q);
    delete t;
    return code;
  };
  if (hipMalloc(&t->tris, sizeof(Tri) * n_tris) != hipSuccess) return bail(fail(MFG_ERR_NOMEM, "tool triangles"));
  if (hipMemcpyAsync(t->tris, tris.data(), sizeof(Tri) * n_tris, hipMemcpyHostToDevice, s) != hipSuccess)
    return bail(fail(MFG_ERR_HIP, "tool upload"));
  int *idx = nullptr;
  rc = tool_bin_triangles(s, t->tris, n_tris, t->gbbmin, t->gdx, t->gnx, t->gny, t->gnz, &t->cells,
                          &idx, &t->n_boxed);
  if (rc != MFG_OK) return bail(fail(rc, "triangle binning failed"));
  t->boxed_idx = idx;
  if (hipMalloc(&t->boxed, sizeof(Tri) * std::max(t->n_boxed, 1)) != hipSuccess)
    return bail(fail(MFG_ERR_NOMEM, "boxed triangles"));
  launch_gather_tris(s, t->tris, idx, t->n_boxed, t->boxed);
  // orient the normals so that a point just outside the grid box tests outside
  int *d_in = nullptr;
  HIP_TRY(hipMalloc(&d_in, sizeof(int)));
  double q[3] = {t->gbbmax[0] + 1e-6, t->gbbmax[1] + 1e-6, t->gbbmax[2] + 1e-6};
  int inside = 0;
  if (t->n_boxed > 0) {
    launch_orient_test(s, t->boxed, t->cells, t->gnx, t->gny, t->gnz, t->gbbmin, t->gdx, q, d_in);
    HIP_TRY(hipMemcpyAsync(&inside, d_in, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  if (inside) {
    launch_flip_normals(s, t->tris, n_tris);
    launch_flip_normals(s, t->boxed, t->n_boxed);
  }
  (void)hipFree(d_in);
  // initial vertex positions of both triangle sets (update_tool moves from these)
  std::vector<Tri> hb(std::max(t->n_boxed, 1));
  HIP_TRY(hipMemcpyAsync(hb.data(), t->boxed, sizeof(Tri) * t->n_boxed, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  std::vector<double> pa(9 * (size_t)n_tris), pb(9 * (size_t)std::max(t->n_boxed, 1));
  for (int k = 0; k < n_tris; k++)
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) pa[9 * k + 3 * a + b] = tris[k].p[a][b];
  for (int k = 0; k < t->n_boxed; k++)
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) pb[9 * k + 3 * a + b] = hb[k].p[a][b];
  if (hipMalloc(&t->p_init_all, sizeof(double) * pa.size()) != hipSuccess ||
      hipMalloc(&t->p_init_boxed, sizeof(double) * pb.size()) != hipSuccess)
    return bail(fail(MFG_ERR_NOMEM, "tool init positions"));
  HIP_TRY(hipMemcpyAsync(t->p_init_all, pa.data(), sizeof(double) * pa.size(), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(t->p_init_boxed, pb.data(), sizeof(double) * pb.size(), hipMemcpyHostToDevice, s));
  HIP_TRY(hipStreamSynchronize(s));
  for (int k = 0; k < 3; k++) {
    t->bbmin[k] = t->gbbmin[k];
    t->bbmax[k] = t->gbbmax[k];
  }
  *out = t;
  return MFG_OK;
}

int mfg_tool_destroy(mfg_tool *t) {
  if (!t) return MFG_OK;
  if (g_lib.ready) (void)hipStreamSynchronize(g_lib.stream);
  void *ptrs[] = {t->tris, t->boxed, t->p_init_all, t->p_init_boxed, t->cells, t->boxed_idx,
                  t->cand_off, t->cand_idx};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  delete t;
  return MFG_OK;
}

int mfg_tool_set_vel(mfg_tool *t, const double v[3]) {
  if (!t || !v) return fail(MFG_ERR_ARG, "null argument");
  for (int k = 0; k < 3; k++) t->vel[k] = t->vel_const[k] = v[k];
  return MFG_OK;
}
int mfg_tool_set_mu(mfg_tool *t, double mu) {
  if (!t || !(mu > 0.)) return fail(MFG_ERR_ARG, "friction coefficient must be positive");
  t->mu = mu;
  return MFG_OK;
}
int mfg_tool_set_contact_alpha(mfg_tool *t, double a) {
  if (!t || !(a > 0.)) return fail(MFG_ERR_ARG, "contact stiffness must be positive");
  t->alpha = a;
  return MFG_OK;
}
int mfg_tool_set_algorithm_type(mfg_tool *t, int algo) {
  if (!t || (algo != MFG_CONTACT_EXHAUSTIVE && algo != MFG_CONTACT_HASHED)) return fail(MFG_ERR_ARG, "algorithm %d", algo);
  t->algo = algo;
  if (algo == MFG_CONTACT_EXHAUSTIVE && !t->cand_off) {
    // candidate lists are built from the untranslated mesh, in the grid's initial frame
    NEED_LIB();
    hipStream_t s = g_lib.stream;
    Tri *tmp = nullptr;
    HIP_TRY(hipMalloc(&tmp, sizeof(Tri) * t->n_tris));
    HIP_TRY(hipMemcpyAsync(tmp, t->tris, sizeof(Tri) * t->n_tris, hipMemcpyDeviceToDevice, s));
    launch_tool_update(s, tmp, t->p_init_all, t->n_tris, 0., t->vel);
    int rc = tool_candidate_lists(s, tmp, t->n_tris, t->gbbmin, t->gdx, t->gnx, t->gny, t->gnz,
                                  &t->cand_off, &t->cand_idx, &t->n_cand);
    (void)hipFree(tmp);
    if (rc != MFG_OK) return fail(rc, "closest-triangle candidate lists");
  }
  return MFG_OK;
}
int mfg_tool_set_birth_death(mfg_tool *t, double b, double d) {
  if (!t) return fail(MFG_ERR_ARG, "null tool");
  t->birth = b;
  t->death = d;
  return MFG_OK;
}
int mfg_tool_set_thermal(mfg_tool *t, int th) {
  if (!t) return fail(MFG_ERR_ARG, "null tool");
  t->thermal = th != 0;
  return MFG_OK;
}
int mfg_tool_set_physics(mfg_tool *t, const mfg_phys_constants *ph) {
  if (!t || !ph) return fail(MFG_ERR_ARG, "null argument");
  t->mass = ph->mass;
  return MFG_OK;
}
int mfg_tool_get_vel(const mfg_tool *t, double v[3]) {
  if (!t || !v) return fail(MFG_ERR_ARG, "null argument");
  for (int k = 0; k < 3; k++) v[k] = t->vel[k];
  return MFG_OK;
}
int mfg_tool_get_bbox(const mfg_tool *t, double mn[3], double mx[3]) {
  if (!t || !mn || !mx) return fail(MFG_ERR_ARG, "null argument");
  for (int k = 0; k < 3; k++) {
    mn[k] = t->bbmin[k];
    mx[k] = t->bbmax[k];
  }
  return MFG_OK;
}
int mfg_tool_get_motion(const mfg_tool *t, double motion7[7]) {
  if (!t || !motion7) return fail(MFG_ERR_ARG, "null argument");
  motion7[0] = t->mesh_tc;
  for (int k = 0; k < 3; k++) {
    motion7[1 + k] = t->bbmin[k];
    motion7[4 + k] = t->bbmax[k];
  }
  return MFG_OK;
}
int mfg_tool_set_motion(mfg_tool *t, const double motion7[7]) {
  NEED_LIB();
  if (!t || !motion7) return fail(MFG_ERR_ARG, "null argument");
  t->mesh_tc = motion7[0];
  for (int k = 0; k < 3; k++) {
    t->bbmin[k] = motion7[1 + k];
    t->bbmax[k] = motion7[4 + k];
  }
  // the triangles follow at the next step (mfg_tool_update moves them to that step's time)
  return MFG_OK;
}
int mfg_tool_get_mu(const mfg_tool *t, double *mu) {
  if (!t || !mu) return fail(MFG_ERR_ARG, "null argument");
  *mu = t->mu;
  return MFG_OK;
}
int mfg_tool_is_alive(const mfg_tool *t, double tc) {
  if (!t) return fail(MFG_ERR_ARG, "null tool");
  return (tc >= t->birth && tc <= t->death) ? 1 : 0;
}
int mfg_tool_get_grid(const mfg_tool *t, long dims[5], double bb[7]) {
  if (!t || !dims || !bb) return fail(MFG_ERR_ARG, "null argument");
  dims[0] = t->gnx; dims[1] = t->gny; dims[2] = t->gnz; dims[3] = t->n_boxed; dims[4] = t->n_tris;
  for (int k = 0; k < 3; k++) {
    bb[k] = t->gbbmin[k];
    bb[3 + k] = t->gbbmax[k];
  }
  bb[6] = t->gdx;
  return MFG_OK;
}
int mfg_tool_download_cells(const mfg_tool *t, unsigned *cells, double *boxed30) {
  NEED_LIB();
  if (!t) return fail(MFG_ERR_ARG, "null tool");
  long nc = (long)t->gnx * t->gny * t->gnz;
  if (cells) HIP_TRY(hipMemcpy(cells, t->cells, sizeof(int) * (nc + 1), hipMemcpyDeviceToHost));
  if (boxed30 && t->n_boxed > 0) {
    std::vector<Tri> hb(t->n_boxed);
    HIP_TRY(hipMemcpy(hb.data(), t->boxed, sizeof(Tri) * t->n_boxed, hipMemcpyDeviceToHost));
    for (int k = 0; k < t->n_boxed; k++) {
      const Tri &r = hb[k];
      double *o = boxed30 + 30 * (size_t)k;
      for (int q = 0; q < 3; q++)
        for (int a = 0; a < 3; a++) {
          o[3 * q + a] = r.p[q][a];
          o[12 + 3 * q + a] = r.np[q][a];
          o[21 + 3 * q + a] = r.ne[q][a];
        }
      for (int a = 0; a < 3; a++) o[9 + a] = r.n[a];
    }
  }
  return MFG_OK;
}
int mfg_tool_update(mfg_tool *t, double tc, double dt) {
  NEED_LIB();
  if (!t) return fail(MFG_ERR_ARG, "null tool");
  if (t->n_boxed == 0) return MFG_OK;
  ProfScope ps("tool_update");
  if (t->algo == MFG_CONTACT_HASHED) launch_tool_update(g_lib.stream, t->boxed, t->p_init_boxed, t->n_boxed, tc, t->vel);
  else {
    // the exhaustive search runs over the boxed copies too (grid-accelerated, same result);
    // both arrays move in one launch
    launch_tool_update(g_lib.stream, t->tris, t->p_init_all, t->n_tris, tc, t->vel, t->boxed,
                       t->p_init_boxed, t->n_boxed);
  }
  t->mesh_tc = tc;
  for (int k = 0; k < 3; k++) {
    t->bbmin[k] += dt * t->vel[k];
    t->bbmax[k] += dt * t->vel[k];
  }
  return MFG_OK;
}

// ---- tool forces ----
int mfg_tool_forces_create(unsigned n, mfg_tool_forces **out) {
  NEED_LIB();
  if (!out || n == 0 || n > MAX_TOOLS) return fail(MFG_ERR_ARG, "tool forces for %u tools (max %d)", n, MAX_TOOLS);
  mfg_tool_forces *f = new mfg_tool_forces();
  f->n = n;
  if (hipMalloc(&f->d, sizeof(float) * 3 * n) != hipSuccess) {
    delete f;
    return fail(MFG_ERR_NOMEM, "tool forces");
  }
  (void)hipMemsetAsync(f->d, 0, sizeof(float) * 3 * n, g_lib.stream);
  *out = f;
  return MFG_OK;
}
int mfg_tool_forces_destroy(mfg_tool_forces *f) {
  if (!f) return MFG_OK;
  if (f->d) (void)hipFree(f->d);
  delete f;
  return MFG_OK;
}
int mfg_tool_forces_reset(mfg_tool_forces *f) {
  NEED_LIB();
  if (!f) return fail(MFG_ERR_ARG, "null forces");
  HIP_TRY(hipMemsetAsync(f->d, 0, sizeof(float) * 3 * f->n, g_lib.stream));
  return MFG_OK;
}
int mfg_tool_forces_get_host(mfg_tool_forces *f, float *xyz) {
  NEED_LIB();
  if (!f || !xyz) return fail(MFG_ERR_ARG, "null argument");
  HIP_TRY(hipMemcpyAsync(xyz, f->d, sizeof(float) * 3 * f->n, hipMemcpyDeviceToHost, g_lib.stream));
  HIP_TRY(hipStreamSynchronize(g_lib.stream));
  return MFG_OK;
}

// ---- actions ----
int mfg_perform_blanking(mfg_particles *p, const mfg_blanking_state *b) {
  NEED_LIB();
  if (!p || !b) return fail(MFG_ERR_ARG, "null argument");
  hipLaunchKernelGGL(k_blanking_only, dim3((p->N_init + 255) / 256), dim3(256), 0, g_lib.stream,
                     p->st(), p->N_init, b->bbmin[0], b->bbmin[1], b->bbmin[2], b->bbmax[0],
                     b->bbmax[1], b->bbmax[2], b->max_vel_squared);
  HIP_TRY(hipGetLastError());
  return MFG_OK;
}

int mfg_debug_check_valid(mfg_particles *p) {
  NEED_LIB();
  if (!p) return fail(MFG_ERR_ARG, "null particles");
  int rc = valid_check(g_lib.stream, p->st(), p->N_host, p->flag_d);
  if (rc == MFG_ERR_NONFINITE) return fail(rc, "found particle with nan values that is not blanked");
  if (rc != MFG_OK) return fail(rc, "valid check failed");
  return MFG_OK;
}

// ---- leap frog ----
int mfg_leap_frog_create(unsigned num_part, unsigned num_cell, mfg_leap_frog **out) {
  if (!out) return fail(MFG_ERR_ARG, "null output");
  mfg_leap_frog *lf = new mfg_leap_frog();
  lf->num_part = num_part;
  lf->num_cell = num_cell;
  *out = lf;
  return MFG_OK;
}
int mfg_leap_frog_destroy(mfg_leap_frog *lf) {
  delete lf;
  return MFG_OK;
}

}  // extern "C"

namespace {

// Work memory of one step beyond the particle set, grid and tools.  A null `base` means the
// step's own buffers, sized on first use (hipMalloc); otherwise the regions are carved out of
// [base, base + bytes) and the step allocates nothing.
struct StepWork {
  unsigned char *base = nullptr;
  size_t bytes = 0;
};

// suspends the per-kernel event timing for the scope (its events are created on demand)
struct ProfOff {
  int saved;
  ProfOff(bool on) : saved(g_lib.prof) {
    if (on) g_lib.prof = 0;
  }
  ~ProfOff() { g_lib.prof = saved; }
};

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// what the contact kernels read of a tool at time t_current (after mfg_tool_update); the
// closest triangle is searched per particle (no search list)
void fill_tool_view(const mfg_tool *t, double t_current, ToolView &v) {
  v.algo = t->algo;
  v.tris = t->algo == MFG_CONTACT_HASHED ? t->boxed : t->tris;
  v.n_tris = t->algo == MFG_CONTACT_HASHED ? t->n_boxed : t->n_tris;
  v.cells = t->cells;
  v.cand_off = t->algo == MFG_CONTACT_EXHAUSTIVE ? t->cand_off : nullptr;
  v.cand_idx = t->cand_idx;
  v.slot = nullptr;
  v.res = nullptr;
  for (int a = 0; a < 3; a++) v.corg[a] = t->gbbmin[a] + t->vel[a] * t->mesh_tc;
  v.nx = t->gnx; v.ny = t->gny; v.nz = t->gnz;
  v.active = (t->n_boxed > 0 && t_current >= t->birth && t_current <= t->death) ? 1 : 0;
  for (int a = 0; a < 3; a++) {
    v.bbmin[a] = t->bbmin[a];
    v.bbmax[a] = t->bbmax[a];
    v.vel_const[a] = t->vel_const[a];
  }
  v.dx = t->gdx;
  v.mu = t->mu;
  v.alpha = t->alpha;
  v.mass = t->mass;
}

int step_impl(mfg_leap_frog *lf, mfg_particles *p, mfg_grid *g, int record,
              mfg_tool *const *tools, int n_tools, const mfg_blanking_state *blanking,
              mfg_tool_forces *forces, double dt, double t_current, const StepWork &work) {
  NEED_LIB();
  if (!lf || !p || !g) return fail(MFG_ERR_ARG, "null argument");
  if (n_tools < 0 || n_tools > MAX_TOOLS || (n_tools > 0 && !tools))
    return fail(MFG_ERR_ARG, "%d tools (max %d)", n_tools, MAX_TOOLS);
  if (record && (!forces || forces->n < (unsigned)n_tools))
    return fail(MFG_ERR_ARG, "recording forces needs a tool_forces of %d tools", n_tools);
  for (int k = 0; k < n_tools; k++)
    if (!tools[k]) return fail(MFG_ERR_ARG, "tool %d is null", k);
  const bool fixed_work = work.base != nullptr;
  ProfOff prof_off(fixed_work);
  hipStream_t s = g_lib.stream;
  Consts c = make_consts(dt, p->precision);
  // spatial sorting
  int rc = grid_update_geometry(g, p, 2.0, t_current);
  if (rc != MFG_OK) return rc;
  const GridInfo gi = g->info();
  // visiting order of the list pair pass: strips of 8 grid columns for 192-byte FP64 records,
  // 16 for 112-byte FP32 records (measured best on the 1M single-grain cut); negative: the
  // sorted order.  Automatic: the sorted order when its window (the records of about five
  // x-rows) fits in half an XCD's 4 MB L2, where strips gain nothing and cost two launches.
  const int tiled = g_lib.opt.pair_pass == MFG_PAIR_TILED ? 1 : 0;
  const double rec_b = p->precision == MFG_FP64 ? 192. : 112.;
  const bool small = 5. * rec_b * p->N_init / (gi.nx > 0 ? gi.nx : 1) <= 2. * 1024 * 1024;
  const int strip = g_lib.opt.pair_strip > 0    ? g_lib.opt.pair_strip
                    : g_lib.opt.pair_strip < 0  ? 0
                    : small                     ? 0
                    : p->precision == MFG_FP64 ? 8
                                               : 16;
  const int nseg = (!tiled && strip > 0) ? pair_order_segments(gi, strip) : 0;
  int n_exh = 0;
  for (int k = 0; k < n_tools; k++) {
    const mfg_tool *t = tools[k];
    n_exh += (t->algo == MFG_CONTACT_EXHAUSTIVE && t->n_boxed > 0 && t_current >= t->birth &&
              t_current <= t->death);
  }
  const bool search = n_exh > 0 && g_lib.opt_contact_search;
  // work memory: the caller's region, or the step's own buffers
  int *seg_pos = nullptr;
  ContactSearch cs{};
  if (fixed_work) {
    // segments | list counter | slots (n_init per tool) | list + results (cap entries).  The
    // list may hold fewer pairs than are in tool boxes: k_contact_bc searches the others per
    // thread, which finds the same closest triangle.
    size_t off = 0;
    const size_t seg_b = align256(sizeof(int) * (size_t)nseg);
    const size_t pairs = (size_t)p->N_init * (size_t)n_tools;
    const size_t slot_b = search ? align256(sizeof(int) * pairs) : 0;
    const size_t entry_b = sizeof(int2) + 6 * sizeof(double);
    const size_t min_cap = search ? std::min<size_t>(pairs, 4096) : 0;
    const size_t need = seg_b + (search ? 256 + slot_b + align256(entry_b * min_cap) : 0);
    if (need > work.bytes)
      return fail(MFG_ERR_ARG,
                  "allocation-free step: %u particles x %d tools and %d visiting-order segments "
                  "need %zu bytes of work memory, more than the %zu the library holds",
                  p->N_init, n_tools, nseg, need, work.bytes);
    if (nseg > 0) seg_pos = (int *)(work.base + off);
    off += seg_b;
    if (search) {
      cs.list_n = (unsigned *)(work.base + off);
      off += 256;
      cs.slot = (int *)(work.base + off);
      off += slot_b;
      const size_t cap = std::min<size_t>(std::min<size_t>(pairs, (size_t)1 << 22),
                                          (work.bytes - off - 256) / entry_b);
      cs.cap = (unsigned)cap;
      cs.list = (int2 *)(work.base + off);
      off += align256(sizeof(int2) * cap);
      cs.res = (double *)(work.base + off);
    }
  } else {
    if (nseg > p->seg_cap) {
      if (p->seg_pos) (void)hipFree(p->seg_pos);
      p->seg_pos = nullptr;
      p->seg_cap = 0;
      if (hipMalloc(&p->seg_pos, sizeof(int) * (size_t)nseg) != hipSuccess)
        return fail(MFG_ERR_NOMEM, "pair-pass order (%d segments)", nseg);
      p->seg_cap = nseg;
    }
    seg_pos = p->seg_pos;
    if (search) {
      rc = contact_search_buffers(p, n_tools, cs);
      if (rc != MFG_OK) return rc;
    }
  }
  if (keep_last_sort(g, p)) {
    p->since_sort++;
    permute(g, p, c, 1, true);
  } else {
    sort_keys(g, p, blanking);
    permute(g, p, c, 1);
  }
  g->scanned = false;  // the permutation is applied: a later mfg_grid_sort hashes again
  // pair pass
  const bool pse = c.heat == MFG_HEAT_PSE && c.pse_on;
  {
    ProfScope ps("neighbours");
    rc = run_neighbours(g, p, c, pse ? 1 : 0, !fixed_work);
    if (rc != MFG_OK) return rc;
  }
  {
    const int *trav = nullptr;
    if (nseg > 0) {
      ProfScope ps("pair_order");
      launch_pair_order(s, p->N_init, gi, g->count, p->key_sorted, p->cnt, strip, seg_pos, p->trav);
      trav = p->trav;
    }
    ProfScope ps(p->precision == MFG_FP64          ? "pair_fp64"
                 : p->precision == MFG_FP32_KAHAN ? "pair_fp32_kahan"
                                                  : "pair_fp32_wide");
    launch_pair(s, p->N_init, p->st(), p->rec, p->nbr, p->nl_stride, p->count, c, p->cnt, gi,
                g->count, p->key_sorted, tiled, p->sums, p->sum_stride, trav, p->cs4,
                g_lib.opt.pair_block);
  }
  {
    ProfScope ps("particle_update");
    launch_particle_update(s, p->N_init, p->st(), p->sums, p->sum_stride, p->poison, c, p->cnt);
  }
  // tools
  ToolSet ts;
  std::memset(&ts, 0, sizeof(ts));
  ts.n = n_tools;
  ts.record = record ? 1 : 0;
  if (record) launch_zero_f32(s, forces->d, 3 * forces->n);
  for (int k = 0; k < n_tools; k++) {
    mfg_tool *t = tools[k];
    rc = mfg_tool_update(t, t_current, dt);
    if (rc != MFG_OK) return rc;
    fill_tool_view(t, t_current, ts.t[k]);
  }
  if (n_tools > 0)
    for (int a = 0; a < 3; a++) ts.fric_vel[a] = tools[0]->vel[a];
  if (search) {
    ProfScope ps("contact_search");
    for (int k = 0; k < n_tools; k++)
      if (ts.t[k].algo == MFG_CONTACT_EXHAUSTIVE) {
        ts.t[k].slot = cs.slot + (size_t)k * p->N_init;
        ts.t[k].res = cs.res;
      }
    launch_contact_search(s, p->st(), p->N_init, ts, p->cnt, cs);
  }
  {
    ProfScope ps("contact_bc");
    launch_contact_bc(s, p->st(), p->N_init, ts, c, p->cnt, p->partials, p->contact_blocks);
  }
  if (record && n_tools > 0) launch_force_reduce(s, p->partials, p->contact_blocks, n_tools, forces->d);
  HIP_TRY(hipGetLastError());
  return MFG_OK;
}

}  // namespace

extern "C" {

int mfg_leap_frog_step(mfg_leap_frog *lf, mfg_particles *p, mfg_grid *g, int record,
                       mfg_tool *const *tools, int n_tools, const mfg_blanking_state *blanking,
                       mfg_tool_forces *forces, double dt, double t_current) {
  return step_impl(lf, p, g, record, tools, n_tools, blanking, forces, dt, t_current, StepWork{});
}

// ---- per-function entry points: the reference's actions, pair passes, one tool's contact and
// the leap-frog phases, each one launch over the particle set's current (sorted) state ----
static int run_action(mfg_particles *p, int which, double dt, const double *vel) {
  NEED_LIB();
  if (!p) return fail(MFG_ERR_ARG, "null particles");
  int rc = ensure_aux(p);
  if (rc != MFG_OK) return rc;
  Consts c = make_consts(dt, p->precision);
  // the reference's host-side switches (src/actions_gpu.cu:594-722)
  if (which == ACT_JOHNSON_COOK && !c.plastic) return MFG_OK;
  if (which == ACT_BC_THERMAL && !c.a_thermal) return MFG_OK;
  if (which == ACT_FRIC_HEAT && !c.a_fric) return MFG_OK;
  if (which == ACT_ART_STRESS && c.a_stresseps == 0.) return MFG_OK;
  ProfScope ps("action");
  launch_action(g_lib.stream, which, p->st(), p->N_init, p->cnt, c, aux_of(p), vel);
  HIP_TRY(hipGetLastError());
  return MFG_OK;
}

int mfg_actions_material_eos(mfg_particles *p) { return run_action(p, ACT_EOS, 0., nullptr); }
int mfg_actions_corrector_artificial_stress(mfg_particles *p) {
  return run_action(p, ACT_ART_STRESS, 0., nullptr);
}
int mfg_actions_material_stress_rate_jaumann(mfg_particles *p) {
  return run_action(p, ACT_JAUMANN, 0., nullptr);
}
int mfg_actions_contmech_continuity(mfg_particles *p) {
  return run_action(p, ACT_CONTINUITY, 0., nullptr);
}
int mfg_actions_contmech_momentum(mfg_particles *p) {
  return run_action(p, ACT_MOMENTUM, 0., nullptr);
}
int mfg_actions_contmech_advection(mfg_particles *p) {
  return run_action(p, ACT_ADVECTION, 0., nullptr);
}
int mfg_actions_plasticity_johnson_cook(mfg_particles *p, double dt) {
  if (!(dt > 0.)) return fail(MFG_ERR_ARG, "time step %g", dt);
  return run_action(p, ACT_JOHNSON_COOK, dt, nullptr);
}
int mfg_actions_material_fric_heat_gen(mfg_particles *p, const double vel[3], double dt) {
  if (!vel) return fail(MFG_ERR_ARG, "null tool velocity");
  return run_action(p, ACT_FRIC_HEAT, dt, vel);
}
int mfg_actions_perform_boundary_conditions(mfg_particles *p) {
  return run_action(p, ACT_BC, 0., nullptr);
}
int mfg_actions_perform_boundary_conditions_thermal(mfg_particles *p) {
  return run_action(p, ACT_BC_THERMAL, 0., nullptr);
}
int mfg_actions_move_tool_particles(mfg_particles *p, const mfg_tool *t, double dt) {
  if (!t) return fail(MFG_ERR_ARG, "null tool");
  return run_action(p, ACT_MOVE_TOOL, dt, t->vel);
}
int mfg_debug_invalidate(mfg_particles *p) { return run_action(p, ACT_INVALIDATE, 0., nullptr); }

int mfg_leap_frog_init(mfg_leap_frog *lf, mfg_particles *p) {
  if (!lf) return fail(MFG_ERR_ARG, "null leap_frog");
  return run_action(p, ACT_LF_INIT, 0., nullptr);
}
int mfg_leap_frog_advance(mfg_leap_frog *lf, mfg_particles *p, double f) {
  NEED_LIB();
  if (!lf || !p) return fail(MFG_ERR_ARG, "null argument");
  int rc = ensure_aux(p);
  if (rc != MFG_OK) return rc;
  launch_lf_advance(g_lib.stream, p->st(), p->N_init, aux_of(p), f);
  HIP_TRY(hipGetLastError());
  return MFG_OK;
}
int mfg_leap_frog_reset_contact_forces(mfg_leap_frog *lf, mfg_particles *p) {
  if (!lf) return fail(MFG_ERR_ARG, "null leap_frog");
  return run_action(p, ACT_RESET_CONTACT, 0., nullptr);
}

// the pair records of the current state; the grid's cell tables must describe its order
static int interaction_records(mfg_particles *p, mfg_grid *g, Consts &c) {
  NEED_LIB();
  if (!p || !g) return fail(MFG_ERR_ARG, "null argument");
  int rc = ensure_aux(p);
  if (rc != MFG_OK) return rc;
  c = make_consts(0., p->precision);
  launch_act_records(g_lib.stream, p->st(), p->N_init, p->cnt, c, aux_of(p), p->rec, p->nl4,
                     p->pse2, p->cs4);
  HIP_TRY(hipGetLastError());
  return MFG_OK;
}

int mfg_interactions_monaghan(mfg_particles *p, mfg_grid *g) {
  Consts c;
  int rc = interaction_records(p, g, c);
  if (rc != MFG_OK) return rc;
  rc = run_neighbours(g, p, c, 0);
  if (rc != MFG_OK) return rc;
  const GridInfo gi = g->info();
  launch_pair(g_lib.stream, p->N_init, p->st(), p->rec, p->nbr, p->nl_stride, p->count, c, p->cnt,
              gi, g->count, p->key_sorted, 0, p->sums, p->sum_stride, nullptr, p->cs4,
              g_lib.opt.pair_block);
  launch_act_sums(g_lib.stream, p->st(), p->N_init, p->cnt, c, aux_of(p), p->sums, p->sum_stride,
                  p->poison);
  HIP_TRY(hipGetLastError());
  return MFG_OK;
}

int mfg_interactions_heat_pse(mfg_particles *p, mfg_grid *g) {
  Consts c;
  int rc = interaction_records(p, g, c);
  if (rc != MFG_OK) return rc;
  if (!c.pse_on) return MFG_OK;  // no workpiece conduction constants (src/interactions_gpu.cu:536)
  return run_neighbours(g, p, c, 1);
}

int mfg_tool_compute_contact_force(mfg_tool *t, mfg_particles *p, int record,
                                   mfg_tool_forces *forces, int slot, double dt,
                                   double t_current) {
  NEED_LIB();
  if (!t || !p) return fail(MFG_ERR_ARG, "null argument");
  if (record && (!forces || slot < 0 || (unsigned)slot >= forces->n))
    return fail(MFG_ERR_ARG, "recording needs a force record with a slot for this tool");
  int rc = ensure_aux(p);
  if (rc != MFG_OK) return rc;
  ToolView v;
  std::memset(&v, 0, sizeof(v));
  fill_tool_view(t, t_current, v);
  hipStream_t s = g_lib.stream;
  launch_act_contact(s, p->st(), p->N_init, v, dt, p->cnt, aux_of(p), record ? 1 : 0, p->partials,
                     p->contact_blocks);
  if (record) launch_force_reduce(s, p->partials, p->contact_blocks, 1, forces->d + 3 * slot);
  HIP_TRY(hipGetLastError());
  return MFG_OK;
}

int mfg_leap_frog_step_noalloc(mfg_leap_frog *lf, mfg_particles *p, mfg_grid *g, int record,
                               mfg_tool *const *tools, int n_tools,
                               const mfg_blanking_state *blanking, mfg_tool_forces *forces,
                               double dt, double t_current) {
  NEED_LIB();
  if (!g_scratch) return fail(MFG_ERR_STATE, "the step's work memory has no device address");
  StepWork w;
  w.base = g_scratch;
  w.bytes = MFG_STEP_SCRATCH_BYTES;
  return step_impl(lf, p, g, record, tools, n_tools, blanking, forces, dt, t_current, w);
}

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// batched steps: many instances, every kernel of the step launched once over all of them
// ---------------------------------------------------------------------------------------------
struct mfg_batch {
  struct Member {
    mfg_particles *p;
    mfg_grid *g;
    std::vector<mfg_tool *> tools;
    mfg_tool_forces *forces;
    int blanking;
    double bb[6], vmax2;
    Consts c;
    double t;    // the instance's clock; the device keeps the same sequence (t += dt per step)
    int search;  // lane-parallel exhaustive contact search
  };
  std::vector<Member> members;
  int precision = -1, pse = -1;
  int cell_tile = 0;   // neighbour pass cells per wavefront; 0: the library option, else automatic
  bool dirty = true;   // the device descriptors and maps must be rebuilt before the next step
  int table_record = -1;
  Inst *d_inst = nullptr;
  size_t inst_cap = 0;
  unsigned *d_maps = nullptr;
  size_t maps_cap = 0;
  BatchLaunch L{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t graph_stream = nullptr;
  int graph_record = -1;
};

namespace {

constexpr int SCAN_TILE_CELLS = 256 * 16;  // cells per scan tile of the counting sort
constexpr unsigned BATCH_SEARCH_BLOCKS = 64;

unsigned round8(unsigned n) { return (n + 7u) / 8u * 8u; }

void batch_drop_graph(mfg_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_record = -1;
}

// descriptor of member m as its objects hold it now
Inst batch_inst(const mfg_batch::Member &m, int record) {
  Inst d;
  std::memset(&d, 0, sizeof(d));
  mfg_particles *p = m.p;
  d.buf[0] = p->buf[0];
  d.buf[1] = p->buf[1];
  d.cur = p->cur;
  d.n_init = p->N_init;
  d.g = m.g->info();
  d.gb = grid_buffers(m.g, p);
  d.nparts = (int)((m.g->num_cell + SCAN_TILE_CELLS - 1) / SCAN_TILE_CELLS);
  d.blanking = m.blanking;
  for (int k = 0; k < 6; k++) d.bb[k] = m.bb[k];
  d.vmax2 = m.vmax2;
  d.c = m.c;
  d.rec = p->rec;
  d.nl4 = p->nl4;
  d.pse2 = p->pse2;
  d.cs4 = p->cs4;
  d.nbr = p->nbr;
  d.count = p->count;
  d.poison = p->poison;
  d.nl_stride = p->nl_stride;
  d.contact_blocks = p->contact_blocks;
  d.sums = p->sums;
  d.sum_stride = p->sum_stride;
  d.cnt = p->cnt;
  d.partials = p->partials;
  d.forces = m.forces ? m.forces->d : nullptr;
  d.n_forces = m.forces ? (int)m.forces->n : 0;
  d.search = m.search;
  if (m.search) {
    d.cs.list = p->cs_list;
    d.cs.list_n = p->cs_n;
    d.cs.cap = p->cs_cap;
    d.cs.slot = p->cs_slot;
    d.cs.res = p->cs_res;
  }
  const int n_tools = (int)m.tools.size();
  d.ts.n = n_tools;
  d.ts.record = record ? 1 : 0;
  if (n_tools > 0)
    for (int a = 0; a < 3; a++) d.ts.fric_vel[a] = m.tools[0]->vel[a];
  for (int k = 0; k < n_tools; k++) {
    const mfg_tool *t = m.tools[k];
    ToolView &v = d.ts.t[k];
    v.algo = t->algo;
    v.tris = t->algo == MFG_CONTACT_HASHED ? t->boxed : t->tris;
    v.n_tris = t->algo == MFG_CONTACT_HASHED ? t->n_boxed : t->n_tris;
    v.cells = t->cells;
    v.cand_off = t->algo == MFG_CONTACT_EXHAUSTIVE ? t->cand_off : nullptr;
    v.cand_idx = t->cand_idx;
    if (m.search && t->algo == MFG_CONTACT_EXHAUSTIVE) {
      v.slot = p->cs_slot + (size_t)k * p->N_init;
      v.res = p->cs_res;
    }
    for (int a = 0; a < 3; a++) v.corg[a] = t->gbbmin[a] + t->vel[a] * t->mesh_tc;
    v.nx = t->gnx; v.ny = t->gny; v.nz = t->gnz;
    v.active = 0;  // set by the tool kernel from the instance's clock every step
    for (int a = 0; a < 3; a++) {
      v.bbmin[a] = t->bbmin[a];
      v.bbmax[a] = t->bbmax[a];
      v.vel_const[a] = t->vel_const[a];
    }
    v.dx = t->gdx;
    v.mu = t->mu;
    v.alpha = t->alpha;
    v.mass = t->mass;
    ToolMotion &tm = d.tm[k];
    tm.tris = t->tris;
    tm.p_init = t->p_init_all;
    tm.boxed = t->boxed;
    tm.p_init_boxed = t->p_init_boxed;
    tm.n_tris = t->n_tris;
    tm.n_boxed = t->n_boxed;
    tm.algo = t->algo;
    for (int a = 0; a < 3; a++) {
      tm.vel[a] = t->vel[a];
      tm.gbbmin[a] = t->gbbmin[a];
    }
    tm.mesh_tc = t->mesh_tc;
    tm.birth = t->birth;
    tm.death = t->death;
  }
  d.t = m.t;
  return d;
}

// descriptors and workgroup maps of every member on the device
int batch_build(mfg_batch *b, int record) {
  const int n = (int)b->members.size();
  batch_drop_graph(b);
  if (n == 0) {
    b->dirty = false;
    return MFG_OK;
  }
  if ((size_t)n > b->inst_cap) {
    if (b->d_inst) (void)hipFree(b->d_inst);
    b->d_inst = nullptr;
    b->inst_cap = 0;
    HIP_TRY(hipMalloc(&b->d_inst, sizeof(Inst) * (size_t)n));
    b->inst_cap = n;
  }
  constexpr int N_MAPS = 9;
  const size_t map_len = (size_t)n + 1;
  if (N_MAPS * map_len > b->maps_cap) {
    if (b->d_maps) (void)hipFree(b->d_maps);
    b->d_maps = nullptr;
    b->maps_cap = 0;
    HIP_TRY(hipMalloc(&b->d_maps, sizeof(unsigned) * N_MAPS * map_len));
    b->maps_cap = N_MAPS * map_len;
  }
  std::vector<Inst> host(n);
  std::vector<unsigned> maps(N_MAPS * map_len, 0u);
  BatchGrid *grids[N_MAPS] = {&b->L.cells1, &b->L.cells, &b->L.tiles, &b->L.one, &b->L.parts,
                              &b->L.parts8, &b->L.nbrs8, &b->L.tools, &b->L.search};
  // cell tile of the batched neighbour pass: by the cells per particle of the whole batch
  long cells = 0, parts = 0;
  for (const mfg_batch::Member &m : b->members) {
    cells += m.g->num_cell;
    parts += m.p->N_init;
  }
  const int tile = resolve_cell_tile(b->cell_tile ? b->cell_tile : g_lib.opt.cell_tile, cells,
                                     (unsigned)std::min<long>(parts, 0xffffffffL));
  for (int i = 0; i < n; i++) {
    const mfg_batch::Member &m = b->members[i];
    host[i] = batch_inst(m, record);
    const GridInfo gi = m.g->info();
    const unsigned nblk = (unsigned)m.p->nblk;
    const unsigned blocks[N_MAPS] = {
        (unsigned)((gi.num_cell + 1 + 255) / 256),
        (unsigned)((gi.num_cell + 255) / 256),
        (unsigned)host[i].nparts,
        1u,
        nblk,
        round8(nblk),
        round8((unsigned)neighbour_blocks(gi, tile)),
        (unsigned)std::max<size_t>(1, m.tools.size()),
        m.search ? std::min(BATCH_SEARCH_BLOCKS, std::max(1u, nblk)) : 0u};
    for (int k = 0; k < N_MAPS; k++) maps[k * map_len + i + 1] = maps[k * map_len + i] + blocks[k];
  }
  for (int k = 0; k < N_MAPS; k++) {
    const unsigned total = maps[k * map_len + n];
    if (total > 0x7fffffffu) return fail(MFG_ERR_ARG, "batched launch of %u workgroups", total);
    grids[k]->m.off = b->d_maps + k * map_len;
    grids[k]->m.n = n;
    grids[k]->blocks = total;
  }
  hipStream_t s = g_lib.stream;
  HIP_TRY(hipMemcpyAsync(b->d_inst, host.data(), sizeof(Inst) * n, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(b->d_maps, maps.data(), sizeof(unsigned) * maps.size(),
                         hipMemcpyHostToDevice, s));
  HIP_TRY(hipStreamSynchronize(s));  // the host arrays go out of scope
  b->L.inst = b->d_inst;
  b->L.n = n;
  b->L.precision = b->precision;
  b->L.pse = b->pse;
  b->L.cell_tile = tile;
  b->table_record = record;
  b->dirty = false;
  return MFG_OK;
}

void batch_launch_step(mfg_batch *b, int record) {
  hipStream_t s = g_lib.stream;
  {
    ProfScope ps("batch_sort");
    launch_batch_sort(s, b->L);
  }
  {
    ProfScope ps("batch_permute_predict");
    launch_batch_permute(s, b->L);
  }
  {
    ProfScope ps("batch_pair");
    launch_batch_pair(s, b->L);
  }
  {
    ProfScope ps("batch_contact");
    launch_batch_contact(s, b->L, record);
  }
}

// what the device does to an instance's host-visible state in one step, mirrored on the host
// with the same arithmetic (tool motion as mfg_tool_update, clock, state buffer)
void batch_mirror_step(mfg_batch::Member &m) {
  const double dt = m.c.dt;
  for (mfg_tool *t : m.tools) {
    if (t->n_boxed == 0) continue;
    t->mesh_tc = m.t;
    for (int k = 0; k < 3; k++) {
      t->bbmin[k] += dt * t->vel[k];
      t->bbmax[k] += dt * t->vel[k];
    }
  }
  m.t = m.t + dt;
  m.p->cur ^= 1;
  m.p->stat_n = 0;
  m.g->scanned = false;
  m.g->sorted_for = m.p;
}

}  // namespace

extern "C" {

int mfg_batch_create(mfg_batch **out) {
  NEED_LIB();
  if (!out) return fail(MFG_ERR_ARG, "null output");
  *out = new mfg_batch();
  return MFG_OK;
}

int mfg_batch_destroy(mfg_batch *b) {
  if (!b) return MFG_OK;
  batch_drop_graph(b);
  if (b->d_inst) (void)hipFree(b->d_inst);
  if (b->d_maps) (void)hipFree(b->d_maps);
  delete b;
  return MFG_OK;
}

int mfg_batch_add(mfg_batch *b, mfg_particles *p, mfg_grid *g, mfg_tool *const *tools,
                  int n_tools, const mfg_blanking_state *blanking, mfg_tool_forces *forces,
                  double dt, double t_current, int *slot) {
  NEED_LIB();
  if (!b || !p || !g) return fail(MFG_ERR_ARG, "null argument");
  if (n_tools < 0 || n_tools > MAX_TOOLS || (n_tools > 0 && !tools))
    return fail(MFG_ERR_ARG, "%d tools (max %d)", n_tools, MAX_TOOLS);
  for (int k = 0; k < n_tools; k++)
    if (!tools[k]) return fail(MFG_ERR_ARG, "tool %d is null", k);
  p->sorted_grid = nullptr;  // batched steps sort on their own: a later single step sorts
  if (forces && forces->n < (unsigned)n_tools)
    return fail(MFG_ERR_ARG, "the force record holds %u tools, the instance has %d", forces->n,
                n_tools);
  if (!g->locked || g->vel_bbox[0] != 0. || g->vel_bbox[1] != 0. || g->vel_bbox[2] != 0.)
    return fail(MFG_ERR_ARG, "a batched instance needs a locked grid that does not move");
  if (g->num_cell <= 0) return fail(MFG_ERR_ARG, "the grid has no cells");
  Consts c = make_consts(dt, p->precision);
  const int pse = (c.heat == MFG_HEAT_PSE && c.pse_on) ? 1 : 0;
  if (b->precision >= 0 && (b->precision != p->precision || b->pse != pse))
    return fail(MFG_ERR_ARG,
                "instances of a batch share the precision mode and the heat model (batch: "
                "precision %d, PSE %d; instance: %d, %d)",
                b->precision, b->pse, p->precision, pse);
  for (const mfg_batch::Member &m : b->members) {
    if (m.p == p || m.g == g)
      return fail(MFG_ERR_ARG, "the particle set or grid is already in the batch");
    for (mfg_tool *t : m.tools)
      for (int k = 0; k < n_tools; k++)
        if (t == tools[k]) return fail(MFG_ERR_ARG, "tool %d already moves with another instance", k);
  }
  mfg_batch::Member m;
  m.p = p;
  m.g = g;
  m.tools.assign(tools, tools + n_tools);
  m.forces = forces;
  m.blanking = blanking ? 1 : 0;
  for (int k = 0; k < 3; k++) {
    m.bb[k] = blanking ? blanking->bbmin[k] : 0.;
    m.bb[3 + k] = blanking ? blanking->bbmax[k] : 0.;
  }
  m.vmax2 = blanking ? blanking->max_vel_squared : FLT_MAX;
  m.c = c;
  m.t = t_current;
  int n_exh = 0;
  for (int k = 0; k < n_tools; k++)
    n_exh += (tools[k]->algo == MFG_CONTACT_EXHAUSTIVE && tools[k]->n_boxed > 0);
  m.search = (n_exh > 0 && g_lib.opt_contact_search) ? 1 : 0;
  if (m.search) {
    ContactSearch cs{};
    int rc = contact_search_buffers(p, n_tools, cs);
    if (rc != MFG_OK) return rc;
  }
  b->precision = p->precision;
  b->pse = pse;
  b->members.push_back(m);
  b->dirty = true;
  if (slot) *slot = (int)b->members.size() - 1;
  return MFG_OK;
}

int mfg_batch_remove(mfg_batch *b, int slot) {
  NEED_LIB();
  if (!b || slot < 0 || slot >= (int)b->members.size())
    return fail(MFG_ERR_ARG, "no instance %d in the batch", slot);
  HIP_TRY(hipStreamSynchronize(g_lib.stream));  // its last steps may still be queued
  b->members.erase(b->members.begin() + slot);
  if (b->members.empty()) {
    b->precision = -1;
    b->pse = -1;
  }
  b->dirty = true;
  return MFG_OK;
}

int mfg_batch_size(const mfg_batch *b) { return b ? (int)b->members.size() : MFG_ERR_ARG; }

int mfg_batch_set_cell_tile(mfg_batch *b, int cell_tile) {
  if (!b) return fail(MFG_ERR_ARG, "null batch");
  if (cell_tile != 0 && cell_tile != 1 && cell_tile != 2 && cell_tile != 4 && cell_tile != 8 &&
      cell_tile != 16)
    return fail(MFG_ERR_ARG, "cell tile %d (0, 1, 2, 4, 8 or 16)", cell_tile);
  if (cell_tile != b->cell_tile) b->dirty = true;
  b->cell_tile = cell_tile;
  return MFG_OK;
}

int mfg_batch_get_time(const mfg_batch *b, int slot, double *t) {
  if (!b || !t || slot < 0 || slot >= (int)b->members.size())
    return fail(MFG_ERR_ARG, "no instance %d in the batch", slot);
  *t = b->members[slot].t;
  return MFG_OK;
}

int mfg_batch_step(mfg_batch *b, int n_steps, int record_forces, int use_graph) {
  NEED_LIB();
  if (!b || n_steps < 0) return fail(MFG_ERR_ARG, "batch step: %d steps", n_steps);
  if (b->members.empty() || n_steps == 0) return MFG_OK;
  const int record = record_forces ? 1 : 0;
  if (record)
    for (const mfg_batch::Member &m : b->members)
      if (!m.forces && !m.tools.empty())
        return fail(MFG_ERR_ARG, "recording forces needs a force record for every instance");
  if (b->dirty || b->table_record != record) {
    int rc = batch_build(b, record);
    if (rc != MFG_OK) return rc;
  }
  hipStream_t s = g_lib.stream;
  if (use_graph) {
    if (!b->exec || b->graph_record != record || b->graph_stream != s) {
      batch_drop_graph(b);
      ProfOff off(true);
      HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      batch_launch_step(b, record);
      hipGraph_t gr = nullptr;
      hipError_t e = hipStreamEndCapture(s, &gr);
      if (e != hipSuccess || !gr) return fail(MFG_ERR_HIP, "capturing the batched step: %s", hipGetErrorString(e));
      b->graph = gr;
      HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->graph_record = record;
      b->graph_stream = s;
    }
    for (int k = 0; k < n_steps; k++) {
      ProfScope ps("batch_step_graph");
      HIP_TRY(hipGraphLaunch(b->exec, s));
    }
  } else {
    for (int k = 0; k < n_steps; k++) batch_launch_step(b, record);
  }
  HIP_TRY(hipGetLastError());
  for (int k = 0; k < n_steps; k++)
    for (mfg_batch::Member &m : b->members) batch_mirror_step(m);
  return MFG_OK;
}


// ---- profiling ----
int mfg_pair_stats_enable(int on) {
  NEED_LIB();
  g_lib.count_pairs = on != 0;
  return MFG_OK;
}
int mfg_profile_enable(int on) {
  NEED_LIB();
  if (on < 0 || on > MFG_PROFILE_PASSES) return fail(MFG_ERR_ARG, "profile level %d", on);
  if (!on) prof_collect();
  g_lib.prof = on;
  return MFG_OK;
}
int mfg_profile_reset(void) {
  NEED_LIB();
  prof_collect();
  g_lib.acc.clear();
  g_lib.names.clear();
  return MFG_OK;
}
int mfg_profile_count(void) {
  if (!g_lib.ready) return 0;
  prof_collect();
  return (int)g_lib.names.size();
}
int mfg_profile_get(int k, char *name48, long *launches, double *total_ms) {
  if (!g_lib.ready) return fail(MFG_ERR_STATE, "mfg_init was not called");
  prof_collect();
  if (k < 0 || k >= (int)g_lib.names.size()) return fail(MFG_ERR_ARG, "profile entry %d", k);
  const std::string &n = g_lib.names[k];
  if (name48) {
    std::strncpy(name48, n.c_str(), 47);
    name48[47] = 0;
  }
  auto &v = g_lib.acc[n];
  if (launches) *launches = v.first;
  if (total_ms) *total_ms = v.second;
  return MFG_OK;
}

// ---- warm starts ----
int mfg_interp_temps(const double *samples4, int n_in, double weights, double h,
                     const double *points3, int n_out, double *values) {
  NEED_LIB();
  if (n_in < 0 || n_out < 0 || (n_in > 0 && !samples4) || (n_out > 0 && (!points3 || !values)))
    return fail(MFG_ERR_ARG, "interp_temps: %d samples, %d points", n_in, n_out);
  if (!(h > 0.)) return fail(MFG_ERR_ARG, "interp_temps: smoothing length %g", h);
  if (n_out == 0) return MFG_OK;
  double *d_in = nullptr, *d_pts = nullptr, *d_out = nullptr;
  auto release = [&](int code) {
    if (d_in) (void)hipFree(d_in);
    if (d_pts) (void)hipFree(d_pts);
    if (d_out) (void)hipFree(d_out);
    return code;
  };
  if (hipMalloc(&d_in, sizeof(double) * 4 * (size_t)std::max(n_in, 1)) != hipSuccess ||
      hipMalloc(&d_pts, sizeof(double) * 3 * (size_t)n_out) != hipSuccess ||
      hipMalloc(&d_out, sizeof(double) * (size_t)n_out) != hipSuccess)
    return release(fail(MFG_ERR_NOMEM, "interp_temps buffers (%d samples, %d points)", n_in, n_out));
  hipStream_t s = g_lib.stream;
  if ((n_in > 0 && hipMemcpyAsync(d_in, samples4, sizeof(double) * 4 * (size_t)n_in,
                                  hipMemcpyHostToDevice, s) != hipSuccess) ||
      hipMemcpyAsync(d_pts, points3, sizeof(double) * 3 * (size_t)n_out, hipMemcpyHostToDevice,
                     s) != hipSuccess)
    return release(fail(MFG_ERR_HIP, "interp_temps upload"));
  launch_interp_temps(s, (const double4 *)d_in, n_in, weights, h, d_pts, n_out, d_out);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(values, d_out, sizeof(double) * (size_t)n_out, hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return release(fail(MFG_ERR_HIP, "interp_temps"));
  return release(MFG_OK);
}

}  // extern "C"
