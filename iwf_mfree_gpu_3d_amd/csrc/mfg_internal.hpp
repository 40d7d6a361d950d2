// mfg_internal.hpp -- device data layout and kernel launchers of the gfx950 SPH cutting step.
//
// Layout in HBM (one particle set):
//   persistent state, double buffered and permuted every step (SoA, FP64):
//     x y z | vx vy vz | vbx vby vbz | fcx fcy fcz ftx fty ftz | h rho | S[6] | eps_pl eps_pl_dot T
//     | xt yt zt | vxt vyt vzt | rho_t T_t | St[6] | flags (u32) | id (u32)
//   symmetric tensors are stored as xx, yy, zz, xy, yz, zx.
//   half-step records (sorted order, rebuilt every step by the permute/predict kernel):
//     nl4   : double4 (x, y, z, blanked)          -- candidate filter of the neighbour-list pass
//     pse2  : double2 (T, rho)                    -- PSE heat pass (only with PSE)
//     cs4   : double4 (x, y, z, vol; CSPM_TOOL_BITS for a tool particle) FP64; the first 32 bytes of
//             the record FP32                     -- CSPM sum of the pair pass
//     rec   : pair record, 192 B (FP64) / 112 B (FP32) -- everything a neighbour contributes
//   neighbour list: int32 nbr[i * cap + k] (row-major), count[i]
//   neighbour sums: 25 doubles per particle (SoA), written by the pair pass and read by the
//     per-particle update
//   grid: key[N_init] (unsorted), key_sorted[N_init], order[N_init], cell counts / offsets
//         [num_cell + 1], fill counters [num_cell]
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/mfree_gpu.h"

namespace mfg {

// particle flag bits (the reference stores these as floating point values and compares them
// with == 1., == 0. or as truthy values; the bits keep each of those predicates)
constexpr uint32_t FL_FIXED = 1u;      // fixed != 0        (src/actions_gpu.cu:470)
constexpr uint32_t FL_BLANKED = 2u;    // blanked == 1      (everywhere)
constexpr uint32_t FL_TOOL = 4u;       // tool_particle == 1 (everywhere)
constexpr uint32_t FL_FIXED_ONE = 8u;  // fixed == 1        (src/actions_gpu.cu:455)
constexpr uint32_t FL_TOOL_NZ = 16u;   // tool_particle != 0 (src/interactions_gpu.cu:167,230)

constexpr int KEY_BLANK = 0x7fffffff;  // INT_MAX, hash of a hard-blanked particle
// volume slot of a tool particle's FP64 CSPM entry: a quiet NaN with a payload that no
// arithmetic result carries (0/0 and inf/inf give the payload-free default NaN)
constexpr long long CSPM_TOOL_BITS = 0x7ff80000cafe7001ll;

struct State {
  double *x, *y, *z;
  double *vx, *vy, *vz;
  double *vbx, *vby, *vbz;
  double *fcx, *fcy, *fcz, *ftx, *fty, *ftz;
  double *h, *rho;
  double *S[6];
  double *eps_pl, *eps_pl_dot, *T;
  double *xt, *yt, *zt;
  double *vxt, *vyt, *vzt;
  double *rho_t, *T_t;
  double *St[6];
  uint32_t *flags;
  uint32_t *id;
};

// number of double arrays in State (for allocation)
constexpr int STATE_DOUBLES = 3 + 3 + 3 + 6 + 2 + 6 + 3 + 3 + 3 + 2 + 6;

// pair record, FP64: 24 doubles
struct alignas(16) Rec64 {
  double x, y, z, h;
  double T, vol, c, tool;  // with x .. h: what the CSPM sum of the per-lane and tiled forms reads
  double vx, vy, vz, rho;
  double A[6];  // (S - p I) / rho^2 for the diagonal, S / rho^2 off the diagonal
  double R[6];  // artificial stress
};

// pair record, FP32 pair arithmetic: positions stay FP64; the first 32 bytes hold everything
// the CSPM pass reads (x, y, z, tool, vol)
struct alignas(16) Rec32 {
  double x, y, z;
  float tool, vol;
  float h, vx, vy, vz;
  float rho, T, c, pad;
  float A[6];
  float R[6];
};

static_assert(sizeof(Rec64) == 192, "Rec64 layout");
static_assert(sizeof(Rec32) == 112, "Rec32 layout");

struct Consts {
  // actions translation unit (src/actions_gpu.cu:30-34)
  double a_rho0, a_K, a_G, a_mass, a_stresseps;
  double jc_A, jc_B, jc_C, jc_n, jc_m, jc_Tmelt, jc_Tref, jc_eps_dot_ref, jc_clamp;
  double twp_cp, twp_tq, twp_T_init, twp_eta;
  int plastic, a_thermal, a_fric, pad0;
  // interactions translation unit (src/interactions_gpu.cu:26-34)
  double i_mass, i_K, wdeltap, v_alpha, v_beta, v_eta, xspheps;
  double alpha_wp, alpha_tool;
  int pse_on, pad1;
  // switches
  int cspm, heat, single_sqrt, precision;
  double dt;
};

struct GridInfo {
  int nx, ny, nz;
  int hard_blank;
  double bbmin[3];
  double dx;
  long num_cell;
};

// device counters of one particle set
struct Counters {
  unsigned N;            // particles in cells (N of the reference)
  unsigned n_blank_key;  // INT_MAX keys
  int err_capacity;      // neighbour list overflow
  int err_grid;          // unblanked particle outside the grid
  int max_count;         // largest neighbour count of the last list build
  int err_list;          // tiled pair pass: a list entry outside the particle's stencil
  unsigned long long candidates;  // candidate pairs visited by the last neighbour pass
  unsigned long long pairs;       // interacting pairs of the last neighbour pass
};

// triangle as stored on the device: 30 doubles (p1 p2 p3 n np1 np2 np3 ne1 ne2 ne3)
struct Tri {
  double p[3][3];
  double n[3];
  double np[3][3];
  double ne[3][3];
};

// per-tool view passed to the contact kernel
struct ToolView {
  const Tri *tris;       // exhaustive: all triangles; hashing: boxed triangles
  const int *cells;      // CSR offsets of the triangle grid (num_cell + 1), hashing
  const int *cand_off;   // exhaustive: closest-triangle candidate lists per grid cell (CSR)
  const int *cand_idx;   //   (triangle indices, ascending within a cell)
  double corg[3];        // exhaustive: grid origin in the mesh's current position
  const int *slot;       // exhaustive: per particle, its entry in the search list (or -1); null
                         //   when the closest triangle is searched per thread
  const double *res;     // per search-list entry: closest point, normal (6 doubles)
  int n_tris;            // triangles in `tris`
  int algo;              // MFG_CONTACT_*
  int nx, ny, nz;
  int active;            // boxed triangles exist and birth <= t <= death
  double bbmin[3], bbmax[3];  // moving bounding box (geometry constant of the reference)
  double dx;
  double mu, alpha, mass;
  double vel_const[3];   // tool_vel constant (only set by set_vel)
};

constexpr int MAX_TOOLS = 8;
struct ToolSet {
  ToolView t[MAX_TOOLS];
  int n;
  int record;
  double fric_vel[3];    // global_tool[0]->get_vel() for friction heat and tool particle motion
};

// ---------------------------------------------------------------------------------------------
// kernel launchers (implemented in k_grid.hip, k_pair.hip, k_tool.hip)
// ---------------------------------------------------------------------------------------------
struct GridBuffers {
  int *key;          // N_init, unsorted
  int *key_sorted;   // N_init
  int *order;        // N_init
  int *count;        // num_cell + 1 (counts, then exclusive offsets)
  int *fill;         // num_cell
  int *blk;          // per-block blanked counts / offsets
  int *blk2;         // scan scratch
  int nblk;          // blocks of the hash kernel
};

void launch_blank_hash_count(hipStream_t s, const State &st, unsigned n_init, const GridInfo &g,
                             int blanking, const double *bb, double vmax2, GridBuffers &gb,
                             Counters *cnt);
void launch_scan_cells(hipStream_t s, const GridInfo &g, GridBuffers &gb, unsigned n_init,
                       Counters *cnt);
void launch_scatter(hipStream_t s, unsigned n_init, const GridInfo &g, GridBuffers &gb,
                    Counters *cnt);
void launch_fixup(hipStream_t s, const GridInfo &g, GridBuffers &gb);
void launch_permute_predict(hipStream_t s, const State &src, const State &dst, unsigned n_init,
                            const GridBuffers &gb, const Counters *cnt, const Consts &c,
                            int predict, void *rec, double4 *nl4, double2 *pse2, double4 *cs4);
void launch_cells_to_tables(hipStream_t s, const GridInfo &g, const int *offsets, int rothlin,
                            int *start, int *end);
void launch_bbox_reduce(hipStream_t s, const State &st, const Counters *cnt, unsigned n_init,
                        double *partials, int nblocks);

// neighbour lists (+ PSE heat) for every cell; stat_part gets 2 totals (candidates, pairs) per
// workgroup, neighbour_blocks(g) workgroups; poison[p] = 1 when a candidate of particle p holds a
// non-finite record (nl4.w == 2), which the reference's sums carry into p's rates
// cell_tile: consecutive cells per wavefront (1, 2, 4 or 8; resolve_cell_tile turns the option,
// 0 for automatic, into one of them)
int neighbour_blocks(const GridInfo &g, int cell_tile);
int resolve_cell_tile(int option, long num_cell, unsigned n_init);
void launch_neighbours(hipStream_t s, const GridInfo &g, const int *offsets, const double4 *nl4,
                       const double2 *pse2, const double *h, const uint32_t *flags,
                       const Consts &c, int *nbr, int cap, int *count, double *T_t, Counters *cnt,
                       int want_pse, unsigned long long *stat_part, int *poison, int cell_tile);
// pair pass (neighbour sums).  tiled: one wavefront per grid column stages the neighbour
// records of the column's particles in LDS (offsets and sorted keys of the grid); otherwise one
// thread per particle gathers them from memory
void launch_pair(hipStream_t s, unsigned n_init, const State &st, const void *rec, const int *nbr,
                 int nl_stride, const int *count, const Consts &c, Counters *cnt,
                 const GridInfo &g, const int *offsets, const int *key_sorted, int tiled,
                 double *sums, size_t sum_stride, const int *trav, const double4 *cs4,
                 int block);
// visiting order of the list pair pass: strips of `strip` grid columns along y, x-row by x-row
// inside a strip (trav[t] = particle visited at position t < N); seg_pos holds
// pair_order_segments(g, strip) ints
int pair_order_segments(const GridInfo &g, int strip);
void launch_pair_order(hipStream_t s, unsigned n_init, const GridInfo &g, const int *offsets,
                       const int *key_sorted, const Counters *cnt, int strip, int *seg_pos,
                       int *trav);
// neighbour sums per particle written by the pair pass (SoA, sum_stride apart)
constexpr int PAIR_SUM_COUNT = 25;
// Jaumann, continuity, momentum, advection, invalidation, leap-frog correction and Johnson-Cook
// radial return from the neighbour sums
void launch_particle_update(hipStream_t s, unsigned n_init, const State &st, const double *sums,
                            size_t sum_stride, const int *poison, const Consts &c,
                            const Counters *cnt);

// moves n triangles (and, in the same launch, n2 more of a second array) to time t
void launch_tool_update(hipStream_t s, Tri *tris, const double *p_init, int n, double t,
                        const double vel[3], Tri *tris2 = nullptr, const double *p_init2 = nullptr,
                        int n2 = 0);
// exhaustive contact, lane-parallel: list the (particle, tool) pairs inside a tool box, then one
// wavefront per pair finds the closest triangle; k_contact_bc reads the results
struct ContactSearch {
  int2 *list;           // (particle, tool)
  unsigned *list_n;     // entries appended this step
  unsigned cap;         // list capacity
  int *slot;            // n_init per tool: entry of the particle, -1 if not listed
  double *res;          // 6 doubles per entry
};
void launch_contact_search(hipStream_t s, const State &st, unsigned n_init, const ToolSet &ts,
                           const Counters *cnt, const ContactSearch &cs);
void launch_contact_bc(hipStream_t s, const State &st, unsigned n_init, const ToolSet &ts,
                       const Consts &c, const Counters *cnt, float *partials, int nblocks);
void launch_force_reduce(hipStream_t s, const float *partials, int nblocks, int n_tools,
                         float *forces);
// forces[0 .. n) = 0 by a kernel
void launch_zero_f32(hipStream_t s, float *a, unsigned n);

// ---------------------------------------------------------------------------------------------
// Batched launches: many independent instances advance together, every kernel of the step
// launched once over all of them.  Instance k owns the workgroups [off[k], off[k + 1]) of a
// launch; a workgroup finds its instance by a binary search of `off` and takes its arguments from
// the instance's descriptor.  Each kernel body is shared with the single-instance launch, so an
// instance computes exactly what it computes alone.
// ---------------------------------------------------------------------------------------------
struct BatchMap {
  const unsigned *off;  // n + 1 ascending workgroup offsets (device)
  int n;
};

// tool motion of one instance (what mfg_tool_update does on the host every step)
struct ToolMotion {
  Tri *tris;                  // all triangles (moved for the exhaustive algorithm)
  const double *p_init;
  Tri *boxed;                 // per-cell copies
  const double *p_init_boxed;
  int n_tris, n_boxed, algo, pad;
  double vel[3];
  double gbbmin[3];           // triangle-grid origin at time 0
  double mesh_tc;             // time the triangles were last moved to
  double birth, death;
};

// descriptor of one instance in device memory; kernels of the batched step read it and the
// step's last kernel advances the clock and flips the state buffers
struct Inst {
  State buf[2];
  int cur;                    // buf[cur]: state at the start of the step
  unsigned n_init;
  GridInfo g;
  GridBuffers gb;
  int nparts;                 // scan tiles of the cell counts
  int blanking;
  double bb[6], vmax2;        // blanking box and squared speed limit
  Consts c;
  void *rec;
  double4 *nl4;
  double2 *pse2;
  double4 *cs4;
  int *nbr, *count, *poison;
  int nl_stride, contact_blocks;
  double *sums;
  size_t sum_stride;
  Counters *cnt;
  float *partials;
  float *forces;              // null: no force record
  int n_forces, search;       // search: lane-parallel exhaustive contact search
  ContactSearch cs;
  ToolSet ts;
  ToolMotion tm[MAX_TOOLS];
  double t;                   // clock (global_time_current of the reference)
};

// host view of one batched launch shape: the device map and its total workgroups
struct BatchGrid {
  BatchMap m;
  unsigned blocks;
};
struct BatchLaunch {
  const Inst *inst;  // device array of n descriptors
  int n;
  int precision, pse;
  int cell_tile;     // cells per wavefront of the neighbour pass
  BatchGrid cells1;  // num_cell + 1 entries, 256 per workgroup
  BatchGrid cells;   // num_cell entries, 256 per workgroup
  BatchGrid tiles;   // scan tiles of the cell counts
  BatchGrid one;     // one workgroup per instance
  BatchGrid parts;   // N_init particles, 256 per workgroup
  BatchGrid parts8;  // the same, padded to a multiple of 8 workgroups (XCD-ordered kernels)
  BatchGrid nbrs8;   // neighbour-pass workgroups, padded to a multiple of 8
  BatchGrid tools;   // one workgroup per tool (at least one per instance)
  BatchGrid search;  // lane-parallel contact search workgroups
};
void launch_batch_sort(hipStream_t s, const BatchLaunch &L);
void launch_batch_permute(hipStream_t s, const BatchLaunch &L);
void launch_batch_pair(hipStream_t s, const BatchLaunch &L);
void launch_batch_contact(hipStream_t s, const BatchLaunch &L, int record);

// ---------------------------------------------------------------------------------------------
// Per-function entry points (the reference's actions, pair passes, one tool's contact and the
// leap-frog phases, one kernel each).  The reference keeps some per-particle arrays that the
// fused step holds only transiently; a particle set allocates them on the first such call.
// ---------------------------------------------------------------------------------------------
struct Aux {
  double *p;     // pressure, N_init
  double *R;     // artificial stress, 6 per particle (xx yy zz xy yz zx), particle-major
  double *vder;  // velocity gradient, 9 x stride (m[a][b] at (3a + b) * stride)
  double *sder;  // stress divergence terms, 9 x stride
  double *nrm;   // contact normal, 3 x stride
  double *x0;    // leap-frog start-of-step copies: x y z, vx vy vz, S[6], rho, T (14 x stride)
  size_t n;      // stride
};
enum {
  ACT_EOS, ACT_ART_STRESS, ACT_JAUMANN, ACT_CONTINUITY, ACT_MOMENTUM, ACT_ADVECTION,
  ACT_INVALIDATE, ACT_JOHNSON_COOK, ACT_FRIC_HEAT, ACT_BC, ACT_BC_THERMAL, ACT_MOVE_TOOL,
  ACT_RESET_CONTACT, ACT_LF_INIT
};
// vel: the tool velocity of ACT_FRIC_HEAT and ACT_MOVE_TOOL
void launch_action(hipStream_t s, int which, const State &st, unsigned n_init, const Counters *cnt,
                   const Consts &c, const Aux &a, const double *vel);
void launch_lf_advance(hipStream_t s, const State &st, unsigned n_init, const Aux &a, double f);
void launch_act_records(hipStream_t s, const State &st, unsigned n_init, const Counters *cnt,
                        const Consts &c, const Aux &a, void *rec, double4 *nl4, double2 *pse2,
                        double4 *cs4);
void launch_act_sums(hipStream_t s, const State &st, unsigned n_init, const Counters *cnt,
                     const Consts &c, const Aux &a, const double *sums, size_t stride,
                     const int *poison);
// one tool's contact; with record, per-workgroup force partials (nblocks x MAX_TOOLS x 3, the
// first three of each used)
void launch_act_contact(hipStream_t s, const State &st, unsigned n_init, const ToolView &tv,
                        double dt, const Counters *cnt, const Aux &a, int record, float *partials,
                        int nblocks);

// triangle binning on the GPU (grid_cpu_tri::get_cells)
int tool_bin_triangles(hipStream_t s, const Tri *tris, int n_tris, const double bbmin[3],
                       double dx, int nx, int ny, int nz, int **cells_out, int **tri_idx_out,
                       int *n_boxed_out);
// closest-triangle candidate lists of every cell of the triangle grid (exhaustive contact)
int tool_candidate_lists(hipStream_t s, const Tri *tris, int n_tris, const double bbmin[3],
                         double dx, int nx, int ny, int nz, int **off_out, int **idx_out,
                         long *total_out);
void launch_gather_tris(hipStream_t s, const Tri *src, const int *idx, int n, Tri *dst);
void launch_orient_test(hipStream_t s, const Tri *boxed, const int *cells, int nx, int ny,
                        int nz, const double bbmin[3], double dx, const double q[3], int *inside);
void launch_flip_normals(hipStream_t s, Tri *tris, int n);

int valid_check(hipStream_t s, const State &st, unsigned n, int *d_flag);

// warm start: v(p) = sum_j weights W(|p - x_j|, h) v_j of samples (x, y, z, v) at n_out points
void launch_interp_temps(hipStream_t s, const double4 *in, int n_in, double weights, double h,
                         const double *pts3, int n_out, double *out);

}  // namespace mfg
