/* mfree_gpu.h -- C ABI of the MI355X (gfx950) SPH cutting time step.
 *
 * Drop-in boundary for the reference's leap_frog::step and the objects it touches
 * (particle_gpu, grid_base / grid_gpu_green / grid_gpu_rothlin, the actions_* and
 * interactions_* constant set-ups, tool_3d_gpu, tool_forces, blanking).  The reference has no
 * FFI of its own; its boundary is a C++ class API plus process globals.  Every entry point below
 * names the reference interface it replaces (file:line relative to the reference tree).
 *
 * Conventions
 *   - plain pointers and sizes only; handles are opaque
 *   - every call returns MFG_OK (0) or a negative MFG_ERR_* code; mfg_last_error() gives the text
 *     (the reference instead prints and calls exit(), src/types.cpp:20-49)
 *   - arrays given to *_create are host arrays in the reference's layout (float4_t positions and
 *     velocities = 4 doubles per particle, mat3x3_t = 9 doubles in GLM column-major order)
 *   - all device work is queued on one HIP stream owned by the library (mfg_get_stream; the
 *     selected stream of a pool, mfg_select_stream); calls that return host data synchronise
 *     that stream
 *   - particle fields are stored in the current sorted order; they move every step
 *     (as the reference's Green grid swaps its buffers, src/grid_gpu_green.cu:242-263), so a
 *     caller never caches device pointers across steps and reads fields with
 *     mfg_particle_gpu_download.
 */
#ifndef MFREE_GPU_H_
#define MFREE_GPU_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------------ */
enum {
  MFG_OK = 0,
  MFG_ERR_ARG = -1,          /* invalid argument or handle */
  MFG_ERR_HIP = -2,          /* HIP runtime error (message in mfg_last_error) */
  MFG_ERR_STATE = -3,        /* call not valid in the current state (e.g. no device) */
  MFG_ERR_CAPACITY = -4,     /* a particle has more neighbours than the neighbour list holds */
  MFG_ERR_ALL_BLANKED = -5,  /* every particle was blanked (src/grid_gpu_base.cu:159-163) */
  MFG_ERR_GRID = -6,         /* unblanked particle outside the grid (src/grid_gpu_green.cu:56-58) */
  MFG_ERR_NOMEM = -7,        /* device allocation failed */
  MFG_ERR_NONFINITE = -8     /* non-finite state found by mfg_debug_check_valid */
};

/* ---- enumerations ------------------------------------------------------------------------ */
/* precision of the pair pass: FP64 throughout; FP32 pair arithmetic with Kahan-compensated FP32
 * neighbour sums; FP32 pair arithmetic with the neighbour sums widened to FP64 */
enum { MFG_FP64 = 0, MFG_FP32_KAHAN = 1, MFG_FP32_WIDE = 2 };
enum { MFG_HEAT_NONE = 0, MFG_HEAT_BROOKSHAW = 1, MFG_HEAT_PSE = 2 }; /* src/types.h:104-105 */
enum { MFG_CONTACT_EXHAUSTIVE = 0, MFG_CONTACT_HASHED = 1 };      /* src/tool_3d_gpu.h:77-79 */
enum { MFG_GRID_GREEN = 0, MFG_GRID_ROTHLIN = 1 };
enum { MFG_TOOL_WEDGE = 0, MFG_TOOL_GRAIN = 1 };

/* per-particle fields for mfg_particle_gpu_download / _upload (width in doubles) */
enum {
  MFG_F_POS = 0,        /* 3 */
  MFG_F_VEL = 1,        /* 3 */
  MFG_F_VEL_BC = 2,     /* 3 */
  MFG_F_FC = 3,         /* 3 */
  MFG_F_FT = 4,         /* 3 */
  MFG_F_H = 5,          /* 1 */
  MFG_F_RHO = 6,        /* 1 */
  MFG_F_P = 7,          /* 1  pressure of the last half step */
  MFG_F_S = 8,          /* 9  deviatoric stress, GLM column-major */
  MFG_F_R = 9,          /* 9  artificial stress of the last half step */
  MFG_F_EPS_PL = 10,    /* 1 */
  MFG_F_EPS_PL_DOT = 11,/* 1 */
  MFG_F_T = 12,         /* 1 */
  MFG_F_FIXED = 13,     /* 1 */
  MFG_F_BLANKED = 14,   /* 1 */
  MFG_F_TOOL = 15,      /* 1 */
  MFG_F_POS_T = 16,     /* 3 */
  MFG_F_VEL_T = 17,     /* 3 */
  MFG_F_RHO_T = 18,     /* 1 */
  MFG_F_S_T = 19,       /* 9 */
  MFG_F_T_T = 20,       /* 1 */
  MFG_F_S_DER = 21,     /* 9  stress divergence terms of the last interactions_monaghan call */
  MFG_F_V_DER = 22,     /* 9  velocity gradient of the last interactions_monaghan call */
  MFG_F_HASH = 23,      /* 1  cell hash (INT_MAX when blanked) */
  MFG_F_IDX = 24,       /* 1  sort permutation of the last sort */
  MFG_F_ID = 25,        /* 1  persistent particle id (creation index) */
  MFG_F_N = 26,         /* 3  contact normal of the last mfg_tool_compute_contact_force call */
  MFG_F_COUNT = 27
};

/* ---- constants: same members and order as src/constants_structs.h:25-92 (FP64 build) ------ */
typedef struct { double E, nu, rho0, K, G, mass; } mfg_phys_constants;
typedef struct { double cp, tq, k, alpha, T_init, eta; } mfg_trml_constants;
typedef struct { double wdeltap, stresseps, xspheps, alpha, beta, eta; } mfg_corr_constants;
typedef struct { double A, B, C, n, m, Tmelt, Tref, eps_dot_ref, clamp_temp; } mfg_joco_constants;
typedef struct {
  int nx, ny, nz;
  double bbmin_x, bbmin_y, bbmin_z, bbmax_x, bbmax_y, bbmax_z, dx;
} mfg_geom_constants;

/* switches that are compile-time macros in the reference (src/types.h:32-108) */
typedef struct {
  int cspm;           /* CSPM (Randles-Libersky) kernel correction, src/types.h:102 */
  int heat;           /* MFG_HEAT_*; Brookshaw is the reference default, src/types.h:104 */
  int single_sqrt;    /* 1: round the sqrtf / fmaxf sites of the reference's FP64 build through
                         float (src/kernels.cuh:48, src/actions_gpu.cu:50,203,217,420-421,
                         src/interactions_gpu.cu:414-415); 0: full double everywhere */
  int max_neighbours; /* neighbour-list capacity per particle; 0 selects 256 */
  int pair_pass;      /* MFG_PAIR_LIST (0): one thread per particle gathers its neighbour records;
                         MFG_PAIR_TILED (1): one wavefront per grid column stages the records of
                         the column's neighbour cells in LDS.  Same sums, same order. */
  int pair_strip;     /* list pair pass: particles are visited in strips of this many grid
                         columns along y, x-row by x-row inside a strip, which keeps the
                         neighbour records in flight within an XCD's L2; 0 selects 8 (FP64) or
                         16 (FP32 modes), or the sorted order when five x-rows of records fit in
                         2 MB (small instances: the strips would only add two launches), a
                         negative value the sorted order.  The sums do not depend on it. */
  int cell_tile;      /* neighbour pass: consecutive grid cells each wavefront takes in turn (1, 2,
                         4, 8 or 16; the tile's empty cells cost one ballot, not a workgroup);
                         0 selects 1, or, where the grid has two or more cells per particle
                         (wide margins around the block), 4 from 256k cells and 16 from 1M.
                         Lists and sums do not depend on it. */
  int sort_interval;  /* opt-in, no counterpart in the reference (which sorts every step,
                         src/leap_frog.cu:139-145): mfg_leap_frog_step sorts a particle set on
                         every k-th step only; in between, the particle order, the cell tables
                         and the visiting order of the last sort are kept (so a particle is
                         searched from the cell it was sorted into, the blanking box and the
                         locked grid's bounds are checked at sorts only, and the neighbour lists
                         and PSE sums differ from the reference's).  Applies to a locked grid that
                         does not move; 0 or 1: every step (the reference's results). */
  int pair_block;     /* list pair pass: threads per workgroup (64, 128 or 256; 0 selects 256).
                         Each wavefront works alone, so the sums do not depend on it; smaller
                         workgroups spread a small instance over more compute units. */
} mfg_options;
enum { MFG_PAIR_LIST = 0, MFG_PAIR_TILED = 1 };

/* opaque handles */
typedef struct mfg_particles mfg_particles;
typedef struct mfg_grid mfg_grid;
typedef struct mfg_tool mfg_tool;
typedef struct mfg_tool_forces mfg_tool_forces;
typedef struct mfg_leap_frog mfg_leap_frog;
typedef struct mfg_scene mfg_scene;
typedef struct mfg_batch mfg_batch;

/* blanking box as blanking::get_bb / get_max_vel_squared return it after blanking::update
 * (src/blanking.h:57-63, src/blanking.cpp:43-67) */
typedef struct {
  double bbmin[3];
  double bbmax[3];
  double max_vel_squared;
} mfg_blanking_state;

/* ---- library ------------------------------------------------------------------------------ */
/* Select the HIP device and create the library stream; replaces cudaFree(0) in
 * src/mfree_gpu_3d.cpp:78.  Must be called once before any other device call. */
int mfg_init(int device);
const char *mfg_last_error(void);
int mfg_synchronize(void);                       /* check_cuda_error, src/types.cpp:20-49 */
void *mfg_get_stream(void);                      /* hipStream_t of the library */
/* Streams for independent instances.  No counterpart in the reference, which runs one instance
 * per process on the default stream: these are for callers that step many independent cutting
 * instances in one process (the served and batched forms), so that the small grids of
 * different instances run on the GPU at the same time.  mfg_stream_pool(n) makes n streams
 * (1 .. MFG_MAX_STREAMS; stream 0 is the library stream, streams are kept until the process
 * ends); mfg_select_stream(k) makes stream k the one every later call queues on, and the one
 * mfg_synchronize and the host-data calls wait for.  All objects of one instance must be used
 * on one stream, and destroyed with that stream selected; instances on different streams share
 * no device state.  mfg_device_synchronize waits for every stream of the pool. */
#define MFG_MAX_STREAMS 16
int mfg_stream_pool(int n);
int mfg_select_stream(int k);
int mfg_device_synchronize(void);
int mfg_set_options(const mfg_options *opt);     /* src/types.h:32-108 */
/* Clears every constant set by the *_setup_* calls below, and the switches they turn on (thermal,
 * friction, plasticity), to the state after mfg_init.  No counterpart in the reference, whose
 * constants live as long as the process: it is for callers that run several set-ups in one
 * process, such as a request loop serving many cutting instances, which re-installs an
 * instance's constants before each of its steps. */
int mfg_constants_reset(void);
int mfg_get_options(mfg_options *opt);
int mfg_version(void);

/* ---- constant set-up (src/actions_gpu.h:71-75, src/interactions_gpu.h:133-141) ------------ */
int mfg_actions_setup_physical_constants(const mfg_phys_constants *c);      /* actions_gpu.cu:659 */
int mfg_actions_setup_corrector_constants(const mfg_corr_constants *c);     /* actions_gpu.cu:667 */
int mfg_actions_setup_johnson_cook_constants(const mfg_joco_constants *c);  /* actions_gpu.cu:676 */
int mfg_actions_setup_thermal_constants_wp(const mfg_trml_constants *c);    /* actions_gpu.cu:681 */
int mfg_actions_setup_thermal_constants_tool(const mfg_trml_constants *c);  /* actions_gpu.cu:700 */
int mfg_interactions_setup_physical_constants(const mfg_phys_constants *c); /* interactions_gpu.cu:564 */
int mfg_interactions_setup_corrector_constants(const mfg_corr_constants *c);/* interactions_gpu.cu:571 */
int mfg_interactions_setup_thermal_constants_workpiece(const mfg_trml_constants *c); /* :575 */
int mfg_interactions_setup_thermal_constants_tool(const mfg_trml_constants *c, mfg_tool *tool); /* :592 */
int mfg_interactions_setup_geometry_constants(const mfg_grid *g);           /* interactions_gpu.cu:552 */

/* ---- particle_gpu (src/particle_gpu.h:36-99, src/particle_gpu.cu:20-125) ------------------- */
/* pos4/vel4: 4 doubles per particle (w ignored); any of the optional arrays may be NULL and
 * then default to zero, as the reference's overloaded constructors do; S9 optional initial stress */
int mfg_particle_gpu_create(unsigned N, const double *pos4, const double *vel4, const double *rho,
                            const double *T_init, const double *h, const double *fixed,
                            const double *tool_p, const double *S9, int precision,
                            mfg_particles **out);
int mfg_particle_gpu_destroy(mfg_particles *p);
/* host view of particle_gpu::N / N_init (synchronises) */
int mfg_particle_gpu_counts(mfg_particles *p, unsigned *N, unsigned *N_init);
int mfg_particle_gpu_precision(const mfg_particles *p);
int mfg_particle_gpu_download(mfg_particles *p, int field, double *host);
int mfg_particle_gpu_upload(mfg_particles *p, int field, const double *host);
/* No counterpart in the reference (its particle_gpu lives as long as the process): a snapshot of
 * a particle set's persistent state (every per-particle field, flags and persistent ids, N) as of
 * its last queued step, into `bytes` of host memory, and its restore into a set of the same
 * N_init and precision (e.g. after the instance was evicted from the GPU and set up again).  The
 * state a step rebuilds (records, lists, cell tables) is not part of it.  Together with the
 * tools' motion state and the caller's clock, the next step continues bit-identically. */
int mfg_particle_gpu_snapshot_bytes(const mfg_particles *p, unsigned long long *bytes);
int mfg_particle_gpu_snapshot(mfg_particles *p, void *host);
int mfg_particle_gpu_restore(mfg_particles *p, const void *host);
int mfg_field_width(int field);
/* work of the last neighbour pass: candidate pairs visited and interacting pairs kept
 * (synchronises; MFG_ERR_STATE if that pass did not count them); no counterpart in the
 * reference, used for the roofline accounting.  Counting is off by default (it adds about 3 %
 * to the neighbour pass) and switched for every later pass by mfg_pair_stats_enable. */
int mfg_particle_gpu_pair_stats(mfg_particles *p, unsigned long long *candidates,
                                unsigned long long *pairs);
int mfg_pair_stats_enable(int on);

/* ---- grid_base / grid_gpu_green / grid_gpu_rothlin (src/grid_gpu_base.h:82-114) ------------ */
/* locked geometry: grid_gpu_green(int num_part, float3_t bbmin, float3_t bbmax, float_t h),
 * src/grid_gpu_green.h:119, src/grid_gpu_rothlin.h:97 */
int mfg_grid_create_locked(int kind, int num_part, const double bbmin[3], const double bbmax[3],
                           double h, mfg_grid **out);
/* geometry follows the particles: grid_gpu_green(unsigned max_cell, unsigned N), :118 */
int mfg_grid_create_unlocked(int kind, long max_cell, int num_part, mfg_grid **out);
int mfg_grid_destroy(mfg_grid *g);
/* nx() .. dx(), is_locked(), num_cell()  (src/grid_gpu_base.h:82-94) */
int mfg_grid_get_geometry(const mfg_grid *g, mfg_geom_constants *geom, long *num_cell, int *locked);
int mfg_grid_set_bbox_vel(mfg_grid *g, const double vel[3]);   /* :102 */
int mfg_grid_set_max_fixed(mfg_grid *g, int fixed);            /* :103 */
int mfg_grid_set_min_fixed(mfg_grid *g, int fixed);            /* :104 */
int mfg_grid_set_hard_blank(mfg_grid *g, int hard_blank);      /* :105 */
/* update_geometry(particle_gpu*, float_t kernel_width) at simulation time t (:97) */
int mfg_grid_update_geometry(mfg_grid *g, mfg_particles *p, double kernel_width, double t);
int mfg_grid_assign_hashes(const mfg_grid *g, mfg_particles *p);        /* :96 */
int mfg_grid_adapt_particle_number(const mfg_grid *g, mfg_particles *p); /* :107 */
int mfg_grid_sort(mfg_grid *g, mfg_particles *p);                        /* :99 */
/* get_cells(particle_gpu*, int *cell_start, int *cell_end): device pointers with num_cell ints
 * each; Green layout (0xffffffff for empty cells) or Rothlin layout (start == end) (:100) */
int mfg_grid_get_cells(mfg_grid *g, mfg_particles *p, int *d_cell_start, int *d_cell_end);
/* host copies of the same tables */
int mfg_grid_download_cells(mfg_grid *g, int *h_cell_start, int *h_cell_end);
/* neighbour lists of the current sorted particles, in the reference's pair-loop order
 * (stencil ii, jj, kk ascending, then sorted index; support r^2 < 4 h_i^2, r^2 >= 1e-8,
 * blanked j skipped; src/interactions_gpu.cu:298-336, src/kernels.cuh:41).
 * counts: N ints; lists: N*cap ints (may be NULL). */
int mfg_grid_neighbours(mfg_grid *g, mfg_particles *p, int *h_counts, int *h_lists, int cap);

/* ---- tool_3d_gpu (src/tool_3d_gpu.h:74-176) ------------------------------------------------- */
/* Builds the tool from an indexed triangle mesh: face, angle-weighted edge and vertex normals
 * (mesh_compute_vertex_normals, src/surface_tri_mesh.cpp:410-568), triangle grid with cell size
 * h_avg (<= 0: mesh_average_edge_length) binned on the GPU (grid_cpu_tri, src/grid_cpu_tri.cpp:
 * 48-175), normals oriented as tool_3d does (src/tool_3d.cpp:318-358).
 * vel / phys may be NULL: tool_3d_gpu(const tool_3d*, vec3_t, phys_constants, float_t alpha) and
 * tool_3d_gpu(const tool_3d*, float_t alpha), src/tool_3d_gpu.cu:763-856 */
int mfg_tool_3d_gpu_create(const double *verts3, int n_verts, const int *tris3, int n_tris,
                           double h_avg, const double *vel, const mfg_phys_constants *phys,
                           double alpha, mfg_tool **out);
int mfg_tool_destroy(mfg_tool *t);
int mfg_tool_set_vel(mfg_tool *t, const double vel[3]);          /* set_vel  :683 */
int mfg_tool_set_mu(mfg_tool *t, double mu);                     /* set_mu   :688 */
int mfg_tool_set_contact_alpha(mfg_tool *t, double alpha);       /* :693 */
int mfg_tool_set_algorithm_type(mfg_tool *t, int algo);          /* :741 */
int mfg_tool_set_birth_death(mfg_tool *t, double birth, double death); /* :698 */
int mfg_tool_set_thermal(mfg_tool *t, int thermal);              /* :719 */
int mfg_tool_set_physics(mfg_tool *t, const mfg_phys_constants *phys); /* :755 */
int mfg_tool_get_vel(const mfg_tool *t, double vel[3]);          /* :703 */
int mfg_tool_get_bbox(const mfg_tool *t, double bbmin[3], double bbmax[3]); /* :727 */
/* motion state of a tool for snapshots (no counterpart in the reference): the time its triangles
 * were last moved to, then its moving box (bbmin[3], bbmax[3]) */
int mfg_tool_get_motion(const mfg_tool *t, double motion7[7]);
int mfg_tool_set_motion(mfg_tool *t, const double motion7[7]);
int mfg_tool_get_mu(const mfg_tool *t, double *mu);              /* :715 */
int mfg_tool_is_alive(const mfg_tool *t, double t_current);      /* :737, returns 0/1 */
/* triangle grid: dims[0..2] = nx, ny, nz, dims[3] = boxed triangles, dims[4] = triangles;
 * bb_dx = bbmin[3], bbmax[3], dx of the grid at construction */
int mfg_tool_get_grid(const mfg_tool *t, long dims[5], double bb_dx[7]);
/* CSR cell offsets (num_cell + 1 unsigned) and boxed triangles as 30 doubles each
 * (p1, p2, p3, n, np1, np2, np3, ne1, ne2, ne3); either output may be NULL */
int mfg_tool_download_cells(const mfg_tool *t, unsigned *cells, double *boxed30);
/* update_tool (:907-948) at time t_current: triangle p = p_init + v t, bbox += dt v */
int mfg_tool_update(mfg_tool *t, double t_current, double dt);

/* ---- tool_forces (src/tool_3d_gpu.h:52-68) ------------------------------------------------- */
int mfg_tool_forces_create(unsigned num_tool, mfg_tool_forces **out);
int mfg_tool_forces_destroy(mfg_tool_forces *f);
int mfg_tool_forces_reset(mfg_tool_forces *f);
int mfg_tool_forces_get_host(mfg_tool_forces *f, float *xyz);   /* 3 floats per tool */

/* ---- actions (src/actions_gpu.h:57,58) ----------------------------------------------------- */
int mfg_perform_blanking(mfg_particles *p, const mfg_blanking_state *b);

/* ---- leap_frog (src/leap_frog.h:37-50) ----------------------------------------------------- */
int mfg_leap_frog_create(unsigned num_part, unsigned num_cell, mfg_leap_frog **out);
int mfg_leap_frog_destroy(mfg_leap_frog *lf);
/* leap_frog::step(particle_gpu*, grid_base*, bool record_forces), src/leap_frog.cu:121-195.
 * The reference reads global_tool, global_blanking, global_tool_forces, global_time_dt and
 * global_time_current (src/mfree_gpu_3d.cpp:41-48); they are arguments here.  `blanking` is the
 * state after blanking::update() for this step, or NULL when there is no global blanking. */
int mfg_leap_frog_step(mfg_leap_frog *lf, mfg_particles *p, mfg_grid *g, int record_forces,
                       mfg_tool *const *tools, int n_tools, const mfg_blanking_state *blanking,
                       mfg_tool_forces *forces, double dt, double t_current);
/* No counterpart in the reference.  mfg_leap_frog_step for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * same arguments, same results, and no hipMalloc / hipFree or allocating library call.  The
 * step's work memory beyond the particle set, grid and tools (the visiting-order segments of the
 * pair pass and the exhaustive contact search's slots, list and results) lives in one static
 * device array of MFG_STEP_SCRATCH_BYTES per GPU, held until the process ends.  The slots take
 * 4 bytes per particle and exhaustive tool, so N_init x tools up to about 16M fit; a step whose
 * work memory does not fit returns MFG_ERR_ARG before any kernel that changes the particles runs.  Every call on one GPU shares
 * the array: calls that overlap on one GPU (on different streams) race.  Per-kernel profiling
 * is off inside the call. */
#define MFG_STEP_SCRATCH_BYTES (64u << 20)
int mfg_leap_frog_step_noalloc(mfg_leap_frog *lf, mfg_particles *p, mfg_grid *g,
                               int record_forces, mfg_tool *const *tools, int n_tools,
                               const mfg_blanking_state *blanking, mfg_tool_forces *forces,
                               double dt, double t_current);
/* ---- batched steps ---------------------------------------------------------------------------
 * No counterpart in the reference, which steps one instance per process (leap_frog::step,
 * src/leap_frog.cu:121-195, driven by src/mfree_gpu_3d.cpp:155-205).  A batch advances many
 * independent instances together: every kernel of the step is launched once over all of them
 * (each instance with its own cell grid, constants, tools and clock in a device descriptor), and
 * a step can be captured into a HIP graph and replayed.  Each instance computes exactly what
 * mfg_leap_frog_step computes for it alone (same kernel bodies; the pair pass in the sorted
 * visiting order, which leaves results unchanged).
 *
 * mfg_batch_add takes the instance's objects, blanking state (NULL: none), force record (NULL:
 * none), time step and clock (global_time_current before its next step); the library's current
 * constants (the *_setup_* calls and mfg_set_options) become the instance's constants.  The
 * instances of a batch share the precision mode and the heat model; each needs a locked grid
 * that does not move, and a blanking box that does not move.  While an instance is in a batch,
 * only mfg_batch_step advances it; the library keeps its objects (state buffers, tool motion) and
 * clock in step on the host, so after mfg_batch_remove, or between batch steps, it can be read
 * (mfg_particle_gpu_download) or stepped alone.  mfg_batch_step queues n_steps steps of every
 * instance on the current stream (record_forces: every instance's force record holds the last
 * step's forces); use_graph: the step is captured once into a graph (again after the batch
 * changes) and replayed.  slot: the instance's index, which later removals shift down. */
int mfg_batch_create(mfg_batch **out);
int mfg_batch_destroy(mfg_batch *b);
int mfg_batch_add(mfg_batch *b, mfg_particles *p, mfg_grid *g, mfg_tool *const *tools,
                  int n_tools, const mfg_blanking_state *blanking, mfg_tool_forces *forces,
                  double dt, double t_current, int *slot);
int mfg_batch_remove(mfg_batch *b, int slot);
int mfg_batch_size(const mfg_batch *b);
/* cells per wavefront of the batch's neighbour pass (1, 2, 4, 8, 16), as mfg_options.cell_tile for a
 * single step; 0 (the default): mfg_options.cell_tile, automatic over the whole batch when that
 * is 0 too.  Results do not depend on it. */
int mfg_batch_set_cell_tile(mfg_batch *b, int cell_tile);
int mfg_batch_get_time(const mfg_batch *b, int slot, double *t);
int mfg_batch_step(mfg_batch *b, int n_steps, int record_forces, int use_graph);

/* ---- per-function entry points ------------------------------------------------------------
 * The reference's actions, pair passes, one tool's contact and the leap-frog phases, each one
 * launch over the particle set's current (sorted) state, so that its leap_frog::step
 * (src/leap_frog.cu:121-195) can be composed call by call.  mfg_leap_frog_step does the same
 * work fused.  The arrays the reference keeps and the fused step does not (pressure, artificial
 * stress, velocity gradient, stress divergence terms, contact normal, the leap-frog start copies)
 * are allocated on the first call; they are step-local and not permuted by mfg_grid_sort.
 * Time steps are arguments (the reference reads global_time_dt). */
int mfg_actions_material_eos(mfg_particles *p);                           /* actions_gpu.h:41 */
int mfg_actions_corrector_artificial_stress(mfg_particles *p);            /* actions_gpu.h:44 */
int mfg_actions_material_stress_rate_jaumann(mfg_particles *p);           /* actions_gpu.h:45 */
int mfg_actions_material_fric_heat_gen(mfg_particles *p, const double vel[3], double dt); /* :48 */
int mfg_actions_contmech_continuity(mfg_particles *p);                    /* actions_gpu.h:51 */
int mfg_actions_contmech_momentum(mfg_particles *p);                      /* actions_gpu.h:52 */
int mfg_actions_contmech_advection(mfg_particles *p);                     /* actions_gpu.h:53 */
int mfg_actions_plasticity_johnson_cook(mfg_particles *p, double dt);     /* actions_gpu.h:61 */
int mfg_actions_perform_boundary_conditions_thermal(mfg_particles *p);   /* actions_gpu.h:64 */
int mfg_actions_perform_boundary_conditions(mfg_particles *p);            /* actions_gpu.h:65 */
int mfg_actions_move_tool_particles(mfg_particles *p, const mfg_tool *t, double dt); /* :68 */
int mfg_debug_invalidate(mfg_particles *p);                               /* actions_gpu.h:79 */
/* interactions_monaghan / interactions_heat_pse (src/interactions_gpu.h:51-52) on the current
 * positions, with the grid's cell tables of the current order (after mfg_grid_sort): v_der,
 * S_der, pos_t (XSPH), vel_t and, with Brookshaw conduction, T_t; T_t of the PSE sum */
int mfg_interactions_monaghan(mfg_particles *p, mfg_grid *g);
int mfg_interactions_heat_pse(mfg_particles *p, mfg_grid *g);
/* tool_3d_gpu::compute_contact_force (src/tool_3d_gpu.h:120) at time t_current (after
 * mfg_tool_update): fc, ft and the normal of every particle in contact; with record, the tool's
 * force is added to slot `slot` of the force record */
int mfg_tool_compute_contact_force(mfg_tool *t, mfg_particles *p, int record,
                                   mfg_tool_forces *forces, int slot, double dt,
                                   double t_current);
/* leap_frog's phases (private in the reference, src/leap_frog.cu:30-119): the start-of-step
 * copies; x = x0 + f x_t for position, velocity, stress, density and temperature (predict
 * f = dt / 2, correct f = dt); the reset of the contact forces and normals */
int mfg_leap_frog_init(mfg_leap_frog *lf, mfg_particles *p);
int mfg_leap_frog_advance(mfg_leap_frog *lf, mfg_particles *p, double f);
int mfg_leap_frog_reset_contact_forces(mfg_leap_frog *lf, mfg_particles *p);

/* debug_check_valid_full (src/actions_gpu.cu:739-745): MFG_ERR_NONFINITE if an unblanked
 * particle holds a non-finite value */
int mfg_debug_check_valid(mfg_particles *p);

/* ---- per-kernel timing (HIP events on the library stream) ----------------------------------- */
/* on: 0 off; MFG_PROFILE_ALL a pair of events around every kernel; MFG_PROFILE_PASSES around the
 * neighbour and pair passes only (each event pair adds a little to a step: a timed run of many
 * steps records only what it needs) */
#define MFG_PROFILE_ALL 1
#define MFG_PROFILE_PASSES 2
int mfg_profile_enable(int on);
int mfg_profile_reset(void);
/* number of kernels recorded; for k < n: name (<= 47 chars), launches and total milliseconds */
int mfg_profile_count(void);
int mfg_profile_get(int k, char *name48, long *launches, double *total_ms);

/* ---- tool mesh readers (set-up time, host only) ------------------------------------------- */
/* Surface of an LS-DYNA keyword solid mesh (quads = 0: tetrahedra, 1: hexahedra), as
 * ldynak_read_triangles_from_tetmesh / _quadmesh (src/ldynak_reader.cpp:308-434), and a MeshLab
 * JSON triangle mesh, as json_read_triangles (src/json_reader.cpp:108-117).  Call once with NULL
 * arrays for the sizes, then with arrays of 3 * n_verts doubles and 3 * n_tris ints.  Vertices
 * come by ascending node id and faces in element order (the reference's order follows
 * std::unordered_map iteration). */
int mfg_mesh_read_ldynak(const char *path, int quads, int *n_verts, int *n_tris, double *verts3,
                         int *tris3);
int mfg_mesh_read_json(const char *path, int *n_verts, int *n_tris, double *verts3, int *tris3);
const char *mfg_mesh_last_error(void);

/* ---- warm starts (set-up time) ------------------------------------------------------------ */
/* Particle positions from a legacy VTK file, as vtk_read_init_pos (src/vtk_reader.cpp:29-69):
 * the POINTS lines of exactly three space-separated fields, up to the CELLS line.  Call once
 * with xyz NULL for n, then with 3 * n doubles.  Host only. */
int mfg_vtk_read_positions(const char *path, int *n, double *xyz);
/* Steady-state temperature samples "x y z T" (one per line, lines of exactly four
 * space-separated fields), as the single-grain set-up reads them
 * (src/Benchmarks/benchmarks_single_grain.cu:281-301).  Two calls as above, 4 * n doubles. */
int mfg_read_steady_temperatures(const char *path, int *n, double *xyzt);
/* interp_temps (src/Benchmarks/interp_utils.cpp:65-99) on the GPU: for each of the n_out points
 * (3 doubles each) values[i] = sum over the n_in samples (x, y, z, v; 4 doubles each), in sample
 * order, of weights * W(|p_i - x_j|, h) * v_j, W the cubic spline with the reference's sqrtf
 * site.  Host arrays in and out; the reference also rewrites the points' w component and dumps
 * debug files, neither of which is done here. */
int mfg_interp_temps(const double *samples4, int n_in, double weights, double h,
                     const double *points3, int n_out, double *values);

/* ---- workload builder (set-up time, host only) --------------------------------------------- */
typedef struct {
  int nx, ny, nz;             /* workpiece lattice */
  double dx;                  /* lattice spacing [cm] */
  double jitter;              /* uniform position jitter, fraction of dx */
  unsigned long long seed;
  double hdx;                 /* h / dx */
  double cfl;
  int tool_kind;              /* MFG_TOOL_* */
  int n_grains;
  double grain_offset_x, grain_offset_y;  /* spacing of grains along / across the cut [cm] */
  double rake_deg, clearance_deg;
  double depth_of_cut;        /* [cm] */
  double speed;               /* tool speed along +x [cm/us] */
  double grain_edge;          /* truncated-octahedron edge length [cm] */
  double edge_radius;         /* edge rounding [cm] */
  double facet_perturb;       /* per-facet offset amplitude [cm] */
  int grain_res;              /* cube-sphere quads per face edge (12 res^2 triangles) */
  int tool_particles;         /* sample the tool interior (thermal contact) */
  double tool_rho, tool_cp, tool_k;
  double E, nu, rho0, cp, k, tq, T_init, eta_fric;
  double jc_A, jc_B, jc_n, jc_C, jc_m, jc_eps_dot_ref, jc_Tref, jc_Tmelt, jc_clamp_temp;
  double visc_alpha, visc_beta, visc_eta, xspheps, stresseps;
  double mu, contact_alpha, max_vel;
  int cspm, heat, algo, steps;
} mfg_cut_config;

typedef struct {
  mfg_phys_constants phys;
  mfg_corr_constants corr;
  mfg_joco_constants joco;
  mfg_trml_constants trml_wp;
  mfg_trml_constants trml_tool;
  double dt;
  int steps;
  double tool_vel[3];
  double h;
  double bbmin[3], bbmax[3];  /* locked grid and blanking box */
  double max_vel;             /* blanking(bbmin, bbmax, max_vel): squared by the constructor */
  int cspm, heat, algo;
  double mu, contact_alpha;
  int n_tools;
  unsigned n_particles, n_workpiece;
} mfg_scene_params;

/* presets: 1 = 8,000-particle wedge cut, 2 = 1M single grain, 3 = 8M three grains;
 * `scale` multiplies the lattice counts (1.0 = full size) */
int mfg_scene_preset(int which, double scale, unsigned long long seed, mfg_cut_config *cfg);
mfg_scene *mfg_scene_build(const mfg_cut_config *cfg);
void mfg_scene_free(mfg_scene *s);
int mfg_scene_get_params(const mfg_scene *s, mfg_scene_params *out);
/* 3 doubles per particle for pos / vel; any output may be NULL */
int mfg_scene_get_particles(const mfg_scene *s, double *pos3, double *vel3, double *rho, double *T,
                            double *h, double *fixed, double *tool_p);
int mfg_scene_tool_size(const mfg_scene *s, int k, int *n_verts, int *n_tris);
int mfg_scene_tool_mesh(const mfg_scene *s, int k, double *verts3, int *tris3);

#ifdef __cplusplus
}
#endif

#endif /* MFREE_GPU_H_ */
