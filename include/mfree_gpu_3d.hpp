// mfree_gpu_3d.hpp -- the reference's class API for the cutting step, as thin C++ classes over
// the C ABI of mfree_gpu.h.  A driver written against the reference (src/mfree_gpu_3d.cpp:75-205)
// keeps its calls: particle_gpu, grid_gpu_green / grid_gpu_rothlin (grid_base), tool_3d_gpu,
// tool_forces, blanking, leap_frog::step and the process globals global_tool,
// global_tool_forces, global_blanking and global_time_*.
//
// Differences a caller sees:
//   - positions / velocities are float4_t-shaped host arrays of 4 doubles, mat3x3_t is 9 doubles
//     in GLM column-major order (the reference's own layouts)
//   - errors throw mfg_shim::error (the reference prints and calls exit(), src/types.cpp:20-49)
//   - particle fields live on the device in the current sorted order and move every step (the
//     Green grid swaps its buffers, src/grid_gpu_green.cu:242-263): read them with
//     particle_gpu::download instead of caching device pointers
//   - the precision of the pair pass is a run-time argument of particle_gpu (the reference
//     selects it at compile time, src/types.h:32)
//   - tool_3d_gpu is built from an indexed triangle mesh (the reference takes a tool_3d, which
//     holds the same mesh, src/tool_3d_gpu.cu:763-856)
//   - the globals are declared here and defined once by the driver (MFG_DEFINE_GLOBALS in one
//     translation unit), as the reference defines them in src/mfree_gpu_3d.cpp:41-48
#ifndef MFREE_GPU_3D_HPP_
#define MFREE_GPU_3D_HPP_

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "mfree_gpu.h"

namespace mfg_shim {

struct error : std::runtime_error {
  int code;
  error(int c, const char *what) : std::runtime_error(what ? what : "mfree_gpu error"), code(c) {}
};

inline void check(int rc) {
  if (rc != MFG_OK) throw error(rc, mfg_last_error());
}

struct vec3 {
  double x = 0., y = 0., z = 0.;
};

}  // namespace mfg_shim

// ---- particle_gpu (src/particle_gpu.h:36-99) ---------------------------------------------------
class particle_gpu {
 public:
  unsigned int N = 0, N_init = 0;
  mfg_particles *handle = nullptr;

  explicit particle_gpu(unsigned int n, int precision = MFG_FP64) {
    build(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, n, precision);
  }
  particle_gpu(const double *pos4, const double *vel4, const double *rho, const double *h,
               unsigned int n, int precision = MFG_FP64) {
    build(pos4, vel4, rho, nullptr, h, nullptr, nullptr, nullptr, n, precision);
  }
  particle_gpu(const double *pos4, const double *vel4, const double *rho, const double *h,
               const double *fixed, unsigned int n, int precision = MFG_FP64) {
    build(pos4, vel4, rho, nullptr, h, fixed, nullptr, nullptr, n, precision);
  }
  particle_gpu(const double *pos4, const double *vel4, const double *rho, const double *T_init,
               const double *h, const double *fixed, unsigned int n, int precision = MFG_FP64) {
    build(pos4, vel4, rho, T_init, h, fixed, nullptr, nullptr, n, precision);
  }
  particle_gpu(const double *pos4, const double *vel4, const double *rho, const double *T_init,
               const double *h, const double *fixed, const double *tool_p, unsigned int n,
               int precision = MFG_FP64) {
    build(pos4, vel4, rho, T_init, h, fixed, tool_p, nullptr, n, precision);
  }
  // (pos, vel, rho, h, Sini): the stress is 9 doubles per particle, GLM column-major
  particle_gpu(const double *pos4, const double *vel4, const double *rho, const double *h,
               const double *S9, unsigned int n, int precision, bool /*with_stress*/) {
    build(pos4, vel4, rho, nullptr, h, nullptr, nullptr, S9, n, precision);
  }
  ~particle_gpu() { mfg_particle_gpu_destroy(handle); }
  particle_gpu(const particle_gpu &) = delete;
  particle_gpu &operator=(const particle_gpu &) = delete;

  // refresh N / N_init from the device (they change when blanking adapts the particle number)
  void sync_counts() { mfg_shim::check(mfg_particle_gpu_counts(handle, &N, &N_init)); }
  // one field (MFG_F_*) of every particle in the current sorted order, mfg_field_width doubles each
  std::vector<double> download(int field) {
    sync_counts();
    std::vector<double> out((size_t)N_init * (size_t)mfg_field_width(field));
    mfg_shim::check(mfg_particle_gpu_download(handle, field, out.data()));
    return out;
  }

 private:
  void build(const double *pos4, const double *vel4, const double *rho, const double *T,
             const double *h, const double *fixed, const double *tool_p, const double *S9,
             unsigned int n, int precision) {
    mfg_shim::check(mfg_particle_gpu_create(n, pos4, vel4, rho, T, h, fixed, tool_p, S9,
                                            precision, &handle));
    N = N_init = n;
  }
};

// ---- grid_base / grid_gpu_green / grid_gpu_rothlin (src/grid_gpu_base.h:82-114) --------------
class grid_base {
 public:
  mfg_grid *handle = nullptr;
  virtual ~grid_base() { mfg_grid_destroy(handle); }
  grid_base(const grid_base &) = delete;
  grid_base &operator=(const grid_base &) = delete;

  int nx() const { return geom().nx; }
  int ny() const { return geom().ny; }
  int nz() const { return geom().nz; }
  double bbmin_x() const { return geom().bbmin_x; }
  double bbmin_y() const { return geom().bbmin_y; }
  double bbmin_z() const { return geom().bbmin_z; }
  double bbmax_x() const { return geom().bbmax_x; }
  double bbmax_y() const { return geom().bbmax_y; }
  double bbmax_z() const { return geom().bbmax_z; }
  double dx() const { return geom().dx; }
  bool is_locked() const {
    mfg_geom_constants g;
    long nc = 0;
    int locked = 0;
    mfg_shim::check(mfg_grid_get_geometry(handle, &g, &nc, &locked));
    return locked != 0;
  }
  long num_cell() const {
    mfg_geom_constants g;
    long nc = 0;
    int locked = 0;
    mfg_shim::check(mfg_grid_get_geometry(handle, &g, &nc, &locked));
    return nc;
  }

  void assign_hashes(particle_gpu *p) const { mfg_shim::check(mfg_grid_assign_hashes(handle, p->handle)); }
  void update_geometry(particle_gpu *p, double kernel_width = 2.0, double t_current = 0.) {
    mfg_shim::check(mfg_grid_update_geometry(handle, p->handle, kernel_width, t_current));
  }
  void sort(particle_gpu *p) { mfg_shim::check(mfg_grid_sort(handle, p->handle)); }
  // cell_start / cell_end: device arrays of num_cell() ints
  void get_cells(particle_gpu *p, int *cell_start, int *cell_end) {
    mfg_shim::check(mfg_grid_get_cells(handle, p->handle, cell_start, cell_end));
  }
  void set_bbox_vel(mfg_shim::vec3 v) {
    const double a[3] = {v.x, v.y, v.z};
    mfg_shim::check(mfg_grid_set_bbox_vel(handle, a));
  }
  void set_max_fixed(bool f) { mfg_shim::check(mfg_grid_set_max_fixed(handle, f ? 1 : 0)); }
  void set_min_fixed(bool f) { mfg_shim::check(mfg_grid_set_min_fixed(handle, f ? 1 : 0)); }
  void set_hard_blank(bool b) { mfg_shim::check(mfg_grid_set_hard_blank(handle, b ? 1 : 0)); }
  void adapt_particle_number(particle_gpu *p) const {
    mfg_shim::check(mfg_grid_adapt_particle_number(handle, p->handle));
    p->sync_counts();
  }

 protected:
  grid_base() = default;
  mfg_geom_constants geom() const {
    mfg_geom_constants g;
    long nc = 0;
    int locked = 0;
    mfg_shim::check(mfg_grid_get_geometry(handle, &g, &nc, &locked));
    return g;
  }
};

class grid_gpu_green : public grid_base {
 public:
  grid_gpu_green(unsigned int max_cell, unsigned int n) {
    mfg_shim::check(mfg_grid_create_unlocked(MFG_GRID_GREEN, (long)max_cell, (int)n, &handle));
  }
  grid_gpu_green(int num_part, mfg_shim::vec3 bbmin, mfg_shim::vec3 bbmax, double h) {
    const double a[3] = {bbmin.x, bbmin.y, bbmin.z}, b[3] = {bbmax.x, bbmax.y, bbmax.z};
    mfg_shim::check(mfg_grid_create_locked(MFG_GRID_GREEN, num_part, a, b, h, &handle));
  }
};

class grid_gpu_rothlin : public grid_base {
 public:
  grid_gpu_rothlin(unsigned int max_cell, unsigned int n) {
    mfg_shim::check(mfg_grid_create_unlocked(MFG_GRID_ROTHLIN, (long)max_cell, (int)n, &handle));
  }
  grid_gpu_rothlin(int num_part, mfg_shim::vec3 bbmin, mfg_shim::vec3 bbmax, double h) {
    const double a[3] = {bbmin.x, bbmin.y, bbmin.z}, b[3] = {bbmax.x, bbmax.y, bbmax.z};
    mfg_shim::check(mfg_grid_create_locked(MFG_GRID_ROTHLIN, num_part, a, b, h, &handle));
  }
};

// ---- tool_3d_gpu (src/tool_3d_gpu.h:74-176) -----------------------------------------------------
class tool_3d_gpu {
 public:
  mfg_tool *handle = nullptr;
  // tool_3d_gpu(const tool_3d*, vec3_t vel, phys_constants, float_t alpha): the tool_3d is given
  // as its indexed mesh (3 doubles per vertex, 3 ints per triangle)
  tool_3d_gpu(const std::vector<double> &verts3, const std::vector<int> &tris3, mfg_shim::vec3 vel,
              const mfg_phys_constants &phys, double alpha = 0.1, double h_avg = 0.) {
    const double v[3] = {vel.x, vel.y, vel.z};
    mfg_shim::check(mfg_tool_3d_gpu_create(verts3.data(), (int)(verts3.size() / 3), tris3.data(),
                                           (int)(tris3.size() / 3), h_avg, v, &phys, alpha,
                                           &handle));
  }
  tool_3d_gpu(const std::vector<double> &verts3, const std::vector<int> &tris3, double alpha = 0.1,
              double h_avg = 0.) {
    mfg_shim::check(mfg_tool_3d_gpu_create(verts3.data(), (int)(verts3.size() / 3), tris3.data(),
                                           (int)(tris3.size() / 3), h_avg, nullptr, nullptr,
                                           alpha, &handle));
  }
  ~tool_3d_gpu() { mfg_tool_destroy(handle); }
  tool_3d_gpu(const tool_3d_gpu &) = delete;
  tool_3d_gpu &operator=(const tool_3d_gpu &) = delete;

  void set_vel(mfg_shim::vec3 v) {
    const double a[3] = {v.x, v.y, v.z};
    mfg_shim::check(mfg_tool_set_vel(handle, a));
  }
  void set_mu(double mu) { mfg_shim::check(mfg_tool_set_mu(handle, mu)); }
  void set_contact_alpha(double a) { mfg_shim::check(mfg_tool_set_contact_alpha(handle, a)); }
  void set_algorithm_type(int algo) { mfg_shim::check(mfg_tool_set_algorithm_type(handle, algo)); }
  void set_birth_death(double b, double d) { mfg_shim::check(mfg_tool_set_birth_death(handle, b, d)); }
  void set_thermal(bool t) { mfg_shim::check(mfg_tool_set_thermal(handle, t ? 1 : 0)); }
  void set_physics(const mfg_phys_constants &p) { mfg_shim::check(mfg_tool_set_physics(handle, &p)); }
  mfg_shim::vec3 get_vel() const {
    double a[3];
    mfg_shim::check(mfg_tool_get_vel(handle, a));
    return {a[0], a[1], a[2]};
  }
  double get_mu() const {
    double mu = 0.;
    mfg_shim::check(mfg_tool_get_mu(handle, &mu));
    return mu;
  }
  bool is_alive(double t_current) const { return mfg_tool_is_alive(handle, t_current) == 1; }
  // update_tool / compute_contact_force (src/tool_3d_gpu.h:82,120), at global_time_current
  inline void update_tool();
  inline void compute_contact_force(particle_gpu *p, bool record_force = false);
};

// ---- tool_forces (src/tool_3d_gpu.h:52-68) ------------------------------------------------------
class tool_forces {
 public:
  mfg_tool_forces *handle = nullptr;
  unsigned int num_tool = 0;
  explicit tool_forces(unsigned int n) : num_tool(n) {
    mfg_shim::check(mfg_tool_forces_create(n, &handle));
  }
  ~tool_forces() { mfg_tool_forces_destroy(handle); }
  tool_forces(const tool_forces &) = delete;
  tool_forces &operator=(const tool_forces &) = delete;
  void reset() { mfg_shim::check(mfg_tool_forces_reset(handle)); }
  // 3 floats per tool, as get_forces_host returns float3 (src/tool_3d_gpu.cu:34-38)
  std::vector<float> get_forces_host() const {
    std::vector<float> f(3 * (size_t)num_tool);
    mfg_shim::check(mfg_tool_forces_get_host(handle, f.data()));
    return f;
  }
  // one line of "%.9g %.9g %.9g " per tool to `fp` (src/tool_3d_gpu.cu:53-69 writes
  // results/forces_new.txt); the reference also prints the forces x 1e7 when verbose
  void report(FILE *fp, bool verbose = false) const {
    const std::vector<float> f = get_forces_host();
    if (verbose) std::printf("Tool Forces:\n");
    for (size_t k = 0; k < f.size() / 3; k++) {
      std::fprintf(fp, "%.9g %.9g %.9g ", f[3 * k], f[3 * k + 1], f[3 * k + 2]);
      if (verbose) std::printf("%f %f %f ", f[3 * k] * 1e7, f[3 * k + 1] * 1e7, f[3 * k + 2] * 1e7);
    }
    std::fprintf(fp, "\n");
    if (verbose) std::printf("\n");
    std::fflush(fp);
  }
};

// ---- blanking (src/blanking.h:44-63, src/blanking.cpp:20-67), host only ------------------------
class blanking {
 public:
  blanking(mfg_shim::vec3 bbmin, mfg_shim::vec3 bbmax) : mn_(bbmin), mx_(bbmax), mn0_(bbmin), mx0_(bbmax) {}
  blanking(mfg_shim::vec3 bbmin, mfg_shim::vec3 bbmax, mfg_shim::vec3 vel)
      : mn_(bbmin), mx_(bbmax), mn0_(bbmin), mx0_(bbmax), vel_(vel) {}
  // stores max_vel unsquared, as the reference does
  explicit blanking(double max_vel) : max_vel2_(max_vel) {}
  blanking(mfg_shim::vec3 bbmin, mfg_shim::vec3 bbmax, double max_vel)
      : mn_(bbmin), mx_(bbmax), mn0_(bbmin), mx0_(bbmax), max_vel2_(max_vel * max_vel) {}
  blanking(mfg_shim::vec3 bbmin, mfg_shim::vec3 bbmax, mfg_shim::vec3 vel, double max_vel)
      : mn_(bbmin), mx_(bbmax), mn0_(bbmin), mx0_(bbmax), vel_(vel), max_vel2_(max_vel * max_vel) {}
  // the two setters are crossed over, as in the reference
  void set_max_fixed(bool f) { min_fixed_ = f; }
  void set_min_fixed(bool f) { max_fixed_ = f; }
  void update(double t_current) {
    if (!min_fixed_) mn_ = {mn0_.x + t_current * vel_.x, mn0_.y + t_current * vel_.y, mn0_.z + t_current * vel_.z};
    if (!max_fixed_) mx_ = {mx0_.x + t_current * vel_.x, mx0_.y + t_current * vel_.y, mx0_.z + t_current * vel_.z};
  }
  void get_bb(mfg_shim::vec3 &bbmin, mfg_shim::vec3 &bbmax) const {
    bbmin = mn_;
    bbmax = mx_;
  }
  double get_max_vel_squared() const { return max_vel2_; }
  mfg_blanking_state state() const {
    mfg_blanking_state s;
    s.bbmin[0] = mn_.x; s.bbmin[1] = mn_.y; s.bbmin[2] = mn_.z;
    s.bbmax[0] = mx_.x; s.bbmax[1] = mx_.y; s.bbmax[2] = mx_.z;
    s.max_vel_squared = max_vel2_;
    return s;
  }

 private:
  mfg_shim::vec3 mn_, mx_, mn0_, mx0_, vel_;
  double max_vel2_ = 1e15;
  bool min_fixed_ = false, max_fixed_ = false;
};

// ---- process globals (src/mfree_gpu_3d.cpp:41-48) ----------------------------------------------
extern std::vector<tool_3d_gpu *> global_tool;
extern tool_forces *global_tool_forces;
extern blanking *global_blanking;
extern double global_time_dt, global_time_final, global_time_current;
extern int global_time_step;

#ifdef MFG_DEFINE_GLOBALS
std::vector<tool_3d_gpu *> global_tool;
tool_forces *global_tool_forces = nullptr;
blanking *global_blanking = nullptr;
double global_time_dt = 0., global_time_final = 0., global_time_current = 0.;
int global_time_step = 0;
#endif

inline void tool_3d_gpu::update_tool() {
  mfg_shim::check(mfg_tool_update(handle, global_time_current, global_time_dt));
}
inline void tool_3d_gpu::compute_contact_force(particle_gpu *p, bool record_force) {
  int slot = 0;
  for (size_t k = 0; k < global_tool.size(); k++)
    if (global_tool[k] == this) slot = (int)k;
  const bool rec = record_force && global_tool_forces;
  mfg_shim::check(mfg_tool_compute_contact_force(handle, p->handle, rec ? 1 : 0,
                                                 rec ? global_tool_forces->handle : nullptr, slot,
                                                 global_time_dt, global_time_current));
}

// ---- actions and pair passes, one call each (src/actions_gpu.h:41-80,
// src/interactions_gpu.h:51-52); time steps from the globals, as the reference reads them.  The
// pair passes take the grid whose cell tables describe the current order (after g->sort).
inline void material_eos(particle_gpu *p) { mfg_shim::check(mfg_actions_material_eos(p->handle)); }
inline void corrector_artificial_stress(particle_gpu *p) {
  mfg_shim::check(mfg_actions_corrector_artificial_stress(p->handle));
}
inline void material_stress_rate_jaumann(particle_gpu *p) {
  mfg_shim::check(mfg_actions_material_stress_rate_jaumann(p->handle));
}
inline void material_fric_heat_gen(particle_gpu *p, mfg_shim::vec3 vel) {
  const double v[3] = {vel.x, vel.y, vel.z};
  mfg_shim::check(mfg_actions_material_fric_heat_gen(p->handle, v, global_time_dt));
}
inline void contmech_continuity(particle_gpu *p) {
  mfg_shim::check(mfg_actions_contmech_continuity(p->handle));
}
inline void contmech_momentum(particle_gpu *p) {
  mfg_shim::check(mfg_actions_contmech_momentum(p->handle));
}
inline void contmech_advection(particle_gpu *p) {
  mfg_shim::check(mfg_actions_contmech_advection(p->handle));
}
inline void plasticity_johnson_cook(particle_gpu *p) {
  mfg_shim::check(mfg_actions_plasticity_johnson_cook(p->handle, global_time_dt));
}
inline void perform_boundary_conditions_thermal(particle_gpu *p) {
  mfg_shim::check(mfg_actions_perform_boundary_conditions_thermal(p->handle));
}
inline void perform_boundary_conditions(particle_gpu *p) {
  mfg_shim::check(mfg_actions_perform_boundary_conditions(p->handle));
}
inline void actions_move_tool_particles(particle_gpu *p, tool_3d_gpu *t) {
  mfg_shim::check(mfg_actions_move_tool_particles(p->handle, t->handle, global_time_dt));
}
inline void debug_invalidate(particle_gpu *p) { mfg_shim::check(mfg_debug_invalidate(p->handle)); }
inline void perform_blanking(particle_gpu *p, blanking *b) {  // actions_gpu.h:57
  const mfg_blanking_state bs = b->state();
  mfg_shim::check(mfg_perform_blanking(p->handle, &bs));
}
inline void interactions_monaghan(particle_gpu *p, grid_base *g) {
  mfg_shim::check(mfg_interactions_monaghan(p->handle, g->handle));
}
inline void interactions_heat_pse(particle_gpu *p, grid_base *g) {
  mfg_shim::check(mfg_interactions_heat_pse(p->handle, g->handle));
}

// ---- leap_frog (src/leap_frog.h:37-50, src/leap_frog.cu:121-206) -------------------------------
class leap_frog {
 public:
  mfg_leap_frog *handle = nullptr;
  leap_frog(unsigned int num_part, unsigned int num_cell) {
    mfg_shim::check(mfg_leap_frog_create(num_part, num_cell, &handle));
  }
  ~leap_frog() { mfg_leap_frog_destroy(handle); }
  leap_frog(const leap_frog &) = delete;
  leap_frog &operator=(const leap_frog &) = delete;

  // one time step; reads the globals as the reference does
  void step(particle_gpu *p, grid_base *g, bool record_forces) {
    std::vector<mfg_tool *> tools;
    for (tool_3d_gpu *t : global_tool) tools.push_back(t->handle);
    mfg_blanking_state bs;
    const mfg_blanking_state *pb = nullptr;
    if (global_blanking) {
      global_blanking->update(global_time_current);
      bs = global_blanking->state();
      pb = &bs;
    }
    mfg_shim::check(mfg_leap_frog_step(handle, p->handle, g->handle, record_forces ? 1 : 0,
                                       tools.data(), (int)tools.size(), pb,
                                       global_tool_forces ? global_tool_forces->handle : nullptr,
                                       global_time_dt, global_time_current));
    p->sync_counts();
  }

  // the same step composed call by call from the per-function entry points, in the order of
  // src/leap_frog.cu:121-195 (the invalidation always on); same results, one launch per call
  void step_by_actions(particle_gpu *p, grid_base *g, bool record_forces) {
    if (global_blanking) {
      global_blanking->update(global_time_current);
      perform_blanking(p, global_blanking);
    }
    g->update_geometry(p, 2.0, global_time_current);
    g->assign_hashes(p);
    if (global_blanking) g->adapt_particle_number(p);
    g->sort(p);
    mfg_shim::check(mfg_leap_frog_init(handle, p->handle));
    mfg_shim::check(mfg_leap_frog_advance(handle, p->handle, 0.5 * global_time_dt));
    material_eos(p);
    corrector_artificial_stress(p);
    interactions_monaghan(p, g);
    mfg_options opt;
    mfg_shim::check(mfg_get_options(&opt));
    if (opt.heat == MFG_HEAT_PSE) interactions_heat_pse(p, g);
    material_stress_rate_jaumann(p);
    contmech_continuity(p);
    contmech_momentum(p);
    contmech_advection(p);
    debug_invalidate(p);
    mfg_shim::check(mfg_leap_frog_advance(handle, p->handle, global_time_dt));
    plasticity_johnson_cook(p);
    mfg_shim::check(mfg_leap_frog_reset_contact_forces(handle, p->handle));
    if (record_forces && global_tool_forces) global_tool_forces->reset();
    for (tool_3d_gpu *t : global_tool) {
      t->update_tool();
      t->compute_contact_force(p, record_forces);
    }
    if (!global_tool.empty()) material_fric_heat_gen(p, global_tool[0]->get_vel());
    perform_boundary_conditions(p);
    perform_boundary_conditions_thermal(p);
    if (!global_tool.empty()) actions_move_tool_particles(p, global_tool[0]);
    p->sync_counts();
  }
};

#endif  // MFREE_GPU_3D_HPP_
