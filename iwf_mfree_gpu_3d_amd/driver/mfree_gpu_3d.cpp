// mfree_gpu_3d.cpp -- host driver of the orthogonal cut in the shape of the reference's main
// (src/mfree_gpu_3d.cpp:75-205): build a workload, set the constants and globals as the
// reference's setup_* functions do (src/Benchmarks/benchmarks_single_grain.cu:674-876), then call
// leap_frog::step with time advancing by dt after every call, reporting progress and tool forces.
//
//   mfree_gpu_3d [preset=1] [scale=1.0] [steps=preset's] [precision=0] [seed=1] [by_actions=0]
//
// preset 1: 8,000-particle wedge cut; 2: 1M-particle single grain; 3: 8M three grains.
// precision: 0 FP64, 1 FP32 + Kahan sums, 2 FP32 + FP64 sums.  by_actions 1: every step composed
// call by call from the actions, pair passes and contact calls, as the reference's leap_frog::step
// makes them (same results).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#define MFG_DEFINE_GLOBALS
#include "mfree_gpu_3d.hpp"

using mfg_shim::check;

int main(int argc, char **argv) {
  const int preset = argc > 1 ? std::atoi(argv[1]) : 1;
  const double scale = argc > 2 ? std::atof(argv[2]) : 1.0;
  const int steps_arg = argc > 3 ? std::atoi(argv[3]) : 0;
  const int precision = argc > 4 ? std::atoi(argv[4]) : MFG_FP64;
  const unsigned long long seed = argc > 5 ? std::strtoull(argv[5], nullptr, 10) : 1ull;
  const bool by_actions = argc > 6 && std::atoi(argv[6]) != 0;
  try {
    check(mfg_init(0));  // cudaFree(0) in the reference

    mfg_cut_config cfg;
    check(mfg_scene_preset(preset, scale, seed, &cfg));
    if (steps_arg > 0) cfg.steps = steps_arg;
    mfg_scene *scene = mfg_scene_build(&cfg);
    if (!scene) {
      std::fprintf(stderr, "scene build failed: %s\n", mfg_last_error());
      return 1;
    }
    mfg_scene_params sp;
    check(mfg_scene_get_params(scene, &sp));
    const unsigned n = sp.n_particles;

    // host arrays in the reference's layout (float4_t pos / vel)
    std::vector<double> pos3(3 * (size_t)n), vel3(3 * (size_t)n), rho(n), T(n), h(n), fixed(n), tool_p(n);
    check(mfg_scene_get_particles(scene, pos3.data(), vel3.data(), rho.data(), T.data(), h.data(),
                                  fixed.data(), tool_p.data()));
    std::vector<double> pos4(4 * (size_t)n, 0.), vel4(4 * (size_t)n, 0.);
    for (size_t i = 0; i < n; i++)
      for (int k = 0; k < 3; k++) {
        pos4[4 * i + k] = pos3[3 * i + k];
        vel4[4 * i + k] = vel3[3 * i + k];
      }

    // switches and constants
    mfg_options opt;
    check(mfg_get_options(&opt));
    opt.cspm = sp.cspm;
    opt.heat = sp.heat;
    opt.single_sqrt = 1;
    check(mfg_set_options(&opt));
    check(mfg_actions_setup_corrector_constants(&sp.corr));
    check(mfg_actions_setup_physical_constants(&sp.phys));
    check(mfg_actions_setup_johnson_cook_constants(&sp.joco));
    check(mfg_actions_setup_thermal_constants_wp(&sp.trml_wp));
    check(mfg_interactions_setup_corrector_constants(&sp.corr));
    check(mfg_interactions_setup_physical_constants(&sp.phys));
    check(mfg_interactions_setup_thermal_constants_workpiece(&sp.trml_wp));

    // tools
    std::vector<std::unique_ptr<tool_3d_gpu>> tools;
    for (int k = 0; k < sp.n_tools; k++) {
      int nv = 0, nt = 0;
      check(mfg_scene_tool_size(scene, k, &nv, &nt));
      std::vector<double> verts(3 * (size_t)nv);
      std::vector<int> tris(3 * (size_t)nt);
      check(mfg_scene_tool_mesh(scene, k, verts.data(), tris.data()));
      tools.emplace_back(new tool_3d_gpu(verts, tris, {sp.tool_vel[0], sp.tool_vel[1], sp.tool_vel[2]},
                                         sp.phys, sp.contact_alpha));
      tools.back()->set_algorithm_type(sp.algo);
      tools.back()->set_mu(sp.mu);
      global_tool.push_back(tools.back().get());
    }
    if (sp.trml_tool.alpha != 0.) {
      check(mfg_actions_setup_thermal_constants_tool(&sp.trml_tool));
      check(mfg_interactions_setup_thermal_constants_tool(&sp.trml_tool,
                                                          tools.empty() ? nullptr : tools[0]->handle));
    }
    mfg_scene_free(scene);

    global_time_dt = sp.dt;
    global_time_final = sp.dt * sp.steps;
    particle_gpu particles(pos4.data(), vel4.data(), rho.data(), T.data(), h.data(), fixed.data(),
                           tool_p.data(), n, precision);
    grid_gpu_green grid((int)n, {sp.bbmin[0], sp.bbmin[1], sp.bbmin[2]},
                        {sp.bbmax[0], sp.bbmax[1], sp.bbmax[2]}, sp.h);
    blanking blank({sp.bbmin[0], sp.bbmin[1], sp.bbmin[2]}, {sp.bbmax[0], sp.bbmax[1], sp.bbmax[2]},
                   sp.max_vel);
    global_blanking = &blank;
    std::unique_ptr<tool_forces> forces;
    if (!tools.empty()) {
      forces.reset(new tool_forces((unsigned)tools.size()));
      global_tool_forces = forces.get();
    }
    leap_frog stepper(particles.N, (unsigned)grid.num_cell());

    std::printf("preset %d scale %.3g: %u particles, %d tools, dt %.4g us, %d steps, precision %d\n",
                preset, scale, n, sp.n_tools, sp.dt, sp.steps, precision);

    // the reference's loop: iteration 0 does not step, time advances every iteration
    const int last_step = sp.steps + 1;
    const int freq = last_step / 10 > 0 ? last_step / 10 : 1;
    check(mfg_synchronize());
    const auto t0 = std::chrono::steady_clock::now();
    double active = 0.;
    for (int step = 0; step < last_step; step++) {
      const bool record = step % freq == 0;
      if (step > 0) {
        if (by_actions)
          stepper.step_by_actions(&particles, &grid, record && global_tool_forces != nullptr);
        else
          stepper.step(&particles, &grid, record && global_tool_forces != nullptr);
        active += particles.N;
      }
      if (record && step > 0) {
        check(mfg_synchronize());
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("%06d of %06d: N %u, %.3f s", step, last_step - 1, particles.N, secs);
        if (global_tool_forces) {
          const std::vector<float> f = global_tool_forces->get_forces_host();
          for (size_t k = 0; k < f.size() / 3; k++)
            std::printf("  F%zu %.9g %.9g %.9g", k, f[3 * k], f[3 * k + 1], f[3 * k + 2]);
        }
        std::printf("\n");
        std::fflush(stdout);
      }
      global_time_step++;
      global_time_current += global_time_dt;
    }
    check(mfg_synchronize());
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    check(mfg_debug_check_valid(particles.handle));
    std::printf("done: %d steps in %.3f s, %.4g particle-steps/s (including host reporting)\n",
                last_step - 1, secs, active / secs);
  } catch (const mfg_shim::error &e) {
    std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
    return 2;
  }
  return 0;
}
