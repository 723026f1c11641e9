// Command-line driver of the HIP backend through the C++ host adapter: builds a dam-break scene
// of the given size (3D: fluid x < 0.4, y < 0.5; 2D: x < 0.4, y < 0.6), runs steps and prints
// per-step statistics.
//   efs_run <n> <steps> [2d] [mg]
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "HipSimulator3D.h"

template <class Sim>
int run(Sim& sim, int steps, const efs_prim& dam) {
    efs::check(efs_scene_rasterise(sim.handle(), &dam, 1), "efs_scene_rasterise");
    efs::check(efs_scene_face_noise(sim.handle(), 0, 1e-3), "efs_scene_face_noise");
    double mass0 = 0;
    for (float r : sim.grid().view(EFS_FIELD_RHO)) mass0 += r;
    for (int s = 0; s < steps; s++) {
        const int rc = sim.update();
        double mass = 0;
        for (float r : sim.grid().view(EFS_FIELD_RHO)) mass += r;
        const efs_stats& st = sim.lastStats();
        std::printf("step %d rc=%d cg_iters=%d rr=%.3e mass=%.6f (%.6f)\n", s, rc, st.iterations, st.rr, mass, mass0);
    }
    return 0;
}

int main(int argc, char** argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 32;
    const int steps = argc > 2 ? std::atoi(argv[2]) : 5;
    bool two_d = false, mg = false;
    for (int a = 3; a < argc; a++) {
        if (!std::strcmp(argv[a], "2d")) two_d = true;
        if (!std::strcmp(argv[a], "mg")) mg = true;
    }
    efs_params p;
    efs_default_params(&p);
    p.dt = two_d ? 1.f / 120.f : 1.f / 60.f;
    p.gravity_force = 9.81f * (float)n;
    p.max_steps = two_d ? 2000 : 500;
    if (mg) p.precond = EFS_PRECOND_MG;
    try {
        if (two_d) {
            efs::HipSimulator2D sim(n + 2, n + 2, &p);
            return run(sim, steps, efs_prim{0, 0, {-4, -4, -4}, {0.4, 0.6, 4}});
        }
        efs::HipSimulator3D sim(n + 2, n + 2, n + 2, &p);
        return run(sim, steps, efs_prim{0, 0, {-4, -4, -4}, {0.4, 0.5, 4}});
    } catch (const efs::Error& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
}
