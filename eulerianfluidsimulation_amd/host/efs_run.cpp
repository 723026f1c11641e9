// Command-line driver of the HIP backend through the C++ host adapter: builds the dam-break
// scene of the given size (fluid x < 0.4, y < 0.5), runs steps and prints per-step statistics.
//   efs_run <n> <steps>
#include <cstdio>
#include <cstdlib>

#include "HipSimulator3D.h"

int main(int argc, char** argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 32;
    const int steps = argc > 2 ? std::atoi(argv[2]) : 5;
    efs_params p;
    efs_default_params(&p);
    p.dt = 1.f / 60.f;
    p.gravity_force = 9.81f * (float)n;
    p.max_steps = 500;
    try {
        efs::HipSimulator3D sim(n + 2, n + 2, n + 2, &p);
        efs_prim dam{0, 0, {-4, -4, -4}, {0.4, 0.5, 4}};
        efs::check(efs_scene_rasterise(sim.handle(), &dam, 1), "efs_scene_rasterise");
        efs::check(efs_scene_face_noise(sim.handle(), 0, 1e-3), "efs_scene_face_noise");
        double mass0 = 0;
        for (float r : sim.grid().rho()) mass0 += r;
        for (int s = 0; s < steps; s++) {
            const int rc = sim.update();
            double mass = 0;
            for (float r : sim.grid().rho()) mass += r;
            const efs_stats& st = sim.lastStats();
            std::printf("step %d rc=%d cg_iters=%d rr=%.3e mass=%.6f (%.6f)\n", s, rc, st.iterations, st.rr, mass, mass0);
        }
    } catch (const efs::Error& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
