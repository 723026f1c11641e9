// Self-check of the C++ adapter's host-mirror protocol (HipSimulator3D.h) on a GPU: after steps
// every mirror equals the device field; an edit made through an accessor (the way the reference's
// DeletionNode writes grid->rho(), DeletionNode.cpp:17-33) reaches the device before the next
// step and keeps every other cell; markModified before the first read after a step keeps the
// device values; fields nobody edited are never uploaded; a raw C-ABI call through handle()
// uploads pending edits first and invalidates the mirrors.  Prints MIRROR_CHECK_OK on success.
//   efs_mirror_check [n] [steps]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "HipSimulator3D.h"

static int g_fail = 0;
#define EXPECT(cond, ...)                       \
    do {                                        \
        if (!(cond)) {                          \
            std::printf("FAIL: " __VA_ARGS__);  \
            std::printf("\n");                  \
            g_fail++;                           \
        }                                       \
    } while (0)

static std::vector<float> raw(efs_handle h, int f, size_t n) {
    std::vector<float> v(n);
    efs::check(efs_download(h, f, v.data(), n * sizeof(float)), "efs_download");
    return v;
}

static size_t ndiff(const std::vector<float>& a, const std::vector<float>& b) {
    size_t k = 0;
    for (size_t i = 0; i < a.size(); i++) k += std::memcmp(&a[i], &b[i], 4) != 0;
    return k;
}

static double sum(const std::vector<float>& a) {
    double s = 0;
    for (float v : a) s += v;
    return s;
}

int main(int argc, char** argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 24;
    const int steps = argc > 2 ? std::atoi(argv[2]) : 3;
    efs_params p;
    efs_default_params(&p);
    p.dt = 1.f / 60.f;
    p.gravity_force = 9.81f * (float)n;
    p.max_steps = 500;
    try {
        efs::HipSimulator3D sim(n + 2, n + 2, n + 2, &p);
        auto& g = sim.grid();
        const size_t N = (size_t)g.size();
        const efs_prim dam{0, 0, {-4, -4, -4}, {0.4, 0.5, 4}};
        efs::check(efs_scene_rasterise(sim.handle(), &dam, 1), "efs_scene_rasterise");
        efs::check(efs_scene_face_noise(sim.handle(), 0, 1e-3), "efs_scene_face_noise");
        efs_handle h = sim.handle();  // kept for raw reads below, which change nothing
        for (int s = 0; s < steps; s++) EXPECT(sim.update() == 0, "step %d did not converge", s);

        // 1. after the steps every mirror equals the device field, bit for bit
        for (int f = 0; f < 8; f++) {
            const size_t len = N * ((f == EFS_FIELD_VEL || f == EFS_FIELD_FORCES) ? 3 : 1);
            const auto dev = raw(h, f, len);
            const auto& mir = g.view(f);
            EXPECT(mir.size() == len && ndiff(mir, dev) == 0, "mirror of field %d differs from the device", f);
            EXPECT(g.uploads(f) == 0, "field %d was uploaded without an edit (%ld)", f, g.uploads(f));
        }

        // 2. an edit through the accessor (no markModified) reaches the device and keeps the rest
        const size_t c = ((size_t)(n / 4) * (n + 2) + (n / 4)) * (n + 2) + (n / 4);  // inside the dam
        const auto before = raw(h, EFS_FIELD_RHO, N);
        g.rho()[c] = 0.25f;
        g.flush();
        const auto after = raw(h, EFS_FIELD_RHO, N);
        EXPECT(after[c] == 0.25f, "edited cell not uploaded (%g)", (double)after[c]);
        EXPECT(ndiff(before, after) == (before[c] != 0.25f ? 1u : 0u), "edit changed %zu cells", ndiff(before, after));
        EXPECT(g.uploads(EFS_FIELD_RHO) == 1, "rho uploads %ld", g.uploads(EFS_FIELD_RHO));
        // the edited mass enters the next step: rho_advect conserves it
        g.rho()[c] = 0.25f;  // pending again (no change of value): uploaded before the step
        const double m_edit = sum(g.view(EFS_FIELD_RHO));
        EXPECT(sim.update() == 0, "step after the edit");
        const double m_next = sum(g.view(EFS_FIELD_RHO));
        EXPECT(std::fabs(m_next - m_edit) < 1e-4 * m_edit, "mass %.6f after the edited step, %.6f edited", m_next, m_edit);
        EXPECT(std::fabs(m_edit - sum(before)) > 0.1, "the edit changed the mass by too little to be seen");
        for (int f = 0; f < 8; f++)
            if (f != EFS_FIELD_RHO) EXPECT(g.uploads(f) == 0, "field %d uploaded (%ld)", f, g.uploads(f));

        // 3. markModified before the first read after a step: the mirror is downloaded first, so a
        //    one-element edit keeps every other value
        const size_t nv = N * 3;
        const auto vdev = raw(h, EFS_FIELD_VEL, nv);
        g.markModified(EFS_FIELD_VEL);
        g.vel()[3 * c + 1] = 1.5f;
        g.flush();
        const auto vafter = raw(h, EFS_FIELD_VEL, nv);
        EXPECT(vafter[3 * c + 1] == 1.5f, "velocity edit not uploaded");
        EXPECT(ndiff(vdev, vafter) == (vdev[3 * c + 1] != 1.5f ? 1u : 0u), "velocity edit changed %zu values",
               ndiff(vdev, vafter));

        // 4. a raw call through handle() uploads the pending edit first, then the mirrors are stale
        g.rho()[c] = 0.75f;  // pending edit ...
        const efs_prim fill{0, 0, {-4, -4, -4}, {4, 0.25, 4}};
        efs::check(efs_scene_rasterise(sim.handle(), &fill, 1), "efs_scene_rasterise");  // ... then the raw call
        const auto rdev = raw(h, EFS_FIELD_RHO, N);
        EXPECT(ndiff(g.view(EFS_FIELD_RHO), rdev) == 0, "mirror not refreshed after a raw call");
        EXPECT(!g.modified(EFS_FIELD_RHO), "edit still pending after handle()");
    } catch (const efs::Error& e) {
        std::printf("FAIL: %s\n", e.what());
        return 1;
    }
    if (g_fail) return 1;
    std::printf("MIRROR_CHECK_OK\n");
    return 0;
}
