// Mesh voxelisation for scene setup: the inside test of the editor's SDF compute pass
// (assets/shaders/sdf.comp:57-115, used by Editor::initFluid / initBoundaries,
// src/core/Editor.cpp:90-258).  A cell centre is inside when the generalized winding number of
// the triangle soup, the sum over triangles of the solid angle 2 atan2(a.(b x c),
// 1 + a.b + b.c + c.a) (a, b, c = unit vectors to the vertices, Van Oosterom and Strackee 1983),
// exceeds pi in magnitude, or when its distance to some triangle is exactly 0 (the pass signs
// the nearest-triangle distance with the winding number and calls "distance > 0" outside, so a
// centre lying on a triangle counts as inside).
//
// One thread per cell; the triangles stream through LDS in chunks shared by the workgroup.
#include <hip/hip_runtime.h>

#include "efs_internal.h"

namespace {

constexpr int VX_NT = 256, VX_CHUNK = 256;

__device__ inline float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float3 cross3(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ inline float3 unit3(float3 a) {
    const float inv = 1.0f / sqrtf(dot3(a, a));
    return make_float3(a.x * inv, a.y * inv, a.z * inv);
}

// solid angle of triangle (v0, v1, v2) seen from p (sdf.comp:53-67)
__device__ inline float solid_angle(float3 p, float3 v0, float3 v1, float3 v2) {
    const float3 a = unit3(sub3(v0, p)), b = unit3(sub3(v1, p)), c = unit3(sub3(v2, p));
    const float num = dot3(a, cross3(b, c));
    const float den = 1.0f + dot3(a, b) + dot3(b, c) + dot3(c, a);
    return 2.0f * atan2f(num, den);
}

__device__ inline float dot2(float3 a) { return dot3(a, a); }
__device__ inline float sgnf(float v) { return (v > 0.f) ? 1.f : ((v < 0.f) ? -1.f : 0.f); }
__device__ inline float clamp01(float v) { return fminf(fmaxf(v, 0.f), 1.f); }

// true when the unsigned point-triangle distance (sdf.comp:29-51) evaluates to exactly 0
__device__ inline bool on_triangle(float3 p, float3 v1, float3 v2, float3 v3) {
    const float3 v21 = sub3(v2, v1), v32 = sub3(v3, v2), v13 = sub3(v1, v3);
    const float3 p1 = sub3(p, v1), p2 = sub3(p, v2), p3 = sub3(p, v3);
    const float3 nor = cross3(v21, v13);
    const float s = sgnf(dot3(cross3(v21, nor), p1)) + sgnf(dot3(cross3(v32, nor), p2)) + sgnf(dot3(cross3(v13, nor), p3));
    if (s < 2.0f) {
        auto edge = [](float3 e, float3 q) {
            const float t = clamp01(dot3(e, q) / dot2(e));
            return dot2(make_float3(e.x * t - q.x, e.y * t - q.y, e.z * t - q.z));
        };
        return fminf(fminf(edge(v21, p1), edge(v32, p2)), edge(v13, p3)) == 0.f;
    }
    const float d = dot3(nor, p1);
    return d * d / dot2(nor) == 0.f;
}

// field[c] = inside ? v_in : v_out for the cells x < W-1, y < H-1, z < D-1 (sdf.comp:93-97
// leaves the last layer of each axis to the cleared buffer, i.e. v_out); cell centres at
// domain_min + cell_size * (i + 0.5) with cell_size = domain_size / (W, H, D).
__global__ __launch_bounds__(VX_NT) void k_voxelise(int W, int H, int D, const float* __restrict__ tris, int ntris, float3 dmin,
                                                    float3 csize, float v_in, float v_out, float* __restrict__ field) {
    __shared__ float st[VX_CHUNK * 9];
    const int n = W * H * D;
    const int c = blockIdx.x * VX_NT + threadIdx.x;
    const int x = c % W, y = (c / W) % H, z = c / (W * H);
    const bool active = c < n && x < W - 1 && y < H - 1 && z < D - 1;
    const float3 p = make_float3(dmin.x + csize.x * (float)x + csize.x * 0.5f, dmin.y + csize.y * (float)y + csize.y * 0.5f,
                                 dmin.z + csize.z * (float)z + csize.z * 0.5f);
    float w = 0.f;
    bool on = false;
    for (int t0 = 0; t0 < ntris; t0 += VX_CHUNK) {
        const int m = min(VX_CHUNK, ntris - t0);
        __syncthreads();
        for (int i = threadIdx.x; i < m * 9; i += VX_NT) st[i] = tris[(size_t)t0 * 9 + i];
        __syncthreads();
        if (active) {
            for (int t = 0; t < m; t++) {
                const float* q = st + 9 * t;
                const float3 v0 = make_float3(q[0], q[1], q[2]), v1 = make_float3(q[3], q[4], q[5]), v2 = make_float3(q[6], q[7], q[8]);
                w += solid_angle(p, v0, v1, v2);
                on = on || on_triangle(p, v0, v1, v2);
            }
        }
    }
    if (c < n) field[c] = (active && (fabsf(w) > 3.1415926f || on)) ? v_in : v_out;
}

}  // namespace

extern "C" int efs_scene_voxelise(efs_handle h, const float* triangles, int32_t ntris, const float domain_center[3],
                                  const float domain_size[3], int32_t field, float value_inside, float value_outside) {
    if (!h || (ntris > 0 && !triangles) || !domain_center || !domain_size || ntris < 0) {
        efs_set_error("efs_scene_voxelise: invalid argument");
        return -1;
    }
    if (field != EFS_FIELD_RHO && field != EFS_FIELD_BOUNDARY) {
        efs_set_error("efs_scene_voxelise: field must be RHO or BOUNDARY");
        return -1;
    }
    h->traces_valid = false;
    float* dtris = nullptr;
    if (ntris > 0) {
        EFS_HIP(hipMalloc(&dtris, (size_t)ntris * 9 * sizeof(float)));
        EFS_HIP(hipMemcpyAsync(dtris, triangles, (size_t)ntris * 9 * sizeof(float), hipMemcpyHostToDevice, h->stream));
    }
    const int W = h->W, H = h->H, D = h->D;
    const float3 csize = make_float3(domain_size[0] / (float)W, domain_size[1] / (float)H, domain_size[2] / (float)D);
    const float3 dmin = make_float3(domain_center[0] - 0.5f * domain_size[0], domain_center[1] - 0.5f * domain_size[1],
                                    domain_center[2] - 0.5f * domain_size[2]);
    float* out = field == EFS_FIELD_RHO ? h->rho : h->bndf;
    hipLaunchKernelGGL(k_voxelise, dim3((h->N + VX_NT - 1) / VX_NT), dim3(VX_NT), 0, h->stream, W, H, D, dtris, ntris, dmin, csize,
                       value_inside, value_outside, out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (dtris) hipFree(dtris);
    if (e != hipSuccess) return efs_hip_fail(e, "k_voxelise");
    if (field == EFS_FIELD_BOUNDARY) {
        const int rc = efs_launch_boundary_flags(h);
        if (rc) return rc;
        EFS_HIP(hipStreamSynchronize(h->stream));
    }
    return 0;
}
