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
//
// The same inside test gives the cells of a force field (Editor::updateForceFields with
// assets/shaders/forcefield_cells.comp), and a box reduction gives QueryNode's per-frame
// statistics without downloading the fields.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <mutex>
#include <vector>

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
                                                    float3 csize, float v_in, float v_out, float* __restrict__ field,
                                                    unsigned int* __restrict__ count) {
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
    const bool inside = active && (fabsf(w) > 3.1415926f || on);
    if (c < n) field[c] = inside ? v_in : v_out;
    if (count && inside) atomicAdd(count, 1u);
}

// forces = f on the cells whose mask is set (the force-field cells, Editor::updateForceFields)
__global__ void k_apply_force(int n, const float* __restrict__ mask, float fx, float fy, float fz, float* __restrict__ fu,
                              float* __restrict__ fv, float* __restrict__ fw) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n || mask[c] != 1.f) return;
    fu[c] = fx;
    fv[c] = fy;
    fw[c] = fz;
}

// QueryNode::update's box statistics (QueryNode.cpp:58-76): per block, over the cells of the
// box [lo, hi) (clipped to [0, n-1)) with rho >= 0.5 when only_fluid: count, min / max cell
// coordinates, and FP64 sums of the coordinates, the face velocities and the pressure.
constexpr int QY_NT = 256, QY_MAXB = 512;
struct QueryPart {
    double sum[7];  // x, y, z, u, v, w, p
    int cnt, mn[3], mx[3], pad;
};
__global__ __launch_bounds__(QY_NT) void k_query_box(DevGrid g, int3 lo, int3 hi, int only_fluid, QueryPart* parts) {
    const int ex = hi.x - lo.x, ey = hi.y - lo.y, ez = hi.z - lo.z;
    const long long n = (long long)ex * ey * ez;
    double sum[7] = {0, 0, 0, 0, 0, 0, 0};
    int cnt = 0, mn[3] = {INT_MAX, INT_MAX, INT_MAX}, mx[3] = {INT_MIN, INT_MIN, INT_MIN};
    for (long long t = (long long)blockIdx.x * QY_NT + threadIdx.x; t < n; t += (long long)gridDim.x * QY_NT) {
        const int x = lo.x + (int)(t % ex), y = lo.y + (int)((t / ex) % ey), z = lo.z + (int)(t / ((long long)ex * ey));
        const int c = cidx(g, x, y, z);
        if (only_fluid && g.rho[c] < 0.5f) continue;
        cnt++;
        const int xyz[3] = {x, y, z};
        for (int k = 0; k < 3; k++) {
            mn[k] = min(mn[k], xyz[k]);
            mx[k] = max(mx[k], xyz[k]);
            sum[k] += (double)xyz[k];
        }
        sum[3] += (double)g.u[c];
        sum[4] += (double)g.v[c];
        sum[5] += g.w ? (double)g.w[c] : 0.0;
        sum[6] += (double)g.p[c];
    }
    __shared__ QueryPart sh[QY_NT / 64];
    // wave reduction, then the waves in order
    for (int o = 32; o > 0; o >>= 1) {
        for (int k = 0; k < 7; k++) sum[k] += __shfl_down(sum[k], o, 64);
        cnt += __shfl_down(cnt, o, 64);
        for (int k = 0; k < 3; k++) {
            mn[k] = min(mn[k], __shfl_down(mn[k], o, 64));
            mx[k] = max(mx[k], __shfl_down(mx[k], o, 64));
        }
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        for (int k = 0; k < 7; k++) sh[wv].sum[k] = sum[k];
        sh[wv].cnt = cnt;
        for (int k = 0; k < 3; k++) { sh[wv].mn[k] = mn[k]; sh[wv].mx[k] = mx[k]; }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        QueryPart q = sh[0];
        for (int w = 1; w < QY_NT / 64; w++) {
            for (int k = 0; k < 7; k++) q.sum[k] += sh[w].sum[k];
            q.cnt += sh[w].cnt;
            for (int k = 0; k < 3; k++) { q.mn[k] = min(q.mn[k], sh[w].mn[k]); q.mx[k] = max(q.mx[k], sh[w].mx[k]); }
        }
        parts[blockIdx.x] = q;
    }
}

// Sums the block partials in block order into *out, the same additions in the same order as the
// host loop of efs_query_box (one thread: at most QY_MAXB partials).
__global__ void k_query_fold(const QueryPart* __restrict__ parts, int nb, QueryPart* __restrict__ out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    QueryPart f;
    for (int k = 0; k < 7; k++) f.sum[k] = 0.0;
    long long cnt = 0;
    for (int k = 0; k < 3; k++) { f.mn[k] = INT_MAX; f.mx[k] = INT_MIN; }
    for (int q = 0; q < nb; q++) {
        for (int k = 0; k < 7; k++) f.sum[k] += parts[q].sum[k];
        cnt += parts[q].cnt;
        for (int k = 0; k < 3; k++) { f.mn[k] = min(f.mn[k], parts[q].mn[k]); f.mx[k] = max(f.mx[k], parts[q].mx[k]); }
    }
    f.cnt = (int)cnt;
    f.pad = 0;
    *out = f;
}

}  // namespace

// Scratch of efs_query_box_noalloc: QY_MAXB block partials and the folded result, one copy per
// GPU, held until the process ends.  Host code never passes this name to a kernel: it is only a
// host-side handle of the symbol, resolved per device by hipGetSymbolAddress below.
static __device__ __align__(256) QueryPart scratch[QY_MAXB + 1];

namespace {

constexpr int QY_MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
QueryPart* g_scratch_ptr[QY_MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up once per device.
int query_scratch(QueryPart** out) {
    int dev = 0;
    EFS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= QY_MAX_DEVICES) {
        efs_set_error("efs_query_box_noalloc: device index out of range");
        return -1;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return efs_hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = (QueryPart*)ptr;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// Clips the box to [max(0, lo), min(n - 1, hi)) per axis as the reference's loops do; returns the
// cell count.
long long query_clip(const efs_sim* h, const int32_t lo_in[3], const int32_t hi_in[3], int lo[3], int hi[3]) {
    const int dims[3] = {h->W, h->H, h->D};
    for (int k = 0; k < 3; k++) {
        lo[k] = std::max(0, lo_in[k]);
        hi[k] = std::max(lo[k], std::min(dims[k] - 1, hi_in[k]));
    }
    if (h->dim == 2) {  // the 2D grid is one plane
        lo[2] = 0;
        hi[2] = 1;
    }
    return (long long)(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
}

// Means from the totals (QueryNode.cpp:82-84: with no cell every value stays 0).
void query_finish(const double sum[7], long long cnt, const int mn[3], const int mx[3], efs_query_result* out) {
    out->count = (int32_t)cnt;
    if (cnt == 0) return;
    for (int k = 0; k < 3; k++) {
        out->max_pos[k] = (float)mx[k];
        out->min_pos[k] = (float)mn[k];
        out->avg_pos[k] = (float)(sum[k] / (double)cnt);
        out->vel[k] = (float)(sum[3 + k] / (double)cnt);
    }
    out->pressure = (float)(sum[6] / (double)cnt);
}

}  // namespace

namespace {
inline float3 f3(const float* v) { return make_float3(v[0], v[1], v[2]); }
}

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
                       value_inside, value_outside, out, (unsigned int*)nullptr);
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

// Editor::updateForceFields for one force field (src/core/Editor.cpp:260-330 and
// assets/shaders/forcefield_cells.comp): the cells inside the field's mesh get force.  A field
// covering more than max_cells cells is skipped, as the reference does (max_cells <= 0: no
// limit); *ncells receives the count either way.
extern "C" int efs_scene_force_region(efs_handle h, const float* triangles, int32_t ntris, const float domain_center[3],
                                      const float domain_size[3], const float force[3], int32_t max_cells, int32_t* ncells) {
    if (!h || (ntris > 0 && !triangles) || !domain_center || !domain_size || !force || ntris < 0) {
        efs_set_error("efs_scene_force_region: invalid argument");
        return -1;
    }
    if (h->dim != 3) {
        efs_set_error("efs_scene_force_region: 3D only");
        return -1;
    }
    if (int rc = ensure_forces(h)) return rc;
    float* dtris = nullptr;
    unsigned int* dcount = nullptr;
    EFS_HIP(hipMalloc(&dcount, sizeof(unsigned int)));
    EFS_HIP(hipMemsetAsync(dcount, 0, sizeof(unsigned int), h->stream));
    if (ntris > 0) {
        EFS_HIP(hipMalloc(&dtris, (size_t)ntris * 9 * sizeof(float)));
        EFS_HIP(hipMemcpyAsync(dtris, triangles, (size_t)ntris * 9 * sizeof(float), hipMemcpyHostToDevice, h->stream));
    }
    const int W = h->W, H = h->H, D = h->D;
    const float3 csize = make_float3(domain_size[0] / (float)W, domain_size[1] / (float)H, domain_size[2] / (float)D);
    const float3 dmin = make_float3(domain_center[0] - 0.5f * domain_size[0], domain_center[1] - 0.5f * domain_size[1],
                                    domain_center[2] - 0.5f * domain_size[2]);
    // the mask goes to the stage scratch s0 (1 inside, 0 outside)
    hipLaunchKernelGGL(k_voxelise, dim3((h->N + VX_NT - 1) / VX_NT), dim3(VX_NT), 0, h->stream, W, H, D, dtris, ntris, dmin, csize,
                       1.f, 0.f, h->s0, dcount);
    unsigned int cnt = 0;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&cnt, dcount, sizeof(cnt), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess && (max_cells <= 0 || (int64_t)cnt <= max_cells)) {
        hipLaunchKernelGGL(k_apply_force, dim3((h->N + 255) / 256), dim3(256), 0, h->stream, h->N, h->s0, force[0], force[1], force[2],
                           h->fu, h->fv, h->fw);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    }
    if (dtris) hipFree(dtris);
    hipFree(dcount);
    if (e != hipSuccess) return efs_hip_fail(e, "efs_scene_force_region");
    if (ncells) *ncells = (int32_t)cnt;
    h->traces_valid = false;
    return 0;
}

// QueryNode::update's statistics of one box (src/simulation/scene/nodes/QueryNode.cpp:45-84),
// reduced on the device (FP64 sums; the reference sums in FP32 in loop order).
extern "C" int efs_query_box(efs_handle h, const int32_t lo_in[3], const int32_t hi_in[3], int32_t only_fluid, efs_query_result* out) {
    if (!h || !lo_in || !hi_in || !out) {
        efs_set_error("efs_query_box: invalid argument");
        return -1;
    }
    int lo[3], hi[3];
    const long long n = query_clip(h, lo_in, hi_in, lo, hi);
    std::memset(out, 0, sizeof(*out));
    if (n == 0) return 0;
    const int nb = (int)std::min<long long>(QY_MAXB, (n + QY_NT - 1) / QY_NT);
    QueryPart* dparts = nullptr;
    EFS_HIP(hipMalloc(&dparts, (size_t)nb * sizeof(QueryPart)));
    hipLaunchKernelGGL(k_query_box, dim3(nb), dim3(QY_NT), 0, h->stream, h->grid(), make_int3(lo[0], lo[1], lo[2]),
                       make_int3(hi[0], hi[1], hi[2]), only_fluid, dparts);
    std::vector<QueryPart> parts((size_t)nb);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(parts.data(), dparts, (size_t)nb * sizeof(QueryPart), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    hipFree(dparts);
    if (e != hipSuccess) return efs_hip_fail(e, "efs_query_box");
    double sum[7] = {0, 0, 0, 0, 0, 0, 0};
    long long cnt = 0;
    int mn[3] = {INT_MAX, INT_MAX, INT_MAX}, mx[3] = {INT_MIN, INT_MIN, INT_MIN};
    for (const QueryPart& q : parts) {
        for (int k = 0; k < 7; k++) sum[k] += q.sum[k];
        cnt += q.cnt;
        for (int k = 0; k < 3; k++) { mn[k] = std::min(mn[k], q.mn[k]); mx[k] = std::max(mx[k], q.mx[k]); }
    }
    query_finish(sum, cnt, mn, mx, out);
    return 0;
}

// efs_query_box without allocation: the block partials and their fold live in the library's
// static device scratch (QY_MAXB + 1 partials, 45 KB per GPU).  The grid-stride kernel needs at
// most QY_MAXB blocks for any box, so only boxes whose cell count does not fit the int32 count are
// refused.  Calls that overlap on one GPU share the scratch and race.
extern "C" int efs_query_box_noalloc(efs_handle h, const int32_t lo_in[3], const int32_t hi_in[3], int32_t only_fluid,
                                     efs_query_result* out) {
    if (!h || !lo_in || !hi_in || !out) {
        efs_set_error("efs_query_box_noalloc: invalid argument");
        return -1;
    }
    int lo[3], hi[3];
    const long long n = query_clip(h, lo_in, hi_in, lo, hi);
    std::memset(out, 0, sizeof(*out));
    if (n > (long long)INT_MAX) {
        efs_set_error("efs_query_box_noalloc: the box holds more than 2^31 - 1 cells");
        return -1;
    }
    if (n == 0) return 0;
    const int nb = (int)std::min<long long>(QY_MAXB, (n + QY_NT - 1) / QY_NT);
    if (nb > QY_MAXB) {  // cannot happen with the clamp above; the scratch holds QY_MAXB partials
        efs_set_error("efs_query_box_noalloc: more block partials than the scratch holds");
        return -1;
    }
    QueryPart* ptr = nullptr;
    if (int rc = query_scratch(&ptr)) return rc;
    // every launched block writes its own partial and the fold overwrites the last slot, so
    // nothing left by an earlier call is read
    hipLaunchKernelGGL(k_query_box, dim3(nb), dim3(QY_NT), 0, h->stream, h->grid(), make_int3(lo[0], lo[1], lo[2]),
                       make_int3(hi[0], hi[1], hi[2]), only_fluid, ptr);
    hipLaunchKernelGGL(k_query_fold, dim3(1), dim3(1), 0, h->stream, ptr, nb, ptr + QY_MAXB);
    QueryPart fin;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&fin, ptr + QY_MAXB, sizeof(QueryPart), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return efs_hip_fail(e, "efs_query_box_noalloc");
    query_finish(fin.sum, (long long)fin.cnt, fin.mn, fin.mx, out);
    return 0;
}
