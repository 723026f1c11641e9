// C ABI of the HIP backend (declared in include/efs.h): handle lifecycle, field transfer in the
// reference layout, and the step orchestration of ISimulator::update (ISimulator.h:12-38) with
// CPUSimulator3D::project's control flow (CPUSimulator3D.cpp:506-566).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>

#include "efs_internal.h"

static thread_local std::string g_err;

void efs_set_error(const std::string& msg) { g_err = msg; }
int efs_hip_fail(hipError_t e, const char* what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return -(1000 + (int)e);
}

int ensure_forces(efs_sim* s) {
    if (s->fu) return 0;
    const size_t b = (size_t)s->N * 4;
    EFS_HIP(hipMalloc(&s->fu, b));
    EFS_HIP(hipMalloc(&s->fv, b));
    EFS_HIP(hipMemsetAsync(s->fu, 0, b, s->stream));
    EFS_HIP(hipMemsetAsync(s->fv, 0, b, s->stream));
    if (s->dim == 3) {
        EFS_HIP(hipMalloc(&s->fw, b));
        EFS_HIP(hipMemsetAsync(s->fw, 0, b, s->stream));
    }
    return 0;
}

namespace {

// AoS float[3] <-> three SoA arrays
__global__ void k_deinterleave(int n, const float* __restrict__ aos, float* __restrict__ a, float* __restrict__ b,
                               float* __restrict__ c) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    a[i] = aos[3 * (size_t)i];
    b[i] = aos[3 * (size_t)i + 1];
    c[i] = aos[3 * (size_t)i + 2];
}
__global__ void k_interleave(int n, const float* __restrict__ a, const float* __restrict__ b, const float* __restrict__ c,
                             float* __restrict__ aos) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    aos[3 * (size_t)i] = a[i];
    aos[3 * (size_t)i + 1] = b[i];
    aos[3 * (size_t)i + 2] = c[i];
}
// AoS float[2] (2D velocity / forces) <-> two SoA arrays
__global__ void k_deinterleave2(int n, const float* __restrict__ aos, float* __restrict__ a, float* __restrict__ b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    a[i] = aos[2 * (size_t)i];
    b[i] = aos[2 * (size_t)i + 1];
}
__global__ void k_interleave2(int n, const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ aos) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    aos[2 * (size_t)i] = a[i];
    aos[2 * (size_t)i + 1] = b[i];
}
// The face velocities of cell i as (u, v, w, 0); w = 0 in 2D (the renderer's vec4 layout)
__global__ void k_pack_vel4(int n, const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                            float4* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float4(u[i], v[i], w ? w[i] : 0.f, 0.f);
}
__global__ void k_fill_f32(int n, float* a, float v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

inline int blocks1d(int n) { return (n + 255) / 256; }

// Seeded primitives at interior cell centres ((i - 0.5) h per axis, in double), in order: a solid
// primitive sets boundary 1 and rho 0, a fluid one rho 1 on non-solid cells; the border is solid
// with rho 0 (Grid3D.cpp:28-42, Editor.cpp:246-257).  One thread per cell.
constexpr int RAST_BATCH = 32;
struct PrimBatch {
    int n;
    efs_prim p[RAST_BATCH];
};
__global__ void k_rasterise(int W, int H, int D, int is3, double hx, double hy, double hz, PrimBatch pb, float* __restrict__ rho,
                            float* __restrict__ bnd) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= W * H * D) return;
    const int x = c % W, y = (c / W) % H, z = c / (W * H);
    float r = rho[c], b = bnd[c];
    if (x == 0 || y == 0 || x == W - 1 || y == H - 1 || (is3 && (z == 0 || z == D - 1))) {
        b = 1.f;
        r = 0.f;
    } else {
        const double cx = (x - 0.5) * hx, cy = (y - 0.5) * hy, cz = is3 ? (z - 0.5) * hz : 0.0;
        for (int k = 0; k < pb.n; k++) {
            const efs_prim& P = pb.p[k];
            bool in;
            if (P.kind == 0) {
                in = cx >= P.a[0] && cx < P.b[0] && cy >= P.a[1] && cy < P.b[1] && (!is3 || (cz >= P.a[2] && cz < P.b[2]));
            } else {
                const double dx = cx - P.a[0], dy = cy - P.a[1], dz = is3 ? cz - P.a[2] : 0.0;
                in = dx * dx + dy * dy + dz * dz <= P.b[0] * P.b[0];
            }
            if (!in) continue;
            if (P.solid) {
                b = 1.f;
                r = 0.f;
            } else if (b != 1.f) {
                r = 1.f;
            }
        }
    }
    rho[c] = r;
    bnd[c] = b;
}

bool valid_handle(efs_sim* s) {
    if (!s) {
        efs_set_error("null handle");
        return false;
    }
    return true;
}

size_t field_bytes(const efs_sim* s, int f) {
    if (f == EFS_FIELD_VEL || f == EFS_FIELD_FORCES) return (size_t)s->N * (s->dim == 2 ? 8 : 12);
    return (size_t)s->N * 4;
}

float* scalar_ptr(efs_sim* s, int f) {
    switch (f) {
        case EFS_FIELD_RHO: return s->rho;
        case EFS_FIELD_GAMMA: return s->gamma;
        case EFS_FIELD_BETA: return s->beta;
        case EFS_FIELD_DIVERGENCE: return s->div;
        case EFS_FIELD_PRESSURE: return s->p;
        case EFS_FIELD_BOUNDARY: return s->bndf;
    }
    return nullptr;
}

// Moves `bytes` of field f between the library and a caller buffer of the given memcpy kind.
int transfer(efs_sim* s, int f, void* other, size_t bytes, bool upload, hipMemcpyKind kind) {
    if (!valid_handle(s)) return -1;
    if (upload) s->traces_valid = false;
    if (f < 0 || f > EFS_FIELD_BOUNDARY) {
        efs_set_error("unknown field id");
        return -1;
    }
    if (!other) {
        efs_set_error("null buffer");
        return -1;
    }
    if (bytes != field_bytes(s, f)) {
        efs_set_error("byte count does not match the field size");
        return -1;
    }
    if (f == EFS_FIELD_VEL || f == EFS_FIELD_FORCES) {
        if (f == EFS_FIELD_FORCES) {
            int rc = ensure_forces(s);
            if (rc) return rc;
        }
        float* a = f == EFS_FIELD_VEL ? s->u : s->fu;
        float* b = f == EFS_FIELD_VEL ? s->v : s->fv;
        float* c = f == EFS_FIELD_VEL ? s->w : s->fw;
        if (s->dim == 2) {
            if (upload) {
                EFS_HIP(hipMemcpyAsync(s->stage_bytes, other, bytes, kind, s->stream));
                hipLaunchKernelGGL(k_deinterleave2, dim3(blocks1d(s->N)), dim3(256), 0, s->stream, s->N, s->stage_bytes, a, b);
                EFS_HIP(hipGetLastError());
            } else {
                hipLaunchKernelGGL(k_interleave2, dim3(blocks1d(s->N)), dim3(256), 0, s->stream, s->N, a, b, s->stage_bytes);
                EFS_HIP(hipGetLastError());
                EFS_HIP(hipMemcpyAsync(other, s->stage_bytes, bytes, kind, s->stream));
            }
        } else if (upload) {
            EFS_HIP(hipMemcpyAsync(s->stage_bytes, other, bytes, kind, s->stream));
            hipLaunchKernelGGL(k_deinterleave, dim3(blocks1d(s->N)), dim3(256), 0, s->stream, s->N, s->stage_bytes, a, b, c);
            EFS_HIP(hipGetLastError());
        } else {
            hipLaunchKernelGGL(k_interleave, dim3(blocks1d(s->N)), dim3(256), 0, s->stream, s->N, a, b, c, s->stage_bytes);
            EFS_HIP(hipGetLastError());
            EFS_HIP(hipMemcpyAsync(other, s->stage_bytes, bytes, kind, s->stream));
        }
    } else {
        float* dst = scalar_ptr(s, f);
        if (upload) EFS_HIP(hipMemcpyAsync(dst, other, bytes, kind, s->stream));
        else EFS_HIP(hipMemcpyAsync(other, dst, bytes, kind, s->stream));
        if (upload && f == EFS_FIELD_BOUNDARY) {
            int rc = efs_launch_boundary_flags(s);
            if (rc) return rc;
        }
    }
    if (kind == hipMemcpyHostToDevice || kind == hipMemcpyDeviceToHost) EFS_HIP(hipStreamSynchronize(s->stream));
    return 0;
}

void free_all(efs_sim* s) {
    float* fl[] = {s->u, s->v, s->w, s->fu, s->fv, s->fw, s->rho, s->gamma, s->beta, s->div, s->p, s->bndf,
                   s->tu, s->tv, s->tw, s->s0, s->s1, s->bwd_w, s->fwd_w, s->stage_bytes, s->render4};
    for (float* p : fl)
        if (p) hipFree(p);
    if (s->bf) hipFree(s->bf);
    if (s->valA) hipFree(s->valA);
    if (s->valB) hipFree(s->valB);
    if (s->bwd_p0) hipFree(s->bwd_p0);
    if (s->fwd_p0) hipFree(s->fwd_p0);
    if (s->acc0) hipFree(s->acc0);
    if (s->acc1) hipFree(s->acc1);
    if (s->acc2) hipFree(s->acc2);
    efs_learned_free(s);
    efs_mg_free(s);
    efs_solver_free(s);
    for (auto& t : s->ktimers) {
        for (auto e : t.start) hipEventDestroy(e);
        for (auto e : t.stop) hipEventDestroy(e);
    }
    for (auto e : s->ev_stage)
        if (e) hipEventDestroy(e);
    if (s->own_stream && s->stream) hipStreamDestroy(s->stream);
}

int alloc_all(efs_sim* s) {
    const size_t b = (size_t)s->N * 4;
    float** fl[] = {&s->u, &s->v, &s->w, &s->rho, &s->gamma, &s->beta, &s->div, &s->p, &s->bndf,
                    &s->tu, &s->tv, &s->tw, &s->s0, &s->s1};
    for (float** p : fl) {
        EFS_HIP(hipMalloc(p, b));
        EFS_HIP(hipMemset(*p, 0, b));
    }
    EFS_HIP(hipMalloc(&s->stage_bytes, b * 3));
    EFS_HIP(hipMalloc(&s->bf, (size_t)s->N));
    EFS_HIP(hipMalloc(&s->valA, (size_t)s->N));
    EFS_HIP(hipMalloc(&s->valB, (size_t)s->N));
    EFS_HIP(hipMemset(s->valA, 0, (size_t)s->N));
    EFS_HIP(hipMemset(s->valB, 0, (size_t)s->N));
    EFS_HIP(hipMalloc(&s->bwd_p0, b));
    EFS_HIP(hipMalloc(&s->fwd_p0, b));
    EFS_HIP(hipMemset(s->bwd_p0, 0xFF, b));
    EFS_HIP(hipMemset(s->fwd_p0, 0xFF, b));
    EFS_HIP(hipMalloc(&s->bwd_w, b * 8));
    EFS_HIP(hipMalloc(&s->fwd_w, b * 8));
    EFS_HIP(hipMalloc(&s->acc0, (size_t)s->N * 8));
    EFS_HIP(hipMalloc(&s->acc1, (size_t)s->N * 8));
    EFS_HIP(hipMalloc(&s->acc2, (size_t)s->N * 8));
    EFS_HIP(hipMemset(s->acc0, 0, (size_t)s->N * 8));
    EFS_HIP(hipMemset(s->acc1, 0, (size_t)s->N * 8));
    EFS_HIP(hipMemset(s->acc2, 0, (size_t)s->N * 8));
    return efs_solver_alloc(s);
}

// Grid3D / Grid2D ctors (Grid3D.cpp:10-48, Grid2D.cpp:11-60): gamma = 1, border solid with rho 0.
int init_grid(efs_sim* s) {
    hipLaunchKernelGGL(k_fill_f32, dim3(blocks1d(s->N)), dim3(256), 0, s->stream, s->N, s->gamma, 1.f);
    std::vector<float> bnd((size_t)s->N, 0.f);
    const bool is3 = s->dim == 3;
    for (int z = 0; z < s->D; z++)
        for (int y = 0; y < s->H; y++)
            for (int x = 0; x < s->W; x++)
                if (x == 0 || y == 0 || x == s->W - 1 || y == s->H - 1 || (is3 && (z == 0 || z == s->D - 1)))
                    bnd[((size_t)z * s->H + y) * s->W + x] = 1.f;
    EFS_HIP(hipMemcpyAsync(s->bndf, bnd.data(), bnd.size() * 4, hipMemcpyHostToDevice, s->stream));
    int rc = efs_launch_boundary_flags(s);
    if (rc) return rc;
    EFS_HIP(hipStreamSynchronize(s->stream));
    return 0;
}

void stage_event(efs_sim* s, int k) {
    if (!s->timing) return;
    if (!s->ev_stage[k]) hipEventCreate(&s->ev_stage[k]);
    hipEventRecord(s->ev_stage[k], s->stream);
}

// The pressure solve ISolver3D::solve dispatches to (CPUSimulator3D.cpp:18-40 picks the solver
// from SOLVER_TYPE): Jacobi, CG, or CG preconditioned by multigrid (the reference's MULTIGRID id,
// Properties.h:481, which no reference backend implements) or by the learned network.
int solve_3d(efs_sim* s, efs_stats* out) {
    if (s->P.solver_type == EFS_SOLVER_JACOBI) return efs_jacobi_solve(s, out);
    if (s->P.solver_type == EFS_SOLVER_MULTIGRID) return efs_mg_cg_solve(s, out);
    if (s->P.precond == EFS_PRECOND_LEARNED) return efs_learned_cg_solve(s, out);
    if (s->P.precond == EFS_PRECOND_MG) return efs_mg_cg_solve(s, out);
    return efs_cg_solve(s, out);
}

// The solvers a backend offers, in the order ISimulator::availableSolvers takes them
// (ISimulator.cpp:5-30: an unavailable SOLVER_TYPE falls back to the first one).
int available_solvers(int dim, int32_t* out) {
    out[0] = EFS_SOLVER_SELF_CG;
    out[1] = EFS_SOLVER_JACOBI;
    if (dim == 2) return 2;
    out[2] = EFS_SOLVER_MULTIGRID;
    return 3;
}

bool check_params(int dim, const efs_params* p) {
    int32_t av[4];
    const int n = available_solvers(dim, av);
    bool ok = false;
    for (int k = 0; k < n; k++) ok |= p->solver_type == av[k];
    if (!ok) {
        efs_set_error(dim == 2 ? "solver_type must be SELF_CG (2) or JACOBI (3) in 2D"
                               : "solver_type must be SELF_CG (2), JACOBI (3) or MULTIGRID (4)");
        return false;
    }
    if (dim == 2 && p->precond != EFS_PRECOND_NONE) {
        efs_set_error("the CG preconditioners are 3D only");
        return false;
    }
    return true;
}

int run_project(efs_sim* s, efs_stats* st) {
    int rc;
    // pressure.clear(): Jacobi starts from it; the CG solves write every cell of p when they end
    // (efs_cg_finalize), so for them the clear would be overwritten unread
    const bool cg = s->P.solver_type == EFS_SOLVER_MULTIGRID ||
                    (s->P.solver_type != EFS_SOLVER_JACOBI && s->P.precond != EFS_PRECOND_LEARNED);
    if (!cg) EFS_HIP(hipMemsetAsync(s->p, 0, (size_t)s->N * 4, s->stream));
    if ((rc = efs_launch_walls_divergence(s))) return rc;
    efs_stats local{};
    efs_stats* out = st ? st : &local;
    const int status = solve_3d(s, out);
    if (status < 0) return status;
    if (status == 1) return 1;  // not converged: velocity keeps the pre-solve state
    if ((rc = efs_launch_gradient_walls(s))) return rc;
    return status;
}

// `fuse`: in a full 3D step, vel_decay runs inside vel_extrapolate and add_force inside vel_advect
// when both stages of the pair are enabled (the same operations in the same order; the second
// stage of each pair is then skipped by the caller).
int run_stage(efs_sim* s, int stage, efs_stats* st, bool fuse = false) {
    if (s->dim == 2) return efs2_run_stage(s, stage, st);
    // every stage but the velocity advection that consumes them invalidates the stored traces
    // (rho_advect stores fresh ones)
    if (stage != EFS_STAGE_VEL_ADVECT) s->traces_valid = false;
    switch (stage) {
        case EFS_STAGE_VEL_DECAY: return efs_launch_decay(s);
        case EFS_STAGE_VEL_EXTRAPOLATE: return efs_launch_extrapolate(s, fuse);
        case EFS_STAGE_RHO_SHARPEN: return efs_launch_sharpen(s);
        case EFS_STAGE_RHO_ADVECT: return efs_launch_rho_advect(s);
        case EFS_STAGE_VEL_ADVECT: return efs_launch_vel_advect(s, fuse);
        case EFS_STAGE_ADD_FORCE: return efs_launch_add_force(s);
        case EFS_STAGE_PROJECT: return run_project(s, st);
    }
    efs_set_error("unknown stage id");
    return -1;
}

// counter-based generator shared with the seeded scenes
inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline double uniform01(uint64_t seed, uint64_t stream, uint64_t ctr) {
    const uint64_t key = mix64(seed) ^ (stream * 0xD1B54A32D192ED03ull);
    return (double)(mix64(key + ctr * 0x9E3779B97F4A7C15ull) >> 11) * 0x1.0p-53;
}
inline double normal01(uint64_t seed, uint64_t stream, uint64_t ctr) {
    const double u1 = uniform01(seed, stream, 2 * ctr);
    const double u2 = uniform01(seed, stream, 2 * ctr + 1);
    return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
}

}  // namespace

extern "C" {

const char* efs_last_error(void) { return g_err.c_str(); }
const char* efs_version(void) { return "efs-hip gfx950 1.0"; }

void efs_default_params(efs_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->dt = 0.0167f;
    p->gravity_force = 10.f;
    p->lambda_minus = 0.5f;
    p->lambda_plus = 0.1f;
    p->eta = 1.f;
    p->liquid_threshold = 0.5f;
    p->rho_d = 3.f;
    p->vel_decay = 0.8f;
    p->gravity_enabled = 1;
    p->solver_type = EFS_SOLVER_SELF_CG;
    p->max_steps = 2000;
    p->solver_threshold = 1e-5f;
    p->disable_mask = 0;
    p->criterion = EFS_CRIT_REFERENCE;
    p->precond = EFS_PRECOND_NONE;
    p->true_residual = 0;
    p->initial_guess = EFS_GUESS_DIVERGENCE;
}

static int create(int dim, int32_t width, int32_t height, int32_t depth, const efs_params* params, efs_handle* out) {
    if (!out) {
        efs_set_error("null output handle");
        return -1;
    }
    *out = nullptr;
    if (width < 3 || height < 3 || (dim == 3 && depth < 3)) {
        efs_set_error("grid must be at least 3 cells per axis (one interior cell plus the border)");
        return -1;
    }
    const long long n = (long long)width * height * depth;
    if (n >= (1ll << 31) / 8) {
        efs_set_error("grid too large for 32-bit cell indexing of the 8-entry advection arrays");
        return -1;
    }
    if (params && !check_params(dim, params)) return -1;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        efs_set_error("no HIP device available");
        return -3;
    }
    efs_sim* s = new efs_sim();
    s->dim = dim;
    s->W = width; s->H = height; s->D = depth; s->N = (int)n;
    if (params) s->P = *params;
    else efs_default_params(&s->P);
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete s;
        return efs_hip_fail(e, "hipStreamCreateWithFlags");
    }
    s->own_stream = true;
    int rc = dim == 2 ? efs2_alloc(s) : alloc_all(s);
    if (rc == 0) rc = init_grid(s);
    if (rc) {
        free_all(s);
        delete s;
        return rc;
    }
    *out = s;
    return 0;
}

int efs_create_3d(int32_t width, int32_t height, int32_t depth, const efs_params* params, efs_handle* out) {
    return create(3, width, height, depth, params, out);
}

int efs_create_2d(int32_t width, int32_t height, const efs_params* params, efs_handle* out) {
    return create(2, width, height, 1, params, out);
}

int efs_destroy(efs_handle h) {
    if (!h) return 0;
    hipStreamSynchronize(h->stream);
    free_all(h);
    delete h;
    return 0;
}

int efs_set_params(efs_handle h, const efs_params* p) {
    if (!valid_handle(h) || !p) return -1;
    if (!check_params(h->dim, p)) return -1;
    h->P = *p;
    h->traces_valid = false;
    return 0;
}
int efs_available_solvers(int32_t dim, int32_t* solvers, int32_t capacity) {
    if (dim != 2 && dim != 3) {
        efs_set_error("dim must be 2 or 3");
        return -1;
    }
    int32_t av[4];
    const int n = available_solvers(dim, av);
    if (solvers)
        for (int k = 0; k < n && k < capacity; k++) solvers[k] = av[k];
    return n;
}
int efs_get_params(efs_handle h, efs_params* p) {
    if (!valid_handle(h) || !p) return -1;
    *p = h->P;
    return 0;
}
int efs_dims(efs_handle h, int32_t dims[3]) {
    if (!valid_handle(h)) return -1;
    dims[0] = h->W; dims[1] = h->H; dims[2] = h->D;
    return 0;
}

int efs_upload(efs_handle h, int32_t f, const void* host, size_t bytes) {
    return transfer(h, f, const_cast<void*>(host), bytes, true, hipMemcpyHostToDevice);
}
int efs_download(efs_handle h, int32_t f, void* host, size_t bytes) {
    return transfer(h, f, host, bytes, false, hipMemcpyDeviceToHost);
}
int efs_upload_device(efs_handle h, int32_t f, const void* dev, size_t bytes) {
    return transfer(h, f, const_cast<void*>(dev), bytes, true, hipMemcpyDeviceToDevice);
}
int efs_download_device(efs_handle h, int32_t f, void* dev, size_t bytes) {
    return transfer(h, f, dev, bytes, false, hipMemcpyDeviceToDevice);
}

// The buffers MarchingCubes::updateMesh feeds its shader (MarchingCubes.cpp:75-105): the density
// (N floats) and the velocity as vec4 (N x 16 bytes), packed on the device.
static int render_feed(efs_sim* s, float* rho, float* vel4, hipMemcpyKind kind) {
    if (!valid_handle(s)) return -1;
    const size_t n = (size_t)s->N;
    if (rho) EFS_HIP(hipMemcpyAsync(rho, s->rho, n * 4, kind, s->stream));
    if (vel4) {
        float4* dst = reinterpret_cast<float4*>(vel4);
        if (kind != hipMemcpyDeviceToDevice) {
            if (!s->render4) EFS_HIP(hipMalloc(&s->render4, n * 16));
            dst = reinterpret_cast<float4*>(s->render4);
        }
        hipLaunchKernelGGL(k_pack_vel4, dim3(blocks1d(s->N)), dim3(256), 0, s->stream, s->N, s->u, s->v,
                           s->dim == 3 ? (const float*)s->w : nullptr, dst);
        EFS_HIP(hipGetLastError());
        if (kind != hipMemcpyDeviceToDevice) EFS_HIP(hipMemcpyAsync(vel4, s->render4, n * 16, kind, s->stream));
    }
    EFS_HIP(hipStreamSynchronize(s->stream));
    return 0;
}
int efs_render_feed(efs_handle h, float* rho, float* vel4) { return render_feed(h, rho, vel4, hipMemcpyDeviceToHost); }
int efs_render_feed_device(efs_handle h, float* rho, float* vel4) {
    return render_feed(h, rho, vel4, hipMemcpyDeviceToDevice);
}

int efs_stage(efs_handle h, int32_t stage, efs_stats* st) {
    if (!valid_handle(h)) return -1;
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->true_rel_residual = -1;
    }
    const int rc = run_stage(h, stage, st);
    if (rc < 0) return rc;
    EFS_HIP(hipStreamSynchronize(h->stream));
    if (st) st->status = rc;
    return rc;
}

int efs_step(efs_handle h, efs_stats* st) {
    if (!valid_handle(h)) return -1;
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->true_rel_residual = -1;
    }
    int status = 0;
    if (h->ktiming_every > 0) h->ktiming = (h->ktiming_steps++ % h->ktiming_every) == 0;
    stage_event(h, 0);
    const unsigned on = ~h->P.disable_mask;
    auto enabled = [&](int k) { return (on >> k) & 1u; };
    const bool fuse_decay = h->dim == 3 && enabled(EFS_STAGE_VEL_DECAY) && enabled(EFS_STAGE_VEL_EXTRAPOLATE);
    const bool fuse_force = h->dim == 3 && enabled(EFS_STAGE_VEL_ADVECT) && enabled(EFS_STAGE_ADD_FORCE);
    for (int k = 0; k <= EFS_STAGE_PROJECT; k++) {
        const bool folded = (k == EFS_STAGE_VEL_DECAY && fuse_decay) || (k == EFS_STAGE_ADD_FORCE && fuse_force);
        if (enabled(k) && !folded) {
            const bool fuse = (k == EFS_STAGE_VEL_EXTRAPOLATE && fuse_decay) || (k == EFS_STAGE_VEL_ADVECT && fuse_force);
            const int rc = run_stage(h, k, st, fuse);
            if (rc < 0) return rc;
            if (k == EFS_STAGE_PROJECT) status = rc;
        }
        stage_event(h, k + 1);
    }
    if (h->timing && st) {
        EFS_HIP(hipEventSynchronize(h->ev_stage[7]));
        for (int k = 0; k < 7; k++) {
            float ms = 0;
            hipEventElapsedTime(&ms, h->ev_stage[k], h->ev_stage[k + 1]);
            st->stage_ms[k] = ms;
        }
    }
    if (st) st->status = status;
    return status;
}

int efs_solve_pressure(efs_handle h, efs_stats* st) {
    if (!valid_handle(h)) return -1;
    efs_stats local{};
    efs_stats* out = st ? st : &local;
    std::memset(out, 0, sizeof(*out));
    out->true_rel_residual = -1;
    EFS_HIP(hipMemsetAsync(h->p, 0, (size_t)h->N * 4, h->stream));
    int rc;
    if (h->dim == 2) rc = efs2_solve(h, out);
    else rc = solve_3d(h, out);
    if (rc < 0) return rc;
    EFS_HIP(hipStreamSynchronize(h->stream));
    out->status = rc;
    return rc;
}

int efs_project_divergence(efs_handle h) {
    if (!valid_handle(h)) return -1;
    h->traces_valid = false;  // the wall enforcement changes the velocity
    const int rc = h->dim == 2 ? efs2_walls_divergence(h) : efs_launch_walls_divergence(h);
    if (rc) return rc;
    EFS_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

int efs_apply_operator(efs_handle h, const float* x, float* y) {
    if (!valid_handle(h) || !x || !y) return -1;
    const int rc = h->dim == 2 ? efs2_spmv(h, x, y) : efs_launch_spmv(h, x, y);
    if (rc) return rc;
    EFS_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

int efs_set_stream(efs_handle h, void* stream) {
    if (!valid_handle(h)) return -1;
    EFS_HIP(hipStreamSynchronize(h->stream));
    if (h->own_stream) hipStreamDestroy(h->stream);
    h->own_stream = false;
    h->stream = (hipStream_t)stream;
    return 0;
}
void* efs_get_stream(efs_handle h) { return h ? (void*)h->stream : nullptr; }
int efs_synchronize(efs_handle h) {
    if (!valid_handle(h)) return -1;
    EFS_HIP(hipStreamSynchronize(h->stream));
    return 0;
}
int efs_set_timing(efs_handle h, int32_t on) {
    if (!valid_handle(h)) return -1;
    h->timing = on != 0;
    return 0;
}
int efs_set_kernel_timing(efs_handle h, int32_t on) {
    if (!valid_handle(h)) return -1;
    EFS_HIP(hipStreamSynchronize(h->stream));
    for (auto& t : h->ktimers) {
        for (auto e : t.start) hipEventDestroy(e);
        for (auto e : t.stop) hipEventDestroy(e);
        t.start.clear();
        t.stop.clear();
        t.launches = 0;
        t.total_ms = 0;
    }
    h->ktiming = on != 0;
    h->ktiming_every = on > 1 ? on : 0;
    h->ktiming_steps = 0;
    return 0;
}
int efs_kernel_time(efs_handle h, int32_t k, int64_t* launches, double* total_ms) {
    if (!valid_handle(h) || k < 0 || k >= KT_COUNT) return -1;
    EFS_HIP(hipStreamSynchronize(h->stream));
    TimerSlot& t = h->ktimers[k];
    for (size_t i = 0; i < t.start.size(); i++) {
        float ms = 0;
        hipEventElapsedTime(&ms, t.start[i], t.stop[i]);
        t.total_ms += ms;
        hipEventDestroy(t.start[i]);
        hipEventDestroy(t.stop[i]);
    }
    t.start.clear();
    t.stop.clear();
    if (launches) *launches = t.launches;
    if (total_ms) *total_ms = t.total_ms;
    return 0;
}

int efs_scene_rasterise(efs_handle h, const efs_prim* prims, int32_t nprims) {
    if (!valid_handle(h) || nprims < 0 || (nprims > 0 && !prims)) return -1;
    h->traces_valid = false;
    const bool is3 = h->dim == 3;
    const double hx = 1.0 / (h->W - 2), hy = 1.0 / (h->H - 2), hz = is3 ? 1.0 / (h->D - 2) : 1.0;
    // primitives go by value in batches; later batches see the cells earlier ones wrote, which is
    // the host order of one loop over all primitives
    for (int k0 = 0; k0 == 0 || k0 < nprims; k0 += RAST_BATCH) {
        PrimBatch pb{};
        pb.n = std::min(RAST_BATCH, nprims - k0);
        for (int k = 0; k < pb.n; k++) pb.p[k] = prims[k0 + k];
        hipLaunchKernelGGL(k_rasterise, dim3(blocks1d(h->N)), dim3(256), 0, h->stream, h->W, h->H, h->D, is3 ? 1 : 0, hx, hy, hz,
                           pb, h->rho, h->bndf);
        EFS_HIP(hipGetLastError());
    }
    int rc = efs_launch_boundary_flags(h);
    if (rc) return rc;
    EFS_HIP(hipStreamSynchronize(h->stream));
    return 0;
}

int efs_scene_face_noise(efs_handle h, uint64_t seed, double sigma) {
    if (!valid_handle(h)) return -1;
    const int W = h->W, H = h->H, D = h->D;
    const int nc = h->dim == 3 ? 3 : 2;  // components per cell
    std::vector<float> bnd((size_t)h->N), vel((size_t)h->N * nc);
    EFS_HIP(hipMemcpyAsync(bnd.data(), h->bndf, bnd.size() * 4, hipMemcpyDeviceToHost, h->stream));
    EFS_HIP(hipStreamSynchronize(h->stream));
    int rc = efs_download(h, EFS_FIELD_VEL, vel.data(), vel.size() * 4);
    if (rc) return rc;
    // every face is independent (counter = cell * nc + component), so host threads split the
    // planes (in 2D the rows) between them; the values do not depend on the split
    const bool three = nc == 3;
    const int lo = 1, hi = three ? D - 1 : H - 1;
    auto work = [&](int a0, int a1) {
        for (int a = a0; a < a1; a++) {
            const int zb = three ? a : 0, ze = three ? a + 1 : 1, yb = three ? 1 : a, ye = three ? H - 1 : a + 1;
            for (int z = zb; z < ze; z++)
                for (int y = yb; y < ye; y++)
                    for (int x = 1; x < W - 1; x++) {
                        const size_t c = ((size_t)z * H + y) * W + x;
                        if (bnd[c] == 1.f) continue;
                        for (int k = 0; k < nc; k++) {
                            const size_t nb = k == 0 ? c - 1 : (k == 1 ? c - (size_t)W : c - (size_t)W * H);
                            if (bnd[nb] == 1.f) continue;
                            vel[c * nc + k] += (float)(sigma * normal01(seed, 1, c * nc + k));
                        }
                    }
        }
    };
    const int span = std::max(0, hi - lo);
    const int nt = (int)std::max(1u, std::min(16u, std::min(std::thread::hardware_concurrency(), (unsigned)std::max(1, span / 8))));
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++) {
        const int a0 = lo + (int)((long long)span * t / nt), a1 = lo + (int)((long long)span * (t + 1) / nt);
        pool.emplace_back(work, a0, a1);
    }
    for (auto& th : pool) th.join();
    return efs_upload(h, EFS_FIELD_VEL, vel.data(), vel.size() * 4);
}

}  // extern "C"
