// Internal declarations of the HIP backend: device field layout, launch helpers and the
// device functions shared by the stage kernels (trace, interpolation, corner weights).
//
// Device layout (HBM): structure-of-arrays FP32, cell (x,y,z) at z*W*H + y*W + x, the
// reference's order (ScalField3D.h:19-25).  The three face-velocity components live in three
// arrays (u = x-face at the cell's low-x side, v, w), so every stencil reads unit-stride rows.
// Boundary classification is folded once per upload into a byte per cell (BF_*), so the
// kernels never re-test the float boundary value.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/efs.h"

// boundary byte: the three ways the reference tests the float boundary value
#define BF_EQ1 1u   // boundary == 1.f       (BOUNDARY tests in CPUSimulator3D.cpp, CGSolver.cpp:471)
#define BF_GT05 2u  // boundary  > 0.5f      (Grid3D.cpp:79, CGSolver.cpp:484, scatter_values)
#define BF_GE05 4u  // boundary >= 0.5f      (JacobiSolver.cpp:172)

// solver row code (u16 per cell), rebuilt from rho and the boundary byte before each solve
#define RC_NB_XM 0x01u
#define RC_NB_XP 0x02u
#define RC_NB_YM 0x04u
#define RC_NB_YP 0x08u
#define RC_NB_ZP 0x10u
#define RC_NB_ZM 0x20u
#define RC_CNT_SHIFT 8
#define RC_CG_ROW 0x8000u
#define RC_JAC_ROW 0x4000u
// the cell is inside the counted region (not on the outer layer, boundary <= 0.5) and fluid:
// exactly the cells whose neighbour bit is set in their neighbours' codes
#define RC_CONTRIB 0x40u

// fixed-point accumulation of scattered float contributions: deterministic regardless of the
// order the atomics land in.  2^36 gives 1.5e-11 resolution and a range of +-1.3e8.
#define FIX_SCALE 68719476736.0
#define FIX_INV (1.0 / 68719476736.0)

struct DevGrid {
    int W, H, D;
    int N;
    float *u, *v, *w;
    float *fu, *fv, *fw;
    float *rho, *gamma, *beta, *div, *p;
    const uint8_t* bf;
};

__host__ __device__ inline int cidx(const DevGrid& g, int x, int y, int z) { return (z * g.H + y) * g.W + x; }

template <class T> __device__ inline T clampv(T v, T lo, T hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }

constexpr int kBufDword3 = 0x00020000;  // gfx9 buffer descriptor word 3 for raw access

// Workgroup i runs on XCD i mod 8 (the hardware deals workgroups round-robin over the 8 XCDs,
// each with its own L2).  The velocity face pass and the pressure gradient, streaming stencils
// that read the -z neighbour, take their block from a 1D grid in which every XCD owns a
// contiguous range of a plane's block columns and walks it plane by plane, so that the plane
// before and most neighbouring block rows were read through the same L2 (gradient 167 -> 118 us,
// face pass 108 -> 87 us at 256^3).  The traces, the row codes and the other per-cell passes
// measured slower so, and so did a division-free variant that gives each XCD whole block rows.
// Blocks past the last column of a range exit at once.  The grid is 2D, NXCD * per blocks along x
// and one row of blocks per plane: the hardware numbers them x fastest, so block x of row z is
// dealt to XCD x mod NXCD exactly as in a 1D grid of planes, and the plane needs no division.
constexpr int NXCD = 8;
struct XcdBlock {
    int bx, by, bz;
};
// q = n / d, r = n % d for 0 <= n < 2^22, d >= 1: a float estimate off by at most one, corrected
// (a runtime integer division costs about thirty scalar instructions)
__device__ inline void divmod_small(int n, int d, int& q, int& r) {
    q = (int)((float)n * __builtin_amdgcn_rcpf((float)d));
    r = n - q * d;
    if (r < 0) {
        q--;
        r += d;
    } else if (r >= d) {
        q++;
        r -= d;
    }
}
__device__ inline bool xcd_block(int gx, int gy, XcdBlock& b) {
    const int ncol = gx * gy, per = (ncol + NXCD - 1) / NXCD;
    const int col = ((int)blockIdx.x % NXCD) * per + (int)blockIdx.x / NXCD;
    b.bz = blockIdx.y;
    divmod_small(col, gx, b.by, b.bx);
    return col < ncol;
}
inline dim3 xcd_grid(int gx, int gy, int nz) { return dim3(NXCD * ((gx * gy + NXCD - 1) / NXCD), nz); }

__device__ inline long long to_fix(float v) { return (long long)rint((double)v * FIX_SCALE); }
// The same integer in three instructions instead of eight when |v| < 2^15: v * 2^36 + 1.5 * 2^52 is
// one rounding of an exact sum (the product is exact), at unit spacing in [2^52, 2^53), ties to
// even, so it rounds v * 2^36 as rint does, and the low bits of the sum are that integer.
__device__ inline long long to_fix_small(float v) {
    const double d = __builtin_fma((double)v, FIX_SCALE, 6755399441055744.0);  // 1.5 * 2^52
    return __double_as_longlong(d) - 0x4338000000000000LL;
}
__device__ inline bool fix_small(float v) { return fabsf(v) < 32768.f; }  // false for NaN
// to_fix of N values, through to_fix_small when every value of every active lane of the wave
// allows it (a wave-uniform choice, so the wave runs one of the two sequences)
template <int N>
__device__ inline void to_fix_n(const float (&v)[N], long long (&q)[N]) {
    bool small = true;
#pragma unroll
    for (int k = 0; k < N; k++) small &= fix_small(v[k]);
    if (__all(small)) {
#pragma unroll
        for (int k = 0; k < N; k++) q[k] = to_fix_small(v[k]);
    } else {
#pragma unroll
        for (int k = 0; k < N; k++) q[k] = to_fix(v[k]);
    }
}
__device__ inline void fix_add(unsigned long long* acc, int i, float v) {
    const long long q = to_fix(v);
    if (q != 0) atomicAdd(acc + i, (unsigned long long)q);
}
__device__ inline double from_fix(unsigned long long a) { return (double)(long long)a * FIX_INV; }

// ---------------------------------------------------------------------------------------
// VecField3D::interpolate on the staggered grid (VecField3D.cpp:20-86).  Low indices are
// clamped to [0, n-2]; for the finite positions the traces produce ([1, n-2] - 0.5) the clamp
// never binds, it only keeps non-finite velocities inside the arrays.
// ---------------------------------------------------------------------------------------
// The reference adds 1e-10f to the floor before truncating; floorf(v) is an integer and 1e-10f is
// below half an ulp of every nonzero integer, so the sum truncates to the same int (0 -> 1e-10
// -> 0) and the addition is left out.
__device__ inline int lowIndex(float coord, float offset) { return (int)floorf(coord - offset); }

__device__ inline float trilin(const DevGrid& g, const float* __restrict__ f, float px, float py, float pz,
                               float ox, float oy, float oz) {
    const int lx = clampv(lowIndex(px, ox), 0, g.W - 2);
    const int ly = clampv(lowIndex(py, oy), 0, g.H - 2);
    const int lz = clampv(lowIndex(pz, oz), 0, g.D - 2);
    const float x0 = (float)lx + ox, y0 = (float)ly + oy, z0 = (float)lz + oz;
    const float tx = px - x0, ty = py - y0, tz = pz - z0;
    const int b = cidx(g, lx, ly, lz);
    const int sy = g.W, sz = g.W * g.H;
    const float v000 = f[b], v001 = f[b + sz], v010 = f[b + sy], v011 = f[b + sy + sz];
    const float v100 = f[b + 1], v101 = f[b + 1 + sz], v110 = f[b + 1 + sy], v111 = f[b + 1 + sy + sz];
    const float v00 = v000 * (1.0f - tz) + v001 * tz;
    const float v01 = v010 * (1.0f - tz) + v011 * tz;
    const float v10 = v100 * (1.0f - tz) + v101 * tz;
    const float v11 = v110 * (1.0f - tz) + v111 * tz;
    const float vLow = v00 * (1.0f - ty) + v01 * ty;
    const float vHigh = v10 * (1.0f - ty) + v11 * ty;
    return vLow * (1.0f - tx) + vHigh * tx;
}

__device__ inline float3 vinterp(const DevGrid& g, float px, float py, float pz) {
    float3 r;
    r.x = trilin(g, g.u, px, py, pz, -1.0f, -0.5f, -0.5f);
    r.y = trilin(g, g.v, px, py, pz, -0.5f, -1.0f, -0.5f);
    r.z = trilin(g, g.w, px, py, pz, -0.5f, -0.5f, -1.0f);
    return r;
}

// ScalField3D::interpolate (ScalField3D.cpp:18-107): cell values sit at i + 0.5.
__device__ inline float sinterp(const DevGrid& g, const float* __restrict__ f, float px, float py, float pz) {
    const int x = (int)px, y = (int)py, z = (int)pz;
    int ix1 = x, ix2 = x + 1;
    if (px - (float)x < 0.5f) { ix1 = x - 1; ix2 = x; }
    int iy1 = y, iy2 = y + 1;
    if (py - (float)y < 0.5f) { iy1 = y - 1; iy2 = y; }
    int iz1 = z, iz2 = z + 1;
    if (pz - (float)z < 0.5f) { iz1 = z - 1; iz2 = z; }
    ix1 = clampv(ix1, 0, g.W - 1); ix2 = clampv(ix2, 0, g.W - 1);
    iy1 = clampv(iy1, 0, g.H - 1); iy2 = clampv(iy2, 0, g.H - 1);
    iz1 = clampv(iz1, 0, g.D - 1); iz2 = clampv(iz2, 0, g.D - 1);
    const float v0 = f[cidx(g, ix1, iy1, iz1)], v1 = f[cidx(g, ix1, iy1, iz2)];
    const float v2 = f[cidx(g, ix1, iy2, iz1)], v3 = f[cidx(g, ix1, iy2, iz2)];
    const float v4 = f[cidx(g, ix2, iy1, iz1)], v5 = f[cidx(g, ix2, iy1, iz2)];
    const float v6 = f[cidx(g, ix2, iy2, iz1)], v7 = f[cidx(g, ix2, iy2, iz2)];
    const float x1 = (float)ix1 + 0.5f, x2 = (float)ix2 + 0.5f;
    const float y1 = (float)iy1 + 0.5f, y2 = (float)iy2 + 0.5f;
    const float z1 = (float)iz1 + 0.5f, z2 = (float)iz2 + 0.5f;
    // The clamped indices of an axis differ by 0 or 1 and i + 0.5f is exact, so each spacing
    // (x2 - x1, ...) is exactly 0 or 1: the reference's divisions by it are divisions by 1.0f,
    // the identity, and are left out.
    const float dz2z1 = z2 - z1;
    float f00, f01, f10, f11;
    if (dz2z1 < 1e-9f) {
        f00 = v1; f01 = v3; f10 = v5; f11 = v7;
    } else {
        const float wz1 = z2 - pz, wz2 = pz - z1;
        f00 = wz1 * v0 + wz2 * v1;
        f01 = wz1 * v2 + wz2 * v3;
        f10 = wz1 * v4 + wz2 * v5;
        f11 = wz1 * v6 + wz2 * v7;
    }
    const float dy2y1 = y2 - y1;
    float fx1, fx2;
    if (dy2y1 < 1e-9f) {
        fx1 = f01; fx2 = f11;
    } else {
        const float wy1 = y2 - py, wy2 = py - y1;
        fx1 = wy1 * f00 + wy2 * f01;
        fx2 = wy1 * f10 + wy2 * f11;
    }
    const float dx2x1 = x2 - x1;
    if (dx2x1 < 1e-9f) return fx2;
    const float wx1 = x2 - px, wx2 = px - x1;
    return wx1 * fx1 + wx2 * fx2;
}

// Grid3D::traceParticle (Grid3D.cpp:66-122): RK2 with `substeps` substeps, velocity sampled at
// p - 0.5, midpoint clamped to [1, n-2], 8-step bisection when a step lands in a wall.
__device__ inline bool inWall(const DevGrid& g, float px, float py, float pz) {
    const float Wf = (float)g.W, Hf = (float)g.H, Df = (float)g.D;
    if (px < 1.f || px > Wf - 2.f || py < 1.f || py > Hf - 2.f || pz < 1.f || pz > Df - 2.f) return true;
    const int ix = clampv((int)px, 0, g.W - 1);
    const int iy = clampv((int)py, 0, g.H - 1);
    const int iz = clampv((int)pz, 0, g.D - 1);
    return (g.bf[cidx(g, ix, iy, iz)] & BF_GT05) != 0;
}

__device__ inline float3 traceParticle(const DevGrid& g, float sx, float sy, float sz, float dt, int substeps) {
    if (inWall(g, sx, sy, sz)) return make_float3(sx, sy, sz);
    const float Wf = (float)g.W, Hf = (float)g.H, Df = (float)g.D;
    const float subDt = dt / (float)substeps;
    const float half = subDt * 0.5f;
    float px = sx, py = sy, pz = sz;
#pragma unroll 1
    for (int s = 0; s < substeps; s++) {
        const float3 k1 = vinterp(g, px - 0.5f, py - 0.5f, pz - 0.5f);
        const float mx = clampv(px + k1.x * half, 1.f, Wf - 2.f);
        const float my = clampv(py + k1.y * half, 1.f, Hf - 2.f);
        const float mz = clampv(pz + k1.z * half, 1.f, Df - 2.f);
        const float3 k2 = vinterp(g, mx - 0.5f, my - 0.5f, mz - 0.5f);
        const float nx = px + k2.x * subDt, ny = py + k2.y * subDt, nz = pz + k2.z * subDt;
        if (inWall(g, nx, ny, nz)) {
            float sx_ = px, sy_ = py, sz_ = pz, ux = nx, uy = ny, uz = nz;
#pragma unroll 1
            for (int b = 0; b < 8; b++) {
                const float hx = (sx_ + ux) * 0.5f, hy = (sy_ + uy) * 0.5f, hz = (sz_ + uz) * 0.5f;
                if (inWall(g, hx, hy, hz)) { ux = hx; uy = hy; uz = hz; }
                else { sx_ = hx; sy_ = hy; sz_ = hz; }
            }
            return make_float3(sx_, sy_, sz_);
        }
        px = nx; py = ny; pz = nz;
    }
    return make_float3(px, py, pz);
}

// Corner weights of a traced position (CPUSimulator3D.cpp:140-163).  m[k] for corners in the
// reference order 000,100,010,110,001,101,011,111 (x fastest); returns the sum.
__device__ inline float cornerWeights(const DevGrid& g, float px, float py, float pz, int& p0, float m[8], int& p0x,
                                     int& p0y, int& p0z) {
    p0x = (int)floorf(px), p0y = (int)floorf(py), p0z = (int)floorf(pz);
    p0x = clampv(p0x, 0, g.W - 2); p0y = clampv(p0y, 0, g.H - 2); p0z = clampv(p0z, 0, g.D - 2);
    const float dx0 = 1 - (px - p0x), dy0 = 1 - (py - p0y), dz0 = 1 - (pz - p0z);
    const float dx1 = 1 - (p0x + 1 - px), dy1 = 1 - (p0y + 1 - py), dz1 = 1 - (p0z + 1 - pz);
    p0 = cidx(g, p0x, p0y, p0z);
    const int sy = g.W, sz = g.W * g.H;
    const int off[8] = {0, 1, sy, sy + 1, sz, sz + 1, sz + sy, sz + sy + 1};
    const float wx[2] = {dx0, dx1}, wy[2] = {dy0, dy1}, wz[2] = {dz0, dz1};
#pragma unroll
    for (int k = 0; k < 8; k++) {
        m[k] = 0.f;
        if (!(g.bf[p0 + off[k]] & BF_EQ1)) m[k] = wx[k & 1] * wy[(k >> 1) & 1] * wz[(k >> 2) & 1];
    }
    float sum = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) sum += m[k];
    return sum;
}

__device__ inline int cornerOffset(const DevGrid& g, int k) {
    return (k & 1) + ((k >> 1) & 1) * g.W + ((k >> 2) & 1) * g.W * g.H;
}

// ---------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------
struct TimerSlot {
    std::vector<hipEvent_t> start, stop;
    int64_t launches = 0;
    double total_ms = 0;
};

struct CGState {
    double gamma;      // r.r of the current residual
    double gam[2];     // r.r by iteration parity (written by block 0 while other blocks read the other slot)
    double thr_in;     // tolerance as given
    double gamma_old;
    double dAd;        // d.Ad
    double alpha;
    double beta;
    double bnorm2;     // b.b
    double thr;        // stop when gamma < thr
    double true_num;   // ||b - A x||^2 (FP64)
    int iter;          // completed iterations
    int done;
    int converged;
    int max_iter;
    unsigned int counter;
    int flags;         // bit0: a border cell is a fluid row
    int stall;         // learned preconditioner: stall detected
    int pad;
    double sigma;      // learned preconditioner: RMS of the restricted vector
    double dq;         // flexible CG: d_k . A d_k of the last iteration
};

struct efs_sim {
    int dim = 3;  // 2: CPUSimulator2D semantics, D = 1 (efs_2d.hip)
    int W, H, D, N;
    efs_params P;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // fields
    float *u = nullptr, *v = nullptr, *w = nullptr, *fu = nullptr, *fv = nullptr, *fw = nullptr;
    float *rho = nullptr, *gamma = nullptr, *beta = nullptr, *div = nullptr, *p = nullptr, *bndf = nullptr;
    uint8_t* bf = nullptr;
    // scratch
    float *tu = nullptr, *tv = nullptr, *tw = nullptr;   // tmpVec (cell-centred advected velocity)
    // rho_advect leaves in tu/tv/tw the velocity sampled at its backward traces from every
    // interior cell centre; vel_advect traces the same cells with the same velocity, walls and dt
    // and samples the same velocity there, so it reuses them when no other stage or upload has
    // run in between (every other stage and every upload clears the flag)
    bool traces_valid = false;
    float *s0 = nullptr, *s1 = nullptr;                   // scalar scratch
    uint8_t *valA = nullptr, *valB = nullptr;             // extrapolation validity
    int *bwd_p0 = nullptr, *fwd_p0 = nullptr;             // advection matrix corner bases (-1 = none)
    float *bwd_w = nullptr, *fwd_w = nullptr;             // 8 entries per cell, entry-major
    unsigned long long *acc0 = nullptr, *acc1 = nullptr, *acc2 = nullptr;  // fixed-point accumulators
    float* stage_bytes = nullptr;                         // AoS staging for upload/download
    float* render4 = nullptr;                             // vec4 velocity of the render feed (allocated on first use)
    // solver
    uint16_t* rowcode = nullptr;
    float *cg_r = nullptr, *cg_d0 = nullptr, *cg_d1 = nullptr, *cg_x1 = nullptr, *cg_z = nullptr;
    float* cg_xd = nullptr;     // FP32 increment of the iterate since the last fold
    double* cg_x64 = nullptr;   // FP64 iterate (reliable-update accumulator)
    bool p64_valid = false;     // the last solve left its FP64 iterate in cg_x64
    bool x64_has_previous = false;  // cg_x64 holds a finished solve (usable as the next initial guess)
    int x0_pending = -1;            // initial iterate still to set by k_cg_init (0: b, 1: previous on CG rows)
    double* partials = nullptr;
    double *p_rr = nullptr, *p_bb = nullptr, *p_dad = nullptr;  // per-block partial sums of the CG kernels
    int max_blocks = 0;
    CGState* st = nullptr;
    CGState* st_host = nullptr;  // pinned
    // timing
    bool timing = false;
    bool ktiming = false;
    int ktiming_every = 0;      // > 0: kernel timing in every ktiming_every-th efs_step (efs_set_kernel_timing)
    long long ktiming_steps = 0;  // efs_step calls since kernel timing was set
    TimerSlot ktimers[9];
    hipEvent_t ev_stage[8] = {};
    // learned preconditioner state (owned by the preconditioner module)
    void* learned = nullptr;
    // multigrid hierarchy (efs_mg.hip)
    void* mg = nullptr;

    DevGrid grid() const {
        DevGrid g;
        g.W = W; g.H = H; g.D = D; g.N = N;
        g.u = u; g.v = v; g.w = w; g.fu = fu; g.fv = fv; g.fw = fw;
        g.rho = rho; g.gamma = gamma; g.beta = beta; g.div = div; g.p = p; g.bf = bf;
        return g;
    }
};

// error reporting
void efs_set_error(const std::string& msg);
int efs_hip_fail(hipError_t e, const char* what);
#define EFS_HIP(call)                                         \
    do {                                                      \
        hipError_t _e = (call);                               \
        if (_e != hipSuccess) return efs_hip_fail(_e, #call); \
    } while (0)

// stage launchers (efs_step.hip)
int efs_launch_decay(efs_sim* s);
int efs_launch_extrapolate(efs_sim* s, bool with_decay = false);  // with_decay: vel_decay first, fused
int efs_launch_sharpen(efs_sim* s);
int efs_launch_rho_advect(efs_sim* s);
int efs_launch_vel_advect(efs_sim* s, bool with_force = false);  // with_force: add_force after, fused
int efs_launch_add_force(efs_sim* s);
int efs_launch_walls_divergence(efs_sim* s);
int efs_launch_gradient_walls(efs_sim* s);
int efs_launch_boundary_flags(efs_sim* s);
int ensure_forces(efs_sim* s);  // allocates the forces field (zero) on first use

// solver launchers (efs_solver.hip)
int efs_solver_alloc(efs_sim* s);
void efs_solver_free(efs_sim* s);
int efs_launch_rowcode(efs_sim* s, int* border_rows);
int efs_cg_solve(efs_sim* s, efs_stats* st);
int efs_cg_run(efs_sim* s, efs_stats* st, bool border, int max_it);
// A preconditioner for efs_cg_run_pc: apply() writes z = M r (r = s->cg_r) into z, in stream
// order; batch = iterations launched between host convergence checks, first_batch (when > 0)
// those launched before the first check.
struct CgPrecond {
    int (*apply)(efs_sim* s, void* ctx);
    void* ctx;
    float* z;
    double* p_rz;
    int batch;
    int first_batch = 0;
    bool writes_rz = false;  // apply also writes the r . z partials (no separate k_cg_rz)
    // stall detection: stop when r.r has not fallen below stall_ratio times its value
    // stall_window iterations earlier (0: off); `stalled` reports it
    int stall_window = 0;
    double stall_ratio = 0.5;
    mutable bool stalled = false;
};
int efs_cg_run_pc(efs_sim* s, efs_stats* st, bool border, int max_it, const CgPrecond* pc);
int efs_cg_setup_x0(efs_sim* s);                   // x64 = b, xd = 0
int efs_cg_defer_x0(efs_sim* s);                   // the same, done by the next efs_cg_init_residual
int efs_cg_init_residual(efs_sim* s, bool border);  // r = b - A x64 (FP64), partials of r.r, b.b
int efs_cg_fold_all(efs_sim* s);                    // x64 += xd, xd = 0
int efs_cg_finalize(efs_sim* s, bool xd_zero = false);  // x64 += xd, pressure = (float)x64
int efs_jacobi_solve(efs_sim* s, efs_stats* st);
int efs_launch_spmv(efs_sim* s, const float* x, float* y);
int efs_true_residual(efs_sim* s, double* num, double* den, bool write_p = false);

// 2D step (efs_2d.hip)
int efs2_alloc(efs_sim* s);
int efs2_run_stage(efs_sim* s, int stage, efs_stats* st);
int efs2_solve(efs_sim* s, efs_stats* st);            // on the current divergence, p cleared by the caller
int efs2_walls_divergence(efs_sim* s);
int efs2_spmv(efs_sim* s, const float* x, float* y);

// multigrid-preconditioned CG (efs_mg.hip)
int efs_mg_cg_solve(efs_sim* s, efs_stats* st);
void efs_mg_free(efs_sim* s);

// learned preconditioner (efs_learned.hip)
int efs_learned_cg_solve(efs_sim* s, efs_stats* st);
void efs_learned_free(efs_sim* s);

// event-timed kernels: KT_OTHER times the fine-level multigrid up sweep, KT_MG_DOWN its down sweep;
// the advection traces, the walls + divergence pass and the gradient pass of the step follow
enum {
    KT_CG_DIR_SPMV = 0, KT_CG_UPDATE = 1, KT_JACOBI = 2, KT_OTHER = 3, KT_MG_DOWN = 4,
    KT_ADV_BWD = 5, KT_ADV_FWD = 6, KT_WALLS_DIV = 7, KT_GRADIENT = 8, KT_COUNT = 9
};
void efs_ktimer_pair(efs_sim* s, hipEvent_t* a, hipEvent_t* b);  // for hipExtLaunchKernelGGL's start/stop events
void efs_ktimer_add(efs_sim* s, int k, hipEvent_t a, hipEvent_t b);
void efs_trim_timer(TimerSlot& t, size_t keep);  // drops the events of launches past `keep`
