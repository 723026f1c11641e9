// Stage kernels of the 3D liquid step on gfx950, following CPUSimulator3D's semantics
// (src/simulation/simulator/CPUSimulator3D.cpp).  All float arithmetic is written in the
// reference's operation order and the library is built without FMA contraction, so the
// per-cell stages reproduce the CPU path bit for bit on the same inputs.  Two stages are
// order-dependent on the CPU and are restated for a parallel device:
//   * rho_sharpen (CPUSimulator3D.cpp:446-455) applies, traces and scatters cell by cell in
//     place; here every change is applied first, then all traces read the same field, and the
//     scatters accumulate in fixed point (deterministic in any atomic order).
//   * rho_advect's sparse matrix A (std::unordered_map, CPUSimulator3D.cpp:133-272) is kept as
//     8 backward entries per row and 8 forward entries per column; sums over the entries a
//     cell does not own are scattered in fixed point, the rest are gathered.
// The density diffusion sweeps (CPUSimulator3D.cpp:276-363) are sequential along a line and
// independent across lines, so one thread per line reproduces them exactly (lines along x are
// transposed through LDS so that the loads stay coalesced).
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <hip/hip_ext.h>

#include <algorithm>

#include "efs_internal.h"

namespace {

constexpr int BX = 64, BY = 4;

inline dim3 interior_grid(const efs_sim* s) { return dim3((s->W - 2 + BX - 1) / BX, (s->H - 2 + BY - 1) / BY, s->D - 2); }
inline dim3 full_grid(const efs_sim* s) { return dim3((s->W + BX - 1) / BX, (s->H + BY - 1) / BY, s->D); }

inline dim3 interior_grid_xcd(const efs_sim* s) { return xcd_grid((s->W - 2 + BX - 1) / BX, (s->H - 2 + BY - 1) / BY, s->D - 2); }
inline dim3 full_grid_xcd(const efs_sim* s) { return xcd_grid((s->W + BX - 1) / BX, (s->H + BY - 1) / BY, s->D); }

// Block shape of the advection passes that scatter (the traces and the row and column sums):
// ABX x ABY cells of one z plane, 256 threads.
#ifndef EFS_ADV_BX
#define EFS_ADV_BX 64
#endif
constexpr int ABX = EFS_ADV_BX, ABY = 256 / ABX;
static_assert(ABX * ABY == 256 && ABX % 16 == 0, "advection blocks are 256 threads, at least 16 wide");
inline dim3 adv_interior_grid(const efs_sim* s) { return dim3((s->W - 2 + ABX - 1) / ABX, (s->H - 2 + ABY - 1) / ABY, s->D - 2); }
inline dim3 adv_full_grid(const efs_sim* s) { return dim3((s->W + ABX - 1) / ABX, (s->H + ABY - 1) / ABY, s->D); }

#define INTERIOR_XYZ(g)                                                             \
    const int x = 1 + blockIdx.x * BX + threadIdx.x, y = 1 + blockIdx.y * BY + threadIdx.y; \
    const int z = 1 + blockIdx.z;                                                   \
    if (x > (g).W - 2 || y > (g).H - 2) return;                                     \
    const int c = cidx((g), x, y, z)

#define INTERIOR_XYZ_XCD(g)                                                              \
    XcdBlock _blk;                                                                       \
    if (!xcd_block(((g).W - 2 + BX - 1) / BX, ((g).H - 2 + BY - 1) / BY, _blk)) return;  \
    const int x = 1 + _blk.bx * BX + threadIdx.x, y = 1 + _blk.by * BY + threadIdx.y;    \
    const int z = 1 + _blk.bz;                                                           \
    if (x > (g).W - 2 || y > (g).H - 2) return;                                          \
    const int c = cidx((g), x, y, z)

#define FULL_XYZ_XCD(g)                                                          \
    XcdBlock _blk;                                                               \
    if (!xcd_block(((g).W + BX - 1) / BX, ((g).H + BY - 1) / BY, _blk)) return;  \
    const int x = _blk.bx * BX + threadIdx.x, y = _blk.by * BY + threadIdx.y;    \
    const int z = _blk.bz;                                                       \
    if (x >= (g).W || y >= (g).H) return;                                        \
    const int c = cidx((g), x, y, z)

#define FULL_XYZ(g)                                                                 \
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y; \
    const int z = blockIdx.z;                                                       \
    if (x >= (g).W || y >= (g).H) return;                                           \
    const int c = cidx((g), x, y, z)

// ---- boundary byte -----------------------------------------------------------------------
__global__ void k_boundary_flags(int n, const float* __restrict__ b, uint8_t* __restrict__ bf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = b[i];
    bf[i] = (uint8_t)((v == 1.f ? BF_EQ1 : 0u) | (v > 0.5f ? BF_GT05 : 0u) | (v >= 0.5f ? BF_GE05 : 0u));
}

__global__ void k_fill_u64(int n, unsigned long long* a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0ull;
}

// ---- vel_decay (CPUSimulator3D.cpp:54-59) --------------------------------------------------
__global__ __launch_bounds__(256) void k_decay(DevGrid g, float decay) {
    INTERIOR_XYZ(g);
    if (g.rho[c] < 0.5f) {
        g.u[c] = g.u[c] * decay;
        g.v[c] = g.v[c] * decay;
        g.w[c] = g.w[c] * decay;
    }
}

// ---- vel_extrapolate (CPUSimulator3D.cpp:61-111) -------------------------------------------
// The reference runs 6 Jacobi passes (velocity and validity ping-ponged between two buffers): an
// invalid, non-solid cell with valid neighbours becomes valid with their mean, every other cell
// is copied.  A cell's velocity therefore changes at most once, in the pass where it becomes
// valid, and what it reads then (the neighbours valid before that pass) never changes again.  So
// the passes run in place: stamp[c] is the pass at which c became valid (0: valid from the
// start, 255: not yet), and pass k updates the cells with stamp 255 that have a neighbour with
// stamp < k.  A neighbour becoming valid in the same pass reads as 255 or k, both "not valid
// before k", so the result does not depend on the order of the threads; a cell that is not on
// the front costs one byte read instead of a copy.
// With DECAY the velocity decay of the step (k_decay) is applied in the same pass: it reads only
// rho, which the stamps read too.  A solid cell (boundary == 1) is stamped 254: it never becomes
// valid (only 255 does) and never counts as a valid neighbour (254 is not below any pass), so the
// passes need not read the boundary.
template <bool DECAY>
__global__ __launch_bounds__(256) void k_extrap_init(DevGrid g, uint8_t* __restrict__ stamp, float decay) {
    FULL_XYZ(g);
    const bool interior = x > 0 && y > 0 && z > 0 && x < g.W - 1 && y < g.H - 1 && z < g.D - 1;
    const float r = g.rho[c];
    const bool solid = (g.bf[c] & BF_EQ1) != 0;
    if constexpr (DECAY) {
        // loaded with rho, not after its test: most cells of a typical state are air and decay
        const float uc = g.u[c], vc = g.v[c], wc = g.w[c];
        if (interior && r < 0.5f) {
            g.u[c] = uc * decay;
            g.v[c] = vc * decay;
            g.w[c] = wc * decay;
        }
    }
    stamp[c] = solid ? 254 : ((interior && r >= 0.5f) ? 0 : 255);
}

// A thread takes EXZ interior planes of one (x, y) column and loads every stamp they read at
// once (the four lateral neighbours of each plane and the column from one plane below to one
// above, all addresses in the array), so a chunk costs one memory round trip; then it tests the
// planes in order.  The column's own stamps are carried as this pass leaves them (a neighbour
// stamped in this pass counts as not valid either way, 255 or the pass, so values loaded before
// another thread stamps them are just as good).  The blocks take (tile, z chunk) in XCD order, so
// neighbouring tiles' rows are read through the same L2.
constexpr int EXZ = 8;
__global__ __launch_bounds__(256) void k_extrap_pass(DevGrid g, uint8_t* stamp, int pass) {
    XcdBlock blk;
    if (!xcd_block((g.W - 2 + BX - 1) / BX, (g.H - 2 + BY - 1) / BY, blk)) return;
    const int x = 1 + blk.bx * BX + threadIdx.x, y = 1 + blk.by * BY + threadIdx.y;
    if (x > g.W - 2 || y > g.H - 2) return;
    const int zb = 1 + blk.bz * EXZ, nz = min(EXZ, g.D - 1 - zb);  // planes [zb, zb + nz)
    const int sy = g.W, sz = g.W * g.H;
    const int c0 = cidx(g, x, y, zb);
    unsigned lat[EXZ][4], col[EXZ + 2];  // col[k]: the column's stamp at plane zb - 1 + k
#pragma unroll
    for (int k = 0; k < EXZ + 2; k++) col[k] = stamp[c0 + (min(k, nz + 1) - 1) * sz];
#pragma unroll
    for (int k = 0; k < EXZ; k++) {
        const int c = c0 + min(k, nz - 1) * sz;
        lat[k][0] = stamp[c - 1];
        lat[k][1] = stamp[c + 1];
        lat[k][2] = stamp[c - sy];
        lat[k][3] = stamp[c + sy];
    }
#pragma unroll
    for (int k = 0; k < EXZ; k++) {
        if (k >= nz) break;
        const int c = c0 + k * sz;
        const unsigned st[6] = {lat[k][0], lat[k][1], lat[k][2], lat[k][3], col[k], col[k + 2]};  // the reference's order
        unsigned m = 0;
#pragma unroll
        for (int j = 0; j < 6; j++) m |= ((int)st[j] < pass ? 1u : 0u) << j;
        if (col[k + 1] == 255 && m != 0) {
            const int nb[6] = {c - 1, c + 1, c - sy, c + sy, c - sz, c + sz};
            float ax = 0.f, ay = 0.f, az = 0.f, n = 0;
#pragma unroll
            for (int j = 0; j < 6; j++) {
                if (m & (1u << j)) {
                    ax = ax + g.u[nb[j]]; ay = ay + g.v[nb[j]]; az = az + g.w[nb[j]];
                    n++;
                }
            }
            g.u[c] = ax / n; g.v[c] = ay / n; g.w[c] = az / n;
            stamp[c] = (uint8_t)pass;
            col[k + 1] = pass;
        }
    }
}

// ---- rho_sharpen (CPUSimulator3D.cpp:366-456) ----------------------------------------------
__device__ inline double dsq(float v) { return (double)v * (double)v; }  // std::pow(float, 2) in double
__device__ inline double dmax(double a, double b) { return (a < b) ? b : a; }
__device__ inline float fmaxr(float a, float b) { return (a < b) ? b : a; }  // std::max
__device__ inline float fminr(float a, float b) { return (b < a) ? b : a; }  // std::min

__global__ __launch_bounds__(256) void k_sharpen_delta(DevGrid g, float delT, float* __restrict__ del) {
    INTERIOR_XYZ(g);
    const int sy = g.W, sz = g.W * g.H;
    const float r = g.rho[c];
    // The tests at the end decide these cells without the rest: rho < 1e-5 gives -rho; rho > 0.5
    // gives 0, since then the weight is >= 0 or NaN and the increment is >= 0 or NaN, so
    // rho + dr < 0 never holds.  Air and liquid waves stop here; NaN goes on as before.
    if (r > 0.5f || (double)r < 1e-5) {
        del[c] = ((double)r < 1e-5) ? -r : 0.f;
        return;
    }
    const float rxm = g.rho[c - 1], rxp = g.rho[c + 1], rym = g.rho[c - sy], ryp = g.rho[c + sy];
    const float rzm = g.rho[c - sz], rzp = g.rho[c + sz];
    const float dpx = -(r - rxm) * delT, dmx = -(rxp - r) * delT;
    const float dpy = -(r - rym) * delT, dmy = -(ryp - r) * delT;
    const float dpz = -(r - rzm) * delT, dmz = -(rzp - r) * delT;
    const float max_inc = (float)sqrt(dmax(dsq(fmaxr(dpx, 0.f)), dsq(fminr(dmx, 0.f))) +
                                      dmax(dsq(fmaxr(dpy, 0.f)), dsq(fminr(dmy, 0.f))) +
                                      dmax(dsq(fmaxr(dpz, 0.f)), dsq(fminr(dmz, 0.f))));
    const float max_dec = (float)sqrt(dmax(dsq(fminr(dpx, 0.f)), dsq(fmaxr(dmx, 0.f))) +
                                      dmax(dsq(fminr(dpy, 0.f)), dsq(fmaxr(dmy, 0.f))) +
                                      dmax(dsq(fminr(dpz, 0.f)), dsq(fmaxr(dmz, 0.f))));
    float mx = -INFINITY;
    mx = fmaxr(mx, fabsf(r - rxm));
    mx = fmaxr(mx, fabsf(r - rxp));
    mx = fmaxr(mx, fabsf(r - rym));
    mx = fmaxr(mx, fabsf(r - ryp));
    mx = fmaxr(mx, fabsf(r - rzm));
    mx = fmaxr(mx, fabsf(r - rzp));
    const float a = r - 0.5f;
    const float cube = (float)((double)a * (double)a * (double)a);  // powf(a, 3)
    const float w = cube * (1 - fminr(1.f, mx / 0.2f));
    float dr = (w >= 0) ? w * max_inc : w * max_dec;
    if (r + dr < 0 || (double)r < 1e-5) dr = -r;
    else if (r > 0.5f) dr = 0;
    del[c] = dr;
}

__global__ __launch_bounds__(256) void k_sharpen_apply(DevGrid g, const float* __restrict__ del) {
    INTERIOR_XYZ(g);
    const float dr = del[c];
    if (fabsf(dr) < 1e-6f) return;
    const float r = g.rho[c] + dr;
    g.rho[c] = fmaxr(r, 0.f);
}

// trace_along_field (CPUSimulator3D.cpp:568-607)
__device__ inline float3 trace_along_field(const DevGrid& g, float sx, float sy, float sz, float Dm, int substeps) {
    const float stepSize = Dm / (float)substeps;
    const float Wf = (float)g.W, Hf = (float)g.H, Df = (float)g.D;
    float px = sx, py = sy, pz = sz;
    const int syo = g.W, szo = g.W * g.H;
    for (int s = 0; s < substeps; s++) {
        if (sinterp(g, g.rho, px, py, pz) >= 0.5f) break;
        const int ix = clampv((int)px, 1, g.W - 2), iy = clampv((int)py, 1, g.H - 2), iz = clampv((int)pz, 1, g.D - 2);
        const int q = cidx(g, ix, iy, iz);
        const float gx = g.rho[q + 1] - g.rho[q - 1];
        const float gy = g.rho[q + syo] - g.rho[q - syo];
        const float gz = g.rho[q + szo] - g.rho[q - szo];
        const float len = sqrtf(gx * gx + gy * gy + gz * gz);
        if (len < 1e-6f) break;
        const float dx = gx / len, dy = gy / len, dz = gz / len;
        float nx = px + dx * stepSize, ny = py + dy * stepSize, nz = pz + dz * stepSize;
        nx = clampv(nx, 1.f, Wf - 2.f);
        ny = clampv(ny, 1.f, Hf - 2.f);
        nz = clampv(nz, 1.f, Df - 2.f);
        if (g.bf[cidx(g, (int)nx, (int)ny, (int)nz)] & BF_GT05) break;
        px = nx; py = ny; pz = nz;
    }
    return make_float3(px, py, pz);
}

// scatter_values (CPUSimulator3D.cpp:609-655) into a fixed-point accumulator
__device__ inline void scatter_fix(const DevGrid& g, float px, float py, float pz, float value, unsigned long long* acc) {
    const int xs[2] = {clampv((int)px, 0, g.W - 1), clampv((int)px + 1, 0, g.W - 1)};
    const int ys[2] = {clampv((int)py, 0, g.H - 1), clampv((int)py + 1, 0, g.H - 1)};
    const int zs[2] = {clampv((int)pz, 0, g.D - 1), clampv((int)pz + 1, 0, g.D - 1)};
    const float fx = px - (float)(int)px, fy = py - (float)(int)py, fz = pz - (float)(int)pz;
    float w[8];
    w[0] = (1.f - fx) * (1.f - fy) * (1.f - fz);
    w[1] = fx * (1.f - fy) * (1.f - fz);
    w[2] = (1.f - fx) * fy * (1.f - fz);
    w[3] = fx * fy * (1.f - fz);
    w[4] = (1.f - fx) * (1.f - fy) * fz;
    w[5] = fx * (1.f - fy) * fz;
    w[6] = (1.f - fx) * fy * fz;
    w[7] = fx * fy * fz;
    int q[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        q[k] = cidx(g, xs[k & 1], ys[(k >> 1) & 1], zs[(k >> 2) & 1]);
        if (g.bf[q[k]] & BF_GT05) w[k] = 0.f;
    }
    const float total = w[0] + w[1] + w[2] + w[3] + w[4] + w[5] + w[6] + w[7];
    if (total < 1e-6f) return;
    const float inv = 1.f / total;
#pragma unroll
    for (int k = 0; k < 8; k++) fix_add(acc, q[k], value * w[k] * inv);
}

// The cells that scatter (|delta| >= 1e-6) are the few near the free surface, spread over many
// waves, and a wave runs the whole trace loop when any of its lanes has one.  So each BX x BY block
// first packs its scattering cells into its lowest lanes through LDS; waves left without a cell
// exit.  The scatters are fixed-point sums, so the order in which lanes land does not matter.
__global__ __launch_bounds__(256) void k_sharpen_scatter(DevGrid g, const float* __restrict__ del, float rhoD,
                                                         unsigned long long* acc) {
    __shared__ unsigned short s_src[BX * BY];
    __shared__ float s_dr[BX * BY];
    __shared__ int s_n;
    const int t = threadIdx.x + BX * threadIdx.y, lane = t & 63;
    if (t == 0) s_n = 0;
    __syncthreads();
    const int x = 1 + blockIdx.x * BX + threadIdx.x, y = 1 + blockIdx.y * BY + threadIdx.y, z = 1 + blockIdx.z;
    bool act = false;
    float dr = 0.f;
    if (x <= g.W - 2 && y <= g.H - 2) {
        dr = del[cidx(g, x, y, z)];
        act = !(fabsf(dr) < 1e-6f);
    }
    const unsigned long long bal = __ballot(act);
    int base = 0;
    if (lane == 0 && bal) base = atomicAdd(&s_n, __popcll(bal));
    base = __shfl(base, 0);
    if (act) {
        const int slot = base + __popcll(bal & ((1ull << lane) - 1ull));
        s_src[slot] = (unsigned short)t;
        s_dr[slot] = dr;
    }
    __syncthreads();
    if (t >= s_n) return;
    const int src = s_src[t];
    const int sx = 1 + blockIdx.x * BX + src % BX, sy = 1 + blockIdx.y * BY + src / BX;
    const float3 p = trace_along_field(g, (float)sx, (float)sy, (float)z, rhoD, 10);
    scatter_fix(g, p.x, p.y, p.z, -s_dr[t], acc);
}

__global__ void k_add_fix(int n, float* __restrict__ f, unsigned long long* __restrict__ acc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long a = acc[i];
    if (a != 0ull) {
        f[i] = (float)((double)f[i] + from_fix(a));
        acc[i] = 0ull;
    }
}

// ---- block-private scatter box -------------------------------------------------------------
// The column and row sums of the advection matrix are fixed-point integer sums (modulo 2^64),
// so their terms may be added in any grouping.  The 8-corner scatters of an ABX x ABY block
// land, where the traced positions move together, on about (ABX + 1) x (ABY + 1) x 2 cells,
// which each receive up to eight terms from the block.  The block first adds its terms into a box of LDS anchored at
// its smallest target corner and then adds each nonzero box cell to HBM with one atomic, which
// cuts the device atomics about threefold.  A lane whose corners reach past the box adds its
// terms to HBM directly.
constexpr int SB_X = ABX + 4, SB_Y = ABY + 4, SB_Z = 4, SB_N = SB_X * SB_Y * SB_Z;
constexpr int SB_WAVES = (ABX * ABY) / 64;

struct ScatterBox {
    unsigned long long cell[SB_N];
    int wmin[SB_WAVES][3];
};

__device__ inline void sbox_clear(ScatterBox& b, int t) {
    for (int i = t; i < SB_N; i += ABX * ABY) b.cell[i] = 0ull;
}

// Each wave's smallest target corner (lanes without terms give none); read after a barrier.
__device__ inline void sbox_vote(ScatterBox& b, int wave, bool live, int px, int py, int pz) {
    int mx = live ? px : INT_MAX, my = live ? py : INT_MAX, mz = live ? pz : INT_MAX;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mx = min(mx, __shfl_xor(mx, o));
        my = min(my, __shfl_xor(my, o));
        mz = min(mz, __shfl_xor(mz, o));
    }
    if (__lane_id() == 0) {
        b.wmin[wave][0] = mx;
        b.wmin[wave][1] = my;
        b.wmin[wave][2] = mz;
    }
}

// Box origin (the block's smallest target corner; INT_MAX when no lane has terms).
__device__ inline int3 sbox_origin(const ScatterBox& b) {
    int3 o = make_int3(INT_MAX, INT_MAX, INT_MAX);
#pragma unroll
    for (int w = 0; w < SB_WAVES; w++) {
        o.x = min(o.x, b.wmin[w][0]);
        o.y = min(o.y, b.wmin[w][1]);
        o.z = min(o.z, b.wmin[w][2]);
    }
    return o;
}

// Adds the 8 corner terms of a lane with target corner (px, py, pz) = linear index p0.
__device__ inline void sbox_add(const DevGrid& g, ScatterBox& b, int3 o, int px, int py, int pz, int p0,
                                const long long q[8], unsigned long long* acc) {
    const int rx = px - o.x, ry = py - o.y, rz = pz - o.z;
    if (rx <= SB_X - 2 && ry <= SB_Y - 2 && rz <= SB_Z - 2) {
        const int s = rx + SB_X * (ry + SB_Y * rz);
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (q[k] != 0) atomicAdd(&b.cell[s + (k & 1) + ((k >> 1) & 1) * SB_X + ((k >> 2) & 1) * SB_X * SB_Y],
                                     (unsigned long long)q[k]);
    } else {
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (q[k] != 0) atomicAdd(acc + p0 + cornerOffset(g, k), (unsigned long long)q[k]);
    }
}

// Adds the nonzero box cells to HBM; after a barrier that follows every sbox_add.  With CLEAR
// each thread zeroes the cells it read, so after a further barrier the box takes new terms.
template <bool CLEAR = false>
__device__ inline void sbox_flush(const DevGrid& g, ScatterBox& b, int3 o, int t, unsigned long long* acc) {
    if (o.x == INT_MAX) return;
    const int base = cidx(g, o.x, o.y, o.z);
    for (int i = t; i < SB_N; i += ABX * ABY) {
        const unsigned long long v = b.cell[i];
        if (v == 0ull) continue;
        if (CLEAR) b.cell[i] = 0ull;
        const int ix = i % SB_X, iy = (i / SB_X) % SB_Y, iz = i / (SB_X * SB_Y);
        atomicAdd(acc + base + ix + g.W * (iy + g.H * iz), v);
    }
}

// (x, y, z) of a linear cell index: quotients by W and W * H from a double estimate corrected
// once (the estimate is within one of the quotient for indices below 2^31).
__device__ inline int3 cell_coords(const DevGrid& g, int i) {
    const int wh = g.W * g.H;
    int z = (int)((double)i * (1.0 / (double)wh));
    int r = i - z * wh;
    if (r < 0) { z--; r += wh; } else if (r >= wh) { z++; r -= wh; }
    int y = (int)((double)r * (1.0 / (double)g.W));
    int x = r - y * g.W;
    if (x < 0) { y--; x += g.W; } else if (x >= g.W) { y++; x -= g.W; }
    return make_int3(x, y, z);
}

// ---- rho_advect (CPUSimulator3D.cpp:113-364) ----------------------------------------------
// bwd_w[k * N + i] = A(i, p0_i + corner k) for row i; fwd_w[k * N + j] = A(q0_j + corner k, j).

// Steps 1-4: backward trace, advected gamma, row weights scaled by it, column sums (beta).
// Also samples the velocity at the traced position, as vel_advect does from the same trace
// (CPUSimulator3D.cpp:486-493), into su/sv/sw: vel_advect reuses those samples when it directly
// follows with the velocity unchanged.
__global__ __launch_bounds__(256) void k_adv_bwd(DevGrid g, float dt, float* __restrict__ gadv, int* __restrict__ bp0,
                                                 float* __restrict__ bw, unsigned long long* beta_acc, float* __restrict__ su,
                                                 float* __restrict__ sv, float* __restrict__ sw) {
    __shared__ ScatterBox box;
    const int t = threadIdx.x + ABX * threadIdx.y;
    sbox_clear(box, t);
    const int x = 1 + blockIdx.x * ABX + threadIdx.x, y = 1 + blockIdx.y * ABY + threadIdx.y, z = 1 + blockIdx.z;
    bool live = false;
    int p0 = 0, px = 0, py = 0, pz = 0;
    long long q[8];
    if (x <= g.W - 2 && y <= g.H - 2) {
        const int c = cidx(g, x, y, z);
        const float3 prev = traceParticle(g, (float)x, (float)y, (float)z, -dt, 4);
        const float3 vs = vinterp(g, prev.x - 0.5f, prev.y - 0.5f, prev.z - 0.5f);
        su[c] = vs.x;
        sv[c] = vs.y;
        sw[c] = vs.z;
        const float ga = (g.bf[c] & BF_EQ1) ? 0.f : sinterp(g, g.gamma, prev.x, prev.y, prev.z);
        gadv[c] = ga;
        float m[8];
        const float sum = cornerWeights(g, prev.x, prev.y, prev.z, p0, m, px, py, pz);
        live = sum != 0;
        bp0[c] = live ? p0 : -1;
        if (live) {
            float a[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                a[k] = m[k] / sum;
                if (ga > 1.e-5f) a[k] /= ga;
                bw[(size_t)k * g.N + c] = a[k];
            }
            to_fix_n(a, q);
        }
    }
    sbox_vote(box, t >> 6, live, px, py, pz);
    __syncthreads();
    const int3 o = sbox_origin(box);
    if (live) sbox_add(g, box, o, px, py, pz, p0, q, beta_acc);
    __syncthreads();
    sbox_flush(g, box, o, t, beta_acc);
}

// Step 5: forward trace for columns with beta < 1; the new entries are scattered into the
// row sums of step 6 (gamma_acc).  A third to a half of the cells have beta < 1, spread over
// most waves, so each workgroup (an ABX x ABY tile) first packs its tracing cells into the
// lowest lanes through LDS.  Every cell's work is independent of the lane it lands on and the
// scatters are fixed-point sums, so the results do not depend on the packing order.
#ifndef EFS_FWD_BOX
#define EFS_FWD_BOX 1
#endif
__global__ __launch_bounds__(256) void k_adv_fwd(DevGrid g, float dt, const unsigned long long* __restrict__ beta_acc,
                                                 int* __restrict__ fq0, float* __restrict__ fw,
                                                 unsigned long long* gamma_acc) {
    __shared__ unsigned short s_slot[256];
    __shared__ float s_bj[256];
    __shared__ int s_n;
#if EFS_FWD_BOX
    __shared__ ScatterBox box;
#endif
    const int t = threadIdx.x + ABX * threadIdx.y, lane = t & 63;
    if (t == 0) s_n = 0;
#if EFS_FWD_BOX
    sbox_clear(box, t);
#endif
    __syncthreads();
    const int x0 = 1 + blockIdx.x * ABX, y0 = 1 + blockIdx.y * ABY, z = 1 + blockIdx.z;
    bool act = false;
    float bj = 0.f;
    if (x0 + (int)threadIdx.x <= g.W - 2 && y0 + (int)threadIdx.y <= g.H - 2) {
        const int c = cidx(g, x0 + threadIdx.x, y0 + threadIdx.y, z);
        bj = (float)from_fix(beta_acc[c]);
        act = bj < 1;
        if (!act) fq0[c] = -1;
    }
    const unsigned long long bal = __ballot(act);
    int base = 0;
    if (lane == 0 && bal) base = atomicAdd(&s_n, __popcll(bal));
    base = __shfl(base, 0);
    if (act) {
        const int slot = base + __popcll(bal & ((1ull << lane) - 1ull));
        s_slot[slot] = (unsigned short)t;
        s_bj[slot] = bj;
    }
    __syncthreads();
#if !EFS_FWD_BOX
    if (t >= s_n) return;
#endif
    bool live = false;
    int q0 = 0, qx = 0, qy = 0, qz = 0;
    long long q[8];
    if (t < s_n) {
        const int src = s_slot[t];
        bj = s_bj[t];
        const int x = x0 + src % ABX, y = y0 + src / ABX;
        const int c = cidx(g, x, y, z);
        const float3 next = traceParticle(g, (float)x, (float)y, (float)z, +dt, 4);
        float m[8];
        const float sum = cornerWeights(g, next.x, next.y, next.z, q0, m, qx, qy, qz);
        live = sum != 0;
        fq0[c] = live ? q0 : -1;
        if (live) {
            float a[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                a[k] = (1 - bj) * (m[k] / sum);
                fw[(size_t)k * g.N + c] = a[k];
            }
            to_fix_n(a, q);
        }
    }
#if EFS_FWD_BOX
    sbox_vote(box, t >> 6, live, qx, qy, qz);
    __syncthreads();
    const int3 o = sbox_origin(box);
    if (live) sbox_add(g, box, o, qx, qy, qz, q0, q, gamma_acc);
    __syncthreads();
    sbox_flush(g, box, o, t, gamma_acc);
#else
    if (live)
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (q[k] != 0) atomicAdd(gamma_acc + q0 + cornerOffset(g, k), (unsigned long long)q[k]);
#endif
}

// The backward entries are stored once, as step 4 leaves them; their row and column scalings
// (steps 7 and 9) are applied where they are consumed, in the same order of float divisions, so
// the values are those of scaling in place.  The forward entries, which few cells have, are
// scaled in place.

// Steps 6-8 for the backward entries (all cells): gamma_i = own backward row + scattered forward
// entries; the row's entries divided by it add to the new column sums (beta).
__global__ __launch_bounds__(256) void k_adv_rows(DevGrid g, const int* __restrict__ bp0, const float* __restrict__ bw,
                                                  unsigned long long* __restrict__ gamma_acc, float* __restrict__ g6,
                                                  unsigned long long* beta_acc) {
    __shared__ ScatterBox box;
    const int t = threadIdx.x + ABX * threadIdx.y;
    sbox_clear(box, t);
    const int x = blockIdx.x * ABX + threadIdx.x, y = blockIdx.y * ABY + threadIdx.y, z = blockIdx.z;
    bool live = false;
    int p0 = -1;
    int3 pc = make_int3(0, 0, 0);
    long long q[8];
    if (x < g.W && y < g.H) {
        const int c = cidx(g, x, y, z);
        long long s = (long long)gamma_acc[c];
        gamma_acc[c] = 0ull;
        p0 = bp0[c];
        live = p0 >= 0;
        float a[8];
        if (live) {
#pragma unroll
            for (int k = 0; k < 8; k++) a[k] = bw[(size_t)k * g.N + c];
            long long qa[8];
            to_fix_n(a, qa);
#pragma unroll
            for (int k = 0; k < 8; k++) s += qa[k];
        }
        const float gi = (float)((double)s * FIX_INV);
        g6[c] = gi;
        if (live) {
            pc = cell_coords(g, p0);
            float v[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                v[k] = a[k];
                if (gi > 1.e-5f) v[k] /= gi;
            }
            to_fix_n(v, q);
        }
    }
    sbox_vote(box, t >> 6, live, pc.x, pc.y, pc.z);
    __syncthreads();
    const int3 o = sbox_origin(box);
    if (live) sbox_add(g, box, o, pc.x, pc.y, pc.z, p0, q, beta_acc);
    __syncthreads();
    sbox_flush(g, box, o, t, beta_acc);
}

// Steps 7-11 for the forward entries, with the beta field.  After k_adv_rows the column sum of
// every column j holds its backward entries; the thread of column j (all cells) adds its own
// forward entries, divided by the gamma of their rows (steps 7-8), stores beta_j and clears the
// accumulator.  A column with forward entries then divides them by beta_j and adds them to the
// next row sums (gamma) and to rho' = A rho (steps 9-11) through the scatter box.  The scaled
// entries stay in registers: nothing after this pass reads them.
__global__ __launch_bounds__(256) void k_adv_fwd_cols(DevGrid g, const float* __restrict__ g6, const int* __restrict__ fq0,
                                                      const float* __restrict__ fw, unsigned long long* __restrict__ beta_acc,
                                                      unsigned long long* gamma_acc, unsigned long long* rho_acc) {
    __shared__ ScatterBox box;
    const int t = threadIdx.x + ABX * threadIdx.y;
    sbox_clear(box, t);
    const int x = blockIdx.x * ABX + threadIdx.x, y = blockIdx.y * ABY + threadIdx.y, z = blockIdx.z;
    bool live = false;
    int q0 = -1;
    int3 qc = make_int3(0, 0, 0);
    long long qg[8], qr[8];
    if (x < g.W && y < g.H) {
        const int c = cidx(g, x, y, z);
        long long sum = (long long)beta_acc[c];
        beta_acc[c] = 0ull;
        q0 = fq0[c];  // -1 on the border and on columns without forward entries
        live = q0 >= 0;
        float a[8];
        if (live) {
            float gk[8];  // every load first, then the divisions in the same order
#pragma unroll
            for (int k = 0; k < 8; k++) {
                gk[k] = g6[q0 + cornerOffset(g, k)];
                a[k] = fw[(size_t)k * g.N + c];
            }
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (gk[k] > 1.e-5f) a[k] /= gk[k];
            long long qa[8];
            to_fix_n(a, qa);
#pragma unroll
            for (int k = 0; k < 8; k++) sum += qa[k];
        }
        const float bj = (float)from_fix((unsigned long long)sum);
        g.beta[c] = bj;
        if (live) {
            const float rj = g.rho[c];
            qc = cell_coords(g, q0);
            float v[8], vr[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                v[k] = a[k];
                if (bj > 1.e-5f) v[k] /= bj;
                vr[k] = v[k] * rj;
            }
            to_fix_n(v, qg);
            to_fix_n(vr, qr);
        }
    }
    // the gamma terms, then the rho terms through the same box
    sbox_vote(box, t >> 6, live, qc.x, qc.y, qc.z);
    __syncthreads();
    const int3 o = sbox_origin(box);
    if (live) sbox_add(g, box, o, qc.x, qc.y, qc.z, q0, qg, gamma_acc);
    __syncthreads();
    sbox_flush<true>(g, box, o, t, gamma_acc);
    __syncthreads();
    if (live) sbox_add(g, box, o, qc.x, qc.y, qc.z, q0, qr, rho_acc);
    __syncthreads();
    sbox_flush(g, box, o, t, rho_acc);
}

// Steps 9-11 for the backward entries, all cells: each entry divided by its row's gamma and its
// column's beta; gamma = row sums; rho' = A rho on the interior.
__global__ __launch_bounds__(256) void k_adv_final(DevGrid g, const int* __restrict__ bp0, const float* __restrict__ bw,
                                                   const float* __restrict__ g6, unsigned long long* __restrict__ gamma_acc,
                                                   unsigned long long* __restrict__ rho_acc, float* __restrict__ rho_out) {
    FULL_XYZ(g);
    long long gs = (long long)gamma_acc[c];
    long long rs = (long long)rho_acc[c];
    gamma_acc[c] = 0ull;
    rho_acc[c] = 0ull;
    const int p0 = bp0[c];
    const bool interior = x > 0 && y > 0 && z > 0 && x < g.W - 1 && y < g.H - 1 && z < g.D - 1;
    if (interior && p0 >= 0) {
        // every load of the row first (the entries, and beta and rho at the 8 columns), then the
        // divisions and sums in the same order
        const float gi = g6[c];
        float e[8], bc[8], rc[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int j = p0 + cornerOffset(g, k);
            e[k] = bw[(size_t)k * g.N + c];
            bc[k] = g.beta[j];
            rc[k] = g.rho[j];
        }
        float a[8], ar[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            a[k] = e[k];
            if (gi > 1.e-5f) a[k] /= gi;
            if (bc[k] > 1.e-5f) a[k] /= bc[k];
            ar[k] = a[k] * rc[k];
        }
        long long qa[8], qr[8];
        to_fix_n(a, qa);
        to_fix_n(ar, qr);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            gs += qa[k];
            rs += qr[k];
        }
    }
    g.gamma[c] = (float)((double)gs * FIX_INV);
    rho_out[c] = interior ? (float)((double)rs * FIX_INV) : g.rho[c];
}

// Step 12: one pair (j = lower neighbour, i = cell) of the in-place diffusion sweep
// (CPUSimulator3D.cpp:281-303), carried in registers along the line.  Written without branches:
// the update is computed for every pair and kept only where the reference does not return early
// (the same float operations on the kept path; the discarded one may divide by zero), since a
// sweep runs about one wave per SIMD and every branch instruction lengthens its serial step.
__device__ inline void diffuse_pair(float& rj, float& gj, unsigned bj, float& ri, float& gi, unsigned bi) {
    const bool skip = (((bi | bj) & BF_EQ1) != 0) | (gj < 1e-5f) | (gi < 1e-5f) | (gj == gi);
    const float avg = (gj + gi) / 2.f;
    // mass moves from the cell of larger gamma (s) to the other (t)
    const bool jg = gj > gi;
    const float rs = jg ? rj : ri, gs = jg ? gj : gi, gt = jg ? gi : gj;
    const float dif = rs * (gs - gt) / (2.f * gs);
    const float nj = jg ? rj - dif : rj + dif;
    const float ni = jg ? ri + dif : ri - dif;
    rj = skip ? rj : nj;
    ri = skip ? ri : ni;
    gj = skip ? gj : avg;
    gi = skip ? gi : avg;
}

// Sweeps along y (AXIS 1, lines over (x, z)) and z (AXIS 2, lines over (x, y)): one thread per
// line, consecutive threads on consecutive x (coalesced), DZ_LANES active lanes per wave.  The
// loads of a line do not depend on the sweep (cell i is read before anything writes it), so they
// are issued DF cells ahead in a register ring indexed by compile-time constants.  Every access
// is a buffer access at the lane's line start (vector offset) plus the step's offset along the
// line, which is the same for all lanes and kept in a scalar register.
#ifndef EFS_DF
#define EFS_DF 8
#endif
#ifndef EFS_DZ_LANES
#define EFS_DZ_LANES 64
#endif
constexpr int DF = EFS_DF, DZ_LANES = EFS_DZ_LANES;
template <int AXIS>
__global__ __launch_bounds__(64) void k_diffuse(DevGrid g) {
    if ((int)threadIdx.x >= DZ_LANES) return;
    const int a = 1 + blockIdx.x * DZ_LANES + threadIdx.x;  // x
    const int b = 1 + blockIdx.y;
    if (a > g.W - 2) return;
    int c0, stride, len;
    if (AXIS == 1) {  // b = z
        c0 = cidx(g, a, 0, b); stride = g.W; len = g.H;
    } else {          // b = y
        c0 = cidx(g, a, b, 0); stride = g.W * g.H; len = g.D;
    }
    const __amdgpu_buffer_rsrc_t rr = __builtin_amdgcn_make_buffer_rsrc(g.rho, (short)0, g.N * 4, kBufDword3);
    const __amdgpu_buffer_rsrc_t rg = __builtin_amdgcn_make_buffer_rsrc(g.gamma, (short)0, g.N * 4, kBufDword3);
    const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc((void*)g.bf, (short)0, g.N, kBufDword3);
    const unsigned vf = (unsigned)c0 * 4u, vb = (unsigned)c0;
    auto ldr = [&](int k) { return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rr, vf, k * stride * 4, 0)); };
    auto ldg = [&](int k) { return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rg, vf, k * stride * 4, 0)); };
    auto ldb = [&](int k) { return (unsigned)__builtin_amdgcn_raw_buffer_load_b8(rb, vb, k * stride, 0); };
    auto st = [&](int k, float r, float gm) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, r), rr, vf, k * stride * 4, 0);
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, gm), rg, vf, k * stride * 4, 0);
    };
    // line positions 1 .. len-2; the pair (k-1, k) for k = 2 .. len-2
    const int last = len - 2;
    float rq[DF], gq[DF];
    unsigned bq[DF];
#pragma unroll
    for (int q = 0; q < DF; q++) {
        const int kk = min(2 + q, last);  // clamped: surplus loads are unused
        rq[q] = ldr(kk);
        gq[q] = ldg(kk);
        bq[q] = ldb(kk);
    }
    float rj = ldr(1), gj = ldg(1);
    unsigned bj = ldb(1);
    int k = 2;  // line position of cell i, paired with j = k - 1
    auto step = [&](int q) {
        float ri = rq[q], gi = gq[q];
        const unsigned bi = bq[q];
        const int kn = min(k + DF, last);
        rq[q] = ldr(kn);
        gq[q] = ldg(kn);
        bq[q] = ldb(kn);
        diffuse_pair(rj, gj, bj, ri, gi, bi);
        st(k - 1, rj, gj);  // cell j is final
        rj = ri; gj = gi; bj = bi;
        k++;
    };
    while (k + DF <= last + 1) {
#pragma unroll
        for (int q = 0; q < DF; q++) step(q);
    }
#pragma unroll
    for (int q = 0; q < DF; q++)
        if (k <= last) step(q);
    st(last, rj, gj);
}

// Sweep along x (lines over (y, z)): a line is contiguous, so consecutive threads would touch
// consecutive rows and every load would hit a different cache line.  A block instead stages a
// tile of DX_ROWS rows x DX_TILE cells through LDS with coalesced row loads; lane r walks row r
// of the tile, and the finished cells are written back coalesced.  The last cell of a tile is
// still paired with the first of the next one, so it stays in registers and is stored on its
// own.
#ifndef EFS_DX_ROWS
#define EFS_DX_ROWS 16
#endif
constexpr int DX_ROWS = EFS_DX_ROWS, DX_TILE = 64;
__global__ __launch_bounds__(64) void k_diffuse_x(DevGrid g) {
    __shared__ float sr[DX_ROWS][DX_TILE + 1], sg[DX_ROWS][DX_TILE + 1];
    __shared__ uint8_t sb[DX_ROWS][DX_TILE + 4];
    const int lane = threadIdx.x;
    const int y0 = 1 + blockIdx.x * DX_ROWS, z = 1 + blockIdx.y;
    const int row = y0 + lane;
    const bool active = lane < DX_ROWS && row <= g.H - 2;
    const int xlast = g.W - 2;
    float rj = 0.f, gj = 0.f;
    unsigned bj = 0;
    int cj = 0;
    for (int x0 = 1; x0 <= xlast; x0 += DX_TILE) {
        const int n = min(DX_TILE, xlast - x0 + 1);
        for (int r = 0; r < DX_ROWS; r++) {
            const int yy = y0 + r;
            if (yy <= g.H - 2 && lane < n) {
                const int c = cidx(g, x0 + lane, yy, z);
                sr[r][lane] = g.rho[c];
                sg[r][lane] = g.gamma[c];
                sb[r][lane] = g.bf[c];
            }
        }
        __syncthreads();
        if (active) {
            for (int t = 0; t < n; t++) {
                const int x = x0 + t;
                float ri = sr[lane][t], gi = sg[lane][t];
                const unsigned bi = sb[lane][t];
                if (x > 1) {
                    diffuse_pair(rj, gj, bj, ri, gi, bi);
                    // cell x - 1 is final
                    if (t == 0) {
                        g.rho[cj] = rj;
                        g.gamma[cj] = gj;
                    } else {
                        sr[lane][t - 1] = rj;
                        sg[lane][t - 1] = gj;
                    }
                }
                rj = ri; gj = gi; bj = bi;
                cj = cidx(g, x, row, z);
            }
        }
        __syncthreads();
        for (int r = 0; r < DX_ROWS; r++) {
            const int yy = y0 + r;
            if (yy <= g.H - 2 && lane < n - 1) {
                const int c = cidx(g, x0 + lane, yy, z);
                g.rho[c] = sr[r][lane];
                g.gamma[c] = sg[r][lane];
            }
        }
        __syncthreads();
    }
    if (active) {
        g.rho[cj] = rj;
        g.gamma[cj] = gj;
    }
}

// ---- vel_advect (CPUSimulator3D.cpp:478-504) ----------------------------------------------
__global__ __launch_bounds__(256) void k_vadv_trace(DevGrid g, float dt, float* __restrict__ tu, float* __restrict__ tv,
                                                    float* __restrict__ tw) {
    INTERIOR_XYZ(g);
    const float3 prev = traceParticle(g, (float)x, (float)y, (float)z, -dt, 4);
    const float3 v = vinterp(g, prev.x - 0.5f, prev.y - 0.5f, prev.z - 0.5f);
    tu[c] = v.x;
    tv[c] = v.y;
    tw[c] = v.z;
}

// With FORCE the add_force stage that follows (k_add_force) is applied to the new face values
// before they are stored, in the same operations.
template <bool FORCE>
__global__ __launch_bounds__(256) void k_vadv_faces(DevGrid g, const float* __restrict__ tu, const float* __restrict__ tv,
                                                    const float* __restrict__ tw, float dt, float gdt, int gravity) {
    INTERIOR_XYZ_XCD(g);
    const int sy = g.W, sz = g.W * g.H;
    // every load first (all in the array for an interior cell), then selects: tests that guard
    // loads would serialise them behind one another
    const unsigned b0 = g.bf[c], bxm = g.bf[c - 1], bym = g.bf[c - sy], bzm = g.bf[c - sz];
    const float u0 = tu[c], um = tu[c - 1], v0 = tv[c], vm = tv[c - sy], w0 = tw[c], wm = tw[c - sz];
    float f0 = 0.f, f1 = 0.f, f2 = 0.f;
    if (FORCE && g.fu) {
        f0 = g.fu[c];
        f1 = g.fv[c];
        f2 = g.fw[c];
    }
    const bool bs = (b0 & BF_EQ1) != 0;
    float u = ((x == 1) | bs | ((bxm & BF_EQ1) != 0)) ? u0 : 0.5f * (u0 + um);
    float v = ((y == 1) | bs | ((bym & BF_EQ1) != 0)) ? v0 : 0.5f * (v0 + vm);
    float w = ((z == 1) | bs | ((bzm & BF_EQ1) != 0)) ? w0 : 0.5f * (w0 + wm);
    if (FORCE) {
        // vel + f dt (vel + 0 dt with no force field), then gravity; solid cells keep the face value
        float fu = u + f0 * dt, fv = v + f1 * dt, fw = w + f2 * dt;
        if (gravity) fv -= gdt;
        u = bs ? u : fu;
        v = bs ? v : fv;
        w = bs ? w : fw;
    }
    g.u[c] = u;
    g.v[c] = v;
    g.w[c] = w;
}

// ---- add_force (CPUSimulator3D.cpp:458-476) -----------------------------------------------
__global__ __launch_bounds__(256) void k_add_force(DevGrid g, float dt, float gdt, int gravity) {
    INTERIOR_XYZ(g);
    if (g.bf[c] & BF_EQ1) return;
    float u = g.u[c], v = g.v[c], w = g.w[c];
    if (g.fu) {
        u = u + g.fu[c] * dt;
        v = v + g.fv[c] * dt;
        w = w + g.fw[c] * dt;
    } else {  // zero forces: vel + 0 * dt
        u = u + 0.f * dt; v = v + 0.f * dt; w = w + 0.f * dt;
    }
    if (gravity) v -= gdt;
    g.u[c] = u; g.v[c] = v; g.w[c] = w;
}

// ---- project (CPUSimulator3D.cpp:506-566, 657-679) ----------------------------------------
// one-sided wall enforcement of the faces stored at cell c (x, y, z >= 1)
__device__ inline float wall_face(float f, bool self_solid, bool lower_solid) {
    if (self_solid && f > 0) f = 0.f;
    if (lower_solid && f < 0) f = 0.f;
    return f;
}

// enforce_wall_boundaries + divergence with density correction (CPUSimulator3D.cpp:516-539,
// 657-679); the outer layer's divergence is 0.  The divergence of cell c reads the +x/+y/+z faces,
// enforced on the fly (enforcement is idempotent, so reading a face another thread has already
// enforced gives the same value).  A face is stored only where the enforcement changed it (a face
// pointing into a solid), which few are.  A thread marches up WD_ZC planes of one (x, y) column:
// the cell's boundary byte and w face come from the previous plane's +z loads, and every load of
// the next plane is issued before the current plane is computed (addresses clamped into the array,
// values that a cell does not use are loaded and ignored), so the loads of a plane wait for nothing
// (131 -> 103 us at 256^3 against one thread per cell, bit-identical).  The blocks take (tile, z
// chunk) in XCD order, so neighbouring tiles' rows are read through the same L2.
constexpr int WD_ZC = 16;
__global__ __launch_bounds__(256) void k_walls_div_z(DevGrid g, float liq, float lp, float lm, float eta) {
    XcdBlock blk;
    if (!xcd_block((g.W + BX - 1) / BX, (g.H + BY - 1) / BY, blk)) return;
    const int x = blk.bx * BX + threadIdx.x, y = blk.by * BY + threadIdx.y;
    if (x >= g.W || y >= g.H) return;
    const int zb = blk.bz * WD_ZC, ze = min(zb + WD_ZC, g.D);
    const int sy = g.W, sz = g.W * g.H;
    if (x == 0 || y == 0) {
        for (int z = zb; z < ze; z++) g.div[cidx(g, x, y, z)] = 0.f;
        return;
    }
    const int last = g.N - 1;
    const int dxp = x < g.W - 1 ? 1 : 0, dyp = y < g.H - 1 ? sy : 0;
    struct Plane {
        float u, up, v, vp, wp, r;
        unsigned bxm, bxp, bym, byp, bzp;
    };
    auto load = [&](int z, Plane& q) {  // plane z (clamped into the array)
        const int c = min(cidx(g, x, y, min(z, g.D - 1)), last);
        const int czp = min(c + sz, last);
        q.u = g.u[c];
        q.up = g.u[c + dxp];
        q.v = g.v[c];
        q.vp = g.v[c + dyp];
        q.wp = g.w[czp];
        q.r = g.rho[c];
        q.bxm = g.bf[c - 1];
        q.bxp = g.bf[c + dxp];
        q.bym = g.bf[c - sy];
        q.byp = g.bf[c + dyp];
        q.bzp = g.bf[czp];
    };
    // carried along z: the cell's boundary byte and w face, and the -z neighbour's boundary byte
    const int c0 = cidx(g, x, y, zb);
    unsigned bc = g.bf[c0], bzm = zb > 0 ? (unsigned)g.bf[c0 - sz] : 0u;
    float wc = g.w[c0];
    Plane q, qn;
    load(zb, q);
    for (int z = zb; z < ze; z++) {
        load(z + 1, qn);
        const int c = cidx(g, x, y, z);
        if (z == 0) {
            g.div[c] = 0.f;
        } else {
            const bool s0 = bc & BF_EQ1;
            const float u0 = wall_face(q.u, s0, q.bxm & BF_EQ1);
            const float v0 = wall_face(q.v, s0, q.bym & BF_EQ1);
            const float w0 = wall_face(wc, s0, bzm & BF_EQ1);
            if (u0 != q.u) g.u[c] = u0;
            if (v0 != q.v) g.v[c] = v0;
            if (w0 != wc) g.w[c] = w0;
            float dv = 0.f;
            const bool interior = x < g.W - 1 && y < g.H - 1 && z < g.D - 1;
            if (interior && !(q.r < liq) && !s0) {
                const float u1 = wall_face(q.up, q.bxp & BF_EQ1, s0);
                const float v1 = wall_face(q.vp, q.byp & BF_EQ1, s0);
                const float w1 = wall_face(q.wp, q.bzp & BF_EQ1, s0);
                const float dx = u1 - u0, dy = v1 - v0, dz = w1 - w0;
                dv = dx + dy + dz;
                if (q.r > 1.f) dv -= fminr(lm * (q.r - 1.f), eta);
                if (q.r < 1.f) dv += fminr(lp * (1.f - q.r), eta);
            }
            g.div[c] = dv;
        }
        bzm = bc;
        bc = q.bzp;
        wc = q.wp;
        q = qn;
    }
}

// u -= grad p on non-solid interior cells (x>1 / y>1 / z>1 guards), then enforce the faces.
// p64 (optional) is the CG solver's FP64 iterate; the FP32 pressure field is its rounding.
__global__ __launch_bounds__(256) void k_gradient_walls(DevGrid g, const double* __restrict__ p64) {
    FULL_XYZ_XCD(g);
    if (x == 0 || y == 0 || z == 0) return;
    const int sy = g.W, sz = g.W * g.H;
    // every load first (the -x/-y/-z neighbours are in the array for x, y, z >= 1), then selects
    const unsigned b0 = g.bf[c], bxm = g.bf[c - 1], bym = g.bf[c - sy], bzm = g.bf[c - sz];
    float u = g.u[c], v = g.v[c], w = g.w[c];
    const bool s0 = (b0 & BF_EQ1) != 0;
    const bool upd = (x < g.W - 1) & (y < g.H - 1) & (z < g.D - 1) & !s0;
    if (p64) {
        const double pc = p64[c], pxm = p64[c - 1], pym = p64[c - sy], pzm = p64[c - sz];
        const double gx = (x > 1) ? pc - pxm : 0.0;
        const double gy = (y > 1) ? pc - pym : 0.0;
        const double gz = (z > 1) ? pc - pzm : 0.0;
        const float nu = (float)((double)u - gx), nv = (float)((double)v - gy), nw = (float)((double)w - gz);
        u = upd ? nu : u; v = upd ? nv : v; w = upd ? nw : w;
    } else {
        const float pc = g.p[c], pxm = g.p[c - 1], pym = g.p[c - sy], pzm = g.p[c - sz];
        const float gx = (x > 1) ? pc - pxm : 0.f;
        const float gy = (y > 1) ? pc - pym : 0.f;
        const float gz = (z > 1) ? pc - pzm : 0.f;
        const float nu = u - gx, nv = v - gy, nw = w - gz;
        u = upd ? nu : u; v = upd ? nv : v; w = upd ? nw : w;
    }
    g.u[c] = wall_face(u, s0, (bxm & BF_EQ1) != 0);
    g.v[c] = wall_face(v, s0, (bym & BF_EQ1) != 0);
    g.w[c] = wall_face(w, s0, (bzm & BF_EQ1) != 0);
}

inline int blocks1d(int n) { return (n + 255) / 256; }

}  // namespace

// =========================================================================================
// launchers
// =========================================================================================
int efs_launch_boundary_flags(efs_sim* s) {
    hipLaunchKernelGGL(k_boundary_flags, dim3(blocks1d(s->N)), dim3(256), 0, s->stream, s->N, s->bndf, s->bf);
    EFS_HIP(hipGetLastError());
    return 0;
}

int efs_launch_decay(efs_sim* s) {
    hipLaunchKernelGGL(k_decay, interior_grid(s), dim3(BX, BY), 0, s->stream, s->grid(), s->P.vel_decay);
    EFS_HIP(hipGetLastError());
    return 0;
}

int efs_launch_extrapolate(efs_sim* s, bool with_decay) {
    const DevGrid g = s->grid();
    const dim3 grd = interior_grid(s), blk(BX, BY);
    if (with_decay)
        hipLaunchKernelGGL(k_extrap_init<true>, full_grid(s), blk, 0, s->stream, g, s->valA, s->P.vel_decay);
    else
        hipLaunchKernelGGL(k_extrap_init<false>, full_grid(s), blk, 0, s->stream, g, s->valA, 1.f);
    for (int pass = 1; pass <= 6; pass++)  // the reference's 3 double passes
        hipLaunchKernelGGL(k_extrap_pass, xcd_grid((s->W - 2 + BX - 1) / BX, (s->H - 2 + BY - 1) / BY, (s->D - 2 + EXZ - 1) / EXZ),
                           blk, 0, s->stream, g, s->valA, pass);
    EFS_HIP(hipGetLastError());
    return 0;
}

int efs_launch_sharpen(efs_sim* s) {
    const DevGrid g = s->grid();
    const dim3 grd = interior_grid(s), blk(BX, BY);
    const float delT = s->P.dt * 3;
    hipLaunchKernelGGL(k_sharpen_delta, grd, blk, 0, s->stream, g, delT, s->s0);
    hipLaunchKernelGGL(k_sharpen_apply, grd, blk, 0, s->stream, g, s->s0);
    hipLaunchKernelGGL(k_sharpen_scatter, grd, blk, 0, s->stream, g, s->s0, s->P.rho_d, s->acc0);
    hipLaunchKernelGGL(k_add_fix, dim3(blocks1d(s->N)), dim3(256), 0, s->stream, s->N, s->rho, s->acc0);
    EFS_HIP(hipGetLastError());
    return 0;
}

int efs_launch_rho_advect(efs_sim* s) {
    const DevGrid g = s->grid();
    const dim3 gf = full_grid(s), blk(BX, BY);
    const dim3 ai = adv_interior_grid(s), af = adv_full_grid(s), ablk(ABX, ABY);
    const dim3 g1(blocks1d(s->N)), b1(256);
    const float dt = s->P.dt;
    // steps 1-4
    hipEvent_t ta, tb;
    efs_ktimer_pair(s, &ta, &tb);
    hipExtLaunchKernelGGL(k_adv_bwd, ai, ablk, 0, s->stream, ta, tb, 0, g, dt, s->s1, s->bwd_p0, s->bwd_w, s->acc0, s->tu, s->tv,
                          s->tw);
    efs_ktimer_add(s, KT_ADV_BWD, ta, tb);
    // step 5 (+ forward half of step 6)
    efs_ktimer_pair(s, &ta, &tb);
    hipExtLaunchKernelGGL(k_adv_fwd, ai, ablk, 0, s->stream, ta, tb, 0, g, dt, s->acc0, s->fwd_p0, s->fwd_w, s->acc1);
    efs_ktimer_add(s, KT_ADV_FWD, ta, tb);
    hipLaunchKernelGGL(k_fill_u64, g1, b1, 0, s->stream, s->N, s->acc0);
    // steps 6-8: gamma6 -> s1 (the advected gamma of step 1 is no longer needed), new beta
    hipLaunchKernelGGL(k_adv_rows, af, ablk, 0, s->stream, g, s->bwd_p0, s->bwd_w, s->acc1, s->s1, s->acc0);
    // steps 7-11 for the forward entries, and beta
    hipLaunchKernelGGL(k_adv_fwd_cols, af, ablk, 0, s->stream, g, s->s1, s->fwd_p0, s->fwd_w, s->acc0, s->acc1, s->acc2);
    // steps 9-11 for the backward entries
    hipLaunchKernelGGL(k_adv_final, gf, blk, 0, s->stream, g, s->bwd_p0, s->bwd_w, s->s1, s->acc1, s->acc2, s->s0);
    EFS_HIP(hipGetLastError());
    s->traces_valid = true;
    std::swap(s->rho, s->s0);
    // step 12: 3 iterations of x, y, z sweeps
    const DevGrid g2 = s->grid();
    for (int it = 0; it < 3; it++) {
        hipLaunchKernelGGL(k_diffuse_x, dim3((s->H - 2 + DX_ROWS - 1) / DX_ROWS, s->D - 2), dim3(64), 0, s->stream, g2);
        hipLaunchKernelGGL(k_diffuse<1>, dim3((s->W - 2 + DZ_LANES - 1) / DZ_LANES, s->D - 2), dim3(64), 0, s->stream, g2);
        hipLaunchKernelGGL(k_diffuse<2>, dim3((s->W - 2 + DZ_LANES - 1) / DZ_LANES, s->H - 2), dim3(64), 0, s->stream, g2);
    }
    EFS_HIP(hipGetLastError());
    return 0;
}

int efs_launch_vel_advect(efs_sim* s, bool with_force) {
    const DevGrid g = s->grid();
    const dim3 grd = interior_grid(s), blk(BX, BY);
    // the samples at the backward traces are in tu/tv/tw already when rho_advect has just run
    if (!s->traces_valid) hipLaunchKernelGGL(k_vadv_trace, grd, blk, 0, s->stream, g, s->P.dt, s->tu, s->tv, s->tw);
    s->traces_valid = false;
    if (with_force)
        hipLaunchKernelGGL(k_vadv_faces<true>, interior_grid_xcd(s), blk, 0, s->stream, g, s->tu, s->tv, s->tw, s->P.dt,
                           s->P.gravity_force * s->P.dt, s->P.gravity_enabled);
    else
        hipLaunchKernelGGL(k_vadv_faces<false>, interior_grid_xcd(s), blk, 0, s->stream, g, s->tu, s->tv, s->tw, 0.f, 0.f, 0);
    EFS_HIP(hipGetLastError());
    return 0;
}

int efs_launch_add_force(efs_sim* s) {
    const float gdt = s->P.gravity_force * s->P.dt;
    hipLaunchKernelGGL(k_add_force, interior_grid(s), dim3(BX, BY), 0, s->stream, s->grid(), s->P.dt, gdt,
                       s->P.gravity_enabled);
    EFS_HIP(hipGetLastError());
    return 0;
}

int efs_launch_walls_divergence(efs_sim* s) {
    hipEvent_t ta, tb;
    efs_ktimer_pair(s, &ta, &tb);
    hipExtLaunchKernelGGL(k_walls_div_z, xcd_grid((s->W + BX - 1) / BX, (s->H + BY - 1) / BY, (s->D + WD_ZC - 1) / WD_ZC),
                          dim3(BX, BY), 0, s->stream, ta, tb, 0, s->grid(), s->P.liquid_threshold, s->P.lambda_plus,
                          s->P.lambda_minus, s->P.eta);
    efs_ktimer_add(s, KT_WALLS_DIV, ta, tb);
    EFS_HIP(hipGetLastError());
    return 0;
}

int efs_launch_gradient_walls(efs_sim* s) {
    hipEvent_t ta, tb;
    efs_ktimer_pair(s, &ta, &tb);
    hipExtLaunchKernelGGL(k_gradient_walls, full_grid_xcd(s), dim3(BX, BY), 0, s->stream, ta, tb, 0, s->grid(),
                          s->p64_valid ? (const double*)s->cg_x64 : nullptr);
    efs_ktimer_add(s, KT_GRADIENT, ta, tb);
    EFS_HIP(hipGetLastError());
    return 0;
}
