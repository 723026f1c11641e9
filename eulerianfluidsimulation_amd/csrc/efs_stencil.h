// Device helpers shared by the pressure solvers: launch domains with XCD-aware block
// numbering, the 7-point operator row, LDS plane staging, and fence-free block reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "efs_internal.h"

namespace efs_stencil {

constexpr int TX = 64;

struct Dom {
    int W, H, D;        // grid dims
    int x0, y0, z0;     // first cell of the launch domain
    int nx, ny, nz;     // extent of the launch domain
    int ty;             // tile height (rows per block)
    int tiles_x, tiles_y, chunks_z, zc;
    int nblocks;        // real blocks (the launch rounds up to a multiple of 8)
};

__device__ inline bool block_coords(const Dom& d, int& bx, int& by, int& bz) {
    // XCD-aware renumbering: hardware block b runs on XCD b % 8; XCD k gets the contiguous
    // logical range [k * per, (k + 1) * per).
    const int b = blockIdx.x;
    const int per = (d.nblocks + 7) / 8;
    const int L = (b % 8) * per + b / 8;
    if (L >= d.nblocks) return false;
    bx = L % d.tiles_x;
    const int r = L / d.tiles_x;
    by = r % d.tiles_y;
    bz = r / d.tiles_y;
    return true;
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Row of the operator.  The reference sums the contributing neighbours in the order
// x-, x+, y-, y+, z+, z- starting from 0.0f, then subtracts nCount * x (no contraction).
__device__ inline float apply_row(unsigned rc, float xc, float xm, float xp, float ym, float yp, float zp, float zm) {
    if (!(rc & RC_CG_ROW)) return xc;
    float sum = 0.0f;
    if (rc & RC_NB_XM) sum += xm;
    if (rc & RC_NB_XP) sum += xp;
    if (rc & RC_NB_YM) sum += ym;
    if (rc & RC_NB_YP) sum += yp;
    if (rc & RC_NB_ZP) sum += zp;
    if (rc & RC_NB_ZM) sum += zm;
    return sum - (float)((rc >> RC_CNT_SHIFT) & 7u) * xc;
}

__device__ inline double apply_row_d(unsigned rc, double xc, double xm, double xp, double ym, double yp, double zp, double zm) {
    if (!(rc & RC_CG_ROW)) return xc;
    double sum = 0.0;
    if (rc & RC_NB_XM) sum += xm;
    if (rc & RC_NB_XP) sum += xp;
    if (rc & RC_NB_YM) sum += ym;
    if (rc & RC_NB_YP) sum += yp;
    if (rc & RC_NB_ZP) sum += zp;
    if (rc & RC_NB_ZM) sum += zm;
    return sum - (double)((rc >> RC_CNT_SHIFT) & 7u) * xc;
}

struct Nb {  // neighbour offsets, clamped into the array on the outer layer (values unused there)
    int xm, xp, ym, yp, zm, zp;
};
__device__ inline Nb nb_offsets(const Dom& d, int x, int y, int z) {
    const int sy = d.W, sz = d.W * d.H;
    Nb n;
    n.xm = (x > 0) ? -1 : 0;
    n.xp = (x < d.W - 1) ? 1 : 0;
    n.ym = (y > 0) ? -sy : 0;
    n.yp = (y < d.H - 1) ? sy : 0;
    n.zm = (z > 0) ? -sz : 0;
    n.zp = (z < d.D - 1) ? sz : 0;
    return n;
}

// ---------------------------------------------------------------------------------------
// Plane staging: thread (tx, ty) owns LDS slot (ty+1, tx+1); edge threads also fill the
// one-cell halo.  Cells outside the array read 0.
// ---------------------------------------------------------------------------------------
template <int TY>
struct Stager {
    int hx_slot, hy_slot;  // LDS column / row of this thread's halo cell (-1 = none)
    int hx_off, hy_off;    // array offset of the halo cell from the own cell
    bool own_ok, hx_ok, hy_ok;
    __device__ void init(const Dom& d, int x, int y) {
        const int tx = threadIdx.x, ty = threadIdx.y;
        own_ok = x < d.W && y < d.H;
        hx_slot = -1; hy_slot = -1; hx_off = 0; hy_off = 0; hx_ok = false; hy_ok = false;
        if (tx == 0) { hx_slot = 0; hx_off = -1; hx_ok = x - 1 >= 0 && x - 1 < d.W && y < d.H; }
        if (tx == TX - 1) { hx_slot = TX + 1; hx_off = 1; hx_ok = x + 1 < d.W && y < d.H; }
        if (ty == 0) { hy_slot = 0; hy_off = -d.W; hy_ok = y - 1 >= 0 && y - 1 < d.H && x < d.W; }
        if (ty == TY - 1) { hy_slot = TY + 1; hy_off = d.W; hy_ok = y + 1 < d.H && x < d.W; }
    }
};

constexpr int CG_TY = 8, CG_NT = TX * CG_TY;

// Workgroup barrier that orders LDS only.  __syncthreads() also fences global memory at
// workgroup scope, which waits for every outstanding global load (s_waitcnt vmcnt(0)) and so
// drains the plane prefetch at each plane; the stencil kernels share data between threads
// through LDS only.
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// sum of the n partials of the previous kernel, identical in every block
__device__ inline double sum_partials(const double* __restrict__ p, int n, double* red) {
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
    double s = 0.0;
    for (int i = lane; i < n; i += CG_NT) s += p[i];
    s = wave_sum(s);
    if ((lane & 63) == 0) red[lane >> 6] = s;
    __syncthreads();
    double tot = 0.0;
#pragma unroll
    for (int w = 0; w < CG_NT / 64; w++) tot += red[w];
    __syncthreads();
    return tot;
}

// this block's total of v, written to p[blockIdx.x] (lane 0)
__device__ inline void write_partial(double v, double* __restrict__ p, double* red) {
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
    v = wave_sum(v);
    if ((lane & 63) == 0) red[lane >> 6] = v;
    __syncthreads();
    if (lane == 0) {
        double tot = 0.0;
#pragma unroll
        for (int w = 0; w < CG_NT / 64; w++) tot += red[w];
        p[blockIdx.x] = tot;
    }
}


inline Dom make_dom(const efs_sim* s, bool full, int ty, int target_blocks) {
    Dom d;
    d.W = s->W; d.H = s->H; d.D = s->D;
    if (full) {
        d.x0 = d.y0 = d.z0 = 0;
        d.nx = s->W; d.ny = s->H; d.nz = s->D;
    } else {
        d.x0 = d.y0 = d.z0 = 1;
        d.nx = s->W - 2; d.ny = s->H - 2; d.nz = s->D - 2;
    }
    d.ty = ty;
    d.tiles_x = (d.nx + TX - 1) / TX;
    d.tiles_y = (d.ny + ty - 1) / ty;
    const int txy = d.tiles_x * d.tiles_y;
    int zc = (int)std::max<long long>(2, std::min<long long>(32, (long long)d.nz * txy / target_blocks));
    zc = std::min(zc, std::max(1, d.nz));
    d.zc = zc;
    d.chunks_z = (d.nz + zc - 1) / zc;
    d.nblocks = txy * d.chunks_z;
    return d;
}
inline Dom dom_simple(const efs_sim* s, bool full) { return make_dom(s, full, 4, 2048); }
inline Dom dom_tiled(const efs_sim* s, bool full) { return make_dom(s, full, CG_TY, 2048); }

inline int launch_blocks(const Dom& d) { return ((d.nblocks + 7) / 8) * 8; }
inline int blocks1d(int n) { return (n + 255) / 256; }

}  // namespace efs_stencil
using namespace efs_stencil;

// flexible CG's direction pass (efs_solver.hip): d_it = z + beta d_{it-1}, q = A d_it, partials of d.q
int efs_fcg_dir(efs_sim* s, const Dom& dt, int it, const float* z, const double* p_zq, float* q);
