// Pressure Poisson solve on gfx950: the matrix-free 7-point operator of CGSolverCPU3D
// (CGSolver.cpp:467-497), CG with FP32 vectors and FP64 dot products / alpha / beta
// (CGSolver.cpp:416-465 restated), Jacobi (JacobiSolver.cpp:154-212) and an FP64 true residual.
//
// Kernel shape: a block is a 64 x 4 tile of an xy plane (one wavefront per row) that marches
// along z over a chunk of planes, keeping the centre column z-1, z, z+1 in registers; lateral
// neighbours come from L1/L2.  Blocks are renumbered so each XCD owns a contiguous run of tiles
// (workgroups are dealt to the 8 XCDs round-robin), which keeps the shared halo rows in one L2.
//
// CG iteration = 2 kernels, each with one grid-level FP64 reduction finished by the last block:
//   K1  d_i = r_i + beta d_{i-1} (computed on the fly for the stencil, stored once), t = A d_i,
//       partial d.t                          -> last block: alpha = gamma / d.t
//   K2  t = A d_i recomputed from d_i, x_{i+1} = x_i + alpha d_i,
//       r_{i+1} = r_i - alpha t  (or b - A x_{i+1} when i % 64 == 0, CGSolver.cpp:444-451),
//       partial r.r                          -> last block: convergence test, beta, iter++
// x and d are double-buffered by iteration parity, so the residual replacement can read
// neighbour x_{i+1} = x_i + alpha d_i without racing with the writes.  Every launch has the same
// arguments (the iteration lives in device memory), and a converged solve turns the remaining
// launches of a batch into no-ops, so the host checks for convergence once per batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "efs_internal.h"

namespace {

constexpr int TX = 64, TY = 4, NT = TX * TY;

struct Dom {
    int W, H, D;        // grid dims
    int x0, y0, z0;     // first cell of the launch domain
    int nx, ny, nz;     // extent of the launch domain
    int tiles_x, tiles_y, chunks_z, zc;  // tiling
    int nblocks;        // real blocks (the launch may round up to a multiple of 8)
    int full;           // 1 = domain touches the outer layer (neighbour indices need clamping)
};

__device__ inline bool block_coords(const Dom& d, int& bx, int& by, int& bz) {
    // XCD-aware renumbering: hardware block b runs on XCD b % 8; give XCD k the contiguous
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

// Block sum (256 threads = 4 waves) deposited in partials[b]; the last block to arrive reduces
// all partials in a fixed order and runs `fin` on thread 0.  Returns true in the last block.
template <int NV>
__device__ inline bool grid_reduce(double (&v)[NV], double* partials, unsigned int* counter, int nb, double (&out)[NV]) {
    __shared__ double red[NV][NT / 64];
    __shared__ bool last;
    const int lane = threadIdx.x + threadIdx.y * TX;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        const double s = wave_sum(v[k]);
        if ((lane & 63) == 0) red[k][lane >> 6] = s;
    }
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double s = 0;
            for (int wv = 0; wv < NT / 64; wv++) s += red[k][wv];
            partials[(size_t)k * gridDim.x + blockIdx.x] = s;
        }
        __threadfence();
        const unsigned int t = atomicAdd(counter, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    // deterministic: each thread sums a fixed strided subset, then a fixed tree
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = 0;
        for (int i = lane; i < (int)gridDim.x; i += NT) s += __builtin_nontemporal_load(&partials[(size_t)k * gridDim.x + i]);
        s = wave_sum(s);
        if ((lane & 63) == 0) red[k][lane >> 6] = s;
    }
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double s = 0;
            for (int wv = 0; wv < NT / 64; wv++) s += red[k][wv];
            out[k] = s;
        }
        *counter = 0;
    }
    (void)nb;
    return true;
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

// Same row in FP64 (true residual).
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
// row codes: neighbour counted = in range, not on the outer layer, boundary <= 0.5;
// contributing = counted and rho > 0.5; CG row = !(rho <= 0.5 || boundary == 1);
// Jacobi row = !(rho <= 0.5 || boundary >= 0.5).
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_rowcode(DevGrid g, uint16_t* rc, CGState* st) {
    const int x = blockIdx.x * TX + threadIdx.x, y = blockIdx.y * TY + threadIdx.y, z = blockIdx.z;
    if (x >= g.W || y >= g.H) return;
    const int c = cidx(g, x, y, z);
    const float r = g.rho[c];
    const unsigned b = g.bf[c];
    unsigned code = 0, cnt = 0;
    auto nb = [&](int nx, int ny, int nz, unsigned bit) {
        if (nx < 0 || nx >= g.W || ny < 0 || ny >= g.H || nz < 0 || nz >= g.D) return;
        if (nx == 0 || nx == g.W - 1 || ny == 0 || ny == g.H - 1 || nz == 0 || nz == g.D - 1) return;
        const int q = cidx(g, nx, ny, nz);
        if (g.bf[q] & BF_GT05) return;
        cnt++;
        if (g.rho[q] > 0.5f) code |= bit;
    };
    nb(x - 1, y, z, RC_NB_XM);
    nb(x + 1, y, z, RC_NB_XP);
    nb(x, y - 1, z, RC_NB_YM);
    nb(x, y + 1, z, RC_NB_YP);
    nb(x, y, z + 1, RC_NB_ZP);
    nb(x, y, z - 1, RC_NB_ZM);
    code |= cnt << RC_CNT_SHIFT;
    if (!(r <= 0.5f || (b & BF_EQ1))) code |= RC_CG_ROW;
    if (!(r <= 0.5f || (b & BF_GE05))) code |= RC_JAC_ROW;
    rc[c] = (uint16_t)code;
    const bool border = x == 0 || y == 0 || z == 0 || x == g.W - 1 || y == g.H - 1 || z == g.D - 1;
    if (border && (code & (RC_CG_ROW | RC_JAC_ROW))) atomicOr(&st->flags, 1);
}

// ---------------------------------------------------------------------------------------
// CG
// ---------------------------------------------------------------------------------------
// init: x0 = b, r0 = b - A b, d0 = r0; reduce r.r and b.b
__global__ __launch_bounds__(256) void k_cg_init(Dom d, const uint16_t* __restrict__ rcode, const float* __restrict__ b,
                                                 float* __restrict__ r, float* __restrict__ d0, CGState* st,
                                                 double* partials) {
    int bx, by, bz;
    double v[2] = {0.0, 0.0};
    if (block_coords(d, bx, by, bz)) {
        const int x = d.x0 + bx * TX + threadIdx.x, y = d.y0 + by * TY + threadIdx.y;
        const int zb = d.z0 + bz * d.zc, ze = min(zb + d.zc, d.z0 + d.nz);
        if (x < d.x0 + d.nx && y < d.y0 + d.ny) {
            for (int z = zb; z < ze; z++) {
                const int c = (z * d.H + y) * d.W + x;
                const Nb n = nb_offsets(d, x, y, z);
                const float bc = b[c];
                const float t = apply_row(rcode[c], bc, b[c + n.xm], b[c + n.xp], b[c + n.ym], b[c + n.yp], b[c + n.zp], b[c + n.zm]);
                const float rr = bc - t;
                r[c] = rr;
                d0[c] = rr;
                v[0] += (double)rr * (double)rr;
                v[1] += (double)bc * (double)bc;
            }
        }
    }
    double out[2];
    if (grid_reduce<2>(v, partials, &st->counter, d.nblocks, out)) {
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            st->gamma = out[0];
            st->bnorm2 = out[1];
            st->thr = (st->flags & 2) ? st->thr * st->thr * out[1] : st->thr;
            st->iter = 0;
            st->converged = (out[0] < st->thr) ? 1 : 0;
            st->done = st->converged || st->max_iter <= 0 || !isfinite(out[0]);
        }
    }
}

// K1: direction update + A d + d.Ad
__global__ __launch_bounds__(256) void k_cg_dir_spmv(Dom d, const uint16_t* __restrict__ rcode, const float* __restrict__ r,
                                                     float* d0, float* d1, CGState* st, double* partials) {
    if (st->done) return;
    const int it = st->iter;
    const double beta = st->beta;
    const bool first = (it == 0);
    float* dnew = (it & 1) ? d1 : d0;
    const float* dold = (it & 1) ? d0 : d1;
    auto D = [&](int c) -> float {
        if (first) return dnew[c];
        return (float)((double)dold[c] * beta + (double)r[c]);
    };
    int bx, by, bz;
    double v[1] = {0.0};
    if (block_coords(d, bx, by, bz)) {
        const int x = d.x0 + bx * TX + threadIdx.x, y = d.y0 + by * TY + threadIdx.y;
        const int zb = d.z0 + bz * d.zc, ze = min(zb + d.zc, d.z0 + d.nz);
        if (x < d.x0 + d.nx && y < d.y0 + d.ny) {
            const int sz = d.W * d.H;
            int c = (zb * d.H + y) * d.W + x;
            float cm = (zb > 0) ? D(c - sz) : 0.f;
            float cc = D(c);
            for (int z = zb; z < ze; z++, c += sz) {
                const Nb n = nb_offsets(d, x, y, z);
                const float cp = D(c + n.zp);
                const unsigned rc = rcode[c];
                float t = cc;
                if (rc & RC_CG_ROW) t = apply_row(rc, cc, D(c + n.xm), D(c + n.xp), D(c + n.ym), D(c + n.yp), cp, cm);
                if (!first) dnew[c] = cc;
                v[0] += (double)cc * (double)t;
                cm = cc;
                cc = cp;
            }
        }
    }
    double out[1];
    if (grid_reduce<1>(v, partials, &st->counter, d.nblocks, out)) {
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            st->dAd = out[0];
            st->alpha = st->gamma / out[0];
        }
    }
}

// K2: x, r update + r.r
__global__ __launch_bounds__(256) void k_cg_update(Dom d, const uint16_t* __restrict__ rcode, const float* __restrict__ b,
                                                   float* r, const float* d0, const float* d1, float* x0b, float* x1b,
                                                   CGState* st, double* partials) {
    if (st->done) return;
    const int it = st->iter;
    const double alpha = st->alpha;
    const float* dd = (it & 1) ? d1 : d0;
    const float* xin = (it & 1) ? x1b : x0b;
    float* xout = (it & 1) ? x0b : x1b;
    const bool replace = (it & 0x3F) == 0;
    auto X = [&](int c) -> float { return (float)((double)xin[c] + alpha * (double)dd[c]); };
    int bx, by, bz;
    double v[1] = {0.0};
    if (block_coords(d, bx, by, bz)) {
        const int x = d.x0 + bx * TX + threadIdx.x, y = d.y0 + by * TY + threadIdx.y;
        const int zb = d.z0 + bz * d.zc, ze = min(zb + d.zc, d.z0 + d.nz);
        if (x < d.x0 + d.nx && y < d.y0 + d.ny) {
            const int sz = d.W * d.H;
            int c = (zb * d.H + y) * d.W + x;
            if (!replace) {
                float cm = (zb > 0) ? dd[c - sz] : 0.f;
                float cc = dd[c];
                for (int z = zb; z < ze; z++, c += sz) {
                    const Nb n = nb_offsets(d, x, y, z);
                    const float cp = dd[c + n.zp];
                    const unsigned rc = rcode[c];
                    float t = cc;
                    if (rc & RC_CG_ROW) t = apply_row(rc, cc, dd[c + n.xm], dd[c + n.xp], dd[c + n.ym], dd[c + n.yp], cp, cm);
                    xout[c] = (float)((double)xin[c] + alpha * (double)cc);
                    const float rn = (float)((double)r[c] - alpha * (double)t);
                    r[c] = rn;
                    v[0] += (double)rn * (double)rn;
                    cm = cc;
                    cc = cp;
                }
            } else {
                float cm = (zb > 0) ? X(c - sz) : 0.f;
                float cc = X(c);
                for (int z = zb; z < ze; z++, c += sz) {
                    const Nb n = nb_offsets(d, x, y, z);
                    const float cp = X(c + n.zp);
                    const unsigned rc = rcode[c];
                    float t = cc;
                    if (rc & RC_CG_ROW) t = apply_row(rc, cc, X(c + n.xm), X(c + n.xp), X(c + n.ym), X(c + n.yp), cp, cm);
                    xout[c] = cc;
                    const float rn = b[c] - t;
                    r[c] = rn;
                    v[0] += (double)rn * (double)rn;
                    cm = cc;
                    cc = cp;
                }
            }
        }
    }
    double out[1];
    if (grid_reduce<1>(v, partials, &st->counter, d.nblocks, out)) {
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            const double gnew = out[0];
            st->iter = it + 1;
            st->gamma_old = st->gamma;
            st->beta = gnew / st->gamma;
            st->gamma = gnew;
            st->converged = (gnew < st->thr) ? 1 : 0;
            st->done = st->converged || (it + 1) >= st->max_iter || !isfinite(gnew);
        }
    }
}

// FP64 true residual ||b - A x||^2 and ||b||^2 over the launch domain.
__global__ __launch_bounds__(256) void k_true_residual(Dom d, const uint16_t* __restrict__ rcode, const float* __restrict__ b,
                                                       const float* __restrict__ xv, CGState* st, double* partials) {
    int bx, by, bz;
    double v[2] = {0.0, 0.0};
    if (block_coords(d, bx, by, bz)) {
        const int x = d.x0 + bx * TX + threadIdx.x, y = d.y0 + by * TY + threadIdx.y;
        const int zb = d.z0 + bz * d.zc, ze = min(zb + d.zc, d.z0 + d.nz);
        if (x < d.x0 + d.nx && y < d.y0 + d.ny) {
            for (int z = zb; z < ze; z++) {
                const int c = (z * d.H + y) * d.W + x;
                const Nb n = nb_offsets(d, x, y, z);
                const double t = apply_row_d(rcode[c], xv[c], xv[c + n.xm], xv[c + n.xp], xv[c + n.ym], xv[c + n.yp],
                                             xv[c + n.zp], xv[c + n.zm]);
                const double rr = (double)b[c] - t;
                v[0] += rr * rr;
                v[1] += (double)b[c] * (double)b[c];
            }
        }
    }
    double out[2];
    if (grid_reduce<2>(v, partials, &st->counter, d.nblocks, out)) {
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            st->true_num = out[0];
            st->bnorm2 = out[1];
        }
    }
}

__global__ __launch_bounds__(256) void k_copy_domain(Dom d, const float* __restrict__ src, float* __restrict__ dst) {
    int bx, by, bz;
    if (!block_coords(d, bx, by, bz)) return;
    const int x = d.x0 + bx * TX + threadIdx.x, y = d.y0 + by * TY + threadIdx.y;
    if (x >= d.x0 + d.nx || y >= d.y0 + d.ny) return;
    const int zb = d.z0 + bz * d.zc, ze = min(zb + d.zc, d.z0 + d.nz);
    for (int z = zb; z < ze; z++) {
        const int c = (z * d.H + y) * d.W + x;
        dst[c] = src[c];
    }
}

// Plain operator application y = A x over the whole grid (efs_apply_operator).
__global__ __launch_bounds__(256) void k_spmv(Dom d, const uint16_t* __restrict__ rcode, const float* __restrict__ xv,
                                              float* __restrict__ yv) {
    int bx, by, bz;
    if (!block_coords(d, bx, by, bz)) return;
    const int x = d.x0 + bx * TX + threadIdx.x, y = d.y0 + by * TY + threadIdx.y;
    if (x >= d.x0 + d.nx || y >= d.y0 + d.ny) return;
    const int zb = d.z0 + bz * d.zc, ze = min(zb + d.zc, d.z0 + d.nz);
    for (int z = zb; z < ze; z++) {
        const int c = (z * d.H + y) * d.W + x;
        const Nb n = nb_offsets(d, x, y, z);
        yv[c] = apply_row(rcode[c], xv[c], xv[c + n.xm], xv[c + n.xp], xv[c + n.ym], xv[c + n.yp], xv[c + n.zp], xv[c + n.zm]);
    }
}

// ---------------------------------------------------------------------------------------
// Jacobi pass (JacobiSolver.cpp:165-196): neighbour order x-, x+, y-, y+, z-, z+;
// p' = (sum - b) / nCount, 0 when nCount == 0; identity rows copy p.
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_jacobi(Dom d, const uint16_t* __restrict__ rcode, const float* __restrict__ b,
                                                const float* __restrict__ pin, float* __restrict__ pout) {
    int bx, by, bz;
    if (!block_coords(d, bx, by, bz)) return;
    const int x = d.x0 + bx * TX + threadIdx.x, y = d.y0 + by * TY + threadIdx.y;
    if (x >= d.x0 + d.nx || y >= d.y0 + d.ny) return;
    const int zb = d.z0 + bz * d.zc, ze = min(zb + d.zc, d.z0 + d.nz);
    const int sz = d.W * d.H;
    int c = (zb * d.H + y) * d.W + x;
    float cm = (zb > 0) ? pin[c - sz] : 0.f;
    float cc = pin[c];
    for (int z = zb; z < ze; z++, c += sz) {
        const Nb n = nb_offsets(d, x, y, z);
        const float cp = pin[c + n.zp];
        const unsigned rc = rcode[c];
        float out = cc;
        if (rc & RC_JAC_ROW) {
            float sum = 0.0f;
            if (rc & RC_NB_XM) sum += pin[c + n.xm];
            if (rc & RC_NB_XP) sum += pin[c + n.xp];
            if (rc & RC_NB_YM) sum += pin[c + n.ym];
            if (rc & RC_NB_YP) sum += pin[c + n.yp];
            if (rc & RC_NB_ZM) sum += cm;
            if (rc & RC_NB_ZP) sum += cp;
            const unsigned cnt = (rc >> RC_CNT_SHIFT) & 7u;
            out = cnt > 0 ? (sum - b[c]) / (float)cnt : 0.0f;
        }
        pout[c] = out;
        cm = cc;
        cc = cp;
    }
}

Dom make_dom(const efs_sim* s, bool full) {
    Dom d;
    d.W = s->W; d.H = s->H; d.D = s->D;
    if (full) {
        d.x0 = d.y0 = d.z0 = 0;
        d.nx = s->W; d.ny = s->H; d.nz = s->D;
    } else {
        d.x0 = d.y0 = d.z0 = 1;
        d.nx = s->W - 2; d.ny = s->H - 2; d.nz = s->D - 2;
    }
    d.full = full ? 1 : 0;
    d.tiles_x = (d.nx + TX - 1) / TX;
    d.tiles_y = (d.ny + TY - 1) / TY;
    const int txy = d.tiles_x * d.tiles_y;
    // aim for >= 2048 blocks (8 per CU) with chunks of 2..32 planes
    int zc = (int)std::max<long long>(2, std::min<long long>(32, (long long)d.nz * txy / 2048));
    zc = std::min(zc, std::max(1, d.nz));
    d.zc = zc;
    d.chunks_z = (d.nz + zc - 1) / zc;
    d.nblocks = txy * d.chunks_z;
    return d;
}

int launch_blocks(const Dom& d) { return ((d.nblocks + 7) / 8) * 8; }

}  // namespace

int efs_solver_alloc(efs_sim* s) {
    const size_t N = (size_t)s->N;
    EFS_HIP(hipMalloc(&s->rowcode, N * sizeof(uint16_t)));
    EFS_HIP(hipMalloc(&s->cg_r, N * 4));
    EFS_HIP(hipMalloc(&s->cg_d0, N * 4));
    EFS_HIP(hipMalloc(&s->cg_d1, N * 4));
    EFS_HIP(hipMalloc(&s->cg_x1, N * 4));
    EFS_HIP(hipMemset(s->cg_r, 0, N * 4));
    EFS_HIP(hipMemset(s->cg_d0, 0, N * 4));
    EFS_HIP(hipMemset(s->cg_d1, 0, N * 4));
    EFS_HIP(hipMemset(s->cg_x1, 0, N * 4));
    const Dom d = make_dom(s, true);
    s->max_blocks = launch_blocks(d) + 64;
    EFS_HIP(hipMalloc(&s->partials, (size_t)s->max_blocks * 2 * sizeof(double)));
    EFS_HIP(hipMalloc(&s->st, sizeof(CGState)));
    EFS_HIP(hipMemset(s->st, 0, sizeof(CGState)));
    EFS_HIP(hipHostMalloc(&s->st_host, sizeof(CGState), hipHostMallocDefault));
    std::memset(s->st_host, 0, sizeof(CGState));
    return 0;
}

void efs_solver_free(efs_sim* s) {
    hipFree(s->rowcode);
    hipFree(s->cg_r);
    hipFree(s->cg_d0);
    hipFree(s->cg_d1);
    hipFree(s->cg_x1);
    hipFree(s->partials);
    hipFree(s->st);
    if (s->st_host) hipHostFree(s->st_host);
}

int efs_launch_rowcode(efs_sim* s, int* border_rows) {
    const DevGrid g = s->grid();
    // flags cleared on the device so the whole solve stays ordered on the stream
    CGState init{};
    init.max_iter = 0;
    std::memcpy(s->st_host, &init, sizeof(CGState));
    EFS_HIP(hipMemcpyAsync(s->st, s->st_host, sizeof(CGState), hipMemcpyHostToDevice, s->stream));
    dim3 blk(TX, TY, 1), grd((g.W + TX - 1) / TX, (g.H + TY - 1) / TY, g.D);
    hipLaunchKernelGGL(k_rowcode, grd, blk, 0, s->stream, g, s->rowcode, s->st);
    EFS_HIP(hipGetLastError());
    EFS_HIP(hipMemcpyAsync(s->st_host, s->st, sizeof(CGState), hipMemcpyDeviceToHost, s->stream));
    EFS_HIP(hipStreamSynchronize(s->stream));
    *border_rows = s->st_host->flags & 1;
    return 0;
}

void efs_ktimer_begin(efs_sim* s, int k, hipEvent_t* a, hipEvent_t* b) {
    *a = *b = nullptr;
    if (!s->ktiming) return;
    if (hipEventCreate(a) != hipSuccess || hipEventCreate(b) != hipSuccess) { *a = *b = nullptr; return; }
    hipEventRecord(*a, s->stream);
    (void)k;
}
void efs_ktimer_end(efs_sim* s, int k, hipEvent_t a, hipEvent_t b) {
    if (!a) return;
    hipEventRecord(b, s->stream);
    s->ktimers[k].start.push_back(a);
    s->ktimers[k].stop.push_back(b);
    s->ktimers[k].launches++;
}

static void trim_timer(TimerSlot& t, size_t keep) {
    if (t.start.size() <= keep) return;
    for (size_t i = keep; i < t.start.size(); i++) {
        hipEventSynchronize(t.stop[i]);
        hipEventDestroy(t.start[i]);
        hipEventDestroy(t.stop[i]);
        t.launches--;
    }
    t.start.resize(keep);
    t.stop.resize(keep);
}

int efs_launch_spmv(efs_sim* s, const float* x, float* y) {
    int border = 0;
    int rc = efs_launch_rowcode(s, &border);
    if (rc) return rc;
    const Dom d = make_dom(s, true);
    hipLaunchKernelGGL(k_spmv, dim3(launch_blocks(d)), dim3(TX, TY), 0, s->stream, d, s->rowcode, x, y);
    EFS_HIP(hipGetLastError());
    return 0;
}

// Prepares the state, the launch domain and the initial guess x0 = b (pressure = divergence,
// CGSolver.cpp:429).  Returns the domain through *dom.
static int cg_prepare(efs_sim* s, Dom* dom) {
    int border = 0;
    int rc = efs_launch_rowcode(s, &border);
    if (rc) return rc;
    *dom = make_dom(s, border != 0);
    EFS_HIP(hipMemcpyAsync(s->p, s->div, (size_t)s->N * 4, hipMemcpyDeviceToDevice, s->stream));
    CGState init{};
    init.thr = (double)s->P.solver_threshold;
    init.flags = (s->P.criterion == EFS_CRIT_RELATIVE) ? 2 : 0;
    init.max_iter = s->P.max_steps;
    std::memcpy(s->st_host, &init, sizeof(CGState));
    EFS_HIP(hipMemcpyAsync(s->st, s->st_host, sizeof(CGState), hipMemcpyHostToDevice, s->stream));
    return 0;
}

int efs_true_residual(efs_sim* s, double* num, double* den) {
    const Dom d = make_dom(s, true);
    hipLaunchKernelGGL(k_true_residual, dim3(launch_blocks(d)), dim3(TX, TY), 0, s->stream, d, s->rowcode, s->div, s->p, s->st,
                       s->partials);
    EFS_HIP(hipGetLastError());
    EFS_HIP(hipMemcpyAsync(s->st_host, s->st, sizeof(CGState), hipMemcpyDeviceToHost, s->stream));
    EFS_HIP(hipStreamSynchronize(s->stream));
    *num = s->st_host->true_num;
    *den = s->st_host->bnorm2;
    return 0;
}

int efs_cg_solve(efs_sim* s, efs_stats* out) {
    Dom d;
    int rc = cg_prepare(s, &d);
    if (rc) return rc;
    const int nb = launch_blocks(d);
    if (nb > s->max_blocks) { efs_set_error("launch exceeds reduction buffer"); return -1; }
    const dim3 blk(TX, TY);
    hipLaunchKernelGGL(k_cg_init, dim3(nb), blk, 0, s->stream, d, s->rowcode, s->div, s->cg_r, s->cg_d0, s->st, s->partials);
    EFS_HIP(hipGetLastError());
    const int max_it = s->P.max_steps;
    const int batch = 16;
    int launched = 0;
    const size_t kt0 = s->ktimers[KT_CG_DIR_SPMV].start.size(), kt1 = s->ktimers[KT_CG_UPDATE].start.size();
    while (true) {
        const int n = std::min(batch, max_it - launched);
        for (int k = 0; k < n; k++) {
            hipEvent_t a, b;
            efs_ktimer_begin(s, KT_CG_DIR_SPMV, &a, &b);
            hipLaunchKernelGGL(k_cg_dir_spmv, dim3(nb), blk, 0, s->stream, d, s->rowcode, s->cg_r, s->cg_d0, s->cg_d1, s->st,
                               s->partials);
            efs_ktimer_end(s, KT_CG_DIR_SPMV, a, b);
            efs_ktimer_begin(s, KT_CG_UPDATE, &a, &b);
            hipLaunchKernelGGL(k_cg_update, dim3(nb), blk, 0, s->stream, d, s->rowcode, s->div, s->cg_r, s->cg_d0, s->cg_d1,
                               s->p, s->cg_x1, s->st, s->partials);
            efs_ktimer_end(s, KT_CG_UPDATE, a, b);
        }
        EFS_HIP(hipGetLastError());
        launched += n;
        EFS_HIP(hipMemcpyAsync(s->st_host, s->st, sizeof(CGState), hipMemcpyDeviceToHost, s->stream));
        EFS_HIP(hipStreamSynchronize(s->stream));
        if (s->st_host->done || launched >= max_it) break;
    }
    const CGState fin = *s->st_host;
    // launches after convergence exit at once; drop their events so kernel timings count real
    // iterations only
    trim_timer(s->ktimers[KT_CG_DIR_SPMV], kt0 + fin.iter);
    trim_timer(s->ktimers[KT_CG_UPDATE], kt1 + fin.iter);
    // the final iterate lives in x[iter & 1]; x[0] is the pressure field
    if (fin.iter & 1)
        hipLaunchKernelGGL(k_copy_domain, dim3(nb), blk, 0, s->stream, d, s->cg_x1, s->p);
    EFS_HIP(hipGetLastError());
    if (out) {
        out->iterations = fin.iter;
        out->rr = fin.gamma;
        out->bnorm = std::sqrt(fin.bnorm2);
        out->rel_residual = fin.bnorm2 > 0 ? std::sqrt(fin.gamma / fin.bnorm2) : 0.0;
        out->true_rel_residual = -1.0;
    }
    if (s->P.true_residual && out) {
        double num = 0, den = 0;
        rc = efs_true_residual(s, &num, &den);
        if (rc) return rc;
        out->true_rel_residual = den > 0 ? std::sqrt(num / den) : std::sqrt(num);
    }
    if (!std::isfinite(fin.gamma)) return 2;
    return fin.converged ? 0 : 1;
}

int efs_jacobi_solve(efs_sim* s, efs_stats* out) {
    int border = 0;
    int rc = efs_launch_rowcode(s, &border);
    if (rc) return rc;
    const Dom d = make_dom(s, border != 0);
    const int nb = launch_blocks(d);
    const size_t bytes = (size_t)s->N * 4;
    // p was cleared by project(); both ping-pong buffers start from it
    EFS_HIP(hipMemcpyAsync(s->cg_x1, s->p, bytes, hipMemcpyDeviceToDevice, s->stream));
    const int iters = s->P.max_steps / 2;
    for (int i = 0; i < iters; i++) {
        const float* pin = (i & 1) ? s->cg_x1 : s->p;
        float* pout = (i & 1) ? s->p : s->cg_x1;
        hipEvent_t a, b;
        efs_ktimer_begin(s, KT_JACOBI, &a, &b);
        hipLaunchKernelGGL(k_jacobi, dim3(nb), dim3(TX, TY), 0, s->stream, d, s->rowcode, s->div, pin, pout);
        efs_ktimer_end(s, KT_JACOBI, a, b);
    }
    EFS_HIP(hipGetLastError());
    if (iters & 1) EFS_HIP(hipMemcpyAsync(s->p, s->cg_x1, bytes, hipMemcpyDeviceToDevice, s->stream));
    if (out) {
        out->iterations = iters;
        out->rr = 0;
        out->rel_residual = -1;
        out->true_rel_residual = -1;
    }
    if (s->P.true_residual && out) {
        double num = 0, den = 0;
        rc = efs_true_residual(s, &num, &den);
        if (rc) return rc;
        out->true_rel_residual = den > 0 ? std::sqrt(num / den) : std::sqrt(num);
        out->bnorm = std::sqrt(den);
    }
    return 0;  // JacobiSolverCPU3D::solve returns 0 (JacobiSolver.cpp:211)
}
