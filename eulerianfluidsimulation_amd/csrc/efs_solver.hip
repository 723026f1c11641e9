// Pressure Poisson solve on gfx950: the matrix-free 7-point operator of CGSolverCPU3D
// (CGSolver.cpp:467-497), a mixed-precision CG (CGSolver.cpp:416-465 restated), Jacobi
// (JacobiSolver.cpp:154-212) and an FP64 true residual.
//
// Precision.  Search direction d, residual r and the operator run in FP32; dot products,
// alpha and beta run in FP64.  The iterate is kept as an FP64 accumulator x64 plus an FP32
// increment xd: each iteration adds alpha d to xd, and at every residual replacement
// (i % 64 == 0, where the reference recomputes r = b - A x) xd is folded into x64 and r is
// recomputed from x64 in FP64.  An FP32 iterate alone cannot reach ||b - A x|| / ||b|| <= 1e-5
// at 256^3: rounding the exact solution to FP32 already leaves a larger residual.
//
// Kernel shape.  A block is a 64 x 8 tile of an xy plane (8 wavefronts, one per row) marching
// along z over a chunk of planes.  The plane's lateral neighbours are staged in LDS with a
// one-cell halo (double-buffered, one barrier per plane); the z neighbours of the thread's own
// column stay in registers; the loads of the next K1_PF / K2_PF planes are in flight while plane z is
// computed.
// Blocks are renumbered so each XCD owns a contiguous run of tiles (workgroups are dealt to
// the 8 XCDs round-robin), keeping the halo rows shared by neighbouring tiles in one L2.
//
// Iteration = 2 kernels.  Neither uses a fence or an atomic: a release fence at device scope
// writes back the XCD's L2 on every block and cost more than the stencil itself.  Each block
// stores its partial sum; every block of the next kernel reduces all partials in the same fixed
// order, so all blocks hold bit-identical FP64 scalars:
//   K1(i)  gamma_i = sum(r.r partials), convergence test, beta = gamma_i / gamma_{i-1},
//          d_i = r_i + beta d_{i-1} (formed while staging, stored once), t = A d_i, partial d.t
//   K2(i)  alpha = gamma_i / sum(d.t partials), xd += alpha d_i, r_{i+1} = r_i - alpha t,
//          partial r.r
// On replacement iterations K2 is replaced by a fold (x64 += xd + alpha d) and an FP64 residual
// kernel.  A converged solve turns the remaining launches of a batch into no-ops, so the host
// synchronises once per batch of 64 iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>
#include <cmath>
#include <cstring>

#include <hip/hip_ext.h>

#include "efs_internal.h"
#include "efs_stencil.h"

namespace {

// ---------------------------------------------------------------------------------------
// row codes: neighbour counted = in range, not on the outer layer, boundary <= 0.5;
// contributing = counted and rho > 0.5; CG row = !(rho <= 0.5 || boundary == 1);
// Jacobi row = !(rho <= 0.5 || boundary >= 0.5).
// ---------------------------------------------------------------------------------------
// A thread marches up RCZ planes of one (x, y) column: the rho and boundary byte of the cell and
// of its -z neighbour come from the planes before, and the loads of the next plane are issued before
// the current plane is computed (clamped into the array; unused values are ignored).  The blocks
// take (tile, z chunk) in XCD order.
constexpr int RCZ = 16;
__global__ __launch_bounds__(256) void k_rowcode(DevGrid g, uint16_t* rc, CGState* st) {
    XcdBlock blk;
    if (!xcd_block((g.W + 63) / 64, (g.H + 3) / 4, blk)) return;
    const int x = blk.bx * 64 + threadIdx.x, y = blk.by * 4 + threadIdx.y;
    if (x >= g.W || y >= g.H) return;
    const int zb = blk.bz * RCZ, ze = min(zb + RCZ, g.D);
    const int sy = g.W, sz = g.W * g.H, last = g.N - 1;
    const bool ox = x == 0 || x == g.W - 1, oy = y == 0 || y == g.H - 1;
    const int dxm = x > 0 ? -1 : 0, dxp = x < g.W - 1 ? 1 : 0, dym = y > 0 ? -sy : 0, dyp = y < g.H - 1 ? sy : 0;
    const unsigned bits[6] = {RC_NB_XM, RC_NB_XP, RC_NB_YM, RC_NB_YP, RC_NB_ZP, RC_NB_ZM};
    struct Plane {
        float r[5];     // the cell's +z neighbour, then x-, x+, y-, y+ of the cell
        unsigned b[5];
    };
    auto load = [&](int z, Plane& q) {  // the loads of plane z (clamped into the array)
        const int c = min(cidx(g, x, y, min(z, g.D - 1)), last);
        const int o[5] = {min(c + sz, last), c + dxm, c + dxp, c + dym, c + dyp};
#pragma unroll
        for (int k = 0; k < 5; k++) {
            q.r[k] = g.rho[o[k]];
            q.b[k] = g.bf[o[k]];
        }
    };
    const int c0 = cidx(g, x, y, zb);
    float r = g.rho[c0], rzm = zb > 0 ? g.rho[c0 - sz] : 0.f;
    unsigned b = g.bf[c0], bzm = zb > 0 ? (unsigned)g.bf[c0 - sz] : 0u;
    Plane q, qn;
    load(zb, q);
    for (int z = zb; z < ze; z++) {
        load(z + 1, qn);
        // neighbour order x-, x+, y-, y+, z+, z- and the tests of CGSolver.cpp:477-490
        const bool oz = z == 0 || z == g.D - 1;
        const bool ok[6] = {x > 0, x < g.W - 1, y > 0, y < g.H - 1, z < g.D - 1, z > 0};
        // a neighbour on the outer layer is never counted: its displaced coordinate is on the
        // layer, or the cell itself is on it along another axis
        const bool outer[6] = {x - 1 == 0 || oy || oz, x + 1 == g.W - 1 || oy || oz, y - 1 == 0 || ox || oz,
                               y + 1 == g.H - 1 || ox || oz, z + 1 == g.D - 1 || ox || oy, z - 1 == 0 || ox || oy};
        const float rq[6] = {q.r[1], q.r[2], q.r[3], q.r[4], q.r[0], rzm};
        const unsigned bq[6] = {q.b[1], q.b[2], q.b[3], q.b[4], q.b[0], bzm};
        unsigned code = 0, cnt = 0;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const bool counted = ok[k] && !outer[k] && !(bq[k] & BF_GT05);
            cnt += counted ? 1u : 0u;
            if (counted && rq[k] > 0.5f) code |= bits[k];
        }
        code |= cnt << RC_CNT_SHIFT;
        if (!(ox || oy || oz) && !(b & BF_GT05) && r > 0.5f) code |= RC_CONTRIB;
        if (!(r <= 0.5f || (b & BF_EQ1))) code |= RC_CG_ROW;
        if (!(r <= 0.5f || (b & BF_GE05))) code |= RC_JAC_ROW;
        rc[cidx(g, x, y, z)] = (uint16_t)code;
        if ((ox || oy || oz) && (code & (RC_CG_ROW | RC_JAC_ROW))) atomicOr(&st->flags, 1);
        rzm = r;
        bzm = b;
        r = q.r[0];
        b = q.b[0];
        q = qn;
    }
}

// ---------------------------------------------------------------------------------------
// CG.  No kernel on the iteration path uses a fence or an atomic: each block writes its
// partial sum with a plain store, and every block of the next kernel reduces all partials in
// the same fixed order (bit-identical scalars in every block).  Block 0 records the scalars
// in CGState for the host; the iteration index is a launch argument.
// ---------------------------------------------------------------------------------------
// whole array: x64 = b, xd = 0 (identity rows keep x0 = b, CGSolver.cpp:429); with `prev`, the
// fluid rows keep the previous iterate (the previous step's pressure)
__global__ void k_cg_setup(int n, const float* __restrict__ b, const uint16_t* __restrict__ rcode, int prev,
                           double* __restrict__ x64, float* __restrict__ xd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!(prev && (rcode[i] & RC_CG_ROW))) x64[i] = (double)b[i];
    xd[i] = 0.f;
}

// Convergence test at the start of iteration `it` on r_it . r_it (from the partials).  The
// scalar that drives alpha and beta, g_it, is r.r for plain CG and r.z (z = M r, partials p_rz)
// for preconditioned CG; it is recorded by parity in gam[].  Returns true when the solve stops;
// block 0 records the outcome.
__device__ inline bool cg_check(CGState* st, int it, const double* __restrict__ p_rr, const double* __restrict__ p_bb,
                                const double* __restrict__ p_rz, double* red, double& g_it) {
    const double rr = sum_partials(p_rr, gridDim.x, red);
    g_it = p_rz ? sum_partials(p_rz, gridDim.x, red) : rr;
    double thr = st->thr;
    double bb = 0.0;
    if (it == 0) {
        thr = st->thr_in;
        bb = sum_partials(p_bb, gridDim.x, red);
        if (st->flags & 2) thr = thr * thr * bb;
    }
    const bool conv = rr < thr || rr == 0.0;
    const bool stop = conv || !isfinite(rr) || !isfinite(g_it) || it >= st->max_iter;
    if (blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
        st->gam[it & 1] = g_it;
        st->gamma = rr;
        if (it == 0) {
            st->thr = thr;
            st->bnorm2 = bb;
        }
        if (stop) {
            st->iter = it;
            st->converged = conv ? 1 : 0;
            st->done = 1;
        }
    }
    return stop;
}

// r0 = b - A x0 (FP64, rounded; x0 = x64, which is b for the reference's initial guess), d0 = r0;
// partials of r.r and b.b
// With x0 >= 0 it also sets the initial iterate on the way (the work of k_cg_setup): x0 = 0 takes
// x0 = b; x0 = 1 keeps the previous iterate on CG rows and takes b elsewhere.  Each cell writes
// only its own x64, and only where the value becomes b, while neighbours read x64 only on CG rows
// (x0 = 1) or not at all (x0 = 0): no thread reads a value another thread writes.
template <int X0>
__global__ __launch_bounds__(CG_NT) void k_cg_init(Dom d, const uint16_t* __restrict__ rcode, const float* __restrict__ b,
                                                   double* __restrict__ x64, float* __restrict__ r, float* __restrict__ d0,
                                                   double* p_rr, double* p_bb, float* __restrict__ xd) {
    constexpr int x0 = X0;
    __shared__ double red[CG_NT / 64];
    int bx, by, bz;
    double vr = 0.0, vb = 0.0;
    if (block_coords(d, bx, by, bz)) {
        const int x = d.x0 + bx * TX + threadIdx.x, y = d.y0 + by * d.ty + threadIdx.y;
        const int zb = d.z0 + bz * d.zc, ze = min(zb + d.zc, d.z0 + d.nz);
        auto X = [&](int q) -> double {
            if (x0 < 0) return x64[q];
            if (x0 == 0) return (double)b[q];
            const double xv = x64[q], bv = (double)b[q];  // loaded together, then selected
            return (rcode[q] & RC_CG_ROW) ? xv : bv;
        };
        // a neighbour that enters a row (its RC_NB bit: counted, so boundary <= 0.5, and
        // rho > 0.5) is a CG row itself, so with x0 = 1 its x0 is the previous iterate
        auto Xn = [&](int q) -> double { return x0 == 0 ? (double)b[q] : x64[q]; };
        if (x < d.x0 + d.nx && y < d.y0 + d.ny) {
            for (int z = zb; z < ze; z++) {
                const int c = (z * d.H + y) * d.W + x;
                const Nb n = nb_offsets(d, x, y, z);
                const double bc = b[c];
                const double xc = X(c);
                if (x0 >= 0) {
                    if (!(x0 == 1 && (rcode[c] & RC_CG_ROW))) x64[c] = xc;
                    xd[c] = 0.f;
                }
                const double t = apply_row_d(rcode[c], xc, Xn(c + n.xm), Xn(c + n.xp), Xn(c + n.ym), Xn(c + n.yp), Xn(c + n.zp),
                                             Xn(c + n.zm));
                const float rr = (float)(bc - t);
                r[c] = rr;
                d0[c] = rr;
                vr += (double)rr * (double)rr;
                vb += bc * bc;
            }
        }
    }
    write_partial(vr, p_rr, red);
    __syncthreads();
    write_partial(vb, p_bb, red);
}

// Loads of one plane zz, issued K1_PF / K2_PF planes ahead of the plane being computed so that
// each wavefront keeps that many planes of loads in flight (one plane per wavefront leaves the
// kernel latency-bound at about a third of HBM bandwidth).  The ring slots are indexed with
// compile-time constants (the plane loop is unrolled by the ring depth), so no slot is ever copied
// while its loads are outstanding.
#ifndef EFS_K1_PF
#define EFS_K1_PF 2
#endif
#ifndef EFS_K2_PF
#define EFS_K2_PF 2
#endif
constexpr int K1_PF = EFS_K1_PF, K2_PF = EFS_K2_PF;

struct K1Plane {
    uint16_t rc;                        // row code of the own cell (widened where it is used)
    float od, orr, hxd, hxr, hyd, hyr;  // d_{i-1} and r of the own cell and of the halo cells
};

// Every load is unconditional: the address is clamped into the array and a value that is not
// needed (a plane past the chunk, a halo cell outside the grid, a lane outside the domain) is
// loaded and never used -- a row code bit never points at such a cell.  A load under a branch,
// or a select right after a load, would make the compiler wait for all outstanding loads.
__device__ inline int clampi(int c, int hi) { return min(max(c, 0), hi); }

__device__ inline void k1_load(K1Plane& q, int c, int last, const Stager<CG_TY>& S, const float* __restrict__ dsrc,
                               const float* __restrict__ r, const uint16_t* __restrict__ rcode) {
    const int co = clampi(c, last), cx = clampi(c + S.hx_off, last), cy = clampi(c + S.hy_off, last);
    q.rc = rcode[co];
    q.od = dsrc[co];
    q.orr = r[co];
    q.hxd = dsrc[cx];
    q.hxr = r[cx];
    q.hyd = dsrc[cy];
    q.hyr = r[cy];
}

// K1 (iteration it): convergence test, d_it = r_it + beta d_{it-1}, t = A d_it, partial d.t.
// Iteration 0 has beta = 0 and stages d_0 = r_0 (k_cg_init stored it in d0 as well).  With a
// preconditioner (p_rz given) `r` is z = M r and d_it = z_it + beta d_{it-1}; iteration 0 then
// stores d_0 = z_0.  FLEX (flexible CG with a nonlinear preconditioner, Notay's FCG(1)): `r` is
// z, beta = -(z_it . q_{it-1}) / (d_{it-1} . q_{it-1}) from the partials p_zq and st->dq, the
// convergence test is a separate kernel, and t = A d_it is also stored in q for the update and
// the next iteration's z . q (qout; p_zq and qout are unused otherwise).
template <bool FLEX>
__global__ __launch_bounds__(CG_NT) void k_cg_dir_spmv(Dom d, int it, const uint16_t* __restrict__ rcode, const float* __restrict__ r,
                                                       float* d0, float* d1, CGState* st, const double* p_rr,
                                                       const double* p_bb, double* p_dad, const double* p_rz,
                                                       const double* p_zq, float* __restrict__ qout) {
    if (st->done) return;
    __shared__ double red[CG_NT / 64];
    double beta;
    if constexpr (FLEX) {
        const double zq = sum_partials(p_zq, gridDim.x, red);
        beta = (it == 0 || st->dq == 0.0) ? 0.0 : -zq / st->dq;
    } else {
        double g_it;
        if (cg_check(st, it, p_rr, p_bb, p_rz, red, g_it)) return;
        beta = (it == 0) ? 0.0 : g_it / st->gam[(it + 1) & 1];
    }
    const bool first = !FLEX && (it == 0) && !p_rz;
    float* dnew = (it & 1) ? d1 : d0;
    // iteration 0: d_0 = r_0 (already in d0) or z_0 (staged from z itself, beta = 0)
    const float* dsrc = (it == 0) ? ((FLEX || p_rz) ? r : dnew) : ((it & 1) ? d0 : d1);
    __shared__ float sh[2][CG_TY + 2][TX + 2];
    const int tx = threadIdx.x, ty = threadIdx.y;
    auto Dv = [&](float dold, float rr) -> float { return (float)((double)dold * beta + (double)rr); };
    int bx, by, bz;
    double v = 0.0;
    if (block_coords(d, bx, by, bz)) {
        const int x = d.x0 + bx * TX + tx, y = d.y0 + by * CG_TY + ty;
        const bool in_dom = x < d.x0 + d.nx && y < d.y0 + d.ny;
        Stager<CG_TY> S;
        S.init(d, x, y);
        const int zb = d.z0 + bz * d.zc, ze = min(zb + d.zc, d.z0 + d.nz);
        const int sz = d.W * d.H, last = sz * d.D - 1;
        int c = (zb * d.H + y) * d.W + x;
        const int cm = clampi(c - sz, last);
        float dm = Dv(dsrc[cm], r[cm]);
        K1Plane q0;
        k1_load(q0, c, last, S, dsrc, r, rcode);
        K1Plane ring[K1_PF];
#pragma unroll
        for (int k = 0; k < K1_PF; k++) k1_load(ring[k], c + (k + 1) * sz, last, S, dsrc, r, rcode);
        float dc = Dv(q0.od, q0.orr);
        unsigned rc_cur = q0.rc;
        sh[0][ty + 1][tx + 1] = dc;
        if (S.hx_slot >= 0) sh[0][ty + 1][S.hx_slot] = Dv(q0.hxd, q0.hxr);
        if (S.hy_slot >= 0) sh[0][S.hy_slot][tx + 1] = Dv(q0.hyd, q0.hyr);
        lds_barrier();
        int cur = 0;
        int z = zb;
        while (z < ze) {
#pragma unroll
            for (int k = 0; k < K1_PF; k++) {
                if (z < ze) {
                    K1Plane& q = ring[k];  // plane z + 1
                    const float np = Dv(q.od, q.orr);
                    const float nhx = Dv(q.hxd, q.hxr), nhy = Dv(q.hyd, q.hyr);
                    const unsigned rc_next = q.rc;
                    k1_load(q, c + (1 + K1_PF) * sz, last, S, dsrc, r, rcode);
                    if (in_dom) {
                        float t = dc;
                        if (rc_cur & RC_CG_ROW)
                            t = apply_row(rc_cur, dc, sh[cur][ty + 1][tx], sh[cur][ty + 1][tx + 2], sh[cur][ty][tx + 1],
                                          sh[cur][ty + 2][tx + 1], np, dm);
                        if (!first) dnew[c] = dc;
                        if constexpr (FLEX) qout[c] = t;
                        v += (double)dc * (double)t;
                    }
                    sh[cur ^ 1][ty + 1][tx + 1] = np;
                    if (S.hx_slot >= 0) sh[cur ^ 1][ty + 1][S.hx_slot] = nhx;
                    if (S.hy_slot >= 0) sh[cur ^ 1][S.hy_slot][tx + 1] = nhy;
                    lds_barrier();
                    cur ^= 1;
                    dm = dc;
                    dc = np;
                    rc_cur = rc_next;
                    z++;
                    c += sz;
                }
            }
        }
    }
    write_partial(v, p_dad, red);
    if (blockIdx.x == 0 && tx == 0 && ty == 0) st->beta = beta;
}

template <class XT>
struct K2Plane {
    uint16_t rc;
    float o, hx, hy;  // d_i of the own cell and of the halo cells
    XT xd;            // iterate increment (FP32) or FP64 iterate of the own cell
    float r;          // residual of the own cell
};

template <class XT>
__device__ inline void k2_load(K2Plane<XT>& q, int c, int last, const Stager<CG_TY>& S, const float* __restrict__ dd,
                               const XT* xd, const float* r, const uint16_t* __restrict__ rcode) {
    const int co = clampi(c, last);
    q.rc = rcode[co];
    q.xd = xd[co];
    q.r = r[co];
    q.o = dd[co];
    q.hx = dd[clampi(c + S.hx_off, last)];
    q.hy = dd[clampi(c + S.hy_off, last)];
}

// K2 (iteration it): alpha = gamma / d.t, xd += alpha d, r -= alpha A d, partial r.r.
// XT = double: the FP64 iterate itself is updated (preconditioned CG, whose few large steps
// would otherwise leave most of the solution in the FP32 increment).
template <class XT>
__global__ __launch_bounds__(CG_NT) void k_cg_update(Dom d, int it, const uint16_t* __restrict__ rcode, float* r, const float* d0,
                                                     const float* d1, XT* xd, CGState* st, const double* p_dad,
                                                     double* p_rr) {
    if (st->done) return;
    __shared__ double red[CG_NT / 64];
    const double alpha = st->gam[it & 1] / sum_partials(p_dad, gridDim.x, red);
    const float* dd = (it & 1) ? d1 : d0;
    __shared__ float sh[2][CG_TY + 2][TX + 2];
    const int tx = threadIdx.x, ty = threadIdx.y;
    int bx, by, bz;
    double v = 0.0;
    if (block_coords(d, bx, by, bz)) {
        const int x = d.x0 + bx * TX + tx, y = d.y0 + by * CG_TY + ty;
        const bool in_dom = x < d.x0 + d.nx && y < d.y0 + d.ny;
        Stager<CG_TY> S;
        S.init(d, x, y);
        const int zb = d.z0 + bz * d.zc, ze = min(zb + d.zc, d.z0 + d.nz);
        const int sz = d.W * d.H, last = sz * d.D - 1;
        int c = (zb * d.H + y) * d.W + x;
        float dm = dd[clampi(c - sz, last)];
        K2Plane<XT> q0;
        k2_load(q0, c, last, S, dd, xd, r, rcode);
        K2Plane<XT> ring[K2_PF];
#pragma unroll
        for (int k = 0; k < K2_PF; k++) k2_load(ring[k], c + (k + 1) * sz, last, S, dd, xd, r, rcode);
        float dc = q0.o;
        unsigned rc_cur = q0.rc;
        XT xd_cur = q0.xd;
        float r_cur = q0.r;
        sh[0][ty + 1][tx + 1] = dc;
        if (S.hx_slot >= 0) sh[0][ty + 1][S.hx_slot] = q0.hx;
        if (S.hy_slot >= 0) sh[0][S.hy_slot][tx + 1] = q0.hy;
        lds_barrier();
        int cur = 0;
        int z = zb;
        while (z < ze) {
#pragma unroll
            for (int k = 0; k < K2_PF; k++) {
                if (z < ze) {
                    K2Plane<XT>& q = ring[k];  // plane z + 1
                    const float np = q.o, nhx = q.hx, nhy = q.hy;
                    const unsigned rc_next = q.rc;
                    const XT xd_next = q.xd;
                    const float r_next = q.r;
                    k2_load(q, c + (1 + K2_PF) * sz, last, S, dd, xd, r, rcode);
                    if (in_dom) {
                        float t = dc;
                        if (rc_cur & RC_CG_ROW)
                            t = apply_row(rc_cur, dc, sh[cur][ty + 1][tx], sh[cur][ty + 1][tx + 2], sh[cur][ty][tx + 1],
                                          sh[cur][ty + 2][tx + 1], np, dm);
                        xd[c] = (XT)((double)xd_cur + alpha * (double)dc);
                        const float rn = (float)((double)r_cur - alpha * (double)t);
                        r[c] = rn;
                        v += (double)rn * (double)rn;
                    }
                    sh[cur ^ 1][ty + 1][tx + 1] = np;
                    if (S.hx_slot >= 0) sh[cur ^ 1][ty + 1][S.hx_slot] = nhx;
                    if (S.hy_slot >= 0) sh[cur ^ 1][S.hy_slot][tx + 1] = nhy;
                    lds_barrier();
                    cur ^= 1;
                    dm = dc;
                    dc = np;
                    rc_cur = rc_next;
                    xd_cur = xd_next;
                    r_cur = r_next;
                    z++;
                    c += sz;
                }
            }
        }
    }
    write_partial(v, p_rr, red);
    if (blockIdx.x == 0 && tx == 0 && ty == 0) st->alpha = alpha;
}

// Replacement iteration, part 1: x64 += xd + alpha d (own cell), xd = 0.
__global__ __launch_bounds__(CG_NT) void k_cg_fold(Dom d, int it, const float* d0, const float* d1, float* __restrict__ xd,
                                                   double* __restrict__ x64, const CGState* st, const double* p_dad) {
    if (st->done) return;
    __shared__ double red[CG_NT / 64];
    const double alpha = st->gam[it & 1] / sum_partials(p_dad, gridDim.x, red);
    const float* dd = (it & 1) ? d1 : d0;
    int bx, by, bz;
    if (!block_coords(d, bx, by, bz)) return;
    const int x = d.x0 + bx * TX + threadIdx.x, y = d.y0 + by * d.ty + threadIdx.y;
    if (x >= d.x0 + d.nx || y >= d.y0 + d.ny) return;
    const int zb = d.z0 + bz * d.zc, ze = min(zb + d.zc, d.z0 + d.nz);
    for (int z = zb; z < ze; z++) {
        const int c = (z * d.H + y) * d.W + x;
        const float xn = (float)((double)xd[c] + alpha * (double)dd[c]);
        x64[c] = x64[c] + (double)xn;
        xd[c] = 0.f;
    }
}

// Replacement iteration, part 2: r = b - A x64 in FP64, rounded to FP32; partial r.r.
__global__ __launch_bounds__(CG_NT) void k_cg_residual(Dom d, const uint16_t* __restrict__ rcode, const float* __restrict__ b,
                                                       const double* __restrict__ x64, float* __restrict__ r, const CGState* st,
                                                       double* p_rr) {
    if (st->done) return;
    __shared__ double red[CG_NT / 64];
    int bx, by, bz;
    double v = 0.0;
    if (block_coords(d, bx, by, bz)) {
        const int x = d.x0 + bx * TX + threadIdx.x, y = d.y0 + by * d.ty + threadIdx.y;
        const int zb = d.z0 + bz * d.zc, ze = min(zb + d.zc, d.z0 + d.nz);
        if (x < d.x0 + d.nx && y < d.y0 + d.ny) {
            for (int z = zb; z < ze; z++) {
                const int c = (z * d.H + y) * d.W + x;
                const Nb n = nb_offsets(d, x, y, z);
                const double t = apply_row_d(rcode[c], x64[c], x64[c + n.xm], x64[c + n.xp], x64[c + n.ym], x64[c + n.yp],
                                             x64[c + n.zp], x64[c + n.zm]);
                const float rn = (float)((double)b[c] - t);
                r[c] = rn;
                v += (double)rn * (double)rn;
            }
        }
    }
    write_partial(v, p_rr, red);
}

// Convergence test for iteration `it` after a batch (one block, the CG grid's partial count).
__global__ __launch_bounds__(CG_NT) void k_cg_check(int nparts, int it, CGState* st, const double* p_rr, const double* p_bb,
                                                    const double* p_rz) {
    if (st->done) return;
    __shared__ double red[CG_NT / 64];
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
    auto sum = [&](const double* p) {
        double s = 0.0;
        for (int i = lane; i < nparts; i += CG_NT) s += p[i];
        s = wave_sum(s);
        if ((lane & 63) == 0) red[lane >> 6] = s;
        __syncthreads();
        double tot = 0.0;
        for (int w = 0; w < CG_NT / 64; w++) tot += red[w];
        __syncthreads();
        return tot;
    };
    const double g = sum(p_rr);
    const double gz = p_rz ? sum(p_rz) : g;
    double thr = st->thr, bb = st->bnorm2;
    if (it == 0) {
        bb = sum(p_bb);
        thr = st->thr_in;
        if (st->flags & 2) thr = thr * thr * bb;
    }
    if (lane == 0) {
        const bool conv = g < thr || g == 0.0;
        st->gamma = g;
        st->gam[it & 1] = gz;
        if (it == 0) {
            st->thr = thr;
            st->bnorm2 = bb;
        }
        if (conv || !isfinite(g) || !isfinite(gz) || it >= st->max_iter) {
            st->iter = it;
            st->converged = conv ? 1 : 0;
            st->done = 1;
        }
    }
}

// partial r.z for preconditioned CG (same grid as K1 / K2)
__global__ __launch_bounds__(CG_NT) void k_cg_rz(Dom d, const float* __restrict__ r, const float* __restrict__ z, const CGState* st,
                                                 double* p_rz) {
    if (st->done) return;
    __shared__ double red[CG_NT / 64];
    int bx, by, bz;
    double v = 0.0;
    if (block_coords(d, bx, by, bz)) {
        const int x = d.x0 + bx * TX + threadIdx.x, y = d.y0 + by * d.ty + threadIdx.y;
        const int zb = d.z0 + bz * d.zc, ze = min(zb + d.zc, d.z0 + d.nz);
        if (x < d.x0 + d.nx && y < d.y0 + d.ny) {
            for (int zz = zb; zz < ze; zz++) {
                const int c = (zz * d.H + y) * d.W + x;
                v += (double)r[c] * (double)z[c];
            }
        }
    }
    write_partial(v, p_rz, red);
}

// whole array: x64 += xd; pressure = (float)x64 (without xd: the increment is known to be 0)
__global__ void k_cg_finalize(int n, double* __restrict__ x64, const float* __restrict__ xd, float* __restrict__ p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!xd) {
        p[i] = (float)x64[i];
        return;
    }
    const double x = x64[i] + (double)xd[i];
    x64[i] = x;
    p[i] = (float)x;
}

// whole array: x64 += xd; xd = 0
__global__ void k_cg_fold_all(int n, double* __restrict__ x64, float* __restrict__ xd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x64[i] = x64[i] + (double)xd[i];
    xd[i] = 0.f;
}

// FP64 true residual ||b - A x||^2 and ||b||^2 (x in FP64 if given, else the FP32 field): block
// partials in partials[block] and partials[gridDim + block], plain stores (no fence), summed in a
// fixed order by k_true_residual_sum.  With pout (and x in FP64) each visited cell also gets its
// FP32 pressure, the rounding of the iterate (the work of k_cg_finalize after a preconditioned
// solve, done in the same pass over the iterate; the launch domain is the whole grid).
constexpr int TR_NT = 256;
// One plane's loads for the true residual: the row code, b and the x of the cell's x, y and +z
// neighbours (the cell's own x and its -z neighbour's are the previous planes' +z values).
struct TRPlane {
    uint16_t rc;
    double b, xm, xp, ym, yp, zp;
};
template <bool XD>
__device__ inline void tr_load(TRPlane& p, const Dom& d, int x, int y, int z, const uint16_t* __restrict__ rcode,
                               const float* __restrict__ b, const float* __restrict__ xf, const double* __restrict__ xdbl) {
    const int zz = min(z, d.D - 1);  // a plane past the chunk: loaded from a valid plane, never used
    const int c = (zz * d.H + y) * d.W + x;
    const Nb n = nb_offsets(d, x, y, zz);
    auto X = [&](int q) -> double {
        if constexpr (XD) return xdbl[q];
        else return (double)xf[q];
    };
    p.rc = rcode[c];
    p.b = (double)b[c];
    p.xm = X(c + n.xm);
    p.xp = X(c + n.xp);
    p.ym = X(c + n.ym);
    p.yp = X(c + n.yp);
    p.zp = X(c + n.zp);
}
// Two planes of loads stay in flight (slots s0 and s1, used alternately): one plane per wavefront
// left the pass waiting on every plane's loads.
template <bool XD>
__global__ __launch_bounds__(TR_NT) void k_true_residual(Dom d, const uint16_t* __restrict__ rcode, const float* __restrict__ b,
                                                         const float* __restrict__ xf, const double* __restrict__ xdbl,
                                                         double* partials, float* __restrict__ pout) {
    int bx, by, bz;
    double v[2] = {0.0, 0.0};
    if (block_coords(d, bx, by, bz)) {
        const int x = d.x0 + bx * TX + threadIdx.x, y = d.y0 + by * d.ty + threadIdx.y;
        const int zb = d.z0 + bz * d.zc, ze = min(zb + d.zc, d.z0 + d.nz);
        if (x < d.x0 + d.nx && y < d.y0 + d.ny && zb < ze) {
            auto X = [&](int q) -> double {
                if constexpr (XD) return xdbl[q];
                else return (double)xf[q];
            };
            // the cell's value and its -z neighbour's are the last plane's cell and +z values
            const int c0 = (zb * d.H + y) * d.W + x;
            double xc = X(c0), xzm = X(c0 + nb_offsets(d, x, y, zb).zm);
            TRPlane s0, s1;
            tr_load<XD>(s0, d, x, y, zb, rcode, b, xf, xdbl);
            tr_load<XD>(s1, d, x, y, zb + 1, rcode, b, xf, xdbl);
            auto step = [&](const TRPlane& p, int z) {
                const int c = (z * d.H + y) * d.W + x;
                const double t = apply_row_d(p.rc, xc, p.xm, p.xp, p.ym, p.yp, p.zp, xzm);
                const double rr = p.b - t;
                v[0] += rr * rr;
                v[1] += p.b * p.b;
                if (pout) pout[c] = (float)xc;
                xzm = xc;
                xc = p.zp;
            };
            int z = zb;
            for (; z + 1 < ze; z += 2) {
                step(s0, z);
                tr_load<XD>(s0, d, x, y, z + 2, rcode, b, xf, xdbl);
                step(s1, z + 1);
                tr_load<XD>(s1, d, x, y, z + 3, rcode, b, xf, xdbl);
            }
            if (z < ze) step(s0, z);
        }
    }
    __shared__ double red[2][TR_NT / 64];
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const double w = wave_sum(v[k]);
        if ((lane & 63) == 0) red[k][lane >> 6] = w;
    }
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 2; k++) {
            double t = 0.0;
            for (int w = 0; w < TR_NT / 64; w++) t += red[k][w];
            partials[(size_t)k * gridDim.x + blockIdx.x] = t;
        }
    }
}
__global__ __launch_bounds__(TR_NT) void k_true_residual_sum(int nparts, const double* __restrict__ partials, CGState* st) {
    __shared__ double red[TR_NT / 64];
    const int lane = threadIdx.x;
    double out[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        double t = 0.0;
        for (int i = lane; i < nparts; i += TR_NT) t += partials[(size_t)k * nparts + i];
        t = wave_sum(t);
        if ((lane & 63) == 0) red[lane >> 6] = t;
        __syncthreads();
        out[k] = 0.0;
        for (int w = 0; w < TR_NT / 64; w++) out[k] += red[w];
        __syncthreads();
    }
    if (lane == 0) {
        st->true_num = out[0];
        st->bnorm2 = out[1];
    }
}

// Plain operator application y = A x over the whole grid (efs_apply_operator).
__global__ __launch_bounds__(256) void k_spmv(Dom d, const uint16_t* __restrict__ rcode, const float* __restrict__ xv,
                                              float* __restrict__ yv) {
    int bx, by, bz;
    if (!block_coords(d, bx, by, bz)) return;
    const int x = d.x0 + bx * TX + threadIdx.x, y = d.y0 + by * d.ty + threadIdx.y;
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
// p' = (sum - b) / nCount, 0 when nCount == 0; identity rows copy p.  Same tiling as K2.
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(CG_NT) void k_jacobi(Dom d, const uint16_t* __restrict__ rcode, const float* __restrict__ b,
                                                  const float* __restrict__ pin, float* __restrict__ pout) {
    __shared__ float sh[2][CG_TY + 2][TX + 2];
    const int tx = threadIdx.x, ty = threadIdx.y;
    int bx, by, bz;
    if (!block_coords(d, bx, by, bz)) return;
    const int x = d.x0 + bx * TX + tx, y = d.y0 + by * CG_TY + ty;
    const bool in_dom = x < d.x0 + d.nx && y < d.y0 + d.ny;
    Stager<CG_TY> S;
    S.init(d, x, y);
    const int zb = d.z0 + bz * d.zc, ze = min(zb + d.zc, d.z0 + d.nz);
    const int sz = d.W * d.H;
    int c = (zb * d.H + y) * d.W + x;
    float pm = (zb > 0 && S.own_ok) ? pin[c - sz] : 0.f;
    float pc = S.own_ok ? pin[c] : 0.f;
    sh[0][ty + 1][tx + 1] = pc;
    if (S.hx_slot >= 0) sh[0][ty + 1][S.hx_slot] = S.hx_ok ? pin[c + S.hx_off] : 0.f;
    if (S.hy_slot >= 0) sh[0][S.hy_slot][tx + 1] = S.hy_ok ? pin[c + S.hy_off] : 0.f;
    __syncthreads();
    int cur = 0;
    for (int z = zb; z < ze; z++, c += sz) {
        const bool has_next = z + 1 < d.D;
        const float np = (has_next && S.own_ok) ? pin[c + sz] : 0.f;
        const float nhx = (has_next && S.hx_ok) ? pin[c + sz + S.hx_off] : 0.f;
        const float nhy = (has_next && S.hy_ok) ? pin[c + sz + S.hy_off] : 0.f;
        if (in_dom) {
            const unsigned rc = rcode[c];
            float out = pc;
            if (rc & RC_JAC_ROW) {
                float sum = 0.0f;
                if (rc & RC_NB_XM) sum += sh[cur][ty + 1][tx];
                if (rc & RC_NB_XP) sum += sh[cur][ty + 1][tx + 2];
                if (rc & RC_NB_YM) sum += sh[cur][ty][tx + 1];
                if (rc & RC_NB_YP) sum += sh[cur][ty + 2][tx + 1];
                if (rc & RC_NB_ZM) sum += pm;
                if (rc & RC_NB_ZP) sum += np;
                const unsigned cnt = (rc >> RC_CNT_SHIFT) & 7u;
                out = cnt > 0 ? (sum - b[c]) / (float)cnt : 0.0f;
            }
            pout[c] = out;
        }
        sh[cur ^ 1][ty + 1][tx + 1] = np;
        if (S.hx_slot >= 0) sh[cur ^ 1][ty + 1][S.hx_slot] = nhx;
        if (S.hy_slot >= 0) sh[cur ^ 1][S.hy_slot][tx + 1] = nhy;
        __syncthreads();
        cur ^= 1;
        pm = pc;
        pc = np;
    }
}

}  // namespace

int efs_solver_alloc(efs_sim* s) {
    const size_t N = (size_t)s->N;
    EFS_HIP(hipMalloc(&s->rowcode, N * sizeof(uint16_t)));
    float** fl[] = {&s->cg_r, &s->cg_d0, &s->cg_d1, &s->cg_x1, &s->cg_xd};
    for (float** p : fl) {
        EFS_HIP(hipMalloc(p, N * 4));
        EFS_HIP(hipMemset(*p, 0, N * 4));
    }
    EFS_HIP(hipMalloc(&s->cg_x64, N * 8));
    EFS_HIP(hipMemset(s->cg_x64, 0, N * 8));
    // the interior launch can have more blocks than the full one (its z chunks round shorter:
    // 2304 against 2200 at 194^3), so both are counted
    s->max_blocks = 64;
    for (bool full : {true, false})
        s->max_blocks = std::max(s->max_blocks, std::max(launch_blocks(dom_simple(s, full)), launch_blocks(dom_tiled(s, full))) + 64);
    EFS_HIP(hipMalloc(&s->partials, (size_t)s->max_blocks * 2 * sizeof(double)));
    double** pp[] = {&s->p_rr, &s->p_bb, &s->p_dad};
    for (double** p : pp) {
        EFS_HIP(hipMalloc(p, (size_t)s->max_blocks * sizeof(double)));
        EFS_HIP(hipMemset(*p, 0, (size_t)s->max_blocks * sizeof(double)));
    }
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
    hipFree(s->cg_xd);
    hipFree(s->cg_x64);
    hipFree(s->partials);
    hipFree(s->p_rr);
    hipFree(s->p_bb);
    hipFree(s->p_dad);
    hipFree(s->st);
    if (s->st_host) hipHostFree(s->st_host);
}

int efs_launch_rowcode(efs_sim* s, int* border_rows) {
    const DevGrid g = s->grid();
    CGState init{};
    std::memcpy(s->st_host, &init, sizeof(CGState));
    EFS_HIP(hipMemcpyAsync(s->st, s->st_host, sizeof(CGState), hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k_rowcode, xcd_grid((g.W + 63) / 64, (g.H + 3) / 4, (g.D + RCZ - 1) / RCZ), dim3(64, 4, 1), 0, s->stream, g,
                       s->rowcode, s->st);
    EFS_HIP(hipGetLastError());
    EFS_HIP(hipMemcpyAsync(s->st_host, s->st, sizeof(CGState), hipMemcpyDeviceToHost, s->stream));
    EFS_HIP(hipStreamSynchronize(s->stream));
    *border_rows = s->st_host->flags & 1;
    return 0;
}

// Event pair of a timed launch (both null when kernel timing is off).  The events go to
// hipExtLaunchKernelGGL, which records the start and the end of the kernels themselves, so timing
// adds no marker packets between the launches.
void efs_ktimer_pair(efs_sim* s, hipEvent_t* a, hipEvent_t* b) {
    *a = *b = nullptr;
    if (!s->ktiming) return;
    if (hipEventCreate(a) != hipSuccess) {
        *a = nullptr;
        return;
    }
    if (hipEventCreate(b) != hipSuccess) {
        hipEventDestroy(*a);
        *a = *b = nullptr;
    }
}
void efs_ktimer_add(efs_sim* s, int k, hipEvent_t a, hipEvent_t b) {
    if (!a) return;
    s->ktimers[k].start.push_back(a);
    s->ktimers[k].stop.push_back(b);
    s->ktimers[k].launches++;
}

void efs_trim_timer(TimerSlot& t, size_t keep) {
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
    const Dom d = dom_simple(s, true);
    hipLaunchKernelGGL(k_spmv, dim3(launch_blocks(d)), dim3(TX, d.ty), 0, s->stream, d, s->rowcode, x, y);
    EFS_HIP(hipGetLastError());
    return 0;
}

int efs_true_residual(efs_sim* s, double* num, double* den, bool write_p) {
    const Dom d = dom_simple(s, true);
    static_assert(TX * 4 == TR_NT, "dom_simple tiles are TX x 4");
    const int nb = launch_blocks(d);
    write_p = write_p && s->p64_valid;
    if (s->p64_valid)
        hipLaunchKernelGGL(k_true_residual<true>, dim3(nb), dim3(TX, d.ty), 0, s->stream, d, s->rowcode, s->div, s->p, s->cg_x64,
                           s->partials, write_p ? s->p : nullptr);
    else
        hipLaunchKernelGGL(k_true_residual<false>, dim3(nb), dim3(TX, d.ty), 0, s->stream, d, s->rowcode, s->div, s->p,
                           (const double*)nullptr, s->partials, (float*)nullptr);
    hipLaunchKernelGGL(k_true_residual_sum, dim3(1), dim3(TR_NT), 0, s->stream, nb, s->partials, s->st);
    EFS_HIP(hipGetLastError());
    EFS_HIP(hipMemcpyAsync(s->st_host, s->st, sizeof(CGState), hipMemcpyDeviceToHost, s->stream));
    EFS_HIP(hipStreamSynchronize(s->stream));
    *num = s->st_host->true_num;
    *den = s->st_host->bnorm2;
    return 0;
}

int efs_cg_setup_x0(efs_sim* s) {
    s->x0_pending = -1;
    // the previous iterate is usable only if the last solve left one (CG, same grid)
    const int prev = (s->P.initial_guess == EFS_GUESS_PREVIOUS && s->x64_has_previous) ? 1 : 0;
    hipLaunchKernelGGL(k_cg_setup, dim3(blocks1d(s->N)), dim3(256), 0, s->stream, s->N, s->div, s->rowcode, prev, s->cg_x64,
                       s->cg_xd);
    EFS_HIP(hipGetLastError());
    return 0;
}

int efs_cg_defer_x0(efs_sim* s) {
    s->x0_pending = (s->P.initial_guess == EFS_GUESS_PREVIOUS && s->x64_has_previous) ? 1 : 0;
    return 0;
}

int efs_cg_init_residual(efs_sim* s, bool border) {
    const Dom dt = dom_tiled(s, border);
    const dim3 grid(launch_blocks(dt)), blk(TX, CG_TY);
    if (s->x0_pending == 0)
        hipLaunchKernelGGL(k_cg_init<0>, grid, blk, 0, s->stream, dt, s->rowcode, s->div, s->cg_x64, s->cg_r, s->cg_d0, s->p_rr,
                           s->p_bb, s->cg_xd);
    else if (s->x0_pending == 1)
        hipLaunchKernelGGL(k_cg_init<1>, grid, blk, 0, s->stream, dt, s->rowcode, s->div, s->cg_x64, s->cg_r, s->cg_d0, s->p_rr,
                           s->p_bb, s->cg_xd);
    else
        hipLaunchKernelGGL(k_cg_init<-1>, grid, blk, 0, s->stream, dt, s->rowcode, s->div, s->cg_x64, s->cg_r, s->cg_d0, s->p_rr,
                           s->p_bb, s->cg_xd);
    s->x0_pending = -1;
    EFS_HIP(hipGetLastError());
    return 0;
}

int efs_cg_fold_all(efs_sim* s) {
    hipLaunchKernelGGL(k_cg_fold_all, dim3(blocks1d(s->N)), dim3(256), 0, s->stream, s->N, s->cg_x64, s->cg_xd);
    EFS_HIP(hipGetLastError());
    return 0;
}

int efs_cg_finalize(efs_sim* s, bool xd_zero) {
    hipLaunchKernelGGL(k_cg_finalize, dim3(blocks1d(s->N)), dim3(256), 0, s->stream, s->N, s->cg_x64,
                       xd_zero ? nullptr : s->cg_xd, s->p);
    EFS_HIP(hipGetLastError());
    s->p64_valid = true;
    s->x64_has_previous = true;
    return 0;
}

int efs_cg_solve(efs_sim* s, efs_stats* out) {
    int border = 0;
    int rc = efs_launch_rowcode(s, &border);
    if (rc) return rc;
    // x0 = b (CGSolver.cpp:429), set by the first residual kernel
    if ((rc = efs_cg_defer_x0(s))) return rc;
    return efs_cg_run(s, out, border != 0, s->P.max_steps);
}

// Mixed-precision CG from the FP64 iterate already in cg_x64 (xd must be 0), with the row
// codes current; runs at most max_it iterations.  With a preconditioner `pc` (z = M r into
// pc_z, then the r.z partials via k_cg_rz) the same kernels run preconditioned CG; the host then
// checks convergence every `batch` iterations instead of 64, since few iterations are expected.
int efs_cg_run_pc(efs_sim* s, efs_stats* out, bool border, int max_it, const CgPrecond* pc) {
    int rc = 0;
    s->p64_valid = false;
    const Dom dt = dom_tiled(s, border);
    const int nbt = launch_blocks(dt);
    if (nbt > s->max_blocks) {
        efs_set_error("launch exceeds the partial-sum buffers");
        return -1;
    }
    CGState init{};
    init.thr_in = (double)s->P.solver_threshold;
    init.thr = init.thr_in;
    init.flags = (s->P.criterion == EFS_CRIT_RELATIVE) ? 2 : 0;
    init.max_iter = max_it;
    std::memcpy(s->st_host, &init, sizeof(CGState));
    EFS_HIP(hipMemcpyAsync(s->st, s->st_host, sizeof(CGState), hipMemcpyHostToDevice, s->stream));
    const dim3 bt(TX, CG_TY);
    if ((rc = efs_cg_init_residual(s, border))) return rc;
    const float* dsrc = s->cg_r;
    const double* p_rz = nullptr;
    if (pc) {
        // z_0 = M r_0, partial r_0 . z_0
        if ((rc = pc->apply(s, pc->ctx))) return rc;
        if (!pc->writes_rz) hipLaunchKernelGGL(k_cg_rz, dim3(nbt), bt, 0, s->stream, dt, s->cg_r, pc->z, s->st, pc->p_rz);
        dsrc = pc->z;
        p_rz = pc->p_rz;
    }
    const int batch = pc ? pc->batch : 64;
    int next_batch = (pc && pc->first_batch > 0) ? pc->first_batch : batch;
    int launched = 0;
    if (pc) pc->stalled = false;
    std::vector<std::pair<int, double>> hist;  // (iteration, r.r) at the host checks
    const size_t kt0 = s->ktimers[KT_CG_DIR_SPMV].start.size(), kt1 = s->ktimers[KT_CG_UPDATE].start.size();
    while (true) {
        const int n = std::min(next_batch, max_it - launched);
        next_batch = batch;
        for (int k = 0; k < n; k++) {
            const int i = launched + k;
            hipEvent_t a, b;
            efs_ktimer_pair(s, &a, &b);
            hipExtLaunchKernelGGL(k_cg_dir_spmv<false>, dim3(nbt), bt, 0, s->stream, a, b, 0, dt, i, s->rowcode, dsrc, s->cg_d0,
                                  s->cg_d1, s->st, s->p_rr, s->p_bb, s->p_dad, p_rz, (const double*)nullptr, (float*)nullptr);
            efs_ktimer_add(s, KT_CG_DIR_SPMV, a, b);
            // residual replacement every 64 iterations (CGSolver.cpp:444-451), including i = 0 as in the
            // reference; preconditioned CG, which has no reference schedule to follow, updates its
            // FP64 iterate at i = 0 like any other iteration
            if ((i & 0x3F) == 0 && !(pc && i == 0)) {
                hipLaunchKernelGGL(k_cg_fold, dim3(nbt), bt, 0, s->stream, dt, i, s->cg_d0, s->cg_d1, s->cg_xd, s->cg_x64, s->st,
                                   s->p_dad);
                hipLaunchKernelGGL(k_cg_residual, dim3(nbt), bt, 0, s->stream, dt, s->rowcode, s->div, s->cg_x64, s->cg_r, s->st,
                                   s->p_rr);
            } else {
                efs_ktimer_pair(s, &a, &b);
                if (pc)
                    hipExtLaunchKernelGGL(k_cg_update<double>, dim3(nbt), bt, 0, s->stream, a, b, 0, dt, i, s->rowcode, s->cg_r,
                                          s->cg_d0, s->cg_d1, s->cg_x64, s->st, s->p_dad, s->p_rr);
                else
                    hipExtLaunchKernelGGL(k_cg_update<float>, dim3(nbt), bt, 0, s->stream, a, b, 0, dt, i, s->rowcode, s->cg_r,
                                          s->cg_d0, s->cg_d1, s->cg_xd, s->st, s->p_dad, s->p_rr);
                efs_ktimer_add(s, KT_CG_UPDATE, a, b);
            }
            if (pc) {
                if ((rc = pc->apply(s, pc->ctx))) return rc;
                if (!pc->writes_rz)
                    hipLaunchKernelGGL(k_cg_rz, dim3(nbt), bt, 0, s->stream, dt, s->cg_r, pc->z, s->st, pc->p_rz);
            }
        }
        launched += n;
        // convergence test of iteration `launched` (no-op if a K1 already stopped the solve)
        hipLaunchKernelGGL(k_cg_check, dim3(1), bt, 0, s->stream, nbt, launched, s->st, s->p_rr, s->p_bb, p_rz);
        EFS_HIP(hipGetLastError());
        EFS_HIP(hipMemcpyAsync(s->st_host, s->st, sizeof(CGState), hipMemcpyDeviceToHost, s->stream));
        EFS_HIP(hipStreamSynchronize(s->stream));
        if (s->st_host->done || launched >= max_it) break;
        if (pc && pc->stall_window > 0) {
            hist.emplace_back(launched, s->st_host->gamma);
            for (auto h = hist.rbegin(); h != hist.rend(); ++h)
                if (h->first <= launched - pc->stall_window) {
                    pc->stalled = !(s->st_host->gamma < pc->stall_ratio * h->second);
                    break;
                }
            if (pc->stalled) break;
        }
    }
    CGState fin = *s->st_host;
    if (pc && pc->stalled) fin.iter = launched;  // stopped by the host between checks
    // launches after convergence exit at once; keep the events of real iterations only
    efs_trim_timer(s->ktimers[KT_CG_DIR_SPMV], kt0 + fin.iter);
    size_t repl = (size_t)(fin.iter + 63) / 64;  // replacement iterations have no K2
    if (pc && repl > 0) repl--;                   // (preconditioned CG: not i = 0)
    efs_trim_timer(s->ktimers[KT_CG_UPDATE], kt1 + (size_t)fin.iter - std::min<size_t>(repl, fin.iter));
    // preconditioned CG updates the FP64 iterate directly and folds clear xd: it is 0 here, and
    // with the true residual requested the FP32 pressure is written by that pass
    const bool fused_p = pc != nullptr && s->P.true_residual && out;
    if (fused_p) {
        s->p64_valid = true;
        s->x64_has_previous = true;
    } else if ((rc = efs_cg_finalize(s, pc != nullptr))) {
        return rc;
    }
    if (out) {
        out->iterations = fin.iter;
        out->rr = fin.gamma;
        out->bnorm = std::sqrt(fin.bnorm2);
        out->rel_residual = fin.bnorm2 > 0 ? std::sqrt(fin.gamma / fin.bnorm2) : 0.0;
        out->true_rel_residual = -1.0;
    }
    if (s->P.true_residual && out) {
        double num = 0, den = 0;
        rc = efs_true_residual(s, &num, &den, fused_p);
        if (rc) return rc;
        out->true_rel_residual = den > 0 ? std::sqrt(num / den) : std::sqrt(num);
    }
    if (!std::isfinite(fin.gamma)) return 2;
    return fin.converged ? 0 : 1;
}

int efs_cg_run(efs_sim* s, efs_stats* out, bool border, int max_it) { return efs_cg_run_pc(s, out, border, max_it, nullptr); }

// Flexible CG's direction pass (the learned preconditioner's solve): d_it = z + beta d_{it-1},
// q = A d_it, partials of d.q into p_dad; a no-op once st->done is set.
int efs_fcg_dir(efs_sim* s, const Dom& dt, int it, const float* z, const double* p_zq, float* q) {
    hipLaunchKernelGGL(k_cg_dir_spmv<true>, dim3(launch_blocks(dt)), dim3(TX, CG_TY), 0, s->stream, dt, it, s->rowcode, z, s->cg_d0,
                       s->cg_d1, s->st, s->p_rr, s->p_bb, s->p_dad, nullptr, p_zq, q);
    EFS_HIP(hipGetLastError());
    return 0;
}

int efs_jacobi_solve(efs_sim* s, efs_stats* out) {
    int border = 0;
    int rc = efs_launch_rowcode(s, &border);
    if (rc) return rc;
    s->p64_valid = false;
    s->x64_has_previous = false;
    const Dom d = dom_tiled(s, border != 0);
    const int nb = launch_blocks(d);
    const size_t bytes = (size_t)s->N * 4;
    // p was cleared by project(); both ping-pong buffers start from it
    EFS_HIP(hipMemcpyAsync(s->cg_x1, s->p, bytes, hipMemcpyDeviceToDevice, s->stream));
    const int iters = s->P.max_steps / 2;
    for (int i = 0; i < iters; i++) {
        const float* pin = (i & 1) ? s->cg_x1 : s->p;
        float* pout = (i & 1) ? s->p : s->cg_x1;
        hipEvent_t a, b;
        efs_ktimer_pair(s, &a, &b);
        hipExtLaunchKernelGGL(k_jacobi, dim3(nb), dim3(TX, CG_TY), 0, s->stream, a, b, 0, d, s->rowcode, s->div, pin, pout);
        efs_ktimer_add(s, KT_JACOBI, a, b);
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
