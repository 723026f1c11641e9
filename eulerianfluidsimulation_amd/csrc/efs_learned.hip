// Learned preconditioner and warm start for the pressure solve (no counterpart in the
// reference; the plain solve it accelerates is CGSolverCPU3D::solve, CGSolver.cpp:416-465).
//
// M(r) = out_scale * sigma * P U(R r / sigma) - w_j * r / nCount   on CG rows, 0 elsewhere
//   R  restriction to the half-resolution grid: per coarse cell, the sum of r over its fluid
//      children / 8, plus the fluid and solid fractions (16-channel BF16 network input)
//   U  the 4-level U-Net (efs_unet.hip), BF16 on the matrix cores, FP32 accumulation
//   P  trilinear prolongation (cell-centred weights 3/4, 1/4 per axis)
//   sigma = RMS of r over the interior, so the network sees unit-scale inputs
// The preconditioner is nonlinear, so the outer solver is flexible CG (Notay's FCG(1)):
//   z_k = M(r_k); d_k = z_k - (z_k . q_{k-1}) / (d_{k-1} . q_{k-1}) d_{k-1}; q_k = A d_k;
//   alpha = (r_k . z_k) / (d_k . q_k); x += alpha d_k; r -= alpha q_k
// with the same FP64 iterate accumulator and FP64 residual replacement every 64 iterations as
// the plain solver.  Warm start: x0 = M(b).  Fallback: when the residual stalls (less than a
// factor `stall_ratio` over `stall_window` iterations), is non-finite, or the flexible solve
// reaches `max_iters`, the plain CG continues from the current iterate.  Convergence is judged
// on the same residual as the plain solver (and on the FP64 true residual when requested).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "efs_internal.h"
#include "efs_stencil.h"
#include "efs_unet.h"

namespace {

struct Learned {
    efs_unet net = nullptr;
    efs_learned_opts opts{};
    float *z = nullptr, *q = nullptr, *y = nullptr;
    double *p_rz = nullptr, *p_zq = nullptr;
    size_t n = 0;
    CGState* slot = nullptr;      // pinned host copies of the solver state, by iteration parity
    hipEvent_t ev[2] = {nullptr, nullptr};
};

__device__ inline uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

struct Coarse {
    int W, H, D;      // fine grid dims
    int nx, ny, nz;   // coarse dims (interior / 2)
};

// sigma = sqrt(sum of squares / interior cells), identical in every block
__device__ inline double sigma_from(const double* parts, int nparts, double ninterior, double* red) {
    const double ss = sum_partials(parts, nparts, red);
    const double s = sqrt(ss / ninterior);
    return (s > 0.0 && isfinite(s)) ? s : 1.0;
}

// sigma of a sum of squares ss over n cells
__device__ inline double sigma_of(double ss, double n) {
    const double s = sqrt(ss / n);
    return (s > 0.0 && isfinite(s)) ? s : 1.0;
}

// R: one thread per coarse cell; writes the 16-channel network input (BF16 or FP32) as two
// 16-byte (BF16) or four (FP32) stores.  sigma from `parts` (the sum-of-squares partials of v,
// stored in st->sigma by block 0), or, with parts null (the flexible iterations), the value the
// convergence check stored; then the pass is a no-op once the solve is done.
template <typename TO>
__global__ __launch_bounds__(CG_NT) void k_restrict(Coarse cg, const uint16_t* __restrict__ rcode, const uint8_t* __restrict__ bf,
                                                    const float* __restrict__ v, const double* parts, int nparts,
                                                    CGState* st, TO* __restrict__ out) {
    __shared__ double red[CG_NT / 64];
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
    double sigma;
    if (parts) {
        sigma = sigma_from(parts, nparts, (double)(cg.W - 2) * (cg.H - 2) * (cg.D - 2), red);
        if (blockIdx.x == 0 && lane == 0) st->sigma = sigma;
    } else {
        if (st->done) return;
        sigma = st->sigma;
    }
    const long long i = (long long)blockIdx.x * CG_NT + lane;
    const long long nc = (long long)cg.nx * cg.ny * cg.nz;
    if (i >= nc) return;
    const int cx = (int)(i % cg.nx), cy = (int)((i / cg.nx) % cg.ny), cz = (int)(i / ((long long)cg.nx * cg.ny));
    float s = 0.f;
    int nf = 0, ns = 0;
    // all 24 loads first, then the tests: a load after its row code's test would be a second
    // round trip (v of a cell that is not a CG row is loaded and not used)
    unsigned rcs[8], bfs[8];
    float vs[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int x = 1 + 2 * cx + (k & 1), y = 1 + 2 * cy + ((k >> 1) & 1), z = 1 + 2 * cz + (k >> 2);
        const int c = (z * cg.H + y) * cg.W + x;
        rcs[k] = rcode[c];
        bfs[k] = bf[c];
        vs[k] = v[c];
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
        if (rcs[k] & RC_CG_ROW) {
            s += vs[k];
            nf++;
        }
        if (bfs[k] & BF_EQ1) ns++;
    }
    const float inv = (float)(1.0 / (8.0 * sigma));
    const float v0 = s * inv, v1 = nf * 0.125f, v2 = ns * 0.125f;  // channels 3-15 are 0
    if constexpr (sizeof(TO) == 2) {
        uint4* o = reinterpret_cast<uint4*>(out + i * 16);
        o[0] = make_uint4((uint32_t)f2bf(v0) | ((uint32_t)f2bf(v1) << 16), (uint32_t)f2bf(v2), 0u, 0u);
        o[1] = make_uint4(0u, 0u, 0u, 0u);
    } else {
        float4* o = reinterpret_cast<float4*>(out + i * 16);
        o[0] = make_float4(v0, v1, v2, 0.f);
        o[1] = o[2] = o[3] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// trilinear prolongation of the coarse prediction at a fine interior cell
__device__ inline float prolong(const Coarse& cg, const float* __restrict__ y, int x, int yy, int z) {
    auto axis = [](int u, int n, int& i0, int& i1, float& w0) {
        // fine interior index u (0-based) sits at coarse coordinate u/2 - 0.25
        const int c = u >> 1;
        if (u & 1) { i0 = c; i1 = min(c + 1, n - 1); w0 = 0.75f; }
        else { i0 = max(c - 1, 0); i1 = c; w0 = 0.25f; }
    };
    int x0, x1, y0, y1, z0, z1;
    float wx, wy, wz;
    axis(x - 1, cg.nx, x0, x1, wx);
    axis(yy - 1, cg.ny, y0, y1, wy);
    axis(z - 1, cg.nz, z0, z1, wz);
    auto Y = [&](int a, int b, int c) { return y[((long long)c * cg.ny + b) * cg.nx + a]; };
    const float c00 = wx * Y(x0, y0, z0) + (1.f - wx) * Y(x1, y0, z0);
    const float c10 = wx * Y(x0, y1, z0) + (1.f - wx) * Y(x1, y1, z0);
    const float c01 = wx * Y(x0, y0, z1) + (1.f - wx) * Y(x1, y0, z1);
    const float c11 = wx * Y(x0, y1, z1) + (1.f - wx) * Y(x1, y1, z1);
    const float c0 = wy * c00 + (1.f - wy) * c10, c1 = wy * c01 + (1.f - wy) * c11;
    return wz * c0 + (1.f - wz) * c1;
}

// z = M(r) on the fine interior; partials r.z and z.q (q = A d of the previous iteration).
// mode 1 writes the warm start into the FP64 iterate instead (x64 = M(b), no partials).
__global__ __launch_bounds__(CG_NT) void k_prolong(Dom d, Coarse cg, const uint16_t* __restrict__ rcode, const float* __restrict__ r,
                                                   const float* __restrict__ y, const float* __restrict__ q, float out_scale,
                                                   float jw, const CGState* st, float* __restrict__ zout, double* __restrict__ x64,
                                                   int mode, double* p_rz, double* p_zq) {
    if (mode == 0 && st->done) return;
    __shared__ double red[CG_NT / 64];
    const float scale = (float)(out_scale * st->sigma);
    int bx, by, bz;
    double vrz = 0.0, vzq = 0.0;
    if (block_coords(d, bx, by, bz)) {
        const int x = d.x0 + bx * TX + threadIdx.x, yy = d.y0 + by * d.ty + threadIdx.y;
        const int zb = d.z0 + bz * d.zc, ze = min(zb + d.zc, d.z0 + d.nz);
        if (x < d.x0 + d.nx && yy < d.y0 + d.ny) {
            for (int z = zb; z < ze; z++) {
                const int c = (z * d.H + yy) * d.W + x;
                const unsigned rc = rcode[c];
                float zv = 0.f;
                if (rc & RC_CG_ROW) {
                    const unsigned cnt = (rc >> RC_CNT_SHIFT) & 7u;
                    const float rv = r[c];
                    zv = y ? scale * prolong(cg, y, x, yy, z) : 0.f;
                    if (cnt) zv -= jw * rv / (float)cnt;
                    vrz += (double)rv * (double)zv;
                    vzq += (double)zv * (double)q[c];
                }
                if (mode == 1) {
                    if (rc & RC_CG_ROW) x64[c] = (double)zv;
                } else {
                    zout[c] = zv;
                }
            }
        }
    }
    if (mode == 0) {
        write_partial(vrz, p_rz, red);
        __syncthreads();
        write_partial(vzq, p_zq, red);
    }
}

// F2: alpha = r.z / d.q; xd += alpha d; r -= alpha q; partial r.r.  fold != 0: replacement
// iteration, x64 += xd + alpha d and xd = 0 (the residual kernel then recomputes r).  (F1, the
// direction d_k = z_k + beta d_{k-1} with q_k = A d_k, is efs_fcg_dir.)  Each thread walks its
// column chunk with the next cell's loads issued before the current cell's stores; a no-op once
// the solve is done.
__global__ __launch_bounds__(CG_NT) void k_fcg_update(Dom d, int it, const float* d0, const float* d1, const float* __restrict__ q,
                                                      float* __restrict__ r, float* __restrict__ xd, double* __restrict__ x64,
                                                      int fold, CGState* st, const double* p_rz, const double* p_dq,
                                                      double* p_rr) {
    if (st->done) return;
    __shared__ double red[CG_NT / 64];
    const double rz = sum_partials(p_rz, gridDim.x, red);
    const double dq = sum_partials(p_dq, gridDim.x, red);
    const double alpha = rz / dq;
    if (blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
        st->dq = dq;
        st->alpha = alpha;
    }
    const float* __restrict__ dd = (it & 1) ? d1 : d0;
    int bx, by, bz;
    double v = 0.0;
    if (block_coords(d, bx, by, bz)) {
        const int x = d.x0 + bx * TX + threadIdx.x, y = d.y0 + by * d.ty + threadIdx.y;
        const int zb = d.z0 + bz * d.zc, ze = min(zb + d.zc, d.z0 + d.nz);
        if (x < d.x0 + d.nx && y < d.y0 + d.ny) {
            const int sz = d.W * d.H, last = sz * d.D - 1;
            int c = (zb * d.H + y) * d.W + x;
            float xn = xd[c], dn = dd[c], rn = r[c], qn = q[c];
            for (int z = zb; z < ze; z++, c += sz) {
                const float xc = xn, dc = dn, rc = rn, qc = qn;
                const int cn = min(c + sz, last);  // past the chunk: loaded, not used
                xn = xd[cn];
                dn = dd[cn];
                rn = r[cn];
                qn = q[cn];
                const float xnew = (float)((double)xc + alpha * (double)dc);
                if (fold) {
                    x64[c] = x64[c] + (double)xnew;
                    xd[c] = 0.f;
                } else {
                    xd[c] = xnew;
                    const float rnew = (float)((double)rc - alpha * (double)qc);
                    r[c] = rnew;
                    v += (double)rnew * (double)rnew;
                }
            }
        }
    }
    if (!fold) write_partial(v, p_rr, red);
}

// FP64 residual r = b - A x64 (replacement iteration); partial r.r
__global__ __launch_bounds__(CG_NT) void k_fcg_residual(Dom d, const uint16_t* __restrict__ rcode, const float* __restrict__ b,
                                                        const double* __restrict__ x64, float* __restrict__ r, double* p_rr,
                                                        const CGState* st) {
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

// partials of v.v over the domain
__global__ __launch_bounds__(CG_NT) void k_sumsq(Dom d, const float* __restrict__ v, double* parts) {
    __shared__ double red[CG_NT / 64];
    int bx, by, bz;
    double s = 0.0;
    if (block_coords(d, bx, by, bz)) {
        const int x = d.x0 + bx * TX + threadIdx.x, y = d.y0 + by * d.ty + threadIdx.y;
        const int zb = d.z0 + bz * d.zc, ze = min(zb + d.zc, d.z0 + d.nz);
        if (x < d.x0 + d.nx && y < d.y0 + d.ny)
            for (int z = zb; z < ze; z++) {
                const float a = v[(z * d.H + y) * d.W + x];
                s += (double)a * (double)a;
            }
    }
    write_partial(s, parts, red);
}

// convergence test of iteration `it` (one block); also stores sigma = RMS of r_it over the
// `ninterior` cells for the restriction.  A no-op once the solve is done, so the state keeps the
// iteration that stopped it.
__global__ __launch_bounds__(CG_NT) void k_fcg_check(int nparts, int it, int max_it, double ninterior, CGState* st,
                                                     const double* p_rr, const double* p_bb) {
    if (st->done) return;
    __shared__ double red[CG_NT / 64];
    const double g = sum_partials(p_rr, nparts, red);
    double thr = st->thr, bb = st->bnorm2;
    if (it == 0) {
        bb = sum_partials(p_bb, nparts, red);
        thr = st->thr_in;
        if (st->flags & 2) thr = thr * thr * bb;
    }
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        st->gamma = g;
        st->sigma = sigma_of(g, ninterior);
        if (it == 0) {
            st->thr = thr;
            st->bnorm2 = bb;
        }
        st->iter = it;
        const bool conv = g < thr || g == 0.0;
        st->converged = conv ? 1 : 0;
        st->done = conv || !isfinite(g) || it >= max_it;
    }
}

// coarse training pair from the current divergence b and the solved FP64 pressure x64:
// input = R b / sigma_b (FP32, 16 channels), target = mean of x64 over the fluid children /
// (out_scale sigma_b), mask = fluid fraction
__global__ __launch_bounds__(256) void k_coarse_target(Coarse cg, const uint16_t* __restrict__ rcode, const double* __restrict__ x64,
                                                       const CGState* st, float out_scale, float* __restrict__ target,
                                                       float* __restrict__ mask) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nc = (long long)cg.nx * cg.ny * cg.nz;
    if (i >= nc) return;
    const int cx = (int)(i % cg.nx), cy = (int)((i / cg.nx) % cg.ny), cz = (int)(i / ((long long)cg.nx * cg.ny));
    double s = 0.0;
    int nf = 0;
    for (int k = 0; k < 8; k++) {
        const int x = 1 + 2 * cx + (k & 1), y = 1 + 2 * cy + ((k >> 1) & 1), z = 1 + 2 * cz + (k >> 2);
        const int c = (z * cg.H + y) * cg.W + x;
        if (rcode[c] & RC_CG_ROW) {
            s += x64[c];
            nf++;
        }
    }
    target[i] = nf ? (float)(s / nf / (out_scale * st->sigma)) : 0.f;
    mask[i] = nf * 0.125f;
}

Learned* get(efs_sim* s) { return (Learned*)s->learned; }

int ensure_buffers(efs_sim* s, Learned* L) {
    if (L->z) return 0;
    const size_t N = (size_t)s->N;
    EFS_HIP(hipMalloc(&L->z, N * 4));
    EFS_HIP(hipMalloc(&L->q, N * 4));
    EFS_HIP(hipMemset(L->z, 0, N * 4));
    EFS_HIP(hipMemset(L->q, 0, N * 4));
    EFS_HIP(hipMalloc(&L->p_rz, (size_t)s->max_blocks * 8));
    EFS_HIP(hipMalloc(&L->p_zq, (size_t)s->max_blocks * 8));
    EFS_HIP(hipMemset(L->p_zq, 0, (size_t)s->max_blocks * 8));
    EFS_HIP(hipHostMalloc((void**)&L->slot, 2 * sizeof(CGState), hipHostMallocDefault));
    for (auto& e : L->ev) EFS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    L->n = N;
    return 0;
}

Coarse coarse_of(const efs_sim* s) { return Coarse{s->W, s->H, s->D, (s->W - 2) / 2, (s->H - 2) / 2, (s->D - 2) / 2}; }

bool coarse_ok(const efs_sim* s) {
    const int n[3] = {s->W - 2, s->H - 2, s->D - 2};
    for (int a : n)
        if (a % 16) return false;
    return true;
}

// M applied to v: network input, forward, prolongation into z (mode 0) or into x64 (mode 1).
// sigma from the sum-of-squares partials of v in `parts`, or (parts null, mode 0) from the
// convergence check; in mode 0 every pass, the network's included, is a no-op once the solve is
// done.  use_net = false keeps the diagonal term only.
int apply_M(efs_sim* s, Learned* L, const Dom& dt, const float* v, const double* parts, int mode, bool use_net = true) {
    const Coarse cg = coarse_of(s);
    const int nbt = launch_blocks(dt);
    const float* y = nullptr;
    if (use_net) {
        uint16_t* inb = nullptr;
        int rc = efs_unet_input_on(L->net, cg.nz, cg.ny, cg.nx, &inb);
        if (rc) return rc;
        const long long nc = (long long)cg.nx * cg.ny * cg.nz;
        hipLaunchKernelGGL(k_restrict<uint16_t>, dim3((unsigned)((nc + CG_NT - 1) / CG_NT)), dim3(TX, CG_TY), 0, s->stream, cg,
                           s->rowcode, s->bf, v, parts, nbt, s->st, inb);
        EFS_HIP(hipGetLastError());
        if ((rc = efs_unet_forward_on(L->net, s->stream, cg.nz, cg.ny, cg.nx, &y, mode == 0 ? &s->st->done : nullptr))) return rc;
    }
    hipLaunchKernelGGL(k_prolong, dim3(nbt), dim3(TX, CG_TY), 0, s->stream, dt, cg, s->rowcode, v, y, L->q,
                       L->opts.out_scale, L->opts.jacobi_weight, s->st, L->z, s->cg_x64, mode, L->p_rz, L->p_zq);
    EFS_HIP(hipGetLastError());
    return 0;
}

}  // namespace

void efs_learned_free(efs_sim* s) {
    Learned* L = get(s);
    if (!L) return;
    hipFree(L->z);
    hipFree(L->q);
    hipFree(L->p_rz);
    hipFree(L->p_zq);
    if (L->slot) hipHostFree(L->slot);
    for (auto e : L->ev)
        if (e) hipEventDestroy(e);
    delete L;
    s->learned = nullptr;
}

int efs_learned_cg_solve(efs_sim* s, efs_stats* out) {
    Learned* L = get(s);
    if (!L || !L->net) {
        efs_set_error("learned preconditioner selected but no network attached (efs_set_learned_preconditioner)");
        return -1;
    }
    int border = 0;
    int rc = efs_launch_rowcode(s, &border);
    if (rc) return rc;
    if (border || !coarse_ok(s)) {
        // the coarse grid needs an interior that is a multiple of 16 per axis and no fluid rows on
        // the outer layer; otherwise the plain solver runs
        if ((rc = efs_cg_setup_x0(s))) return rc;
        rc = efs_cg_run(s, out, border != 0, s->P.max_steps);
        if (out) out->fallback = 2;
        return rc;
    }
    if ((rc = ensure_buffers(s, L))) return rc;
    s->p64_valid = false;
    const Dom dt = dom_tiled(s, false);
    const int nbt = launch_blocks(dt);
    const dim3 bt(TX, CG_TY);
    CGState init{};
    init.thr_in = (double)s->P.solver_threshold;
    init.thr = init.thr_in;
    init.flags = (s->P.criterion == EFS_CRIT_RELATIVE) ? 2 : 0;
    init.max_iter = s->P.max_steps;
    std::memcpy(s->st_host, &init, sizeof(CGState));
    EFS_HIP(hipMemcpyAsync(s->st, s->st_host, sizeof(CGState), hipMemcpyHostToDevice, s->stream));
    EFS_HIP(hipMemsetAsync(L->q, 0, (size_t)s->N * 4, s->stream));
    // with the previous pressure as the initial guess (time stepping) the solve starts from it and
    // its first flexible iteration applies the network to that residual; otherwise x0 = M(b)
    const bool from_previous = s->P.initial_guess == EFS_GUESS_PREVIOUS && s->x64_has_previous;
    if ((rc = efs_cg_setup_x0(s))) return rc;
    if (L->opts.warm_start && !from_previous) {
        // x0 = M(b) on the fluid rows (identity rows keep x0 = b)
        hipLaunchKernelGGL(k_sumsq, dim3(nbt), bt, 0, s->stream, dt, s->div, s->p_bb);
        if ((rc = apply_M(s, L, dt, s->div, s->p_bb, 1))) return rc;
    }
    if (L->opts.every <= 0) {
        // warm start only: the plain solver continues from x0 = M(b)
        rc = efs_cg_run(s, out, false, s->P.max_steps);
        if (out) out->warm_started = L->opts.warm_start && !from_previous;
        return rc;
    }
    // r0 = b - A x0, d0 unused (the flexible solve forms d from z)
    if ((rc = efs_cg_init_residual(s, false))) return rc;
    const int max_it = s->P.max_steps;
    const int fmax = L->opts.max_iters > 0 ? std::min(L->opts.max_iters, max_it) : max_it;
    const double ninterior = (double)(s->W - 2) * (s->H - 2) * (s->D - 2);
    // Iteration it queues the convergence check of r_it, then reads the check of iteration it - 1
    // (copied to a pinned slot) and queues z = M r_it, the direction and the update: the host
    // decides one iteration behind, so the GPU does not wait for it between iterations.  Once a
    // check sets done, the passes queued after it, the network's included, return at once, and
    // later checks keep the state of the iteration that stopped.  A stall seen in iteration
    // it - 1's check stops the flexible solve after iteration it - 1.
    std::vector<double> hist;
    int it = 0, fcg_iters = 0;
    bool stalled = false;
    CGState fin{};
    for (;; it++) {
        hipLaunchKernelGGL(k_fcg_check, dim3(1), bt, 0, s->stream, nbt, it, max_it, ninterior, s->st, s->p_rr, s->p_bb);
        EFS_HIP(hipMemcpyAsync(L->slot + (it & 1), s->st, sizeof(CGState), hipMemcpyDeviceToHost, s->stream));
        EFS_HIP(hipEventRecord(L->ev[it & 1], s->stream));
        if (it > 0) {
            EFS_HIP(hipEventSynchronize(L->ev[(it - 1) & 1]));
            const CGState h = L->slot[(it - 1) & 1];
            if (h.done) {
                fin = h;
                fcg_iters = it - 1;
                break;
            }
            hist.push_back(h.gamma);
            const int w = L->opts.stall_window;
            if (w > 0 && (int)hist.size() > w && h.gamma > L->opts.stall_ratio * hist[hist.size() - 1 - w]) {
                stalled = true;
                fcg_iters = it;
                break;
            }
        }
        if (it >= fmax) {  // no flexible iteration past fmax: settle with this iteration's check
            EFS_HIP(hipEventSynchronize(L->ev[it & 1]));
            fin = L->slot[it & 1];
            fcg_iters = it;
            stalled = !fin.done;
            break;
        }
        if ((rc = apply_M(s, L, dt, s->cg_r, nullptr, 0, it % L->opts.every == 0))) return rc;
        if ((rc = efs_fcg_dir(s, dt, it, L->z, L->p_zq, L->q))) return rc;
        const bool fold = (it & 0x3F) == 0;
        hipLaunchKernelGGL(k_fcg_update, dim3(nbt), bt, 0, s->stream, dt, it, s->cg_d0, s->cg_d1, L->q, s->cg_r, s->cg_xd,
                           s->cg_x64, fold ? 1 : 0, s->st, L->p_rz, s->p_dad, s->p_rr);
        if (fold)
            hipLaunchKernelGGL(k_fcg_residual, dim3(nbt), bt, 0, s->stream, dt, s->rowcode, s->div, s->cg_x64, s->cg_r, s->p_rr,
                               s->st);
        EFS_HIP(hipGetLastError());
    }
    if (stalled) {
        // plain CG continues from the current iterate: fold xd, then r = b - A x
        if ((rc = efs_cg_fold_all(s))) return rc;
        efs_stats plain{};
        rc = efs_cg_run(s, &plain, false, std::max(0, max_it - fcg_iters));
        if (rc < 0) return rc;
        if (out) {
            *out = plain;
            out->iterations += fcg_iters;
            out->fallback = 1;
            out->warm_started = L->opts.warm_start && !from_previous;
        }
        return rc;
    }
    if ((rc = efs_cg_finalize(s))) return rc;
    if (out) {
        out->iterations = fin.iter;
        out->rr = fin.gamma;
        out->bnorm = std::sqrt(fin.bnorm2);
        out->rel_residual = fin.bnorm2 > 0 ? std::sqrt(fin.gamma / fin.bnorm2) : 0.0;
        out->true_rel_residual = -1.0;
        out->fallback = 0;
        out->warm_started = L->opts.warm_start && !from_previous;
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

extern "C" {

void efs_default_learned_opts(efs_learned_opts* o) {
    if (!o) return;
    o->out_scale = 1.f;
    o->jacobi_weight = 0.6f;
    o->warm_start = 1;
    o->max_iters = 0;
    o->every = 1;
    o->stall_window = 20;
    o->stall_ratio = 0.9f;
}

int efs_set_learned_preconditioner(efs_handle s, efs_unet net, const efs_learned_opts* opts) {
    if (!s) return -1;
    if (s->dim != 3) {
        efs_set_error("the learned preconditioner is 3D only");
        return -1;
    }
    Learned* L = get(s);
    if (!L) {
        L = new Learned();
        s->learned = L;
    }
    L->net = net;
    if (opts) L->opts = *opts;
    else efs_default_learned_opts(&L->opts);
    // allocate here, not in the first solve: the solver's vectors, its pinned check slots and
    // the network's buffers for this coarse grid (a grid the network cannot take runs plain CG)
    if (net && coarse_ok(s)) {
        int rc = ensure_buffers(s, L);
        if (rc) return rc;
        const Coarse c = coarse_of(s);
        uint16_t* inb = nullptr;
        if ((rc = efs_unet_input_on(net, c.nz, c.ny, c.nx, &inb))) return rc;
    }
    return 0;
}

int efs_learned_coarse_dims(efs_handle s, int32_t dims[3]) {
    if (!s || !dims) return -1;
    if (s->dim != 3 || !coarse_ok(s)) {
        efs_set_error("the learned preconditioner needs an interior that is a multiple of 16 per axis");
        return -1;
    }
    const Coarse c = coarse_of(s);
    dims[0] = c.nx; dims[1] = c.ny; dims[2] = c.nz;
    return 0;
}

// Training pair for the current divergence and the pressure of the last CG solve (device
// buffers: in16 (nz, ny, nx, 16), target and mask (nz, ny, nx)); *sigma = RMS of b.
int efs_learned_coarse_pair(efs_handle s, float out_scale, float* in16, float* target, float* mask, double* sigma) {
    if (!s || !in16 || !target || !mask) return -1;
    if (s->dim != 3 || !coarse_ok(s)) {
        efs_set_error("the learned preconditioner needs an interior that is a multiple of 16 per axis");
        return -1;
    }
    if (!s->p64_valid) {
        efs_set_error("no CG solution to build a training pair from (run a CG solve first)");
        return -1;
    }
    const Dom dt = dom_tiled(s, false);
    const int nbt = launch_blocks(dt);
    const Coarse cg = coarse_of(s);
    const long long nc = (long long)cg.nx * cg.ny * cg.nz;
    hipLaunchKernelGGL(k_sumsq, dim3(nbt), dim3(TX, CG_TY), 0, s->stream, dt, s->div, s->p_bb);
    hipLaunchKernelGGL(k_restrict<float>, dim3((unsigned)((nc + CG_NT - 1) / CG_NT)), dim3(TX, CG_TY), 0, s->stream, cg, s->rowcode,
                       s->bf, s->div, s->p_bb, nbt, s->st, in16);
    hipLaunchKernelGGL(k_coarse_target, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, s->stream, cg, s->rowcode, s->cg_x64, s->st,
                       out_scale, target, mask);
    EFS_HIP(hipGetLastError());
    EFS_HIP(hipMemcpyAsync(s->st_host, s->st, sizeof(CGState), hipMemcpyDeviceToHost, s->stream));
    EFS_HIP(hipStreamSynchronize(s->stream));
    if (sigma) *sigma = s->st_host->sigma;
    return 0;
}

}  // extern "C"
