// Geometric multigrid preconditioner for the pressure CG (no counterpart in the reference: it
// accelerates the solve of CGSolverCPU3D::solve, CGSolver.cpp:416-465, on the same operator).
//
// One V-cycle is z = M r for the fluid rows of the 7-point operator of CGSolver.cpp:467-497
// (Dirichlet p = 0 at air neighbours, Neumann at solids and the outer layer).
//   Levels.  Level l+1 halves every interior extent (ceil).  A coarse cell is air if any of its
//     eight children is air, else fluid if any child is fluid, else solid (the Dirichlet-first
//     labelling of McAdams, Sifakis and Teran 2010).  The coarse operator is the same 7-point
//     stencil on the coarse labels, scaled by s_{l+1} = s_l / 4, the scaling of the continuous
//     Laplacian at twice the spacing.  (The Galerkin product R A P for piecewise-constant
//     prolongation P and restriction R = P^T / 8 would be s_l / 2; the stiffer-than-Galerkin
//     coarse correction is twice as large and, measured on the sphere-over-pool scene, halves the
//     iteration count at 256^3: 35 -> 16.)
//   Smoother.  Red-black Gauss-Seidel, one sweep down (red, black) and one sweep up (black, red)
//     so that M is symmetric.  The down sweep starts from x = 0, so both of its passes are one
//     kernel (k_mg_presmooth); the first up pass reads the coarse correction of its neighbours on
//     the fly (the prolongation is never stored).
//   Coarsest level (<= 8 cells per axis): 16 symmetric sweeps in one workgroup, in LDS.
// Non-fluid rows of the fine level return z = 0 (their residual is 0: identity rows).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "efs_internal.h"
#include "efs_stencil.h"

#ifndef EFS_MG_CFAC
#define EFS_MG_CFAC 0.25f
#endif
#ifndef EFS_MG_COARSE_SWEEPS
#define EFS_MG_COARSE_SWEEPS 16
#endif
#ifndef EFS_MG_MIN_N
#define EFS_MG_MIN_N 8
#endif
#ifndef EFS_MG_BATCH
#define EFS_MG_BATCH 8
#endif

namespace {

enum : uint8_t { CLS_SOLID = 0, CLS_AIR = 1, CLS_FLUID = 2 };

struct Lvl {
    int W, H, D;     // with the border layer
    int nx, ny, nz;  // interior
    float s;         // operator scale
    uint16_t* rc;    // row code (level 0: the solver's)
    uint8_t* cls;    // cell class (levels >= 1)
    float* x;        // correction (level 0: z)
    float* b;        // right-hand side (level 0: the CG residual)
};

struct MG {
    int W = 0, H = 0, D = 0;
    std::vector<Lvl> L;
    float* z = nullptr;
    double* p_rz = nullptr;
};

MG* get(efs_sim* s) { return (MG*)s->mg; }

constexpr int MB = 64, MBY = 4;
constexpr int MG_CO_NT = 1024;

__device__ inline int lidx(const Lvl& l, int x, int y, int z) { return (z * l.H + y) * l.W + x; }

// class of a fine-level cell (level 0) from its row code and boundary byte
__device__ inline uint8_t class0(unsigned rc, unsigned bf, bool outer) {
    if (rc & RC_CG_ROW) return CLS_FLUID;
    if (outer || (bf & BF_GT05)) return CLS_SOLID;
    return CLS_AIR;
}

// coarse labels from the children: air if any air, else fluid if any fluid, else solid
__global__ __launch_bounds__(256) void k_mg_label(Lvl f, Lvl c, const uint8_t* __restrict__ bf0, int from_level0) {
    const int x = blockIdx.x * MB + threadIdx.x, y = blockIdx.y * MBY + threadIdx.y, z = blockIdx.z;
    if (x >= c.W || y >= c.H) return;
    const int ci = lidx(c, x, y, z);
    if (x == 0 || y == 0 || z == 0 || x > c.nx || y > c.ny || z > c.nz) {
        c.cls[ci] = CLS_SOLID;
        return;
    }
    bool any_air = false, any_fluid = false;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int fx = 2 * x - 1 + (k & 1), fy = 2 * y - 1 + ((k >> 1) & 1), fz = 2 * z - 1 + (k >> 2);
        if (fx > f.nx || fy > f.ny || fz > f.nz) continue;  // past the fine interior: solid
        const int fi = lidx(f, fx, fy, fz);
        const uint8_t k_cls = from_level0 ? class0(f.rc[fi], bf0[fi], false) : f.cls[fi];
        any_air |= (k_cls == CLS_AIR);
        any_fluid |= (k_cls == CLS_FLUID);
    }
    c.cls[ci] = any_air ? CLS_AIR : (any_fluid ? CLS_FLUID : CLS_SOLID);
}

// row code of a coarse level from its labels (neighbour counted if interior and not solid,
// contributing if fluid; the bit layout of the fine row code)
__global__ __launch_bounds__(256) void k_mg_rowcode(Lvl c) {
    const int x = blockIdx.x * MB + threadIdx.x, y = blockIdx.y * MBY + threadIdx.y, z = blockIdx.z;
    if (x >= c.W || y >= c.H) return;
    const int ci = lidx(c, x, y, z);
    unsigned code = 0, cnt = 0;
    if (c.cls[ci] == CLS_FLUID) {
        auto nb = [&](int nx, int ny, int nz, unsigned bit) {
            const uint8_t k = c.cls[lidx(c, nx, ny, nz)];  // the border is solid
            if (k == CLS_SOLID) return;
            cnt++;
            if (k == CLS_FLUID) code |= bit;
        };
        nb(x - 1, y, z, RC_NB_XM);
        nb(x + 1, y, z, RC_NB_XP);
        nb(x, y - 1, z, RC_NB_YM);
        nb(x, y + 1, z, RC_NB_YP);
        nb(x, y, z + 1, RC_NB_ZP);
        nb(x, y, z - 1, RC_NB_ZM);
        code |= (cnt << RC_CNT_SHIFT) | RC_CG_ROW;
    }
    c.rc[ci] = (uint16_t)code;
}

// sum of the contributing neighbours of interior cell i (x values, plus the coarse correction
// of each neighbour's parent when cl is given).  Every load is issued unconditionally (the six
// neighbours of an interior cell are inside the array) and masked afterwards: loads under a
// branch would serialise on their latency.
__device__ inline float nb_sum(const Lvl& l, unsigned rc, int i, int x, int y, int z, const Lvl* cl) {
    const int sy = l.W, sz = l.W * l.H;
    float v[6] = {l.x[i - 1], l.x[i + 1], l.x[i - sy], l.x[i + sy], l.x[i + sz], l.x[i - sz]};
    if (cl) {
        const int px0 = x >> 1, px1 = (x + 2) >> 1, px = (x + 1) >> 1;
        const int py0 = y >> 1, py1 = (y + 2) >> 1, py = (y + 1) >> 1;
        const int pz0 = z >> 1, pz1 = (z + 2) >> 1, pz = (z + 1) >> 1;
        v[0] += cl->x[lidx(*cl, px0, py, pz)];
        v[1] += cl->x[lidx(*cl, px1, py, pz)];
        v[2] += cl->x[lidx(*cl, px, py0, pz)];
        v[3] += cl->x[lidx(*cl, px, py1, pz)];
        v[4] += cl->x[lidx(*cl, px, py, pz1)];
        v[5] += cl->x[lidx(*cl, px, py, pz0)];
    }
    // the reference order x-, x+, y-, y+, z+, z-
    const unsigned bits[6] = {RC_NB_XM, RC_NB_XP, RC_NB_YM, RC_NB_YP, RC_NB_ZP, RC_NB_ZM};
    float sum = 0.f;
#pragma unroll
    for (int k = 0; k < 6; k++) sum += (rc & bits[k]) ? v[k] : 0.f;
    return sum;
}

// One red-black Gauss-Seidel pass over the cells of `color` ((x + y + z) & 1).  mode 0: normal;
// mode 2: neighbours read x + the coarse correction of their parent (the first pass after the
// coarse-grid correction).
__global__ __launch_bounds__(256) void k_mg_smooth(Lvl l, int color, int mode, Lvl cl, const CGState* st) {
    if (st->done) return;
    const int y = 1 + blockIdx.y * MBY + threadIdx.y, z = 1 + blockIdx.z;
    if (y > l.ny) return;
    // thread k owns the pair of cells (1 + 2k, 2 + 2k) of its row: one of each colour
    const int xa = 1 + 2 * (int)(blockIdx.x * MB + threadIdx.x);
    if (xa > l.nx) return;
    const int x = (((xa + y + z) & 1) == color) ? xa : xa + 1;
    if (x > l.nx) return;
    const int i = lidx(l, x, y, z);
    const unsigned rc = l.rc[i];
    const float bi = l.b[i];
    const float sum = nb_sum(l, rc, i, x, y, z, mode == 2 ? &cl : nullptr);
    const unsigned cnt = (rc >> RC_CNT_SHIFT) & 7u;
    l.x[i] = ((rc & RC_CG_ROW) && cnt > 0) ? (sum - bi / l.s) / (float)cnt : 0.f;
}

// The first down sweep from x = 0 (red pass, then black pass) in one kernel: a red cell gets
// -b / (s n); a black cell gets (sum of its fluid red neighbours' -b_j / (s n_j) - b / s) / n,
// the values the two passes would compute, in the same operations.
__global__ __launch_bounds__(256) void k_mg_presmooth(Lvl l, const CGState* st) {
    if (st->done) return;
    const int y = 1 + blockIdx.y * MBY + threadIdx.y, z = 1 + blockIdx.z;
    if (y > l.ny) return;
    const int xa = 1 + 2 * (int)(blockIdx.x * MB + threadIdx.x);
    if (xa > l.nx) return;
    const int sy = l.W, sz = l.W * l.H;
    auto red_value = [&](unsigned rc, float b) -> float {
        const unsigned cnt = (rc >> RC_CNT_SHIFT) & 7u;
        return ((rc & RC_CG_ROW) && cnt > 0) ? (0.f - b / l.s) / (float)cnt : 0.f;
    };
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int x = xa + h;
        if (x > l.nx) break;
        const int i = lidx(l, x, y, z);
        const unsigned rc = l.rc[i];
        const float bi = l.b[i];
        if (((x + y + z) & 1) == 0) {
            l.x[i] = red_value(rc, bi);
        } else {
            const int nb[6] = {i - 1, i + 1, i - sy, i + sy, i + sz, i - sz};
            const unsigned bits[6] = {RC_NB_XM, RC_NB_XP, RC_NB_YM, RC_NB_YP, RC_NB_ZP, RC_NB_ZM};
            float v[6];
#pragma unroll
            for (int k = 0; k < 6; k++) v[k] = red_value(l.rc[nb[k]], l.b[nb[k]]);
            float sum = 0.f;
#pragma unroll
            for (int k = 0; k < 6; k++) sum += (rc & bits[k]) ? v[k] : 0.f;
            const unsigned cnt = (rc >> RC_CNT_SHIFT) & 7u;
            l.x[i] = ((rc & RC_CG_ROW) && cnt > 0) ? (sum - bi / l.s) / (float)cnt : 0.f;
        }
    }
}

// residual of level f restricted to level c: b_c(I) = 1/8 sum over the fluid children of
// (b - s A x); zero on non-fluid coarse cells.
__global__ __launch_bounds__(256) void k_mg_restrict(Lvl f, Lvl c, const CGState* st) {
    if (st->done) return;
    const int x = 1 + blockIdx.x * MB + threadIdx.x, y = 1 + blockIdx.y * MBY + threadIdx.y, z = 1 + blockIdx.z;
    if (x > c.nx || y > c.ny) return;
    const int ci = lidx(c, x, y, z);
    float acc = 0.f;
    if (c.cls[ci] == CLS_FLUID) {
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int fx = 2 * x - 1 + (k & 1), fy = 2 * y - 1 + ((k >> 1) & 1), fz = 2 * z - 1 + (k >> 2);
            if (fx > f.nx || fy > f.ny || fz > f.nz) continue;
            const int fi = lidx(f, fx, fy, fz);
            const unsigned rc = f.rc[fi];
            const float sum = nb_sum(f, rc, fi, fx, fy, fz, nullptr);
            const float ax = f.s * (sum - (float)((rc >> RC_CNT_SHIFT) & 7u) * f.x[fi]);
            const float res = f.b[fi] - ax;
            acc += (rc & RC_CG_ROW) ? res : 0.f;
        }
        acc *= 0.125f;
    }
    c.b[ci] = acc;
}

// coarsest level: symmetric red-black sweeps from x = 0 in one workgroup, on a copy of the level
// (values, right-hand side, row codes, with the border) held in LDS
constexpr int MG_CO_MAX = 1728;  // (10 + 2)^3: interiors up to 10 cells per axis
__global__ __launch_bounds__(MG_CO_NT) void k_mg_coarsest(Lvl l, int sweeps, const CGState* st) {
    if (st->done) return;
    __shared__ float xs[MG_CO_MAX], bs[MG_CO_MAX];
    __shared__ uint16_t rs[MG_CO_MAX];
    const int N = l.W * l.H * l.D;
    for (int t = threadIdx.x; t < N; t += MG_CO_NT) {
        xs[t] = 0.f;
        bs[t] = l.b[t];
        rs[t] = l.rc[t];
    }
    __syncthreads();
    const int n = l.nx * l.ny * l.nz;
    const int sy = l.W, sz = l.W * l.H;
    for (int sw = 0; sw < 2 * sweeps; sw++) {
        for (int h = 0; h < 2; h++) {
            // forward sweeps red then black, backward sweeps black then red
            const int color = (sw & 1) ? 1 - h : h;
            for (int t = threadIdx.x; t < n; t += MG_CO_NT) {
                const int x = 1 + t % l.nx, y = 1 + (t / l.nx) % l.ny, z = 1 + t / (l.nx * l.ny);
                if (((x + y + z) & 1) != color) continue;
                const int i = lidx(l, x, y, z);
                const unsigned rc = rs[i];
                const float v[6] = {xs[i - 1], xs[i + 1], xs[i - sy], xs[i + sy], xs[i + sz], xs[i - sz]};
                const unsigned bits[6] = {RC_NB_XM, RC_NB_XP, RC_NB_YM, RC_NB_YP, RC_NB_ZP, RC_NB_ZM};
                float sum = 0.f;
#pragma unroll
                for (int k = 0; k < 6; k++) sum += (rc & bits[k]) ? v[k] : 0.f;
                const unsigned cnt = (rc >> RC_CNT_SHIFT) & 7u;
                xs[i] = ((rc & RC_CG_ROW) && cnt > 0) ? (sum - bs[i] / l.s) / (float)cnt : 0.f;
            }
            __syncthreads();
        }
    }
    for (int t = threadIdx.x; t < N; t += MG_CO_NT) l.x[t] = xs[t];
}

inline dim3 cell_grid(const Lvl& l) { return dim3((l.W + MB - 1) / MB, (l.H + MBY - 1) / MBY, l.D); }
inline dim3 color_grid(const Lvl& l) { return dim3(((l.nx + 1) / 2 + MB - 1) / MB, (l.ny + MBY - 1) / MBY, l.nz); }

void free_mg(MG* m) {
    if (!m) return;
    for (size_t k = 1; k < m->L.size(); k++) {
        hipFree(m->L[k].rc);
        hipFree(m->L[k].cls);
        hipFree(m->L[k].x);
        hipFree(m->L[k].b);
    }
    hipFree(m->z);
    hipFree(m->p_rz);
    delete m;
}

int ensure(efs_sim* s) {
    MG* m = get(s);
    if (m && m->W == s->W && m->H == s->H && m->D == s->D) return 0;
    free_mg(m);
    s->mg = nullptr;
    m = new MG();
    s->mg = m;
    m->W = s->W; m->H = s->H; m->D = s->D;
    EFS_HIP(hipMalloc(&m->z, (size_t)s->N * 4));
    EFS_HIP(hipMemset(m->z, 0, (size_t)s->N * 4));
    EFS_HIP(hipMalloc(&m->p_rz, (size_t)s->max_blocks * sizeof(double)));
    EFS_HIP(hipMemset(m->p_rz, 0, (size_t)s->max_blocks * sizeof(double)));
    Lvl l0{};
    l0.W = s->W; l0.H = s->H; l0.D = s->D;
    l0.nx = s->W - 2; l0.ny = s->H - 2; l0.nz = s->D - 2;
    l0.s = 1.f;
    l0.rc = s->rowcode;
    l0.cls = nullptr;
    l0.x = m->z;
    l0.b = s->cg_r;
    m->L.push_back(l0);
    while (true) {
        const Lvl& f = m->L.back();
        if (std::max(f.nx, std::max(f.ny, f.nz)) <= EFS_MG_MIN_N || m->L.size() >= 10) break;
        Lvl c{};
        c.nx = (f.nx + 1) / 2; c.ny = (f.ny + 1) / 2; c.nz = (f.nz + 1) / 2;
        c.W = c.nx + 2; c.H = c.ny + 2; c.D = c.nz + 2;
        c.s = f.s * EFS_MG_CFAC;
        const size_t n = (size_t)c.W * c.H * c.D;
        EFS_HIP(hipMalloc(&c.rc, n * sizeof(uint16_t)));
        EFS_HIP(hipMalloc(&c.cls, n));
        EFS_HIP(hipMalloc(&c.x, n * 4));
        EFS_HIP(hipMalloc(&c.b, n * 4));
        EFS_HIP(hipMemset(c.x, 0, n * 4));  // the border stays 0
        EFS_HIP(hipMemset(c.b, 0, n * 4));
        m->L.push_back(c);
    }
    const Lvl& cl = m->L.back();
    if ((long long)cl.W * cl.H * cl.D > MG_CO_MAX) {
        efs_set_error("multigrid coarsest level too large for the LDS solve");
        return -1;
    }
    return 0;
}

// labels and row codes of every coarse level for the current fluid configuration
int setup_levels(efs_sim* s) {
    MG* m = get(s);
    for (size_t k = 1; k < m->L.size(); k++) {
        const Lvl& f = m->L[k - 1];
        const Lvl& c = m->L[k];
        hipLaunchKernelGGL(k_mg_label, cell_grid(c), dim3(MB, MBY), 0, s->stream, f, c, s->bf, k == 1 ? 1 : 0);
        hipLaunchKernelGGL(k_mg_rowcode, cell_grid(c), dim3(MB, MBY), 0, s->stream, c);
    }
    EFS_HIP(hipGetLastError());
    return 0;
}

// z = M r: one V-cycle
int vcycle(efs_sim* s, void*) {
    MG* m = get(s);
    const int nl = (int)m->L.size();
    const dim3 blk(MB, MBY);
    const CGState* st = s->st;
    // the fine-level passes of the up sweep are event-timed when kernel timing is on
    auto smooth = [&](int k, int color, int mode, const Lvl& cl) {
        const Lvl& l = m->L[k];
        hipEvent_t a = nullptr, b = nullptr;
        if (k == 0) efs_ktimer_begin(s, KT_OTHER, &a, &b);
        hipLaunchKernelGGL(k_mg_smooth, color_grid(l), blk, 0, s->stream, l, color, mode, cl, st);
        if (k == 0) efs_ktimer_end(s, KT_OTHER, a, b);
    };
    for (int k = 0; k < nl - 1; k++) {
        const Lvl& l = m->L[k];
        hipLaunchKernelGGL(k_mg_presmooth, color_grid(l), blk, 0, s->stream, l, st);
        const Lvl& c = m->L[k + 1];
        hipLaunchKernelGGL(k_mg_restrict, dim3((c.nx + MB - 1) / MB, (c.ny + MBY - 1) / MBY, c.nz), blk, 0, s->stream, l, c, st);
    }
    hipLaunchKernelGGL(k_mg_coarsest, dim3(1), dim3(MG_CO_NT), 0, s->stream, m->L[nl - 1], EFS_MG_COARSE_SWEEPS, st);
    for (int k = nl - 2; k >= 0; k--) {
        const Lvl& l = m->L[k];
        smooth(k, 1, 2, m->L[k + 1]);
        smooth(k, 0, 0, l);
    }
    EFS_HIP(hipGetLastError());
    return 0;
}

}  // namespace

void efs_mg_free(efs_sim* s) {
    free_mg(get(s));
    s->mg = nullptr;
}

int efs_mg_cg_solve(efs_sim* s, efs_stats* out) {
    int border = 0;
    int rc = efs_launch_rowcode(s, &border);
    if (rc) return rc;
    if ((rc = efs_cg_setup_x0(s))) return rc;
    // fluid rows on the outer layer (never in practice) or a grid too small to coarsen: plain CG
    if (border || std::min(s->W, std::min(s->H, s->D)) < 6) return efs_cg_run(s, out, border != 0, s->P.max_steps);
    if ((rc = ensure(s))) return rc;
    if ((rc = setup_levels(s))) return rc;
    MG* m = get(s);
    CgPrecond pc;
    pc.apply = vcycle;
    pc.ctx = nullptr;
    pc.z = m->z;
    pc.p_rz = m->p_rz;
    pc.batch = EFS_MG_BATCH;
    // The recurrence residual stops the iteration; with the true residual enabled and the relative
    // criterion, a solve whose FP64 true residual is still above the tolerance restarts from its
    // iterate (at most twice).
    efs_stats local{};
    efs_stats* o = out ? out : &local;
    int total = 0;
    for (int attempt = 0; attempt < 3; attempt++) {
        const size_t kt0 = s->ktimers[KT_OTHER].start.size();
        rc = efs_cg_run_pc(s, o, false, std::max(1, s->P.max_steps - total), &pc);
        // launches after convergence return at once: keep the 2 timed fine passes of each real V-cycle
        efs_trim_timer(s->ktimers[KT_OTHER], kt0 + 2 * ((size_t)s->st_host->iter + 1));
        total += o->iterations;
        const bool verify = s->P.true_residual && s->P.criterion == EFS_CRIT_RELATIVE;
        if (rc != 0 || !verify || o->true_rel_residual <= (double)s->P.solver_threshold) break;
    }
    o->iterations = total;
    return rc;
}
