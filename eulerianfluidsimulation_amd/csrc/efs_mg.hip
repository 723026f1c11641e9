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
//     so that M is symmetric.  Each level costs two kernels (k_mg_sweep): the down sweep from
//     x = 0 fused with the residual and its restriction to the next level, and the up sweep,
//     whose black pass reads the coarse correction of its neighbours' parents on the fly (the
//     prolongation is never stored).  x is never read: a red cell's value before the pass that
//     updates it is always the down sweep's, recomputed from b and the row code.
//   Row arithmetic.  A row's update is (sum - b / s) * (1 / n): s is a power of two, so b / s is
//     the exact product b * (1 / s); 1 / n is the hardware reciprocal (within an ulp of the
//     IEEE quotient; the same in every pass, so the V-cycle stays symmetric).
//   Coarsest level (<= 8 cells per axis): 16 symmetric sweeps in one workgroup, in LDS.  The
//     level above it runs its whole V-cycle in the same workgroup (k_mg_tail).
// Non-fluid rows of the fine level return z = 0 (their residual is 0: identity rows).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <hip/hip_ext.h>

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
// Iterations launched before the first host convergence check: the previous solve's count
// (EFS_MG_FIRST_BATCH for the first solve), since consecutive steps need about as many; then
// EFS_MG_BATCH between checks.  Iterations launched past convergence return at once but still
// dispatch every kernel of a V-cycle, so a batch sized to the expected count saves their launches.
// Shortest z chunk of a sweep block.  Levels with few tiles get short chunks, so that more
// blocks run shorter plane marches (a march step costs about as much on a small level as on the
// fine one, and the halo planes are recomputed exactly whatever the chunk).
#ifndef EFS_MG_MIN_ZC
#define EFS_MG_MIN_ZC 2
#endif
#ifndef EFS_MG_FIRST_BATCH
#define EFS_MG_FIRST_BATCH 8
#endif
#ifndef EFS_MG_BATCH
#define EFS_MG_BATCH 2
#endif
#ifndef EFS_MG_IEEE_DIV
#define EFS_MG_IEEE_DIV 0
#endif
#ifndef EFS_MG_FUSED_RZ
#define EFS_MG_FUSED_RZ 1
#endif
#ifndef EFS_MG_NU
#define EFS_MG_NU 2  // red-black Gauss-Seidel sweeps down and up per level: V(1,1) or V(2,2)
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
    float* x2;       // second buffer of the V(2,2) sweeps
};

struct MG {
    int W = 0, H = 0, D = 0;
    std::vector<Lvl> L;
    float* z = nullptr;
    double* p_rz = nullptr;
    bool fused_rz = false;  // the fine up sweep writes the r . z partials
    bool tail = false;      // the last two levels run in one workgroup (k_mg_tail)
    bool down2 = false;     // the two down sweeps of a level run in one kernel (k_mg_down2)
    int last_its = 0;       // iterations of the previous solve
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
        code |= (cnt << RC_CNT_SHIFT) | RC_CG_ROW | RC_CONTRIB;
    }
    c.rc[ci] = (uint16_t)code;
}

// Over-relaxation of the V(2,2) smoother: x <- x + w (gs - x) with the same w in every pass,
// so the symmetric cycle (red-black down, black-red up) stays symmetric; w = 1 is Gauss-Seidel.
#ifndef EFS_MG_OMEGA
#define EFS_MG_OMEGA 1.2f
#endif
// Relaxed update of a row whose old value is xo and Gauss-Seidel value gs.  Only the rows the
// Gauss-Seidel pass updates (fluid rows with counted neighbours) are relaxed; every other cell
// keeps the pass's value, 0, as in Gauss-Seidel (a fluid row with no counted neighbour is a zero
// row of A, and a nonzero value there would enter r.z but not A d).  Fluid rows without
// RC_CONTRIB (next to a boundary value in (0.5, 1), possible only on the fine level) are stored
// masked, so their old value is not kept: they take the plain Gauss-Seidel value, which keeps
// the cycle symmetric with a per-row factor.
template <bool SOR>
__device__ inline float relax(unsigned rc, float gs, float xo) {
    if (!SOR) return gs;
    const bool upd = (rc & RC_CONTRIB) && (rc & RC_CG_ROW) && ((rc >> RC_CNT_SHIFT) & 7u) > 0;
    return upd ? xo + (float)EFS_MG_OMEGA * (gs - xo) : gs;
}

// Gauss-Seidel row update: (sum - b / s) / n on fluid rows with neighbours, else 0.
__device__ inline float row_update(unsigned rc, float sum, float b, float inv_s) {
    const unsigned cnt = (rc >> RC_CNT_SHIFT) & 7u;
    const bool row = (rc & RC_CG_ROW) && cnt > 0;
#if EFS_MG_IEEE_DIV
    return row ? (sum - b * inv_s) / (float)cnt : 0.f;
#else
    return row ? (sum - b * inv_s) * __builtin_amdgcn_rcpf((float)cnt) : 0.f;
#endif
}
// the masked neighbour sum in the reference order x-, x+, y-, y+, z+, z-
__device__ inline float masked_sum6(unsigned rc, const float (&v)[6]) {
    const unsigned bits[6] = {RC_NB_XM, RC_NB_XP, RC_NB_YM, RC_NB_YP, RC_NB_ZP, RC_NB_ZM};
    float sum = 0.f;
#pragma unroll
    for (int k = 0; k < 6; k++) sum += (rc & bits[k]) ? v[k] : 0.f;
    return sum;
}

// One smoothing sweep of a level in one kernel, in two modes.
//   SW_UP (and SW_UP_RZ): the up sweep.  The black pass, whose neighbours (all red) read their value
//     plus the coarse correction of their parent, then the red pass from the new black values;
//     writes x.  With SW_UP_RZ (level 0: x = z, b = r) the block also writes its partial of r . z.
//   SW_DOWN: the down sweep from x = 0 (red, then black), then the residual b - s A x of every cell
//     and its restriction to the coarse level: b_c = 1/8 of the sum of the fluid rows' residuals
//     over the 8 children (in the child order x, y, z), 0 on non-fluid coarse cells.  x itself is
//     not stored: nothing reads it before the up sweep overwrites it.
// In both modes a red cell's value before the pass that updates it is the down sweep's
// (0 - b / s) / n, recomputed from b and the row code, so x is never read.
//
// A block owns a TX x PTY column tile and marches up a chunk of planes.  Per plane it loads b, the
// row codes (and for the up sweep the red cells' parent correction) once for the tile and a halo
// of 2, and keeps in LDS three planes of red values (halo 2), three planes of black values
// (halo 1) and four planes of b and row codes.  Each lane owns two x-adjacent cells, one of each
// colour, so every phase does one row update per lane.  The LDS values are stored masked by the
// cell's RC_CONTRIB bit, which is the neighbour bit of every row that reads them, so the neighbour
// sums need no masks (the same additions in the same order as masked sums).
#ifndef EFS_MG_PTY
#define EFS_MG_PTY 8
#endif
constexpr int PTY = EFS_MG_PTY;
constexpr int PX = TX + 4, PY = PTY + 4, PP = PX * PY;  // a plane of the tile with its halo of 2
constexpr int POST_NT = TX * PTY;
constexpr int NF = (PX / 2) * PY, NB = ((PX - 2) / 2) * (PY - 2), NT2 = (TX / 2) * PTY;  // pair lanes per phase
constexpr int NC = (TX / 2) * (PTY / 2);  // coarse cells of a tile plane pair
static_assert(POST_NT <= PP && PP <= 2 * POST_NT, "one or two fetch cells per lane");
static_assert(PTY % 2 == 0, "tiles cover whole coarse cells");
// Modes.  C0 is the colour whose values enter from the commit (halo 2), C1 the colour updated
// first (halo 1); the finish pass then updates C0 on the tile (or, SW_DOWN, takes the residuals;
// SW_DOWN0, stores the first pass's C0 values).
//   V(1,1) cycle: SW_DOWN, then SW_UP / SW_UP_RZ (C0 red, recomputed from b; C1 black).
//   V(2,2) cycle, x read from `xin` and written to `xout` (another buffer, so that no block reads
//   values a neighbouring block has already replaced):
//     SW_DOWN0  first down sweep from 0 (red from b, then black), stores both colours;
//     SW_RB     second down sweep (red from the stored black, then black), stores both;
//     then k_mg_restrict_x restricts the residual of the stored x;
//     SW_BR_E   first up sweep (black from the stored red plus its parent's correction, then red);
//     SW_BR / SW_BR_RZ  second up sweep (black from the stored red, then red).
enum { SW_UP = 0, SW_UP_RZ = 1, SW_DOWN = 2, SW_DOWN0 = 3, SW_RB = 4, SW_BR_E = 5, SW_BR = 6, SW_BR_RZ = 7 };
template <int MODE>
__global__ __launch_bounds__(POST_NT) void k_mg_sweep(Lvl l, Lvl cl, float inv_s, Dom d, const CGState* st, double* p_rz,
                                                      const float* __restrict__ xin, float* __restrict__ xout) {
    constexpr bool UP = MODE != SW_DOWN, RZ = MODE == SW_UP_RZ || MODE == SW_BR_RZ;
    constexpr bool USE_E = MODE == SW_UP || MODE == SW_UP_RZ || MODE == SW_BR_E;
    constexpr bool FROM_X = MODE == SW_RB || MODE == SW_BR_E || MODE == SW_BR || MODE == SW_BR_RZ;
    constexpr int C1 = MODE == SW_RB ? 0 : 1, C0 = 1 - C1;
    // the V(2,2) sweeps over-relax; their old values come from the stored x (or 0 from the start)
    constexpr bool SOR = MODE >= SW_DOWN0 && EFS_MG_OMEGA != 1.0f;
    if (st->done) return;
    __shared__ float ro[3][PP], bn[3][PP], bs[4][PP];
    __shared__ uint16_t rs[4][PP];
    __shared__ float res[UP ? 1 : 4][UP ? 1 : TX * PTY];  // residuals of the tile (down sweep)
    __shared__ double red[POST_NT / 64];
    const int lane = threadIdx.x + threadIdx.y * TX;
    double acc = 0.0;
    int bx, by, bz;
    if (block_coords(d, bx, by, bz)) {
        const int x0 = 1 + bx * TX, y0 = 1 + by * PTY;
        const int zb = 1 + bz * d.zc, ze = min(zb + d.zc, 1 + l.nz);  // planes [zb, ze)
        const int plane = l.W * l.H, cplane = cl.W * cl.H;
        // fetch lanes: one cell of the halo-2 plane each, cells k = lane and k = lane + POST_NT in
        // row-major order, so that every load of a plane is coalesced along the rows
        int fk[2], fo[2], fc[2], fpar[2];
        bool fv[2];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int k = lane + i * POST_NT;
            fv[i] = k < PP;
            const int kk = fv[i] ? k : 0;
            const int hx = kk % PX, hy = kk / PX;
            const int xu = x0 - 2 + hx, yu = y0 - 2 + hy;  // unclamped (colour)
            const int fx = min(max(xu, 0), l.W - 1), fy = min(max(yu, 0), l.H - 1);
            fk[i] = kk;
            fo[i] = fy * l.W + fx;
            fc[i] = ((fy + 1) >> 1) * cl.W + ((fx + 1) >> 1);
            fpar[i] = (xu + yu) & 1;
        }
        // black lanes: pair (hx, hx + 1), hx = 1 + 2 j, in row hy = 1 .. PY - 2
        const bool blk = lane < NB;
        const int bj = lane % ((PX - 2) / 2), bhy = 1 + min(lane / ((PX - 2) / 2), PY - 3);
        const int bk = bhy * PX + 1 + 2 * bj;
        const int bxu = x0 - 1 + 2 * bj, byu = y0 - 2 + bhy;
        const bool brow_tile = bhy >= 2 && bhy < 2 + PTY;
        // finish lanes: pair (hx, hx + 1), hx = 2 + 2 j, in row hy = 2 .. PY - 3
        const bool fin = lane < NT2;
        const int tj = lane % (TX / 2), thy = 2 + min(lane / (TX / 2), PTY - 1);
        const int tk = thy * PX + 2 + 2 * tj;
        const int txu = x0 + 2 * tj, tyu = y0 - 2 + thy;

        struct Fetched {
            unsigned rc[2];
            float b[2], e[2], xv[2];
        };
        // buffer loads: the lane's cell offset is the vector offset, the plane's the scalar one
        const int nb = l.W * l.H * l.D;
        const __amdgpu_buffer_rsrc_t r_rc = __builtin_amdgcn_make_buffer_rsrc(l.rc, (short)0, nb * 2, kBufDword3);
        const __amdgpu_buffer_rsrc_t r_b = __builtin_amdgcn_make_buffer_rsrc(l.b, (short)0, nb * 4, kBufDword3);
        const __amdgpu_buffer_rsrc_t r_x = __builtin_amdgcn_make_buffer_rsrc((void*)xin, (short)0, nb * 4, kBufDword3);
        const __amdgpu_buffer_rsrc_t r_e =
            __builtin_amdgcn_make_buffer_rsrc(cl.x, (short)0, cl.W * cl.H * cl.D * 4, kBufDword3);
        auto fetch = [&](int p, Fetched& f) {  // loads of plane p (clamped into the array)
            const int pz = min(max(p, 0), l.D - 1);
            const int zo = pz * plane, czo = ((pz + 1) >> 1) * cplane;
#pragma unroll
            for (int i = 0; i < 2; i++) {
                if (i == 1 && !fv[1]) continue;  // every lane has a first cell (POST_NT <= PP)
                f.rc[i] = __builtin_amdgcn_raw_buffer_load_b16(r_rc, (unsigned)fo[i] * 2u, zo * 2, 0);
                f.b[i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r_b, (unsigned)fo[i] * 4u, zo * 4, 0));
                if (USE_E)
                    f.e[i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r_e, (unsigned)fc[i] * 4u, czo * 4, 0));
                if (FROM_X)
                    f.xv[i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r_x, (unsigned)fo[i] * 4u, zo * 4, 0));
            }
        };
        // Ring slots: plane p sits in slot p mod 3 of the three-plane rings (red, black) and p mod 4
        // of the four-plane rings (b, codes); the mod-3 slots are carried along the march.
        auto commit = [&](int p, const Fetched& f, int sp) {  // into the ring slots of plane p (sp: its mod-3 slot)
            float* ro_p = ro[sp];
            float* bs_p = bs[p & 3];
            uint16_t* rs_p = rs[p & 3];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                if (i == 1 && !fv[1]) continue;
                const unsigned rc = f.rc[i];
                float v = FROM_X ? f.xv[i] : relax<SOR>(rc, row_update(rc, 0.f, f.b[i], inv_s), 0.f);
                if (USE_E) v = v + f.e[i];
                // the C0 cells of the red ring are the neighbours of C1 cells; with over-relaxation
                // the C1 cells keep their own old value there too (0 when sweeping from x = 0)
                const bool c0 = ((fpar[i] + p) & 1) == C0;
                ro_p[fk[i]] = ((c0 || (SOR && FROM_X)) && (rc & RC_CONTRIB)) ? v : 0.f;
                bs_p[fk[i]] = f.b[i];
                rs_p[fk[i]] = (uint16_t)rc;
            }
        };
        // residual of a fluid row after the down sweep (b - s (sum - n x)), else 0
        auto residual = [&](unsigned rc, float sum, float x, float b) -> float {
            const float ax = l.s * (sum - (float)((rc >> RC_CNT_SHIFT) & 7u) * x);
            return (rc & RC_CG_ROW) ? b - ax : 0.f;
        };
        // the C1 value of the lane's pair in plane p; sm, s, sp: the mod-3 slots of planes p - 1, p, p + 1
        auto black = [&](int p, int sm, int s, int sp) {
            if (!blk) return;
            const int q = ((bxu + byu + p) & 1) == C1 ? 0 : 1;
            const int k = bk + q, km = k - 1, kp = k + 1, x = bxu + q;
            const float *r0 = ro[sm], *r1 = ro[s], *r2 = ro[sp];
            const unsigned rc = rs[p & 3][k];
            const float b = bs[p & 3][k];
            const float sum = 0.f + r1[km] + r1[kp] + r1[k - PX] + r1[k + PX] + r2[k] + r0[k];
            const bool inside = x >= 1 && x <= l.nx && byu >= 1 && byu <= l.ny && p >= 1 && p <= l.nz;
            const float val = inside ? relax<SOR>(rc, row_update(rc, sum, b, inv_s), r1[k]) : 0.f;
            bn[s][k] = (rc & RC_CONTRIB) ? val : 0.f;
            const bool tile = brow_tile && x >= x0 && x < x0 + TX && p >= zb && p < ze;
            if (UP) {
                if (tile && inside) {
                    xout[p * plane + byu * l.W + x] = val;
                    if (RZ) acc += (double)b * (double)val;
                }
            } else if (tile) {
                res[p & 3][(byu - y0) * TX + (x - x0)] = inside ? residual(rc, sum, val, b) : 0.f;
            }
        };
        // the C0 value (or its residual, SW_DOWN) of the lane's pair in plane q; tm, t, tp: the mod-3
        // slots of planes q - 1, q, q + 1
        // xo: the cell's old value, read from the red ring by the caller before the barrier that
        // precedes the next commit (the ring has three planes, so plane q's slot is refilled then)
        auto finish = [&](int q, int tm, int t, int tp, float xo) {
            if (!fin) return;
            const int h = ((txu + tyu + q) & 1) == C0 ? 0 : 1;
            const int k = tk + h, km = k - 1, kp = k + 1, x = txu + h;
            const bool inside = x <= l.nx && tyu <= l.ny;
            const float *n0 = bn[tm], *n1 = bn[t], *n2 = bn[tp];
            const unsigned rc = rs[q & 3][k];
            const float b = bs[q & 3][k];
            const float sum = 0.f + n1[km] + n1[kp] + n1[k - PX] + n1[k + PX] + n2[k] + n0[k];
            if (MODE == SW_DOWN0) {
                if (inside) xout[q * plane + tyu * l.W + x] = relax<SOR>(rc, row_update(rc, 0.f, b, inv_s), 0.f);
            } else if (UP) {
                if (!inside) return;
                const float val = relax<SOR>(rc, row_update(rc, sum, b, inv_s), xo);
                xout[q * plane + tyu * l.W + x] = val;
                if (RZ) acc += (double)b * (double)val;
            } else {
                res[q & 3][(tyu - y0) * TX + (x - x0)] = inside ? residual(rc, sum, row_update(rc, 0.f, b, inv_s), b) : 0.f;
            }
        };
        // coarse right-hand sides of the pair of planes (q0, q0 + 1), q0 - zb even
        auto restrict_pair = [&](int q0) {
            if (lane >= NC) return;
            const int cx = lane % (TX / 2), cy = lane / (TX / 2);
            const int X = (x0 + 1) / 2 + cx, Y = (y0 + 1) / 2 + cy, Z = (q0 + 1) / 2;
            if (X > cl.nx || Y > cl.ny || Z > cl.nz) return;
            const int ci = lidx(cl, X, Y, Z);
            float a = 0.f;
            if (cl.cls[ci] == CLS_FLUID) {
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
                    if (2 * cx + dx + x0 > l.nx || 2 * cy + dy + y0 > l.ny || q0 + dz > l.nz) continue;
                    a += res[(q0 + dz) & 3][(2 * cy + dy) * TX + 2 * cx + dx];
                }
                a *= 0.125f;
            }
            cl.b[ci] = a;
        };
        // Planes p + 2 and p + 3 are in flight while plane p is processed (two register sets,
        // alternating with the plane).  The barriers order LDS accesses only, so they do not wait
        // for those loads (a full __syncthreads would drain them).
        Fetched fa, fb;
        auto next3 = [](int m) { return m == 2 ? 0 : m + 1; };
        auto body = [&](int p, Fetched& f, int m) {  // m = p mod 3
            const int mp = next3(m), mm = m == 0 ? 2 : m - 1;
            commit(p + 1, f, mp);  // ring slots of planes p - 2 (red) and p - 3 (b, codes): no longer read
            if (p + 3 <= ze + 1) fetch(p + 3, f);
            lds_barrier();
            const int q = p - 1;
            float xo = 0.f;  // the finish cell's old value in plane q (over-relaxation)
            if (SOR && fin && q >= zb) xo = ro[mm][tk + (((txu + tyu + q) & 1) == C0 ? 0 : 1)];
            black(p, mm, m, mp);
            lds_barrier();
            if (q >= zb) {
                finish(q, mp, mm, m, xo);  // planes q - 1 = p - 2 (slot of p + 1), q, q + 1
                if (!UP && (((q - zb) & 1) == 1 || q == ze - 1)) {
                    lds_barrier();
                    restrict_pair(q - ((q - zb) & 1));
                }
            }
        };
        fetch(zb - 2, fa);
        commit(zb - 2, fa, (zb + 1) % 3);
        fetch(zb - 1, fb);
        commit(zb - 1, fb, (zb + 2) % 3);
        fetch(zb, fa);
        fetch(zb + 1, fb);
        int m = (zb - 1) % 3;  // zb >= 1
        for (int p = zb - 1; p <= ze; p += 2) {
            body(p, fa, m);
            m = next3(m);
            if (p + 1 <= ze) body(p + 1, fb, m);
            m = next3(m);
        }
    }
    if (RZ) {  // every launched block writes its partial (0 when idle)
        acc = wave_sum(acc);
        if ((lane & 63) == 0) red[lane >> 6] = acc;
        __syncthreads();
        if (lane == 0) {
            double tot = 0.0;
#pragma unroll
            for (int w = 0; w < POST_NT / 64; w++) tot += red[w];
            p_rz[blockIdx.x] = tot;
        }
    }
}

// Both down sweeps of the V(2,2) cycle in one plane march: the work of k_mg_sweep<SW_DOWN0> and then
// k_mg_sweep<SW_RB>, with the same operations in the same order (bit-identical), without their
// intermediate x buffer and second fetch of b and the row codes.
//   red0   the first sweep's red values: pointwise in b and the row code, since it starts from 0;
//   black0 from the red0 neighbours (its old value 0);
//   red1   from the black0 neighbours (its old value red0);
//   black1 from the red1 neighbours (its old value black0).
// The output x holds red1 and black1.  The tile's black1 needs red1 on a halo of 1, which needs
// black0 on a halo of 2, which needs red0 (and so b and the row codes) on a halo of 3.  Iteration i
// of the march commits red0 of plane i + 1 (from registers loaded two planes earlier), then computes
// black0 of plane i, red1 of plane i - 1 and black1 of plane i - 2, with a barrier between stages;
// each ring holds the three planes its next stage reads.  The row code and b of the tile's black
// cells are copied aside in the red1 stage, since the three-plane ring of b and codes has moved on
// when black1 reads them.  LDS values are masked by RC_CONTRIB as in k_mg_sweep.
constexpr int D3X = TX + 6, D3Y = PTY + 6, D3P = D3X * D3Y;  // halo 3 (fetch, red0)
constexpr int D2X = TX + 4, D2Y = PTY + 4, D2P = D2X * D2Y;  // halo 2 (black0, b, codes)
constexpr int D1X = TX + 2, D1Y = PTY + 2, D1P = D1X * D1Y;  // halo 1 (red1)
static_assert(D3P <= 2 * POST_NT && (D2X / 2) * D2Y <= POST_NT, "two fetch cells and one black0 pair per lane");
__global__ __launch_bounds__(POST_NT) void k_mg_down2(Lvl l, float inv_s, Dom d, const CGState* st, float* __restrict__ xout) {
    constexpr bool SOR = EFS_MG_OMEGA != 1.0f;
    if (st->done) return;
    __shared__ float r0[3][D3P], b0[3][D2P], r1[3][D1P], bs[3][D2P], tb[2][NT2];
    __shared__ uint16_t rs[3][D2P], tr[2][NT2];
    const int lane = threadIdx.x + threadIdx.y * TX;
    int bx, by, bz;
    if (!block_coords(d, bx, by, bz)) return;
    const int x0 = 1 + bx * TX, y0 = 1 + by * PTY;
    const int zb = 1 + bz * d.zc, ze = min(zb + d.zc, 1 + l.nz);  // output planes [zb, ze)
    const int plane = l.W * l.H;
    auto slot3 = [](int p) { return (p + 3) % 3; };  // p >= -2
    // fetch lanes: cells k = lane and lane + POST_NT of the halo-3 plane, row-major
    int fk2[2], fo[2], fpar[2];
    bool fv[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const int k = lane + i * POST_NT;
        fv[i] = k < D3P;
        const int kk = fv[i] ? k : 0;
        const int hx = kk % D3X, hy = kk / D3X;
        const int xu = x0 - 3 + hx, yu = y0 - 3 + hy;
        fo[i] = min(max(yu, 0), l.H - 1) * l.W + min(max(xu, 0), l.W - 1);
        fpar[i] = (xu + yu) & 1;
        fk2[i] = (hx >= 1 && hx < D3X - 1 && hy >= 1 && hy < D3Y - 1) ? (hy - 1) * D2X + (hx - 1) : -1;
    }
    const int nb = l.W * l.H * l.D;
    const __amdgpu_buffer_rsrc_t r_rc = __builtin_amdgcn_make_buffer_rsrc(l.rc, (short)0, nb * 2, kBufDword3);
    const __amdgpu_buffer_rsrc_t r_b = __builtin_amdgcn_make_buffer_rsrc(l.b, (short)0, nb * 4, kBufDword3);
    struct Fetched {
        unsigned rc[2];
        float b[2];
    };
    auto fetch = [&](int p, Fetched& f) {
        const int zo = min(max(p, 0), l.D - 1) * plane;
#pragma unroll
        for (int i = 0; i < 2; i++) {
            if (i == 1 && !fv[1]) continue;
            f.rc[i] = __builtin_amdgcn_raw_buffer_load_b16(r_rc, (unsigned)fo[i] * 2u, zo * 2, 0);
            f.b[i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r_b, (unsigned)fo[i] * 4u, zo * 4, 0));
        }
    };
    auto commit = [&](int p, const Fetched& f) {  // red0 of plane p; b and codes of its halo-2 part
        const int s = slot3(p);
#pragma unroll
        for (int i = 0; i < 2; i++) {
            if (i == 1 && !fv[1]) continue;
            const unsigned rc = f.rc[i];
            const float v = relax<SOR>(rc, row_update(rc, 0.f, f.b[i], inv_s), 0.f);
            const bool red = ((fpar[i] + p) & 1) == 0;
            r0[s][lane + i * POST_NT] = (red && (rc & RC_CONTRIB)) ? v : 0.f;
            if (fk2[i] >= 0) {
                bs[s][fk2[i]] = f.b[i];
                rs[s][fk2[i]] = (uint16_t)rc;
            }
        }
    };
    auto in_level = [&](int x, int y, int p) { return x >= 1 && x <= l.nx && y >= 1 && y <= l.ny && p >= 1 && p <= l.nz; };
    // black0 of plane p on the halo-2 region: lane pair (2 j, 2 j + 1) of row hy
    const bool l_b0 = lane < (D2X / 2) * D2Y;
    const int b0j = lane % (D2X / 2), b0y = min(lane / (D2X / 2), D2Y - 1);
    auto black0 = [&](int p) {
        if (!l_b0) return;
        const int xu = x0 - 2 + 2 * b0j, yu = y0 - 2 + b0y;
        const int q = ((xu + yu + p) & 1) == 1 ? 0 : 1;
        const int k2 = b0y * D2X + 2 * b0j + q, k3 = (b0y + 1) * D3X + 2 * b0j + q + 1;
        const float *rm = r0[slot3(p - 1)], *rc0 = r0[slot3(p)], *rp = r0[slot3(p + 1)];
        const float sum = 0.f + rc0[k3 - 1] + rc0[k3 + 1] + rc0[k3 - D3X] + rc0[k3 + D3X] + rp[k3] + rm[k3];
        const unsigned rc = rs[slot3(p)][k2];
        const float val = in_level(xu + q, yu, p) ? relax<SOR>(rc, row_update(rc, sum, bs[slot3(p)][k2], inv_s), 0.f) : 0.f;
        b0[slot3(p)][k2] = (rc & RC_CONTRIB) ? val : 0.f;
    };
    // red1 of plane p on the halo-1 region (stored for the tile's cells); the b and code of the
    // tile's black cells of plane p are copied aside for black1
    const bool l_r1 = lane < (D1X / 2) * D1Y;
    const int r1j = lane % (D1X / 2), r1y = min(lane / (D1X / 2), D1Y - 1);
    const bool l_t = lane < NT2;
    const int tj = lane % (TX / 2), ty = lane / (TX / 2);
    auto red1 = [&](int p) {
        if (l_r1) {
            const int xu = x0 - 1 + 2 * r1j, yu = y0 - 1 + r1y;
            const int q = ((xu + yu + p) & 1) == 0 ? 0 : 1;
            const int k1 = r1y * D1X + 2 * r1j + q, k2 = (r1y + 1) * D2X + 2 * r1j + q + 1;
            const int k3 = (r1y + 2) * D3X + 2 * r1j + q + 2;
            const float *bm = b0[slot3(p - 1)], *bc = b0[slot3(p)], *bp = b0[slot3(p + 1)];
            const float sum = 0.f + bc[k2 - 1] + bc[k2 + 1] + bc[k2 - D2X] + bc[k2 + D2X] + bp[k2] + bm[k2];
            const unsigned rc = rs[slot3(p)][k2];
            const int x = xu + q;
            const bool inside = in_level(x, yu, p);
            const float val = inside ? relax<SOR>(rc, row_update(rc, sum, bs[slot3(p)][k2], inv_s), r0[slot3(p)][k3]) : 0.f;
            r1[slot3(p)][k1] = (rc & RC_CONTRIB) ? val : 0.f;
            if (inside && x < x0 + TX && yu >= y0 && yu < y0 + PTY && p >= zb && p < ze) xout[p * plane + yu * l.W + x] = val;
        }
        if (l_t) {
            const int q = ((x0 + 2 * tj + y0 + ty + p) & 1) == 1 ? 0 : 1;
            const int k2 = (ty + 2) * D2X + 2 * tj + q + 2;
            tb[p & 1][lane] = bs[slot3(p)][k2];
            tr[p & 1][lane] = rs[slot3(p)][k2];
        }
    };
    // black1 of plane p on the tile
    auto black1 = [&](int p) {
        if (!l_t) return;
        const int x = x0 + 2 * tj, y = y0 + ty;
        const int q = ((x + y + p) & 1) == 1 ? 0 : 1;
        const int k1 = (ty + 1) * D1X + 2 * tj + q + 1, k2 = (ty + 2) * D2X + 2 * tj + q + 2;
        const float *rm = r1[slot3(p - 1)], *rc1 = r1[slot3(p)], *rp = r1[slot3(p + 1)];
        const float sum = 0.f + rc1[k1 - 1] + rc1[k1 + 1] + rc1[k1 - D1X] + rc1[k1 + D1X] + rp[k1] + rm[k1];
        const unsigned rc = tr[p & 1][lane];
        const float val = relax<SOR>(rc, row_update(rc, sum, tb[p & 1][lane], inv_s), b0[slot3(p)][k2]);
        if (x + q <= l.nx && y <= l.ny) xout[p * plane + y * l.W + x + q] = val;
    };
    // planes zb - 3 .. ze + 2 are committed; black0 runs on zb - 2 .. ze + 1, red1 on zb - 1 .. ze,
    // black1 on zb .. ze - 1.  Two register sets alternate: plane i + 1 is committed while plane
    // i + 2 is in flight, then plane i + 3 is loaded into the set just committed.
    Fetched fa, fb;
    fetch(zb - 3, fa);
    commit(zb - 3, fa);
    fetch(zb - 2, fb);
    commit(zb - 2, fb);
    fetch(zb - 1, fa);
    fetch(zb, fb);
    auto body = [&](int i, Fetched& f) {
        commit(i + 1, f);
        if (i + 3 <= ze + 2) fetch(i + 3, f);
        lds_barrier();
        black0(i);
        lds_barrier();
        if (i >= zb) red1(i - 1);
        lds_barrier();
        if (i >= zb + 2) black1(i - 2);
    };
    for (int i = zb - 2; i <= ze + 1; i += 2) {
        body(i, fa);
        if (i + 1 <= ze + 1) body(i + 1, fb);
    }
}

// residual b - s A x of level f (x stored, FP32) restricted to level c: b_c = 1/8 of the sum of
// the fluid children's residuals (child order x, y, z), 0 on non-fluid coarse cells.  A block
// takes a 64 x 4 tile of fine cells in the two fine planes of one coarse plane: each thread
// computes the residuals of its two cells with row-coalesced loads, and the tile's 32 x 2 coarse
// cells then add their children from LDS in the child order.  (One thread per coarse cell reading
// its 8 children itself made every load a stride-2 gather.)
__global__ __launch_bounds__(256) void k_mg_restrict_x(Lvl f, const float* __restrict__ xf, Lvl c, const CGState* st) {
    if (st->done) return;
    __shared__ float res[2][MBY][MB];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int fx = 1 + blockIdx.x * MB + tx, fy = 1 + blockIdx.y * MBY + ty, Z = 1 + blockIdx.z;
    const int sy = f.W, sz = f.W * f.H;
#pragma unroll
    for (int dz = 0; dz < 2; dz++) {
        const int fz = 2 * Z - 1 + dz;
        float r = 0.f;
        const int i = lidx(f, fx, fy, fz);
        const unsigned rc = (fx <= f.nx && fy <= f.ny && fz <= f.nz) ? f.rc[i] : 0u;
        if (rc & RC_CG_ROW) {  // other rows add 0 (air waves skip the loads)
            const float v[6] = {xf[i - 1], xf[i + 1], xf[i - sy], xf[i + sy], xf[i + sz], xf[i - sz]};
            const float ax = f.s * (masked_sum6(rc, v) - (float)((rc >> RC_CNT_SHIFT) & 7u) * xf[i]);
            r = f.b[i] - ax;
        }
        res[dz][ty][tx] = r;
    }
    __syncthreads();
    const int t = tx + MB * ty;
    if (t >= (MB / 2) * (MBY / 2)) return;
    const int cx = t % (MB / 2), cy = t / (MB / 2);
    const int X = 1 + blockIdx.x * (MB / 2) + cx, Y = 1 + blockIdx.y * (MBY / 2) + cy;
    if (X > c.nx || Y > c.ny) return;
    const int ci = lidx(c, X, Y, Z);
    float acc = 0.f;
    if (c.cls[ci] == CLS_FLUID) {
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
            if (2 * X - 1 + dx > f.nx || 2 * Y - 1 + dy > f.ny || 2 * Z - 1 + dz > f.nz) continue;
            acc += res[dz][2 * cy + dy][2 * cx + dx];
        }
        acc *= 0.125f;
    }
    c.b[ci] = acc;
}

// coarsest level: symmetric red-black sweeps from x = 0 in one workgroup, on a copy of the level
// (values, right-hand side, row codes, with the border) held in LDS
constexpr int MG_CO_MAX = 1728;  // (10 + 2)^3: interiors up to 10 cells per axis
__global__ __launch_bounds__(MG_CO_NT) void k_mg_coarsest(Lvl l, float inv_s, int sweeps, const CGState* st) {
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
    // each thread owns at most one interior cell (at most 10^3 <= MG_CO_NT): its index and colour
    // are computed once, outside the sweeps
    const int n = l.nx * l.ny * l.nz;
    const int sy = l.W, sz = l.W * l.H;
    const int t = threadIdx.x;
    const bool own = t < n;
    const int cx = 1 + t % l.nx, cy = 1 + (t / l.nx) % l.ny, cz = 1 + t / (l.nx * l.ny);
    const int ci = own ? lidx(l, cx, cy, cz) : 0, ccol = (cx + cy + cz) & 1;
    for (int sw = 0; sw < 2 * sweeps; sw++) {
        for (int h = 0; h < 2; h++) {
            // forward sweeps red then black, backward sweeps black then red
            const int color = (sw & 1) ? 1 - h : h;
            if (own && ccol == color) {
                const unsigned rc = rs[ci];
                const float v[6] = {xs[ci - 1], xs[ci + 1], xs[ci - sy], xs[ci + sy], xs[ci + sz], xs[ci - sz]};
                xs[ci] = relax<EFS_MG_NU == 2 && EFS_MG_OMEGA != 1.0f>(rc, row_update(rc, masked_sum6(rc, v), bs[ci], inv_s),
                                                                     xs[ci]);
            }
            __syncthreads();
        }
    }
    for (int t = threadIdx.x; t < N; t += MG_CO_NT) l.x[t] = xs[t];
}

// The last two levels of the V-cycle in one workgroup (replaces the five sweep and restriction
// launches of level T and the coarsest solve of level C, which cost more in launch gaps than in
// work at 16^3 and 8^3).  Both levels sit in LDS with their borders: T's x, b and row codes (up
// to 18^3 cells, 58 KB) and C's (up to 12^3).  The arithmetic is that of the separate kernels,
// in the same order, so the result is bit-identical:
//   down: two red-black sweeps from x = 0 (red, then black), over-relaxed as in k_mg_sweep;
//   the residual of T's fluid rows restricted to C (child order x, y, z, times 1/8);
//   C: `sweeps` symmetric sweep pairs from 0 (k_mg_coarsest);
//   up: x += the parent's correction on T's fluid cells, then two sweeps (black, then red).
// Within one workgroup the colour passes run in place between barriers: a pass reads only the
// other colour, which is what the halo-buffered multi-block sweeps read too.  Cells that are not
// fluid rows hold +0, so the neighbour sums need no masks (on coarse levels a row's neighbour bit
// is set exactly for its fluid neighbours, and the others add +0 in the same order).
constexpr int MG_TAIL_MAX = 5832;  // (16 + 2)^3
constexpr int MG_TAIL_PER = 2;     // cells of one colour per thread (interiors up to 16 cells per axis)
__global__ __launch_bounds__(MG_CO_NT) void k_mg_tail(Lvl T, Lvl C, float inv_st, float inv_sc, int sweeps, const CGState* st) {
    if (st->done) return;
    __shared__ float tx[MG_TAIL_MAX], tb[MG_TAIL_MAX];
    __shared__ uint16_t tr[MG_TAIL_MAX];
    __shared__ float cx[MG_CO_MAX], cb[MG_CO_MAX];
    __shared__ uint16_t cr[MG_CO_MAX];
    constexpr bool SOR = EFS_MG_NU == 2 && EFS_MG_OMEGA != 1.0f;
    const int t = threadIdx.x;
    const int NTc = T.W * T.H * T.D, NCc = C.W * C.H * C.D;
    for (int i = t; i < NTc; i += MG_CO_NT) {
        tx[i] = 0.f;
        tb[i] = T.b[i];
        tr[i] = T.rc[i];
    }
    for (int i = t; i < NCc; i += MG_CO_NT) {
        cx[i] = 0.f;
        cb[i] = 0.f;
        cr[i] = C.rc[i];
    }
    // this thread's cells of each colour on T (colour = (x + y + z) & 1; -1: none)
    int own[2][MG_TAIL_PER];
    {
        const int hw = (T.nx + 1) / 2, slots = hw * T.ny * T.nz;
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int j = 0; j < MG_TAIL_PER; j++) {
                const int sl = t + j * MG_CO_NT;
                own[c][j] = -1;
                if (sl < slots) {
                    const int k = sl % hw, r = sl / hw;
                    const int y = 1 + r % T.ny, z = 1 + r / T.ny;
                    const int x = 1 + (((1 + y + z) & 1) != c) + 2 * k;
                    if (x <= T.nx) own[c][j] = lidx(T, x, y, z);
                }
            }
    }
    const int sy = T.W, sz = T.W * T.H;
    __syncthreads();
    auto tpass = [&](int c) {
#pragma unroll
        for (int j = 0; j < MG_TAIL_PER; j++) {
            const int i = own[c][j];
            if (i < 0) continue;
            const unsigned rc = tr[i];
            const float sum = 0.f + tx[i - 1] + tx[i + 1] + tx[i - sy] + tx[i + sy] + tx[i + sz] + tx[i - sz];
            tx[i] = relax<SOR>(rc, row_update(rc, sum, tb[i], inv_st), tx[i]);
        }
        __syncthreads();
    };
    // down: two sweeps, red then black
    for (int sw = 0; sw < EFS_MG_NU; sw++) {
        tpass(0);
        tpass(1);
    }
    // restriction of T's residual to C (k_mg_restrict_x)
    const int cn = C.nx * C.ny * C.nz;
    for (int q = t; q < cn; q += MG_CO_NT) {
        const int X = 1 + q % C.nx, Y = 1 + (q / C.nx) % C.ny, Z = 1 + q / (C.nx * C.ny);
        const int ci = lidx(C, X, Y, Z);
        float acc = 0.f;
        if (cr[ci] & RC_CG_ROW) {
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int fx = 2 * X - 1 + (k & 1), fy = 2 * Y - 1 + ((k >> 1) & 1), fz = 2 * Z - 1 + (k >> 2);
                if (fx > T.nx || fy > T.ny || fz > T.nz) continue;
                const int i = lidx(T, fx, fy, fz);
                const unsigned rc = tr[i];
                float r = 0.f;
                if (rc & RC_CG_ROW) {
                    const float v[6] = {tx[i - 1], tx[i + 1], tx[i - sy], tx[i + sy], tx[i + sz], tx[i - sz]};
                    const float ax = T.s * (masked_sum6(rc, v) - (float)((rc >> RC_CNT_SHIFT) & 7u) * tx[i]);
                    r = tb[i] - ax;
                }
                acc += r;
            }
            acc *= 0.125f;
        }
        cb[ci] = acc;
    }
    __syncthreads();
    // coarsest: symmetric sweeps from 0, one cell per thread (k_mg_coarsest)
    {
        const int csy = C.W, csz = C.W * C.H;
        const bool co = t < cn;
        const int ccx = 1 + t % C.nx, ccy = 1 + (t / C.nx) % C.ny, ccz = 1 + t / (C.nx * C.ny);
        const int ci = co ? lidx(C, ccx, ccy, ccz) : 0, ccol = (ccx + ccy + ccz) & 1;
        for (int sw = 0; sw < 2 * sweeps; sw++) {
            for (int h = 0; h < 2; h++) {
                const int color = (sw & 1) ? 1 - h : h;
                if (co && ccol == color) {
                    const unsigned rc = cr[ci];
                    const float v[6] = {cx[ci - 1], cx[ci + 1], cx[ci - csy], cx[ci + csy], cx[ci + csz], cx[ci - csz]};
                    cx[ci] = relax<SOR>(rc, row_update(rc, masked_sum6(rc, v), cb[ci], inv_sc), cx[ci]);
                }
                __syncthreads();
            }
        }
    }
    // up: the parent's correction on T's fluid cells, then two sweeps, black then red
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int j = 0; j < MG_TAIL_PER; j++) {
            const int i = own[c][j];
            if (i < 0 || !(tr[i] & RC_CONTRIB)) continue;
            const int x = i % T.W, y = (i / T.W) % T.H, z = i / (T.W * T.H);
            tx[i] = tx[i] + cx[lidx(C, (x + 1) >> 1, (y + 1) >> 1, (z + 1) >> 1)];
        }
    __syncthreads();
    for (int sw = 0; sw < EFS_MG_NU; sw++) {
        tpass(1);
        tpass(0);
    }
    for (int i = t; i < NTc; i += MG_CO_NT) T.x[i] = tx[i];
}

inline dim3 cell_grid(const Lvl& l) { return dim3((l.W + MB - 1) / MB, (l.H + MBY - 1) / MBY, l.D); }

void free_mg(MG* m) {
    if (!m) return;
    for (size_t k = 0; k < m->L.size(); k++) {
        if (m->L[k].x2) hipFree(m->L[k].x2);
        if (k == 0) continue;  // level 0 uses the solver's buffers
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
    l0.x2 = nullptr;
    if (EFS_MG_NU == 2) {
        EFS_HIP(hipMalloc(&l0.x2, (size_t)s->N * 4));
        EFS_HIP(hipMemset(l0.x2, 0, (size_t)s->N * 4));
    }
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
        c.x2 = nullptr;
        if (EFS_MG_NU == 2) {
            EFS_HIP(hipMalloc(&c.x2, n * 4));
            EFS_HIP(hipMemset(c.x2, 0, n * 4));
        }
        m->L.push_back(c);
    }
    const Lvl& cl = m->L.back();
    if ((long long)cl.W * cl.H * cl.D > MG_CO_MAX) {
        efs_set_error("multigrid coarsest level too large for the LDS solve");
        return -1;
    }
    // the tail kernel takes the last two levels when the second-to-last is a coarse level that
    // fits its LDS arrays and thread slots (EFS_MG_TAIL=0 in the environment keeps the separate
    // kernels, which compute the same bits)
    const char* env = std::getenv("EFS_MG_TAIL");
    m->tail = !(env && env[0] == '0');
    const char* env2 = std::getenv("EFS_MG_DOWN2");  // EFS_MG_DOWN2=0: the two separate sweep kernels
    m->down2 = EFS_MG_NU == 2 && !(env2 && env2[0] == '0');
    if (m->L.size() >= 3) {
        const Lvl& t = m->L[m->L.size() - 2];
        const long long hw = (t.nx + 1) / 2;
        m->tail = m->tail && (long long)t.W * t.H * t.D <= MG_TAIL_MAX && hw * t.ny * t.nz <= (long long)MG_TAIL_PER * MG_CO_NT;
    } else {
        m->tail = false;
    }
    return 0;
}

Dom post_dom(const Lvl& l);

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
    // The CG kernels sum the r . z partials of its own launch (nbt blocks).  The fine up sweep can
    // write them when its launch is no larger; the slots past it stay 0.
    const int nbt = launch_blocks(dom_tiled(s, false));
    const int nbp = launch_blocks(post_dom(m->L[0]));
    m->fused_rz = EFS_MG_FUSED_RZ && m->L.size() >= 2 && nbp <= nbt && nbt <= s->max_blocks;
    EFS_HIP(hipMemsetAsync(m->p_rz, 0, (size_t)s->max_blocks * sizeof(double), s->stream));
    return 0;
}

// launch domain of k_mg_sweep on level l: TX x PTY tiles, chunks of EFS_MG_MIN_ZC..32 planes
// (even, so that no coarse cell straddles two chunks)
Dom post_dom(const Lvl& l) {
    Dom d{};
    d.W = l.W; d.H = l.H; d.D = l.D;
    d.x0 = d.y0 = d.z0 = 1;
    d.nx = l.nx; d.ny = l.ny; d.nz = l.nz;
    d.ty = PTY;
    d.tiles_x = (l.nx + TX - 1) / TX;
    d.tiles_y = (l.ny + PTY - 1) / PTY;
    const int txy = d.tiles_x * d.tiles_y;
    d.zc = std::min(l.nz, std::max(EFS_MG_MIN_ZC, std::min(32, l.nz * txy / 1024)) & ~1);
    d.chunks_z = (l.nz + d.zc - 1) / d.zc;
    d.nblocks = txy * d.chunks_z;
    return d;
}

// z = M r: one V-cycle.  With fused_rz the fine up sweep also writes the r . z partials.  The fine
// down and up sweeps are event-timed when kernel timing is on.
int vcycle(efs_sim* s, void*) {
    MG* m = get(s);
    const int nl = (int)m->L.size();
    const CGState* st = s->st;
    const dim3 pb(TX, PTY);
    hipEvent_t ta = nullptr, tb = nullptr;
    const int nsep = m->tail ? nl - 2 : nl - 1;  // levels whose down sweeps run as separate kernels
    for (int k = 0; k < nsep; k++) {
        const Lvl& l = m->L[k];
        const Lvl& c = m->L[k + 1];
        const Dom d = post_dom(l);
        const dim3 grid(launch_blocks(d));
        const float is = 1.f / l.s;
        // the fine level's down bracket is timed from the start of its first kernel to the end of its last
        ta = tb = nullptr;
        if (k == 0) efs_ktimer_pair(s, &ta, &tb);
        if (EFS_MG_NU == 2) {
            if (m->down2) {
                hipExtLaunchKernelGGL(k_mg_down2, grid, pb, 0, s->stream, ta, nullptr, 0, l, is, d, st, l.x);
            } else {
                hipExtLaunchKernelGGL(k_mg_sweep<SW_DOWN0>, grid, pb, 0, s->stream, ta, nullptr, 0, l, c, is, d, st,
                                      (double*)nullptr, (const float*)nullptr, l.x2);
                hipLaunchKernelGGL(k_mg_sweep<SW_RB>, grid, pb, 0, s->stream, l, c, is, d, st, nullptr, l.x2, l.x);
            }
            hipExtLaunchKernelGGL(k_mg_restrict_x, dim3((l.nx + MB - 1) / MB, (l.ny + MBY - 1) / MBY, c.nz), dim3(MB, MBY), 0,
                                  s->stream, nullptr, tb, 0, l, (const float*)l.x, c, st);
        } else {
            hipExtLaunchKernelGGL(k_mg_sweep<SW_DOWN>, grid, pb, 0, s->stream, ta, tb, 0, l, c, is, d, st, (double*)nullptr,
                                  (const float*)nullptr, (float*)nullptr);
        }
        if (k == 0) efs_ktimer_add(s, KT_MG_DOWN, ta, tb);
    }
    if (m->tail)
        hipLaunchKernelGGL(k_mg_tail, dim3(1), dim3(MG_CO_NT), 0, s->stream, m->L[nl - 2], m->L[nl - 1], 1.f / m->L[nl - 2].s,
                           1.f / m->L[nl - 1].s, EFS_MG_COARSE_SWEEPS, st);
    else
        hipLaunchKernelGGL(k_mg_coarsest, dim3(1), dim3(MG_CO_NT), 0, s->stream, m->L[nl - 1], 1.f / m->L[nl - 1].s,
                           EFS_MG_COARSE_SWEEPS, st);
    for (int k = nsep - 1; k >= 0; k--) {
        const Lvl& l = m->L[k];
        const Lvl& c = m->L[k + 1];
        const Dom d = post_dom(l);
        const dim3 grid(launch_blocks(d));
        const float is = 1.f / l.s;
        const bool rz = k == 0 && m->fused_rz;
        double* prz = rz ? m->p_rz : nullptr;
        ta = tb = nullptr;
        if (k == 0) efs_ktimer_pair(s, &ta, &tb);
        if (EFS_MG_NU == 2) {
            hipExtLaunchKernelGGL(k_mg_sweep<SW_BR_E>, grid, pb, 0, s->stream, ta, nullptr, 0, l, c, is, d, st, (double*)nullptr,
                                  (const float*)l.x, l.x2);
            if (rz)
                hipExtLaunchKernelGGL(k_mg_sweep<SW_BR_RZ>, grid, pb, 0, s->stream, nullptr, tb, 0, l, c, is, d, st, prz,
                                      (const float*)l.x2, l.x);
            else
                hipExtLaunchKernelGGL(k_mg_sweep<SW_BR>, grid, pb, 0, s->stream, nullptr, tb, 0, l, c, is, d, st, (double*)nullptr,
                                      (const float*)l.x2, l.x);
        } else {
            if (rz)
                hipExtLaunchKernelGGL(k_mg_sweep<SW_UP_RZ>, grid, pb, 0, s->stream, ta, tb, 0, l, c, is, d, st, prz,
                                      (const float*)nullptr, l.x);
            else
                hipExtLaunchKernelGGL(k_mg_sweep<SW_UP>, grid, pb, 0, s->stream, ta, tb, 0, l, c, is, d, st, (double*)nullptr,
                                      (const float*)nullptr, l.x);
        }
        if (k == 0) efs_ktimer_add(s, KT_OTHER, ta, tb);
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
    if ((rc = efs_cg_defer_x0(s))) return rc;  // x0 is set by the first residual kernel
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
    pc.first_batch = m->last_its > 0 ? m->last_its : EFS_MG_FIRST_BATCH;
    pc.writes_rz = m->fused_rz;
    // A V-cycle normally cuts r.r by an order of magnitude per iteration.  On a fluid region with
    // no air neighbour (a pure-Neumann block, A 1 = 0) its output carries a null-space component,
    // d.Ad collapses and FP32 rounding stalls the residual (a closed box full of liquid: r.r stayed
    // at 0.3 for 2000 iterations where plain CG met 1e-5 in 15); such a solve restarts as plain CG.
    pc.stall_window = 20;
    pc.stall_ratio = 0.5;
    // The recurrence residual stops the iteration; with the true residual enabled and the relative
    // criterion, a solve whose FP64 true residual is still above the tolerance restarts from its
    // iterate (at most twice) while iterations remain under max_steps.  A restart never runs past
    // the cap: once the cap is spent the solve keeps the status of the attempt that met the
    // recurrence criterion, with its true residual in the stats.
    efs_stats local{};
    efs_stats* o = out ? out : &local;
    o->fallback = 0;
    int total = 0, first_rc = 0;
    for (int attempt = 0; attempt < 3; attempt++) {
        if (attempt > 0 && total >= s->P.max_steps) break;
        const size_t kt0 = s->ktimers[KT_OTHER].start.size(), kt1 = s->ktimers[KT_MG_DOWN].start.size();
        rc = efs_cg_run_pc(s, o, false, std::max(1, s->P.max_steps - total), &pc);
        // launches after convergence return at once: keep the timed fine up sweep of each real V-cycle
        efs_trim_timer(s->ktimers[KT_OTHER], kt0 + ((size_t)s->st_host->iter + 1));
        efs_trim_timer(s->ktimers[KT_MG_DOWN], kt1 + ((size_t)s->st_host->iter + 1));
        if (rc < 0) return rc;
        total += o->iterations;
        if (pc.stalled) {  // plain CG from x0 = b with the iterations left
            if ((rc = efs_cg_defer_x0(s))) return rc;
            rc = efs_cg_run(s, o, false, std::max(1, s->P.max_steps - total));
            if (rc < 0) return rc;
            total += o->iterations;
            o->fallback = 1;
            first_rc = rc;
            break;
        }
        if (attempt == 0) {
            m->last_its = o->iterations;
            first_rc = rc;
        }
        const bool verify = s->P.true_residual && s->P.criterion == EFS_CRIT_RELATIVE;
        if (rc != 0 || !verify || o->true_rel_residual <= (double)s->P.solver_threshold) break;
    }
    o->iterations = total;
    if (rc == 1 && first_rc == 0) rc = 0;  // a restart stopped by the cap: the first attempt converged
    return rc;
}
