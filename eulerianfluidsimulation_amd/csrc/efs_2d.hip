// The 2D liquid step on gfx950, following CPUSimulator2D (src/simulation/simulator/
// CPUSimulator2D.cpp) and the 2D solvers (CGSolver.cpp:32-111, JacobiSolver.cpp:18-102).
//
// Layout: the same SoA FP32 fields as 3D with D = 1 (cell (x, y) at y*W + x, Grid2D's order);
// u is the x-face at the cell's low-x side, v the y-face.  Per-cell stages are written in the
// reference's operation order (bit-exact without FMA contraction); rho_sharpen and rho_advect
// use the same parallel restatement as the 3D kernels (every change applied, then traced; the
// sparse advection matrix held as 4 backward entries per row and 4 forward entries per
// column, with fixed-point scatters).
//
// Solvers.  The CG runs as two multi-block kernels per iteration and Jacobi as one kernel per
// pass (k2m_*).  Dots, alpha and beta are FP64, the iterate is FP64, the operator FP32 -- the
// same mixed precision as the 3D CG.  (A whole solve in one 1024-thread workgroup, which avoids
// the launches, measured slower at every size from 130^2 -- 21.6 against 12.6 us per CG
// iteration -- and grows with the grid on its one CU: 1.59 ms against 21 us per iteration at
// 1026^2.)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "efs_internal.h"

namespace {

constexpr int BX = 64, BY = 4;

inline dim3 grid2_interior(const efs_sim* s) { return dim3((s->W - 2 + BX - 1) / BX, (s->H - 2 + BY - 1) / BY); }
inline dim3 grid2_full(const efs_sim* s) { return dim3((s->W + BX - 1) / BX, (s->H + BY - 1) / BY); }
inline int blocks1d(int n) { return (n + 255) / 256; }

#define INTERIOR2(g)                                                                        \
    const int x = 1 + blockIdx.x * BX + threadIdx.x, y = 1 + blockIdx.y * BY + threadIdx.y; \
    if (x > (g).W - 2 || y > (g).H - 2) return;                                             \
    const int c = y * (g).W + x

#define FULL2(g)                                                                    \
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y; \
    if (x >= (g).W || y >= (g).H) return;                                           \
    const int c = y * (g).W + x

__device__ inline float fmaxr(float a, float b) { return (a < b) ? b : a; }  // std::max
__device__ inline float fminr(float a, float b) { return (b < a) ? b : a; }  // std::min
__device__ inline double dsq(float v) { return (double)v * (double)v; }     // std::pow(float, 2)
__device__ inline double dmax(double a, double b) { return (a < b) ? b : a; }

// ScalField2D::interpolate (ScalField2D.cpp:20-69): position shifted by +1 first (:21), cell
// values at i + 0.5.
__device__ inline float sinterp2(const DevGrid& g, const float* __restrict__ f, float px, float py) {
    px = px + 1.f;
    py = py + 1.f;
    const int x = (int)px, y = (int)py;
    int ix1 = x, ix2 = x + 1;
    if (px - (float)x < 0.5f) { ix1 = x - 1; ix2 = x; }
    int iy1 = y, iy2 = y + 1;
    if (py - (float)y < 0.5f) { iy1 = y - 1; iy2 = y; }
    ix1 = clampv(ix1, 0, g.W - 1); ix2 = clampv(ix2, 0, g.W - 1);
    iy1 = clampv(iy1, 0, g.H - 1); iy2 = clampv(iy2, 0, g.H - 1);
    const float v0 = f[iy1 * g.W + ix1], v1 = f[iy2 * g.W + ix1];
    const float v2 = f[iy1 * g.W + ix2], v3 = f[iy2 * g.W + ix2];
    const float x1 = (float)ix1 + 0.5f, x2 = (float)ix2 + 0.5f;
    const float y1 = (float)iy1 + 0.5f, y2 = (float)iy2 + 0.5f;
    // the clamped indices of an axis differ by 0 or 1 and i + 0.5f is exact, so each spacing is
    // exactly 0 or 1: the reference's divisions by it are divisions by 1.0f and are left out
    const float dx2x1 = x2 - x1;
    float fxy1, fxy2;
    if (dx2x1 < 1e-9f) {
        fxy1 = v1;
        fxy2 = v3;
    } else {
        fxy1 = (x2 - px) * v0 + (px - x1) * v2;
        fxy2 = (x2 - px) * v1 + (px - x1) * v3;
    }
    const float dy2y1 = y2 - y1;
    if (dy2y1 < 1e-9f) return fxy1;
    return (y2 - py) * fxy1 + (py - y1) * fxy2;
}

// VecField2D::bilinear_interpolation / interpolate (VecField2D.cpp:21-64).  Low indices are
// clamped to [0, n-2]; finite traced positions never reach the clamp.
__device__ inline float bilin2(const DevGrid& g, const float* __restrict__ f, float px, float py, float ox, float oy) {
    const int lx = clampv((int)(floorf(px - ox) + 1.0e-10f), 0, g.W - 2);
    const int ly = clampv((int)(floorf(py - oy) + 1.0e-10f), 0, g.H - 2);
    const float x0 = (float)lx + ox, y0 = (float)ly + oy;
    const float tx = px - x0, ty = py - y0;
    const int b = ly * g.W + lx;
    const float v00 = f[b], v01 = f[b + g.W], v10 = f[b + 1], v11 = f[b + 1 + g.W];
    const float vLow = v00 * (1.0f - ty) + v01 * ty;
    const float vHigh = v10 * (1.0f - ty) + v11 * ty;
    return vLow * (1.0f - tx) + vHigh * tx;
}
__device__ inline float2 vinterp2(const DevGrid& g, float px, float py) {
    return make_float2(bilin2(g, g.u, px, py, -1.0f, -0.5f), bilin2(g, g.v, px, py, -0.5f, -1.0f));
}

// Grid2D::traceParticle (Grid2D.cpp:84-136)
__device__ inline bool inWall2(const DevGrid& g, float px, float py) {
    const float Wf = (float)g.W, Hf = (float)g.H;
    if (px < 1.f || px > Wf - 2.f || py < 1.f || py > Hf - 2.f) return true;
    const int ix = clampv((int)px, 0, g.W - 1), iy = clampv((int)py, 0, g.H - 1);
    return (g.bf[iy * g.W + ix] & BF_GT05) != 0;
}
__device__ inline float2 trace2(const DevGrid& g, float sx, float sy, float dt, int substeps) {
    if (inWall2(g, sx, sy)) return make_float2(sx, sy);
    const float Wf = (float)g.W, Hf = (float)g.H;
    const float subDt = dt / (float)substeps;
    const float half = subDt * 0.5f;
    float px = sx, py = sy;
    for (int s = 0; s < substeps; s++) {
        const float2 k1 = vinterp2(g, px - 0.5f, py - 0.5f);
        const float mx = clampv(px + k1.x * half, 1.f, Wf - 2.f);
        const float my = clampv(py + k1.y * half, 1.f, Hf - 2.f);
        const float2 k2 = vinterp2(g, mx - 0.5f, my - 0.5f);
        const float nx = px + k2.x * subDt, ny = py + k2.y * subDt;
        if (inWall2(g, nx, ny)) {
            float ax = px, ay = py, bx = nx, by = ny;
            for (int b = 0; b < 8; b++) {
                const float hx = (ax + bx) * 0.5f, hy = (ay + by) * 0.5f;
                if (inWall2(g, hx, hy)) { bx = hx; by = hy; }
                else { ax = hx; ay = hy; }
            }
            return make_float2(ax, ay);
        }
        px = nx;
        py = ny;
    }
    return make_float2(px, py);
}

// Corner weights of a traced position (CPUSimulator2D.cpp:140-165): corners 00, 10, 01, 11.
__device__ inline float cornerWeights2(const DevGrid& g, float px, float py, int& p0, float m[4]) {
    int p0x = (int)floorf(px), p0y = (int)floorf(py);
    p0x = clampv(p0x, 0, g.W - 2);
    p0y = clampv(p0y, 0, g.H - 2);
    const float dx0 = 1 - (px - p0x), dy0 = 1 - (py - p0y);
    const float dx1 = 1 - (p0x + 1 - px), dy1 = 1 - (p0y + 1 - py);
    p0 = p0y * g.W + p0x;
    const int off[4] = {0, 1, g.W, g.W + 1};
    const float wx[2] = {dx0, dx1}, wy[2] = {dy0, dy1};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        m[k] = 0.f;
        if (!(g.bf[p0 + off[k]] & BF_EQ1)) m[k] = wx[k & 1] * wy[k >> 1];
    }
    return m[0] + m[1] + m[2] + m[3];
}
__device__ inline int cornerOffset2(const DevGrid& g, int k) { return (k & 1) + (k >> 1) * g.W; }

// ---- vel_decay (CPUSimulator2D.cpp:65-70) ---------------------------------------------------
__global__ __launch_bounds__(256) void k2_decay(DevGrid g, float decay) {
    INTERIOR2(g);
    if (g.rho[c] < 0.5f) {
        g.u[c] = g.u[c] * decay;
        g.v[c] = g.v[c] * decay;
    }
}

// ---- vel_extrapolate (CPUSimulator2D.cpp:72-120) --------------------------------------------
__global__ __launch_bounds__(256) void k2_extrap_init(DevGrid g, uint8_t* valid) {
    INTERIOR2(g);
    valid[c] = (g.rho[c] >= 0.5f && !(g.bf[c] & BF_EQ1)) ? 1 : 0;
}
__global__ __launch_bounds__(256) void k2_extrap_pass(DevGrid g, const float* __restrict__ su, const float* __restrict__ sv,
                                                      const uint8_t* __restrict__ sval, float* __restrict__ du,
                                                      float* __restrict__ dv, uint8_t* __restrict__ dval) {
    INTERIOR2(g);
    if ((g.bf[c] & BF_EQ1) || sval[c]) {
        du[c] = su[c];
        dv[c] = sv[c];
        dval[c] = (g.bf[c] & BF_EQ1) ? 0 : 1;
        return;
    }
    const int nb[4] = {c - 1, c + 1, c - g.W, c + g.W};
    float ax = 0.f, ay = 0.f, n = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (sval[nb[k]]) {
            ax = ax + su[nb[k]];
            ay = ay + sv[nb[k]];
            n++;
        }
    }
    if (n > 0) {
        du[c] = ax / n;
        dv[c] = ay / n;
        dval[c] = 1;
    } else {
        du[c] = su[c];
        dv[c] = sv[c];
        dval[c] = 0;
    }
}

// ---- rho_sharpen (CPUSimulator2D.cpp:320-402, 526-595) --------------------------------------
__global__ __launch_bounds__(256) void k2_sharpen_delta(DevGrid g, float delT, float* __restrict__ del) {
    INTERIOR2(g);
    const float r = g.rho[c];
    const float rxm = g.rho[c - 1], rxp = g.rho[c + 1], rym = g.rho[c - g.W], ryp = g.rho[c + g.W];
    const float dpx = -(r - rxm) * delT, dmx = -(rxp - r) * delT;
    const float dpy = -(r - rym) * delT, dmy = -(ryp - r) * delT;
    const float max_inc = (float)sqrt(dmax(dsq(fmaxr(dpx, 0.f)), dsq(fminr(dmx, 0.f))) +
                                      dmax(dsq(fmaxr(dpy, 0.f)), dsq(fminr(dmy, 0.f))));
    const float max_dec = (float)sqrt(dmax(dsq(fminr(dpx, 0.f)), dsq(fmaxr(dmx, 0.f))) +
                                      dmax(dsq(fminr(dpy, 0.f)), dsq(fmaxr(dmy, 0.f))));
    float mx = -INFINITY;
    mx = fmaxr(mx, fabsf(r - rxm));
    mx = fmaxr(mx, fabsf(r - rxp));
    mx = fmaxr(mx, fabsf(r - rym));
    mx = fmaxr(mx, fabsf(r - ryp));
    const float a = r - 0.5f;
    const float cube = (float)((double)a * (double)a * (double)a);  // std::pow(float, 3.f)
    const float w = cube * (1 - fminr(1.f, mx / 0.2f));
    float dr = (w >= 0) ? w * max_inc : w * max_dec;
    if (r + dr < 0 || (double)r < 1e-5) dr = -r;
    else if (r > 0.5f) dr = 0;
    del[c] = dr;
}
__global__ __launch_bounds__(256) void k2_sharpen_apply(DevGrid g, const float* __restrict__ del) {
    INTERIOR2(g);
    const float dr = del[c];
    if (fabsf(dr) < 1e-6f) return;
    g.rho[c] = fmaxr(g.rho[c] + dr, 0.f);
}
__device__ inline float2 trace_along_field2(const DevGrid& g, float sx, float sy, float Dm, int substeps) {
    const float stepSize = Dm / (float)substeps;
    float px = sx, py = sy;
    for (int s = 0; s < substeps; s++) {
        if (sinterp2(g, g.rho, px, py) >= 0.5f) break;
        const int ix = clampv((int)px, 1, g.W - 2), iy = clampv((int)py, 1, g.H - 2);
        const int q = iy * g.W + ix;
        const float gx = g.rho[q + 1] - g.rho[q - 1];
        const float gy = g.rho[q + g.W] - g.rho[q - g.W];
        const float len = sqrtf(gx * gx + gy * gy);
        if (len < 1e-6f) break;
        const float dx = gx / len, dy = gy / len;
        float nx = px + dx * stepSize, ny = py + dy * stepSize;
        nx = clampv(nx, 1.f, (float)g.W - 2.f);
        ny = clampv(ny, 1.f, (float)g.H - 2.f);
        if (g.bf[(int)ny * g.W + (int)nx] & BF_GT05) break;
        px = nx;
        py = ny;
    }
    return make_float2(px, py);
}
__global__ __launch_bounds__(256) void k2_sharpen_scatter(DevGrid g, const float* __restrict__ del, float rhoD,
                                                          unsigned long long* acc) {
    INTERIOR2(g);
    const float dr = del[c];
    if (fabsf(dr) < 1e-6f) return;
    const float2 p = trace_along_field2(g, (float)x, (float)y, rhoD, 10);
    const float value = -dr;
    const int x0 = clampv((int)p.x, 0, g.W - 1), y0 = clampv((int)p.y, 0, g.H - 1);
    const int x1 = clampv((int)p.x + 1, 0, g.W - 1), y1 = clampv((int)p.y + 1, 0, g.H - 1);
    const float fx = p.x - (float)(int)p.x, fy = p.y - (float)(int)p.y;
    float w[4] = {(1.f - fx) * (1.f - fy), fx * (1.f - fy), (1.f - fx) * fy, fx * fy};
    const int q[4] = {y0 * g.W + x0, y0 * g.W + x1, y1 * g.W + x0, y1 * g.W + x1};
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (g.bf[q[k]] & BF_GT05) w[k] = 0.f;
    const float total = w[0] + w[1] + w[2] + w[3];
    if (total < 1e-6f) return;
    const float inv = 1.f / total;
#pragma unroll
    for (int k = 0; k < 4; k++) fix_add(acc, q[k], value * w[k] * inv);
}
__global__ void k2_add_fix(int n, float* __restrict__ f, unsigned long long* __restrict__ acc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long a = acc[i];
    if (a != 0ull) {
        f[i] = (float)((double)f[i] + from_fix(a));
        acc[i] = 0ull;
    }
}

// ---- rho_advect (CPUSimulator2D.cpp:122-318) ------------------------------------------------
// bw[k * N + i] = A(i, p0_i + corner k); fw[k * N + j] = A(q0_j + corner k, j).
__global__ __launch_bounds__(256) void k2_adv_bwd(DevGrid g, float dt, float* __restrict__ gadv, int* __restrict__ bp0,
                                                  float* __restrict__ bw, unsigned long long* beta_acc) {
    INTERIOR2(g);
    const float2 prev = trace2(g, (float)x, (float)y, -dt, 4);
    const float ga = (g.bf[c] & BF_EQ1) ? 0.f : sinterp2(g, g.gamma, prev.x, prev.y);
    gadv[c] = ga;
    float m[4];
    int p0;
    const float sum = cornerWeights2(g, prev.x, prev.y, p0, m);
    if (sum == 0) {
        bp0[c] = -1;
        return;
    }
    bp0[c] = p0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        float a = m[k] / sum;
        if (ga > 1.e-5f) a /= ga;
        bw[(size_t)k * g.N + c] = a;
        fix_add(beta_acc, p0 + cornerOffset2(g, k), a);
    }
}
__global__ __launch_bounds__(256) void k2_adv_fwd(DevGrid g, float dt, const unsigned long long* __restrict__ beta_acc,
                                                  int* __restrict__ fq0, float* __restrict__ fw, unsigned long long* gamma_acc) {
    INTERIOR2(g);
    const float bj = (float)from_fix(beta_acc[c]);
    if (bj >= 1) {
        fq0[c] = -1;
        return;
    }
    const float2 next = trace2(g, (float)x, (float)y, +dt, 4);
    float m[4];
    int q0;
    const float sum = cornerWeights2(g, next.x, next.y, q0, m);
    if (sum == 0) {
        fq0[c] = -1;
        return;
    }
    fq0[c] = q0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float a = (1 - bj) * (m[k] / sum);
        fw[(size_t)k * g.N + c] = a;
        fix_add(gamma_acc, q0 + cornerOffset2(g, k), a);
    }
}
// gamma_i = own backward row + scattered forward entries (all cells)
__global__ __launch_bounds__(256) void k2_adv_rowsum(DevGrid g, const int* __restrict__ bp0, const float* __restrict__ bw,
                                                     unsigned long long* __restrict__ acc, float* __restrict__ out) {
    FULL2(g);
    long long s = (long long)acc[c];
    acc[c] = 0ull;
    if (bp0[c] >= 0) {
#pragma unroll
        for (int k = 0; k < 4; k++) s += to_fix(bw[(size_t)k * g.N + c]);
    }
    out[c] = (float)((double)s * FIX_INV);
}
__global__ __launch_bounds__(256) void k2_adv_scale_rows(DevGrid g, const float* __restrict__ g6, const int* __restrict__ bp0,
                                                         float* __restrict__ bw, const int* __restrict__ fq0,
                                                         float* __restrict__ fw, unsigned long long* beta_acc) {
    INTERIOR2(g);
    const int p0 = bp0[c];
    if (p0 >= 0) {
        const float gi = g6[c];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            float a = bw[(size_t)k * g.N + c];
            if (gi > 1.e-5f) a /= gi;
            bw[(size_t)k * g.N + c] = a;
            fix_add(beta_acc, p0 + cornerOffset2(g, k), a);
        }
    }
    const int q0 = fq0[c];
    if (q0 >= 0) {
        long long own = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float gi = g6[q0 + cornerOffset2(g, k)];
            float a = fw[(size_t)k * g.N + c];
            if (gi > 1.e-5f) a /= gi;
            fw[(size_t)k * g.N + c] = a;
            own += to_fix(a);
        }
        if (own != 0) atomicAdd(beta_acc + c, (unsigned long long)own);
    }
}
__global__ void k2_fix_to_field(int n, unsigned long long* __restrict__ acc, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = (float)from_fix(acc[i]);
    acc[i] = 0ull;
}
__global__ __launch_bounds__(256) void k2_adv_scale_cols(DevGrid g, const int* __restrict__ bp0, float* __restrict__ bw,
                                                         const int* __restrict__ fq0, float* __restrict__ fw,
                                                         unsigned long long* gamma_acc, unsigned long long* rho_acc) {
    INTERIOR2(g);
    const int p0 = bp0[c];
    if (p0 >= 0) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float bj = g.beta[p0 + cornerOffset2(g, k)];
            float a = bw[(size_t)k * g.N + c];
            if (bj > 1.e-5f) a /= bj;
            bw[(size_t)k * g.N + c] = a;
        }
    }
    const int q0 = fq0[c];
    if (q0 >= 0) {
        const float bj = g.beta[c];
        const float rj = g.rho[c];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            float a = fw[(size_t)k * g.N + c];
            if (bj > 1.e-5f) a /= bj;
            const int i = q0 + cornerOffset2(g, k);
            fix_add(gamma_acc, i, a);
            fix_add(rho_acc, i, a * rj);
        }
    }
}
__global__ __launch_bounds__(256) void k2_adv_final(DevGrid g, const int* __restrict__ bp0, const float* __restrict__ bw,
                                                    unsigned long long* __restrict__ gamma_acc,
                                                    unsigned long long* __restrict__ rho_acc, float* __restrict__ rho_out) {
    FULL2(g);
    long long gs = (long long)gamma_acc[c];
    long long rs = (long long)rho_acc[c];
    gamma_acc[c] = 0ull;
    rho_acc[c] = 0ull;
    const int p0 = bp0[c];
    const bool interior = x > 0 && y > 0 && x < g.W - 1 && y < g.H - 1;
    if (interior && p0 >= 0) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float a = bw[(size_t)k * g.N + c];
            gs += to_fix(a);
            rs += to_fix(a * g.rho[p0 + cornerOffset2(g, k)]);
        }
    }
    g.gamma[c] = (float)((double)gs * FIX_INV);
    rho_out[c] = interior ? (float)((double)rs * FIX_INV) : g.rho[c];
}
// Diffusion sweep pair (CPUSimulator2D.cpp:266-315), sequential along each line.
__device__ inline void diffuse_pair2(float& rj, float& gj, unsigned bj, float& ri, float& gi, unsigned bi) {
    if ((bi & BF_EQ1) || (bj & BF_EQ1)) return;
    if (gj < 1e-5f || gi < 1e-5f) return;
    if (gj == gi) return;
    const float avg = (gj + gi) / 2.f;
    // mass moves from the cell of larger gamma (s) to the other (t); written on values, since
    // branches that update rj / ri through their references make the compiler select between
    // the two addresses and keep both in scratch memory
    const bool jg = gj > gi;
    const float rs = jg ? rj : ri, gs = jg ? gj : gi, gt = jg ? gi : gj;
    const float dif = rs * (gs - gt) / (2.f * gs);
    const float nj = jg ? rj - dif : rj + dif;
    const float ni = jg ? ri + dif : ri - dif;
    rj = nj;
    ri = ni;
    gj = avg;
    gi = avg;
}
// AXIS 0: one thread per row y (sweep in x); AXIS 1: one thread per column x (sweep in y)
template <int AXIS>
__global__ __launch_bounds__(64) void k2_diffuse(DevGrid g) {
    const int a = 1 + blockIdx.x * blockDim.x + threadIdx.x;
    int c0, stride, len;
    if (AXIS == 0) {
        if (a > g.H - 2) return;
        c0 = a * g.W + 1; stride = 1; len = g.W;
    } else {
        if (a > g.W - 2) return;
        c0 = g.W + a; stride = g.W; len = g.H;
    }
    int cj = c0;
    float rj = g.rho[cj], gj = g.gamma[cj];
    unsigned bj = g.bf[cj];
    for (int k = 2; k < len - 1; k++) {
        const int ci = cj + stride;
        float ri = g.rho[ci], gi = g.gamma[ci];
        const unsigned bi = g.bf[ci];
        diffuse_pair2(rj, gj, bj, ri, gi, bi);
        g.rho[cj] = rj;
        g.gamma[cj] = gj;
        cj = ci; rj = ri; gj = gi; bj = bi;
    }
    g.rho[cj] = rj;
    g.gamma[cj] = gj;
}

// ---- vel_advect (CPUSimulator2D.cpp:427-448) ------------------------------------------------
__global__ __launch_bounds__(256) void k2_vadv_trace(DevGrid g, float dt, float* __restrict__ tu, float* __restrict__ tv) {
    INTERIOR2(g);
    const float2 prev = trace2(g, (float)x, (float)y, -dt, 4);
    const float2 v = vinterp2(g, prev.x - 0.5f, prev.y - 0.5f);
    tu[c] = v.x;
    tv[c] = v.y;
}
__global__ __launch_bounds__(256) void k2_vadv_faces(DevGrid g, const float* __restrict__ tu, const float* __restrict__ tv) {
    INTERIOR2(g);
    const bool bs = g.bf[c] & BF_EQ1;
    g.u[c] = (x == 1 || bs || (g.bf[c - 1] & BF_EQ1)) ? tu[c] : 0.5f * (tu[c] + tu[c - 1]);
    g.v[c] = (y == 1 || bs || (g.bf[c - g.W] & BF_EQ1)) ? tv[c] : 0.5f * (tv[c] + tv[c - g.W]);
}

// ---- add_force (CPUSimulator2D.cpp:404-425): face-wise ---------------------------------------
__global__ __launch_bounds__(256) void k2_add_force(DevGrid g, float dt, float gdt, int gravity) {
    INTERIOR2(g);
    const bool s0 = g.bf[c] & BF_EQ1;
    const float fx = g.fu ? g.fu[c] : 0.f, fy = g.fv ? g.fv[c] : 0.f;
    if (!s0 || !(g.bf[c - 1] & BF_EQ1)) g.u[c] = g.u[c] + fx * dt;
    else g.u[c] = 0.f;
    if (!s0 || !(g.bf[c - g.W] & BF_EQ1)) {
        float v = g.v[c] + fy * dt;
        if (gravity) v -= gdt;
        g.v[c] = v;
    } else {
        g.v[c] = 0.f;
    }
}

// ---- project (CPUSimulator2D.cpp:450-524, 597-607) --------------------------------------------
// two-sided wall enforcement of the faces stored at cell c (every cell, x > 0 / y > 0 guards)
__device__ inline void walls2_at(const DevGrid& g, int x, int y, int c) {
    const bool s0 = g.bf[c] & BF_EQ1;
    if (s0 || (x > 0 && (g.bf[c - 1] & BF_EQ1))) g.u[c] = 0.f;
    if (s0 || (y > 0 && (g.bf[c - g.W] & BF_EQ1))) g.v[c] = 0.f;
}
__global__ __launch_bounds__(256) void k2_walls(DevGrid g) {
    FULL2(g);
    walls2_at(g, x, y, c);
}
// divergence with the density correction on liquid, non-solid interior cells; reads faces
// already enforced by k2_walls
__global__ __launch_bounds__(256) void k2_divergence(DevGrid g, float liq, float lp, float lm, float eta) {
    INTERIOR2(g);
    const float r = g.rho[c];
    if (r < liq) return;
    if (g.bf[c] & BF_EQ1) return;
    const float dx = g.u[c + 1] - g.u[c];
    const float dy = g.v[c + g.W] - g.v[c];
    float dv = dx + dy;
    if (r > 1.f) dv -= fminr(lm * (r - 1.f), eta);
    if (r < 1.f) dv += fminr(lp * (1.f - r), eta);
    g.div[c] = dv;
}
// u -= grad p on faces between two non-solid cells, 0 otherwise (interior cells)
__global__ __launch_bounds__(256) void k2_gradient(DevGrid g, const double* __restrict__ p64) {
    INTERIOR2(g);
    const bool s0 = g.bf[c] & BF_EQ1;
    if (!s0 && !(g.bf[c - 1] & BF_EQ1)) {
        if (p64) g.u[c] = (float)((double)g.u[c] - (p64[c] - p64[c - 1]));
        else g.u[c] = g.u[c] - (g.p[c] - g.p[c - 1]);
    } else {
        g.u[c] = 0.f;
    }
    if (!s0 && !(g.bf[c - g.W] & BF_EQ1)) {
        if (p64) g.v[c] = (float)((double)g.v[c] - (p64[c] - p64[c - g.W]));
        else g.v[c] = g.v[c] - (g.p[c] - g.p[c - g.W]);
    } else {
        g.v[c] = 0.f;
    }
}

// ---- solvers -------------------------------------------------------------------------------
// row code: neighbours x-, x+, y-, y+ (bits as in 3D); counted = in range, not on the outer
// layer, boundary <= 0.5; contributing = counted and rho > 0.5 (CGSolver.cpp:85-111)
__global__ __launch_bounds__(256) void k2_rowcode(DevGrid g, uint16_t* rc) {
    FULL2(g);
    const float r = g.rho[c];
    const unsigned b = g.bf[c];
    unsigned code = 0, cnt = 0;
    auto nb = [&](int nx, int ny, unsigned bit) {
        if (nx < 0 || nx >= g.W || ny < 0 || ny >= g.H) return;
        if (nx == 0 || nx == g.W - 1 || ny == 0 || ny == g.H - 1) return;
        const int q = ny * g.W + nx;
        if (g.bf[q] & BF_GT05) return;
        cnt++;
        if (g.rho[q] > 0.5f) code |= bit;
    };
    nb(x - 1, y, RC_NB_XM);
    nb(x + 1, y, RC_NB_XP);
    nb(x, y - 1, RC_NB_YM);
    nb(x, y + 1, RC_NB_YP);
    code |= cnt << RC_CNT_SHIFT;
    if (!(r <= 0.5f || (b & BF_EQ1))) code |= RC_CG_ROW;
    if (!(r <= 0.5f || (b & BF_GE05))) code |= RC_JAC_ROW;
    rc[c] = (uint16_t)code;
}

__device__ inline float row2(unsigned rc, const float* __restrict__ v, int c, int W) {
    const float xc = v[c];
    if (!(rc & RC_CG_ROW)) return xc;
    float sum = 0.0f;
    if (rc & RC_NB_XM) sum += v[c - 1];
    if (rc & RC_NB_XP) sum += v[c + 1];
    if (rc & RC_NB_YM) sum += v[c - W];
    if (rc & RC_NB_YP) sum += v[c + W];
    return sum - (float)((rc >> RC_CNT_SHIFT) & 7u) * xc;
}
__device__ inline double row2d(unsigned rc, const double* __restrict__ v, int c, int W) {
    const double xc = v[c];
    if (!(rc & RC_CG_ROW)) return xc;
    double sum = 0.0;
    if (rc & RC_NB_XM) sum += v[c - 1];
    if (rc & RC_NB_XP) sum += v[c + 1];
    if (rc & RC_NB_YM) sum += v[c - W];
    if (rc & RC_NB_YP) sum += v[c + W];
    return sum - (double)((rc >> RC_CNT_SHIFT) & 7u) * xc;
}

struct Cg2Args {
    int N, W;
    int max_it;
    int prev;          // keep the previous FP64 iterate on CG rows
    int relative;      // stop on ||r|| / ||b|| <= thr instead of r.r < thr
    int true_residual;
    double thr;
};

__global__ __launch_bounds__(256) void k2_spmv(int N, int W, const uint16_t* __restrict__ rcode, const float* __restrict__ xv,
                                               float* __restrict__ yv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    yv[i] = row2(rcode[i], xv, i, W);
}

__global__ void k2_fill_u64(int n, unsigned long long* a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0ull;
}

// ---- solvers (CGSolver.cpp:32-111, JacobiSolver.cpp:18-70) ---------------------------------
// The CG runs as two kernels per iteration over a fixed grid of at most MB_MAXB blocks
// (grid-stride loops), and Jacobi as one kernel per pass.  Each CG kernel ends with the per-block
// partial sums of its dot product; every block of the next kernel adds all partials in the same
// order, so all blocks hold the same r.r, d.t, alpha and beta and take the same convergence
// decision without a grid-wide barrier.
constexpr int MB_NT = 256, MB_MAXB = 1024;

#define MB_CELLS(i, N) for (int i = blockIdx.x * MB_NT + threadIdx.x; i < (N); i += gridDim.x * MB_NT)

__device__ inline double mb_wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// this block's total of v into p[blockIdx.x] (waves in order)
__device__ inline void mb_partial(double v, double* __restrict__ p, double* red) {
    v = mb_wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < MB_NT / 64; w++) t += red[w];
        p[blockIdx.x] = t;
    }
}

// the sum of n partials, in a fixed order: the same value in every block
__device__ inline double mb_total(const double* __restrict__ p, int n, double* red) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += MB_NT) s += p[i];
    s = mb_wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < MB_NT / 64; w++) t += red[w];
    __syncthreads();
    return t;
}

// x_0: b (CGSolver.cpp:47), or the previous FP64 iterate on CG rows
__global__ __launch_bounds__(MB_NT) void k2m_init_x(int N, int prev, const uint16_t* __restrict__ rcode,
                                                   const float* __restrict__ b, double* __restrict__ x64) {
    MB_CELLS(i, N) {
        if (!(prev && (rcode[i] & RC_CG_ROW))) x64[i] = (double)b[i];
    }
}

// r = b - A x (FP64 operator on the FP64 iterate, rounded to FP32) with the partials of r.r and,
// for the initial residual, of b.b.  Also the residual replacement of iterations i % 64 == 0.
__global__ __launch_bounds__(MB_NT) void k2m_residual(int N, int W, const uint16_t* __restrict__ rcode,
                                                     const float* __restrict__ b, const double* __restrict__ x64,
                                                     float* __restrict__ r, const CGState* st, double* __restrict__ p_rr,
                                                     double* __restrict__ p_bb) {
    __shared__ double red[MB_NT / 64];
    if (st->done) return;
    double vr = 0.0, vb = 0.0;
    MB_CELLS(i, N) {
        const double bi = (double)b[i];
        const float ri = (float)(bi - row2d(rcode[i], x64, i, W));
        r[i] = ri;
        vr += (double)ri * (double)ri;
        vb += bi * bi;
    }
    mb_partial(vr, p_rr, red);
    if (p_bb) {
        __syncthreads();
        mb_partial(vb, p_bb, red);
    }
}

// b.b and the stopping threshold (relative criterion: thr^2 b.b), once before the iterations
__global__ __launch_bounds__(MB_NT) void k2m_setup(int nb, double thr_in, int relative, const double* __restrict__ p_bb,
                                                  CGState* st) {
    __shared__ double red[MB_NT / 64];
    const double bb = mb_total(p_bb, nb, red);
    if (threadIdx.x == 0) {
        st->bnorm2 = bb;
        st->thr = relative ? thr_in * thr_in * bb : thr_in;
    }
}

// Iteration i, first kernel: r_i . r_i from the partials and the convergence test (which ends
// the solve: later launches return at once); beta_i = rr_i / rr_{i-1} and the direction
// d_i = d_{i-1} beta_i + r_i (d_0 = r_0), recomputed at each cell's neighbours from r and
// d_{i-1} so that no block waits for another; t = A d_i and the partials of d_i . t.
__global__ __launch_bounds__(MB_NT) void k2m_dir_spmv(int N, int W, int nb, int i, const uint16_t* __restrict__ rcode,
                                                     const float* __restrict__ r, const float* __restrict__ dprev,
                                                     float* __restrict__ dnew, float* __restrict__ t, CGState* st,
                                                     const double* __restrict__ p_rr, double* __restrict__ p_dt) {
    __shared__ double red[MB_NT / 64];
    if (st->done) return;
    const double rr = mb_total(p_rr, nb, red);
    const bool conv = rr < st->thr || rr == 0.0;
    if (conv || !isfinite(rr)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->iter = i;
            st->gamma = rr;
            st->converged = conv ? 1 : 0;
            st->done = 1;
        }
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) st->gam[i & 1] = rr;  // other blocks read the other slot
    const double beta = i == 0 ? 0.0 : rr / st->gam[(i - 1) & 1];
    auto dn = [&](int j) -> float { return i == 0 ? r[j] : (float)((double)dprev[j] * beta + (double)r[j]); };
    double vd = 0.0;
    MB_CELLS(c, N) {
        const unsigned rc = rcode[c];
        const float dc = dn(c);
        float tc = dc;
        if (rc & RC_CG_ROW) {
            float sum = 0.0f;
            if (rc & RC_NB_XM) sum += dn(c - 1);
            if (rc & RC_NB_XP) sum += dn(c + 1);
            if (rc & RC_NB_YM) sum += dn(c - W);
            if (rc & RC_NB_YP) sum += dn(c + W);
            tc = sum - (float)((rc >> RC_CNT_SHIFT) & 7u) * dc;
        }
        dnew[c] = dc;
        t[c] = tc;
        vd += (double)dc * (double)tc;
    }
    mb_partial(vd, p_dt, red);
}

// Iteration i, second kernel: alpha = rr_i / d.t; x += alpha d; r -= alpha t with the partials
// of r.r.  On replacement iterations only x is updated: k2m_residual then sets r = b - A x.
__global__ __launch_bounds__(MB_NT) void k2m_update(int N, int nb, int i, int replace, const float* __restrict__ d,
                                                   const float* __restrict__ t, double* __restrict__ x64,
                                                   float* __restrict__ r, const CGState* st, const double* __restrict__ p_dt,
                                                   double* __restrict__ p_rr) {
    __shared__ double red[MB_NT / 64];
    if (st->done) return;
    const double alpha = st->gam[i & 1] / mb_total(p_dt, nb, red);
    double vn = 0.0;
    MB_CELLS(c, N) {
        x64[c] = x64[c] + alpha * (double)d[c];
        if (!replace) {
            const float rn = (float)((double)r[c] - alpha * (double)t[c]);
            r[c] = rn;
            vn += (double)rn * (double)rn;
        }
    }
    if (!replace) mb_partial(vn, p_rr, red);
}

// After `launched` iterations: the convergence test of r_launched; the solve also ends at max_it.
__global__ __launch_bounds__(MB_NT) void k2m_check(int nb, int launched, int max_it, CGState* st,
                                                  const double* __restrict__ p_rr) {
    __shared__ double red[MB_NT / 64];
    if (st->done) return;
    const double rr = mb_total(p_rr, nb, red);
    const bool conv = rr < st->thr || rr == 0.0;
    if (threadIdx.x == 0) {
        st->iter = launched;
        st->gamma = rr;
        st->converged = conv ? 1 : 0;
        if (conv || !isfinite(rr) || launched >= max_it) st->done = 1;
    }
}

// pressure = (float) x and, optionally, the partials of ||b - A x||^2 (FP64)
__global__ __launch_bounds__(MB_NT) void k2m_final(int N, int W, int true_res, const uint16_t* __restrict__ rcode,
                                                  const float* __restrict__ b, const double* __restrict__ x64,
                                                  float* __restrict__ p, double* __restrict__ p_tn) {
    __shared__ double red[MB_NT / 64];
    double tn = 0.0;
    MB_CELLS(c, N) {
        p[c] = (float)x64[c];
        if (true_res) {
            const double e = (double)b[c] - row2d(rcode[c], x64, c, W);
            tn += e * e;
        }
    }
    if (true_res) mb_partial(tn, p_tn, red);
}

__global__ __launch_bounds__(MB_NT) void k2m_true_total(int nb, CGState* st, const double* __restrict__ p_tn) {
    __shared__ double red[MB_NT / 64];
    const double t = mb_total(p_tn, nb, red);
    if (threadIdx.x == 0) st->true_num = t;
}

// One Jacobi pass (JacobiSolver.cpp:18-70) from src into dst: on Jacobi rows (sum of the
// contributing neighbours - b) / nCount, 0 without neighbours; other cells keep their value.
__global__ __launch_bounds__(MB_NT) void k2m_jacobi_pass(int N, int W, const uint16_t* __restrict__ rcode,
                                                        const float* __restrict__ b, const float* __restrict__ src,
                                                        float* __restrict__ dst) {
    MB_CELLS(i, N) {
        const unsigned rc = rcode[i];
        float out = src[i];
        if (rc & RC_JAC_ROW) {
            float sum = 0.0f;
            if (rc & RC_NB_XM) sum += src[i - 1];
            if (rc & RC_NB_XP) sum += src[i + 1];
            if (rc & RC_NB_YM) sum += src[i - W];
            if (rc & RC_NB_YP) sum += src[i + W];
            const unsigned cnt = (rc >> RC_CNT_SHIFT) & 7u;
            out = cnt > 0 ? (sum - b[i]) / (float)cnt : 0.0f;
        }
        dst[i] = out;
    }
}

// ---- stage launchers -----------------------------------------------------------------------
int launch_decay(efs_sim* s) {
    hipLaunchKernelGGL(k2_decay, grid2_interior(s), dim3(BX, BY), 0, s->stream, s->grid(), s->P.vel_decay);
    EFS_HIP(hipGetLastError());
    return 0;
}

int launch_extrapolate(efs_sim* s) {
    const DevGrid g = s->grid();
    const dim3 grd = grid2_interior(s), blk(BX, BY);
    hipLaunchKernelGGL(k2_extrap_init, grd, blk, 0, s->stream, g, s->valA);
    for (int i = 0; i < 3; i++) {
        hipLaunchKernelGGL(k2_extrap_pass, grd, blk, 0, s->stream, g, s->u, s->v, s->valA, s->tu, s->tv, s->valB);
        hipLaunchKernelGGL(k2_extrap_pass, grd, blk, 0, s->stream, g, s->tu, s->tv, s->valB, s->u, s->v, s->valA);
    }
    EFS_HIP(hipGetLastError());
    return 0;
}

int launch_sharpen(efs_sim* s) {
    const DevGrid g = s->grid();
    const dim3 grd = grid2_interior(s), blk(BX, BY);
    hipLaunchKernelGGL(k2_sharpen_delta, grd, blk, 0, s->stream, g, s->P.dt * 3, s->s0);
    hipLaunchKernelGGL(k2_sharpen_apply, grd, blk, 0, s->stream, g, s->s0);
    hipLaunchKernelGGL(k2_sharpen_scatter, grd, blk, 0, s->stream, g, s->s0, s->P.rho_d, s->acc0);
    hipLaunchKernelGGL(k2_add_fix, dim3(blocks1d(s->N)), dim3(256), 0, s->stream, s->N, s->rho, s->acc0);
    EFS_HIP(hipGetLastError());
    return 0;
}

int launch_rho_advect(efs_sim* s) {
    const DevGrid g = s->grid();
    const dim3 gi = grid2_interior(s), gf = grid2_full(s), blk(BX, BY);
    const dim3 g1(blocks1d(s->N)), b1(256);
    const float dt = s->P.dt;
    hipLaunchKernelGGL(k2_adv_bwd, gi, blk, 0, s->stream, g, dt, s->s1, s->bwd_p0, s->bwd_w, s->acc0);
    hipLaunchKernelGGL(k2_adv_fwd, gi, blk, 0, s->stream, g, dt, s->acc0, s->fwd_p0, s->fwd_w, s->acc1);
    hipLaunchKernelGGL(k2_fill_u64, g1, b1, 0, s->stream, s->N, s->acc0);
    hipLaunchKernelGGL(k2_adv_rowsum, gf, blk, 0, s->stream, g, s->bwd_p0, s->bwd_w, s->acc1, s->s1);
    hipLaunchKernelGGL(k2_adv_scale_rows, gi, blk, 0, s->stream, g, s->s1, s->bwd_p0, s->bwd_w, s->fwd_p0, s->fwd_w, s->acc0);
    hipLaunchKernelGGL(k2_fix_to_field, g1, b1, 0, s->stream, s->N, s->acc0, s->beta);
    hipLaunchKernelGGL(k2_adv_scale_cols, gi, blk, 0, s->stream, g, s->bwd_p0, s->bwd_w, s->fwd_p0, s->fwd_w, s->acc1, s->acc2);
    hipLaunchKernelGGL(k2_adv_final, gf, blk, 0, s->stream, g, s->bwd_p0, s->bwd_w, s->acc1, s->acc2, s->s0);
    EFS_HIP(hipGetLastError());
    std::swap(s->rho, s->s0);
    // 7 iterations of x then y sweeps (CPUSimulator2D.cpp:261)
    const DevGrid g2 = s->grid();
    for (int it = 0; it < 7; it++) {
        hipLaunchKernelGGL(k2_diffuse<0>, dim3((s->H - 2 + 63) / 64), dim3(64), 0, s->stream, g2);
        hipLaunchKernelGGL(k2_diffuse<1>, dim3((s->W - 2 + 63) / 64), dim3(64), 0, s->stream, g2);
    }
    EFS_HIP(hipGetLastError());
    return 0;
}

int launch_vel_advect(efs_sim* s) {
    const DevGrid g = s->grid();
    const dim3 grd = grid2_interior(s), blk(BX, BY);
    hipLaunchKernelGGL(k2_vadv_trace, grd, blk, 0, s->stream, g, s->P.dt, s->tu, s->tv);
    hipLaunchKernelGGL(k2_vadv_faces, grd, blk, 0, s->stream, g, s->tu, s->tv);
    EFS_HIP(hipGetLastError());
    return 0;
}

int launch_add_force(efs_sim* s) {
    const float gdt = s->P.gravity_force * s->P.dt;
    hipLaunchKernelGGL(k2_add_force, grid2_interior(s), dim3(BX, BY), 0, s->stream, s->grid(), s->P.dt, gdt, s->P.gravity_enabled);
    EFS_HIP(hipGetLastError());
    return 0;
}

int launch_divergence(efs_sim* s) {
    const DevGrid g = s->grid();
    hipLaunchKernelGGL(k2_divergence, grid2_interior(s), dim3(BX, BY), 0, s->stream, g, s->P.liquid_threshold,
                       s->P.lambda_plus, s->P.lambda_minus, s->P.eta);
    EFS_HIP(hipGetLastError());
    return 0;
}

int launch_walls(efs_sim* s) {
    hipLaunchKernelGGL(k2_walls, grid2_full(s), dim3(BX, BY), 0, s->stream, s->grid());
    EFS_HIP(hipGetLastError());
    return 0;
}

int launch_rowcode(efs_sim* s) {
    hipLaunchKernelGGL(k2_rowcode, grid2_full(s), dim3(BX, BY), 0, s->stream, s->grid(), s->rowcode);
    EFS_HIP(hipGetLastError());
    return 0;
}

}  // namespace

int efs2_walls_divergence(efs_sim* s) {
    EFS_HIP(hipMemsetAsync(s->div, 0, (size_t)s->N * 4, s->stream));  // divergence.clear()
    int rc = launch_walls(s);
    if (rc) return rc;
    return launch_divergence(s);
}

// The CG solve (CGSolverCPU2D::solve, CGSolver.cpp:32-83) from x0 = b or the previous iterate.
// The host launches the iterations in batches of 32 and reads the state between batches;
// launches past convergence return at once.
int cg2_solve(efs_sim* s, const Cg2Args& a) {
    const int nb = std::min(MB_MAXB, (a.N + MB_NT - 1) / MB_NT);
    if (nb > s->max_blocks) {
        efs_set_error("2D solve: launch exceeds the partial-sum buffers");
        return -1;
    }
    CGState init{};
    std::memcpy(s->st_host, &init, sizeof(CGState));
    EFS_HIP(hipMemcpyAsync(s->st, s->st_host, sizeof(CGState), hipMemcpyHostToDevice, s->stream));
    const dim3 grd(nb), blk(MB_NT);
    hipLaunchKernelGGL(k2m_init_x, grd, blk, 0, s->stream, a.N, a.prev, s->rowcode, s->div, s->cg_x64);
    hipLaunchKernelGGL(k2m_residual, grd, blk, 0, s->stream, a.N, a.W, s->rowcode, s->div, s->cg_x64, s->cg_r, s->st, s->p_rr,
                       s->p_bb);
    hipLaunchKernelGGL(k2m_setup, dim3(1), blk, 0, s->stream, nb, a.thr, a.relative, s->p_bb, s->st);
    constexpr int BATCH = 32;
    int launched = 0;
    while (true) {
        const int n = std::min(BATCH, a.max_it - launched);
        for (int k = 0; k < n; k++) {
            const int i = launched + k;
            const float* dprev = (i & 1) ? s->cg_d0 : s->cg_d1;
            float* dnew = (i & 1) ? s->cg_d1 : s->cg_d0;
            hipLaunchKernelGGL(k2m_dir_spmv, grd, blk, 0, s->stream, a.N, a.W, nb, i, s->rowcode, s->cg_r, dprev, dnew, s->cg_x1,
                               s->st, s->p_rr, s->p_dad);
            const int replace = (i & 0x3F) == 0;  // residual replacement every 64 iterations (CGSolver.cpp:62-69)
            hipLaunchKernelGGL(k2m_update, grd, blk, 0, s->stream, a.N, nb, i, replace, (const float*)dnew, (const float*)s->cg_x1,
                               s->cg_x64, s->cg_r, s->st, s->p_dad, s->p_rr);
            if (replace)
                hipLaunchKernelGGL(k2m_residual, grd, blk, 0, s->stream, a.N, a.W, s->rowcode, s->div, s->cg_x64, s->cg_r, s->st,
                                   s->p_rr, (double*)nullptr);
        }
        launched += n;
        hipLaunchKernelGGL(k2m_check, dim3(1), blk, 0, s->stream, nb, launched, a.max_it, s->st, s->p_rr);
        EFS_HIP(hipGetLastError());
        EFS_HIP(hipMemcpyAsync(s->st_host, s->st, sizeof(CGState), hipMemcpyDeviceToHost, s->stream));
        EFS_HIP(hipStreamSynchronize(s->stream));
        if (s->st_host->done) break;
    }
    hipLaunchKernelGGL(k2m_final, grd, blk, 0, s->stream, a.N, a.W, a.true_residual, s->rowcode, s->div, s->cg_x64, s->p,
                       s->p_dad);
    if (a.true_residual) hipLaunchKernelGGL(k2m_true_total, dim3(1), blk, 0, s->stream, nb, s->st, s->p_dad);
    EFS_HIP(hipGetLastError());
    EFS_HIP(hipMemcpyAsync(s->st_host, s->st, sizeof(CGState), hipMemcpyDeviceToHost, s->stream));
    EFS_HIP(hipStreamSynchronize(s->stream));
    return 0;
}

// Pressure solve on the current divergence (p cleared by the caller, as project() does).
int efs2_solve(efs_sim* s, efs_stats* out) {
    int rc = launch_rowcode(s);
    if (rc) return rc;
    if (s->P.solver_type == EFS_SOLVER_JACOBI) {
        s->p64_valid = false;
        s->x64_has_previous = false;
        const int iters = s->P.max_steps / 2;
        // ping-pong between p and a scratch vector; the result ends in p
        float* src = s->p;
        float* dst = s->cg_x1;
        for (int it = 0; it < iters; it++) {
            hipLaunchKernelGGL(k2m_jacobi_pass, dim3((s->N + MB_NT - 1) / MB_NT), dim3(MB_NT), 0, s->stream, s->N, s->W,
                               s->rowcode, s->div, (const float*)src, dst);
            std::swap(src, dst);
        }
        if (src != s->p) EFS_HIP(hipMemcpyAsync(s->p, src, (size_t)s->N * 4, hipMemcpyDeviceToDevice, s->stream));
        EFS_HIP(hipGetLastError());
        EFS_HIP(hipStreamSynchronize(s->stream));
        if (out) {
            out->iterations = iters;
            out->rr = 0;
            out->rel_residual = 0;
            out->true_rel_residual = -1;
        }
        return 0;  // JacobiSolver2D::solve returns 0: it never reports non-convergence
    }
    Cg2Args a;
    a.N = s->N;
    a.W = s->W;
    a.max_it = s->P.max_steps;
    a.prev = (s->P.initial_guess == EFS_GUESS_PREVIOUS && s->x64_has_previous) ? 1 : 0;
    a.relative = s->P.criterion == EFS_CRIT_RELATIVE;
    a.true_residual = s->P.true_residual != 0;
    a.thr = (double)s->P.solver_threshold;
    if ((rc = cg2_solve(s, a))) return rc;
    const CGState fin = *s->st_host;
    s->p64_valid = true;
    s->x64_has_previous = true;
    if (out) {
        out->iterations = fin.iter;
        out->rr = fin.gamma;
        out->bnorm = std::sqrt(fin.bnorm2);
        out->rel_residual = fin.bnorm2 > 0 ? std::sqrt(fin.gamma / fin.bnorm2) : 0.0;
        out->true_rel_residual = -1.0;
        if (s->P.true_residual)
            out->true_rel_residual = fin.bnorm2 > 0 ? std::sqrt(fin.true_num / fin.bnorm2) : std::sqrt(fin.true_num);
    }
    if (!std::isfinite(fin.gamma)) return 2;
    return fin.converged ? 0 : 1;
}

int efs2_spmv(efs_sim* s, const float* x, float* y) {
    int rc = launch_rowcode(s);
    if (rc) return rc;
    hipLaunchKernelGGL(k2_spmv, dim3(blocks1d(s->N)), dim3(256), 0, s->stream, s->N, s->W, s->rowcode, x, y);
    EFS_HIP(hipGetLastError());
    return 0;
}

static int project2(efs_sim* s, efs_stats* st) {
    EFS_HIP(hipMemsetAsync(s->p, 0, (size_t)s->N * 4, s->stream));  // pressure.clear()
    int rc = efs2_walls_divergence(s);
    if (rc) return rc;
    efs_stats local{};
    const int status = efs2_solve(s, st ? st : &local);
    if (status < 0) return status;
    if (status == 1) return 1;  // not converged: the velocity keeps its pre-solve state
    hipLaunchKernelGGL(k2_gradient, grid2_interior(s), dim3(BX, BY), 0, s->stream, s->grid(),
                       s->p64_valid ? (const double*)s->cg_x64 : nullptr);
    EFS_HIP(hipGetLastError());
    if (s->P.re_calc_div && (rc = launch_divergence(s))) return rc;  // RE_CALC_DIV (CPUSimulator2D.cpp:503-521)
    if ((rc = launch_walls(s))) return rc;
    return status;
}

int efs2_run_stage(efs_sim* s, int stage, efs_stats* st) {
    switch (stage) {
        case EFS_STAGE_VEL_DECAY: return launch_decay(s);
        case EFS_STAGE_VEL_EXTRAPOLATE: return launch_extrapolate(s);
        case EFS_STAGE_RHO_SHARPEN: return launch_sharpen(s);
        case EFS_STAGE_RHO_ADVECT: return launch_rho_advect(s);
        case EFS_STAGE_VEL_ADVECT: return launch_vel_advect(s);
        case EFS_STAGE_ADD_FORCE: return launch_add_force(s);
        case EFS_STAGE_PROJECT: return project2(s, st);
    }
    efs_set_error("unknown stage id");
    return -1;
}

int efs2_alloc(efs_sim* s) {
    const size_t N = (size_t)s->N, b = N * 4;
    float** fl[] = {&s->u, &s->v, &s->rho, &s->gamma, &s->beta, &s->div, &s->p, &s->bndf, &s->tu, &s->tv, &s->s0, &s->s1,
                    &s->cg_r, &s->cg_d0, &s->cg_d1, &s->cg_x1};
    for (float** p : fl) {
        EFS_HIP(hipMalloc(p, b));
        EFS_HIP(hipMemset(*p, 0, b));
    }
    EFS_HIP(hipMalloc(&s->stage_bytes, b * 2));
    uint8_t** by[] = {&s->bf, &s->valA, &s->valB};
    for (uint8_t** p : by) {
        EFS_HIP(hipMalloc(p, N));
        EFS_HIP(hipMemset(*p, 0, N));
    }
    EFS_HIP(hipMalloc(&s->bwd_p0, b));
    EFS_HIP(hipMalloc(&s->fwd_p0, b));
    EFS_HIP(hipMemset(s->bwd_p0, 0xFF, b));
    EFS_HIP(hipMemset(s->fwd_p0, 0xFF, b));
    EFS_HIP(hipMalloc(&s->bwd_w, b * 4));
    EFS_HIP(hipMalloc(&s->fwd_w, b * 4));
    unsigned long long** ac[] = {&s->acc0, &s->acc1, &s->acc2};
    for (unsigned long long** p : ac) {
        EFS_HIP(hipMalloc(p, N * 8));
        EFS_HIP(hipMemset(*p, 0, N * 8));
    }
    EFS_HIP(hipMalloc(&s->rowcode, N * sizeof(uint16_t)));
    EFS_HIP(hipMalloc(&s->cg_x64, N * 8));
    EFS_HIP(hipMemset(s->cg_x64, 0, N * 8));
    EFS_HIP(hipMalloc(&s->st, sizeof(CGState)));
    EFS_HIP(hipMemset(s->st, 0, sizeof(CGState)));
    // per-block partial sums of the CG kernels
    s->max_blocks = MB_MAXB;
    double** pp[] = {&s->p_rr, &s->p_bb, &s->p_dad};
    for (double** p : pp) {
        EFS_HIP(hipMalloc(p, MB_MAXB * sizeof(double)));
        EFS_HIP(hipMemset(*p, 0, MB_MAXB * sizeof(double)));
    }
    EFS_HIP(hipHostMalloc(&s->st_host, sizeof(CGState), hipHostMallocDefault));
    std::memset(s->st_host, 0, sizeof(CGState));
    return 0;
}
