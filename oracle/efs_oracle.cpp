// CPU ORACLE — test infrastructure only.
//
// A restatement of the reference CPU simulation path of MarQMark/EulerianFluidSimulation
// (CPUSimulator3D / CPUSimulator2D, Grid3D / Grid2D, ScalField*, VecField*, and the CPU parts
// of CGSolver / JacobiSolver).  It is the checker for the HIP backend: only the tests, the
// smoke entry and the benchmark's CPU-baseline leg load it.  Nothing in the product links it.
//
// Two instantiations:
//   ref32 (T = float)  : bit-faithful to the reference: same loop orders, the same
//                         std::unordered_map (3D) / std::map (2D) accumulation order for the
//                         mass-conserving advection matrix, std::pow(float,int) promoted to
//                         double, no FMA contraction (built with -ffp-contract=off, no -march).
//   ref64 (T = double) : the same algorithm in FP64; the error oracle for pressure/divergence.
//
// Switches that are NOT reference behaviour are off by default:
//   key64        : key the advection matrix by a 64-bit i*N+j (the reference keys by a 32-bit
//                  unsigned that wraps for N > 65536, Grid3D.cpp:169).
//   sharpen_mode : 1 = apply every density change first, then trace and scatter (the order a
//                  parallel device can reproduce); 0 = the reference's in-place sequential loop.
//
// Parity status of each function (how it is pinned without running the reference, whose
// build needs the empty glm / nlohmann submodules) is listed in the tests under tests/.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <unordered_map>
#include <vector>

namespace {

// ----------------------------------------------------------------------------------------
// Small vector types with glm's element-wise semantics (glm::vec3 / glm::vec2 are plain
// per-component IEEE operations; glm::length(v) = sqrt(dot(v,v)), dot = (x*x + y*y) + z*z).
// ----------------------------------------------------------------------------------------
template <class T> struct V3 {
    T x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(T a, T b, T c) : x(a), y(b), z(c) {}
    explicit V3(T s) : x(s), y(s), z(s) {}
    T& operator[](int c) { return c == 0 ? x : (c == 1 ? y : z); }
    T operator[](int c) const { return c == 0 ? x : (c == 1 ? y : z); }
};
template <class T> V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }

template <class T> struct V2 {
    T x, y;
    V2() : x(0), y(0) {}
    V2(T a, T b) : x(a), y(b) {}
    explicit V2(T s) : x(s), y(s) {}
    T& operator[](int c) { return c == 0 ? x : y; }
    T operator[](int c) const { return c == 0 ? x : y; }
};
template <class T> V2<T> operator+(V2<T> a, V2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T> V2<T> operator-(V2<T> a, V2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T> V2<T> operator*(V2<T> a, T s) { return {a.x * s, a.y * s}; }
template <class T> V2<T> operator/(V2<T> a, T s) { return {a.x / s, a.y / s}; }

// std::clamp(v, lo, hi) == (v < lo) ? lo : (hi < v) ? hi : v
template <class T> inline T clampv(T v, T lo, T hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }

}  // namespace

extern "C" {

// Step parameters (Properties.h:136-218 names; gravity_force is GRAVITY_FORCE in cells/s^2,
// i.e. GRAVITY_FORCE_PHYSICAL / CELL_SIZE as Application.cpp:89 sets it each frame).
typedef struct {
    float dt;
    float gravity_force;
    float lambda_minus;
    float lambda_plus;
    float eta;
    float liquid_threshold;
    float rho_d;
    float vel_decay;
    int32_t gravity_enabled;
    int32_t solver_type;   // 2 = SELF_CG, 3 = JACOBI (Properties.h:475-481 enum values)
    int32_t max_steps;
    float solver_threshold;
    uint32_t disable_mask; // bit s disables stage s (0 decay .. 6 project), ISimulator.h:15-35
    int32_t key64;         // 0 = reference 32-bit key wrap; 1 = 64-bit key
    int32_t sharpen_mode;  // 0 = reference sequential; 1 = apply-all-then-trace
    int32_t re_calc_div;   // 2D only (RE_CALC_DIV, CPUSimulator2D.cpp:503)
} ora_params;

typedef struct {
    int32_t status;      // 0 ok, 1 solver did not converge (PAUSED set), 2 non-finite
    int32_t iterations;  // SOLVE_STEPS
    double err;          // value the solver returned (r.r for CG, 0 for Jacobi)
} ora_stats;

}  // extern "C"

namespace {

enum { ST_DECAY = 0, ST_EXTRAP, ST_SHARPEN, ST_RHO_ADVECT, ST_VEL_ADVECT, ST_FORCE, ST_PROJECT };
enum { F_VEL = 0, F_RHO, F_GAMMA, F_BETA, F_DIV, F_PRESSURE, F_FORCES, F_BOUNDARY, F_TMPSCAL, F_TMPVEC };

struct AdvCell {  // Grid3D.h _DisAdvCell3D: i, j, val; += adds val only
    unsigned int i = 0;
    unsigned int j = 0;
    float val = 0.f;
};
struct AdvCellD {
    unsigned int i = 0;
    unsigned int j = 0;
    double val = 0.0;
};

// ========================================================================================
// 3D
// ========================================================================================
template <class T> struct Sim3D {
    using Cell = typename std::conditional<std::is_same<T, float>::value, AdvCell, AdvCellD>::type;
    int W, H, D;
    size_t N;
    ora_params P;
    std::vector<V3<T>> vel, tmpVec, forces;
    std::vector<T> rho, tmpScal, boundary, gamma, beta, divergence, pressure;
    // CG / Jacobi scratch (CGSolver.cpp:401-414, JacobiSolver.cpp:145-152)
    std::vector<T> cg_t, cg_tmp, cg_r, cg_d, jac_tmp;
    std::unordered_map<unsigned int, Cell> A32;
    std::unordered_map<uint64_t, Cell> A64;
    int solve_steps = 0;
    bool paused = false;

    Sim3D(int w, int h, int d) : W(w), H(h), D(d) {
        N = (size_t)w * h * d;
        vel.assign(N, V3<T>());
        tmpVec.assign(N, V3<T>());
        forces.assign(N, V3<T>());
        rho.assign(N, T(0));
        tmpScal.assign(N, T(0));
        boundary.assign(N, T(0));
        gamma.assign(N, T(1));  // Grid3D.cpp:26
        beta.assign(N, T(0));
        divergence.assign(N, T(0));
        pressure.assign(N, T(0));
        cg_t.assign(N, T(0));
        cg_tmp.assign(N, T(0));
        cg_r.assign(N, T(0));
        cg_d.assign(N, T(0));
        jac_tmp.assign(N, T(0));
        // Grid3D.cpp:28-42: border cells solid, rho 0.
        for (int z = 0; z < D; z++)
            for (int y = 0; y < H; y++)
                for (int x = 0; x < W; x++) {
                    if (x == 0 || y == 0 || z == 0 || x == W - 1 || y == H - 1 || z == D - 1) {
                        rho[id(x, y, z)] = T(0);
                        boundary[id(x, y, z)] = T(1);
                    }
                }
        std::memset(&P, 0, sizeof(P));
    }

    inline size_t id(size_t x, size_t y, size_t z) const { return z * (size_t)W * H + y * (size_t)W + x; }
    T& R(std::vector<T>& f, size_t x, size_t y, size_t z) { return f[id(x, y, z)]; }

    // ---- ScalField3D::interpolate (ScalField3D.cpp:18-107) ----
    T sinterp(const std::vector<T>& f, V3<T> pos) const {
        const int x = (int)pos.x;
        const int y = (int)pos.y;
        const int z = (int)pos.z;
        int ix1 = x, ix2 = x + 1;
        if (pos.x - (T)x < T(0.5f)) { ix1 = x - 1; ix2 = x; }
        int iy1 = y, iy2 = y + 1;
        if (pos.y - (T)y < T(0.5f)) { iy1 = y - 1; iy2 = y; }
        int iz1 = z, iz2 = z + 1;
        if (pos.z - (T)z < T(0.5f)) { iz1 = z - 1; iz2 = z; }
        ix1 = clampv(ix1, 0, W - 1); ix2 = clampv(ix2, 0, W - 1);
        iy1 = clampv(iy1, 0, H - 1); iy2 = clampv(iy2, 0, H - 1);
        iz1 = clampv(iz1, 0, D - 1); iz2 = clampv(iz2, 0, D - 1);
        const T v[8] = {f[id(ix1, iy1, iz1)], f[id(ix1, iy1, iz2)], f[id(ix1, iy2, iz1)], f[id(ix1, iy2, iz2)],
                        f[id(ix2, iy1, iz1)], f[id(ix2, iy1, iz2)], f[id(ix2, iy2, iz1)], f[id(ix2, iy2, iz2)]};
        const T x1 = (T)ix1 + T(0.5f), x2 = (T)ix2 + T(0.5f);
        const T y1 = (T)iy1 + T(0.5f), y2 = (T)iy2 + T(0.5f);
        const T z1 = (T)iz1 + T(0.5f), z2 = (T)iz2 + T(0.5f);
        const T dz2z1 = z2 - z1;
        T f00, f01, f10, f11;
        if (dz2z1 < T(1e-9f)) {
            f00 = v[1]; f01 = v[3]; f10 = v[5]; f11 = v[7];
        } else {
            const T wz1 = (z2 - pos.z) / dz2z1;
            const T wz2 = (pos.z - z1) / dz2z1;
            f00 = wz1 * v[0] + wz2 * v[1];
            f01 = wz1 * v[2] + wz2 * v[3];
            f10 = wz1 * v[4] + wz2 * v[5];
            f11 = wz1 * v[6] + wz2 * v[7];
        }
        const T dy2y1 = y2 - y1;
        T fx1, fx2;
        if (dy2y1 < T(1e-9f)) {
            fx1 = f01; fx2 = f11;
        } else {
            const T wy1 = (y2 - pos.y) / dy2y1;
            const T wy2 = (pos.y - y1) / dy2y1;
            fx1 = wy1 * f00 + wy2 * f01;
            fx2 = wy1 * f10 + wy2 * f11;
        }
        const T dx2x1 = x2 - x1;
        T fxyz;
        if (dx2x1 < T(1e-9f)) {
            fxyz = fx2;
        } else {
            const T wx1 = (x2 - pos.x) / dx2x1;
            const T wx2 = (pos.x - x1) / dx2x1;
            fxyz = wx1 * fx1 + wx2 * fx2;
        }
        return fxyz;
    }

    // ---- VecField3D::trilinear_interpolation / interpolate (VecField3D.cpp:20-86) ----
    static V3<T> stagOff(int comp) {
        switch (comp) {
            case 0: return V3<T>(T(-1.0f), T(-0.5f), T(-0.5f));
            case 1: return V3<T>(T(-0.5f), T(-1.0f), T(-0.5f));
            default: return V3<T>(T(-0.5f), T(-0.5f), T(-1.0f));
        }
    }
    T trilin(const std::vector<V3<T>>& f, V3<T> pos, int lx, int ly, int lz, int comp) const {
        const V3<T> off = stagOff(comp);
        const T x0 = (T)lx + off.x;
        const T y0 = (T)ly + off.y;
        const T z0 = (T)lz + off.z;
        const T tx = pos.x - x0, ty = pos.y - y0, tz = pos.z - z0;
        const T v000 = f[id(lx, ly, lz)][comp];
        const T v001 = f[id(lx, ly, lz + 1)][comp];
        const T v010 = f[id(lx, ly + 1, lz)][comp];
        const T v011 = f[id(lx, ly + 1, lz + 1)][comp];
        const T v100 = f[id(lx + 1, ly, lz)][comp];
        const T v101 = f[id(lx + 1, ly, lz + 1)][comp];
        const T v110 = f[id(lx + 1, ly + 1, lz)][comp];
        const T v111 = f[id(lx + 1, ly + 1, lz + 1)][comp];
        const T v00 = v000 * (T(1.0f) - tz) + v001 * tz;
        const T v01 = v010 * (T(1.0f) - tz) + v011 * tz;
        const T v10 = v100 * (T(1.0f) - tz) + v101 * tz;
        const T v11 = v110 * (T(1.0f) - tz) + v111 * tz;
        const T vLow = v00 * (T(1.0f) - ty) + v01 * ty;
        const T vHigh = v10 * (T(1.0f) - ty) + v11 * ty;
        return vLow * (T(1.0f) - tx) + vHigh * tx;
    }
    static int lowIndex(T coord, T offset) { return (int)(std::floor(coord - offset) + T(1.0e-10f)); }
    V3<T> vinterp(const std::vector<V3<T>>& f, V3<T> pos) const {
        V3<T> r;
        for (int c = 0; c < 3; c++) {
            const V3<T> o = stagOff(c);
            r[c] = trilin(f, pos, lowIndex(pos.x, o.x), lowIndex(pos.y, o.y), lowIndex(pos.z, o.z), c);
        }
        return r;
    }

    // ---- Grid3D::traceParticle (Grid3D.cpp:66-122) ----
    bool inWall(V3<T> p) const {
        const T Wf = (T)(float)W, Hf = (T)(float)H, Df = (T)(float)D;
        if (p.x < T(1.f) || p.x > Wf - T(2.f) || p.y < T(1.f) || p.y > Hf - T(2.f) || p.z < T(1.f) ||
            p.z > Df - T(2.f))
            return true;
        const int ix = clampv((int)p.x, 0, W - 1);
        const int iy = clampv((int)p.y, 0, H - 1);
        const int iz = clampv((int)p.z, 0, D - 1);
        return boundary[id(ix, iy, iz)] > T(0.5f);
    }
    V3<T> traceParticle(V3<T> start, T dt, int substeps) const {
        const T Wf = (T)(float)W, Hf = (T)(float)H, Df = (T)(float)D;
        auto clampP = [&](V3<T> p) {
            return V3<T>(clampv(p.x, T(1.f), Wf - T(2.f)), clampv(p.y, T(1.f), Hf - T(2.f)),
                         clampv(p.z, T(1.f), Df - T(2.f)));
        };
        auto sampleVel = [&](V3<T> p) { return vinterp(vel, p - V3<T>(T(0.5f))); };
        if (inWall(start)) return start;
        const T subDt = dt / (T)substeps;
        V3<T> pos = start;
        for (int s = 0; s < substeps; s++) {
            const V3<T> k1 = sampleVel(pos);
            const V3<T> mid = clampP(pos + k1 * (subDt * T(0.5f)));
            const V3<T> k2 = sampleVel(mid);
            const V3<T> next = pos + k2 * subDt;
            if (inWall(next)) {
                V3<T> safe = pos, unsafe = next;
                for (int b = 0; b < 8; b++) {
                    const V3<T> halfway = (safe + unsafe) * T(0.5f);
                    if (inWall(halfway)) unsafe = halfway;
                    else safe = halfway;
                }
                return safe;
            }
            pos = next;
        }
        return pos;
    }

    // ---- CPUSimulator3D::vel_decay (CPUSimulator3D.cpp:54-59) ----
    void vel_decay() {
        const T decay = (T)P.vel_decay;
        for (int z = 1; z < D - 1; z++)
            for (int y = 1; y < H - 1; y++)
                for (int x = 1; x < W - 1; x++)
                    if (rho[id(x, y, z)] < 0.5) vel[id(x, y, z)] = vel[id(x, y, z)] * decay;
    }

    // ---- CPUSimulator3D::vel_extrapolate (CPUSimulator3D.cpp:61-111) ----
    void vel_extrapolate() {
        std::vector<T> validA(N, T(0)), validB(N, T(0));
        for (int z = 1; z < D - 1; z++)
            for (int y = 1; y < H - 1; y++)
                for (int x = 1; x < W - 1; x++)
                    validA[id(x, y, z)] = (rho[id(x, y, z)] >= T(0.5f) && boundary[id(x, y, z)] != T(1.f)) ? T(1) : T(0);
        auto pass = [&](std::vector<V3<T>>& sv, std::vector<T>& sval, std::vector<V3<T>>& dv, std::vector<T>& dval) {
            for (int z = 1; z < D - 1; z++)
                for (int y = 1; y < H - 1; y++)
                    for (int x = 1; x < W - 1; x++) {
                        const size_t c = id(x, y, z);
                        if (boundary[c] == T(1.f)) { dv[c] = sv[c]; dval[c] = T(0); continue; }
                        if (sval[c] > T(0.5f)) { dv[c] = sv[c]; dval[c] = T(1); continue; }
                        V3<T> sum(T(0));
                        T n = 0;
                        auto sample = [&](int nx, int ny, int nz) {
                            const size_t q = id(nx, ny, nz);
                            if (sval[q] > T(0.5f)) { sum = sum + sv[q]; n++; }
                        };
                        sample(x - 1, y, z); sample(x + 1, y, z);
                        sample(x, y - 1, z); sample(x, y + 1, z);
                        sample(x, y, z - 1); sample(x, y, z + 1);
                        dv[c] = (n > 0) ? sum / n : sv[c];
                        dval[c] = (n > 0) ? T(1) : T(0);
                    }
        };
        for (int i = 0; i < 3; i++) {
            pass(vel, validA, tmpVec, validB);
            pass(tmpVec, validB, vel, validA);
        }
    }

    // ---- Grid3D::addA (Grid3D.cpp:168-174) ----
    void addA(unsigned int i, unsigned int j, T val) {
        if (P.key64) {
            const uint64_t k = (uint64_t)i * (uint64_t)N + j;
            auto it = A64.find(k);
            if (it != A64.end()) it->second.val += val;
            else A64[k] = Cell{i, j, val};
        } else {
            const unsigned int k = i * (unsigned int)N + j;  // wraps mod 2^32 like the reference
            auto it = A32.find(k);
            if (it != A32.end()) it->second.val += val;
            else A32[k] = Cell{i, j, val};
        }
    }
    template <class F> void forA(F fn) {
        if (P.key64) for (auto& kv : A64) fn(kv.second);
        else for (auto& kv : A32) fn(kv.second);
    }

    // corner weights of a traced position (CPUSimulator3D.cpp:140-163); returns false if sum == 0
    bool cornerWeights(V3<T> p, int& p0x, int& p0y, int& p0z, T m[8], T& sum) const {
        p0x = (int)std::floor(p.x);
        p0y = (int)std::floor(p.y);
        p0z = (int)std::floor(p.z);
        const T dx0 = 1 - (p.x - p0x), dy0 = 1 - (p.y - p0y), dz0 = 1 - (p.z - p0z);
        const T dx1 = 1 - (p0x + 1 - p.x), dy1 = 1 - (p0y + 1 - p.y), dz1 = 1 - (p0z + 1 - p.z);
        for (int k = 0; k < 8; k++) m[k] = 0;
        if (boundary[id(p0x + 0, p0y + 0, p0z + 0)] != T(1.f)) m[0] = dx0 * dy0 * dz0;
        if (boundary[id(p0x + 1, p0y + 0, p0z + 0)] != T(1.f)) m[1] = dx1 * dy0 * dz0;
        if (boundary[id(p0x + 0, p0y + 1, p0z + 0)] != T(1.f)) m[2] = dx0 * dy1 * dz0;
        if (boundary[id(p0x + 1, p0y + 1, p0z + 0)] != T(1.f)) m[3] = dx1 * dy1 * dz0;
        if (boundary[id(p0x + 0, p0y + 0, p0z + 1)] != T(1.f)) m[4] = dx0 * dy0 * dz1;
        if (boundary[id(p0x + 1, p0y + 0, p0z + 1)] != T(1.f)) m[5] = dx1 * dy0 * dz1;
        if (boundary[id(p0x + 0, p0y + 1, p0z + 1)] != T(1.f)) m[6] = dx0 * dy1 * dz1;
        if (boundary[id(p0x + 1, p0y + 1, p0z + 1)] != T(1.f)) m[7] = dx1 * dy1 * dz1;
        sum = 0;
        for (int k = 0; k < 8; k++) sum += m[k];
        return !(sum == 0);
    }

    // ---- CPUSimulator3D::rho_advect (CPUSimulator3D.cpp:113-364) ----
    void rho_advect() {
        const T dt = (T)P.dt;
        auto idx = [&](int x, int y, int z) -> unsigned int {
            return (unsigned int)z * (unsigned int)W * (unsigned int)H + (unsigned int)y * (unsigned int)W + (unsigned int)x;
        };
        // Step 1: advect gamma (a zero-initialised copy, ScalField3D(gamma,false) + vector resize)
        std::vector<T> adv(N, T(0));
        for (int z = 1; z < D - 1; z++)
            for (int y = 1; y < H - 1; y++)
                for (int x = 1; x < W - 1; x++) {
                    if (boundary[id(x, y, z)] == T(1.f)) continue;
                    const V3<T> prev = traceParticle(V3<T>((T)x, (T)y, (T)z), -dt, 4);
                    adv[id(x, y, z)] = sinterp(gamma, prev);
                }
        gamma = adv;
        // Step 2: A from backward traces
        A32.clear();
        A64.clear();
        for (int z = 1; z < D - 1; z++)
            for (int y = 1; y < H - 1; y++)
                for (int x = 1; x < W - 1; x++) {
                    const V3<T> prev = traceParticle(V3<T>((T)x, (T)y, (T)z), -dt, 4);
                    int px, py, pz; T m[8], sum;
                    if (!cornerWeights(prev, px, py, pz, m, sum)) continue;
                    const unsigned int i = idx(x, y, z);
                    addA(i, idx(px + 0, py + 0, pz + 0), m[0] / sum);
                    addA(i, idx(px + 1, py + 0, pz + 0), m[1] / sum);
                    addA(i, idx(px + 0, py + 1, pz + 0), m[2] / sum);
                    addA(i, idx(px + 1, py + 1, pz + 0), m[3] / sum);
                    addA(i, idx(px + 0, py + 0, pz + 1), m[4] / sum);
                    addA(i, idx(px + 1, py + 0, pz + 1), m[5] / sum);
                    addA(i, idx(px + 0, py + 1, pz + 1), m[6] / sum);
                    addA(i, idx(px + 1, py + 1, pz + 1), m[7] / sum);
                }
        // Step 3
        forA([&](Cell& c) { if (gamma[c.i] > T(1.e-5f)) c.val /= gamma[c.i]; });
        // Step 4
        std::fill(beta.begin(), beta.end(), T(0));
        forA([&](Cell& c) { beta[c.j] += c.val; });
        // Step 5: forward traces for cells with beta < 1
        for (int z = 1; z < D - 1; z++)
            for (int y = 1; y < H - 1; y++)
                for (int x = 1; x < W - 1; x++) {
                    const unsigned int j = idx(x, y, z);
                    if (beta[j] >= 1) continue;
                    const V3<T> next = traceParticle(V3<T>((T)x, (T)y, (T)z), +dt, 4);
                    int px, py, pz; T m[8], sum;
                    if (!cornerWeights(next, px, py, pz, m, sum)) continue;
                    addA(idx(px + 0, py + 0, pz + 0), j, (1 - beta[j]) * (m[0] / sum));
                    addA(idx(px + 1, py + 0, pz + 0), j, (1 - beta[j]) * (m[1] / sum));
                    addA(idx(px + 0, py + 1, pz + 0), j, (1 - beta[j]) * (m[2] / sum));
                    addA(idx(px + 1, py + 1, pz + 0), j, (1 - beta[j]) * (m[3] / sum));
                    addA(idx(px + 0, py + 0, pz + 1), j, (1 - beta[j]) * (m[4] / sum));
                    addA(idx(px + 1, py + 0, pz + 1), j, (1 - beta[j]) * (m[5] / sum));
                    addA(idx(px + 0, py + 1, pz + 1), j, (1 - beta[j]) * (m[6] / sum));
                    addA(idx(px + 1, py + 1, pz + 1), j, (1 - beta[j]) * (m[7] / sum));
                }
        // Steps 6-10
        std::fill(gamma.begin(), gamma.end(), T(0));
        forA([&](Cell& c) { gamma[c.i] += c.val; });
        forA([&](Cell& c) { if (gamma[c.i] > T(1.e-5f)) c.val /= gamma[c.i]; });
        std::fill(beta.begin(), beta.end(), T(0));
        forA([&](Cell& c) { beta[c.j] += c.val; });
        forA([&](Cell& c) { if (beta[c.j] > T(1.e-5f)) c.val /= beta[c.j]; });
        std::fill(gamma.begin(), gamma.end(), T(0));
        forA([&](Cell& c) { gamma[c.i] += c.val; });
        // Step 11: rho' = A rho on the interior
        std::fill(tmpScal.begin(), tmpScal.end(), T(0));
        forA([&](Cell& c) { tmpScal[c.i] += c.val * rho[c.j]; });
        for (int z = 1; z < D - 1; z++)
            for (int y = 1; y < H - 1; y++)
                for (int x = 1; x < W - 1; x++) rho[id(x, y, z)] = tmpScal[id(x, y, z)];
        // Step 12: 3 iterations of in-place pairwise diffusion, x then y then z
        for (int it = 0; it < 3; it++) {
            for (int z = 1; z < D - 1; z++)
                for (int y = 1; y < H - 1; y++)
                    for (int x = 2; x < W - 1; x++) diffusePair(id(x - 1, y, z), id(x, y, z));
            for (int z = 1; z < D - 1; z++)
                for (int x = 1; x < W - 1; x++)
                    for (int y = 2; y < H - 1; y++) diffusePair(id(x, y - 1, z), id(x, y, z));
            for (int y = 1; y < H - 1; y++)
                for (int x = 1; x < W - 1; x++)
                    for (int z = 2; z < D - 1; z++) diffusePair(id(x, y, z - 1), id(x, y, z));
        }
    }
    // one pair of the diffusion sweep (CPUSimulator3D.cpp:281-303): j = lower neighbour, i = cell
    void diffusePair(size_t jj, size_t ii) {
        const T g_j = gamma[jj];
        const T g_i = gamma[ii];
        if (boundary[ii] == T(1.f) || boundary[jj] == T(1.f)) return;
        if (g_j < T(1e-5f) || g_i < T(1e-5f)) return;
        if (g_j == g_i) return;
        const T avg = (g_j + g_i) / T(2.f);
        if (g_j > g_i) {
            const T dif = rho[jj] * (g_j - g_i) / (T(2.f) * g_j);
            rho[jj] -= dif;
            rho[ii] += dif;
        } else {
            const T dif = rho[ii] * (g_i - g_j) / (T(2.f) * g_i);
            rho[ii] -= dif;
            rho[jj] += dif;
        }
        gamma[jj] = avg;
        gamma[ii] = avg;
    }

    // ---- CPUSimulator3D::rho_sharpen (CPUSimulator3D.cpp:366-456) ----
    // Delta rho of every interior cell (CPUSimulator3D.cpp:366-444), before any is applied.
    void sharpen_delta(std::vector<T>& del_rho) const {
        const T delT = (T)(P.dt * 3);  // PROPERTY_F(DT) * 3 is a float product
        del_rho.assign(N, T(0));
        std::vector<T> max_inc(N, T(0)), max_dec(N, T(0)), w(N, T(0));
        std::vector<V3<T>> dp(N), dm(N);
        for (int z = 1; z < D - 1; z++)
            for (int y = 1; y < H - 1; y++)
                for (int x = 1; x < W - 1; x++) {
                    const size_t c = id(x, y, z);
                    dp[c].x = -(rho[c] - rho[id(x - 1, y, z)]) * delT;
                    dm[c].x = -(rho[id(x + 1, y, z)] - rho[c]) * delT;
                    dp[c].y = -(rho[c] - rho[id(x, y - 1, z)]) * delT;
                    dm[c].y = -(rho[id(x, y + 1, z)] - rho[c]) * delT;
                    dp[c].z = -(rho[c] - rho[id(x, y, z - 1)]) * delT;
                    dm[c].z = -(rho[id(x, y, z + 1)] - rho[c]) * delT;
                }
        // std::pow(float, int) promotes to double (C++11 [c.math]); the sum and sqrt are double.
        auto sq = [](T v) -> double { return std::pow((double)v, 2.0); };
        for (int z = 1; z < D - 1; z++)
            for (int y = 1; y < H - 1; y++)
                for (int x = 1; x < W - 1; x++) {
                    const size_t c = id(x, y, z);
                    max_inc[c] = (T)std::sqrt(std::max(sq(std::max(dp[c].x, T(0))), sq(std::min(dm[c].x, T(0)))) +
                                              std::max(sq(std::max(dp[c].y, T(0))), sq(std::min(dm[c].y, T(0)))) +
                                              std::max(sq(std::max(dp[c].z, T(0))), sq(std::min(dm[c].z, T(0)))));
                    max_dec[c] = (T)std::sqrt(std::max(sq(std::min(dp[c].x, T(0))), sq(std::max(dm[c].x, T(0)))) +
                                              std::max(sq(std::min(dp[c].y, T(0))), sq(std::max(dm[c].y, T(0)))) +
                                              std::max(sq(std::min(dp[c].z, T(0))), sq(std::max(dm[c].z, T(0)))));
                }
        for (int z = 1; z < D - 1; z++)
            for (int y = 1; y < H - 1; y++)
                for (int x = 1; x < W - 1; x++) {
                    const size_t c = id(x, y, z);
                    T mx = -INFINITY;
                    mx = std::max(mx, std::abs(rho[c] - rho[id(x - 1, y, z)]));
                    mx = std::max(mx, std::abs(rho[c] - rho[id(x + 1, y, z)]));
                    mx = std::max(mx, std::abs(rho[c] - rho[id(x, y - 1, z)]));
                    mx = std::max(mx, std::abs(rho[c] - rho[id(x, y + 1, z)]));
                    mx = std::max(mx, std::abs(rho[c] - rho[id(x, y, z - 1)]));
                    mx = std::max(mx, std::abs(rho[c] - rho[id(x, y, z + 1)]));
                    // std::pow(float, float) is powf
                    w[c] = std::pow(rho[c] - T(0.5f), T(3.f)) * (1 - std::min(T(1.f), mx / T(0.2f)));
                }
        for (int z = 1; z < D - 1; z++)
            for (int y = 1; y < H - 1; y++)
                for (int x = 1; x < W - 1; x++) {
                    const size_t c = id(x, y, z);
                    if (w[c] >= 0) del_rho[c] = w[c] * max_inc[c];
                    else del_rho[c] = w[c] * max_dec[c];
                    if (rho[c] + del_rho[c] < 0 || rho[c] < 1e-5) del_rho[c] = -rho[c];
                    else if (rho[c] > T(0.5f)) del_rho[c] = 0;
                }
    }
    void rho_sharpen() {
        std::vector<T> del_rho;
        sharpen_delta(del_rho);
        const T rhoD = (T)P.rho_d;
        if (P.sharpen_mode == 0) {
            for (int z = 1; z < D - 1; z++)
                for (int y = 1; y < H - 1; y++)
                    for (int x = 1; x < W - 1; x++) {
                        const size_t c = id(x, y, z);
                        const T dr = del_rho[c];
                        if (std::abs(dr) < T(1e-6f)) continue;
                        rho[c] += dr;
                        rho[c] = std::max(rho[c], T(0.f));
                        const V3<T> p = trace_along_field(V3<T>((T)x, (T)y, (T)z), rhoD, 10);
                        scatter_values(rho, p, -dr);
                    }
        } else {
            // apply-all-then-trace: every change is applied, then every trace reads the same field,
            // and the scatters accumulate into a separate buffer that is added at the end.
            for (size_t c = 0; c < N; c++) {
                const T dr = del_rho[c];
                if (std::abs(dr) < T(1e-6f)) continue;
                rho[c] += dr;
                rho[c] = std::max(rho[c], T(0.f));
            }
            std::vector<double> acc(N, 0.0);
            for (int z = 1; z < D - 1; z++)
                for (int y = 1; y < H - 1; y++)
                    for (int x = 1; x < W - 1; x++) {
                        const size_t c = id(x, y, z);
                        const T dr = del_rho[c];
                        if (std::abs(dr) < T(1e-6f)) continue;
                        const V3<T> p = trace_along_field(V3<T>((T)x, (T)y, (T)z), rhoD, 10);
                        scatter_acc(acc, p, -dr);
                    }
            for (size_t c = 0; c < N; c++)
                if (acc[c] != 0.0) rho[c] = (T)((double)rho[c] + acc[c]);
        }
    }

    // ---- CPUSimulator3D::trace_along_field (CPUSimulator3D.cpp:568-607) ----
    V3<T> trace_along_field(V3<T> start, T Dm, int substeps) const {
        const T stepSize = Dm / (T)substeps;
        V3<T> pos = start;
        for (int s = 0; s < substeps; s++) {
            if (sinterp(rho, pos) >= T(0.5f)) break;
            const int ix = clampv((int)pos.x, 1, W - 2);
            const int iy = clampv((int)pos.y, 1, H - 2);
            const int iz = clampv((int)pos.z, 1, D - 2);
            const V3<T> grad(rho[id(ix + 1, iy, iz)] - rho[id(ix - 1, iy, iz)],
                             rho[id(ix, iy + 1, iz)] - rho[id(ix, iy - 1, iz)],
                             rho[id(ix, iy, iz + 1)] - rho[id(ix, iy, iz - 1)]);
            const T len = std::sqrt(grad.x * grad.x + grad.y * grad.y + grad.z * grad.z);
            if (len < T(1e-6f)) break;
            const V3<T> dir = grad / len;
            V3<T> next = pos + dir * stepSize;
            next.x = clampv(next.x, T(1.f), (T)(float)W - T(2.f));
            next.y = clampv(next.y, T(1.f), (T)(float)H - T(2.f));
            next.z = clampv(next.z, T(1.f), (T)(float)D - T(2.f));
            if (boundary[id((size_t)next.x, (size_t)next.y, (size_t)next.z)] > T(0.5f)) break;
            pos = next;
        }
        return pos;
    }

    // ---- CPUSimulator3D::scatter_values (CPUSimulator3D.cpp:609-655) ----
    void scatter_values(std::vector<T>& f, V3<T> pos, T value) const {
        int xs[2], ys[2], zs[2];
        T w[8];
        xs[0] = clampv((int)pos.x, 0, W - 1); ys[0] = clampv((int)pos.y, 0, H - 1); zs[0] = clampv((int)pos.z, 0, D - 1);
        xs[1] = clampv((int)pos.x + 1, 0, W - 1); ys[1] = clampv((int)pos.y + 1, 0, H - 1); zs[1] = clampv((int)pos.z + 1, 0, D - 1);
        const T fx = pos.x - (T)(int)pos.x, fy = pos.y - (T)(int)pos.y, fz = pos.z - (T)(int)pos.z;
        w[0] = (T(1.f) - fx) * (T(1.f) - fy) * (T(1.f) - fz);
        w[1] = fx * (T(1.f) - fy) * (T(1.f) - fz);
        w[2] = (T(1.f) - fx) * fy * (T(1.f) - fz);
        w[3] = fx * fy * (T(1.f) - fz);
        w[4] = (T(1.f) - fx) * (T(1.f) - fy) * fz;
        w[5] = fx * (T(1.f) - fy) * fz;
        w[6] = (T(1.f) - fx) * fy * fz;
        w[7] = fx * fy * fz;
        for (int k = 0; k < 8; k++)
            if (boundary[id(xs[k & 1], ys[(k >> 1) & 1], zs[(k >> 2) & 1])] > T(0.5f)) w[k] = T(0.f);
        const T total = w[0] + w[1] + w[2] + w[3] + w[4] + w[5] + w[6] + w[7];
        if (total < T(1e-6f)) return;
        const T inv = T(1.f) / total;
        for (int k = 0; k < 8; k++) f[id(xs[k & 1], ys[(k >> 1) & 1], zs[(k >> 2) & 1])] += value * w[k] * inv;
    }
    void scatter_acc(std::vector<double>& acc, V3<T> pos, T value) const {
        int xs[2], ys[2], zs[2];
        T w[8];
        xs[0] = clampv((int)pos.x, 0, W - 1); ys[0] = clampv((int)pos.y, 0, H - 1); zs[0] = clampv((int)pos.z, 0, D - 1);
        xs[1] = clampv((int)pos.x + 1, 0, W - 1); ys[1] = clampv((int)pos.y + 1, 0, H - 1); zs[1] = clampv((int)pos.z + 1, 0, D - 1);
        const T fx = pos.x - (T)(int)pos.x, fy = pos.y - (T)(int)pos.y, fz = pos.z - (T)(int)pos.z;
        w[0] = (T(1.f) - fx) * (T(1.f) - fy) * (T(1.f) - fz);
        w[1] = fx * (T(1.f) - fy) * (T(1.f) - fz);
        w[2] = (T(1.f) - fx) * fy * (T(1.f) - fz);
        w[3] = fx * fy * (T(1.f) - fz);
        w[4] = (T(1.f) - fx) * (T(1.f) - fy) * fz;
        w[5] = fx * (T(1.f) - fy) * fz;
        w[6] = (T(1.f) - fx) * fy * fz;
        w[7] = fx * fy * fz;
        for (int k = 0; k < 8; k++)
            if (boundary[id(xs[k & 1], ys[(k >> 1) & 1], zs[(k >> 2) & 1])] > T(0.5f)) w[k] = T(0.f);
        const T total = w[0] + w[1] + w[2] + w[3] + w[4] + w[5] + w[6] + w[7];
        if (total < T(1e-6f)) return;
        const T inv = T(1.f) / total;
        for (int k = 0; k < 8; k++) acc[id(xs[k & 1], ys[(k >> 1) & 1], zs[(k >> 2) & 1])] += (double)(value * w[k] * inv);
    }

    // ---- CPUSimulator3D::add_force (CPUSimulator3D.cpp:458-476) ----
    void add_force() {
        const T dt = (T)P.dt;
        const T gdt = (T)(P.gravity_force * P.dt);  // PROPERTY_F(GRAVITY_FORCE) * PROPERTY_F(DT) in float
        for (int z = 1; z < D - 1; z++)
            for (int y = 1; y < H - 1; y++)
                for (int x = 1; x < W - 1; x++) {
                    const size_t c = id(x, y, z);
                    if (boundary[c] == T(1.f)) continue;
                    vel[c] = vel[c] + forces[c] * dt;
                    if (P.gravity_enabled) vel[c].y -= gdt;
                }
    }

    // ---- CPUSimulator3D::vel_advect (CPUSimulator3D.cpp:478-504) ----
    void vel_advect() {
        const T dt = (T)P.dt;
        for (int z = 1; z < D - 1; z++)
            for (int y = 1; y < H - 1; y++)
                for (int x = 1; x < W - 1; x++) {
                    const V3<T> prev = traceParticle(V3<T>((T)x, (T)y, (T)z), -dt, 4);
                    tmpVec[id(x, y, z)] = vinterp(vel, prev - V3<T>(T(0.5f)));
                }
        for (int z = 1; z < D - 1; z++)
            for (int y = 1; y < H - 1; y++)
                for (int x = 1; x < W - 1; x++) {
                    const size_t c = id(x, y, z);
                    if (x == 1 || boundary[c] == T(1.f) || boundary[id(x - 1, y, z)] == T(1.f)) vel[c].x = tmpVec[c].x;
                    else vel[c].x = T(0.5f) * (tmpVec[c].x + tmpVec[id(x - 1, y, z)].x);
                    if (y == 1 || boundary[c] == T(1.f) || boundary[id(x, y - 1, z)] == T(1.f)) vel[c].y = tmpVec[c].y;
                    else vel[c].y = T(0.5f) * (tmpVec[c].y + tmpVec[id(x, y - 1, z)].y);
                    if (z == 1 || boundary[c] == T(1.f) || boundary[id(x, y, z - 1)] == T(1.f)) vel[c].z = tmpVec[c].z;
                    else vel[c].z = T(0.5f) * (tmpVec[c].z + tmpVec[id(x, y, z - 1)].z);
                }
    }

    // ---- CPUSimulator3D::enforce_wall_boundaries (CPUSimulator3D.cpp:657-679) ----
    void enforce_wall_boundaries() {
        for (int z = 0; z < D; z++)
            for (int y = 0; y < H; y++)
                for (int x = 0; x < W; x++) {
                    if (x == 0 || y == 0 || z == 0) continue;
                    V3<T>& v = vel[id(x, y, z)];
                    const T b = boundary[id(x, y, z)];
                    if (b == T(1.f) && v.x > 0) v.x = T(0.f);
                    if (boundary[id(x - 1, y, z)] == T(1.f) && v.x < 0) v.x = T(0.f);
                    if (b == T(1.f) && v.y > 0) v.y = T(0.f);
                    if (boundary[id(x, y - 1, z)] == T(1.f) && v.y < 0) v.y = T(0.f);
                    if (b == T(1.f) && v.z > 0) v.z = T(0.f);
                    if (boundary[id(x, y, z - 1)] == T(1.f) && v.z < 0) v.z = T(0.f);
                }
    }

    // divergence with density correction (CPUSimulator3D.cpp:518-539)
    void compute_divergence() {
        const T liq = (T)P.liquid_threshold, lp = (T)P.lambda_plus, lm = (T)P.lambda_minus, eta = (T)P.eta;
        for (int z = 1; z < D - 1; z++)
            for (int y = 1; y < H - 1; y++)
                for (int x = 1; x < W - 1; x++) {
                    const size_t c = id(x, y, z);
                    if (rho[c] < liq) continue;
                    if (boundary[c] == T(1.f)) continue;
                    const T dx = vel[id(x + 1, y, z)].x - vel[c].x;
                    const T dy = vel[id(x, y + 1, z)].y - vel[c].y;
                    const T dz = vel[id(x, y, z + 1)].z - vel[c].z;
                    divergence[c] = dx + dy + dz;
                    if (rho[c] > T(1.f)) divergence[c] -= std::min(lm * (rho[c] - T(1.f)), eta);
                    if (rho[c] < T(1.f)) divergence[c] += std::min(lp * (T(1.f) - rho[c]), eta);
                }
    }

    // ---- CPUSimulator3D::project (CPUSimulator3D.cpp:506-566) ----
    int project(ora_stats* st) {
        std::fill(pressure.begin(), pressure.end(), T(0));
        std::fill(divergence.begin(), divergence.end(), T(0));
        enforce_wall_boundaries();
        compute_divergence();
        T err;
        if (P.solver_type == 3) err = jacobi_solve();
        else err = cg_solve();
        if (st) { st->err = (double)err; st->iterations = solve_steps; }
        if (err > (T)P.solver_threshold) {
            paused = true;
            return 1;
        }
        subtract_pressure_gradient();
        enforce_wall_boundaries();
        return std::isfinite((double)err) ? 0 : 2;
    }
    void subtract_pressure_gradient() {
        for (int z = 1; z < D - 1; z++)
            for (int y = 1; y < H - 1; y++)
                for (int x = 1; x < W - 1; x++) {
                    const size_t c = id(x, y, z);
                    if (boundary[c] == T(1.f)) continue;
                    V3<T> dP(T(0));
                    if (x > 1) dP.x = pressure[c] - pressure[id(x - 1, y, z)];
                    if (y > 1) dP.y = pressure[c] - pressure[id(x, y - 1, z)];
                    if (z > 1) dP.z = pressure[c] - pressure[id(x, y, z - 1)];
                    vel[c] = vel[c] - dP;
                }
    }

    // ---- CGSolverCPU3D::impl_mat_mul (CGSolver.cpp:467-497) ----
    void mat_mul(std::vector<T>& out, const std::vector<T>& xf) const {
        for (int z = 0; z < D; z++)
            for (int y = 0; y < H; y++)
                for (int x = 0; x < W; x++) {
                    const size_t c = id(x, y, z);
                    if (rho[c] <= T(0.5f) || boundary[c] == T(1)) { out[c] = xf[c]; continue; }
                    T sum = T(0.0f);
                    int nCount = 0;
                    auto nb = [&](int nx, int ny, int nz) {
                        if (nx < 0 || nx >= W || ny < 0 || ny >= H || nz < 0 || nz >= D) return;
                        if (nx == 0 || nx == W - 1 || ny == 0 || ny == H - 1 || nz == 0 || nz == D - 1) return;
                        const size_t q = id(nx, ny, nz);
                        if (boundary[q] > T(0.5f)) return;
                        nCount++;
                        if (rho[q] > T(0.5f)) sum += xf[q];
                    };
                    nb(x - 1, y, z); nb(x + 1, y, z);
                    nb(x, y - 1, z); nb(x, y + 1, z);
                    nb(x, y, z + 1); nb(x, y, z - 1);
                    out[c] = sum - (T)nCount * xf[c];
                }
    }
    static T dot(const std::vector<T>& a, const std::vector<T>& b) {
        T d = 0;
        for (size_t i = 0; i < b.size(); i++) d += a[i] * b[i];
        return d;
    }

    // ---- CGSolverCPU3D::solve (CGSolver.cpp:416-465) ----
    T cg_solve() {
        const int steps = P.max_steps;
        const T thr = (T)P.solver_threshold;
        solve_steps = 0;  // the reference leaves SOLVE_STEPS stale on the early exit below
        auto& t = cg_t; std::fill(t.begin(), t.end(), T(0));
        auto& tmp = cg_tmp; std::fill(tmp.begin(), tmp.end(), T(0));
        auto& r = cg_r; r = divergence;
        pressure = divergence;
        mat_mul(t, pressure);
        for (size_t i = 0; i < N; i++) r[i] -= t[i];
        auto& d = cg_d; d = r;
        T sq = dot(r, r);
        if (sq < thr) return sq;  // SOLVE_STEPS is left unchanged on this early exit
        int i = 0;
        for (i = 0; i < steps; i++) {
            mat_mul(t, d);
            const T alpha = sq / dot(d, t);
            for (size_t k = 0; k < N; k++) pressure[k] += alpha * d[k];
            if (i & 0x3F) {
                for (size_t k = 0; k < N; k++) r[k] += -alpha * t[k];
            } else {
                mat_mul(t, pressure);
                r = divergence;
                for (size_t k = 0; k < N; k++) r[k] -= t[k];
            }
            const T old = sq;
            sq = dot(r, r);
            if (sq < thr) break;
            const T beta_ = sq / old;
            for (size_t k = 0; k < N; k++) d[k] *= beta_;
            for (size_t k = 0; k < N; k++) d[k] += r[k];
        }
        solve_steps = i + 1;
        return sq;
    }

    // ---- JacobiSolverCPU3D::solve (JacobiSolver.cpp:154-212) ----
    T jacobi_solve() {
        const int iter = P.max_steps / 2;
        solve_steps = iter;
        std::vector<T>& p = pressure;
        std::vector<T>& tmp = jac_tmp;
        for (int i = 0; i < iter; i++) {
            for (int z = 0; z < D; z++)
                for (int y = 0; y < H; y++)
                    for (int x = 0; x < W; x++) {
                        const size_t c = id(x, y, z);
                        if (rho[c] <= T(0.5f) || boundary[c] >= T(0.5f)) { tmp[c] = p[c]; continue; }
                        T sum = T(0.0f);
                        int nCount = 0;
                        auto nb = [&](int nx, int ny, int nz) {
                            if (nx < 0 || nx >= W || ny < 0 || ny >= H || nz < 0 || nz >= D) return;
                            if (nx == 0 || nx == W - 1 || ny == 0 || ny == H - 1 || nz == 0 || nz == D - 1) return;
                            const size_t q = id(nx, ny, nz);
                            if (boundary[q] > T(0.5f)) return;
                            nCount++;
                            if (rho[q] > T(0.5f)) sum += p[q];
                        };
                        nb(x - 1, y, z); nb(x + 1, y, z);
                        nb(x, y - 1, z); nb(x, y + 1, z);
                        nb(x, y, z - 1); nb(x, y, z + 1);
                        tmp[c] = nCount > 0 ? (sum - divergence[c]) / (T)nCount : T(0.0f);
                    }
            std::swap(p, tmp);  // the reference swaps contents (copy-based std::swap)
        }
        return 0;
    }

    int stage(int s, ora_stats* st) {
        switch (s) {
            case ST_DECAY: vel_decay(); return 0;
            case ST_EXTRAP: vel_extrapolate(); return 0;
            case ST_SHARPEN: rho_sharpen(); return 0;
            case ST_RHO_ADVECT: rho_advect(); return 0;
            case ST_VEL_ADVECT: vel_advect(); return 0;
            case ST_FORCE: add_force(); return 0;
            case ST_PROJECT: return project(st);
        }
        return -1;
    }
    // ---- ISimulator::update (ISimulator.h:12-38) ----
    int step(ora_stats* st) {
        int status = 0;
        if (st) { st->status = 0; st->iterations = 0; st->err = 0; }
        for (int s = 0; s <= ST_PROJECT; s++) {
            if (P.disable_mask & (1u << s)) continue;
            const int rc = stage(s, st);
            if (s == ST_PROJECT) status = rc;
        }
        if (st) st->status = status;
        return status;
    }
};

// ========================================================================================
// 2D
// ========================================================================================
template <class T> struct Sim2D {
    using Cell = typename std::conditional<std::is_same<T, float>::value, AdvCell, AdvCellD>::type;
    int W, H;
    size_t N;
    ora_params P;
    std::vector<V2<T>> vel, tmpVec, forces;
    std::vector<T> rho, tmpScal, boundary, gamma, beta, divergence, pressure;
    std::vector<T> cg_t, cg_tmp, cg_r, cg_d, jac_tmp;
    std::map<unsigned int, Cell> A;  // Grid2D.h: std::map keyed by 32-bit i*N+j
    int solve_steps = 0;
    bool paused = false;

    Sim2D(int w, int h) : W(w), H(h) {
        N = (size_t)w * h;
        vel.assign(N, V2<T>());
        tmpVec.assign(N, V2<T>());
        forces.assign(N, V2<T>());
        rho.assign(N, T(0));
        tmpScal.assign(N, T(0));
        boundary.assign(N, T(0));
        gamma.assign(N, T(1));
        beta.assign(N, T(0));
        divergence.assign(N, T(0));
        pressure.assign(N, T(0));
        cg_t.assign(N, T(0)); cg_tmp.assign(N, T(0)); cg_r.assign(N, T(0)); cg_d.assign(N, T(0));
        jac_tmp.assign(N, T(0));
        // Grid2D.cpp:50-60 (the INIT_SHAPE fill is replaced by the caller's scene)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++)
                if (x == 0 || y == 0 || x == W - 1 || y == H - 1) {
                    rho[id(x, y)] = T(0);
                    boundary[id(x, y)] = T(1);
                }
        std::memset(&P, 0, sizeof(P));
    }
    inline size_t id(size_t x, size_t y) const { return y * (size_t)W + x; }

    // ScalField2D::interpolate (ScalField2D.cpp:20-69), including the +1 shift at :21
    T sinterp(const std::vector<T>& f, V2<T> pos) const {
        pos = pos + V2<T>(T(1.f));
        const int x = (int)pos.x, y = (int)pos.y;
        int ix1 = x, ix2 = x + 1;
        if (pos.x - (T)x < T(0.5f)) { ix1 = x - 1; ix2 = x; }
        int iy1 = y, iy2 = y + 1;
        if (pos.y - (T)y < T(0.5f)) { iy1 = y - 1; iy2 = y; }
        ix1 = clampv(ix1, 0, W - 1); ix2 = clampv(ix2, 0, W - 1);
        iy1 = clampv(iy1, 0, H - 1); iy2 = clampv(iy2, 0, H - 1);
        const T v[4] = {f[id(ix1, iy1)], f[id(ix1, iy2)], f[id(ix2, iy1)], f[id(ix2, iy2)]};
        const T x1 = (T)ix1 + T(0.5f), x2 = (T)ix2 + T(0.5f);
        const T y1 = (T)iy1 + T(0.5f), y2 = (T)iy2 + T(0.5f);
        const T dx2x1 = x2 - x1;
        T fxy1, fxy2;
        if (dx2x1 < 1e-9) { fxy1 = v[1]; fxy2 = v[3]; }
        else {
            fxy1 = ((x2 - pos.x) / dx2x1) * v[0] + ((pos.x - x1) / dx2x1) * v[2];
            fxy2 = ((x2 - pos.x) / dx2x1) * v[1] + ((pos.x - x1) / dx2x1) * v[3];
        }
        const T dy2y1 = y2 - y1;
        T fxy;
        if (dy2y1 < 1e-9) fxy = fxy1;
        else fxy = ((y2 - pos.y) / dy2y1) * fxy1 + ((pos.y - y1) / dy2y1) * fxy2;
        return fxy;
    }
    // VecField2D::bilinear_interpolation / interpolate (VecField2D.cpp:21-64)
    static V2<T> stagOff(int comp) { return comp == 0 ? V2<T>(T(-1.0f), T(-0.5f)) : V2<T>(T(-0.5f), T(-1.0f)); }
    T bilin(const std::vector<V2<T>>& f, V2<T> pos, int lx, int ly, int comp) const {
        const V2<T> off = stagOff(comp);
        const T x0 = (T)lx + off.x, y0 = (T)ly + off.y;
        const T tx = pos.x - x0, ty = pos.y - y0;
        const T v00 = f[id(lx, ly)][comp], v01 = f[id(lx, ly + 1)][comp];
        const T v10 = f[id(lx + 1, ly)][comp], v11 = f[id(lx + 1, ly + 1)][comp];
        const T vLow = v00 * (T(1.0f) - ty) + v01 * ty;
        const T vHigh = v10 * (T(1.0f) - ty) + v11 * ty;
        return vLow * (T(1.0f) - tx) + vHigh * tx;
    }
    static int lowIndex(T coord, T offset) { return (int)(std::floor(coord - offset) + T(1.0e-10f)); }
    V2<T> vinterp(const std::vector<V2<T>>& f, V2<T> pos) const {
        V2<T> r;
        for (int c = 0; c < 2; c++) {
            const V2<T> o = stagOff(c);
            r[c] = bilin(f, pos, lowIndex(pos.x, o.x), lowIndex(pos.y, o.y), c);
        }
        return r;
    }
    // Grid2D::traceParticle (Grid2D.cpp:84-136)
    bool inWall(V2<T> p) const {
        const T Wf = (T)(float)W, Hf = (T)(float)H;
        if (p.x < T(1.f) || p.x > Wf - T(2.f) || p.y < T(1.f) || p.y > Hf - T(2.f)) return true;
        const int ix = clampv((int)p.x, 0, W - 1), iy = clampv((int)p.y, 0, H - 1);
        return boundary[id(ix, iy)] > T(0.5f);
    }
    V2<T> traceParticle(V2<T> start, T dt, int substeps) const {
        const T Wf = (T)(float)W, Hf = (T)(float)H;
        auto clampP = [&](V2<T> p) { return V2<T>(clampv(p.x, T(1.f), Wf - T(2.f)), clampv(p.y, T(1.f), Hf - T(2.f))); };
        auto sampleVel = [&](V2<T> p) { return vinterp(vel, p - V2<T>(T(0.5f))); };
        if (inWall(start)) return start;
        const T subDt = dt / (T)substeps;
        V2<T> pos = start;
        for (int s = 0; s < substeps; s++) {
            const V2<T> k1 = sampleVel(pos);
            const V2<T> mid = clampP(pos + k1 * (subDt * T(0.5f)));
            const V2<T> k2 = sampleVel(mid);
            const V2<T> next = pos + k2 * subDt;
            if (inWall(next)) {
                V2<T> safe = pos, unsafe = next;
                for (int b = 0; b < 8; b++) {
                    const V2<T> hw = (safe + unsafe) * T(0.5f);
                    if (inWall(hw)) unsafe = hw;
                    else safe = hw;
                }
                return safe;
            }
            pos = next;
        }
        return pos;
    }

    void vel_decay() {  // CPUSimulator2D.cpp:65-70
        for (int y = 1; y < H - 1; y++)
            for (int x = 1; x < W - 1; x++)
                if (rho[id(x, y)] < 0.5) vel[id(x, y)] = vel[id(x, y)] * (T)P.vel_decay;
    }
    void vel_extrapolate() {  // CPUSimulator2D.cpp:72-120
        std::vector<T> validA(N, T(0)), validB(N, T(0));
        for (int y = 1; y < H - 1; y++)
            for (int x = 1; x < W - 1; x++)
                validA[id(x, y)] = (rho[id(x, y)] >= T(0.5f) && boundary[id(x, y)] != T(1.f)) ? T(1) : T(0);
        auto pass = [&](std::vector<V2<T>>& sv, std::vector<T>& sval, std::vector<V2<T>>& dv, std::vector<T>& dval) {
            for (int y = 1; y < H - 1; y++)
                for (int x = 1; x < W - 1; x++) {
                    const size_t c = id(x, y);
                    if (boundary[c] == T(1.f)) { dv[c] = sv[c]; dval[c] = 0; continue; }
                    if (sval[c] > T(0.5f)) { dv[c] = sv[c]; dval[c] = 1; continue; }
                    V2<T> sum(T(0));
                    T n = 0;
                    auto sample = [&](int nx, int ny) {
                        const size_t q = id(nx, ny);
                        if (sval[q] > T(0.5f)) { sum = sum + sv[q]; n++; }
                    };
                    sample(x - 1, y); sample(x + 1, y); sample(x, y - 1); sample(x, y + 1);
                    dv[c] = (n > 0) ? sum / n : sv[c];
                    dval[c] = (n > 0) ? T(1) : T(0);
                }
        };
        for (int i = 0; i < 3; i++) { pass(vel, validA, tmpVec, validB); pass(tmpVec, validB, vel, validA); }
    }
    void addA(unsigned int i, unsigned int j, T val) {  // Grid2D.cpp: std::map, 32-bit key
        const unsigned int k = i * (unsigned int)N + j;
        auto it = A.find(k);
        if (it != A.end()) it->second.val += val;
        else A[k] = Cell{i, j, val};
    }
    bool cornerWeights(V2<T> p, int& px, int& py, T m[4], T& sum) const {
        px = (int)std::floor(p.x); py = (int)std::floor(p.y);
        const T dx0 = 1 - (p.x - px), dy0 = 1 - (p.y - py);
        const T dx1 = 1 - (px + 1 - p.x), dy1 = 1 - (py + 1 - p.y);
        m[0] = m[1] = m[2] = m[3] = 0;
        if (boundary[id(px + 0, py + 0)] != T(1.f)) m[0] = dx0 * dy0;
        if (boundary[id(px + 1, py + 0)] != T(1.f)) m[1] = dx1 * dy0;
        if (boundary[id(px + 0, py + 1)] != T(1.f)) m[2] = dx0 * dy1;
        if (boundary[id(px + 1, py + 1)] != T(1.f)) m[3] = dx1 * dy1;
        sum = m[0] + m[1] + m[2] + m[3];
        return !(sum == 0);
    }
    void rho_advect() {  // CPUSimulator2D.cpp:122-318
        const T dt = (T)P.dt;
        auto idx = [&](int x, int y) -> unsigned int { return (unsigned int)y * (unsigned int)W + (unsigned int)x; };
        std::vector<T> adv(N, T(0));
        for (int y = 1; y < H - 1; y++)
            for (int x = 1; x < W - 1; x++) {
                if (boundary[id(x, y)] == T(1.f)) continue;
                adv[id(x, y)] = sinterp(gamma, traceParticle(V2<T>((T)x, (T)y), -dt, 4));
            }
        gamma = adv;
        A.clear();
        for (int y = 1; y < H - 1; y++)
            for (int x = 1; x < W - 1; x++) {
                const V2<T> prev = traceParticle(V2<T>((T)x, (T)y), -dt, 4);
                int px, py; T m[4], sum;
                if (!cornerWeights(prev, px, py, m, sum)) continue;
                addA(idx(x, y), idx(px, py), m[0] / sum);
                addA(idx(x, y), idx(px + 1, py), m[1] / sum);
                addA(idx(x, y), idx(px, py + 1), m[2] / sum);
                addA(idx(x, y), idx(px + 1, py + 1), m[3] / sum);
            }
        for (auto& kv : A) if (gamma[kv.second.i] > T(1.e-5f)) kv.second.val /= gamma[kv.second.i];
        std::fill(beta.begin(), beta.end(), T(0));
        for (auto& kv : A) beta[kv.second.j] += kv.second.val;
        for (int y = 1; y < H - 1; y++)
            for (int x = 1; x < W - 1; x++) {
                const unsigned int j = idx(x, y);
                if (beta[j] >= 1) continue;
                const V2<T> next = traceParticle(V2<T>((T)x, (T)y), +dt, 4);
                int px, py; T m[4], sum;
                if (!cornerWeights(next, px, py, m, sum)) continue;
                addA(idx(px, py), j, (1 - beta[j]) * (m[0] / sum));
                addA(idx(px + 1, py), j, (1 - beta[j]) * (m[1] / sum));
                addA(idx(px, py + 1), j, (1 - beta[j]) * (m[2] / sum));
                addA(idx(px + 1, py + 1), j, (1 - beta[j]) * (m[3] / sum));
            }
        std::fill(gamma.begin(), gamma.end(), T(0));
        for (auto& kv : A) gamma[kv.second.i] += kv.second.val;
        for (auto& kv : A) if (gamma[kv.second.i] > T(1.e-5f)) kv.second.val /= gamma[kv.second.i];
        std::fill(beta.begin(), beta.end(), T(0));
        for (auto& kv : A) beta[kv.second.j] += kv.second.val;
        for (auto& kv : A) if (beta[kv.second.j] > T(1.e-5f)) kv.second.val /= beta[kv.second.j];
        std::fill(gamma.begin(), gamma.end(), T(0));
        for (auto& kv : A) gamma[kv.second.i] += kv.second.val;
        std::fill(tmpScal.begin(), tmpScal.end(), T(0));
        for (auto& kv : A) tmpScal[kv.second.i] += kv.second.val * rho[kv.second.j];
        for (int y = 1; y < H - 1; y++)
            for (int x = 1; x < W - 1; x++) rho[id(x, y)] = tmpScal[id(x, y)];
        for (int it = 0; it < 7; it++) {
            for (int y = 1; y < H - 1; y++)
                for (int x = 2; x < W - 1; x++) diffusePair(id(x - 1, y), id(x, y));
            for (int x = 1; x < W - 1; x++)
                for (int y = 2; y < H - 1; y++) diffusePair(id(x, y - 1), id(x, y));
        }
    }
    void diffusePair(size_t jj, size_t ii) {
        const T g_j = gamma[jj], g_i = gamma[ii];
        if (boundary[ii] == T(1.f) || boundary[jj] == T(1.f)) return;
        if (g_j < T(1e-5f) || g_i < T(1e-5f)) return;
        if (g_j == g_i) return;
        const T avg = (g_j + g_i) / T(2.f);
        if (g_j > g_i) {
            const T dif = rho[jj] * (g_j - g_i) / (T(2.f) * g_j);
            rho[jj] -= dif; rho[ii] += dif;
        } else {
            const T dif = rho[ii] * (g_i - g_j) / (T(2.f) * g_i);
            rho[ii] -= dif; rho[jj] += dif;
        }
        gamma[jj] = avg; gamma[ii] = avg;
    }
    // Delta rho of every interior cell (CPUSimulator2D.cpp:320-390), before any is applied.
    void sharpen_delta(std::vector<T>& del_rho) const {
        const T delT = (T)(P.dt * 3);
        del_rho.assign(N, T(0));
        std::vector<T> mi(N, 0), md(N, 0), w(N, 0);
        std::vector<V2<T>> dp(N), dm(N);
        for (int y = 1; y < H - 1; y++)
            for (int x = 1; x < W - 1; x++) {
                const size_t c = id(x, y);
                dp[c].x = -(rho[c] - rho[id(x - 1, y)]) * delT;
                dm[c].x = -(rho[id(x + 1, y)] - rho[c]) * delT;
                dp[c].y = -(rho[c] - rho[id(x, y - 1)]) * delT;
                dm[c].y = -(rho[id(x, y + 1)] - rho[c]) * delT;
            }
        auto sq = [](T v) -> double { return std::pow((double)v, 2.0); };
        for (int y = 1; y < H - 1; y++)
            for (int x = 1; x < W - 1; x++) {
                const size_t c = id(x, y);
                mi[c] = (T)std::sqrt(std::max(sq(std::max(dp[c].x, T(0))), sq(std::min(dm[c].x, T(0)))) +
                                     std::max(sq(std::max(dp[c].y, T(0))), sq(std::min(dm[c].y, T(0)))));
                md[c] = (T)std::sqrt(std::max(sq(std::min(dp[c].x, T(0))), sq(std::max(dm[c].x, T(0)))) +
                                     std::max(sq(std::min(dp[c].y, T(0))), sq(std::max(dm[c].y, T(0)))));
            }
        for (int y = 1; y < H - 1; y++)
            for (int x = 1; x < W - 1; x++) {
                const size_t c = id(x, y);
                T mx = -INFINITY;
                mx = std::max(mx, std::abs(rho[c] - rho[id(x - 1, y)]));
                mx = std::max(mx, std::abs(rho[c] - rho[id(x + 1, y)]));
                mx = std::max(mx, std::abs(rho[c] - rho[id(x, y - 1)]));
                mx = std::max(mx, std::abs(rho[c] - rho[id(x, y + 1)]));
                w[c] = std::pow(rho[c] - T(0.5f), T(3.f)) * (1 - std::min(T(1.f), mx / T(0.2f)));
            }
        for (int y = 1; y < H - 1; y++)
            for (int x = 1; x < W - 1; x++) {
                const size_t c = id(x, y);
                del_rho[c] = (w[c] >= 0) ? w[c] * mi[c] : w[c] * md[c];
                if (rho[c] + del_rho[c] < 0 || rho[c] < 1e-5) del_rho[c] = -rho[c];
                else if (rho[c] > T(0.5f)) del_rho[c] = 0;
            }
    }
    void rho_sharpen() {  // CPUSimulator2D.cpp:320-402
        std::vector<T> del_rho;
        sharpen_delta(del_rho);
        if (P.sharpen_mode == 0) {
            for (int y = 1; y < H - 1; y++)
                for (int x = 1; x < W - 1; x++) {
                    const size_t c = id(x, y);
                    const T dr = del_rho[c];
                    if (std::abs(dr) < T(1e-6f)) continue;
                    rho[c] += dr;
                    rho[c] = std::max(rho[c], T(0.f));
                    const V2<T> p = trace_along_field(V2<T>((T)x, (T)y), (T)P.rho_d, 10);
                    scatter_values(p, -dr);
                }
        } else {
            // apply-all-then-trace (see the 3D restatement): scatters into a double buffer
            for (int y = 1; y < H - 1; y++)
                for (int x = 1; x < W - 1; x++) {
                    const size_t c = id(x, y);
                    const T dr = del_rho[c];
                    if (std::abs(dr) < T(1e-6f)) continue;
                    rho[c] += dr;
                    rho[c] = std::max(rho[c], T(0.f));
                }
            std::vector<double> acc(N, 0.0);
            for (int y = 1; y < H - 1; y++)
                for (int x = 1; x < W - 1; x++) {
                    const size_t c = id(x, y);
                    const T dr = del_rho[c];
                    if (std::abs(dr) < T(1e-6f)) continue;
                    const V2<T> p = trace_along_field(V2<T>((T)x, (T)y), (T)P.rho_d, 10);
                    scatter_acc(acc, p, -dr);
                }
            for (size_t c = 0; c < N; c++)
                if (acc[c] != 0.0) rho[c] = (T)((double)rho[c] + acc[c]);
        }
    }
    V2<T> trace_along_field(V2<T> start, T Dm, int substeps) const {  // CPUSimulator2D.cpp:526-562
        const T stepSize = Dm / (T)substeps;
        V2<T> pos = start;
        for (int s = 0; s < substeps; s++) {
            if (sinterp(rho, pos) >= T(0.5f)) break;
            const int ix = clampv((int)pos.x, 1, W - 2), iy = clampv((int)pos.y, 1, H - 2);
            const V2<T> grad(rho[id(ix + 1, iy)] - rho[id(ix - 1, iy)], rho[id(ix, iy + 1)] - rho[id(ix, iy - 1)]);
            const T len = std::sqrt(grad.x * grad.x + grad.y * grad.y);
            if (len < T(1e-6f)) break;
            const V2<T> dir = grad / len;
            V2<T> next = pos + dir * stepSize;
            next.x = clampv(next.x, T(1.f), (T)(float)W - T(2.f));
            next.y = clampv(next.y, T(1.f), (T)(float)H - T(2.f));
            if (boundary[id((size_t)next.x, (size_t)next.y)] > T(0.5f)) break;
            pos = next;
        }
        return pos;
    }
    // scatter_values' weights, accumulated in double (the parallel restatement's order)
    void scatter_acc(std::vector<double>& acc, V2<T> pos, T value) const {
        const int x0 = clampv((int)pos.x, 0, W - 1), y0 = clampv((int)pos.y, 0, H - 1);
        const int x1 = clampv((int)pos.x + 1, 0, W - 1), y1 = clampv((int)pos.y + 1, 0, H - 1);
        const T fx = pos.x - (T)(int)pos.x, fy = pos.y - (T)(int)pos.y;
        T w00 = (T(1.f) - fx) * (T(1.f) - fy), w10 = fx * (T(1.f) - fy), w01 = (T(1.f) - fx) * fy, w11 = fx * fy;
        if (boundary[id(x0, y0)] > T(0.5f)) w00 = 0;
        if (boundary[id(x1, y0)] > T(0.5f)) w10 = 0;
        if (boundary[id(x0, y1)] > T(0.5f)) w01 = 0;
        if (boundary[id(x1, y1)] > T(0.5f)) w11 = 0;
        const T total = w00 + w10 + w01 + w11;
        if (total < T(1e-6f)) return;
        const T inv = T(1.f) / total;
        acc[id(x0, y0)] += (double)(value * w00 * inv);
        acc[id(x1, y0)] += (double)(value * w10 * inv);
        acc[id(x0, y1)] += (double)(value * w01 * inv);
        acc[id(x1, y1)] += (double)(value * w11 * inv);
    }
    void scatter_values(V2<T> pos, T value) {  // CPUSimulator2D.cpp:564-595
        const int x0 = clampv((int)pos.x, 0, W - 1), y0 = clampv((int)pos.y, 0, H - 1);
        const int x1 = clampv((int)pos.x + 1, 0, W - 1), y1 = clampv((int)pos.y + 1, 0, H - 1);
        const T fx = pos.x - (T)(int)pos.x, fy = pos.y - (T)(int)pos.y;
        T w00 = (T(1.f) - fx) * (T(1.f) - fy), w10 = fx * (T(1.f) - fy), w01 = (T(1.f) - fx) * fy, w11 = fx * fy;
        if (boundary[id(x0, y0)] > T(0.5f)) w00 = 0;
        if (boundary[id(x1, y0)] > T(0.5f)) w10 = 0;
        if (boundary[id(x0, y1)] > T(0.5f)) w01 = 0;
        if (boundary[id(x1, y1)] > T(0.5f)) w11 = 0;
        const T total = w00 + w10 + w01 + w11;
        if (total < T(1e-6f)) return;
        const T inv = T(1.f) / total;
        rho[id(x0, y0)] += value * w00 * inv;
        rho[id(x1, y0)] += value * w10 * inv;
        rho[id(x0, y1)] += value * w01 * inv;
        rho[id(x1, y1)] += value * w11 * inv;
    }
    void add_force() {  // CPUSimulator2D.cpp:404-425 (face-wise)
        const T dt = (T)P.dt;
        const T gdt = (T)(P.gravity_force * P.dt);
        for (int y = 1; y < H - 1; y++)
            for (int x = 1; x < W - 1; x++) {
                const size_t c = id(x, y);
                if (boundary[c] != T(1.f) || boundary[id(x - 1, y)] != T(1.f)) vel[c].x += forces[c].x * dt;
                else vel[c].x = 0;
                if (boundary[c] != T(1.f) || boundary[id(x, y - 1)] != T(1.f)) {
                    vel[c].y += forces[c].y * dt;
                    if (P.gravity_enabled) vel[c].y -= gdt;
                } else vel[c].y = 0;
            }
    }
    void vel_advect() {  // CPUSimulator2D.cpp:427-448
        const T dt = (T)P.dt;
        for (int y = 1; y < H - 1; y++)
            for (int x = 1; x < W - 1; x++)
                tmpVec[id(x, y)] = vinterp(vel, traceParticle(V2<T>((T)x, (T)y), -dt, 4) - V2<T>(T(0.5f)));
        for (int y = 1; y < H - 1; y++)
            for (int x = 1; x < W - 1; x++) {
                const size_t c = id(x, y);
                if (x == 1 || boundary[c] == T(1.f) || boundary[id(x - 1, y)] == T(1.f)) vel[c].x = tmpVec[c].x;
                else vel[c].x = T(0.5f) * (tmpVec[c].x + tmpVec[id(x - 1, y)].x);
                if (y == 1 || boundary[c] == T(1.f) || boundary[id(x, y - 1)] == T(1.f)) vel[c].y = tmpVec[c].y;
                else vel[c].y = T(0.5f) * (tmpVec[c].y + tmpVec[id(x, y - 1)].y);
            }
    }
    void enforce_wall_boundaries() {  // CPUSimulator2D.cpp:597-607 (two-sided)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                if (boundary[id(x, y)] == T(1.f) || (x > 0 && boundary[id(x - 1, y)] == T(1.f))) vel[id(x, y)].x = 0;
                if (boundary[id(x, y)] == T(1.f) || (y > 0 && boundary[id(x, y - 1)] == T(1.f))) vel[id(x, y)].y = 0;
            }
    }
    void compute_divergence() {
        const T liq = (T)P.liquid_threshold, lp = (T)P.lambda_plus, lm = (T)P.lambda_minus, eta = (T)P.eta;
        for (int y = 1; y < H - 1; y++)
            for (int x = 1; x < W - 1; x++) {
                const size_t c = id(x, y);
                if (rho[c] < liq) continue;
                if (boundary[c] == T(1)) continue;
                const T dx = vel[id(x + 1, y)].x - vel[c].x;
                const T dy = vel[id(x, y + 1)].y - vel[c].y;
                divergence[c] = dx + dy;
                if (rho[c] > T(1.f)) divergence[c] -= std::min(lm * (rho[c] - T(1.f)), eta);
                if (rho[c] < T(1.f)) divergence[c] += std::min(lp * (T(1.f) - rho[c]), eta);
            }
    }
    // face-wise gradient (CPUSimulator2D.cpp:491-501): a face touching a solid is set to 0
    void subtract_pressure_gradient() {
        for (int y = 1; y < H - 1; y++)
            for (int x = 1; x < W - 1; x++) {
                const size_t c = id(x, y);
                if (boundary[c] != T(1.f) && boundary[id(x - 1, y)] != T(1.f)) vel[c].x -= (pressure[c] - pressure[id(x - 1, y)]);
                else vel[c].x = 0;
                if (boundary[c] != T(1.f) && boundary[id(x, y - 1)] != T(1.f)) vel[c].y -= (pressure[c] - pressure[id(x, y - 1)]);
                else vel[c].y = 0;
            }
    }
    int project(ora_stats* st) {  // CPUSimulator2D.cpp:450-524
        std::fill(pressure.begin(), pressure.end(), T(0));
        std::fill(divergence.begin(), divergence.end(), T(0));
        enforce_wall_boundaries();
        compute_divergence();
        const T err = (P.solver_type == 3) ? jacobi_solve() : cg_solve();
        if (st) { st->err = (double)err; st->iterations = solve_steps; }
        if (err > (T)P.solver_threshold) { paused = true; return 1; }
        subtract_pressure_gradient();
        if (P.re_calc_div) compute_divergence();
        enforce_wall_boundaries();
        return std::isfinite((double)err) ? 0 : 2;
    }
    void mat_mul(std::vector<T>& out, const std::vector<T>& xf) const {  // CGSolver.cpp:85-111
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                const size_t c = id(x, y);
                if (rho[c] <= T(0.5f) || boundary[c] == T(1)) { out[c] = xf[c]; continue; }
                T sum = T(0.0f);
                int nCount = 0;
                auto nb = [&](int nx, int ny) {
                    if (nx < 0 || nx >= W || ny < 0 || ny >= H) return;
                    if (nx == 0 || nx == W - 1 || ny == 0 || ny == H - 1) return;
                    if (boundary[id(nx, ny)] > T(0.5f)) return;
                    nCount++;
                    if (rho[id(nx, ny)] > T(0.5f)) sum += xf[id(nx, ny)];
                };
                nb(x - 1, y); nb(x + 1, y); nb(x, y - 1); nb(x, y + 1);
                out[c] = sum - (T)nCount * xf[c];
            }
    }
    static T dot(const std::vector<T>& a, const std::vector<T>& b) {
        T d = 0;
        for (size_t i = 0; i < b.size(); i++) d += a[i] * b[i];
        return d;
    }
    T cg_solve() {  // CGSolver.cpp:32-83
        const T thr = (T)P.solver_threshold;
        solve_steps = 0;
        auto& t = cg_t; std::fill(t.begin(), t.end(), T(0));
        auto& r = cg_r; r = divergence;
        pressure = divergence;
        mat_mul(t, pressure);
        for (size_t i = 0; i < N; i++) r[i] -= t[i];
        auto& d = cg_d; d = r;
        T sq = dot(r, r);
        if (sq < thr) return sq;
        int i = 0;
        for (i = 0; i < P.max_steps; i++) {
            mat_mul(t, d);
            const T alpha = sq / dot(d, t);
            for (size_t k = 0; k < N; k++) pressure[k] += alpha * d[k];
            if (i & 0x3F) { for (size_t k = 0; k < N; k++) r[k] += -alpha * t[k]; }
            else { mat_mul(t, pressure); r = divergence; for (size_t k = 0; k < N; k++) r[k] -= t[k]; }
            const T old = sq;
            sq = dot(r, r);
            if (sq < thr) break;
            const T b_ = sq / old;
            for (size_t k = 0; k < N; k++) d[k] *= b_;
            for (size_t k = 0; k < N; k++) d[k] += r[k];
        }
        solve_steps = i + 1;
        return sq;
    }
    T jacobi_solve() {  // JacobiSolver.cpp:18-70
        std::vector<T>& p = pressure;
        std::vector<T>& tmp = jac_tmp;
        solve_steps = P.max_steps / 2;
        for (int i = 0; i < P.max_steps / 2; i++) {
            for (int y = 0; y < H; y++)
                for (int x = 0; x < W; x++) {
                    const size_t c = id(x, y);
                    if (rho[c] <= T(0.5f) || boundary[c] >= T(0.5f)) { tmp[c] = p[c]; continue; }
                    T sum = T(0.0f);
                    int nCount = 0;
                    auto nb = [&](int nx, int ny) {
                        if (nx < 0 || nx >= W || ny < 0 || ny >= H) return;
                        if (nx == 0 || nx == W - 1 || ny == 0 || ny == H - 1) return;
                        if (boundary[id(nx, ny)] > T(0.5f)) return;
                        nCount++;
                        if (rho[id(nx, ny)] > T(0.5f)) sum += p[id(nx, ny)];
                    };
                    nb(x - 1, y); nb(x + 1, y); nb(x, y - 1); nb(x, y + 1);
                    tmp[c] = nCount > 0 ? (sum - divergence[c]) / T(nCount) : T(0.0f);
                }
            std::swap(p, tmp);
        }
        return 0;
    }
    int stage(int s, ora_stats* st) {
        switch (s) {
            case ST_DECAY: vel_decay(); return 0;
            case ST_EXTRAP: vel_extrapolate(); return 0;
            case ST_SHARPEN: rho_sharpen(); return 0;
            case ST_RHO_ADVECT: rho_advect(); return 0;
            case ST_VEL_ADVECT: vel_advect(); return 0;
            case ST_FORCE: add_force(); return 0;
            case ST_PROJECT: return project(st);
        }
        return -1;
    }
    int step(ora_stats* st) {
        int status = 0;
        if (st) { st->status = 0; st->iterations = 0; st->err = 0; }
        for (int s = 0; s <= ST_PROJECT; s++) {
            if (P.disable_mask & (1u << s)) continue;
            const int rc = stage(s, st);
            if (s == ST_PROJECT) status = rc;
        }
        if (st) st->status = status;
        return status;
    }
};

// ========================================================================================
// Seeded scene inputs: counter-based generator and primitive rasterisation.
// ========================================================================================
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

// ========================================================================================
// C API (ctypes)
// ========================================================================================
extern "C" {

typedef struct {
    int32_t kind;      // 0 box, 1 sphere
    int32_t solid;     // 0 fluid (rho=1), 1 solid (boundary=1, rho=0)
    double a[3];       // box lo / sphere centre (unit coordinates of the interior)
    double b[3];       // box hi / sphere radius in b[0]
} ora_prim;

void* ora3_create(int w, int h, int d, int dbl) {
    if (dbl) return new Sim3D<double>(w, h, d);
    return new Sim3D<float>(w, h, d);
}
void ora3_destroy(void* s, int dbl) {
    if (dbl) delete (Sim3D<double>*)s;
    else delete (Sim3D<float>*)s;
}
void ora3_set_params(void* s, int dbl, const ora_params* p) {
    if (dbl) ((Sim3D<double>*)s)->P = *p;
    else ((Sim3D<float>*)s)->P = *p;
}
void* ora3_field(void* s, int dbl, int f) {
    auto pick = [&](auto* S) -> void* {
        switch (f) {
            case F_VEL: return S->vel.data();
            case F_RHO: return S->rho.data();
            case F_GAMMA: return S->gamma.data();
            case F_BETA: return S->beta.data();
            case F_DIV: return S->divergence.data();
            case F_PRESSURE: return S->pressure.data();
            case F_FORCES: return S->forces.data();
            case F_BOUNDARY: return S->boundary.data();
            case F_TMPSCAL: return S->tmpScal.data();
            case F_TMPVEC: return S->tmpVec.data();
        }
        return nullptr;
    };
    if (dbl) return pick((Sim3D<double>*)s);
    return pick((Sim3D<float>*)s);
}
int ora3_step(void* s, int dbl, ora_stats* st) {
    if (dbl) return ((Sim3D<double>*)s)->step(st);
    return ((Sim3D<float>*)s)->step(st);
}
int ora3_stage(void* s, int dbl, int stage, ora_stats* st) {
    const int rc = dbl ? ((Sim3D<double>*)s)->stage(stage, st) : ((Sim3D<float>*)s)->stage(stage, st);
    if (st) st->status = rc;
    return rc;
}
// Pressure solve on the current divergence (pressure cleared first, as project() does).
int ora3_solve(void* s, int dbl, ora_stats* st) {
    auto run = [&](auto* S) {
        std::fill(S->pressure.begin(), S->pressure.end(), 0);
        S->solve_steps = 0;
        const double err = (S->P.solver_type == 3) ? (double)S->jacobi_solve() : (double)S->cg_solve();
        if (st) { st->err = err; st->iterations = S->solve_steps; st->status = err > S->P.solver_threshold ? 1 : 0; }
        return st ? st->status : 0;
    };
    if (dbl) return run((Sim3D<double>*)s);
    return run((Sim3D<float>*)s);
}
void ora3_divergence(void* s, int dbl) {
    auto run = [&](auto* S) {
        std::fill(S->divergence.begin(), S->divergence.end(), 0);
        S->enforce_wall_boundaries();
        S->compute_divergence();
    };
    if (dbl) run((Sim3D<double>*)s);
    else run((Sim3D<float>*)s);
}
void ora3_enforce_walls(void* s, int dbl) {
    if (dbl) ((Sim3D<double>*)s)->enforce_wall_boundaries();
    else ((Sim3D<float>*)s)->enforce_wall_boundaries();
}
void ora3_matmul(void* s, int dbl, const void* x, void* y) {
    auto run = [&](auto* S, auto tag) {
        using T = decltype(tag);
        std::vector<T> xv((const T*)x, (const T*)x + S->N), yv(S->N);
        S->mat_mul(yv, xv);
        std::memcpy(y, yv.data(), S->N * sizeof(T));
    };
    if (dbl) run((Sim3D<double>*)s, 0.0);
    else run((Sim3D<float>*)s, 0.0f);
}
// One backward/forward trace (Grid3D::traceParticle) for tests of the trace.
void ora3_trace(void* s, int dbl, const double* start, double dt, int substeps, double* out) {
    auto run = [&](auto* S, auto tag) {
        using T = decltype(tag);
        const V3<T> r = S->traceParticle(V3<T>((T)start[0], (T)start[1], (T)start[2]), (T)dt, substeps);
        out[0] = r.x; out[1] = r.y; out[2] = r.z;
    };
    if (dbl) run((Sim3D<double>*)s, 0.0);
    else run((Sim3D<float>*)s, 0.0f);
}
void ora3_vinterp(void* s, int dbl, const double* pos, double* out) {
    auto run = [&](auto* S, auto tag) {
        using T = decltype(tag);
        const V3<T> r = S->vinterp(S->vel, V3<T>((T)pos[0], (T)pos[1], (T)pos[2]));
        out[0] = r.x; out[1] = r.y; out[2] = r.z;
    };
    if (dbl) run((Sim3D<double>*)s, 0.0);
    else run((Sim3D<float>*)s, 0.0f);
}
double ora3_sinterp(void* s, int dbl, int field, const double* pos) {
    auto run = [&](auto* S, auto tag) -> double {
        using T = decltype(tag);
        auto* f = (T*)ora3_field(s, dbl, field);
        std::vector<T> v(f, f + S->N);
        return (double)S->sinterp(v, V3<T>((T)pos[0], (T)pos[1], (T)pos[2]));
    };
    if (dbl) return run((Sim3D<double>*)s, 0.0);
    return run((Sim3D<float>*)s, 0.0f);
}
size_t ora3_num_entries(void* s, int dbl) {
    if (dbl) { auto* S = (Sim3D<double>*)s; return S->P.key64 ? S->A64.size() : S->A32.size(); }
    auto* S = (Sim3D<float>*)s;
    return S->P.key64 ? S->A64.size() : S->A32.size();
}

// Pieces of the step in isolation, for the oracle's own pins.
// project()'s gradient subtraction (CPUSimulator3D.cpp:550-563) on the current pressure.
void ora3_gradient(void* s, int dbl) {
    if (dbl) ((Sim3D<double>*)s)->subtract_pressure_gradient();
    else ((Sim3D<float>*)s)->subtract_pressure_gradient();
}
// rho_sharpen's delta rho (CPUSimulator3D.cpp:366-444) into tmpScal; rho is not changed.
void ora3_sharpen_delta(void* s, int dbl) {
    auto run = [&](auto* S) { S->sharpen_delta(S->tmpScal); };
    if (dbl) run((Sim3D<double>*)s);
    else run((Sim3D<float>*)s);
}
// trace_along_field (CPUSimulator3D.cpp:568-607) from start with the RHO_D parameter, 10 steps.
void ora3_trace_field(void* s, int dbl, const double* start, double* out) {
    auto run = [&](auto* S, auto tag) {
        using T = decltype(tag);
        const V3<T> r = S->trace_along_field(V3<T>((T)start[0], (T)start[1], (T)start[2]), (T)S->P.rho_d, 10);
        out[0] = r.x; out[1] = r.y; out[2] = r.z;
    };
    if (dbl) run((Sim3D<double>*)s, 0.0);
    else run((Sim3D<float>*)s, 0.0f);
}
// scatter_values (CPUSimulator3D.cpp:609-655) of value at pos into rho.
void ora3_scatter(void* s, int dbl, const double* pos, double value) {
    auto run = [&](auto* S, auto tag) {
        using T = decltype(tag);
        S->scatter_values(S->rho, V3<T>((T)pos[0], (T)pos[1], (T)pos[2]), (T)value);
    };
    if (dbl) run((Sim3D<double>*)s, 0.0);
    else run((Sim3D<float>*)s, 0.0f);
}

// FP64 solution of A p = divergence for the error statements of the full-size configurations:
// the same operator as CGSolverCPU3D::impl_mat_mul (CGSolver.cpp:467-497), conjugate gradients in
// FP64 with OpenMP-parallel loops (reduction order differs from the sequential solve only in the
// last bits of FP64), stopped at ||b - A p|| / ||b|| <= rel_tol (recomputed residual, checked every
// 50 iterations and at the end).  The error of the result is set by that final residual, not by
// the initial guess, so x0 (N doubles, or null for x0 = b) may be any guess.  Writes the
// FP64 instance's pressure; returns the iterations and the final relative residual.
int ora3_cg64(void* s, const double* x0, double rel_tol, int max_iter, int nthreads, double* out_rel) {
    auto* S = (Sim3D<double>*)s;
    const long long N = (long long)S->N;
    const int W = S->W, H = S->H, D = S->D;
    std::vector<int8_t> row((size_t)N);  // -1: identity row; else neighbour count
    std::vector<uint8_t> nbm((size_t)N);  // bit k: neighbour k counted and fluid (enters the sum)
    const long long off[6] = {-1, 1, -(long long)W, (long long)W, (long long)W * H, -(long long)W * H};
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long long c = 0; c < N; c++) {
        const int x = (int)(c % W), y = (int)((c / W) % H), z = (int)(c / ((long long)W * H));
        if (S->rho[c] <= 0.5 || S->boundary[c] == 1.0) { row[c] = -1; nbm[c] = 0; continue; }
        const int nx[6] = {x - 1, x + 1, x, x, x, x}, ny[6] = {y, y, y - 1, y + 1, y, y}, nz[6] = {z, z, z, z, z + 1, z - 1};
        int cnt = 0;
        uint8_t m = 0;
        for (int k = 0; k < 6; k++) {
            if (nx[k] < 0 || nx[k] >= W || ny[k] < 0 || ny[k] >= H || nz[k] < 0 || nz[k] >= D) continue;
            if (nx[k] == 0 || nx[k] == W - 1 || ny[k] == 0 || ny[k] == H - 1 || nz[k] == 0 || nz[k] == D - 1) continue;
            const long long q = c + off[k];
            if (S->boundary[q] > 0.5) continue;
            cnt++;
            if (S->rho[q] > 0.5) m |= (uint8_t)(1u << k);
        }
        row[c] = (int8_t)cnt;
        nbm[c] = m;
    }
    auto matmul = [&](std::vector<double>& out, const std::vector<double>& v) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (long long c = 0; c < N; c++) {
            if (row[c] < 0) { out[c] = v[c]; continue; }
            double sum = 0.0;
            for (int k = 0; k < 6; k++)
                if (nbm[c] & (1u << k)) sum += v[c + off[k]];
            out[c] = sum - (double)row[c] * v[c];
        }
    };
    auto dot = [&](const std::vector<double>& a, const std::vector<double>& b) {
        double d = 0.0;
#pragma omp parallel for num_threads(nthreads) reduction(+ : d) schedule(static)
        for (long long c = 0; c < N; c++) d += a[c] * b[c];
        return d;
    };
    const std::vector<double>& b = S->divergence;
    std::vector<double>& x = S->pressure;
    if (x0) x.assign(x0, x0 + N);
    else x = b;
    std::vector<double> r((size_t)N), d((size_t)N), t((size_t)N);
    const double bb = dot(b, b);
    auto residual = [&]() {
        matmul(t, x);
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (long long c = 0; c < N; c++) r[c] = b[c] - t[c];
        return dot(r, r);
    };
    double rr = residual();
    d = r;
    int it = 0;
    const double stop = rel_tol * rel_tol * bb;
    while (it < max_iter && rr > stop) {
        matmul(t, d);
        const double alpha = rr / dot(d, t);
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (long long c = 0; c < N; c++) {
            x[c] += alpha * d[c];
            r[c] -= alpha * t[c];
        }
        it++;
        double rn = dot(r, r);
        if (it % 50 == 0 || rn <= stop) {
            rn = residual();  // recomputed: the recurrence residual drifts in long solves
            if (rn <= stop) { rr = rn; break; }
        }
        const double beta = rn / rr;
        rr = rn;
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (long long c = 0; c < N; c++) d[c] = r[c] + beta * d[c];
    }
    rr = residual();
    if (out_rel) *out_rel = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
    return it;
}

// ---- 2D ----
void* ora2_create(int w, int h, int dbl) {
    if (dbl) return new Sim2D<double>(w, h);
    return new Sim2D<float>(w, h);
}
void ora2_destroy(void* s, int dbl) {
    if (dbl) delete (Sim2D<double>*)s;
    else delete (Sim2D<float>*)s;
}
void ora2_set_params(void* s, int dbl, const ora_params* p) {
    if (dbl) ((Sim2D<double>*)s)->P = *p;
    else ((Sim2D<float>*)s)->P = *p;
}
void* ora2_field(void* s, int dbl, int f) {
    auto pick = [&](auto* S) -> void* {
        switch (f) {
            case F_VEL: return S->vel.data();
            case F_RHO: return S->rho.data();
            case F_GAMMA: return S->gamma.data();
            case F_BETA: return S->beta.data();
            case F_DIV: return S->divergence.data();
            case F_PRESSURE: return S->pressure.data();
            case F_FORCES: return S->forces.data();
            case F_BOUNDARY: return S->boundary.data();
            case F_TMPSCAL: return S->tmpScal.data();
            case F_TMPVEC: return S->tmpVec.data();
        }
        return nullptr;
    };
    if (dbl) return pick((Sim2D<double>*)s);
    return pick((Sim2D<float>*)s);
}
int ora2_step(void* s, int dbl, ora_stats* st) {
    if (dbl) return ((Sim2D<double>*)s)->step(st);
    return ((Sim2D<float>*)s)->step(st);
}
int ora2_stage(void* s, int dbl, int stage, ora_stats* st) {
    const int rc = dbl ? ((Sim2D<double>*)s)->stage(stage, st) : ((Sim2D<float>*)s)->stage(stage, st);
    if (st) st->status = rc;
    return rc;
}
int ora2_solve(void* s, int dbl, ora_stats* st) {
    auto run = [&](auto* S) {
        std::fill(S->pressure.begin(), S->pressure.end(), 0);
        S->solve_steps = 0;
        const double err = (S->P.solver_type == 3) ? (double)S->jacobi_solve() : (double)S->cg_solve();
        if (st) { st->err = err; st->iterations = S->solve_steps; st->status = err > S->P.solver_threshold ? 1 : 0; }
        return st ? st->status : 0;
    };
    if (dbl) return run((Sim2D<double>*)s);
    return run((Sim2D<float>*)s);
}
void ora2_matmul(void* s, int dbl, const void* x, void* y) {
    auto run = [&](auto* S, auto tag) {
        using T = decltype(tag);
        std::vector<T> xv((const T*)x, (const T*)x + S->N), yv(S->N);
        S->mat_mul(yv, xv);
        std::memcpy(y, yv.data(), S->N * sizeof(T));
    };
    if (dbl) run((Sim2D<double>*)s, 0.0);
    else run((Sim2D<float>*)s, 0.0f);
}

void ora2_gradient(void* s, int dbl) {
    if (dbl) ((Sim2D<double>*)s)->subtract_pressure_gradient();
    else ((Sim2D<float>*)s)->subtract_pressure_gradient();
}
void ora2_enforce_walls(void* s, int dbl) {
    if (dbl) ((Sim2D<double>*)s)->enforce_wall_boundaries();
    else ((Sim2D<float>*)s)->enforce_wall_boundaries();
}
void ora2_divergence(void* s, int dbl) {
    auto run = [&](auto* S) {
        std::fill(S->divergence.begin(), S->divergence.end(), 0);
        S->enforce_wall_boundaries();
        S->compute_divergence();
    };
    if (dbl) run((Sim2D<double>*)s);
    else run((Sim2D<float>*)s);
}
void ora2_sharpen_delta(void* s, int dbl) {
    auto run = [&](auto* S) { S->sharpen_delta(S->tmpScal); };
    if (dbl) run((Sim2D<double>*)s);
    else run((Sim2D<float>*)s);
}
double ora2_sinterp(void* s, int dbl, int field, const double* pos) {
    auto run = [&](auto* S, auto tag) -> double {
        using T = decltype(tag);
        auto* f = (T*)ora2_field(s, dbl, field);
        std::vector<T> v(f, f + S->N);
        return (double)S->sinterp(v, V2<T>((T)pos[0], (T)pos[1]));
    };
    if (dbl) return run((Sim2D<double>*)s, 0.0);
    return run((Sim2D<float>*)s, 0.0f);
}
void ora2_trace(void* s, int dbl, const double* start, double dt, int substeps, double* out) {
    auto run = [&](auto* S, auto tag) {
        using T = decltype(tag);
        const V2<T> r = S->traceParticle(V2<T>((T)start[0], (T)start[1]), (T)dt, substeps);
        out[0] = r.x; out[1] = r.y;
    };
    if (dbl) run((Sim2D<double>*)s, 0.0);
    else run((Sim2D<float>*)s, 0.0f);
}

// ---- scene inputs (float fields, reference layout) ----
double ora_uniform(uint64_t seed, uint64_t stream, uint64_t ctr) { return uniform01(seed, stream, ctr); }
double ora_normal(uint64_t seed, uint64_t stream, uint64_t ctr) { return normal01(seed, stream, ctr); }

// Rasterise primitives at interior cell centres ((i - 0.5) h, h = 1 / (n - 2)) in order.
// dims[2] == 0 means 2D.  rho and boundary are float arrays of the grid size; the domain
// border is solid and holds rho 0 (Grid3D.cpp:28-42, Editor.cpp:246-257).
void ora_rasterise(const int* dims, const ora_prim* prims, int nprims, float* rho, float* boundary) {
    const int W = dims[0], H = dims[1], D = dims[2] > 0 ? dims[2] : 1;
    const bool is3 = dims[2] > 0;
    const double hx = 1.0 / (W - 2), hy = 1.0 / (H - 2), hz = is3 ? 1.0 / (D - 2) : 1.0;
    for (int z = 0; z < D; z++)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                const size_t c = ((size_t)z * H + y) * W + x;
                const bool border = x == 0 || y == 0 || x == W - 1 || y == H - 1 || (is3 && (z == 0 || z == D - 1));
                float r = rho[c], b = boundary[c];
                if (border) { b = 1.f; r = 0.f; }
                else {
                    const double cx = (x - 0.5) * hx, cy = (y - 0.5) * hy, cz = is3 ? (z - 0.5) * hz : 0.0;
                    for (int k = 0; k < nprims; k++) {
                        const ora_prim& P = prims[k];
                        bool in;
                        if (P.kind == 0) {
                            in = cx >= P.a[0] && cx < P.b[0] && cy >= P.a[1] && cy < P.b[1] &&
                                 (!is3 || (cz >= P.a[2] && cz < P.b[2]));
                        } else {
                            const double dx = cx - P.a[0], dy = cy - P.a[1], dz = is3 ? cz - P.a[2] : 0.0;
                            in = dx * dx + dy * dy + dz * dz <= P.b[0] * P.b[0];
                        }
                        if (!in) continue;
                        if (P.solid) { b = 1.f; r = 0.f; }
                        else if (b != 1.f) r = 1.f;
                    }
                }
                rho[c] = r;
                boundary[c] = b;
            }
}

// N(0, sigma) noise on the faces that touch no solid, reference AoS layout (ncomp floats per
// cell); counter = cell index * ncomp + component, stream 1.
void ora_face_noise(const int* dims, const float* boundary, uint64_t seed, double sigma, float* vel) {
    const int W = dims[0], H = dims[1], D = dims[2] > 0 ? dims[2] : 1;
    const int nc = dims[2] > 0 ? 3 : 2;
    const int z0 = nc == 3 ? 1 : 0, z1 = nc == 3 ? D - 1 : 1;
    for (int z = z0; z < z1; z++)
        for (int y = 1; y < H - 1; y++)
            for (int x = 1; x < W - 1; x++) {
                const size_t c = ((size_t)z * H + y) * W + x;
                if (boundary[c] == 1.f) continue;
                const size_t nb[3] = {c - 1, c - (size_t)W, c - (size_t)W * H};
                for (int k = 0; k < nc; k++) {
                    if (boundary[nb[k]] == 1.f) continue;
                    vel[c * nc + k] += (float)(sigma * normal01(seed, 1, c * nc + k));
                }
            }
}

}  // extern "C"
