// 3D U-Net of the learned pressure preconditioner: forward, backward and Adam on gfx950.
//
// Layout: activations channels-last (NDHWC), channels a multiple of 16.  Convolutions are
// implicit GEMMs on the matrix cores with BF16 operands and FP32 accumulation: the forward and
// backward-data convolutions stage each block's input halo in LDS and run gfx950's
// v_mfma_f32_16x16x32_bf16 (k_conv3), and so do the weight gradients (k_conv_wgrad).
//   forward        y[v][co] = b[co] + sum_tap sum_ci x[v + off(tap)][ci] * W[tap][co][ci]
//   backward data  dx = the same kernel on dy with W flipped in space and transposed in channels
//   backward W     dW[tap][co][ci] = sum_v dy[v][co] * x[v + off(tap)][ci]: per-block partials,
//                  then a fixed-order reduction (deterministic)
// Master weights, gradients and Adam moments are FP32; BF16 copies feed the matrix cores.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "efs_internal.h"
#include "efs_unet.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

namespace {

__device__ inline uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }


// ---- 3x3x3 convolution (zero padding) as an implicit GEMM on v_mfma_f32_16x16x32_bf16 --------
// A 256-thread block computes a 16 x 8 x ZT tile of output voxels (ZT = 1, 2 or 4) for NT
// 16-channel output tiles.  Per 16-channel input chunk the block stages the tile's input halo (18 x
// 10 x (ZT + 2) voxels x 16 channels, 17, 23 or 35 KB) in LDS once and reuses it for all 27 taps;
// the GEMM's K runs over (tap, channel) in 15 steps of 32 (pairs of taps, see k_conv3).  Each wave
// owns 2 ZT rows of 16 voxels: rows y0 + 2 ZT (w % (4 / ZT)) + r (r < 2 ZT) of plane
// z0 + w / (4 / ZT).  The deeper tiles stage fewer halo voxels per output voxel (2.1 for ZT = 4,
// 2.8 for 2, 4.2 for 1); the shallower ones make more blocks for the coarse levels.
// Operand layout of the 16x16x32 BF16 MFMA: lane l supplies A row l % 16 and B column l % 16 for K
// elements 8 (l / 16) .. 8 (l / 16) + 7 and receives D rows 4 (l / 16) + i, i = 0..3, of column
// l % 16.  A holds the weights (row = output channel) and B the voxels (column = voxel, 8
// consecutive channels of one tap, a 16-byte LDS load), so D is output channel x voxel: lane l ends
// with 4 consecutive output channels of voxel l % 16, which it writes as one 8-byte (BF16) or
// 16-byte (FP32) store.
// The halo is held as two planes of 16-byte halves, channels 0-7 of every voxel, then channels
// 8-15: the 16 lanes that read one half for 16 consecutive voxels then read 256 contiguous bytes.
// Global loads go through raw buffer resources with 32-bit byte offsets (tensor sizes are checked
// on the host to stay below 2^31 bytes); a load at offset OOB returns zeros, which reads the halo's
// padding and the missing tap's zero weights without branches.
// Each chunk's weights are staged in LDS once per block (the 4 waves share them) beside the halo.
constexpr int CT_X = 16, CT_Y = 8;
template <int ZT>
struct Tile {
    static constexpr int HX = CT_X + 2, HY = CT_Y + 2, HZ = ZT + 2, HN = HX * HY * HZ;  // the input halo
    // LDS stride of the halves: a multiple of 256 bytes, so that the lanes reading the two halves
    // of one voxel fall in the same banks only in different cycles (others conflict 2-way)
    static constexpr int HS = (HN + 15) / 16 * 16;
    static constexpr int NI = (HN * 2 + 255) / 256;  // 16-byte halo halves per thread
    static constexpr int RW = 2 * ZT;                // rows per wave
    // LDS offset (in halo voxels) of tap t from the tile position it applies to
    static constexpr int tap_off(int t) { return (t / 9) * HY * HX + ((t / 3) % 3) * HX + t % 3; }
};
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

constexpr unsigned OOB = 0x80000000u;  // a byte offset past every buffer the kernels address

// raw buffer resource over `bytes` bytes at p
__device__ inline __amdgpu_buffer_rsrc_t bufres(const void* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, kBufDword3);
}

// two FP32 values as a packed BF16 pair (lo in bits 0-15), rounded to nearest even
__device__ inline uint32_t pk_bf16(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{lo, hi}, bf16x2));
}

// A tile's halo for channels [c0, c0 + 16) of a source with Csrc channels (zero outside the
// D x H x W volume): thread t holds half t & 1 (8 channels, BF16) of voxels t / 2 + 128 k.
// halo_issue loads it into registers (FP32 sources rounded to nearest even), halo_put writes
// tile[half * HS + voxel].  With `up`, the source is at half resolution and each fine voxel reads
// its parent (nearest upsampling).
template <int ZT>
struct Halo {
    uint4 a[Tile<ZT>::NI];
};
template <int ZT, typename TS>
__device__ inline void halo_issue(Halo<ZT>& h, const TS* __restrict__ src, int Csrc, int c0, int x0, int y0, int z0, int D, int H,
                                  int W, bool up) {
    using T = Tile<ZT>;
    const int sh = up ? 1 : 0, sD = D >> sh, sH = H >> sh, sW = W >> sh;
    const __amdgpu_buffer_rsrc_t r = bufres(src, (unsigned)sD * sH * sW * Csrc * (unsigned)sizeof(TS));
    const int half = threadIdx.x & 1;
#pragma unroll
    for (int k = 0; k < T::NI; k++) {
        const int v = (threadIdx.x >> 1) + 128 * k;
        const int hx = v % T::HX, hy = (v / T::HX) % T::HY, hz = v / (T::HX * T::HY);
        const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
        const bool inside = v < T::HN && gx >= 0 && gx < W && gy >= 0 && gy < H && gz >= 0 && gz < D;
        const unsigned vox = (unsigned)(((gz >> sh) * sH + (gy >> sh)) * sW + (gx >> sh));
        const unsigned off = inside ? (vox * Csrc + c0 + half * 8) * (unsigned)sizeof(TS) : OOB;
        if constexpr (std::is_same<TS, float>::value) {
            const f32x4 lo = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
            const f32x4 hi = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 16, 0));
            h.a[k] = make_uint4(pk_bf16(lo.x, lo.y), pk_bf16(lo.z, lo.w), pk_bf16(hi.x, hi.y), pk_bf16(hi.z, hi.w));
        } else {
            h.a[k] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
        }
    }
}
template <int ZT>
__device__ inline void halo_put(uint4* tile, const Halo<ZT>& h) {
    using T = Tile<ZT>;
    const int half = threadIdx.x & 1;
#pragma unroll
    for (int k = 0; k < T::NI; k++) {
        const int v = (threadIdx.x >> 1) + 128 * k;
        if (v < T::HN) tile[half * T::HS + v] = h.a[k];
    }
}

// The block's position (bx, by, bz) in its grid, renumbered so that each XCD (the hardware deals
// workgroup i to XCD i mod 8, each with its own L2) takes a contiguous run of the x-fastest block
// order: neighbouring tiles, whose input halos overlap, then share an L2 instead of each XCD
// fetching the overlap again.
__device__ inline void xcd_tile(int& bx, int& by, int& bz) {
    const int gx = gridDim.x, gy = gridDim.y, n = gx * gy * gridDim.z;
    const int b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const int t = n % NXCD == 0 ? (b % NXCD) * (n / NXCD) + b / NXCD : b;
    bx = t % gx;
    by = (t / gx) % gy;
    bz = t / (gx * gy);
}

// out[v][co] = bias[co] + sum_tap sum_ci in[v + off(tap)][ci] * w[tap][co][ci], then ReLU (relu),
// written as FP32 (out_f) or BF16 (out_b).  With `up` (Cup channels at half resolution) the input
// is the decoder's concatenation, channels [0, Cup) the upsampled `up` and the rest `in` (Cin - Cup
// channels), read in place.  The block's z index (after the XCD renumbering) = z tile * ngroups +
// output-channel group; a group covers NT 16-channel tiles (tiles past Cout are skipped).  Cin, Cup, Cout: multiples of 16; D, H, W:
// any (partial tiles are masked; with `up`, even).  `in` is BF16, or FP32 rounded to BF16 as it is
// staged (TIN = float).  With out_pool (D, H, W even) the block also writes the 2x2x2 average of
// its BF16 outputs, the next encoder level's input.
template <int NT, int ZT, typename TIN>
__global__ __launch_bounds__(256) void k_conv3(const uint16_t* __restrict__ up, int Cup, const TIN* __restrict__ in, int Cin,
                                               const uint16_t* __restrict__ w, const float* __restrict__ bias, int Cout, int D,
                                               int H, int W, int relu, int ngroups, float* __restrict__ out_f,
                                               uint16_t* __restrict__ out_b, uint16_t* __restrict__ out_pool,
                                               const int* __restrict__ skip) {
    if (skip && *skip) return;
    using T = Tile<ZT>;
    constexpr int RW = T::RW, NFAM = 5;
    __shared__ uint4 tile[T::HS * 2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, g = lane >> 4, hsel = g >> 1;
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    const int x0 = bx * CT_X, y0 = by * CT_Y;
    const int zt = bz / ngroups, nt0 = (bz % ngroups) * NT;
    const int rz = wave / (4 / ZT), ry = (wave % (4 / ZT)) * RW;  // the wave's plane and first row
    floatx4 acc[RW][NT];
    // K steps.  The 27 taps are 9 (dz, dx) combos c = 3 dz + dx times 3 dy.  A K step holds two
    // taps with the same dy, one per lane half hsel = g >> 1: combos 2 f and 2 f + 1 of family f
    // (f = 0..4; family 4 pairs combo 8 with nothing, zero weights).  The B operand of row r, step
    // (f, dy) is voxel row r + dy of the family's combos, so a family's 3 steps over RW rows read
    // only RW + 2 rows from LDS: row q feeds the MFMAs of (r, dy) = (q - dy, dy).  15 steps of 32
    // (7 % more MFMAs than 14) for 2.2 (RW = 8) or 1.9 (RW = 4) times fewer LDS reads.
    const int hbase = (g & 1) * T::HS + (rz * T::HY + ry) * T::HX + m;
    auto fam_base = [&](int f) {  // LDS index of the lane's row q = 0 in family f
        const int c0 = 2 * f, c1 = 2 * f + 1 < 9 ? 2 * f + 1 : 8;  // the missing combo reads valid LDS
        const int o0 = (c0 / 3) * T::HY * T::HX + c0 % 3, o1 = (c1 / 3) * T::HY * T::HX + c1 % 3;
        return hbase + o0 + hsel * (o1 - o0);
    };
    const __amdgpu_buffer_rsrc_t rw = bufres(w, 27u * Cout * Cin * 2u);
    auto mfma_fam = [&](int f, const bf16x8 (&a)[3][NT]) {
        const int base = fam_base(f);
#pragma unroll
        for (int q = 0; q < RW + 2; q++) {
            const bf16x8 b = __builtin_bit_cast(bf16x8, tile[base + q * T::HX]);
#pragma unroll
            for (int dy = 0; dy < 3; dy++) {
                const int r = q - dy;
                if (r < 0 || r >= RW) continue;
#pragma unroll
                for (int n = 0; n < NT; n++) acc[r][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[dy][n], b, acc[r][n], 0, 0, 0);
            }
        }
    };
    auto zero_acc = [&]() {
#pragma unroll
        for (int r = 0; r < RW; r++)
#pragma unroll
            for (int n = 0; n < NT; n++) acc[r][n] = floatx4{0.f, 0.f, 0.f, 0.f};
    };
    const int nchunks = Cin >> 4;
    auto issue = [&](Halo<ZT>& h, int zt, int ch) {
        if (up && ch * 16 < Cup) halo_issue<ZT>(h, up, Cup, ch * 16, x0, y0, zt * ZT, D, H, W, true);
        else halo_issue<ZT>(h, in, Cin - Cup, ch * 16 - Cup, x0, y0, zt * ZT, D, H, W, false);
    };
    // lane (m, g) holds, for row r and tile n, output channels (nt0 + n) 16 + 4 g + i (i = 0..3) of
    // voxel (x0 + m, y0 + ry + r, z0 + rz)
    f32x4 bv[NT];
#pragma unroll
    for (int n = 0; n < NT; n++) {
        const int co = (nt0 + n) * 16 + 4 * g;
        bv[n] = f32x4{0.f, 0.f, 0.f, 0.f};
        if (bias)
            bv[n] = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(bufres(bias, Cout * 4u), co < Cout ? co * 4u : OOB, 0, 0));
    }
    auto value = [&](int n, int r, int i) {  // the output before rounding
        const float v = acc[r][n][i] + bv[n][i];
        return relu ? (v > 0.f ? v : 0.f) : v;
    };
    // the outputs of the tile at z0 (and with out_pool its 2x2x2 averages); uses the LDS tile
    auto finish = [&](int z0) {
        const int x = x0 + m, z = z0 + rz;
        const size_t row0 = (((size_t)z * H + y0 + ry) * W + x) * Cout, rowstride = (size_t)W * Cout;
#pragma unroll
        for (int n = 0; n < NT; n++) {
            const int co = (nt0 + n) * 16 + 4 * g;
#pragma unroll
            for (int r = 0; r < RW; r++) {
                if (co >= Cout || z >= D || y0 + ry + r >= H || x >= W) continue;
                const size_t o = row0 + r * rowstride + co;
                if (out_f)
                    *reinterpret_cast<f32x4*>(out_f + o) = f32x4{value(n, r, 0), value(n, r, 1), value(n, r, 2), value(n, r, 3)};
                if (out_b)
                    *reinterpret_cast<u32x2*>(out_b + o) =
                        u32x2{pk_bf16(value(n, r, 0), value(n, r, 1)), pk_bf16(value(n, r, 2), value(n, r, 3))};
            }
        }
        if (!out_pool || ZT < 2) return;  // pooling pairs planes within the tile (never asked with ZT = 1)
        // A pooled cell averages its 8 children's BF16 values, summed z-major, then y, then x.  The
        // x pairs sit in lanes m and m ^ 1 (exchanged by a DPP quad permutation), the y pairs in rows
        // 2 j and 2 j + 1 of one lane, the z pairs in the waves of planes rz = 2 p and 2 p + 1 with
        // the same rows: the even plane's wave adds its four from 0 and hands the partial through
        // LDS to the odd plane's wave, which adds its four.  The even-m lanes hold the cells' sums.
        auto bfr = [](float v) { return __uint_as_float(pk_bf16(v, 0.f) << 16); };
        auto add4 = [&](float sum, int n, int j, int i) {
            const float a0 = bfr(value(n, 2 * j, i)), a1 = bfr(value(n, 2 * j + 1, i));
            const float b0 = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(a0), 0xB1, 0xF, 0xF, false));
            const float b1 = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(a1), 0xB1, 0xF, 0xF, false));
            sum = sum + a0;
            sum = sum + b0;
            sum = sum + a1;
            return sum + b1;
        };
        // NT x 2 (wave pairs) x RW / 2 (j) x 4 channels x 64 lanes
        float* part = reinterpret_cast<float*>(tile);
        const int pair = (rz >> 1) * (4 / ZT) + wave % (4 / ZT);
        auto pidx = [&](int n, int j, int i) { return (((n * 2 + pair) * (RW / 2) + j) * 4 + i) * 64 + lane; };
        static_assert(NT * 2 * (RW / 2) * 4 * 64 <= T::HS * 2 * 4, "pool partials must fit the halo tile");
        __syncthreads();  // every wave is done with the halo
        if (!(rz & 1)) {
#pragma unroll
            for (int n = 0; n < NT; n++)
#pragma unroll
                for (int j = 0; j < RW / 2; j++)
#pragma unroll
                    for (int i = 0; i < 4; i++) part[pidx(n, j, i)] = add4(0.f, n, j, i);
        }
        __syncthreads();
        if (rz & 1) {
            const int D2 = D >> 1, H2 = H >> 1, W2 = W >> 1, pz = (z0 + rz) >> 1, px = (x0 >> 1) + (m >> 1);
#pragma unroll
            for (int n = 0; n < NT; n++) {
                const int co = (nt0 + n) * 16 + 4 * g;
#pragma unroll
                for (int j = 0; j < RW / 2; j++) {
                    float sum[4];
#pragma unroll
                    for (int i = 0; i < 4; i++) sum[i] = add4(part[pidx(n, j, i)], n, j, i) * 0.125f;
                    const int py = ((y0 + ry) >> 1) + j;
                    if (!(m & 1) && co < Cout && pz < D2 && py < H2 && px < W2)
                        *reinterpret_cast<u32x2*>(out_pool + (((size_t)pz * H2 + py) * W2 + px) * Cout + co) =
                            u32x2{pk_bf16(sum[0], sum[1]), pk_bf16(sum[2], sum[3])};
                }
            }
        }
    };
    zero_acc();
    {
        // The chunk's weights are staged in LDS once per block, so a wave reads its A operands
        // from LDS instead of each wave loading all of them: wl[((f * 3 + dy) * NT + n) * 64 + lane]
        // is the lane's A operand of step (f, dy), tile n.
        constexpr int WL_N = NFAM * 3 * NT * 64, WL_NI = (WL_N + 255) / 256;
        __shared__ uint4 wl[WL_N];
        for (int ch = 0; ch < nchunks; ch++) {
            uint4 wv[WL_NI];
#pragma unroll
            for (int k = 0; k < WL_NI; k++) {
                const int i = threadIdx.x + 256 * k;
                const int ln = i & 63, q = i >> 6;
                const int n = q % NT, f = (q / NT) / 3, dy = (q / NT) % 3;
                const int gg = ln >> 4, c = 2 * f + (gg >> 1), co = (nt0 + n) * 16 + (ln & 15);
                const int tap = (c / 3) * 9 + dy * 3 + c % 3;
                const bool ok = i < WL_N && c < 9 && co < Cout;
                const unsigned off = ok ? (((unsigned)tap * Cout + co) * Cin + ch * 16 + (gg & 1) * 8) * 2u : OOB;
                wv[k] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rw, off, 0, 0));
            }
            Halo<ZT> h;
            issue(h, zt, ch);
            if (ch) __syncthreads();  // the previous chunk is done with the LDS tiles
            halo_put(tile, h);
#pragma unroll
            for (int k = 0; k < WL_NI; k++) {
                const int i = threadIdx.x + 256 * k;
                if (i < WL_N) wl[i] = wv[k];
            }
            __syncthreads();
#pragma unroll
            for (int f = 0; f < NFAM; f++) {
                bf16x8 a[3][NT];
#pragma unroll
                for (int dy = 0; dy < 3; dy++)
#pragma unroll
                    for (int n = 0; n < NT; n++) a[dy][n] = __builtin_bit_cast(bf16x8, wl[((f * 3 + dy) * NT + n) * 64 + lane]);
                mfma_fam(f, a);
            }
        }
    }
    finish(zt * ZT);
}

// Output channel 0 only of a 16 -> Cout 3x3x3 convolution without activation (the U-Net's last
// layer at inference, whose other output channels are unused), on 16 x 8 x ZT output tiles.
// out[z][y][x] = sum over (dz, dy) of Q_{dz,dy}[z + dz][y + dy][x], where for each of the halo's
// (ZT + 2) x 10 voxel rows Q_{dz,dy}[row][x] = sum over dx and the 16 channels of w[dz][dy][dx] x[row][x + dx].
// The 9 (dz, dy) combinations of one row are the rows of one MFMA product (A = weights, combos
// x channels; B = the row's 16 voxels shifted by dx, channels x voxels; three products for dx =
// 0..2 accumulate in FP32).  The partial sums go to LDS in place of the halo (9 x rows x 16
// floats fit in the halo's size) and each output voxel adds 9 of them: for ZT = 2, 61 + 23 + 9 KB
// of LDS traffic per tile instead of the 221 KB that 27 taps x 32 bytes per voxel read.
constexpr int CH0_ZT = 4;  // planes per output tile
template <int ZT>
__global__ __launch_bounds__(256) void k_conv3_ch0(const uint16_t* __restrict__ in, const uint16_t* __restrict__ w, int Cout,
                                                   const float* __restrict__ bias, int D, int H, int W, float* __restrict__ out0,
                                                   const int* __restrict__ skip) {
    if (skip && *skip) return;
    using T = Tile<ZT>;
    constexpr int NROW = T::HZ * T::HY;  // halo rows
    constexpr int RPW = NROW / 4;        // rows per wave
    static_assert(NROW % 4 == 0, "the halo rows split evenly over the 4 waves");
    static_assert(9 * NROW * CT_X * 4 <= T::HS * 2 * 16, "the partial sums must fit in the halo's LDS");
    __shared__ uint4 tile[T::HS * 2];
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    const int x0 = bx * CT_X, y0 = by * CT_Y, z0 = bz * ZT;
    {
        Halo<ZT> h;
        halo_issue<ZT>(h, in, 16, 0, x0, y0, z0, D, H, W, false);
        halo_put(tile, h);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = lane & 15, q = lane >> 4;  // MFMA column (voxel) and k group (channels 4q..4q+3)
    // A operands: row m = (dz, dy) = (m / 3, m % 3) for m < 9, zero rows above
    bf16x4 a[3];
#pragma unroll
    for (int dx = 0; dx < 3; dx++) {
        a[dx] = bf16x4{0, 0, 0, 0};
        if (n < 9) {
            const int tap = (n / 3) * 9 + (n % 3) * 3 + dx;
            a[dx] = *reinterpret_cast<const bf16x4*>(w + (size_t)tap * Cout * 16 + q * 4);
        }
    }
    __syncthreads();
    floatx4 acc[RPW];
    const uint2* t2 = reinterpret_cast<const uint2*>(tile);
    const int hoff = (q >> 1) * T::HS * 2 + (q & 1);  // the lane's 8 bytes within a voxel's two halves
#pragma unroll
    for (int i = 0; i < RPW; i++) {
        const int row = wave + 4 * i;
        acc[i] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int dx = 0; dx < 3; dx++) {
            const uint2 v = t2[hoff + (row * T::HX + n + dx) * 2];
            acc[i] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[dx], __builtin_bit_cast(bf16x4, v), acc[i], 0, 0, 0);
        }
    }
    __syncthreads();  // every wave has read the halo: the partial sums replace it
    float* qs = reinterpret_cast<float*>(tile);
    // D: lane (n, q) holds combos 4q + e of voxel n
#pragma unroll
    for (int i = 0; i < RPW; i++) {
        const int row = wave + 4 * i;
#pragma unroll
        for (int e = 0; e < 4; e++)
            if (4 * q + e < 9) qs[((4 * q + e) * NROW + row) * CT_X + n] = acc[i][e];
    }
    __syncthreads();
    const int tx = threadIdx.x & 15, ty = (threadIdx.x >> 4) & 7;
    const int x = x0 + tx, y = y0 + ty;
    if (x >= W || y >= H) return;
#pragma unroll
    for (int tz = threadIdx.x >> 7; tz < ZT; tz += 2) {
        const int z = z0 + tz;
        if (z >= D) return;
        float sum = 0.f;
#pragma unroll
        for (int m = 0; m < 9; m++) sum += qs[(m * NROW + (tz + m / 3) * T::HY + ty + m % 3) * CT_X + tx];
        out0[((size_t)z * H + y) * W + x] = sum + (bias ? bias[0] : 0.f);
    }
}

// Weight gradient partials of a 3x3x3 convolution: dW[tap][co][ci] = sum_v dy[v][co] x[v + off(tap)][ci]
// on v_mfma_f32_16x16x32_bf16 with the voxels as the K dimension (D = co x ci).  A block takes one
// 16-channel group of co and of ci and a strided share of the 16 x 8 x 2 voxel tiles; its 4 waves
// take the 27 taps in turn (tap = wave + 4 j) and keep their sums in registers over all its tiles.
// Per tile, dy and the input halo are staged in LDS channel-major (each channel a run of 16 voxels
// per row), the halo in three copies shifted by dx = 0, 1, 2, so that every operand of every tap
// is one aligned 16-byte LDS read: a lane's 8 K values are 8 voxels along x of one row and one
// channel.  Each block writes its sums as partial `b` (part[b][tap][co][ci]) for the fixed-order
// reduction; the blocks of the first ci group also sum their dy tiles per channel (the bias
// gradient), as partial b of bpart (bpart[b][co]).
constexpr int WG_ZT = 2;                               // planes per voxel tile
constexpr int WG_ROWS = CT_Y * WG_ZT;                  // 16 voxel rows of 16 per tile
constexpr int WG_HROWS = (CT_Y + 2) * (WG_ZT + 2);     // 40 halo rows
constexpr int WG_DYS = WG_ROWS * 16 + 8;               // BF16 elements per co (16 bytes of padding:
constexpr int WG_XS = WG_HROWS * 16 + 8;               // the 16 channels of a read fall in different banks)
constexpr int WG_TAPS = 7;                             // taps per wave (27 = 4 waves x 7, the last 6)

// dword of channel c (0..7) of two voxels' 8-channel halves: voxel a in the low half, b in the high
__device__ inline unsigned pair_ch(const uint4& a, const uint4& b, int c) {
    const unsigned wa = c < 2 ? a.x : (c < 4 ? a.y : (c < 6 ? a.z : a.w));
    const unsigned wb = c < 2 ? b.x : (c < 4 ? b.y : (c < 6 ? b.z : b.w));
    return (c & 1) ? ((wa >> 16) | (wb & 0xFFFF0000u)) : ((wa & 0xFFFFu) | (wb << 16));
}

template <typename TX>
__global__ __launch_bounds__(256) void k_conv_wgrad(const uint16_t* __restrict__ dy, int Cout, const TX* __restrict__ x, int Cin,
                                                    int D, int H, int W, int nper, float* __restrict__ part,
                                                    float* __restrict__ bpart) {
    __shared__ uint4 dyT4[16 * WG_DYS / 8];      // [co][row][16 voxels]
    __shared__ uint4 xT4[3 * 16 * WG_XS / 8];    // [dx][ci][halo row][i]: the halo at column i + dx
    unsigned* dyT32 = reinterpret_cast<unsigned*>(dyT4);
    unsigned* xT32 = reinterpret_cast<unsigned*>(xT4);
    const uint16_t* dyT = reinterpret_cast<const uint16_t*>(dyT4);
    const uint16_t* xT = reinterpret_cast<const uint16_t*>(xT4);
    const int ncg = Cout >> 4;
    const int g = blockIdx.x / nper, b = blockIdx.x % nper;
    const int cog = g % ncg, cig = g / ncg;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n16 = lane & 15, kq = lane >> 4;
    const int tiles_x = (W + 15) / 16, tiles_y = (H + CT_Y - 1) / CT_Y, tiles_z = (D + WG_ZT - 1) / WG_ZT;
    const int ntiles = tiles_x * tiles_y * tiles_z;
    const __amdgpu_buffer_rsrc_t rdy = bufres(dy, (unsigned)((size_t)D * H * W * Cout * 2));
    const __amdgpu_buffer_rsrc_t rx = bufres(x, (unsigned)((size_t)D * H * W * Cin * sizeof(TX)));
    floatx4 acc[WG_TAPS];
#pragma unroll
    for (int j = 0; j < WG_TAPS; j++) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
    float bacc = 0.f;  // bias: thread (co = tid / 16, row = tid % 16) of the first ci group
    // The next tile's global loads are issued into registers before the current tile's MFMAs
    // (LDS, not registers, limits the blocks per CU).  dy: thread = (row, half of the 16 channels,
    // voxel pair along x).  Halo: task = (halo row, half, pair j) for tasks tid, tid + 256 and
    // tid + 512 (< 640, waves 0 and 1 only), halo columns 2j .. 2j + 3 (global x0 - 1 + ...).
    constexpr int NTASK = WG_HROWS * 16, KT = (NTASK + 255) / 256;
    constexpr bool F32 = std::is_same<TX, float>::value;
    constexpr int NL = F32 ? 2 : 1;  // 16-byte loads per voxel half
    uint4 qd[2];
    uint4 qh[KT][4][NL];
    auto issue = [&](int t) {
        const int x0 = (t % tiles_x) * 16, y0 = ((t / tiles_x) % tiles_y) * CT_Y, z0 = (t / (tiles_x * tiles_y)) * WG_ZT;
        {
            const int j = tid & 7, h = (tid >> 3) & 1, row = tid >> 4;
            const int gy = y0 + row % CT_Y, gz = z0 + row / CT_Y;
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const int gx = x0 + 2 * j + e;
                const bool in = gx < W && gy < H && gz < D;
                const unsigned off = in ? (unsigned)(((((size_t)gz * H + gy) * W + gx) * Cout + cog * 16 + h * 8) * 2) : OOB;
                qd[e] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rdy, off, 0, 0));
            }
        }
#pragma unroll
        for (int k = 0; k < KT; k++) {
            const int task = tid + 256 * k;
            if (task >= NTASK) break;  // wave-uniform
            const int j = task & 7, h = (task >> 3) & 1, hrow = task >> 4;
            const int gy = y0 - 1 + hrow % (CT_Y + 2), gz = z0 - 1 + hrow / (CT_Y + 2);
            const bool rowin = gy >= 0 && gy < H && gz >= 0 && gz < D;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int gx = x0 - 1 + 2 * j + e;
                const bool in = rowin && gx >= 0 && gx < W;
                const unsigned off =
                    in ? (unsigned)(((((size_t)gz * H + gy) * W + gx) * Cin + cig * 16 + h * 8) * sizeof(TX)) : OOB;
#pragma unroll
                for (int l = 0; l < NL; l++)
                    qh[k][e][l] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rx, off, 16 * l, 0));
            }
        }
    };
    auto commit = [&]() {
        {
            const int j = tid & 7, h = (tid >> 3) & 1, row = tid >> 4;
#pragma unroll
            for (int c = 0; c < 8; c++) dyT32[((8 * h + c) * WG_DYS + row * 16 + 2 * j) / 2] = pair_ch(qd[0], qd[1], c);
        }
#pragma unroll
        for (int k = 0; k < KT; k++) {
            const int task = tid + 256 * k;
            if (task >= NTASK) break;
            const int j = task & 7, h = (task >> 3) & 1, hrow = task >> 4;
            uint4 q[4];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                if constexpr (F32) {
                    const f32x4 lo = __builtin_bit_cast(f32x4, qh[k][e][0]), hi = __builtin_bit_cast(f32x4, qh[k][e][NL - 1]);
                    q[e] = make_uint4(pk_bf16(lo.x, lo.y), pk_bf16(lo.z, lo.w), pk_bf16(hi.x, hi.y), pk_bf16(hi.z, hi.w));
                } else {
                    q[e] = qh[k][e][0];
                }
            }
#pragma unroll
            for (int dx = 0; dx < 3; dx++)
#pragma unroll
                for (int c = 0; c < 8; c++)
                    xT32[((dx * 16 + 8 * h + c) * WG_XS + hrow * 16 + 2 * j) / 2] = pair_ch(q[dx], q[dx + 1], c);
        }
    };
    if (b < ntiles) issue(b);
    for (int t = b; t < ntiles; t += nper) {
        __syncthreads();  // every wave is done with the previous tile
        commit();
        __syncthreads();
        if (t + nper < ntiles) issue(t + nper);
        if (cig == 0) {
            const uint16_t* rowp = dyT + (tid >> 4) * WG_DYS + (tid & 15) * 16;
#pragma unroll
            for (int k = 0; k < 16; k++) bacc += bf2f(rowp[k]);
        }
        // K step p: output rows 2p (lanes kq 0, 1) and 2p + 1 (kq 2, 3), voxels 8 (kq & 1) .. + 7
#pragma unroll 2
        for (int p = 0; p < WG_ROWS / 2; p++) {
            const int r = 2 * p + (kq >> 1), xo = (kq & 1) * 8;
            const int ty = r % CT_Y, tz = r / CT_Y;
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(dyT + n16 * WG_DYS + r * 16 + xo);
#pragma unroll
            for (int j = 0; j < WG_TAPS; j++) {
                const int tap = wave + 4 * j;
                if (tap >= 27) break;  // wave-uniform
                const int dz = tap / 9, dyy = (tap / 3) % 3, dx = tap % 3;
                const int hrow = (tz + dz) * (CT_Y + 2) + ty + dyy;
                const bf16x8 bb = *reinterpret_cast<const bf16x8*>(xT + (dx * 16 + n16) * WG_XS + hrow * 16 + xo);
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc[j], 0, 0, 0);
            }
        }
    }
    // D[co = 4 kq + i][ci = n16] of each tap
#pragma unroll
    for (int j = 0; j < WG_TAPS; j++) {
        const int tap = wave + 4 * j;
        if (tap >= 27) break;
#pragma unroll
        for (int i = 0; i < 4; i++)
            part[(((size_t)b * 27 + tap) * Cout + cog * 16 + 4 * kq + i) * Cin + cig * 16 + n16] = acc[j][i];
    }
    if (cig == 0) {
        __syncthreads();  // the last tile's dy is no longer read: its LDS takes the row sums
        float* red = reinterpret_cast<float*>(dyT4);
        red[tid] = bacc;
        __syncthreads();
        if (tid < 16) {
            float a = 0.f;
            for (int k = 0; k < 16; k++) a += red[tid * 16 + k];
            bpart[(size_t)b * Cout + cog * 16 + tid] = a;
        }
    }
}

__global__ void k_reduce_chunks(int n, int nchunks, const float* __restrict__ part, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int c = 0; c < nchunks; c++) s += part[(size_t)c * n + i];
    out[i] = s;
}

// First pass of a two-pass fixed-order sum over chunks: block (x, seg) adds chunks
// [seg * nchunks / nseg, (seg + 1) * nchunks / nseg) in order into tmp[seg]; k_reduce_chunks then
// adds the nseg segment sums in order.
__global__ void k_reduce_seg(int n, int nchunks, int nseg, const float* __restrict__ part, float* __restrict__ tmp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, seg = blockIdx.y;
    if (i >= n) return;
    const int c0 = (int)((long long)seg * nchunks / nseg), c1 = (int)((long long)(seg + 1) * nchunks / nseg);
    float s = 0.f;
    for (int c = c0; c < c1; c++) s += part[(size_t)c * n + i];
    tmp[(size_t)seg * n + i] = s;
}
constexpr int RED_SEG = 32;  // segments of the chunk sums

// dpre = dy * (y > 0) (y is the post-ReLU activation), as BF16 MFMA operand
__global__ void k_relu_bwd(long long n, const float* __restrict__ dy, const uint16_t* __restrict__ y, int relu,
                           uint16_t* __restrict__ dpre) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = dy[i];
    dpre[i] = f2bf((!relu || bf2f(y[i]) > 0.f) ? g : 0.f);
}

// Backward of the 2x2x2 average pool (the forward is fused into k_conv3), channels-last:
// d_in[child] += d_out[parent] / 8   (dims of the input / fine level)
__global__ void k_pool_bwd(const float* __restrict__ dout, int C, int D, int H, int W, float* __restrict__ din) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)D * H * W * C;
    if (i >= n) return;
    const int c = (int)(i % C);
    long long v = i / C;
    const int x = (int)(v % W); v /= W;
    const int y = (int)(v % H);
    const int z = (int)(v / H);
    din[i] += 0.125f * dout[(((long long)(z >> 1) * (H >> 1) + (y >> 1)) * (W >> 1) + (x >> 1)) * C + c];
}
// nearest 2x upsample of `a` (coarse, Ca channels) concatenated with `b` (fine, Cb channels)
__global__ void k_up_concat(const uint16_t* __restrict__ a, int Ca, const uint16_t* __restrict__ b, int Cb, int D, int H, int W,
                            uint16_t* __restrict__ out) {
    const int C = Ca + Cb;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)D * H * W * C;
    if (i >= n) return;
    const int c = (int)(i % C);
    long long v = i / C;
    const int x = (int)(v % W); v /= W;
    const int y = (int)(v % H);
    const int z = (int)(v / H);
    if (c < Ca) out[i] = a[(((long long)(z >> 1) * (H >> 1) + (y >> 1)) * (W >> 1) + (x >> 1)) * Ca + c];
    else out[i] = b[(((long long)z * H + y) * W + x) * Cb + (c - Ca)];
}
// backward of k_up_concat: d_a[parent] += sum of children, d_b += own
__global__ void k_up_concat_bwd_a(const float* __restrict__ dout, int Ca, int Cb, int D, int H, int W, float* __restrict__ da) {
    // one thread per coarse element
    const int C = Ca + Cb;
    const int Dc = D >> 1, Hc = H >> 1, Wc = W >> 1;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)Dc * Hc * Wc * Ca;
    if (i >= n) return;
    const int c = (int)(i % Ca);
    long long v = i / Ca;
    const int x = (int)(v % Wc); v /= Wc;
    const int y = (int)(v % Hc);
    const int z = (int)(v / Hc);
    float s = 0.f;
    for (int k = 0; k < 8; k++) {
        const long long q = (((long long)(2 * z + (k >> 2)) * H + (2 * y + ((k >> 1) & 1))) * W + (2 * x + (k & 1))) * C + c;
        s += dout[q];
    }
    da[i] += s;
}
__global__ void k_concat_bwd_b(const float* __restrict__ dout, int Ca, int Cb, long long nvox, float* __restrict__ db) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nvox * Cb) return;
    const int c = (int)(i % Cb);
    const long long v = i / Cb;
    db[i] += dout[v * (Ca + Cb) + Ca + c];
}

__global__ void k_to_bf16(long long n, const float* __restrict__ a, uint16_t* __restrict__ b) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = f2bf(a[i]);
}
__global__ void k_zero(long long n, float* a) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.f;
}
// W[tap][co][ci] (fp32) -> BF16 copy and the flipped/transposed copy Wt[26 - tap][ci][co]
__global__ void k_prep_weights(int ntap, int Cout, int Cin, const float* __restrict__ w, uint16_t* __restrict__ wb,
                               uint16_t* __restrict__ wt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = ntap * Cout * Cin;
    if (i >= n) return;
    const int ci = i % Cin, co = (i / Cin) % Cout, tap = i / (Cin * Cout);
    const uint16_t v = f2bf(w[i]);
    wb[i] = v;
    wt[((size_t)(ntap - 1 - tap) * Cin + ci) * Cout + co] = v;
}
// loss on output channel 0: sum over voxels with mask of (y - t)^2 / count; dy = 2 (y - t) / count
__global__ void k_mse(long long nvox, int C, const float* __restrict__ y, const float* __restrict__ t,
                      const float* __restrict__ m, float inv_count, float* __restrict__ dy, double* __restrict__ part) {
    __shared__ double red[256];
    double s = 0.0;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox; v += (long long)gridDim.x * blockDim.x) {
        const float w = m ? m[v] : 1.f;
        const float e = (y[v * C] - t[v]) * w;
        s += (double)e * e;
        for (int c = 0; c < C; c++) dy[v * C + c] = (c == 0) ? 2.f * e * w * inv_count : 0.f;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void k_adam(int n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       float lr, float b1, float b2, float eps, float c1, float c2, const float* __restrict__ frozen) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (frozen && frozen[i] != 0.f) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

inline unsigned nb(long long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace

// ======================================================================================
// host side
// ======================================================================================
namespace efsu {

static int hipfail(hipError_t e, const char* what) { return efs_hip_fail(e, what); }
#define UHIP(call)                                            \
    do {                                                      \
        hipError_t _e = (call);                               \
        if (_e != hipSuccess) return hipfail(_e, #call);      \
    } while (0)

struct Conv {
    int cin, cout, ks, relu;
    size_t woff, boff;  // offsets into the parameter vector (weights [tap][co][ci], bias [co])
    size_t nw() const { return (size_t)(ks == 3 ? 27 : 1) * cout * cin; }
};

// Layer table of the 4-level U-Net (channels widths c0..c3, input 16 channels, output
// channel 0 of the last layer).
static std::vector<Conv> make_layers(const int c[4]) {
    std::vector<Conv> L;
    auto add = [&](int ci, int co, int ks, int relu) { L.push_back({ci, co, ks, relu, 0, 0}); };
    add(16, c[0], 3, 1);            // 0 enc0a
    add(c[0], c[0], 3, 1);          // 1 enc0b  -> e0
    add(c[0], c[1], 3, 1);          // 2 enc1a (after pool)
    add(c[1], c[1], 3, 1);          // 3 enc1b  -> e1
    add(c[1], c[2], 3, 1);          // 4 enc2a
    add(c[2], c[2], 3, 1);          // 5 enc2b  -> e2
    add(c[2], c[3], 3, 1);          // 6 enc3a
    add(c[3], c[3], 3, 1);          // 7 enc3b  -> e3
    add(c[3] + c[2], c[2], 3, 1);   // 8 dec2 (up(e3) ++ e2)
    add(c[2] + c[1], c[1], 3, 1);   // 9 dec1 (up(d2) ++ e1)
    add(c[1] + c[0], c[0], 3, 1);   // 10 dec0 (up(d1) ++ e0)
    add(c[0], 16, 3, 0);            // 11 out (channel 0 used)
    size_t off = 0;
    for (auto& l : L) {
        l.woff = off;
        off += l.nw();
        l.boff = off;
        off += l.cout;
    }
    return L;
}

}  // namespace efsu

using namespace efsu;

struct efs_unet_s {
    int widths[4];
    std::vector<Conv> layers;
    size_t nparams = 0;
    hipStream_t stream = nullptr;
    const int* skip = nullptr;  // device flag: the forward's convolutions return at once when it is set
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr, *frozen = nullptr;  // FP32 params / grads / moments
    uint16_t *wb = nullptr, *wt = nullptr;                                          // BF16 copies (fwd / transposed)
    bool wdirty = true;
    long long adam_t = 0;
    // activation buffers for the current dims
    int D = 0, H = 0, W = 0;
    std::vector<void*> bufs;
    uint16_t* act[12] = {};     // post-activation outputs of layers 0..10 (BF16)
    float* out = nullptr;       // layer 11 output (FP32, 16 channels)
    float* out0 = nullptr;      // layer 11 channel 0, compact (inference)
    uint16_t *p1 = nullptr, *p2 = nullptr, *p3 = nullptr;      // pooled inputs of enc1a/2a/3a
    uint16_t *cat2 = nullptr, *cat1 = nullptr, *cat0 = nullptr;  // decoder inputs
    uint16_t* inb = nullptr;    // BF16 input (16 channels)
    // backward scratch
    float *ga = nullptr, *gb = nullptr;  // FP32 gradient buffers (largest activation)
    float *g_e[4] = {};                  // accumulated gradients of the skip tensors e0..e3
    uint16_t* dpre = nullptr;
    float* wpart = nullptr;
    size_t wpart_n = 0;
    double* lpart = nullptr;
};

namespace {

int unet_alloc_dims(efs_unet_s* u, int D, int H, int W) {
    if (u->D == D && u->H == H && u->W == W) return 0;
    if (D % 8 || H % 8 || W % 8) {
        efs_set_error("U-Net dims must be multiples of 8");
        return -1;
    }
    {  // the convolutions address their input tensors with 32-bit byte offsets
        const size_t nv = (size_t)D * H * W;
        size_t big = nv * 16 * 4;  // the FP32 network input
        for (auto& l : u->layers) big = std::max(big, nv * (size_t)std::max(l.cin, l.cout) * 2);
        if (big > 0x7fffffffull) {
            efs_set_error("U-Net grid too large: a convolution input would exceed 2 GiB");
            return -1;
        }
    }
    for (void* p : u->bufs) hipFree(p);
    u->bufs.clear();
    u->D = D; u->H = H; u->W = W;
    const int* c = u->widths;
    auto A = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        u->bufs.push_back(p);
        return p;
    };
    const size_t v0 = (size_t)D * H * W, v1 = v0 / 8, v2 = v1 / 8, v3 = v2 / 8;
    const size_t lv[12] = {v0, v0, v1, v1, v2, v2, v3, v3, v2, v1, v0, v0};
    for (int i = 0; i < 11; i++) u->act[i] = (uint16_t*)A(lv[i] * u->layers[i].cout * 2);
    u->out = (float*)A(v0 * 16 * 4);
    u->out0 = (float*)A(v0 * 4);
    u->inb = (uint16_t*)A(v0 * 16 * 2);
    u->p1 = (uint16_t*)A(v1 * c[0] * 2);
    u->p2 = (uint16_t*)A(v2 * c[1] * 2);
    u->p3 = (uint16_t*)A(v3 * c[2] * 2);
    u->cat2 = (uint16_t*)A(v2 * (c[3] + c[2]) * 2);
    u->cat1 = (uint16_t*)A(v1 * (c[2] + c[1]) * 2);
    u->cat0 = (uint16_t*)A(v0 * (c[1] + c[0]) * 2);
    size_t maxel = 0;
    for (int i = 0; i < 12; i++) maxel = std::max(maxel, lv[i] * (size_t)u->layers[i].cin);
    maxel = std::max(maxel, v0 * 16);
    u->ga = (float*)A(maxel * 4);
    u->gb = (float*)A(maxel * 4);
    u->dpre = (uint16_t*)A(maxel * 2);
    u->g_e[0] = (float*)A(v0 * c[0] * 4);
    u->g_e[1] = (float*)A(v1 * c[1] * 4);
    u->g_e[2] = (float*)A(v2 * c[2] * 4);
    u->g_e[3] = (float*)A(v3 * c[3] * 4);
    u->lpart = (double*)A(1024 * 8);
    for (void* p : u->bufs)
        if (!p) {
            efs_set_error("U-Net activation allocation failed");
            return -2;
        }
    return 0;
}

int prep_weights(efs_unet_s* u) {
    if (!u->wdirty) return 0;
    for (auto& l : u->layers) {
        const int nt = l.ks == 3 ? 27 : 1;
        hipLaunchKernelGGL(k_prep_weights, dim3(nb(l.nw())), dim3(256), 0, u->stream, nt, l.cout, l.cin, u->w + l.woff,
                           u->wb + l.woff, u->wt + l.woff);
    }
    UHIP(hipGetLastError());
    u->wdirty = false;
    return 0;
}

// A 3x3x3 convolution of BF16 channels-last input with `cin` channels into `cout` channels
// (k_conv3: NT output tiles per block, the rest in further block groups).
template <typename TIN>
void conv3(hipStream_t st, const TIN* in, int cin, const uint16_t* w, const float* bias, int cout, int D, int H, int W,
           int relu, float* of, uint16_t* ob, const uint16_t* up = nullptr, int cup = 0, uint16_t* opool = nullptr,
           const int* skip = nullptr) {
    const int ntiles = cout / 16;
    const int nxy = ((W + CT_X - 1) / CT_X) * ((H + CT_Y - 1) / CT_Y);
    // Up to 4 output tiles per block, which reuses each staged halo for more outputs; few voxel
    // tiles (the coarse levels): one per block, so that more blocks share the work, and one plane
    // per tile where no pooling pairs planes.  Many blocks with one output tile: 4 planes.
    int NT = ntiles <= 4 ? ntiles : (ntiles % 4 == 0 ? 4 : (ntiles % 3 == 0 ? 3 : 4));
    const bool few = nxy * ((D + 1) / 2) * ((ntiles + NT - 1) / NT) < 1024;
    if (few) NT = 1;
    const int ZT = few ? (opool ? 2 : 1) : (NT == 1 ? 4 : 2);
    const int ngroups = (ntiles + NT - 1) / NT, nz = (D + ZT - 1) / ZT;
    const dim3 grid((W + CT_X - 1) / CT_X, (H + CT_Y - 1) / CT_Y, nz * ngroups);
#define EFS_CONV3(nt, zt) \
    hipLaunchKernelGGL((k_conv3<nt, zt, TIN>), grid, dim3(256), 0, st, up, cup, in, cin, w, bias, cout, D, H, W, relu, ngroups, of, \
                       ob, opool, skip)
    if (ZT == 1) EFS_CONV3(1, 1);
    else if (ZT == 4) EFS_CONV3(1, 4);
    else if (NT == 1) EFS_CONV3(1, 2);
    else if (NT == 2) EFS_CONV3(2, 2);
    else if (NT == 3) EFS_CONV3(3, 2);
    else EFS_CONV3(4, 2);
#undef EFS_CONV3
}

// `up` (cup channels at half resolution): the layer's input is up(up) ++ in, read in place.
template <typename TI>
void conv_fwd(efs_unet_s* u, const Conv& l, const TI* in, int D, int H, int W, float* of, uint16_t* ob, int ch0 = 0,
              const uint16_t* up = nullptr, int cup = 0, uint16_t* opool = nullptr) {
    if constexpr (std::is_same<TI, uint16_t>::value) {
        if (ch0) {  // channel 0 only (inference output); the layer has 16 input channels
            const dim3 grid((W + CT_X - 1) / CT_X, (H + CT_Y - 1) / CT_Y, (D + CH0_ZT - 1) / CH0_ZT);
            hipLaunchKernelGGL(k_conv3_ch0<CH0_ZT>, grid, dim3(256), 0, u->stream, in, u->wb + l.woff, l.cout, u->w + l.boff, D, H, W, of,
                               u->skip);
            return;
        }
    }
    conv3(u->stream, in, l.cin, u->wb + l.woff, u->w + l.boff, l.cout, D, H, W, l.relu, of, ob, up, cup, opool, u->skip);
}

// forward from the BF16 input already in u->inb, or from the FP32 input in32 (rounded to BF16
// as the first layer stages it); compact = write only channel 0, into out0 (default u->out0)
int forward(efs_unet_s* u, bool compact = false, const float* in32 = nullptr, float* out0 = nullptr) {
    int rc = prep_weights(u);
    if (rc) return rc;
    const int D = u->D, H = u->H, W = u->W;
    const int* c = u->widths;
    const auto& L = u->layers;
    if (in32) conv_fwd<float>(u, L[0], in32, D, H, W, nullptr, u->act[0]);
    else conv_fwd<uint16_t>(u, L[0], u->inb, D, H, W, nullptr, u->act[0]);
    // the encoder's second convolutions also write the pooled input of the next level
    conv_fwd<uint16_t>(u, L[1], u->act[0], D, H, W, nullptr, u->act[1], 0, nullptr, 0, u->p1);
    conv_fwd<uint16_t>(u, L[2], u->p1, D / 2, H / 2, W / 2, nullptr, u->act[2]);
    conv_fwd<uint16_t>(u, L[3], u->act[2], D / 2, H / 2, W / 2, nullptr, u->act[3], 0, nullptr, 0, u->p2);
    conv_fwd<uint16_t>(u, L[4], u->p2, D / 4, H / 4, W / 4, nullptr, u->act[4]);
    conv_fwd<uint16_t>(u, L[5], u->act[4], D / 4, H / 4, W / 4, nullptr, u->act[5], 0, nullptr, 0, u->p3);
    conv_fwd<uint16_t>(u, L[6], u->p3, D / 8, H / 8, W / 8, nullptr, u->act[6]);
    conv_fwd<uint16_t>(u, L[7], u->act[6], D / 8, H / 8, W / 8, nullptr, u->act[7]);
    // decoder: each layer reads up(coarser) ++ skip in place; training also materialises the
    // concatenations, which the weight gradients read
    conv_fwd<uint16_t>(u, L[8], u->act[5], D / 4, H / 4, W / 4, nullptr, u->act[8], 0, u->act[7], c[3]);
    conv_fwd<uint16_t>(u, L[9], u->act[3], D / 2, H / 2, W / 2, nullptr, u->act[9], 0, u->act[8], c[2]);
    conv_fwd<uint16_t>(u, L[10], u->act[1], D, H, W, nullptr, u->act[10], 0, u->act[9], c[1]);
    if (!compact) {
        hipLaunchKernelGGL(k_up_concat, dim3(nb((long long)D / 4 * H / 4 * W / 4 * (c[3] + c[2]))), dim3(256), 0, u->stream,
                           u->act[7], c[3], u->act[5], c[2], D / 4, H / 4, W / 4, u->cat2);
        hipLaunchKernelGGL(k_up_concat, dim3(nb((long long)D / 2 * H / 2 * W / 2 * (c[2] + c[1]))), dim3(256), 0, u->stream,
                           u->act[8], c[2], u->act[3], c[1], D / 2, H / 2, W / 2, u->cat1);
        hipLaunchKernelGGL(k_up_concat, dim3(nb((long long)D * H * W * (c[1] + c[0]))), dim3(256), 0, u->stream, u->act[9], c[1],
                           u->act[1], c[0], D, H, W, u->cat0);
    }
    if (compact) conv_fwd<uint16_t>(u, L[11], u->act[10], D, H, W, out0 ? out0 : u->out0, nullptr, 1);
    else conv_fwd<uint16_t>(u, L[11], u->act[10], D, H, W, u->out, nullptr);
    UHIP(hipGetLastError());
    return 0;
}

// out[i] = sum over c of part[c * n + i] (i < n) in a fixed order, in two passes; part must have
// room for RED_SEG * n more floats after its nchunks * n.
void reduce_chunks(efs_unet_s* u, int n, int nchunks, float* part, float* out) {
    const int nseg = std::min(nchunks, RED_SEG);
    float* tmp = part + (size_t)nchunks * n;
    hipLaunchKernelGGL(k_reduce_seg, dim3(nb(n), nseg), dim3(256), 0, u->stream, n, nchunks, nseg, part, tmp);
    hipLaunchKernelGGL(k_reduce_chunks, dim3(nb(n)), dim3(256), 0, u->stream, n, nseg, tmp, out);
}

// Backward of one conv layer given dL/dy (FP32, post-activation gradient) in `gy`:
// weight/bias gradients accumulate into u->g; if gx != null, dL/dx is written (not accumulated).
template <typename TX>
int conv_bwd(efs_unet_s* u, const Conv& l, const TX* x, const uint16_t* y_act, const float* gy, int D, int H, int W, float* gx) {
    const long long nvox = (long long)D * H * W;
    hipLaunchKernelGGL(k_relu_bwd, dim3(nb(nvox * l.cout)), dim3(256), 0, u->stream, nvox * l.cout, gy, y_act, l.relu, u->dpre);
    if (l.ks != 3) {
        efs_set_error("U-Net weight gradients are implemented for 3x3x3 layers only");
        return -1;
    }
    // about 1024 blocks over the channel groups, each with at least one voxel tile
    const long long tiles = (long long)((W + 15) / 16) * ((H + CT_Y - 1) / CT_Y) * ((D + WG_ZT - 1) / WG_ZT);
    const int groups = (l.cout / 16) * (l.cin / 16);
    const int nchunks = (int)std::max<long long>(1, std::min<long long>(tiles, (1024 + groups - 1) / groups));
    // the weight partials and the bias partials, each followed by room for the segment sums of
    // its reduction
    const size_t wn = (size_t)(nchunks + RED_SEG) * l.nw(), need = wn + (size_t)(nchunks + RED_SEG) * l.cout;
    if (need > u->wpart_n) {
        if (u->wpart) hipFree(u->wpart);
        UHIP(hipMalloc(&u->wpart, need * 4));
        u->wpart_n = need;
    }
    float* bpart = u->wpart + wn;
    hipLaunchKernelGGL(k_conv_wgrad<TX>, dim3(groups * nchunks), dim3(256), 0, u->stream, u->dpre, l.cout, x, l.cin, D, H, W,
                       nchunks, u->wpart, bpart);
    reduce_chunks(u, (int)l.nw(), nchunks, u->wpart, u->g + l.woff);
    reduce_chunks(u, l.cout, nchunks, bpart, u->g + l.boff);
    if (gx)  // dx = conv(dpre, W flipped / transposed), no bias, no activation
        conv3(u->stream, u->dpre, l.cout, u->wt + l.woff, nullptr, l.cin, D, H, W, 0, gx, nullptr);
    UHIP(hipGetLastError());
    return 0;
}

// backward from dL/dout (FP32, 16 channels, in u->ga); fills u->g
int backward(efs_unet_s* u) {
    const int D = u->D, H = u->H, W = u->W;
    const int* c = u->widths;
    const auto& L = u->layers;
    const size_t v0 = (size_t)D * H * W, v1 = v0 / 8, v2 = v1 / 8, v3 = v2 / 8;
    auto zero = [&](float* p, size_t n) { hipLaunchKernelGGL(k_zero, dim3(nb(n)), dim3(256), 0, u->stream, (long long)n, p); };
    zero(u->g_e[0], v0 * c[0]);
    zero(u->g_e[1], v1 * c[1]);
    zero(u->g_e[2], v2 * c[2]);
    zero(u->g_e[3], v3 * c[3]);
    int rc;
    // out layer (11): x = act[10]
    if ((rc = conv_bwd<uint16_t>(u, L[11], u->act[10], nullptr, u->ga, D, H, W, u->gb))) return rc;
    // dec0 (10): y = act[10], x = cat0 (up(d1) ++ e0); gb = dL/d act[10]
    if ((rc = conv_bwd<uint16_t>(u, L[10], u->cat0, u->act[10], u->gb, D, H, W, u->ga))) return rc;
    zero(u->gb, v1 * c[1]);
    hipLaunchKernelGGL(k_up_concat_bwd_a, dim3(nb((long long)v1 * c[1])), dim3(256), 0, u->stream, u->ga, c[1], c[0], D, H, W, u->gb);
    hipLaunchKernelGGL(k_concat_bwd_b, dim3(nb((long long)v0 * c[0])), dim3(256), 0, u->stream, u->ga, c[1], c[0], (long long)v0,
                       u->g_e[0]);
    // dec1 (9): y = act[9] (gb holds dL/d act[9]), x = cat1 (up(d2) ++ e1)
    if ((rc = conv_bwd<uint16_t>(u, L[9], u->cat1, u->act[9], u->gb, D / 2, H / 2, W / 2, u->ga))) return rc;
    zero(u->gb, v2 * c[2]);
    hipLaunchKernelGGL(k_up_concat_bwd_a, dim3(nb((long long)v2 * c[2])), dim3(256), 0, u->stream, u->ga, c[2], c[1], D / 2, H / 2,
                       W / 2, u->gb);
    hipLaunchKernelGGL(k_concat_bwd_b, dim3(nb((long long)v1 * c[1])), dim3(256), 0, u->stream, u->ga, c[2], c[1], (long long)v1,
                       u->g_e[1]);
    // dec2 (8): y = act[8], x = cat2 (up(e3) ++ e2)
    if ((rc = conv_bwd<uint16_t>(u, L[8], u->cat2, u->act[8], u->gb, D / 4, H / 4, W / 4, u->ga))) return rc;
    hipLaunchKernelGGL(k_up_concat_bwd_a, dim3(nb((long long)v3 * c[3])), dim3(256), 0, u->stream, u->ga, c[3], c[2], D / 4, H / 4,
                       W / 4, u->g_e[3]);
    hipLaunchKernelGGL(k_concat_bwd_b, dim3(nb((long long)v2 * c[2])), dim3(256), 0, u->stream, u->ga, c[3], c[2], (long long)v2,
                       u->g_e[2]);
    // enc3b (7): y = act[7] = e3, x = act[6]
    if ((rc = conv_bwd<uint16_t>(u, L[7], u->act[6], u->act[7], u->g_e[3], D / 8, H / 8, W / 8, u->ga))) return rc;
    // enc3a (6): y = act[6], x = p3
    if ((rc = conv_bwd<uint16_t>(u, L[6], u->p3, u->act[6], u->ga, D / 8, H / 8, W / 8, u->gb))) return rc;
    hipLaunchKernelGGL(k_pool_bwd, dim3(nb((long long)v2 * c[2])), dim3(256), 0, u->stream, u->gb, c[2], D / 4, H / 4, W / 4,
                       u->g_e[2]);
    // enc2b (5), enc2a (4)
    if ((rc = conv_bwd<uint16_t>(u, L[5], u->act[4], u->act[5], u->g_e[2], D / 4, H / 4, W / 4, u->ga))) return rc;
    if ((rc = conv_bwd<uint16_t>(u, L[4], u->p2, u->act[4], u->ga, D / 4, H / 4, W / 4, u->gb))) return rc;
    hipLaunchKernelGGL(k_pool_bwd, dim3(nb((long long)v1 * c[1])), dim3(256), 0, u->stream, u->gb, c[1], D / 2, H / 2, W / 2,
                       u->g_e[1]);
    // enc1b (3), enc1a (2)
    if ((rc = conv_bwd<uint16_t>(u, L[3], u->act[2], u->act[3], u->g_e[1], D / 2, H / 2, W / 2, u->ga))) return rc;
    if ((rc = conv_bwd<uint16_t>(u, L[2], u->p1, u->act[2], u->ga, D / 2, H / 2, W / 2, u->gb))) return rc;
    hipLaunchKernelGGL(k_pool_bwd, dim3(nb((long long)v0 * c[0])), dim3(256), 0, u->stream, u->gb, c[0], D, H, W, u->g_e[0]);
    // enc0b (1), enc0a (0) (no input gradient needed)
    if ((rc = conv_bwd<uint16_t>(u, L[1], u->act[0], u->act[1], u->g_e[0], D, H, W, u->ga))) return rc;
    if ((rc = conv_bwd<uint16_t>(u, L[0], u->inb, u->act[0], u->ga, D, H, W, nullptr))) return rc;
    UHIP(hipGetLastError());
    return 0;
}

}  // namespace

// ======================================================================================
// C ABI
// ======================================================================================
extern "C" {

int efs_unet_create(const int32_t* widths, uint64_t seed, efs_unet* out) {
    if (!out) return -1;
    *out = nullptr;
    efs_unet_s* u = new efs_unet_s();
    const int def[4] = {16, 32, 48, 64};
    for (int i = 0; i < 4; i++) {
        u->widths[i] = widths ? widths[i] : def[i];
        if (u->widths[i] <= 0 || u->widths[i] % 16 || u->widths[i] > 128) {
            delete u;
            efs_set_error("U-Net widths must be positive multiples of 16, at most 128");
            return -1;
        }
    }
    u->layers = make_layers(u->widths);
    for (auto& l : u->layers)
        if (l.cin > 128 || l.cout > 128) {
            delete u;
            efs_set_error("U-Net layer wider than 128 channels");
            return -1;
        }
    u->nparams = u->layers.back().boff + u->layers.back().cout;
    hipError_t e = hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete u;
        return efs_hip_fail(e, "hipStreamCreateWithFlags");
    }
    const size_t n = u->nparams;
    if (hipMalloc(&u->w, n * 4) || hipMalloc(&u->g, n * 4) || hipMalloc(&u->m, n * 4) || hipMalloc(&u->v, n * 4) ||
        hipMalloc(&u->frozen, n * 4) || hipMalloc(&u->wb, n * 2) || hipMalloc(&u->wt, n * 2)) {
        efs_set_error("U-Net parameter allocation failed");
        return -2;
    }
    // He-style init from the counter generator; the unused output channels 1..15 of the last
    // layer start at zero and are frozen (their gradients are zero anyway)
    std::vector<float> h(n, 0.f), fr(n, 0.f);
    uint64_t ctr = 0;
    auto uni = [&]() {
        uint64_t z = seed * 0x9E3779B97F4A7C15ull + (++ctr) * 0xD1B54A32D192ED03ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * 0x1.0p-53;
    };
    for (size_t li = 0; li < u->layers.size(); li++) {
        const Conv& l = u->layers[li];
        const double fan_in = (double)(l.ks == 3 ? 27 : 1) * l.cin;
        const double a = std::sqrt(6.0 / fan_in) * (l.relu ? 1.0 : 0.5);
        for (size_t i = 0; i < l.nw(); i++) {
            const int co = (int)((i / l.cin) % l.cout);
            const bool dead = (li == u->layers.size() - 1) && co != 0;
            h[l.woff + i] = dead ? 0.f : (float)((2.0 * uni() - 1.0) * a);
            if (dead) fr[l.woff + i] = 1.f;
        }
        for (int co = 0; co < l.cout; co++) {
            const bool dead = (li == u->layers.size() - 1) && co != 0;
            h[l.boff + co] = 0.f;
            if (dead) fr[l.boff + co] = 1.f;
        }
    }
    hipMemcpy(u->w, h.data(), n * 4, hipMemcpyHostToDevice);
    hipMemcpy(u->frozen, fr.data(), n * 4, hipMemcpyHostToDevice);
    hipMemset(u->g, 0, n * 4);
    hipMemset(u->m, 0, n * 4);
    hipMemset(u->v, 0, n * 4);
    u->wdirty = true;
    *out = u;
    return 0;
}

int efs_unet_destroy(efs_unet u) {
    if (!u) return 0;
    hipStreamSynchronize(u->stream);
    for (void* p : u->bufs) hipFree(p);
    hipFree(u->w); hipFree(u->g); hipFree(u->m); hipFree(u->v); hipFree(u->frozen); hipFree(u->wb); hipFree(u->wt);
    if (u->wpart) hipFree(u->wpart);
    hipStreamDestroy(u->stream);
    delete u;
    return 0;
}

int64_t efs_unet_num_params(efs_unet u) { return u ? (int64_t)u->nparams : -1; }

int efs_unet_get_params(efs_unet u, float* host) {
    if (!u || !host) return -1;
    UHIP(hipMemcpyAsync(host, u->w, u->nparams * 4, hipMemcpyDeviceToHost, u->stream));
    UHIP(hipStreamSynchronize(u->stream));
    return 0;
}
int efs_unet_set_params(efs_unet u, const float* host) {
    if (!u || !host) return -1;
    UHIP(hipMemcpyAsync(u->w, host, u->nparams * 4, hipMemcpyHostToDevice, u->stream));
    UHIP(hipStreamSynchronize(u->stream));
    u->wdirty = true;
    return 0;
}
int efs_unet_get_grads(efs_unet u, float* host) {
    if (!u || !host) return -1;
    UHIP(hipMemcpyAsync(host, u->g, u->nparams * 4, hipMemcpyDeviceToHost, u->stream));
    UHIP(hipStreamSynchronize(u->stream));
    return 0;
}

// Input: device FP32 channels-last (D, H, W, 16).  Output: device FP32 (D, H, W) = channel 0.
int efs_unet_forward_device(efs_unet u, const float* in, int32_t D, int32_t H, int32_t W, float* out) {
    if (!u || !in || !out) return -1;
    int rc = unet_alloc_dims(u, D, H, W);
    if (rc) return rc;
    if ((rc = forward(u, true, in, out))) return rc;
    UHIP(hipStreamSynchronize(u->stream));
    return 0;
}

// Loss (mean over voxels with mask > 0 of mask * (y - t)^2) and its parameter gradient.
// in: (D,H,W,16), target: (D,H,W), mask: (D,H,W) or null; device FP32.  Gradients replace u->g.
int efs_unet_loss_grad_device(efs_unet u, const float* in, const float* target, const float* mask, int32_t D, int32_t H,
                              int32_t W, double* loss) {
    if (!u || !in || !target) return -1;
    int rc = unet_alloc_dims(u, D, H, W);
    if (rc) return rc;
    const long long nvox = (long long)D * H * W;
    hipLaunchKernelGGL(k_to_bf16, dim3(nb(nvox * 16)), dim3(256), 0, u->stream, nvox * 16, in, u->inb);
    if ((rc = forward(u))) return rc;
    const float inv = 1.f / (float)nvox;
    hipLaunchKernelGGL(k_mse, dim3(1024), dim3(256), 0, u->stream, nvox, 16, u->out, target, mask, inv, u->ga, u->lpart);
    if ((rc = backward(u))) return rc;
    std::vector<double> lp(1024);
    UHIP(hipMemcpyAsync(lp.data(), u->lpart, 1024 * 8, hipMemcpyDeviceToHost, u->stream));
    UHIP(hipStreamSynchronize(u->stream));
    double s = 0;
    for (double x : lp) s += x;
    if (loss) *loss = s / (double)nvox;
    return 0;
}

int efs_unet_adam_step(efs_unet u, float lr) {
    if (!u) return -1;
    u->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)u->adam_t), c2 = 1.f - std::pow(b2, (float)u->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(nb(u->nparams)), dim3(256), 0, u->stream, (int)u->nparams, u->w, u->g, u->m, u->v, lr, b1,
                       b2, 1e-8f, c1, c2, u->frozen);
    UHIP(hipGetLastError());
    u->wdirty = true;
    return 0;
}

int efs_unet_synchronize(efs_unet u) {
    if (!u) return -1;
    UHIP(hipStreamSynchronize(u->stream));
    return 0;
}

void* efs_unet_stream(efs_unet u) { return u ? (void*)u->stream : nullptr; }

}  // extern "C"

int efs_unet_forward_on(efs_unet u, hipStream_t stream, int D, int H, int W, const float** out16, const int* skip) {
    int rc = unet_alloc_dims(u, D, H, W);
    if (rc) return rc;
    hipStream_t saved = u->stream;
    u->stream = stream;
    u->skip = skip;
    rc = forward(u, true);
    u->skip = nullptr;
    u->stream = saved;
    if (rc) return rc;
    *out16 = u->out0;
    return 0;
}
int efs_unet_input_on(efs_unet u, int D, int H, int W, uint16_t** inb) {
    int rc = unet_alloc_dims(u, D, H, W);
    if (rc) return rc;
    *inb = u->inb;
    return 0;
}
