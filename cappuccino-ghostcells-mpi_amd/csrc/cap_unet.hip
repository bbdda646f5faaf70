// cap_unet.hip -- learned initial guess for the p' solves: a 3-D U-Net.
//
// No counterpart in the reference.  The network maps three channels on the
// interior cells -- the source, seen as the residual su - A p' of the
// previous p' (which is where a new source enters), the diagonal ap and the
// previous p' (each normalised) -- to the change of p'; the unchanged
// PCG-SIP solve then iterates from previous p' + change, and a guess that
// does not lower the L1 residual below the default start's is discarded.
//
// Architecture: 4 levels with 16 / 32 / 48 / 64 channels, two 3x3x3
// convolutions + ReLU per level, 2x2x2 average pooling down, nearest
// upsampling up with the encoder output concatenated, a 3x3x3 head to one
// channel.  Activations and weights are BF16, channels-last (voxel-major,
// channel stride padded to 8); every convolution accumulates in FP32.
//
// Kernels:
//  * k_conv: implicit-GEMM convolution on MFMA (v_mfma_f32_16x16x32_bf16):
//    a workgroup computes a 16 x 4 x 4 voxel brick for all output channels
//    from an LDS-staged halo brick of 32 input channels at a time; K = 27
//    taps x input channels in steps of 32.  Used forward, and with the
//    weights mirrored in the taps and transposed in the channels for the
//    gradient with respect to the input.
//  * k_wgrad: weight and bias gradients on MFMA (K = voxels of 16 x 4 x 4
//    bricks, staged transposed in LDS once for all 27 taps), one partial sum
//    per weight and block, added up in block order by k_wgrad_sum.
//  * pooling, upsampling, gradient combination with the ReLU mask, loss,
//    Adam.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <vector>
#include "cap_kernels.h"
#include "cap_amd.h"
#include "cap_ctx.h"

using namespace cap;

typedef unsigned short bf16_t;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__host__ __device__ inline bf16_t f2bf(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
    return (bf16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

namespace {

constexpr int NL = 15;   // convolution layers

// convolution layer: cin real input channels (stride cinp, padded to 8),
// cout outputs; kp = 27*cinp rounded up to 32 (MFMA K step); coutp = cout
// rounded up to 16 (MFMA N tile)
struct ConvL {
    int cin, cout, cinp, kp, coutp, relu, level;
    size_t woff, boff;      // offsets of weights [cout][27][cin] and bias in the parameters
    // backward-data packing: input channels cout (stride coutp8), outputs cin
    int coutp8, kpb, cinp16;
};

__device__ __forceinline__ void tap_off(int tap, int &dz, int &dy, int &dx)
{
    dz = tap / 9 - 1;
    dy = (tap / 3) % 3 - 1;
    dx = tap % 3 - 1;
}

// Y = conv(X, W) (+ bias, ReLU) on a D x H x W grid as an implicit GEMM on
// the matrix cores, with the input staged in LDS.
//  X: bf16 [nvox][cx], channels xoff .. xoff + cinp - 1 (cinp a multiple of 8)
//  Wt: bf16 [ntile*16][kp], row n = output channel, k = tap*cinp + ci
//  Y: bf16 [nvox][cy] at channel offset yoff (yf: fp32 [nvox][cout] instead)
//  mask (optional): Y *= (M[v][moff + co] > 0), M of stride cm
// A workgroup (4 waves) computes an output brick of BX x BY x BZ voxels; wave
// w the plane z = w of it as BY M-tiles of 16 voxels (one row y each), for
// all 16*NT output channels.  The input channels go in chunks of CC = 32:
// the chunk's halo brick ((BX+2) x (BY+2) x (BZ+2) voxels, zero outside the
// grid and beyond cinp) is staged in LDS once and read by all 27 taps, so
// each input value is read from memory about 1.7 times instead of 27.
constexpr int BX = 16, BY = 4;
constexpr int HX = BX + 2, HY = BY + 2;
constexpr int CC = 32;       // input channels per chunk (one MFMA K step)
constexpr int WP = 4;        // taps of weights in flight in k_conv
constexpr int VST = 40;      // LDS stride of a voxel in bf16 (80 bytes: spreads the banks)

template <int NT, int BZ>
__global__ __launch_bounds__(64 * BZ) void k_conv(int D, int H, int W, const bf16_t *__restrict__ X,
                                              int cx, int xoff, int cinp, const bf16_t *__restrict__ Wt,
                                              int kp, const float *__restrict__ bias, int relu,
                                              bf16_t *__restrict__ Y, int cy, int yoff, int cout,
                                              float *__restrict__ yf, const bf16_t *__restrict__ M,
                                              int cm, int moff)
{
    constexpr int HZ = BZ + 2, HV = HX * HY * HZ, NTH = 64 * BZ;
    __shared__ __attribute__((aligned(16))) bf16_t lds[HV * VST];
    const int nbx = (W + BX - 1) / BX, nby = (H + BY - 1) / BY;
    const int bid = blockIdx.x;
    const int x0 = (bid % nbx) * BX, y0 = ((bid / nbx) % nby) * BY, z0 = (bid / (nbx * nby)) * BZ;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int kg = lane >> 4, mrow = lane & 15;
    f32x4 acc[BY][NT];
#pragma unroll
    for (int m = 0; m < BY; ++m)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int nchunk = (cinp + CC - 1) / CC;
    for (int ch = 0; ch < nchunk; ++ch) {
        const int c0 = ch * CC;
        __syncthreads();   // the previous chunk's readers are done
        // the halo brick, in batches of up to 12 loads per thread, all issued
        // before their stores: every lane loads (the first element when its
        // item lies outside the grid or beyond cinp) and zeros replace those
        // values at the store
        constexpr int NQ = HV * (CC / 8), NQT = (NQ + NTH - 1) / NTH, NB = NQT < 12 ? NQT : 12;
        for (int q0 = 0; q0 < NQ; q0 += NB * NTH) {
            uint4 raw[NB];
            unsigned okb = 0u;
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                const int q = q0 + u * NTH + (int)threadIdx.x;
                const int hv = q / (CC / 8), g = q % (CC / 8);
                const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
                const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
                const int ci = c0 + 8 * g;
                const bool ok = q < NQ && ci < cinp && gx >= 0 && gx < W && gy >= 0 && gy < H && gz >= 0 && gz < D;
                okb |= (unsigned)ok << u;
                raw[u] = *reinterpret_cast<const uint4 *>(ok ? X + (((long)gz * H + gy) * W + gx) * cx + xoff + ci : X);
            }
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                const int q = q0 + u * NTH + (int)threadIdx.x;
                if (q < NQ)
                    *reinterpret_cast<uint4 *>(lds + (q / (CC / 8)) * VST + 8 * (q % (CC / 8))) =
                        (okb >> u) & 1u ? raw[u] : make_uint4(0u, 0u, 0u, 0u);
            }
        }
        __syncthreads();
        const bool kin = c0 + 8 * kg < cinp;
        // the weights of tap + WP are loaded while tap's MFMAs run (a ring of
        // WP taps in registers).  Every lane loads, from column 0 of its row
        // when its channel group lies beyond cinp, so the loads need no branch;
        // those lanes' activations in LDS are zero, so the weight they read
        // (a finite one) adds nothing
        uint4 wring[WP][NT];
        auto wload = [&](int tap, uint4 (&dst)[NT]) {
            const int k = kin ? tap * cinp + c0 + 8 * kg : 0;
#pragma unroll
            for (int t = 0; t < NT; ++t) dst[t] = *reinterpret_cast<const uint4 *>(Wt + (size_t)(t * 16 + mrow) * kp + k);
        };
#pragma unroll
        for (int q = 0; q < WP; ++q) wload(q, wring[q]);
#pragma unroll
        for (int tap = 0; tap < 27; ++tap) {
            int dz, dy, dx;
            tap_off(tap, dz, dy, dx);
            bf16x8 bv[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) memcpy(&bv[t], &wring[tap % WP][t], 16);
            if (tap + WP < 27) wload(tap + WP, wring[tap % WP]);
#pragma unroll
            for (int m = 0; m < BY; ++m) {
                const int hv = ((w + 1 + dz) * HY + (m + 1 + dy)) * HX + (mrow + 1 + dx);
                const uint4 ra = *reinterpret_cast<const uint4 *>(lds + hv * VST + 8 * kg);
                bf16x8 av;
                memcpy(&av, &ra, 16);
#pragma unroll
                for (int t = 0; t < NT; ++t)
                    acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv[t], acc[m][t], 0, 0, 0);
            }
        }
    }
    // D[row][n]: lane holds rows 4*(lane/16) + j (voxel x = x0 + row), column n = lane%16.
    // The bias and mask values are read for every lane (at clamped indices)
    // before they are used, a row m at a time, so their loads overlap.
    const int z = z0 + w;
    float bvv[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int co = t * 16 + mrow;
        bvv[t] = bias ? bias[co < cout ? co : cout - 1] : 0.f;
    }
    const int zc = z < D ? z : D - 1;
#pragma unroll
    for (int m = 0; m < BY; ++m) {
        const int y = y0 + m, yc = y < H ? y : H - 1;
        bf16_t mv[NT][4];
        if (M) {
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int co = t * 16 + mrow, coc = co < cout ? co : cout - 1;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int x = x0 + 4 * kg + j, xc = x < W ? x : W - 1;
                    mv[t][j] = M[(((long)zc * H + yc) * W + xc) * cm + moff + coc];
                }
            }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int co = t * 16 + mrow;
            if (co >= cout || y >= H || z >= D) continue;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int x = x0 + 4 * kg + j;
                if (x >= W) continue;
                const long v = ((long)z * H + y) * W + x;
                float r = acc[m][t][j] + bvv[t];
                if (relu) r = r > 0.f ? r : 0.f;
                if (M) r = bf2f(mv[t][j]) > 0.f ? r : 0.f;
                if (yf) yf[v * cout + co] = r;
                else Y[v * cy + yoff + co] = f2bf(r);
            }
        }
    }
}

int conv_grid(int D, int H, int W, int bz)
{
    return ((W + BX - 1) / BX) * ((H + BY - 1) / BY) * ((D + bz - 1) / bz);
}

// bricks 4 planes deep (4 waves) where they give enough workgroups, else 1
// plane (1 wave): small grids are latency-bound and want more workgroups
int conv_bz(int D, int H, int W) { return conv_grid(D, H, W, 4) >= 512 ? 4 : 1; }

// dW[co][tap][ci] += sum_v G[v][co] X[v + off(tap)][ci] on the matrix cores
// (C = A B with M = co, N = ci, K = voxels).  The voxels go in bricks of
// 16 x 4 x 4 (x fastest; K = the brick's 256 voxels in 8 steps of 32).  A
// block (4 waves) takes one tile of 16 input channels and a strided set of
// bricks; for each brick it stages G transposed (gt[co][voxel]) and the X
// halo brick of its 16 channels transposed, in three copies shifted by
// dx = -1, 0, +1 (xs[dx + 1][ci][z][y][x], x = 0..15), so that for every tap a
// lane's 8 consecutive voxels of one channel are one aligned 16-byte read
// (one copy read at 2-byte offsets, in the LDS's unaligned mode, took twice
// as long).
// Each brick is read from memory once for all 27 taps.  Wave w accumulates
// the taps w, w + 4, ... for all NCO output tiles in registers across its
// bricks, and writes its sums to its own slot of a partials buffer, which
// k_wgrad_sum adds up in block order (no atomics, the same sums every run).
// Wave 0 of the blocks of the first channel tile also sums G for the bias
// gradient db[co], as G^T times a column of ones on the matrix cores.
constexpr int WBX = 16, WBY = 4, WBZ = 4, WBV = WBX * WBY * WBZ;   // brick, 256 voxels
constexpr int WHY = WBY + 2, WHZ = WBZ + 2, WHX = WBX + 2;          // halo extents
constexpr int XCS = WHZ * WHY * WBX + 8;   // LDS stride (bf16) of one channel of one copy (584:
                                           // 292 dwords, 4 mod 32, so 8 channels' reads miss each other's banks)
constexpr int GST = WBV + 8;               // LDS stride (bf16) of one co row of gt
constexpr int WTAPS = 7;                   // taps per wave (27 over 4 waves)

struct WBricks {
    int nbx, nby, nbz, nb;
};
__host__ __device__ inline WBricks wbricks(int D, int H, int W)
{
    WBricks r;
    r.nbx = (W + WBX - 1) / WBX;
    r.nby = (H + WBY - 1) / WBY;
    r.nbz = (D + WBZ - 1) / WBZ;
    r.nb = r.nbx * r.nby * r.nbz;
    return r;
}

// keep the first n of the eight bf16 values of v (n <= 0: none)
__device__ __forceinline__ uint4 keep8(uint4 v, int n)
{
    unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] &= (2 * k < n ? 0xFFFFu : 0u) | (2 * k + 1 < n ? 0xFFFF0000u : 0u);
    return make_uint4(w[0], w[1], w[2], w[3]);
}
__device__ __forceinline__ void put8(bf16_t *p, int stride, uint4 v)
{
    const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        p[(2 * k) * stride] = (bf16_t)(w[k] & 0xFFFFu);
        p[(2 * k + 1) * stride] = (bf16_t)(w[k] >> 16);
    }
}
constexpr int WXI = (WHZ * WHY * WHX * 2 + 255) / 256;   // halo items (voxel, 8 channels) per thread

template <int NCO>
__global__ __launch_bounds__(256) void k_wgrad(int D, int H, int W, const bf16_t *__restrict__ X,
                                               int cx, int xoff, int cin,
                                               const bf16_t *__restrict__ G, int cg, int cout,
                                               int ncit, int nbg, int want_bias,
                                               float *__restrict__ part)
{
    constexpr int NXI = WHZ * WHY * WHX * 2;          // halo items: voxel x two groups of 8 channels
    constexpr int GIT = WBV * NCO * 2 / 256;          // G items (voxel, 8 channels) per thread
    static_assert(WXI + GIT <= 32, "one validity bit per item in okm");
    __shared__ __attribute__((aligned(16))) bf16_t gt[NCO * 16 * GST];
    __shared__ __attribute__((aligned(16))) bf16_t xs[3 * 16 * XCS];
    const WBricks B = wbricks(D, H, W);
    const int cit = blockIdx.x % ncit, grp = blockIdx.x / ncit;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, kg = lane >> 4, col = lane & 15;
    const bool bias = want_bias && cit == 0;
    // this wave's taps w, w + 4, ...; wave 3 has six, its seventh repeats
    // tap 26 and is not added
    int toff[WTAPS];
#pragma unroll
    for (int i = 0; i < WTAPS; ++i) {
        const int tap = w + 4 * i < 27 ? w + 4 * i : 26;
        int dz, dy, dx;
        tap_off(tap, dz, dy, dx);
        toff[i] = (dx + 1) * 16 * XCS + (dz * WHY + dy) * WBX;
    }
    f32x4 acc[WTAPS][NCO], accb[NCO];
#pragma unroll
    for (int t = 0; t < NCO; ++t) {
        accb[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int i = 0; i < WTAPS; ++i) acc[i][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    bf16x8 ones;
#pragma unroll
    for (int j = 0; j < 8; ++j) ones[j] = (__bf16)1.0f;

    // a brick's values, read into registers one brick ahead of their use.
    // Every lane loads (an item outside the grid or the channels reads the
    // first element instead), so that the loads of a brick are in flight
    // together; okm marks the items that hold data, and the zeros (outside
    // the grid, from cin / cout on) are put in when the values go to LDS.
    uint4 xr[WXI], gr[GIT];
    unsigned okm = 0u;
    auto load = [&](int br) {
        const int bx = br % B.nbx, by = (br / B.nbx) % B.nby, bz = br / (B.nbx * B.nby);
        const int x0 = bx * WBX, y0 = by * WBY, z0 = bz * WBZ;
        okm = 0u;
#pragma unroll
        for (int i = 0; i < WXI; ++i) {
            const int q = threadIdx.x + 256 * i;
            const int hx = q % WHX, hy = (q / WHX) % WHY, hz = (q / (WHX * WHY)) % WHZ, g = q / (WHX * WHY * WHZ);
            const int c0 = 16 * cit + 8 * g;
            const int xx = x0 + hx - 1, yy = y0 + hy - 1, zz = z0 + hz - 1;
            const bool ok = q < NXI && c0 < cin && xx >= 0 && xx < W && yy >= 0 && yy < H && zz >= 0 && zz < D;
            okm |= (unsigned)ok << i;
            xr[i] = *reinterpret_cast<const uint4 *>(ok ? X + ((long)(zz * H + yy) * W + xx) * cx + xoff + c0 : X);
        }
#pragma unroll
        for (int i = 0; i < GIT; ++i) {
            const int q = threadIdx.x + 256 * i;
            const int v = q % WBV, c0 = 8 * (q / WBV);
            const int x = x0 + v % WBX, y = y0 + (v / WBX) % WBY, z = z0 + v / (WBX * WBY);
            const bool ok = c0 < cout && x < W && y < H && z < D;
            okm |= (unsigned)ok << (WXI + i);
            gr[i] = *reinterpret_cast<const uint4 *>(ok ? G + ((long)(z * H + y) * W + x) * cg + c0 : G);
        }
    };
    // registers -> LDS: G transposed (gt[co][voxel]); each halo voxel into
    // the copies whose rows hold it (copy c at row position hx - c)
    auto store = [&]() {
        const uint4 zero = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
        for (int i = 0; i < WXI; ++i) {
            const int q = threadIdx.x + 256 * i;
            if (q >= NXI) break;
            const int hx = q % WHX, hy = (q / WHX) % WHY, hz = (q / (WHX * WHY)) % WHZ, g = q / (WHX * WHY * WHZ);
            const uint4 val = (okm >> i) & 1u ? keep8(xr[i], cin - 16 * cit - 8 * g) : zero;
            const int row = (hz * WHY + hy) * WBX;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int x = hx - c;
                if (x >= 0 && x < WBX) put8(xs + (c * 16 + 8 * g) * XCS + row + x, XCS, val);
            }
        }
#pragma unroll
        for (int i = 0; i < GIT; ++i) {
            const int q = threadIdx.x + 256 * i;
            const int c0 = 8 * (q / WBV);
            const uint4 val = (okm >> (WXI + i)) & 1u ? keep8(gr[i], cout - c0) : zero;
            put8(gt + c0 * GST + q % WBV, GST, val);
        }
    };

    int br = grp;
    if (br < B.nb) load(br);
    for (; br < B.nb; br += nbg) {
        __syncthreads();   // the previous brick's readers are done
        store();
        __syncthreads();
        if (br + nbg < B.nb) load(br + nbg);   // in flight during this brick's products
#pragma unroll 2
        for (int st = 0; st < WBV / 32; ++st) {
            // K step st: voxels 32 st .. 32 st + 31, i.e. rows r = 2 st and
            // 2 st + 1 of the brick (r = 4 z + y); lane group kg takes row
            // 2 st + kg / 2, x = 8 (kg % 2) .. + 7
            bf16x8 av[NCO], bv[WTAPS];
#pragma unroll
            for (int t = 0; t < NCO; ++t) {
                const uint4 ra = *reinterpret_cast<const uint4 *>(gt + (t * 16 + col) * GST + 32 * st + 8 * kg);
                memcpy(&av[t], &ra, 16);
            }
            const int r = 2 * st + (kg >> 1), z = r / WBY, y = r % WBY;
            const bf16_t *xb = xs + col * XCS + ((z + 1) * WHY + (y + 1)) * WBX + 8 * (kg & 1);
#pragma unroll
            for (int i = 0; i < WTAPS; ++i) {
                const uint4 rb = *reinterpret_cast<const uint4 *>(xb + toff[i]);
                memcpy(&bv[i], &rb, 16);
            }
            if (bias && w == 0) {   // db: G^T times a column of ones
#pragma unroll
                for (int t = 0; t < NCO; ++t)
                    accb[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[t], ones, accb[t], 0, 0, 0);
            }
#pragma unroll
            for (int i = 0; i < WTAPS; ++i)
#pragma unroll
                for (int t = 0; t < NCO; ++t)
                    acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[t], bv[i], acc[i][t], 0, 0, 0);
        }
    }
    // D[row][n]: lane holds rows 4*(lane/16) + j (co), column lane%16 (ci);
    // the block's sums go to its own slot of the partials
    float *pw = part + (size_t)blockIdx.x * 27 * NCO * 256;
#pragma unroll
    for (int i = 0; i < WTAPS; ++i) {
        const int tap = w + 4 * i;
        if (tap >= 27) break;
#pragma unroll
        for (int t = 0; t < NCO; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) pw[((size_t)tap * NCO * 16 + t * 16 + 4 * kg + j) * 16 + col] = acc[i][t][j];
    }
    if (bias && w == 0 && col == 0) {
        float *pb = part + (size_t)ncit * nbg * 27 * NCO * 256 + (size_t)grp * NCO * 16;
#pragma unroll
        for (int t = 0; t < NCO; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) pb[t * 16 + 4 * kg + j] = accb[t][j];
    }
}

// dW[co][tap][ci] += the blocks' partial sums, db[co] += the bias partials of
// the first channel tile's blocks.  One block per (co, tap, channel tile) and
// per 16 bias rows: 16 columns x 16 slices of the partials (slice k sums the
// blocks k, k + 16, ... in order), then the slices in order; the same sums
// every run.
__global__ __launch_bounds__(256) void k_wgrad_sum(int cout, int cin, int nco16, int ncit, int nbg,
                                                   const float *__restrict__ part, float *__restrict__ dW,
                                                   float *__restrict__ db)
{
    __shared__ float red[16][17];
    const int col = threadIdx.x & 15, sl = threadIdx.x >> 4;
    const int nwb = cout * 27 * ncit;
    const int b = blockIdx.x;
    float sum = 0.f;
    if (b < nwb) {
        const int cit = b % ncit, tap = (b / ncit) % 27, co = b / (ncit * 27);
        const float *p = part + ((size_t)tap * nco16 + co) * 16 + col;
        const size_t gs = (size_t)27 * nco16 * 16;   // floats per block slot
#pragma unroll 8
        for (int g = sl; g < nbg; g += 16) sum += p[(size_t)(g * ncit + cit) * gs];
        red[sl][col] = sum;
        __syncthreads();
        if (sl == 0) {
            float s = 0.f;
            for (int k = 0; k < 16; ++k) s += red[k][col];
            const int ci = cit * 16 + col;
            if (ci < cin) dW[((size_t)co * 27 + tap) * cin + ci] += s;
        }
    } else {
        const int co = (b - nwb) * 16 + col;
        const float *pb = part + (size_t)ncit * nbg * 27 * nco16 * 16;
        if (co < cout) {
#pragma unroll 8
            for (int g = sl; g < nbg; g += 16) sum += pb[(size_t)g * nco16 + co];
        }
        red[sl][col] = sum;
        __syncthreads();
        if (sl == 0 && co < cout && db) {
            float s = 0.f;
            for (int k = 0; k < 16; ++k) s += red[k][col];
            db[co] += s;
        }
    }
}

// 2x2x2 average pooling: out[Dc][Hc][Wc][c] from in (stride ci, offset ioff)
__global__ void k_pool(int Dc, int Hc, int Wc, int C, const bf16_t *in, int cin_, int ioff,
                       bf16_t *out, int cout_)
{
    const long n = (long)Dc * Hc * Wc * C;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int c = (int)(q % C);
        const long v = q / C;
        const int x = (int)(v % Wc), y = (int)((v / Wc) % Hc), z = (int)(v / ((long)Wc * Hc));
        const int W = 2 * Wc, H = 2 * Hc;
        float s = 0.f;
        for (int a = 0; a < 8; ++a) {
            const long f = ((long)(2 * z + (a >> 2)) * H + (2 * y + ((a >> 1) & 1))) * W + 2 * x + (a & 1);
            s += bf2f(in[f * cin_ + ioff + c]);
        }
        out[v * cout_ + c] = f2bf(0.125f * s);
    }
}

// nearest upsampling: out[fine v][ooff + c] = in[coarse v][c]
__global__ void k_up(int D, int H, int W, int C, const bf16_t *in, int cin_, bf16_t *out, int cout_,
                     int ooff)
{
    // 8 channels (16 bytes) per thread: C, cin_, cout_ and ooff are multiples of 8
    const int C8 = C / 8;
    const long n = (long)D * H * W * C8;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int c = 8 * (int)(q % C8);
        const long v = q / C8;
        const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)(v / ((long)W * H));
        const long cv = ((long)(z >> 1) * (H >> 1) + (y >> 1)) * (W >> 1) + (x >> 1);
        *reinterpret_cast<uint4 *>(out + v * cout_ + ooff + c) =
            *reinterpret_cast<const uint4 *>(in + cv * cin_ + c);
    }
}

// gradient of a coarse tensor: g[v][c] = (A? sum of the 8 fine children of
// A[..][aoff + c] (upsampling) : 0) + (B? B[v][boff + c] : 0) + (P? pooled
// gradient P[parent][c] / 8 : 0), times (mask[v][moff + c] > 0)
__global__ void k_gcombine(int D, int H, int W, int C, bf16_t *g, int cg_, const bf16_t *A, int ca,
                           int aoff, const bf16_t *B, int cb, int boff, const bf16_t *P, int cp,
                           const bf16_t *mask, int cm, int moff)
{
    const long n = (long)D * H * W * C;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int c = (int)(q % C);
        const long v = q / C;
        const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)(v / ((long)W * H));
        float s = 0.f;
        if (A) {   // A lives on the grid twice as fine
            for (int a = 0; a < 8; ++a) {
                const long f = ((long)(2 * z + (a >> 2)) * (2 * H) + (2 * y + ((a >> 1) & 1))) * (2 * W) +
                               2 * x + (a & 1);
                s += bf2f(A[f * ca + aoff + c]);
            }
        }
        if (B) s += bf2f(B[v * cb + boff + c]);
        if (P) {   // P lives on the grid twice as coarse
            const long pv = ((long)(z >> 1) * (H >> 1) + (y >> 1)) * (W >> 1) + (x >> 1);
            s += 0.125f * bf2f(P[pv * cp + c]);
        }
        if (mask && !(bf2f(mask[v * cm + moff + c]) > 0.f)) s = 0.f;
        g[v * cg_ + c] = f2bf(s);
    }
}

// network input: 3 fp32 channel-major arrays -> bf16 [nvox][8]
__global__ void k_pack_in(long nvox, const float *in, float s0, float s1, float s2, bf16_t *x)
{
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox; v += (long)gridDim.x * blockDim.x) {
        x[v * 8 + 0] = f2bf(in[v] * s0);
        x[v * 8 + 1] = f2bf(in[nvox + v] * s1);
        x[v * 8 + 2] = f2bf(in[2 * nvox + v] * s2);
        for (int c = 3; c < 8; ++c) x[v * 8 + c] = 0;
    }
}

// loss = mean((out - target)^2) (per-block partials), gradient 2 (out - target) / n in bf16
__global__ void k_loss(long n, const float *out, const float *tgt, bf16_t *g, int cg_, float *part)
{
    __shared__ float sh[256];
    float s = 0.f;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x) {
        const float d = out[v] - tgt[v];
        s += d * d;
        g[v * cg_] = f2bf(2.f * d / (float)n);
        for (int c = 1; c < cg_; ++c) g[v * cg_ + c] = 0;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Adam on fp32 master parameters
__global__ void k_adam(long n, float *p, const float *g, float *m, float *v, float lr, float b1,
                       float b2, float eps, float c1, float c2)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// forward packing Wt[n][k] (k = tap*cinp + ci) and backward-data packing
// Wb[n = ci][k' = tap'*coutp8 + co] with tap' = 26 - tap
__global__ void k_pack_w(const float *w, int cin, int cout, int cinp, int kp, int coutp, bf16_t *Wt,
                         int coutp8, int kpb, int cinp16, bf16_t *Wb)
{
    const long nf = (long)coutp * kp, nb = (long)cinp16 * kpb;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < nf + nb; q += (long)gridDim.x * blockDim.x) {
        if (q < nf) {
            const int n = (int)(q / kp), k = (int)(q % kp);
            const int tap = k / cinp, ci = k % cinp;
            float val = 0.f;
            if (n < cout && tap < 27 && ci < cin) val = w[((size_t)n * 27 + tap) * cin + ci];
            Wt[q] = f2bf(val);
        } else {
            const long r = q - nf;
            const int n = (int)(r / kpb), k = (int)(r % kpb);
            const int tapp = k / coutp8, co = k % coutp8;
            float val = 0.f;
            if (n < cin && tapp < 27 && co < cout) val = w[((size_t)co * 27 + (26 - tapp)) * cin + n];
            Wb[r] = f2bf(val);
        }
    }
}

__global__ void k_fill_f(float *p, long n, float v)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

// L1 norm of the 7-point residual su - A fi on the solve range (reference
// layout, k, i, j = 3..n?mm, operand order of pcg-sip.f90:53-55), compensated
__global__ __launch_bounds__(256) void k_res_l1(int ni, int nj, int nk, const double *ae,
                                                const double *aw, const double *an, const double *as,
                                                const double *at, const double *ab, const double *ap,
                                                const double *su, const double *fi, double *part)
{
    const int nij = ni * nj;
    const long n = (long)(ni - 4) * (nj - 4) * (nk - 4);
    Acc s;
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        const int j = 3 + (int)(q % (nj - 4)), i = 3 + (int)((q / (nj - 4)) % (ni - 4)),
                  k = 3 + (int)(q / ((long)(nj - 4) * (ni - 4)));
        const size_t c = (size_t)(k - 1) * nij + (size_t)(i - 1) * nj + (size_t)(j - 1);
        const double r = ae[c] * fi[c + nj] + aw[c] * fi[c - nj] + an[c] * fi[c + 1] +
                         as[c] * fi[c - 1] + at[c] * fi[c + nij] + ab[c] * fi[c - nij] + su[c] -
                         ap[c] * fi[c];
        s += fabs(r);
    }
    Acc t = wave_sum(s);
    __shared__ double hh[4], ll[4];
    if ((threadIdx.x & 63) == 0) { hh[threadIdx.x >> 6] = t.hi; ll[threadIdx.x >> 6] = t.lo; }
    __syncthreads();
    if (threadIdx.x == 0) {
        Acc b;
        for (int w = 0; w < 4; ++w) { Acc o; o.hi = hh[w]; o.lo = ll[w]; b += o; }
        part[2 * blockIdx.x] = b.hi;
        part[2 * blockIdx.x + 1] = b.lo;
    }
}

// interior cells of the reference layout <-> network voxels (x = j fastest,
// y = i, z = k).  mode 0: the three raw input channels -- the residual
// su - M x (M x = ap x - sum a_nb x_nb) of x = the previous-step p' (of the
// start fi when prev is null), ap and the previous-step p' (zero when prev
// is null); mode 1: the network's output
// into the padded direction array g (whose other cells stay zero); mode 2:
// fi += beta (prev - fi) + scale * g
__global__ void k_ref_io(int ni, int nj, int nk, int mode, const double *ae, const double *aw,
                         const double *an, const double *as, const double *at, const double *ab,
                         const double *ap, const double *su, double *fi, const double *prev,
                         float *in3, const float *out, double *g, double scale, double beta)
{
    const int nx = nj - 4, ny = ni - 4, nz = nk - 4, nij = ni * nj;
    const long n = (long)nx * ny * nz;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x) {
        const int x = (int)(v % nx), y = (int)((v / nx) % ny), z = (int)(v / ((long)nx * ny));
        const size_t c = (size_t)(z + 2) * nij + (size_t)(y + 2) * nj + (size_t)(x + 2);
        if (mode == 0) {
            const double *x = prev ? prev : fi;
            const double r = ae[c] * x[c + nj] + aw[c] * x[c - nj] + an[c] * x[c + 1] +
                             as[c] * x[c - 1] + at[c] * x[c + nij] + ab[c] * x[c - nij] + su[c] -
                             ap[c] * x[c];
            in3[v] = (float)r;
            in3[n + v] = (float)ap[c];
            in3[2 * n + v] = prev ? (float)prev[c] : 0.f;
        } else if (mode == 1) {
            g[c] = (double)out[v];
        } else {
            fi[c] = fi[c] + (prev ? beta * (prev[c] - fi[c]) : 0.0) + scale * g[c];
        }
    }
}

// block partials (compensated) of the normalisers: sum r^2, sum ap, sum prev^2
__global__ __launch_bounds__(256) void k_in_stats(long n, const float *in3, double *part)
{
    Acc s[3];
    for (long v = (long)blockIdx.x * 256 + threadIdx.x; v < n; v += (long)gridDim.x * 256) {
        const double r = in3[v], a = in3[n + v], q = in3[2 * n + v];
        s[0] += r * r;
        s[1] += a;
        s[2] += q * q;
    }
    __shared__ double hh[3][4], ll[3][4];
    for (int m = 0; m < 3; ++m) {
        const Acc t = wave_sum(s[m]);
        if ((threadIdx.x & 63) == 0) { hh[m][threadIdx.x >> 6] = t.hi; ll[m][threadIdx.x >> 6] = t.lo; }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        Acc b;
        for (int w = 0; w < 4; ++w) { Acc o; o.hi = hh[threadIdx.x][w]; o.lo = ll[threadIdx.x][w]; b += o; }
        part[6 * blockIdx.x + 2 * threadIdx.x] = b.hi;
        part[6 * blockIdx.x + 2 * threadIdx.x + 1] = b.lo;
    }
}

// block partials (compensated) of g.r and g.Mg over the solve range, r =
// su - M x: the energy-norm scale of the direction g as a correction of x
__global__ __launch_bounds__(256) void k_edots(int ni, int nj, int nk, const double *ae,
                                               const double *aw, const double *an, const double *as,
                                               const double *at, const double *ab, const double *ap,
                                               const double *su, const double *x, const double *g,
                                               double *part)
{
    const int nij = ni * nj;
    const long n = (long)(ni - 4) * (nj - 4) * (nk - 4);
    Acc s[2];
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        const int j = 3 + (int)(q % (nj - 4)), i = 3 + (int)((q / (nj - 4)) % (ni - 4)),
                  k = 3 + (int)(q / ((long)(nj - 4) * (ni - 4)));
        const size_t c = (size_t)(k - 1) * nij + (size_t)(i - 1) * nj + (size_t)(j - 1);
        const double r = ae[c] * x[c + nj] + aw[c] * x[c - nj] + an[c] * x[c + 1] +
                         as[c] * x[c - 1] + at[c] * x[c + nij] + ab[c] * x[c - nij] + su[c] -
                         ap[c] * x[c];
        const double mg = ap[c] * g[c] - (ae[c] * g[c + nj] + aw[c] * g[c - nj] + an[c] * g[c + 1] +
                                          as[c] * g[c - 1] + at[c] * g[c + nij] + ab[c] * g[c - nij]);
        s[0] += g[c] * r;
        s[1] += g[c] * mg;
    }
    __shared__ double hh[2][4], ll[2][4];
    for (int m = 0; m < 2; ++m) {
        const Acc t = wave_sum(s[m]);
        if ((threadIdx.x & 63) == 0) { hh[m][threadIdx.x >> 6] = t.hi; ll[m][threadIdx.x >> 6] = t.lo; }
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        Acc b;
        for (int w = 0; w < 4; ++w) { Acc o; o.hi = hh[threadIdx.x][w]; o.lo = ll[threadIdx.x][w]; b += o; }
        part[4 * blockIdx.x + 2 * threadIdx.x] = b.hi;
        part[4 * blockIdx.x + 2 * threadIdx.x + 1] = b.lo;
    }
}

// one damped-Jacobi sweep on M g = su - M x for the correction g of x, from
// sc * gin: gout = sc gin + w (su - M x - sc M gin) / ap on the solve range
// (the other cells of gout stay zero)
__global__ void k_jacobi(int ni, int nj, int nk, const double *ae, const double *aw,
                         const double *an, const double *as, const double *at, const double *ab,
                         const double *ap, const double *su, const double *x, const double *gin,
                         double *gout, double sc, double w)
{
    const int nij = ni * nj;
    const long n = (long)(ni - 4) * (nj - 4) * (nk - 4);
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int j = 3 + (int)(q % (nj - 4)), i = 3 + (int)((q / (nj - 4)) % (ni - 4)),
                  k = 3 + (int)(q / ((long)(nj - 4) * (ni - 4)));
        const size_t c = (size_t)(k - 1) * nij + (size_t)(i - 1) * nj + (size_t)(j - 1);
        const double r = ae[c] * x[c + nj] + aw[c] * x[c - nj] + an[c] * x[c + 1] +
                         as[c] * x[c - 1] + at[c] * x[c + nij] + ab[c] * x[c - nij] + su[c] -
                         ap[c] * x[c];
        const double *g = gin;
        const double mg = ap[c] * g[c] - (ae[c] * g[c + nj] + aw[c] * g[c - nj] + an[c] * g[c + 1] +
                                          as[c] * g[c - 1] + at[c] * g[c + nij] + ab[c] * g[c - nij]);
        gout[c] = sc * g[c] + w * (r - sc * mg) / ap[c];
    }
}

// block partials (compensated) over the solve range of the five sums of the
// least-squares fit of fi + beta (prev - fi) + alpha g to M x = su:
// r.Mg, Mg.Mg, r.Md, Md.Md, Md.Mg with r = su - M fi, M g = ap g - sum a_nb g_nb
// and M d = M (prev - fi) = r - r(prev) (d = 0 when prev is null)
__global__ __launch_bounds__(256) void k_alpha(int ni, int nj, int nk, const double *ae,
                                               const double *aw, const double *an, const double *as,
                                               const double *at, const double *ab, const double *ap,
                                               const double *su, const double *fi, const double *prev,
                                               const double *g, double *part)
{
    const int nij = ni * nj;
    const long n = (long)(ni - 4) * (nj - 4) * (nk - 4);
    Acc s[5];
    auto res = [&](const double *x, size_t c) {
        return ae[c] * x[c + nj] + aw[c] * x[c - nj] + an[c] * x[c + 1] + as[c] * x[c - 1] +
               at[c] * x[c + nij] + ab[c] * x[c - nij] + su[c] - ap[c] * x[c];
    };
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        const int j = 3 + (int)(q % (nj - 4)), i = 3 + (int)((q / (nj - 4)) % (ni - 4)),
                  k = 3 + (int)(q / ((long)(nj - 4) * (ni - 4)));
        const size_t c = (size_t)(k - 1) * nij + (size_t)(i - 1) * nj + (size_t)(j - 1);
        const double r = res(fi, c);
        const double mg = ap[c] * g[c] - (ae[c] * g[c + nj] + aw[c] * g[c - nj] + an[c] * g[c + 1] +
                                          as[c] * g[c - 1] + at[c] * g[c + nij] + ab[c] * g[c - nij]);
        const double md = prev ? r - res(prev, c) : 0.0;
        s[0] += r * mg;
        s[1] += mg * mg;
        s[2] += r * md;
        s[3] += md * md;
        s[4] += md * mg;
    }
    __shared__ double hh[5][4], ll[5][4];
    for (int m = 0; m < 5; ++m) {
        const Acc t = wave_sum(s[m]);
        if ((threadIdx.x & 63) == 0) { hh[m][threadIdx.x >> 6] = t.hi; ll[m][threadIdx.x >> 6] = t.lo; }
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        Acc b;
        for (int w = 0; w < 4; ++w) { Acc o; o.hi = hh[threadIdx.x][w]; o.lo = ll[threadIdx.x][w]; b += o; }
        part[10 * blockIdx.x + 2 * threadIdx.x] = b.hi;
        part[10 * blockIdx.x + 2 * threadIdx.x + 1] = b.lo;
    }
}

// damped-Jacobi sweeps (an even number: the result lands in gpad) and weight
constexpr int GUESS_SWEEPS = 4;
constexpr double GUESS_OMEGA = 0.8;

int grid_of(long n, int bs = 256) { return (int)std::max<long>(1, std::min<long>((n + bs - 1) / bs, 16384)); }

}  // namespace

struct cap_unet {
    cap_ctx *ctx;
    int nx, ny, nz;
    long nvox[4];
    ConvL L[NL];
    size_t nparam;
    float *p, *g, *m, *v;              // parameters, gradients, Adam moments (device)
    bf16_t *Wt[NL], *Wb[NL];           // packed weights
    int packed;                        // Wt/Wb match p
    long step;                         // Adam step count
    // activations (bf16, channels-last)
    bf16_t *in0, *a1, *c1, *p1, *a3, *c2, *p2, *a5, *c3, *p3, *a7, *a8, *a9, *a10, *a11, *a12,
        *a13, *a14;
    float *out, *lossp;
    // gradients
    bf16_t *g15, *g14, *g13, *gc1, *g12, *g11, *gc2, *g10, *g9, *gc3, *g8, *g7, *gp3, *g6, *g5,
        *gp2, *g4, *g3, *gp1, *g2, *g1;
    float *in3, *tgt;                  // staging of caller inputs (fp32)
    double *resp;                      // reduction partials
    float *wpart;                      // weight-gradient partials, one slot per block
    double *gpad, *gtmp;               // guess correction and a sweep buffer, reference layout
                                       // (allocated by the first guess)
    std::vector<void *> allocs;
};

namespace {

int ualloc(cap_unet *n, void **p, size_t bytes)
{
    if (hipMalloc(p, bytes) != hipSuccess) return CAP_ERR_NOMEM;
    n->allocs.push_back(*p);
    return CAP_OK;
}

#define UCHK(call)                                                 \
    do {                                                           \
        int _rc = (call);                                          \
        if (_rc) return _rc;                                       \
    } while (0)
#define HCHK(call)                                                 \
    do {                                                           \
        if ((call) != hipSuccess) return CAP_ERR_HIP;              \
    } while (0)

void conv_launch(cap_unet *n, int l, const bf16_t *X, int cx, int xoff, bf16_t *Y, int cy, int yoff,
                 float *yf)
{
    const ConvL &c = n->L[l];
    const int lv = c.level;
    const int D = n->nz >> lv, H = n->ny >> lv, W = n->nx >> lv;
    const int bz = conv_bz(D, H, W);
    const int grid = conv_grid(D, H, W, bz);
    const float *bias = n->p + c.boff;
    hipStream_t s = n->ctx->stream;
#define CONV_FWD(NT, BZ)                                                                            \
    hipLaunchKernelGGL((k_conv<NT, BZ>), dim3(grid), dim3(64 * BZ), 0, s, D, H, W, X, cx, xoff, c.cinp, \
                       n->Wt[l], c.kp, bias, c.relu, Y, cy, yoff, c.cout, yf, (const bf16_t *)nullptr, 0, 0)
    const int nt = c.coutp / 16;
    if (bz == 4) {
        switch (nt) { case 1: CONV_FWD(1, 4); break; case 2: CONV_FWD(2, 4); break;
                      case 3: CONV_FWD(3, 4); break; case 4: CONV_FWD(4, 4); break; }
    } else {
        switch (nt) { case 1: CONV_FWD(1, 1); break; case 2: CONV_FWD(2, 1); break;
                      case 3: CONV_FWD(3, 1); break; case 4: CONV_FWD(4, 1); break; }
    }
#undef CONV_FWD
}

// gradient with respect to the layer input: G (layer output gradient, stride
// coutp8) -> dX (cin channels, stride cdx), optionally masked
void conv_bwd_launch(cap_unet *n, int l, const bf16_t *G, bf16_t *dX, int cdx, const bf16_t *M,
                     int cm, int moff)
{
    const ConvL &c = n->L[l];
    const int lv = c.level;
    const int D = n->nz >> lv, H = n->ny >> lv, W = n->nx >> lv;
    const int bz = conv_bz(D, H, W);
    const int grid = conv_grid(D, H, W, bz);
    hipStream_t s = n->ctx->stream;
#define CONV_BWD(NT, BZ)                                                                            \
    hipLaunchKernelGGL((k_conv<NT, BZ>), dim3(grid), dim3(64 * BZ), 0, s, D, H, W, G, c.coutp8, 0,      \
                       c.coutp8, n->Wb[l], c.kpb, (const float *)nullptr, 0, dX, cdx, 0, c.cin,         \
                       (float *)nullptr, M, cm, moff)
    const int nt = c.cinp16 / 16;
    if (bz == 4) {
        switch (nt) { case 1: CONV_BWD(1, 4); break; case 2: CONV_BWD(2, 4); break;
                      case 3: CONV_BWD(3, 4); break; case 4: CONV_BWD(4, 4); break;
                      case 5: CONV_BWD(5, 4); break; case 7: CONV_BWD(7, 4); break; }
    } else {
        switch (nt) { case 1: CONV_BWD(1, 1); break; case 2: CONV_BWD(2, 1); break;
                      case 3: CONV_BWD(3, 1); break; case 4: CONV_BWD(4, 1); break;
                      case 5: CONV_BWD(5, 1); break; case 7: CONV_BWD(7, 1); break; }
    }
#undef CONV_BWD
}

// the weight-gradient launch of layer l: ncit channel tiles x nbg brick
// groups, about two blocks a CU (one for four output tiles, whose LDS holds
// one block a CU); returns the floats of partials it writes
size_t wgrad_shape(const cap_unet *n, int l, int &nco, int &ncit, int &nbg)
{
    const ConvL &c = n->L[l];
    const int D = n->nz >> c.level, H = n->ny >> c.level, W = n->nx >> c.level;
    nco = (c.cout + 15) / 16;
    ncit = (c.cin + 15) / 16;
    const WBricks B = wbricks(D, H, W);
    const int target = nco <= 2 ? 512 : 256;
    nbg = (target + ncit - 1) / ncit;
    nbg = nbg < B.nb ? nbg : B.nb;
    return (size_t)ncit * nbg * 27 * nco * 256 + (size_t)nbg * nco * 16;
}

void wgrad_launch(cap_unet *n, int l, const bf16_t *X, int cx, int xoff, const bf16_t *G)
{
    const ConvL &c = n->L[l];
    const int lv = c.level;
    const int D = n->nz >> lv, H = n->ny >> lv, W = n->nx >> lv;
    int nco, ncit, nbg;
    (void)wgrad_shape(n, l, nco, ncit, nbg);
    const dim3 grid((unsigned)(ncit * nbg));
    hipStream_t s = n->ctx->stream;
#define WGRAD(NCO)                                                                                  \
    hipLaunchKernelGGL(k_wgrad<NCO>, grid, dim3(256), 0, s, D, H, W, X, cx, xoff, c.cin, G, c.coutp8, \
                       c.cout, ncit, nbg, 1, n->wpart)
    switch (nco) {
    case 1: WGRAD(1); break;
    case 2: WGRAD(2); break;
    case 3: WGRAD(3); break;
    case 4: WGRAD(4); break;
    }
#undef WGRAD
    hipLaunchKernelGGL(k_wgrad_sum, dim3((unsigned)(c.cout * 27 * ncit + nco)), dim3(256), 0, s, c.cout,
                       c.cin, nco * 16, ncit, nbg, (const float *)n->wpart, n->g + c.woff, n->g + c.boff);
}

int pack_weights(cap_unet *n)
{
    if (n->packed) return CAP_OK;
    for (int l = 0; l < NL; ++l) {
        const ConvL &c = n->L[l];
        const long tot = (long)c.coutp * c.kp + (long)c.cinp16 * c.kpb;
        hipLaunchKernelGGL(k_pack_w, dim3(grid_of(tot)), dim3(256), 0, n->ctx->stream, n->p + c.woff,
                           c.cin, c.cout, c.cinp, c.kp, c.coutp, n->Wt[l], c.coutp8, c.kpb, c.cinp16,
                           n->Wb[l]);
    }
    HCHK(hipGetLastError());
    n->packed = 1;
    return CAP_OK;
}

void pool_launch(cap_unet *n, int lvc, int C, const bf16_t *in, int cin_, int ioff, bf16_t *out)
{
    const int D = n->nz >> lvc, H = n->ny >> lvc, W = n->nx >> lvc;
    hipLaunchKernelGGL(k_pool, dim3(grid_of((long)D * H * W * C)), dim3(256), 0, n->ctx->stream, D, H,
                       W, C, in, cin_, ioff, out, C);
}

void up_launch(cap_unet *n, int lvf, int C, const bf16_t *in, int cin_, bf16_t *out, int cout_)
{
    const int D = n->nz >> lvf, H = n->ny >> lvf, W = n->nx >> lvf;
    hipLaunchKernelGGL(k_up, dim3(grid_of((long)D * H * W * C)), dim3(256), 0, n->ctx->stream, D, H, W,
                       C, in, cin_, out, cout_, 0);
}

void gcomb_launch(cap_unet *n, int lv, int C, bf16_t *g, int cg_, const bf16_t *A, int ca, int aoff,
                  const bf16_t *B, int cb, int boff, const bf16_t *P, int cp, const bf16_t *mask,
                  int cm, int moff)
{
    const int D = n->nz >> lv, H = n->ny >> lv, W = n->nx >> lv;
    hipLaunchKernelGGL(k_gcombine, dim3(grid_of((long)D * H * W * C)), dim3(256), 0, n->ctx->stream, D,
                       H, W, C, g, cg_, A, ca, aoff, B, cb, boff, P, cp, mask, cm, moff);
}

// input staging: device pointers as given, host pointers copied
int stage_f32(cap_unet *n, const float *src, float *dst, long count, const float **out)
{
    if (n->ctx->flags & CAP_DEVICE_POINTERS) { *out = src; return CAP_OK; }
    HCHK(hipMemcpyAsync(dst, src, count * sizeof(float), hipMemcpyHostToDevice, n->ctx->stream));
    *out = dst;
    return CAP_OK;
}

int forward(cap_unet *n, const float *in, float s0 = 1.f, float s1 = 1.f, float s2 = 1.f)
{
    UCHK(pack_weights(n));
    const long n0 = n->nvox[0];
    hipStream_t s = n->ctx->stream;
    // the channels are multiplied by s0, s1, s2 on the way in
    hipLaunchKernelGGL(k_pack_in, dim3(grid_of(n0)), dim3(256), 0, s, n0, in, s0, s1, s2, n->in0);
    conv_launch(n, 0, n->in0, 8, 0, n->a1, 16, 0, nullptr);           // 3 -> 16
    conv_launch(n, 1, n->a1, 16, 0, n->c1, 48, 32, nullptr);           // 16 -> 16 (skip 1)
    pool_launch(n, 1, 16, n->c1, 48, 32, n->p1);
    conv_launch(n, 2, n->p1, 16, 0, n->a3, 32, 0, nullptr);            // 16 -> 32
    conv_launch(n, 3, n->a3, 32, 0, n->c2, 80, 48, nullptr);           // 32 -> 32 (skip 2)
    pool_launch(n, 2, 32, n->c2, 80, 48, n->p2);
    conv_launch(n, 4, n->p2, 32, 0, n->a5, 48, 0, nullptr);            // 32 -> 48
    conv_launch(n, 5, n->a5, 48, 0, n->c3, 112, 64, nullptr);          // 48 -> 48 (skip 3)
    pool_launch(n, 3, 48, n->c3, 112, 64, n->p3);
    conv_launch(n, 6, n->p3, 48, 0, n->a7, 64, 0, nullptr);            // 48 -> 64
    conv_launch(n, 7, n->a7, 64, 0, n->a8, 64, 0, nullptr);            // 64 -> 64
    up_launch(n, 2, 64, n->a8, 64, n->c3, 112);
    conv_launch(n, 8, n->c3, 112, 0, n->a9, 48, 0, nullptr);           // 112 -> 48
    conv_launch(n, 9, n->a9, 48, 0, n->a10, 48, 0, nullptr);           // 48 -> 48
    up_launch(n, 1, 48, n->a10, 48, n->c2, 80);
    conv_launch(n, 10, n->c2, 80, 0, n->a11, 32, 0, nullptr);          // 80 -> 32
    conv_launch(n, 11, n->a11, 32, 0, n->a12, 32, 0, nullptr);         // 32 -> 32
    up_launch(n, 0, 32, n->a12, 32, n->c1, 48);
    conv_launch(n, 12, n->c1, 48, 0, n->a13, 16, 0, nullptr);          // 48 -> 16
    conv_launch(n, 13, n->a13, 16, 0, n->a14, 16, 0, nullptr);         // 16 -> 16
    conv_launch(n, 14, n->a14, 16, 0, nullptr, 0, 0, n->out);          // 16 -> 1 (fp32)
    HCHK(hipGetLastError());
    return CAP_OK;
}

// gradients of mean((out - tgt)^2) into n->g; partial losses into n->lossp
int backward(cap_unet *n, const float *tgt, int *nlossp)
{
    const long n0 = n->nvox[0];
    hipStream_t s = n->ctx->stream;
    HCHK(hipMemsetAsync(n->g, 0, n->nparam * sizeof(float), s));
    const int lb = grid_of(n0);
    hipLaunchKernelGGL(k_loss, dim3(lb), dim3(256), 0, s, n0, n->out, tgt, n->g15, 8, n->lossp);
    *nlossp = lb;
    // head
    wgrad_launch(n, 14, n->a14, 16, 0, n->g15);
    conv_bwd_launch(n, 14, n->g15, n->g14, 16, n->a14, 16, 0);
    wgrad_launch(n, 13, n->a13, 16, 0, n->g14);
    conv_bwd_launch(n, 13, n->g14, n->g13, 16, n->a13, 16, 0);
    wgrad_launch(n, 12, n->c1, 48, 0, n->g13);
    conv_bwd_launch(n, 12, n->g13, n->gc1, 48, nullptr, 0, 0);
    gcomb_launch(n, 1, 32, n->g12, 32, n->gc1, 48, 0, nullptr, 0, 0, nullptr, 0, n->a12, 32, 0);
    wgrad_launch(n, 11, n->a11, 32, 0, n->g12);
    conv_bwd_launch(n, 11, n->g12, n->g11, 32, n->a11, 32, 0);
    wgrad_launch(n, 10, n->c2, 80, 0, n->g11);
    conv_bwd_launch(n, 10, n->g11, n->gc2, 80, nullptr, 0, 0);
    gcomb_launch(n, 2, 48, n->g10, 48, n->gc2, 80, 0, nullptr, 0, 0, nullptr, 0, n->a10, 48, 0);
    wgrad_launch(n, 9, n->a9, 48, 0, n->g10);
    conv_bwd_launch(n, 9, n->g10, n->g9, 48, n->a9, 48, 0);
    wgrad_launch(n, 8, n->c3, 112, 0, n->g9);
    conv_bwd_launch(n, 8, n->g9, n->gc3, 112, nullptr, 0, 0);
    gcomb_launch(n, 3, 64, n->g8, 64, n->gc3, 112, 0, nullptr, 0, 0, nullptr, 0, n->a8, 64, 0);
    wgrad_launch(n, 7, n->a7, 64, 0, n->g8);
    conv_bwd_launch(n, 7, n->g8, n->g7, 64, n->a7, 64, 0);
    wgrad_launch(n, 6, n->p3, 48, 0, n->g7);
    conv_bwd_launch(n, 6, n->g7, n->gp3, 48, nullptr, 0, 0);
    gcomb_launch(n, 2, 48, n->g6, 48, nullptr, 0, 0, n->gc3, 112, 64, n->gp3, 48, n->c3, 112, 64);
    wgrad_launch(n, 5, n->a5, 48, 0, n->g6);
    conv_bwd_launch(n, 5, n->g6, n->g5, 48, n->a5, 48, 0);
    wgrad_launch(n, 4, n->p2, 32, 0, n->g5);
    conv_bwd_launch(n, 4, n->g5, n->gp2, 32, nullptr, 0, 0);
    gcomb_launch(n, 1, 32, n->g4, 32, nullptr, 0, 0, n->gc2, 80, 48, n->gp2, 32, n->c2, 80, 48);
    wgrad_launch(n, 3, n->a3, 32, 0, n->g4);
    conv_bwd_launch(n, 3, n->g4, n->g3, 32, n->a3, 32, 0);
    wgrad_launch(n, 2, n->p1, 16, 0, n->g3);
    conv_bwd_launch(n, 2, n->g3, n->gp1, 16, nullptr, 0, 0);
    gcomb_launch(n, 0, 16, n->g2, 16, nullptr, 0, 0, n->gc1, 48, 32, n->gp1, 16, n->c1, 48, 32);
    wgrad_launch(n, 1, n->a1, 16, 0, n->g2);
    conv_bwd_launch(n, 1, n->g2, n->g1, 16, n->a1, 16, 0);
    wgrad_launch(n, 0, n->in0, 8, 0, n->g1);
    HCHK(hipGetLastError());
    return CAP_OK;
}

int read_loss(cap_unet *n, int np, float *loss)
{
    std::vector<float> h(np);
    HCHK(hipMemcpyAsync(h.data(), n->lossp, np * sizeof(float), hipMemcpyDeviceToHost, n->ctx->stream));
    HCHK(hipStreamSynchronize(n->ctx->stream));
    double s = 0.0;
    for (int i = 0; i < np; ++i) s += h[i];
    *loss = (float)(s / (double)n->nvox[0]);
    return CAP_OK;
}

uint64_t smx(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

}  // namespace

extern "C" int cap_unet_create(cap_ctx *ctx, int nx, int ny, int nz, unsigned long long seed,
                               cap_unet **out)
{
    if (!ctx || !out || nx < 8 || ny < 8 || nz < 8 || nx % 8 || ny % 8 || nz % 8) return CAP_ERR_ARG;
    cap_unet *n = new (std::nothrow) cap_unet();
    if (!n) return CAP_ERR_NOMEM;
    n->ctx = ctx;
    n->nx = nx; n->ny = ny; n->nz = nz;
    for (int l = 0; l < 4; ++l) n->nvox[l] = (long)(nx >> l) * (ny >> l) * (nz >> l);
    // layer table: cin, cout, level, relu
    const int spec[NL][4] = {{3, 16, 0, 1},  {16, 16, 0, 1}, {16, 32, 1, 1}, {32, 32, 1, 1},
                             {32, 48, 2, 1}, {48, 48, 2, 1}, {48, 64, 3, 1}, {64, 64, 3, 1},
                             {112, 48, 2, 1}, {48, 48, 2, 1}, {80, 32, 1, 1}, {32, 32, 1, 1},
                             {48, 16, 0, 1}, {16, 16, 0, 1}, {16, 1, 0, 0}};
    size_t off = 0;
    for (int l = 0; l < NL; ++l) {
        ConvL &c = n->L[l];
        c.cin = spec[l][0]; c.cout = spec[l][1]; c.level = spec[l][2]; c.relu = spec[l][3];
        c.cinp = (c.cin + 7) / 8 * 8;
        c.kp = (27 * c.cinp + 31) / 32 * 32;
        c.coutp = (c.cout + 15) / 16 * 16;
        c.coutp8 = (c.cout + 7) / 8 * 8;
        c.kpb = (27 * c.coutp8 + 31) / 32 * 32;
        c.cinp16 = (c.cin + 15) / 16 * 16;
        c.woff = off;
        off += (size_t)c.cout * 27 * c.cin;
        c.boff = off;
        off += (size_t)c.cout;
    }
    n->nparam = off;
    int rc;
    auto A = [&](void **p, size_t bytes) { return ualloc(n, p, bytes); };
    const long n0 = n->nvox[0], n1 = n->nvox[1], n2 = n->nvox[2], n3 = n->nvox[3];
    const size_t B = sizeof(bf16_t);
    if ((rc = A((void **)&n->p, off * 4)) || (rc = A((void **)&n->g, off * 4)) ||
        (rc = A((void **)&n->m, off * 4)) || (rc = A((void **)&n->v, off * 4)))
        goto fail;
    for (int l = 0; l < NL; ++l) {
        const ConvL &c = n->L[l];
        if ((rc = A((void **)&n->Wt[l], (size_t)c.coutp * c.kp * B)) ||
            (rc = A((void **)&n->Wb[l], (size_t)c.cinp16 * c.kpb * B)))
            goto fail;
    }
    if ((rc = A((void **)&n->in0, n0 * 8 * B)) || (rc = A((void **)&n->a1, n0 * 16 * B)) ||
        (rc = A((void **)&n->c1, n0 * 48 * B)) || (rc = A((void **)&n->p1, n1 * 16 * B)) ||
        (rc = A((void **)&n->a3, n1 * 32 * B)) || (rc = A((void **)&n->c2, n1 * 80 * B)) ||
        (rc = A((void **)&n->p2, n2 * 32 * B)) || (rc = A((void **)&n->a5, n2 * 48 * B)) ||
        (rc = A((void **)&n->c3, n2 * 112 * B)) || (rc = A((void **)&n->p3, n3 * 48 * B)) ||
        (rc = A((void **)&n->a7, n3 * 64 * B)) || (rc = A((void **)&n->a8, n3 * 64 * B)) ||
        (rc = A((void **)&n->a9, n2 * 48 * B)) || (rc = A((void **)&n->a10, n2 * 48 * B)) ||
        (rc = A((void **)&n->a11, n1 * 32 * B)) || (rc = A((void **)&n->a12, n1 * 32 * B)) ||
        (rc = A((void **)&n->a13, n0 * 16 * B)) || (rc = A((void **)&n->a14, n0 * 16 * B)) ||
        (rc = A((void **)&n->out, n0 * 4)) || (rc = A((void **)&n->lossp, 16384 * 4)) ||
        (rc = A((void **)&n->g15, n0 * 8 * B)) || (rc = A((void **)&n->g14, n0 * 16 * B)) ||
        (rc = A((void **)&n->g13, n0 * 16 * B)) || (rc = A((void **)&n->gc1, n0 * 48 * B)) ||
        (rc = A((void **)&n->g12, n1 * 32 * B)) || (rc = A((void **)&n->g11, n1 * 32 * B)) ||
        (rc = A((void **)&n->gc2, n1 * 80 * B)) || (rc = A((void **)&n->g10, n2 * 48 * B)) ||
        (rc = A((void **)&n->g9, n2 * 48 * B)) || (rc = A((void **)&n->gc3, n2 * 112 * B)) ||
        (rc = A((void **)&n->g8, n3 * 64 * B)) || (rc = A((void **)&n->g7, n3 * 64 * B)) ||
        (rc = A((void **)&n->gp3, n3 * 48 * B)) || (rc = A((void **)&n->g6, n2 * 48 * B)) ||
        (rc = A((void **)&n->g5, n2 * 48 * B)) || (rc = A((void **)&n->gp2, n2 * 32 * B)) ||
        (rc = A((void **)&n->g4, n1 * 32 * B)) || (rc = A((void **)&n->g3, n1 * 32 * B)) ||
        (rc = A((void **)&n->gp1, n1 * 16 * B)) || (rc = A((void **)&n->g2, n0 * 16 * B)) ||
        (rc = A((void **)&n->g1, n0 * 16 * B)) || (rc = A((void **)&n->in3, n0 * 3 * 4)) ||
        (rc = A((void **)&n->tgt, n0 * 4)) || (rc = A((void **)&n->resp, 16384 * 2 * 8)))
        goto fail;
    {
        size_t wp = 0;
        for (int l = 0; l < NL; ++l) {
            int nco, ncit, nbg;
            const size_t f = wgrad_shape(n, l, nco, ncit, nbg);
            wp = f > wp ? f : wp;
        }
        if ((rc = A((void **)&n->wpart, wp * 4))) goto fail;
    }
    {   // He initialisation (counter-based normal deviates), zero biases
        std::vector<float> h(off, 0.f);
        for (int l = 0; l < NL; ++l) {
            const ConvL &c = n->L[l];
            const double sd = (l == NL - 1 ? 0.1 : 1.0) * sqrt(2.0 / (27.0 * c.cin));
            for (size_t i = 0; i < (size_t)c.cout * 27 * c.cin; ++i) {
                const uint64_t k = smx(seed ^ smx((uint64_t)l * 0x100000001ull + i));
                const double u1 = ((k >> 11) + 0.5) * 0x1.0p-53, u2 = ((smx(k) >> 11) + 0.5) * 0x1.0p-53;
                h[c.woff + i] = (float)(sd * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
            }
        }
        if (hipMemcpy(n->p, h.data(), off * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemset(n->m, 0, off * 4) != hipSuccess || hipMemset(n->v, 0, off * 4) != hipSuccess) {
            rc = CAP_ERR_HIP;
            goto fail;
        }
    }
    // concat buffers are written piecewise: clear them (padding channels too)
    if (hipMemset(n->c1, 0, n0 * 48 * B) != hipSuccess || hipMemset(n->c2, 0, n1 * 80 * B) != hipSuccess ||
        hipMemset(n->c3, 0, n2 * 112 * B) != hipSuccess || hipMemset(n->in0, 0, n0 * 8 * B) != hipSuccess) {
        rc = CAP_ERR_HIP;
        goto fail;
    }
    n->packed = 0;
    n->step = 0;
    *out = n;
    return CAP_OK;
fail:
    for (void *p : n->allocs) (void)hipFree(p);
    delete n;
    return rc;
}

extern "C" int cap_unet_destroy(cap_unet *n)
{
    if (!n) return CAP_OK;
    (void)hipStreamSynchronize(n->ctx->stream);
    for (void *p : n->allocs) (void)hipFree(p);
    delete n;
    return CAP_OK;
}

extern "C" int cap_unet_num_params(const cap_unet *n, long *count)
{
    if (!n || !count) return CAP_ERR_ARG;
    *count = (long)n->nparam;
    return CAP_OK;
}

extern "C" int cap_unet_get_params(cap_unet *n, float *p)
{
    if (!n || !p) return CAP_ERR_ARG;
    HCHK(hipMemcpyAsync(p, n->p, n->nparam * 4, hipMemcpyDeviceToHost, n->ctx->stream));
    HCHK(hipStreamSynchronize(n->ctx->stream));
    return CAP_OK;
}

extern "C" int cap_unet_set_params(cap_unet *n, const float *p)
{
    if (!n || !p) return CAP_ERR_ARG;
    HCHK(hipMemcpyAsync(n->p, p, n->nparam * 4, hipMemcpyHostToDevice, n->ctx->stream));
    HCHK(hipStreamSynchronize(n->ctx->stream));
    n->packed = 0;
    return CAP_OK;
}

extern "C" int cap_unet_forward(cap_unet *n, const float *in, float *out)
{
    if (!n || !in || !out) return CAP_ERR_ARG;
    const long n0 = n->nvox[0];
    const float *din;
    UCHK(stage_f32(n, in, n->in3, 3 * n0, &din));
    UCHK(forward(n, din));
    if (n->ctx->flags & CAP_DEVICE_POINTERS)
        HCHK(hipMemcpyAsync(out, n->out, n0 * 4, hipMemcpyDeviceToDevice, n->ctx->stream));
    else
        HCHK(hipMemcpyAsync(out, n->out, n0 * 4, hipMemcpyDeviceToHost, n->ctx->stream));
    HCHK(hipStreamSynchronize(n->ctx->stream));
    return CAP_OK;
}

extern "C" int cap_unet_gradients(cap_unet *n, const float *in, const float *target, float *loss,
                                  float *grads)
{
    if (!n || !in || !target || !loss) return CAP_ERR_ARG;
    const long n0 = n->nvox[0];
    const float *din, *dt;
    UCHK(stage_f32(n, in, n->in3, 3 * n0, &din));
    UCHK(stage_f32(n, target, n->tgt, n0, &dt));
    UCHK(forward(n, din));
    int np;
    UCHK(backward(n, dt, &np));
    UCHK(read_loss(n, np, loss));
    if (grads) {
        HCHK(hipMemcpyAsync(grads, n->g, n->nparam * 4,
                            (n->ctx->flags & CAP_DEVICE_POINTERS) ? hipMemcpyDeviceToDevice
                                                                  : hipMemcpyDeviceToHost,
                            n->ctx->stream));
        HCHK(hipStreamSynchronize(n->ctx->stream));
    }
    return CAP_OK;
}

extern "C" int cap_unet_train_step(cap_unet *n, const float *in, const float *target, float lr,
                                   float *loss)
{
    if (!n || !in || !target || !loss || !(lr > 0.f)) return CAP_ERR_ARG;
    UCHK(cap_unet_gradients(n, in, target, loss, nullptr));
    n->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)n->step), c2 = 1.f - powf(b2, (float)n->step);
    hipLaunchKernelGGL(k_adam, dim3(grid_of((long)n->nparam)), dim3(256), 0, n->ctx->stream,
                       (long)n->nparam, n->p, (const float *)n->g, n->m, n->v, lr, b1, b2, 1e-8f, c1, c2);
    HCHK(hipGetLastError());
    n->packed = 0;
    HCHK(hipStreamSynchronize(n->ctx->stream));
    return CAP_OK;
}

// ---------------------------------------------------------- guess + solve
namespace {
// L1 residual of fi on the solve range (part: 2 doubles per block of the
// grid_of(cells) grid, at most 32768)
int res_l1_in(cap_ctx *ctx, double *part, const cap_coef *c, const double *fi, double *out)
{
    const long cells = (long)(ctx->ni - 4) * (ctx->nj - 4) * (ctx->nk - 4);
    const int nb = grid_of(cells);
    hipLaunchKernelGGL(k_res_l1, dim3(nb), dim3(256), 0, ctx->stream, ctx->ni, ctx->nj, ctx->nk,
                       (const double *)c->ae, (const double *)c->aw, (const double *)c->an,
                       (const double *)c->as, (const double *)c->at, (const double *)c->ab,
                       (const double *)c->ap, (const double *)c->su, fi, part);
    HCHK(hipGetLastError());
    std::vector<double> h(2 * nb);
    HCHK(hipMemcpyAsync(h.data(), part, 2 * nb * 8, hipMemcpyDeviceToHost, ctx->stream));
    HCHK(hipStreamSynchronize(ctx->stream));
    long double s = 0.0L;
    for (int i = 0; i < 2 * nb; ++i) s += h[i];
    *out = (double)s;
    return CAP_OK;
}

int res_l1(cap_ctx *ctx, cap_unet *n, const cap_coef *c, const double *fi, double *out)
{
    return res_l1_in(ctx, n->resp, c, fi, out);
}
}  // namespace

namespace {
// sums of the per-block compensated partials (nb blocks x nv values)
int sum_partials(cap_unet *n, int nb, int nv, double *out)
{
    std::vector<double> h((size_t)2 * nv * nb);
    HCHK(hipMemcpyAsync(h.data(), n->resp, h.size() * 8, hipMemcpyDeviceToHost, n->ctx->stream));
    HCHK(hipStreamSynchronize(n->ctx->stream));
    for (int m = 0; m < nv; ++m) {
        long double t = 0.0L;
        for (int b = 0; b < nb; ++b) t += (long double)h[(size_t)2 * nv * b + 2 * m] + h[(size_t)2 * nv * b + 2 * m + 1];
        out[m] = (double)t;
    }
    return CAP_OK;
}
}  // namespace

extern "C" int cap_unet_guess(cap_ctx *ctx, cap_unet *n, double *fi, const double *prev,
                              const cap_coef *c, double *scales)
{
    if (!ctx || !n || !fi || !c || n->ctx != ctx) return CAP_ERR_ARG;
    if (!(ctx->flags & CAP_DEVICE_POINTERS)) return CAP_ERR_ARG;   // device-resident fields
    if (ctx->nj - 4 != n->nx || ctx->ni - 4 != n->ny || ctx->nk - 4 != n->nz) return CAP_ERR_ARG;
    const double *A[8] = {(const double *)c->ae, (const double *)c->aw, (const double *)c->an,
                          (const double *)c->as, (const double *)c->at, (const double *)c->ab,
                          (const double *)c->ap, (const double *)c->su};
    for (int q = 0; q < 8; ++q) if (!A[q]) return CAP_ERR_ARG;
    if (!n->gpad) {
        UCHK(ualloc(n, (void **)&n->gpad, ctx->nstd * sizeof(double)));
        UCHK(ualloc(n, (void **)&n->gtmp, ctx->nstd * sizeof(double)));
        HCHK(hipMemsetAsync(n->gpad, 0, ctx->nstd * sizeof(double), ctx->stream));
        HCHK(hipMemsetAsync(n->gtmp, 0, ctx->nstd * sizeof(double), ctx->stream));
    }
    const long n0 = n->nvox[0];
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(k_ref_io, dim3(grid_of(n0)), dim3(256), 0, st, ctx->ni, ctx->nj, ctx->nk, 0,
                       A[0], A[1], A[2], A[3], A[4], A[5], A[6], A[7], fi, prev, n->in3,
                       (const float *)nullptr, (double *)nullptr, 0.0, 0.0);
    // normalisers: the residual's RMS s, the mean a of ap, the RMS ps of the
    // previous p' (1 when it is zero or absent)
    const int nbs = std::min(grid_of(n0, 256 * 4), 4096);   // 6 doubles each fit resp
    hipLaunchKernelGGL(k_in_stats, dim3(nbs), dim3(256), 0, st, n0, (const float *)n->in3, n->resp);
    HCHK(hipGetLastError());
    double m[3];
    UCHK(sum_partials(n, nbs, 3, m));
    const double s = m[0] > 0 ? sqrt(m[0] / n0) : 1.0, a = m[1] != 0 ? m[1] / n0 : 1.0,
                 ps = m[2] > 0 ? sqrt(m[2] / n0) : 1.0;
    UCHK(forward(n, n->in3, (float)(1.0 / s), (float)(1.0 / a), (float)(1.0 / ps)));
    hipLaunchKernelGGL(k_ref_io, dim3(grid_of(n0)), dim3(256), 0, st, ctx->ni, ctx->nj, ctx->nk, 1,
                       A[0], A[1], A[2], A[3], A[4], A[5], A[6], A[7], fi, prev, (float *)nullptr,
                       (const float *)n->out, n->gpad, 0.0, 0.0);
    // the network gives the direction g of the correction of x (the previous
    // p', or the start): scale it by its energy-norm step g.r / g.Mg, then
    // damp its high-frequency error -- which the residual weights most --
    // with GUESS_SWEEPS damped-Jacobi sweeps on M g = su - M x
    const double *x = prev ? prev : fi;
    const int nbe = std::min(grid_of(n0, 256 * 4), 4096);
    hipLaunchKernelGGL(k_edots, dim3(nbe), dim3(256), 0, st, ctx->ni, ctx->nj, ctx->nk, A[0], A[1],
                       A[2], A[3], A[4], A[5], A[6], A[7], x, (const double *)n->gpad, n->resp);
    HCHK(hipGetLastError());
    double e[2];
    UCHK(sum_partials(n, nbe, 2, e));
    double sc = e[1] > 0 ? e[0] / e[1] : 0.0;
    for (int sw = 0; sw < GUESS_SWEEPS; ++sw) {
        const double *gi = (sw & 1) ? n->gtmp : n->gpad;
        double *go = (sw & 1) ? n->gpad : n->gtmp;
        hipLaunchKernelGGL(k_jacobi, dim3(grid_of(n0)), dim3(256), 0, st, ctx->ni, ctx->nj, ctx->nk,
                           A[0], A[1], A[2], A[3], A[4], A[5], A[6], A[7], x, gi, go, sc, GUESS_OMEGA);
        sc = 1.0;
    }
    HCHK(hipGetLastError());
    // the guess is fi + beta (prev - fi) + alpha g with the least-squares
    // (beta, alpha) of the residual's 2-norm (alpha alone when prev is null),
    // so it never raises that norm above the start's or the previous p''s
    const int nba = std::min(grid_of(n0, 256 * 4), 3072);   // 10 doubles each fit resp
    hipLaunchKernelGGL(k_alpha, dim3(nba), dim3(256), 0, st, ctx->ni, ctx->nj, ctx->nk, A[0], A[1],
                       A[2], A[3], A[4], A[5], A[6], A[7], (const double *)fi, prev,
                       (const double *)n->gpad, n->resp);
    HCHK(hipGetLastError());
    double q[5];
    UCHK(sum_partials(n, nba, 5, q));
    double alpha = q[1] > 0 ? q[0] / q[1] : 0.0, beta = 0.0;
    if (prev && q[3] > 0) {
        const double det = q[1] * q[3] - q[4] * q[4];
        if (det > 1e-12 * q[1] * q[3]) {
            alpha = (q[0] * q[3] - q[2] * q[4]) / det;
            beta = (q[2] * q[1] - q[0] * q[4]) / det;
        } else {    // g and prev - fi are (nearly) parallel
            alpha = 0.0;
            beta = q[2] / q[3];
        }
    }
    hipLaunchKernelGGL(k_ref_io, dim3(grid_of(n0)), dim3(256), 0, st, ctx->ni, ctx->nj, ctx->nk, 2,
                       A[0], A[1], A[2], A[3], A[4], A[5], A[6], A[7], fi, prev, (float *)nullptr,
                       (const float *)nullptr, n->gpad, alpha, beta);
    HCHK(hipGetLastError());
    HCHK(hipStreamSynchronize(st));
    if (scales) { scales[0] = s; scales[1] = a; scales[2] = ps; scales[3] = alpha; scales[4] = beta; }
    return CAP_OK;
}

extern "C" int cap_pcgsip_guess(cap_ctx *ctx, cap_unet *n, double *fi, const double *prev,
                                const cap_coef *c, double alfa, double sor, int nsw, cap_report *rep,
                                int *accepted)
{
    if (!ctx || !n || !fi || !c || !rep || !accepted) return CAP_ERR_ARG;
    if (!(ctx->flags & CAP_DEVICE_POINTERS)) return CAP_ERR_ARG;
    const size_t bytes = ctx->nstd * sizeof(double);
    double *keep;
    HCHK(hipMalloc(&keep, bytes));
    int rc = CAP_OK;
    double r0 = 0.0, rg = 0.0;
    if (hipMemcpyAsync(keep, fi, bytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess) rc = CAP_ERR_HIP;
    if (!rc) rc = res_l1(ctx, n, c, fi, &r0);
    if (!rc) rc = cap_unet_guess(ctx, n, fi, prev, c, nullptr);
    if (!rc) rc = res_l1(ctx, n, c, fi, &rg);
    // the guess is used only if it lowers the L1 residual of the start
    *accepted = !rc && rg < r0;
    if (!rc && !*accepted &&
        hipMemcpyAsync(fi, keep, bytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
        rc = CAP_ERR_HIP;
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(keep);
    if (rc) return rc;
    // the unchanged solver stops on its L1 residual relative to its own
    // start; from the guess, the tolerance is rescaled so that it stops at
    // the absolute residual the solve from the caller's start stops at
    const double sor_eff = *accepted && rg > 0 ? sor * (r0 / rg) : sor;
    rc = cap_pcgsip(ctx, fi, c, alfa, sor_eff, nsw, rep);
    if (rc >= 0 && !*accepted && rep->status == CAP_OK) return CAP_GUESS_REJECTED;
    return rc;
}

// ------------------------------------------------- PISO with a guessed start
namespace {
// dst = src on the solve range k, i, j = 3..n?mm
__global__ void k_copy_range(int ni, int nj, int nk, const double *src, double *dst)
{
    const int nij = ni * nj;
    const long n = (long)(ni - 4) * (nj - 4) * (nk - 4);
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int j = 3 + (int)(q % (nj - 4)), i = 3 + (int)((q / (nj - 4)) % (ni - 4)),
                  k = 3 + (int)(q / ((long)(nj - 4) * (ni - 4)));
        const size_t c = (size_t)(k - 1) * nij + (size_t)(i - 1) * nj + (size_t)(j - 1);
        dst[c] = src[c];
    }
}

struct GuessStart {
    cap_unet *net;
    const double *prev;
    int accepted;
    double l1_ref, l1_start;
};

// the start of the last corrector's p solve: the network's guess (with net)
// or prev on the solve range (without), kept when it lowers the L1 residual
// of the reference's start x; sor then scales by l1_ref / l1_start, so the
// solve stops at the absolute residual the reference's start stops at
int guess_start(void *user, cap_ctx *ctx, const cap_coef *c, double *x, double *sor)
{
    GuessStart *g = (GuessStart *)user;
    double *keep = cap_slot_buffer(ctx, 46);
    if (!keep) return CAP_ERR_NOMEM;
    const size_t bytes = ctx->nstd * sizeof(double);
    double r0 = 0.0, rg = 0.0;
    UCHK(res_l1_in(ctx, ctx->partials, c, x, &r0));
    HCHK(hipMemcpyAsync(keep, x, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    // the warm start: prev on the solve range, x's ghost values (the solve
    // holds them fixed, and they enter the residual of the boundary cells)
    const long cells = (long)(ctx->ni - 4) * (ctx->nj - 4) * (ctx->nk - 4);
    if (g->net) {
        double *pw = cap_slot_buffer(ctx, 45);
        if (!pw) return CAP_ERR_NOMEM;
        HCHK(hipMemcpyAsync(pw, x, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        hipLaunchKernelGGL(k_copy_range, dim3(grid_of(cells)), dim3(256), 0, ctx->stream, ctx->ni,
                           ctx->nj, ctx->nk, g->prev, pw);
        HCHK(hipGetLastError());
        UCHK(cap_unet_guess(ctx, g->net, x, pw, c, nullptr));
    } else {
        hipLaunchKernelGGL(k_copy_range, dim3(grid_of(cells)), dim3(256), 0, ctx->stream, ctx->ni,
                           ctx->nj, ctx->nk, g->prev, x);
        HCHK(hipGetLastError());
    }
    UCHK(res_l1_in(ctx, ctx->partials, c, x, &rg));
    g->l1_ref = r0;
    g->accepted = rg < r0;
    g->l1_start = g->accepted ? rg : r0;
    if (g->accepted) *sor = *sor * (r0 / rg);
    else HCHK(hipMemcpyAsync(x, keep, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return CAP_OK;
}
}  // namespace

extern "C" int cap_piso_guess(cap_ctx *ctx, cap_flow *f, const cap_drv_opts *o, int pimple,
                              cap_unet *net, const double *prev, cap_report *reps, double *cont,
                              int *accepted, double *l1)
{
    if (!ctx || !prev || !accepted) return CAP_ERR_ARG;
    if (!(ctx->flags & CAP_DEVICE_POINTERS)) return CAP_ERR_ARG;
    if (net && (net->ctx != ctx || ctx->nj - 4 != net->nx || ctx->ni - 4 != net->ny ||
                ctx->nk - 4 != net->nz))
        return CAP_ERR_ARG;
    GuessStart g = {net, prev, 0, 0.0, 0.0};
    const int rc = cap_piso_run(ctx, f, o, pimple, reps, cont, guess_start, &g);
    *accepted = g.accepted;
    if (l1) { l1[0] = g.l1_ref; l1[1] = g.l1_start; }
    return rc;
}
