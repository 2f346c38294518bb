// fs_learned.hip -- learned preconditioner network on gfx950 (see include/fluidsim_learned.h).
//
// Activations are BF16, channels-last (16 channels = 32 bytes per voxel).  The
// 16 -> 16 convolutions are implicit GEMMs on v_mfma_f32_16x16x16bf16_1k:
// M = 16 voxels along x, N = 16 output channels, K = 16 input channels of one
// tap, 27 taps accumulated in FP32.  Operand lane layout (checked on gfx950):
// A: lane l holds row l%16, k = 4*(l/16)+e; B: column l%16, k = 4*(l/16)+e;
// D: column l%16, row 4*(l/16)+e.  A 16x4x4 voxel output tile per 256-thread
// block stages its 18x6x6 input halo in LDS; zero padding at the grid edge.
// The data gradient of a convolution is the same kernel with the weights
// flipped in space and transposed in channels; the weight gradient is another
// MFMA GEMM with K running over the voxels of a tile.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/fluidsim_learned.h"
#include "fs_device.hpp"
#include "fs_learned.hpp"

using namespace fsd;

namespace fsk {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int TX = 16, TY = 4, TZ = 4;                 // output tile
constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2;   // input halo
constexpr int HALO = HX * HY * HZ;

// round-to-nearest-even FP32 -> BF16
FS_HD uint16_t f2bf(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
FS_DEV float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

FS_DEV void tile_origin_of(const LGrid& g, int b, int& x0, int& y0, int& z0) {
    const int ntx = (g.W + TX - 1) / TX, nty = (g.H + TY - 1) / TY;
    x0 = (b % ntx) * TX;
    y0 = ((b / ntx) % nty) * TY;
    z0 = (b / (ntx * nty)) * TZ;
}
FS_DEV void tile_origin(const LGrid& g, int& x0, int& y0, int& z0) { tile_origin_of(g, blockIdx.x, x0, y0, z0); }
static unsigned ntiles(const LGrid& g) {
    return (unsigned)(((g.W + TX - 1) / TX) * ((g.H + TY - 1) / TY) * ((g.D + TZ - 1) / TZ));
}
static unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }

// LDS layout of the staged halo: 16 BF16 channels (32 bytes) per voxel.
// Offset in uint16 of channel group `grp` (0..3, 4 channels) of halo voxel hv
// (affine, so that the 108 reads of a tile are one base plus immediate
// offsets; permuting the groups per voxel against bank conflicts measured no
// faster)
FS_DEV int lds_grp(int hv, int grp) { return hv * 16 + grp * 4; }

// stage the 16-channel halo of a tile into LDS (uint4 = 8 channels per load);
// a 256-thread block issues all of a thread's loads before its first store
constexpr int STAGE16_U = (HALO * 2 + 255) / 256;
FS_DEV void halo16_load(const LGrid& g, const uint16_t* __restrict__ in, int x0, int y0, int z0,
                        uint4 (&val)[STAGE16_U]) {
#pragma unroll
    for (int u = 0; u < STAGE16_U; ++u) {
        const int q = threadIdx.x + 256 * u;
        const int hv = q >> 1, half = q & 1;
        const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
        const int x = x0 + hx - 1, y = y0 + hy - 1, z = z0 + hz - 1;
        val[u] = make_uint4(0, 0, 0, 0);
        if (q < HALO * 2 && x >= 0 && x < g.W && y >= 0 && y < g.H && z >= 0 && z < g.D) {
            const long long v = (long long)x + (long long)y * g.W + (long long)z * g.W * g.H;
            val[u] = reinterpret_cast<const uint4*>(in + v * 16)[half];
        }
    }
}
FS_DEV void halo16_store(uint16_t* sh, const uint4 (&val)[STAGE16_U]) {
#pragma unroll
    for (int u = 0; u < STAGE16_U; ++u) {
        const int q = threadIdx.x + 256 * u;
        if (q < HALO * 2) {
            const int hv = q >> 1, g0 = 2 * (q & 1);
            *reinterpret_cast<uint2*>(sh + lds_grp(hv, g0)) = make_uint2(val[u].x, val[u].y);
            *reinterpret_cast<uint2*>(sh + lds_grp(hv, g0 + 1)) = make_uint2(val[u].z, val[u].w);
        }
    }
}
FS_DEV void stage_halo16(const LGrid& g, const uint16_t* __restrict__ in, int x0, int y0, int z0, uint16_t* sh) {
    uint4 val[STAGE16_U];
    halo16_load(g, in, x0, y0, z0, val);
    halo16_store(sh, val);
}

// ---------------------------------------------------------------- MFMA conv
// out = act(conv(in, w) + bias).  OUT 0: BF16 16 channels (ReLU if relu).
// OUT 1: channel 0 only, as FP32: zn = acc + bias[0]; z = zn * scale * (type==1).
template <int OUT>
__global__ __launch_bounds__(256) void k_conv16(LGrid g, const uint16_t* __restrict__ in,
                                                const uint16_t* __restrict__ w, const float* __restrict__ bias,
                                                int relu, uint16_t* __restrict__ out, float* __restrict__ z,
                                                float* __restrict__ zn, const uint8_t* __restrict__ type,
                                                const float* __restrict__ scale, const int* __restrict__ done,
                                                int ntile) {
    if (done && *done) return;
    __shared__ __attribute__((aligned(16))) uint16_t sh[HALO * 16];
    // a block walks tiles blockIdx.x, + gridDim.x, ...: the weights are loaded
    // once, and the next tile's halo is loaded into registers while the
    // current one is computed
    int tile = blockIdx.x;
    if (tile >= ntile) return;
    int x0, y0, z0;
    tile_origin_of(g, tile, x0, y0, z0);
    uint4 hal[STAGE16_U];
    halo16_load(g, in, x0, y0, z0, hal);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int col = lane & 15, kg = lane >> 4;
    v4s wreg[LN_TAPS];
#pragma unroll
    for (int t = 0; t < LN_TAPS; ++t) wreg[t] = *reinterpret_cast<const v4s*>(w + t * 256 + col * 16 + kg * 4);
    const float bcol = bias ? bias[col] : 0.0f;
    for (;;) {
    __syncthreads();   // the previous tile's MFMAs have read sh
    halo16_store(sh, hal);
    __syncthreads();
    const int next = tile + (int)gridDim.x;
    int nx0 = 0, ny0 = 0, nz0 = 0;
    if (next < ntile) {
        tile_origin_of(g, next, nx0, ny0, nz0);
        halo16_load(g, in, nx0, ny0, nz0, hal);
    }
    v4f acc[TY];
#pragma unroll
    for (int r = 0; r < TY; ++r) acc[r] = v4f{0.f, 0.f, 0.f, 0.f};
    // opaque per tile: keeps the 108 LDS addresses below from being hoisted
    // out of the tile loop and held in registers
    int hv0 = col + 1 + HX * (1 + HY * (wv + 1));
    asm volatile("" : "+v"(hv0));
#pragma unroll
    for (int t = 0; t < LN_TAPS; ++t) {
        const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            const int hv = hv0 + dx + HX * (r + dy + HY * dz);
            v4s a = *reinterpret_cast<const v4s*>(sh + lds_grp(hv, kg));
            acc[r] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, wreg[t], acc[r], 0, 0, 0);
        }
    }
    const int zz = z0 + wv;
#pragma unroll
    for (int r = 0; r < TY; ++r) {
        const int yy = y0 + r;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int xx = x0 + kg * 4 + e;
            if (xx >= g.W || yy >= g.H || zz >= g.D) continue;
            const long long v = (long long)xx + (long long)yy * g.W + (long long)zz * g.W * g.H;
            float val = acc[r][e] + bcol;
            if (OUT == 0) {
                if (relu) val = fmaxf(val, 0.0f);
                out[v * 16 + col] = f2bf(val);
            } else if (col == 0) {
                if (zn) zn[v] = val;
                z[v] = (type[v] == 1) ? val * scale[0] : 0.0f;
            }
        }
    }
    if (next >= ntile) break;
    tile = next;
    x0 = nx0; y0 = ny0; z0 = nz0;
    }
}
// blocks of a k_conv16 launch: at most 3 per CU (its registers allow 3 waves per SIMD)
static unsigned conv_blocks(const LGrid& g) { return std::min(ntiles(g), 768u); }

// weight gradient dW[t][ci][co] += sum_v in(v+t)[ci] dout(v)[co]; dbias[co] += sum_v dout(v)[co]
__global__ __launch_bounds__(256) void k_wgrad16(LGrid g, const uint16_t* __restrict__ in,
                                                 const uint16_t* __restrict__ dout, float* __restrict__ dw,
                                                 float* __restrict__ db) {
    __shared__ __attribute__((aligned(16))) uint16_t sh[HALO * 16];
    __shared__ uint16_t sd[TX * TY * TZ * 16];
    int x0, y0, z0;
    tile_origin(g, x0, y0, z0);
    stage_halo16(g, in, x0, y0, z0, sh);
    for (int q = threadIdx.x; q < TX * TY * TZ * 2; q += blockDim.x) {
        int tv = q >> 1, half = q & 1;
        int x = x0 + tv % TX, y = y0 + (tv / TX) % TY, zq = z0 + tv / (TX * TY);
        uint4 val = make_uint4(0, 0, 0, 0);
        if (x < g.W && y < g.H && zq < g.D) {
            long long v = (long long)x + (long long)y * g.W + (long long)zq * g.W * g.H;
            val = reinterpret_cast<const uint4*>(dout + v * 16)[half];
        }
        reinterpret_cast<uint4*>(sd + tv * 16)[half] = val;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int col = lane & 15, kg = lane >> 4;
    for (int t = wv; t < LN_TAPS; t += 4) {
        const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
        v4f acc = {0.f, 0.f, 0.f, 0.f};
        for (int chunk = 0; chunk < TX * TY * TZ / 16; ++chunk) {
            v4s a, b;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int tv = chunk * 16 + kg * 4 + e;          // voxel of the tile (K index)
                const int tx = tv % TX, ty = (tv / TX) % TY, tz = tv / (TX * TY);
                const int hv = (tx + 1 + dx) + HX * ((ty + 1 + dy) + HY * (tz + 1 + dz));
                a[e] = (short)sh[lds_grp(hv, col >> 2) + (col & 3)];   // A[ci = col][k]
                b[e] = (short)sd[tv * 16 + col];                  // B[k][co = col]
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int ci = kg * 4 + e;                            // D row
            atomicAdd(&dw[t * 256 + ci * 16 + col], acc[e]);
        }
    }
    if (db && threadIdx.x < 16) {
        float s = 0.0f;
        for (int tv = 0; tv < TX * TY * TZ; ++tv) s += bf2f(sd[tv * 16 + threadIdx.x]);
        atomicAdd(&db[threadIdx.x], s);
    }
}

// ---------------------------------------------------------------- VALU layers
// layer A: a0 = ReLU(bA + conv([rn, fluid, air], wA)), wA [tap][ci 3][co 16] FP32
// (type: 1 unknown (FLUID), 2 Dirichlet (AIR), anything else Neumann/none)
// layer A on the FP32 MFMA (v_mfma_f32_16x16x4_f32): per row of 16 voxels an
// implicit GEMM with M = 16 voxels along x, N = 16 output channels, K = 27
// taps x 3 inputs (81, padded to 84 with zero weights) in 21 steps of 4.
// Operand layout: A lane l row l % 16, k = l / 16; B lane l column l % 16,
// k = l / 16; D lane l column l % 16, rows 4 (l / 16) + e.  The products are
// FP32 as in a VALU evaluation; the sum runs in the MFMA's order, and the
// bias is added last.
constexpr int A_KSTEPS = 21;
__global__ __launch_bounds__(256) void k_layerA(LGrid g, const float* __restrict__ rn,
                                                const uint8_t* __restrict__ type, const float* __restrict__ prm,
                                                uint16_t* __restrict__ out, const int* __restrict__ done, int ntile) {
    if (done && *done) return;
    __shared__ float sx[HALO * 3];
    // a block walks tiles blockIdx.x, + gridDim.x, ...: the weights are loaded
    // once, the next tile's inputs while the current one is computed
    int tile = blockIdx.x;
    if (tile >= ntile) return;
    constexpr int U = (HALO + 255) / 256;
    float hr[U];
    uint8_t ht[U];
    auto load = [&](int x0, int y0, int z0) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int hv = threadIdx.x + 256 * u;
            const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
            const int x = x0 + hx - 1, y = y0 + hy - 1, z = z0 + hz - 1;
            hr[u] = 0.0f;
            ht[u] = 0;
            if (hv < HALO && x >= 0 && x < g.W && y >= 0 && y < g.H && z >= 0 && z < g.D) {
                const long long v = (long long)x + (long long)y * g.W + (long long)z * g.W * g.H;
                hr[u] = rn[v];
                ht[u] = type[v];
            }
        }
    };
    int x0, y0, z0;
    tile_origin_of(g, tile, x0, y0, z0);
    load(x0, y0, z0);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int col = lane & 15, kg = lane >> 4;
    // this lane's (tap, input) of each K step: LDS offset relative to the
    // row's centre voxel, and the weight of output channel col
    int koff[A_KSTEPS];
    float wk[A_KSTEPS];
    const float* wA = prm + LN_AW;
#pragma unroll
    for (int j = 0; j < A_KSTEPS; ++j) {
        const int k = 4 * j + kg;
        const int t = k / 3, ci = k - 3 * (k / 3);
        const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
        koff[j] = k < 81 ? (dx + HX * (dy + HY * dz)) * 3 + ci : 0;
        wk[j] = k < 81 ? wA[t * 48 + ci * 16 + col] : 0.0f;
    }
    const float bcol = prm[LN_AB + col];
    for (;;) {
        __syncthreads();   // the previous tile's MFMAs have read sx
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int hv = threadIdx.x + 256 * u;
            if (hv >= HALO) break;
            sx[hv * 3] = hr[u];
            sx[hv * 3 + 1] = ht[u] == 1 ? 1.0f : 0.0f;
            sx[hv * 3 + 2] = ht[u] == 2 ? 1.0f : 0.0f;
        }
        __syncthreads();
        const int next = tile + (int)gridDim.x;
        int nx0 = 0, ny0 = 0, nz0 = 0;
        if (next < ntile) {
            tile_origin_of(g, next, nx0, ny0, nz0);
            load(nx0, ny0, nz0);
        }
        v4f acc[TY];
#pragma unroll
        for (int r = 0; r < TY; ++r) acc[r] = v4f{0.f, 0.f, 0.f, 0.f};
        // opaque per tile, so that the LDS addresses are not hoisted out of the loop
        int c0 = ((col + 1) + HX * (1 + HY * (wv + 1))) * 3;
        asm volatile("" : "+v"(c0));
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            const int c3 = c0 + HX * 3 * r;
#pragma unroll
            for (int j = 0; j < A_KSTEPS; ++j)
                acc[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(sx[c3 + koff[j]], wk[j], acc[r], 0, 0, 0);
        }
        const int zz = z0 + wv;
#pragma unroll
        for (int r = 0; r < TY; ++r) {
            const int yy = y0 + r;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int xx = x0 + kg * 4 + e;
                if (xx >= g.W || yy >= g.H || zz >= g.D) continue;
                const long long v = (long long)xx + (long long)yy * g.W + (long long)zz * g.W * g.H;
                out[v * 16 + col] = f2bf(fmaxf(acc[r][e] + bcol, 0.0f));
            }
        }
        if (next >= ntile) break;
        tile = next;
        x0 = nx0; y0 = ny0; z0 = nz0;
    }
}

// rn = r / scale with scale = rms of r over the unknowns, sqrt(sc->rr / sc->n_unknown)
// when sc is given, else scale_host; the scale is kept for layer E
__global__ __launch_bounds__(256) void k_normalise(long long n, const float* __restrict__ r,
                                                   const SolveScalars* __restrict__ sc, float scale_host,
                                                   const float* __restrict__ scale_dev,
                                                   float* __restrict__ rn, float* __restrict__ scale_out) {
    if (sc && sc->done) return;
    float s = scale_host;
    if (scale_dev) {
        s = scale_dev[0];
    } else if (sc) {
        double v = sc->rr / (double)(sc->n_unknown > 0 ? sc->n_unknown : 1);
        s = v > 0.0 ? (float)sqrt(v) : 1.0f;
    }
    if (!(s > 0.0f) || !isfinite(s)) s = 1.0f;
    long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c == 0) scale_out[0] = s;
    if (c >= n) return;
    rn[c] = r[c] / s;
}

// eight BF16 channels of a uint4 as floats, and back
FS_DEV void bf8_unpack(const uint4& u, float (&f)[8]) {
    const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        f[2 * e] = __uint_as_float(w[e] << 16);
        f[2 * e + 1] = __uint_as_float(w[e] & 0xffff0000u);
    }
}
FS_DEV uint4 bf8_pack(const float (&f)[8]) {
    uint32_t w[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) w[e] = (uint32_t)f2bf(f[2 * e]) | ((uint32_t)f2bf(f[2 * e + 1]) << 16);
    return make_uint4(w[0], w[1], w[2], w[3]);
}

// average pooling 2x (zero padding), 16 channels; a thread takes 8 channels
// of one coarse voxel (16-byte loads and stores), each summed in the order
// dk, dj, di
__global__ __launch_bounds__(256) void k_pool16(LGrid f, LGrid c, const uint16_t* __restrict__ in,
                                                uint16_t* __restrict__ out, const int* __restrict__ done) {
    if (done && *done) return;
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= c.n * 2) return;
    const long long V = q >> 1;
    const int half = (int)(q & 1);
    int I, J, K;
    split3(V, c.W, c.H, I, J, K);
    float s[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) s[e] = 0.0f;
    for (int dk = 0; dk < 2; ++dk)
        for (int dj = 0; dj < 2; ++dj)
            for (int di = 0; di < 2; ++di) {
                const int i = 2 * I + di, j = 2 * J + dj, k = 2 * K + dk;
                if (i < f.W && j < f.H && k < f.D) {
                    float v[8];
                    bf8_unpack(reinterpret_cast<const uint4*>(
                                   in + ((long long)i + (long long)j * f.W + (long long)k * f.W * f.H) * 16)[half], v);
#pragma unroll
                    for (int e = 0; e < 8; ++e) s[e] += v[e];
                }
            }
#pragma unroll
    for (int e = 0; e < 8; ++e) s[e] = 0.125f * s[e];
    reinterpret_cast<uint4*>(out + V * 16)[half] = bf8_pack(s);
}

// out = fine + up(coarse) (nearest), 16 channels; a thread takes 8 channels of
// one fine voxel
__global__ __launch_bounds__(256) void k_upadd16(LGrid f, LGrid c, const uint16_t* __restrict__ fine,
                                                 const uint16_t* __restrict__ coarse, uint16_t* __restrict__ out,
                                                 const int* __restrict__ done) {
    if (done && *done) return;
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= f.n * 2) return;
    const long long v = q >> 1;
    const int half = (int)(q & 1);
    int i, j, k;
    split3(v, f.W, f.H, i, j, k);
    const long long V = (long long)(i >> 1) + (long long)(j >> 1) * c.W + (long long)(k >> 1) * c.W * c.H;
    float a[8], b[8];
    bf8_unpack(reinterpret_cast<const uint4*>(fine + v * 16)[half], a);
    bf8_unpack(reinterpret_cast<const uint4*>(coarse + V * 16)[half], b);
#pragma unroll
    for (int e = 0; e < 8; ++e) a[e] = a[e] + b[e];
    reinterpret_cast<uint4*>(out + v * 16)[half] = bf8_pack(a);
}

// ---------------------------------------------------------------- backward
// A-norm loss of the unscaled output: d = zn - target on the unknowns,
// loss = 0.5 d.(A d) / n_fluid, dz = A d / n_fluid (A symmetric), with A the
// pressure matrix of the types: diag = #in-grid unknown/Dirichlet neighbours,
// -1 per unknown neighbour
__global__ __launch_bounds__(256) void k_loss_grad(LGrid g, const float* __restrict__ zn,
                                                   const float* __restrict__ target, const uint8_t* __restrict__ type,
                                                   double inv_nf, float* __restrict__ dz, double* __restrict__ loss) {
    __shared__ double sm[4];
    long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (c < g.n) {
        double ad = 0.0;
        if (type[c] == 1) {
            int x, y, z;
            split3(c, g.W, g.H, x, y, z);
            const double d = (double)zn[c] - (double)target[c];
            const long long WH = (long long)g.W * g.H;
            const int nx[6] = {x + 1, x - 1, x, x, x, x}, ny[6] = {y, y, y + 1, y - 1, y, y},
                      nz[6] = {z, z, z, z, z + 1, z - 1};
            const long long off[6] = {1, -1, g.W, -(long long)g.W, WH, -WH};
            for (int q = 0; q < 6; ++q) {
                if (nx[q] < 0 || nx[q] >= g.W || ny[q] < 0 || ny[q] >= g.H || nz[q] < 0 || nz[q] >= g.D) continue;
                const uint8_t t = type[c + off[q]];
                if (t == 1 || t == 2) ad += d;
                if (t == 1) ad -= (double)zn[c + off[q]] - (double)target[c + off[q]];
            }
            l = 0.5 * d * ad * inv_nf;
        }
        dz[c] = (float)(ad * inv_nf);
    }
    double t = block_sum(l, sm);
    if (threadIdx.x == 0 && t != 0.0) atomicAdd(loss, t);
}

// layer E: ds0[v][c] = sum_t wE[t][c] dz(v - t); wE FP32 [tap][16]
__global__ __launch_bounds__(256) void k_layerE_dx(LGrid g, const float* __restrict__ dz,
                                                   const float* __restrict__ prm, uint16_t* __restrict__ ds0) {
    __shared__ float sz[HALO];
    int x0, y0, z0;
    tile_origin(g, x0, y0, z0);
    for (int hv = threadIdx.x; hv < HALO; hv += blockDim.x) {
        int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
        int x = x0 + hx - 1, y = y0 + hy - 1, z = z0 + hz - 1;
        sz[hv] = (x >= 0 && x < g.W && y >= 0 && y < g.H && z >= 0 && z < g.D)
                     ? dz[(long long)x + (long long)y * g.W + (long long)z * g.W * g.H] : 0.0f;
    }
    __syncthreads();
    const int tx = threadIdx.x % TX, ty = (threadIdx.x / TX) % TY, tz = threadIdx.x / (TX * TY);
    const int x = x0 + tx, y = y0 + ty, z = z0 + tz;
    if (x >= g.W || y >= g.H || z >= g.D) return;
    float acc[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) acc[c] = 0.0f;
    for (int t = 0; t < LN_TAPS; ++t) {
        const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dzz = t / 9 - 1;
        const float d = sz[(tx + 1 - dx) + HX * ((ty + 1 - dy) + HY * (tz + 1 - dzz))];
#pragma unroll
        for (int c = 0; c < 16; ++c) acc[c] += prm[LN_EW + t * 16 + c] * d;
    }
    const long long v = (long long)x + (long long)y * g.W + (long long)z * g.W * g.H;
    uint16_t o[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) o[c] = f2bf(acc[c]);
    uint4* dst = reinterpret_cast<uint4*>(ds0 + v * 16);
    dst[0] = *reinterpret_cast<uint4*>(&o[0]);
    dst[1] = *reinterpret_cast<uint4*>(&o[8]);
}

// weight gradient of layer E (16 -> 1, FP32 dz) or A (2 -> 16, FP32 inputs):
// generic VALU reduction over a tile; MODE 0: E (in = s0 16ch, dout = dz);
// MODE 1: A (in = [rn, fluid, air], dout = ga0 16ch)
template <int MODE>
__global__ __launch_bounds__(256) void k_wgrad_valu(LGrid g, const uint16_t* __restrict__ in16,
                                                    const float* __restrict__ rn, const uint8_t* __restrict__ type,
                                                    const float* __restrict__ dzf, const uint16_t* __restrict__ d16,
                                                    float* __restrict__ dw, float* __restrict__ db) {
    __shared__ float sin_[HALO * 16];
    __shared__ float sout[TX * TY * TZ * 16];
    int x0, y0, z0;
    tile_origin(g, x0, y0, z0);
    constexpr int CI = MODE == 0 ? 16 : 3;
    constexpr int CO = MODE == 0 ? 1 : 16;
    for (int q = threadIdx.x; q < HALO * CI; q += blockDim.x) {
        int hv = q / CI, ci = q % CI;
        int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
        int x = x0 + hx - 1, y = y0 + hy - 1, z = z0 + hz - 1;
        float val = 0.0f;
        if (x >= 0 && x < g.W && y >= 0 && y < g.H && z >= 0 && z < g.D) {
            long long v = (long long)x + (long long)y * g.W + (long long)z * g.W * g.H;
            if (MODE == 0) val = bf2f(in16[v * 16 + ci]);
            else val = ci == 0 ? rn[v] : (type[v] == ci ? 1.0f : 0.0f);
        }
        sin_[q] = val;
    }
    for (int q = threadIdx.x; q < TX * TY * TZ * CO; q += blockDim.x) {
        int tv = q / CO, co = q % CO;
        int x = x0 + tv % TX, y = y0 + (tv / TX) % TY, z = z0 + tv / (TX * TY);
        float val = 0.0f;
        if (x < g.W && y < g.H && z < g.D) {
            long long v = (long long)x + (long long)y * g.W + (long long)z * g.W * g.H;
            val = MODE == 0 ? dzf[v] : bf2f(d16[v * 16 + co]);
        }
        sout[q] = val;
    }
    __syncthreads();
    const int np = LN_TAPS * CI * CO;
    for (int pidx = threadIdx.x; pidx < np; pidx += blockDim.x) {
        int t = pidx / (CI * CO), ci = (pidx / CO) % CI, co = pidx % CO;
        const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
        float s = 0.0f;
        for (int tv = 0; tv < TX * TY * TZ; ++tv) {
            int tx = tv % TX, ty = (tv / TX) % TY, tz = tv / (TX * TY);
            int hv = (tx + 1 + dx) + HX * ((ty + 1 + dy) + HY * (tz + 1 + dz));
            s += sin_[hv * CI + ci] * sout[tv * CO + co];
        }
        atomicAdd(&dw[pidx], s);
    }
    if (threadIdx.x < CO) {
        float s = 0.0f;
        for (int tv = 0; tv < TX * TY * TZ; ++tv) s += sout[tv * CO + threadIdx.x];
        atomicAdd(&db[threadIdx.x], s);
    }
}

// coarse gradient of a nearest upsampling: sum of the 8 children; then the
// ReLU mask of the coarse activation: out = (act > 0) ? sum : 0
__global__ __launch_bounds__(256) void k_upT_relu16(LGrid f, LGrid c, const uint16_t* __restrict__ gf,
                                                    const uint16_t* __restrict__ act, uint16_t* __restrict__ out) {
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= c.n * 16) return;
    long long V = q >> 4;
    int ch = (int)(q & 15);
    int I, J, K;
    split3(V, c.W, c.H, I, J, K);
    float s = 0.0f;
    for (int dk = 0; dk < 2; ++dk)
        for (int dj = 0; dj < 2; ++dj)
            for (int di = 0; di < 2; ++di) {
                int i = 2 * I + di, j = 2 * J + dj, k = 2 * K + dk;
                if (i < f.W && j < f.H && k < f.D)
                    s += bf2f(gf[((long long)i + (long long)j * f.W + (long long)k * f.W * f.H) * 16 + ch]);
            }
    out[q] = f2bf(bf2f(act[q]) > 0.0f ? s : 0.0f);
}

// fine gradient through a skip and an average pooling: out = (act > 0) ? gs + up(gc)/8 : 0
// (gs and out may be the same buffer: each thread reads its element before writing it)
__global__ __launch_bounds__(256) void k_poolT_add_relu16(LGrid f, LGrid c, const uint16_t* gs,
                                                          const uint16_t* __restrict__ gc,
                                                          const uint16_t* __restrict__ act, uint16_t* out) {
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= f.n * 16) return;
    long long v = q >> 4;
    int ch = (int)(q & 15);
    int i, j, k;
    split3(v, f.W, f.H, i, j, k);
    long long V = (long long)(i >> 1) + (long long)(j >> 1) * c.W + (long long)(k >> 1) * c.W * c.H;
    float s = bf2f(gs[q]) + 0.125f * bf2f(gc[V * 16 + ch]);
    out[q] = f2bf(bf2f(act[q]) > 0.0f ? s : 0.0f);
}

// ---------------------------------------------------------------- parameters
// forward layout [tap][cout][cin] BF16 from [tap][cin][cout]; the data-gradient
// copy is flipped in space and transposed in channels
__global__ __launch_bounds__(256) void k_pack16(const float* __restrict__ w, uint16_t* __restrict__ fwd,
                                                uint16_t* __restrict__ bwd) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= LN_TAPS * 256) return;
    int t = q / 256, ci = (q / 16) % 16, co = q % 16;
    uint16_t h = f2bf(w[q]);
    fwd[t * 256 + co * 16 + ci] = h;
    if (bwd) bwd[(LN_TAPS - 1 - t) * 256 + ci * 16 + co] = h;
}

// layer E (16 -> 1) as a 16 -> 16 MFMA conv with only output channel 0 non-zero
__global__ __launch_bounds__(256) void k_packE(const float* __restrict__ w, uint16_t* __restrict__ fwd) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= LN_TAPS * 256) return;
    int t = q / 256, co = (q / 16) % 16, ci = q % 16;
    fwd[q] = co == 0 ? f2bf(w[t * 16 + ci]) : (uint16_t)0;
}

__global__ __launch_bounds__(256) void k_adam(int n, float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, float lr, float c1,
                                              float c2) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float gq = g[q];
    if (!isfinite(gq)) gq = 0.0f;
    float mq = b1 * m[q] + (1.0f - b1) * gq;
    float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    p[q] -= lr * (mq * c1) / (sqrtf(vq * c2) + eps);
}

// ---------------------------------------------------------------- host side
LNet::~LNet() {
    free_acts();
    for (void* p : {(void*)d_params, (void*)d_grad, (void*)d_m, (void*)d_v, (void*)d_loss, (void*)wB, (void*)wC,
                    (void*)wD, (void*)wE, (void*)wBt, (void*)wCt, (void*)wDt, (void*)d_scale})
        if (p) hipFree(p);
}

void LNet::free_acts() {
    for (void* p : {(void*)rn, (void*)dz, (void*)a0, (void*)s0, (void*)ga0, (void*)a1, (void*)b1, (void*)s1,
                    (void*)d1, (void*)gd1, (void*)gs1, (void*)gb1, (void*)ga1, (void*)b2, (void*)c2, (void*)gc2,
                    (void*)gb2})
        if (p) hipFree(p);
    rn = dz = nullptr;
    a0 = s0 = ga0 = a1 = b1 = s1 = d1 = gd1 = gs1 = gb1 = ga1 = b2 = c2 = gc2 = gb2 = nullptr;
    cap0 = cap1 = cap2 = 0;
}

static uint64_t sm64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

int LNet::init_params(uint64_t seed, hipStream_t st) {
    std::vector<float> h(LN_NPARAM, 0.0f);
    auto fill = [&](int off, int n, int fan_in) {
        const double sd = std::sqrt(2.0 / fan_in);
        for (int i = 0; i < n; ++i) {
            // Box-Muller from two counter-based uniforms
            double u1 = ((sm64(seed * 1315423911ULL + (uint64_t)(off + i) * 2) >> 11) + 0.5) / 9007199254740992.0;
            double u2 = ((sm64(seed * 1315423911ULL + (uint64_t)(off + i) * 2 + 1) >> 11) + 0.5) / 9007199254740992.0;
            h[off + i] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
        }
    };
    fill(LN_AW, 27 * 3 * 16, 27 * 3);
    fill(LN_BW, 27 * 256, 27 * 16);
    fill(LN_CW, 27 * 256, 27 * 16);
    fill(LN_DW, 27 * 256, 27 * 16);
    fill(LN_EW, 27 * 16, 27 * 16);
    for (int i = 0; i < 27 * 16; ++i) h[LN_EW + i] *= 0.5f;
    return set_params(h.data(), st);
}

int LNet::set_params(const float* host, hipStream_t st) {
    if (!d_params) {
        bool ok = true;
        auto A = [&](void** p, size_t b) { if (hipMalloc(p, b) != hipSuccess) ok = false; };
        A((void**)&d_params, sizeof(float) * LN_NPARAM);
        A((void**)&d_grad, sizeof(float) * LN_NPARAM);
        A((void**)&d_m, sizeof(float) * LN_NPARAM);
        A((void**)&d_v, sizeof(float) * LN_NPARAM);
        A((void**)&d_loss, sizeof(double) * 2);
        A((void**)&d_scale, sizeof(float) * 2);
        for (uint16_t** p : {&wB, &wC, &wD, &wE, &wBt, &wCt, &wDt}) A((void**)p, sizeof(uint16_t) * LN_TAPS * 256);
        if (!ok) return -1;
    }
    if (hipMemcpyAsync(d_params, host, sizeof(float) * LN_NPARAM, hipMemcpyHostToDevice, st) != hipSuccess) return -1;
    hipMemsetAsync(d_m, 0, sizeof(float) * LN_NPARAM, st);
    hipMemsetAsync(d_v, 0, sizeof(float) * LN_NPARAM, st);
    adam_t = 0;
    refresh_weights(st);
    return hipStreamSynchronize(st) == hipSuccess ? 0 : -1;
}

int LNet::get_params(float* host, hipStream_t st) {
    if (!d_params) return -1;
    hipMemcpyAsync(host, d_params, sizeof(float) * LN_NPARAM, hipMemcpyDeviceToHost, st);
    return hipStreamSynchronize(st) == hipSuccess ? 0 : -1;
}

void LNet::refresh_weights(hipStream_t st) {
    const unsigned nb = (LN_TAPS * 256 + 255) / 256;
    hipLaunchKernelGGL(k_pack16, dim3(nb), dim3(256), 0, st, d_params + LN_BW, wB, wBt);
    hipLaunchKernelGGL(k_pack16, dim3(nb), dim3(256), 0, st, d_params + LN_CW, wC, wCt);
    hipLaunchKernelGGL(k_pack16, dim3(nb), dim3(256), 0, st, d_params + LN_DW, wD, wDt);
    hipLaunchKernelGGL(k_packE, dim3(nb), dim3(256), 0, st, d_params + LN_EW, wE);
}

int LNet::ensure(int W, int H, int D) {
    g0 = LGrid{W, H, D, (long long)W * H * D};
    g1 = LGrid{(W + 1) / 2, (H + 1) / 2, (D + 1) / 2, 0};
    g1.n = (long long)g1.W * g1.H * g1.D;
    g2 = LGrid{(g1.W + 1) / 2, (g1.H + 1) / 2, (g1.D + 1) / 2, 0};
    g2.n = (long long)g2.W * g2.H * g2.D;
    if (g0.n <= cap0 && g1.n <= cap1 && g2.n <= cap2) return 0;
    free_acts();
    bool ok = true;
    auto A = [&](void** p, size_t b) { if (hipMalloc(p, b) != hipSuccess) ok = false; };
    A((void**)&rn, sizeof(float) * g0.n);
    A((void**)&dz, sizeof(float) * g0.n);
    for (uint16_t** p : {&a0, &s0, &ga0}) A((void**)p, sizeof(uint16_t) * 16 * g0.n);
    for (uint16_t** p : {&a1, &b1, &s1, &d1, &gd1, &gs1, &gb1, &ga1}) A((void**)p, sizeof(uint16_t) * 16 * g1.n);
    for (uint16_t** p : {&b2, &c2, &gc2, &gb2}) A((void**)p, sizeof(uint16_t) * 16 * g2.n);
    if (!ok) { free_acts(); return -1; }
    cap0 = g0.n; cap1 = g1.n; cap2 = g2.n;
    ++epoch;
    return 0;
}

void LNet::forward(const uint8_t* type, const float* r, float* z, const SolveScalars* sc, float scale_host, float* zn,
                   hipStream_t st, const float* scale_dev) {
    const int* done = sc ? &sc->done : nullptr;
    const float* nf = nullptr;
    const uint8_t* nt = nullptr;
    float* nz = nullptr;
    hipLaunchKernelGGL(k_normalise, dim3(nblk(g0.n)), dim3(256), 0, st, g0.n, r, sc, scale_host, scale_dev, rn,
                       d_scale);
    hipLaunchKernelGGL(k_layerA, dim3(conv_blocks(g0)), dim3(256), 0, st, g0, rn, type, d_params, a0, done,
                       (int)ntiles(g0));
    hipLaunchKernelGGL(k_pool16, dim3(nblk(g1.n * 2)), dim3(256), 0, st, g0, g1, a0, a1, done);
    hipLaunchKernelGGL(k_conv16<0>, dim3(conv_blocks(g1)), dim3(256), 0, st, g1, a1, wB, d_params + LN_BB, 1, b1, nz,
                       nz, nt, nf, done, (int)ntiles(g1));
    hipLaunchKernelGGL(k_pool16, dim3(nblk(g2.n * 2)), dim3(256), 0, st, g1, g2, b1, b2, done);
    hipLaunchKernelGGL(k_conv16<0>, dim3(conv_blocks(g2)), dim3(256), 0, st, g2, b2, wC, d_params + LN_CB, 1, c2, nz,
                       nz, nt, nf, done, (int)ntiles(g2));
    hipLaunchKernelGGL(k_upadd16, dim3(nblk(g1.n * 2)), dim3(256), 0, st, g1, g2, b1, c2, s1, done);
    hipLaunchKernelGGL(k_conv16<0>, dim3(conv_blocks(g1)), dim3(256), 0, st, g1, s1, wD, d_params + LN_DB, 1, d1, nz,
                       nz, nt, nf, done, (int)ntiles(g1));
    hipLaunchKernelGGL(k_upadd16, dim3(nblk(g0.n * 2)), dim3(256), 0, st, g0, g1, a0, d1, s0, done);
    hipLaunchKernelGGL(k_conv16<1>, dim3(conv_blocks(g0)), dim3(256), 0, st, g0, s0, wE, d_params + LN_EB, 0,
                       (uint16_t*)nullptr, z, zn, type, (const float*)d_scale, done, (int)ntiles(g0));
}

void LNet::backward(const uint8_t* type, const float* target, const float* zn, int n_fluid, hipStream_t st) {
    hipMemsetAsync(d_grad, 0, sizeof(float) * LN_NPARAM, st);
    hipMemsetAsync(d_loss, 0, sizeof(double), st);
    const double inv_nf = n_fluid > 0 ? 1.0 / n_fluid : 0.0;
    hipLaunchKernelGGL(k_loss_grad, dim3(nblk(g0.n)), dim3(256), 0, st, g0, zn, target, type, inv_nf, dz, d_loss);
    // E
    hipLaunchKernelGGL((k_wgrad_valu<0>), dim3(ntiles(g0)), dim3(256), 0, st, g0, s0, (const float*)nullptr,
                       (const uint8_t*)nullptr, dz, (const uint16_t*)nullptr, d_grad + LN_EW, d_grad + LN_EB);
    hipLaunchKernelGGL(k_layerE_dx, dim3(ntiles(g0)), dim3(256), 0, st, g0, dz, d_params, ga0);   // ga0 = ds0
    // D: gd1 = relu'(d1) * upT(ds0)
    hipLaunchKernelGGL(k_upT_relu16, dim3(nblk(g1.n * 16)), dim3(256), 0, st, g0, g1, ga0, d1, gd1);
    hipLaunchKernelGGL(k_wgrad16, dim3(ntiles(g1)), dim3(256), 0, st, g1, s1, gd1, d_grad + LN_DW, d_grad + LN_DB);
    hipLaunchKernelGGL(k_conv16<0>, dim3(conv_blocks(g1)), dim3(256), 0, st, g1, gd1, wDt, (const float*)nullptr, 0,
                       gs1, (float*)nullptr, (float*)nullptr, (const uint8_t*)nullptr, (const float*)nullptr,
                       (const int*)nullptr, (int)ntiles(g1));
    // C: gc2 = relu'(c2) * upT(gs1)
    hipLaunchKernelGGL(k_upT_relu16, dim3(nblk(g2.n * 16)), dim3(256), 0, st, g1, g2, gs1, c2, gc2);
    hipLaunchKernelGGL(k_wgrad16, dim3(ntiles(g2)), dim3(256), 0, st, g2, b2, gc2, d_grad + LN_CW, d_grad + LN_CB);
    hipLaunchKernelGGL(k_conv16<0>, dim3(conv_blocks(g2)), dim3(256), 0, st, g2, gc2, wCt, (const float*)nullptr, 0,
                       gb2, (float*)nullptr, (float*)nullptr, (const uint8_t*)nullptr, (const float*)nullptr,
                       (const int*)nullptr, (int)ntiles(g2));
    // B: gb1 = relu'(b1) * (gs1 + poolT(gb2))
    hipLaunchKernelGGL(k_poolT_add_relu16, dim3(nblk(g1.n * 16)), dim3(256), 0, st, g1, g2, gs1, gb2, b1, gb1);
    hipLaunchKernelGGL(k_wgrad16, dim3(ntiles(g1)), dim3(256), 0, st, g1, a1, gb1, d_grad + LN_BW, d_grad + LN_BB);
    hipLaunchKernelGGL(k_conv16<0>, dim3(conv_blocks(g1)), dim3(256), 0, st, g1, gb1, wBt, (const float*)nullptr, 0,
                       ga1, (float*)nullptr, (float*)nullptr, (const uint8_t*)nullptr, (const float*)nullptr,
                       (const int*)nullptr, (int)ntiles(g1));
    // A: ga0 = relu'(a0) * (ds0 + poolT(ga1)) (in place over ds0)
    hipLaunchKernelGGL(k_poolT_add_relu16, dim3(nblk(g0.n * 16)), dim3(256), 0, st, g0, g1, ga0, ga1, a0, ga0);
    hipLaunchKernelGGL((k_wgrad_valu<1>), dim3(ntiles(g0)), dim3(256), 0, st, g0, (const uint16_t*)nullptr, rn, type,
                       (const float*)nullptr, ga0, d_grad + LN_AW, d_grad + LN_AB);
}

void LNet::adam(float lr, hipStream_t st) {
    ++adam_t;
    const float c1 = 1.0f / (1.0f - std::pow(0.9f, (float)adam_t));
    const float c2 = 1.0f / (1.0f - std::pow(0.999f, (float)adam_t));
    hipLaunchKernelGGL(k_adam, dim3((LN_NPARAM + 255) / 256), dim3(256), 0, st, LN_NPARAM, d_params, d_grad, d_m, d_v,
                       lr, c1, c2);
    refresh_weights(st);
}

int LearnedPre::prepare_box() {
    MGLevel L0 = s_->fine();
    return net_->ensure(L0.W, L0.H, L0.D);
}

void LearnedPre::apply(const float* r, float* z, double* partial, SolveScalars* sc, hipStream_t st) {
    MGLevel L0 = s_->fine();
    net_->forward(L0.type, r, z, sc, 1.0f, nullptr, st);
    s_->dot_rz_zq(r, z, partial, sc, st);
}

HybridPre::~HybridPre() {
    for (void* p : {(void*)r1_, (void*)z2_, (void*)scale_, (void*)part_})
        if (p) hipFree(p);
}

int HybridPre::prepare_box() {
    MGLevel L0 = s_->fine();
    if (net_->ensure(L0.W, L0.H, L0.D) != 0) return -1;
    if (L0.n <= cap_ && part_) return 0;
    for (void* p : {(void*)r1_, (void*)z2_, (void*)scale_, (void*)part_})
        if (p) hipFree(p);
    r1_ = z2_ = scale_ = nullptr;
    part_ = nullptr;
    bool ok = hipMalloc((void**)&r1_, sizeof(float) * L0.n) == hipSuccess &&
              hipMalloc((void**)&z2_, sizeof(float) * L0.n) == hipSuccess &&
              hipMalloc((void**)&scale_, sizeof(float) * 2) == hipSuccess &&
              hipMalloc((void**)&part_, sizeof(double) * 4 * 2048) == hipSuccess;
    if (!ok) return -1;
    cap_ = L0.n;
    ++epoch_;
    return 0;
}

void HybridPre::apply(const float* r, float* z, double* partial, SolveScalars* sc, hipStream_t st) {
    MGLevel L0 = s_->fine();
    double* p_scratch = part_;
    double* p_rr = part_ + 2048;
    double* p_num = part_ + 2 * 2048;
    double* p_den = part_ + 3 * 2048;
    s_->vcycle(r, z, p_scratch, sc, st);                  // z1 = V(r)
    s_->resid_sq(r, z, r1_, p_rr, sc, st);                // r1 = r - A z1
    s_->partial_rms(p_rr, scale_, sc, st);
    net_->forward(L0.type, r1_, z2_, sc, 1.0f, nullptr, st, scale_);   // z2 = N(r1)
    s_->line_dots(z2_, r1_, p_num, p_den, sc, st);
    s_->add_scaled(z, z2_, p_num, p_den, sc, st);         // z = z1 + c z2
    s_->dot_rz_zq(r, z, partial, sc, st);
}

}  // namespace fsk
