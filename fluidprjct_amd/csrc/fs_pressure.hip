// fs_pressure.hip -- matrix-free MGPCG with FP64 dots and FP64 iterative
// refinement for the reference's pressure system (see fs_pressure.hpp).
//
// Multigrid follows McAdams, Sifakis & Teran 2010 ("A parallel multigrid
// Poisson solver for fluids simulation on large grids"): cell-centred levels
// coarsened 2x per axis, a coarse cell is Dirichlet if any child is Dirichlet,
// else an unknown if any child is an unknown, else inactive; the 7-point
// operator is re-discretised on every level with unit coefficients; transfers
// are cell-centred trilinear prolongation P and restriction R = P^T / 2 (the
// factor 4 of the unit-coefficient coarse operator times the 1/8 average);
// a degree-4 Chebyshev polynomial smoother in D^-1 A (four Jacobi-form sweeps
// with the steps of fs_pressure.hpp) before and after the coarse correction and
// a fixed number of damped-Jacobi sweeps on the coarsest level keep the V-cycle
// a symmetric positive definite linear operator, as CG requires.
//
// Kernel count per iteration is what limits small levels on MI355X (~10 us of
// launch-to-launch latency each), so the V-cycle fuses two pre-smoothing sweeps
// from zero everywhere and, on levels of at most FUSE_MAX cells, prolongation +
// first post-smoothing sweep (it recomputes prolongations at the stencil points
// and is slower than separate passes on large levels); the CG scalars are derived
// inside the consumer kernels from <= 2048 FP64 partial sums, summed in the same
// fixed order by every block (deterministic, no extra launch).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "fs_kernels.hpp"
#include "fs_pressure.hpp"

using namespace fsd;

namespace fsk {

constexpr int RED_BLOCKS = 2048;      // partial-sum slots of every reducing kernel
constexpr int COARSE_MAX = 4096;      // cells of the LDS-resident coarsest level
constexpr long long FUSE_MAX = 1LL << 16;  // levels up to this size use the fused (recomputing) transfer kernels
constexpr long long TB_MAX = 1LL << 18;
constexpr long long RESTRICT_V2_MIN = 1LL << 18;   // coarse levels above this restrict two cells per thread    // coarse levels up to this size use the temporally blocked smoothers
constexpr uint32_t SELF = 0x8000u;
constexpr uint32_t WBIT = 0x4000u;   // level 0: the cell has a partly open face (weighted stencil)

// partial-sum slots (each RED_BLOCKS doubles) in MGPCG::partial
enum { PS_RZ0 = 0, PS_RZ1, PS_PQ0, PS_PQ1, PS_RR, PS_B2, PS_B1, PS_RES, PS_ZQ, PS_RP, PS_MEAN, PS_COUNT };

enum { FIN_BNORM = 0, FIN_OUTER, FIN_OUTER_REL, FIN_CHECK };

// 1/d for the smoothers (d = 1..6): the hardware reciprocal (1 ulp) instead of
// a correctly rounded division.  The preconditioner only needs one fixed
// diagonal scaling, the same in every sweep, which keeps it symmetric.
FS_DEV float inv_diag(uint32_t d) { return __builtin_amdgcn_rcpf((float)d); }
// the diagonal of cell c as a float: the weighted sum for a WBIT cell of level 0
FS_DEV float diag_L(const float* __restrict__ wd, long long c, uint32_t code) {
    return (code & WBIT) ? wd[c] : (float)((code >> 8) & 7u);
}

FS_DEV float stencil_Ax(const float* __restrict__ x, long long c, uint32_t code, int W, long long WH) {
    float s = (float)((code >> 8) & 7u) * x[c];
    if (code & 1u) s -= x[c + 1];
    if (code & 2u) s -= x[c - 1];
    if (code & 4u) s -= x[c + W];
    if (code & 8u) s -= x[c - W];
    if (code & 16u) s -= x[c + WH];
    if (code & 32u) s -= x[c - WH];
    return s;
}

FS_DEV double stencil_Ax64(const double* __restrict__ x, long long c, uint32_t code, int W, long long WH) {
    double s = (double)((code >> 8) & 7u) * x[c];
    if (code & 1u) s -= x[c + 1];
    if (code & 2u) s -= x[c - 1];
    if (code & 4u) s -= x[c + W];
    if (code & 8u) s -= x[c - W];
    if (code & 16u) s -= x[c + WH];
    if (code & 32u) s -= x[c - WH];
    return s;
}

// A x at an unknown c with the face weights: d_c x_c minus the open fraction
// times the neighbour's value on each coupled face (nb(q): the neighbour's
// value across face q = +x, -x, +y, -y, +z, -z, as a double).  With every
// fraction 1 it is stencil_Ax64's sum, term for term.
template <class Nb>
FS_DEV double stencil_Ax64_w(const MGLevel& L, long long c, uint32_t code, long long WH, double xc, Nb nb) {
    double s = (double)L.wd[c] * xc;
    if (code & 1u) s -= (double)L.wx[c] * nb(0);
    if (code & 2u) s -= (double)L.wx[c - 1] * nb(1);
    if (code & 4u) s -= (double)L.wy[c] * nb(2);
    if (code & 8u) s -= (double)L.wy[c - L.W] * nb(3);
    if (code & 16u) s -= (double)L.wz[c] * nb(4);
    if (code & 32u) s -= (double)L.wz[c - WH] * nb(5);
    return s;
}

// FP32 A x and 1/diag at cell c of level L: the unit stencil, or the weighted
// one for a cell with WBIT (level 0 with face fractions)
FS_DEV float stencil_Ax_L(const MGLevel& L, const float* __restrict__ x, long long c, uint32_t code, long long WH) {
    if (!(code & WBIT)) return stencil_Ax(x, c, code, L.W, WH);
    float s = L.wd[c] * x[c];
    if (code & 1u) s -= L.wx[c] * x[c + 1];
    if (code & 2u) s -= L.wx[c - 1] * x[c - 1];
    if (code & 4u) s -= L.wy[c] * x[c + L.W];
    if (code & 8u) s -= L.wy[c - L.W] * x[c - L.W];
    if (code & 16u) s -= L.wz[c] * x[c + WH];
    if (code & 32u) s -= L.wz[c - WH] * x[c - WH];
    return s;
}

// Sum of n (<= RED_BLOCKS) partials in a fixed order, result in every thread.
// A thread's slots (threadIdx.x + u * blockDim.x) are all loaded before they
// are added, in the order of u.
constexpr int PART_UNROLL = RED_BLOCKS / 256;
FS_DEV double sum_partials(const double* __restrict__ p, int n, double* sm) {
    double v[PART_UNROLL];
#pragma unroll
    for (int u = 0; u < PART_UNROLL; ++u) {
        const int i = threadIdx.x + u * blockDim.x;
        v[u] = i < n ? p[i] : 0.0;
    }
    double a = 0.0;
#pragma unroll
    for (int u = 0; u < PART_UNROLL; ++u)
        if ((int)threadIdx.x + u * (int)blockDim.x < n) a += v[u];
    for (int i = threadIdx.x + PART_UNROLL * blockDim.x; i < n; i += blockDim.x) a += p[i];
    a = wave_sum(a);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sm[w] = a;
    __syncthreads();
    double t = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += sm[q];
    __syncthreads();   // sm is reused by the caller's next reduction
    return t;
}

// Sums of two or three partial arrays, each in sum_partials' order (so the
// results are the same bits), with the loads of all of them in flight at
// once.  sm holds at least 3 * 4 doubles.
template <int K>
FS_DEV void sum_partials_k(const double* const (&p)[K], int n, double* sm, double (&out)[K]) {
    double a[K];
#pragma unroll
    for (int q = 0; q < K; ++q) a[q] = 0.0;
    {
        double v[K][PART_UNROLL];
#pragma unroll
        for (int u = 0; u < PART_UNROLL; ++u) {
            const int i = threadIdx.x + u * blockDim.x;
#pragma unroll
            for (int q = 0; q < K; ++q) v[q][u] = i < n ? p[q][i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < PART_UNROLL; ++u) {
            if ((int)threadIdx.x + u * (int)blockDim.x >= n) continue;
#pragma unroll
            for (int q = 0; q < K; ++q) a[q] += v[q][u];
        }
    }
    for (int i = threadIdx.x + PART_UNROLL * blockDim.x; i < n; i += blockDim.x) {
#pragma unroll
        for (int q = 0; q < K; ++q) a[q] += p[q][i];
    }
    const int w = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
#pragma unroll
    for (int q = 0; q < K; ++q) a[q] = wave_sum(a[q]);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int q = 0; q < K; ++q) sm[q * 4 + w] = a[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < K; ++q) {
        double t = 0.0;
        for (int r = 0; r < nw; ++r) t += sm[q * 4 + r];
        out[q] = t;
    }
    __syncthreads();
}

FS_DEV void store_partial(double v, double* sm, double* __restrict__ partial) {
    double t = block_sum(v, sm);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// ---------------------------------------------------------------- level setup
// Fluid bounding box: per block [min i, max i, min j, max j, min k, max k]
// into part[6 * block] (no global atomics: the host folds the blocks).  A
// thread scans chunks of 16 consecutive cells (one 16-byte load when rows are
// a multiple of 16 cells), splitting only a chunk's first index into (i, j, k)
// and stepping along the rows from there.
constexpr int BBOX_BLOCKS = 256;
static_assert(MGPCG::HOST_COUNTERS == MGPCG::COUNTERS_BBOX + 6 * BBOX_BLOCKS, "host counter mirror size");
template <bool VEC16>
__global__ __launch_bounds__(256) void k_bbox(const uint8_t* __restrict__ labels, int W, int H, long long n,
                                              int* __restrict__ part) {
    __shared__ int sb[6];
    if (threadIdx.x < 6) sb[threadIdx.x] = (threadIdx.x & 1) ? -1 : 0x7fffffff;
    __syncthreads();
    int mn[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, mx[3] = {-1, -1, -1};
    for (long long c0 = 16 * ((long long)blockIdx.x * blockDim.x + threadIdx.x); c0 < n;
         c0 += 16LL * gridDim.x * blockDim.x) {
        uint8_t v[16];
        if (VEC16) {
            const uint4 q = *reinterpret_cast<const uint4*>(labels + c0);
            const uint32_t w4[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int t = 0; t < 16; ++t) v[t] = (uint8_t)(w4[t >> 2] >> (8 * (t & 3)));
        } else {
#pragma unroll
            for (int t = 0; t < 16; ++t) v[t] = c0 + t < n ? labels[c0 + t] : (uint8_t)FS_SOLID;
        }
        int i, j, k;
        split3(c0, W, H, i, j, k);
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            if (v[t] == FS_FLUID) {
                mn[0] = min(mn[0], i); mx[0] = max(mx[0], i);
                mn[1] = min(mn[1], j); mx[1] = max(mx[1], j);
                mn[2] = min(mn[2], k); mx[2] = max(mx[2], k);
            }
            if (++i == W) {
                i = 0;
                if (++j == H) { j = 0; ++k; }
            }
        }
    }
    for (int q = 0; q < 3; ++q) {
        for (int off = 32; off > 0; off >>= 1) {
            mn[q] = min(mn[q], __shfl_xor(mn[q], off, 64));
            mx[q] = max(mx[q], __shfl_xor(mx[q], off, 64));
        }
    }
    if ((threadIdx.x & 63) == 0) {
        for (int q = 0; q < 3; ++q) { atomicMin(&sb[2 * q], mn[q]); atomicMax(&sb[2 * q + 1], mx[q]); }
    }
    __syncthreads();
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = sb[threadIdx.x];
}

// level-0 types of the solve box (offset ox, oy, oz in the full W x H grid)
// and the right-hand side gathered into the box
__global__ __launch_bounds__(256) void k_type0(const uint8_t* __restrict__ labels, const float* __restrict__ bfull,
                                               MGLevel L, int W, int H, int ox, int oy, int oz,
                                               float* __restrict__ bbox) {
    long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (c >= L.n) return;
    int i, j, k;
    split3(c, L.W, L.H, i, j, k);
    i = i + ox;
    j = j + oy;
    k = k + oz;
    const size_t g = (size_t)i + (size_t)j * W + (size_t)k * W * H;
    const uint8_t lab = labels[g];
    L.type[c] = lab == FS_FLUID ? 1 : (lab == FS_AIR ? 2 : 0);
    bbox[c] = bfull[g];
}

// full-grid pressure (FP64 and FP32, 0 off fluid) from the solve box
// p = the box solution on FLUID cells, 0 elsewhere, FP64 and FP32, over the
// full grid; 4 consecutive cells per thread (vector loads and stores when
// the group is whole), one index split per group.
__global__ __launch_bounds__(256) void k_scatter_out(const uint8_t* __restrict__ labels, long long n, int W, int H,
                                                     MGLevel L, int ox, int oy, int oz, const double* __restrict__ xb,
                                                     double* __restrict__ p64, float* __restrict__ p) {
    const long long c0 = 4 * ((long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x);
    if (c0 >= n) return;
    int i, j, k;
    split3(c0, W, H, i, j, k);
    const bool whole = c0 + 3 < n;
    uint8_t lab[4];
    if (whole) {
        const uint32_t lw = *reinterpret_cast<const uint32_t*>(labels + c0);
#pragma unroll
        for (int q = 0; q < 4; ++q) lab[q] = (uint8_t)(lw >> (8 * q));
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) lab[q] = c0 + q < n ? labels[c0 + q] : (uint8_t)FS_SOLID;
    }
    double v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int bi = i - ox, bj = j - oy, bk = k - oz;
        v[q] = 0.0;
        if (lab[q] == FS_FLUID && bi >= 0 && bi < L.W && bj >= 0 && bj < L.H && bk >= 0 && bk < L.D)
            v[q] = xb[(size_t)bi + (size_t)bj * L.W + (size_t)bk * L.W * L.H];
        if (++i == W) {
            i = 0;
            if (++j == H) { j = 0; ++k; }
        }
    }
    if (whole) {
        *reinterpret_cast<double2*>(p64 + c0) = make_double2(v[0], v[1]);
        *reinterpret_cast<double2*>(p64 + c0 + 2) = make_double2(v[2], v[3]);
        *reinterpret_cast<float4*>(p + c0) = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
    } else {
        for (int q = 0; q < 4 && c0 + q < n; ++q) {
            p64[c0 + q] = v[q];
            p[c0 + q] = (float)v[q];
        }
    }
}

// type of coarse cell c: Dirichlet if a child is, else unknown if a child is
FS_DEV void type_coarsen_at(const MGLevel& F, const MGLevel& Cc, long long c) {
    int I, J, K;
    split3(c, Cc.W, Cc.H, I, J, K);
    int any1 = 0, any2 = 0;
    for (int dk = 0; dk < 2; ++dk) {
        int k = 2 * K + dk;
        if (k >= F.D) continue;
        for (int dj = 0; dj < 2; ++dj) {
            int j = 2 * J + dj;
            if (j >= F.H) continue;
            for (int di = 0; di < 2; ++di) {
                int i = 2 * I + di;
                if (i >= F.W) continue;
                uint8_t t = F.type[(size_t)i + (size_t)j * F.W + (size_t)k * F.W * F.H];
                any1 |= (t == 1);
                any2 |= (t == 2);
            }
        }
    }
    Cc.type[c] = any2 ? 2 : (any1 ? 1 : 0);
}

__global__ __launch_bounds__(256) void k_type_coarsen(MGLevel F, MGLevel Cc) {
    long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (c >= Cc.n) return;
    type_coarsen_at(F, Cc, c);
}

// per-cell stencil code; counters[0] += unknowns, counters[1] |= Dirichlet contact
// (grid-strided, one pair of global atomics per block).  COARSEN: the types of
// the next level Cc in the same launch (both read only this level's types).
template <bool COARSEN>
__global__ __launch_bounds__(256) void k_code(MGLevel L, MGLevel Cc, int* __restrict__ counters) {
    __shared__ int s_cnt[2];
    if (COARSEN)
        for (long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; c < Cc.n;
             c += (long long)gridDim.x * blockDim.x)
            type_coarsen_at(L, Cc, c);
    if (threadIdx.x == 0) { s_cnt[0] = 0; s_cnt[1] = 0; }
    __syncthreads();
    int unk = 0, dir = 0;
    const long long WH = (long long)L.W * L.H;
    const long long stride = (long long)gridDim.x * blockDim.x;
    // four cells per thread and pass, their type and neighbour-type loads
    // issued together (the grid is capped, so each thread makes many passes)
    for (long long c0 = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; c0 < L.n;
         c0 += 4 * stride) {
        uint8_t t[4], nt[4][6];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const long long c = c0 + u * stride;
            t[u] = 0;
#pragma unroll
            for (int q = 0; q < 6; ++q) nt[u][q] = 0;
            if (c >= L.n) continue;
            t[u] = L.type[c];
            int i, j, k;
            split3(c, L.W, L.H, i, j, k);
            if (i + 1 < L.W) nt[u][0] = L.type[c + 1];
            if (i > 0) nt[u][1] = L.type[c - 1];
            if (j + 1 < L.H) nt[u][2] = L.type[c + L.W];
            if (j > 0) nt[u][3] = L.type[c - L.W];
            if (k + 1 < L.D) nt[u][4] = L.type[c + WH];
            if (k > 0) nt[u][5] = L.type[c - WH];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const long long c = c0 + u * stride;
            if (c >= L.n) continue;
            uint32_t code = 0;
            if (t[u] == 1) {
                uint32_t diag = 0;
                int air = 0;
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    if (nt[u][q] == 1) { code |= (1u << q); ++diag; }
                    else if (nt[u][q] == 2) { ++diag; air = 1; }
                }
                // a FLUID cell with no FLUID or AIR neighbour has an empty row:
                // it is no unknown (p = 0, its b is dropped)
                if (diag) {
                    code |= SELF | (diag << 8);
                    ++unk;
                    dir |= air;
                } else {
                    code = 0;
                }
            }
            L.code[c] = (uint16_t)code;
        }
    }
    unk = wave_sum(unk);
    dir = wave_sum(dir);
    if ((threadIdx.x & 63) == 0) {
        if (unk) atomicAdd(&s_cnt[0], unk);
        if (dir) atomicOr(&s_cnt[1], 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_cnt[0]) atomicAdd(&counters[0], s_cnt[0]);
        if (s_cnt[1]) atomicOr(&counters[1], 1);
    }
}

// ---------------------------------------------------------------- smoothing
// one damped-Jacobi sweep; DOT: grid-strided, partial sums of b.x_new
template <bool DOT>
__global__ __launch_bounds__(256) void k_jacobi(MGLevel L, const float* __restrict__ b, const float* __restrict__ xin,
                                                float* __restrict__ xout, float omega, double* __restrict__ partial,
                                                const SolveScalars* __restrict__ sc) {
    if (sc && sc->done) return;
    __shared__ double sm[12];
    const long long WH = (long long)L.W * L.H;
    double acc = 0.0;
    for (long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; c < L.n;
         c += (long long)gridDim.x * blockDim.x) {
        uint32_t code = L.code[c];
        float xn = 0.0f;
        uint32_t d = (code >> 8) & 7u;
        if ((code & SELF) && d) {
            float bc = b[c];
            const float id = (code & WBIT) ? __builtin_amdgcn_rcpf(L.wd[c]) : inv_diag(d);
            xn = xin[c] + omega * (bc - stencil_Ax_L(L, xin, c, code, WH)) * id;
            if (DOT) acc += (double)bc * (double)xn;
        }
        xout[c] = xn;
    }
    if (DOT) store_partial(acc, sm, partial);
}

// diagonal preconditioner z = r / d with partial sums of r.z
__global__ __launch_bounds__(256) void k_diag(MGLevel L, const float* __restrict__ r, float* __restrict__ z,
                                              double* __restrict__ partial, const SolveScalars* __restrict__ sc) {
    if (sc->done) return;
    __shared__ double sm[12];
    double acc = 0.0;
    for (long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; c < L.n;
         c += (long long)gridDim.x * blockDim.x) {
        uint32_t code = L.code[c];
        uint32_t d = (code >> 8) & 7u;
        const float id = (code & WBIT) ? __builtin_amdgcn_rcpf(L.wd[c]) : inv_diag(d);
        float zn = ((code & SELF) && d) ? r[c] * id : 0.0f;
        z[c] = zn;
        acc += (double)r[c] * (double)zn;
    }
    store_partial(acc, sm, partial);
}

// Two damped-Jacobi sweeps from x = 0 in one pass: the first sweep's value at
// a neighbour, omega b_n / d_n, is recomputed from b and the level code.
// omega2 = 0 gives a single sweep.
FS_DEV float jacobi_first(const float* __restrict__ b, const uint16_t* __restrict__ code, long long c, float omega) {
    uint32_t cd = code[c];
    uint32_t d = (cd >> 8) & 7u;
    return ((cd & SELF) && d) ? omega * b[c] * inv_diag(d) : 0.0f;
}
// the same with the weighted diagonal of a WBIT cell (level 0 with face fractions)
FS_DEV float jacobi_first_L(const MGLevel& L, const float* __restrict__ b, long long c, float omega) {
    uint32_t cd = L.code[c];
    if (!(cd & WBIT)) return jacobi_first(b, L.code, c, omega);
    return (cd & SELF) ? omega * b[c] * __builtin_amdgcn_rcpf(L.wd[c]) : 0.0f;
}

__global__ __launch_bounds__(256) void k_jacobi2(MGLevel L, const float* __restrict__ b, float* __restrict__ xout,
                                                 float omega, float omega2, const SolveScalars* __restrict__ sc) {
    if (sc && sc->done) return;
    long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (c >= L.n) return;
    uint32_t code = L.code[c];
    float xn = 0.0f;
    uint32_t d = (code >> 8) & 7u;
    if ((code & SELF) && d && !(code & WBIT)) {
        const long long WH = (long long)L.W * L.H;
        float bc = b[c];
        float x1 = omega * bc * inv_diag(d);
        float ax = (float)d * x1;
        if (L.wd) {   // neighbours may be WBIT cells
            if (code & 1u) ax -= jacobi_first_L(L, b, c + 1, omega);
            if (code & 2u) ax -= jacobi_first_L(L, b, c - 1, omega);
            if (code & 4u) ax -= jacobi_first_L(L, b, c + L.W, omega);
            if (code & 8u) ax -= jacobi_first_L(L, b, c - L.W, omega);
            if (code & 16u) ax -= jacobi_first_L(L, b, c + WH, omega);
            if (code & 32u) ax -= jacobi_first_L(L, b, c - WH, omega);
        } else {
            if (code & 1u) ax -= jacobi_first(b, L.code, c + 1, omega);
            if (code & 2u) ax -= jacobi_first(b, L.code, c - 1, omega);
            if (code & 4u) ax -= jacobi_first(b, L.code, c + L.W, omega);
            if (code & 8u) ax -= jacobi_first(b, L.code, c - L.W, omega);
            if (code & 16u) ax -= jacobi_first(b, L.code, c + WH, omega);
            if (code & 32u) ax -= jacobi_first(b, L.code, c - WH, omega);
        }
        xn = x1 + omega2 * (bc - ax) * inv_diag(d);
    } else if ((code & SELF) && d) {
        // a WBIT cell: the weighted diagonal and couplings
        const long long WH = (long long)L.W * L.H;
        float bc = b[c];
        const float id = __builtin_amdgcn_rcpf(L.wd[c]);
        float x1 = omega * bc * id;
        float ax = L.wd[c] * x1;
        if (code & 1u) ax -= L.wx[c] * jacobi_first_L(L, b, c + 1, omega);
        if (code & 2u) ax -= L.wx[c - 1] * jacobi_first_L(L, b, c - 1, omega);
        if (code & 4u) ax -= L.wy[c] * jacobi_first_L(L, b, c + L.W, omega);
        if (code & 8u) ax -= L.wy[c - L.W] * jacobi_first_L(L, b, c - L.W, omega);
        if (code & 16u) ax -= L.wz[c] * jacobi_first_L(L, b, c + WH, omega);
        if (code & 32u) ax -= L.wz[c - WH] * jacobi_first_L(L, b, c - WH, omega);
        xn = x1 + omega2 * (bc - ax) * id;
    }
    xout[c] = xn;
}

__global__ __launch_bounds__(256) void k_residual(MGLevel L, const float* __restrict__ b, const float* __restrict__ x,
                                                  float* __restrict__ r, const SolveScalars* __restrict__ sc) {
    if (sc && sc->done) return;
    long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (c >= L.n) return;
    uint32_t code = L.code[c];
    float bc = b ? b[c] : 0.0f;
    r[c] = (code & SELF) ? bc - stencil_Ax_L(L, x, c, code, (long long)L.W * L.H) : 0.0f;
}

// coarse b = 1/2 * P^T r with trilinear weights 1/4,3/4,3/4,1/4 per axis.
// Summation order (shared by every kernel that restricts, so that they agree
// bit for bit): one partial sum per fine z-slice dk = 0..3, each over dj then
// di, and the four partials added in dk order.
FS_DEV float restrict_slice(int FW, int FH, int FD, const float* __restrict__ rf, int I, int J, int K, int dk) {
    const float wt[4] = {0.25f, 0.75f, 0.75f, 0.25f};
    const int k = 2 * K - 1 + dk;
    float s = 0.0f;
    if (k < 0 || k >= FD) return s;
    const long long WH = (long long)FW * FH;
    for (int dj = 0; dj < 4; ++dj) {
        int j = 2 * J - 1 + dj;
        if (j < 0 || j >= FH) continue;
        float wjk = wt[dj] * wt[dk];
        const long long row = (long long)j * FW + (long long)k * WH;
        for (int di = 0; di < 4; ++di) {
            int i = 2 * I - 1 + di;
            if (i < 0 || i >= FW) continue;
            s += wt[di] * wjk * rf[row + i];
        }
    }
    return s;
}

FS_DEV float restrict_sum(float p0, float p1, float p2, float p3) { return 0.5f * ((((0.0f + p0) + p1) + p2) + p3); }

// Four lanes per coarse cell (one fine z-slice each): the 64 fine reads of a
// coarse cell are spread over 4 lanes instead of serialised in one.
__global__ __launch_bounds__(256) void k_restrict(MGLevel F, MGLevel Cc, const float* __restrict__ rf,
                                                  float* __restrict__ bc, const SolveScalars* __restrict__ sc) {
    if (sc && sc->done) return;
    const long long t = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const long long c = t >> 2;
    const int dk = (int)(t & 3);
    const bool in = c < Cc.n;
    const bool unk = in && (Cc.code[c] & SELF);
    float p = 0.0f;
    if (unk) {
        int I, J, K;
        split3(c, Cc.W, Cc.H, I, J, K);
        p = restrict_slice(F.W, F.H, F.D, rf, I, J, K, dk);
    }
    // the four lanes of a cell are consecutive lanes of one wave
    const float p1 = __shfl_down(p, 1, 64), p2 = __shfl_down(p, 2, 64), p3 = __shfl_down(p, 3, 64);
    if (in && dk == 0) bc[c] = unk ? restrict_sum(p, p1, p2, p3) : 0.0f;
}

// (P xc) at fine cell (i,j,k)
FS_DEV float prolong_at(const MGLevel& Cc, const float* __restrict__ xc, int i, int j, int k) {
    int I0 = i >> 1, J0 = j >> 1, K0 = k >> 1;
    int Is[2] = {I0, (i & 1) ? I0 + 1 : I0 - 1}, Js[2] = {J0, (j & 1) ? J0 + 1 : J0 - 1},
        Ks[2] = {K0, (k & 1) ? K0 + 1 : K0 - 1};
    const float w[2] = {0.75f, 0.25f};
    float s = 0.0f;
    for (int a = 0; a < 2; ++a) {
        if (Ks[a] < 0 || Ks[a] >= Cc.D) continue;
        for (int bb = 0; bb < 2; ++bb) {
            if (Js[bb] < 0 || Js[bb] >= Cc.H) continue;
            for (int cc = 0; cc < 2; ++cc) {
                if (Is[cc] < 0 || Is[cc] >= Cc.W) continue;
                s += w[cc] * w[bb] * w[a] * xc[(size_t)Is[cc] + (size_t)Js[bb] * Cc.W + (size_t)Ks[a] * Cc.W * Cc.H];
            }
        }
    }
    return s;
}

// fine x += P xc on unknowns
__global__ __launch_bounds__(256) void k_prolong(MGLevel F, MGLevel Cc, const float* __restrict__ xc,
                                                 float* __restrict__ xf, const SolveScalars* __restrict__ sc) {
    if (sc && sc->done) return;
    long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (c >= F.n) return;
    if (!(F.code[c] & SELF)) return;
    int i, j, k;
    split3(c, F.W, F.H, i, j, k);
    xf[c] += prolong_at(Cc, xc, i, j, k);
}

// x' = x + P xc (fine unknowns), then one damped-Jacobi sweep on x' -> xout.
// DOT: grid-strided, partial sums of b.xout.
template <bool DOT>
__global__ __launch_bounds__(256) void k_prolong_smooth(MGLevel F, MGLevel Cc, const float* __restrict__ b,
                                                        const float* __restrict__ x, const float* __restrict__ xc,
                                                        float* __restrict__ xout, float omega,
                                                        double* __restrict__ partial,
                                                        const SolveScalars* __restrict__ sc) {
    if (sc && sc->done) return;
    __shared__ double sm[12];
    const long long WH = (long long)F.W * F.H;
    double acc = 0.0;
    for (long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; c < F.n;
         c += (long long)gridDim.x * blockDim.x) {
        uint32_t code = F.code[c];
        float xn = 0.0f;
        uint32_t d = (code >> 8) & 7u;
        if ((code & SELF) && d) {
            int i, j, k;
            split3(c, F.W, F.H, i, j, k);
            float xp = x[c] + prolong_at(Cc, xc, i, j, k);
            float bc = b[c];
            if (code & WBIT) {   // level 0 with face fractions: the weighted stencil
                float ax = F.wd[c] * xp;
                if (code & 1u) ax -= F.wx[c] * (x[c + 1] + prolong_at(Cc, xc, i + 1, j, k));
                if (code & 2u) ax -= F.wx[c - 1] * (x[c - 1] + prolong_at(Cc, xc, i - 1, j, k));
                if (code & 4u) ax -= F.wy[c] * (x[c + F.W] + prolong_at(Cc, xc, i, j + 1, k));
                if (code & 8u) ax -= F.wy[c - F.W] * (x[c - F.W] + prolong_at(Cc, xc, i, j - 1, k));
                if (code & 16u) ax -= F.wz[c] * (x[c + WH] + prolong_at(Cc, xc, i, j, k + 1));
                if (code & 32u) ax -= F.wz[c - WH] * (x[c - WH] + prolong_at(Cc, xc, i, j, k - 1));
                xn = xp + omega * (bc - ax) * __builtin_amdgcn_rcpf(F.wd[c]);
            } else {
                float ax = (float)d * xp;
                if (code & 1u) ax -= x[c + 1] + prolong_at(Cc, xc, i + 1, j, k);
                if (code & 2u) ax -= x[c - 1] + prolong_at(Cc, xc, i - 1, j, k);
                if (code & 4u) ax -= x[c + F.W] + prolong_at(Cc, xc, i, j + 1, k);
                if (code & 8u) ax -= x[c - F.W] + prolong_at(Cc, xc, i, j - 1, k);
                if (code & 16u) ax -= x[c + WH] + prolong_at(Cc, xc, i, j, k + 1);
                if (code & 32u) ax -= x[c - WH] + prolong_at(Cc, xc, i, j, k - 1);
                xn = xp + omega * (bc - ax) * inv_diag(d);
            }
            if (DOT) acc += (double)bc * (double)xn;
        }
        xout[c] = xn;
    }
    if (DOT) store_partial(acc, sm, partial);
}

// ---------------------------------------------------------------- coarse levels, temporally blocked
// On the small levels every kernel is bound by its launch latency, so all the
// pre-smoothing sweeps and the residual run as one kernel, and the
// prolongation and all the post-smoothing sweeps as another: a block owns an
// 8x8x4-cell tile and recomputes what its tile needs from a halo of NS cells
// in LDS (the sweep before the last on the tile plus one cell, and so on).
// Every cell value is computed with the operations of k_jacobi2, k_jacobi,
// k_residual and k_prolong_smooth in the same order, so the results are the
// same bits as those kernels'.
constexpr int TB_TX = 8, TB_TY = 8, TB_TZ = 4;
struct SweepSteps {
    float w[SMOOTH_STEPS];
};

// PRE:  x_1 = w_0 b / d pointwise on the tile + NS, then NS - 1 sweeps (the
//       valid region shrinking by one cell each), x -> xout and the residual
//       b - A x -> rout on the tile.
// !PRE: x + P xc on the tile + NS (xin is the level's iterate, xc the coarse
//       correction), NS sweeps, x -> xout on the tile.
constexpr int TB_THREADS = 1024;
template <int NS, bool PRE>
__global__ __launch_bounds__(TB_THREADS) void k_smooth_tb(MGLevel L, const float* __restrict__ b,
                                                   const float* __restrict__ xin, float* __restrict__ xout,
                                                   float* __restrict__ rout, MGLevel Cc, const float* __restrict__ xc,
                                                   SweepSteps st, const SolveScalars* __restrict__ sc) {
    if (sc && sc->done) return;
    constexpr int G = NS, RX = TB_TX + 2 * G, RY = TB_TY + 2 * G, RZ = TB_TZ + 2 * G, RN = RX * RY * RZ;
    constexpr int RXY = RX * RY;
    // coarse values the region's prolongation reads: coarse cells
    // (o >> 1) - 1 .. ((o + R - 1) >> 1) + 1 per axis
    constexpr int CX = RX / 2 + 3, CY = RY / 2 + 3, CZ = RZ / 2 + 3, CN = CX * CY * CZ;
    __shared__ float sx0[RN], sx1[RN], sbv[RN];
    __shared__ float sxc[PRE ? 1 : CN];
    __shared__ uint16_t scd[RN];
    const int W = L.W, H = L.H, D = L.D;
    const long long WH = (long long)W * H;
    const int ntx = (W + TB_TX - 1) / TB_TX, nty = (H + TB_TY - 1) / TB_TY;
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int ox = (bid % ntx) * TB_TX - G, oy = ((bid / ntx) % nty) * TB_TY - G, oz = (bid / (ntx * nty)) * TB_TZ - G;
    // (ox, oy, oz are even: tile sizes and G are even)
    const int cx0 = (ox >> 1) - 1, cy0 = (oy >> 1) - 1, cz0 = (oz >> 1) - 1;
    if (!PRE) {
        for (int l = threadIdx.x; l < CN; l += blockDim.x) {
            const int I = cx0 + l % CX, J = cy0 + (l / CX) % CY, K = cz0 + l / (CX * CY);
            const bool in = I >= 0 && I < Cc.W && J >= 0 && J < Cc.H && K >= 0 && K < Cc.D;
            sxc[l] = in ? xc[(size_t)I + (size_t)J * Cc.W + (size_t)K * Cc.W * Cc.H] : 0.0f;
        }
        __syncthreads();
    }
    for (int l = threadIdx.x; l < RN; l += blockDim.x) {
        const int i = ox + l % RX, j = oy + (l / RX) % RY, k = oz + l / RXY;
        const bool in = i >= 0 && i < W && j >= 0 && j < H && k >= 0 && k < D;
        const long long g = in ? (long long)i + (long long)j * W + (long long)k * WH : 0;
        const uint32_t cd = in ? L.code[g] : 0u;
        scd[l] = (uint16_t)cd;
        sbv[l] = in ? b[g] : 0.0f;
        if (!PRE) {
            float xv = 0.0f;
            if (cd & SELF) {
                // prolong_at with the coarse values from LDS (same terms, same order)
                const int I0 = i >> 1, J0 = j >> 1, K0 = k >> 1;
                const int Is[2] = {I0, (i & 1) ? I0 + 1 : I0 - 1}, Js[2] = {J0, (j & 1) ? J0 + 1 : J0 - 1},
                          Ks[2] = {K0, (k & 1) ? K0 + 1 : K0 - 1};
                const float w[2] = {0.75f, 0.25f};
                float sp = 0.0f;
                for (int a = 0; a < 2; ++a) {
                    if (Ks[a] < 0 || Ks[a] >= Cc.D) continue;
                    for (int bb = 0; bb < 2; ++bb) {
                        if (Js[bb] < 0 || Js[bb] >= Cc.H) continue;
                        for (int cc = 0; cc < 2; ++cc) {
                            if (Is[cc] < 0 || Is[cc] >= Cc.W) continue;
                            sp += w[cc] * w[bb] * w[a] *
                                  sxc[(Is[cc] - cx0) + CX * ((Js[bb] - cy0) + CY * (Ks[a] - cz0))];
                        }
                    }
                }
                xv = xin[g] + sp;
            }
            sx0[l] = xv;
        }
    }
    __syncthreads();
    float* cur = sx0;
    float* nxt = sx1;
    int m = G;   // margin (cells around the tile) on which cur is valid
    int k0 = 0;
    if (PRE) {
        for (int l = threadIdx.x; l < RN; l += blockDim.x) {
            const uint32_t cd = scd[l], d = (cd >> 8) & 7u;
            cur[l] = ((cd & SELF) && d) ? st.w[0] * sbv[l] * inv_diag(d) : 0.0f;
        }
        __syncthreads();
        k0 = 1;
    }
    for (int k = k0; k < NS; ++k) {
        --m;
        const int ex = TB_TX + 2 * m, ey = TB_TY + 2 * m, ez = TB_TZ + 2 * m, en = ex * ey * ez;
        const float w = st.w[k];
        for (int q = threadIdx.x; q < en; q += blockDim.x) {
            const int l = (G - m + q % ex) + RX * (G - m + (q / ex) % ey) + RXY * (G - m + q / (ex * ey));
            const uint32_t cd = scd[l], d = (cd >> 8) & 7u;
            float xn = 0.0f;
            if ((cd & SELF) && d) xn = cur[l] + w * (sbv[l] - stencil_Ax(cur, l, cd, RX, RXY)) * inv_diag(d);
            nxt[l] = xn;
        }
        __syncthreads();
        float* t = cur; cur = nxt; nxt = t;
    }
    // tile cells inside the level: x (and the residual)
    for (int q = threadIdx.x; q < TB_TX * TB_TY * TB_TZ; q += blockDim.x) {
        const int li = q % TB_TX, lj = (q / TB_TX) % TB_TY, lk = q / (TB_TX * TB_TY);
        const int i = ox + G + li, j = oy + G + lj, k = oz + G + lk;
        if (i >= W || j >= H || k >= D) continue;
        const int l = (G + li) + RX * (G + lj) + RXY * (G + lk);
        const long long g = (long long)i + (long long)j * W + (long long)k * WH;
        xout[g] = cur[l];
        if (PRE) {
            const uint32_t cd = scd[l];
            rout[g] = (cd & SELF) ? sbv[l] - stencil_Ax(cur, l, cd, RX, RXY) : 0.0f;
        }
    }
}

// coarsest level: fixed number of damped-Jacobi sweeps from zero in LDS
__global__ __launch_bounds__(1024) void k_coarse_solve(MGLevel L, const float* __restrict__ b, float* __restrict__ x,
                                                       int sweeps, float omega, const SolveScalars* __restrict__ sc) {
    if (sc && sc->done) return;
    __shared__ float sx[2][COARSE_MAX];
    __shared__ float sb[COARSE_MAX];
    __shared__ uint16_t scode[COARSE_MAX];
    const int n = (int)L.n;
    const long long WH = (long long)L.W * L.H;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        sb[c] = b[c];
        scode[c] = L.code[c];
        sx[0][c] = 0.0f;
    }
    __syncthreads();
    int cur = 0;
    for (int s = 0; s < sweeps; ++s) {
        for (int c = threadIdx.x; c < n; c += blockDim.x) {
            uint32_t code = scode[c];
            float xn = 0.0f;
            uint32_t d = (code >> 8) & 7u;
            if ((code & SELF) && d) xn = sx[cur][c] + omega * (sb[c] - stencil_Ax(sx[cur], c, code, L.W, WH)) * inv_diag(d);
            sx[cur ^ 1][c] = xn;
        }
        __syncthreads();
        cur ^= 1;
    }
    for (int c = threadIdx.x; c < n; c += blockDim.x) x[c] = sx[cur][c];
}

// ---------------------------------------------------------------- fine level, 4 cells per thread
// Used when the level's row length W is a multiple of 4 (the solve box is
// widened to one when the grid allows): every row of 4 cells is 16-byte
// aligned, so the centre and the +-y, +-z neighbours of 4 cells are single
// float4 (ushort4 for the codes) loads; the +-x neighbours outside the group
// are two scalar loads.  The arithmetic per cell is that of the scalar kernels.
struct Codes4 {
    uint32_t c[4];
    uint32_t any;   // union of the neighbour bits
};

FS_DEV Codes4 load_codes4(const uint16_t* __restrict__ code, long long c0) {
    const ushort4 v = *reinterpret_cast<const ushort4*>(code + c0);
    Codes4 k;
    k.c[0] = v.x; k.c[1] = v.y; k.c[2] = v.z; k.c[3] = v.w;
    k.any = (k.c[0] | k.c[1] | k.c[2] | k.c[3]) & 63u;
    return k;
}

FS_DEV float4 ld4(const float* __restrict__ p, long long c) { return *reinterpret_cast<const float4*>(p + c); }
FS_DEV float f4(const float4& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w)); }

// The 7-point neighbourhood of 4 cells of a field given by Fld: vec(c) -> values
// at c..c+3 (c a multiple of 4), one(c) -> value at c.  Loads only what a code
// bit asks for (a neighbour whose bit is set is inside the level).
struct Nbhd4 {
    float ctr[4], xm[4], xp[4], ym[4], yp[4], zm[4], zp[4];
};

template <class Fld>
FS_DEV void gather4(const Fld& f, long long c0, const Codes4& k, int W, long long WH, Nbhd4& o) {
    const float4 c = f.vec(c0);
    o.ctr[0] = c.x; o.ctr[1] = c.y; o.ctr[2] = c.z; o.ctr[3] = c.w;
    const float left = (k.c[0] & 2u) ? f.one(c0 - 1) : 0.0f;
    const float right = (k.c[3] & 1u) ? f.one(c0 + 4) : 0.0f;
    o.xm[0] = left; o.xm[1] = c.x; o.xm[2] = c.y; o.xm[3] = c.z;
    o.xp[0] = c.y; o.xp[1] = c.z; o.xp[2] = c.w; o.xp[3] = right;
    float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 ym = (k.any & 8u) ? f.vec(c0 - W) : z4;
    const float4 yp = (k.any & 4u) ? f.vec(c0 + W) : z4;
    const float4 zm = (k.any & 32u) ? f.vec(c0 - WH) : z4;
    const float4 zp = (k.any & 16u) ? f.vec(c0 + WH) : z4;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        o.ym[q] = f4(ym, q); o.yp[q] = f4(yp, q); o.zm[q] = f4(zm, q); o.zp[q] = f4(zp, q);
    }
}

// (A v)_q of cell q of the group, in the operation order of stencil_Ax
FS_DEV float ax4(const Nbhd4& v, uint32_t code, int q) {
    float s = (float)((code >> 8) & 7u) * v.ctr[q];
    if (code & 1u) s -= v.xp[q];
    if (code & 2u) s -= v.xm[q];
    if (code & 4u) s -= v.yp[q];
    if (code & 8u) s -= v.ym[q];
    if (code & 16u) s -= v.zp[q];
    if (code & 32u) s -= v.zm[q];
    return s;
}

// the same with the weights of a WBIT cell c (level 0 with face fractions),
// in stencil_Ax_L's order
FS_DEV float ax4_L(const MGLevel& L, const Nbhd4& v, uint32_t code, int q, long long c, long long WH) {
    if (!(code & WBIT)) return ax4(v, code, q);
    float s = L.wd[c] * v.ctr[q];
    if (code & 1u) s -= L.wx[c] * v.xp[q];
    if (code & 2u) s -= L.wx[c - 1] * v.xm[q];
    if (code & 4u) s -= L.wy[c] * v.yp[q];
    if (code & 8u) s -= L.wy[c - L.W] * v.ym[q];
    if (code & 16u) s -= L.wz[c] * v.zp[q];
    if (code & 32u) s -= L.wz[c - WH] * v.zm[q];
    return s;
}

// 1/diagonal of cell c: the weighted one of a WBIT cell
FS_DEV float inv_diag_L(const MGLevel& L, uint32_t code, long long c) {
    return (code & WBIT) ? __builtin_amdgcn_rcpf(L.wd[c]) : inv_diag((code >> 8) & 7u);
}

struct PlainField {
    const float* __restrict__ p;
    FS_DEV float4 vec(long long c) const { return ld4(p, c); }
    FS_DEV float one(long long c) const { return p[c]; }
};

// omega b / d on unknowns (the first sweep from zero), 0 elsewhere (wd: the
// weighted diagonals of WBIT cells, null without face fractions)
struct FirstSweepField {
    const float* __restrict__ b;
    const uint16_t* __restrict__ code;
    float omega;
    const float* __restrict__ wd;
    FS_DEV float f(float bb, uint32_t cd, long long c) const {
        if (cd & WBIT) return (cd & SELF) ? omega * bb * __builtin_amdgcn_rcpf(wd[c]) : 0.0f;
        uint32_t d = (cd >> 8) & 7u;
        return ((cd & SELF) && d) ? omega * bb * inv_diag(d) : 0.0f;
    }
    FS_DEV float one(long long c) const { return f(b[c], code[c], c); }
    FS_DEV float4 vec(long long c) const {
        const float4 bv = ld4(b, c);
        const ushort4 cv = *reinterpret_cast<const ushort4*>(code + c);
        return make_float4(f(bv.x, cv.x, c), f(bv.y, cv.y, c + 1), f(bv.z, cv.z, c + 2), f(bv.w, cv.w, c + 3));
    }
};

// p = z + beta pold, the new direction as the scalar kernel forms it
struct DirField {
    const float* __restrict__ z;
    const float* __restrict__ pold;
    float beta;
    FS_DEV float one(long long c) const { return z[c] + beta * pold[c]; }
    FS_DEV float4 vec(long long c) const {
        const float4 a = ld4(z, c), b = ld4(pold, c);
        return make_float4(a.x + beta * b.x, a.y + beta * b.y, a.z + beta * b.z, a.w + beta * b.w);
    }
};

__global__ __launch_bounds__(256) void k_jacobi2_v4(MGLevel L, const float* __restrict__ b, float* __restrict__ xout,
                                                    float omega, float omega2, const SolveScalars* __restrict__ sc) {
    if (sc && sc->done) return;
    const long long c0 = 4 * ((long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x);
    if (c0 >= L.n) return;
    const Codes4 k = load_codes4(L.code, c0);
    const long long WH = (long long)L.W * L.H;
    Nbhd4 v;
    gather4(FirstSweepField{b, L.code, omega, L.wd}, c0, k, L.W, WH, v);
    const float4 bv = ld4(b, c0);
    float out[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t code = k.c[q], d = (code >> 8) & 7u;
        float xn = 0.0f;
        if ((code & SELF) && d) {
            const float bc = f4(bv, q);
            const float x1 = v.ctr[q];
            const float ax = ax4_L(L, v, code, q, c0 + q, WH);
            xn = x1 + omega2 * (bc - ax) * inv_diag_L(L, code, c0 + q);
        }
        out[q] = xn;
    }
    *reinterpret_cast<float4*>(xout + c0) = make_float4(out[0], out[1], out[2], out[3]);
}

template <bool DOT>
__global__ __launch_bounds__(256) void k_jacobi_v4(MGLevel L, const float* __restrict__ b, const float* __restrict__ xin,
                                                   float* __restrict__ xout, float omega, double* __restrict__ partial,
                                                   const SolveScalars* __restrict__ sc) {
    if (sc && sc->done) return;
    __shared__ double sm[12];
    const long long WH = (long long)L.W * L.H;
    double acc = 0.0;
    for (long long c0 = 4 * ((long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x); c0 < L.n;
         c0 += 4LL * gridDim.x * blockDim.x) {
        const Codes4 k = load_codes4(L.code, c0);
        Nbhd4 v;
        gather4(PlainField{xin}, c0, k, L.W, WH, v);
        const float4 bv = ld4(b, c0);
        float out[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t code = k.c[q], d = (code >> 8) & 7u;
            float xn = 0.0f;
            if ((code & SELF) && d) {
                const float bc = f4(bv, q);
                xn = v.ctr[q] + omega * (bc - ax4_L(L, v, code, q, c0 + q, WH)) * inv_diag_L(L, code, c0 + q);
                if (DOT) acc += (double)bc * (double)xn;
            }
            out[q] = xn;
        }
        *reinterpret_cast<float4*>(xout + c0) = make_float4(out[0], out[1], out[2], out[3]);
    }
    if (DOT) store_partial(acc, sm, partial);
}

__global__ __launch_bounds__(256) void k_residual_v4(MGLevel L, const float* __restrict__ b, const float* __restrict__ x,
                                                     float* __restrict__ r, const SolveScalars* __restrict__ sc) {
    if (sc && sc->done) return;
    const long long c0 = 4 * ((long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x);
    if (c0 >= L.n) return;
    const Codes4 k = load_codes4(L.code, c0);
    Nbhd4 v;
    gather4(PlainField{x}, c0, k, L.W, (long long)L.W * L.H, v);
    const float4 bv = b ? ld4(b, c0) : make_float4(0.f, 0.f, 0.f, 0.f);
    float out[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        out[q] = (k.c[q] & SELF) ? f4(bv, q) - ax4_L(L, v, k.c[q], q, c0 + q, (long long)L.W * L.H) : 0.0f;
    *reinterpret_cast<float4*>(r + c0) = make_float4(out[0], out[1], out[2], out[3]);
}

// fine x += P xc for 4 fine cells along x (they share their coarse columns)
// The four fine cells 4m..4m+3 of a row share their coarse columns 2m-1 ..
// 2m+2 (16 coarse values for the group instead of 8 per cell); each cell's
// sum is formed in prolong_at's order.  Coarse indices fit in 32 bits.
__global__ __launch_bounds__(256) void k_prolong_v4(MGLevel F, MGLevel Cc, const float* __restrict__ xc,
                                                    float* __restrict__ xf, const SolveScalars* __restrict__ sc) {
    if (sc && sc->done) return;
    const long long c0 = 4 * ((long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x);
    if (c0 >= F.n) return;
    const Codes4 k = load_codes4(F.code, c0);
    if (!((k.c[0] | k.c[1] | k.c[2] | k.c[3]) & SELF)) return;
    int i0, j, kk;
    split3(c0, F.W, F.H, i0, j, kk);
    const int m2 = i0 >> 1;   // even; columns m2 - 1 .. m2 + 2
    const int J0 = j >> 1, K0 = kk >> 1;
    const int Js[2] = {J0, (j & 1) ? J0 + 1 : J0 - 1}, Ks[2] = {K0, (kk & 1) ? K0 + 1 : K0 - 1};
    const bool jok[2] = {true, Js[1] >= 0 && Js[1] < Cc.H}, kok[2] = {true, Ks[1] >= 0 && Ks[1] < Cc.D};
    const bool lok = m2 >= 1, rok = m2 + 2 < Cc.W;
    float col[2][2][4];   // [K][J][column m2 - 1 + t]
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int bb = 0; bb < 2; ++bb) {
            col[a][bb][0] = col[a][bb][1] = col[a][bb][2] = col[a][bb][3] = 0.0f;
            if (!(kok[a] && jok[bb])) continue;
            const int row = Js[bb] * Cc.W + Ks[a] * Cc.W * Cc.H;
            const float2 mid = *reinterpret_cast<const float2*>(xc + row + m2);
            col[a][bb][1] = mid.x;
            col[a][bb][2] = mid.y;
            if (lok) col[a][bb][0] = xc[row + m2 - 1];
            if (rok) col[a][bb][3] = xc[row + m2 + 2];
        }
    float4 xv = ld4(xf, c0);
    float out[4] = {xv.x, xv.y, xv.z, xv.w};
    const float w[2] = {0.75f, 0.25f};
    // cell q: its own column and its second column (t indices into col), and
    // whether the second one exists
    const int t0[4] = {1, 1, 2, 2}, t1[4] = {0, 2, 1, 3};
    const bool ok1[4] = {lok, true, true, rok};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (!(k.c[q] & SELF)) continue;
        float s = 0.0f;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            if (!kok[a]) continue;
#pragma unroll
            for (int bb = 0; bb < 2; ++bb) {
                if (!jok[bb]) continue;
                s += w[0] * w[bb] * w[a] * col[a][bb][t0[q]];
                if (ok1[q]) s += w[1] * w[bb] * w[a] * col[a][bb][t1[q]];
            }
        }
        out[q] += s;
    }
    *reinterpret_cast<float4*>(xf + c0) = make_float4(out[0], out[1], out[2], out[3]);
}

// coarse b = 1/2 P^T r for two coarse cells along x per thread (fine W a
// multiple of 4): their fine columns 4m-1 .. 4m+4 are one float4 and two
// scalars per fine row, shared by both cells
__global__ __launch_bounds__(256) void k_restrict_v2(MGLevel F, MGLevel Cc, const float* __restrict__ rf,
                                                     float* __restrict__ bc, const SolveScalars* __restrict__ sc) {
    if (sc && sc->done) return;
    const long long pidx = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int half = Cc.W >> 1;
    if (pidx >= Cc.n / 2) return;
    const int m = (int)(pidx % half);
    const int J = (int)((pidx / half) % Cc.H), K = (int)(pidx / ((long long)half * Cc.H));
    const long long c = (long long)(2 * m) + (long long)J * Cc.W + (long long)K * Cc.W * Cc.H;
    const bool s0 = (Cc.code[c] & SELF) != 0, s1 = (Cc.code[c + 1] & SELF) != 0;
    float q0[4] = {0.0f, 0.0f, 0.0f, 0.0f}, q1[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (s0 || s1) {
        const float wt[4] = {0.25f, 0.75f, 0.75f, 0.25f};
        const long long WH = (long long)F.W * F.H;
        const int x0 = 4 * m;
#pragma unroll
        for (int dk = 0; dk < 4; ++dk) {
            int k = 2 * K - 1 + dk;
            if (k < 0 || k >= F.D) continue;
            float sum0 = 0.0f, sum1 = 0.0f;
            for (int dj = 0; dj < 4; ++dj) {
                int j = 2 * J - 1 + dj;
                if (j < 0 || j >= F.H) continue;
                const float wjk = wt[dj] * wt[dk];
                const long long row = (long long)j * F.W + (long long)k * WH;
                const float4 v = ld4(rf, row + x0);
                const float l = x0 > 0 ? rf[row + x0 - 1] : 0.0f;
                const float r = x0 + 4 < F.W ? rf[row + x0 + 4] : 0.0f;
                // cell 2m: fine 4m-1 .. 4m+2 ; cell 2m+1: fine 4m+1 .. 4m+4
                if (x0 > 0) sum0 += wt[0] * wjk * l;
                sum0 += wt[1] * wjk * v.x;
                sum0 += wt[2] * wjk * v.y;
                sum0 += wt[3] * wjk * v.z;
                sum1 += wt[0] * wjk * v.y;
                sum1 += wt[1] * wjk * v.z;
                sum1 += wt[2] * wjk * v.w;
                if (x0 + 4 < F.W) sum1 += wt[3] * wjk * r;
            }
            q0[dk] = sum0;
            q1[dk] = sum1;
        }
    }
    bc[c] = s0 ? restrict_sum(q0[0], q0[1], q0[2], q0[3]) : 0.0f;
    bc[c + 1] = s1 ? restrict_sum(q1[0], q1[1], q1[2], q1[3]) : 0.0f;
}

// k_pupdate_spmv for 4 cells per thread (same sums, FP64 A p)
__global__ __launch_bounds__(256) void k_pupdate_spmv_v4(MGLevel L, const float* __restrict__ z,
                                                         const float* __restrict__ pold, float* __restrict__ pnew,
                                                         float* __restrict__ q, const double* __restrict__ rz_new,
                                                         const double* __restrict__ rz_old,
                                                         const double* __restrict__ zq,
                                                         const double* __restrict__ pq_old, int np,
                                                         double* __restrict__ ppq, const float* __restrict__ r,
                                                         double* __restrict__ prp,
                                                         const SolveScalars* __restrict__ sc) {
    if (sc->done) return;
    __shared__ double sm[12];
    const bool first = sc->first != 0;
    float beta = 0.0f;
    if (!first) {
        if (zq) {
            double r2[2];
            sum_partials_k<2>({zq, pq_old}, np, sm, r2);
            const double a = r2[0], bsum = r2[1];
            beta = bsum > 0.0 ? (float)(-a / bsum) : 0.0f;
        } else {
            double r2[2];
            sum_partials_k<2>({rz_new, rz_old}, np, sm, r2);
            const double bn = r2[0], bo = r2[1];
            beta = (float)(bn / bo);
        }
    }
    const long long WH = (long long)L.W * L.H;
    double acc = 0.0, acc_rp = 0.0;
    for (long long c0 = 4 * ((long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x); c0 < L.n;
         c0 += 4LL * gridDim.x * blockDim.x) {
        const Codes4 k = load_codes4(L.code, c0);
        Nbhd4 v;
        gather4(DirField{z, pold, beta}, c0, k, L.W, WH, v);
        float pn[4], qc[4];
        float4 rv = make_float4(0.f, 0.f, 0.f, 0.f);
        if (prp) rv = ld4(r, c0);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t code = k.c[t];
            pn[t] = 0.0f;
            qc[t] = 0.0f;
            if (code & SELF) {
                pn[t] = v.ctr[t];
                double qd;
                if (code & WBIT) {   // the weighted row, in stencil_Ax64_w's order
                    const float nb[6] = {v.xp[t], v.xm[t], v.yp[t], v.ym[t], v.zp[t], v.zm[t]};
                    qd = stencil_Ax64_w(L, c0 + t, code, WH, (double)pn[t], [&](int m) { return (double)nb[m]; });
                } else {
                    qd = (double)((code >> 8) & 7u) * (double)pn[t];
                    if (code & 1u) qd -= (double)v.xp[t];
                    if (code & 2u) qd -= (double)v.xm[t];
                    if (code & 4u) qd -= (double)v.yp[t];
                    if (code & 8u) qd -= (double)v.ym[t];
                    if (code & 16u) qd -= (double)v.zp[t];
                    if (code & 32u) qd -= (double)v.zm[t];
                }
                qc[t] = (float)qd;
                acc += (double)pn[t] * qd;
                if (prp) acc_rp += (double)f4(rv, t) * (double)pn[t];
            }
        }
        *reinterpret_cast<float4*>(pnew + c0) = make_float4(pn[0], pn[1], pn[2], pn[3]);
        *reinterpret_cast<float4*>(q + c0) = make_float4(qc[0], qc[1], qc[2], qc[3]);
    }
    store_partial(acc, sm, ppq);
    if (prp) store_partial(acc_rp, sm, prp);
}

__global__ __launch_bounds__(256) void k_update_v4(long long n, const float* __restrict__ p, const float* __restrict__ q,
                                                   double* __restrict__ x64, float* __restrict__ r,
                                                   const double* __restrict__ rz, const double* __restrict__ pq, int np,
                                                   double* __restrict__ prr, const SolveScalars* __restrict__ sc) {
    if (sc->done) return;
    __shared__ double sm[12];
    double r2[2];
    sum_partials_k<2>({rz, pq}, np, sm, r2);
    const double srz = r2[0], spq = r2[1];
    const float alpha = (spq > 0.0) ? (float)(srz / spq) : 0.0f;
    double acc = 0.0;
    for (long long c0 = 4 * ((long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x); c0 < n;
         c0 += 4LL * gridDim.x * blockDim.x) {
        const float4 pv = ld4(p, c0), qv = ld4(q, c0), rv0 = ld4(r, c0);
        double2 xa = *reinterpret_cast<const double2*>(x64 + c0), xb2 = *reinterpret_cast<const double2*>(x64 + c0 + 2);
        const float rc0 = rv0.x - alpha * qv.x, rc1 = rv0.y - alpha * qv.y, rc2 = rv0.z - alpha * qv.z,
                    rc3 = rv0.w - alpha * qv.w;
        xa.x += (double)alpha * (double)pv.x; xa.y += (double)alpha * (double)pv.y;
        xb2.x += (double)alpha * (double)pv.z; xb2.y += (double)alpha * (double)pv.w;
        *reinterpret_cast<double2*>(x64 + c0) = xa;
        *reinterpret_cast<double2*>(x64 + c0 + 2) = xb2;
        *reinterpret_cast<float4*>(r + c0) = make_float4(rc0, rc1, rc2, rc3);
        acc += (double)rc0 * (double)rc0 + (double)rc1 * (double)rc1 + (double)rc2 * (double)rc2 +
               (double)rc3 * (double)rc3;
    }
    store_partial(acc, sm, prr);
}

// ---------------------------------------------------------------- small levels
// The V-cycle from the first level of at most TAIL_MAX cells down to the
// coarsest, in one 1024-thread workgroup with every iterate in LDS.  Each of
// these levels is a few thousand cells, so as separate launches every phase
// would cost a kernel's fixed latency; here a phase is a workgroup barrier.
// The operations and their floating-point order are those of the per-level
// kernels (k_jacobi2, k_jacobi, k_residual, k_restrict, prolong_at,
// k_coarse_solve), so the tail gives the same bits as the launch sequence.
// LDS layout (floats): top level x, xt (its b and codes stay in registers,
// TAIL_CPT cells per thread); each lower level x, xt, b and its codes
// (uint16) at off[q].  Only the top level's result is written back (to x).
constexpr int TAIL_LEVELS = 6;
constexpr int TAIL_THREADS = 1024;
constexpr int TAIL_CPT = 4;
constexpr long long TAIL_MAX = (long long)TAIL_THREADS * TAIL_CPT;
constexpr int TAIL_LDS_FLOATS = 38912;   // 152 KiB of the 160 KiB per workgroup
struct TailLevels {
    MGLevel lv[TAIL_LEVELS];
    int off[TAIL_LEVELS];
    int n;        // lv[n - 1] is the coarsest level
    float w[SMOOTH_STEPS];   // smoothing step of pre/post sweep k
};

// keeps the scheduler from interleaving the unrolled per-cell bodies of the
// top level (which would hold every cell's operands at once and spill)
#define CELL_FENCE() __builtin_amdgcn_sched_barrier(0)
// makes the row and plane strides opaque at the start of a top-level pass, so
// that the neighbour addresses of the TAIL_CPT cells are formed inside the
// pass instead of being hoisted out of the sweep loop and held in registers
#define OPAQUE_STRIDES(W, WH) asm volatile("" : "+s"(W), "+s"(WH))
// the same for the register-held b and codes: values derived from them (1/d,
// the first sweep's x) are recomputed in each pass rather than kept live
#define OPAQUE_TOP(b, c)                                        \
    _Pragma("unroll") for (int q_ = 0; q_ < TAIL_CPT; ++q_)     \
        asm volatile("" : "+v"(b[q_]), "+v"(c[q_]))

// floats of LDS for a level below the top one: x, xt, b and the codes
FS_HD int tail_level_floats(long long n) { return (int)((3 * n + (n + 1) / 2 + 3) & ~3LL); }

// Jacobi sweep on an LDS level: xout = xin + omega (b - A xin) / d
FS_DEV void lds_sweep(const MGLevel& L, const float* xin, float* xout, const float* b, const uint16_t* code,
                      float omega) {
    const int W = L.W, WH = L.W * L.H, n = (int)L.n;
    for (int c = threadIdx.x; c < n; c += TAIL_THREADS) {
        uint32_t cd = code[c];
        float xn = 0.0f;
        uint32_t d = (cd >> 8) & 7u;
        if ((cd & SELF) && d) xn = xin[c] + omega * (b[c] - stencil_Ax(xin, c, cd, W, WH)) * inv_diag(d);
        xout[c] = xn;
    }
}

// b of the coarse level Cc from the fine residual rf (both in LDS)
FS_DEV void lds_restrict(const MGLevel& F, const float* rf, const MGLevel& Cc, float* bc, const uint16_t* ccode) {
    const int n = (int)Cc.n;
    for (int c = threadIdx.x; c < n; c += TAIL_THREADS) {
        float v = 0.0f;
        if (ccode[c] & SELF) {
            int I, J, K;
            split3(c, Cc.W, Cc.H, I, J, K);
            v = restrict_sum(restrict_slice(F.W, F.H, F.D, rf, I, J, K, 0),
                             restrict_slice(F.W, F.H, F.D, rf, I, J, K, 1),
                             restrict_slice(F.W, F.H, F.D, rf, I, J, K, 2),
                             restrict_slice(F.W, F.H, F.D, rf, I, J, K, 3));
        }
        bc[c] = v;
    }
}

// x += P xc on the unknowns of an LDS level
FS_DEV void lds_prolong(const MGLevel& L, float* x, const uint16_t* code, const MGLevel& Cc, const float* xc) {
    const int n = (int)L.n;
    for (int c = threadIdx.x; c < n; c += TAIL_THREADS) {
        if (!(code[c] & SELF)) continue;
        int i, j, k;
        split3(c, L.W, L.H, i, j, k);
        x[c] += prolong_at(Cc, xc, i, j, k);
    }
}

__global__ __launch_bounds__(TAIL_THREADS) void k_tail(TailLevels t, int npre, int npost, int coarse_sweeps,
                                                       float omega, const SolveScalars* __restrict__ sc) {
    if (sc && sc->done) return;
    __shared__ float lds[TAIL_LDS_FLOATS];
    const int tid = threadIdx.x;
    // omega: the coarsest level's sweeps; t.w[k % SMOOTH_STEPS]: pre/post sweep k
    const float w0 = t.w[0], omega2 = npre >= 2 ? t.w[1] : 0.0f;
    const MGLevel T = t.lv[0];
    const int nT = (int)T.n;
    int WT = T.W, WHT = T.W * T.H;
    // top level: b and codes in registers, cells tid + q * TAIL_THREADS
    float bT[TAIL_CPT];
    uint32_t cT[TAIL_CPT];
#pragma unroll
    for (int q = 0; q < TAIL_CPT; ++q) {
        const int c = tid + q * TAIL_THREADS;
        bT[q] = 0.0f;
        cT[q] = 0u;
        if (c < nT) {
            bT[q] = T.b[c];
            cT[q] = T.code[c];
        }
    }
    // codes of the lower levels
    for (int l = 1; l < t.n; ++l) {
        const MGLevel& L = t.lv[l];
        uint16_t* code = reinterpret_cast<uint16_t*>(lds + t.off[l] + 3 * L.n);
        for (int c = tid; c < (int)L.n; c += TAIL_THREADS) code[c] = L.code[c];
    }
    float* tx = lds;
    float* txt = lds + nT;
    // ---- top level down: two sweeps from zero (k_jacobi2), extra pre sweeps, residual
#pragma unroll
    for (int q = 0; q < TAIL_CPT; ++q) {
        const int c = tid + q * TAIL_THREADS;
        const uint32_t d = (cT[q] >> 8) & 7u;
        if (c < nT) tx[c] = ((cT[q] & SELF) && d) ? w0 * bT[q] * inv_diag(d) : 0.0f;
    }
    __syncthreads();
    OPAQUE_TOP(bT, cT);
#pragma unroll
    for (int q = 0; q < TAIL_CPT; ++q) {
        const int c = tid + q * TAIL_THREADS;
        if (c >= nT) continue;
        const uint32_t code = cT[q], d = (code >> 8) & 7u;
        float xn = 0.0f;
        if ((code & SELF) && d) {
            const float x1 = w0 * bT[q] * inv_diag(d);
            float ax = (float)d * x1;
            if (code & 1u) ax -= tx[c + 1];
            if (code & 2u) ax -= tx[c - 1];
            if (code & 4u) ax -= tx[c + WT];
            if (code & 8u) ax -= tx[c - WT];
            if (code & 16u) ax -= tx[c + WHT];
            if (code & 32u) ax -= tx[c - WHT];
            xn = x1 + omega2 * (bT[q] - ax) * inv_diag(d);
        }
        txt[c] = xn;
        CELL_FENCE();
    }
    __syncthreads();
    float* cur = txt;
    float* oth = tx;
    for (int k = 2; k < npre; ++k) {
        OPAQUE_STRIDES(WT, WHT);
        OPAQUE_TOP(bT, cT);
#pragma unroll
        for (int q = 0; q < TAIL_CPT; ++q) {
            const int c = tid + q * TAIL_THREADS;
            if (c >= nT) continue;
            const uint32_t code = cT[q], d = (code >> 8) & 7u;
            float xn = 0.0f;
            if ((code & SELF) && d)
                xn = cur[c] + t.w[k % SMOOTH_STEPS] * (bT[q] - stencil_Ax(cur, c, code, WT, WHT)) * inv_diag(d);
            oth[c] = xn;
            CELL_FENCE();
        }
        __syncthreads();
        float* tmp = cur; cur = oth; oth = tmp;
    }
    OPAQUE_STRIDES(WT, WHT);
    OPAQUE_TOP(bT, cT);
#pragma unroll
    for (int q = 0; q < TAIL_CPT; ++q) {
        const int c = tid + q * TAIL_THREADS;
        if (c < nT) oth[c] = (cT[q] & SELF) ? bT[q] - stencil_Ax(cur, c, cT[q], WT, WHT) : 0.0f;
        CELL_FENCE();
    }
    __syncthreads();
    {
        const MGLevel& C1 = t.lv[1];
        lds_restrict(T, oth, C1, lds + t.off[1] + 2 * C1.n,
                     reinterpret_cast<const uint16_t*>(lds + t.off[1] + 3 * C1.n));
    }
    __syncthreads();
    float* top = cur;   // the top level's iterate
    // ---- lower levels down; bit l of in_xt: level l's iterate is in its xt
    unsigned in_xt = 0;
    for (int l = 1; l < t.n - 1; ++l) {
        const MGLevel& L = t.lv[l];
        float* x = lds + t.off[l];
        float* xt = x + L.n;
        const float* b = x + 2 * L.n;
        const uint16_t* code = reinterpret_cast<const uint16_t*>(x + 3 * L.n);
        const int W = L.W, WH = L.W * L.H, n = (int)L.n;
        for (int c = tid; c < n; c += TAIL_THREADS) {
            const uint32_t cd = code[c], d = (cd >> 8) & 7u;
            x[c] = ((cd & SELF) && d) ? w0 * b[c] * inv_diag(d) : 0.0f;
        }
        __syncthreads();
        for (int c = tid; c < n; c += TAIL_THREADS) {
            const uint32_t cd = code[c], d = (cd >> 8) & 7u;
            float xn = 0.0f;
            if ((cd & SELF) && d) {
                const float x1 = w0 * b[c] * inv_diag(d);
                float ax = (float)d * x1;
                if (cd & 1u) ax -= x[c + 1];
                if (cd & 2u) ax -= x[c - 1];
                if (cd & 4u) ax -= x[c + W];
                if (cd & 8u) ax -= x[c - W];
                if (cd & 16u) ax -= x[c + WH];
                if (cd & 32u) ax -= x[c - WH];
                xn = x1 + omega2 * (b[c] - ax) * inv_diag(d);
            }
            xt[c] = xn;
        }
        __syncthreads();
        float* lc = xt;
        float* lo = x;
        for (int k = 2; k < npre; ++k) {
            lds_sweep(L, lc, lo, b, code, t.w[k % SMOOTH_STEPS]);
            __syncthreads();
            float* tmp = lc; lc = lo; lo = tmp;
        }
        for (int c = tid; c < n; c += TAIL_THREADS) {
            const uint32_t cd = code[c];
            lo[c] = (cd & SELF) ? b[c] - stencil_Ax(lc, c, cd, W, WH) : 0.0f;
        }
        __syncthreads();
        const MGLevel& Cn = t.lv[l + 1];
        lds_restrict(L, lo, Cn, lds + t.off[l + 1] + 2 * Cn.n,
                     reinterpret_cast<const uint16_t*>(lds + t.off[l + 1] + 3 * Cn.n));
        __syncthreads();
        if (lc == xt) in_xt |= 1u << l;
    }
    // ---- coarsest: damped-Jacobi sweeps from zero (k_coarse_solve)
    {
        const int l = t.n - 1;
        const MGLevel& L = t.lv[l];
        float* x = lds + t.off[l];
        float* xt = x + L.n;
        const float* b = x + 2 * L.n;
        const uint16_t* code = reinterpret_cast<const uint16_t*>(x + 3 * L.n);
        for (int c = tid; c < (int)L.n; c += TAIL_THREADS) x[c] = 0.0f;
        __syncthreads();
        float* lc = x;
        float* lo = xt;
        for (int sw = 0; sw < coarse_sweeps; ++sw) {
            lds_sweep(L, lc, lo, b, code, omega);
            __syncthreads();
            float* tmp = lc; lc = lo; lo = tmp;
        }
        if (lc == xt) in_xt |= 1u << l;
    }
    // ---- lower levels up: prolongation, post sweeps
    for (int l = t.n - 2; l >= 1; --l) {
        const MGLevel& L = t.lv[l];
        const MGLevel& Cn = t.lv[l + 1];
        float* x = lds + t.off[l];
        float* xt = x + L.n;
        const float* b = x + 2 * L.n;
        const uint16_t* code = reinterpret_cast<const uint16_t*>(x + 3 * L.n);
        const float* xc = lds + t.off[l + 1] + (((in_xt >> (l + 1)) & 1u) ? Cn.n : 0);
        float* lc = ((in_xt >> l) & 1u) ? xt : x;
        float* lo = (lc == x) ? xt : x;
        lds_prolong(L, lc, code, Cn, xc);
        __syncthreads();
        for (int k = 0; k < npost; ++k) {
            lds_sweep(L, lc, lo, b, code, t.w[k % SMOOTH_STEPS]);
            __syncthreads();
            float* tmp = lc; lc = lo; lo = tmp;
        }
        in_xt = (in_xt & ~(1u << l)) | ((lc == xt ? 1u : 0u) << l);
    }
    // ---- top level up
    {
        const MGLevel& C1 = t.lv[1];
        const float* xc = lds + t.off[1] + (((in_xt >> 1) & 1u) ? C1.n : 0);
        cur = top;
        oth = (cur == tx) ? txt : tx;
        OPAQUE_TOP(bT, cT);
#pragma unroll
        for (int q = 0; q < TAIL_CPT; ++q) {
            const int c = tid + q * TAIL_THREADS;
            if (c >= nT || !(cT[q] & SELF)) continue;
            int i, j, k;
            split3(c, T.W, T.H, i, j, k);
            cur[c] += prolong_at(C1, xc, i, j, k);
            CELL_FENCE();
        }
        __syncthreads();
        for (int k = 0; k < npost; ++k) {
            OPAQUE_STRIDES(WT, WHT);
        OPAQUE_TOP(bT, cT);
#pragma unroll
            for (int q = 0; q < TAIL_CPT; ++q) {
                const int c = tid + q * TAIL_THREADS;
                if (c >= nT) continue;
                const uint32_t code = cT[q], d = (code >> 8) & 7u;
                float xn = 0.0f;
                if ((code & SELF) && d)
                    xn = cur[c] + t.w[k % SMOOTH_STEPS] * (bT[q] - stencil_Ax(cur, c, code, WT, WHT)) * inv_diag(d);
                oth[c] = xn;
            }
            __syncthreads();
            float* tmp = cur; cur = oth; oth = tmp;
        }
#pragma unroll
        for (int q = 0; q < TAIL_CPT; ++q) {
            const int c = tid + q * TAIL_THREADS;
            if (c < nT) T.x[c] = cur[c];
        }
    }
}

// ---------------------------------------------------------------- CG kernels
// beta = (r.z)_new / (r.z)_old from the partial sums (0 on the first iteration of
// a refinement pass); flexible variant (zq given): beta = -(z.A p_old) / (p_old.A p_old),
// which makes p_new A-conjugate to p_old.  p_new = z + beta p_old ; q = A p_new ;
// partials of p_new.q, and of r.p_new when prp is given (the exact line search
// of the flexible variant, alpha = r.p / p.q)
__global__ __launch_bounds__(256) void k_pupdate_spmv(MGLevel L, const float* __restrict__ z,
                                                      const float* __restrict__ pold, float* __restrict__ pnew,
                                                      float* __restrict__ q, const double* __restrict__ rz_new,
                                                      const double* __restrict__ rz_old,
                                                      const double* __restrict__ zq,
                                                      const double* __restrict__ pq_old, int np,
                                                      double* __restrict__ ppq, const float* __restrict__ r,
                                                      double* __restrict__ prp, const SolveScalars* __restrict__ sc) {
    if (sc->done) return;
    __shared__ double sm[12];
    const bool first = sc->first != 0;
    float beta = 0.0f;
    if (!first) {
        if (zq) {
            double r2[2];
            sum_partials_k<2>({zq, pq_old}, np, sm, r2);
            const double a = r2[0], b = r2[1];
            beta = b > 0.0 ? (float)(-a / b) : 0.0f;
        } else {
            double r2[2];
            sum_partials_k<2>({rz_new, rz_old}, np, sm, r2);
            const double bn = r2[0], bo = r2[1];
            beta = (float)(bn / bo);
        }
    }
    const long long WH = (long long)L.W * L.H;
    double acc = 0.0, acc_rp = 0.0;
    for (long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; c < L.n;
         c += (long long)gridDim.x * blockDim.x) {
        uint32_t code = L.code[c];
        float pn = 0.0f, qc = 0.0f;
        if (code & SELF) {
            // A p of the stored (rounded) p, summed in FP64: p is smooth and large
            // next to A p, and an FP32 sum would lose A p to cancellation
            pn = z[c] + beta * pold[c];
            double qd;
            if (code & WBIT) {
                const long long off[6] = {1, -1, L.W, -(long long)L.W, WH, -WH};
                qd = stencil_Ax64_w(L, c, code, WH, (double)pn,
                                    [&](int k) { return (double)(z[c + off[k]] + beta * pold[c + off[k]]); });
            } else {
                qd = (double)((code >> 8) & 7u) * (double)pn;
                if (code & 1u) qd -= (double)(z[c + 1] + beta * pold[c + 1]);
                if (code & 2u) qd -= (double)(z[c - 1] + beta * pold[c - 1]);
                if (code & 4u) qd -= (double)(z[c + L.W] + beta * pold[c + L.W]);
                if (code & 8u) qd -= (double)(z[c - L.W] + beta * pold[c - L.W]);
                if (code & 16u) qd -= (double)(z[c + WH] + beta * pold[c + WH]);
                if (code & 32u) qd -= (double)(z[c - WH] + beta * pold[c - WH]);
            }
            qc = (float)qd;
            acc += (double)pn * qd;
            if (prp) acc_rp += (double)r[c] * (double)pn;
        }
        pnew[c] = pn;
        q[c] = qc;
    }
    store_partial(acc, sm, ppq);
    if (prp) store_partial(acc_rp, sm, prp);
}

// alpha = (r.z) / (p.q) (flexible: rz holds r.p) ; x64 += alpha p (FP64: the
// product of two floats is exact in double, so the solution follows the
// recurrence instead of losing the true residual at the FP32 resolution of a
// large correction) ; r -= alpha q ; partials of r.r
__global__ __launch_bounds__(256) void k_update(long long n, const float* __restrict__ p, const float* __restrict__ q,
                                                double* __restrict__ x64, float* __restrict__ r,
                                                const double* __restrict__ rz, const double* __restrict__ pq, int np,
                                                double* __restrict__ prr, const SolveScalars* __restrict__ sc) {
    if (sc->done) return;
    __shared__ double sm[12];
    double r2[2];
    sum_partials_k<2>({rz, pq}, np, sm, r2);
    const double srz = r2[0], spq = r2[1];
    const float alpha = (spq > 0.0) ? (float)(srz / spq) : 0.0f;
    double acc = 0.0;
    for (long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; c < n;
         c += (long long)gridDim.x * blockDim.x) {
        float rc = r[c] - alpha * q[c];
        x64[c] += (double)alpha * (double)p[c];
        r[c] = rc;
        acc += (double)rc * (double)rc;
    }
    store_partial(acc, sm, prr);
}

// end of an iteration: ||r||^2, iteration count, stop flags.  flex: a flexible
// preconditioner (r.z may take either sign; only p.q <= 0 is a breakdown);
// stall_window > 0: done = 5 after that many iterations without a new smallest
// ||r||^2 in the refinement pass.
__global__ __launch_bounds__(256) void k_fin_iter(const double* __restrict__ rz, const double* __restrict__ pq,
                                                  const double* __restrict__ rr, int np, int flex, int stall_window,
                                                  SolveScalars* sc) {
    if (sc->done) return;
    __shared__ double sm[12];
    double r3[3];
    sum_partials_k<3>({rz, pq, rr}, np, sm, r3);
    const double srz = r3[0], spq = r3[1], srr = r3[2];
    if (threadIdx.x != 0) return;
    sc->rz = srz;
    sc->rr = srr;
    sc->iters += 1;
    sc->first = 0;
    const bool broken = flex ? (!(spq > 0.0) || isnan(srz)) : (!(srz > 0.0) || !(spq > 0.0));
    if (broken || isnan(srr)) sc->done = 3;   // breakdown: restart from the true residual
    else if (srr <= sc->target) sc->done = 1;
    else if (sc->iters >= sc->maxit) sc->done = 4;
    else if (stall_window > 0) {
        if (srr < sc->best_rr) {
            sc->best_rr = srr;
            sc->stall = 0;
        } else if (++sc->stall >= stall_window) {
            sc->done = 5;
        }
    }
}

// r.z over the solve box (for preconditioners whose last kernel does not
// accumulate it)
__global__ __launch_bounds__(256) void k_dot_rz(long long n, const float* __restrict__ r, const float* __restrict__ z,
                                                double* __restrict__ partial, const SolveScalars* __restrict__ sc) {
    if (sc->done) return;
    __shared__ double sm[12];
    double acc = 0.0;
    for (long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x)
        acc += (double)r[c] * (double)z[c];
    store_partial(acc, sm, partial);
}

// r.z and z.q over the solve box
__global__ __launch_bounds__(256) void k_dot_rz_zq(long long n, const float* __restrict__ r,
                                                   const float* __restrict__ z, const float* __restrict__ q,
                                                   double* __restrict__ prz, double* __restrict__ pzq,
                                                   const SolveScalars* __restrict__ sc) {
    if (sc->done) return;
    __shared__ double sm[12];
    double a = 0.0, b = 0.0;
    for (long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
        const double zc = z[c];
        a += (double)r[c] * zc;
        b += zc * (double)q[c];
    }
    store_partial(a, sm, prz);
    store_partial(b, sm, pzq);
}

// sums of b^2 and b over unknowns
__global__ __launch_bounds__(256) void k_bnorm(MGLevel L, const float* __restrict__ b, double* __restrict__ p2,
                                               double* __restrict__ p1) {
    __shared__ double sm[12];
    double a2 = 0.0, a1 = 0.0;
    for (long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; c < L.n;
         c += (long long)gridDim.x * blockDim.x) {
        if (L.code[c] & SELF) {
            double v = b[c];
            a2 += v * v;
            a1 += v;
        }
    }
    store_partial(a2, sm, p2);
    store_partial(a1, sm, p1);
}

// Level-0 set-up in one pass over the solve box (offset ox, oy, oz in the full
// W x H grid): the type and stencil code of each cell from the labels (as
// k_type0 + k_code on level 0: neighbours outside the box count as none), b
// gathered into the box, the unknown count and Dirichlet flag (counters[0],
// counters[1]) and the partial sums of b^2 and b over the unknowns, with the
// cells visited in k_bnorm's order (same partial sums, same bits).
// FF: a face to a FLUID/AIR neighbour counts only when its open fraction
// (face_open_sdf) is positive; a cell with such a face partly open gets WBIT
// and its diagonal in L.wd (sum of the open fractions of its faces); every
// FLUID cell writes the fractions of its +x/+y/+z faces to L.wx/wy/wz.
template <bool FF>
__global__ __launch_bounds__(256) void k_setup0(const uint8_t* __restrict__ labels, const float* __restrict__ bfull,
                                                MGLevel L, int W, int H, int ox, int oy, int oz,
                                                float* __restrict__ bbox, double* __restrict__ p2,
                                                double* __restrict__ p1, int* __restrict__ counters, FaceFrac ff) {
    __shared__ double sm[12];
    __shared__ int s_cnt[2];
    if (threadIdx.x == 0) { s_cnt[0] = 0; s_cnt[1] = 0; }
    __syncthreads();
    const long long WHf = (long long)W * H;
    double a2 = 0.0, a1 = 0.0;
    int unk = 0, dir = 0;
    auto kind = [](uint8_t lab) { return lab == FS_FLUID ? 1 : (lab == FS_AIR ? 2 : 0); };
    for (long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; c < L.n;
         c += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        split3(c, L.W, L.H, i, j, k);
        const long long g = (long long)(i + ox) + (long long)(j + oy) * W + (long long)(k + oz) * WHf;
        const int t = kind(labels[g]);
        const float bv = bfull[g];
        L.type[c] = (uint8_t)t;
        bbox[c] = bv;
        uint32_t code = 0;
        if (t == 1) {
            int nt[6] = {0, 0, 0, 0, 0, 0};
            if (i + 1 < L.W) nt[0] = kind(labels[g + 1]);
            if (i > 0) nt[1] = kind(labels[g - 1]);
            if (j + 1 < L.H) nt[2] = kind(labels[g + W]);
            if (j > 0) nt[3] = kind(labels[g - W]);
            if (k + 1 < L.D) nt[4] = kind(labels[g + WHf]);
            if (k > 0) nt[5] = kind(labels[g - WHf]);
            float th[6] = {1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 1.0f};
            if (FF) {
                const int gi = i + ox, gj = j + oy, gk = k + oz;
                if (nt[0]) th[0] = face_open_box(ff, 0, gi + 1, gj, gk);
                if (nt[1]) th[1] = face_open_box(ff, 0, gi, gj, gk);
                if (nt[2]) th[2] = face_open_box(ff, 1, gi, gj + 1, gk);
                if (nt[3]) th[3] = face_open_box(ff, 1, gi, gj, gk);
                if (nt[4]) th[4] = face_open_box(ff, 2, gi, gj, gk + 1);
                if (nt[5]) th[5] = face_open_box(ff, 2, gi, gj, gk);
#pragma unroll
                for (int q = 0; q < 6; ++q)
                    if (!nt[q] || th[q] == 0.0f) { nt[q] = 0; th[q] = 0.0f; }
            }
            uint32_t diag = 0;
            int air = 0;
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                if (nt[q] == 1) { code |= (1u << q); ++diag; }
                else if (nt[q] == 2) { ++diag; air = 1; }
            }
            bool part = false;
            if (FF) {
#pragma unroll
                for (int q = 0; q < 6; ++q) part |= th[q] > 0.0f && th[q] < 1.0f;
                // the plus faces of the FLUID cells that a WBIT row reads (its
                // own and its minus neighbours'), the diagonal of the WBIT cells.
                // A WBIT cell has a face in the fraction box, so it and its
                // neighbours lie within two cells of the box: only those are written.
                const int gi = i + ox, gj = j + oy, gk = k + oz;
                const bool near = (unsigned)(gi - ff.lo[0] + 2) < (unsigned)(ff.n[0] + 4) &&
                                  (unsigned)(gj - ff.lo[1] + 2) < (unsigned)(ff.n[1] + 4) &&
                                  (unsigned)(gk - ff.lo[2] + 2) < (unsigned)(ff.n[2] + 4);
                if (near) {
                    const_cast<float*>(L.wx)[c] = th[0];
                    const_cast<float*>(L.wy)[c] = th[2];
                    const_cast<float*>(L.wz)[c] = th[4];
                }
                if (part) const_cast<float*>(L.wd)[c] = ((((th[0] + th[1]) + th[2]) + th[3]) + th[4]) + th[5];
            }
            // an empty row (no FLUID or AIR neighbour) is no unknown
            if (diag) {
                if (part) code |= WBIT;
                code |= SELF | (diag << 8);
                ++unk;
                dir |= air;
                const double v = bv;
                a2 += v * v;
                a1 += v;
            } else {
                code = 0;
            }
        }
        L.code[c] = (uint16_t)code;
    }
    unk = wave_sum(unk);
    dir = wave_sum(dir);
    if ((threadIdx.x & 63) == 0) {
        if (unk) atomicAdd(&s_cnt[0], unk);
        if (dir) atomicOr(&s_cnt[1], 1);
    }
    store_partial(a2, sm, p2);
    store_partial(a1, sm, p1);
    if (threadIdx.x == 0) {
        if (s_cnt[0]) atomicAdd(&counters[0], s_cnt[0]);
        if (s_cnt[1]) atomicOr(&counters[1], 1);
    }
}

// The warm start's first residual: x = x0 (full grid, box offset ox, oy, oz) on
// the unknowns, 0 elsewhere, written to x64, and r = b - shift - A x with the
// operations of k_resid64 (the same bits as gathering x0 first)
__global__ __launch_bounds__(256) void k_resid64_x0(MGLevel L, const float* __restrict__ b,
                                                    const double* __restrict__ x0, int W, int H, int ox, int oy,
                                                    int oz, double shift, float* __restrict__ r,
                                                    double* __restrict__ partial, double* __restrict__ x64) {
    __shared__ double sm[12];
    const long long WHf = (long long)W * H;
    double acc = 0.0;
    for (long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; c < L.n;
         c += (long long)gridDim.x * blockDim.x) {
        const uint32_t code = L.code[c];
        float rf = 0.0f;
        double xc = 0.0;
        if (code & SELF) {
            int i, j, k;
            split3(c, L.W, L.H, i, j, k);
            const long long g = (long long)(i + ox) + (long long)(j + oy) * W + (long long)(k + oz) * WHf;
            xc = x0[g];
            double sx;
            if (code & WBIT) {
                const long long off[6] = {1, -1, W, -(long long)W, WHf, -WHf};
                sx = stencil_Ax64_w(L, c, code, (long long)L.W * L.H, xc, [&](int k) { return x0[g + off[k]]; });
            } else {
                sx = (double)((code >> 8) & 7u) * xc;
                if (code & 1u) sx -= x0[g + 1];
                if (code & 2u) sx -= x0[g - 1];
                if (code & 4u) sx -= x0[g + W];
                if (code & 8u) sx -= x0[g - W];
                if (code & 16u) sx -= x0[g + WHf];
                if (code & 32u) sx -= x0[g - WHf];
            }
            const double rv = ((double)b[c] - shift) - sx;
            acc += rv * rv;
            rf = (float)rv;
        }
        r[c] = rf;
        x64[c] = xc;
    }
    store_partial(acc, sm, partial);
}

// r = (b - shift) - A x64 in FP64 on unknowns; r stored FP32; partials of r.r (FP64)
// ZERO_X: the first pass, which also sets x64 = 0, so r = b - shift
template <bool ZERO_X>
__global__ __launch_bounds__(256) void k_resid64(MGLevel L, const float* __restrict__ b, double* __restrict__ x64,
                                                 double shift, float* __restrict__ r, double* __restrict__ partial,
                                                 const SolveScalars* __restrict__ gate) {
    // gate: run only once the inner iterations have converged (the check
    // that closes an inner batch), as every block reads the same flag
    if (gate && gate->done != 1) return;
    __shared__ double sm[12];
    const long long WH = (long long)L.W * L.H;
    double acc = 0.0;
    for (long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; c < L.n;
         c += (long long)gridDim.x * blockDim.x) {
        uint32_t code = L.code[c];
        float rf = 0.0f;
        if (code & SELF) {
            double rv = (double)b[c] - shift;
            if (!ZERO_X) {
                if (code & WBIT) {
                    const long long off[6] = {1, -1, L.W, -(long long)L.W, WH, -WH};
                    rv -= stencil_Ax64_w(L, c, code, WH, x64[c], [&](int k) { return x64[c + off[k]]; });
                } else {
                    rv -= stencil_Ax64(x64, c, code, L.W, WH);
                }
            }
            acc += rv * rv;
            rf = (float)rv;
        }
        r[c] = rf;
        if (ZERO_X) x64[c] = 0.0;
    }
    store_partial(acc, sm, partial);
}

// Pure-Neumann (singular) systems: the constant vector on the unknowns spans
// the null space.  The preconditioned residual and the solution are kept
// orthogonal to it, so the iterates do not drift along it (a drift that grows
// as the FP32 residual's rounding, which is not orthogonal to the constant,
// is amplified by the near-singular coarse levels).
// Partial sums of v over the unknowns.
template <typename T>
__global__ __launch_bounds__(256) void k_sum_unknowns(MGLevel L, const T* __restrict__ v, double* __restrict__ partial,
                                                      const SolveScalars* __restrict__ sc) {
    if (sc && sc->done) return;
    __shared__ double sm[12];
    double acc = 0.0;
    for (long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; c < L.n;
         c += (long long)gridDim.x * blockDim.x)
        if (L.code[c] & SELF) acc += (double)v[c];
    store_partial(acc, sm, partial);
}

// v -= mean of v over the unknowns (the sum from the partials, folded by
// every block in the same order)
template <typename T>
__global__ __launch_bounds__(256) void k_sub_mean(MGLevel L, T* __restrict__ v, const double* __restrict__ partial,
                                                  int np, int n_unknown, const SolveScalars* __restrict__ sc) {
    if (sc && sc->done) return;
    __shared__ double sm[12];
    const double m = sum_partials(partial, np, sm) / (double)n_unknown;
    const T mt = (T)m;
    for (long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; c < L.n;
         c += (long long)gridDim.x * blockDim.x)
        if (L.code[c] & SELF) v[c] -= mt;
}

// r -= mean of r over the unknowns, and the partial sums of the projected r.r
// (they replace those of the residual update, for the convergence test)
__global__ __launch_bounds__(256) void k_sub_mean_rr(MGLevel L, float* __restrict__ r, const double* __restrict__ partial,
                                                     int np, int n_unknown, double* __restrict__ prr,
                                                     const SolveScalars* __restrict__ sc) {
    if (sc && sc->done) return;
    __shared__ double sm[12];
    const float m = (float)(sum_partials(partial, np, sm) / (double)n_unknown);
    double acc = 0.0;
    for (long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; c < L.n;
         c += (long long)gridDim.x * blockDim.x)
        if (L.code[c] & SELF) {
            const float v = r[c] - m;
            r[c] = v;
            acc += (double)v * (double)v;
        }
    store_partial(acc, sm, prr);
}

// x = x0 (full grid) on the unknowns of the box, 0 elsewhere
__global__ __launch_bounds__(256) void k_gather_x0(MGLevel L, const double* __restrict__ x0, int W, int H, int ox,
                                                   int oy, int oz, double* __restrict__ x) {
    const long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (c >= L.n) return;
    double v = 0.0;
    if (L.code[c] & SELF) {
        int i, j, k;
        split3(c, L.W, L.H, i, j, k);
        v = x0[(size_t)(i + ox) + (size_t)(j + oy) * W + (size_t)(k + oz) * W * H];
    }
    x[c] = v;
}

__global__ __launch_bounds__(256) void k_fill0(long long n, double* __restrict__ x64) {
    long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (c >= n) return;
    x64[c] = 0.0;
}

__global__ __launch_bounds__(256) void k_finalize(int mode, const double* __restrict__ p1,
                                                  const double* __restrict__ p2, int np, SolveScalars* sc,
                                                  double tol2b) {
    // FIN_CHECK: FIN_OUTER with the tolerance of the solve (sc->tol2b), only
    // after inner iterations that converged
    if (mode == FIN_CHECK && sc->done != 1) return;
    __shared__ double sm[12];
    double s1 = sum_partials(p1, np, sm);
    double s2 = p2 ? sum_partials(p2, np, sm) : 0.0;
    if (threadIdx.x != 0) return;
    if (mode == FIN_BNORM) {
        sc->bnorm2 = s1;
        sc->bsum = s2;
    } else {
        // FIN_OUTER_REL: tol2b holds tol^2, relative to the b norm of FIN_BNORM
        if (mode == FIN_OUTER_REL) tol2b *= sc->bnorm2;
        if (mode == FIN_CHECK) tol2b = sc->tol2b;
        sc->rr_true = s1;
        sc->tol2b = tol2b;
        if (s1 <= tol2b) {
            sc->done = 2;
        } else {
            sc->done = 0;
            sc->first = 1;
            sc->rr = s1;
            sc->rr_start = s1;
            sc->best_rr = s1;
            sc->stall = 0;
            sc->target = fmax(sc->inner_rtol2 * s1, 0.25 * tol2b);
        }
    }
}

// ---------------------------------------------------------------- host side
static inline unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }

MGPCG::~MGPCG() {
    free_all();
    if (bbox_ev) hipEventDestroy(bbox_ev);
}

void MGPCG::free_all() {
    clear_graphs();
    if (owns_memory) {
        for (size_t l = 0; l < lv.size(); ++l) {
            MGLevel& L = lv[l];
            hipFree(L.code); hipFree(L.type);
            if (l > 0) { hipFree(L.b); hipFree(L.x); }
            hipFree(L.xt); hipFree(L.r);
        }
        hipFree(r); hipFree(z); hipFree(pa); hipFree(pb); hipFree(q); hipFree(bb); hipFree(xb);
        hipFree(partial); hipFree(sc); hipFree(counters);
        if (sc_host) hipHostFree(sc_host);
        if (host_counters) hipHostFree(host_counters);
    }
    lv.clear();
    nlev = 0;
    r = z = pa = pb = q = bb = nullptr;
    xb = nullptr;
    if (fw) hipFree(fw);
    fw = nullptr;
    partial = nullptr;
    sc = nullptr;
    sc_host = nullptr;
    counters = host_counters = nullptr;
}

// Level dimensions for a W x H x D box: halve (rounding up) until a level has at
// most coarse_max cells.
static int level_dims(int W, int H, int D, long long coarse_max, int dims[][3], int cap) {
    int n = 0, w = W, h = H, d = D;
    while (n < cap) {
        dims[n][0] = w; dims[n][1] = h; dims[n][2] = d;
        ++n;
        if ((long long)w * h * d <= coarse_max || (w == 1 && h == 1 && d == 1)) break;
        w = (w + 1) / 2; h = (h + 1) / 2; d = (d + 1) / 2;
    }
    return n;
}

int MGPCG::init(int W_, int H_, int D_) { return init_impl(W_, H_, D_, nullptr, 0, nullptr, nullptr, nullptr); }

int MGPCG::init_in(int W_, int H_, int D_, void* base, size_t bytes, SolveScalars* sc_host_buf,
                   int* host_counters_buf, size_t* used) {
    if (!base || !sc_host_buf || !host_counters_buf) return -1;
    return init_impl(W_, H_, D_, base, bytes, sc_host_buf, host_counters_buf, used);
}

// Device buffers from hipMalloc (base == nullptr), or carved in order from
// [base, base + cap) at 256-byte alignment with the host mirrors given by the
// caller (nothing allocated; *used = the bytes the carving needs, also when
// they exceed cap, in which case nothing is carved and -2 is returned).
int MGPCG::init_impl(int W_, int H_, int D_, void* base, size_t cap, SolveScalars* sc_host_buf,
                     int* host_counters_buf, size_t* used) {
    free_all();
    owns_memory = base == nullptr;
    coarse_max = std::min<long long>(coarse_max, COARSE_MAX);
    W = W_; H = H_; D = D_;
    N = (long long)W * H * D;
    int dims[40][3];
    int n = level_dims(W, H, D, coarse_max, dims, 40);
    bool ok = true;
    size_t off = 0;
    if (base) {
        // the size of the carving first, so that a workspace that is too small
        // leaves every pointer null
        size_t need = 0;
        auto add = [&](size_t b) { need = ((need + 255) & ~(size_t)255) + b; };
        for (int l = 0; l < n; ++l) {
            const size_t ln = (size_t)dims[l][0] * dims[l][1] * dims[l][2];
            add(sizeof(uint16_t) * ln);
            add(ln);
            if (l > 0) { add(sizeof(float) * ln); add(sizeof(float) * ln); }
            add(sizeof(float) * ln);
            add(sizeof(float) * ln);
        }
        for (int v = 0; v < 6; ++v) add(sizeof(float) * (size_t)N);
        add(sizeof(double) * (size_t)N);
        add(sizeof(double) * RED_BLOCKS * PS_COUNT);
        add(sizeof(SolveScalars));
        add(sizeof(int) * (COUNTERS_BBOX + 6 * BBOX_BLOCKS));
        if (used) *used = need;
        if (need > cap) return -2;
    }
    auto A = [&](void** p, size_t bytes) {
        if (base) {
            off = (off + 255) & ~(size_t)255;
            *p = (char*)base + off;
            off += bytes;
        } else if (hipMalloc(p, bytes) != hipSuccess) {
            ok = false;
        }
    };
    lv.resize(n);
    for (int l = 0; l < n; ++l) {
        MGLevel& L = lv[l];
        L = MGLevel{};
        L.W = dims[l][0]; L.H = dims[l][1]; L.D = dims[l][2];
        L.n = (long long)L.W * L.H * L.D;
        A((void**)&L.code, sizeof(uint16_t) * L.n);
        A((void**)&L.type, L.n);
        if (l > 0) {
            A((void**)&L.b, sizeof(float) * L.n);
            A((void**)&L.x, sizeof(float) * L.n);
        }
        A((void**)&L.xt, sizeof(float) * L.n);
        A((void**)&L.r, sizeof(float) * L.n);
    }
    for (float** p : {&r, &z, &pa, &pb, &q, &bb}) A((void**)p, sizeof(float) * N);
    A((void**)&xb, sizeof(double) * N);
    A((void**)&partial, sizeof(double) * RED_BLOCKS * PS_COUNT);
    A((void**)&sc, sizeof(SolveScalars));
    A((void**)&counters, sizeof(int) * (COUNTERS_BBOX + 6 * BBOX_BLOCKS));
    if (base) {
        sc_host = sc_host_buf;
        host_counters = host_counters_buf;
    } else {
        if (hipHostMalloc((void**)&sc_host, sizeof(SolveScalars)) != hipSuccess) ok = false;
        if (hipHostMalloc((void**)&host_counters, sizeof(int) * (COUNTERS_BBOX + 6 * BBOX_BLOCKS)) != hipSuccess)
            ok = false;
    }
    if (!ok) { free_all(); return -1; }
    // CG direction buffers start finite: the first update multiplies them by beta = 0
    // (carved memory holds the previous call's contents: cleared the same way)
    hipMemset(pa, 0, sizeof(float) * N);
    hipMemset(pb, 0, sizeof(float) * N);
    hipMemset(partial, 0, sizeof(double) * RED_BLOCKS * PS_COUNT);
    configure(W, H, D);
    return 0;
}

void MGPCG::configure(int Wb, int Hb, int Db) {
    int dims[40][3];
    int n = level_dims(Wb, Hb, Db, coarse_max, dims, (int)lv.size());
    for (int l = 0; l < n; ++l) {
        lv[l].W = dims[l][0]; lv[l].H = dims[l][1]; lv[l].D = dims[l][2];
        lv[l].n = (long long)lv[l].W * lv[l].H * lv[l].D;
    }
    nlev = n;
    n0 = lv[0].n;
    nred = (int)std::min<long long>(RED_BLOCKS, (n0 + 255) / 256);
    vec0 = use_vec && (lv[0].W % 4 == 0);
}

double* MGPCG::slot(int s) { return partial + (size_t)s * RED_BLOCKS; }

void MGPCG::setup(const uint8_t* labels, const float* b, hipStream_t st) {
    // level 0: types, codes, b and the partial sums of the b norm in one pass
    MGLevel& L0 = lv[0];
    const bool weighted = face.active && fw;
    L0.wd = weighted ? fw : nullptr;
    L0.wx = weighted ? fw + N : nullptr;
    L0.wy = weighted ? fw + 2 * N : nullptr;
    L0.wz = weighted ? fw + 3 * N : nullptr;
    if (weighted)
        hipLaunchKernelGGL(k_setup0<true>, dim3(nred), dim3(256), 0, st, labels, b, L0, W, H, box[0], box[1], box[2],
                           bb, slot(PS_B2), slot(PS_B1), counters + 8, face);
    else
        hipLaunchKernelGGL(k_setup0<false>, dim3(nred), dim3(256), 0, st, labels, b, L0, W, H, box[0], box[1],
                           box[2], bb, slot(PS_B2), slot(PS_B1), counters + 8, face);
    // level 1's types, then per level its codes and the next level's types
    if (nlev > 1) hipLaunchKernelGGL(k_type_coarsen, dim3(nblk(lv[1].n)), dim3(256), 0, st, lv[0], lv[1]);
    for (int l = 1; l < nlev; ++l) {
        const dim3 g((unsigned)std::min<long long>(1024, nblk(lv[l].n)));
        if (l + 1 < nlev)
            hipLaunchKernelGGL(k_code<true>, g, dim3(256), 0, st, lv[l], lv[l + 1], counters + 10);
        else
            hipLaunchKernelGGL(k_code<false>, g, dim3(256), 0, st, lv[l], lv[l], counters + 10);
    }
}

MGLevel MGPCG::fine() const { return lv[0]; }



float* MGPCG::vcycle_level(int l, hipStream_t st, double* pdot, SolveScalars* s) {
    MGLevel& L = lv[l];
    // small levels: one workgroup (a fine level that small, as in a thin slab,
    // too; its r.z partial sums then come from a separate pass)
    if (use_tail && l < nlev - 1 && L.n <= TAIL_MAX && nlev - l <= TAIL_LEVELS && !L.wd) {
        TailLevels t{};
        t.n = nlev - l;
        for (int q = 0; q < SMOOTH_STEPS; ++q) t.w[q] = sweep_omega[q];
        long long off = (2 * L.n + 3) & ~3LL;
        for (int q = 0; q < t.n; ++q) {
            t.lv[q] = lv[l + q];
            if (q == 0) continue;
            t.off[q] = (int)off;
            off += tail_level_floats(lv[l + q].n);
        }
        if (off <= TAIL_LDS_FLOATS) {
            hipLaunchKernelGGL(k_tail, dim3(1), dim3(TAIL_THREADS), 0, st, t, std::max(nu_pre, 1),
                               std::max(nu_post, 1), coarse_sweeps, omega, (const SolveScalars*)s);
            if (l == 0 && pdot)
                hipLaunchKernelGGL(k_dot_rz, dim3(nred), dim3(256), 0, st, L.n, L.b, L.x, pdot, s);
            return L.x;
        }
    }
    if (l == nlev - 1) {
        hipLaunchKernelGGL(k_coarse_solve, dim3(1), dim3(1024), 0, st, L, L.b, L.x, coarse_sweeps, omega, s);
        if (l == 0 && pdot)   // single-level hierarchy: r.z of the coarse result
            hipLaunchKernelGGL(k_dot_rz, dim3(nred), dim3(256), 0, st, L.n, L.b, L.x, pdot, s);
        return L.x;
    }
    // small coarse levels: pre-smoothing + residual and prolongation +
    // post-smoothing as one temporally blocked kernel each
    if (use_tb && l > 0 && L.n <= TB_MAX && nu_pre == SMOOTH_STEPS && nu_post == SMOOTH_STEPS) {
        MGLevel& Cn = lv[l + 1];
        const unsigned ntb = (unsigned)(((L.W + TB_TX - 1) / TB_TX) * ((L.H + TB_TY - 1) / TB_TY) *
                                        ((L.D + TB_TZ - 1) / TB_TZ));
        SweepSteps ss;
        for (int q = 0; q < SMOOTH_STEPS; ++q) ss.w[q] = sweep_omega[q];
        hipLaunchKernelGGL((k_smooth_tb<SMOOTH_STEPS, true>), dim3(ntb), dim3(TB_THREADS), 0, st, L, (const float*)L.b,
                           (const float*)nullptr, L.x, L.r, Cn, (const float*)nullptr, ss, (const SolveScalars*)s);
        hipLaunchKernelGGL(k_restrict, dim3(nblk(4 * Cn.n)), dim3(256), 0, st, L, Cn, (const float*)L.r, Cn.b,
                           (const SolveScalars*)s);
        const float* xc = vcycle_level(l + 1, st, nullptr, s);
        hipLaunchKernelGGL((k_smooth_tb<SMOOTH_STEPS, false>), dim3(ntb), dim3(TB_THREADS), 0, st, L, (const float*)L.b,
                           (const float*)L.x, L.xt, (float*)nullptr, Cn, xc, ss, (const SolveScalars*)s);
        return L.xt;
    }
    const unsigned nb = nblk(L.n);
    const unsigned ng = (l == 0) ? (unsigned)nred : nb;   // grid of the r.z-accumulating sweep
    // a fine level whose sweeps are bandwidth-bound passes (above the fused
    // transfer kernels' size): its own sweep count and polynomial
    const bool fine = l == 0 && L.n > FUSE_MAX && nu_pre == SMOOTH_STEPS && nu_post == SMOOTH_STEPS;
    const float* om = fine ? fine_omega : sweep_omega;
    const int msteps = fine ? NU_FINE : SMOOTH_STEPS;
    // Buffer writes after the first: the extra pre sweeps plus the post sweeps;
    // the first buffer is chosen so that the last sweep lands in L.x.
    const int npre = fine ? NU_FINE : std::max(nu_pre, 1), npost = fine ? NU_FINE : std::max(nu_post, 1);
    // the fused prolongation+sweep writes the other buffer; the plain
    // prolongation adds in place, so either way npost sweeps follow the pre-smoothing
    const int extra = (npre >= 2 ? npre - 2 : 0) + npost;
    float* cur = (extra & 1) ? L.xt : L.x;
    float* oth = (cur == L.x) ? L.xt : L.x;
    // 4 cells per thread on every level this path handles whose rows are a
    // multiple of 4 cells (the fine level when the box allows, the next ones
    // above the blocked-smoothing size)
    const bool v4 = use_vec && L.W % 4 == 0;
    const unsigned nb4 = nblk(L.n / 4);
    if (v4)
        hipLaunchKernelGGL(k_jacobi2_v4, dim3(nb4), dim3(256), 0, st, L, L.b, cur, om[0],
                           npre >= 2 ? om[1 % msteps] : 0.0f, s);
    else
        hipLaunchKernelGGL(k_jacobi2, dim3(nb), dim3(256), 0, st, L, L.b, cur, om[0],
                           npre >= 2 ? om[1 % msteps] : 0.0f, s);
    for (int k = 2; k < npre; ++k) {
        if (v4)
            hipLaunchKernelGGL(k_jacobi_v4<false>, dim3(nb4), dim3(256), 0, st, L, L.b, cur, oth, om[k % msteps],
                               (double*)nullptr, s);
        else
            hipLaunchKernelGGL(k_jacobi<false>, dim3(nb), dim3(256), 0, st, L, L.b, cur, oth, om[k % msteps],
                               (double*)nullptr, s);
        std::swap(cur, oth);
    }
    MGLevel& Cn = lv[l + 1];
    const bool fuse = L.n <= FUSE_MAX;
    // (a restriction that recomputes the fine residuals is slower at every level
    // size: one long serial thread per coarse cell)
    if (v4)
        hipLaunchKernelGGL(k_residual_v4, dim3(nb4), dim3(256), 0, st, L, (const float*)L.b, (const float*)cur, L.r,
                           (const SolveScalars*)s);
    else
        hipLaunchKernelGGL(k_residual, dim3(nb), dim3(256), 0, st, L, L.b, cur, L.r, s);
    // two coarse cells per thread when the coarse level is large; four lanes
    // per coarse cell (more threads in flight) when it is small (same bits)
    if (v4 && Cn.n > RESTRICT_V2_MIN)
        hipLaunchKernelGGL(k_restrict_v2, dim3(nblk(Cn.n / 2)), dim3(256), 0, st, L, Cn, (const float*)L.r, Cn.b,
                           (const SolveScalars*)s);
    else
        hipLaunchKernelGGL(k_restrict, dim3(nblk(4 * Cn.n)), dim3(256), 0, st, L, Cn, (const float*)L.r, Cn.b,
                           (const SolveScalars*)s);
    float* xc = vcycle_level(l + 1, st, nullptr, s);
    int k0;
    if (fuse) {
        if (pdot && npost == 1)
            hipLaunchKernelGGL(k_prolong_smooth<true>, dim3(ng), dim3(256), 0, st, L, Cn, L.b, cur, xc, oth, om[0],
                               pdot, s);
        else
            hipLaunchKernelGGL(k_prolong_smooth<false>, dim3(nb), dim3(256), 0, st, L, Cn, L.b, cur, xc, oth, om[0],
                               (double*)nullptr, s);
        std::swap(cur, oth);
        k0 = 1;
    } else {
        if (v4)
            hipLaunchKernelGGL(k_prolong_v4, dim3(nb4), dim3(256), 0, st, L, Cn, (const float*)xc, cur,
                               (const SolveScalars*)s);
        else
            hipLaunchKernelGGL(k_prolong, dim3(nb), dim3(256), 0, st, L, Cn, xc, cur, s);
        k0 = 0;
    }
    for (int k = k0; k < npost; ++k) {
        const bool dot = k == npost - 1 && pdot;
        if (v4) {
            if (dot)
                hipLaunchKernelGGL(k_jacobi_v4<true>, dim3(ng), dim3(256), 0, st, L, L.b, cur, oth, om[k % msteps],
                                   pdot, s);
            else
                hipLaunchKernelGGL(k_jacobi_v4<false>, dim3(nb4), dim3(256), 0, st, L, L.b, cur, oth,
                                   om[k % msteps], (double*)nullptr, s);
        } else if (dot) {
            hipLaunchKernelGGL(k_jacobi<true>, dim3(ng), dim3(256), 0, st, L, L.b, cur, oth, om[k % msteps], pdot,
                               s);
        } else {
            hipLaunchKernelGGL(k_jacobi<false>, dim3(nb), dim3(256), 0, st, L, L.b, cur, oth, om[k % msteps],
                               (double*)nullptr, s);
        }
        std::swap(cur, oth);
    }
    return cur;
}

void MGPCG::vcycle(const float* rin, float* zout, double* pdot, SolveScalars* s, hipStream_t st) {
    MGLevel& L0 = lv[0];
    L0.b = const_cast<float*>(rin);
    L0.x = zout;
    float* res = vcycle_level(0, st, pdot, s);
    if (res != zout) hipMemcpyAsync(zout, res, sizeof(float) * L0.n, hipMemcpyDeviceToDevice, st);
}

void MGPCG::apply_A(const float* x, float* y, hipStream_t st) {
    MGLevel& L = lv[0];
    hipLaunchKernelGGL(k_residual, dim3(nblk(L.n)), dim3(256), 0, st, L, (const float*)nullptr, x, y,
                       (const SolveScalars*)nullptr);
}

void MGPCG::dot_rz(const float* rr, const float* zz, double* pdot, SolveScalars* s, hipStream_t st) {
    hipLaunchKernelGGL(k_dot_rz, dim3(nred), dim3(256), 0, st, n0, rr, zz, pdot, s);
}

void MGPCG::dot_rz_zq(const float* rr, const float* zz, double* pdot, SolveScalars* s, hipStream_t st) {
    hipLaunchKernelGGL(k_dot_rz_zq, dim3(nred), dim3(256), 0, st, n0, rr, zz, (const float*)q, pdot, slot(PS_ZQ), s);
}

// ---------------------------------------------------------------- hybrid preconditioner pieces
__global__ __launch_bounds__(256) void k_resid_sq(MGLevel L, const float* __restrict__ r, const float* __restrict__ z,
                                                  float* __restrict__ r1, double* __restrict__ p_rr,
                                                  const SolveScalars* __restrict__ sc) {
    if (sc && sc->done) return;
    __shared__ double sm[12];
    const long long WH = (long long)L.W * L.H;
    double acc = 0.0;
    for (long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; c < L.n;
         c += (long long)gridDim.x * blockDim.x) {
        const uint32_t code = L.code[c];
        const float v = (code & SELF) ? r[c] - stencil_Ax_L(L, z, c, code, WH) : 0.0f;
        r1[c] = v;
        acc += (double)v * (double)v;
    }
    store_partial(acc, sm, p_rr);
}

__global__ __launch_bounds__(256) void k_partial_rms(const double* __restrict__ p_rr, int np, float* __restrict__ scale,
                                                     const SolveScalars* __restrict__ sc) {
    if (sc->done) return;
    __shared__ double sm[12];
    double o[1];
    sum_partials_k<1>({p_rr}, np, sm, o);
    if (threadIdx.x != 0) return;
    const double v = o[0] / (double)(sc->n_unknown > 0 ? sc->n_unknown : 1);
    float s = v > 0.0 ? (float)sqrt(v) : 1.0f;
    if (!(s > 0.0f) || !isfinite(s)) s = 1.0f;
    scale[0] = s;
}

__global__ __launch_bounds__(256) void k_line_dots(MGLevel L, const float* __restrict__ z2, const float* __restrict__ r1,
                                                   double* __restrict__ p_num, double* __restrict__ p_den,
                                                   const SolveScalars* __restrict__ sc) {
    if (sc->done) return;
    __shared__ double sm[12];
    const long long WH = (long long)L.W * L.H;
    double a = 0.0, b = 0.0;
    for (long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; c < L.n;
         c += (long long)gridDim.x * blockDim.x) {
        const uint32_t code = L.code[c];
        if (!(code & SELF)) continue;
        const double zc = z2[c];
        a += zc * (double)r1[c];
        b += zc * (double)stencil_Ax_L(L, z2, c, code, WH);
    }
    store_partial(a, sm, p_num);
    store_partial(b, sm, p_den);
}

__global__ __launch_bounds__(256) void k_add_scaled(long long n, float* __restrict__ z, const float* __restrict__ z2,
                                                    const double* __restrict__ p_num, const double* __restrict__ p_den,
                                                    int np, const SolveScalars* __restrict__ sc) {
    if (sc->done) return;
    __shared__ double sm[12];
    double o[2];
    sum_partials_k<2>({p_num, p_den}, np, sm, o);
    const float cf = o[1] > 0.0 ? (float)(o[0] / o[1]) : 0.0f;
    for (long long c = (long long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; c < n;
         c += (long long)gridDim.x * blockDim.x)
        z[c] = fmaf(cf, z2[c], z[c]);
}

__global__ __launch_bounds__(256) void k_box_to_full(long long n, int W, int H, MGLevel L, int ox, int oy, int oz,
                                                     const float* __restrict__ zb, float* __restrict__ z) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    int i, j, k;
    split3(c, W, H, i, j, k);
    const int bi = i - ox, bj = j - oy, bk = k - oz;
    float v = 0.0f;
    if (bi >= 0 && bi < L.W && bj >= 0 && bj < L.H && bk >= 0 && bk < L.D) {
        const long long q = (long long)bi + (long long)bj * L.W + (long long)bk * L.W * L.H;
        if (L.code[q] & SELF) v = zb[q];
    }
    z[c] = v;
}

void MGPCG::resid_sq(const float* rr, const float* zz, float* r1, double* p_rr, SolveScalars* s, hipStream_t st) {
    hipLaunchKernelGGL(k_resid_sq, dim3(nred), dim3(256), 0, st, lv[0], rr, zz, r1, p_rr, (const SolveScalars*)s);
}

void MGPCG::partial_rms(const double* p_rr, float* scale, SolveScalars* s, hipStream_t st) {
    hipLaunchKernelGGL(k_partial_rms, dim3(1), dim3(256), 0, st, p_rr, nred, scale, (const SolveScalars*)s);
}

void MGPCG::line_dots(const float* z2, const float* r1, double* p_num, double* p_den, SolveScalars* s,
                      hipStream_t st) {
    hipLaunchKernelGGL(k_line_dots, dim3(nred), dim3(256), 0, st, lv[0], z2, r1, p_num, p_den, (const SolveScalars*)s);
}

void MGPCG::add_scaled(float* zz, const float* z2, const double* p_num, const double* p_den, SolveScalars* s,
                       hipStream_t st) {
    hipLaunchKernelGGL(k_add_scaled, dim3(nred), dim3(256), 0, st, n0, zz, z2, p_num, p_den, nred,
                       (const SolveScalars*)s);
}

int MGPCG::vcycle_system(const uint8_t* labels, const float* b, float* zfull, hipStream_t st) {
    face = fsd::FaceFrac{};
    int rc = prepare(labels, b, st);
    if (rc) {
        if (rc == -3) hipMemsetAsync(zfull, 0, sizeof(float) * (size_t)N, st);
        return rc;
    }
    hipMemsetAsync(sc, 0, sizeof(SolveScalars), st);
    vcycle(bb, z, slot(PS_RES), sc, st);
    hipLaunchKernelGGL(k_box_to_full, dim3(nblk(N)), dim3(256), 0, st, N, W, H, lv[0], box[0], box[1], box[2],
                       (const float*)z, zfull);
    return hipStreamSynchronize(st) == hipSuccess ? 0 : -5;
}

class JacobiPre : public Preconditioner {
public:
    explicit JacobiPre(MGPCG* s) : s_(s) {}
    void apply(const float* r, float* z, double* partial, SolveScalars* sc, hipStream_t st) override {
        hipLaunchKernelGGL(k_diag, dim3(s_->nred), dim3(256), 0, st, s_->lv[0], r, z, partial, sc);
    }

private:
    MGPCG* s_;
};

Preconditioner* make_jacobi_preconditioner(MGPCG* s) { return new JacobiPre(s); }

// One batch of inner PCG iterations (an even number, so the two direction
// buffers and the two r.z partial slots are in the same roles at the start of
// every batch).
void MGPCG::record_batch(Preconditioner* pre, int nb, hipStream_t st) {
    MGLevel& L0 = lv[0];
    const bool flex = pre->flexible();
    float* pold = pa;
    float* pnew = pb;
    int cur = 0;
    for (int it = 0; it < nb; ++it) {
        double* rz_new = slot(PS_RZ0 + cur);
        double* rz_old = slot(PS_RZ0 + (cur ^ 1));
        double* pq_new = slot(PS_PQ0 + cur);
        double* pq_old = slot(PS_PQ0 + (cur ^ 1));
        pre->apply(r, z, rz_new, sc, st);   // z = M r and the partial sums of r.z (and z.q when flexible)
        if (singular_mode) {
            // z orthogonal to the null space (r.z keeps its partial sums: r is
            // orthogonal to it up to rounding)
            hipLaunchKernelGGL(k_sum_unknowns<float>, dim3(nred), dim3(256), 0, st, L0, (const float*)z,
                               slot(PS_MEAN), (const SolveScalars*)sc);
            hipLaunchKernelGGL(k_sub_mean<float>, dim3(nred), dim3(256), 0, st, L0, z, (const double*)slot(PS_MEAN),
                               nred, (int)n_unk, (const SolveScalars*)sc);
        }
        const double* zq = flex ? (const double*)slot(PS_ZQ) : (const double*)nullptr;
        double* prp = flex ? slot(PS_RP) : (double*)nullptr;
        const double* num = flex ? (const double*)slot(PS_RP) : (const double*)rz_new;
        if (vec0) {
            hipLaunchKernelGGL(k_pupdate_spmv_v4, dim3(nred), dim3(256), 0, st, L0, (const float*)z,
                               (const float*)pold, pnew, q, (const double*)rz_new, (const double*)rz_old, zq,
                               (const double*)pq_old, nred, pq_new, (const float*)r, prp, (const SolveScalars*)sc);
            hipLaunchKernelGGL(k_update_v4, dim3(nred), dim3(256), 0, st, n0, (const float*)pnew, (const float*)q, xb,
                               r, num, (const double*)pq_new, nred, slot(PS_RR), (const SolveScalars*)sc);
        } else {
            hipLaunchKernelGGL(k_pupdate_spmv, dim3(nred), dim3(256), 0, st, L0, z, pold, pnew, q, rz_new, rz_old, zq,
                               (const double*)pq_old, nred, pq_new, (const float*)r, prp, sc);
            hipLaunchKernelGGL(k_update, dim3(nred), dim3(256), 0, st, n0, pnew, q, xb, r, num, pq_new, nred,
                               slot(PS_RR), sc);
        }
        if (singular_mode) {
            // the FP32 residual update leaves a rounding component along the
            // null space that no iteration can remove: project it out before
            // the convergence test
            hipLaunchKernelGGL(k_sum_unknowns<float>, dim3(nred), dim3(256), 0, st, L0, (const float*)r,
                               slot(PS_MEAN), (const SolveScalars*)sc);
            hipLaunchKernelGGL(k_sub_mean_rr, dim3(nred), dim3(256), 0, st, L0, r, (const double*)slot(PS_MEAN), nred,
                               (int)n_unk, slot(PS_RR), (const SolveScalars*)sc);
        }
        hipLaunchKernelGGL(k_fin_iter, dim3(1), dim3(256), 0, st, rz_new, pq_new, slot(PS_RR), nred, flex ? 1 : 0,
                           flex ? stall_window : 0, sc);
        std::swap(pold, pnew);
        cur ^= 1;
    }
    // Iterations that converged are checked here against the FP64 residual
    // (done = 2, or a restart from the true residual), saving the host round
    // trip of a separate check.  Not for singular systems (their check
    // shifts b by its mean, a value of the solve).
    if (!singular_mode) {
        hipLaunchKernelGGL(k_resid64<false>, dim3(nred), dim3(256), 0, st, L0, (const float*)bb, xb, 0.0, r,
                           slot(PS_RES), (const SolveScalars*)sc);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, (int)FIN_CHECK, (const double*)slot(PS_RES),
                           (const double*)nullptr, nred, sc, 0.0);
    }
}

// The batch as a HIP graph, cached per solve-box shape and preconditioner
// (kernel arguments depend only on those).  The graph holds kernel nodes only.
hipGraphExec_t MGPCG::batch_graph(Preconditioner* pre, int nb, hipStream_t st) {
    const long long key[10] = {lv[0].W, lv[0].H, lv[0].D, nlev, (long long)(uintptr_t)pre->graph_key(), stall_window,
                               pre->graph_epoch(), nb, singular_mode ? (long long)n_unk : -1, face.active ? 1 : 0};
    for (auto& g : graphs)
        if (std::equal(key, key + 10, g.key)) return g.exec;
    hipGraph_t graph;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
    record_batch(pre, nb, st);
    if (hipStreamEndCapture(st, &graph) != hipSuccess) return nullptr;
    hipGraphExec_t ex;
    if (hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0) != hipSuccess) {
        hipGraphDestroy(graph);
        return nullptr;
    }
    hipGraphDestroy(graph);
    if (graphs.size() >= 32) {
        hipGraphExecDestroy(graphs.front().exec);
        graphs.erase(graphs.begin());
    }
    GraphEntry g;
    std::copy(key, key + 10, g.key);
    g.exec = ex;
    graphs.push_back(g);
    return ex;
}

void MGPCG::clear_graphs() {
    for (auto& g : graphs) hipGraphExecDestroy(g.exec);
    graphs.clear();
}

int MGPCG::launch_bbox(const uint8_t* labels, hipStream_t st) {
    const int nbb = (int)std::min<long long>(BBOX_BLOCKS, (N + 16 * 256 - 1) / (16 * 256));
    int* part = counters + COUNTERS_BBOX;
    if (W % 16 == 0)
        hipLaunchKernelGGL(k_bbox<true>, dim3(nbb), dim3(256), 0, st, labels, W, H, N, part);
    else
        hipLaunchKernelGGL(k_bbox<false>, dim3(nbb), dim3(256), 0, st, labels, W, H, N, part);
    hipMemcpyAsync(host_counters + COUNTERS_BBOX, part, sizeof(int) * 6 * nbb, hipMemcpyDeviceToHost, st);
    return nbb;
}

int MGPCG::queue_bbox(const uint8_t* labels, hipStream_t st) {
    if (!owns_memory) return 0;
    if (!bbox_ev && hipEventCreateWithFlags(&bbox_ev, hipEventDisableTiming) != hipSuccess) {
        bbox_ev = nullptr;
        return -5;
    }
    launch_bbox(labels, st);
    if (hipEventRecord(bbox_ev, st) != hipSuccess) return -5;
    bbox_labels = labels;
    return 0;
}

int MGPCG::prepare(const uint8_t* labels, const float* b, hipStream_t st) {
    // fluid bounding box, expanded by one cell so that every neighbour of a
    // FLUID cell (AIR/SOLID/BODY or out of domain) is inside the solve box;
    // queued earlier by queue_bbox (the host then waits for that copy only),
    // else computed here
    const int nbb = (int)std::min<long long>(BBOX_BLOCKS, (N + 16 * 256 - 1) / (16 * 256));
    if (bbox_labels && bbox_labels == labels) {
        if (hipEventSynchronize(bbox_ev) != hipSuccess) return -5;
    } else {
        launch_bbox(labels, st);
        if (hipStreamSynchronize(st) != hipSuccess) return -5;
    }
    bbox_labels = nullptr;
    int hb[6] = {0x7fffffff, -1, 0x7fffffff, -1, 0x7fffffff, -1};
    for (int blk = 0; blk < nbb; ++blk)
        for (int q = 0; q < 3; ++q) {
            hb[2 * q] = std::min(hb[2 * q], host_counters[COUNTERS_BBOX + 6 * blk + 2 * q]);
            hb[2 * q + 1] = std::max(hb[2 * q + 1], host_counters[COUNTERS_BBOX + 6 * blk + 2 * q + 1]);
        }
    if (hb[1] < 0) return -3;   // no fluid
    int lo[3], hi[3];
    const int dimsz[3] = {W, H, D};
    for (int q = 0; q < 3; ++q) {
        lo[q] = std::max(0, hb[2 * q] - 1);
        hi[q] = std::min(dimsz[q] - 1, hb[2 * q + 1] + 1);
    }
    // widen the box along x to a multiple of 4 cells (aligned 4-cell groups on
    // the fine level) when the grid is wide enough; the extra cells are not
    // unknowns
    if (use_vec) {
        int w = hi[0] - lo[0] + 1;
        const int w4 = (w + 3) & ~3;
        if (w4 <= W) {
            hi[0] = std::min(W - 1, lo[0] + w4 - 1);
            lo[0] = hi[0] - w4 + 1;
        }
    }
    box[0] = lo[0]; box[1] = lo[1]; box[2] = lo[2];
    configure(hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1);
    hipMemsetAsync(counters + 8, 0, sizeof(int) * 4, st);
    setup(labels, b, st);
    return 0;
}

int MGPCG::solve(const uint8_t* labels, const float* b, double* x64, float* p, double tol, int maxit,
                 Preconditioner* pre, hipStream_t st, fs_solve_stats* stats, const double* x0) {
    auto t0 = std::chrono::steady_clock::now();
    MGPreconditioner mgp(this);
    if (!pre) pre = &mgp;
    fs_solve_stats loc{};
    if (face.active && !fw) {
        if (!owns_memory || hipMalloc((void**)&fw, sizeof(float) * 4 * (size_t)N) != hipSuccess) {
            fw = nullptr;
            return -5;
        }
    }
    // r = (b - shift) - A x64 (zero: x64 = 0 first)
    auto launch_resid64 = [&](bool zero, double sh, hipStream_t s) {
        if (zero)
            hipLaunchKernelGGL(k_resid64<true>, dim3(nred), dim3(256), 0, s, lv[0], bb, xb, sh, r, slot(PS_RES),
                               (const SolveScalars*)nullptr);
        else
            hipLaunchKernelGGL(k_resid64<false>, dim3(nred), dim3(256), 0, s, lv[0], bb, xb, sh, r, slot(PS_RES),
                               (const SolveScalars*)nullptr);
    };
    int rc = prepare(labels, b, st);
    if (rc == -5) return -5;
    if (rc == -3) {
        box[0] = box[1] = box[2] = 0;
        configure(W, H, D);
        hipLaunchKernelGGL(k_fill0, dim3(nblk(N)), dim3(256), 0, st, N, xb);
        hipLaunchKernelGGL(k_scatter_out, dim3(nblk((N + 3) / 4)), dim3(256), 0, st, labels, N, W, H, lv[0], 0, 0, 0, xb, x64, p);
        if (stats) *stats = loc;
        return -3;
    }
    if (pre->prepare_box() != 0) return -5;
    MGLevel& L0 = lv[0];
    SolveScalars init{};
    init.maxit = maxit;
    init.inner_rtol2 = inner_rtol * inner_rtol;
    hipMemcpyAsync(sc, &init, sizeof(SolveScalars), hipMemcpyHostToDevice, st);
    // the b norm's partial sums came from the level-0 set-up pass
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, (int)FIN_BNORM, slot(PS_B2), slot(PS_B1), nred, sc, 0.0);
    // The first residual is launched before the host knows whether the system
    // is singular (no Dirichlet cell: then b is shifted by its mean and the
    // warm start's mean removed, and this pass is redone below), with the
    // tolerance relative to the b norm on the device: one host round trip.
    hipMemcpyAsync(&sc->n_unknown, counters + 8, sizeof(int), hipMemcpyDeviceToDevice, st);
    if (x0)
        hipLaunchKernelGGL(k_resid64_x0, dim3(nred), dim3(256), 0, st, L0, bb, x0, W, H, box[0], box[1], box[2], 0.0,
                           r, slot(PS_RES), xb);
    else
        launch_resid64(true, 0.0, st);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, (int)FIN_OUTER_REL, slot(PS_RES),
                       (const double*)nullptr, nred, sc, tol * tol);
    // The first batch of iterations is launched before the host sees the
    // first pass (sized from the previous solve's count): it stands unless
    // the system turns out singular or b zero, and is undone below in those
    // cases.  Not for flexible
    // preconditioners (their fallback needs the host between batches).
    const bool spec = use_graph && !singular_mode && !pre->flexible() && last_iters > 0 && maxit > 0;
    if (spec) {
        const int nbs = std::max(2, std::min(2 * batch, 2 * ((last_iters + 1) / 2)));
        hipGraphExec_t ex = batch_graph(pre, nbs, st);
        if (!ex || hipGraphLaunch(ex, st) != hipSuccess) return -5;
    }
    hipMemcpyAsync(host_counters + 8, counters + 8, sizeof(int) * 2, hipMemcpyDeviceToHost, st);
    hipMemcpyAsync(sc_host, sc, sizeof(SolveScalars), hipMemcpyDeviceToHost, st);
    if (hipStreamSynchronize(st) != hipSuccess) return -5;
    const int nunk = host_counters[8];
    const bool singular = (nunk > 0) && !host_counters[9];
    loc.fluid_cells = nunk;
    loc.singular = singular ? 1 : 0;
    n_unk = nunk;
    singular_mode = singular;
    const double shift = singular ? sc_host->bsum / nunk : 0.0;
    const double bn2 =
        singular ? std::max(0.0, sc_host->bnorm2 - sc_host->bsum * sc_host->bsum / nunk) : sc_host->bnorm2;
    const double tol2b = tol * tol * bn2;
    // the first pass is redone when the system is singular, when b is zero
    // (the solution is then 0, whatever the initial guess), and when the
    // initial guess is no better than 0
    // (after a speculative batch a start worse than 0 stands too: the batch
    // iterated from it, and starting over would cost more than it saves)
    const bool first_done = !singular && bn2 > 0.0 && (spec || !(x0 && sc_host->rr_true >= bn2));
    // the speculative batch stands when the first pass does
    bool spec_pending = spec && first_done;
    if (spec && !spec_pending) {
        // undo it: the passes below restart the iteration (x, r and the
        // scalars are rewritten; the directions are cleared, since beta = 0
        // times a non-finite direction would not clear it; the count restarts)
        hipMemcpyAsync(&sc->iters, &zero_count, sizeof(int), hipMemcpyHostToDevice, st);
        hipMemsetAsync(pa, 0, sizeof(float) * (size_t)n0, st);
        hipMemsetAsync(pb, 0, sizeof(float) * (size_t)n0, st);
    }
    int total_iters = 0, outer = 0;
    bool converged = false;
    bool checked = false;   // the last batch ran the FP64 check (sc_host current)
    double rr_true = 0.0, rr_first = 0.0;
    for (outer = 0;; ++outer) {
        if (outer == 0 && first_done) {
            // the speculative first pass stands
        } else if (checked) {
            // the batch closed with the check (done = 2: converged)
        } else if (outer == 0 && x0 && bn2 > 0.0 && singular) {
            // warm start: the initial guess on the unknowns of the box (0 elsewhere)
            hipLaunchKernelGGL(k_gather_x0, dim3(nblk(n0)), dim3(256), 0, st, L0, x0, W, H, box[0], box[1], box[2],
                               xb);
            if (singular) {
                hipLaunchKernelGGL(k_sum_unknowns<double>, dim3(nred), dim3(256), 0, st, L0, (const double*)xb,
                                   slot(PS_MEAN), (const SolveScalars*)nullptr);
                hipLaunchKernelGGL(k_sub_mean<double>, dim3(nred), dim3(256), 0, st, L0, xb,
                                   (const double*)slot(PS_MEAN), nred, nunk, (const SolveScalars*)nullptr);
            }
            launch_resid64(false, shift, st);
        } else if (outer == 0)
            launch_resid64(true, shift, st);
        else
            launch_resid64(false, shift, st);
        if (!(outer == 0 && first_done) && !checked) {
            hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, (int)FIN_OUTER, slot(PS_RES),
                               (const double*)nullptr, nred, sc, tol2b);
            hipMemcpyAsync(sc_host, sc, sizeof(SolveScalars), hipMemcpyDeviceToHost, st);
            if (hipStreamSynchronize(st) != hipSuccess) return -5;
        }
        if (outer == 0 && x0 && bn2 > 0.0 && singular && sc_host->rr_true >= bn2) {
            // the initial guess is no better than 0: start from 0
            launch_resid64(true, shift, st);
            hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, st, (int)FIN_OUTER, slot(PS_RES),
                               (const double*)nullptr, nred, sc, tol2b);
            hipMemcpyAsync(sc_host, sc, sizeof(SolveScalars), hipMemcpyDeviceToHost, st);
            if (hipStreamSynchronize(st) != hipSuccess) return -5;
        }
        rr_true = sc_host->rr_true;
        // (after the speculative batch, rr_true may already be its final check)
        if (outer == 0) rr_first = spec_pending ? sc_host->rr_start : rr_true;
        total_iters = sc_host->iters;
        if (sc_host->done == 2 || bn2 == 0.0) { converged = true; break; }
        if (outer >= max_outer || total_iters >= maxit) break;
        if (singular) {
            // solution orthogonal to the null space before the next pass
            hipLaunchKernelGGL(k_sum_unknowns<double>, dim3(nred), dim3(256), 0, st, L0, (const double*)xb,
                               slot(PS_MEAN), (const SolveScalars*)nullptr);
            hipLaunchKernelGGL(k_sub_mean<double>, dim3(nred), dim3(256), 0, st, L0, xb,
                               (const double*)slot(PS_MEAN), nred, nunk, (const SolveScalars*)nullptr);
        }
        // Batches of at most `batch` iterations, each sized (an even count, so
        // that every batch starts from the same direction buffer) to the
        // iterations still expected: from the rate measured between two checks
        // of this solve, and for the first batch from the previous solve's
        // rate.  An iteration launched after convergence still costs its
        // launches (a few microseconds each), one more host check tens.
        auto batch_for = [&](double need) {
            if (!(need > 0.0) || !std::isfinite(need)) return batch;
            const int n = 2 * (int)std::ceil(need / 2.0);
            return std::max(2, std::min(batch, n));
        };
        // after the speculative batch (sc_host already current) the rate is
        // measured from the first pass
        double rr_prev = spec_pending ? sc_host->rr_start : sc_host->rr;
        int it_prev = spec_pending ? 0 : sc_host->iters;
        int nb = batch;
        if (rate_per_iter < 0.0 && sc_host->rr > 0.0 && sc_host->target > 0.0)
            nb = batch_for(std::log(sc_host->target / sc_host->rr) / rate_per_iter);
        while (true) {
            if (spec_pending) {
                spec_pending = false;
            } else {
                if (use_graph) {
                    hipGraphExec_t ex = batch_graph(pre, nb, st);
                    if (!ex) return -5;
                    if (hipGraphLaunch(ex, st) != hipSuccess) return -5;
                } else {
                    record_batch(pre, nb, st);
                }
                hipMemcpyAsync(sc_host, sc, sizeof(SolveScalars), hipMemcpyDeviceToHost, st);
                if (hipStreamSynchronize(st) != hipSuccess) return -5;
            }
            if (sc_host->done) break;
            nb = batch;
            const int dit = sc_host->iters - it_prev;
            if (dit > 0 && sc_host->rr > 0.0 && rr_prev > sc_host->rr && sc_host->target > 0.0) {
                const double per_it = std::log(sc_host->rr / rr_prev) / dit;   // < 0
                nb = batch_for(std::log(sc_host->target / sc_host->rr) / per_it);
            }
            rr_prev = sc_host->rr;
            it_prev = sc_host->iters;
        }
        checked = sc_host->done == 2;
        // a flexible preconditioner that stalls or breaks down hands over to the V-cycle
        if (pre->flexible() && (sc_host->done == 5 || sc_host->done == 3)) {
            pre = &mgp;
            loc.fallback = 1;
        }
    }
    if (singular) {
        // the solution orthogonal to the null space (zero mean over the unknowns)
        hipLaunchKernelGGL(k_sum_unknowns<double>, dim3(nred), dim3(256), 0, st, L0, (const double*)xb,
                           slot(PS_MEAN), (const SolveScalars*)nullptr);
        hipLaunchKernelGGL(k_sub_mean<double>, dim3(nred), dim3(256), 0, st, L0, xb, (const double*)slot(PS_MEAN),
                           nred, nunk, (const SolveScalars*)nullptr);
    }
    hipLaunchKernelGGL(k_scatter_out, dim3(nblk((N + 3) / 4)), dim3(256), 0, st, labels, N, W, H, L0, box[0], box[1], box[2], xb,
                       x64, p);
    // mean reduction of ||r||^2 per iteration, for sizing the next solve's first batch
    if (converged && total_iters > 0 && rr_true > 0.0 && rr_first > 0.0 && rr_true < rr_first)
        rate_per_iter = std::log(rr_true / rr_first) / total_iters;
    last_iters = total_iters;
    last_outer = outer + 1;
    loc.iterations = total_iters;
    loc.outer = outer + 1;
    loc.converged = converged ? 1 : 0;
    loc.rel_residual = bn2 > 0 ? std::sqrt(rr_true / bn2) : 0.0;
    loc.solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = loc;
    return converged ? 0 : -4;
}

}  // namespace fsk
