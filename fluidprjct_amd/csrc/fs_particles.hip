// fs_particles.hip -- particle-side kernels of the PIC/FLIP step (gfx950).
//
//   k_seed              seedParticles           FluidSolver3D.cu:213-297
//   k_key_count         MarkFluidCellsFunctor   FluidSolver3D.cu:306-337 (key part)
//   k_scan_*            cell-count scan fused with labelGrid :340-397
//   k_scatter           counting-sort scatter (no reference counterpart)
//   k_p2g_tile          particlesToGrid :607-685 + saveVelocities :399-415 + applyForces :700-761
//   k_g2p_advect_tile   gridToParticles :983-1039 + advectParticles :1334-1357
//
// Particles are SoA (px,py,pz,vx,vy,vz,id), kept sorted by cell index so that
// each P2G tile reads contiguous rows and G2P gathers hit neighbouring faces.
#include <algorithm>
#include "fs_kernels.hpp"

using namespace fsd;

namespace fsk {

// ---------------------------------------------------------------------------
// wave-aggregated per-cell counting: lanes holding the same key (particles are
// sorted, so they are adjacent) issue one atomic per run instead of one each.
FS_DEV int wave_agg_rank(int key, bool act, int* __restrict__ count) {
    const int lane = __lane_id();
    int prev = __shfl_up(key, 1, 64);
    bool head = act && (lane == 0 || prev != key);
    uint64_t hm = __ballot(head);
    uint64_t am = __ballot(act);
    int nact = __popcll(am);  // active lanes form a prefix of the wave
    int base = 0, h = 0;
    if (act) {
        uint64_t below = (lane == 63) ? ~0ull : ((1ull << (lane + 1)) - 1ull);
        h = 63 - __clzll(hm & below);
        uint64_t above = hm & ~below;
        int next = above ? (__ffsll((unsigned long long)above) - 1) : 64;
        int end = next < nact ? next : nact;
        if (head) base = atomicAdd(&count[key], end - lane);
    }
    base = __shfl(base, h, 64);
    return base + (lane - h);
}

__global__ __launch_bounds__(256) void k_key_count(const float* __restrict__ px, const float* __restrict__ py,
                                                   const float* __restrict__ pz, int n, float dx, int W, int H,
                                                   int D, int* __restrict__ key, int* __restrict__ rank,
                                                   int* __restrict__ count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool act = i < n;
    int k = act ? cell_key(px[i], py[i], pz[i], dx, W, H, D) : -1;
    int r = wave_agg_rank(k, act, count);
    if (act) {
        key[i] = k;
        rank[i] = r;
    }
}

// ---------------------------------------------------------------------------
// Exclusive scan over cells, 4 cells per thread (int4), 256 threads per block.
constexpr int SCAN_T = 256;
constexpr int SCAN_PER_BLOCK = SCAN_T * 4;

FS_DEV int block_excl_scan(int v, int* sm, int* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) sm[w] = x;
    __syncthreads();
    int wpre = 0, tot = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
        int s = sm[q];
        if (q < w) wpre += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return wpre + x - v;
}

// mode 0: values are cell counts; mode 1: values are (label == FLUID)
template <int MODE>
FS_DEV int4 scan_load(const int* __restrict__ cnt, const uint8_t* __restrict__ lab, long long c4, long long C) {
    int4 v = make_int4(0, 0, 0, 0);
    if (MODE != 1) {
        if (c4 + 3 < C) {
            v = *reinterpret_cast<const int4*>(cnt + c4);
        } else {
            if (c4 < C) v.x = cnt[c4];
            if (c4 + 1 < C) v.y = cnt[c4 + 1];
            if (c4 + 2 < C) v.z = cnt[c4 + 2];
        }
    } else {
        if (c4 < C) v.x = lab[c4] == FS_FLUID;
        if (c4 + 1 < C) v.y = lab[c4 + 1] == FS_FLUID;
        if (c4 + 2 < C) v.z = lab[c4 + 2] == FS_FLUID;
        if (c4 + 3 < C) v.w = lab[c4 + 3] == FS_FLUID;
    }
    return v;
}

template <int MODE>
__global__ __launch_bounds__(SCAN_T) void k_scan_reduce(const int* __restrict__ cnt, const uint8_t* __restrict__ lab,
                                                        long long C, int* __restrict__ bsum) {
    __shared__ int sm[SCAN_T / 64];
    long long c4 = (long long)blockIdx.x * SCAN_PER_BLOCK + threadIdx.x * 4;
    int4 v = scan_load<MODE>(cnt, lab, c4, C);
    int s = v.x + v.y + v.z + v.w;
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int q = 0; q < SCAN_T / 64; ++q) t += sm[q];
        bsum[blockIdx.x] = t;
    }
}

// single block: exclusive scan of nb block sums in place; total -> *total_out
// Exclusive scan of the nb block sums in place (one 1024-thread block): chunks
// of SCAN_TOP_CHUNK sums staged in LDS with coalesced loads, each thread
// scanning SCAN_TOP_PER consecutive ones, the chunk totals carried over.
constexpr int SCAN_TOP_PER = 8;
constexpr int SCAN_TOP_CHUNK = 1024 * SCAN_TOP_PER;
__global__ __launch_bounds__(1024) void k_scan_top(int* __restrict__ bsum, int nb, int* __restrict__ total_out) {
    __shared__ int sm[16];
    __shared__ int chunk[SCAN_TOP_CHUNK];
    int carry = 0;
    for (int base = 0; base < nb; base += SCAN_TOP_CHUNK) {
        const int m = min(SCAN_TOP_CHUNK, nb - base);
        for (int q = threadIdx.x; q < m; q += 1024) chunk[q] = bsum[base + q];
        __syncthreads();
        const int b0 = threadIdx.x * SCAN_TOP_PER;
        int v[SCAN_TOP_PER];
        int local = 0;
#pragma unroll
        for (int e = 0; e < SCAN_TOP_PER; ++e) {
            v[e] = b0 + e < m ? chunk[b0 + e] : 0;
            local += v[e];
        }
        int tot;
        int run = carry + block_excl_scan(local, sm, &tot);
#pragma unroll
        for (int e = 0; e < SCAN_TOP_PER; ++e) {
            if (b0 + e < m) chunk[b0 + e] = run;
            run += v[e];
        }
        __syncthreads();
        for (int q = threadIdx.x; q < m; q += 1024) bsum[base + q] = chunk[q];
        __syncthreads();
        carry += tot;
    }
    if (threadIdx.x == 0) *total_out = carry;
}

// Final pass.  MODE 0 (cell counts): writes start[], labels (labelGrid :340-397
// order: SOLID kept, BODY marked, then FLUID where particles are) and clears
// the count for the next step.
// MODE 1 (fluid flags for seeding): writes start[] only.
// MODE 2 (cell counts, re-sort without relabelling): start[] and cleared counts.
template <int MODE>
__global__ __launch_bounds__(SCAN_T) void k_scan_final(int* __restrict__ cnt, const uint8_t* __restrict__ solid,
                                                       uint8_t* __restrict__ labels, long long C,
                                                       const int* __restrict__ bsum, int* __restrict__ start,
                                                       LabelArgs la) {
    __shared__ int sm[SCAN_T / 64];
    long long c4 = (long long)blockIdx.x * SCAN_PER_BLOCK + threadIdx.x * 4;
    int4 v = scan_load<MODE>(cnt, labels, c4, C);
    int s = v.x + v.y + v.z + v.w;
    int tot;
    int pre = block_excl_scan(s, sm, &tot) + bsum[blockIdx.x];
    int e[4] = {pre, pre + v.x, pre + v.x + v.y, pre + v.x + v.y + v.z};
    int vv[4] = {v.x, v.y, v.z, v.w};
    for (int q = 0; q < 4; ++q) {
        long long c = c4 + q;
        if (c >= C) break;
        start[c] = e[q];
        if (MODE == 0) {
            uint8_t L;
            if (solid[c]) {
                L = FS_SOLID;
            } else {
                L = FS_AIR;
                if (la.body.active) {
                    int i, j, k;
                    split3(c, la.W, la.H, i, j, k);
                    f3 cc = mk((i + 0.5f) * la.dx, (j + 0.5f) * la.dx, (k + 0.5f) * la.dx);
                    if (body_sphere_has(la.body, cc) && body_contains(la.body, la.sdf, cc)) L = FS_BODY;
                }
                if (L == FS_AIR && vv[q] > 0) L = FS_FLUID;
            }
            labels[c] = L;
        }
    }
    if (MODE != 1) {
        if (c4 + 3 < C) {
            *reinterpret_cast<int4*>(cnt + c4) = make_int4(0, 0, 0, 0);
        } else {
            for (int q = 0; q < 4; ++q)
                if (c4 + q < C) cnt[c4 + q] = 0;
        }
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) start[C] = bsum[blockIdx.x] + tot;
}

// ---------------------------------------------------------------------------
// dkey (optional): the cell of each destination slot, for k_canonical
__global__ __launch_bounds__(256) void k_scatter(ParticlesView src, ParticlesView dst, int n,
                                                 const int* __restrict__ key, const int* __restrict__ rank,
                                                 const int* __restrict__ start, int* __restrict__ dkey) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = key[i];
    int d = start[k] + rank[i];
    if (dkey) dkey[d] = k;
    dst.px[d] = src.px[i];
    dst.py[d] = src.py[i];
    dst.pz[d] = src.pz[i];
    dst.vx[d] = src.vx[i];
    dst.vy[d] = src.vy[i];
    dst.vz[d] = src.vz[i];
    dst.id[d] = src.id[i];
}

// Canonical order within each cell: the atomic ranks of the counting pass
// place a cell's particles in an order that can change from run to run; this
// pass moves each particle of the cell-sorted src to its rank by original id
// within its cell's segment of dst, so that the particle order - and with it
// every floating-point sum over a cell's particles - depends on the state
// only.  One thread per slot; a cell's ids are read from the cache.
__global__ __launch_bounds__(256) void k_canonical(ParticlesView src, ParticlesView dst, int n,
                                                   const int* __restrict__ dkey, const int* __restrict__ start) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n) return;
    const int k = dkey[d];
    const int s = start[k], e = start[k + 1];
    const int id = src.id[d];
    int r = 0;
    for (int q = s; q < e; ++q) r += src.id[q] < id;
    const int t = s + r;
    dst.px[t] = src.px[d];
    dst.py[t] = src.py[d];
    dst.pz[t] = src.pz[d];
    dst.vx[t] = src.vx[d];
    dst.vy[t] = src.vy[d];
    dst.vz[t] = src.vz[d];
    dst.id[t] = id;
}

// ---------------------------------------------------------------------------
// Counter-based seeding: cell c (FLUID) with fluid rank r gets particles
// r*ppc .. r*ppc+ppc-1, each from draws (seed, c, slot, d).
__global__ __launch_bounds__(256) void k_seed(const uint8_t* __restrict__ labels, const int* __restrict__ frank,
                                              long long C, int W, int H, float dx, int ppc, uint64_t seed,
                                              ParticlesView P) {
    long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C || labels[c] != FS_FLUID) return;
    int i, j, k;
    split3(c, W, H, i, j, k);
    const double lo = (double)(-0.24f), hi = (double)(0.24f);
    f3 cc = mk(((float)i + 0.5f) * dx, ((float)j + 0.5f) * dx, ((float)k + 0.5f) * dx);
    const float off = 0.25f * dx;
    const uint64_t key = splitmix64(seed);
    long long base = (long long)frank[c] * ppc;
    for (int q = 0; q < ppc; ++q) {
        uint64_t ctr = ((uint64_t)c << 8) | ((uint64_t)q << 2);
        int oct = (int)(splitmix64(key ^ ctr) >> 61);
        f3 sc = mk(cc.x + ((oct & 1) ? off : -off), cc.y + ((oct & 2) ? off : -off), cc.z + ((oct & 4) ? off : -off));
        float jit[3];
        for (int a = 0; a < 3; ++a) {
            double uu = (double)(splitmix64(key ^ (ctr | (uint64_t)(1 + a))) >> 11) * (1.0 / 9007199254740992.0);
            double val = lo + (hi - lo) * uu;
            jit[a] = (float)(val * (double)dx);
        }
        f3 p = mk(sc.x + jit[0], sc.y + jit[1], sc.z + jit[2]);
        float lx = i * dx, hx = (i + 1) * dx, ly = j * dx, hy = (j + 1) * dx, lz = k * dx, hz = (k + 1) * dx;
        p.x = p.x < lx ? lx : (hx < p.x ? hx : p.x);
        p.y = p.y < ly ? ly : (hy < p.y ? hy : p.y);
        p.z = p.z < lz ? lz : (hz < p.z ? hz : p.z);
        long long d = base + q;
        P.px[d] = p.x; P.py[d] = p.y; P.pz[d] = p.z;
        P.vx[d] = 0.f; P.vy[d] = 0.f; P.vz[d] = 0.f;
        P.id[d] = (int)d;
    }
}

// ---------------------------------------------------------------------------
// P2G over an LDS tile.  A block owns the faces of a TX x TY x TZ cell tile;
// the particles that reach them are those of the tile's cells plus a one-cell
// halo.  A particle of cell (ci,cj,ck) reaches, per component, a 2x3x3 window
// of faces around its cell (u: i in {ci,ci+1}, j in {cj-1..cj+1}, k in
// {ck-1..ck+1}; v and w alike with the 2-wide axis along y or z).  Along each
// axis a particle's base slot s gets 1 - f and slot s+1 gets f; a slot beyond
// the window is a face beyond the domain, which the reference skips.
// The halo is walked plane by plane along z; per plane kc:
//   S. the plane's particles (one contiguous range per halo row, the cells
//      being sorted) are copied to LDS by direct global->LDS loads, issued
//      while the previous plane is still being reduced (phase B), so their
//      latency is hidden; a plane with more than CAP particles takes further
//      batches;
//   A. one thread per halo cell sums the weights and weighted velocities of
//      its particles, in their sorted order, into its three windows in
//      registers (weights and products as AccumulateVelocities :468-594);
//   B. per component, the windows are summed along x across neighbouring
//      lanes (a halo row is one 16-lane DPP row), the sums go to LDS, and one
//      thread per owned face column adds those of the 2 or 3 cells in y that
//      land on its faces into registers that hold the column's faces of the
//      face planes the current plane reaches.
// A face plane is complete once the cell planes either side of it are done;
// its faces are then normalised (ComputeVelocityFunc :598-605), saved
// (saveVelocities) and given gravity (ApplyScalarForce :687-698) in one
// write.  No atomics: every face sum has one writer and a fixed order, so the
// result is also the same from run to run.
constexpr int P2G_SLOTS = 18;
typedef __attribute__((address_space(1))) const void gvoid_t;
typedef __attribute__((address_space(3))) void lvoid_t;

// Barrier that orders LDS reads and writes only: the direct global->LDS loads
// of the next plane stay in flight (a __syncthreads would wait for them).
FS_DEV void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Weights of one axis of a window: base slot s gets a0 = 1 - f and slot s + 1
// gets a1 = f.  s is 0 or 1 for every particle: the cell index and the base
// index are floors of the same quotient (or of it minus 1/2), clamped alike.
template <int N>
FS_DEV void p2g_axis(int s, float f, float (&w)[N]) {
    const bool s0 = s == 0;
    const float a0 = 1.0f - f, a1 = f;
    w[0] = s0 ? a0 : 0.0f;
    w[1] = s0 ? a1 : a0;
    if (N == 3) w[N - 1] = s0 ? 0.0f : a1;
}

// Adds one particle (rx, ry, rz = position / dx) of cell (ci,cj,ck) to that
// cell's window of one component: slot x + NX (y + NY z) gains
// (wx wy)(vc wz) in num and (wx wy) wz in den.
template <int COMP>
FS_DEV void p2g_add(const P2GArgs& a, int ci, int cj, int ck, float rx, float ry, float rz, float vc,
                    float (&num)[P2G_SLOTS], float (&den)[P2G_SLOTS]) {
    constexpr int NX = (COMP == 0) ? 2 : 3, NY = (COMP == 1) ? 2 : 3, NZ = (COMP == 2) ? 2 : 3;
    const int ox = (COMP == 0) ? ci : ci - 1, oy = (COMP == 1) ? cj : cj - 1, oz = (COMP == 2) ? ck : ck - 1;
    const Base b = mac_base<COMP>(rx, ry, rz, a.W, a.H, a.D);
    float wx[NX], wy[NY], wz[NZ], vz[NZ];
    p2g_axis<NX>(b.i - ox, b.fx, wx);
    p2g_axis<NY>(b.j - oy, b.fy, wy);
    p2g_axis<NZ>(b.k - oz, b.fz, wz);
#pragma unroll
    for (int z = 0; z < NZ; ++z) vz[z] = vc * wz[z];
#pragma unroll
    for (int y = 0; y < NY; ++y)
#pragma unroll
        for (int x = 0; x < NX; ++x) {
            const float wxy = wx[x] * wy[y];
#pragma unroll
            for (int z = 0; z < NZ; ++z) {
                const int sl = x + NX * (y + NY * z);
                num[sl] = fmaf(wxy, vz[z], num[sl]);
                den[sl] = fmaf(wxy, wz[z], den[sl]);
            }
        }
}

// Owned faces of one component: columns (li, lj) of [x0, hix] x [y0, hiy],
// planes [z0, hiz]; LX x LY columns, one per thread.
template <int COMP, int TX, int TY, int TZ>
struct P2GFaces {
    static constexpr int NX = (COMP == 0) ? 2 : 3, NY = (COMP == 1) ? 2 : 3, NZ = (COMP == 2) ? 2 : 3;
    static constexpr int LX = TX + (COMP == 0), LY = TY + (COMP == 1);
    int i, j, hiz;
    bool own;
    FS_DEV P2GFaces(const P2GArgs& a, int x0, int y0, int z0) {
        const int t = threadIdx.x;
        const int hix = (COMP == 0) ? ((x0 + TX >= a.W) ? a.W : x0 + TX - 1) : min(x0 + TX, a.W) - 1;
        const int hiy = (COMP == 1) ? ((y0 + TY >= a.H) ? a.H : y0 + TY - 1) : min(y0 + TY, a.H) - 1;
        hiz = (COMP == 2) ? ((z0 + TZ >= a.D) ? a.D : z0 + TZ - 1) : min(z0 + TZ, a.D) - 1;
        i = x0 + t % LX;
        j = y0 + t / LX;
        own = t < LX * LY && i <= hix && j <= hiy;
    }
    // normalise, save and add gravity to face (i, j, k)
    FS_DEV void store(const P2GArgs& a, int k, float num, float den) const {
        const int nx = a.W + (COMP == 0), ny = a.H + (COMP == 1);
        float* __restrict__ g = COMP == 0 ? a.u : COMP == 1 ? a.v : a.w;
        float* __restrict__ gs = COMP == 0 ? a.us : COMP == 1 ? a.vs : a.ws;
        const float acc = COMP == 1 ? a.gravity_y : 0.0f;
        const float vel_unknown = -2147483648.0f;  // (float)INT_MIN, utility.cuh:56
        const float val = (den > 1e-8f) ? num / den : 0.0f;
        const size_t idx = (size_t)i + (size_t)j * nx + (size_t)k * nx * ny;
        gs[idx] = val;
        g[idx] = (val > vel_unknown) ? val + a.dt * acc : val;
    }
};

// Lane l gets v of lane l - 1 (SH = 1) or l + 1 (SH = -1) within its row of 16
// lanes; 0 past the row's ends.
template <int SH>
FS_DEV float row_shift(float v) {
    constexpr int ctrl = SH > 0 ? 0x110 + SH : 0x100 - SH;  // row_shr:SH / row_shl:-SH
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, 0xF, 0xF, false));
}

// Phase B of one component for the current cell plane, in two halves with a
// block barrier between them.  The halo rows are 16 cells wide, one DPP row of
// lanes each, so p2g_xsum first sums the windows along x across neighbouring
// lanes: the lane of cell (ci, cj) then holds, per window slot (y, z), the sum
// for face column i = ci, and these NY x NZ sums go to LDS (x: [num | den]
// [y z slot][halo cell]).  p2g_ysum then lets each owned face column add the
// sums of the NY cells in y that land on its face in slice z, face plane
// kc - (NZ == 3) + z, to rn[z] / rd[z].
template <int COMP, int TX, int TY, int TZ>
FS_DEV void p2g_xsum(const float (&num)[P2G_SLOTS], const float (&den)[P2G_SLOTS], float* __restrict__ x) {
    typedef P2GFaces<COMP, TX, TY, TZ> FT;
    constexpr int NX = FT::NX, NYZ = FT::NY * FT::NZ;
    constexpr int HX = TX + 2, HY = TY + 2, HP = HX * HY;
    static_assert(HX == 16 && HP == 256, "halo rows are DPP rows");
    const int t = threadIdx.x;
    // u face i gets slot 0 of cell i and slot 1 of cell i - 1; v and w face i
    // slot 0 of cell i + 1, slot 1 of cell i and slot 2 of cell i - 1
#pragma unroll
    for (int yz = 0; yz < NYZ; ++yz) {
        float sn, sd;
        if (NX == 2) {
            sn = num[NX * yz] + row_shift<1>(num[NX * yz + 1]);
            sd = den[NX * yz] + row_shift<1>(den[NX * yz + 1]);
        } else {
            sn = row_shift<-1>(num[NX * yz]) + num[NX * yz + 1] + row_shift<1>(num[NX * yz + 2]);
            sd = row_shift<-1>(den[NX * yz]) + den[NX * yz + 1] + row_shift<1>(den[NX * yz + 2]);
        }
        x[yz * HP + t] = sn;
        x[(NYZ + yz) * HP + t] = sd;
    }
}

template <int COMP, int TX, int TY, int TZ>
FS_DEV void p2g_ysum(const P2GFaces<COMP, TX, TY, TZ>& F, int x0, int y0, const float* __restrict__ x,
                     float (&rn)[3], float (&rd)[3]) {
    typedef P2GFaces<COMP, TX, TY, TZ> FT;
    constexpr int NY = FT::NY, NZ = FT::NZ, NYZ = NY * NZ;
    constexpr int HX = TX + 2, HY = TY + 2, HP = HX * HY;
    if (!F.own) return;
    // cells with window slot y = j - oy(cj), oy = cj - (NY == 3)
    const int hx = F.i - (x0 - 1);
    float sn[NZ], sd[NZ];
#pragma unroll
    for (int z = 0; z < NZ; ++z) { sn[z] = 0.0f; sd[z] = 0.0f; }
#pragma unroll
    for (int y = 0; y < NY; ++y) {
        const int hy = F.j - y + (NY == 3 ? 1 : 0) - (y0 - 1);
        if (hy < 0 || hy >= HY) continue;
        const int cell = hx + HX * hy;
#pragma unroll
        for (int z = 0; z < NZ; ++z) {
            sn[z] += x[(y + NY * z) * HP + cell];
            sd[z] += x[(NYZ + y + NY * z) * HP + cell];
        }
    }
#pragma unroll
    for (int z = 0; z < NZ; ++z) { rn[z] += sn[z]; rd[z] += sd[z]; }
}

// Halo rows of one cell plane: lanes q < HY of every wave hold row q's first
// particle (rb) and count (cnt); all other lanes hold 0.
struct P2GRows {
    int rb, cnt;
};

template <int TX, int TY>
FS_DEV P2GRows p2g_rows_load(const P2GArgs& a, int x0, int y0, int kc) {
    constexpr int HY = TY + 2;
    P2GRows r{0, 0};
    const int q = threadIdx.x & 63, rj = y0 - 1 + q;
    if (q < HY && rj >= 0 && rj < a.H) {
        const int lo = max(x0 - 1, 0), hi = min(x0 + TX, a.W - 1);
        const long long row = (long long)rj * a.W + (long long)kc * a.W * a.H;
        r.rb = a.start[row + lo];
        r.cnt = a.start[row + hi + 1] - r.rb;
    }
    return r;
}

// Copies particles [c0, c0 + CAP) of the plane's list (rows back to back) to
// stg by direct global->LDS loads: each wave takes every fourth row, 64
// particles per instruction.  excl: row offsets (exclusive scan, lane q).
template <int TY, int CAP>
FS_DEV void p2g_stage(const P2GArgs& a, float* __restrict__ stg, int c0, const P2GRows& R, int excl) {
    constexpr int HY = TY + 2;
    const float* __restrict__ src[6] = {a.P.px, a.P.py, a.P.pz, a.P.vx, a.P.vy, a.P.vz};
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nw = blockDim.x >> 6;
    for (int q = wave; q < HY; q += nw) {
        const int ex = __builtin_amdgcn_readlane(excl, q), cn = __builtin_amdgcn_readlane(R.cnt, q);
        const int rb = __builtin_amdgcn_readlane(R.rb, q);
        const int s0 = max(ex, c0), s1 = min(ex + cn, c0 + CAP);
        for (int g0 = s0; g0 < s1; g0 += 64) {
            const int ge = g0 + lane;
            if (ge < s1) {
                const int gi = rb + (ge - ex);
                float* dst = stg + (g0 - c0);
#pragma unroll
                for (int s = 0; s < 6; ++s)
                    __builtin_amdgcn_global_load_lds((gvoid_t*)(src[s] + gi), (lvoid_t*)(dst + s * CAP), 4, 0, 0);
            }
        }
    }
}

// CAP: particles of a plane staged per batch.
template <int TX, int TY, int TZ, int CAP, bool P2>
__global__ __launch_bounds__(256) void k_p2g_tile(P2GArgs a) {
    constexpr int HX = TX + 2, HY = TY + 2, HP = HX * HY;
    static_assert(HP <= 256 && HY <= 64, "one thread per halo cell of a plane; halo rows in one wave");
    static_assert((TX + 1) * TY <= 256 && TX * (TY + 1) <= 256, "one thread per owned face column");
    __shared__ float stg[6 * CAP];
    __shared__ float xch[2 * (9 + 6) * HP];

    const int W = a.W, H = a.H, D = a.D;
    const int ntx = (W + TX - 1) / TX, nty = (H + TY - 1) / TY;
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int tx = bid % ntx, ty = (bid / ntx) % nty, tz = bid / (ntx * nty);
    const int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
    const P2GFaces<0, TX, TY, TZ> FU(a, x0, y0, z0);
    const P2GFaces<1, TX, TY, TZ> FV(a, x0, y0, z0);
    const P2GFaces<2, TX, TY, TZ> FW(a, x0, y0, z0);
    const int lo = max(x0 - 1, 0), hi = min(x0 + TX, W - 1);
    // particles in the halo box, over its halo rows (cj, ck)
    bool any = false;
    for (int r = threadIdx.x; r < HY * (TZ + 2); r += blockDim.x) {
        const int cj = y0 - 1 + r % HY, ck = z0 - 1 + r / HY;
        if (cj >= 0 && cj < H && ck >= 0 && ck < D) {
            const long long row = (long long)cj * W + (long long)ck * W * H;
            any = any || a.start[row + hi + 1] > a.start[row + lo];
        }
    }
    if (!__syncthreads_or(any)) {
        // no particle reaches the tile: every owned face is empty
        for (int k = z0; k <= FU.hiz; ++k) {
            if (FU.own) FU.store(a, k, 0.0f, 0.0f);
            if (FV.own) FV.store(a, k, 0.0f, 0.0f);
        }
        if (FW.own)
            for (int k = z0; k <= FW.hiz; ++k) FW.store(a, k, 0.0f, 0.0f);
        return;
    }
    const int t = threadIdx.x;
    const int ci = x0 - 1 + t % HX, cj = y0 - 1 + t / HX;
    const bool cell_in = t < HP && ci >= 0 && ci < W && cj >= 0 && cj < H;
    const int kc0 = max(z0 - 1, 0), kc1 = min(z0 + TZ, D - 1);
    const float dx = a.dx, inv = a.inv_dx;

    // this cell's particles in plane kc as positions in the plane's list;
    // returns the plane's particle count (uniform)
    auto locate = [&](int kc, const P2GRows& R, int& excl, int& beg, int& end) -> int {
        int inc = R.cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(inc, off, 64);
            if ((t & 63) >= off) inc += y;
        }
        excl = inc - R.cnt;
        const int sh = __shfl(excl - R.rb, t / HX, 64);
        beg = end = 0;
        if (cell_in) {
            const long long c = (long long)ci + (long long)cj * W + (long long)kc * W * H;
            beg = a.start[c] + sh;
            end = a.start[c + 1] + sh;
        }
        return __builtin_amdgcn_readlane(inc, HY - 1);
    };

    // face sums of the owned columns: u, v planes kc - 1, kc, kc + 1; w planes
    // kc, kc + 1 (relative to the current cell plane)
    float un[3] = {0.f, 0.f, 0.f}, ud[3] = {0.f, 0.f, 0.f}, vn[3] = {0.f, 0.f, 0.f}, vd[3] = {0.f, 0.f, 0.f};
    float wn[3] = {0.f, 0.f, 0.f}, wd[3] = {0.f, 0.f, 0.f};

    // first plane: rows, ranges and its first batch; rows of the second
    P2GRows R = p2g_rows_load<TX, TY>(a, x0, y0, kc0);
    int excl, beg, end;
    int total = locate(kc0, R, excl, beg, end);
    p2g_stage<TY, CAP>(a, stg, 0, R, excl);
    P2GRows Rn = kc0 < kc1 ? p2g_rows_load<TX, TY>(a, x0, y0, kc0 + 1) : P2GRows{0, 0};
    for (int kc = kc0; kc <= kc1; ++kc) {
        float nu[P2G_SLOTS], du[P2G_SLOTS], nv[P2G_SLOTS], dv[P2G_SLOTS], nw[P2G_SLOTS], dw[P2G_SLOTS];
#pragma unroll
        for (int q = 0; q < P2G_SLOTS; ++q) {
            nu[q] = 0.f; du[q] = 0.f; nv[q] = 0.f; dv[q] = 0.f; nw[q] = 0.f; dw[q] = 0.f;
        }
        for (int c0 = 0; c0 < total; c0 += CAP) {
            if (c0 > 0) {  // a further batch of a crowded plane
                lds_barrier();
                p2g_stage<TY, CAP>(a, stg, c0, R, excl);
            }
            __syncthreads();  // this batch has landed
            // A: this cell's particles of the batch, in order
            const int pb = max(beg, c0), pe = min(end, c0 + CAP);
            for (int p = pb; p < pe; ++p) {
                const int e = p - c0;
                const float rx = div_dx<P2>(stg[e], dx, inv), ry = div_dx<P2>(stg[CAP + e], dx, inv),
                            rz = div_dx<P2>(stg[2 * CAP + e], dx, inv);
                p2g_add<0>(a, ci, cj, kc, rx, ry, rz, stg[3 * CAP + e], nu, du);
                p2g_add<1>(a, ci, cj, kc, rx, ry, rz, stg[4 * CAP + e], nv, dv);
                p2g_add<2>(a, ci, cj, kc, rx, ry, rz, stg[5 * CAP + e], nw, dw);
            }
        }
        lds_barrier();  // stg is free
        if (kc < kc1) {  // the next plane's first batch flies during phase B
            R = Rn;
            total = locate(kc + 1, R, excl, beg, end);
            p2g_stage<TY, CAP>(a, stg, 0, R, excl);
            Rn = kc + 1 < kc1 ? p2g_rows_load<TX, TY>(a, x0, y0, kc + 2) : P2GRows{0, 0};
        }
        // B: u and v, then w (xch holds 2 x (9 + 6) sums per halo cell); the
        // barriers of the next plane's phase A order its exchange after these
        // reads
        p2g_xsum<0, TX, TY, TZ>(nu, du, xch);
        p2g_xsum<1, TX, TY, TZ>(nv, dv, xch + 18 * HP);
        lds_barrier();
        p2g_ysum(FU, x0, y0, xch, un, ud);
        p2g_ysum(FV, x0, y0, xch + 18 * HP, vn, vd);
        lds_barrier();
        p2g_xsum<2, TX, TY, TZ>(nw, dw, xch);
        lds_barrier();
        p2g_ysum(FW, x0, y0, xch, wn, wd);
        // u, v face plane kc - 1 and w face plane kc are complete
        if (FU.own && kc - 1 >= z0 && kc - 1 <= FU.hiz) FU.store(a, kc - 1, un[0], ud[0]);
        if (FV.own && kc - 1 >= z0 && kc - 1 <= FV.hiz) FV.store(a, kc - 1, vn[0], vd[0]);
        if (FW.own && kc >= z0 && kc <= FW.hiz) FW.store(a, kc, wn[0], wd[0]);
        un[0] = un[1]; un[1] = un[2]; un[2] = 0.f;
        ud[0] = ud[1]; ud[1] = ud[2]; ud[2] = 0.f;
        vn[0] = vn[1]; vn[1] = vn[2]; vn[2] = 0.f;
        vd[0] = vd[1]; vd[1] = vd[2]; vd[2] = 0.f;
        wn[0] = wn[1]; wn[1] = 0.f;
        wd[0] = wd[1]; wd[1] = 0.f;
    }
    if (FU.own && kc1 >= z0 && kc1 <= FU.hiz) FU.store(a, kc1, un[0], ud[0]);
    if (FV.own && kc1 >= z0 && kc1 <= FV.hiz) FV.store(a, kc1, vn[0], vd[0]);
    if (FW.own && kc1 + 1 >= z0 && kc1 + 1 <= FW.hiz) FW.store(a, kc1 + 1, wn[0], wd[0]);
}

// ---------------------------------------------------------------------------
// G2P + advection, one thread per particle.

// PIC and FLIP samples of one component: du = u - uSaved per corner (:990-1012)
// GLOBAL_F: read f through a global-address-space pointer, so that the
// compiler cannot merge these reads with the LDS reads of the same corners on
// the staged path into generic (flat) loads that serve both address spaces
typedef __attribute__((address_space(1))) const float gfloat_t;
typedef __attribute__((address_space(1))) const uint8_t gbyte_t;
typedef __attribute__((address_space(3))) const uint8_t lbyte_t;
template <int COMP, bool GLOBAL_F = false>
FS_DEV void sample_pic_flip(const float* __restrict__ f, const float* __restrict__ fs, int n, const Base& b, int sx,
                            int sxy, float& pic, float& flip) {
    int i0 = b.i + b.j * sx + b.k * sxy;
    int ix[8] = {i0, i0 + 1, i0 + sxy, i0 + 1 + sxy, i0 + sx, i0 + 1 + sx, i0 + sx + sxy, i0 + 1 + sx + sxy};
    float c[8], d[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        int id = ix[q] < 0 ? 0 : (ix[q] >= n ? n - 1 : ix[q]);
        c[q] = GLOBAL_F ? *(gfloat_t*)(f + id) : f[id];
        d[q] = c[q] - fs[id];
    }
    pic = trilin(b.fx, b.fy, b.fz, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7]);
    flip = trilin(b.fx, b.fy, b.fz, d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7]);
}

FS_DEV f3 interp_vel(const G2PArgs& a, f3 pos) {
    return interp_vel3(a.u, a.v, a.w, a.nu, a.nv, a.nw, a.W, a.H, a.D, a.dx, pos);
}

__constant__ signed char c_proj_off[34][3] = {
    {1, 0, 0},    {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},   {0, 0, 1},  {0, 0, -1},   {1, 1, 0},
    {-1, 1, 0},   {1, -1, 0},  {-1, -1, 0}, {1, 0, 1},    {-1, 0, 1}, {1, 0, -1},   {-1, 0, -1},
    {0, 1, 1},    {0, -1, 1},  {0, 1, -1},  {0, -1, -1},  {1, 1, 1},  {-1, 1, 1},   {1, -1, 1},
    {1, 1, -1},   {-1, -1, 1}, {-1, 1, -1}, {1, -1, -1},  {-1, -1, -1}, {-2, -1, 0}, {0, -1, -2},
    {-2, -2, 0},  {0, -2, -2}, {-2, -2, -2}, {-2, 0, -2}, {2, -1, 0},  {0, -1, 2}};

// projectParticleDevice3D :1161-1239 (FLUID pass only, as the reference's loop runs once)
template <bool P2>
FS_DEV bool project_particle(f3& pos, const uint8_t* __restrict__ labels, int W, int H, int D, float dx, float inv) {
    int cx = (int)floorf(div_dx<P2>(pos.x, dx, inv)), cy = (int)floorf(div_dx<P2>(pos.y, dx, inv)),
        cz = (int)floorf(div_dx<P2>(pos.z, dx, inv));
    float bestD = 1e10f;
    int best = -1;
    for (int n = 0; n < 34; ++n) {
        int nx = cx + c_proj_off[n][0], ny = cy + c_proj_off[n][1], nz = cz + c_proj_off[n][2];
        if (nx < 0 || nx >= W || ny < 0 || ny >= H || nz < 0 || nz >= D) continue;
        if (labels[(size_t)nx + (size_t)ny * W + (size_t)nz * W * H] != FS_FLUID) continue;
        f3 cc = mk((nx + 0.5f) * dx, (ny + 0.5f) * dx, (nz + 0.5f) * dx);
        float d = (cc.x - pos.x) * (cc.x - pos.x) + (cc.y - pos.y) * (cc.y - pos.y) + (cc.z - pos.z) * (cc.z - pos.z);
        if (d < bestD) { bestD = d; best = n; }
    }
    if (best < 0) return false;
    pos = pos + mk((float)c_proj_off[best][0], (float)c_proj_off[best][1], (float)c_proj_off[best][2]) * 0.5f * dx;
    return true;
}

// max |v| over a block, folded into a device scalar with one atomic per block
// (float bits of a non-negative value order like unsigned integers).  Atomics
// on one address serialise in L2, so one per wave would cost milliseconds at
// tens of millions of particles.  Every thread of the block must call it.
FS_DEV void block_max_speed(float sp, unsigned* umax) {
    __shared__ float s_wmax[16];
    for (int off = 32; off > 0; off >>= 1) sp = fmaxf(sp, __shfl_xor(sp, off, 64));
    const int wv = threadIdx.x >> 6;
    if (__lane_id() == 0) s_wmax[wv] = sp;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = 0.0f;
        for (int q = 0; q < (int)((blockDim.x + 63) >> 6); ++q) m = fmaxf(m, s_wmax[q]);
        if (m > 0.0f) atomicMax(umax, __float_as_uint(m));
    }
}

__global__ __launch_bounds__(256) void k_max_speed(ParticlesView P, int n, unsigned* umax) {
    float sp = 0.0f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float vx = P.vx[i], vy = P.vy[i], vz = P.vz[i];
        float s = sqrtf(vx * vx + vy * vy + vz * vz);
        if (!isnan(s)) sp = fmaxf(sp, s);
    }
    block_max_speed(sp, umax);
}

// G2P + advection over a tile of TX x TY x TZ cells.  The particles of a tile
// are the rows of the cell-sorted arrays; u, v, w and the labels of the tile
// plus a margin of M cells are staged in LDS once, and every interpolation of
// the RK3 substeps reads LDS.  A sample whose 8 corners are not all inside the
// staged box (a particle that left the margin) reads global memory instead,
// with the same indices, so both paths return the same values.
// LDS layout of one staged face component: row and plane pitches padded so
// that the corners a half-wave reads (lanes spread along x over a few cells,
// differing by one row or plane) fall in different bank ranges.
constexpr int lds_row_pitch(int lx) {
    int p = lx;
    while (p % 32 < 9 || p % 32 > 23) ++p;
    return p;
}
constexpr int lds_plane_pitch(int lx, int ly) {
    const int pr = lds_row_pitch(lx), want = (2 * (pr % 32)) % 32;
    int p = pr * ly;
    while (p % 32 != want) ++p;
    return p;
}
template <int BX, int BY, int BZ, int COMP>
struct LdsFace {
    static constexpr int LX = BX + (COMP == 0), LY = BY + (COMP == 1), LZ = BZ + (COMP == 2);
    static constexpr int PR = lds_row_pitch(LX), PP = lds_plane_pitch(LX, LY), SIZE = PP * LZ;
};

template <int BX, int BY, int BZ, bool P2>
struct TileFields {
    int X0, Y0, Z0;
    const float* su;
    const float* sv;
    const float* sw;
    const uint8_t* sl;
    const G2PArgs* a;
    // fast-path window of r = x / dx per axis, [lo, hi): see interp()
    float lox, hix, loy, hiy, loz, hiz;

    FS_DEV void set_window() {
        lox = (float)(max(X0, 0) + 1); hix = (float)(min(X0 + BX, a->W) - 1);
        loy = (float)(max(Y0, 0) + 1); hiy = (float)(min(Y0 + BY, a->H) - 1);
        loz = (float)(max(Z0, 0) + 1); hiz = (float)(min(Z0 + BZ, a->D) - 1);
    }

    template <int COMP>
    FS_DEV bool in_box(const Base& b, int& i0) const {
        using F = LdsFace<BX, BY, BZ, COMP>;
        const int nx = a->W + (COMP == 0), ny = a->H + (COMP == 1), nz = a->D + (COMP == 2);
        int li = b.i - X0, lj = b.j - Y0, lk = b.k - Z0;
        bool ok = li >= 0 && li + 1 < F::LX && lj >= 0 && lj + 1 < F::LY && lk >= 0 && lk + 1 < F::LZ &&
                  b.i + 1 < nx && b.j + 1 < ny && b.k + 1 < nz;
        i0 = li + F::PR * lj + F::PP * lk;
        return ok;
    }

    template <int COMP>
    FS_DEV float lds_sample(const Base& b, int i0) const {
        using F = LdsFace<BX, BY, BZ, COMP>;
        const float* s = COMP == 0 ? su : (COMP == 1 ? sv : sw);
        constexpr int R = F::PR, P = F::PP;
        return trilin(b.fx, b.fy, b.fz, s[i0], s[i0 + 1], s[i0 + P], s[i0 + 1 + P], s[i0 + R], s[i0 + 1 + R],
                      s[i0 + R + P], s[i0 + 1 + R + P]);
    }

    template <int COMP>
    FS_DEV float global_sample(const Base& b) const {
        const int W = a->W, H = a->H;
        if (COMP == 0) return sample_comp(a->u, a->nu, b, W + 1, (W + 1) * H);
        if (COMP == 1) return sample_comp(a->v, a->nv, b, W, W * (H + 1));
        return sample_comp(a->w, a->nw, b, W, W * H);
    }

    // PIC value and FLIP delta (u - uSaved per corner, FluidSolver3D.cu:990-1012)
    template <int COMP>
    FS_DEV void pic_flip(const Base& b, float& pic, float& flip) const {
        constexpr int LX = LdsFace<BX, BY, BZ, COMP>::PR, SXY = LdsFace<BX, BY, BZ, COMP>::PP;
        const int W = a->W, H = a->H;
        const int sx = COMP == 0 ? W + 1 : W;
        const int sxy = COMP == 0 ? (W + 1) * H : (COMP == 1 ? W * (H + 1) : W * H);
        const float* fs = COMP == 0 ? a->us : (COMP == 1 ? a->vs : a->ws);
        const int n = COMP == 0 ? a->nu : (COMP == 1 ? a->nv : a->nw);
        int i0;
        if (in_box<COMP>(b, i0)) {
            const float* s = COMP == 0 ? su : (COMP == 1 ? sv : sw);
            const int lo[8] = {0, 1, SXY, 1 + SXY, LX, 1 + LX, LX + SXY, 1 + LX + SXY};
            const int g0 = b.i + b.j * sx + b.k * sxy;
            const int go[8] = {0, 1, sxy, 1 + sxy, sx, 1 + sx, sx + sxy, 1 + sx + sxy};
            float c[8], d[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                c[q] = s[i0 + lo[q]];
                d[q] = c[q] - fs[g0 + go[q]];
            }
            pic = trilin(b.fx, b.fy, b.fz, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7]);
            flip = trilin(b.fx, b.fy, b.fz, d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7]);
            return;
        }
        const float* f = COMP == 0 ? a->u : (COMP == 1 ? a->v : a->w);
        sample_pic_flip<COMP, true>(f, fs, n, b, sx, sxy, pic, flip);
    }

    FS_DEV f3 interp(f3 pos) const {
        const float dx = a->dx, inv = a->inv_dx;
        return interp_r(div_dx<P2>(pos.x, dx, inv), div_dx<P2>(pos.y, dx, inv), div_dx<P2>(pos.z, dx, inv));
    }

    // the same sample from r = pos / dx per axis
    FS_DEV f3 interp_r(float rx, float ry, float rz) const {
        // Fast path: floor(r) in [lo0 + 1, hi - 2] on every axis, lo0 = max(X0, 0)
        // and hi = min(X0 + BX, W).  Then both bases per axis, floor(r) and
        // floor(r - 1/2) (which is floor(r) or floor(r) - 1), lie in the staged
        // box and the domain with their +1 corners, mac_base's clamps are
        // identities, r >= 1 makes truncation equal floor, and the fractions
        // r - floor(r) are exact (so the hardware fract gives the same bits).
        // NaN or infinite r fails the comparisons.  Other cases take the
        // general path below, which returns the same values.
        if (rx >= lox && rx < hix && ry >= loy && ry < hiy && rz >= loz && rz < hiz) {
            const float hx = rx - 0.5f, hy = ry - 0.5f, hz = rz - 0.5f;
            const int fx = (int)rx, fy = (int)ry, fz = (int)rz;
            const int gx = (int)hx, gy = (int)hy, gz = (int)hz;
            const float frx = __builtin_amdgcn_fractf(rx), fry = __builtin_amdgcn_fractf(ry),
                        frz = __builtin_amdgcn_fractf(rz);
            const float fhx = __builtin_amdgcn_fractf(hx), fhy = __builtin_amdgcn_fractf(hy),
                        fhz = __builtin_amdgcn_fractf(hz);
            Base bu{fx, gy, gz, frx, fhy, fhz};
            Base bv{gx, fy, gz, fhx, fry, fhz};
            Base bw{gx, gy, fz, fhx, fhy, frz};
            using FU = LdsFace<BX, BY, BZ, 0>;
            using FV = LdsFace<BX, BY, BZ, 1>;
            using FW = LdsFace<BX, BY, BZ, 2>;
            // box-local offsets are small and non-negative here: 24-bit multiplies
            // (full rate) instead of 32-bit ones (quarter rate)
            const int iu = (fx - X0) + __mul24(FU::PR, gy - Y0) + __mul24(FU::PP, gz - Z0);
            const int iv = (gx - X0) + __mul24(FV::PR, fy - Y0) + __mul24(FV::PP, gz - Z0);
            const int iw = (gx - X0) + __mul24(FW::PR, gy - Y0) + __mul24(FW::PP, fz - Z0);
            return mk(lds_sample<0>(bu, iu), lds_sample<1>(bv, iv), lds_sample<2>(bw, iw));
        }
        Base bu = mac_base<0>(rx, ry, rz, a->W, a->H, a->D);
        Base bv = mac_base<1>(rx, ry, rz, a->W, a->H, a->D);
        Base bw = mac_base<2>(rx, ry, rz, a->W, a->H, a->D);
        // one branch for all 24 corners, so that their LDS reads issue together
        int iu, iv, iw;
        const bool in = (int)in_box<0>(bu, iu) + (int)in_box<1>(bv, iv) + (int)in_box<2>(bw, iw) == 3;
        if (in) return mk(lds_sample<0>(bu, iu), lds_sample<1>(bv, iv), lds_sample<2>(bw, iw));
        return mk(global_sample<0>(bu), global_sample<1>(bv), global_sample<2>(bw));
    }

    // label of an in-domain cell: an LDS read (index clamped to the box), and a
    // global read only for a cell outside the staged box, so that the common
    // case is a plain LDS load rather than one generic load for both spaces
    FS_DEV uint8_t label(int cx, int cy, int cz) const {
        const unsigned li = cx - X0, lj = cy - Y0, lk = cz - Z0;
        const bool in = (li < (unsigned)BX) & (lj < (unsigned)BY) & (lk < (unsigned)BZ);
        // 32-bit LDS address arithmetic
        const unsigned off = li + __umul24(BX, lj) + __umul24(BX * BY, lk);
        uint8_t L = ((lbyte_t*)sl)[in ? off : 0u];
        if (!in) L = *(gbyte_t*)(a->labels + (size_t)cx + (size_t)cy * a->W + (size_t)cz * a->W * a->H);
        return L;
    }
};

template <int TX, int TY, int TZ, int M, bool P2>
__global__ __launch_bounds__(256) void k_g2p_advect_tile(G2PArgs a, const int* __restrict__ start) {
    constexpr int BX = TX + 2 * M, BY = TY + 2 * M, BZ = TZ + 2 * M;
    using FU = LdsFace<BX, BY, BZ, 0>;
    using FV = LdsFace<BX, BY, BZ, 1>;
    using FW = LdsFace<BX, BY, BZ, 2>;
    constexpr int NU = (BX + 1) * BY * BZ, NV = BX * (BY + 1) * BZ, NW = BX * BY * (BZ + 1), NL = BX * BY * BZ;
    constexpr int ROWS = TY * TZ;
    __shared__ float su[FU::SIZE], sv[FV::SIZE], sw[FW::SIZE];
    __shared__ uint8_t sl[NL];
    __shared__ int s_beg[ROWS], s_pref[ROWS + 1], s_scan[4];

    const int W = a.W, H = a.H, D = a.D;
    const int ntx = (W + TX - 1) / TX, nty = (H + TY - 1) / TY;
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int x0 = (bid % ntx) * TX, y0 = ((bid / ntx) % nty) * TY, z0 = (bid / (ntx * nty)) * TZ;

    int cnt = 0;
    if ((int)threadIdx.x < ROWS) {
        int r = threadIdx.x;
        int cy = y0 + r % TY, cz = z0 + r / TY;
        if (cy < H && cz < D) {
            long long rowc = (long long)cy * W + (long long)cz * W * H;
            int beg = start[rowc + x0], end = start[rowc + min(x0 + TX, W)];
            s_beg[r] = beg;
            cnt = end - beg;
        } else {
            s_beg[r] = 0;
        }
    }
    int tot;
    int pre = block_excl_scan(cnt, s_scan, &tot);
    if (tot == 0) return;   // uniform over the block
    if ((int)threadIdx.x < ROWS) s_pref[threadIdx.x] = pre;
    if (threadIdx.x == 0) s_pref[ROWS] = tot;

    // stage the velocity faces and labels of the tile + margin
    const int X0 = x0 - M, Y0 = y0 - M, Z0 = z0 - M;
    for (int l = threadIdx.x; l < NU; l += blockDim.x) {
        int li = l % (BX + 1), lj = (l / (BX + 1)) % BY, lk = l / ((BX + 1) * BY);
        int i = X0 + li, j = Y0 + lj, k = Z0 + lk;
        su[li + FU::PR * lj + FU::PP * lk] = (i >= 0 && i <= W && j >= 0 && j < H && k >= 0 && k < D)
                    ? a.u[(size_t)i + (size_t)j * (W + 1) + (size_t)k * (W + 1) * H] : 0.0f;
    }
    for (int l = threadIdx.x; l < NV; l += blockDim.x) {
        int li = l % BX, lj = (l / BX) % (BY + 1), lk = l / (BX * (BY + 1));
        int i = X0 + li, j = Y0 + lj, k = Z0 + lk;
        sv[li + FV::PR * lj + FV::PP * lk] = (i >= 0 && i < W && j >= 0 && j <= H && k >= 0 && k < D)
                    ? a.v[(size_t)i + (size_t)j * W + (size_t)k * W * (H + 1)] : 0.0f;
    }
    for (int l = threadIdx.x; l < NW; l += blockDim.x) {
        int li = l % BX, lj = (l / BX) % BY, lk = l / (BX * BY);
        int i = X0 + li, j = Y0 + lj, k = Z0 + lk;
        sw[li + FW::PR * lj + FW::PP * lk] = (i >= 0 && i < W && j >= 0 && j < H && k >= 0 && k <= D)
                    ? a.w[(size_t)i + (size_t)j * W + (size_t)k * W * H] : 0.0f;
    }
    for (int l = threadIdx.x; l < NL; l += blockDim.x) {
        int i = X0 + l % BX, j = Y0 + (l / BX) % BY, k = Z0 + l / (BX * BY);
        sl[l] = (i >= 0 && i < W && j >= 0 && j < H && k >= 0 && k < D)
                    ? a.labels[(size_t)i + (size_t)j * W + (size_t)k * W * H] : (uint8_t)FS_SOLID;
    }
    __syncthreads();

    TileFields<BX, BY, BZ, P2> F{X0, Y0, Z0, su, sv, sw, sl, &a};
    F.set_window();
    const float dx = a.dx, dt = a.dt, C = a.C, al = a.alpha, inv = a.inv_dx;
    const float vfloor = sqrtf(5.0f * dx * 9.8f);
    const float xmax = (W - 1) * dx, ymax = (H - 1) * dx, zmax = (D - 1) * dx;

    // The particle velocity update (G2P) and the advection of the position are
    // independent (advection reads the grid, not the particle velocity), so
    // they run as two passes over the block's particles: the register peak of
    // each pass alone is lower than that of the two fused.
    auto particle_of = [&](int t) {
        int lo = 0, hi = ROWS - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (s_pref[mid] <= t) lo = mid; else hi = mid - 1;
        }
        return s_beg[lo] + (t - s_pref[lo]);
    };
    float speed_max = 0.0f;   // of this thread's particles, folded once per block below
    // ---- G2P (GridToParticleFunctor :822-979)
    for (int base = 0; base < tot; base += blockDim.x) {
        const int t = base + threadIdx.x;
        if (t < tot) {
            const int p = particle_of(t);
            f3 pos = mk(a.P.px[p], a.P.py[p], a.P.pz[p]);
            f3 vel = mk(a.P.vx[p], a.P.vy[p], a.P.vz[p]);
            float rx = div_dx<P2>(pos.x, dx, inv), ry = div_dx<P2>(pos.y, dx, inv), rz = div_dx<P2>(pos.z, dx, inv);
            Base bu = mac_base<0>(rx, ry, rz, W, H, D);
            Base bv = mac_base<1>(rx, ry, rz, W, H, D);
            Base bw = mac_base<2>(rx, ry, rz, W, H, D);
            float uP, uF, vP, vF, wP, wF;
            F.template pic_flip<0>(bu, uP, uF);
            F.template pic_flip<1>(bv, vP, vF);
            F.template pic_flip<2>(bw, wP, wF);
            vel.x = al * uP + (1.0f - al) * (vel.x + uF);
            vel.y = al * vP + (1.0f - al) * (vel.y + vF);
            vel.z = al * wP + (1.0f - al) * (vel.z + wF);
            a.P.vx[p] = vel.x; a.P.vy[p] = vel.y; a.P.vz[p] = vel.z;
            float sp = sqrtf(vel.x * vel.x + vel.y * vel.y + vel.z * vel.z);
            if (!isnan(sp)) speed_max = fmaxf(speed_max, sp);
        }
    }
    // ---- advection (AdvectParticlesFunctor :1260-1330)
    for (int base = 0; base < tot; base += blockDim.x) {
        const int t = base + threadIdx.x;
        const bool act = t < tot;
        int key = -1, p = 0;
        if (act) {
            p = particle_of(t);
            f3 pos = mk(a.P.px[p], a.P.py[p], a.P.pz[p]);
            // pos / dx, carried from the cell index of one substep to the first
            // sample of the next
            float rx = div_dx<P2>(pos.x, dx, inv), ry = div_dx<P2>(pos.y, dx, inv), rz = div_dx<P2>(pos.z, dx, inv);
            float subT = 0.0f;
            bool finished = false;
            int iter = 0;
            while (!finished && iter++ < 1000) {
                f3 k1 = F.interp_r(rx, ry, rz);
                float speed = sqrtf(k1.x * k1.x + k1.y * k1.y + k1.z * k1.z) + 1e-37f + vfloor;
                float dT = (C * dx) / speed;
                if (subT + dT >= dt) {
                    dT = dt - subT;
                    finished = true;
                } else if (subT + 2 * dT >= dt) {
                    dT *= 0.5f;
                }
                f3 k2 = F.interp(pos + k1 * 0.5f * dT);
                f3 k3 = F.interp(pos + k2 * 0.75f * dT);
                pos = pos + (2.0f / 9.0f) * dT * k1 + (3.0f / 9.0f) * dT * k2 + (4.0f / 9.0f) * dT * k3;
                subT += dT;
                pos.x = fmaxf(pos.x, 0.0f); pos.x = fminf(pos.x, xmax);
                pos.y = fmaxf(pos.y, 0.0f); pos.y = fminf(pos.y, ymax);
                pos.z = fmaxf(pos.z, 0.0f); pos.z = fminf(pos.z, zmax);
                // the clamps leave pos finite and >= 0 (fmaxf returns the non-NaN
                // operand), so truncation is the floor
                rx = div_dx<P2>(pos.x, dx, inv); ry = div_dx<P2>(pos.y, dx, inv); rz = div_dx<P2>(pos.z, dx, inv);
                const int cx = (int)rx, cy = (int)ry, cz = (int)rz;
                // one predicate and one branch (the rare projection) per substep
                const bool inside = ((unsigned)cx < (unsigned)W) & ((unsigned)cy < (unsigned)H) &
                                    ((unsigned)cz < (unsigned)D);
                if (inside) {
                    const uint8_t L = F.label(cx, cy, cz);
                    // L is SOLID or BODY, as a bit test rather than two branches
                    const bool blocked = (((1u << (L & 3u)) & ((1u << FS_SOLID) | (1u << FS_BODY))) != 0) & (L < 4);
                    if (blocked) {
                        if (!project_particle<P2>(pos, a.labels, W, H, D, dx, inv)) break;
                        rx = div_dx<P2>(pos.x, dx, inv); ry = div_dx<P2>(pos.y, dx, inv); rz = div_dx<P2>(pos.z, dx, inv);
                    }
                }
                if (pos.x < 0.0f || pos.y < 0.0f || pos.z < 0.0f || isnan(pos.x) || isnan(pos.y) || isnan(pos.z))
                    break;
            }
            a.P.px[p] = pos.x; a.P.py[p] = pos.y; a.P.pz[p] = pos.z;
            if (a.next_count) key = cell_key_t<P2>(pos.x, pos.y, pos.z, dx, inv, W, H, D);
        }
        if (a.next_count) {
            int r = wave_agg_rank(key, act, a.next_count);
            if (act) {
                a.next_key[p] = key;
                a.next_rank[p] = r;
            }
        }
    }
    if (a.umax) block_max_speed(speed_max, a.umax);
}

// ---------------------------------------------------------------------------
// host launchers
void launch_key_count(const ParticlesView& P, int n, float dx, int W, int H, int D, int* key, int* rank, int* count,
                      hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_key_count, dim3((n + 255) / 256), dim3(256), 0, st, P.px, P.py, P.pz, n, dx, W, H, D, key,
                       rank, count);
}

int scan_blocks(long long C) { return (int)((C + SCAN_PER_BLOCK - 1) / SCAN_PER_BLOCK); }

void launch_scan_label(int* count, const uint8_t* solid, uint8_t* labels, long long C, int* bsum, int* total,
                       int* start, const LabelArgs& la, hipStream_t st) {
    int nb = scan_blocks(C);
    hipLaunchKernelGGL(k_scan_reduce<0>, dim3(nb), dim3(SCAN_T), 0, st, count, labels, C, bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, bsum, nb, total);
    hipLaunchKernelGGL(k_scan_final<0>, dim3(nb), dim3(SCAN_T), 0, st, count, solid, labels, C, bsum, start, la);
}

void launch_scan_sort(int* count, long long C, int* bsum, int* total, int* start, hipStream_t st) {
    int nb = scan_blocks(C);
    LabelArgs la{};
    hipLaunchKernelGGL(k_scan_reduce<2>, dim3(nb), dim3(SCAN_T), 0, st, count, nullptr, C, bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, bsum, nb, total);
    hipLaunchKernelGGL(k_scan_final<2>, dim3(nb), dim3(SCAN_T), 0, st, count, nullptr, nullptr, C, bsum, start, la);
}

void launch_scan_fluid(const uint8_t* labels, long long C, int* bsum, int* total, int* start, hipStream_t st) {
    int nb = scan_blocks(C);
    LabelArgs la{};
    hipLaunchKernelGGL(k_scan_reduce<1>, dim3(nb), dim3(SCAN_T), 0, st, nullptr, labels, C, bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, st, bsum, nb, total);
    hipLaunchKernelGGL(k_scan_final<1>, dim3(nb), dim3(SCAN_T), 0, st, nullptr, nullptr, const_cast<uint8_t*>(labels),
                       C, bsum, start, la);
}

void launch_seed(const uint8_t* labels, const int* frank, long long C, int W, int H, float dx, int ppc, uint64_t seed,
                 const ParticlesView& P, hipStream_t st) {
    hipLaunchKernelGGL(k_seed, dim3((unsigned)((C + 255) / 256)), dim3(256), 0, st, labels, frank, C, W, H, dx, ppc,
                       seed, P);
}

void launch_scatter(const ParticlesView& src, const ParticlesView& dst, int n, const int* key, const int* rank,
                    const int* start, int* dkey, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_scatter, dim3((n + 255) / 256), dim3(256), 0, st, src, dst, n, key, rank, start, dkey);
}

void launch_canonical(const ParticlesView& src, const ParticlesView& dst, int n, const int* dkey, const int* start,
                      hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_canonical, dim3((n + 255) / 256), dim3(256), 0, st, src, dst, n, dkey, start);
}

void launch_p2g(const P2GArgs& a, hipStream_t st) {
    constexpr int TX = P2G_TX, TY = P2G_TY, TZ = P2G_TZ, CAP = P2G_CAP;
    int nb = ((a.W + TX - 1) / TX) * ((a.H + TY - 1) / TY) * ((a.D + TZ - 1) / TZ);
    if (a.inv_dx > 0.0f)
        hipLaunchKernelGGL((k_p2g_tile<TX, TY, TZ, CAP, true>), dim3(nb), dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((k_p2g_tile<TX, TY, TZ, CAP, false>), dim3(nb), dim3(256), 0, st, a);
}

void launch_max_speed(const ParticlesView& P, int n, unsigned* umax, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_max_speed, dim3(std::min(2048, (n + 255) / 256)), dim3(256), 0, st, P, n, umax);
}

void launch_g2p_advect(const G2PArgs& a, const int* start, hipStream_t st) {
    if (a.n <= 0) return;
    constexpr int TX = ADV_TX, TY = ADV_TY, TZ = ADV_TZ;
    static_assert(TY * TZ <= 256, "one tile row per thread");
    int nb = ((a.W + TX - 1) / TX) * ((a.H + TY - 1) / TY) * ((a.D + TZ - 1) / TZ);
    if (a.inv_dx > 0.0f)
        hipLaunchKernelGGL((k_g2p_advect_tile<TX, TY, TZ, ADV_MARGIN, true>), dim3(nb), dim3(256), 0, st, a, start);
    else
        hipLaunchKernelGGL((k_g2p_advect_tile<TX, TY, TZ, ADV_MARGIN, false>), dim3(nb), dim3(256), 0, st, a, start);
}

}  // namespace fsk
