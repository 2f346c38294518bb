// fs_device.hpp -- device-side helpers shared by the gfx950 kernels.
//
// All FP32 arithmetic here keeps the reference's operation order
// (src/FluidSim/FluidSolver3D.cu, utility.cu) and the library is compiled with
// -ffp-contract=off, so per-element stages reproduce the reference's FP32 bits
// wherever the reference's own result is defined.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fluidsim.h"

#define FS_DEV __device__ __forceinline__
#define FS_HD __host__ __device__ __forceinline__

namespace fsd {

struct f3 {
    float x, y, z;
};
FS_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
// utility.cu:3-25
FS_HD f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
FS_HD f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
FS_HD f3 operator*(f3 a, float b) { return mk(a.x * b, a.y * b, a.z * b); }
FS_HD f3 operator*(float b, f3 a) { return a * b; }
FS_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
FS_HD f3 operator/(f3 a, float b) { return a * (1.0f / b); }
FS_HD f3 cross(f3 a, f3 b) {  // utility.cu:208-215
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Particle -> cell index, MarkFluidCellsFunctor FluidSolver3D.cu:322-332.
// IEEE division and floorf: bit-exact with the reference.
FS_DEV int cell_key(float px, float py, float pz, float dx, int W, int H, int D) {
    int i = (int)floorf(px / dx);
    int j = (int)floorf(py / dx);
    int k = (int)floorf(pz / dx);
    i = i < 0 ? 0 : (i >= W ? W - 1 : i);
    j = j < 0 ? 0 : (j >= H ? H - 1 : j);
    k = k < 0 ? 0 : (k >= D ? D - 1 : k);
    return i + j * W + k * (W * H);
}

// x / dx.  P2: dx is a power of two and inv = 1/dx, so the product is the
// same bits as the IEEE quotient (both are exact scalings by 2^k) at a fraction
// of the cost of a correctly rounded division.
template <bool P2>
FS_DEV float div_dx(float x, float dx, float inv) {
    return P2 ? x * inv : x / dx;
}

template <bool P2>
FS_DEV int cell_key_t(float px, float py, float pz, float dx, float inv, int W, int H, int D) {
    int i = (int)floorf(div_dx<P2>(px, dx, inv));
    int j = (int)floorf(div_dx<P2>(py, dx, inv));
    int k = (int)floorf(div_dx<P2>(pz, dx, inv));
    i = i < 0 ? 0 : (i >= W ? W - 1 : i);
    j = j < 0 ? 0 : (j >= H ? H - 1 : j);
    k = k < 0 ? 0 : (k >= D ? D - 1 : k);
    return i + j * W + k * (W * H);
}

FS_DEV float clamp01(float f) { return fminf(fmaxf(f, 0.0f), 1.0f); }

// Base corner and fractions of one MAC component (FluidSolver3D.cu:476-571, :833-939).
struct Base {
    int i, j, k;
    float fx, fy, fz;
};
template <int COMP>
FS_DEV Base mac_base(float rx, float ry, float rz, int W, int H, int D) {
    Base b;
    if (COMP == 0) {
        b.i = (int)floorf(rx); b.j = (int)floorf(ry - 0.5f); b.k = (int)floorf(rz - 0.5f);
        b.i = min(max(b.i, 0), W); b.j = min(max(b.j, 0), H - 1); b.k = min(max(b.k, 0), D - 1);
        b.fx = rx - b.i; b.fy = (ry - 0.5f) - b.j; b.fz = (rz - 0.5f) - b.k;
    } else if (COMP == 1) {
        b.i = (int)floorf(rx - 0.5f); b.j = (int)floorf(ry); b.k = (int)floorf(rz - 0.5f);
        b.i = min(max(b.i, 0), W - 1); b.j = min(max(b.j, 0), H); b.k = min(max(b.k, 0), D - 1);
        b.fx = (rx - 0.5f) - b.i; b.fy = ry - b.j; b.fz = (rz - 0.5f) - b.k;
    } else {
        b.i = (int)floorf(rx - 0.5f); b.j = (int)floorf(ry - 0.5f); b.k = (int)floorf(rz);
        b.i = min(max(b.i, 0), W - 1); b.j = min(max(b.j, 0), H - 1); b.k = min(max(b.k, 0), D);
        b.fx = (rx - 0.5f) - b.i; b.fy = (ry - 0.5f) - b.j; b.fz = rz - b.k;
    }
    b.fx = clamp01(b.fx); b.fy = clamp01(b.fy); b.fz = clamp01(b.fz);
    return b;
}

// trilinearInterpolation FluidSolver3D.cu:784-801; each lerp a*(1-t) + b*t is
// fma(a, 1-t, b*t) (the contraction a CUDA build makes by default; the oracle
// uses the same expression, so G2P and advection stay bit-exact against it)
FS_DEV float trilin(float xd, float yd, float zd, float c000, float c100, float c001, float c101,
                    float c010, float c110, float c011, float c111) {
    float c00 = fmaf(c000, 1.0f - xd, c100 * xd);
    float c01 = fmaf(c001, 1.0f - xd, c101 * xd);
    float c10 = fmaf(c010, 1.0f - xd, c110 * xd);
    float c11 = fmaf(c011, 1.0f - xd, c111 * xd);
    float c0 = fmaf(c00, 1.0f - yd, c10 * yd);
    float c1 = fmaf(c01, 1.0f - yd, c11 * yd);
    return fmaf(c0, 1.0f - zd, c1 * zd);
}

// Rigid-body pose used by the label pass and the face boundary conditions.
struct BodyParams {
    int active;
    int sw, sh, sd;
    float cs;
    f3 pos, vel, omega;
    f3 file_origin, local_com;
    float R[9];
    // a world sphere that holds every point body_contains can accept (the
    // SDF box's circumsphere plus two SDF cells), for a cheap early reject
    f3 bcen;
    float brad2;
};

FS_DEV bool body_sphere_has(const BodyParams& b, f3 wp) {
    const float dx = wp.x - b.bcen.x, dy = wp.y - b.bcen.y, dz = wp.z - b.bcen.z;
    return dx * dx + dy * dy + dz * dz <= b.brad2;
}

// Utility::contains utility.cu:163-206 (the reference passes fileOrigin as
// sdf_origin, FluidSolver3D.cu:183,359).
FS_DEV bool body_contains(const BodyParams& b, const float* __restrict__ sdf, f3 wp) {
    f3 pr = wp - b.pos;
    f3 plc;
    plc.x = b.R[0] * pr.x + b.R[3] * pr.y + b.R[6] * pr.z;
    plc.y = b.R[1] * pr.x + b.R[4] * pr.y + b.R[7] * pr.z;
    plc.z = b.R[2] * pr.x + b.R[5] * pr.y + b.R[8] * pr.z;
    f3 pl = plc + b.local_com;
    f3 sc = mk((pl.x - b.file_origin.x) / b.cs, (pl.y - b.file_origin.y) / b.cs,
               (pl.z - b.file_origin.z) / b.cs);
    if (sc.x < 0 || sc.x >= b.sw || sc.y < 0 || sc.y >= b.sh || sc.z < 0 || sc.z >= b.sd) return false;
    int i = min((int)sc.x, b.sw - 1);
    int j = min((int)sc.y, b.sh - 1);
    int k = min((int)sc.z, b.sd - 1);
    return sdf[(size_t)i + (size_t)j * b.sw + (size_t)k * b.sw * b.sh] <= 0.0f;
}

// ---- face fractions of the rigid body (an option with no reference
// counterpart: the reference voxelises the body at cell centres)
//
// Signed distance of the body at a world point: the SDF samples (at the
// points file_origin + cs (i, j, k) of the body frame) interpolated
// trilinearly; FACE_FAR (outside) beyond the sampled box.
constexpr float FACE_FAR = 1.0e30f;
FS_DEV float body_phi(const BodyParams& b, const float* __restrict__ sdf, f3 wp) {
    f3 pr = wp - b.pos;
    f3 plc;
    plc.x = b.R[0] * pr.x + b.R[3] * pr.y + b.R[6] * pr.z;
    plc.y = b.R[1] * pr.x + b.R[4] * pr.y + b.R[7] * pr.z;
    plc.z = b.R[2] * pr.x + b.R[5] * pr.y + b.R[8] * pr.z;
    f3 pl = plc + b.local_com;
    const float sx = (pl.x - b.file_origin.x) / b.cs, sy = (pl.y - b.file_origin.y) / b.cs,
                sz = (pl.z - b.file_origin.z) / b.cs;
    if (!(sx >= 0.0f && sx <= (float)(b.sw - 1) && sy >= 0.0f && sy <= (float)(b.sh - 1) && sz >= 0.0f &&
          sz <= (float)(b.sd - 1)))
        return FACE_FAR;
    const int i = min((int)sx, b.sw - 2 < 0 ? 0 : b.sw - 2), j = min((int)sy, b.sh - 2 < 0 ? 0 : b.sh - 2),
              k = min((int)sz, b.sd - 2 < 0 ? 0 : b.sd - 2);
    const float fx = sx - (float)i, fy = sy - (float)j, fz = sz - (float)k;
    const size_t sw = (size_t)b.sw, swh = (size_t)b.sw * b.sh;
    const size_t o = (size_t)i + (size_t)j * sw + (size_t)k * swh;
    const size_t ox = b.sw > 1 ? 1 : 0, oy = b.sh > 1 ? sw : 0, oz = b.sd > 1 ? swh : 0;
    const float c00 = sdf[o] * (1.0f - fx) + sdf[o + ox] * fx;
    const float c10 = sdf[o + oy] * (1.0f - fx) + sdf[o + oy + ox] * fx;
    const float c01 = sdf[o + oz] * (1.0f - fx) + sdf[o + oz + ox] * fx;
    const float c11 = sdf[o + oz + oy] * (1.0f - fx) + sdf[o + oz + oy + ox] * fx;
    const float c0 = c00 * (1.0f - fy) + c10 * fy, c1 = c01 * (1.0f - fy) + c11 * fy;
    return c0 * (1.0f - fz) + c1 * fz;
}

// Fraction of a triangle inside the body (phi < 0) for the signed distances
// a, b, c at its corners, phi linear over the triangle.
FS_HD float tri_inside(float a, float b, float c) {
    const int n = (a < 0.0f) + (b < 0.0f) + (c < 0.0f);
    if (n == 0) return 0.0f;
    if (n == 3) return 1.0f;
    // order so that the corner that differs is first
    if (n == 1) {
        const float p = a < 0.0f ? a : (b < 0.0f ? b : c);
        const float q = a < 0.0f ? b : (b < 0.0f ? c : a);
        const float r = a < 0.0f ? c : (b < 0.0f ? a : b);
        return (p / (p - q)) * (p / (p - r));
    }
    const float p = a >= 0.0f ? a : (b >= 0.0f ? b : c);
    const float q = a >= 0.0f ? b : (b >= 0.0f ? c : a);
    const float r = a >= 0.0f ? c : (b >= 0.0f ? a : b);
    return 1.0f - (p / (p - q)) * (p / (p - r));
}

// Open (fluid) fraction of a square face from the distances at its corners
// (p00, p10, p01, p11 in the face's two axes): the square as two triangles
// (00, 10, 11) and (00, 11, 01), phi linear on each; below FACE_MIN_OPEN the
// face counts as closed.
constexpr float FACE_MIN_OPEN = 1.0e-3f;
FS_HD float face_open(float p00, float p10, float p01, float p11) {
    const float inside = 0.5f * (tri_inside(p00, p10, p11) + tri_inside(p00, p11, p01));
    const float open = 1.0f - inside;
    return open < FACE_MIN_OPEN ? 0.0f : open;
}

// What the face-fraction kernels need of the body.  active = 0 makes every
// face between two FLUID/AIR cells fully open (the reference's operator).
// far2: faces whose centre lies farther than sqrt(far2) from body.bcen have
// all four corners outside the SDF box (body_sphere_has's radius plus one
// cell), so they are open without a lookup.  th[c]: the open fractions of
// the faces of component c in the box of cells [lo, lo + n) (the faces of
// those cells, (n0 + (c == 0)) x (n1 + (c == 1)) x (n2 + (c == 2)) of them,
// i-fastest), computed once per body position by face_open_sdf; the box
// holds every face within sqrt(far2) of bcen.
struct FaceFrac {
    int active;
    float dx, far2;
    const float* sdf;
    BodyParams body;
    int lo[3], n[3];
    const float* th[3];
};

// Open fraction of the face of component comp (0: u at x = fi dx, 1: v at
// y = fj dx, 2: w at z = fk dx) from the body's SDF at its four corners.
// Only faces between two FLUID/AIR cells are asked; wall faces are closed by
// the callers' label tests.
FS_DEV float face_open_sdf(const FaceFrac& f, int comp, int fi, int fj, int fk) {
    if (!f.active) return 1.0f;
    const float dx = f.dx;
    const float x0 = (float)fi * dx, y0 = (float)fj * dx, z0 = (float)fk * dx;
    const float x1 = (float)(fi + 1) * dx, y1 = (float)(fj + 1) * dx, z1 = (float)(fk + 1) * dx;
    const f3 c = comp == 0 ? mk(x0, (fj + 0.5f) * dx, (fk + 0.5f) * dx)
               : comp == 1 ? mk((fi + 0.5f) * dx, y0, (fk + 0.5f) * dx)
                           : mk((fi + 0.5f) * dx, (fj + 0.5f) * dx, z0);
    const float ex = c.x - f.body.bcen.x, ey = c.y - f.body.bcen.y, ez = c.z - f.body.bcen.z;
    if (ex * ex + ey * ey + ez * ez > f.far2) return 1.0f;
    f3 p00, p10, p01, p11;
    if (comp == 0) {
        p00 = mk(x0, y0, z0); p10 = mk(x0, y1, z0); p01 = mk(x0, y0, z1); p11 = mk(x0, y1, z1);
    } else if (comp == 1) {
        p00 = mk(x0, y0, z0); p10 = mk(x1, y0, z0); p01 = mk(x0, y0, z1); p11 = mk(x1, y0, z1);
    } else {
        p00 = mk(x0, y0, z0); p10 = mk(x1, y0, z0); p01 = mk(x0, y1, z0); p11 = mk(x1, y1, z0);
    }
    return face_open(body_phi(f.body, f.sdf, p00), body_phi(f.body, f.sdf, p10), body_phi(f.body, f.sdf, p01),
                     body_phi(f.body, f.sdf, p11));
}

// The same fraction read from the box table th (1 outside the box)
FS_DEV float face_open_box(const FaceFrac& f, int comp, int fi, int fj, int fk) {
    if (!f.active) return 1.0f;
    const int nx = f.n[0] + (comp == 0), ny = f.n[1] + (comp == 1), nz = f.n[2] + (comp == 2);
    const unsigned a = (unsigned)(fi - f.lo[0]), b = (unsigned)(fj - f.lo[1]), c = (unsigned)(fk - f.lo[2]);
    if (a >= (unsigned)nx || b >= (unsigned)ny || c >= (unsigned)nz) return 1.0f;
    return f.th[comp][a + (size_t)nx * (b + (size_t)ny * c)];
}

// Rigid velocity V + Omega x r at a face (FluidSolver3D.cu:1386-1393, :2125-2135).
FS_DEV f3 body_velocity(const BodyParams& b, f3 pos) {
    f3 r = mk(pos.x - b.pos.x, pos.y - b.pos.y, pos.z - b.pos.z);
    return mk(b.vel.x + b.omega.y * r.z - b.omega.z * r.y, b.vel.y + b.omega.z * r.x - b.omega.x * r.z,
              b.vel.z + b.omega.x * r.y - b.omega.y * r.x);
}

// 64-bit SplitMix finaliser used by the counter-based particle seeding.
FS_HD uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

// Wave-level helpers (64-wide wavefronts).
FS_DEV int lane_id() { return __lane_id(); }

template <typename T>
FS_DEV T wave_sum(T v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block sum of a double, result valid in thread 0.  `sm` needs blockDim/64 slots.
FS_DEV double block_sum(double v, double* sm) {
    v = wave_sum(v);
    int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sm[w] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0) {
        int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; ++i) t += sm[i];
    }
    __syncthreads();
    return t;
}

// (i, j, k) of the linear index c of a W x H x D grid, with 32-bit division
// (every grid of the library has fewer than 2^31 cells or faces; 64-bit
// division is a long instruction sequence on the GPU)
FS_DEV void split3(long long c, int W, int H, int& i, int& j, int& k) {
    const unsigned u = (unsigned)c, t = u / (unsigned)W;
    i = (int)(u - t * (unsigned)W);
    j = (int)(t % (unsigned)H);
    k = (int)(t / (unsigned)H);
}

// Remap a linear block id so that consecutive tiles land on the same XCD
// (block ids are dealt round-robin over the 8 XCDs of MI355X).
FS_DEV int xcd_remap(int bid, int nblocks) {
    const int nx = 8;
    int per = nblocks / nx;
    if (per == 0 || bid >= per * nx) return bid;
    return (bid % nx) * per + bid / nx;
}

}  // namespace fsd

namespace fsd {

// Corner reads clamped into the array; the reference reads these corners
// unguarded (FluidSolver3D.cu:851-858), which stays in range whenever positions
// are clamped to (N-1)dx as advection does (:1298-1303).
FS_DEV float ld_clamped(const float* __restrict__ f, int n, int idx) {
    idx = idx < 0 ? 0 : (idx >= n ? n - 1 : idx);
    return f[idx];
}

FS_DEV float sample_comp(const float* __restrict__ f, int n, const Base& b, int sx, int sxy) {
    int i0 = b.i + b.j * sx + b.k * sxy;
    return trilin(b.fx, b.fy, b.fz, ld_clamped(f, n, i0), ld_clamped(f, n, i0 + 1), ld_clamped(f, n, i0 + sxy),
                  ld_clamped(f, n, i0 + 1 + sxy), ld_clamped(f, n, i0 + sx), ld_clamped(f, n, i0 + 1 + sx),
                  ld_clamped(f, n, i0 + sx + sxy), ld_clamped(f, n, i0 + 1 + sx + sxy));
}

// interpVelDevice3D FluidSolver3D.cu:1051-1157
FS_DEV f3 interp_vel3(const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w, int nu,
                      int nv, int nw, int W, int H, int D, float dx, f3 pos) {
    float rx = pos.x / dx, ry = pos.y / dx, rz = pos.z / dx;
    Base bu = mac_base<0>(rx, ry, rz, W, H, D);
    Base bv = mac_base<1>(rx, ry, rz, W, H, D);
    Base bw = mac_base<2>(rx, ry, rz, W, H, D);
    return mk(sample_comp(u, nu, bu, W + 1, (W + 1) * H), sample_comp(v, nv, bv, W, W * (H + 1)),
              sample_comp(w, nw, bw, W, W * H));
}

}  // namespace fsd
