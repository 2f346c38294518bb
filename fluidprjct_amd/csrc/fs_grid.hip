// fs_grid.hip -- grid-side kernels of the PIC/FLIP step (gfx950).
//
//   k_rhs           RHSCalculator3D        FluidSolver3D.cu:1360-1474 (scale :1520)
//   k_project       U/V/WFunctor           FluidSolver3D.cu:2072-2336 (scale :2341)
//   k_drag_lift     DragLiftForceCalculator FluidSolver3D.cu:2674-2791
//   k_buoyancy      applyBuoyancy           FluidSolver3D.cu:2917-3026 (moved to the GPU)
//
// Label reads that the reference performs past a row/slab end (:1406-1455,
// :2109-2110) use the same wrapped linear index; an index outside the array
// reads SOLID with p = 0.
#include <algorithm>
#include <cmath>

#include "fs_kernels.hpp"

using namespace fsd;

namespace fsk {

FS_DEV int label_at(const uint8_t* __restrict__ L, long long idx, long long C) {
    return (idx < 0 || idx >= C) ? FS_SOLID : (int)L[idx];
}
FS_DEV float p_at(const float* __restrict__ p, long long idx, long long C) {
    return (idx < 0 || idx >= C) ? 0.0f : p[idx];
}

// The RHS and projection kernels walk columns: thread (i, j) of a 64 x 4
// block visits COL_KC cells along z and carries the label (and pressure) of
// the cell below from the previous step.  No index division, and the loads
// of one step are independent of each other.  Per cell and face the
// operations and their order are those of one thread per element.
constexpr int COL_TX = 64, COL_TY = 4;

FS_DEV bool is_wall(int L) { return L == FS_SOLID || L == FS_BODY; }

// FF: faces between two FLUID/AIR cells that the body's SDF covers in part
// (open fraction th < 1) carry th u + (1 - th) u_body into the divergence;
// wall faces are those of the reference (th = 0).
template <bool FF>
__global__ __launch_bounds__(256) void k_rhs(GridArgs a, FaceFrac ff, float* __restrict__ rhs, int kc) {
    const int W = a.W, H = a.H, D = a.D;
    const int i = blockIdx.x * COL_TX + (int)(threadIdx.x % COL_TX);
    const int j = blockIdx.y * COL_TY + (int)(threadIdx.x / COL_TX);
    if (i >= W || j >= H) return;
    const int k0 = blockIdx.z * kc, k1 = min(k0 + kc, D);
    const long long WH = (long long)W * H, C = WH * D;
    const float dx = a.dx;
    const float scale = (a.density * dx) / a.dt;
    const float *u = a.u, *v = a.v, *w = a.w;
    long long idx = i + (long long)j * W + (long long)k0 * WH;
    int Lm = label_at(a.labels, idx - WH, C);   // cell (i, j, k - 1)
    int Lc = label_at(a.labels, idx, C);
    for (int k = k0; k < k1; ++k, idx += WH) {
        const int Lp = label_at(a.labels, idx + WH, C);
        if (Lc != FS_FLUID) {
            rhs[idx] = 0.0f;
        } else {
            size_t ui0 = (size_t)i + (size_t)j * (W + 1) + (size_t)k * (W + 1) * H;
            size_t vi0 = (size_t)i + (size_t)j * W + (size_t)k * W * (H + 1);
            size_t wi0 = (size_t)idx;
            float div = u[ui0 + 1] - u[ui0] + v[vi0 + W] - v[vi0] + w[wi0 + (size_t)WH] - w[wi0];
            float r = -scale * div;
            if (i - 1 >= 0) {
                int L = label_at(a.labels, idx - 1, C);
                if (is_wall(L)) {
                    float s = (L == FS_BODY) ? body_velocity(a.body, mk(i * dx, (j + 0.5f) * dx, (k + 0.5f) * dx)).x : 0.0f;
                    r -= scale * (u[ui0] - s);
                } else if (FF) {
                    const float th = face_open_box(ff, 0, i, j, k);
                    if (th < 1.0f) {
                        float s = body_velocity(a.body, mk(i * dx, (j + 0.5f) * dx, (k + 0.5f) * dx)).x;
                        r -= scale * ((1.0f - th) * (u[ui0] - s));
                    }
                }
            }
            {
                int L = label_at(a.labels, idx + 1, C);
                if (is_wall(L)) {
                    float s = (L == FS_BODY) ? body_velocity(a.body, mk((i + 1) * dx, (j + 0.5f) * dx, (k + 0.5f) * dx)).x : 0.0f;
                    r += scale * (u[ui0 + 1] - s);
                } else if (FF) {
                    const float th = face_open_box(ff, 0, i + 1, j, k);
                    if (th < 1.0f) {
                        float s = body_velocity(a.body, mk((i + 1) * dx, (j + 0.5f) * dx, (k + 0.5f) * dx)).x;
                        r += scale * ((1.0f - th) * (u[ui0 + 1] - s));
                    }
                }
            }
            if (j - 1 >= 0) {
                int L = label_at(a.labels, idx - W, C);
                if (is_wall(L)) {
                    float s = (L == FS_BODY) ? body_velocity(a.body, mk((i + 0.5f) * dx, j * dx, (k + 0.5f) * dx)).y : 0.0f;
                    r -= scale * (v[vi0] - s);
                } else if (FF) {
                    const float th = face_open_box(ff, 1, i, j, k);
                    if (th < 1.0f) {
                        float s = body_velocity(a.body, mk((i + 0.5f) * dx, j * dx, (k + 0.5f) * dx)).y;
                        r -= scale * ((1.0f - th) * (v[vi0] - s));
                    }
                }
            }
            {
                int L = label_at(a.labels, idx + W, C);
                if (is_wall(L)) {
                    float s = (L == FS_BODY) ? body_velocity(a.body, mk((i + 0.5f) * dx, (j + 1) * dx, (k + 0.5f) * dx)).y : 0.0f;
                    r += scale * (v[vi0 + W] - s);
                } else if (FF) {
                    const float th = face_open_box(ff, 1, i, j + 1, k);
                    if (th < 1.0f) {
                        float s = body_velocity(a.body, mk((i + 0.5f) * dx, (j + 1) * dx, (k + 0.5f) * dx)).y;
                        r += scale * ((1.0f - th) * (v[vi0 + W] - s));
                    }
                }
            }
            if (k - 1 >= 0 && is_wall(Lm)) {
                float s = (Lm == FS_BODY) ? body_velocity(a.body, mk((i + 0.5f) * dx, (j + 0.5f) * dx, k * dx)).z : 0.0f;
                r -= scale * (w[wi0] - s);
            } else if (FF && k - 1 >= 0) {
                const float th = face_open_box(ff, 2, i, j, k);
                if (th < 1.0f) {
                    float s = body_velocity(a.body, mk((i + 0.5f) * dx, (j + 0.5f) * dx, k * dx)).z;
                    r -= scale * ((1.0f - th) * (w[wi0] - s));
                }
            }
            if (is_wall(Lp)) {
                float s = (Lp == FS_BODY) ? body_velocity(a.body, mk((i + 0.5f) * dx, (j + 0.5f) * dx, (k + 1) * dx)).z : 0.0f;
                r += scale * (w[wi0 + (size_t)WH] - s);
            } else if (FF) {
                const float th = face_open_box(ff, 2, i, j, k + 1);
                if (th < 1.0f) {
                    float s = body_velocity(a.body, mk((i + 0.5f) * dx, (j + 0.5f) * dx, (k + 1) * dx)).z;
                    r += scale * ((1.0f - th) * (w[wi0 + (size_t)WH] - s));
                }
            }
            rhs[idx] = r;
        }
        Lm = Lc;
        Lc = Lp;
    }
}

// One face of component COMP at face index (fi, fj, fk) between the cells
// with labels Ll, Lr and pressures pl, pr.
// FF: a face between two FLUID/AIR cells that the body's SDF closes
// completely (open fraction 0) takes the body's velocity, as a BODY wall face
// on its plus side would.
template <int COMP, bool FF>
FS_DEV void project_face(const GridArgs& a, const FaceFrac& ff, float* __restrict__ G, size_t idx, int Ll, float pl,
                         int Lr, float pr, int fi, int fj, int fk, float scale, float invScale) {
    const float dx = a.dx;
    const float gv = G[idx];
    float p0, p1;
    if ((Ll == FS_FLUID || Ll == FS_AIR) && (Lr == FS_FLUID || Lr == FS_AIR)) {
        p0 = pl;
        p1 = pr;
        if (FF && face_open_box(ff, COMP, fi, fj, fk) == 0.0f) {
            const f3 fp = COMP == 0 ? mk(fi * dx, (fj + 0.5f) * dx, (fk + 0.5f) * dx)
                        : COMP == 1 ? mk((fi + 0.5f) * dx, fj * dx, (fk + 0.5f) * dx)
                                    : mk((fi + 0.5f) * dx, (fj + 0.5f) * dx, fk * dx);
            const f3 vb = body_velocity(a.body, fp);
            const float sol = COMP == 0 ? vb.x : COMP == 1 ? vb.y : vb.z;
            p1 = pl + invScale * (gv - sol);
        }
    } else if (is_wall(Ll) || is_wall(Lr)) {
        const int Lw = is_wall(Ll) ? Ll : Lr;
        float sol = 0.0f;
        if (Lw == FS_BODY) {
            const f3 fp = COMP == 0 ? mk(fi * dx, (fj + 0.5f) * dx, (fk + 0.5f) * dx)
                        : COMP == 1 ? mk((fi + 0.5f) * dx, fj * dx, (fk + 0.5f) * dx)
                                    : mk((fi + 0.5f) * dx, (fj + 0.5f) * dx, fk * dx);
            const f3 vb = body_velocity(a.body, fp);
            sol = COMP == 0 ? vb.x : COMP == 1 ? vb.y : vb.z;
        }
        if (is_wall(Ll)) {
            p0 = pr - invScale * (gv - sol);
            p1 = pr;
        } else {
            p0 = pl;
            p1 = pl + invScale * (gv - sol);
        }
    } else {
        p0 = pl;
        p1 = pr;
    }
    G[idx] = gv - scale * (p1 - p0);
}

// Faces (i, j, k) of u, v and w for every cell, plus the faces i = W, j = H
// and k = D by the threads of the last column, row and plane.  The cells on
// either side of a face are at linear indices c - 1 | c, c - W | c and
// c - WH | c (wrapped past a row or slab end as the reference indexes them).
template <bool FF>
__global__ __launch_bounds__(256) void k_project(GridArgs a, FaceFrac ff, const float* __restrict__ p, int kc) {
    const int W = a.W, H = a.H, D = a.D;
    const int i = blockIdx.x * COL_TX + (int)(threadIdx.x % COL_TX);
    const int j = blockIdx.y * COL_TY + (int)(threadIdx.x / COL_TX);
    if (i >= W || j >= H) return;
    const int k0 = blockIdx.z * kc, k1 = min(k0 + kc, D);
    const long long WH = (long long)W * H, C = WH * D;
    const float scale = a.dt / (a.density * a.dx);
    const float invScale = 1.0f / scale;
    long long c = i + (long long)j * W + (long long)k0 * WH;
    int Lb = label_at(a.labels, c - WH, C);
    float pb = p_at(p, c - WH, C);
    for (int k = k0; k < k1; ++k, c += WH) {
        const int Lc = label_at(a.labels, c, C), Lx = label_at(a.labels, c - 1, C), Ly = label_at(a.labels, c - W, C);
        const float pc = p_at(p, c, C), px = p_at(p, c - 1, C), py = p_at(p, c - W, C);
        const size_t ui = (size_t)i + (size_t)j * (W + 1) + (size_t)k * (W + 1) * H;
        const size_t vi = (size_t)i + (size_t)j * W + (size_t)k * W * (H + 1);
        project_face<0, FF>(a, ff, a.u, ui, Lx, px, Lc, pc, i, j, k, scale, invScale);
        project_face<1, FF>(a, ff, a.v, vi, Ly, py, Lc, pc, i, j, k, scale, invScale);
        project_face<2, FF>(a, ff, a.w, (size_t)c, Lb, pb, Lc, pc, i, j, k, scale, invScale);
        if (i == W - 1)
            project_face<0, FF>(a, ff, a.u, ui + 1, Lc, pc, label_at(a.labels, c + 1, C), p_at(p, c + 1, C), i + 1, j, k,
                            scale, invScale);
        if (j == H - 1)
            project_face<1, FF>(a, ff, a.v, vi + W, Lc, pc, label_at(a.labels, c + W, C), p_at(p, c + W, C), i, j + 1, k,
                            scale, invScale);
        if (k == D - 1)
            project_face<2, FF>(a, ff, a.w, (size_t)(c + WH), Lc, pc, label_at(a.labels, c + WH, C), p_at(p, c + WH, C), i,
                            j, k + 1, scale, invScale);
        Lb = Lc;
        pb = pc;
    }
}


// Drag and lift on BODY-FLUID faces.  The lift term is skipped when
// |n x u_rel| == 0, where the reference divides by zero (:2774-2775).
// Over the box of cells [lo, lo + n) that holds every BODY cell (the body's
// bounding sphere): each block writes the sums of its cells, component q at
// partial[q * gridDim.x + block], for a fixed-order final sum (k_body_sum).
__global__ __launch_bounds__(256) void k_drag_lift(GridArgs a, int3 lo, int3 n, double* __restrict__ partial) {
    __shared__ double sm[4];
    const int W = a.W, H = a.H, D = a.D;
    const long long nbox = (long long)n.x * n.y * n.z;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    int ci = 0, cj = 0, ck = 0;
    if (t < nbox) {
        split3(t, n.x, n.y, ci, cj, ck);
        ci += lo.x; cj += lo.y; ck += lo.z;
    }
    const bool body = t < nbox && a.labels[(size_t)ci + (size_t)cj * W + (size_t)ck * W * H] == FS_BODY;
    // blocks without a body cell skip their six block reductions
    if (!__syncthreads_or(body)) {
        if (threadIdx.x < 6) partial[threadIdx.x * gridDim.x + blockIdx.x] = 0.0;
        return;
    }
    if (body) {
        const float dx = a.dx, rho = a.density;
        f3 cc = mk((ci + 0.5f) * dx, (cj + 0.5f) * dx, (ck + 0.5f) * dx);
        const float area = dx * dx;
        const int NUc = (W + 1) * H * D, NVc = W * (H + 1) * D, NWc = W * H * (D + 1);
        for (int dir = 0; dir < 6; ++dir) {
            int ax = dir >> 1;
            int sgn = (dir & 1) ? -1 : 1;
            int ni = ci + (ax == 0 ? sgn : 0), nj = cj + (ax == 1 ? sgn : 0), nk = ck + (ax == 2 ? sgn : 0);
            if (ni < 0 || ni >= W || nj < 0 || nj >= H || nk < 0 || nk >= D) continue;
            if (a.labels[(size_t)ni + (size_t)nj * W + (size_t)nk * W * H] != FS_FLUID) continue;
            f3 n = mk(ax == 0 ? (float)sgn : 0.0f, ax == 1 ? (float)sgn : 0.0f, ax == 2 ? (float)sgn : 0.0f);
            f3 fo = mk(ax == 0 ? 0.5f * sgn : 0.0f, ax == 1 ? 0.5f * sgn : 0.0f, ax == 2 ? 0.5f * sgn : 0.0f);
            f3 fc = cc + fo * dx;
            f3 r = fc - a.body.pos;
            f3 up = interp_vel3(a.u, a.v, a.w, NUc, NVc, NWc, W, H, D, dx, fc);
            f3 urel = a.body.vel - up;
            float ura = sqrtf(dot(urel, urel));
            const float alpha_s = 0.5f;
            float Aeff = 0.0f;
            float nur = dot(n, urel);
            if (nur > 0) Aeff = (nur / ura * alpha_s + (1 - alpha_s)) * area;
            const float Cd = 0.1f, Cl = 0.1f;
            float sd = (float)(-0.5 * (double)rho * (double)Cd * (double)Aeff * (double)ura);
            f3 Fd = sd * urel;
            f3 Td = cross(r, Fd);
            f3 nxu = cross(n, urel);
            float nxa = sqrtf(dot(nxu, nxu));
            f3 Fl = mk(0, 0, 0), Tl = mk(0, 0, 0);
            if (nxa != 0.0f) {
                Fl = -0.5f * rho * Cl * Aeff * ura * cross(urel, nxu / nxa);
                Tl = cross(r, Fl);
            }
            f3 Ft = Fd + Fl, Tt = Td + Tl;
            acc[0] += Ft.x; acc[1] += Ft.y; acc[2] += Ft.z;
            acc[3] += Tt.x; acc[4] += Tt.y; acc[5] += Tt.z;
        }
    }
    for (int q = 0; q < 6; ++q) {
        double s = block_sum(acc[q], sm);
        if (threadIdx.x == 0) partial[q * gridDim.x + blockIdx.x] = s;
    }
}

// Submerged SDF volume: a voxel with phi <= 0 counts (partially when phi > -cs)
// if its cell is FLUID or a particle binned into that cell lies within 2dx.
// partial[block]: the block's submerged volume (k_body_sum adds them in order)
__global__ __launch_bounds__(256) void k_buoyancy(GridArgs a, const float* __restrict__ sdf,
                                                  const int* __restrict__ start, ParticlesView P,
                                                  double* __restrict__ partial) {
    __shared__ double sm[4];
    const BodyParams& b = a.body;
    long long nvox = (long long)b.sw * b.sh * b.sd;
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double vol = 0.0;
    if (q < nvox) {
        float sv = sdf[q];
        if (sv <= 0.0f) {
            int i, j, k;
            split3(q, b.sw, b.sh, i, j, k);
            const float cs = b.cs;
            f3 pl = b.file_origin + mk(i * cs, j * cs, k * cs);
            f3 pc = pl - b.local_com;
            f3 pr = mk(b.R[0] * pc.x + b.R[1] * pc.y + b.R[2] * pc.z, b.R[3] * pc.x + b.R[4] * pc.y + b.R[5] * pc.z,
                       b.R[6] * pc.x + b.R[7] * pc.y + b.R[8] * pc.z);
            f3 vc = b.pos + pr;
            float v = cs * cs * cs;
            if (sv > -cs) v *= 0.5f - 0.5f * sv / cs;
            const float dx = a.dx;
            int ii = min(max((int)(vc.x / dx), 0), a.W - 1);
            int jj = min(max((int)(vc.y / dx), 0), a.H - 1);
            int kk = min(max((int)(vc.z / dx), 0), a.D - 1);
            size_t gi = (size_t)ii + (size_t)jj * a.W + (size_t)kk * a.W * a.H;
            bool subm = a.labels[gi] == FS_FLUID;
            if (!subm) {
                const float search = 2.0f * dx;
                for (int e = start[gi]; e < start[gi + 1]; ++e) {
                    float ddx = P.px[e] - vc.x, ddy = P.py[e] - vc.y, ddz = P.pz[e] - vc.z;
                    if (sqrtf(ddx * ddx + ddy * ddy + ddz * ddz) < search) { subm = true; break; }
                }
            }
            if (subm) vol = v;
        }
    }
    double s = block_sum(vol, sm);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// out[0..5] = drag/lift force and torque, out[6] = submerged volume, out[7] = 0:
// the per-block partial sums added in a fixed order (one block), so that the
// body's forces are the same bits from run to run
__global__ __launch_bounds__(256) void k_body_sum(const double* __restrict__ pdrag, int nd,
                                                  const double* __restrict__ pbuoy, int nbu, double* __restrict__ out) {
    __shared__ double sm[4];
    double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int nmax = nd > nbu ? nd : nbu;
    for (int i = threadIdx.x; i < nmax; i += blockDim.x) {
        if (i < nd)
#pragma unroll
            for (int q = 0; q < 6; ++q) v[q] += pdrag[(size_t)q * nd + i];
        if (i < nbu) v[6] += pbuoy[i];
    }
    for (int q = 0; q < 7; ++q) {
        const double s = block_sum(v[q], sm);
        if (threadIdx.x == 0) out[q] = s;
    }
    if (threadIdx.x == 0) out[7] = 0.0;
}

// z cells per column thread (4 measured faster than 1, 2, 8 and 16 at 256^3:
// longer columns serialise more dependent loads per thread)
constexpr int COL_KC = 4;

void launch_rhs(const GridArgs& a, const FaceFrac& ff, float* rhs, hipStream_t st) {
    dim3 grid((a.W + COL_TX - 1) / COL_TX, (a.H + COL_TY - 1) / COL_TY, (a.D + COL_KC - 1) / COL_KC);
    if (ff.active)
        hipLaunchKernelGGL(k_rhs<true>, grid, dim3(256), 0, st, a, ff, rhs, COL_KC);
    else
        hipLaunchKernelGGL(k_rhs<false>, grid, dim3(256), 0, st, a, ff, rhs, COL_KC);
}

void launch_project(const GridArgs& a, const FaceFrac& ff, const float* p, hipStream_t st) {
    dim3 grid((a.W + COL_TX - 1) / COL_TX, (a.H + COL_TY - 1) / COL_TY, (a.D + COL_KC - 1) / COL_KC);
    if (ff.active)
        hipLaunchKernelGGL(k_project<true>, grid, dim3(256), 0, st, a, ff, p, COL_KC);
    else
        hipLaunchKernelGGL(k_project<false>, grid, dim3(256), 0, st, a, ff, p, COL_KC);
}

// Open fractions of every face of the three components (1 between FLUID/AIR
// cells away from the body, 0 where a wall cell or the domain boundary
// closes the face), for read-back.
__global__ __launch_bounds__(256) void k_face_open(GridArgs a, FaceFrac ff, int comp, float* __restrict__ out,
                                                   long long n) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int W = a.W, H = a.H, D = a.D;
    const int FW = W + (comp == 0), FH = H + (comp == 1);
    const int fi = (int)(q % FW), fj = (int)((q / FW) % FH), fk = (int)(q / ((long long)FW * FH));
    int li = fi, lj = fj, lk = fk;   // cell on the minus side
    if (comp == 0) li -= 1; else if (comp == 1) lj -= 1; else lk -= 1;
    float th = 0.0f;
    if (li >= 0 && lj >= 0 && lk >= 0 && fi < W && fj < H && fk < D) {
        const long long WH = (long long)W * H;
        const int Ll = a.labels[li + (long long)lj * W + lk * WH], Lr = a.labels[fi + (long long)fj * W + fk * WH];
        if ((Ll == FS_FLUID || Ll == FS_AIR) && (Lr == FS_FLUID || Lr == FS_AIR)) th = face_open_box(ff, comp, fi, fj, fk);
    }
    out[q] = th;
}

// the box table of face_open_box: face (a, b, c) of component comp in the box
__global__ __launch_bounds__(256) void k_face_theta(FaceFrac ff, int comp, float* __restrict__ th, long long n) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int nx = ff.n[0] + (comp == 0), ny = ff.n[1] + (comp == 1);
    const int a = (int)(q % nx), b = (int)((q / nx) % ny), c = (int)(q / ((long long)nx * ny));
    th[q] = face_open_sdf(ff, comp, ff.lo[0] + a, ff.lo[1] + b, ff.lo[2] + c);
}

void launch_face_theta(const FaceFrac& ff, float* const th[3], hipStream_t st) {
    for (int comp = 0; comp < 3; ++comp) {
        const long long n = (long long)(ff.n[0] + (comp == 0)) * (ff.n[1] + (comp == 1)) * (ff.n[2] + (comp == 2));
        if (n > 0)
            hipLaunchKernelGGL(k_face_theta, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ff, comp, th[comp], n);
    }
}

void launch_face_open(const GridArgs& a, const FaceFrac& ff, int comp, float* out, hipStream_t st) {
    const long long n = (long long)(a.W + (comp == 0)) * (a.H + (comp == 1)) * (a.D + (comp == 2));
    hipLaunchKernelGGL(k_face_open, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, ff, comp, out, n);
}

// the cells whose centres lie in the bounding sphere of the pose that marked the
// BODY cells (the label pass marks only cells there), one cell of margin,
// clamped to the grid; n = 0 when none; the whole grid without a pose
static void body_box(const GridArgs& a, const BodyParams* pose, int3& lo, int3& n) {
    if (!pose) {
        lo = make_int3(0, 0, 0);
        n = make_int3(a.W, a.H, a.D);
        return;
    }
    const float r = sqrtf(pose->brad2);
    const float c[3] = {pose->bcen.x, pose->bcen.y, pose->bcen.z};
    const int dim[3] = {a.W, a.H, a.D};
    int l[3], m[3];
    for (int q = 0; q < 3; ++q) {
        // centre (i + 1/2) dx within [c - r, c + r]
        const int i0 = (int)std::floor((c[q] - r) / a.dx - 0.5f) - 1, i1 = (int)std::ceil((c[q] + r) / a.dx - 0.5f) + 1;
        l[q] = std::max(i0, 0);
        m[q] = std::max(0, std::min(i1, dim[q] - 1) - l[q] + 1);
    }
    lo = make_int3(l[0], l[1], l[2]);
    n = make_int3(m[0], m[1], m[2]);
    if (!(r >= 0.0f) || m[0] == 0 || m[1] == 0 || m[2] == 0) n = make_int3(0, 0, 0);
}

long long body_force_partials(const GridArgs& a, const BodyParams* pose) {
    int3 lo, n;
    body_box(a, pose, lo, n);
    const long long nb = ((long long)n.x * n.y * n.z + 255) / 256;
    const long long nv = (long long)a.body.sw * a.body.sh * a.body.sd;
    return 6 * nb + (nv + 255) / 256;
}

void launch_body_forces(const GridArgs& a, const BodyParams* pose, const float* sdf, const int* start,
                        const ParticlesView& P, double* partial, double* out, hipStream_t st) {
    int3 lo, n;
    body_box(a, pose, lo, n);
    const unsigned nd = (unsigned)(((long long)n.x * n.y * n.z + 255) / 256);
    const long long nv = (long long)a.body.sw * a.body.sh * a.body.sd;
    const unsigned nbu = (unsigned)((nv + 255) / 256);
    if (nd) hipLaunchKernelGGL(k_drag_lift, dim3(nd), dim3(256), 0, st, a, lo, n, partial);
    if (nbu) hipLaunchKernelGGL(k_buoyancy, dim3(nbu), dim3(256), 0, st, a, sdf, start, P, partial + 6 * (size_t)nd);
    hipLaunchKernelGGL(k_body_sum, dim3(1), dim3(256), 0, st, (const double*)partial, (int)nd,
                       (const double*)(partial + 6 * (size_t)nd), (int)nbu, out);
}

}  // namespace fsk

namespace fsk {

// MarkBodyCellsFunctor FluidSolver3D.cu:45-78 (scene set-up: BODY where the SDF
// is <= 0 at the cell centre, unless SOLID)
__global__ __launch_bounds__(256) void k_mark_body(LabelArgs la, uint8_t* __restrict__ labels, long long C) {
    long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    int i, j, k;
    split3(c, la.W, la.H, i, j, k);
    f3 cc = mk((i + 0.5f) * la.dx, (j + 0.5f) * la.dx, (k + 0.5f) * la.dx);
    if (labels[c] != FS_SOLID && body_contains(la.body, la.sdf, cc)) labels[c] = FS_BODY;
}

// per-particle cell index of the current positions (read-back only)
__global__ __launch_bounds__(256) void k_cell_keys(ParticlesView P, int n, float dx, int W, int H, int D,
                                                   int* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = cell_key(P.px[i], P.py[i], P.pz[i], dx, W, H, D);
}

void launch_mark_body(const LabelArgs& la, uint8_t* labels, long long C, hipStream_t st) {
    hipLaunchKernelGGL(k_mark_body, dim3((unsigned)((C + 255) / 256)), dim3(256), 0, st, la, labels, C);
}

void launch_cell_keys(const ParticlesView& P, int n, float dx, int W, int H, int D, int* out, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cell_keys, dim3((n + 255) / 256), dim3(256), 0, st, P, n, dx, W, H, D, out);
}

}  // namespace fsk
