// fs_kernels.hpp -- argument blocks and host launchers of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fs_device.hpp"

// G2P/advection LDS tile (cells) and staged margin; one 256-thread block per
// tile.  Stage ms at 256^3 / in a developed 128^3 dam break: 8x4x8 4.64 /
// 0.28, 16x4x4 4.68 / 0.31, 8x8x4 4.69 / 0.28, 8x4x4 4.77 / 0.25, 32x2x2
// 5.02 / 0.34, 32x4x4 5.08, 64x2x1 5.54.  An 8x4x8 tile stages 4.9 faces per
// cell and component (tile + margin over tile), 32x2x2 10.4.  A 1-cell margin
// is faster at rest but slower once particles cross tiles.
#ifndef ADV_TX
#define ADV_TX 8
#endif
#ifndef ADV_TY
#define ADV_TY 4
#endif
#ifndef ADV_TZ
#define ADV_TZ 8
#endif
#ifndef ADV_MARGIN
#define ADV_MARGIN 2
#endif

// P2G LDS tile (cells); one 256-thread block per tile.  (TX + 2) x (TY + 2)
// = 16 x 16 halo cells per z-plane, one per thread of the window phase, each
// halo row one 16-lane DPP row; CAP particles of a plane staged in LDS per
// batch (24 bytes each; with the window exchange two blocks fit a CU).  Stage
// ms at 256^3, 8 ppc at rest: 14x14x16 CAP 2048 0.85, CAP 2560 0.85; with the
// exchange of whole windows (no lane sums) 18x10x16 CAP 1920 0.99, 18x10x32
// 0.99, 18x10x8 1.08, 18x10x16 CAP 1280 1.31.
#ifndef P2G_TX
#define P2G_TX 14
#endif
#ifndef P2G_TY
#define P2G_TY 14
#endif
#ifndef P2G_TZ
#define P2G_TZ 16
#endif
#ifndef P2G_CAP
#define P2G_CAP 2048
#endif

namespace fsk {

struct ParticlesView {
    float* px;
    float* py;
    float* pz;
    float* vx;
    float* vy;
    float* vz;
    int* id;
};

struct LabelArgs {
    int W, H;
    float dx;
    fsd::BodyParams body;
    const float* sdf;
};

struct P2GArgs {
    ParticlesView P;
    const int* start;
    int W, H, D;
    float dx, dt, gravity_y;
    float inv_dx;     // 1/dx when dx is a power of two, else 0
    float *u, *v, *w, *us, *vs, *ws;
};

struct G2PArgs {
    ParticlesView P;
    int n;
    int W, H, D;
    float dx, dt, alpha, C;
    float inv_dx;     // 1/dx when dx is a power of two (then x * inv_dx == x / dx exactly), else 0
    const float *u, *v, *w, *us, *vs, *ws;
    int nu, nv, nw;
    const uint8_t* labels;
    int* next_count;  // null: do not compute next-step keys
    int* next_key;
    int* next_rank;
    unsigned* umax;   // null, or max |v| over the new particle velocities (float bits)
};

struct GridArgs {
    int W, H, D;
    float dx, dt, density;
    const uint8_t* labels;
    float *u, *v, *w;
    fsd::BodyParams body;
};

// particles
void launch_key_count(const ParticlesView& P, int n, float dx, int W, int H, int D, int* key, int* rank, int* count,
                      hipStream_t st);
int scan_blocks(long long C);
void launch_scan_label(int* count, const uint8_t* solid, uint8_t* labels, long long C, int* bsum, int* total,
                       int* start, const LabelArgs& la, hipStream_t st);
void launch_scan_sort(int* count, long long C, int* bsum, int* total, int* start, hipStream_t st);
void launch_scan_fluid(const uint8_t* labels, long long C, int* bsum, int* total, int* start, hipStream_t st);
void launch_seed(const uint8_t* labels, const int* frank, long long C, int W, int H, float dx, int ppc, uint64_t seed,
                 const ParticlesView& P, hipStream_t st);
void launch_scatter(const ParticlesView& src, const ParticlesView& dst, int n, const int* key, const int* rank,
                    const int* start, int* dkey, hipStream_t st);
// particles of each cell segment of src to their rank by id in dst (dkey: cell of each slot)
void launch_canonical(const ParticlesView& src, const ParticlesView& dst, int n, const int* dkey, const int* start,
                      hipStream_t st);
void launch_p2g(const P2GArgs& a, hipStream_t st);
void launch_g2p_advect(const G2PArgs& a, const int* start, hipStream_t st);
void launch_max_speed(const ParticlesView& P, int n, unsigned* umax, hipStream_t st);

// grid stages
void launch_rhs(const GridArgs& a, const fsd::FaceFrac& ff, float* rhs, hipStream_t st);
void launch_project(const GridArgs& a, const fsd::FaceFrac& ff, const float* p, hipStream_t st);
void launch_face_open(const GridArgs& a, const fsd::FaceFrac& ff, int comp, float* out, hipStream_t st);
void launch_face_theta(const fsd::FaceFrac& ff, float* const th[3], hipStream_t st);
// drag/lift + buoyancy: out[0..2] force, [3..5] torque, [6] submerged volume (FP64), from
// per-block partial sums (body_force_partials doubles of scratch) added in a fixed order.
// pose: the pose whose label pass marked the BODY cells (drag/lift visits the cells of its
// bounding sphere), null to visit every cell (labels from elsewhere)
long long body_force_partials(const GridArgs& a, const fsd::BodyParams* pose);
void launch_body_forces(const GridArgs& a, const fsd::BodyParams* pose, const float* sdf, const int* start,
                        const ParticlesView& P, double* partial, double* out, hipStream_t st);
void launch_mark_body(const LabelArgs& la, uint8_t* labels, long long C, hipStream_t st);
void launch_cell_keys(const ParticlesView& P, int n, float dx, int W, int H, int D, int* out, hipStream_t st);

}  // namespace fsk
