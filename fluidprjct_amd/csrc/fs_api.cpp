// fs_api.cpp -- the fs_* C ABI (include/fluidsim.h) over the gfx950 kernels.
//
// One handle owns one HIP stream and every device buffer of the step; nothing
// is allocated per step.  Host work is limited to what the reference also does
// on the host: scene/config/SDF parsing, the rigid-body integrator
// (RigidBody::integrate utility.cuh:408-467) and file output.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fluidsim.h"
#include "../../include/fluidsim_checkpoint.h"
#include "../../include/fluidsim_learned.h"
#include "../../include/fluidsim_noalloc.h"
#include "fs_kernels.hpp"
#include "fs_learned.hpp"
#include "fs_noalloc.hpp"
#include "fs_pressure.hpp"

using fsd::f3;
using fsd::mk;
using fsd::FaceFrac;
using fsd::BodyParams;
using namespace fsk;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& m) {
    g_err = m;
    return code;
}

// Runs the body of a C entry point: exceptions (a failed host allocation for
// sizes read from a file or passed in, a parse error) become error codes and
// never cross the C ABI.
template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return fail(FS_ERR_ARG, "host allocation failed (sizes out of range?)");
    } catch (const std::exception& e) {
        return fail(FS_ERR_IO, std::string("unexpected error: ") + e.what());
    } catch (...) {
        return fail(FS_ERR_IO, "unexpected error");
    }
}

#define HIPCHK(expr)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(FS_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));      \
    } while (0)

template <typename T>
void dfree(T*& p) {
    if (p) hipFree((void*)p);
    p = nullptr;
}

struct HostBody {
    bool active = false;
    std::vector<float> sdf;
    int sw = 0, sh = 0, sd = 0;
    f3 file_origin{0, 0, 0}, sdf_origin{0, 0, 0}, local_com{0, 0, 0}, size{0, 0, 0};
    float cs = 0;
    f3 pos{0, 0, 0}, vel{0, 0, 0}, force{0, 0, 0}, omega{0, 0, 0}, torque{0, 0, 0};
    float mass = 0, inertia = 0, inv_inertia = 0, theta = 0, psi = 0, phi = 0;
    float R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    f3 last_force{0, 0, 0}, last_torque{0, 0, 0};
    float last_sub = 0;
    // FS_STEP_BODY_INERTIA: body-frame inertia tensor about local_com, its
    // inverse, and the orientation quaternion (w, x, y, z), body to world
    double Ib[9] = {0}, Ibinv[9] = {0};
    double q[4] = {1, 0, 0, 0};

    // RigidBody::update_rotation_matrix utility.cuh:470-484
    void update_rotation() {
        double th = theta, ps = psi, ph = phi;
        R[0] = (float)(std::cos(th) * std::cos(ps));
        R[1] = (float)(std::sin(th) * std::sin(ph) - std::cos(th) * std::cos(ph) * std::sin(ps));
        R[2] = (float)(std::cos(ph) * std::sin(th) + std::cos(th) * std::sin(ph) * std::sin(ps));
        R[3] = (float)std::sin(ps);
        R[4] = (float)(std::cos(ph) * std::cos(ps));
        R[5] = (float)(-std::cos(ps) * std::sin(ph));
        R[6] = (float)(-std::cos(ps) * std::sin(th));
        R[7] = (float)(std::cos(th) * std::sin(ph) + std::cos(ph) * std::sin(th) * std::sin(ps));
        R[8] = (float)(std::cos(th) * std::cos(ph) - std::sin(th) * std::sin(ph) * std::sin(ps));
    }

    // RigidBody::integrate utility.cuh:408-467 (semi-implicit Euler + Euler-angle rates)
    void integrate(float dt) {
        f3 accel = force / mass;
        vel = vel + accel * dt;
        pos = pos + vel * dt;
        const float maxAngle = 3.14159265358979323846f;
        theta = std::clamp(theta, -maxAngle, maxAngle);
        update_rotation();
        f3 tl = mk(R[0] * torque.x + R[3] * torque.y + R[6] * torque.z, R[1] * torque.x + R[4] * torque.y + R[7] * torque.z,
                   R[2] * torque.x + R[5] * torque.y + R[8] * torque.z);
        f3 aa = inv_inertia * tl;
        omega = omega + aa * dt;
        float dth = (float)(omega.y * std::cos((double)phi) - omega.z * std::sin((double)phi));
        float dps = (float)((omega.y * std::sin((double)phi) + omega.z * std::cos((double)phi)) / std::cos((double)theta));
        float dph = (float)(omega.x + dps * std::sin((double)theta));
        theta += dth * dt;
        psi += dps * dt;
        phi += dph * dt;
        f3 lo = file_origin;
        f3 ro = mk(R[0] * lo.x + R[1] * lo.y + R[2] * lo.z, R[3] * lo.x + R[4] * lo.y + R[5] * lo.z,
                   R[6] * lo.x + R[7] * lo.y + R[8] * lo.z);
        sdf_origin = pos + ro;
        force = mk(0, 0, 0);
        torque = mk(0, 0, 0);
    }

    // ---- FS_STEP_BODY_INERTIA (no reference counterpart) -----------------
    static void quat_mul(const double* a, const double* b, double* o) {
        o[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
        o[1] = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
        o[2] = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
        o[3] = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
    }
    void rotation_of_quaternion() {
        const double w = q[0], x = q[1], y = q[2], z = q[3];
        R[0] = (float)(1.0 - 2.0 * (y * y + z * z));
        R[1] = (float)(2.0 * (x * y - w * z));
        R[2] = (float)(2.0 * (x * z + w * y));
        R[3] = (float)(2.0 * (x * y + w * z));
        R[4] = (float)(1.0 - 2.0 * (x * x + z * z));
        R[5] = (float)(2.0 * (y * z - w * x));
        R[6] = (float)(2.0 * (x * z - w * y));
        R[7] = (float)(2.0 * (y * z + w * x));
        R[8] = (float)(1.0 - 2.0 * (x * x + y * y));
    }
    // theta, psi, phi of R = Ry(theta) Rz(psi) Rx(phi), the reference's convention
    void angles_of_rotation() {
        psi = (float)std::asin(std::clamp((double)R[3], -1.0, 1.0));
        theta = (float)std::atan2(-(double)R[6], (double)R[0]);
        phi = (float)std::atan2(-(double)R[5], (double)R[4]);
    }
    void quaternion_of_angles() {
        const double a = 0.5 * theta, b = 0.5 * psi, c = 0.5 * phi;
        const double qy[4] = {std::cos(a), 0, std::sin(a), 0}, qz[4] = {std::cos(b), 0, 0, std::sin(b)},
                     qx[4] = {std::cos(c), std::sin(c), 0, 0};
        double t[4];
        quat_mul(qy, qz, t);
        quat_mul(t, qx, q);
    }
    // second moments of the SDF's inside voxels (partial weights as the
    // buoyancy's, FluidSolver3D.cu:2979-2983) about local_com, scaled to the
    // mass; inverse by cofactors
    void inertia_tensor() {
        double wsum = 0, A[9] = {0};
        for (int k = 0; k < sd; ++k)
            for (int j = 0; j < sh; ++j)
                for (int i = 0; i < sw; ++i) {
                    const float v = sdf[(size_t)i + (size_t)j * sw + (size_t)k * sw * sh];
                    if (v > 0.0f) continue;
                    double wv = 1.0;
                    if (v > -cs) wv = 0.5 - 0.5 * (double)v / cs;
                    const double rx = file_origin.x + (double)i * cs - local_com.x;
                    const double ry = file_origin.y + (double)j * cs - local_com.y;
                    const double rz = file_origin.z + (double)k * cs - local_com.z;
                    const double r2 = rx * rx + ry * ry + rz * rz;
                    A[0] += wv * (r2 - rx * rx); A[1] -= wv * rx * ry; A[2] -= wv * rx * rz;
                    A[4] += wv * (r2 - ry * ry); A[5] -= wv * ry * rz;
                    A[8] += wv * (r2 - rz * rz);
                    wsum += wv;
                }
        A[3] = A[1]; A[6] = A[2]; A[7] = A[5];
        for (int e = 0; e < 9; ++e) Ib[e] = wsum > 0 ? (double)mass * A[e] / wsum : 0.0;
        const double* M = Ib;
        const double c0 = M[4] * M[8] - M[5] * M[7], c1 = M[5] * M[6] - M[3] * M[8], c2 = M[3] * M[7] - M[4] * M[6];
        const double det = M[0] * c0 + M[1] * c1 + M[2] * c2;
        if (!(std::fabs(det) > 0)) {
            for (double& e : Ibinv) e = 0;
            return;
        }
        Ibinv[0] = c0 / det;
        Ibinv[1] = (M[2] * M[7] - M[1] * M[8]) / det;
        Ibinv[2] = (M[1] * M[5] - M[2] * M[4]) / det;
        Ibinv[3] = c1 / det;
        Ibinv[4] = (M[0] * M[8] - M[2] * M[6]) / det;
        Ibinv[5] = (M[2] * M[3] - M[0] * M[5]) / det;
        Ibinv[6] = c2 / det;
        Ibinv[7] = (M[1] * M[6] - M[0] * M[7]) / det;
        Ibinv[8] = (M[0] * M[4] - M[1] * M[3]) / det;
    }
    // translation as integrate(); rotation by Euler's equations in the body
    // frame (omega held in the world frame), quaternion step then normalised
    void integrate_tensor(float dt) {
        f3 accel = force / mass;
        vel = vel + accel * dt;
        pos = pos + vel * dt;
        rotation_of_quaternion();
        double Rm[9];
        for (int e = 0; e < 9; ++e) Rm[e] = R[e];
        const double tw[3] = {torque.x, torque.y, torque.z}, ow[3] = {omega.x, omega.y, omega.z};
        double tb[3], wb[3], Iw[3], g[3], d[3];
        for (int a = 0; a < 3; ++a) {
            tb[a] = Rm[a] * tw[0] + Rm[3 + a] * tw[1] + Rm[6 + a] * tw[2];
            wb[a] = Rm[a] * ow[0] + Rm[3 + a] * ow[1] + Rm[6 + a] * ow[2];
        }
        for (int a = 0; a < 3; ++a) Iw[a] = Ib[3 * a] * wb[0] + Ib[3 * a + 1] * wb[1] + Ib[3 * a + 2] * wb[2];
        g[0] = wb[1] * Iw[2] - wb[2] * Iw[1];
        g[1] = wb[2] * Iw[0] - wb[0] * Iw[2];
        g[2] = wb[0] * Iw[1] - wb[1] * Iw[0];
        for (int a = 0; a < 3; ++a) d[a] = tb[a] - g[a];
        for (int a = 0; a < 3; ++a)
            wb[a] += (double)dt * (Ibinv[3 * a] * d[0] + Ibinv[3 * a + 1] * d[1] + Ibinv[3 * a + 2] * d[2]);
        const double wq[4] = {0, wb[0], wb[1], wb[2]};
        double dq[4];
        quat_mul(q, wq, dq);
        for (int e = 0; e < 4; ++e) q[e] += 0.5 * (double)dt * dq[e];
        const double qn = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        for (double& e : q) e /= qn;
        omega = mk((float)(Rm[0] * wb[0] + Rm[1] * wb[1] + Rm[2] * wb[2]),
                   (float)(Rm[3] * wb[0] + Rm[4] * wb[1] + Rm[5] * wb[2]),
                   (float)(Rm[6] * wb[0] + Rm[7] * wb[1] + Rm[8] * wb[2]));
        rotation_of_quaternion();
        angles_of_rotation();
        f3 lo = file_origin;
        f3 ro = mk(R[0] * lo.x + R[1] * lo.y + R[2] * lo.z, R[3] * lo.x + R[4] * lo.y + R[5] * lo.z,
                   R[6] * lo.x + R[7] * lo.y + R[8] * lo.z);
        sdf_origin = pos + ro;
        force = mk(0, 0, 0);
        torque = mk(0, 0, 0);
    }
    // FluidSolver3D::handleBodyWallCollisions FluidSolver3D.cu:2848-2895 (off
    // in the reference, :2841); its assert dx == SDF cell size is not enforced
    void collide_walls(int W, int H, int D, float dx) {
        const float r = 0.5f * std::fmax(std::fmax(size.x, size.y), size.z);
        const float half = 0.5f * dx;
        const f3 lo = mk(half, half, half), hi = mk(W * dx - half, H * dx - half, D * dx - half);
        auto axis = [&](float& p, float pMin, float pMax, f3 n) {
            const float distMin = p - pMin, distMax = pMax - p;
            const float dist = std::fmin(distMin, distMax);
            const float pen = r - dist;
            if (pen > 0.0f) {
                p = p + pen * (distMin < distMax ? +1.0f : -1.0f);
                const float vn = vel.x * n.x + vel.y * n.y + vel.z * n.z;
                if (vn < 0.0f) vel = vel - vn * n;
            }
        };
        axis(pos.x, lo.x + r, hi.x - r, mk(1, 0, 0));
        axis(pos.y, lo.y + r, hi.y - r, mk(0, 1, 0));
        axis(pos.z, lo.z + r, hi.z - r, mk(0, 0, 1));
        sdf_origin = pos - size * 0.5f;
    }

    fsd::BodyParams params() const {
        fsd::BodyParams b{};
        b.active = active ? 1 : 0;
        b.sw = sw; b.sh = sh; b.sd = sd;
        b.cs = cs;
        b.pos = pos; b.vel = vel; b.omega = omega;
        b.file_origin = file_origin;
        b.local_com = local_com;
        for (int i = 0; i < 9; ++i) b.R[i] = R[i];
        // body_contains maps a world point w to local Rt (w - pos) + local_com
        // and accepts it inside the SDF box; R is a rotation, so the box's
        // circumsphere maps to a world sphere of the same radius
        const double hx = 0.5 * cs * sw, hy = 0.5 * cs * sh, hz = 0.5 * cs * sd;
        const double lx = file_origin.x + hx - local_com.x, ly = file_origin.y + hy - local_com.y,
                     lz = file_origin.z + hz - local_com.z;
        b.bcen = mk((float)(pos.x + R[0] * lx + R[1] * ly + R[2] * lz),
                    (float)(pos.y + R[3] * lx + R[4] * ly + R[5] * lz),
                    (float)(pos.z + R[6] * lx + R[7] * ly + R[8] * lz));
        const double r = std::sqrt(hx * hx + hy * hy + hz * hz) + 2.0 * cs;
        b.brad2 = (float)(r * r * 1.0001);
        return b;
    }
};

// 1/x when x is a positive normal power of two whose inverse is also normal
// (then v * (1/x) and v / x are the same bits for every v), else 0
float pow2_inverse(float x) {
    if (!(x > 0.0f) || !std::isnormal(x)) return 0.0f;
    int e = 0;
    const float m = std::frexp(x, &e);
    if (m != 0.5f) return 0.0f;
    const float inv = std::ldexp(1.0f, 1 - e);
    return std::isnormal(inv) ? inv : 0.0f;
}

struct Solver {
    fs_config cfg{};
    int W = 0, H = 0, D = 0;
    long long C = 0;
    hipStream_t st = nullptr;
    // stage events of two steps (8 each): a step's times are collected once
    // its events completed, during the next step or when the stream drains
    hipEvent_t ev[16] = {};
    int ev_set = 0;             // the set the current step records into
    int times_pending = -1;     // the set whose times are not collected yet
    // grids
    uint8_t *d_solid = nullptr, *d_labels = nullptr;
    float *d_u = nullptr, *d_v = nullptr, *d_w = nullptr, *d_us = nullptr, *d_vs = nullptr, *d_ws = nullptr;
    float *d_p = nullptr, *d_rhs = nullptr;
    double* d_p64 = nullptr;
    int *d_count = nullptr, *d_start = nullptr, *d_bsum = nullptr, *d_total = nullptr;
    // particles
    long long np = 0;
    ParticlesView PA{}, PB{};
    int *d_key = nullptr, *d_rank = nullptr, *d_tmpkey = nullptr;
    int* d_dkey = nullptr;      // cell of each sorted slot (deterministic order only)
    bool keys_valid = false;   // d_key/d_rank/d_count hold the counting pass of the current order
    bool sorted_valid = false; // d_start matches the current particle order
    bool count_clean = true;   // d_count is all zero
    // body
    HostBody body;
    float* d_sdf = nullptr;
    double* d_bacc = nullptr;
    double* h_bacc = nullptr;   // pinned copy of d_bacc, read after body_done
    double* d_bpart = nullptr;  // per-block partial sums of the body forces
    long long bpart_cap = 0;
    // the pose of the label pass that marked the current BODY cells (false:
    // labels written from elsewhere, BODY cells anywhere)
    bool body_box_pose = false;
    BodyParams box_pose{};
    // labels written from the host: BODY cells anywhere, a queued bounding box stale
    void labels_from_outside() {
        body_box_pose = false;
        mg.forget_bbox();
    }
    hipEvent_t body_done = nullptr;
    MGPCG mg;
    std::unique_ptr<Preconditioner> jacobi;
    std::unique_ptr<LNet> net;             // learned preconditioner (created on first use)
    std::unique_ptr<LearnedPre> learned;
    std::unique_ptr<HybridPre> hybrid;
    fs_solve_stats stats{};
    double times[FS_NUM_STAGES] = {0};
    float dt_cur = 0.0f;        // dt of the current step
    double sim_time = 0.0;
    int64_t nsteps = 0;
    int64_t nonconv = 0;        // steps whose pressure solve missed the tolerance
    double nonconv_res = 0.0;
    bool p64_from_step = false; // d_p64 holds the last step's pressure (the next solve's initial guess)
    unsigned* d_umax = nullptr; // max |v| of the particle velocities (float bits)
    bool umax_valid = false;
    unsigned* h_umax = nullptr; // pinned copy of d_umax
    hipEvent_t umax_ev = nullptr;
    bool umax_copied = false;   // h_umax receives d_umax of the last G2P launch (umax_ev after the copy)
    // face fractions: box tables of the current body pose (FaceFrac::th)
    float* d_theta[3] = {nullptr, nullptr, nullptr};
    long long theta_cap[3] = {0, 0, 0};
    bool theta_valid = false;   // cleared whenever the SDF is replaced
    FaceFrac theta_of{};        // the pose, box and SDF the tables were computed for

    long long NU() const { return (long long)(W + 1) * H * D; }
    long long NV() const { return (long long)W * (H + 1) * D; }
    long long NW() const { return (long long)W * H * (D + 1); }

    ~Solver() {
        free_grids();
        free_particles();
        dfree(d_sdf);
        dfree(d_bacc);
        dfree(d_bpart);
        dfree(d_umax);
        for (float*& t : d_theta) dfree(t);
        if (h_bacc) hipHostFree(h_bacc);
        if (h_umax) hipHostFree(h_umax);
        if (umax_ev) hipEventDestroy(umax_ev);
        if (body_done) hipEventDestroy(body_done);
        for (auto& e : ev) if (e) hipEventDestroy(e);
        if (st) hipStreamDestroy(st);
    }
    void free_grids() {
        dfree(d_solid); dfree(d_labels); dfree(d_u); dfree(d_v); dfree(d_w); dfree(d_us); dfree(d_vs); dfree(d_ws);
        dfree(d_p); dfree(d_rhs); dfree(d_p64); dfree(d_count); dfree(d_start); dfree(d_bsum); dfree(d_total);
    }
    static void free_view(ParticlesView& P) {
        dfree(P.px); dfree(P.py); dfree(P.pz); dfree(P.vx); dfree(P.vy); dfree(P.vz); dfree(P.id);
    }
    void free_particles() {
        free_view(PA);
        free_view(PB);
        dfree(d_key); dfree(d_rank); dfree(d_tmpkey); dfree(d_dkey);
        np = 0;
    }

    int alloc_grids(int W_, int H_, int D_) {
        // the kernels index cells and faces with 32-bit arithmetic
        const long long faces = std::max({(long long)(W_ + 1) * H_ * D_, (long long)W_ * (H_ + 1) * D_,
                                          (long long)W_ * H_ * (D_ + 1)});
        if (W_ <= 0 || H_ <= 0 || D_ <= 0 || faces >= (1LL << 31))
            return fail(FS_ERR_ARG, "grid dimensions out of range (at most 2^31 - 1 faces per component)");
        free_grids();
        W = W_; H = H_; D = D_;
        C = (long long)W * H * D;
        HIPCHK(hipMalloc(&d_solid, C));
        HIPCHK(hipMalloc(&d_labels, C));
        for (float** p : {&d_u, &d_us}) HIPCHK(hipMalloc(p, sizeof(float) * NU()));
        for (float** p : {&d_v, &d_vs}) HIPCHK(hipMalloc(p, sizeof(float) * NV()));
        for (float** p : {&d_w, &d_ws}) HIPCHK(hipMalloc(p, sizeof(float) * NW()));
        HIPCHK(hipMalloc(&d_p, sizeof(float) * C));
        HIPCHK(hipMalloc(&d_rhs, sizeof(float) * C));
        HIPCHK(hipMalloc(&d_p64, sizeof(double) * C));
        HIPCHK(hipMalloc(&d_count, sizeof(int) * (C + 4)));
        HIPCHK(hipMalloc(&d_start, sizeof(int) * (C + 1)));
        HIPCHK(hipMalloc(&d_bsum, sizeof(int) * (scan_blocks(C) + 1)));
        HIPCHK(hipMalloc(&d_total, sizeof(int) * 2));
        HIPCHK(hipMemset(d_count, 0, sizeof(int) * (C + 4)));
        count_clean = true;
        HIPCHK(hipMemset(d_start, 0, sizeof(int) * (C + 1)));
        for (float* p : {d_u, d_us}) HIPCHK(hipMemset(p, 0, sizeof(float) * NU()));
        for (float* p : {d_v, d_vs}) HIPCHK(hipMemset(p, 0, sizeof(float) * NV()));
        for (float* p : {d_w, d_ws}) HIPCHK(hipMemset(p, 0, sizeof(float) * NW()));
        HIPCHK(hipMemset(d_p, 0, sizeof(float) * C));
        HIPCHK(hipMemset(d_rhs, 0, sizeof(float) * C));
        HIPCHK(hipMemset(d_p64, 0, sizeof(double) * C));
        if (mg.init(W, H, D) != 0) return fail(FS_ERR_DEVICE, "pressure solver allocation failed");
        p64_from_step = false;
        jacobi.reset(make_jacobi_preconditioner(&mg));
        keys_valid = sorted_valid = false;
        return FS_OK;
    }

    static int alloc_view(ParticlesView& P, long long n) {
        size_t b = sizeof(float) * (size_t)std::max<long long>(n, 1);
        for (float** p : {&P.px, &P.py, &P.pz, &P.vx, &P.vy, &P.vz})
            if (hipMalloc(p, b) != hipSuccess) return fail(FS_ERR_DEVICE, "particle allocation failed");
        if (hipMalloc(&P.id, sizeof(int) * (size_t)std::max<long long>(n, 1)) != hipSuccess)
            return fail(FS_ERR_DEVICE, "particle allocation failed");
        return FS_OK;
    }

    int alloc_particles(long long n) {
        if (n > 0x7fffffffLL) return fail(FS_ERR_ARG, "particle count exceeds 2^31-1");
        free_particles();
        np = n;
        int rc = alloc_view(PA, n);
        if (rc) return rc;
        rc = alloc_view(PB, n);
        if (rc) return rc;
        size_t b = sizeof(int) * (size_t)std::max<long long>(n, 1);
        HIPCHK(hipMalloc(&d_key, b));
        HIPCHK(hipMalloc(&d_rank, b));
        keys_valid = sorted_valid = umax_valid = false;
        if (C > 0) {
            HIPCHK(hipMemset(d_count, 0, sizeof(int) * (C + 4)));
            count_clean = true;
        }
        return FS_OK;
    }

    // CFL-limited dt (no reference counterpart): dt = min(dt_max, cfl*dx/max|v|)
    int choose_dt() {
        dt_cur = cfg.dt;
        if (!(cfg.cfl > 0.0f) || np == 0) return FS_OK;
        if (!umax_valid) {
            HIPCHK(hipMemsetAsync(d_umax, 0, sizeof(unsigned), st));
            launch_max_speed(PA, (int)np, d_umax, st);
        }
        unsigned bits = 0;
        if (umax_valid && umax_copied) {
            // copied after the last G2P launch; the work queued since then keeps running
            HIPCHK(hipEventSynchronize(umax_ev));
            bits = *h_umax;
        } else {
            HIPCHK(hipMemcpyAsync(h_umax, d_umax, sizeof(unsigned), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            bits = *h_umax;
        }
        float umax;
        std::memcpy(&umax, &bits, sizeof(float));
        if (umax > 0.0f) dt_cur = std::fmin(cfg.dt, (cfg.cfl * cfg.dx) / umax);
        return FS_OK;
    }

    LabelArgs label_args() {
        LabelArgs la{};
        la.W = W; la.H = H;
        la.dx = cfg.dx;
        la.body = body.params();
        la.sdf = d_sdf;
        return la;
    }

    // The body's face fractions (FS_STEP_FACE_FRACTIONS with a body loaded),
    // with the box table of the current pose (computed when the pose, the box
    // or the SDF changed); inactive otherwise: the reference's operator.
    int face_frac(FaceFrac& f) {
        f = FaceFrac{};
        f.body = body.params();
        if (!((cfg.step_flags & FS_STEP_FACE_FRACTIONS) && body.active && d_sdf)) return FS_OK;
        f.active = 1;
        f.dx = cfg.dx;
        f.sdf = d_sdf;
        const double far = std::sqrt((double)f.body.brad2) + (double)cfg.dx;
        f.far2 = (float)(far * far * 1.0001);
        // every face whose centre lies within far of bcen, with a cell of margin
        const double c[3] = {f.body.bcen.x, f.body.bcen.y, f.body.bcen.z};
        const int dims[3] = {W, H, D};
        for (int q = 0; q < 3; ++q) {
            const double a = std::floor((c[q] - far) / cfg.dx) - 1.0, b = std::ceil((c[q] + far) / cfg.dx) + 1.0;
            const int lo = (int)std::max(0.0, std::min(a, (double)dims[q]));
            const int hi = (int)std::max((double)lo, std::min(b, (double)dims[q]));
            f.lo[q] = lo;
            f.n[q] = hi - lo;
        }
        for (int comp = 0; comp < 3; ++comp) {
            const long long need =
                std::max(1LL, (long long)(f.n[0] + (comp == 0)) * (f.n[1] + (comp == 1)) * (f.n[2] + (comp == 2)));
            if (need > theta_cap[comp]) {
                dfree(d_theta[comp]);
                HIPCHK(hipMalloc(&d_theta[comp], sizeof(float) * need));
                theta_cap[comp] = need;
                theta_valid = false;
            }
            f.th[comp] = d_theta[comp];
        }
        const bool same = theta_valid && std::memcmp(&theta_of.body, &f.body, sizeof(f.body)) == 0 &&
                          theta_of.dx == f.dx && theta_of.far2 == f.far2 && theta_of.sdf == f.sdf &&
                          std::equal(f.lo, f.lo + 3, theta_of.lo) && std::equal(f.n, f.n + 3, theta_of.n);
        if (!same) {
            launch_face_theta(f, d_theta, st);
            theta_of = f;
            theta_valid = true;
        }
        return FS_OK;
    }

    GridArgs grid_args() {
        GridArgs g{};
        g.W = W; g.H = H; g.D = D;
        g.dx = cfg.dx; g.dt = dt_cur; g.density = cfg.density;
        g.labels = d_labels;
        g.u = d_u; g.v = d_v; g.w = d_w;
        g.body = body.params();
        return g;
    }

    // ---- stages -----------------------------------------------------------
    int stage_sort_label() {
        int rc = ensure_keys();
        if (rc) return rc;
        LabelArgs la = label_args();
        mg.forget_bbox();   // the labels change
        launch_scan_label(d_count, d_solid, d_labels, C, d_bsum, d_total, d_start, la, st);
        box_pose = la.body;
        body_box_pose = true;
        return finish_sort();
    }

    // the counting pass of the current particle order (keys, ranks, counts)
    int ensure_keys() {
        if (keys_valid) return FS_OK;
        if (!count_clean) HIPCHK(hipMemsetAsync(d_count, 0, sizeof(int) * (C + 4), st));
        launch_key_count(PA, (int)np, cfg.dx, W, H, D, d_key, d_rank, d_count, st);
        keys_valid = true;
        count_clean = false;
        return FS_OK;
    }

    int finish_sort() {
        if (cfg.step_flags & FS_STEP_DETERMINISTIC) {
            // scatter into PB, then each cell's particles back into PA in id order
            if (!d_dkey) HIPCHK(hipMalloc(&d_dkey, sizeof(int) * (size_t)std::max<long long>(np, 1)));
            launch_scatter(PA, PB, (int)np, d_key, d_rank, d_start, d_dkey, st);
            launch_canonical(PB, PA, (int)np, d_dkey, d_start, st);
        } else {
            launch_scatter(PA, PB, (int)np, d_key, d_rank, d_start, nullptr, st);
            std::swap(PA, PB);
        }
        keys_valid = false;  // counts were consumed (and cleared) by the scan
        count_clean = true;
        sorted_valid = true;
        return FS_OK;
    }

    // sort by cell without relabelling (for a G2P/advection stage called on its own)
    int sort_only() {
        int rc = ensure_keys();
        if (rc) return rc;
        launch_scan_sort(d_count, C, d_bsum, d_total, d_start, st);
        return finish_sort();
    }

    int stage_p2g() {
        if (!sorted_valid) return fail(FS_ERR_STATE, "P2G needs the label/sort stage of this step");
        P2GArgs a{};
        a.P = PA;
        a.start = d_start;
        a.W = W; a.H = H; a.D = D;
        a.dx = cfg.dx; a.dt = dt_cur; a.gravity_y = cfg.gravity_y;
        a.inv_dx = pow2_inverse(cfg.dx);
        a.u = d_u; a.v = d_v; a.w = d_w; a.us = d_us; a.vs = d_vs; a.ws = d_ws;
        launch_p2g(a, st);
        return FS_OK;
    }

    int stage_rhs() {
        FaceFrac ff;
        if (int rc = face_frac(ff)) return rc;
        launch_rhs(grid_args(), ff, d_rhs, st);
        return FS_OK;
    }

    // the network, with random parameters from cfg.seed until others are set
    LNet* get_net() {
        if (!net) {
            std::unique_ptr<LNet> n(new LNet());
            if (n->init_params(cfg.seed, st) != 0) return nullptr;
            net = std::move(n);
        }
        return net.get();
    }

    // warm: start from the previous step's pressure (d_p64) when it is one;
    // faces: the operator of the step (face fractions when the option is on),
    // else the reference's matrix of the labels
    int stage_pressure(bool warm = false, bool faces = true) {
        Preconditioner* pre = nullptr;
        if (cfg.precond == FS_PRECOND_JACOBI) pre = jacobi.get();
        if (cfg.precond == FS_PRECOND_LEARNED) {
            if (!learned) {
                LNet* n = get_net();
                // activations for the whole grid, so that no solve box reallocates them
                if (!n || n->ensure(W, H, D) != 0) return fail(FS_ERR_DEVICE, "learned preconditioner allocation failed");
                learned.reset(new LearnedPre(&mg, n));
            }
            pre = learned.get();
        }
        if (cfg.precond == FS_PRECOND_HYBRID) {
            if (!hybrid) {
                LNet* n = get_net();
                if (!n || n->ensure(W, H, D) != 0) return fail(FS_ERR_DEVICE, "learned preconditioner allocation failed");
                hybrid.reset(new HybridPre(&mg, n));
            }
            pre = hybrid.get();
        }
        mg.stall_window = cfg.stall_window > 0 ? cfg.stall_window : 5;
        mg.use_graph = !(cfg.solver_flags & FS_SOLVER_NO_GRAPH);
        mg.use_tail = !(cfg.solver_flags & FS_SOLVER_NO_TAIL);
        mg.use_tb = !(cfg.solver_flags & FS_SOLVER_NO_FUSE);
        const bool use_x0 = warm && p64_from_step && !(cfg.solver_flags & FS_SOLVER_COLD_START);
        mg.face = FaceFrac{};
        if (faces) {
            if (int rc0 = face_frac(mg.face)) return rc0;
        }
        int rc = mg.solve(d_labels, d_rhs, d_p64, d_p, cfg.solver_tol, cfg.solver_max_iter, pre, st, &stats,
                          use_x0 ? d_p64 : nullptr);
        mg.face.active = 0;
        p64_from_step = warm && rc != -5;
        if (rc == -5) return fail(FS_ERR_DEVICE, std::string("pressure solve: ") + hipGetErrorString(hipGetLastError()));
        if (rc == -3) return fail(FS_ERR_NO_FLUID, "No fluid cells found!");
        if (rc == -4) return fail(FS_ERR_NOT_CONVERGED, "pressure solve did not reach the tolerance");
        return FS_OK;
    }

    int stage_project() {
        FaceFrac ff;
        if (int rc = face_frac(ff)) return rc;
        launch_project(grid_args(), ff, d_p, st);
        return FS_OK;
    }

    // updateBody FluidSolver3D.cu:2793-2846 with drag/lift and buoyancy reduced on the GPU
    // The force sums are copied to pinned memory; the integration waits for
    // that copy only, so the step can launch G2P (which does not read the
    // body) before it.
    int stage_body_launch() {
        if (!body.active) return FS_OK;
        if (!sorted_valid) {
            // buoyancy looks particles up by cell: sort them (labels unchanged)
            int rc = sort_only();
            if (rc) return rc;
        }
        const GridArgs ga = grid_args();
        const BodyParams* pose = body_box_pose ? &box_pose : nullptr;
        const long long need = body_force_partials(ga, pose);
        if (need > bpart_cap) {
            dfree(d_bpart);
            bpart_cap = 0;
            HIPCHK(hipMalloc(&d_bpart, sizeof(double) * (size_t)need));
            bpart_cap = need;
        }
        launch_body_forces(ga, pose, d_sdf, d_start, PA, d_bpart, d_bacc, st);
        HIPCHK(hipMemcpyAsync(h_bacc, d_bacc, sizeof(double) * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipEventRecord(body_done, st));
        return FS_OK;
    }

    int stage_body_finish() {
        if (!body.active) return FS_OK;
        HIPCHK(hipEventSynchronize(body_done));
        const double* acc = h_bacc;
        const f3 G = mk(0.0f, cfg.gravity_y, 0.0f);
        f3 totalF = mk((float)acc[0], (float)acc[1], (float)acc[2]);
        f3 totalT = mk((float)acc[3], (float)acc[4], (float)acc[5]);
        body.force = body.force + body.mass * G;
        body.force = body.force + totalF;
        body.torque = body.torque + totalT;
        float sub = (float)acc[6];
        f3 Farch = (-cfg.density * sub) * G;
        body.force = body.force + Farch;
        body.last_force = body.force;
        body.last_torque = body.torque;
        body.last_sub = sub;
        if (cfg.step_flags & FS_STEP_BODY_INERTIA)
            body.integrate_tensor(dt_cur);
        else
            body.integrate(dt_cur);
        if (cfg.step_flags & FS_STEP_BODY_WALLS) body.collide_walls(W, H, D, cfg.dx);
        return FS_OK;
    }

    int stage_body() {
        int rc = stage_body_launch();
        return rc ? rc : stage_body_finish();
    }

    int stage_g2p_advect() {
        if (!sorted_valid) {
            int rc = sort_only();
            if (rc) return rc;
        }
        G2PArgs a{};
        a.P = PA;
        a.n = (int)np;
        a.W = W; a.H = H; a.D = D;
        a.dx = cfg.dx; a.dt = dt_cur; a.alpha = cfg.pic_weight; a.C = cfg.advect_c;
        a.inv_dx = pow2_inverse(cfg.dx);
        a.u = d_u; a.v = d_v; a.w = d_w; a.us = d_us; a.vs = d_vs; a.ws = d_ws;
        a.nu = (int)NU(); a.nv = (int)NV(); a.nw = (int)NW();
        a.labels = d_labels;
        // the counts of the next label pass are accumulated here
        if (!count_clean) HIPCHK(hipMemsetAsync(d_count, 0, sizeof(int) * (C + 4), st));
        a.next_count = d_count;
        a.next_key = d_key;
        a.next_rank = d_rank;
        HIPCHK(hipMemsetAsync(d_umax, 0, sizeof(unsigned), st));
        a.umax = d_umax;
        launch_g2p_advect(a, d_start, st);
        umax_valid = true;
        umax_copied = false;
        keys_valid = true;
        count_clean = false;
        sorted_valid = false;
        return FS_OK;
    }

    // FS_ERR_NOT_CONVERGED when a solve of the steps since `before` missed the tolerance
    int report_nonconv(int64_t before, int64_t n) const {
        if (nonconv == before) return FS_OK;
        return fail(FS_ERR_NOT_CONVERGED, "pressure solve missed the tolerance in " + std::to_string(nonconv - before) +
                                              " of " + std::to_string(n) + " steps (last relative residual " +
                                              std::to_string(nonconv_res) + ")");
    }

    // max |v| after the G2P launch, to pinned memory: the next step's dt waits
    // for this copy only
    int copy_umax() {
        HIPCHK(hipMemcpyAsync(h_umax, d_umax, sizeof(unsigned), hipMemcpyDeviceToHost, st));
        HIPCHK(hipEventRecord(umax_ev, st));
        umax_copied = true;
        return FS_OK;
    }

    // add the stage times of the last completed step (waits for its last event)
    int collect_times() {
        if (times_pending < 0) return FS_OK;
        hipEvent_t* e = ev + 8 * times_pending;
        HIPCHK(hipEventSynchronize(e[6]));
        float ms;
        const int slot[6] = {FS_STAGE_LABEL, FS_STAGE_P2G, FS_STAGE_PRESSURE, FS_STAGE_PROJECT, FS_STAGE_BODY,
                             FS_STAGE_G2P};
        for (int i = 0; i < 6; ++i) {
            HIPCHK(hipEventElapsedTime(&ms, e[i], e[i + 1]));
            times[slot[i]] += ms;
        }
        times_pending = -1;
        return FS_OK;
    }

    // wait for every queued step and collect its times
    int drain() {
        HIPCHK(hipStreamSynchronize(st));
        return collect_times();
    }

    // One step, launched without draining the stream: the label/sort stage of
    // the step is queued behind the previous step's G2P before the host waits
    // for that G2P's max speed (which sets dt), so the GPU does not idle
    // between steps.  Callers drain() before reading device state.
    int step_once() {
        if (C == 0) return fail(FS_ERR_STATE, "no scene loaded");
        int rc;
        hipEvent_t* e = ev + 8 * ev_set;
        HIPCHK(hipEventRecord(e[0], st));
        if ((rc = stage_sort_label())) return rc;
        // the solve's bounding box of these labels, copied while P2G and RHS run
        if (mg.queue_bbox(d_labels, st) != 0) return fail(FS_ERR_DEVICE, "bounding box launch failed");
        HIPCHK(hipEventRecord(e[1], st));
        if ((rc = choose_dt())) return rc;
        if ((rc = collect_times())) return rc;
        if ((rc = stage_p2g())) return rc;
        HIPCHK(hipEventRecord(e[2], st));
        if ((rc = stage_rhs())) return rc;
        rc = stage_pressure(true);
        // a solve that misses the tolerance still completes the step (the
        // reference ignores the solver status); fs_step reports it afterwards
        if (rc == FS_ERR_NOT_CONVERGED) {
            nonconv += 1;
            nonconv_res = stats.rel_residual;
        } else if (rc && rc != FS_ERR_NO_FLUID) {
            return rc;
        }
        HIPCHK(hipEventRecord(e[3], st));
        if ((rc = stage_project())) return rc;
        HIPCHK(hipEventRecord(e[4], st));
        if ((rc = stage_body_launch())) return rc;
        HIPCHK(hipEventRecord(e[5], st));
        if ((rc = stage_g2p_advect())) return rc;
        HIPCHK(hipEventRecord(e[6], st));
        if ((rc = copy_umax())) return rc;
        if ((rc = stage_body_finish())) return rc;
        times_pending = ev_set;
        ev_set ^= 1;
        sim_time += dt_cur;
        nsteps += 1;
        return FS_OK;
    }
};

// ---------------- host IO --------------------------------------------------
bool json_value(const std::string& t, const char* key, std::string& out) {
    const std::string k = std::string("\"") + key + "\"";
    size_t p = 0;
    while ((p = t.find(k, p)) != std::string::npos) {
        size_t q = p + k.size();
        while (q < t.size() && std::isspace((unsigned char)t[q])) ++q;
        if (q < t.size() && t[q] == ':') {
            ++q;
            while (q < t.size() && std::isspace((unsigned char)t[q])) ++q;
            if (q < t.size() && t[q] == '"') {
                size_t e = t.find('"', q + 1);
                if (e == std::string::npos) return false;
                out = t.substr(q + 1, e - q - 1);
            } else {
                size_t e = q;
                while (e < t.size() && t[e] != ',' && t[e] != '}' && !std::isspace((unsigned char)t[e])) ++e;
                out = t.substr(q, e - q);
            }
            return true;
        }
        p = q;
    }
    return false;
}

int read_label_file(const char* path, int& W, int& H, int& D, std::vector<uint8_t>& L) {
    std::ifstream f(path);
    if (!f.is_open()) return fail(FS_ERR_IO, std::string("Cannot open file: ") + path);
    std::string header;
    if (!std::getline(f, header)) return fail(FS_ERR_IO, std::string("empty label file: ") + path);
    std::istringstream hs(header);
    W = H = 0;
    D = 1;
    hs >> W >> H;
    int d;
    if (hs >> d) D = d;
    if (W <= 0 || H <= 0 || D <= 0) return fail(FS_ERR_IO, std::string("bad label header in ") + path);
    L.assign((size_t)W * H * D, FS_SOLID);
    // characters are assigned in reverse order: the first is cell (W-1,H-1,D-1)
    // (FluidSolver3D.cu:139-152); unknown characters leave SOLID
    for (long long idx = (long long)W * H * D - 1; idx >= 0; --idx) {
        char ch;
        if (!(f >> ch)) break;
        if (ch == 'S') L[(size_t)idx] = FS_SOLID;
        else if (ch == 'F') L[(size_t)idx] = FS_FLUID;
        else if (ch == 'A') L[(size_t)idx] = FS_AIR;
    }
    return FS_OK;
}

int upload_scene(Solver& s, int W, int H, int D, const std::vector<uint8_t>& L) {
    int rc = s.alloc_grids(W, H, D);
    if (rc) return rc;
    std::vector<uint8_t> solid(L.size());
    for (size_t c = 0; c < L.size(); ++c) solid[c] = (L[c] == FS_SOLID);
    HIPCHK(hipMemcpy(s.d_labels, L.data(), L.size(), hipMemcpyHostToDevice));
    s.labels_from_outside();
    HIPCHK(hipMemcpy(s.d_solid, solid.data(), L.size(), hipMemcpyHostToDevice));
    if (s.body.active) {
        launch_mark_body(s.label_args(), s.d_labels, s.C, s.st);
        s.labels_from_outside();
        HIPCHK(hipStreamSynchronize(s.st));
    }
    return FS_OK;
}

// positions in original particle order (D2H of the current order and its ids)
int gather_positions(Solver& s, std::vector<float>& pos) {
    pos.assign((size_t)s.np * 3, 0.0f);
    std::vector<float> px(s.np), py(s.np), pz(s.np);
    std::vector<int> id(s.np);
    if (s.np) {
        HIPCHK(hipStreamSynchronize(s.st));
        HIPCHK(hipMemcpy(px.data(), s.PA.px, sizeof(float) * s.np, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(py.data(), s.PA.py, sizeof(float) * s.np, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(pz.data(), s.PA.pz, sizeof(float) * s.np, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(id.data(), s.PA.id, sizeof(int) * s.np, hipMemcpyDeviceToHost));
    }
    for (long long i = 0; i < s.np; ++i) {
        size_t d = (size_t)id[i];
        pos[3 * d] = px[i]; pos[3 * d + 1] = py[i]; pos[3 * d + 2] = pz[i];
    }
    return FS_OK;
}

// save3dParticlesToPLY / ...ToOFF utility.cu:91-127 (ASCII, default ostream
// precision), or binary little-endian PLY; host only
int format_particle_file(const std::vector<float>& pos, const std::string& path, int fmt, std::string* err) {
    const long long n = (long long)(pos.size() / 3);
    std::ofstream f(path, fmt == FS_FMT_PLY_BINARY ? std::ios::binary : std::ios::out);
    if (!f.is_open()) {
        *err = "Error opening file: " + path;
        return FS_ERR_IO;
    }
    if (fmt == FS_FMT_OFF) {
        f << "OFF\n" << n << " 0 0\n";
    } else {
        f << "ply\nformat " << (fmt == FS_FMT_PLY_BINARY ? "binary_little_endian" : "ascii") << " 1.0\nelement vertex "
          << n << "\nproperty float x\nproperty float y\nproperty float z\nend_header\n";
    }
    if (fmt == FS_FMT_PLY_BINARY) {
        f.write((const char*)pos.data(), (std::streamsize)(sizeof(float) * pos.size()));
    } else {
        for (long long q = 0; q < n; ++q) f << pos[3 * q] << " " << pos[3 * q + 1] << " " << pos[3 * q + 2] << "\n";
    }
    if (!f) {
        *err = "write failed: " + path;
        return FS_ERR_IO;
    }
    return FS_OK;
}

int write_particle_file(Solver& s, const char* path, int fmt) {
    std::vector<float> pos;
    int rc = gather_positions(s, pos);
    if (rc) return rc;
    std::string err;
    rc = format_particle_file(pos, path, fmt, &err);
    return rc ? fail(rc, err) : FS_OK;
}

// RigidBody::generateSurfacePoints + exportToPLY utility.cuh:341-400
void write_body_ply(const Solver& s, const std::string& path) {
    const HostBody& b = s.body;
    std::vector<f3> pts;
    const float density = 0.5f * s.cfg.dx;
    for (int k = 0; k < b.sd; ++k)
        for (int j = 0; j < b.sh; ++j)
            for (int i = 0; i < b.sw; ++i) {
                float v = b.sdf[(size_t)i + (size_t)j * b.sw + (size_t)k * b.sw * b.sh];
                if (!(std::fabs(v) < density)) continue;
                f3 pl = b.file_origin + mk(i * b.cs, j * b.cs, k * b.cs);
                f3 pc = pl - b.local_com;
                f3 pr = mk(b.R[0] * pc.x + b.R[1] * pc.y + b.R[2] * pc.z, b.R[3] * pc.x + b.R[4] * pc.y + b.R[5] * pc.z,
                           b.R[6] * pc.x + b.R[7] * pc.y + b.R[8] * pc.z);
                pts.push_back(b.pos + pr);
            }
    std::ofstream f(path);
    if (!f.is_open()) return;
    f << "ply\nformat ascii 1.0\nelement vertex " << pts.size()
      << "\nproperty float x\nproperty float y\nproperty float z\nend_header\n";
    for (auto& p : pts) f << p.x << " " << p.y << " " << p.z << "\n";
}

}  // namespace

// ============================== C ABI ======================================
extern "C" {

const char* fs_last_error(void) { return g_err.c_str(); }
const char* fs_backend(void) { return "hip"; }

void fs_config_default(fs_config* c) {
    std::memset(c, 0, sizeof(*c));
    c->dt = 0.02f;
    c->dx = 0.1f;
    c->pic_weight = 0.01f;
    c->iter_per_frame = 1;
    c->output_format = FS_FMT_PLY;
    std::snprintf(c->output_template, sizeof(c->output_template), "%s", "InputData/particles_");
    std::snprintf(c->input_label_data, sizeof(c->input_label_data), "%s", "InputData/labels.txt");
    c->frame_amount = 100;
    c->active_body = 0;
    c->body_x = c->body_y = c->body_z = 0.5f;
    c->body_mass = 50.0f;
    std::snprintf(c->body_sdf_path, sizeof(c->body_sdf_path), "%s", "InputData/ball.sdf");
    c->particles_per_cell = 4;
    c->seed = 1;
    c->gravity_y = -9.81f;
    c->density = 1000.0f;
    c->advect_c = 0.1f;
    c->solver_tol = 1e-6;
    c->solver_max_iter = 500;
    c->precond = FS_PRECOND_MG;
    c->stall_window = 5;
    c->cfl = 0.0f;
}

int fs_config_load_json(fs_config* c, const char* path) {
    return guarded([&]() -> int {
        std::ifstream f(path);
        if (!f.is_open()) return fail(FS_ERR_IO, std::string("cannot open ") + path);
        std::stringstream ss;
        ss << f.rdbuf();
        const std::string t = ss.str();
        std::string v;
        try {
            if (json_value(t, "dt", v)) c->dt = std::stof(v);
            if (json_value(t, "dx", v)) c->dx = std::stof(v);
            if (json_value(t, "PIC_WEIGHT", v)) c->pic_weight = std::stof(v);
            if (json_value(t, "iterPerFrame", v)) c->iter_per_frame = std::stoi(v);
            if (json_value(t, "outputFormat", v)) c->output_format = (v == "PLY") ? FS_FMT_PLY : FS_FMT_OFF;
            if (json_value(t, "outputTemplate", v)) std::snprintf(c->output_template, 256, "%s", v.c_str());
            if (json_value(t, "inputLabelData", v)) std::snprintf(c->input_label_data, 256, "%s", v.c_str());
            if (json_value(t, "frameAmount", v)) c->frame_amount = std::stoi(v);
            if (json_value(t, "activeBody", v)) c->active_body = (v == "true" || v == "1");
            if (json_value(t, "body_x", v)) c->body_x = std::stof(v);
            if (json_value(t, "body_y", v)) c->body_y = std::stof(v);
            if (json_value(t, "body_z", v)) c->body_z = std::stof(v);
            if (json_value(t, "body_mass", v)) c->body_mass = std::stof(v);
            if (json_value(t, "bodySdf", v)) std::snprintf(c->body_sdf_path, 256, "%s", v.c_str());
        } catch (const std::exception& e) {
            return fail(FS_ERR_IO, std::string("bad value in ") + path + ": " + e.what());
        }
        return FS_OK;
    });
}

int fs_create(const fs_config* cfg, void** out) {
    return guarded([&]() -> int {
        if (!cfg || !out) return fail(FS_ERR_ARG, "null argument");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(FS_ERR_DEVICE, "no HIP device");
        std::unique_ptr<Solver> s(new Solver());
        s->cfg = *cfg;
        HIPCHK(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
        for (auto& e : s->ev) HIPCHK(hipEventCreate(&e));
        HIPCHK(hipEventCreateWithFlags(&s->umax_ev, hipEventDisableTiming));
        HIPCHK(hipHostMalloc((void**)&s->h_umax, sizeof(unsigned)));
        HIPCHK(hipMalloc(&s->d_bacc, sizeof(double) * 8));
        HIPCHK(hipHostMalloc((void**)&s->h_bacc, sizeof(double) * 8));
        HIPCHK(hipEventCreateWithFlags(&s->body_done, hipEventDisableTiming));
        HIPCHK(hipMalloc(&s->d_umax, sizeof(unsigned)));
        s->dt_cur = cfg->dt;
        *out = s.release();
        return FS_OK;
    });
}

void fs_destroy(void* h) { delete (Solver*)h; }

int fs_load_labels_file(void* h, const char* path) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        int W, H, D;
        std::vector<uint8_t> L;
        int rc = read_label_file(path, W, H, D, L);
        if (rc) return rc;
        return upload_scene(s, W, H, D, L);
    });
}

int fs_set_labels(void* h, const uint8_t* labels, int W, int H, int D) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        if (!labels || W <= 0 || H <= 0 || D <= 0) return fail(FS_ERR_ARG, "bad label grid");
        std::vector<uint8_t> L(labels, labels + (size_t)W * H * D);
        return upload_scene(s, W, H, D, L);
    });
}

int fs_set_level_sets(void* h, const float* ps, const float* pf, int W, int H, int D) {
    return guarded([&]() -> int {
        if (!ps || !pf || W <= 0 || H <= 0 || D <= 0) return fail(FS_ERR_ARG, "bad level sets");
        std::vector<uint8_t> L((size_t)W * H * D);
        for (size_t c = 0; c < L.size(); ++c) L[c] = ps[c] < 0 ? FS_SOLID : (pf[c] < 0 ? FS_FLUID : FS_AIR);
        return fs_set_labels(h, L.data(), W, H, D);
    });
}

int fs_get_dims(void* h, int* W, int* H, int* D) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        *W = s.W; *H = s.H; *D = s.D;
        return FS_OK;
    });
}

int fs_set_body_sdf(void* h, const float* sdf, int w, int hh, int d, float ox, float oy, float oz, float cs, float x,
                    float y, float z, float mass) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        if (!sdf || w <= 0 || hh <= 0 || d <= 0 || !(cs > 0)) return fail(FS_ERR_ARG, "bad SDF grid");
        HostBody& b = s.body;
        b = HostBody();
        b.active = true;
        b.sdf.assign(sdf, sdf + (size_t)w * hh * d);
        b.sw = w; b.sh = hh; b.sd = d;
        b.file_origin = mk(ox, oy, oz);
        b.cs = cs;
        b.size = mk(w * cs, hh * cs, d * cs);
        f3 center = b.file_origin + b.size * 0.5f;
        f3 T = mk(x, y, z) - center;
        b.pos = mk(x, y, z);
        b.sdf_origin = b.file_origin + T;
        b.local_com = center;
        b.mass = mass;
        float radius = b.size.x / 2.0f;
        b.inertia = 0.4f * b.mass * radius * radius;
        b.inv_inertia = 1.0f / b.inertia;
        b.update_rotation();
        b.inertia_tensor();
        dfree(s.d_sdf);
        s.theta_valid = false;
        HIPCHK(hipMalloc(&s.d_sdf, sizeof(float) * b.sdf.size()));
        HIPCHK(hipMemcpy(s.d_sdf, b.sdf.data(), sizeof(float) * b.sdf.size(), hipMemcpyHostToDevice));
        if (s.C > 0) {
            launch_mark_body(s.label_args(), s.d_labels, s.C, s.st);
            s.labels_from_outside();
            HIPCHK(hipStreamSynchronize(s.st));
        }
        return FS_OK;
    });
}

// RigidBody::loadSDF utility.cuh:231-338: "w h d", origin, cell size, then k->j->i values
int fs_load_body_sdf(void* h, const char* path, float x, float y, float z, float mass) {
    return guarded([&]() -> int {
        std::ifstream f(path);
        if (!f.is_open()) return fail(FS_ERR_IO, std::string("Cannot open SDF file: ") + path);
        std::string line;
        int w = 0, hh = 0, d = 0;
        float ox = 0, oy = 0, oz = 0, cs = 0;
        if (!std::getline(f, line)) return fail(FS_ERR_IO, "SDF: missing dimensions");
        { std::istringstream is(line); is >> w >> hh >> d; }
        if (!std::getline(f, line)) return fail(FS_ERR_IO, "SDF: missing origin");
        { std::istringstream is(line); is >> ox >> oy >> oz; }
        if (!std::getline(f, line)) return fail(FS_ERR_IO, "SDF: missing cell size");
        try { cs = std::stof(line); } catch (...) { return fail(FS_ERR_IO, "SDF: bad cell size"); }
        if (w <= 0 || hh <= 0 || d <= 0 || (long long)w * hh * d > (1LL << 31))
            return fail(FS_ERR_IO, "SDF: bad dimensions");
        std::vector<float> data((size_t)w * hh * d);
        for (auto& v : data)
            if (!(f >> v)) return fail(FS_ERR_IO, "Unexpected end of SDF data");
        return fs_set_body_sdf(h, data.data(), w, hh, d, ox, oy, oz, cs, x, y, z, mass);
    });
}

int fs_get_body_state(void* h, fs_body_state* o) {
    return guarded([&]() -> int {
        const HostBody& b = ((Solver*)h)->body;
        std::memset(o, 0, sizeof(*o));
        o->pos[0] = b.pos.x; o->pos[1] = b.pos.y; o->pos[2] = b.pos.z;
        o->vel[0] = b.vel.x; o->vel[1] = b.vel.y; o->vel[2] = b.vel.z;
        o->omega[0] = b.omega.x; o->omega[1] = b.omega.y; o->omega[2] = b.omega.z;
        o->theta = b.theta; o->psi = b.psi; o->phi = b.phi;
        o->force[0] = b.last_force.x; o->force[1] = b.last_force.y; o->force[2] = b.last_force.z;
        o->torque[0] = b.last_torque.x; o->torque[1] = b.last_torque.y; o->torque[2] = b.last_torque.z;
        o->submerged_volume = b.last_sub;
        for (int i = 0; i < 9; ++i) o->rotation[i] = b.R[i];
        return FS_OK;
    });
}

int fs_set_body_state(void* h, const fs_body_state* in) {
    return guarded([&]() -> int {
        HostBody& b = ((Solver*)h)->body;
        b.pos = mk(in->pos[0], in->pos[1], in->pos[2]);
        b.vel = mk(in->vel[0], in->vel[1], in->vel[2]);
        b.omega = mk(in->omega[0], in->omega[1], in->omega[2]);
        b.theta = in->theta; b.psi = in->psi; b.phi = in->phi;
        b.update_rotation();
        b.quaternion_of_angles();
        return FS_OK;
    });
}

int fs_seed_particles(void* h, int ppc, uint64_t seed) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        if (s.C == 0) return fail(FS_ERR_STATE, "no scene loaded");
        if (ppc <= 0 || ppc > 63) return fail(FS_ERR_ARG, "particles per cell must be in [1,63]");
        // fluid-cell ranks by an exclusive scan over (label == FLUID) in k->j->i order
        launch_scan_fluid(s.d_labels, s.C, s.d_bsum, s.d_total, s.d_start, s.st);
        int nfluid = 0;
        HIPCHK(hipMemcpyAsync(&nfluid, s.d_total, sizeof(int), hipMemcpyDeviceToHost, s.st));
        HIPCHK(hipStreamSynchronize(s.st));
        int rc = s.alloc_particles((long long)nfluid * ppc);
        if (rc) return rc;
        launch_seed(s.d_labels, s.d_start, s.C, s.W, s.H, s.cfg.dx, ppc, seed, s.PA, s.st);
        HIPCHK(hipStreamSynchronize(s.st));
        return FS_OK;
    });
}

int fs_set_particles(void* h, const float* pos, const float* vel, int64_t n) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        if (n < 0 || (n > 0 && !pos)) return fail(FS_ERR_ARG, "bad particle arrays");
        int rc = s.alloc_particles(n);
        if (rc) return rc;
        std::vector<float> a((size_t)n);
        std::vector<int> id((size_t)n);
        float* dst[6] = {s.PA.px, s.PA.py, s.PA.pz, s.PA.vx, s.PA.vy, s.PA.vz};
        for (int c = 0; c < 6; ++c) {
            const float* src = c < 3 ? pos : vel;
            for (int64_t i = 0; i < n; ++i) a[(size_t)i] = src ? src[3 * i + (c % 3)] : 0.0f;
            if (n) HIPCHK(hipMemcpy(dst[c], a.data(), sizeof(float) * n, hipMemcpyHostToDevice));
        }
        for (int64_t i = 0; i < n; ++i) id[(size_t)i] = (int)i;
        if (n) HIPCHK(hipMemcpy(s.PA.id, id.data(), sizeof(int) * n, hipMemcpyHostToDevice));
        return FS_OK;
    });
}

int fs_get_particles(void* h, float* pos, float* vel, int64_t cap, int64_t* n) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        if (n) *n = s.np;
        if (cap < s.np) return fail(FS_ERR_ARG, "buffer too small");
        if (s.np == 0) return FS_OK;
        HIPCHK(hipStreamSynchronize(s.st));
        std::vector<float> a((size_t)s.np);
        std::vector<int> id((size_t)s.np);
        HIPCHK(hipMemcpy(id.data(), s.PA.id, sizeof(int) * s.np, hipMemcpyDeviceToHost));
        const float* src[6] = {s.PA.px, s.PA.py, s.PA.pz, s.PA.vx, s.PA.vy, s.PA.vz};
        for (int c = 0; c < 6; ++c) {
            float* out = c < 3 ? pos : vel;
            if (!out) continue;
            HIPCHK(hipMemcpy(a.data(), src[c], sizeof(float) * s.np, hipMemcpyDeviceToHost));
            for (long long i = 0; i < s.np; ++i) out[3 * (size_t)id[i] + (c % 3)] = a[i];
        }
        return FS_OK;
    });
}

int64_t fs_num_particles(void* h) { return ((Solver*)h)->np; }

int fs_step(void* h, int nsteps) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        const int64_t before = s.nonconv;
        for (int i = 0; i < nsteps; ++i) {
            int rc = s.step_once();
            if (rc) {
                s.drain();
                return rc;
            }
        }
        if (int rc = s.drain()) return rc;
        return s.report_nonconv(before, nsteps);
    });
}

// run() FluidSolver3D.cu:3100-3163: max_steps * iterPerFrame + 1 steps, a
// particle file every iterPerFrame steps.  The positions of a frame are copied
// to the host after its step; formatting and writing the file runs on a host
// thread while the following steps run on the GPU (one frame in flight).
int fs_run(void* h, int frames) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        const int ipf = std::max(1, s.cfg.iter_per_frame);
        std::ofstream csv("OutputData/csvRes.csv");
        if (csv.is_open()) csv << "frame,centerX1,centerX2,centerX3,angleX1,angleX2,angleX3\n";
        std::thread writer;
        std::vector<float> frame_pos;
        int writer_rc = FS_OK;
        std::string writer_err;
        auto join = [&]() {
            if (writer.joinable()) writer.join();
            return writer_rc;
        };
        int rc = FS_OK;
        const int64_t before = s.nonconv;
        for (int i = 0; i <= frames * ipf && rc == FS_OK; ++i) {
            rc = s.step_once();
            if (rc) break;
            if (i % ipf == 0) {
                if ((rc = s.drain())) break;
                if ((rc = join())) break;
                if ((rc = gather_positions(s, frame_pos))) break;
                const std::string fn = std::string(s.cfg.output_template) + std::to_string(i / ipf) +
                                       (s.cfg.output_format == FS_FMT_OFF ? ".off" : ".ply");
                const int fmt = s.cfg.output_format;
                writer = std::thread([&frame_pos, fn, fmt, &writer_rc, &writer_err]() {
                    writer_rc = format_particle_file(frame_pos, fn, fmt, &writer_err);
                });
                if (s.body.active) write_body_ply(s, "OutputData/body_" + std::to_string(i / ipf) + ".ply");
                if (csv.is_open())
                    csv << i / ipf << "," << s.body.pos.x << "," << s.body.pos.y << "," << s.body.pos.z << ","
                        << s.body.theta << "," << s.body.psi << "," << s.body.phi << "\n";
            }
        }
        const int drc = s.drain();
        const int wrc = join();
        if (rc) return rc;
        if (drc) return drc;
        if (wrc) return fail(wrc, writer_err);
        return s.report_nonconv(before, (int64_t)frames * ipf + 1);
    });
}

#define STAGE_SYNC(call)                                                  \
    do {                                                                  \
        int rc_ = (call);                                                 \
        if (rc_) return rc_;                                              \
        HIPCHK(hipStreamSynchronize(s.st));                               \
        return FS_OK;                                                     \
    } while (0)

int fs_stage_label(void* h) {
    return guarded([&]() -> int { Solver& s = *(Solver*)h; STAGE_SYNC(s.stage_sort_label());
    });
}
int fs_stage_p2g(void* h) {
    return guarded([&]() -> int { Solver& s = *(Solver*)h; STAGE_SYNC(s.stage_p2g());
    });
}
int fs_stage_rhs(void* h) {
    return guarded([&]() -> int { Solver& s = *(Solver*)h; STAGE_SYNC(s.stage_rhs());
    });
}
int fs_stage_pressure(void* h) {
    return guarded([&]() -> int { Solver& s = *(Solver*)h; STAGE_SYNC(s.stage_pressure(true));
    });
}
int fs_stage_project(void* h) {
    return guarded([&]() -> int { Solver& s = *(Solver*)h; STAGE_SYNC(s.stage_project());
    });
}
int fs_stage_body(void* h) {
    return guarded([&]() -> int { Solver& s = *(Solver*)h; STAGE_SYNC(s.stage_body());
    });
}
int fs_stage_g2p_advect(void* h) {
    return guarded([&]() -> int { Solver& s = *(Solver*)h; STAGE_SYNC(s.stage_g2p_advect());
    });
}

int fs_get_grid(void* h, int which, void* out, int64_t count) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        HIPCHK(hipStreamSynchronize(s.st));
        const void* src = nullptr;
        long long n = 0;
        size_t esz = 4;
        switch (which) {
            case FS_GRID_U: src = s.d_u; n = s.NU(); break;
            case FS_GRID_V: src = s.d_v; n = s.NV(); break;
            case FS_GRID_W: src = s.d_w; n = s.NW(); break;
            case FS_GRID_USAVED: src = s.d_us; n = s.NU(); break;
            case FS_GRID_VSAVED: src = s.d_vs; n = s.NV(); break;
            case FS_GRID_WSAVED: src = s.d_ws; n = s.NW(); break;
            case FS_GRID_P: src = s.d_p; n = s.C; break;
            case FS_GRID_RHS: src = s.d_rhs; n = s.C; break;
            case FS_GRID_LABELS: src = s.d_labels; n = s.C; esz = 1; break;
            case FS_GRID_SOLID: src = s.d_solid; n = s.C; esz = 1; break;
            case FS_GRID_P64: src = s.d_p64; n = s.C; esz = 8; break;
            case FS_GRID_OPEN_U:
            case FS_GRID_OPEN_V:
            case FS_GRID_OPEN_W: {
                const int comp = which - FS_GRID_OPEN_U;
                n = comp == 0 ? s.NU() : comp == 1 ? s.NV() : s.NW();
                if (count < n) return fail(FS_ERR_ARG, "buffer too small");
                float* tmp = nullptr;
                HIPCHK(hipMalloc(&tmp, sizeof(float) * n));
                FaceFrac ff;
                if (int rc = s.face_frac(ff)) {
                    hipFree(tmp);
                    return rc;
                }
                launch_face_open(s.grid_args(), ff, comp, tmp, s.st);
                const hipError_t e1 = hipStreamSynchronize(s.st);
                const hipError_t e2 = e1 == hipSuccess ? hipMemcpy(out, tmp, sizeof(float) * n, hipMemcpyDeviceToHost) : e1;
                hipFree(tmp);
                if (e2 != hipSuccess) return fail(FS_ERR_DEVICE, hipGetErrorString(e2));
                return FS_OK;
            }
            case FS_GRID_CELLKEY: {
                n = s.np;
                if (count < n) return fail(FS_ERR_ARG, "buffer too small");
                if (n == 0) return FS_OK;
                if (!s.d_tmpkey) HIPCHK(hipMalloc(&s.d_tmpkey, sizeof(int) * n));
                launch_cell_keys(s.PA, (int)n, s.cfg.dx, s.W, s.H, s.D, s.d_tmpkey, s.st);
                std::vector<int> k((size_t)n), id((size_t)n);
                HIPCHK(hipMemcpyAsync(k.data(), s.d_tmpkey, sizeof(int) * n, hipMemcpyDeviceToHost, s.st));
                HIPCHK(hipMemcpyAsync(id.data(), s.PA.id, sizeof(int) * n, hipMemcpyDeviceToHost, s.st));
                HIPCHK(hipStreamSynchronize(s.st));
                int* o = (int*)out;
                for (long long i = 0; i < n; ++i) o[id[i]] = k[i];
                return FS_OK;
            }
            default: return fail(FS_ERR_ARG, "unknown grid");
        }
        if (count < n) return fail(FS_ERR_ARG, "buffer too small");
        if (n) HIPCHK(hipMemcpy(out, src, esz * n, hipMemcpyDeviceToHost));
        return FS_OK;
    });
}

int fs_set_grid(void* h, int which, const void* in, int64_t count) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        void* dst = nullptr;
        long long n = 0;
        size_t esz = 4;
        switch (which) {
            case FS_GRID_U: dst = s.d_u; n = s.NU(); break;
            case FS_GRID_V: dst = s.d_v; n = s.NV(); break;
            case FS_GRID_W: dst = s.d_w; n = s.NW(); break;
            case FS_GRID_USAVED: dst = s.d_us; n = s.NU(); break;
            case FS_GRID_VSAVED: dst = s.d_vs; n = s.NV(); break;
            case FS_GRID_WSAVED: dst = s.d_ws; n = s.NW(); break;
            case FS_GRID_P: dst = s.d_p; n = s.C; break;
            case FS_GRID_RHS: dst = s.d_rhs; n = s.C; break;
            case FS_GRID_LABELS: dst = s.d_labels; n = s.C; esz = 1; break;
            case FS_GRID_P64: dst = s.d_p64; n = s.C; esz = 8; break;
            default: return fail(FS_ERR_ARG, "grid not settable");
        }
        if (count != n) return fail(FS_ERR_ARG, "size mismatch");
        HIPCHK(hipStreamSynchronize(s.st));
        if (n) HIPCHK(hipMemcpy(dst, in, esz * n, hipMemcpyHostToDevice));
        if (dst == s.d_labels) s.labels_from_outside();
        return FS_OK;
    });
}

int fs_get_stage_times(void* h, double* ms, int n) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        for (int i = 0; i < n && i < FS_NUM_STAGES; ++i) ms[i] = s.times[i];
        return FS_OK;
    });
}

int fs_get_clock(void* h, double* t, float* dt, int64_t* n) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        if (t) *t = s.sim_time;
        if (dt) *dt = s.dt_cur;
        if (n) *n = s.nsteps;
        return FS_OK;
    });
}

int fs_get_solve_stats(void* h, fs_solve_stats* o) {
    return guarded([&]() -> int {
        *o = ((Solver*)h)->stats;
        return FS_OK;
    });
}

int fs_write_particles(void* h, const char* path, int fmt) {
    return guarded([&]() -> int { return write_particle_file(*(Solver*)h, path, fmt);
    });
}

int fs_solve_pressure_system(void* h, const uint8_t* labels, const float* b, double* p, int W, int H, int D,
                             double tol, fs_solve_stats* stats) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        if (!labels || !b || !p || W <= 0 || H <= 0 || D <= 0) return fail(FS_ERR_ARG, "bad system");
        if (s.W != W || s.H != H || s.D != D) {
            std::vector<uint8_t> L(labels, labels + (size_t)W * H * D);
            int rc = upload_scene(s, W, H, D, L);
            if (rc) return rc;
        } else {
            HIPCHK(hipMemcpy(s.d_labels, labels, (size_t)W * H * D, hipMemcpyHostToDevice));
            s.labels_from_outside();
        }
        HIPCHK(hipMemcpy(s.d_rhs, b, sizeof(float) * (size_t)W * H * D, hipMemcpyHostToDevice));
        double save = s.cfg.solver_tol;
        s.cfg.solver_tol = tol;
        int rc = s.stage_pressure(false, false);
        s.cfg.solver_tol = save;
        HIPCHK(hipStreamSynchronize(s.st));
        HIPCHK(hipMemcpy(p, s.d_p64, sizeof(double) * (size_t)W * H * D, hipMemcpyDeviceToHost));
        if (stats) *stats = s.stats;
        return rc;
    });
}

}  // extern "C"

// ---------------------------------------------------------------- learned preconditioner
namespace {

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) hipFree(p); }
    bool alloc(size_t b) { return hipMalloc(&p, std::max<size_t>(b, 1)) == hipSuccess; }
};

// rms of x over the FLUID cells (1 when there are none or x is zero there)
double fluid_rms(const uint8_t* labels, const float* x, size_t n, int* nf) {
    double s = 0.0;
    int c = 0;
    for (size_t i = 0; i < n; ++i)
        if (labels[i] == FS_FLUID) { s += (double)x[i] * x[i]; ++c; }
    *nf = c;
    double r = c > 0 ? std::sqrt(s / c) : 0.0;
    return r > 0.0 && std::isfinite(r) ? r : 1.0;
}

// p.(A p) of the pressure matrix of the labels (FluidSolver3D.cu:1598-1716):
// on FLUID cells diag = #in-domain FLUID/AIR neighbours, -1 per FLUID neighbour
double a_norm2(const uint8_t* labels, const double* p, int W, int H, int D) {
    const long long WH = (long long)W * H;
    double s = 0.0;
    for (int z = 0; z < D; ++z)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const long long c = x + (long long)W * y + WH * z;
                if (labels[c] != FS_FLUID) continue;
                double ap = 0.0;
                const int nx[6] = {x + 1, x - 1, x, x, x, x}, ny[6] = {y, y, y + 1, y - 1, y, y},
                          nz[6] = {z, z, z, z, z + 1, z - 1};
                for (int q = 0; q < 6; ++q) {
                    if (nx[q] < 0 || nx[q] >= W || ny[q] < 0 || ny[q] >= H || nz[q] < 0 || nz[q] >= D) continue;
                    const long long n = nx[q] + (long long)W * ny[q] + WH * nz[q];
                    if (labels[n] == FS_FLUID || labels[n] == FS_AIR) ap += p[c];
                    if (labels[n] == FS_FLUID) ap -= p[n];
                }
                s += p[c] * ap;
            }
    return s;
}

int learned_forward_host(Solver& s, const uint8_t* labels, const float* r, int W, int H, int D, DevBuf& dl,
                         DevBuf& dr, DevBuf& dz, DevBuf& dzn, double* scale, int* nf) {
    if (!labels || !r || W <= 0 || H <= 0 || D <= 0 || (long long)W * H * D >= (1LL << 31))
        return fail(FS_ERR_ARG, "bad grid");
    LNet* n = s.get_net();
    if (!n) return fail(FS_ERR_DEVICE, "learned preconditioner allocation failed");
    const size_t C = (size_t)W * H * D;
    *scale = fluid_rms(labels, r, C, nf);
    if (!dl.alloc(C) || !dr.alloc(sizeof(float) * C) || !dz.alloc(sizeof(float) * C) ||
        !dzn.alloc(sizeof(float) * C) || n->ensure(W, H, D) != 0)
        return fail(FS_ERR_DEVICE, "allocation failed");
    HIPCHK(hipMemcpyAsync(dl.p, labels, C, hipMemcpyHostToDevice, s.st));
    HIPCHK(hipMemcpyAsync(dr.p, r, sizeof(float) * C, hipMemcpyHostToDevice, s.st));
    // FLUID = 1 and AIR = 2 are also the network's unknown / Dirichlet types
    n->forward((const uint8_t*)dl.p, (const float*)dr.p, (float*)dz.p, nullptr, (float)*scale, (float*)dzn.p, s.st);
    return FS_OK;
}

int learned_loss_grad(Solver& s, const uint8_t* labels, const float* b, const double* p, int W, int H, int D,
                      double* loss, float* grad, float lr, bool step) {
    if (!p) return fail(FS_ERR_ARG, "null target");
    DevBuf dl, dr, dz, dzn, dt;
    double scale;
    int nf;
    int rc = learned_forward_host(s, labels, b, W, H, D, dl, dr, dz, dzn, &scale, &nf);
    if (rc) return rc;
    const size_t C = (size_t)W * H * D;
    // target p / sqrt(p.Ap / n_fluid): flexible PCG is invariant to the scale
    // of z, so only the shape of p is learned, with a loss of order one
    double sp = nf > 0 ? std::sqrt(std::max(0.0, a_norm2(labels, p, W, H, D)) / nf) : 0.0;
    if (!(sp > 0.0) || !std::isfinite(sp)) sp = 1.0;
    std::vector<float> t(C);
    for (size_t i = 0; i < C; ++i) t[i] = labels[i] == FS_FLUID ? (float)(p[i] / sp) : 0.0f;
    if (!dt.alloc(sizeof(float) * C)) return fail(FS_ERR_DEVICE, "allocation failed");
    HIPCHK(hipMemcpyAsync(dt.p, t.data(), sizeof(float) * C, hipMemcpyHostToDevice, s.st));
    LNet* n = s.net.get();
    n->backward((const uint8_t*)dl.p, (const float*)dt.p, (const float*)dzn.p, nf, s.st);
    double l = 0.0;
    HIPCHK(hipMemcpyAsync(&l, n->d_loss, sizeof(double), hipMemcpyDeviceToHost, s.st));
    if (grad) HIPCHK(hipMemcpyAsync(grad, n->d_grad, sizeof(float) * LN_NPARAM, hipMemcpyDeviceToHost, s.st));
    if (step) n->adam(lr, s.st);
    HIPCHK(hipStreamSynchronize(s.st));
    if (loss) *loss = l;
    return FS_OK;
}

}  // namespace

extern "C" {

int fs_learned_init(void* h, uint64_t seed) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        LNet* n = s.get_net();
        if (!n || n->init_params(seed, s.st) != 0) return fail(FS_ERR_DEVICE, "parameter upload failed");
        return FS_OK;
    });
}

int fs_learned_get_params(void* h, float* params, int n) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        if (!params || n != LN_NPARAM) return fail(FS_ERR_ARG, "expected FSL_NUM_PARAMS parameters");
        LNet* net = s.get_net();
        if (!net || net->get_params(params, s.st) != 0) return fail(FS_ERR_DEVICE, "parameter download failed");
        return FS_OK;
    });
}

int fs_learned_set_params(void* h, const float* params, int n) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        if (!params || n != LN_NPARAM) return fail(FS_ERR_ARG, "expected FSL_NUM_PARAMS parameters");
        for (int i = 0; i < n; ++i)
            if (!std::isfinite(params[i])) return fail(FS_ERR_ARG, "non-finite parameter");
        LNet* net = s.get_net();
        if (!net || net->set_params(params, s.st) != 0) return fail(FS_ERR_DEVICE, "parameter upload failed");
        return FS_OK;
    });
}

int fs_learned_apply(void* h, const uint8_t* labels, const float* r, float* z, int W, int H, int D) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        if (!z) return fail(FS_ERR_ARG, "null output");
        DevBuf dl, dr, dz, dzn;
        double scale;
        int nf;
        int rc = learned_forward_host(s, labels, r, W, H, D, dl, dr, dz, dzn, &scale, &nf);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(z, dz.p, sizeof(float) * (size_t)W * H * D, hipMemcpyDeviceToHost, s.st));
        HIPCHK(hipStreamSynchronize(s.st));
        return FS_OK;
    });
}

int fs_learned_loss_grad(void* h, const uint8_t* labels, const float* b, const double* p, int W, int H, int D,
                         double* loss, float* grad) {
    return guarded([&]() -> int {
        return learned_loss_grad(*(Solver*)h, labels, b, p, W, H, D, loss, grad, 0.0f, false);
    });
}

int fs_learned_train_step(void* h, const uint8_t* labels, const float* b, const double* p, int W, int H, int D,
                          float lr, double* loss) {
    return guarded([&]() -> int {
        if (!(lr > 0.0f)) return fail(FS_ERR_ARG, "learning rate must be positive");
        return learned_loss_grad(*(Solver*)h, labels, b, p, W, H, D, loss, nullptr, lr, true);
    });
}

int fs_mg_vcycle(void* h, const uint8_t* labels, const float* r, float* z, int W, int H, int D) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        if (!labels || !r || !z || W <= 0 || H <= 0 || D <= 0) return fail(FS_ERR_ARG, "bad system");
        if (s.W != W || s.H != H || s.D != D) {
            std::vector<uint8_t> L(labels, labels + (size_t)W * H * D);
            int rc = upload_scene(s, W, H, D, L);
            if (rc) return rc;
        } else {
            HIPCHK(hipMemcpy(s.d_labels, labels, (size_t)W * H * D, hipMemcpyHostToDevice));
            s.labels_from_outside();
        }
        const size_t C = (size_t)W * H * D;
        HIPCHK(hipMemcpy(s.d_rhs, r, sizeof(float) * C, hipMemcpyHostToDevice));
        // the full-grid float pressure buffer receives z
        int rc = s.mg.vcycle_system(s.d_labels, s.d_rhs, s.d_p, s.st);
        s.p64_from_step = false;
        if (rc == -5) return fail(FS_ERR_DEVICE, std::string("V-cycle: ") + hipGetErrorString(hipGetLastError()));
        if (rc == -3) return fail(FS_ERR_NO_FLUID, "No fluid cells found!");
        HIPCHK(hipMemcpy(z, s.d_p, sizeof(float) * C, hipMemcpyDeviceToHost));
        return FS_OK;
    });
}

int fs_set_preconditioner(void* h, int precond, int stall_window) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        if (precond < FS_PRECOND_MG || precond > FS_PRECOND_HYBRID) return fail(FS_ERR_ARG, "unknown preconditioner");
        s.cfg.precond = precond;
        if (stall_window > 0) s.cfg.stall_window = stall_window;
        return FS_OK;
    });
}

}  // extern "C"

// ---------------------------------------------------------------- checkpoint
namespace {

constexpr char kCkptMagic[8] = {'F', 'S', 'C', 'K', 'P', 'T', '0', '2'};

struct CkptWriter {
    std::ofstream f;
    bool ok = true;
    void raw(const void* p, size_t n) {
        if (n) f.write((const char*)p, (std::streamsize)n);
        ok = ok && (bool)f;
    }
    template <class T>
    void put(const T& v) { raw(&v, sizeof(T)); }
};

struct CkptReader {
    std::ifstream f;
    bool ok = true;
    // bytes left after the current position
    long long remaining() {
        const std::streampos here = f.tellg();
        f.seekg(0, std::ios::end);
        const std::streampos end = f.tellg();
        f.seekg(here);
        return (here < 0 || end < 0) ? -1 : (long long)(end - here);
    }
    void raw(void* p, size_t n) {
        if (n) f.read((char*)p, (std::streamsize)n);
        ok = ok && (bool)f;
    }
    template <class T>
    void get(T& v) { raw(&v, sizeof(T)); }
};

// the body fields a step reads (everything but the SDF values), each passed
// to io(pointer, bytes)
template <class IO>
void body_fields(HostBody& b, IO io) {
    auto f = [&](auto& v) { io((void*)&v, sizeof(v)); };
    f(b.sw); f(b.sh); f(b.sd);
    f(b.file_origin); f(b.sdf_origin); f(b.local_com); f(b.size); f(b.cs);
    f(b.pos); f(b.vel); f(b.force); f(b.omega); f(b.torque);
    f(b.mass); f(b.inertia); f(b.inv_inertia); f(b.theta); f(b.psi); f(b.phi);
    f(b.R); f(b.last_force); f(b.last_torque); f(b.last_sub);
    f(b.Ib); f(b.Ibinv); f(b.q);
}

}  // namespace

extern "C" {

int fs_save_checkpoint(void* h, const char* path) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        if (!path) return fail(FS_ERR_ARG, "null path");
        if (s.C == 0) return fail(FS_ERR_STATE, "nothing to save: no scene");
        HIPCHK(hipStreamSynchronize(s.st));
        CkptWriter w;
        w.f.open(path, std::ios::binary);
        if (!w.f.is_open()) return fail(FS_ERR_IO, std::string("cannot write ") + path);
        w.raw(kCkptMagic, 8);
        const int32_t cfg_size = (int32_t)sizeof(fs_config);
        w.put(cfg_size);
        w.put(s.cfg);
        w.put(s.W); w.put(s.H); w.put(s.D);
        w.put(s.dt_cur); w.put(s.sim_time); w.put(s.nsteps);
        const int64_t np = s.np;
        w.put(np);
        std::vector<uint8_t> g((size_t)s.C);
        HIPCHK(hipMemcpy(g.data(), s.d_solid, (size_t)s.C, hipMemcpyDeviceToHost));
        w.raw(g.data(), g.size());
        HIPCHK(hipMemcpy(g.data(), s.d_labels, (size_t)s.C, hipMemcpyDeviceToHost));
        w.raw(g.data(), g.size());
        std::vector<float> a((size_t)np);
        for (float* src : {s.PA.px, s.PA.py, s.PA.pz, s.PA.vx, s.PA.vy, s.PA.vz}) {
            if (np) HIPCHK(hipMemcpy(a.data(), src, sizeof(float) * np, hipMemcpyDeviceToHost));
            w.raw(a.data(), sizeof(float) * np);
        }
        std::vector<int> ids((size_t)np);
        if (np) HIPCHK(hipMemcpy(ids.data(), s.PA.id, sizeof(int) * np, hipMemcpyDeviceToHost));
        w.raw(ids.data(), sizeof(int) * np);
        const int32_t active = s.body.active ? 1 : 0;
        w.put(active);
        if (active) {
            body_fields(s.body, [&](void* p, size_t n) { w.raw(p, n); });
            w.raw(s.body.sdf.data(), sizeof(float) * s.body.sdf.size());
        }
        // the last step's pressure, the next solve's initial guess
        const int32_t has_p = s.p64_from_step ? 1 : 0;
        w.put(has_p);
        if (has_p) {
            std::vector<double> pv((size_t)s.C);
            HIPCHK(hipMemcpy(pv.data(), s.d_p64, sizeof(double) * (size_t)s.C, hipMemcpyDeviceToHost));
            w.raw(pv.data(), sizeof(double) * pv.size());
        }
        // the learned preconditioner's parameters (not its optimiser state)
        const int32_t nparam = s.net ? (int32_t)LN_NPARAM : 0;
        w.put(nparam);
        if (nparam) {
            std::vector<float> prm((size_t)nparam);
            if (s.net->get_params(prm.data(), s.st) != 0) return fail(FS_ERR_DEVICE, "parameter download failed");
            w.raw(prm.data(), sizeof(float) * prm.size());
        }
        w.f.close();
        if (!w.ok || w.f.fail()) return fail(FS_ERR_IO, std::string("write failed: ") + path);
        return FS_OK;
    });
}

int fs_load_checkpoint(void* h, const char* path) {
    return guarded([&]() -> int {
        Solver& s = *(Solver*)h;
        if (!path) return fail(FS_ERR_ARG, "null path");
        CkptReader r;
        r.f.open(path, std::ios::binary);
        if (!r.f.is_open()) return fail(FS_ERR_IO, std::string("cannot open ") + path);
        char magic[8];
        r.raw(magic, 8);
        int32_t cfg_size = 0;
        r.get(cfg_size);
        if (!r.ok || std::memcmp(magic, kCkptMagic, 8) != 0 || cfg_size != (int32_t)sizeof(fs_config))
            return fail(FS_ERR_IO, std::string("not a checkpoint of this version: ") + path);
        fs_config cfg;
        r.get(cfg);
        int W = 0, H = 0, D = 0;
        r.get(W); r.get(H); r.get(D);
        float dt_cur = 0.0f;
        double sim_time = 0.0;
        int64_t nsteps = 0, np = 0;
        r.get(dt_cur); r.get(sim_time); r.get(nsteps); r.get(np);
        if (!r.ok || W <= 0 || H <= 0 || D <= 0 || np < 0 || np > 0x7fffffffLL)
            return fail(FS_ERR_IO, std::string("corrupt checkpoint header: ") + path);
        // sizes from the header are checked against the file before anything is allocated
        const long long left = r.remaining();
        const long long cells = (long long)W * H * D;
        if (left < 0 || cells > (1LL << 40) || (long long)W > left || (long long)H > left ||
            2 * cells + 28 * np + 8 > left)
            return fail(FS_ERR_IO, std::string("checkpoint shorter than its header says: ") + path);
        const size_t C = (size_t)W * H * D;
        std::vector<uint8_t> solid(C), labels(C);
        r.raw(solid.data(), C);
        r.raw(labels.data(), C);
        std::vector<float> arr[6];
        for (auto& a : arr) {
            a.resize((size_t)np);
            r.raw(a.data(), sizeof(float) * np);
        }
        std::vector<int> ids((size_t)np);
        r.raw(ids.data(), sizeof(int) * np);
        int32_t active = 0;
        r.get(active);
        HostBody body;
        if (r.ok && active) {
            body.active = true;
            body_fields(body, [&](void* p, size_t n) { r.raw(p, n); });
            if (!(body.sw > 0 && body.sh > 0 && body.sd > 0)) return fail(FS_ERR_IO, "corrupt checkpoint body");
            const long long nv = (long long)body.sw * body.sh * body.sd;
            if (!r.ok || 4 * nv + 4 > r.remaining()) return fail(FS_ERR_IO, "corrupt checkpoint body");
            body.sdf.resize((size_t)nv);
            r.raw(body.sdf.data(), sizeof(float) * body.sdf.size());
        }
        int32_t has_p = 0;
        r.get(has_p);
        std::vector<double> pv;
        if (r.ok && has_p) {
            if (8LL * (long long)C + 4 > r.remaining()) return fail(FS_ERR_IO, std::string("corrupt checkpoint: ") + path);
            pv.resize(C);
            r.raw(pv.data(), sizeof(double) * C);
        }
        int32_t nparam = 0;
        r.get(nparam);
        std::vector<float> prm;
        if (r.ok && nparam) {
            if (nparam != LN_NPARAM || 4LL * nparam > r.remaining())
                return fail(FS_ERR_IO, std::string("corrupt checkpoint parameters: ") + path);
            prm.resize((size_t)nparam);
            r.raw(prm.data(), sizeof(float) * prm.size());
        }
        if (!r.ok) return fail(FS_ERR_IO, std::string("truncated checkpoint: ") + path);
        // the state of the handle is replaced only once the whole file has been read
        s.cfg = cfg;
        int rc = s.alloc_grids(W, H, D);
        if (rc) return rc;
        HIPCHK(hipMemcpy(s.d_solid, solid.data(), C, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(s.d_labels, labels.data(), C, hipMemcpyHostToDevice));
        s.labels_from_outside();
        rc = s.alloc_particles(np);
        if (rc) return rc;
        float* dst[6] = {s.PA.px, s.PA.py, s.PA.pz, s.PA.vx, s.PA.vy, s.PA.vz};
        for (int c = 0; c < 6; ++c)
            if (np) HIPCHK(hipMemcpy(dst[c], arr[c].data(), sizeof(float) * np, hipMemcpyHostToDevice));
        if (np) HIPCHK(hipMemcpy(s.PA.id, ids.data(), sizeof(int) * np, hipMemcpyHostToDevice));
        s.body = std::move(body);
        dfree(s.d_sdf);
        s.theta_valid = false;
        if (s.body.active) {
            HIPCHK(hipMalloc(&s.d_sdf, sizeof(float) * s.body.sdf.size()));
            HIPCHK(hipMemcpy(s.d_sdf, s.body.sdf.data(), sizeof(float) * s.body.sdf.size(), hipMemcpyHostToDevice));
        }
        if (!pv.empty()) {
            HIPCHK(hipMemcpy(s.d_p64, pv.data(), sizeof(double) * C, hipMemcpyHostToDevice));
            s.p64_from_step = true;
        }
        if (!prm.empty()) {
            LNet* n = s.get_net();
            if (!n || n->set_params(prm.data(), s.st) != 0) return fail(FS_ERR_DEVICE, "parameter upload failed");
        }
        s.learned.reset();
        s.dt_cur = dt_cur;
        s.sim_time = sim_time;
        s.nsteps = nsteps;
        return FS_OK;
    });
}

}  // extern "C"

// ---------------------------------------------------------------- allocation-free solve
namespace {

// Host state of the allocation-free solve on one device: the solver object
// (its level list reserved at set-up) and the host mirrors of its scalars.
// Created on the first call on a device and kept until the process ends.
struct NoAllocState {
    MGPCG mg;
    SolveScalars sc_host{};
    int host_counters[MGPCG::HOST_COUNTERS] = {};
};
constexpr int kNoAllocDevices = 64;
NoAllocState* g_noalloc[kNoAllocDevices] = {};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

int fs_solve_pressure_system_noalloc(const uint8_t* labels, const float* b, double* p, int W, int H, int D,
                                     double tol, fs_solve_stats* stats) {
    return guarded([&]() -> int {
        if (!labels || !b || !p || W <= 0 || H <= 0 || D <= 0) return fail(FS_ERR_ARG, "bad system");
        const long long C = (long long)W * H * D;
        if (C > FS_NOALLOC_MAX_CELLS)
            return fail(FS_ERR_ARG, "allocation-free solve: " + std::to_string(C) + " cells, at most " +
                                        std::to_string(FS_NOALLOC_MAX_CELLS));
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        if (dev < 0 || dev >= kNoAllocDevices) return fail(FS_ERR_DEVICE, "device index out of range");
        hipError_t e = hipSuccess;
        char* base = (char*)noalloc_scratch(dev, &e);
        if (!base) return fail(FS_ERR_DEVICE, std::string("workspace address: ") + hipGetErrorString(e));
        if (!g_noalloc[dev]) {
            // set-up, once per device: host objects only
            NoAllocState* st = new NoAllocState();
            st->mg.lv.reserve(40);
            g_noalloc[dev] = st;
        }
        NoAllocState& st = *g_noalloc[dev];
        // staging of labels, b, p (FP64) and p (FP32), then the solver's workspace
        size_t off = 0;
        auto carve = [&](size_t bytes) {
            off = align256(off);
            char* q = base + off;
            off += bytes;
            return (void*)q;
        };
        uint8_t* dl = (uint8_t*)carve((size_t)C);
        float* db = (float*)carve(sizeof(float) * (size_t)C);
        double* dp64 = (double*)carve(sizeof(double) * (size_t)C);
        float* dp = (float*)carve(sizeof(float) * (size_t)C);
        off = align256(off);
        size_t used = 0;
        int rc = FS_NOALLOC_SCRATCH_BYTES > off
                     ? st.mg.init_in(W, H, D, base + off, FS_NOALLOC_SCRATCH_BYTES - off, &st.sc_host,
                                     st.host_counters, &used)
                     : -2;
        if (rc == -2)
            return fail(FS_ERR_ARG, "allocation-free solve: a " + std::to_string(W) + " x " + std::to_string(H) +
                                        " x " + std::to_string(D) + " system needs " + std::to_string(off + used) +
                                        " bytes of workspace, the limit is " +
                                        std::to_string((size_t)FS_NOALLOC_SCRATCH_BYTES));
        if (rc) return fail(FS_ERR_DEVICE, "allocation-free solve: workspace set-up failed");
        HIPCHK(hipMemcpy(dl, labels, (size_t)C, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(db, b, sizeof(float) * (size_t)C, hipMemcpyHostToDevice));
        // the settings of a default handle (fs_config_default), launched eagerly
        // (the same bits as the graph replays of fs_solve_pressure_system)
        st.mg.stall_window = 5;
        st.mg.use_graph = false;
        st.mg.use_tail = true;
        st.mg.use_tb = true;
        fs_solve_stats loc{};
        rc = st.mg.solve(dl, db, dp64, dp, tol, 500, nullptr, (hipStream_t)0, &loc);
        if (rc == -5) return fail(FS_ERR_DEVICE, std::string("pressure solve: ") + hipGetErrorString(hipGetLastError()));
        HIPCHK(hipStreamSynchronize((hipStream_t)0));
        HIPCHK(hipMemcpy(p, dp64, sizeof(double) * (size_t)C, hipMemcpyDeviceToHost));
        if (stats) *stats = loc;
        if (rc == -3) return fail(FS_ERR_NO_FLUID, "No fluid cells found!");
        if (rc == -4) return fail(FS_ERR_NOT_CONVERGED, "pressure solve did not reach the tolerance");
        return FS_OK;
    });
}

}  // extern "C"
