/*
 * fs_oracle.cpp -- CPU restatement of the reference PIC/FLIP step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * (fluidprjct_amd/libfluidsim_hip.so).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it.  It exports the same C ABI
 * (include/fluidsim.h) so a test can drive both implementations with the same
 * calls.
 *
 * Every stage follows the reference (Ship-Vano/FluidPrjct) line by line in
 * scalar FP32 wherever the reference computes in FP32, with the same operation
 * order; it must be compiled with -ffp-contract=off and without fast-math.
 * Deliberate differences (each one is where the reference is undefined or
 * nondeterministic):
 *   - seeding uses a counter-based SplitMix64 stream keyed by (seed, cell, slot)
 *     instead of std::random_device + mt19937 (FluidSolver3D.cu:215-218);
 *   - P2G accumulates num/den in FP64 in particle order, then rounds each sum to
 *     FP32 (the reference's FP32 atomics have no fixed order, :513-590);
 *   - out-of-bounds label/pressure reads of the reference (:1406-1455,
 *     :2109-2110, :2201-2202, :2288-2289) read the wrapped linear index exactly
 *     as the reference does when that index lies inside the array and read SOLID
 *     (p = 0) when it lies outside;
 *   - G2P/advection corner reads that would leave the array are clamped to the
 *     last element (never reached when positions are clamped to (N-1)dx, :1298);
 *   - the pressure system of constructA/constructRHS is solved in FP64 (band
 *     LDL^T direct solve for small systems, else PCG) instead of cuDSS FP32;
 *   - lift is skipped when |n x u_rel| == 0 (the reference divides by zero,
 *     :2774-2775).
 * Options with no reference counterpart (FS_STEP_* flags) are restated here
 * too; the face-fraction option (FS_STEP_FACE_FRACTIONS) is pinned by
 * tests/test_oracle_face_fractions.py (plane cuts against exact polygon
 * areas, the hydrostatic balance with a resting body, the weighted divergence
 * after the projection, a plate thinner than a cell, scipy on the weighted
 * matrix).
 *
 * Parity pins (tests/test_oracle_*.py): label-file fixtures (counts), the
 * particles_-1.ply seeding fixture (layout/order/offset bounds), the cuDSS 5x5
 * known-answer test (test.cu:262-290), hydrostatic closed forms, P2G/G2P
 * partition-of-unity and uniform-field closed forms, constant-field RK3,
 * brute-force tiny cases, scipy spsolve of the assembled system.
 */
#include "../include/fluidsim.h"
#include "../include/fluidsim_checkpoint.h"

#include <algorithm>
#include <functional>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>
#include <climits>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

struct f3 { float x, y, z; };
inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
/* utility.cu:3-25 */
inline f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline f3 operator*(f3 a, float b) { return mk(a.x * b, a.y * b, a.z * b); }
inline f3 operator*(float b, f3 a) { return a * b; }
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline f3 operator/(f3 a, float b) { return a * (1.0f / b); }
/* utility.cu:208-215 */
inline f3 cross(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

/* ---------------- counter-based seeding generator -------------------------- */
inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}
/* draw d (0 = octant, 1..3 = jitter x,y,z) of particle slot s in cell c */
inline uint64_t seed_draw(uint64_t seed, uint64_t cell, uint64_t slot, uint64_t d) {
    uint64_t ctr = (cell << 8) | (slot << 2) | d;
    return splitmix64(splitmix64(seed) ^ ctr);
}

struct Body {
    bool active = false;
    std::vector<float> sdf;
    int sw = 0, sh = 0, sd = 0;
    f3 fileOrigin{0, 0, 0}, sdf_origin{0, 0, 0}, local_com{0, 0, 0}, size{0, 0, 0};
    float cs = 0;
    f3 pos{0, 0, 0}, vel{0, 0, 0}, force{0, 0, 0}, omega{0, 0, 0}, torque{0, 0, 0};
    float mass = 0, inertia = 0, inv_inertia = 0;
    float theta = 0, psi = 0, phi = 0;
    float R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    f3 last_force{0, 0, 0}, last_torque{0, 0, 0};
    float last_sub = 0;
    /* FS_STEP_BODY_INERTIA (no reference counterpart): inertia tensor of the
     * SDF's inside voxels about local_com in the body frame, its inverse, and
     * the orientation as a unit quaternion (w, x, y, z), body to world */
    double Ib[9] = {0}, Ibinv[9] = {0};
    double q[4] = {1, 0, 0, 0};

    /* RigidBody::update_rotation_matrix utility.cuh:470-484 */
    void update_rotation_matrix() {
        double th = theta, ps = psi, ph = phi;
        R[0] = (float)(std::cos(th) * std::cos(ps));
        R[1] = (float)(std::sin(th) * std::sin(ph) - std::cos(th) * std::cos(ph) * std::sin(ps));
        R[2] = (float)(std::cos(ph) * std::sin(th) + std::cos(th) * std::sin(ph) * std::sin(ps));
        R[3] = (float)(std::sin(ps));
        R[4] = (float)(std::cos(ph) * std::cos(ps));
        R[5] = (float)(-std::cos(ps) * std::sin(ph));
        R[6] = (float)(-std::cos(ps) * std::sin(th));
        R[7] = (float)(std::cos(th) * std::sin(ph) + std::cos(ph) * std::sin(th) * std::sin(ps));
        R[8] = (float)(std::cos(th) * std::cos(ph) - std::sin(th) * std::sin(ph) * std::sin(ps));
    }

    /* RigidBody::integrate utility.cuh:408-467 */
    void integrate(float dt) {
        f3 accel = force / mass;
        vel = vel + accel * dt;
        pos = pos + vel * dt;
        const float maxAngle = 3.14159265358979323846f;
        theta = std::clamp(theta, -maxAngle, maxAngle);
        update_rotation_matrix();
        f3 tau_local = mk(R[0] * torque.x + R[3] * torque.y + R[6] * torque.z,
                          R[1] * torque.x + R[4] * torque.y + R[7] * torque.z,
                          R[2] * torque.x + R[5] * torque.y + R[8] * torque.z);
        f3 angAccel = inv_inertia * tau_local;
        omega = omega + angAccel * dt;
        float dtheta_dt = (float)(omega.y * std::cos((double)phi) - omega.z * std::sin((double)phi));
        float dpsi_dt = (float)((omega.y * std::sin((double)phi) + omega.z * std::cos((double)phi)) /
                                std::cos((double)theta));
        float dphi_dt = (float)(omega.x + dpsi_dt * std::sin((double)theta));
        theta += dtheta_dt * dt;
        psi += dpsi_dt * dt;
        phi += dphi_dt * dt;
        f3 lo = fileOrigin;
        f3 ro = mk(R[0] * lo.x + R[1] * lo.y + R[2] * lo.z,
                   R[3] * lo.x + R[4] * lo.y + R[5] * lo.z,
                   R[6] * lo.x + R[7] * lo.y + R[8] * lo.z);
        sdf_origin = pos + ro;
        force = mk(0, 0, 0);
        torque = mk(0, 0, 0);
    }

    /* ---- FS_STEP_BODY_INERTIA ---------------------------------------------- */
    /* R of the quaternion (body to world), row-major as R[] */
    void rotation_from_quaternion() {
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
    /* the reference's angles of R = Ry(theta) Rz(psi) Rx(phi) (utility.cuh:470-484):
     * R[3] = sin psi, R[0] = cos theta cos psi, R[6] = -sin theta cos psi,
     * R[4] = cos phi cos psi, R[5] = -sin phi cos psi */
    void angles_from_rotation() {
        psi = (float)std::asin(std::clamp((double)R[3], -1.0, 1.0));
        theta = (float)std::atan2(-(double)R[6], (double)R[0]);
        phi = (float)std::atan2(-(double)R[5], (double)R[4]);
    }
    /* quaternion of Ry(theta) Rz(psi) Rx(phi): q_y(theta) q_z(psi) q_x(phi) */
    void quaternion_from_angles() {
        const double a = 0.5 * theta, b = 0.5 * psi, c = 0.5 * phi;
        const double qy[4] = {std::cos(a), 0, std::sin(a), 0}, qz[4] = {std::cos(b), 0, 0, std::sin(b)},
                     qx[4] = {std::cos(c), std::sin(c), 0, 0};
        double t[4];
        qmul(qy, qz, t);
        qmul(t, qx, q);
    }
    static void qmul(const double* a, const double* b, double* o) {
        o[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
        o[1] = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
        o[2] = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
        o[3] = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
    }
    /* Ib = sum over voxels with phi <= 0 (weights as applyBuoyancy,
     * FluidSolver3D.cu:2979-2983) of m_v (|r|^2 I - r r^T), r from local_com,
     * m_v = mass * weight / total weight; inverted by cofactors */
    void compute_inertia_tensor() {
        double Wsum = 0, A[9] = {0};
        for (int k = 0; k < sd; ++k)
            for (int j = 0; j < sh; ++j)
                for (int i = 0; i < sw; ++i) {
                    const float v = sdf[(size_t)i + (size_t)j * sw + (size_t)k * sw * sh];
                    if (v > 0.0f) continue;
                    double wv = 1.0;
                    if (v > -cs) wv = 0.5 - 0.5 * (double)v / cs;
                    const double rx = fileOrigin.x + (double)i * cs - local_com.x;
                    const double ry = fileOrigin.y + (double)j * cs - local_com.y;
                    const double rz = fileOrigin.z + (double)k * cs - local_com.z;
                    const double r2 = rx * rx + ry * ry + rz * rz;
                    A[0] += wv * (r2 - rx * rx); A[1] -= wv * rx * ry; A[2] -= wv * rx * rz;
                    A[4] += wv * (r2 - ry * ry); A[5] -= wv * ry * rz;
                    A[8] += wv * (r2 - rz * rz);
                    Wsum += wv;
                }
        A[3] = A[1]; A[6] = A[2]; A[7] = A[5];
        for (int e = 0; e < 9; ++e) Ib[e] = Wsum > 0 ? (double)mass * A[e] / Wsum : 0.0;
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
    /* semi-implicit Euler as integrate() for the translation; rotation by
     * Euler's equations in the body frame, w_b += dt Ib^-1 (tau_b - w_b x Ib w_b),
     * with omega kept in the world frame (the frame V + omega x r uses), and
     * the quaternion advanced by q += dt q (0, w_b) / 2, then normalised */
    void integrate_tensor(float dt) {
        f3 accel = force / mass;
        vel = vel + accel * dt;
        pos = pos + vel * dt;
        rotation_from_quaternion();
        double Rm[9];
        for (int e = 0; e < 9; ++e) Rm[e] = R[e];
        const double tw[3] = {torque.x, torque.y, torque.z}, ow[3] = {omega.x, omega.y, omega.z};
        double tb[3], wb[3], Iw[3], g[3], d[3];
        for (int a = 0; a < 3; ++a) {   /* R^T */
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
        qmul(q, wq, dq);
        for (int e = 0; e < 4; ++e) q[e] += 0.5 * (double)dt * dq[e];
        const double qn = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        for (double& e : q) e /= qn;
        /* omega in the world frame with the rotation the step started from */
        omega = mk((float)(Rm[0] * wb[0] + Rm[1] * wb[1] + Rm[2] * wb[2]),
                   (float)(Rm[3] * wb[0] + Rm[4] * wb[1] + Rm[5] * wb[2]),
                   (float)(Rm[6] * wb[0] + Rm[7] * wb[1] + Rm[8] * wb[2]));
        rotation_from_quaternion();
        angles_from_rotation();
        f3 lo = fileOrigin;
        f3 ro = mk(R[0] * lo.x + R[1] * lo.y + R[2] * lo.z, R[3] * lo.x + R[4] * lo.y + R[5] * lo.z,
                   R[6] * lo.x + R[7] * lo.y + R[8] * lo.z);
        sdf_origin = pos + ro;
        force = mk(0, 0, 0);
        torque = mk(0, 0, 0);
    }

    /* FluidSolver3D::handleBodyWallCollisions FluidSolver3D.cu:2848-2895 (the
     * reference has the call switched off, :2841): per axis, with r = half the
     * largest SDF box side and walls at the centres of the border cells, a
     * body closer than r to its clamp position is pushed back by the
     * difference and a negative velocity along the axis is removed; the
     * reference's assert dx == sdf cell size is not enforced */
    void collide_walls(int W, int H, int D, float dx) {
        const float r = 0.5f * fmaxf(fmaxf(size.x, size.y), size.z);
        const float half = 0.5f * dx;
        const f3 minWall = mk(half, half, half);
        const f3 maxWall = mk(W * dx - half, H * dx - half, D * dx - half);
        auto axis = [&](float& p, float pMin, float pMax, f3 n) {
            const float distMin = p - pMin, distMax = pMax - p;
            const float dist = fminf(distMin, distMax);
            const float penetration = r - dist;
            if (penetration > 0.0f) {
                p = p + penetration * (distMin < distMax ? +1.0f : -1.0f);
                const float vn = dot(vel, n);
                if (vn < 0.0f) vel = vel - vn * n;
            }
        };
        axis(pos.x, minWall.x + r, maxWall.x - r, mk(1, 0, 0));
        axis(pos.y, minWall.y + r, maxWall.y - r, mk(0, 1, 0));
        axis(pos.z, minWall.z + r, maxWall.z - r, mk(0, 0, 1));
        sdf_origin = pos - size * 0.5f;
    }
};

struct Solver {
    fs_config cfg;
    int W = 0, H = 0, D = 0;
    std::vector<uint8_t> solid;    /* 1 = SOLID (static) */
    std::vector<uint8_t> labels;
    std::vector<float> u, v, w, us, vs, ws, p, rhs;
    std::vector<double> p64;
    std::vector<float> pos, vel;   /* AoS xyz */
    std::vector<int> key;
    Body body;
    fs_solve_stats stats{};
    double times[FS_NUM_STAGES] = {0};
    float dt_cur = 0.0f;     /* dt of the current step */
    double sim_time = 0.0;
    int64_t nsteps = 0;
    int64_t nonconv = 0;     /* steps whose pressure solve missed the tolerance */
    double nonconv_res = 0.0;

    size_t C() const { return (size_t)W * H * D; }
    size_t NU() const { return (size_t)(W + 1) * H * D; }
    size_t NV() const { return (size_t)W * (H + 1) * D; }
    size_t NW() const { return (size_t)W * H * (D + 1); }
    int64_t np() const { return (int64_t)(pos.size() / 3); }

    void alloc_grids() {
        labels.assign(C(), FS_AIR);
        solid.assign(C(), 0);
        p.assign(C(), 0.f);
        p64.assign(C(), 0.0);
        rhs.assign(C(), 0.f);
        u.assign(NU(), 0.f); us.assign(NU(), 0.f);
        v.assign(NV(), 0.f); vs.assign(NV(), 0.f);
        w.assign(NW(), 0.f); ws.assign(NW(), 0.f);
    }
};

using clk = std::chrono::steady_clock;
inline double ms_since(clk::time_point t0) {
    return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
}

/* ---------------- label reads with the reference's wrap semantics ------------ */
inline int label_at(const Solver& s, long long idx) {
    if (idx < 0 || idx >= (long long)s.C()) return FS_SOLID;
    return s.labels[(size_t)idx];
}
inline float p_at(const Solver& s, long long idx) {
    if (idx < 0 || idx >= (long long)s.C()) return 0.0f;
    return s.p[(size_t)idx];
}

/* ---------------- Utility::contains utility.cu:163-206 ----------------------- */
bool body_contains(const Body& b, f3 world_pos) {
    f3 p_rel = world_pos - b.pos;
    const float* R = b.R;
    f3 plc;
    plc.x = R[0] * p_rel.x + R[3] * p_rel.y + R[6] * p_rel.z;
    plc.y = R[1] * p_rel.x + R[4] * p_rel.y + R[7] * p_rel.z;
    plc.z = R[2] * p_rel.x + R[5] * p_rel.y + R[8] * p_rel.z;
    f3 pl = plc + b.local_com;
    /* the reference passes body.fileOrigin as sdf_origin (FluidSolver3D.cu:183,359) */
    f3 sc = mk((pl.x - b.fileOrigin.x) / b.cs, (pl.y - b.fileOrigin.y) / b.cs,
               (pl.z - b.fileOrigin.z) / b.cs);
    if (sc.x < 0 || sc.x >= b.sw || sc.y < 0 || sc.y >= b.sh || sc.z < 0 || sc.z >= b.sd)
        return false;
    int i = std::min((int)sc.x, b.sw - 1);
    int j = std::min((int)sc.y, b.sh - 1);
    int k = std::min((int)sc.z, b.sd - 1);
    return b.sdf[(size_t)i + (size_t)j * b.sw + (size_t)k * b.sw * b.sh] <= 0.0f;
}

/* ---------------- SDF face fractions (FS_STEP_FACE_FRACTIONS) ----------------
 * No reference counterpart: the reference voxelises the body at cell centres
 * (MarkBodyCellsFunctor, below).  The option's algorithm, stated once here:
 *   phi(x): the body's SDF samples, taken at the points fileOrigin + cs (i, j, k)
 *           of the body frame, interpolated trilinearly (lerps a (1 - t) + b t);
 *           outside the sampled box the point is outside (phi = FACE_FAR);
 *   a face's open fraction th: the square split along its (00)-(11) diagonal
 *           into two triangles, phi linear on each; the inside part of a
 *           triangle with one corner p on its own side of zero is
 *           (p / (p - q)) (p / (p - r)) (that corner inside) or 1 minus it
 *           (that corner outside); th = 1 - mean of the two; th < 1e-3 -> 0.
 *   Faces between two FLUID/AIR cells get th; faces next to SOLID/BODY cells
 *   keep the reference's wall treatment (th = 0). */
const float FACE_FAR = 1.0e30f, FACE_MIN_OPEN = 1.0e-3f;

float body_phi(const Body& b, f3 wp) {
    f3 pr = wp - b.pos;
    const float* R = b.R;
    f3 plc;
    plc.x = R[0] * pr.x + R[3] * pr.y + R[6] * pr.z;
    plc.y = R[1] * pr.x + R[4] * pr.y + R[7] * pr.z;
    plc.z = R[2] * pr.x + R[5] * pr.y + R[8] * pr.z;
    f3 pl = plc + b.local_com;
    const float sx = (pl.x - b.fileOrigin.x) / b.cs, sy = (pl.y - b.fileOrigin.y) / b.cs,
                sz = (pl.z - b.fileOrigin.z) / b.cs;
    if (!(sx >= 0.0f && sx <= (float)(b.sw - 1) && sy >= 0.0f && sy <= (float)(b.sh - 1) && sz >= 0.0f &&
          sz <= (float)(b.sd - 1)))
        return FACE_FAR;
    const int i = std::min((int)sx, std::max(b.sw - 2, 0)), j = std::min((int)sy, std::max(b.sh - 2, 0)),
              k = std::min((int)sz, std::max(b.sd - 2, 0));
    const float fx = sx - (float)i, fy = sy - (float)j, fz = sz - (float)k;
    auto S = [&](int a, int bb, int c) {
        a = std::min(a, b.sw - 1); bb = std::min(bb, b.sh - 1); c = std::min(c, b.sd - 1);
        return b.sdf[(size_t)a + (size_t)bb * b.sw + (size_t)c * b.sw * b.sh];
    };
    const float c00 = S(i, j, k) * (1.0f - fx) + S(i + 1, j, k) * fx;
    const float c10 = S(i, j + 1, k) * (1.0f - fx) + S(i + 1, j + 1, k) * fx;
    const float c01 = S(i, j, k + 1) * (1.0f - fx) + S(i + 1, j, k + 1) * fx;
    const float c11 = S(i, j + 1, k + 1) * (1.0f - fx) + S(i + 1, j + 1, k + 1) * fx;
    const float c0 = c00 * (1.0f - fy) + c10 * fy, c1 = c01 * (1.0f - fy) + c11 * fy;
    return c0 * (1.0f - fz) + c1 * fz;
}

/* inside fraction of a triangle with distances a, b, c at its corners */
float tri_inside(float a, float b, float c) {
    const float v[3] = {a, b, c};
    int neg = 0;
    for (float e : v) neg += e < 0.0f;
    if (neg == 0) return 0.0f;
    if (neg == 3) return 1.0f;
    /* the corner alone on its side of zero, then the other two in cyclic order */
    int o = 0;
    for (int t = 0; t < 3; ++t)
        if ((v[t] < 0.0f) == (neg == 1)) { o = t; break; }
    const float p = v[o], q = v[(o + 1) % 3], r = v[(o + 2) % 3];
    const float part = (p / (p - q)) * (p / (p - r));
    return neg == 1 ? part : 1.0f - part;
}

float face_open(float p00, float p10, float p01, float p11) {
    const float inside = 0.5f * (tri_inside(p00, p10, p11) + tri_inside(p00, p11, p01));
    const float open = 1.0f - inside;
    return open < FACE_MIN_OPEN ? 0.0f : open;
}

bool face_fractions_on(const Solver& s) { return (s.cfg.step_flags & FS_STEP_FACE_FRACTIONS) && s.body.active; }

/* open fraction of face (fi, fj, fk) of component comp from the SDF at its corners */
float face_open_sdf(const Solver& s, int comp, int fi, int fj, int fk) {
    if (!face_fractions_on(s)) return 1.0f;
    const float dx = s.cfg.dx;
    const float x0 = (float)fi * dx, y0 = (float)fj * dx, z0 = (float)fk * dx;
    const float x1 = (float)(fi + 1) * dx, y1 = (float)(fj + 1) * dx, z1 = (float)(fk + 1) * dx;
    f3 c[4];
    if (comp == 0) { c[0] = mk(x0, y0, z0); c[1] = mk(x0, y1, z0); c[2] = mk(x0, y0, z1); c[3] = mk(x0, y1, z1); }
    else if (comp == 1) { c[0] = mk(x0, y0, z0); c[1] = mk(x1, y0, z0); c[2] = mk(x0, y0, z1); c[3] = mk(x1, y0, z1); }
    else { c[0] = mk(x0, y0, z0); c[1] = mk(x1, y0, z0); c[2] = mk(x0, y1, z0); c[3] = mk(x1, y1, z0); }
    return face_open(body_phi(s.body, c[0]), body_phi(s.body, c[1]), body_phi(s.body, c[2]), body_phi(s.body, c[3]));
}

/* the operator's open fraction of every face of component comp */
std::vector<float> open_fractions(const Solver& s, int comp) {
    const int W = s.W, H = s.H, D = s.D;
    const int FW = W + (comp == 0), FH = H + (comp == 1), FD = D + (comp == 2);
    std::vector<float> th((size_t)FW * FH * FD, 0.0f);
#pragma omp parallel for schedule(static)
    for (long long q = 0; q < (long long)th.size(); ++q) {
        const int fi = (int)(q % FW), fj = (int)((q / FW) % FH), fk = (int)(q / ((long long)FW * FH));
        const int li = fi - (comp == 0), lj = fj - (comp == 1), lk = fk - (comp == 2);
        if (li < 0 || lj < 0 || lk < 0 || fi >= W || fj >= H || fk >= D) continue;
        const int Ll = s.labels[(size_t)li + (size_t)lj * W + (size_t)lk * W * H];
        const int Lr = s.labels[(size_t)fi + (size_t)fj * W + (size_t)fk * W * H];
        if ((Ll == FS_FLUID || Ll == FS_AIR) && (Lr == FS_FLUID || Lr == FS_AIR))
            th[(size_t)q] = face_open_sdf(s, comp, fi, fj, fk);
    }
    return th;
}

/* MarkBodyCellsFunctor FluidSolver3D.cu:45-78 */
void mark_body(Solver& s) {
    const float dx = s.cfg.dx;
    const int W = s.W, H = s.H;
#pragma omp parallel for schedule(static)
    for (long long idx = 0; idx < (long long)s.C(); ++idx) {
        int i = (int)(idx % W), j = (int)((idx / W) % H), k = (int)(idx / ((long long)W * H));
        f3 cc = mk((i + 0.5f) * dx, (j + 0.5f) * dx, (k + 0.5f) * dx);
        if (body_contains(s.body, cc) && s.labels[idx] != FS_SOLID) s.labels[idx] = FS_BODY;
    }
}

/* MarkFluidCellsFunctor key computation FluidSolver3D.cu:322-332 */
inline int cell_key(float px, float py, float pz, float dx, int W, int H, int D) {
    int i = (int)floorf(px / dx);
    int j = (int)floorf(py / dx);
    int k = (int)floorf(pz / dx);
    if (i < 0) i = 0;
    if (i >= W) i = W - 1;
    if (j < 0) j = 0;
    if (j >= H) j = H - 1;
    if (k < 0) k = 0;
    if (k >= D) k = D - 1;
    return i + j * W + k * (W * H);
}

/* labelGrid FluidSolver3D.cu:340-397 */
void stage_label(Solver& s) {
    for (size_t c = 0; c < s.C(); ++c) s.labels[c] = s.solid[c] ? FS_SOLID : FS_AIR;
    if (s.body.active) mark_body(s);
    const int64_t n = s.np();
    s.key.resize((size_t)n);
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < n; ++q)
        s.key[(size_t)q] = cell_key(s.pos[3 * q], s.pos[3 * q + 1], s.pos[3 * q + 2], s.cfg.dx, s.W, s.H, s.D);
    for (int64_t q = 0; q < n; ++q) {
        int idx = s.key[(size_t)q];
        if (s.labels[idx] != FS_SOLID && s.labels[idx] != FS_BODY) s.labels[idx] = FS_FLUID;
    }
}

/* ---------------- MAC base index / fraction (FluidSolver3D.cu:476-571) ------ */
struct Base { int i, j, k; float fx, fy, fz; };
inline float clamp01(float f) { return fminf(fmaxf(f, 0.0f), 1.0f); }
/* comp 0 = u, 1 = v, 2 = w */
inline Base mac_base(int comp, float rx, float ry, float rz, int W, int H, int D) {
    Base b;
    if (comp == 0) {
        b.i = (int)floorf(rx); b.j = (int)floorf(ry - 0.5f); b.k = (int)floorf(rz - 0.5f);
        b.i = std::min(std::max(b.i, 0), W);
        b.j = std::min(std::max(b.j, 0), H - 1);
        b.k = std::min(std::max(b.k, 0), D - 1);
        b.fx = rx - b.i; b.fy = (ry - 0.5f) - b.j; b.fz = (rz - 0.5f) - b.k;
    } else if (comp == 1) {
        b.i = (int)floorf(rx - 0.5f); b.j = (int)floorf(ry); b.k = (int)floorf(rz - 0.5f);
        b.i = std::min(std::max(b.i, 0), W - 1);
        b.j = std::min(std::max(b.j, 0), H);
        b.k = std::min(std::max(b.k, 0), D - 1);
        b.fx = (rx - 0.5f) - b.i; b.fy = ry - b.j; b.fz = (rz - 0.5f) - b.k;
    } else {
        b.i = (int)floorf(rx - 0.5f); b.j = (int)floorf(ry - 0.5f); b.k = (int)floorf(rz);
        b.i = std::min(std::max(b.i, 0), W - 1);
        b.j = std::min(std::max(b.j, 0), H - 1);
        b.k = std::min(std::max(b.k, 0), D);
        b.fx = (rx - 0.5f) - b.i; b.fy = (ry - 0.5f) - b.j; b.fz = rz - b.k;
    }
    b.fx = clamp01(b.fx); b.fy = clamp01(b.fy); b.fz = clamp01(b.fz);
    return b;
}

/* particlesToGrid FluidSolver3D.cu:607-685 (AccumulateVelocities :468-594,
 * ComputeVelocityFunc :598-605), then saveVelocities :399-415 and
 * applyForces :700-761 in frameStep order. */
void stage_p2g(Solver& s) {
    const int W = s.W, H = s.H, D = s.D;
    const float dx = s.cfg.dx;
    std::vector<double> nu(s.NU(), 0.0), du(s.NU(), 0.0), nv(s.NV(), 0.0), dv(s.NV(), 0.0),
        nw(s.NW(), 0.0), dw(s.NW(), 0.0);
    const int nx[3] = {W + 1, W, W};
    const int ny[3] = {H, H + 1, H};
    const int nz[3] = {D, D, D + 1};
    std::vector<double>* num[3] = {&nu, &nv, &nw};
    std::vector<double>* den[3] = {&du, &dv, &dw};
    const int64_t n = s.np();
    for (int64_t q = 0; q < n; ++q) {
        float px = s.pos[3 * q], py = s.pos[3 * q + 1], pz = s.pos[3 * q + 2];
        float rx = px / dx, ry = py / dx, rz = pz / dx;
        for (int c = 0; c < 3; ++c) {
            float vc = s.vel[3 * q + c];
            Base b = mac_base(c, rx, ry, rz, W, H, D);
            for (int dz = 0; dz < 2; ++dz) {
                int k = b.k + dz;
                if (k < 0 || k >= nz[c]) continue;
                for (int dy = 0; dy < 2; ++dy) {
                    int j = b.j + dy;
                    if (j < 0 || j >= ny[c]) continue;
                    for (int dxi = 0; dxi < 2; ++dxi) {
                        int i = b.i + dxi;
                        if (i < 0 || i >= nx[c]) continue;
                        float wx = (dxi == 0 ? (1.0f - b.fx) : b.fx);
                        float wy = (dy == 0 ? (1.0f - b.fy) : b.fy);
                        float wz = (dz == 0 ? (1.0f - b.fz) : b.fz);
                        float wgt = wx * wy * wz;
                        size_t idx = (size_t)i + (size_t)j * nx[c] + (size_t)k * nx[c] * ny[c];
                        (*num[c])[idx] += (double)(vc * wgt);
                        (*den[c])[idx] += (double)wgt;
                    }
                }
            }
        }
    }
    std::vector<float>* g[3] = {&s.u, &s.v, &s.w};
    std::vector<float>* sv[3] = {&s.us, &s.vs, &s.ws};
    const float acc[3] = {0.0f, s.cfg.gravity_y, 0.0f};
    const float vel_unknown = (float)INT_MIN;
    for (int c = 0; c < 3; ++c) {
        std::vector<float>& G = *g[c];
        for (size_t idx = 0; idx < G.size(); ++idx) {
            float fn = (float)(*num[c])[idx], fd = (float)(*den[c])[idx];
            float val = (fd > 1e-8f) ? fn / fd : 0.0f;
            (*sv[c])[idx] = val;                                          /* saveVelocities */
            G[idx] = (val > vel_unknown) ? val + s.dt_cur * acc[c] : val;  /* ApplyScalarForce :695-697 */
        }
    }
}

/* RHSCalculator3D FluidSolver3D.cu:1360-1474 with scale = rho*dx/dt (:1520) */
void stage_rhs(Solver& s) {
    const int W = s.W, H = s.H;
    const bool ff = face_fractions_on(s);
    const float dx = s.cfg.dx;
    const float scale = (s.cfg.density * dx) / s.dt_cur;
    const f3 vel_com = s.body.vel, omega = s.body.omega, com = s.body.pos;
    auto body_vel = [&](f3 pos) -> f3 {
        f3 r = mk(pos.x - com.x, pos.y - com.y, pos.z - com.z);
        return mk(vel_com.x + omega.y * r.z - omega.z * r.y,
                  vel_com.y + omega.z * r.x - omega.x * r.z,
                  vel_com.z + omega.x * r.y - omega.y * r.x);
    };
    const std::vector<float>&u = s.u, &v = s.v, &w = s.w;
#pragma omp parallel for schedule(static)
    for (long long idx = 0; idx < (long long)s.C(); ++idx) {
        int i = (int)(idx % W), j = (int)((idx / W) % H), k = (int)(idx / ((long long)W * H));
        if (s.labels[idx] != FS_FLUID) { s.rhs[idx] = 0.0f; continue; }
        size_t ui0 = (size_t)i + (size_t)j * (W + 1) + (size_t)k * (W + 1) * H;
        size_t vi0 = (size_t)i + (size_t)j * W + (size_t)k * W * (H + 1);
        size_t wi0 = (size_t)i + (size_t)j * W + (size_t)k * W * H;
        float div = u[ui0 + 1] - u[ui0] + v[vi0 + W] - v[vi0] + w[wi0 + (size_t)W * H] - w[wi0];
        float rhs_val = -scale * div;
        if (i - 1 >= 0) {
            int L = label_at(s, idx - 1);
            if (L == FS_SOLID || L == FS_BODY) {
                f3 fp = mk(i * dx, (j + 0.5f) * dx, (k + 0.5f) * dx);
                float us_ = (L == FS_BODY) ? body_vel(fp).x : 0.0f;
                rhs_val -= scale * (u[ui0] - us_);
            } else if (ff) {   /* a FLUID/AIR neighbour: the SDF's open fraction of the face */
                const float th = face_open_sdf(s, 0, i, j, k);
                if (th < 1.0f) {
                    f3 fp = mk(i * dx, (j + 0.5f) * dx, (k + 0.5f) * dx);
                    rhs_val -= scale * ((1.0f - th) * (u[ui0] - body_vel(fp).x));
                }
            }
        }
        {   /* i < W always holds: the plus-side read is unguarded (:1406) */
            int L = label_at(s, idx + 1);
            if (L == FS_SOLID || L == FS_BODY) {
                f3 fp = mk((i + 1) * dx, (j + 0.5f) * dx, (k + 0.5f) * dx);
                float us_ = (L == FS_BODY) ? body_vel(fp).x : 0.0f;
                rhs_val += scale * (u[ui0 + 1] - us_);
            } else if (ff) {
                const float th = face_open_sdf(s, 0, i + 1, j, k);
                if (th < 1.0f) {
                    f3 fp = mk((i + 1) * dx, (j + 0.5f) * dx, (k + 0.5f) * dx);
                    rhs_val += scale * ((1.0f - th) * (u[ui0 + 1] - body_vel(fp).x));
                }
            }
        }
        if (j - 1 >= 0) {
            int L = label_at(s, idx - W);
            if (L == FS_SOLID || L == FS_BODY) {
                f3 fp = mk((i + 0.5f) * dx, j * dx, (k + 0.5f) * dx);
                float vs_ = (L == FS_BODY) ? body_vel(fp).y : 0.0f;
                rhs_val -= scale * (v[vi0] - vs_);
            } else if (ff) {
                const float th = face_open_sdf(s, 1, i, j, k);
                if (th < 1.0f) {
                    f3 fp = mk((i + 0.5f) * dx, j * dx, (k + 0.5f) * dx);
                    rhs_val -= scale * ((1.0f - th) * (v[vi0] - body_vel(fp).y));
                }
            }
        }
        {
            int L = label_at(s, idx + W);
            if (L == FS_SOLID || L == FS_BODY) {
                f3 fp = mk((i + 0.5f) * dx, (j + 1) * dx, (k + 0.5f) * dx);
                float vs_ = (L == FS_BODY) ? body_vel(fp).y : 0.0f;
                rhs_val += scale * (v[vi0 + W] - vs_);
            } else if (ff) {
                const float th = face_open_sdf(s, 1, i, j + 1, k);
                if (th < 1.0f) {
                    f3 fp = mk((i + 0.5f) * dx, (j + 1) * dx, (k + 0.5f) * dx);
                    rhs_val += scale * ((1.0f - th) * (v[vi0 + W] - body_vel(fp).y));
                }
            }
        }
        if (k - 1 >= 0) {
            int L = label_at(s, idx - (long long)W * H);
            if (L == FS_SOLID || L == FS_BODY) {
                f3 fp = mk((i + 0.5f) * dx, (j + 0.5f) * dx, k * dx);
                float ws_ = (L == FS_BODY) ? body_vel(fp).z : 0.0f;
                rhs_val -= scale * (w[wi0] - ws_);
            } else if (ff) {
                const float th = face_open_sdf(s, 2, i, j, k);
                if (th < 1.0f) {
                    f3 fp = mk((i + 0.5f) * dx, (j + 0.5f) * dx, k * dx);
                    rhs_val -= scale * ((1.0f - th) * (w[wi0] - body_vel(fp).z));
                }
            }
        }
        {
            int L = label_at(s, idx + (long long)W * H);
            if (L == FS_SOLID || L == FS_BODY) {
                f3 fp = mk((i + 0.5f) * dx, (j + 0.5f) * dx, (k + 1) * dx);
                float ws_ = (L == FS_BODY) ? body_vel(fp).z : 0.0f;
                rhs_val += scale * (w[wi0 + (size_t)W * H] - ws_);
            } else if (ff) {
                const float th = face_open_sdf(s, 2, i, j, k + 1);
                if (th < 1.0f) {
                    f3 fp = mk((i + 0.5f) * dx, (j + 0.5f) * dx, (k + 1) * dx);
                    rhs_val += scale * ((1.0f - th) * (w[wi0 + (size_t)W * H] - body_vel(fp).z));
                }
            }
        }
        s.rhs[idx] = rhs_val;
    }
}

/* ---------------- pressure system (constructA :1718-1777) ------------------- */
struct Csr {
    int n = 0;
    std::vector<int> off, col;      /* full symmetric pattern */
    std::vector<double> val;
    std::vector<int> cell;          /* fluid number -> cell index */
};

/* Open fraction of the face between cell (i, j, k) and its neighbour n (order
 * +x, -x, +y, -y, +z, -z) for the face-fraction operator; null: every face 1. */
using FaceFn = std::function<float(int i, int j, int k, int n)>;

/* exclusive-scan numbering (:1830-1853) and the SPD matrix of MatrixFiller3D
 * (:1656-1716): diag = #in-domain FLUID/AIR neighbours, -1 per FLUID neighbour.
 * With open fractions th (faces to FLUID/AIR neighbours with th > 0 only):
 * diag = their FP32 sum in the neighbour order above, -th per FLUID neighbour. */
Csr assemble(const std::vector<uint8_t>& L, int W, int H, int D, std::vector<int>& num,
             const FaceFn& open = nullptr) {
    size_t C = (size_t)W * H * D;
    num.assign(C, -1);
    Csr A;
    for (size_t c = 0; c < C; ++c)
        if (L[c] == FS_FLUID) { num[c] = A.n++; A.cell.push_back((int)c); }
    A.off.assign((size_t)A.n + 1, 0);
    const int off[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (int r = 0; r < A.n; ++r) {
        int c = A.cell[r];
        int i = c % W, j = (c / W) % H, k = c / (W * H);
        double diag = 0;
        float fdiag = 0.0f;
        std::vector<std::pair<int, double>> row;
        for (int n = 0; n < 6; ++n) {
            int ni = i + off[n][0], nj = j + off[n][1], nk = k + off[n][2];
            if (ni < 0 || ni >= W || nj < 0 || nj >= H || nk < 0 || nk >= D) continue;
            size_t nidx = (size_t)ni + (size_t)nj * W + (size_t)nk * W * H;
            if (L[nidx] != FS_FLUID && L[nidx] != FS_AIR) continue;
            const float th = open ? open(i, j, k, n) : 1.0f;
            if (th == 0.0f) continue;
            if (L[nidx] == FS_FLUID) row.push_back({num[nidx], -(double)th});
            diag += 1.0;
            fdiag = fdiag + th;
        }
        row.push_back({r, open ? (double)fdiag : diag});
        std::sort(row.begin(), row.end());
        for (auto& e : row) { A.col.push_back(e.first); A.val.push_back(e.second); }
        A.off[r + 1] = (int)A.col.size();
    }
    return A;
}

void spmv(const Csr& A, const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int r = 0; r < A.n; ++r) {
        double acc = 0;
        for (int e = A.off[r]; e < A.off[r + 1]; ++e) acc += A.val[e] * x[A.col[e]];
        y[r] = acc;
    }
}

double dotd(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long long i = 0; i < (long long)a.size(); ++i) s += a[i] * b[i];
    return s;
}

/* Dense-band LDL^T of a symmetric matrix given in full CSR; returns false if a
 * pivot is not positive (singular / indefinite). */
bool band_ldlt_solve(const Csr& A, const std::vector<double>& b, std::vector<double>& x) {
    int n = A.n, bw = 0;
    for (int r = 0; r < n; ++r)
        for (int e = A.off[r]; e < A.off[r + 1]; ++e) bw = std::max(bw, std::abs(A.col[e] - r));
    /* L stored row-wise: Lb[r*(bw+1) + (r - c)] for c in [r-bw, r] */
    std::vector<double> M((size_t)n * (bw + 1), 0.0);
    auto at = [&](int r, int c) -> double& { return M[(size_t)r * (bw + 1) + (r - c)]; };
    for (int r = 0; r < n; ++r)
        for (int e = A.off[r]; e < A.off[r + 1]; ++e)
            if (A.col[e] <= r) at(r, A.col[e]) = A.val[e];
    std::vector<double> d(n);
    for (int j = 0; j < n; ++j) {
        double dj = at(j, j);
        int k0 = std::max(0, j - bw);
        for (int k = k0; k < j; ++k) dj -= at(j, k) * at(j, k) * d[k];
        if (!(dj > 0)) return false;
        d[j] = dj;
        int iend = std::min(n - 1, j + bw);
        for (int i = j + 1; i <= iend; ++i) {
            double lij = at(i, j);
            int kk = std::max(k0, i - bw);
            for (int k = kk; k < j; ++k) lij -= at(i, k) * at(j, k) * d[k];
            at(i, j) = lij / dj;
        }
    }
    x = b;
    for (int i = 0; i < n; ++i)
        for (int k = std::max(0, i - bw); k < i; ++k) x[i] -= at(i, k) * x[k];
    for (int i = 0; i < n; ++i) x[i] /= d[i];
    for (int i = n - 1; i >= 0; --i)
        for (int k = i + 1; k <= std::min(n - 1, i + bw); ++k) x[i] -= at(k, i) * x[k];
    return true;
}

/* FP64 PCG with a modified-incomplete-Cholesky MIC(0) preconditioner on the
 * structured 7-point pattern.  Stops at ||r|| <= tol ||b||.  When the system is
 * singular (no Dirichlet contact), b and r are kept orthogonal to constants. */
int pcg64(const Csr& A, const std::vector<double>& b, std::vector<double>& x, double tol,
          int maxit, bool singular, double* relres) {
    int n = A.n;
    x.assign(n, 0.0);
    std::vector<double> r = b, z(n), pdir(n), q(n);
    auto project = [&](std::vector<double>& vv) {
        if (!singular) return;
        double m = 0;
        for (double e : vv) m += e;
        m /= n;
        for (double& e : vv) e -= m;
    };
    project(r);
    double bn = std::sqrt(dotd(b, b));
    if (bn == 0) { *relres = 0; return 0; }
    /* MIC(0) factor: diagonal of L D L^T with L sharing A's strict lower pattern */
    std::vector<double> pd(n, 0.0);
    {
        const double tau = 0.97, sigma = 0.25;
        for (int r0 = 0; r0 < n; ++r0) {
            double e = 0, dg = 0;
            for (int k = A.off[r0]; k < A.off[r0 + 1]; ++k) if (A.col[k] == r0) dg = A.val[k];
            e = dg;
            for (int k = A.off[r0]; k < A.off[r0 + 1]; ++k) {
                int c = A.col[k];
                if (c >= r0) continue;
                double a = A.val[k];
                e -= (a * pd[c]) * (a * pd[c]);
                /* modified term: fill that would land on other upper neighbours of c */
                double fill = 0;
                for (int k2 = A.off[c]; k2 < A.off[c + 1]; ++k2) {
                    int c2 = A.col[k2];
                    if (c2 > c && c2 != r0) fill += A.val[k2];
                }
                e -= tau * a * fill * pd[c] * pd[c];
            }
            if (e < sigma * dg) e = dg;
            pd[r0] = (e > 0) ? 1.0 / std::sqrt(e) : 0.0;
        }
    }
    auto apply_prec = [&](const std::vector<double>& rr, std::vector<double>& zz) {
        std::vector<double> t(n);
        for (int r0 = 0; r0 < n; ++r0) {
            double acc = rr[r0];
            for (int k = A.off[r0]; k < A.off[r0 + 1]; ++k) {
                int c = A.col[k];
                if (c < r0) acc -= A.val[k] * pd[c] * t[c];
            }
            t[r0] = acc * pd[r0];
        }
        for (int r0 = n - 1; r0 >= 0; --r0) {
            double acc = t[r0];
            for (int k = A.off[r0]; k < A.off[r0 + 1]; ++k) {
                int c = A.col[k];
                if (c > r0) acc -= A.val[k] * pd[r0] * zz[c];
            }
            zz[r0] = acc * pd[r0];
        }
        project(zz);
    };
    apply_prec(r, z);
    pdir = z;
    double rz = dotd(r, z);
    int it = 0;
    double rn = std::sqrt(dotd(r, r));
    while (rn > tol * bn && it < maxit) {
        spmv(A, pdir.data(), q.data());
        double pq = dotd(pdir, q);
        if (pq <= 0) break;
        double alpha = rz / pq;
        for (int i = 0; i < n; ++i) { x[i] += alpha * pdir[i]; r[i] -= alpha * q[i]; }
        project(r);
        rn = std::sqrt(dotd(r, r));
        ++it;
        if (rn <= tol * bn) break;
        apply_prec(r, z);
        double rz2 = dotd(r, z);
        double beta = rz2 / rz;
        rz = rz2;
        for (int i = 0; i < n; ++i) pdir[i] = z[i] + beta * pdir[i];
    }
    /* true residual */
    spmv(A, x.data(), q.data());
    std::vector<double> tr(n);
    for (int i = 0; i < n; ++i) tr[i] = b[i] - q[i];
    project(tr);
    *relres = std::sqrt(dotd(tr, tr)) / bn;
    return it;
}

bool system_singular(const std::vector<uint8_t>& L, int W, int H, int D, const FaceFn& open = nullptr) {
    /* the constant vector is in the null space iff no fluid cell has an AIR
     * neighbour (through an open face) */
    for (int k = 0; k < D; ++k)
        for (int j = 0; j < H; ++j)
            for (int i = 0; i < W; ++i) {
                size_t c = (size_t)i + (size_t)j * W + (size_t)k * W * H;
                if (L[c] != FS_FLUID) continue;
                const int off[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
                for (int n = 0; n < 6; ++n) {
                    int ni = i + off[n][0], nj = j + off[n][1], nk = k + off[n][2];
                    if (ni < 0 || ni >= W || nj < 0 || nj >= H || nk < 0 || nk >= D) continue;
                    if (L[(size_t)ni + (size_t)nj * W + (size_t)nk * W * H] == FS_AIR &&
                        (!open || open(i, j, k, n) > 0.0f))
                        return false;
                }
            }
    return true;
}

/* Rows of FLUID cells with no FLUID or AIR neighbour are empty (diagonal 0,
 * no off-diagonal, and no other row refers to them): they are dropped from
 * the system, so those cells keep p = 0 and their b is not part of the
 * residual.  The reference would hand such a singular matrix to cuDSS. */
void drop_empty_rows(Csr& A) {
    std::vector<int> keep(A.n, -1);
    int m = 0;
    for (int r = 0; r < A.n; ++r) {
        bool empty = true;
        for (int e = A.off[r]; e < A.off[r + 1]; ++e)
            if (A.val[e] != 0.0) empty = false;
        if (!empty) keep[r] = m++;
    }
    if (m == A.n) return;
    Csr B;
    B.n = m;
    B.off.assign((size_t)m + 1, 0);
    for (int r = 0; r < A.n; ++r) {
        if (keep[r] < 0) continue;
        for (int e = A.off[r]; e < A.off[r + 1]; ++e) {
            B.col.push_back(keep[A.col[e]]);
            B.val.push_back(A.val[e]);
        }
        B.off[keep[r] + 1] = (int)B.col.size();
        B.cell.push_back(A.cell[r]);
    }
    A = std::move(B);
}

/* Solve A p = b for the full-grid b; p64 full grid (0 off fluid). */
int solve_system(const std::vector<uint8_t>& L, const float* bfull, double* pfull, int W, int H, int D,
                 double tol, int maxit, fs_solve_stats* st, const FaceFn& open = nullptr) {
    std::vector<int> num;
    Csr A = assemble(L, W, H, D, num, open);
    drop_empty_rows(A);
    size_t C = (size_t)W * H * D;
    fs_solve_stats loc{};
    loc.fluid_cells = A.n;
    for (size_t c = 0; c < C; ++c) pfull[c] = 0.0;
    if (A.n == 0) {
        if (st) *st = loc;
        if (num.empty() || std::find_if(num.begin(), num.end(), [](int v) { return v >= 0; }) == num.end())
            return fail(FS_ERR_NO_FLUID, "No fluid cells found!");
        loc.converged = 1;
        if (st) *st = loc;
        return FS_OK;   /* FLUID cells, but none is an unknown */
    }
    std::vector<double> b(A.n), x;
    for (int r = 0; r < A.n; ++r) b[r] = (double)bfull[A.cell[r]];
    bool sing = system_singular(L, W, H, D, open);
    loc.singular = sing ? 1 : 0;
    auto t0 = clk::now();
    double rr = 0;
    bool direct_ok = false;
    if (!sing && A.n <= 20000) {
        direct_ok = band_ldlt_solve(A, b, x);
        if (direct_ok) {
            std::vector<double> q(A.n);
            spmv(A, x.data(), q.data());
            double bn = std::sqrt(dotd(b, b)), rn = 0;
            for (int r = 0; r < A.n; ++r) rn += (b[r] - q[r]) * (b[r] - q[r]);
            rr = bn > 0 ? std::sqrt(rn) / bn : 0;
            loc.iterations = 0;
        }
    }
    if (!direct_ok) {
        if (sing) {
            double m = 0;
            for (double e : b) m += e;
            m /= A.n;
            for (double& e : b) e -= m;
        }
        loc.iterations = pcg64(A, b, x, tol, maxit, sing, &rr);
    }
    loc.solve_ms = ms_since(t0);
    loc.rel_residual = rr;
    loc.converged = (rr <= std::max(tol, 1e-14) * 1.0000001 || direct_ok) ? 1 : 0;
    loc.outer = 1;
    for (int r = 0; r < A.n; ++r) pfull[A.cell[r]] = x[r];
    if (st) *st = loc;
    if (!loc.converged) return fail(FS_ERR_NOT_CONVERGED, "pressure solve did not reach the tolerance");
    return FS_OK;
}

/* pressureSolve minus RHS (FluidSolver3D.cu:1815-2062): p = solution on FLUID, 0 elsewhere */
/* the face-fraction operator's open fraction of the face between cell
 * (i, j, k) and its neighbour n (+x, -x, +y, -y, +z, -z) */
FaceFn face_fn(const Solver& s) {
    if (!face_fractions_on(s)) return nullptr;
    return [&s](int i, int j, int k, int n) {
        switch (n) {
            case 0: return face_open_sdf(s, 0, i + 1, j, k);
            case 1: return face_open_sdf(s, 0, i, j, k);
            case 2: return face_open_sdf(s, 1, i, j + 1, k);
            case 3: return face_open_sdf(s, 1, i, j, k);
            case 4: return face_open_sdf(s, 2, i, j, k + 1);
            default: return face_open_sdf(s, 2, i, j, k);
        }
    };
}

int stage_pressure(Solver& s) {
    int rc = solve_system(s.labels, s.rhs.data(), s.p64.data(), s.W, s.H, s.D, s.cfg.solver_tol,
                          s.cfg.solver_max_iter, &s.stats, face_fn(s));
    for (size_t c = 0; c < s.C(); ++c) s.p[c] = (s.labels[c] == FS_FLUID) ? (float)s.p64[c] : 0.0f;
    return rc;
}

/* UFunctor/VFunctor/WFunctor FluidSolver3D.cu:2072-2336, scale = dt/(rho*dx) (:2341) */
void stage_project(Solver& s) {
    const int W = s.W, H = s.H;
    const bool ff = face_fractions_on(s);
    const float dx = s.cfg.dx;
    const float scale = s.dt_cur / (s.cfg.density * dx);
    const f3 vel_com = s.body.vel, omega = s.body.omega, com = s.body.pos;
    for (int comp = 0; comp < 3; ++comp) {
        std::vector<float>& G = comp == 0 ? s.u : comp == 1 ? s.v : s.w;
        const long long n = (long long)G.size();
#pragma omp parallel for schedule(static)
        for (long long idx = 0; idx < n; ++idx) {
            int fi, fj, fk;
            long long left, right;
            f3 fp;
            if (comp == 0) {
                long long slice = (long long)(W + 1) * H;
                fk = (int)(idx / slice); long long rem = idx % slice;
                fj = (int)(rem / (W + 1)); fi = (int)(rem % (W + 1));
                left = (long long)(fi - 1) + (long long)fj * W + (long long)fk * W * H;
                right = (long long)fi + (long long)fj * W + (long long)fk * W * H;
                fp = mk(fi * dx, (fj + 0.5f) * dx, (fk + 0.5f) * dx);
            } else if (comp == 1) {
                long long slice = (long long)W * (H + 1);
                fk = (int)(idx / slice); long long rem = idx % slice;
                fj = (int)(rem / W); fi = (int)(rem % W);
                left = (long long)fi + (long long)(fj - 1) * W + (long long)fk * W * H;
                right = (long long)fi + (long long)fj * W + (long long)fk * W * H;
                fp = mk((fi + 0.5f) * dx, fj * dx, (fk + 0.5f) * dx);
            } else {
                long long slice = (long long)W * H;
                fk = (int)(idx / slice); long long rem = idx % slice;
                fj = (int)(rem / W); fi = (int)(rem % W);
                left = (long long)fi + (long long)fj * W + (long long)(fk - 1) * W * H;
                right = (long long)fi + (long long)fj * W + (long long)fk * W * H;
                fp = mk((fi + 0.5f) * dx, (fj + 0.5f) * dx, fk * dx);
            }
            float invScale = 1.0f / scale;
            f3 r = mk(fp.x - com.x, fp.y - com.y, fp.z - com.z);
            f3 vb = mk(vel_com.x + omega.y * r.z - omega.z * r.y,
                       vel_com.y + omega.z * r.x - omega.x * r.z,
                       vel_com.z + omega.x * r.y - omega.y * r.x);
            float vbc = comp == 0 ? vb.x : comp == 1 ? vb.y : vb.z;
            int Ll = label_at(s, left), Lr = label_at(s, right);
            float p0, p1, g = G[idx];
            if ((Ll == FS_FLUID || Ll == FS_AIR) && (Lr == FS_FLUID || Lr == FS_AIR)) {
                p0 = p_at(s, left); p1 = p_at(s, right);
                /* face fractions: a face the SDF closes completely takes the body's velocity */
                if (ff && face_open_sdf(s, comp, fi, fj, fk) == 0.0f) p1 = p0 + invScale * (g - vbc);
            } else if (Ll == FS_SOLID || Ll == FS_BODY) {
                float sol = (Ll == FS_BODY) ? vbc : 0.0f;
                p0 = p_at(s, right) - invScale * (g - sol);
                p1 = p_at(s, right);
            } else if (Lr == FS_SOLID || Lr == FS_BODY) {
                float sol = (Lr == FS_BODY) ? vbc : 0.0f;
                p0 = p_at(s, left);
                p1 = p_at(s, left) + invScale * (g - sol);
            } else {
                p0 = p_at(s, left); p1 = p_at(s, right);
            }
            G[idx] = g - scale * (p1 - p0);
        }
    }
}

/* trilinearInterpolation FluidSolver3D.cu:784-801.  Each lerp a*(1-t) + b*t is
   one fused multiply-add over the rounded product b*t, the form a CUDA build
   contracts it to by default; the HIP kernels use the same expression. */
inline float trilin(float xd, float yd, float zd, float c000, float c100, float c001, float c101,
                    float c010, float c110, float c011, float c111) {
    float c00 = fmaf(c000, 1.0f - xd, c100 * xd);
    float c01 = fmaf(c001, 1.0f - xd, c101 * xd);
    float c10 = fmaf(c010, 1.0f - xd, c110 * xd);
    float c11 = fmaf(c011, 1.0f - xd, c111 * xd);
    float c0 = fmaf(c00, 1.0f - yd, c10 * yd);
    float c1 = fmaf(c01, 1.0f - yd, c11 * yd);
    return fmaf(c0, 1.0f - zd, c1 * zd);
}

/* 8-corner sample of one MAC component at base b (reads clamped into the array) */
inline float sample(const float* f, size_t n, const Base& b, int sx, int sxy) {
    auto at = [&](int i, int j, int k) -> float {
        long long idx = (long long)i + (long long)j * sx + (long long)k * sxy;
        if (idx < 0) idx = 0;
        if (idx >= (long long)n) idx = (long long)n - 1;
        return f[idx];
    };
    return trilin(b.fx, b.fy, b.fz, at(b.i, b.j, b.k), at(b.i + 1, b.j, b.k), at(b.i, b.j, b.k + 1),
                  at(b.i + 1, b.j, b.k + 1), at(b.i, b.j + 1, b.k), at(b.i + 1, b.j + 1, b.k),
                  at(b.i, b.j + 1, b.k + 1), at(b.i + 1, b.j + 1, b.k + 1));
}

struct Fields {
    const float *u, *v, *w;
    size_t nu, nv, nw;
    int W, H, D;
    float dx;
};

/* interpVelDevice3D FluidSolver3D.cu:1051-1157 */
inline f3 interp_vel(const Fields& F, f3 pos) {
    float rx = pos.x / F.dx, ry = pos.y / F.dx, rz = pos.z / F.dx;
    Base bu = mac_base(0, rx, ry, rz, F.W, F.H, F.D);
    Base bv = mac_base(1, rx, ry, rz, F.W, F.H, F.D);
    Base bw = mac_base(2, rx, ry, rz, F.W, F.H, F.D);
    return mk(sample(F.u, F.nu, bu, F.W + 1, (F.W + 1) * F.H),
              sample(F.v, F.nv, bv, F.W, F.W * (F.H + 1)),
              sample(F.w, F.nw, bw, F.W, F.W * F.H));
}

/* projectParticleDevice3D FluidSolver3D.cu:1161-1239 */
const int kProjOff[34][3] = {
    {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
    {1, 1, 0}, {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0},
    {1, 0, 1}, {-1, 0, 1}, {1, 0, -1}, {-1, 0, -1},
    {0, 1, 1}, {0, -1, 1}, {0, 1, -1}, {0, -1, -1},
    {1, 1, 1}, {-1, 1, 1}, {1, -1, 1}, {1, 1, -1},
    {-1, -1, 1}, {-1, 1, -1}, {1, -1, -1}, {-1, -1, -1},
    {-2, -1, 0}, {0, -1, -2}, {-2, -2, 0}, {0, -2, -2},
    {-2, -2, -2}, {-2, 0, -2}, {2, -1, 0}, {0, -1, 2}};

bool project_particle(f3& pos, const uint8_t* labels, int W, int H, int D, float dx) {
    int cx = (int)floorf(pos.x / dx), cy = (int)floorf(pos.y / dx), cz = (int)floorf(pos.z / dx);
    float bestD = 1e10f;
    bool found = false;
    int best = 0;
    for (int n = 0; n < 34; ++n) {
        int nx = cx + kProjOff[n][0], ny = cy + kProjOff[n][1], nz = cz + kProjOff[n][2];
        if (nx < 0 || nx >= W || ny < 0 || ny >= H || nz < 0 || nz >= D) continue;
        if (labels[(size_t)nx + (size_t)ny * W + (size_t)nz * W * H] != FS_FLUID) continue;
        f3 cc = mk((nx + 0.5f) * dx, (ny + 0.5f) * dx, (nz + 0.5f) * dx);
        float d = (cc.x - pos.x) * (cc.x - pos.x) + (cc.y - pos.y) * (cc.y - pos.y) +
                  (cc.z - pos.z) * (cc.z - pos.z);
        if (d < bestD) { bestD = d; found = true; best = n; }
    }
    if (!found) return false;
    pos = pos + mk((float)kProjOff[best][0], (float)kProjOff[best][1], (float)kProjOff[best][2]) * 0.5f * dx;
    return true;
}

/* AdvectParticlesFunctor FluidSolver3D.cu:1260-1330 */
f3 advect_one(const Fields& F, const uint8_t* labels, f3 pos, float dt, float C) {
    const int W = F.W, H = F.H, D = F.D;
    const float dx = F.dx;
    float subT = 0.0f;
    bool finished = false;
    int iter = 0;
    while (!finished && iter++ < 1000) {
        f3 vel = interp_vel(F, pos);
        float speed = sqrtf(vel.x * vel.x + vel.y * vel.y + vel.z * vel.z) + 1e-37f + sqrtf(5.0f * dx * 9.8f);
        float dT = (C * dx) / speed;
        if (subT + dT >= dt) {
            dT = dt - subT;
            finished = true;
        } else if (subT + 2 * dT >= dt) {
            dT *= 0.5f;
        }
        f3 k1 = vel;
        f3 k2 = interp_vel(F, pos + k1 * 0.5f * dT);
        f3 k3 = interp_vel(F, pos + k2 * 0.75f * dT);
        pos = pos + (2.0f / 9.0f) * dT * k1 + (3.0f / 9.0f) * dT * k2 + (4.0f / 9.0f) * dT * k3;
        subT += dT;
        pos.x = fmaxf(pos.x, 0.0f); pos.x = fminf(pos.x, (W - 1) * dx);
        pos.y = fmaxf(pos.y, 0.0f); pos.y = fminf(pos.y, (H - 1) * dx);
        pos.z = fmaxf(pos.z, 0.0f); pos.z = fminf(pos.z, (D - 1) * dx);
        int cx = (int)floorf(pos.x / dx), cy = (int)floorf(pos.y / dx), cz = (int)floorf(pos.z / dx);
        if (cx >= 0 && cx < W && cy >= 0 && cy < H && cz >= 0 && cz < D) {
            uint8_t L = labels[(size_t)cx + (size_t)cy * W + (size_t)cz * W * H];
            if (L == FS_SOLID || L == FS_BODY) {
                if (!project_particle(pos, labels, W, H, D, dx)) {
                    project_particle(pos, labels, W, H, D, dx);
                    break;
                }
            }
        }
        if (pos.x < 0.0f || pos.y < 0.0f || pos.z < 0.0f || std::isnan(pos.x) || std::isnan(pos.y) ||
            std::isnan(pos.z))
            break;
    }
    return pos;
}

/* gridToParticles :983-1039 (GridToParticleFunctor :822-979) then advectParticles(C) :1334-1357 */
void stage_g2p_advect(Solver& s) {
    const int W = s.W, H = s.H, D = s.D;
    std::vector<float> du(s.NU()), dv(s.NV()), dw(s.NW());
    for (size_t i = 0; i < du.size(); ++i) du[i] = s.u[i] - s.us[i];
    for (size_t i = 0; i < dv.size(); ++i) dv[i] = s.v[i] - s.vs[i];
    for (size_t i = 0; i < dw.size(); ++i) dw[i] = s.w[i] - s.ws[i];
    Fields F{s.u.data(), s.v.data(), s.w.data(), s.NU(), s.NV(), s.NW(), W, H, D, s.cfg.dx};
    Fields dF{du.data(), dv.data(), dw.data(), s.NU(), s.NV(), s.NW(), W, H, D, s.cfg.dx};
    const float alpha = s.cfg.pic_weight;
    const int64_t n = s.np();
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t q = 0; q < n; ++q) {
        f3 pos = mk(s.pos[3 * q], s.pos[3 * q + 1], s.pos[3 * q + 2]);
        f3 pic = interp_vel(F, pos);
        f3 flip = interp_vel(dF, pos);
        float nu = alpha * pic.x + (1.0f - alpha) * (s.vel[3 * q] + flip.x);
        float nv = alpha * pic.y + (1.0f - alpha) * (s.vel[3 * q + 1] + flip.y);
        float nw = alpha * pic.z + (1.0f - alpha) * (s.vel[3 * q + 2] + flip.z);
        s.vel[3 * q] = nu; s.vel[3 * q + 1] = nv; s.vel[3 * q + 2] = nw;
        f3 np = advect_one(F, s.labels.data(), pos, s.dt_cur, s.cfg.advect_c);
        s.pos[3 * q] = np.x; s.pos[3 * q + 1] = np.y; s.pos[3 * q + 2] = np.z;
    }
}

/* DragLiftForceCalculator FluidSolver3D.cu:2674-2791 + applyBuoyancy :2917-3026 +
 * updateBody :2793-2846 */
void stage_body(Solver& s) {
    if (!s.body.active) return;
    Body& b = s.body;
    const int W = s.W, H = s.H, D = s.D;
    const float dx = s.cfg.dx, rho = s.cfg.density;
    Fields F{s.u.data(), s.v.data(), s.w.data(), s.NU(), s.NV(), s.NW(), W, H, D, dx};
    double fa[3] = {0, 0, 0}, ta[3] = {0, 0, 0};
    const int di[6] = {1, -1, 0, 0, 0, 0}, dj[6] = {0, 0, 1, -1, 0, 0}, dk[6] = {0, 0, 0, 0, 1, -1};
    const f3 normals[6] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    const f3 fofs[6] = {{0.5f, 0, 0}, {-0.5f, 0, 0}, {0, 0.5f, 0}, {0, -0.5f, 0}, {0, 0, 0.5f}, {0, 0, -0.5f}};
    const f3 cm = b.pos, bvel = b.vel;
    for (size_t idx = 0; idx < s.C(); ++idx) {
        if (s.labels[idx] != FS_BODY) continue;
        int ci = (int)(idx % W), cj = (int)((idx / W) % H), ck = (int)(idx / ((size_t)W * H));
        f3 cc = mk((ci + 0.5f) * dx, (cj + 0.5f) * dx, (ck + 0.5f) * dx);
        float area = dx * dx;
        for (int dir = 0; dir < 6; ++dir) {
            int ni = ci + di[dir], nj = cj + dj[dir], nk = ck + dk[dir];
            if (ni < 0 || ni >= W || nj < 0 || nj >= H || nk < 0 || nk >= D) continue;
            if (s.labels[(size_t)ni + (size_t)nj * W + (size_t)nk * W * H] != FS_FLUID) continue;
            f3 fc = cc + fofs[dir] * dx;
            f3 n = normals[dir];
            f3 r = fc - cm;
            f3 up = interp_vel(F, fc);
            f3 urel = bvel - up;
            float ura = sqrtf(dot(urel, urel));
            float alpha_s = 0.5f, Aeff = 0.0f;
            float nur = dot(n, urel);
            if (nur > 0) Aeff = (nur / ura * alpha_s + (1 - alpha_s)) * area;
            float Cd = 0.1f;
            float sd_ = (float)(-0.5 * (double)rho * (double)Cd * (double)Aeff * (double)ura);
            f3 Fd = sd_ * urel;
            f3 Td = cross(r, Fd);
            float Cl = 0.1f;
            f3 nxu = cross(n, urel);
            float nxa = sqrtf(dot(nxu, nxu));
            f3 Fl = mk(0, 0, 0), Tl = mk(0, 0, 0);
            if (nxa != 0.0f) {
                Fl = -0.5f * rho * Cl * Aeff * ura * cross(urel, nxu / nxa);
                Tl = cross(r, Fl);
            }
            f3 Ft = Fd + Fl, Tt = Td + Tl;
            fa[0] += Ft.x; fa[1] += Ft.y; fa[2] += Ft.z;
            ta[0] += Tt.x; ta[1] += Tt.y; ta[2] += Tt.z;
        }
    }
    f3 totalF = mk((float)fa[0], (float)fa[1], (float)fa[2]);
    f3 totalT = mk((float)ta[0], (float)ta[1], (float)ta[2]);
    const f3 G = mk(0.0f, s.cfg.gravity_y, 0.0f);
    b.force = b.force + b.mass * G;
    b.force = b.force + totalF;
    b.torque = b.torque + totalT;
    /* applyBuoyancy: binning by truncated cell index (:2938-2944) of the
     * start-of-step particles, submerged if the voxel's cell is FLUID or a
     * particle of that cell lies within 2dx (:2993-3004). */
    std::vector<int> cstart(s.C() + 1, 0), order((size_t)s.np());
    auto trunc_cell = [&](f3 p) {
        int i = std::clamp((int)(p.x / dx), 0, W - 1);
        int j = std::clamp((int)(p.y / dx), 0, H - 1);
        int k = std::clamp((int)(p.z / dx), 0, D - 1);
        return (size_t)i + (size_t)j * W + (size_t)k * W * H;
    };
    for (int64_t q = 0; q < s.np(); ++q) {
        cstart[s.key[(size_t)q] + 1]++;
    }
    for (size_t c = 0; c < s.C(); ++c) cstart[c + 1] += cstart[c];
    {
        std::vector<int> cur(cstart.begin(), cstart.end() - 1);
        for (int64_t q = 0; q < s.np(); ++q) order[(size_t)cur[s.key[(size_t)q]]++] = (int)q;
    }
    const std::vector<float>& P0 = s.pos;   /* start-of-step positions held by the caller */
    double sub = 0;
    const float cs = b.cs;
    const float search = 2.0f * dx;
    for (int k = 0; k < b.sd; ++k)
        for (int j = 0; j < b.sh; ++j)
            for (int i = 0; i < b.sw; ++i) {
                float sv = b.sdf[(size_t)i + (size_t)j * b.sw + (size_t)k * b.sw * b.sh];
                if (sv > 0.0f) continue;
                f3 pl = b.fileOrigin + mk(i * cs, j * cs, k * cs);
                f3 pc = pl - b.local_com;
                f3 pr = mk(b.R[0] * pc.x + b.R[1] * pc.y + b.R[2] * pc.z,
                           b.R[3] * pc.x + b.R[4] * pc.y + b.R[5] * pc.z,
                           b.R[6] * pc.x + b.R[7] * pc.y + b.R[8] * pc.z);
                f3 vc = b.pos + pr;
                float vol = cs * cs * cs;
                if (sv > -cs) vol *= 0.5f - 0.5f * sv / cs;
                size_t gi = trunc_cell(vc);
                bool subm = s.labels[gi] == FS_FLUID;
                if (!subm) {
                    for (int e = cstart[gi]; e < cstart[gi + 1]; ++e) {
                        int q = order[(size_t)e];
                        float ddx = P0[3 * q] - vc.x, ddy = P0[3 * q + 1] - vc.y, ddz = P0[3 * q + 2] - vc.z;
                        if (sqrtf(ddx * ddx + ddy * ddy + ddz * ddz) < search) { subm = true; break; }
                    }
                }
                if (subm) sub += vol;
            }
    float subf = (float)sub;
    f3 Farch = (-rho * subf) * G;
    b.force = b.force + Farch;
    b.last_force = b.force;
    b.last_torque = b.torque;
    b.last_sub = subf;
    if (s.cfg.step_flags & FS_STEP_BODY_INERTIA)
        b.integrate_tensor(s.dt_cur);
    else
        b.integrate(s.dt_cur);
    if (s.cfg.step_flags & FS_STEP_BODY_WALLS) b.collide_walls(s.W, s.H, s.D, s.cfg.dx);
}

/* ---------------- scene IO ------------------------------------------------------ */
int read_labels_file(Solver& s, const char* path) {
    std::ifstream f(path);
    if (!f.is_open()) return fail(FS_ERR_IO, std::string("Cannot open file: ") + path);
    std::string first;
    std::getline(f, first);
    std::istringstream hs(first);
    int W = 0, H = 0, D = 1;
    hs >> W >> H;
    if (!(hs >> D)) D = 1;
    if (W <= 0 || H <= 0 || D <= 0) return fail(FS_ERR_IO, "bad label header");
    s.W = W; s.H = H; s.D = D;
    s.alloc_grids();
    std::vector<uint8_t> L(s.C(), FS_SOLID);
    for (int k = D - 1; k > -1; --k)
        for (int j = H - 1; j > -1; --j)
            for (int i = W - 1; i > -1; --i) {
                char ch;
                if (!(f >> ch)) goto done;
                size_t idx = (size_t)i + (size_t)j * W + (size_t)k * W * H;
                switch (ch) {
                    case 'S': L[idx] = FS_SOLID; break;
                    case 'F': L[idx] = FS_FLUID; break;
                    case 'A': L[idx] = FS_AIR; break;
                    default: break;
                }
            }
done:
    s.labels = L;
    for (size_t c = 0; c < s.C(); ++c) s.solid[c] = (L[c] == FS_SOLID);
    return FS_OK;
}

/* RigidBody init (FluidSolver3D.cu:157-196, loadSDF utility.cuh:231-338) */
void init_body(Solver& s, const float* sdf, int w, int h, int d, f3 origin, float cs, f3 pos, float mass) {
    Body& b = s.body;
    b = Body();
    b.active = true;
    b.sdf.assign(sdf, sdf + (size_t)w * h * d);
    b.sw = w; b.sh = h; b.sd = d;
    b.fileOrigin = origin;
    b.cs = cs;
    b.size = mk(w * cs, h * cs, d * cs);
    f3 centerInFile = b.fileOrigin + b.size * 0.5f;
    f3 T = pos - centerInFile;
    b.pos = pos;
    b.sdf_origin = b.fileOrigin + T;
    b.local_com = centerInFile;
    b.mass = mass;
    b.vel = mk(0, 0, 0);
    float radius = b.size.x / 2.0f;
    b.inertia = 0.4f * b.mass * radius * radius;
    b.inv_inertia = 1.0f / b.inertia;
    b.force = mk(0, 0, 0);
    b.omega = mk(0, 0, 0);
    b.torque = mk(0, 0, 0);
    b.theta = b.psi = b.phi = 0;
    b.update_rotation_matrix();
    b.compute_inertia_tensor();
    if (!s.labels.empty()) mark_body(s);
}

int write_particles(const Solver& s, const char* path, int fmt) {
    std::ofstream f(path, fmt == FS_FMT_PLY_BINARY ? std::ios::binary : std::ios::out);
    if (!f.is_open()) return fail(FS_ERR_IO, std::string("Error opening file: ") + path);
    int64_t n = s.np();
    if (fmt == FS_FMT_OFF) {
        f << "OFF\n" << n << " 0 0\n";
    } else {
        f << "ply\nformat " << (fmt == FS_FMT_PLY_BINARY ? "binary_little_endian" : "ascii") << " 1.0\nelement vertex "
          << n << "\nproperty float x\nproperty float y\nproperty float z\nend_header\n";
    }
    if (fmt == FS_FMT_PLY_BINARY)
        f.write((const char*)s.pos.data(), (std::streamsize)(sizeof(float) * 3 * n));
    else
        for (int64_t q = 0; q < n; ++q) f << s.pos[3 * q] << " " << s.pos[3 * q + 1] << " " << s.pos[3 * q + 2] << "\n";
    return FS_OK;
}

/* RigidBody::generateSurfacePoints(0.5 dx) + exportToPLY utility.cuh:341-400,
 * written by run() for an active body (FluidSolver3D.cu:3129-3131): the SDF
 * voxels with |phi| < 0.5 dx, mapped to the world frame, as ASCII PLY */
void write_body_ply(const Solver& s, const std::string& path) {
    const Body& b = s.body;
    const float density = 0.5f * s.cfg.dx;
    std::vector<f3> pts;
    for (int k = 0; k < b.sd; ++k)
        for (int j = 0; j < b.sh; ++j)
            for (int i = 0; i < b.sw; ++i) {
                float v = b.sdf[(size_t)i + (size_t)j * b.sw + (size_t)k * b.sw * b.sh];
                if (!(std::fabs(v) < density)) continue;
                f3 pl = b.fileOrigin + mk(i * b.cs, j * b.cs, k * b.cs);
                f3 pc = pl - b.local_com;
                f3 pr = mk(b.R[0] * pc.x + b.R[1] * pc.y + b.R[2] * pc.z,
                           b.R[3] * pc.x + b.R[4] * pc.y + b.R[5] * pc.z,
                           b.R[6] * pc.x + b.R[7] * pc.y + b.R[8] * pc.z);
                pts.push_back(b.pos + pr);
            }
    std::ofstream f(path);
    if (!f.is_open()) return;
    f << "ply\nformat ascii 1.0\nelement vertex " << pts.size()
      << "\nproperty float x\nproperty float y\nproperty float z\nend_header\n";
    for (auto& p : pts) f << p.x << " " << p.y << " " << p.z << "\n";
}

/* CFL-limited dt (no reference counterpart): dt = min(dt_max, cfl*dx/max|v|) */
void choose_dt(Solver& s) {
    s.dt_cur = s.cfg.dt;
    if (!(s.cfg.cfl > 0.0f)) return;
    float umax = 0.0f;
    for (int64_t q = 0; q < s.np(); ++q) {
        float vx = s.vel[3 * q], vy = s.vel[3 * q + 1], vz = s.vel[3 * q + 2];
        float sp = sqrtf(vx * vx + vy * vy + vz * vz);
        if (sp > umax) umax = sp;
    }
    if (umax > 0.0f) s.dt_cur = fminf(s.cfg.dt, (s.cfg.cfl * s.cfg.dx) / umax);
}

int step_once(Solver& s) {
    choose_dt(s);
    auto t = clk::now();
    stage_label(s);
    s.times[FS_STAGE_LABEL] += ms_since(t);
    t = clk::now();
    stage_p2g(s);
    s.times[FS_STAGE_P2G] += ms_since(t);
    t = clk::now();
    stage_rhs(s);
    int rc = stage_pressure(s);
    s.times[FS_STAGE_PRESSURE] += ms_since(t);
    /* a solve that misses the tolerance still completes the step (the
     * reference ignores the solver status); fs_step reports it afterwards */
    if (rc == FS_ERR_NOT_CONVERGED) {
        s.nonconv += 1;
        s.nonconv_res = s.stats.rel_residual;
    } else if (rc != FS_OK && rc != FS_ERR_NO_FLUID) {
        return rc;
    }
    t = clk::now();
    stage_project(s);
    s.times[FS_STAGE_PROJECT] += ms_since(t);
    t = clk::now();
    stage_body(s);
    s.times[FS_STAGE_BODY] += ms_since(t);
    t = clk::now();
    stage_g2p_advect(s);
    s.times[FS_STAGE_G2P] += ms_since(t);
    s.sim_time += s.dt_cur;
    s.nsteps += 1;
    return FS_OK;
}

/* FS_ERR_NOT_CONVERGED when a solve of the steps since `before` missed the tolerance */
int report_nonconv(Solver& s, int64_t before, int64_t nsteps) {
    if (s.nonconv == before) return FS_OK;
    return fail(FS_ERR_NOT_CONVERGED, "pressure solve missed the tolerance in " + std::to_string(s.nonconv - before) +
                                          " of " + std::to_string(nsteps) + " steps (last relative residual " +
                                          std::to_string(s.nonconv_res) + ")");
}

/* ---------------- minimal flat-JSON reader for config.json ------------------------ */
bool json_find(const std::string& txt, const char* key, std::string& val) {
    std::string k = std::string("\"") + key + "\"";
    size_t p = txt.find(k);
    if (p == std::string::npos) return false;
    p = txt.find(':', p + k.size());
    if (p == std::string::npos) return false;
    ++p;
    while (p < txt.size() && isspace((unsigned char)txt[p])) ++p;
    if (p < txt.size() && txt[p] == '"') {
        size_t e = txt.find('"', p + 1);
        val = txt.substr(p + 1, e - p - 1);
        return true;
    }
    size_t e = p;
    while (e < txt.size() && txt[e] != ',' && txt[e] != '}' && !isspace((unsigned char)txt[e])) ++e;
    val = txt.substr(p, e - p);
    return true;
}

}  // namespace

/* ======================= C ABI ==================================================== */
extern "C" {

const char* fs_last_error(void) { return g_err.c_str(); }
const char* fs_backend(void) { return "oracle"; }

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
    std::ifstream f(path);
    if (!f.is_open()) return fail(FS_ERR_IO, std::string("cannot open ") + path);
    std::stringstream ss;
    ss << f.rdbuf();
    std::string t = ss.str(), v;
    try {
    if (json_find(t, "dt", v)) c->dt = std::stof(v);
    if (json_find(t, "dx", v)) c->dx = std::stof(v);
    if (json_find(t, "PIC_WEIGHT", v)) c->pic_weight = std::stof(v);
    if (json_find(t, "iterPerFrame", v)) c->iter_per_frame = std::stoi(v);
    if (json_find(t, "outputFormat", v)) c->output_format = (v == "PLY") ? FS_FMT_PLY : FS_FMT_OFF;
    if (json_find(t, "outputTemplate", v)) std::snprintf(c->output_template, 256, "%s", v.c_str());
    if (json_find(t, "inputLabelData", v)) std::snprintf(c->input_label_data, 256, "%s", v.c_str());
    if (json_find(t, "frameAmount", v)) c->frame_amount = std::stoi(v);
    if (json_find(t, "activeBody", v)) c->active_body = (v == "true" || v == "1");
    if (json_find(t, "body_x", v)) c->body_x = std::stof(v);
    if (json_find(t, "body_y", v)) c->body_y = std::stof(v);
    if (json_find(t, "body_z", v)) c->body_z = std::stof(v);
    if (json_find(t, "body_mass", v)) c->body_mass = std::stof(v);
    /* the reference hard-codes InputData/ball.sdf (FluidSolver3D.cu:160); this key overrides it */
    if (json_find(t, "bodySdf", v)) std::snprintf(c->body_sdf_path, sizeof(c->body_sdf_path), "%s", v.c_str());
    } catch (const std::exception& e) {
        return fail(FS_ERR_IO, std::string("bad value in ") + path + ": " + e.what());
    }
    return FS_OK;
}

int fs_create(const fs_config* cfg, void** out) {
    if (!cfg || !out) return fail(FS_ERR_ARG, "null argument");
    Solver* s = new Solver();
    s->cfg = *cfg;
    s->dt_cur = cfg->dt;
    *out = s;
    return FS_OK;
}

void fs_destroy(void* s) { delete (Solver*)s; }

int fs_load_labels_file(void* h, const char* path) { return read_labels_file(*(Solver*)h, path); }

int fs_set_labels(void* h, const uint8_t* L, int W, int H, int D) {
    Solver& s = *(Solver*)h;
    if (W <= 0 || H <= 0 || D <= 0 || !L) return fail(FS_ERR_ARG, "bad dims");
    s.W = W; s.H = H; s.D = D;
    s.alloc_grids();
    for (size_t c = 0; c < s.C(); ++c) {
        s.labels[c] = L[c];
        s.solid[c] = (L[c] == FS_SOLID);
    }
    return FS_OK;
}

int fs_set_level_sets(void* h, const float* ps, const float* pf, int W, int H, int D) {
    std::vector<uint8_t> L((size_t)W * H * D);
    for (size_t c = 0; c < L.size(); ++c) L[c] = ps[c] < 0 ? FS_SOLID : (pf[c] < 0 ? FS_FLUID : FS_AIR);
    return fs_set_labels(h, L.data(), W, H, D);
}

int fs_get_dims(void* h, int* W, int* H, int* D) {
    Solver& s = *(Solver*)h;
    *W = s.W; *H = s.H; *D = s.D;
    return FS_OK;
}

/* RigidBody::loadSDF utility.cuh:231-338.  The reference throws
 * std::runtime_error (or std::invalid_argument from std::stof) on a bad file;
 * here each of those cases returns FS_ERR_IO, and dimensions are checked
 * before the value buffer is sized. */
int fs_load_body_sdf(void* h, const char* path, float x, float y, float z, float mass) {
    std::ifstream f(path);
    if (!f.is_open()) return fail(FS_ERR_IO, std::string("Cannot open SDF file: ") + path);
    std::string line;
    int w = 0, hh = 0, d = 0;
    f3 o{0, 0, 0};
    float cs = 0.0f;
    if (!std::getline(f, line)) return fail(FS_ERR_IO, "SDF: missing dimensions");
    { std::istringstream d1(line); d1 >> w >> hh >> d; }
    if (!std::getline(f, line)) return fail(FS_ERR_IO, "SDF: missing origin");
    { std::istringstream d2(line); d2 >> o.x >> o.y >> o.z; }
    if (!std::getline(f, line)) return fail(FS_ERR_IO, "SDF: missing cell size");
    try {
        cs = std::stof(line);
    } catch (const std::exception&) {
        return fail(FS_ERR_IO, "SDF: bad cell size");
    }
    if (w <= 0 || hh <= 0 || d <= 0 || (long long)w * hh * d > (1LL << 31))
        return fail(FS_ERR_IO, "SDF: bad dimensions");
    std::vector<float> data((size_t)w * hh * d);
    for (size_t q = 0; q < data.size(); ++q)
        if (!(f >> data[q])) return fail(FS_ERR_IO, "Unexpected end of SDF data");
    return fs_set_body_sdf(h, data.data(), w, hh, d, o.x, o.y, o.z, cs, x, y, z, mass);
}

int fs_set_body_sdf(void* h, const float* sdf, int w, int hh, int d, float ox, float oy, float oz, float cs,
                    float x, float y, float z, float mass) {
    Solver& s = *(Solver*)h;
    if (w <= 0 || hh <= 0 || d <= 0 || cs <= 0) return fail(FS_ERR_ARG, "bad sdf dims");
    init_body(s, sdf, w, hh, d, mk(ox, oy, oz), cs, mk(x, y, z), mass);
    return FS_OK;
}

int fs_get_body_state(void* h, fs_body_state* o) {
    Solver& s = *(Solver*)h;
    Body& b = s.body;
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
}

int fs_set_body_state(void* h, const fs_body_state* in) {
    Solver& s = *(Solver*)h;
    Body& b = s.body;
    b.pos = mk(in->pos[0], in->pos[1], in->pos[2]);
    b.vel = mk(in->vel[0], in->vel[1], in->vel[2]);
    b.omega = mk(in->omega[0], in->omega[1], in->omega[2]);
    b.theta = in->theta; b.psi = in->psi; b.phi = in->phi;
    b.update_rotation_matrix();
    b.quaternion_from_angles();
    return FS_OK;
}

/* seedParticles FluidSolver3D.cu:213-297 */
int fs_seed_particles(void* h, int ppc, uint64_t seed) {
    Solver& s = *(Solver*)h;
    if (ppc <= 0 || ppc > 63) return fail(FS_ERR_ARG, "particles per cell must be in [1,63]");
    const float dx = s.cfg.dx;
    const double lo = (double)(-0.24f), hi = (double)(0.24f);
    s.pos.clear();
    s.vel.clear();
    for (int k = 0; k < s.D; ++k)
        for (int j = 0; j < s.H; ++j)
            for (int i = 0; i < s.W; ++i) {
                size_t c = (size_t)i + (size_t)j * s.W + (size_t)k * s.W * s.H;
                if (s.labels[c] != FS_FLUID) continue;
                f3 cc = mk(((float)i + 0.5f) * dx, ((float)j + 0.5f) * dx, ((float)k + 0.5f) * dx);
                const float off = 0.25f * dx;
                for (int q = 0; q < ppc; ++q) {
                    int oct = (int)(seed_draw(seed, c, q, 0) >> 61);
                    f3 sc = mk(cc.x + ((oct & 1) ? off : -off), cc.y + ((oct & 2) ? off : -off),
                               cc.z + ((oct & 4) ? off : -off));
                    float jit[3];
                    for (int a = 0; a < 3; ++a) {
                        double uu = (double)(seed_draw(seed, c, q, 1 + a) >> 11) * (1.0 / 9007199254740992.0);
                        double val = lo + (hi - lo) * uu;
                        jit[a] = (float)(val * (double)dx);
                    }
                    f3 p = mk(sc.x + jit[0], sc.y + jit[1], sc.z + jit[2]);
                    p.x = std::clamp(p.x, i * dx, (i + 1) * dx);
                    p.y = std::clamp(p.y, j * dx, (j + 1) * dx);
                    p.z = std::clamp(p.z, k * dx, (k + 1) * dx);
                    s.pos.push_back(p.x); s.pos.push_back(p.y); s.pos.push_back(p.z);
                    s.vel.push_back(0); s.vel.push_back(0); s.vel.push_back(0);
                }
            }
    return FS_OK;
}

int fs_set_particles(void* h, const float* pos, const float* vel, int64_t n) {
    Solver& s = *(Solver*)h;
    if (n < 0) return fail(FS_ERR_ARG, "negative count");
    s.pos.assign(pos, pos + 3 * n);
    if (vel) s.vel.assign(vel, vel + 3 * n); else s.vel.assign((size_t)3 * n, 0.f);
    return FS_OK;
}

int fs_get_particles(void* h, float* pos, float* vel, int64_t cap, int64_t* n) {
    Solver& s = *(Solver*)h;
    int64_t m = s.np();
    if (n) *n = m;
    if (cap < m) return fail(FS_ERR_ARG, "buffer too small");
    if (pos) std::memcpy(pos, s.pos.data(), sizeof(float) * 3 * m);
    if (vel) std::memcpy(vel, s.vel.data(), sizeof(float) * 3 * m);
    return FS_OK;
}

int64_t fs_num_particles(void* h) { return ((Solver*)h)->np(); }

int fs_step(void* h, int nsteps) {
    Solver& s = *(Solver*)h;
    if (s.C() == 0) return fail(FS_ERR_STATE, "no scene loaded");
    const int64_t before = s.nonconv;
    for (int i = 0; i < nsteps; ++i) {
        int rc = step_once(s);
        if (rc != FS_OK) return rc;
    }
    return report_nonconv(s, before, nsteps);
}

int fs_run(void* h, int frames) {
    Solver& s = *(Solver*)h;
    const int ipf = std::max(1, s.cfg.iter_per_frame);
    std::ofstream csv("OutputData/csvRes.csv");
    if (csv.is_open()) csv << "frame,centerX1,centerX2,centerX3,angleX1,angleX2,angleX3\n";
    const int64_t before = s.nonconv;
    for (int i = 0; i <= frames * ipf; ++i) {
        int rc = step_once(s);
        if (rc != FS_OK) return rc;
        if (i % ipf == 0) {
            std::string fn = std::string(s.cfg.output_template) + std::to_string(i / ipf) +
                             (s.cfg.output_format == FS_FMT_OFF ? ".off" : ".ply");
            write_particles(s, fn.c_str(), s.cfg.output_format);
            if (s.body.active) write_body_ply(s, "OutputData/body_" + std::to_string(i / ipf) + ".ply");
            if (csv.is_open())
                csv << i / ipf << "," << s.body.pos.x << "," << s.body.pos.y << "," << s.body.pos.z << ","
                    << s.body.theta << "," << s.body.psi << "," << s.body.phi << "\n";
        }
    }
    return report_nonconv(s, before, (int64_t)frames * ipf + 1);
}

int fs_stage_label(void* h) { stage_label(*(Solver*)h); return FS_OK; }
int fs_stage_p2g(void* h) { stage_p2g(*(Solver*)h); return FS_OK; }
int fs_stage_rhs(void* h) { stage_rhs(*(Solver*)h); return FS_OK; }
int fs_stage_pressure(void* h) { return stage_pressure(*(Solver*)h); }
int fs_stage_project(void* h) { stage_project(*(Solver*)h); return FS_OK; }
int fs_stage_body(void* h) { stage_body(*(Solver*)h); return FS_OK; }
int fs_stage_g2p_advect(void* h) { stage_g2p_advect(*(Solver*)h); return FS_OK; }

int fs_get_grid(void* h, int which, void* out, int64_t count) {
    Solver& s = *(Solver*)h;
    auto cp = [&](const void* src, size_t n, size_t esz) -> int {
        if ((size_t)count < n) return fail(FS_ERR_ARG, "buffer too small");
        std::memcpy(out, src, n * esz);
        return FS_OK;
    };
    switch (which) {
        case FS_GRID_U: return cp(s.u.data(), s.u.size(), 4);
        case FS_GRID_V: return cp(s.v.data(), s.v.size(), 4);
        case FS_GRID_W: return cp(s.w.data(), s.w.size(), 4);
        case FS_GRID_USAVED: return cp(s.us.data(), s.us.size(), 4);
        case FS_GRID_VSAVED: return cp(s.vs.data(), s.vs.size(), 4);
        case FS_GRID_WSAVED: return cp(s.ws.data(), s.ws.size(), 4);
        case FS_GRID_P: return cp(s.p.data(), s.p.size(), 4);
        case FS_GRID_LABELS: return cp(s.labels.data(), s.labels.size(), 1);
        case FS_GRID_RHS: return cp(s.rhs.data(), s.rhs.size(), 4);
        case FS_GRID_P64: return cp(s.p64.data(), s.p64.size(), 8);
        case FS_GRID_CELLKEY: {
            std::vector<int> k((size_t)s.np());
            for (int64_t q = 0; q < s.np(); ++q)
                k[(size_t)q] = cell_key(s.pos[3 * q], s.pos[3 * q + 1], s.pos[3 * q + 2], s.cfg.dx, s.W, s.H, s.D);
            return cp(k.data(), k.size(), 4);
        }
        case FS_GRID_SOLID: return cp(s.solid.data(), s.solid.size(), 1);
        case FS_GRID_OPEN_U:
        case FS_GRID_OPEN_V:
        case FS_GRID_OPEN_W: {
            std::vector<float> th = open_fractions(s, which - FS_GRID_OPEN_U);
            return cp(th.data(), th.size(), 4);
        }
        default: return fail(FS_ERR_ARG, "unknown grid");
    }
}

int fs_set_grid(void* h, int which, const void* in, int64_t count) {
    Solver& s = *(Solver*)h;
    auto cp = [&](void* dst, size_t n, size_t esz) -> int {
        if ((size_t)count != n) return fail(FS_ERR_ARG, "size mismatch");
        std::memcpy(dst, in, n * esz);
        return FS_OK;
    };
    switch (which) {
        case FS_GRID_U: return cp(s.u.data(), s.u.size(), 4);
        case FS_GRID_V: return cp(s.v.data(), s.v.size(), 4);
        case FS_GRID_W: return cp(s.w.data(), s.w.size(), 4);
        case FS_GRID_USAVED: return cp(s.us.data(), s.us.size(), 4);
        case FS_GRID_VSAVED: return cp(s.vs.data(), s.vs.size(), 4);
        case FS_GRID_WSAVED: return cp(s.ws.data(), s.ws.size(), 4);
        case FS_GRID_P: return cp(s.p.data(), s.p.size(), 4);
        case FS_GRID_LABELS: return cp(s.labels.data(), s.labels.size(), 1);
        case FS_GRID_RHS: return cp(s.rhs.data(), s.rhs.size(), 4);
        case FS_GRID_P64: return cp(s.p64.data(), s.p64.size(), 8);
        default: return fail(FS_ERR_ARG, "grid not settable");
    }
}

int fs_get_stage_times(void* h, double* ms, int n) {
    Solver& s = *(Solver*)h;
    for (int i = 0; i < n && i < FS_NUM_STAGES; ++i) ms[i] = s.times[i];
    return FS_OK;
}

int fs_get_clock(void* h, double* t, float* dt, int64_t* n) {
    Solver& s = *(Solver*)h;
    if (t) *t = s.sim_time;
    if (dt) *dt = s.dt_cur;
    if (n) *n = s.nsteps;
    return FS_OK;
}

int fs_get_solve_stats(void* h, fs_solve_stats* o) {
    *o = ((Solver*)h)->stats;
    return FS_OK;
}

int fs_write_particles(void* h, const char* path, int fmt) { return write_particles(*(Solver*)h, path, fmt); }

int fs_solve_pressure_system(void* h, const uint8_t* labels, const float* b, double* p, int W, int H, int D,
                             double tol, fs_solve_stats* st) {
    (void)h;
    std::vector<uint8_t> L(labels, labels + (size_t)W * H * D);
    return solve_system(L, b, p, W, H, D, tol, 100000, st);
}

/* Oracle-only helper: solve an SPD system given in upper-triangular CSR (the
 * cuDSS input of FluidSolver3D.cu:1932-1946) by the band LDL^T used above. */
int oracle_spd_solve_upper_csr(int n, const int* off, const int* col, const double* val, const double* b,
                               double* x) {
    Csr A;
    A.n = n;
    std::vector<std::vector<std::pair<int, double>>> rows((size_t)n);
    for (int r = 0; r < n; ++r)
        for (int e = off[r]; e < off[r + 1]; ++e) {
            rows[(size_t)r].push_back({col[e], val[e]});
            if (col[e] != r) rows[(size_t)col[e]].push_back({r, val[e]});
        }
    A.off.assign((size_t)n + 1, 0);
    for (int r = 0; r < n; ++r) {
        std::sort(rows[(size_t)r].begin(), rows[(size_t)r].end());
        for (auto& e : rows[(size_t)r]) { A.col.push_back(e.first); A.val.push_back(e.second); }
        A.off[r + 1] = (int)A.col.size();
    }
    std::vector<double> bb(b, b + n), xx;
    if (!band_ldlt_solve(A, bb, xx)) return fail(FS_ERR_NOT_CONVERGED, "matrix not SPD");
    std::memcpy(x, xx.data(), sizeof(double) * n);
    return FS_OK;
}

/* Oracle-only helper: the reference CSR (upper triangle, FluidSolver3D.cu:1598-1777)
 * of a label grid.  Returns nnz; arrays may be null to query sizes. */
int64_t oracle_assemble_upper_csr(const uint8_t* labels, int W, int H, int D, int* off, int* col, float* val) {
    std::vector<uint8_t> L(labels, labels + (size_t)W * H * D);
    std::vector<int> num;
    Csr A = assemble(L, W, H, D, num);
    int64_t nnz = 0;
    if (off) off[0] = 0;
    for (int r = 0; r < A.n; ++r) {
        /* reference row order: diagonal first, then neighbours in offset order
         * (+x,-x,+y,-y,+z,-z) with col > row */
        int c = A.cell[r];
        int i = c % W, j = (c / W) % H, k = c / (W * H);
        double diag = 0;
        for (int e = A.off[r]; e < A.off[r + 1]; ++e) if (A.col[e] == r) diag = A.val[e];
        if (col) col[nnz] = r;
        if (val) val[nnz] = (float)diag;
        ++nnz;
        const int of[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
        for (auto& o : of) {
            int ni = i + o[0], nj = j + o[1], nk = k + o[2];
            if (ni < 0 || ni >= W || nj < 0 || nj >= H || nk < 0 || nk >= D) continue;
            size_t nidx = (size_t)ni + (size_t)nj * W + (size_t)nk * W * H;
            if (L[nidx] == FS_FLUID && num[nidx] > r) {
                if (col) col[nnz] = num[nidx];
                if (val) val[nnz] = -1.0f;
                ++nnz;
            }
        }
        if (off) off[r + 1] = (int)nnz;
    }
    return nnz;
}

}  // extern "C"

/* ---- checkpoint (include/fluidsim_checkpoint.h; no reference counterpart):
 * the file format of the HIP library, so that a run saved by either library
 * continues in the other.  Particles are stored in the oracle's order (by
 * original id); a loaded file's particles are put back into id order. */
namespace {
const char kCkptMagic[8] = {'F', 'S', 'C', 'K', 'P', 'T', '0', '2'};

template <class IO>
void body_ckpt_fields(Body& b, IO io) {
    auto f = [&](auto& v) { io((void*)&v, sizeof(v)); };
    f(b.sw); f(b.sh); f(b.sd);
    f(b.fileOrigin); f(b.sdf_origin); f(b.local_com); f(b.size); f(b.cs);
    f(b.pos); f(b.vel); f(b.force); f(b.omega); f(b.torque);
    f(b.mass); f(b.inertia); f(b.inv_inertia); f(b.theta); f(b.psi); f(b.phi);
    f(b.R); f(b.last_force); f(b.last_torque); f(b.last_sub);
    f(b.Ib); f(b.Ibinv); f(b.q);
}
}  // namespace

extern "C" {


int fs_save_checkpoint(void* h, const char* path) {
    Solver& s = *(Solver*)h;
    if (!path) return fail(FS_ERR_ARG, "null path");
    if (s.C() == 0) return fail(FS_ERR_STATE, "nothing to save: no scene");
    std::ofstream f(path, std::ios::binary);
    if (!f.is_open()) return fail(FS_ERR_IO, std::string("cannot write ") + path);
    auto raw = [&](const void* p, size_t n) { if (n) f.write((const char*)p, (std::streamsize)n); };
    raw(kCkptMagic, 8);
    const int32_t cfg_size = (int32_t)sizeof(fs_config);
    raw(&cfg_size, 4);
    raw(&s.cfg, sizeof(fs_config));
    raw(&s.W, 4); raw(&s.H, 4); raw(&s.D, 4);
    raw(&s.dt_cur, 4); raw(&s.sim_time, 8); raw(&s.nsteps, 8);
    const int64_t np = s.np();
    raw(&np, 8);
    raw(s.solid.data(), s.C());
    raw(s.labels.data(), s.C());
    std::vector<float> a((size_t)np);
    for (int c = 0; c < 6; ++c) {
        const std::vector<float>& src = c < 3 ? s.pos : s.vel;
        for (int64_t q = 0; q < np; ++q) a[(size_t)q] = src[3 * (size_t)q + (c % 3)];
        raw(a.data(), sizeof(float) * (size_t)np);
    }
    std::vector<int> ids((size_t)np);
    for (int64_t q = 0; q < np; ++q) ids[(size_t)q] = (int)q;
    raw(ids.data(), sizeof(int) * (size_t)np);
    const int32_t active = s.body.active ? 1 : 0;
    raw(&active, 4);
    if (active) {
        body_ckpt_fields(s.body, [&](void* p, size_t n) { raw(p, n); });
        raw(s.body.sdf.data(), sizeof(float) * s.body.sdf.size());
    }
    const int32_t has_p = 1;   /* the last solution (the HIP library's next initial guess) */
    raw(&has_p, 4);
    raw(s.p64.data(), sizeof(double) * s.C());
    const int32_t nparam = 0;
    raw(&nparam, 4);
    f.close();
    if (f.fail()) return fail(FS_ERR_IO, std::string("write failed: ") + path);
    return FS_OK;
}

int fs_load_checkpoint(void* h, const char* path) {
    Solver& s = *(Solver*)h;
    if (!path) return fail(FS_ERR_ARG, "null path");
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) return fail(FS_ERR_IO, std::string("cannot open ") + path);
    f.seekg(0, std::ios::end);
    const long long size = (long long)f.tellg();
    f.seekg(0);
    bool ok = true;
    auto raw = [&](void* p, size_t n) {
        if (n) f.read((char*)p, (std::streamsize)n);
        ok = ok && (bool)f;
    };
    auto left = [&]() { return size - (long long)f.tellg(); };
    char magic[8];
    int32_t cfg_size = 0;
    raw(magic, 8);
    raw(&cfg_size, 4);
    if (!ok || std::memcmp(magic, kCkptMagic, 8) != 0 || cfg_size != (int32_t)sizeof(fs_config))
        return fail(FS_ERR_IO, std::string("not a checkpoint of this version: ") + path);
    fs_config cfg;
    raw(&cfg, sizeof(cfg));
    int W = 0, H = 0, D = 0;
    raw(&W, 4); raw(&H, 4); raw(&D, 4);
    float dt_cur = 0;
    double sim_time = 0;
    int64_t nsteps = 0, np = 0;
    raw(&dt_cur, 4); raw(&sim_time, 8); raw(&nsteps, 8); raw(&np, 8);
    const long long cells = (long long)W * H * D;
    if (!ok || W <= 0 || H <= 0 || D <= 0 || np < 0 || np > 0x7fffffffLL || cells > (1LL << 40) ||
        2 * cells + 28 * np + 8 > left())
        return fail(FS_ERR_IO, std::string("corrupt checkpoint header: ") + path);
    std::vector<uint8_t> solid((size_t)cells), labels((size_t)cells);
    raw(solid.data(), solid.size());
    raw(labels.data(), labels.size());
    std::vector<float> arr[6];
    for (auto& a : arr) {
        a.resize((size_t)np);
        raw(a.data(), sizeof(float) * (size_t)np);
    }
    std::vector<int> ids((size_t)np);
    raw(ids.data(), sizeof(int) * (size_t)np);
    int32_t active = 0;
    raw(&active, 4);
    Body body;
    if (ok && active) {
        body.active = true;
        body_ckpt_fields(body, [&](void* p, size_t n) { raw(p, n); });
        const long long nv = (long long)body.sw * body.sh * body.sd;
        if (!ok || body.sw <= 0 || body.sh <= 0 || body.sd <= 0 || 4 * nv + 4 > left())
            return fail(FS_ERR_IO, "corrupt checkpoint body");
        body.sdf.resize((size_t)nv);
        raw(body.sdf.data(), sizeof(float) * (size_t)nv);
    }
    int32_t has_p = 0;
    raw(&has_p, 4);
    std::vector<double> pv;
    if (ok && has_p) {
        if (8 * cells + 4 > left()) return fail(FS_ERR_IO, "corrupt checkpoint pressure");
        pv.resize((size_t)cells);
        raw(pv.data(), sizeof(double) * (size_t)cells);
    }
    int32_t nparam = 0;
    raw(&nparam, 4);
    if (!ok) return fail(FS_ERR_IO, std::string("truncated checkpoint: ") + path);
    if (nparam > 0) {   /* learned parameters: skipped (the oracle has no network in its step) */
        if (4LL * nparam > left()) return fail(FS_ERR_IO, "corrupt checkpoint parameters");
        f.seekg(4LL * nparam, std::ios::cur);
    }
    /* particles back into original-id order (ids are a permutation of 0..np-1) */
    std::vector<int64_t> where((size_t)np, -1);
    for (int64_t q = 0; q < np; ++q) {
        const int id = ids[(size_t)q];
        if (id < 0 || id >= np || where[(size_t)id] >= 0) return fail(FS_ERR_IO, "corrupt checkpoint particle ids");
        where[(size_t)id] = q;
    }
    s.cfg = cfg;
    s.W = W; s.H = H; s.D = D;
    s.alloc_grids();
    s.solid = solid;
    s.labels = labels;
    s.pos.assign(3 * (size_t)np, 0.f);
    s.vel.assign(3 * (size_t)np, 0.f);
    for (int64_t id = 0; id < np; ++id) {
        const size_t q = (size_t)where[(size_t)id];
        for (int c = 0; c < 3; ++c) {
            s.pos[3 * (size_t)id + c] = arr[c][q];
            s.vel[3 * (size_t)id + c] = arr[3 + c][q];
        }
    }
    s.key.clear();
    if (!pv.empty()) s.p64 = pv;
    s.body = std::move(body);
    s.dt_cur = dt_cur;
    s.sim_time = sim_time;
    s.nsteps = nsteps;
    return FS_OK;
}

/* Sets the body's rotation matrix without touching its angles: the matrix
 * lags the Euler angles by one integration (RigidBody::integrate
 * utility.cuh:408-467 forms R before it advances the angles), so a state read
 * back mid-run is restored exactly only with both.  Test infrastructure. */
int oracle_set_body_rotation(void* h, const float* R) {
    Solver& s = *(Solver*)h;
    if (!R) return fail(FS_ERR_ARG, "null rotation");
    for (int i = 0; i < 9; ++i) s.body.R[i] = R[i];
    return FS_OK;
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
