// fs_pressure.hpp -- matrix-free multigrid-preconditioned CG for A p = b.
//
// A is the reference's pressure matrix (constructA FluidSolver3D.cu:1718-1777):
// on FLUID cells, diag = #in-domain FLUID/AIR neighbours, -1 per FLUID
// neighbour, SOLID/BODY neighbours Neumann, AIR Dirichlet (p = 0).  It replaces
// the cuDSS FP32 direct solve (:1930-2008).
//
// Precision: stencil, smoother and transfers in FP32; dot products in FP64;
// the solution is accumulated in FP64 and the residual b - A x is recomputed in
// FP64 by an outer iterative-refinement loop until ||r|| <= tol ||b||.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fs_device.hpp"

#include <vector>

namespace fsk {

struct SolveScalars {
    double rz, alpha, beta, rr, target, rr_true, bnorm2, bsum, tol2b, inner_rtol2, rr_start;
    double best_rr;   // smallest ||r||^2 of the refinement pass (stall detection)
    int done, iters, first, maxit, n_unknown, dirichlet, stall, pad;
};

struct MGLevel {
    int W, H, D;
    long long n;
    uint16_t* code;   // bits 0-5 neighbour unknown (+x,-x,+y,-y,+z,-z), 8-10 diag, 14 weighted, 15 self unknown
    uint8_t* type;    // 0 none, 1 unknown, 2 Dirichlet
    float* b;         // right-hand side (level 0: aliases the CG residual)
    float* x;         // current iterate
    float* xt;        // Jacobi ping-pong partner
    float* r;         // residual scratch
    // level 0 with the body's face fractions: the open fractions of the +x,
    // +y and +z faces of the FLUID cells and, for the cells whose code has
    // bit 14 (a face to a FLUID/AIR neighbour partly open), the diagonal (sum
    // of the open fractions of the faces to FLUID/AIR neighbours); null otherwise
    const float *wd, *wx, *wy, *wz;
};

// A preconditioner that maps r -> z on the fine grid.  `apply` must leave z = 0
// on non-unknown cells and add the partial sums of r.z into `partial`.
struct Preconditioner {
    virtual ~Preconditioner() {}
    virtual void apply(const float* r, float* z, double* partial, SolveScalars* sc, hipStream_t st) = 0;
    // identifies the launch sequence of apply() for graph caching
    virtual const void* graph_key() const { return this; }
    // a preconditioner that is not a fixed SPD operator (flexible CG: beta makes
    // the new direction A-conjugate to the previous one; stall detection on)
    virtual bool flexible() const { return false; }
    // called once the solve box is configured, before apply() is recorded
    // (allocations belong here, never inside apply); 0 on success
    virtual int prepare_box() { return 0; }
    // changes whenever buffers that apply() records are reallocated
    virtual long long graph_epoch() const { return 0; }
};

// pre and post smoothing sweeps per level (and the degree of their polynomial)
constexpr int SMOOTH_STEPS = 4;

class MGPCG {
public:
    MGPCG() {}
    ~MGPCG();
    // allocate the hierarchy for a W x H x D grid
    int init(int W, int H, int D);
    // the same hierarchy carved from a caller's device workspace of `bytes`
    // bytes, with caller-owned host mirrors; nothing is allocated (the level
    // list only grows its host vector up to the capacity reserved before).
    // *used: the bytes needed.  -2 when they exceed `bytes`.
    int init_in(int W, int H, int D, void* base, size_t bytes, SolveScalars* sc_host_buf, int* host_counters_buf,
                size_t* used);
    // Open fractions of the faces between FLUID/AIR cells from the body's SDF
    // (face.active = 0: every such face open, the reference's matrix).  When
    // active, A's coefficient on a face is its open fraction: the CG operator,
    // the FP64 residuals and the fine level of the V-cycle use the fractions
    // (on the cells that have a partly open face; the others keep unit
    // coefficients); the coarse levels are re-discretised with unit
    // coefficients on the faces that are open at all.
    fsd::FaceFrac face{};
    // Solve A x = b (b: full-grid float, 0 off fluid).  x64 (full grid, double)
    // and p (float, 0 off fluid) are written.  Returns 0 on convergence.
    // x0 (optional, full grid, may be x64 itself): initial guess on the unknowns
    // (the previous step's pressure); null starts from 0.
    int solve(const uint8_t* labels, const float* b, double* x64, float* p, double tol, int maxit,
              Preconditioner* pre, hipStream_t st, struct fs_solve_stats* stats, const double* x0 = nullptr);
    // Fluid bounding box of `labels` (+1 cell), level set-up for that box and
    // the gathered right-hand side bb.  Returns -3 when there is no fluid.
    int prepare(const uint8_t* labels, const float* b, hipStream_t st);
    // level dimensions for a solve box of Wb x Hb x Db cells
    void configure(int Wb, int Hb, int Db);
    // level codes of the configured box and its right-hand side (from b)
    void setup(const uint8_t* labels, const float* b, hipStream_t st);
    MGLevel fine() const;
    // One V-cycle z = M r on level 0, accumulating r.z partial sums.
    void vcycle(const float* r, float* z, double* partial, SolveScalars* sc, hipStream_t st);
    // y = A x on the fine grid (FP32), for tests and the learned path.
    void apply_A(const float* x, float* y, hipStream_t st);
    // Pieces of the hybrid preconditioner z = V(r) + c N(r - A V(r)) over the
    // solve box (N the learned network, c the exact line search
    // (z2.r1) / (z2.A z2) along its output z2):
    // r1 = r - A z on the unknowns (0 elsewhere) and partial sums of r1.r1
    void resid_sq(const float* r, const float* z, float* r1, double* p_rr, SolveScalars* sc, hipStream_t st);
    // scale = sqrt(sum p_rr / n_unknown), 1 when that is not a positive finite number
    void partial_rms(const double* p_rr, float* scale, SolveScalars* sc, hipStream_t st);
    // partial sums of z2.r1 and z2.(A z2)
    void line_dots(const float* z2, const float* r1, double* p_num, double* p_den, SolveScalars* sc, hipStream_t st);
    // z += c z2 with c = sum p_num / sum p_den (0 unless the denominator is positive)
    void add_scaled(float* z, const float* z2, const double* p_num, const double* p_den, SolveScalars* sc,
                    hipStream_t st);
    // One V-cycle from zero for the system (labels, b) of the full grid (the
    // reference's matrix), z on the full grid (0 off the unknowns).  Returns -3
    // without fluid, -5 on a device error.
    int vcycle_system(const uint8_t* labels, const float* b, float* zfull, hipStream_t st);
    int red_blocks() const { return nred; }

    int W = 0, H = 0, D = 0;
    long long N = 0;               // full grid cells
    std::vector<MGLevel> lv;       // allocated for the full grid; the first nlev are in use
    int nlev = 0;
    long long n0 = 0;              // cells of the solve box (level 0)
    int box[3] = {0, 0, 0};        // solve-box origin in the full grid
    int nu_pre = SMOOTH_STEPS, nu_post = SMOOTH_STEPS, coarse_sweeps = 8;
    long long coarse_max = 512;   // coarsen until a level has at most this many cells (LDS-resident)
    float omega = 6.0f / 7.0f;          // coarsest-level sweeps (damped Jacobi)
    // Step of pre/post sweep k: x += w_k D^-1 (b - A x).  The SMOOTH_STEPS
    // steps are the reciprocal roots of the Chebyshev polynomial of that
    // degree on [b/40, b], b = 2 (Gershgorin bound of the spectrum of D^-1 A
    // for this operator), theta_k = (a + b)/2 + (b - a)/2 cos((2k - 1) pi / 8).
    // The polynomial is in D^-1 A, so pre and post smoothing commute with
    // their A-adjoints and the V-cycle stays symmetric (a valid CG
    // preconditioner).  Measured at 256^3 (CG iterations / solve time):
    // damped Jacobi w = 6/7 V(3,3) 16 / 6.7 ms; Chebyshev V(3,3) on [b/20, b]
    // 12 / 5.4 ms; V(4,4) on [b/40, b] 10 / 5.1 ms (at 512^3 19 and 17
    // iterations, within 3% in time); V(2,2), V(5,5) and other intervals
    // were slower.  Equal steps give damped Jacobi.
    float sweep_omega[SMOOTH_STEPS] = {0.5192694f, 0.7152481f, 1.5340161f, 8.0503983f};
    // A fine level of more than 2^16 cells smooths with a degree-2 Chebyshev
    // polynomial on [b/10, b] (V(2,2)): its sweeps are the expensive HBM
    // passes, the coarse levels' sweeps are cheap (on a smaller, launch-bound
    // fine level fewer sweeps only add iterations: the 128^2 slab took 19
    // instead of 16 per solve and 1.27 instead of 1.17 ms).  Measured over
    // steps 61-90 of the 256^3 sphere drop (face fractions): V(4,4) everywhere
    // 12.0 iterations, 546 us each, 6.54 ms per solve; fine V(2,2) on
    // [b/10, b] 13.2 iterations, 432 us, 5.71 ms; on [b/20, b] 13.4, 5.75 ms;
    // [b/7, b] 13.7, 5.90 ms; [b/5, b] 14.0, 5.96 ms; fine V(3,3) on [b/40, b]
    // 12.6, 6.16 ms; fine V(1,1) 17.2 iterations, 7.11 ms.
    static constexpr int NU_FINE = 2;
    float fine_omega[NU_FINE] = {0.57590544f, 2.157014f};
    int batch = 8;              // inner iterations launched between host checks (even)
    double inner_rtol = 0.0;    // inner relative reduction per refinement pass (0: straight to the
                                // tolerance; FP64 A p sums and solution keep r with the true residual)
    int max_outer = 12;
    // fine-grid CG vectors
    float *r = nullptr, *z = nullptr, *pa = nullptr, *pb = nullptr, *q = nullptr;
    float* bb = nullptr;           // b on the solve box
    double* xb = nullptr;          // FP64 solution on the solve box
    static constexpr int COUNTERS_BBOX = 16;   // per-block bounding boxes from here on
    static constexpr int HOST_COUNTERS = COUNTERS_BBOX + 6 * 256;   // ints in the host mirror
    int* counters = nullptr;       // device: unknown count [8], Dirichlet flag [9], per-block bbox [16..]
    int* host_counters = nullptr;  // pinned mirror
    double* partial = nullptr;     // partial-sum slots of the reducing kernels
    double* slot(int s);
    SolveScalars* sc = nullptr;       // device
    SolveScalars* sc_host = nullptr;  // pinned
    int nred = 0;   // number of partial-sum slots used by reducing kernels
    int last_iters = 0, last_outer = 0;
    int zero_count = 0;           // host source of the iteration count's reset
    // The fluid bounding box of `labels` queued ahead of prepare() (the
    // host then waits for its copy only, while the work queued since runs);
    // the caller clears bbox_labels (forget_bbox) when the labels change
    // before the solve.
    int queue_bbox(const uint8_t* labels, hipStream_t st);
    void forget_bbox() { bbox_labels = nullptr; }
    double rate_per_iter = 0.0;   // log(||r||^2 reduction) per iteration of the last converged solve (< 0)
    bool use_graph = true;      // replay inner batches as HIP graphs
    bool use_tail = true;
    bool use_tb = true;         // temporally blocked smoothing kernels on small coarse levels
    bool singular_mode = false; // the current system has no Dirichlet cell (pure Neumann)
    bool use_vec = true;        // 4-cells-per-thread kernels on levels whose width is a multiple of 4
    bool vec0 = false;          // the configured fine level uses them       // levels of at most 32K cells in one workgroup launch
    // partial sums of r.z over the solve box (for preconditioners that do not fuse it)
    void dot_rz(const float* r, const float* z, double* partial, SolveScalars* sc, hipStream_t st);
    // r.z and z.q (q = A p of the previous iteration) for flexible preconditioners
    void dot_rz_zq(const float* r, const float* z, double* partial, SolveScalars* sc, hipStream_t st);
    int stall_window = 5;       // flexible path: iterations without a new smallest residual before falling back
    long long n_unk = 0;        // unknowns of the current solve

private:
    struct GraphEntry {
        long long key[10];
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
    void record_batch(Preconditioner* pre, int nb, hipStream_t st);
    hipGraphExec_t batch_graph(Preconditioner* pre, int nb, hipStream_t st);
    void clear_graphs();
    void free_all();
    int init_impl(int W, int H, int D, void* base, size_t bytes, SolveScalars* sc_host_buf, int* host_counters_buf,
                  size_t* used);
    bool owns_memory = true;   // buffers from hipMalloc (else carved from a caller's workspace)
    float* fw = nullptr;       // face weights of the solve box, 4 x N floats (allocated on first use)
    float* vcycle_level(int l, hipStream_t st, double* partial_dot, SolveScalars* sc);
    int launch_bbox(const uint8_t* labels, hipStream_t st);
    hipEvent_t bbox_ev = nullptr;
    const uint8_t* bbox_labels = nullptr;   // labels of the queued bounding box (null: none)
};

Preconditioner* make_jacobi_preconditioner(MGPCG* s);

class MGPreconditioner : public Preconditioner {
public:
    explicit MGPreconditioner(MGPCG* s) : s_(s) {}
    void apply(const float* r, float* z, double* partial, SolveScalars* sc, hipStream_t st) override {
        s_->vcycle(r, z, partial, sc, st);
    }
    const void* graph_key() const override { return s_; }

private:
    MGPCG* s_;
};

}  // namespace fsk
