// fs_learned.hpp -- the learned preconditioner network (see include/fluidsim_learned.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fs_pressure.hpp"

namespace fsk {

constexpr int LN_C = 16;           // channels
constexpr int LN_TAPS = 27;
// parameter offsets in the FP32 vector
constexpr int LN_AW = 0, LN_AB = LN_AW + 27 * 3 * 16;
constexpr int LN_BW = LN_AB + 16, LN_BB = LN_BW + 27 * 16 * 16;
constexpr int LN_CW = LN_BB + 16, LN_CB = LN_CW + 27 * 16 * 16;
constexpr int LN_DW = LN_CB + 16, LN_DB = LN_DW + 27 * 16 * 16;
constexpr int LN_EW = LN_DB + 16, LN_EB = LN_EW + 27 * 16;
constexpr int LN_NPARAM = LN_EB + 1;

struct LGrid {
    int W, H, D;
    long long n;
};

class LNet {
public:
    ~LNet();
    int init_params(uint64_t seed, hipStream_t st);
    int set_params(const float* host, hipStream_t st);
    int get_params(float* host, hipStream_t st);
    // activations for a grid of W x H x D (reallocated only when it grows)
    int ensure(int W, int H, int D);
    // z = M(r) on the grid set by ensure(): type (1 unknown, 2 Dirichlet) and r.
    // scale = rms(r) over the unknowns, sqrt(sc->rr / sc->n_unknown) when sc is
    // given (every kernel then returns at once when sc->done is set), else
    // scale_host.  zn (optional): the unscaled output (training).
    // scale_dev (optional): the scale as a device float, in place of the two above.
    void forward(const uint8_t* type, const float* r, float* z, const SolveScalars* sc, float scale_host, float* zn,
                 hipStream_t st, const float* scale_dev = nullptr);
    // gradient of the A-norm loss 0.5 (zn - target).A(zn - target) / n_fluid into
    // d_grad (zeroed first); the forward pass must have run with zn.  The loss
    // -> d_loss[0].
    void backward(const uint8_t* type, const float* target, const float* zn, int n_fluid, hipStream_t st);
    void adam(float lr, hipStream_t st);
    void refresh_weights(hipStream_t st);   // FP32 params -> BF16 kernel layouts

    long long epoch = 0;   // incremented when the activations are reallocated
    float* d_params = nullptr;
    float* d_grad = nullptr;
    double* d_loss = nullptr;
    int adam_t = 0;

private:
    float *d_m = nullptr, *d_v = nullptr;
    uint16_t *wB = nullptr, *wC = nullptr, *wD = nullptr, *wE = nullptr;      // [tap][cout][cin] BF16
    uint16_t *wBt = nullptr, *wCt = nullptr, *wDt = nullptr;                 // flipped, transposed
    LGrid g0{}, g1{}, g2{};
    long long cap0 = 0, cap1 = 0, cap2 = 0;
    float* rn = nullptr;          // normalised input (level 0)
    float* dz = nullptr;          // loss gradient (level 0)
    uint16_t *a0 = nullptr, *s0 = nullptr, *ga0 = nullptr;                    // level 0, 16 ch
    uint16_t *a1 = nullptr, *b1 = nullptr, *s1 = nullptr, *d1 = nullptr;      // level 1
    uint16_t *gd1 = nullptr, *gs1 = nullptr, *gb1 = nullptr, *ga1 = nullptr;
    uint16_t *b2 = nullptr, *c2 = nullptr, *gc2 = nullptr, *gb2 = nullptr;  // level 2
    float* d_scale = nullptr;     // rms used by the last forward
    void free_acts();
};

// Preconditioner z = M(r) on the fine level of an MGPCG box (flexible PCG).
class LearnedPre : public Preconditioner {
public:
    // the network's activations must be sized for the solve box (LNet::ensure)
    // before apply() is recorded
    LearnedPre(MGPCG* s, LNet* net) : s_(s), net_(net) {}
    void apply(const float* r, float* z, double* partial, SolveScalars* sc, hipStream_t st) override;
    bool flexible() const override { return true; }
    int prepare_box() override;
    long long graph_epoch() const override { return net_->epoch; }

private:
    MGPCG* s_;
    LNet* net_;
};

// Hybrid preconditioner on the fine level of an MGPCG box (flexible PCG):
// z1 = V(r) (one multigrid V-cycle), r1 = r - A z1, z2 = N(r1) (the network on
// the residual the V-cycle leaves), z = z1 + c z2 with the exact line search
// c = (z2.r1) / (z2.A z2), which minimises the A-norm of the error of z1 + c z2
// as an approximation of A^-1 r.  The network is trained on pairs (r1, A^-1 r1)
// of V-cycle residuals.
class HybridPre : public Preconditioner {
public:
    HybridPre(MGPCG* s, LNet* net) : s_(s), net_(net) {}
    ~HybridPre();
    void apply(const float* r, float* z, double* partial, SolveScalars* sc, hipStream_t st) override;
    bool flexible() const override { return true; }
    int prepare_box() override;
    long long graph_epoch() const override { return net_->epoch * 1000003LL + epoch_; }

private:
    MGPCG* s_;
    LNet* net_;
    float *r1_ = nullptr, *z2_ = nullptr, *scale_ = nullptr;
    double* part_ = nullptr;   // 4 partial-sum arrays of RED_BLOCKS doubles
    long long cap_ = 0, epoch_ = 0;
};

}  // namespace fsk
