/*
 * lnet_oracle.cpp -- FP64 CPU restatement of the learned pressure
 * preconditioner network (include/fluidsim_learned.h).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for fs_learned_* of the HIP library.
 * The network has no reference counterpart (the reference solves the pressure
 * system with cuDSS, src/FluidSim/FluidSolver3D.cu:1930-2008), so this file is
 * pinned by closed forms and finite differences (tests/test_learned_oracle.py):
 * single-tap weights reduce the network to shifts / pooling / upsampling of
 * ReLU(r) that the test computes independently, and central differences of the
 * FP64 loss check every parameter block of the gradient.
 *
 * bf16 != 0 rounds values to BF16 (round-to-nearest-even) at the points where
 * the HIP implementation stores BF16: activations, gradients of activations,
 * and the weights of the 16-channel convolutions.  bf16 == 0 is the exact
 * FP64 network.
 */
#include <stdint.h>

#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr int C16 = 16, TAPS = 27;
constexpr int AW = 0, AB = AW + TAPS * 3 * 16;
constexpr int BW = AB + 16, BB = BW + TAPS * 256;
constexpr int CW = BB + 16, CB = CW + TAPS * 256;
constexpr int DW = CB + 16, DB = DW + TAPS * 256;
constexpr int EW = DB + 16, EB = EW + TAPS * 16;
constexpr int NPARAM = EB + 1;

double bfr(double x) {
    float f = (float)x;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return x;
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return f;
}

struct Grid {
    int W, H, D;
    long long n() const { return (long long)W * H * D; }
    long long at(int x, int y, int z) const { return x + (long long)W * (y + (long long)H * z); }
    bool in(int x, int y, int z) const { return x >= 0 && y >= 0 && z >= 0 && x < W && y < H && z < D; }
    Grid half() const { return Grid{(W + 1) / 2, (H + 1) / 2, (D + 1) / 2}; }
};

void tap_offset(int t, int& dx, int& dy, int& dz) {
    dx = t % 3 - 1;
    dy = (t / 3) % 3 - 1;
    dz = t / 9 - 1;
}

// out[v][co] = bias[co] + sum_t sum_ci w[t][ci][co] in[v + d_t][ci]
std::vector<double> conv(const Grid& g, const std::vector<double>& in, int ci_n, const double* w, const double* bias,
                         int co_n) {
    std::vector<double> out(g.n() * co_n, 0.0);
    for (int z = 0; z < g.D; ++z)
        for (int y = 0; y < g.H; ++y)
            for (int x = 0; x < g.W; ++x) {
                double* o = &out[g.at(x, y, z) * co_n];
                for (int co = 0; co < co_n; ++co) o[co] = bias ? bias[co] : 0.0;
                for (int t = 0; t < TAPS; ++t) {
                    int dx, dy, dz;
                    tap_offset(t, dx, dy, dz);
                    if (!g.in(x + dx, y + dy, z + dz)) continue;
                    const double* s = &in[g.at(x + dx, y + dy, z + dz) * ci_n];
                    for (int ci = 0; ci < ci_n; ++ci)
                        for (int co = 0; co < co_n; ++co) o[co] += w[(t * ci_n + ci) * co_n + co] * s[ci];
                }
            }
    return out;
}

// din[u][ci] = sum_t sum_co w[t][ci][co] dout[u - d_t][co]
std::vector<double> conv_t(const Grid& g, const std::vector<double>& dout, int co_n, const double* w, int ci_n) {
    std::vector<double> din(g.n() * ci_n, 0.0);
    for (int z = 0; z < g.D; ++z)
        for (int y = 0; y < g.H; ++y)
            for (int x = 0; x < g.W; ++x) {
                double* o = &din[g.at(x, y, z) * ci_n];
                for (int t = 0; t < TAPS; ++t) {
                    int dx, dy, dz;
                    tap_offset(t, dx, dy, dz);
                    if (!g.in(x - dx, y - dy, z - dz)) continue;
                    const double* s = &dout[g.at(x - dx, y - dy, z - dz) * co_n];
                    for (int ci = 0; ci < ci_n; ++ci)
                        for (int co = 0; co < co_n; ++co) o[ci] += w[(t * ci_n + ci) * co_n + co] * s[co];
                }
            }
    return din;
}

// dw[t][ci][co] += sum_v in[v + d_t][ci] dout[v][co] ; db[co] += sum_v dout[v][co]
void conv_wgrad(const Grid& g, const std::vector<double>& in, int ci_n, const std::vector<double>& dout, int co_n,
                double* dw, double* db) {
    for (int z = 0; z < g.D; ++z)
        for (int y = 0; y < g.H; ++y)
            for (int x = 0; x < g.W; ++x) {
                const double* d = &dout[g.at(x, y, z) * co_n];
                for (int co = 0; co < co_n; ++co) db[co] += d[co];
                for (int t = 0; t < TAPS; ++t) {
                    int dx, dy, dz;
                    tap_offset(t, dx, dy, dz);
                    if (!g.in(x + dx, y + dy, z + dz)) continue;
                    const double* s = &in[g.at(x + dx, y + dy, z + dz) * ci_n];
                    for (int ci = 0; ci < ci_n; ++ci)
                        for (int co = 0; co < co_n; ++co) dw[(t * ci_n + ci) * co_n + co] += s[ci] * d[co];
                }
            }
}

// average of the 2x2x2 children (zero outside the fine grid), C channels
std::vector<double> pool(const Grid& f, const Grid& c, const std::vector<double>& in, int C) {
    std::vector<double> out(c.n() * C, 0.0);
    for (int K = 0; K < c.D; ++K)
        for (int J = 0; J < c.H; ++J)
            for (int I = 0; I < c.W; ++I)
                for (int ch = 0; ch < C; ++ch) {
                    double s = 0.0;
                    for (int q = 0; q < 8; ++q) {
                        int i = 2 * I + (q & 1), j = 2 * J + ((q >> 1) & 1), k = 2 * K + (q >> 2);
                        if (f.in(i, j, k)) s += in[f.at(i, j, k) * C + ch];
                    }
                    out[c.at(I, J, K) * C + ch] = 0.125 * s;
                }
    return out;
}

// nearest-neighbour upsampling value of the coarse parent
double up_at(const Grid& c, const std::vector<double>& coarse, int x, int y, int z, int ch, int C) {
    return coarse[c.at(x >> 1, y >> 1, z >> 1) * C + ch];
}

// sum over the children of each coarse cell (adjoint of up)
std::vector<double> up_t(const Grid& f, const Grid& c, const std::vector<double>& fine, int C) {
    std::vector<double> out(c.n() * C, 0.0);
    for (int z = 0; z < f.D; ++z)
        for (int y = 0; y < f.H; ++y)
            for (int x = 0; x < f.W; ++x)
                for (int ch = 0; ch < C; ++ch) out[c.at(x >> 1, y >> 1, z >> 1) * C + ch] += fine[f.at(x, y, z) * C + ch];
    return out;
}

// A x of the pressure matrix (FluidSolver3D.cu:1598-1716) on FLUID cells: diag
// = #in-grid FLUID/AIR neighbours, -1 per FLUID neighbour; 0 elsewhere
std::vector<double> apply_A(const Grid& g, const uint8_t* labels, const std::vector<double>& x) {
    std::vector<double> out(g.n(), 0.0);
    const int off[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (int z = 0; z < g.D; ++z)
        for (int y = 0; y < g.H; ++y)
            for (int xx = 0; xx < g.W; ++xx) {
                const long long c = g.at(xx, y, z);
                if (labels[c] != 1) continue;
                double s = 0.0;
                for (const auto& o : off) {
                    if (!g.in(xx + o[0], y + o[1], z + o[2])) continue;
                    const long long n = g.at(xx + o[0], y + o[1], z + o[2]);
                    if (labels[n] == 1 || labels[n] == 2) s += x[c];
                    if (labels[n] == 1) s -= x[n];
                }
                out[c] = s;
            }
    return out;
}

struct Net {
    Grid g0, g1, g2;
    bool bf;
    std::vector<double> prm;          // FP64 parameters
    std::vector<double> wb, wc, wd, we;   // conv weights as used by the 16-channel kernels
    std::vector<double> x0;           // [rn, fluid, air]
    std::vector<double> a0, a1, b1, b2, c2, s1, d1, s0, zn;
    double scale = 1.0;

    double R(double v) const { return bf ? bfr(v) : v; }
    void Rv(std::vector<double>& v) const {
        if (bf)
            for (double& x : v) x = bfr(x);
    }

    Net(const float* p, int W, int H, int D, int bf16) : bf(bf16 != 0) {
        g0 = Grid{W, H, D};
        g1 = g0.half();
        g2 = g1.half();
        prm.assign(p, p + NPARAM);
        auto cw = [&](int off, int n) {
            std::vector<double> w(prm.begin() + off, prm.begin() + off + n);
            Rv(w);
            return w;
        };
        wb = cw(BW, TAPS * 256);
        wc = cw(CW, TAPS * 256);
        wd = cw(DW, TAPS * 256);
        we = cw(EW, TAPS * 16);
    }

    void forward(const uint8_t* labels, const float* r) {
        const long long n0 = g0.n();
        double ss = 0.0;
        long long nf = 0;
        for (long long v = 0; v < n0; ++v)
            if (labels[v] == 1) { ss += (double)r[v] * r[v]; ++nf; }
        scale = nf > 0 ? std::sqrt(ss / nf) : 0.0;
        if (!(scale > 0.0) || !std::isfinite(scale)) scale = 1.0;
        if (bf) scale = (float)scale;
        x0.assign(n0 * 3, 0.0);
        for (long long v = 0; v < n0; ++v) {
            x0[v * 3] = bf ? (double)(float)(r[v] / (float)scale) : r[v] / scale;
            x0[v * 3 + 1] = labels[v] == 1 ? 1.0 : 0.0;
            x0[v * 3 + 2] = labels[v] == 2 ? 1.0 : 0.0;
        }
        a0 = conv(g0, x0, 3, &prm[AW], &prm[AB], 16);
        for (double& x : a0) x = R(std::max(x, 0.0));
        a1 = pool(g0, g1, a0, 16);
        Rv(a1);
        b1 = conv(g1, a1, 16, wb.data(), &prm[BB], 16);
        for (double& x : b1) x = R(std::max(x, 0.0));
        b2 = pool(g1, g2, b1, 16);
        Rv(b2);
        c2 = conv(g2, b2, 16, wc.data(), &prm[CB], 16);
        for (double& x : c2) x = R(std::max(x, 0.0));
        s1 = b1;
        for (int z = 0; z < g1.D; ++z)
            for (int y = 0; y < g1.H; ++y)
                for (int x = 0; x < g1.W; ++x)
                    for (int ch = 0; ch < 16; ++ch) {
                        double& o = s1[g1.at(x, y, z) * 16 + ch];
                        o = R(o + up_at(g2, c2, x, y, z, ch, 16));
                    }
        d1 = conv(g1, s1, 16, wd.data(), &prm[DB], 16);
        for (double& x : d1) x = R(std::max(x, 0.0));
        s0 = a0;
        for (int z = 0; z < g0.D; ++z)
            for (int y = 0; y < g0.H; ++y)
                for (int x = 0; x < g0.W; ++x)
                    for (int ch = 0; ch < 16; ++ch) {
                        double& o = s0[g0.at(x, y, z) * 16 + ch];
                        o = R(o + up_at(g1, d1, x, y, z, ch, 16));
                    }
        zn = conv(g0, s0, 16, we.data(), &prm[EB], 1);
    }

    // gradient of the A-norm loss 0.5 d.(A d) / nf, d = zn - target on FLUID cells
    double backward(const uint8_t* labels, const double* target, double* grad) {
        const long long n0 = g0.n();
        long long nf = 0;
        for (long long v = 0; v < n0; ++v) nf += labels[v] == 1;
        const double inv = nf > 0 ? 1.0 / nf : 0.0;
        std::vector<double> d(n0, 0.0), dz(n0, 0.0);
        for (long long v = 0; v < n0; ++v)
            if (labels[v] == 1) d[v] = zn[v] - target[v];
        std::vector<double> ad = apply_A(g0, labels, d);
        double loss = 0.0;
        for (long long v = 0; v < n0; ++v) {
            loss += 0.5 * d[v] * ad[v] * inv;
            dz[v] = ad[v] * inv;
        }
        std::vector<double> gw(NPARAM, 0.0);
        // E (data gradient with the FP32 weights)
        conv_wgrad(g0, s0, 16, dz, 1, &gw[EW], &gw[EB]);
        std::vector<double> ds0 = conv_t(g0, dz, 1, &prm[EW], 16);
        Rv(ds0);
        // D
        std::vector<double> gd1 = up_t(g0, g1, ds0, 16);
        for (size_t i = 0; i < gd1.size(); ++i) gd1[i] = R(d1[i] > 0.0 ? gd1[i] : 0.0);
        conv_wgrad(g1, s1, 16, gd1, 16, &gw[DW], &gw[DB]);
        std::vector<double> gs1 = conv_t(g1, gd1, 16, wd.data(), 16);
        Rv(gs1);
        // C
        std::vector<double> gc2 = up_t(g1, g2, gs1, 16);
        for (size_t i = 0; i < gc2.size(); ++i) gc2[i] = R(c2[i] > 0.0 ? gc2[i] : 0.0);
        conv_wgrad(g2, b2, 16, gc2, 16, &gw[CW], &gw[CB]);
        std::vector<double> gb2 = conv_t(g2, gc2, 16, wc.data(), 16);
        Rv(gb2);
        // B
        std::vector<double> gb1(g1.n() * 16);
        for (int z = 0; z < g1.D; ++z)
            for (int y = 0; y < g1.H; ++y)
                for (int x = 0; x < g1.W; ++x)
                    for (int ch = 0; ch < 16; ++ch) {
                        long long i = g1.at(x, y, z) * 16 + ch;
                        double s = gs1[i] + 0.125 * up_at(g2, gb2, x, y, z, ch, 16);
                        gb1[i] = R(b1[i] > 0.0 ? s : 0.0);
                    }
        conv_wgrad(g1, a1, 16, gb1, 16, &gw[BW], &gw[BB]);
        std::vector<double> ga1 = conv_t(g1, gb1, 16, wb.data(), 16);
        Rv(ga1);
        // A
        std::vector<double> ga0(n0 * 16);
        for (int z = 0; z < g0.D; ++z)
            for (int y = 0; y < g0.H; ++y)
                for (int x = 0; x < g0.W; ++x)
                    for (int ch = 0; ch < 16; ++ch) {
                        long long i = g0.at(x, y, z) * 16 + ch;
                        double s = ds0[i] + 0.125 * up_at(g1, ga1, x, y, z, ch, 16);
                        ga0[i] = R(a0[i] > 0.0 ? s : 0.0);
                    }
        conv_wgrad(g0, x0, 3, ga0, 16, &gw[AW], &gw[AB]);
        if (grad)
            for (int i = 0; i < NPARAM; ++i) grad[i] = gw[i];
        return loss;
    }
};

}  // namespace

extern "C" {

int oracle_lnet_num_params(void) { return NPARAM; }

/* z = M(r) (FP64): zn the unscaled output, z = zn * rms * fluid mask */
int oracle_lnet_forward(const float* params, const uint8_t* labels, const float* r, int W, int H, int D, int bf16,
                        double* z, double* zn) {
    if (!params || !labels || !r || W <= 0 || H <= 0 || D <= 0) return -1;
    Net net(params, W, H, D, bf16);
    net.forward(labels, r);
    const long long n = net.g0.n();
    for (long long v = 0; v < n; ++v) {
        if (zn) zn[v] = net.zn[v];
        if (z) z[v] = labels[v] == 1 ? net.zn[v] * net.scale : 0.0;
    }
    return 0;
}

/* A-norm loss 0.5 d.(A d) / nf, d = zn - p / sqrt(p.Ap / nf) on the FLUID cells,
   and its gradient (NPARAM doubles); zn is the unscaled output for the input b
   (normalised by rms(b)) */
int oracle_lnet_loss_grad(const float* params, const uint8_t* labels, const float* b, const double* p, int W, int H,
                          int D, int bf16, double* loss, double* grad) {
    if (!params || !labels || !b || !p || W <= 0 || H <= 0 || D <= 0) return -1;
    Net net(params, W, H, D, bf16);
    net.forward(labels, b);
    const long long n = net.g0.n();
    long long nf = 0;
    std::vector<double> pv(n, 0.0);
    for (long long v = 0; v < n; ++v)
        if (labels[v] == 1) { pv[v] = p[v]; ++nf; }
    std::vector<double> ap = apply_A(net.g0, labels, pv);
    double pap = 0.0;
    for (long long v = 0; v < n; ++v) pap += pv[v] * ap[v];
    double sp = nf > 0 ? std::sqrt(std::max(0.0, pap) / nf) : 0.0;
    if (!(sp > 0.0) || !std::isfinite(sp)) sp = 1.0;
    std::vector<double> t(n, 0.0);
    for (long long v = 0; v < n; ++v)
        if (labels[v] == 1) t[v] = bf16 ? (double)(float)(p[v] / sp) : p[v] / sp;
    double l = net.backward(labels, t.data(), grad);
    if (loss) *loss = l;
    return 0;
}

}  // extern "C"
