// FP64 CPU oracle of the learned luminance denoiser (TEST INFRASTRUCTURE ONLY).
//
// The network has no counterpart in the reference: this file restates the
// library's own definition so that the GPU forward, backward and Adam kernels
// can be checked.  It is pinned independently by the tests (torch's FP64
// conv2d and autograd on the CPU, and central finite differences).
//
// Network on an n x H x W batch of 1-channel inputs x, zero padding:
//     a1 = relu(conv(x; Win[32][9], bin))
//     a(k+1) = relu(conv(ak; Wk[32][9][32], bk)),  k = 1..6
//     out = x + conv(a7; Wout[9][32], bout)
// Parameter layout (one flat array): Win, bin, then per middle layer
// W[co][tap][ci], b[co], then Wout[tap][ci], bout; tap = 3*(dy+1) + (dx+1).
// Loss: mean over the pixels at least `margin` from the crop border of
// (out - tgt)^2.
//
// With bf16 != 0 the forward rounds what the GPU feeds its BF16 MFMAs to the
// nearest bf16, ties to even: every layer's weights, the input layer's x, and
// every activation a1..a7; biases, the residual x and all sums stay FP64.
// The backward is the exact FP64 gradient of that forward, with the
// roundings taken as identities.

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int C = 32;
constexpr size_t OFF_IN = 0, MID = 320, MID_STRIDE = 32 * 9 * 32 + 32, OFF_OUT = MID + 6 * MID_STRIDE, NPARAM = OFF_OUT + 9 * 32 + 1;

double bf16_round(double v) {
    float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return v;  // inf / nan unchanged
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return (double)f;
}

struct Net {
    int n, H, W;
    bool bf;
    std::vector<double> wmid;  // (possibly rounded) middle weights, 6 x [co][tap][ci]
    std::vector<double> win, wout;  // (possibly rounded) input / output layer weights
    const double* p;
    size_t npx() const { return (size_t)n * H * W; }
    size_t at(int img, int y, int x) const { return ((size_t)img * H + y) * W + x; }
};

// a (npx x C) -> o (npx x C): o = conv(a; w[co][tap][ci]) + b, no activation
void conv_mid(const Net& N, const double* a, const double* w, const double* b, double* o) {
    for (int img = 0; img < N.n; img++)
        for (int y = 0; y < N.H; y++)
            for (int x = 0; x < N.W; x++) {
                double* op = o + N.at(img, y, x) * C;
                for (int co = 0; co < C; co++) op[co] = b[co];
                for (int t = 0; t < 9; t++) {
                    const int yy = y + t / 3 - 1, xx = x + t % 3 - 1;
                    if (yy < 0 || yy >= N.H || xx < 0 || xx >= N.W) continue;
                    const double* ap = a + N.at(img, yy, xx) * C;
                    for (int co = 0; co < C; co++) {
                        const double* wp = w + ((size_t)co * 9 + t) * C;
                        double s = 0.0;
                        for (int ci = 0; ci < C; ci++) s += wp[ci] * ap[ci];
                        op[co] += s;
                    }
                }
            }
}

// Forward; acts[k] (k = 0..6) hold a(k+1) after the activation (and rounding).
void forward(Net& N, const float* x, std::vector<std::vector<double>>& acts, double* out) {
    const size_t np = N.npx();
    const double* p = N.p;
    acts.assign(7, std::vector<double>(np * C));
    std::vector<double> pre(np * C);
    for (int img = 0; img < N.n; img++)
        for (int y = 0; y < N.H; y++)
            for (int xx = 0; xx < N.W; xx++) {
                double* ap = &acts[0][N.at(img, y, xx) * C];
                for (int co = 0; co < C; co++) {
                    double s = p[OFF_IN + 9 * C + co];
                    for (int t = 0; t < 9; t++) {
                        const int yy = y + t / 3 - 1, xc = xx + t % 3 - 1;
                        if (yy < 0 || yy >= N.H || xc < 0 || xc >= N.W) continue;
                        const double xv = (double)x[N.at(img, yy, xc)];
                        s += N.win[co * 9 + t] * (N.bf ? bf16_round(xv) : xv);
                    }
                    s = s > 0.0 ? s : 0.0;
                    ap[co] = N.bf ? bf16_round(s) : s;
                }
            }
    for (int l = 0; l < 6; l++) {
        conv_mid(N, acts[l].data(), &N.wmid[(size_t)l * C * 9 * C], p + MID + l * MID_STRIDE + C * 9 * C, pre.data());
        for (size_t i = 0; i < np * C; i++) {
            const double v = pre[i] > 0.0 ? pre[i] : 0.0;
            acts[l + 1][i] = N.bf ? bf16_round(v) : v;
        }
    }
    for (int img = 0; img < N.n; img++)
        for (int y = 0; y < N.H; y++)
            for (int xx = 0; xx < N.W; xx++) {
                double s = p[OFF_OUT + 9 * C];
                for (int t = 0; t < 9; t++) {
                    const int yy = y + t / 3 - 1, xc = xx + t % 3 - 1;
                    if (yy < 0 || yy >= N.H || xc < 0 || xc >= N.W) continue;
                    const double* ap = &acts[6][N.at(img, yy, xc) * C];
                    for (int ci = 0; ci < C; ci++) s += N.wout[t * C + ci] * ap[ci];
                }
                const size_t q = N.at(img, y, xx);
                out[q] = (double)x[q] + s;
            }
}

void init_net(Net& N, const double* params, int n, int H, int W, int bf16) {
    N.n = n;
    N.H = H;
    N.W = W;
    N.bf = bf16 != 0;
    N.p = params;
    N.win.resize(C * 9);
    N.wout.resize(9 * C);
    for (int i = 0; i < C * 9; i++) N.win[i] = N.bf ? bf16_round(params[OFF_IN + i]) : params[OFF_IN + i];
    for (int i = 0; i < 9 * C; i++) N.wout[i] = N.bf ? bf16_round(params[OFF_OUT + i]) : params[OFF_OUT + i];
    N.wmid.resize((size_t)6 * C * 9 * C);
    for (int l = 0; l < 6; l++)
        for (size_t i = 0; i < (size_t)C * 9 * C; i++) {
            const double v = params[MID + l * MID_STRIDE + i];
            N.wmid[(size_t)l * C * 9 * C + i] = N.bf ? bf16_round(v) : v;
        }
}

}  // namespace

extern "C" {

size_t or_cnn_param_count(void) { return NPARAM; }

double or_cnn_bf16_round(double v) { return bf16_round(v); }

// out (n*H*W) = network(x); FP64.
void or_cnn_forward_f64(const double* params, const float* x, int n, int H, int W, int bf16, double* out) {
    Net N;
    init_net(N, params, n, H, W, bf16);
    std::vector<std::vector<double>> acts;
    forward(N, x, acts, out);
}

// Returns the loss; grad (NPARAM, nullable) receives d loss / d params.
double or_cnn_loss_grad_f64(const double* params, const float* x, const float* tgt, int n, int H, int W, int margin, int bf16, double* grad) {
    Net N;
    init_net(N, params, n, H, W, bf16);
    const size_t np = N.npx();
    std::vector<std::vector<double>> acts;
    std::vector<double> out(np);
    forward(N, x, acts, out.data());
    const long nvalid = (long)n * (H - 2 * margin) * (W - 2 * margin);
    const double inv_n = 1.0 / (double)(nvalid > 0 ? nvalid : 1);
    double loss = 0.0;
    std::vector<double> dout(np, 0.0);
    for (int img = 0; img < n; img++)
        for (int y = margin; y < H - margin; y++)
            for (int xx = margin; xx < W - margin; xx++) {
                const size_t q = N.at(img, y, xx);
                const double e = out[q] - (double)tgt[q];
                loss += e * e;
                dout[q] = 2.0 * e * inv_n;
            }
    loss *= inv_n;
    if (!grad) return loss;
    std::memset(grad, 0, NPARAM * sizeof(double));
    // output layer
    std::vector<double> g(np * C, 0.0);  // d loss / d a7
    for (int img = 0; img < n; img++)
        for (int y = 0; y < H; y++)
            for (int xx = 0; xx < W; xx++) {
                const double d = dout[N.at(img, y, xx)];
                if (d == 0.0) continue;
                grad[OFF_OUT + 9 * C] += d;
                for (int t = 0; t < 9; t++) {
                    const int yy = y + t / 3 - 1, xc = xx + t % 3 - 1;
                    if (yy < 0 || yy >= H || xc < 0 || xc >= W) continue;
                    const size_t q = N.at(img, yy, xc) * C;
                    for (int ci = 0; ci < C; ci++) {
                        grad[OFF_OUT + t * C + ci] += d * acts[6][q + ci];
                        g[q + ci] += d * N.wout[t * C + ci];
                    }
                }
            }
    // middle layers, l = 5..0: acts[l] -> acts[l+1]
    std::vector<double> gin(np * C);
    for (int l = 5; l >= 0; l--) {
        // through the relu of acts[l+1]
        for (size_t i = 0; i < np * C; i++)
            if (!(acts[l + 1][i] > 0.0)) g[i] = 0.0;
        double* gw = grad + MID + l * MID_STRIDE;
        double* gb = gw + C * 9 * C;
        const double* w = &N.wmid[(size_t)l * C * 9 * C];
        std::fill(gin.begin(), gin.end(), 0.0);
        for (int img = 0; img < n; img++)
            for (int y = 0; y < H; y++)
                for (int xx = 0; xx < W; xx++) {
                    const double* gp = &g[N.at(img, y, xx) * C];
                    for (int co = 0; co < C; co++) gb[co] += gp[co];
                    for (int t = 0; t < 9; t++) {
                        const int yy = y + t / 3 - 1, xc = xx + t % 3 - 1;
                        if (yy < 0 || yy >= H || xc < 0 || xc >= W) continue;
                        const size_t q = N.at(img, yy, xc) * C;
                        const double* ap = &acts[l][q];
                        double* gi = &gin[q];
                        for (int co = 0; co < C; co++) {
                            const double d = gp[co];
                            if (d == 0.0) continue;
                            double* gwp = gw + ((size_t)co * 9 + t) * C;
                            const double* wp = w + ((size_t)co * 9 + t) * C;
                            for (int ci = 0; ci < C; ci++) {
                                gwp[ci] += d * ap[ci];
                                gi[ci] += d * wp[ci];
                            }
                        }
                    }
                }
        g.swap(gin);
    }
    // input layer: g = d loss / d a1 -> through its relu
    for (size_t i = 0; i < np * C; i++)
        if (!(acts[0][i] > 0.0)) g[i] = 0.0;
    for (int img = 0; img < n; img++)
        for (int y = 0; y < H; y++)
            for (int xx = 0; xx < W; xx++) {
                const double* gp = &g[N.at(img, y, xx) * C];
                for (int co = 0; co < C; co++) {
                    grad[OFF_IN + 9 * C + co] += gp[co];
                    for (int t = 0; t < 9; t++) {
                        const int yy = y + t / 3 - 1, xc = xx + t % 3 - 1;
                        if (yy < 0 || yy >= H || xc < 0 || xc >= W) continue;
                        const double xv = (double)x[N.at(img, yy, xc)];
                        grad[OFF_IN + co * 9 + t] += gp[co] * (N.bf ? bf16_round(xv) : xv);
                    }
                }
            }
    return loss;
}

// One Adam step (b1 0.9, b2 0.999, eps 1e-8) on FP64 state; step counts from 1.
void or_cnn_adam_f64(double* w, const double* g, double* m, double* v, size_t n, int step, double lr) {
    const double b1 = 0.9, b2 = 0.999, eps = 1e-8;
    const double bc1 = 1.0 - std::pow(b1, step), bc2 = 1.0 - std::pow(b2, step);
    for (size_t i = 0; i < n; i++) {
        m[i] = b1 * m[i] + (1.0 - b1) * g[i];
        v[i] = b2 * v[i] + (1.0 - b2) * g[i] * g[i];
        w[i] -= lr * (m[i] / bc1) / (std::sqrt(v[i] / bc2) + eps);
    }
}

// Network input from a luminance plane: x = clamp((L - mean5x5(L)) / s, -64, 64),
// s = max(sigma, 1e-4), the 5x5 mean over clamped coordinates.
void or_cnn_input_f64(const float* L, int w, int h, double sigma, double* x) {
    const double s = sigma > 1e-4 ? sigma : 1e-4;
    for (int y = 0; y < h; y++)
        for (int xx = 0; xx < w; xx++) {
            double m = 0.0;
            for (int dy = -2; dy <= 2; dy++)
                for (int dx = -2; dx <= 2; dx++) {
                    const int yy = y + dy < 0 ? 0 : (y + dy >= h ? h - 1 : y + dy);
                    const int xc = xx + dx < 0 ? 0 : (xx + dx >= w ? w - 1 : xx + dx);
                    m += (double)L[(size_t)yy * w + xc];
                }
            double v = ((double)L[(size_t)y * w + xx] - m / 25.0) / s;
            x[(size_t)y * w + xx] = v < -64.0 ? -64.0 : (v > 64.0 ? 64.0 : v);
        }
}

}  // extern "C"
