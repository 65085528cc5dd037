/*
 * cnn_oracle.hpp -- CPU oracle of the learned Richardson-Lucy warm start
 * (test infrastructure only: used by tests/ to check the GPU network).
 *
 * The network has no counterpart in the reference.  It is fixed by the task:
 * four 3x3x3 convolutions (1 -> 16 -> 16 -> 16 -> 1 channels), ReLU after the
 * first three, and a residual output y = x + conv4(...).  Convolutions are
 * "valid": a tile of side S + 8 gives an output of side S, each layer
 * shrinking the side by 2.
 *
 * Parameter vector (float, 14737 values, the order the library uses):
 *   W1[27][16] (tap, cout)   b1[16]
 *   W2[27][16][16] (tap, cin, cout)   b2[16]
 *   W3[27][16][16]           b3[16]
 *   W4[27][16] (tap, cin)    b4[1]
 * with tap = (dz * 3 + dx) * 3 + dy.
 *
 * Arithmetic is double.  With emulate_bf16 the values the GPU keeps in bf16
 * (the weights of layers 1-3, the input taps of layer 1, the activations of
 * layers 1-3, the back-propagated gradients of layers 1-3 and the layer-4
 * output gradient as a GEMM operand) are rounded to bf16 (round to nearest
 * even) at the same points, so the remaining GPU/oracle difference is FP32
 * accumulation only.  Without it, the backward pass is the exact gradient of
 * the double-precision loss (checked by finite differences).
 */
#ifndef APR_CNN_ORACLE_HPP
#define APR_CNN_ORACLE_HPP

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace cnn_oracle {

constexpr int C = 16;
constexpr int NT = 27;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NT * C, OFF_W2 = OFF_B1 + C, OFF_B2 = OFF_W2 + NT * C * C, OFF_W3 = OFF_B2 + C,
              OFF_B3 = OFF_W3 + NT * C * C, OFF_W4 = OFF_B3 + C, OFF_B4 = OFF_W4 + NT * C, N_PARAMS = OFF_B4 + 1;

inline double bf16r(double v) {
    /* float -> bf16 round to nearest even, back to double */
    float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return v;   /* inf / nan unchanged */
    const uint32_t lsb = (u >> 16) & 1u;
    u = (u + 0x7fffu + lsb) & 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return (double)f;
}

struct Net {
    bool bf = false;
    std::vector<double> W;   /* parameters as used in the forward pass */
    explicit Net(const float* p, bool emulate) : bf(emulate), W(p, p + N_PARAMS) {
        if (bf) {
            for (int i = OFF_W1; i < OFF_B1; ++i) W[i] = bf16r(W[i]);
            for (int i = OFF_W2; i < OFF_B2; ++i) W[i] = bf16r(W[i]);
            for (int i = OFF_W3; i < OFF_B3; ++i) W[i] = bf16r(W[i]);
        }
    }
    double r(double v) const { return bf ? bf16r(v) : v; }
};

/* activations of one tile; sides S+6, S+4, S+2 */
struct Acts {
    int S = 0;
    std::vector<double> a1, a2, a3, y;   /* channels-last for a1..a3 */
};

inline int64_t vid(int64_t z, int64_t x, int64_t y, int64_t s) { return (z * s + x) * s + y; }

/* forward of one tile x[(S+8)^3] */
inline void forward(const Net& n, const float* x, int S, Acts& A) {
    A.S = S;
    const int s0 = S + 8, s1 = S + 6, s2 = S + 4, s3 = S + 2;
    A.a1.assign((size_t)s1 * s1 * s1 * C, 0.0);
    for (int z = 0; z < s1; ++z)
        for (int xx = 0; xx < s1; ++xx)
            for (int y = 0; y < s1; ++y)
                for (int c = 0; c < C; ++c) {
                    double acc = 0;
                    for (int t = 0; t < NT; ++t) {
                        const int dz = t / 9, dx = (t / 3) % 3, dy = t % 3;
                        acc += n.W[OFF_W1 + t * C + c] * n.r(x[vid(z + dz, xx + dx, y + dy, s0)]);
                    }
                    acc += n.W[OFF_B1 + c];
                    A.a1[vid(z, xx, y, s1) * C + c] = n.r(acc > 0 ? acc : 0.0);
                }
    auto conv16 = [&](const std::vector<double>& in, int sin, std::vector<double>& out, int sout, int offw, int offb) {
        out.assign((size_t)sout * sout * sout * C, 0.0);
        for (int z = 0; z < sout; ++z)
            for (int xx = 0; xx < sout; ++xx)
                for (int y = 0; y < sout; ++y)
                    for (int co = 0; co < C; ++co) {
                        double acc = 0;
                        for (int t = 0; t < NT; ++t) {
                            const int dz = t / 9, dx = (t / 3) % 3, dy = t % 3;
                            const double* ip = &in[vid(z + dz, xx + dx, y + dy, sin) * C];
                            for (int ci = 0; ci < C; ++ci) acc += n.W[offw + (t * C + ci) * C + co] * ip[ci];
                        }
                        acc += n.W[offb + co];
                        out[vid(z, xx, y, sout) * C + co] = n.r(acc > 0 ? acc : 0.0);
                    }
    };
    conv16(A.a1, s1, A.a2, s2, OFF_W2, OFF_B2);
    conv16(A.a2, s2, A.a3, s3, OFF_W3, OFF_B3);
    A.y.assign((size_t)S * S * S, 0.0);
    for (int z = 0; z < S; ++z)
        for (int xx = 0; xx < S; ++xx)
            for (int y = 0; y < S; ++y) {
                double acc = 0;
                for (int t = 0; t < NT; ++t) {
                    const int dz = t / 9, dx = (t / 3) % 3, dy = t % 3;
                    const double* ip = &A.a3[vid(z + dz, xx + dx, y + dy, s3) * C];
                    for (int ci = 0; ci < C; ++ci) acc += n.W[OFF_W4 + t * C + ci] * ip[ci];
                }
                A.y[vid(z, xx, y, S)] = (double)x[vid(z + 4, xx + 4, y + 4, s0)] + n.W[OFF_B4] + acc;
            }
}

/* loss = sum over the batch of (y - t)^2 * inv_n; adds d loss / d params into grad */
inline double backward(const Net& n, const float* x, const float* tgt, const Acts& A, double inv_n, std::vector<double>& grad) {
    const int S = A.S, s0 = S + 8, s1 = S + 6, s2 = S + 4, s3 = S + 2;
    double loss = 0;
    std::vector<double> g4((size_t)S * S * S), g4r((size_t)S * S * S);
    for (size_t i = 0; i < g4.size(); ++i) {
        const double d = A.y[i] - (double)tgt[i];
        loss += d * d * inv_n;
        g4[i] = 2.0 * d * inv_n;
        g4r[i] = n.r(g4[i]);
    }
    /* layer 4 */
    std::vector<double> g3((size_t)s3 * s3 * s3 * C, 0.0);
    for (int z = 0; z < S; ++z)
        for (int xx = 0; xx < S; ++xx)
            for (int y = 0; y < S; ++y) {
                const double g = g4[vid(z, xx, y, S)], gr = g4r[vid(z, xx, y, S)];
                grad[OFF_B4] += g;
                for (int t = 0; t < NT; ++t) {
                    const int dz = t / 9, dx = (t / 3) % 3, dy = t % 3;
                    const int64_t u = vid(z + dz, xx + dx, y + dy, s3) * C;
                    for (int ci = 0; ci < C; ++ci) {
                        grad[OFF_W4 + t * C + ci] += A.a3[u + ci] * gr;
                        g3[u + ci] += n.W[OFF_W4 + t * C + ci] * g;
                    }
                }
            }
    for (size_t i = 0; i < g3.size(); ++i) g3[i] = n.r(A.a3[i] > 0 ? g3[i] : 0.0);
    /* layers 3, 2: weight gradients and the gradient of their input */
    auto back16 = [&](const std::vector<double>& ain, int sin, const std::vector<double>& gout, int sout, int offw, int offb,
                      std::vector<double>* gin) {
        if (gin) gin->assign((size_t)sin * sin * sin * C, 0.0);
        for (int z = 0; z < sout; ++z)
            for (int xx = 0; xx < sout; ++xx)
                for (int y = 0; y < sout; ++y) {
                    const double* g = &gout[vid(z, xx, y, sout) * C];
                    for (int co = 0; co < C; ++co) grad[offb + co] += g[co];
                    for (int t = 0; t < NT; ++t) {
                        const int dz = t / 9, dx = (t / 3) % 3, dy = t % 3;
                        const int64_t u = vid(z + dz, xx + dx, y + dy, sin) * C;
                        for (int ci = 0; ci < C; ++ci)
                            for (int co = 0; co < C; ++co) {
                                grad[offw + (t * C + ci) * C + co] += ain[u + ci] * g[co];
                                if (gin) (*gin)[u + ci] += n.W[offw + (t * C + ci) * C + co] * g[co];
                            }
                    }
                }
        if (gin)
            for (size_t i = 0; i < gin->size(); ++i) (*gin)[i] = n.r(ain[i] > 0 ? (*gin)[i] : 0.0);
    };
    std::vector<double> g2, g1;
    back16(A.a2, s2, g3, s3, OFF_W3, OFF_B3, &g2);
    back16(A.a1, s1, g2, s2, OFF_W2, OFF_B2, &g1);
    /* layer 1 */
    for (int z = 0; z < s1; ++z)
        for (int xx = 0; xx < s1; ++xx)
            for (int y = 0; y < s1; ++y) {
                const double* g = &g1[vid(z, xx, y, s1) * C];
                for (int co = 0; co < C; ++co) grad[OFF_B1 + co] += g[co];
                for (int t = 0; t < NT; ++t) {
                    const int dz = t / 9, dx = (t / 3) % 3, dy = t % 3;
                    const double xv = n.r(x[vid(z + dz, xx + dx, y + dy, s0)]);
                    for (int co = 0; co < C; ++co) grad[OFF_W1 + t * C + co] += xv * g[co];
                }
            }
    return loss;
}

}  // namespace cnn_oracle

#endif
