// denoise.cpp — oracle for the learned depth denoiser: forward pass, loss and
// backward pass in double precision.  Test infrastructure only (see oracle.h).
//
// Network (no reference counterpart): for output pixel p with depth z_p
// (metres), inverse depth w_p = 1000 / d_p and inverse-depth noise
// sigma_w = sigma_z(z_p) w_p^2 (sigma_z = (1.2 + 1.9 (z - 0.4)^2) mm), the
// input channels at tap q are x0 = clamp((w_q - w_p) / sigma_w, -8, 8) (0
// where q or p has no depth) and x1 = valid mask of q; four 3x3
// convolutions with zero padding, 2 -> 16 -> 16 -> 16 -> 1 channels, ReLU
// after the first three; output y = 1 / max(w_p + sigma_w conv4, w_p / 4) on
// valid pixels (residual in inverse-depth noise units), 0 elsewhere; the
// clamp's derivative is taken as 1 (as on the GPU).  The normalisation is evaluated in
// single precision exactly as the GPU does.  Parameters (5,089, flat fp32):
// per layer the weights [co][ci][ky][kx] then the biases [co].  Loss: mean
// over pixels with a valid input, a valid target and |z_p - target| <
// outlier of (y - target)^2.
//
// With bf16 != 0 the oracle rounds what the GPU stores in bfloat16 (weights
// and the three hidden activations, round to nearest even) before using it,
// so the comparison with the GPU isolates accumulation-order differences;
// with bf16 == 0 everything is double (for finite-difference checks).  The
// GPU also stores the hidden layers' gradients in bfloat16; it propagates the
// unscaled output gradient 2 (y - t) and applies 1 / count at the end, which
// the bf16 mode follows.
#include <math.h>
#include <string.h>

#include <vector>

#include "oracle.h"

namespace {

const int kC = 16;
const int kLayerCin[4] = {2, 16, 16, 16};
const int kLayerCout[4] = {16, 16, 16, 1};

int layer_offset(int l) {
    int o = 0;
    for (int k = 0; k < l; ++k) o += kLayerCout[k] * kLayerCin[k] * 9 + kLayerCout[k];
    return o;
}

double bf16_round(double v) {
    float f = (float)v;
    uint32_t b;
    memcpy(&b, &f, 4);
    b += 0x7FFFu + ((b >> 16) & 1u);
    b &= 0xFFFF0000u;
    memcpy(&f, &b, 4);
    return f;
}

float noise_sigma_m(float z) {
    const float dz = z - 0.4f;
    return (1.2f + 1.9f * (dz * dz)) * 0.001f;
}

struct Net {
    int W, H;
    bool bf16;
    std::vector<double> w;                  // parameters (rounded if bf16)
    std::vector<double> x;                  // [p][2]: depth (m), valid
    std::vector<float> wf, inv_sigma, sigma; // single-precision inverse depth, 1 / sigma_w, sigma_w per pixel
    std::vector<double> a[3];               // hidden activations [p][16]
    std::vector<double> y;                  // output
    std::vector<uint8_t> valid;
};

// Input of layer l at tap (px, py) for the output pixel c (layer 0 inputs
// are relative to c).
double in_val(const Net& n, int l, const std::vector<double>* src, int px, int py, int ci, int c) {
    if (px < 0 || py < 0 || px >= n.W || py >= n.H) return 0.0;
    const int p = py * n.W + px;
    if (l == 0) {
        if (ci == 1) return n.x[2 * p + 1];
        if (!n.valid[p] || !n.valid[c]) return 0.0;
        float r = (n.wf[p] - n.wf[c]) * n.inv_sigma[c];
        r = fminf(fmaxf(r, -8.0f), 8.0f);
        return (double)r;
    }
    return (*src)[(size_t)p * kC + ci];
}

void forward(Net& n) {
    const int P = n.W * n.H;
    for (int l = 0; l < 4; ++l) {
        const int cin = kLayerCin[l], cout = kLayerCout[l], off = layer_offset(l);
        const std::vector<double>* src = l == 0 ? nullptr : &n.a[l - 1];
        std::vector<double> out((size_t)P * cout);
        for (int py = 0; py < n.H; ++py)
            for (int px = 0; px < n.W; ++px)
                for (int co = 0; co < cout; ++co) {
                    double s = n.w[off + cout * cin * 9 + co];
                    for (int ci = 0; ci < cin; ++ci)
                        for (int ky = 0; ky < 3; ++ky)
                            for (int kx = 0; kx < 3; ++kx)
                                s += n.w[off + ((co * cin + ci) * 3 + ky) * 3 + kx] *
                                     in_val(n, l, src, px + kx - 1, py + ky - 1, ci, py * n.W + px);
                    out[(size_t)(py * n.W + px) * cout + co] = s;
                }
        if (l < 3) {
            for (double& v : out) {
                v = v > 0.0 ? v : 0.0;
                if (n.bf16) v = bf16_round(v);
            }
            n.a[l] = out;
        } else {
            n.y.assign(P, 0.0);
            for (int p = 0; p < P; ++p) {
                if (!n.valid[p]) { n.y[p] = 0.0; continue; }
                const double w = fmax((double)n.wf[p] + (double)n.sigma[p] * out[p], 0.25 * (double)n.wf[p]);
                n.y[p] = 1.0 / w;
            }
        }
    }
}

void setup(Net& n, const uint16_t* depth, int W, int H, const float* params, int bf16) {
    n.W = W;
    n.H = H;
    n.bf16 = bf16 != 0;
    const int np = layer_offset(4);
    n.w.resize(np);
    for (int i = 0; i < np; ++i) n.w[i] = n.bf16 ? bf16_round(params[i]) : (double)params[i];
    // Biases stay fp32 on the GPU.
    for (int l = 0; l < 4; ++l) {
        const int off = layer_offset(l) + kLayerCout[l] * kLayerCin[l] * 9;
        for (int co = 0; co < kLayerCout[l]; ++co) n.w[off + co] = params[off + co];
    }
    const int P = W * H;
    n.x.assign(2 * (size_t)P, 0.0);
    n.valid.assign(P, 0);
    n.wf.assign(P, 0.0f);
    n.sigma.assign(P, 0.0f);
    n.inv_sigma.assign(P, 0.0f);
    for (int p = 0; p < P; ++p) {
        if (!depth[p]) continue;
        const float w = 1000.0f / (float)depth[p];
        n.wf[p] = w;
        n.sigma[p] = noise_sigma_m((float)depth[p] * 0.001f) * (w * w);
        n.inv_sigma[p] = 1.0f / n.sigma[p];
        n.valid[p] = 1;
        n.x[2 * p] = (double)((float)depth[p] * 0.001f);
        n.x[2 * p + 1] = 1.0;
    }
}

}  // namespace

extern "C" int orc_dn_param_count(void) { return layer_offset(4); }

// Forward pass: y (metres, 0 on invalid pixels); hidden activations optional
// (3 * W * H * 16 doubles, [layer][p][c]).
extern "C" void orc_dn_forward(const uint16_t* depth, int W, int H, const float* params, int bf16, double* y,
                               double* hidden) {
    Net n;
    setup(n, depth, W, H, params, bf16);
    forward(n);
    memcpy(y, n.y.data(), sizeof(double) * W * H);
    if (hidden)
        for (int l = 0; l < 3; ++l) memcpy(hidden + (size_t)l * W * H * kC, n.a[l].data(), sizeof(double) * W * H * kC);
}

// Loss and its gradient with respect to every parameter (5,089 doubles).
// Returns the loss; *count receives the number of pixels in it.
extern "C" double orc_dn_loss_grad(const uint16_t* depth, const float* target_m, int W, int H, const float* params,
                                   int bf16, double outlier_m, double* grad, int* count) {
    Net n;
    setup(n, depth, W, H, params, bf16);
    forward(n);
    const int P = W * H;
    std::vector<double> dy(P, 0.0);
    int cnt = 0;
    double loss = 0.0;
    for (int p = 0; p < P; ++p) {
        const double t = target_m[p];
        if (!n.valid[p] || !(t > 0.0) || !(fabs(n.x[2 * p] - t) < outlier_m)) continue;
        ++cnt;
    }
    for (int p = 0; p < P; ++p) {
        const double t = target_m[p];
        if (!n.valid[p] || !(t > 0.0) || !(fabs(n.x[2 * p] - t) < outlier_m)) continue;
        const double e = n.y[p] - t;
        loss += e * e;
        // dL/d(conv4) = -2 e y^2 sigma_w (y = 1 / (w_p + sigma_w conv4))
        if (n.bf16) {
            const float yf = (float)n.y[p], ef = (float)e;
            dy[p] = (double)(-(2.0f * ef) * (yf * yf) * n.sigma[p]);
        } else {
            dy[p] = -2.0 * e * n.y[p] * n.y[p] * (double)n.sigma[p];
        }
    }
    if (count) *count = cnt;
    if (cnt == 0) {
        memset(grad, 0, sizeof(double) * layer_offset(4));
        return 0.0;
    }
    loss /= cnt;
    memset(grad, 0, sizeof(double) * layer_offset(4));
    // dz of the output layer: dy on valid pixels.
    std::vector<double> dz(dy);
    for (int p = 0; p < P; ++p)
        if (!n.valid[p]) dz[p] = 0.0;
    for (int l = 3; l >= 0; --l) {
        const int cin = kLayerCin[l], cout = kLayerCout[l], off = layer_offset(l);
        const std::vector<double>* src = l == 0 ? nullptr : &n.a[l - 1];
        // parameter gradients
        for (int py = 0; py < H; ++py)
            for (int px = 0; px < W; ++px)
                for (int co = 0; co < cout; ++co) {
                    const double g = dz[(size_t)(py * W + px) * cout + co];
                    if (g == 0.0) continue;
                    grad[off + cout * cin * 9 + co] += g;
                    for (int ci = 0; ci < cin; ++ci)
                        for (int ky = 0; ky < 3; ++ky)
                            for (int kx = 0; kx < 3; ++kx)
                                grad[off + ((co * cin + ci) * 3 + ky) * 3 + kx] +=
                                    g * in_val(n, l, src, px + kx - 1, py + ky - 1, ci, py * W + px);
                }
        if (l == 0) break;
        // input gradient, through the ReLU of the previous layer
        std::vector<double> da((size_t)P * cin, 0.0);
        for (int py = 0; py < H; ++py)
            for (int px = 0; px < W; ++px)
                for (int co = 0; co < cout; ++co) {
                    const double g = dz[(size_t)(py * W + px) * cout + co];
                    if (g == 0.0) continue;
                    for (int ky = 0; ky < 3; ++ky)
                        for (int kx = 0; kx < 3; ++kx) {
                            const int qx = px + kx - 1, qy = py + ky - 1;
                            if (qx < 0 || qy < 0 || qx >= W || qy >= H) continue;
                            for (int ci = 0; ci < cin; ++ci)
                                da[(size_t)(qy * W + qx) * cin + ci] += g * n.w[off + ((co * cin + ci) * 3 + ky) * 3 + kx];
                        }
                }
        const std::vector<double>& a = n.a[l - 1];
        for (size_t i = 0; i < da.size(); ++i) {
            da[i] = a[i] > 0.0 ? da[i] : 0.0;
            if (n.bf16) da[i] = bf16_round(da[i]);
        }
        dz.swap(da);
    }
    for (int i = 0; i < layer_offset(4); ++i) grad[i] /= cnt;
    return loss;
}
