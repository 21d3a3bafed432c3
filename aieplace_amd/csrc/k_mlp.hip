// k_mlp.hip -- the step-length MLP (16 -> 64 -> 64 -> 1, GELU): training
// (forward, backward, Adam), batch inference, and the in-loop proposal kernel
// that writes the predicted step into the optimiser's device scalars.
// No counterpart in the reference (fixed learning rate, AIEplace.cpp:1020-1040).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/eplace_mlp.h"
#include "ep_ctx.h"
#include "ep_mlp.h"

namespace ep {

namespace {

constexpr int MI = EP_MLP_IN, MH = EP_MLP_HID;
constexpr int OFF_W1 = 34, OFF_B1 = OFF_W1 + MH * MI, OFF_W2 = OFF_B1 + MH;
constexpr int OFF_B2 = OFF_W2 + MH * MH, OFF_W3 = OFF_B2 + MH, OFF_B3 = OFF_W3 + MH;
constexpr int NW = OFF_B3 + 1;
static_assert(NW == EP_MLP_NWEIGHTS, "weight blob layout");
constexpr int ACT = MI + 4 * MH;  // per-sample activations: xn, z1, a1, z2, a2

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_d(float z)
{
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * __expf(-0.5f * z * z) * 0.3989422804014327f;
}

// Forward, one thread per sample.  idx (nullable) maps minibatch rows to records.
// dout (nullable) = d loss / d output = 2 (o - t) / n for MSE.
__global__ void k_mlp_fwd(const float *__restrict__ w, const float *__restrict__ X,
                          const float *__restrict__ Y, const int32_t *__restrict__ idx, int n,
                          float *__restrict__ act, float *__restrict__ out,
                          float *__restrict__ dout, double *__restrict__ loss_acc)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int r = idx ? idx[s] : s;
    float *a = act + (size_t)s * ACT;
    float *xn = a, *z1 = a + MI, *a1 = a + MI + MH, *z2 = a + MI + 2 * MH, *a2 = a + MI + 3 * MH;
    for (int i = 0; i < MI; ++i) xn[i] = (X[(size_t)r * MI + i] - w[i]) / w[16 + i];
    for (int j = 0; j < MH; ++j) {
        float acc = w[OFF_B1 + j];
        for (int i = 0; i < MI; ++i) acc += w[OFF_W1 + j * MI + i] * xn[i];
        z1[j] = acc;
        a1[j] = gelu(acc);
    }
    for (int j = 0; j < MH; ++j) {
        float acc = w[OFF_B2 + j];
        for (int k = 0; k < MH; ++k) acc += w[OFF_W2 + j * MH + k] * a1[k];
        z2[j] = acc;
        a2[j] = gelu(acc);
    }
    float o = w[OFF_B3];
    for (int k = 0; k < MH; ++k) o += w[OFF_W3 + k] * a2[k];
    if (out) out[s] = o * w[33] + w[32];
    if (Y && dout) {
        const float t = (Y[r] - w[32]) / w[33];
        const float e = o - t;
        dout[s] = 2.0f * e / (float)n;
        atomicAdd(loss_acc, (double)e * e / n);
    }
}

// Backward deltas, one thread per sample: d2 = dL/dz2, d1 = dL/dz1.
__global__ void k_mlp_delta(const float *__restrict__ w, const float *__restrict__ act,
                            const float *__restrict__ dout, int n, float *__restrict__ d2,
                            float *__restrict__ d1)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float *a = act + (size_t)s * ACT;
    const float *z1 = a + MI, *z2 = a + MI + 2 * MH;
    float *e2 = d2 + (size_t)s * MH, *e1 = d1 + (size_t)s * MH;
    const float g = dout[s];
    for (int k = 0; k < MH; ++k) e2[k] = g * w[OFF_W3 + k] * gelu_d(z2[k]);
    for (int k = 0; k < MH; ++k) {
        float acc = 0.0f;
        for (int j = 0; j < MH; ++j) acc += e2[j] * w[OFF_W2 + j * MH + k];
        e1[k] = acc * gelu_d(z1[k]);
    }
}

// Weight gradient, one thread per parameter, summed over the minibatch in a
// fixed order (deterministic).
__global__ void k_mlp_wgrad(const float *__restrict__ act, const float *__restrict__ dout,
                            const float *__restrict__ d2, const float *__restrict__ d1, int n,
                            float *__restrict__ grad)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NW) return;
    float acc = 0.0f;
    if (p < OFF_W1) {
        acc = 0.0f;  // normalisation constants are not trained
    } else if (p < OFF_B1) {
        const int j = (p - OFF_W1) / MI, i = (p - OFF_W1) % MI;
        for (int s = 0; s < n; ++s) acc += d1[(size_t)s * MH + j] * act[(size_t)s * ACT + i];
    } else if (p < OFF_W2) {
        const int j = p - OFF_B1;
        for (int s = 0; s < n; ++s) acc += d1[(size_t)s * MH + j];
    } else if (p < OFF_B2) {
        const int j = (p - OFF_W2) / MH, k = (p - OFF_W2) % MH;
        for (int s = 0; s < n; ++s) acc += d2[(size_t)s * MH + j] * act[(size_t)s * ACT + MI + MH + k];
    } else if (p < OFF_W3) {
        const int j = p - OFF_B2;
        for (int s = 0; s < n; ++s) acc += d2[(size_t)s * MH + j];
    } else if (p < OFF_B3) {
        const int k = p - OFF_W3;
        for (int s = 0; s < n; ++s) acc += dout[s] * act[(size_t)s * ACT + MI + 3 * MH + k];
    } else {
        for (int s = 0; s < n; ++s) acc += dout[s];
    }
    grad[p] = acc;
}

__global__ void k_mlp_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                           float *__restrict__ v, int t, float lr, float b1, float b2, float eps)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < OFF_W1 || p >= NW) return;
    const float gi = g[p];
    const float mi = b1 * m[p] + (1.0f - b1) * gi;
    const float vi = b2 * v[p] + (1.0f - b2) * gi * gi;
    m[p] = mi;
    v[p] = vi;
    const float mh = mi / (1.0f - powf(b1, (float)t));
    const float vh = vi / (1.0f - powf(b2, (float)t));
    w[p] -= lr * mh / (sqrtf(vh) + eps);
}

// In-loop proposal (host-controlled steps): the model on the feature vector
// staged in the device scalars overwrites the trial step when use_model is set.
// Device-controlled runs make the same proposal inside k_finalize's decision.
__global__ __launch_bounds__(MH) void k_mlp_propose(const float *__restrict__ w,
                                                    DevScalars *__restrict__ sc)
{
    mlp_propose_block(w, sc->feat, sc, sc->use_model && !sc->halt);
}

struct Dev {
    float *w = nullptr, *X = nullptr, *Y = nullptr, *act = nullptr, *out = nullptr, *dout = nullptr;
    float *d1 = nullptr, *d2 = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    int32_t *idx = nullptr;
    double *loss = nullptr;
    ~Dev()
    {
        void *p[] = {w, X, Y, act, out, dout, d1, d2, g, m, v, idx, loss};
        for (void *q : p)
            if (q) (void)hipFree(q);
    }
};

template <class T>
ep_status dmalloc(T **p, size_t n)
{
    if (hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
        return fail(EP_ENOMEM, "mlp: hipMalloc failed");
    return EP_OK;
}

ep_status h2d(void *d, const void *h, size_t bytes)
{
    if (bytes && hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) != hipSuccess)
        return fail(EP_EDEVICE, "mlp: copy to device failed");
    return EP_OK;
}

ep_status d2h(void *h, const void *d, size_t bytes)
{
    if (bytes && hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(EP_EDEVICE, "mlp: copy to host failed");
    return EP_OK;
}

ep_status check_device()
{
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1) return fail(EP_EDEVICE, "mlp: no HIP device");
    return EP_OK;
}

ep_status check_blob(const float *w)
{
    for (int i = 0; i < 16; ++i)
        if (!(w[16 + i] > 0.0f)) return fail(EP_EINVAL, "mlp: feature std must be > 0");
    if (!(w[33] > 0.0f)) return fail(EP_EINVAL, "mlp: target std must be > 0");
    return EP_OK;
}

// forward (+ optional backward into d.g) of rows [0, n) of X/Y through idx
ep_status run_batch(Dev &d, const int32_t *idx, int n, bool backward)
{
    const int tb = 128, gb = (n + tb - 1) / tb;
    hipLaunchKernelGGL(k_mlp_fwd, dim3(gb), dim3(tb), 0, 0, d.w, d.X, backward ? d.Y : nullptr, idx, n,
                       d.act, d.out, backward ? d.dout : nullptr, d.loss);
    if (backward) {
        hipLaunchKernelGGL(k_mlp_delta, dim3(gb), dim3(tb), 0, 0, d.w, d.act, d.dout, n, d.d2, d.d1);
        hipLaunchKernelGGL(k_mlp_wgrad, dim3((NW + 255) / 256), dim3(256), 0, 0, d.act, d.dout, d.d2,
                           d.d1, n, d.g);
    }
    if (hipGetLastError() != hipSuccess) return fail(EP_EDEVICE, "mlp: kernel launch failed");
    return EP_OK;
}

uint64_t mix(uint64_t &s)
{
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double unif(uint64_t &s) { return (double)(mix(s) >> 11) * (1.0 / 9007199254740992.0); }
double normal(uint64_t &s)
{
    double u1 = std::max(unif(s), 1e-300), u2 = unif(s);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

}  // namespace

void launch_mlp_propose(Ctx &c, hipStream_t s)
{
    if (!c.d_mlp_w) return;
    hipLaunchKernelGGL(k_mlp_propose, dim3(1), dim3(MH), 0, s, c.d_mlp_w, c.d_scal);
}

}  // namespace ep

using namespace ep;

extern "C" void ep_mlp_default_train_params(ep_mlp_train_params *p)
{
    p->epochs = 40;
    p->batch = 512;
    p->lr = 2e-3f;
    p->beta1 = 0.9f;
    p->beta2 = 0.999f;
    p->eps = 1e-8f;
    p->seed = 1;
}

extern "C" ep_status ep_mlp_forward(const float *w, const float *X, int32_t n, float *out)
{
    if (!w || !X || !out || n < 0) return fail(EP_EINVAL, "ep_mlp_forward: bad argument");
    ep_status st;
    if ((st = check_blob(w)) != EP_OK || (st = check_device()) != EP_OK) return st;
    Dev d;
    if ((st = dmalloc(&d.w, NW)) || (st = dmalloc(&d.X, (size_t)n * MI)) ||
        (st = dmalloc(&d.act, (size_t)n * ACT)) || (st = dmalloc(&d.out, n)) || (st = dmalloc(&d.loss, 1)))
        return st;
    if ((st = h2d(d.w, w, NW * 4)) || (st = h2d(d.X, X, (size_t)n * MI * 4))) return st;
    if (n > 0 && (st = run_batch(d, nullptr, n, false)) != EP_OK) return st;
    return d2h(out, d.out, (size_t)n * 4);
}

extern "C" ep_status ep_mlp_grad(const float *w, const float *X, const float *Y, int32_t n, float *grad,
                                 float *loss)
{
    if (!w || !X || !Y || !grad || n < 1) return fail(EP_EINVAL, "ep_mlp_grad: bad argument");
    ep_status st;
    if ((st = check_blob(w)) != EP_OK || (st = check_device()) != EP_OK) return st;
    Dev d;
    if ((st = dmalloc(&d.w, NW)) || (st = dmalloc(&d.X, (size_t)n * MI)) || (st = dmalloc(&d.Y, n)) ||
        (st = dmalloc(&d.act, (size_t)n * ACT)) || (st = dmalloc(&d.out, n)) || (st = dmalloc(&d.dout, n)) ||
        (st = dmalloc(&d.d1, (size_t)n * MH)) || (st = dmalloc(&d.d2, (size_t)n * MH)) ||
        (st = dmalloc(&d.g, NW)) || (st = dmalloc(&d.loss, 1)))
        return st;
    if ((st = h2d(d.w, w, NW * 4)) || (st = h2d(d.X, X, (size_t)n * MI * 4)) || (st = h2d(d.Y, Y, (size_t)n * 4)))
        return st;
    if (hipMemset(d.loss, 0, sizeof(double)) != hipSuccess) return fail(EP_EDEVICE, "mlp: memset failed");
    if ((st = run_batch(d, nullptr, n, true)) != EP_OK) return st;
    double l = 0;
    if ((st = d2h(grad, d.g, NW * 4)) || (st = d2h(&l, d.loss, sizeof(double)))) return st;
    if (loss) *loss = (float)l;
    return EP_OK;
}

extern "C" ep_status ep_mlp_train(const float *X, const float *Y, int32_t n, const ep_mlp_train_params *p,
                                  float *w_out, float *loss_curve)
{
    if (!X || !Y || !p || !w_out || n < 2 || p->epochs < 1 || p->batch < 1)
        return fail(EP_EINVAL, "ep_mlp_train: bad argument");
    ep_status st;
    if ((st = check_device()) != EP_OK) return st;
    // normalisation from the records (FP64 on the host)
    std::vector<float> w(NW, 0.0f);
    for (int i = 0; i < MI; ++i) {
        double s = 0, s2 = 0;
        for (int r = 0; r < n; ++r) {
            const double v = X[(size_t)r * MI + i];
            s += v;
            s2 += v * v;
        }
        const double mean = s / n, var = std::max(0.0, s2 / n - mean * mean);
        w[i] = (float)mean;
        w[16 + i] = (float)std::max(std::sqrt(var), 1e-6);
    }
    {
        double s = 0, s2 = 0;
        for (int r = 0; r < n; ++r) {
            s += Y[r];
            s2 += (double)Y[r] * Y[r];
        }
        const double mean = s / n, var = std::max(0.0, s2 / n - mean * mean);
        w[32] = (float)mean;
        w[33] = (float)std::max(std::sqrt(var), 1e-6);
    }
    // He-style initialisation, zero biases
    uint64_t rs = p->seed;
    for (int k = 0; k < MH * MI; ++k) w[OFF_W1 + k] = (float)(normal(rs) * std::sqrt(2.0 / MI));
    for (int k = 0; k < MH * MH; ++k) w[OFF_W2 + k] = (float)(normal(rs) * std::sqrt(2.0 / MH));
    for (int k = 0; k < MH; ++k) w[OFF_W3 + k] = (float)(normal(rs) * std::sqrt(1.0 / MH));
    const int B = std::min(p->batch, n);
    Dev d;
    if ((st = dmalloc(&d.w, NW)) || (st = dmalloc(&d.X, (size_t)n * MI)) || (st = dmalloc(&d.Y, n)) ||
        (st = dmalloc(&d.act, (size_t)B * ACT)) || (st = dmalloc(&d.out, B)) || (st = dmalloc(&d.dout, B)) ||
        (st = dmalloc(&d.d1, (size_t)B * MH)) || (st = dmalloc(&d.d2, (size_t)B * MH)) ||
        (st = dmalloc(&d.g, NW)) || (st = dmalloc(&d.m, NW)) || (st = dmalloc(&d.v, NW)) ||
        (st = dmalloc(&d.idx, n)) || (st = dmalloc(&d.loss, 1)))
        return st;
    if ((st = h2d(d.w, w.data(), NW * 4)) || (st = h2d(d.X, X, (size_t)n * MI * 4)) ||
        (st = h2d(d.Y, Y, (size_t)n * 4)))
        return st;
    if (hipMemset(d.m, 0, NW * 4) != hipSuccess || hipMemset(d.v, 0, NW * 4) != hipSuccess)
        return fail(EP_EDEVICE, "mlp: memset failed");
    std::vector<int32_t> perm(n);
    for (int r = 0; r < n; ++r) perm[r] = r;
    int t = 0;
    for (int e = 0; e < p->epochs; ++e) {
        for (int r = n - 1; r > 0; --r) std::swap(perm[r], perm[(int)(unif(rs) * (r + 1))]);
        if ((st = h2d(d.idx, perm.data(), (size_t)n * 4)) != EP_OK) return st;
        if (hipMemset(d.loss, 0, sizeof(double)) != hipSuccess) return fail(EP_EDEVICE, "mlp: memset failed");
        int nb = 0;
        for (int s0 = 0; s0 < n; s0 += B) {
            const int bn = std::min(B, n - s0);
            if ((st = run_batch(d, d.idx + s0, bn, true)) != EP_OK) return st;
            ++t;
            hipLaunchKernelGGL(k_mlp_adam, dim3((NW + 255) / 256), dim3(256), 0, 0, d.w, d.g, d.m, d.v, t,
                               p->lr, p->beta1, p->beta2, p->eps);
            ++nb;
        }
        if (loss_curve) {
            double l = 0;
            if ((st = d2h(&l, d.loss, sizeof(double))) != EP_OK) return st;
            loss_curve[e] = (float)(l / nb);
        }
    }
    if (hipDeviceSynchronize() != hipSuccess) return fail(EP_EDEVICE, "ep_mlp_train: device error");
    return d2h(w_out, d.w, NW * 4);
}
