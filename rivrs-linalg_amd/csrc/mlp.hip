// Per-front update-precision classifier: a 16 -> 64 -> 64 -> 3 multilayer perceptron (ReLU
// hidden layers, softmax output over {BF16, FP32, FP64}), FP32, with its training on the GPU.
// The reference has no counterpart; the solver uses it to choose each front's Schur-update
// arithmetic from front features, with full FP64 as the fallback.
//
// Parameters, one FP32 array of RV_MLP_PARAMS entries: W1[16][64], b1[64], W2[64][64], b2[64],
// W3[64][3], b3[3] (row = input unit).  Training is full-batch Adam on the class-weighted mean
// cross-entropy L = sum_n w[y_n] (-log p[n][y_n]) / sum_n w[y_n].  Every reduction runs in a
// fixed order (one thread per parameter summing over samples), so training is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "rv_mlp.h"

namespace rv {

namespace {

constexpr int NI = MLP_IN, NH = MLP_HID, NO = MLP_OUT;
constexpr int O_W1 = 0, O_B1 = O_W1 + NI * NH, O_W2 = O_B1 + NH, O_B2 = O_W2 + NH * NH, O_W3 = O_B2 + NH,
              O_B3 = O_W3 + NH * NO, NPAR = O_B3 + NO;
static_assert(NPAR == MLP_PARAMS, "parameter layout");

#define MLP_HIP(x)                                                                        \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP: ") + hipGetErrorString(e_)); \
    } while (0)

// Forward pass, one thread per sample; parameters staged in LDS.  Stores the hidden
// activations and the class probabilities for the backward pass.
__global__ __launch_bounds__(256) void k_mlp_forward(const float *__restrict__ par, int n,
                                                     const float *__restrict__ x, float *__restrict__ h1,
                                                     float *__restrict__ h2, float *__restrict__ prob) {
    __shared__ float P[NPAR];
    for (int i = threadIdx.x; i < NPAR; i += blockDim.x) P[i] = par[i];
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float xi[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) xi[i] = x[(size_t)s * NI + i];
    float a[NH];
#pragma unroll
    for (int j = 0; j < NH; ++j) {
        float acc = P[O_B1 + j];
#pragma unroll
        for (int i = 0; i < NI; ++i) acc += xi[i] * P[O_W1 + i * NH + j];
        a[j] = fmaxf(acc, 0.0f);
        if (h1) h1[(size_t)s * NH + j] = a[j];
    }
    // second hidden layer streamed into the output logits (its units are not kept in registers)
    float z[NO];
#pragma unroll
    for (int c = 0; c < NO; ++c) z[c] = P[O_B3 + c];
#pragma unroll
    for (int j = 0; j < NH; ++j) {
        float acc = P[O_B2 + j];
#pragma unroll
        for (int i = 0; i < NH; ++i) acc += a[i] * P[O_W2 + i * NH + j];
        const float bj = fmaxf(acc, 0.0f);
        if (h2) h2[(size_t)s * NH + j] = bj;
#pragma unroll
        for (int c = 0; c < NO; ++c) z[c] += bj * P[O_W3 + j * NO + c];
    }
    float zm = z[0];
#pragma unroll
    for (int c = 1; c < NO; ++c) zm = fmaxf(zm, z[c]);
    float e[NO], se = 0.0f;
#pragma unroll
    for (int c = 0; c < NO; ++c) { e[c] = expf(z[c] - zm); se += e[c]; }
#pragma unroll
    for (int c = 0; c < NO; ++c) prob[(size_t)s * NO + c] = e[c] / se;
}

// Backward pass to the layer inputs, one thread per sample: output delta w_y (p - onehot) / W,
// then through W3 and W2 with the ReLU masks.  Per-sample loss terms for the fixed-order sum.
__global__ __launch_bounds__(256) void k_mlp_delta(const float *__restrict__ par, int n,
                                                   const int *__restrict__ y, const float *__restrict__ cw,
                                                   float inv_wsum, const float *__restrict__ h1,
                                                   const float *__restrict__ h2, const float *__restrict__ prob,
                                                   float *__restrict__ d1, float *__restrict__ d2,
                                                   float *__restrict__ d3, float *__restrict__ lossn) {
    __shared__ float P[NPAR];
    for (int i = threadIdx.x; i < NPAR; i += blockDim.x) P[i] = par[i];
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int ys = y[s];
    const float w = cw[ys] * inv_wsum;
    float g3[NO];
#pragma unroll
    for (int c = 0; c < NO; ++c) {
        const float p = prob[(size_t)s * NO + c];
        g3[c] = w * (p - (c == ys ? 1.0f : 0.0f));
        d3[(size_t)s * NO + c] = g3[c];
    }
    lossn[s] = -w * logf(fmaxf(prob[(size_t)s * NO + ys], 1e-30f));
    float g2[NH];
#pragma unroll
    for (int j = 0; j < NH; ++j) {
        float acc = 0.0f;
#pragma unroll
        for (int c = 0; c < NO; ++c) acc += g3[c] * P[O_W3 + j * NO + c];
        g2[j] = h2[(size_t)s * NH + j] > 0.0f ? acc : 0.0f;
        d2[(size_t)s * NH + j] = g2[j];
    }
#pragma unroll
    for (int i = 0; i < NH; ++i) {
        float acc = 0.0f;
#pragma unroll
        for (int j = 0; j < NH; ++j) acc += g2[j] * P[O_W2 + i * NH + j];
        d1[(size_t)s * NH + i] = h1[(size_t)s * NH + i] > 0.0f ? acc : 0.0f;
    }
}

// Gradient of every parameter: one thread per parameter summing its outer-product terms over
// the samples in order; the last thread also sums the loss terms.
__global__ __launch_bounds__(256) void k_mlp_grad(int n, const float *__restrict__ x,
                                                  const float *__restrict__ h1, const float *__restrict__ h2,
                                                  const float *__restrict__ d1, const float *__restrict__ d2,
                                                  const float *__restrict__ d3, const float *__restrict__ lossn,
                                                  float *__restrict__ grad, float *__restrict__ loss) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q == NPAR) {
        float acc = 0.0f;
        for (int s = 0; s < n; ++s) acc += lossn[s];
        *loss = acc;
        return;
    }
    if (q > NPAR) return;
    const float *in;
    const float *dl;
    int ld_in, ld_d, i, j;
    if (q < O_B1) { i = q / NH; j = q % NH; in = x; ld_in = NI; dl = d1; ld_d = NH; }
    else if (q < O_W2) { i = -1; j = q - O_B1; in = nullptr; ld_in = 0; dl = d1; ld_d = NH; }
    else if (q < O_B2) { i = (q - O_W2) / NH; j = (q - O_W2) % NH; in = h1; ld_in = NH; dl = d2; ld_d = NH; }
    else if (q < O_W3) { i = -1; j = q - O_B2; in = nullptr; ld_in = 0; dl = d2; ld_d = NH; }
    else if (q < O_B3) { i = (q - O_W3) / NO; j = (q - O_W3) % NO; in = h2; ld_in = NH; dl = d3; ld_d = NO; }
    else { i = -1; j = q - O_B3; in = nullptr; ld_in = 0; dl = d3; ld_d = NO; }
    float acc = 0.0f;
    if (in)
        for (int s = 0; s < n; ++s) acc += in[(size_t)s * ld_in + i] * dl[(size_t)s * ld_d + j];
    else
        for (int s = 0; s < n; ++s) acc += dl[(size_t)s * ld_d + j];
    grad[q] = acc;
}

// Adam (Kingma & Ba 2015) with bias correction.
__global__ __launch_bounds__(256) void k_mlp_adam(float *__restrict__ par, const float *__restrict__ grad,
                                                  float *__restrict__ m1, float *__restrict__ m2, float lr,
                                                  float c1, float c2) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= NPAR) return;
    const float g = grad[q];
    const float a = 0.9f * m1[q] + 0.1f * g;
    const float b = 0.999f * m2[q] + 0.001f * g * g;
    m1[q] = a;
    m2[q] = b;
    par[q] -= lr * (a / c1) / (sqrtf(b / c2) + 1e-8f);
}

template <typename T>
struct DevBuf {
    T *p = nullptr;
    explicit DevBuf(size_t n) { MLP_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T))); }
    ~DevBuf() { if (p) (void)hipFree(p); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
};

struct Workspace {
    int n;
    DevBuf<float> par, x, h1, h2, prob, d1, d2, d3, lossn, grad, loss, m1, m2, cw;
    DevBuf<int> y;
    explicit Workspace(int n_)
        : n(n_), par(NPAR), x((size_t)n_ * NI), h1((size_t)n_ * NH), h2((size_t)n_ * NH), prob((size_t)n_ * NO),
          d1((size_t)n_ * NH), d2((size_t)n_ * NH), d3((size_t)n_ * NO), lossn(n_), grad(NPAR), loss(1),
          m1(NPAR), m2(NPAR), cw(NO), y(n_) {}
};

float weight_sum(int n, const int *y, const float *cw) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        if (y[i] < 0 || y[i] >= NO) throw std::invalid_argument("label out of range");
        s += cw ? cw[y[i]] : 1.0f;
    }
    if (!(s > 0.0)) throw std::invalid_argument("class weights sum to zero");
    return (float)(1.0 / s);
}

void upload(Workspace &w, const float *par, const float *x, const int *y, const float *cw) {
    MLP_HIP(hipMemcpy(w.par.p, par, NPAR * sizeof(float), hipMemcpyHostToDevice));
    MLP_HIP(hipMemcpy(w.x.p, x, (size_t)w.n * NI * sizeof(float), hipMemcpyHostToDevice));
    if (y) {
        MLP_HIP(hipMemcpy(w.y.p, y, (size_t)w.n * sizeof(int), hipMemcpyHostToDevice));
        float ones[NO];
        for (float &o : ones) o = 1.0f;
        MLP_HIP(hipMemcpy(w.cw.p, cw ? cw : ones, NO * sizeof(float), hipMemcpyHostToDevice));
    }
}

void grad_step(Workspace &w, float inv_wsum) {
    const dim3 bs(256), gs((unsigned)((w.n + 255) / 256)), gp((unsigned)((NPAR + 1 + 255) / 256));
    hipLaunchKernelGGL(k_mlp_forward, gs, bs, 0, 0, w.par.p, w.n, w.x.p, w.h1.p, w.h2.p, w.prob.p);
    hipLaunchKernelGGL(k_mlp_delta, gs, bs, 0, 0, w.par.p, w.n, w.y.p, w.cw.p, inv_wsum, w.h1.p, w.h2.p,
                       w.prob.p, w.d1.p, w.d2.p, w.d3.p, w.lossn.p);
    hipLaunchKernelGGL(k_mlp_grad, gp, bs, 0, 0, w.n, w.x.p, w.h1.p, w.h2.p, w.d1.p, w.d2.p, w.d3.p,
                       w.lossn.p, w.grad.p, w.loss.p);
    MLP_HIP(hipGetLastError());
}

}  // namespace

// Inference buffers (parameters, features, probabilities) and a non-blocking stream (default
// priority) per device, made on first use and kept until the process ends: a forward pass per
// factorization allocates nothing after the first and does not wait for other streams' work (a
// solve running on the device meanwhile).
struct FwdBufs {
    float *par = nullptr, *x = nullptr, *prob = nullptr;
    size_t cap = 0;
    hipStream_t st = nullptr;
};

void mlp_forward(const float *par, int n, const float *x, float *prob) {
    if (n <= 0) return;
    static std::mutex mu;
    static std::map<int, FwdBufs> bufs;
    std::lock_guard<std::mutex> lock(mu);
    int dev = 0;
    MLP_HIP(hipGetDevice(&dev));
    FwdBufs &b = bufs[dev];
    if (!b.st) MLP_HIP(hipStreamCreateWithFlags(&b.st, hipStreamNonBlocking));
    if (!b.par) MLP_HIP(hipMalloc(&b.par, NPAR * sizeof(float)));
    if ((size_t)n > b.cap) {
        if (b.x) (void)hipFree(b.x);
        if (b.prob) (void)hipFree(b.prob);
        b.x = b.prob = nullptr;
        b.cap = 0;
        const size_t cap = std::max<size_t>((size_t)n, 4096);
        MLP_HIP(hipMalloc(&b.x, cap * NI * sizeof(float)));
        MLP_HIP(hipMalloc(&b.prob, cap * NO * sizeof(float)));
        b.cap = cap;
    }
    MLP_HIP(hipMemcpyAsync(b.par, par, NPAR * sizeof(float), hipMemcpyHostToDevice, b.st));
    MLP_HIP(hipMemcpyAsync(b.x, x, (size_t)n * NI * sizeof(float), hipMemcpyHostToDevice, b.st));
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b.st, b.par, n, b.x,
                       (float *)nullptr, (float *)nullptr, b.prob);
    MLP_HIP(hipGetLastError());
    MLP_HIP(hipMemcpyAsync(prob, b.prob, (size_t)n * NO * sizeof(float), hipMemcpyDeviceToHost, b.st));
    MLP_HIP(hipStreamSynchronize(b.st));
}

void mlp_loss_grad(const float *par, int n, const float *x, const int *y, const float *cw, float *loss,
                   float *grad) {
    if (n <= 0) throw std::invalid_argument("no samples");
    const float iw = weight_sum(n, y, cw);
    Workspace w(n);
    upload(w, par, x, y, cw);
    grad_step(w, iw);
    MLP_HIP(hipMemcpy(grad, w.grad.p, NPAR * sizeof(float), hipMemcpyDeviceToHost));
    MLP_HIP(hipMemcpy(loss, w.loss.p, sizeof(float), hipMemcpyDeviceToHost));
}

void mlp_train(float *par, int n, const float *x, const int *y, const float *cw, int epochs, float lr,
               float *final_loss) {
    if (n <= 0) throw std::invalid_argument("no samples");
    if (epochs < 0 || !(lr > 0.0f)) throw std::invalid_argument("epochs >= 0 and lr > 0 required");
    const float iw = weight_sum(n, y, cw);
    Workspace w(n);
    upload(w, par, x, y, cw);
    MLP_HIP(hipMemset(w.m1.p, 0, NPAR * sizeof(float)));
    MLP_HIP(hipMemset(w.m2.p, 0, NPAR * sizeof(float)));
    double c1 = 1.0, c2 = 1.0;
    for (int e = 0; e < epochs; ++e) {
        grad_step(w, iw);
        c1 *= 0.9;
        c2 *= 0.999;
        hipLaunchKernelGGL(k_mlp_adam, dim3((NPAR + 255) / 256), dim3(256), 0, 0, w.par.p, w.grad.p, w.m1.p,
                           w.m2.p, lr, (float)(1.0 - c1), (float)(1.0 - c2));
        MLP_HIP(hipGetLastError());
    }
    // loss of the final parameters
    grad_step(w, iw);
    MLP_HIP(hipMemcpy(par, w.par.p, NPAR * sizeof(float), hipMemcpyDeviceToHost));
    if (final_loss) MLP_HIP(hipMemcpy(final_loss, w.loss.p, sizeof(float), hipMemcpyDeviceToHost));
}

}  // namespace rv
