// Training and stand-alone inference of the troubled-face network (include/fvg.h, fvg_trainer_* and fvg_mlp_logits).
//
// Training runs in float on the vector units: a block takes 128 samples, each thread one sample forward and backward
// (weighted binary cross-entropy with logits), keeps the per-sample activations and back-propagated errors in LDS,
// then the block's threads sum each gradient entry over the block's samples in sample order. Blocks write partial
// gradients that one more kernel adds up in block order (in double), so a gradient is the same bits on every run.
// Adam (Kingma and Ba 2015) updates the float parameters. Stand-alone inference runs the same wavefront routine as
// the face kernels (bfloat16 matrix instructions).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <random>
#include <string>
#include <vector>

#include "fvg_host.hpp"

namespace fvg {

constexpr int TB = 128;        // samples per training block
constexpr int NPG = MLP_NPARAM + 1;   // partial gradient row: parameters, then the loss

__device__ __forceinline__ float softplus_neg_abs(float z) { return log1pf(expf(-fabsf(z))); }

// Partial gradients of one block of samples idx[s0 .. s0 + TB) (all n when idx is null).
__global__ __launch_bounds__(TB) void k_mlp_grad(const float* __restrict__ X, const float* __restrict__ y, const int* __restrict__ idx,
                                                 long long s0, long long count, long long n_norm, const float* __restrict__ P,
                                                 float pos_w, float* __restrict__ part) {
    __shared__ float sw[MLP_NPARAM];
    __shared__ float sx[TB][16];
    __shared__ float sh1[TB][33], sg1[TB][33], sg2[TB][33], sh2[TB][33];
    __shared__ float sgz[TB], sloss[TB];
    for (int e = threadIdx.x; e < MLP_NPARAM; e += TB) sw[e] = P[e];
    __syncthreads();
    const float* W1 = sw;
    const float* b1 = sw + 480;
    const float* W2 = sw + 512;
    const float* b2 = sw + 1536;
    const float* w3 = sw + 1568;
    const float b3 = sw[1600];
    const int tid = threadIdx.x;
    const long long s = (long long)blockIdx.x * TB + tid;
    const bool valid = s < count;
    float x[15];
    float yy = 0.0f;
    if (valid) {
        const long long r = idx ? (long long)idx[s0 + s] : s0 + s;
#pragma unroll
        for (int k = 0; k < 15; ++k) x[k] = X[r * 16 + k];
        yy = y[r];
    } else {
#pragma unroll
        for (int k = 0; k < 15; ++k) x[k] = 0.0f;
    }
    float z1[32], h1[32], z2[32];
#pragma unroll
    for (int h = 0; h < 32; ++h) {
        float a = b1[h];
#pragma unroll
        for (int k = 0; k < 15; ++k) a += W1[h * 15 + k] * x[k];
        z1[h] = a;
        h1[h] = fmaxf(a, 0.0f);
    }
    float z = b3;
#pragma unroll
    for (int h = 0; h < 32; ++h) {
        float a = b2[h];
#pragma unroll
        for (int k = 0; k < 32; ++k) a += W2[h * 32 + k] * h1[k];
        z2[h] = a;
        z += w3[h] * fmaxf(a, 0.0f);
    }
    const float w = yy > 0.5f ? pos_w : 1.0f;
    const float loss = valid ? w * (fmaxf(z, 0.0f) - z * yy + softplus_neg_abs(z)) : 0.0f;
    const float gz = valid ? w * (1.0f / (1.0f + expf(-z)) - yy) / (float)n_norm : 0.0f;
    float g2[32];
#pragma unroll
    for (int h = 0; h < 32; ++h) g2[h] = z2[h] > 0.0f ? gz * w3[h] : 0.0f;
#pragma unroll
    for (int k = 0; k < 32; ++k) {
        float a = 0.0f;
#pragma unroll
        for (int h = 0; h < 32; ++h) a += g2[h] * W2[h * 32 + k];
        sg1[tid][k] = z1[k] > 0.0f ? a : 0.0f;
        sh1[tid][k] = h1[k];
        sg2[tid][k] = g2[k];
        sh2[tid][k] = fmaxf(z2[k], 0.0f);
    }
#pragma unroll
    for (int k = 0; k < 15; ++k) sx[tid][k] = x[k];
    sgz[tid] = gz;
    sloss[tid] = loss;
    __syncthreads();
    float* out = part + (long long)blockIdx.x * NPG;
    for (int e = tid; e < NPG; e += TB) {
        float a = 0.0f;
        if (e < 480) {
            const int h = e / 15, k = e % 15;
            for (int q = 0; q < TB; ++q) a += sg1[q][h] * sx[q][k];
        } else if (e < 512) {
            const int h = e - 480;
            for (int q = 0; q < TB; ++q) a += sg1[q][h];
        } else if (e < 1536) {
            const int h = (e - 512) / 32, k = (e - 512) % 32;
            for (int q = 0; q < TB; ++q) a += sg2[q][h] * sh1[q][k];
        } else if (e < 1568) {
            const int h = e - 1536;
            for (int q = 0; q < TB; ++q) a += sg2[q][h];
        } else if (e < 1600) {
            const int h = e - 1568;
            for (int q = 0; q < TB; ++q) a += sgz[q] * sh2[q][h];
        } else if (e == 1600) {
            for (int q = 0; q < TB; ++q) a += sgz[q];
        } else {
            for (int q = 0; q < TB; ++q) a += sloss[q];
        }
        out[e] = a;
    }
}

// Sum of the blocks' partial gradients in block order: grad[e] (e < NPARAM) and the summed loss at grad[NPARAM].
__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ part, int nb, float* __restrict__ grad) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NPG) return;
    double a = 0.0;
    for (int b = 0; b < nb; ++b) a += (double)part[(long long)b * NPG + e];
    grad[e] = (float)a;
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ P, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                                              float lr, float b1, float b2, float c1, float c2, float eps) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= MLP_NPARAM) return;
    const float gg = g[e];
    const float mm = b1 * m[e] + (1.0f - b1) * gg;
    const float vv = b2 * v[e] + (1.0f - b2) * gg * gg;
    m[e] = mm;
    v[e] = vv;
    P[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// Logits of n feature rows (16 floats each, the 16th ignored) through the wavefront inference routine.
__global__ __launch_bounds__(256) void k_mlp_logits(const MlpDev* __restrict__ m, const float* __restrict__ X, long long n, float* __restrict__ z) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float x[16];
#pragma unroll
    for (int k = 0; k < 15; ++k) x[k] = s < n ? X[s * 16 + k] : 0.0f;
    x[15] = 1.0f;
    const float r = mlp_logit_wave(m, x);
    if (s < n) z[s] = r;
}

}   // namespace fvg

using namespace fvg;

namespace {
std::string g_trainer_error;
}

struct fvg_trainer {
    int device = 0;
    hipStream_t stream = nullptr;
    long long cap = 0, n = 0;
    float *X = nullptr, *y = nullptr, *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr, *part = nullptr;
    int* idx = nullptr;
    int max_blocks = 0;
    std::string err;
};

namespace {

int tset(fvg_trainer* t, int code, const std::string& m) {
    if (t) t->err = m;
    else g_trainer_error = m;
    return code;
}

#define TCHK(t, call)                                                                                      \
    do {                                                                                                   \
        const hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess) return tset(t, FVG_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

void free_trainer(fvg_trainer* t) {
    if (!t) return;
    hipFree(t->X);
    hipFree(t->y);
    hipFree(t->P);
    hipFree(t->G);
    hipFree(t->M);
    hipFree(t->V);
    hipFree(t->part);
    hipFree(t->idx);
    if (t->stream) hipStreamDestroy(t->stream);
    delete t;
}

// gradient (and summed loss) of the samples idx[s0 .. s0 + count) into t->G; the loss is normalised by count
int enqueue_grad(fvg_trainer* t, const int* idx, long long s0, long long count, float pos_w) {
    const int nb = (int)((count + TB - 1) / TB);
    if (nb > t->max_blocks) return tset(t, FVG_EINVAL, "batch larger than the trainer's capacity");
    hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(TB), 0, t->stream, t->X, t->y, idx, s0, count, count, t->P, pos_w, t->part);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((NPG + 255) / 256), dim3(256), 0, t->stream, t->part, nb, t->G);
    TCHK(t, hipGetLastError());
    return FVG_OK;
}

}   // namespace

extern "C" {

const char* fvg_trainer_last_error(const fvg_trainer* t) { return t ? t->err.c_str() : g_trainer_error.c_str(); }

int fvg_trainer_create(int device, long long capacity, fvg_trainer** out) {
    if (!out) return tset(nullptr, FVG_EINVAL, "null argument");
    *out = nullptr;
    if (capacity <= 0 || capacity > (1ll << 30)) return tset(nullptr, FVG_EINVAL, "capacity out of range");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return tset(nullptr, FVG_EHIP, "no HIP device");
    if (device < 0 || device >= ndev) return tset(nullptr, FVG_EINVAL, "device index out of range");
    fvg_trainer* t = new fvg_trainer();
    t->device = device;
    t->cap = capacity;
    t->max_blocks = (int)((capacity + TB - 1) / TB);
    auto fail = [&](int code, const std::string& m) {
        free_trainer(t);
        return tset(nullptr, code, m);
    };
    if (hipSetDevice(device) != hipSuccess) return fail(FVG_EHIP, "hipSetDevice failed");
    if (hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess) return fail(FVG_EHIP, "hipStreamCreate failed");
    if (hipMalloc(&t->X, sizeof(float) * 16 * capacity) != hipSuccess || hipMalloc(&t->y, sizeof(float) * capacity) != hipSuccess ||
        hipMalloc(&t->idx, sizeof(int) * capacity) != hipSuccess || hipMalloc(&t->P, sizeof(float) * MLP_NPARAM) != hipSuccess ||
        hipMalloc(&t->G, sizeof(float) * NPG) != hipSuccess || hipMalloc(&t->M, sizeof(float) * MLP_NPARAM) != hipSuccess ||
        hipMalloc(&t->V, sizeof(float) * MLP_NPARAM) != hipSuccess ||
        hipMalloc(&t->part, sizeof(float) * NPG * (size_t)t->max_blocks) != hipSuccess)
        return fail(FVG_ENOMEM, "hipMalloc of the trainer buffers");
    *out = t;
    return FVG_OK;
}

int fvg_trainer_destroy(fvg_trainer* t) {
    free_trainer(t);
    return FVG_OK;
}

int fvg_trainer_set_data(fvg_trainer* t, const float* X, const float* y, long long n) {
    if (!t || !X || !y) return tset(t, FVG_EINVAL, "null argument");
    if (n <= 0 || n > t->cap) return tset(t, FVG_EINVAL, "sample count out of range");
    TCHK(t, hipSetDevice(t->device));
    TCHK(t, hipMemcpy(t->X, X, sizeof(float) * 16 * n, hipMemcpyHostToDevice));
    TCHK(t, hipMemcpy(t->y, y, sizeof(float) * n, hipMemcpyHostToDevice));
    t->n = n;
    return FVG_OK;
}

int fvg_trainer_grad(fvg_trainer* t, const float* params, float pos_weight, float* grad, double* loss) {
    if (!t || !params) return tset(t, FVG_EINVAL, "null argument");
    if (t->n <= 0) return tset(t, FVG_ELOGIC, "no data set");
    TCHK(t, hipSetDevice(t->device));
    TCHK(t, hipMemcpyAsync(t->P, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, t->stream));
    int rc = enqueue_grad(t, nullptr, 0, t->n, pos_weight);
    if (rc) return rc;
    std::vector<float> g(NPG);
    TCHK(t, hipMemcpyAsync(g.data(), t->G, sizeof(float) * NPG, hipMemcpyDeviceToHost, t->stream));
    TCHK(t, hipStreamSynchronize(t->stream));
    if (grad) std::memcpy(grad, g.data(), sizeof(float) * MLP_NPARAM);
    if (loss) *loss = (double)g[MLP_NPARAM] / (double)t->n;
    return FVG_OK;
}

int fvg_trainer_fit(fvg_trainer* t, float* params, int epochs, int batch, float lr, float pos_weight, unsigned long long seed,
                    double* loss_out) {
    if (!t || !params) return tset(t, FVG_EINVAL, "null argument");
    if (t->n <= 0) return tset(t, FVG_ELOGIC, "no data set");
    if (epochs <= 0 || batch <= 0 || !(lr > 0.0f)) return tset(t, FVG_EINVAL, "epochs, batch and lr must be positive");
    TCHK(t, hipSetDevice(t->device));
    TCHK(t, hipMemcpy(t->P, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
    TCHK(t, hipMemset(t->M, 0, sizeof(float) * MLP_NPARAM));
    TCHK(t, hipMemset(t->V, 0, sizeof(float) * MLP_NPARAM));
    std::vector<int> order((size_t)t->n);
    for (long long i = 0; i < t->n; ++i) order[i] = (int)i;
    std::mt19937_64 rng(seed);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    long long step = 0;
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(order.begin(), order.end(), rng);
        TCHK(t, hipMemcpyAsync(t->idx, order.data(), sizeof(int) * t->n, hipMemcpyHostToDevice, t->stream));
        for (long long s0 = 0; s0 < t->n; s0 += batch) {
            const long long cnt = std::min<long long>(batch, t->n - s0);
            int rc = enqueue_grad(t, t->idx, s0, cnt, pos_weight);
            if (rc) return rc;
            ++step;
            const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
            hipLaunchKernelGGL(k_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, t->stream, t->P, t->G, t->M, t->V, lr, b1, b2, c1, c2, eps);
            TCHK(t, hipGetLastError());
        }
        TCHK(t, hipStreamSynchronize(t->stream));   // the host order vector is reused by the next epoch
    }
    TCHK(t, hipMemcpy(params, t->P, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost));
    if (loss_out) {
        std::vector<float> g(NPG);
        int rc = enqueue_grad(t, nullptr, 0, t->n, pos_weight);
        if (rc) return rc;
        TCHK(t, hipMemcpy(g.data(), t->G, sizeof(float) * NPG, hipMemcpyDeviceToHost));
        *loss_out = (double)g[MLP_NPARAM] / (double)t->n;
    }
    return FVG_OK;
}

int fvg_mlp_logits(int device, const float* params, float threshold, const float* X, long long n, float* z) {
    if (!params || !X || !z) return tset(nullptr, FVG_EINVAL, "null argument");
    if (n <= 0) return tset(nullptr, FVG_EINVAL, "n must be positive");
    if (hipSetDevice(device) != hipSuccess) return tset(nullptr, FVG_EHIP, "hipSetDevice failed");
    const MlpDev m = pack_mlp(params, threshold);
    MlpDev* dm = nullptr;
    float *dX = nullptr, *dz = nullptr;
    int rc = FVG_OK;
    if (hipMalloc(&dm, sizeof m) != hipSuccess || hipMalloc(&dX, sizeof(float) * 16 * n) != hipSuccess ||
        hipMalloc(&dz, sizeof(float) * n) != hipSuccess)
        rc = tset(nullptr, FVG_ENOMEM, "hipMalloc of the inference buffers");
    if (!rc && (hipMemcpy(dm, &m, sizeof m, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(dX, X, sizeof(float) * 16 * n, hipMemcpyHostToDevice) != hipSuccess))
        rc = tset(nullptr, FVG_EHIP, "hipMemcpy of the inference inputs");
    if (!rc) {
        hipLaunchKernelGGL(k_mlp_logits, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, dm, dX, n, dz);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = tset(nullptr, FVG_EHIP, "inference kernel failed");
    }
    if (!rc && hipMemcpy(z, dz, sizeof(float) * n, hipMemcpyDeviceToHost) != hipSuccess) rc = tset(nullptr, FVG_EHIP, "hipMemcpy of the logits");
    hipFree(dm);
    hipFree(dX);
    hipFree(dz);
    return rc;
}

}   // extern "C"
