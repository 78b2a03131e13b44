// mlp.hip -- the selector's 3-layer MLP (in -> H -> H -> out, ReLU, FP32)
// trained and evaluated on the GPU by hand-written kernels: fused
// linear+bias+ReLU forward, input- and weight-gradient kernels, a masked MSE
// loss with a fixed-order (deterministic) reduction, and Adam.
//
// Parameters are one flat FP32 array: W1 [H x in], b1 [H], W2 [H x H], b2 [H],
// W3 [out x H], b3 [out]; weights are row-major with one row per output unit.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "pava_hip.h"
#include "pava_internal.h"

namespace mlpk {

// Y[b, j] = act(sum_k X[b, k] W[j, k] + bias[j])
__global__ void linear_fwd(const float *__restrict__ X, const float *__restrict__ W, const float *__restrict__ bias,
                           float *__restrict__ Z, float *__restrict__ Y, int B, int K, int N, int relu) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * N) return;
    const int b = t / N, j = t % N;
    const float *x = X + (size_t)b * K;
    const float *w = W + (size_t)j * K;
    float s = bias[j];
    for (int k = 0; k < K; ++k) s = __builtin_fmaf(x[k], w[k], s);
    if (Z) Z[t] = s;
    Y[t] = (relu && s < 0.f) ? 0.f : s;
}

// dX[b, k] = (sum_j dY[b, j] W[j, k]) * (Zprev[b, k] > 0)
__global__ void linear_bwd_input(const float *__restrict__ dY, const float *__restrict__ W,
                                 const float *__restrict__ Zprev, float *__restrict__ dX, int B, int K, int N) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * K) return;
    const int b = t / K, k = t % K;
    float s = 0.f;
    for (int j = 0; j < N; ++j) s = __builtin_fmaf(dY[(size_t)b * N + j], W[(size_t)j * K + k], s);
    dX[t] = (Zprev[t] > 0.f) ? s : 0.f;
}

// dW[j, k] = sum_b dY[b, j] X[b, k]; db[j] = sum_b dY[b, j] (thread (j, K) does db)
__global__ void linear_bwd_weight(const float *__restrict__ dY, const float *__restrict__ X, float *__restrict__ dW,
                                  float *__restrict__ db, int B, int K, int N) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * (K + 1)) return;
    const int j = t / (K + 1), k = t % (K + 1);
    float s = 0.f;
    if (k < K) {
        for (int b = 0; b < B; ++b) s = __builtin_fmaf(dY[(size_t)b * N + j], X[(size_t)b * K + k], s);
        dW[(size_t)j * K + k] = s;
    } else {
        for (int b = 0; b < B; ++b) s += dY[(size_t)b * N + j];
        db[j] = s;
    }
}

// masked MSE: loss = sum M (Y - T)^2 / sum M; dY = 2 M (Y - T) / sum M.
// One 256-thread block, strided partial sums then a fixed tree: deterministic.
__global__ __launch_bounds__(256) void mse_loss_grad(const float *__restrict__ Y, const float *__restrict__ T,
                                                     const float *__restrict__ M, float *__restrict__ dY,
                                                     float *__restrict__ loss, int n) {
    __shared__ float sl[256], sm[256];
    const int tid = threadIdx.x;
    float l = 0.f, w = 0.f;
    for (int i = tid; i < n; i += 256) {
        const float d = Y[i] - T[i];
        l = __builtin_fmaf(M[i] * d, d, l);
        w += M[i];
    }
    sl[tid] = l;
    sm[tid] = w;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            sl[tid] += sl[tid + s];
            sm[tid] += sm[tid + s];
        }
        __syncthreads();
    }
    const float inv = sm[0] > 0.f ? 1.f / sm[0] : 0.f;
    for (int i = tid; i < n; i += 256) dY[i] = 2.f * M[i] * (Y[i] - T[i]) * inv;
    if (tid == 0) *loss = sl[0] * inv;
}

__global__ void adam_step(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                          float *__restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void gather_rows(const float *__restrict__ src, const int *__restrict__ idx, float *__restrict__ dst,
                            int rows, int cols) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows * cols) return;
    const int r = t / cols, c = t % cols;
    dst[t] = src[(size_t)idx[r] * cols + c];
}

static inline dim3 g1(long long n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace mlpk

using namespace pava;

struct pava_mlp {
    int device = 0, I = 0, H = 0, O = 0, np = 0, cap = 0;
    float *P = nullptr, *G = nullptr, *Mo = nullptr, *Vo = nullptr;
    // activations for a batch of up to `cap` rows
    float *X = nullptr, *T = nullptr, *Mk = nullptr, *Z1 = nullptr, *A1 = nullptr, *Z2 = nullptr, *A2 = nullptr,
          *Y = nullptr, *dY = nullptr, *dA2 = nullptr, *dA1 = nullptr, *loss = nullptr;
    int *idx = nullptr;
    long long step = 0;
};

#define MLP_TRY(expr)                                                                                       \
    do {                                                                                                    \
        hipError_t _e = (expr);                                                                             \
        if (_e != hipSuccess) return set_error(PAVA_HIP_EDEVICE, "%s: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

static int mlp_reserve(pava_mlp *m, int batch) {
    if (batch <= m->cap) return 0;
    float **bufs[] = {&m->X, &m->T, &m->Mk, &m->Z1, &m->A1, &m->Z2, &m->A2, &m->Y, &m->dY, &m->dA2, &m->dA1};
    for (auto b : bufs)
        if (*b) (void)hipFree(*b), *b = nullptr;
    if (m->idx) (void)hipFree(m->idx), m->idx = nullptr;
    const size_t B = (size_t)batch;
    const size_t sz[] = {B * m->I, B * m->O, B * m->O, B * m->H, B * m->H, B * m->H, B * m->H,
                         B * m->O, B * m->O, B * m->H, B * m->H};
    for (int i = 0; i < 11; ++i) MLP_TRY(hipMalloc(bufs[i], sz[i] * 4 + 16));
    MLP_TRY(hipMalloc(&m->idx, B * 4 + 16));
    m->cap = batch;
    return 0;
}

extern "C" int pava_mlp_create(int n_in, int n_hidden, int n_out, unsigned long long seed, int device, pava_mlp **out) {
    if (!out || n_in <= 0 || n_hidden <= 0 || n_out <= 0) return set_error(PAVA_HIP_EINVAL, "bad MLP shape");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_error(PAVA_HIP_EDEVICE, "no HIP device");
    MLP_TRY(hipSetDevice(device));
    pava_mlp *m = new pava_mlp();
    m->device = device;
    m->I = n_in;
    m->H = n_hidden;
    m->O = n_out;
    m->np = n_hidden * n_in + n_hidden + n_hidden * n_hidden + n_hidden + n_out * n_hidden + n_out;
    if (hipMalloc(&m->P, (size_t)m->np * 4) != hipSuccess || hipMalloc(&m->G, (size_t)m->np * 4) != hipSuccess ||
        hipMalloc(&m->Mo, (size_t)m->np * 4) != hipSuccess || hipMalloc(&m->Vo, (size_t)m->np * 4) != hipSuccess ||
        hipMalloc(&m->loss, 64) != hipSuccess) {
        delete m;
        return set_error(PAVA_HIP_ENOMEM, "MLP allocation failed");
    }
    // He-uniform weights from the counter-based generator, zero biases
    std::vector<float> p((size_t)m->np, 0.f);
    size_t o = 0;
    auto fill = [&](int rows, int fan_in) {
        const double a = std::sqrt(6.0 / fan_in);
        for (int i = 0; i < rows * fan_in; ++i, ++o) p[o] = (float)(a * (2.0 * u01(seed, o) - 1.0));
        o += rows;  // biases stay zero
    };
    fill(n_hidden, n_in);
    fill(n_hidden, n_hidden);
    fill(n_out, n_hidden);
    (void)hipMemcpy(m->P, p.data(), p.size() * 4, hipMemcpyHostToDevice);
    (void)hipMemset(m->Mo, 0, (size_t)m->np * 4);
    (void)hipMemset(m->Vo, 0, (size_t)m->np * 4);
    *out = m;
    return 0;
}

extern "C" int pava_mlp_num_params(const pava_mlp *m) { return m ? m->np : 0; }

extern "C" int pava_mlp_get_params(const pava_mlp *m, float *host) {
    if (!m || !host) return set_error(PAVA_HIP_EINVAL, "null argument");
    MLP_TRY(hipSetDevice(m->device));
    MLP_TRY(hipMemcpy(host, m->P, (size_t)m->np * 4, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int pava_mlp_set_params(pava_mlp *m, const float *host) {
    if (!m || !host) return set_error(PAVA_HIP_EINVAL, "null argument");
    MLP_TRY(hipSetDevice(m->device));
    MLP_TRY(hipMemcpy(m->P, host, (size_t)m->np * 4, hipMemcpyHostToDevice));
    MLP_TRY(hipMemset(m->Mo, 0, (size_t)m->np * 4));
    MLP_TRY(hipMemset(m->Vo, 0, (size_t)m->np * 4));
    m->step = 0;
    return 0;
}

// forward of the rows already in m->X
static int mlp_forward_dev(pava_mlp *m, int B) {
    const int I = m->I, H = m->H, O = m->O;
    float *W1 = m->P, *b1 = W1 + (size_t)H * I, *W2 = b1 + H, *b2 = W2 + (size_t)H * H, *W3 = b2 + H, *b3 = W3 + (size_t)O * H;
    hipLaunchKernelGGL(mlpk::linear_fwd, mlpk::g1((long long)B * H), dim3(256), 0, nullptr, m->X, W1, b1, m->Z1, m->A1, B,
                       I, H, 1);
    hipLaunchKernelGGL(mlpk::linear_fwd, mlpk::g1((long long)B * H), dim3(256), 0, nullptr, m->A1, W2, b2, m->Z2, m->A2,
                       B, H, H, 1);
    hipLaunchKernelGGL(mlpk::linear_fwd, mlpk::g1((long long)B * O), dim3(256), 0, nullptr, m->A2, W3, b3, nullptr, m->Y,
                       B, H, O, 0);
    MLP_TRY(hipGetLastError());
    return 0;
}

// loss and gradient of the rows in m->X / m->T / m->Mk into m->G
static int mlp_backward_dev(pava_mlp *m, int B) {
    const int I = m->I, H = m->H, O = m->O;
    float *W1 = m->P, *b1 = W1 + (size_t)H * I, *W2 = b1 + H, *b2 = W2 + (size_t)H * H, *W3 = b2 + H;
    (void)b1;
    (void)b2;
    float *gW1 = m->G, *gb1 = gW1 + (size_t)H * I, *gW2 = gb1 + H, *gb2 = gW2 + (size_t)H * H, *gW3 = gb2 + H,
          *gb3 = gW3 + (size_t)O * H;
    hipLaunchKernelGGL(mlpk::mse_loss_grad, dim3(1), dim3(256), 0, nullptr, m->Y, m->T, m->Mk, m->dY, m->loss, B * O);
    hipLaunchKernelGGL(mlpk::linear_bwd_weight, mlpk::g1((long long)O * (H + 1)), dim3(256), 0, nullptr, m->dY, m->A2,
                       gW3, gb3, B, H, O);
    hipLaunchKernelGGL(mlpk::linear_bwd_input, mlpk::g1((long long)B * H), dim3(256), 0, nullptr, m->dY, W3, m->Z2,
                       m->dA2, B, H, O);
    hipLaunchKernelGGL(mlpk::linear_bwd_weight, mlpk::g1((long long)H * (H + 1)), dim3(256), 0, nullptr, m->dA2, m->A1,
                       gW2, gb2, B, H, H);
    hipLaunchKernelGGL(mlpk::linear_bwd_input, mlpk::g1((long long)B * H), dim3(256), 0, nullptr, m->dA2, W2, m->Z1,
                       m->dA1, B, H, H);
    hipLaunchKernelGGL(mlpk::linear_bwd_weight, mlpk::g1((long long)H * (I + 1)), dim3(256), 0, nullptr, m->dA1, m->X,
                       gW1, gb1, B, I, H);
    MLP_TRY(hipGetLastError());
    return 0;
}

extern "C" int pava_mlp_predict(pava_mlp *m, int batch, const float *X, float *Y) {
    if (!m || batch < 0 || (batch && (!X || !Y))) return set_error(PAVA_HIP_EINVAL, "bad predict arguments");
    if (batch == 0) return 0;
    MLP_TRY(hipSetDevice(m->device));
    int rc = mlp_reserve(m, batch);
    if (rc) return rc;
    MLP_TRY(hipMemcpy(m->X, X, (size_t)batch * m->I * 4, hipMemcpyHostToDevice));
    if ((rc = mlp_forward_dev(m, batch))) return rc;
    MLP_TRY(hipMemcpy(Y, m->Y, (size_t)batch * m->O * 4, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int pava_mlp_loss_grad(pava_mlp *m, int batch, const float *X, const float *T, const float *M, float *loss,
                                  float *grad) {
    if (!m || batch <= 0 || !X || !T || !M) return set_error(PAVA_HIP_EINVAL, "bad loss arguments");
    MLP_TRY(hipSetDevice(m->device));
    int rc = mlp_reserve(m, batch);
    if (rc) return rc;
    MLP_TRY(hipMemcpy(m->X, X, (size_t)batch * m->I * 4, hipMemcpyHostToDevice));
    MLP_TRY(hipMemcpy(m->T, T, (size_t)batch * m->O * 4, hipMemcpyHostToDevice));
    MLP_TRY(hipMemcpy(m->Mk, M, (size_t)batch * m->O * 4, hipMemcpyHostToDevice));
    if ((rc = mlp_forward_dev(m, batch)) || (rc = mlp_backward_dev(m, batch))) return rc;
    if (loss) MLP_TRY(hipMemcpy(loss, m->loss, 4, hipMemcpyDeviceToHost));
    if (grad) MLP_TRY(hipMemcpy(grad, m->G, (size_t)m->np * 4, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int pava_mlp_train(pava_mlp *m, int n, const float *X, const float *T, const float *M, int epochs,
                              int batch, float lr, unsigned long long seed, float *loss_hist) {
    if (!m || n <= 0 || !X || !T || !M || epochs <= 0 || batch <= 0) return set_error(PAVA_HIP_EINVAL, "bad train arguments");
    MLP_TRY(hipSetDevice(m->device));
    batch = std::min(batch, n);
    int rc = mlp_reserve(m, batch);
    if (rc) return rc;
    float *dX = nullptr, *dT = nullptr, *dM = nullptr;
    auto fail = [&](int code) {
        (void)hipFree(dX);
        (void)hipFree(dT);
        (void)hipFree(dM);
        return code;
    };
    if (hipMalloc(&dX, (size_t)n * m->I * 4) != hipSuccess || hipMalloc(&dT, (size_t)n * m->O * 4) != hipSuccess ||
        hipMalloc(&dM, (size_t)n * m->O * 4) != hipSuccess)
        return fail(set_error(PAVA_HIP_ENOMEM, "training set allocation failed"));
    if (hipMemcpy(dX, X, (size_t)n * m->I * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dT, T, (size_t)n * m->O * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dM, M, (size_t)n * m->O * 4, hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_error(PAVA_HIP_EDEVICE, "training set upload failed"));
    std::vector<int> perm((size_t)n);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs; ++ep) {
        std::iota(perm.begin(), perm.end(), 0);
        for (int i = n - 1; i > 0; --i) {  // Fisher-Yates with the counter generator
            const int j = (int)(rand64(seed ^ (0x9E37ull * (unsigned long long)(ep + 1)), (unsigned long long)i) % (unsigned long long)(i + 1));
            std::swap(perm[(size_t)i], perm[(size_t)j]);
        }
        double esum = 0;
        int nb = 0;
        for (int s0 = 0; s0 < n; s0 += batch) {
            const int B = std::min(batch, n - s0);
            if (hipMemcpy(m->idx, perm.data() + s0, (size_t)B * 4, hipMemcpyHostToDevice) != hipSuccess)
                return fail(set_error(PAVA_HIP_EDEVICE, "index upload failed"));
            hipLaunchKernelGGL(mlpk::gather_rows, mlpk::g1((long long)B * m->I), dim3(256), 0, nullptr, dX, m->idx, m->X, B, m->I);
            hipLaunchKernelGGL(mlpk::gather_rows, mlpk::g1((long long)B * m->O), dim3(256), 0, nullptr, dT, m->idx, m->T, B, m->O);
            hipLaunchKernelGGL(mlpk::gather_rows, mlpk::g1((long long)B * m->O), dim3(256), 0, nullptr, dM, m->idx, m->Mk, B, m->O);
            if ((rc = mlp_forward_dev(m, B)) || (rc = mlp_backward_dev(m, B))) return fail(rc);
            ++m->step;
            const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
            hipLaunchKernelGGL(mlpk::adam_step, mlpk::g1(m->np), dim3(256), 0, nullptr, m->P, m->G, m->Mo, m->Vo, m->np, lr,
                               b1, b2, eps, c1, c2);
            float l = 0.f;
            if (hipMemcpy(&l, m->loss, 4, hipMemcpyDeviceToHost) != hipSuccess)
                return fail(set_error(PAVA_HIP_EDEVICE, "loss download failed"));
            esum += l;
            ++nb;
        }
        if (loss_hist) loss_hist[ep] = (float)(esum / std::max(nb, 1));
    }
    return fail(0);
}

extern "C" int pava_mlp_destroy(pava_mlp *m) {
    if (!m) return 0;
    (void)hipSetDevice(m->device);
    (void)hipDeviceSynchronize();
    float *bufs[] = {m->P, m->G, m->Mo, m->Vo, m->X, m->T, m->Mk, m->Z1, m->A1, m->Z2, m->A2, m->Y, m->dY, m->dA2, m->dA1, m->loss};
    for (float *b : bufs)
        if (b) (void)hipFree(b);
    if (m->idx) (void)hipFree(m->idx);
    delete m;
    return 0;
}
