// mlp.hip -- the small multilayer perceptron of the learned SCF guess:
// forward, backward and Adam on the device, GEMMs on the MFMA units in BF16
// with FP32 accumulation (master weights, activations and gradients FP32).
//
// Layer l: Z_l = A_{l-1} W_l^T + b_l, A_l = silu(Z_l) (identity on the last
// layer); A_0 = X (batch x d_0), W_l is d_l x d_{l-1} row-major.  The loss is
// sum_{r,o} m_{r,o} (A_L - T)_{r,o}^2 / sum m (m: per-element weights, 0 for
// padding), so every output of a variable-size density block can share one
// network width.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/polyeri.h"

// polyeri.hip: records the message returned by polyeri_last_error()
int polyeri_record_error(int code, const char* msg);

namespace {

int mfail(int code, const std::string& m) { return polyeri_record_error(code, m.c_str()); }

#define MCHK(x)                                                                          \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) return mfail(POLYERI_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

// FP32 -> BF16 bits, round to nearest even (finite inputs).
__device__ __forceinline__ short to_bf16(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// C[m][n] = alpha * sum_k A(m,k) B(k,n) + beta * C[m][n], A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn], C row-major (ldc).  Operands are rounded to BF16
// and multiplied on v_mfma_f32_16x16x16bf16_1k (FP32 accumulation).  One
// wavefront per 16x16 tile of C; lane l holds row l%16 of the A slice and
// column l%16 of the B slice for k = 4*(l/16) .. +3, and receives C rows
// 4*(l/16) .. +3 of column l%16.  Out-of-range rows, columns and k are zero.
__global__ void __launch_bounds__(256) k_gemm_bf16(int M, int N, int K, const float* __restrict__ A, long sam,
                                                   long sak, const float* __restrict__ B, long sbk, long sbn,
                                                   float alpha, float beta, float* __restrict__ C, long ldc) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tiles_n = (N + 15) >> 4;
    const long tile = (long)blockIdx.x * 4 + wave;
    const int tm = (int)(tile / tiles_n), tn = (int)(tile % tiles_n);
    if ((long)tm * 16 >= M) return;
    const int r = lane & 15, g = lane >> 4;
    const int m = tm * 16 + r, n = tn * 16 + r;
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = k0 + 4 * g + j;
            const float av = (m < M && k < K) ? A[(long)m * sam + (long)k * sak] : 0.f;
            const float bv = (n < N && k < K) ? B[(long)k * sbk + (long)n * sbn] : 0.f;
            a[j] = to_bf16(av);
            b[j] = to_bf16(bv);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int row = tm * 16 + 4 * g + j, col = tn * 16 + r;
        if (row < M && col < N) {
            float* c = C + (long)row * ldc + col;
            *c = alpha * acc[j] + (beta != 0.f ? beta * *c : 0.f);
        }
    }
}

__device__ __forceinline__ float silu(float z) { return z / (1.f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
    const float s = 1.f / (1.f + __expf(-z));
    return s * (1.f + z * (1.f - s));
}

// Z += b (row broadcast); A = silu(Z) or Z.
__global__ void k_bias_act(float* __restrict__ Z, float* __restrict__ Aout, const float* __restrict__ b, int rows,
                           int cols, int act) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)rows * cols) return;
    const float z = Z[t] + b[t % cols];
    Z[t] = z;
    Aout[t] = act ? silu(z) : z;
}

// dZ = dA * silu'(Z) in place on dA.
__global__ void k_act_back(float* __restrict__ dA, const float* __restrict__ Z, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dA[t] *= silu_grad(Z[t]);
}

// db[c] = sum_r dZ[r][c] (one thread per column; batches are small).
__global__ void k_colsum(const float* __restrict__ dZ, int rows, int cols, float* __restrict__ db) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int r = 0; r < rows; ++r) s += dZ[(long)r * cols + c];
    db[c] = s;
}

// Weighted squared error: partial sums of m (y - t)^2 per block into part[],
// and dY = 2 m (y - t) / wsum.
__global__ void k_mse(const float* __restrict__ Y, const float* __restrict__ T, const float* __restrict__ Wm, long n,
                      float inv_wsum, float* __restrict__ dY, float* __restrict__ part) {
    __shared__ float sh[256];
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float e = 0.f;
    if (t < n) {
        const float d = Y[t] - T[t], w = Wm[t];
        e = w * d * d;
        dY[t] = 2.f * w * d * inv_wsum;
    }
    sh[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, long n, float lr, float b1, float b2, float c1, float c2, float eps) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gt = g[t];
    const float a = b1 * m1[t] + (1.f - b1) * gt;
    const float v = b2 * m2[t] + (1.f - b2) * gt * gt;
    m1[t] = a;
    m2[t] = v;
    p[t] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

unsigned nb(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace

struct polyeri_mlp {
    int nl = 0;                       // layers
    std::vector<int> dim;             // nl + 1
    std::vector<long> woff, boff;     // offsets into the parameter vector
    long nparam = 0;
    float *d_p = nullptr, *d_g = nullptr, *d_m1 = nullptr, *d_m2 = nullptr;
    int cap = 0;                      // batch capacity of the activation buffers
    std::vector<float*> d_z, d_a;     // per layer: pre-activations, activations (batch x dim)
    float *d_part = nullptr;
    long step = 0;
    hipStream_t stream = nullptr;
};

static int ensure_batch(polyeri_mlp* m, int batch) {
    if (batch <= m->cap) return 0;
    for (float* p : m->d_z) (void)hipFree(p);
    for (float* p : m->d_a) (void)hipFree(p);
    (void)hipFree(m->d_part);
    m->d_z.assign(m->nl + 1, nullptr);
    m->d_a.assign(m->nl + 1, nullptr);
    for (int l = 1; l <= m->nl; ++l) {
        MCHK(hipMalloc((void**)&m->d_z[l], sizeof(float) * (size_t)batch * m->dim[l]));
        MCHK(hipMalloc((void**)&m->d_a[l], sizeof(float) * (size_t)batch * m->dim[l]));
    }
    MCHK(hipMalloc((void**)&m->d_part, sizeof(float) * (size_t)nb((long)batch * m->dim[m->nl])));
    m->cap = batch;
    return 0;
}

static void gemm(hipStream_t s, int M, int N, int K, const float* A, long sam, long sak, const float* B, long sbk,
                 long sbn, float alpha, float beta, float* C, long ldc) {
    const long tiles = (long)((M + 15) / 16) * ((N + 15) / 16);
    hipLaunchKernelGGL(k_gemm_bf16, dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, s, M, N, K, A, sam, sak, B, sbk,
                       sbn, alpha, beta, C, ldc);
}

extern "C" {

int polyeri_mlp_create(int nlayers, const int* dims, unsigned long long seed, polyeri_mlp** out) {
    if (!dims || !out || nlayers < 1 || nlayers > 16) return mfail(POLYERI_E_ARG, "polyeri_mlp_create: bad arguments");
    for (int l = 0; l <= nlayers; ++l)
        if (dims[l] < 1 || dims[l] > 4096) return mfail(POLYERI_E_ARG, "polyeri_mlp_create: layer width out of range");
    polyeri_mlp* m = new polyeri_mlp;
    m->nl = nlayers;
    m->dim.assign(dims, dims + nlayers + 1);
    for (int l = 1; l <= nlayers; ++l) {
        m->woff.push_back(m->nparam);
        m->nparam += (long)dims[l] * dims[l - 1];
        m->boff.push_back(m->nparam);
        m->nparam += dims[l];
    }
    // He-style initialisation from a fixed xorshift stream
    std::vector<float> p(m->nparam, 0.f);
    unsigned long long s = seed * 0x9E3779B97F4A7C15ull + 1;
    auto uni = [&]() {
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        return (double)(s >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int l = 1; l <= nlayers; ++l) {
        const double sc = std::sqrt(6.0 / dims[l - 1]) * (l == nlayers ? 0.1 : 1.0);
        for (long i = 0; i < (long)dims[l] * dims[l - 1]; ++i) p[m->woff[l - 1] + i] = (float)((2.0 * uni() - 1.0) * sc);
    }
    int rc = 0;
    hipError_t e = hipMalloc((void**)&m->d_p, sizeof(float) * m->nparam);
    if (e == hipSuccess) e = hipMalloc((void**)&m->d_g, sizeof(float) * m->nparam);
    if (e == hipSuccess) e = hipMalloc((void**)&m->d_m1, sizeof(float) * m->nparam);
    if (e == hipSuccess) e = hipMalloc((void**)&m->d_m2, sizeof(float) * m->nparam);
    if (e == hipSuccess) e = hipMemcpy(m->d_p, p.data(), sizeof(float) * m->nparam, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(m->d_m1, 0, sizeof(float) * m->nparam);
    if (e == hipSuccess) e = hipMemset(m->d_m2, 0, sizeof(float) * m->nparam);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e != hipSuccess) rc = mfail(POLYERI_E_HIP, std::string("polyeri_mlp_create: ") + hipGetErrorString(e));
    if (rc) { polyeri_mlp_destroy(m); return rc; }
    *out = m;
    return 0;
}

void polyeri_mlp_destroy(polyeri_mlp* m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    (void)hipFree(m->d_p); (void)hipFree(m->d_g); (void)hipFree(m->d_m1); (void)hipFree(m->d_m2);
    for (float* p : m->d_z) (void)hipFree(p);
    for (float* p : m->d_a) (void)hipFree(p);
    (void)hipFree(m->d_part);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

long polyeri_mlp_nparam(const polyeri_mlp* m) { return m ? m->nparam : -1; }

int polyeri_mlp_params(polyeri_mlp* m, float* host, int set) {
    if (!m || !host) return mfail(POLYERI_E_ARG, "polyeri_mlp_params: null argument");
    MCHK(hipStreamSynchronize(m->stream));
    if (set) {
        MCHK(hipMemcpy(m->d_p, host, sizeof(float) * m->nparam, hipMemcpyHostToDevice));
        MCHK(hipMemset(m->d_m1, 0, sizeof(float) * m->nparam));
        MCHK(hipMemset(m->d_m2, 0, sizeof(float) * m->nparam));
        m->step = 0;
    } else {
        MCHK(hipMemcpy(host, m->d_p, sizeof(float) * m->nparam, hipMemcpyDeviceToHost));
    }
    return 0;
}

int polyeri_mlp_grads(polyeri_mlp* m, float* host) {
    if (!m || !host) return mfail(POLYERI_E_ARG, "polyeri_mlp_grads: null argument");
    MCHK(hipStreamSynchronize(m->stream));
    MCHK(hipMemcpy(host, m->d_g, sizeof(float) * m->nparam, hipMemcpyDeviceToHost));
    return 0;
}

static int forward(polyeri_mlp* m, const float* d_x, int batch) {
    int rc = ensure_batch(m, batch);
    if (rc) return rc;
    const float* a = d_x;
    for (int l = 1; l <= m->nl; ++l) {
        const int din = m->dim[l - 1], dout = m->dim[l];
        // Z = A W^T: A(b,k) = a[b*din + k], W^T(k,o) = W[o*din + k]
        gemm(m->stream, batch, dout, din, a, din, 1, m->d_p + m->woff[l - 1], 1, din, 1.f, 0.f, m->d_z[l], dout);
        hipLaunchKernelGGL(k_bias_act, dim3(nb((long)batch * dout)), dim3(256), 0, m->stream, m->d_z[l], m->d_a[l],
                           m->d_p + m->boff[l - 1], batch, dout, l < m->nl ? 1 : 0);
        a = m->d_a[l];
    }
    MCHK(hipGetLastError());
    return 0;
}

int polyeri_mlp_forward(polyeri_mlp* m, const float* d_x, int batch, float* d_y) {
    if (!m || !d_x || !d_y || batch < 1) return mfail(POLYERI_E_ARG, "polyeri_mlp_forward: bad arguments");
    int rc = forward(m, d_x, batch);
    if (rc) return rc;
    MCHK(hipMemcpyAsync(d_y, m->d_a[m->nl], sizeof(float) * (size_t)batch * m->dim[m->nl], hipMemcpyDeviceToDevice,
                        m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    return 0;
}

int polyeri_mlp_train_step(polyeri_mlp* m, const float* d_x, const float* d_t, const float* d_w, int batch,
                           float lr, int update, float* loss) {
    if (!m || !d_x || !d_t || !d_w || batch < 1) return mfail(POLYERI_E_ARG, "polyeri_mlp_train_step: bad arguments");
    int rc = forward(m, d_x, batch);
    if (rc) return rc;
    const int L = m->nl, dl = m->dim[L];
    const long n = (long)batch * dl;
    // weight sum on the host side of the call: the caller's weights are fixed per batch
    std::vector<float> w(n);
    MCHK(hipMemcpyAsync(w.data(), d_w, sizeof(float) * n, hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    double ws = 0;
    for (float x : w) ws += x;
    if (!(ws > 0)) return mfail(POLYERI_E_ARG, "polyeri_mlp_train_step: all loss weights are zero");
    // dA_L (reuse Z_L's buffer after the loss: the last layer is linear)
    float* dA = m->d_z[L];
    hipLaunchKernelGGL(k_mse, dim3(nb(n)), dim3(256), 0, m->stream, m->d_a[L], d_t, d_w, n, (float)(1.0 / ws), dA,
                       m->d_part);
    for (int l = L; l >= 1; --l) {
        const int din = m->dim[l - 1], dout = m->dim[l];
        const float* ain = l > 1 ? m->d_a[l - 1] : d_x;
        // dW = dZ^T A_in: (dout x din), dZ(b,o) = dA[b*dout + o]
        gemm(m->stream, dout, din, batch, dA, 1, dout, ain, din, 1, 1.f, 0.f, m->d_g + m->woff[l - 1], din);
        hipLaunchKernelGGL(k_colsum, dim3(nb(dout)), dim3(256), 0, m->stream, dA, batch, dout, m->d_g + m->boff[l - 1]);
        if (l > 1) {
            // dA_{l-1} = dZ W (batch x din), into A_{l-1}'s pre-activation-sized scratch
            float* dprev = m->d_a[l - 1];
            // Z_{l-1} is still needed for the activation derivative: dA goes to A's buffer
            gemm(m->stream, batch, din, dout, dA, dout, 1, m->d_p + m->woff[l - 1], din, 1, 1.f, 0.f, dprev, din);
            hipLaunchKernelGGL(k_act_back, dim3(nb((long)batch * din)), dim3(256), 0, m->stream, dprev,
                               m->d_z[l - 1], (long)batch * din);
            dA = dprev;
        }
    }
    MCHK(hipGetLastError());
    std::vector<float> part(nb(n));
    MCHK(hipMemcpyAsync(part.data(), m->d_part, sizeof(float) * part.size(), hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    double s = 0;
    for (float x : part) s += x;
    if (loss) *loss = (float)(s / ws);
    if (update) {
        m->step += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
        hipLaunchKernelGGL(k_adam, dim3(nb(m->nparam)), dim3(256), 0, m->stream, m->d_p, m->d_g, m->d_m1, m->d_m2,
                           m->nparam, lr, b1, b2, c1, c2, 1e-8f);
        MCHK(hipGetLastError());
        MCHK(hipStreamSynchronize(m->stream));
    }
    return 0;
}

int polyeri_gemm_bf16(int M, int N, int K, const float* d_a, long sam, long sak, const float* d_b, long sbk,
                      long sbn, float alpha, float beta, float* d_c, long ldc) {
    if (M < 1 || N < 1 || K < 1 || !d_a || !d_b || !d_c) return mfail(POLYERI_E_ARG, "polyeri_gemm_bf16: bad arguments");
    gemm(nullptr, M, N, K, d_a, sam, sak, d_b, sbk, sbn, alpha, beta, d_c, ldc);
    MCHK(hipGetLastError());
    MCHK(hipDeviceSynchronize());
    return 0;
}

}  // extern "C"
