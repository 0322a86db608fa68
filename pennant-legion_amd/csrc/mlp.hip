// The learned QCS closure's network on the device: parameters, BF16 MFMA
// GEMM, training (forward, backward, Adam) and batched inference.
//
// Training step on a minibatch of B tuples (all matrices row-major, BF16
// operands, FP32 accumulation; H1, H2 and the gradients fed to the next GEMM
// are rounded to BF16 where the MFMA consumes them):
//   Z1 = X W1^T          H1 = relu(Z1)
//   Z2 = H1 W2^T + b2    H2 = relu(Z2)
//   Y  = H2 W3^T + b3    loss = sum_b sum_o (Y - T)^2 / B
//   dY = 2 (Y - T) / B
//   dW3 = dY^T H2   db3 = colsum dY    dZ2 = (dY W3) * [H2 > 0]
//   dW2 = dZ2^T H1  db2 = colsum dZ2   dZ1 = (dZ2 W2) * [H1 > 0]
//   dW1 = dZ1^T X   (column 10 of X is 1, so dW1[:,10] is the bias gradient)
// Every reduction runs in a fixed order (one wave per output tile, split-K
// partials summed in order), so training is deterministic.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "closure.h"
#include "common.h"

using namespace pn;

struct pennant_mlp {
    float* p = nullptr;        // FP32 master parameters (PN_MLP_NPARAM)
    float* g = nullptr;        // gradient
    float* m = nullptr;        // Adam moments
    float* v = nullptr;
    __bf16* w1b = nullptr;     // BF16 copies, row-major [out][in]
    __bf16* w2b = nullptr;
    __bf16* w3b = nullptr;
    ClosurePack* pack = nullptr;
    float lo[PN_MLP_NF], hi[PN_MLP_NF];
    int adam_t = 0;
    // minibatch buffers (capacity cap rows)
    int64_t cap = 0;
    int* idx = nullptr;
    int* idx_host = nullptr;   // pinned
    __bf16 *xb = nullptr, *h1 = nullptr, *h2 = nullptr, *dyb = nullptr, *dz2 = nullptr, *dz1 = nullptr;
    float *t = nullptr, *z = nullptr, *y = nullptr, *dh = nullptr, *part = nullptr, *lossp = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
};

namespace {

#define MCK(x) do { const int e_ = (int)(x); if (e_ != 0) return e_; } while (0)

// ---------------------------------------------------------------------------
// GEMM: C[M][N] = sum_k A(i,k) B(k,j), A(i,k) = ta ? A[k*lda+i] : A[i*lda+k],
// B(k,j) = tb ? B[j*ldb+k] : B[k*ldb+j].  One wave per 32x32 tile of C;
// blockIdx.z splits K into gridDim.z contiguous ranges whose partial tiles go
// to C + z*M*N (summed afterwards in order by k_sum_parts).
__global__ void __launch_bounds__(64) k_gemm(int M, int N, int K, int kper, const __bf16* __restrict__ A,
                                             int lda, int ta, const __bf16* __restrict__ B, int ldb,
                                             int tb, float* __restrict__ C) {
    const int lane = threadIdx.x, h = lane >> 5, r = lane & 31;
    const int i = blockIdx.y * 32 + r;        // A row of this lane
    const int j = blockIdx.x * 32 + r;        // B column of this lane
    const int kb = blockIdx.z * kper;
    const int ke = (kb + kper < K) ? kb + kper : K;
    const __bf16 zero = (__bf16)0.f;
    f32x16 acc = {};
    for (int k0 = kb; k0 < ke; k0 += 16) {
        bf16x8 a, b;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int k = k0 + 8 * h + t;
            const bool kin = k < ke;
            a[t] = (kin && i < M) ? (ta ? A[(int64_t)k * lda + i] : A[(int64_t)i * lda + k]) : zero;
            b[t] = (kin && j < N) ? (tb ? B[(int64_t)j * ldb + k] : B[(int64_t)k * ldb + j]) : zero;
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    float* Cz = C + (int64_t)blockIdx.z * M * N;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int ci = blockIdx.y * 32 + 8 * (q >> 2) + 4 * h + (q & 3);
        const int cj = blockIdx.x * 32 + r;
        if (ci < M && cj < N) Cz[(int64_t)ci * N + cj] = acc[q];
    }
}

__global__ void k_sum_parts(int n, int nparts, const float* __restrict__ part, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int z = 0; z < nparts; ++z) s += part[(int64_t)z * n + i];
    out[i] = s;
}

// C (row-major M x N, FP32) = op(A) op(B).  Uses `part` (M*N*ksplit floats)
// when K is split.
int gemm(int M, int N, int K, const __bf16* A, int lda, int ta, const __bf16* B, int ldb, int tb,
         float* C, float* part, hipStream_t s) {
    // split long reductions (the weight gradients over the batch) into
    // chunks of 512 so that more waves share them
    int ksplit = 1;
    if (K > 1024 && part) ksplit = (K + 511) / 512;
    if (ksplit > 64) ksplit = 64;
    int kper = (K + ksplit - 1) / ksplit;
    kper = (kper + 15) / 16 * 16;
    ksplit = (K + kper - 1) / kper;
    const dim3 grid((N + 31) / 32, (M + 31) / 32, ksplit);
    if (ksplit == 1) {
        hipLaunchKernelGGL(k_gemm, grid, dim3(64), 0, s, M, N, K, kper, A, lda, ta, B, ldb, tb, C);
    } else {
        hipLaunchKernelGGL(k_gemm, grid, dim3(64), 0, s, M, N, K, kper, A, lda, ta, B, ldb, tb, part);
        const int n = M * N;
        hipLaunchKernelGGL(k_sum_parts, dim3((n + 255) / 256), dim3(256), 0, s, n, ksplit, part, C);
    }
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// elementwise and small kernels

// minibatch rows idx[b] of the tuple array: X (BF16, slot 10 = 1) and T
__global__ void k_gather(int B, const int* __restrict__ idx, const float* __restrict__ tup,
                         __bf16* __restrict__ X, float* __restrict__ T) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const float* r = tup + (int64_t)idx[b] * PN_MLP_TUPLE;
#pragma unroll
    for (int k = 0; k < PN_MLP_K1; ++k)
        X[(int64_t)b * PN_MLP_K1 + k] = (__bf16)(k < PN_MLP_NF ? r[k] : (k == PN_MLP_NF ? 1.f : 0.f));
    T[2 * b] = r[PN_MLP_NF];
    T[2 * b + 1] = r[PN_MLP_NF + 1];
}

// X rows (BF16, slot 10 = 1) from feature rows x (B x 10)
__global__ void k_x_rows(int B, const float* __restrict__ x, __bf16* __restrict__ X) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
#pragma unroll
    for (int k = 0; k < PN_MLP_K1; ++k)
        X[(int64_t)b * PN_MLP_K1 + k] =
            (__bf16)(k < PN_MLP_NF ? x[(int64_t)b * PN_MLP_NF + k] : (k == PN_MLP_NF ? 1.f : 0.f));
}

// Y = Z + b3 (inference)
__global__ void k_out(int B, const float* __restrict__ Z, const float* __restrict__ b3, float* __restrict__ Y) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    Y[2 * b] = Z[2 * b] + b3[0];
    Y[2 * b + 1] = Z[2 * b + 1] + b3[1];
}

// H = bf16(relu(Z + bias)), Z is B x 64
__global__ void k_bias_relu(int n, const float* __restrict__ Z, const float* __restrict__ bias,
                            __bf16* __restrict__ H) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float zb = Z[i] + (bias ? bias[i & (PN_MLP_NH - 1)] : 0.f);
    H[i] = (__bf16)fmaxf(zb, 0.f);
}

// Y = Z + b3; dY = 2 (Y - T) / B as BF16; per-block partial losses
__global__ void __launch_bounds__(256) k_out_loss(int B, const float* __restrict__ Z,
                                                  const float* __restrict__ b3,
                                                  const float* __restrict__ T, float* __restrict__ Y,
                                                  __bf16* __restrict__ dYb, float* __restrict__ lossp) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    float l = 0.f;
    if (b < B) {
        const float inv = 2.f / (float)B;
#pragma unroll
        for (int o = 0; o < 2; ++o) {
            const float y = Z[2 * b + o] + b3[o];
            Y[2 * b + o] = y;
            const float d = y - T[2 * b + o];
            l += d * d;
            dYb[2 * b + o] = (__bf16)(inv * d);
        }
    }
    __shared__ float sh[256];
    sh[threadIdx.x] = l;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) lossp[blockIdx.x] = sh[0];
}

// dZ = bf16(dH * [H > 0])
__global__ void k_relu_grad(int n, const float* __restrict__ dH, const __bf16* __restrict__ H,
                            __bf16* __restrict__ dZ) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dZ[i] = (__bf16)(((float)H[i] > 0.f) ? dH[i] : 0.f);
}

// out[j] = sum over rows of X[:, j] (X is rows x ncol BF16); one block per
// column, fixed-order tree
__global__ void __launch_bounds__(256) k_colsum(int rows, int ncol, const __bf16* __restrict__ X,
                                                float* __restrict__ out) {
    const int j = blockIdx.x;
    float s = 0.f;
    for (int r = threadIdx.x; r < rows; r += 256) s += (float)X[(int64_t)r * ncol + j];
    __shared__ float sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = sh[0];
}

__global__ void k_sum_small(int n, const float* __restrict__ x, float* __restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        float s = 0.f;
        for (int i = 0; i < n; ++i) s += x[i];
        *out = s;
    }
}

// Adam (bias-corrected); W1 input columns 11-15 stay zero
__global__ void k_adam(int n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i < PN_MLP_OFF_W2 && (i % PN_MLP_K1) > PN_MLP_NF) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 row-major copies and the A-operand pack of the fused inference
__global__ void __launch_bounds__(256) k_pack(const float* __restrict__ p, __bf16* __restrict__ w1b,
                                              __bf16* __restrict__ w2b, __bf16* __restrict__ w3b,
                                              ClosurePack* __restrict__ pk) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const float* W1 = p + PN_MLP_OFF_W1;
    const float* W2 = p + PN_MLP_OFF_W2;
    const float* b2 = p + PN_MLP_OFF_B2;
    const float* W3 = p + PN_MLP_OFF_W3;
    for (int i = tid; i < PN_MLP_NH * PN_MLP_K1; i += gridDim.x * blockDim.x) w1b[i] = (__bf16)W1[i];
    for (int i = tid; i < PN_MLP_NH * PN_MLP_NH; i += gridDim.x * blockDim.x) w2b[i] = (__bf16)W2[i];
    for (int i = tid; i < PN_MLP_NO * PN_MLP_NH; i += gridDim.x * blockDim.x) w3b[i] = (__bf16)W3[i];
    // fragments: 2 (w1) + 8 (w2) + 4 (w3) per lane
    for (int f = tid; f < 14 * 64; f += gridDim.x * blockDim.x) {
        const int fr = f >> 6, lane = f & 63, h = lane >> 5, m = lane & 31;
        bf16x8 v;
        if (fr < 2) {
            const int row = 32 * fr + m;
            for (int t = 0; t < 8; ++t) v[t] = (__bf16)W1[row * PN_MLP_K1 + 8 * h + t];
            pk->w1[fr][lane] = v;
        } else if (fr < 10) {
            const int mb = (fr - 2) >> 2, c = (fr - 2) & 3, row = 32 * mb + m;
            for (int t = 0; t < 8; ++t) v[t] = (__bf16)W2[row * PN_MLP_NH + 16 * c + 8 * (t >> 2) + 4 * h + (t & 3)];
            pk->w2[mb][c][lane] = v;
        } else {
            const int c = fr - 10;
            for (int t = 0; t < 8; ++t)
                v[t] = (__bf16)(m < PN_MLP_NO ? W3[m * PN_MLP_NH + 16 * c + 8 * (t >> 2) + 4 * h + (t & 3)] : 0.f);
            pk->w3[c][lane] = v;
        }
    }
    for (int i = tid; i < 64; i += gridDim.x * blockDim.x) {
        const int mb = i >> 5, h = (i >> 4) & 1, r = i & 15;
        pk->b2c[mb][h][r] = b2[32 * mb + 8 * (r >> 2) + 4 * h + (r & 3)];
    }
    if (tid < PN_MLP_NO) pk->b3[tid] = p[PN_MLP_OFF_B3 + tid];
}

// Batched inference through the same wave routine the fused predictor uses:
// each lane packs one row of x (n x 10) and receives its two outputs.  Every
// lane of a wave takes part (rows past n are evaluated on zero features and
// discarded).
__global__ void __launch_bounds__(256) k_mlp_wave(const ClosurePack* __restrict__ pk, int64_t n,
                                                  const float* __restrict__ x, float* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool valid = i < n;
    float f[PN_MLP_NF];
#pragma unroll
    for (int k = 0; k < PN_MLP_NF; ++k) f[k] = valid ? x[i * PN_MLP_NF + k] : 0.f;
    u32x4 lo, hi;
    pack_features(f, lo, hi);
    const uint64_t need = __ballot(valid);
    float y0, y1;
    mlp_wave(pk, lo, hi, (uint32_t)need != 0u, (need >> 32) != 0u, y0, y1);
    if (valid) {
        y[2 * i] = y0;
        y[2 * i + 1] = y1;
    }
}

int ensure_batch(pennant_mlp* m, int64_t B) {
    if (B <= m->cap) return 0;
    auto fr = [](void* p) { if (p) (void)hipFree(p); };
    fr(m->idx); fr(m->xb); fr(m->h1); fr(m->h2); fr(m->dyb); fr(m->dz2); fr(m->dz1);
    fr(m->t); fr(m->z); fr(m->y); fr(m->dh); fr(m->part); fr(m->lossp);
    if (m->idx_host) (void)hipHostFree(m->idx_host);
    m->cap = B;
    MCK(hipMalloc(&m->idx, B * 4));
    MCK(hipHostMalloc((void**)&m->idx_host, B * 4));
    MCK(hipMalloc(&m->xb, B * PN_MLP_K1 * 2));
    MCK(hipMalloc(&m->h1, B * PN_MLP_NH * 2));
    MCK(hipMalloc(&m->h2, B * PN_MLP_NH * 2));
    MCK(hipMalloc(&m->dyb, B * PN_MLP_NO * 2));
    MCK(hipMalloc(&m->dz2, B * PN_MLP_NH * 2));
    MCK(hipMalloc(&m->dz1, B * PN_MLP_NH * 2));
    MCK(hipMalloc(&m->t, B * PN_MLP_NO * 4));
    MCK(hipMalloc(&m->z, B * PN_MLP_NH * 4));
    MCK(hipMalloc(&m->y, B * PN_MLP_NO * 4));
    MCK(hipMalloc(&m->dh, B * PN_MLP_NH * 4));
    // split-K partials: at most 64 parts of the largest gradient (64 x 64)
    MCK(hipMalloc(&m->part, (size_t)64 * PN_MLP_NH * PN_MLP_NH * 4));
    MCK(hipMalloc(&m->lossp, ((B + 255) / 256 + 1) * 4));
    return 0;
}

int repack(pennant_mlp* m) {
    hipLaunchKernelGGL(k_pack, dim3(8), dim3(256), 0, m->stream, m->p, m->w1b, m->w2b, m->w3b, m->pack);
    MCK(hipGetLastError());
    MCK(hipMemcpyAsync(m->pack->lo, m->lo, sizeof m->lo, hipMemcpyHostToDevice, m->stream));
    MCK(hipMemcpyAsync(m->pack->hi, m->hi, sizeof m->hi, hipMemcpyHostToDevice, m->stream));
    return 0;
}

inline int nb(int64_t n) { return (int)((n + 255) / 256); }

// forward + backward on the gathered minibatch (m->xb, m->t, B rows);
// leaves the gradient in m->g and the loss in m->lossp[nb(B)]
int fwd_bwd(pennant_mlp* m, int B) {
    hipStream_t s = m->stream;
    float* g = m->g;
    const float* b2 = m->p + PN_MLP_OFF_B2;
    const float* b3 = m->p + PN_MLP_OFF_B3;
    const int nh = B * PN_MLP_NH;
    // forward
    MCK(gemm(B, PN_MLP_NH, PN_MLP_K1, m->xb, PN_MLP_K1, 0, m->w1b, PN_MLP_K1, 1, m->z, m->part, s));
    hipLaunchKernelGGL(k_bias_relu, dim3(nb(nh)), dim3(256), 0, s, nh, m->z, (const float*)nullptr, m->h1);
    MCK(gemm(B, PN_MLP_NH, PN_MLP_NH, m->h1, PN_MLP_NH, 0, m->w2b, PN_MLP_NH, 1, m->z, m->part, s));
    hipLaunchKernelGGL(k_bias_relu, dim3(nb(nh)), dim3(256), 0, s, nh, m->z, b2, m->h2);
    MCK(gemm(B, PN_MLP_NO, PN_MLP_NH, m->h2, PN_MLP_NH, 0, m->w3b, PN_MLP_NH, 1, m->z, m->part, s));
    hipLaunchKernelGGL(k_out_loss, dim3(nb(B)), dim3(256), 0, s, B, m->z, b3, m->t, m->y, m->dyb, m->lossp);
    hipLaunchKernelGGL(k_sum_small, dim3(1), dim3(64), 0, s, nb(B), m->lossp, m->lossp + nb(B));
    // backward
    MCK(gemm(PN_MLP_NO, PN_MLP_NH, B, m->dyb, PN_MLP_NO, 1, m->h2, PN_MLP_NH, 0, g + PN_MLP_OFF_W3, m->part, s));
    hipLaunchKernelGGL(k_colsum, dim3(PN_MLP_NO), dim3(256), 0, s, B, PN_MLP_NO, m->dyb, g + PN_MLP_OFF_B3);
    MCK(gemm(B, PN_MLP_NH, PN_MLP_NO, m->dyb, PN_MLP_NO, 0, m->w3b, PN_MLP_NH, 0, m->dh, m->part, s));
    hipLaunchKernelGGL(k_relu_grad, dim3(nb(nh)), dim3(256), 0, s, nh, m->dh, m->h2, m->dz2);
    MCK(gemm(PN_MLP_NH, PN_MLP_NH, B, m->dz2, PN_MLP_NH, 1, m->h1, PN_MLP_NH, 0, g + PN_MLP_OFF_W2, m->part, s));
    hipLaunchKernelGGL(k_colsum, dim3(PN_MLP_NH), dim3(256), 0, s, B, PN_MLP_NH, m->dz2, g + PN_MLP_OFF_B2);
    MCK(gemm(B, PN_MLP_NH, PN_MLP_NH, m->dz2, PN_MLP_NH, 0, m->w2b, PN_MLP_NH, 0, m->dh, m->part, s));
    hipLaunchKernelGGL(k_relu_grad, dim3(nb(nh)), dim3(256), 0, s, nh, m->dh, m->h1, m->dz1);
    MCK(gemm(PN_MLP_NH, PN_MLP_K1, B, m->dz1, PN_MLP_NH, 1, m->xb, PN_MLP_K1, 0, g + PN_MLP_OFF_W1, m->part, s));
    return (int)hipGetLastError();
}

}  // namespace

extern "C" {

int pennant_mlp_num_params(void) { return PN_MLP_NPARAM; }

pennant_mlp* pennant_mlp_create(uint64_t seed) {
    pennant_mlp* m = new pennant_mlp();
    bool ok = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc(&m->p, PN_MLP_NPARAM * 4) == hipSuccess;
    ok = ok && hipMalloc(&m->g, PN_MLP_NPARAM * 4) == hipSuccess;
    ok = ok && hipMalloc(&m->m, PN_MLP_NPARAM * 4) == hipSuccess;
    ok = ok && hipMalloc(&m->v, PN_MLP_NPARAM * 4) == hipSuccess;
    ok = ok && hipMalloc(&m->w1b, PN_MLP_NH * PN_MLP_K1 * 2) == hipSuccess;
    ok = ok && hipMalloc(&m->w2b, PN_MLP_NH * PN_MLP_NH * 2) == hipSuccess;
    ok = ok && hipMalloc(&m->w3b, PN_MLP_NO * PN_MLP_NH * 2) == hipSuccess;
    ok = ok && hipMalloc(&m->pack, sizeof(ClosurePack)) == hipSuccess;
    if (!ok) { pennant_mlp_destroy(m); return nullptr; }
    // He-uniform initial weights from a seeded generator; zero biases
    std::vector<float> h(PN_MLP_NPARAM, 0.f);
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    const double a1 = std::sqrt(6.0 / (PN_MLP_NF + 1)), a2 = std::sqrt(6.0 / PN_MLP_NH);
    for (int o = 0; o < PN_MLP_NH; ++o)
        for (int k = 0; k < PN_MLP_NF; ++k) h[PN_MLP_OFF_W1 + o * PN_MLP_K1 + k] = (float)(a1 * u(rng));
    for (int i = 0; i < PN_MLP_NH * PN_MLP_NH; ++i) h[PN_MLP_OFF_W2 + i] = (float)(a2 * u(rng));
    for (int i = 0; i < PN_MLP_NO * PN_MLP_NH; ++i) h[PN_MLP_OFF_W3 + i] = (float)(a2 * u(rng));
    for (int k = 0; k < PN_MLP_NF; ++k) { m->lo[k] = -INFINITY; m->hi[k] = INFINITY; }
    if (pennant_mlp_set_params(m, h.data(), PN_MLP_NPARAM) != 0) { pennant_mlp_destroy(m); return nullptr; }
    return m;
}

void pennant_mlp_destroy(pennant_mlp* m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    void* bufs[] = {m->p, m->g, m->m, m->v, m->w1b, m->w2b, m->w3b, m->pack, m->idx, m->xb, m->h1,
                    m->h2, m->dyb, m->dz2, m->dz1, m->t, m->z, m->y, m->dh, m->part, m->lossp};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (m->idx_host) (void)hipHostFree(m->idx_host);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

int pennant_mlp_set_params(pennant_mlp* m, const float* host, int n) {
    if (n != PN_MLP_NPARAM) return PENNANT_EINVAL;
    MCK(hipMemcpyAsync(m->p, host, PN_MLP_NPARAM * 4, hipMemcpyHostToDevice, m->stream));
    MCK(hipMemsetAsync(m->m, 0, PN_MLP_NPARAM * 4, m->stream));
    MCK(hipMemsetAsync(m->v, 0, PN_MLP_NPARAM * 4, m->stream));
    m->adam_t = 0;
    MCK(repack(m));
    return (int)hipStreamSynchronize(m->stream);
}

int pennant_mlp_get_params(pennant_mlp* m, float* host, int n) {
    if (n != PN_MLP_NPARAM) return PENNANT_EINVAL;
    MCK(hipMemcpyAsync(host, m->p, PN_MLP_NPARAM * 4, hipMemcpyDeviceToHost, m->stream));
    return (int)hipStreamSynchronize(m->stream);
}

int pennant_mlp_set_envelope(pennant_mlp* m, const float* lo, const float* hi) {
    for (int k = 0; k < PN_MLP_NF; ++k) { m->lo[k] = lo[k]; m->hi[k] = hi[k]; }
    MCK(repack(m));
    return (int)hipStreamSynchronize(m->stream);
}

int pennant_mlp_get_envelope(pennant_mlp* m, float* lo, float* hi) {
    for (int k = 0; k < PN_MLP_NF; ++k) { lo[k] = m->lo[k]; hi[k] = m->hi[k]; }
    return 0;
}

const void* pennant_mlp_device_pack(pennant_mlp* m) { return m->pack; }

int pennant_mlp_forward(pennant_mlp* m, const float* x, int64_t n, float* y, int path) {
    if (n < 0 || (path != 0 && path != 1)) return PENNANT_EINVAL;
    if (n == 0) return 0;
    if (path == 1) {
        hipLaunchKernelGGL(k_mlp_wave, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->stream, m->pack,
                           n, x, y);
        MCK(hipGetLastError());
        return (int)hipStreamSynchronize(m->stream);
    }
    // GEMM path, in chunks of the batch capacity
    const int64_t chunk = 65536;
    MCK(ensure_batch(m, chunk));
    hipStream_t s = m->stream;
    for (int64_t b0 = 0; b0 < n; b0 += chunk) {
        const int B = (int)((n - b0 < chunk) ? n - b0 : chunk);
        const int nh = B * PN_MLP_NH;
        hipLaunchKernelGGL(k_x_rows, dim3(nb(B)), dim3(256), 0, s, B, x + b0 * PN_MLP_NF, m->xb);
        MCK(gemm(B, PN_MLP_NH, PN_MLP_K1, m->xb, PN_MLP_K1, 0, m->w1b, PN_MLP_K1, 1, m->z, m->part, s));
        hipLaunchKernelGGL(k_bias_relu, dim3(nb(nh)), dim3(256), 0, s, nh, m->z, (const float*)nullptr, m->h1);
        MCK(gemm(B, PN_MLP_NH, PN_MLP_NH, m->h1, PN_MLP_NH, 0, m->w2b, PN_MLP_NH, 1, m->z, m->part, s));
        hipLaunchKernelGGL(k_bias_relu, dim3(nb(nh)), dim3(256), 0, s, nh, m->z, m->p + PN_MLP_OFF_B2, m->h2);
        MCK(gemm(B, PN_MLP_NO, PN_MLP_NH, m->h2, PN_MLP_NH, 0, m->w3b, PN_MLP_NH, 1, m->z, m->part, s));
        hipLaunchKernelGGL(k_out, dim3(nb(B)), dim3(256), 0, s, B, m->z, m->p + PN_MLP_OFF_B3, y + 2 * b0);
        MCK(hipGetLastError());
    }
    return (int)hipStreamSynchronize(s);
}

int pennant_mlp_loss_grad(pennant_mlp* m, const float* tuples, int64_t n, float* grad, float* loss) {
    if (n <= 0 || n > (1 << 30)) return PENNANT_EINVAL;
    MCK(ensure_batch(m, n));
    for (int64_t b = 0; b < n; ++b) m->idx_host[b] = (int)b;
    MCK(hipMemcpyAsync(m->idx, m->idx_host, n * 4, hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_gather, dim3(nb(n)), dim3(256), 0, m->stream, (int)n, m->idx, tuples, m->xb, m->t);
    MCK(fwd_bwd(m, (int)n));
    MCK(hipMemcpyAsync(grad, m->g, PN_MLP_NPARAM * 4, hipMemcpyDeviceToHost, m->stream));
    MCK(hipMemcpyAsync(loss, m->lossp + nb(n), 4, hipMemcpyDeviceToHost, m->stream));
    MCK(hipStreamSynchronize(m->stream));
    *loss /= (float)n;
    return 0;
}

int pennant_mlp_train(pennant_mlp* m, const float* tuples, int64_t n, int steps, int batch, float lr,
                      uint64_t seed, float* loss_out) {
    if (n <= 0 || n > (1 << 30) || steps < 0 || batch <= 0 || batch > (1 << 22)) return PENNANT_EINVAL;
    MCK(ensure_batch(m, batch));
    MCK(hipMemsetAsync(m->g, 0, PN_MLP_NPARAM * 4, m->stream));
    std::mt19937_64 rng(seed);
    std::uniform_int_distribution<int64_t> pick(0, n - 1);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int st = 0; st < steps; ++st) {
        MCK(hipStreamSynchronize(m->stream));    // the pinned index buffer is reused
        for (int b = 0; b < batch; ++b) m->idx_host[b] = (int)pick(rng);
        MCK(hipMemcpyAsync(m->idx, m->idx_host, (size_t)batch * 4, hipMemcpyHostToDevice, m->stream));
        hipLaunchKernelGGL(k_gather, dim3(nb(batch)), dim3(256), 0, m->stream, batch, m->idx, tuples, m->xb, m->t);
        MCK(fwd_bwd(m, batch));
        m->adam_t += 1;
        const float c1 = 1.f - std::pow(b1, (float)m->adam_t), c2 = 1.f - std::pow(b2, (float)m->adam_t);
        // cosine decay from lr to lr/50 over the call's steps
        const float lrt = lr * (0.02f + 0.98f * 0.5f * (1.f + std::cos(3.14159265f * (float)st / (float)steps)));
        hipLaunchKernelGGL(k_adam, dim3(nb(PN_MLP_NPARAM)), dim3(256), 0, m->stream, PN_MLP_NPARAM, m->p, m->g,
                           m->m, m->v, lrt, b1, b2, eps, c1, c2);
        hipLaunchKernelGGL(k_pack, dim3(8), dim3(256), 0, m->stream, m->p, m->w1b, m->w2b, m->w3b, m->pack);
        if (loss_out) MCK(hipMemcpyAsync(loss_out + st, m->lossp + nb(batch), 4, hipMemcpyDeviceToHost, m->stream));
    }
    MCK(hipStreamSynchronize(m->stream));
    if (loss_out)
        for (int st = 0; st < steps; ++st) loss_out[st] /= (float)batch;
    return (int)hipGetLastError();
}

}  // extern "C"
