// mlp.hip - training kernels of the learned proposal's network (ldo_mlp.h).
//
// One sample is a 7x7x7 one-hot box (13 channels per site) around a staple
// end domain's anchor plus the domain's identity; the network is
//   z1 = b1 + sum over sites of W1[site*13 + channel] + emb[id] . W1e
//   h1 = relu(z1); z2 = h1 W2 + b2; h2 = relu(z2); z3 = h2 W3 + b3;
//   h3 = relu(z3); z4 = h3 W4 + b4 (36 logits), masked softmax.
// The first layer is a gather (its input is one-hot); the products with W2,
// W3, W4 and their backward products run on gfx950's 16x16x32 bf16 MFMA with
// FP32 accumulation and LDS-staged bf16 operand tiles.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/ldo.h"
#include "../../include/ldo_mlp.h"
#include "plan.h"

namespace {

constexpr int H = LDO_MLP_HIDDEN;
constexpr int NO = LDO_MLP_OUT;
constexpr int NS = LDO_MLP_SITES;
constexpr int NE = LDO_MLP_EMBED;

typedef float v4f __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (finite inputs)
__device__ inline short to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// C(i, j) [+]= sum_k A(i, k) B(k, j) with A(i, k) = A[i*as0 + k*as1] and
// B(k, j) = B[k*bs0 + j*bs1] (strides express the transposes).
struct GemmArgs {
    int M, N, K;
    const float* A;
    int64_t as0, as1;
    const float* B;
    int64_t bs0, bs1;
    float* C;
    int64_t cs0;
    int accumulate;
};

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// One 32 x 32 output tile per workgroup of four wavefronts, each owning a
// 16 x 16 quarter computed with gfx950's v_mfma_f32_16x16x32_bf16 (32 k per
// instruction).  Every k step stages the A tile (row i, k contiguous) and
// the B tile (column j, k contiguous) in LDS, rounded to bf16 once, with rows
// padded by 8 elements; a lane's operand is then one 16-byte LDS read.
// Operand layout: A(i, k) in lane i + 16*(k/8), element k%8; B(k, j) in
// lane j + 16*(k/8), element k%8; D(i, j) in lane j + 16*(i/4), element i%4.
// (Only the row and column mappings matter: A and B share the k mapping, so
// any order of k inside the instruction gives the same dot products.)
constexpr int GT = 32;      // output tile edge
constexpr int GK = 32;      // k per step
constexpr int GP = GK + 8;  // padded LDS row (elements)

__global__ void __launch_bounds__(256) mlp_gemm_kernel(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) unsigned short As[GT * GP];
    __shared__ __attribute__((aligned(16))) unsigned short Bs[GT * GP];
    int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int i0 = blockIdx.y * GT, j0 = blockIdx.x * GT;
    int wi = (w >> 1) * 16, wj = (w & 1) * 16;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.K; k0 += GK) {
        for (int e = tid; e < GT * GK; e += 256) {
            int r = e / GK, k = e % GK;
            int kk = k0 + k;
            int i = i0 + r, j = j0 + r;
            float av = (i < g.M && kk < g.K) ? g.A[i * g.as0 + kk * g.as1] : 0.f;
            float bv = (j < g.N && kk < g.K) ? g.B[kk * g.bs0 + j * g.bs1] : 0.f;
            As[r * GP + k] = (unsigned short)to_bf16(av);
            Bs[r * GP + k] = (unsigned short)to_bf16(bv);
        }
        __syncthreads();
        v8bf a, b;
        __builtin_memcpy(&a, &As[(wi + (lane & 15)) * GP + (lane >> 4) * 8], sizeof a);
        __builtin_memcpy(&b, &Bs[(wj + (lane & 15)) * GP + (lane >> 4) * 8], sizeof b);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        __syncthreads();
    }
    for (int t = 0; t < 4; t++) {
        int ii = i0 + wi + (lane >> 4) * 4 + t, jj = j0 + wj + (lane & 15);
        if (ii < g.M && jj < g.N) {
            float* c = g.C + ii * g.cs0 + jj;
            *c = g.accumulate ? *c + acc[t] : acc[t];
        }
    }
}

// z1 and h1 of one sample per workgroup, one hidden unit per lane
__global__ void __launch_bounds__(256) mlp_layer1_fwd(const float* P, ldo_mlp_layout L,
                                                      const uint8_t* ch, const int32_t* id,
                                                      int batch, float* z1, float* h1) {
    int b = blockIdx.x, j = threadIdx.x;
    if (b >= batch) {
        return;
    }
    const uint8_t* c = ch + (size_t)b * NS;
    float s = P[L.b1 + j];
    for (int site = 0; site < NS; site++) {
        int f = site * LDO_MLP_CHANNELS + c[site];
        s += P[L.w1 + (int64_t)f * H + j];
    }
    const float* e = P + L.emb + (int64_t)id[b] * NE;
    for (int k = 0; k < NE; k++) {
        s = fmaf(e[k], P[L.w1e + k * H + j], s);
    }
    z1[b * H + j] = s;
    h1[b * H + j] = fmaxf(s, 0.f);
}

__global__ void mlp_bias_act(float* z, const float* bias, float* h, int M, int N, int relu) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= M * N) {
        return;
    }
    float v = z[n] + bias[n % N];
    z[n] = v;
    if (h) {
        h[n] = relu ? fmaxf(v, 0.f) : v;
    }
}

// masked log-softmax, per-sample loss and d(mean loss)/dz4
__global__ void mlp_softmax_ce(const float* z4, const uint8_t* legal, const int32_t* tgt, int batch,
                               float* dz4, double* loss, float* logp) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) {
        return;
    }
    const float* z = z4 + b * NO;
    const uint8_t* lg = legal + b * NO;
    float m = -INFINITY;
    for (int k = 0; k < NO; k++) {
        if (lg[k]) {
            m = fmaxf(m, z[k]);
        }
    }
    double sum = 0;
    for (int k = 0; k < NO; k++) {
        if (lg[k]) {
            sum += exp((double)(z[k] - m));
        }
    }
    double lse = (double)m + log(sum);
    if (logp) {
        for (int k = 0; k < NO; k++) {
            logp[b * NO + k] = lg[k] ? (float)((double)z[k] - lse) : -INFINITY;
        }
    }
    if (tgt) {
        int t = tgt[b];
        loss[b] = lse - (double)z[t];
        for (int k = 0; k < NO; k++) {
            double p = lg[k] ? exp((double)z[k] - lse) : 0.0;
            dz4[b * NO + k] = (float)((p - (k == t ? 1.0 : 0.0)) / batch);
        }
    }
}

__global__ void mlp_relu_back(const float* dh, const float* z, float* dz, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        dz[i] = z[i] > 0.f ? dh[i] : 0.f;
    }
}

__global__ void mlp_colsum(const float* x, int M, int N, float* out) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) {
        return;
    }
    float s = 0.f;
    for (int i = 0; i < M; i++) {
        s += x[i * N + j];
    }
    out[j] = s;
}

// gradients of W1 (scatter over the sample's active features), W1e and the
// embedding row of the sample's identity
__global__ void __launch_bounds__(256) mlp_layer1_bwd(const float* P, ldo_mlp_layout L,
                                                      const uint8_t* ch, const int32_t* id,
                                                      int batch, const float* dz1, float* G) {
    __shared__ float part[NE][H];
    int b = blockIdx.x, j = threadIdx.x;
    if (b >= batch) {
        return;
    }
    float d = dz1[b * H + j];
    const uint8_t* c = ch + (size_t)b * NS;
    for (int site = 0; site < NS; site++) {
        int f = site * LDO_MLP_CHANNELS + c[site];
        atomicAdd(&G[L.w1 + (int64_t)f * H + j], d);
    }
    const float* e = P + L.emb + (int64_t)id[b] * NE;
    for (int k = 0; k < NE; k++) {
        atomicAdd(&G[L.w1e + k * H + j], e[k] * d);
        part[k][j] = P[L.w1e + k * H + j] * d;
    }
    __syncthreads();
    for (int w = H / 2; w > 0; w /= 2) {
        if (j < w) {
            for (int k = 0; k < NE; k++) {
                part[k][j] += part[k][j + w];
            }
        }
        __syncthreads();
    }
    if (j < NE) {
        atomicAdd(&G[L.emb + (int64_t)id[b] * NE + j], part[j][0]);
    }
}

__global__ void mlp_fill(float* x, int64_t n, float v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        x[i] = v;
    }
}

__global__ void mlp_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr,
                         float b1, float b2, float eps, float c1, float c2) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    float gi = g[i];
    float mi = b1 * m[i] + (1.f - b1) * gi;
    float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

struct MlpError {
    int code;
    std::string msg;
};

void hcheck(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        throw MlpError {LDO_E_HIP, std::string(what) + ": " + hipGetErrorString(e)};
    }
}

struct MlpCtx {
    int n_ident = 0;
    int cap = 0;
    ldo_mlp_layout L {};
    int64_t adam_t = 0;
    float* d_p = nullptr;
    float* d_g = nullptr;
    float* d_m = nullptr;
    float* d_v = nullptr;
    uint8_t* d_ch = nullptr;
    int32_t* d_id = nullptr;
    uint8_t* d_legal = nullptr;
    int32_t* d_tgt = nullptr;
    float *z1 = nullptr, *h1 = nullptr, *z2 = nullptr, *h2 = nullptr, *z3 = nullptr, *h3 = nullptr;
    float *z4 = nullptr, *dz4 = nullptr, *dh = nullptr, *dz = nullptr, *logp = nullptr;
    double* d_loss = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
};

std::string g_mlp_create_error;

template <class F>
int mguarded(MlpCtx* c, F f) {
    std::string& err = c ? c->err : g_mlp_create_error;
    try {
        f();
        err.clear();
        return LDO_OK;
    } catch (MlpError& e) {
        err = e.msg;
        return e.code;
    } catch (std::exception& e) {
        err = e.what();
        return LDO_E_HIP;
    }
}

template <class T>
void dalloc(T*& p, size_t n, const char* what) {
    hcheck(hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T)), what);
}

void free_mlp(MlpCtx* c) {
    void* ptrs[] = {c->d_p, c->d_g, c->d_m, c->d_v, c->d_ch, c->d_id, c->d_legal, c->d_tgt,
                    c->z1, c->h1, c->z2, c->h2, c->z3, c->h3, c->z4, c->dz4, c->dh, c->dz,
                    c->logp, c->d_loss};
    for (void* p: ptrs) {
        if (p) {
            (void)hipFree(p);
        }
    }
    if (c->stream) {
        (void)hipStreamDestroy(c->stream);
    }
    delete c;
}

void fill(MlpCtx& c, float* x, int64_t n, float v) {
    hipLaunchKernelGGL(mlp_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, x, n,
                       v);
}

void gemm(MlpCtx& c, int M, int N, int K, const float* A, int64_t as0, int64_t as1, const float* B,
          int64_t bs0, int64_t bs1, float* C, int64_t cs0, int accumulate) {
    GemmArgs g {M, N, K, A, as0, as1, B, bs0, bs1, C, cs0, accumulate};
    hipLaunchKernelGGL(mlp_gemm_kernel, dim3((N + GT - 1) / GT, (M + GT - 1) / GT), dim3(256), 0,
                       c.stream, g);
}

void bias_act(MlpCtx& c, float* z, const float* bias, float* h, int M, int N, int relu) {
    int n = M * N;
    hipLaunchKernelGGL(mlp_bias_act, dim3((n + 255) / 256), dim3(256), 0, c.stream, z, bias, h, M,
                       N, relu);
}

// validates a batch and copies it to the device
void upload_batch(MlpCtx& c, int batch, const uint8_t* ch, const int32_t* id, const uint8_t* legal,
                  const int32_t* tgt) {
    if (batch < 1 || batch > c.cap || !ch || !id || !legal) {
        throw MlpError {LDO_E_ARG, "batch must be 1..max_batch with channels, ident and legal"};
    }
    for (int64_t i = 0; i < (int64_t)batch * NS; i++) {
        if (ch[i] >= LDO_MLP_CHANNELS) {
            throw MlpError {LDO_E_ARG, "channel out of range"};
        }
    }
    for (int b = 0; b < batch; b++) {
        if (id[b] < 0 || id[b] >= c.n_ident) {
            throw MlpError {LDO_E_ARG, "identity index out of range"};
        }
        int n_legal = 0;
        for (int k = 0; k < NO; k++) {
            n_legal += legal[b * NO + k] != 0;
        }
        if (n_legal == 0) {
            throw MlpError {LDO_E_ARG, "a sample has no legal output"};
        }
        if (tgt && (tgt[b] < 0 || tgt[b] >= NO || !legal[b * NO + tgt[b]])) {
            throw MlpError {LDO_E_ARG, "target is not a legal output"};
        }
    }
    hcheck(hipMemcpyAsync(c.d_ch, ch, (size_t)batch * NS, hipMemcpyHostToDevice, c.stream),
           "copy channels");
    hcheck(hipMemcpyAsync(c.d_id, id, (size_t)batch * 4, hipMemcpyHostToDevice, c.stream),
           "copy ident");
    hcheck(hipMemcpyAsync(c.d_legal, legal, (size_t)batch * NO, hipMemcpyHostToDevice, c.stream),
           "copy legal");
    if (tgt) {
        hcheck(hipMemcpyAsync(c.d_tgt, tgt, (size_t)batch * 4, hipMemcpyHostToDevice, c.stream),
               "copy target");
    }
}

void forward(MlpCtx& c, int batch) {
    const float* P = c.d_p;
    const ldo_mlp_layout& L = c.L;
    hipLaunchKernelGGL(mlp_layer1_fwd, dim3(batch), dim3(H), 0, c.stream, P, L, c.d_ch, c.d_id,
                       batch, c.z1, c.h1);
    gemm(c, batch, H, H, c.h1, H, 1, P + L.w2, H, 1, c.z2, H, 0);
    bias_act(c, c.z2, P + L.b2, c.h2, batch, H, 1);
    gemm(c, batch, H, H, c.h2, H, 1, P + L.w3, H, 1, c.z3, H, 0);
    bias_act(c, c.z3, P + L.b3, c.h3, batch, H, 1);
    gemm(c, batch, NO, H, c.h3, H, 1, P + L.w4, NO, 1, c.z4, NO, 0);
    bias_act(c, c.z4, P + L.b4, nullptr, batch, NO, 0);
}

void relu_back(MlpCtx& c, const float* dh, const float* z, float* dz, int n) {
    hipLaunchKernelGGL(mlp_relu_back, dim3((n + 255) / 256), dim3(256), 0, c.stream, dh, z, dz, n);
}

void colsum(MlpCtx& c, const float* x, int M, int N, float* out) {
    hipLaunchKernelGGL(mlp_colsum, dim3((N + 63) / 64), dim3(64), 0, c.stream, x, M, N, out);
}

void backward(MlpCtx& c, int batch) {
    const float* P = c.d_p;
    float* G = c.d_g;
    const ldo_mlp_layout& L = c.L;
    fill(c, G, L.total, 0.f);
    // output layer
    gemm(c, H, NO, batch, c.h3, 1, H, c.dz4, NO, 1, G + L.w4, NO, 0);
    colsum(c, c.dz4, batch, NO, G + L.b4);
    gemm(c, batch, H, NO, c.dz4, NO, 1, P + L.w4, 1, NO, c.dh, H, 0);
    relu_back(c, c.dh, c.z3, c.dz, batch * H);
    // third hidden layer
    gemm(c, H, H, batch, c.h2, 1, H, c.dz, H, 1, G + L.w3, H, 0);
    colsum(c, c.dz, batch, H, G + L.b3);
    gemm(c, batch, H, H, c.dz, H, 1, P + L.w3, 1, H, c.dh, H, 0);
    relu_back(c, c.dh, c.z2, c.dz, batch * H);
    // second hidden layer
    gemm(c, H, H, batch, c.h1, 1, H, c.dz, H, 1, G + L.w2, H, 0);
    colsum(c, c.dz, batch, H, G + L.b2);
    gemm(c, batch, H, H, c.dz, H, 1, P + L.w2, 1, H, c.dh, H, 0);
    relu_back(c, c.dh, c.z1, c.dz, batch * H);
    // first layer
    colsum(c, c.dz, batch, H, G + L.b1);
    hipLaunchKernelGGL(mlp_layer1_bwd, dim3(batch), dim3(H), 0, c.stream, P, L, c.d_ch, c.d_id,
                       batch, c.dz, G);
}

} // namespace

extern "C" {

int ldo_mlp_layout_of(int32_t n_ident, ldo_mlp_layout* out) {
    if (n_ident < 1 || !out) {
        return LDO_E_ARG;
    }
    ldo::mlp_layout(n_ident, *out);
    return LDO_OK;
}

int ldo_mlp_create(int32_t n_ident, int32_t max_batch, const float* params, uint64_t seed,
                   void** out) {
    MlpCtx* c = new MlpCtx;
    int rc = mguarded(nullptr, [&] {
        if (n_ident < 1 || max_batch < 1 || max_batch > 65536 || !out) {
            throw MlpError {LDO_E_ARG, "n_ident >= 1, 1 <= max_batch <= 65536 and out required"};
        }
        c->n_ident = n_ident;
        c->cap = max_batch;
        ldo::mlp_layout(n_ident, c->L);
        hcheck(hipStreamCreate(&c->stream), "stream");
        int64_t n = c->L.total;
        dalloc(c->d_p, n, "params");
        dalloc(c->d_g, n, "grads");
        dalloc(c->d_m, n, "adam m");
        dalloc(c->d_v, n, "adam v");
        size_t bh = (size_t)max_batch * H, bo = (size_t)max_batch * NO;
        dalloc(c->d_ch, (size_t)max_batch * NS, "channels");
        dalloc(c->d_id, max_batch, "ident");
        dalloc(c->d_legal, bo, "legal");
        dalloc(c->d_tgt, max_batch, "target");
        for (float** p: {&c->z1, &c->h1, &c->z2, &c->h2, &c->z3, &c->h3, &c->dh, &c->dz}) {
            dalloc(*p, bh, "activations");
        }
        dalloc(c->z4, bo, "logits");
        dalloc(c->dz4, bo, "logit grads");
        dalloc(c->logp, bo, "log-probs");
        dalloc(c->d_loss, max_batch, "loss");
        std::vector<float> h(n, 0.f);
        if (params) {
            std::memcpy(h.data(), params, n * sizeof(float));
        }
        else {
            std::mt19937_64 g(seed);
            auto uni = [&](int64_t off, int64_t cnt, float lim) {
                std::uniform_real_distribution<float> u(-lim, lim);
                for (int64_t i = 0; i < cnt; i++) {
                    h[off + i] = u(g);
                }
            };
            const ldo_mlp_layout& L = c->L;
            uni(L.w1, (int64_t)LDO_MLP_FEATURES * H, std::sqrt(6.f / NS));
            uni(L.emb, (int64_t)n_ident * NE, 1.f);
            uni(L.w1e, (int64_t)NE * H, std::sqrt(6.f / NE));
            uni(L.w2, (int64_t)H * H, std::sqrt(6.f / H));
            uni(L.w3, (int64_t)H * H, std::sqrt(6.f / H));
            uni(L.w4, (int64_t)H * NO, std::sqrt(6.f / H));
        }
        hcheck(hipMemcpy(c->d_p, h.data(), n * sizeof(float), hipMemcpyHostToDevice), "upload");
        fill(*c, c->d_m, n, 0.f);
        fill(*c, c->d_v, n, 0.f);
        fill(*c, c->d_g, n, 0.f);
        hcheck(hipGetLastError(), "init kernels");
        hcheck(hipStreamSynchronize(c->stream), "init sync");
    });
    if (rc != LDO_OK) {
        free_mlp(c);
        return rc;
    }
    *out = c;
    return LDO_OK;
}

void ldo_mlp_destroy(void* mlp) {
    if (mlp) {
        free_mlp(static_cast<MlpCtx*>(mlp));
    }
}

const char* ldo_mlp_last_error(void* mlp) {
    return mlp ? static_cast<MlpCtx*>(mlp)->err.c_str() : g_mlp_create_error.c_str();
}

int ldo_mlp_get_params(void* mlp, float* out) {
    MlpCtx* c = static_cast<MlpCtx*>(mlp);
    return mguarded(c, [&] {
        if (!out) {
            throw MlpError {LDO_E_ARG, "out required"};
        }
        hcheck(hipStreamSynchronize(c->stream), "sync");
        hcheck(hipMemcpy(out, c->d_p, c->L.total * sizeof(float), hipMemcpyDeviceToHost),
               "download params");
    });
}

int ldo_mlp_set_params(void* mlp, const float* in) {
    MlpCtx* c = static_cast<MlpCtx*>(mlp);
    return mguarded(c, [&] {
        if (!in) {
            throw MlpError {LDO_E_ARG, "in required"};
        }
        hcheck(hipStreamSynchronize(c->stream), "sync");
        hcheck(hipMemcpy(c->d_p, in, c->L.total * sizeof(float), hipMemcpyHostToDevice),
               "upload params");
    });
}

int ldo_mlp_loss_grad(void* mlp, int32_t batch, const uint8_t* channels, const int32_t* ident,
                      const uint8_t* legal, const int32_t* target, double* loss, float* grads) {
    MlpCtx* c = static_cast<MlpCtx*>(mlp);
    return mguarded(c, [&] {
        if (!target || !loss) {
            throw MlpError {LDO_E_ARG, "target and loss required"};
        }
        upload_batch(*c, batch, channels, ident, legal, target);
        forward(*c, batch);
        hipLaunchKernelGGL(mlp_softmax_ce, dim3((batch + 63) / 64), dim3(64), 0, c->stream, c->z4,
                           c->d_legal, c->d_tgt, batch, c->dz4, c->d_loss, nullptr);
        backward(*c, batch);
        hcheck(hipGetLastError(), "loss/grad kernels");
        std::vector<double> l(batch);
        hcheck(hipMemcpyAsync(l.data(), c->d_loss, batch * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream),
               "download loss");
        if (grads) {
            hcheck(hipMemcpyAsync(grads, c->d_g, c->L.total * sizeof(float),
                                  hipMemcpyDeviceToHost, c->stream),
                   "download grads");
        }
        hcheck(hipStreamSynchronize(c->stream), "loss/grad sync");
        double s = 0;
        for (double v: l) {
            s += v;
        }
        *loss = s / batch;
    });
}

int ldo_mlp_adam_step(void* mlp, float lr, float beta1, float beta2, float eps) {
    MlpCtx* c = static_cast<MlpCtx*>(mlp);
    return mguarded(c, [&] {
        c->adam_t++;
        float c1 = 1.f - std::pow(beta1, (float)c->adam_t);
        float c2 = 1.f - std::pow(beta2, (float)c->adam_t);
        int64_t n = c->L.total;
        hipLaunchKernelGGL(mlp_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                           c->d_p, c->d_g, c->d_m, c->d_v, n, lr, beta1, beta2, eps, c1, c2);
        hcheck(hipGetLastError(), "adam");
        hcheck(hipStreamSynchronize(c->stream), "adam sync");
    });
}

int ldo_mlp_log_probs(void* mlp, int32_t batch, const uint8_t* channels, const int32_t* ident,
                      const uint8_t* legal, float* out) {
    MlpCtx* c = static_cast<MlpCtx*>(mlp);
    return mguarded(c, [&] {
        if (!out) {
            throw MlpError {LDO_E_ARG, "out required"};
        }
        upload_batch(*c, batch, channels, ident, legal, nullptr);
        forward(*c, batch);
        hipLaunchKernelGGL(mlp_softmax_ce, dim3((batch + 63) / 64), dim3(64), 0, c->stream, c->z4,
                           c->d_legal, nullptr, batch, nullptr, nullptr, c->logp);
        hcheck(hipGetLastError(), "forward kernels");
        hcheck(hipMemcpyAsync(out, c->logp, (size_t)batch * NO * sizeof(float),
                              hipMemcpyDeviceToHost, c->stream),
               "download log-probs");
        hcheck(hipStreamSynchronize(c->stream), "forward sync");
    });
}

} // extern "C"
