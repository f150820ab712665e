// Per atom-pair density network of the learned SCF guess, on the gfx950 matrix
// cores: n_in -> n_hidden -> n_hidden -> n_out with SiLU, BF16 weights and
// activations, FP32 accumulation, FP32 master weights and Adam.
//
// No reference counterpart: the reference starts every SCF from the SAD
// guess (src/pymodule/sadguessdriver.py:841-971); this network predicts the
// starting density block by block, and SAD stays the fallback.
//
// GEMM: C[M][N] = sum_k op(A)[m][k] op(B)[k][n], op(A) = A (M x K row-major)
// or A^T (A stored K x M), op(B) = B (K x N) or B^T (B stored N x K).  A
// 256-thread block owns a 64 x 64 tile of C; each wavefront a 32 x 32
// quarter, as 2 x 2 v_mfma_f32_16x16x32_bf16 tiles.  K advances by 32: the
// block stages op(A) as [m][k] and op(B) as [n][k] (k contiguous) in LDS, so
// that every lane reads its 8 consecutive k of one row or column with one
// 16-byte LDS load.  Out-of-range rows, columns and k are staged as zeros.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/vlxfock.h"
#include "../../include/vlxguess.h"
#include "vlx_types.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

namespace {

// BF16 as its 16 bits; conversions round to nearest even (finite inputs)
__host__ __device__ __forceinline__ uint16_t f2bf(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
[[maybe_unused]] __host__ __device__ __forceinline__ float bf2f(uint16_t h)
{
    const uint32_t u = static_cast<uint32_t>(h) << 16;
    float x;
    std::memcpy(&x, &u, 4);
    return x;
}

__device__ __forceinline__ float sigmoid(float x) { return 1.0f / (1.0f + __expf(-x)); }

constexpr int TILE = 64;
constexpr int KT = 32;
constexpr int LDS_ROW = KT + 8;  // bf16 per staged row (padding against bank conflicts)

enum Epilogue {
    EPI_STORE = 0,       // C (fp32)
    EPI_BIAS = 1,        // C + bias (fp32)
    EPI_BIAS_SILU = 2,   // H = C + bias (fp32) and silu(H) (bf16)
    EPI_SILU_GRAD = 3    // dH = C silu'(H) with H given (fp32 and bf16)
};

struct GemmArgs {
    const uint16_t* A;
    const uint16_t* B;
    int M, N, K;
    const float* bias;   // EPI_BIAS, EPI_BIAS_SILU
    const float* H;      // EPI_SILU_GRAD: pre-activations, M x N
    float* out_f;        // M x N fp32
    uint16_t* out_b;     // M x N bf16 (EPI_BIAS_SILU, EPI_SILU_GRAD)
};

template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) mlp_gemm_kernel(const GemmArgs g)
{
    __shared__ __attribute__((aligned(16))) uint16_t sa[TILE * LDS_ROW];
    __shared__ __attribute__((aligned(16))) uint16_t sb[TILE * LDS_ROW];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int m0 = blockIdx.y * TILE, n0 = blockIdx.x * TILE;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;  // the wave's quarter of the tile
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < g.K; k0 += KT) {
        // stage op(A)[m0 .. m0+63][k0 .. k0+31] as sa[m][k] and op(B)[k][n0 ..] as sb[n][k]:
        // 2048 elements each, 8 per thread
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int e = tid + 256 * t;
            int r, c;  // r: tile row (m or n), c: k within the step
            if (TA) { c = e >> 6; r = e & 63; }   // A stored K x M: consecutive threads along m
            else { r = e >> 5; c = e & 31; }      // A stored M x K: consecutive threads along k
            const int gm = m0 + r, gk = k0 + c;
            uint16_t v = 0;
            if (gm < g.M && gk < g.K) v = TA ? g.A[(size_t)gk * g.M + gm] : g.A[(size_t)gm * g.K + gk];
            sa[r * LDS_ROW + c] = v;
            int rb, cb;
            if (TB) { rb = e >> 5; cb = e & 31; }  // B stored N x K
            else { cb = e >> 6; rb = e & 63; }     // B stored K x N
            const int gn = n0 + rb, gkb = k0 + cb;
            uint16_t w = 0;
            if (gn < g.N && gkb < g.K) w = TB ? g.B[(size_t)gn * g.K + gkb] : g.B[(size_t)gkb * g.N + gn];
            sb[rb * LDS_ROW + cb] = w;
        }
        __syncthreads();
        // lane l: row (or column) l % 16 of a 16-wide tile, k = 8 (l / 16) .. + 7
        const int fr = lane & 15, fk = (lane >> 4) * 8;
        bf16x8 af[2], bfr[2];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            af[i] = *reinterpret_cast<const bf16x8*>(&sa[(wm + 16 * i + fr) * LDS_ROW + fk]);
            bfr[i] = *reinterpret_cast<const bf16x8*>(&sb[(wn + 16 * i + fr) * LDS_ROW + fk]);
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    // D of a 16 x 16 tile: lane l holds rows 4 (l / 16) + r, r = 0..3, of column l % 16
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int gm = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
                const int gn = n0 + wn + 16 * j + (lane & 15);
                if (gm >= g.M || gn >= g.N) continue;
                const size_t o = (size_t)gm * g.N + gn;
                float v = acc[i][j][r];
                if (EPI == EPI_STORE) {
                    g.out_f[o] = v;
                } else if (EPI == EPI_BIAS) {
                    g.out_f[o] = v + g.bias[gn];
                } else if (EPI == EPI_BIAS_SILU) {
                    v += g.bias[gn];
                    g.out_f[o] = v;
                    g.out_b[o] = f2bf(v * sigmoid(v));
                } else {
                    const float h = g.H[o], s = sigmoid(h);
                    const float d = v * s * (1.0f + h * (1.0f - s));
                    g.out_f[o] = d;
                    g.out_b[o] = f2bf(d);
                }
            }
}

__global__ void to_bf16_kernel(const float* __restrict__ x, uint16_t* __restrict__ y, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = f2bf(x[i]);
}

// Per-sample mean squared error over the valid entries of a block (rows < ni, columns < nj of
// a block_cols-wide flattening), averaged over the m samples:
//   loss = (1/m) sum_s (1/(ni nj)) sum_valid (y - t)^2,  dY = 2 (y - t) / (m ni nj) (0 elsewhere).
__global__ void loss_grad_kernel(const float* __restrict__ y, const float* __restrict__ t,
                                 const int2* __restrict__ shape, int m, int n_out, int block_cols,
                                 float* __restrict__ dy_f, uint16_t* __restrict__ dy_b, float* __restrict__ loss)
{
    float part = 0.f;
    const size_t total = (size_t)m * n_out;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int s = static_cast<int>(i / n_out), o = static_cast<int>(i % n_out);
        const int2 sh = shape[s];
        float d = 0.f;
        if (o / block_cols < sh.x && o % block_cols < sh.y) {
            const float w = 1.0f / (static_cast<float>(m) * sh.x * sh.y);
            const float r = y[i] - t[i];
            part += r * r * w;
            d = 2.0f * r * w;
        }
        dy_f[i] = d;
        dy_b[i] = f2bf(d);
    }
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, part);
}

// column sums of an m x n fp32 matrix into out[n] (bias gradients)
__global__ void colsum_kernel(const float* __restrict__ x, int m, int n, float* __restrict__ out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    float s = 0.f;
    for (int r = 0; r < m; r++) s += x[(size_t)r * n + c];
    out[c] = s;
}

// Adam on fp32 master weights, with their bf16 copies refreshed (bias vectors have no copy)
__global__ void adam_kernel(float* __restrict__ w, uint16_t* __restrict__ wb, const float* __restrict__ grad,
                            float* __restrict__ mom, float* __restrict__ vel, size_t n, float lr, float b1, float b2,
                            float eps, float c1, float c2)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float gi = grad[i];
        const float mi = b1 * mom[i] + (1.0f - b1) * gi;
        const float vi = b2 * vel[i] + (1.0f - b2) * gi * gi;
        mom[i] = mi;
        vel[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (wb) wb[i] = f2bf(wi);
    }
}

__global__ void clear_kernel(float* __restrict__ p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

int mfail(int code, const char* msg)
{
    vlx_set_last_error(msg);
    return code;
}

}  // namespace

// Layer l (0..2): W_l (fan_in x fan_out, row-major) and b_l (fan_out).
struct vlx_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    int n_in = 0, n_hidden = 0, n_out = 0, max_batch = 0, block_cols = 1;
    int fan_in[3], fan_out[3];
    size_t woff[3], boff[3], nparams = 0;  // offsets into the packed parameter vector
    float* p = nullptr;      // packed W0, b0, W1, b1, W2, b2 (fp32 master)
    float* grad = nullptr;   // same packing
    float* mom = nullptr;
    float* vel = nullptr;
    uint16_t* wb[3] = {};    // bf16 copies of the weight matrices
    int step = 0;
    // activations of one batch
    uint16_t *xb = nullptr, *a1b = nullptr, *a2b = nullptr, *dyb = nullptr, *dh2b = nullptr, *dh1b = nullptr;
    float *h1 = nullptr, *h2 = nullptr, *dyf = nullptr, *dh2f = nullptr, *dh1f = nullptr, *loss = nullptr;
};

namespace {

template <typename T>
hipError_t dalloc(T** p, size_t n)
{
    return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(n, 1) * sizeof(T));
}

#define MLP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) return mfail(VLX_ERR_HIP, hipGetErrorString(e_));         \
    } while (0)

template <bool TA, bool TB, int EPI>
int gemm(hipStream_t s, const GemmArgs& g)
{
    if (g.M <= 0 || g.N <= 0 || g.K <= 0) return VLX_OK;
    dim3 grid((g.N + TILE - 1) / TILE, (g.M + TILE - 1) / TILE);
    hipLaunchKernelGGL((mlp_gemm_kernel<TA, TB, EPI>), grid, dim3(256), 0, s, g);
    MLP_TRY(hipGetLastError());
    return VLX_OK;
}

int grid_for(size_t n) { return static_cast<int>(std::min<size_t>((n + 255) / 256, 4096)); }

int forward(vlx_mlp* net, const float* d_x, int m, float* d_y)
{
    hipStream_t s = net->stream;
    to_bf16_kernel<<<grid_for((size_t)m * net->n_in), 256, 0, s>>>(d_x, net->xb, (size_t)m * net->n_in);
    MLP_TRY(hipGetLastError());
    GemmArgs g{};
    g = {net->xb, net->wb[0], m, net->n_hidden, net->n_in, net->p + net->boff[0], nullptr, net->h1, net->a1b};
    int rc = gemm<false, false, EPI_BIAS_SILU>(s, g);
    if (rc) return rc;
    g = {net->a1b, net->wb[1], m, net->n_hidden, net->n_hidden, net->p + net->boff[1], nullptr, net->h2, net->a2b};
    rc = gemm<false, false, EPI_BIAS_SILU>(s, g);
    if (rc) return rc;
    g = {net->a2b, net->wb[2], m, net->n_out, net->n_hidden, net->p + net->boff[2], nullptr, d_y, nullptr};
    return gemm<false, false, EPI_BIAS>(s, g);
}

}  // namespace

extern "C" {

int vlx_mlp_create(vlx_context* ctx, int n_in, int n_hidden, int n_out, int block_cols, int max_batch, vlx_mlp** out)
{
    if (!ctx || !out) return mfail(VLX_ERR_ARG, "NULL argument to vlx_mlp_create");
    if (n_in <= 0 || n_hidden <= 0 || n_out <= 0 || max_batch <= 0 || block_cols <= 0)
        return mfail(VLX_ERR_ARG, "vlx_mlp_create: sizes must be positive");
    auto net = new vlx_mlp();
    net->stream = static_cast<hipStream_t>(vlx_context_stream(ctx));
    MLP_TRY(hipGetDevice(&net->device));
    net->n_in = n_in;
    net->n_hidden = n_hidden;
    net->n_out = n_out;
    net->block_cols = block_cols;
    net->max_batch = max_batch;
    const int fi[3] = {n_in, n_hidden, n_hidden}, fo[3] = {n_hidden, n_hidden, n_out};
    size_t off = 0;
    for (int l = 0; l < 3; l++) {
        net->fan_in[l] = fi[l];
        net->fan_out[l] = fo[l];
        net->woff[l] = off;
        off += (size_t)fi[l] * fo[l];
        net->boff[l] = off;
        off += fo[l];
    }
    net->nparams = off;
    const size_t mb = static_cast<size_t>(max_batch);
    hipError_t e = hipSuccess;
    for (float** q : {&net->p, &net->grad, &net->mom, &net->vel})
        if (e == hipSuccess) e = dalloc(q, off);
    for (int l = 0; l < 3 && e == hipSuccess; l++) e = dalloc(&net->wb[l], (size_t)fi[l] * fo[l]);
    if (e == hipSuccess) e = dalloc(&net->xb, mb * n_in);
    for (uint16_t** q : {&net->a1b, &net->a2b, &net->dh2b, &net->dh1b})
        if (e == hipSuccess) e = dalloc(q, mb * n_hidden);
    if (e == hipSuccess) e = dalloc(&net->dyb, mb * n_out);
    for (float** q : {&net->h1, &net->h2, &net->dh2f, &net->dh1f})
        if (e == hipSuccess) e = dalloc(q, mb * n_hidden);
    if (e == hipSuccess) e = dalloc(&net->dyf, mb * n_out);
    if (e == hipSuccess) e = dalloc(&net->loss, 1);
    if (e != hipSuccess) {
        vlx_mlp_destroy(net);
        return mfail(VLX_ERR_NOMEM, "vlx_mlp_create: device allocation failed");
    }
    *out = net;
    return VLX_OK;
}

void vlx_mlp_destroy(vlx_mlp* net)
{
    if (!net) return;
    if (net->stream) (void)hipStreamSynchronize(net->stream);
    for (float* q : {net->p, net->grad, net->mom, net->vel, net->h1, net->h2, net->dyf, net->dh2f, net->dh1f, net->loss})
        if (q) (void)hipFree(q);
    for (uint16_t* q : {net->wb[0], net->wb[1], net->wb[2], net->xb, net->a1b, net->a2b, net->dyb, net->dh2b, net->dh1b})
        if (q) (void)hipFree(q);
    delete net;
}

int vlx_mlp_num_params(vlx_mlp* net, long long* n)
{
    if (!net || !n) return mfail(VLX_ERR_ARG, "NULL argument");
    *n = static_cast<long long>(net->nparams);
    return VLX_OK;
}

int vlx_mlp_set_params(vlx_mlp* net, const float* params)
{
    if (!net || !params) return mfail(VLX_ERR_ARG, "NULL argument");
    MLP_TRY(hipMemcpyAsync(net->p, params, net->nparams * sizeof(float), hipMemcpyHostToDevice, net->stream));
    for (int l = 0; l < 3; l++) {
        const size_t n = (size_t)net->fan_in[l] * net->fan_out[l];
        to_bf16_kernel<<<grid_for(n), 256, 0, net->stream>>>(net->p + net->woff[l], net->wb[l], n);
        MLP_TRY(hipGetLastError());
    }
    // a fresh parameter set restarts the optimiser
    clear_kernel<<<(int)((net->nparams + 255) / 256), 256, 0, net->stream>>>(net->mom, (int)net->nparams);
    clear_kernel<<<(int)((net->nparams + 255) / 256), 256, 0, net->stream>>>(net->vel, (int)net->nparams);
    MLP_TRY(hipGetLastError());
    net->step = 0;
    MLP_TRY(hipStreamSynchronize(net->stream));
    return VLX_OK;
}

int vlx_mlp_get_params(vlx_mlp* net, float* params)
{
    if (!net || !params) return mfail(VLX_ERR_ARG, "NULL argument");
    MLP_TRY(hipMemcpyAsync(params, net->p, net->nparams * sizeof(float), hipMemcpyDeviceToHost, net->stream));
    MLP_TRY(hipStreamSynchronize(net->stream));
    return VLX_OK;
}

int vlx_mlp_forward(vlx_mlp* net, const float* d_x, int m, float* d_y)
{
    if (!net || !d_x || !d_y) return mfail(VLX_ERR_ARG, "NULL argument");
    if (m < 0 || m > net->max_batch) return mfail(VLX_ERR_ARG, "vlx_mlp_forward: batch outside 0..max_batch");
    return forward(net, d_x, m, d_y);
}

int vlx_mlp_loss_grad(vlx_mlp* net, const float* d_x, const float* d_target, const int* d_shape, int m, float* d_y,
                      float* loss, float* grads)
{
    if (!net || !d_x || !d_target || !d_shape || !d_y || !loss) return mfail(VLX_ERR_ARG, "NULL argument");
    if (m <= 0 || m > net->max_batch) return mfail(VLX_ERR_ARG, "vlx_mlp_loss_grad: batch outside 1..max_batch");
    hipStream_t s = net->stream;
    int rc = forward(net, d_x, m, d_y);
    if (rc) return rc;
    clear_kernel<<<1, 64, 0, s>>>(net->loss, 1);
    const size_t ny = (size_t)m * net->n_out;
    loss_grad_kernel<<<grid_for(ny), 256, 0, s>>>(d_y, d_target, reinterpret_cast<const int2*>(d_shape), m, net->n_out,
                                                   net->block_cols, net->dyf, net->dyb, net->loss);
    MLP_TRY(hipGetLastError());
    const int H = net->n_hidden;
    float* gw[3] = {net->grad + net->woff[0], net->grad + net->woff[1], net->grad + net->woff[2]};
    float* gb[3] = {net->grad + net->boff[0], net->grad + net->boff[1], net->grad + net->boff[2]};
    GemmArgs g{};
    // layer 3: dW2 = A2^T dY, db2 = colsum dY, dH2 = (dY W2^T) silu'(H2)
    g = {net->a2b, net->dyb, H, net->n_out, m, nullptr, nullptr, gw[2], nullptr};
    if ((rc = gemm<true, false, EPI_STORE>(s, g))) return rc;
    colsum_kernel<<<(net->n_out + 255) / 256, 256, 0, s>>>(net->dyf, m, net->n_out, gb[2]);
    g = {net->dyb, net->wb[2], m, H, net->n_out, nullptr, net->h2, net->dh2f, net->dh2b};
    if ((rc = gemm<false, true, EPI_SILU_GRAD>(s, g))) return rc;
    // layer 2: dW1 = A1^T dH2, db1, dH1 = (dH2 W1^T) silu'(H1)
    g = {net->a1b, net->dh2b, H, H, m, nullptr, nullptr, gw[1], nullptr};
    if ((rc = gemm<true, false, EPI_STORE>(s, g))) return rc;
    colsum_kernel<<<(H + 255) / 256, 256, 0, s>>>(net->dh2f, m, H, gb[1]);
    g = {net->dh2b, net->wb[1], m, H, H, nullptr, net->h1, net->dh1f, net->dh1b};
    if ((rc = gemm<false, true, EPI_SILU_GRAD>(s, g))) return rc;
    // layer 1: dW0 = X^T dH1, db0
    g = {net->xb, net->dh1b, net->n_in, H, m, nullptr, nullptr, gw[0], nullptr};
    if ((rc = gemm<true, false, EPI_STORE>(s, g))) return rc;
    colsum_kernel<<<(H + 255) / 256, 256, 0, s>>>(net->dh1f, m, H, gb[0]);
    MLP_TRY(hipGetLastError());
    MLP_TRY(hipMemcpyAsync(loss, net->loss, sizeof(float), hipMemcpyDeviceToHost, s));
    if (grads) MLP_TRY(hipMemcpyAsync(grads, net->grad, net->nparams * sizeof(float), hipMemcpyDeviceToHost, s));
    MLP_TRY(hipStreamSynchronize(s));
    return VLX_OK;
}

int vlx_mlp_adam_step(vlx_mlp* net, float lr, float beta1, float beta2, float eps)
{
    if (!net) return mfail(VLX_ERR_ARG, "NULL argument");
    net->step++;
    const float c1 = 1.0f - std::pow(beta1, static_cast<float>(net->step));
    const float c2 = 1.0f - std::pow(beta2, static_cast<float>(net->step));
    for (int l = 0; l < 3; l++) {
        const size_t nw = (size_t)net->fan_in[l] * net->fan_out[l];
        adam_kernel<<<grid_for(nw), 256, 0, net->stream>>>(net->p + net->woff[l], net->wb[l], net->grad + net->woff[l],
                                                          net->mom + net->woff[l], net->vel + net->woff[l], nw, lr,
                                                          beta1, beta2, eps, c1, c2);
        const size_t nb = net->fan_out[l];
        adam_kernel<<<grid_for(nb), 256, 0, net->stream>>>(net->p + net->boff[l], nullptr, net->grad + net->boff[l],
                                                          net->mom + net->boff[l], net->vel + net->boff[l], nb, lr,
                                                          beta1, beta2, eps, c1, c2);
    }
    MLP_TRY(hipGetLastError());
    return VLX_OK;
}

}  // extern "C"
