// The learned warm start of multi-direct forcing: a 17 -> 64 -> 64 -> 2 GELU
// network that predicts the sum of a node's first plain MDF corrections -- as
// two gains on the first one -- from its slip velocity, its membrane + contact
// force, the local curvature, the distances to other capsules and walls, and
// its slips at the previous steps.  The reference has no counterpart; its
// only network (ml/simple_nn.cpp) is not on the stepping path.
//
// Everything runs on the device in FP32: inference inside the step, and
// training on sample rows the plain solver wrote (forward, backward, Adam).
// Gradient sums are deterministic: each 64-sample workgroup writes its own
// partial gradient and one kernel adds the partials in a fixed order.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <vector>

#include "sf_internal.h"

namespace sf {

constexpr int W1_OFF = 0;
constexpr int B1_OFF = W1_OFF + MLP_W * MLP_IN;
constexpr int W2_OFF = B1_OFF + MLP_W;
constexpr int B2_OFF = W2_OFF + MLP_W * MLP_W;
constexpr int W3_OFF = B2_OFF + MLP_W;
constexpr int B3_OFF = W3_OFF + MLP_OUT * MLP_W;
static_assert(B3_OFF + MLP_OUT == MLP_P, "parameter layout");

constexpr int ROW = MLP_IN + MLP_OUT;     // a sample row: inputs, then targets
constexpr int GB = 64;                    // samples per workgroup (one wavefront)
constexpr int LDP = MLP_W + 1;            // padded LDS row

__device__ __forceinline__ float gelu(float z) {
    return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_d(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
           z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// The network on the matrix cores.  A wavefront evaluates 32 nodes at once:
// each layer is a [32 x K] x [K x 64] product done as two 32x32 output tiles
// of v_mfma_f32_32x32x2f32 (K in steps of 2), with the weights staged once
// per workgroup in LDS and the activations passed between layers through
// LDS.  Operand layout of that instruction (wave64): lane l supplies
// A[l % 32][l / 32] and B[l / 32][l % 32]; accumulator r of lane l holds
// D[r % 4 + 8 (r / 4) + 4 (l / 32)][l % 32].  Rows of odd length keep the
// LDS accesses of a 32-lane half free of bank conflicts.  The sums run in the
// matrix cores' order; the training forward (k_mlp_grad) runs them in index
// order, so the two agree to FP32 rounding.
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int NODES_PER_WAVE = 32;
constexpr int X_LD = 19;                 // 17 normalised inputs + zero padding
constexpr int W1_LD = 19;
constexpr int H_LD = MLP_W + 1;
constexpr int MLP_WAVES = 4;             // wavefronts per workgroup
constexpr int MLP_NODES_WG = MLP_WAVES * NODES_PER_WAVE;

struct MlpStage {
    float w1[MLP_W * W1_LD], b1[MLP_W], w2[MLP_W * H_LD], b2[MLP_W];
    float w3[MLP_OUT * MLP_W], b3[MLP_OUT], norm[MLP_NORM];
};
struct MlpWave {
    float x[NODES_PER_WAVE * X_LD], h1[NODES_PER_WAVE * H_LD], h2[NODES_PER_WAVE * H_LD];
    int bad[NODES_PER_WAVE];
};

// Every thread issues all its loads before its LDS stores (fixed trip counts),
// so a workgroup waits for the weights about once, not once per row.
constexpr int STAGE_THREADS = MLP_WAVES * 64;
constexpr int W1_PER = (MLP_W * MLP_IN + STAGE_THREADS - 1) / STAGE_THREADS;
constexpr int W2_PER = (MLP_W * MLP_W + STAGE_THREADS - 1) / STAGE_THREADS;

__device__ void mlp_stage(MlpStage& s, const float* P, const float* norm) {
    const int t = threadIdx.x;
    float r1[W1_PER], r2[W2_PER];
#pragma unroll
    for (int u = 0; u < W1_PER; ++u) {
        const int i = t + u * STAGE_THREADS;
        r1[u] = i < MLP_W * MLP_IN ? P[W1_OFF + i] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < W2_PER; ++u) {
        const int i = t + u * STAGE_THREADS;
        r2[u] = i < MLP_W * MLP_W ? P[W2_OFF + i] : 0.f;
    }
    float rb1 = 0.f, rb2 = 0.f, rw3 = 0.f, rb3 = 0.f, rn = 0.f;
    if (t < MLP_W) { rb1 = P[B1_OFF + t]; rb2 = P[B2_OFF + t]; }
    if (t < MLP_OUT * MLP_W) rw3 = P[W3_OFF + t];
    if (t < MLP_OUT) rb3 = P[B3_OFF + t];
    if (t < MLP_NORM) rn = norm[t];
#pragma unroll
    for (int u = 0; u < W1_PER; ++u) {
        const int i = t + u * STAGE_THREADS;
        if (i < MLP_W * MLP_IN) s.w1[(i / MLP_IN) * W1_LD + i % MLP_IN] = r1[u];
    }
    if (t < MLP_W)
        for (int k = MLP_IN; k < W1_LD; ++k) s.w1[t * W1_LD + k] = 0.f;   // the padding the inputs meet
#pragma unroll
    for (int u = 0; u < W2_PER; ++u) {
        const int i = t + u * STAGE_THREADS;
        if (i < MLP_W * MLP_W) s.w2[(i / MLP_W) * H_LD + i % MLP_W] = r2[u];
    }
    if (t < MLP_W) { s.b1[t] = rb1; s.b2[t] = rb2; }
    if (t < MLP_OUT * MLP_W) s.w3[t] = rw3;
    if (t < MLP_OUT) s.b3[t] = rb3;
    if (t < MLP_NORM) s.norm[t] = rn;
}

// Inputs of nodes k0 .. k0+31 (rows of x, clipped at n) normalised into the
// wave's LDS; bad[node] is set where an input is not finite.  The caller
// synchronises the workgroup before the values are read.
constexpr int LOAD_PER = (NODES_PER_WAVE * MLP_IN + 63) / 64;

__device__ void mlp_load(const MlpStage& s, MlpWave& w, const float* x, int k0, int n) {
    const int l = threadIdx.x & 63;
    if (l < NODES_PER_WAVE) {
        w.bad[l] = 0;
#pragma unroll
        for (int k = MLP_IN; k < X_LD; ++k) w.x[l * X_LD + k] = 0.f;
    }
    const int avail = max(0, min(NODES_PER_WAVE, n - k0)) * MLP_IN;   // contiguous input values
    const float* src = x + static_cast<size_t>(max(k0, 0)) * MLP_IN;
    float v[LOAD_PER];
#pragma unroll
    for (int u = 0; u < LOAD_PER; ++u) {
        const int idx = l + 64 * u;
        v[u] = idx < avail ? src[idx] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < LOAD_PER; ++u) {
        const int idx = l + 64 * u;
        if (idx < NODES_PER_WAVE * MLP_IN) {
            const int node = idx / MLP_IN, f = idx % MLP_IN;
            w.x[node * X_LD + f] = (v[u] - s.norm[f]) / s.norm[MLP_IN + f];
            if (!isfinite(v[u])) w.bad[node] = 1;
        }
    }
}

// The network on the wave's 32 loaded nodes.  Returns, on lane l, output l / 32
// of node l % 32 (before denormalisation).  Every wave of the workgroup calls it.
__device__ float mlp_wave(const MlpStage& s, MlpWave& w) {
    const int l = threadIdx.x & 63, lo = l & 31, kh = l >> 5;
    f32x16 c0 = {}, c1 = {};
    for (int kk = 0; kk < (MLP_IN + 1) / 2; ++kk) {
        const int k = 2 * kk + kh;
        const float a = w.x[lo * X_LD + k];
        c0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, s.w1[lo * W1_LD + k], c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, s.w1[(32 + lo) * W1_LD + k], c1, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = (r & 3) + 8 * (r >> 2) + 4 * kh;
        w.h1[i * H_LD + lo] = gelu(c0[r] + s.b1[lo]);
        w.h1[i * H_LD + 32 + lo] = gelu(c1[r] + s.b1[32 + lo]);
    }
    __syncthreads();
    c0 = f32x16{};
    c1 = f32x16{};
    for (int kk = 0; kk < MLP_W / 2; ++kk) {
        const int k = 2 * kk + kh;
        const float a = w.h1[lo * H_LD + k];
        c0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, s.w2[lo * H_LD + k], c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, s.w2[(32 + lo) * H_LD + k], c1, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = (r & 3) + 8 * (r >> 2) + 4 * kh;
        w.h2[i * H_LD + lo] = gelu(c0[r] + s.b2[lo]);
        w.h2[i * H_LD + 32 + lo] = gelu(c1[r] + s.b2[32 + lo]);
    }
    __syncthreads();
    float out = s.b3[kh];
    for (int k = 0; k < MLP_W; ++k) out += s.w3[kh * MLP_W + k] * w.h2[lo * H_LD + k];
    return out;
}

// Inference inside the step: the network's two outputs are the gains of the
// node's total correction in the frame of its first plain correction dF1,
// T = g_par dF1 + g_perp rot90(dF1).  They go to gain[2k], held to
// g_par in [0, 4] and g_perp in [-2, 2], and default to (1, 0), the plain
// correction, where an input or an output is not finite.  Nothing
// happens when the residual is already within tolerance.
__global__ __launch_bounds__(MLP_WAVES * 64) void k_mlp_infer(const float* __restrict__ P,
                                                              const float* __restrict__ norm,
                                                              const float* feat, double* gain, int K,
                                                              const MdfDev* md, double tol) {
    if (tol > 0.0 && mdf_res(md, 1) <= tol) return;
    __shared__ MlpStage s;
    __shared__ MlpWave waves[MLP_WAVES];
    mlp_stage(s, P, norm);
    __syncthreads();
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int k0 = blockIdx.x * MLP_NODES_WG + wv * NODES_PER_WAVE;
    mlp_load(s, waves[wv], feat, k0, K);
    __syncthreads();
    const bool ok = l < NODES_PER_WAVE && waves[wv].bad[l] == 0;
    const float out = mlp_wave(s, waves[wv]);
    const float out1 = __shfl(out, (l & 31) + 32);
    const int k = k0 + l;
    if (l < NODES_PER_WAVE && k < K) {
        double g0 = 1.0, g1 = 0.0;
        if (ok) {
            const float gp = s.norm[2 * MLP_IN] + s.norm[2 * MLP_IN + MLP_OUT] * out;
            const float gq = s.norm[2 * MLP_IN + 1] + s.norm[2 * MLP_IN + MLP_OUT + 1] * out1;
            if (isfinite(gp) && isfinite(gq)) {
                g0 = fminf(fmaxf(gp, 0.f), 4.f);
                g1 = fminf(fmaxf(gq, -2.f), 2.f);
            }
        }
        gain[2 * k] = g0;
        gain[2 * k + 1] = g1;
    }
}

// The network alone on n inputs, on the matrix cores: y = out_mean + out_scale * net(x^).
__global__ __launch_bounds__(MLP_WAVES * 64) void k_mlp_predict(const float* __restrict__ P,
                                                                const float* __restrict__ norm,
                                                                const float* x, float* y, int n) {
    __shared__ MlpStage s;
    __shared__ MlpWave waves[MLP_WAVES];
    mlp_stage(s, P, norm);
    __syncthreads();
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int k0 = blockIdx.x * MLP_NODES_WG + wv * NODES_PER_WAVE;
    mlp_load(s, waves[wv], x, k0, n);
    __syncthreads();
    const float out = mlp_wave(s, waves[wv]);
    const int k = k0 + (l & 31), o = l >> 5;
    if (k < n) y[static_cast<size_t>(k) * MLP_OUT + o] = s.norm[2 * MLP_IN + o] + s.norm[2 * MLP_IN + MLP_OUT + o] * out;
}

// The gains, low-passed along each membrane with the binomial weights
// (1 4 6 4 1)/16, scale the first plain correction dF1 (smooth, since both
// velocities it compares are interpolated).  Node-by-node prediction errors
// would otherwise alternate between neighbours: that zig-zag barely moves the
// interpolated velocity, so the residual cannot see it, yet it pushes the
// fluid.
__global__ void k_mdf_smooth(const double* gain, NodeSpread* rec, const int* off, const int* owner,
                             const double* x, const double* y, int K, int ny, const MdfDev* md, double tol) {
    if (tol > 0.0 && mdf_res(md, 1) <= tol) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int cap = owner[k];
    const int o = off[cap], n = off[cap + 1] - o, i = k - o;
    double gp = gain[2 * k], gq = gain[2 * k + 1];
    if (n >= 5) {
        const int j[5] = {o + (i - 2 + n) % n, o + (i - 1 + n) % n, k, o + (i + 1) % n, o + (i + 2) % n};
        const double w[5] = {1.0, 4.0, 6.0, 4.0, 1.0};
        gp = 0.0;
        gq = 0.0;
        for (int t = 0; t < 5; ++t) { gp += w[t] * gain[2 * j[t]]; gq += w[t] * gain[2 * j[t] + 1]; }
        gp *= 0.0625;
        gq *= 0.0625;
    }
    const NodeSpread r = rec[k];    // dF1
    // the whole record, valid where the correction is finite
    rec[k] = node_record(x[k], y[k], gp * r.fx - gq * r.fy, gp * r.fy + gq * r.fx, r.ds, k, ny);
}

// Forward + backward of GB samples; the workgroup's gradient of
// L = 1/(2 n) sum_s |out_s - t^_s|^2 goes to partial[blockIdx.x][MLP_P] and its
// summed squared error to partial_loss[blockIdx.x].
// rows: sample rows; idx (optional) picks rows; [first, first + count) is
// this launch's slice of idx (or of the rows).
__global__ __launch_bounds__(GB) void k_mlp_grad(const float* __restrict__ P,
                                                 const float* __restrict__ norm, const float* rows,
                                                 const int* idx, int64_t first, int count,
                                                 float inv_n, float* partial, double* partial_loss) {
    __shared__ float xs[GB][MLP_IN + 1];
    __shared__ float a1s[GB][LDP], a2s[GB][LDP], d1s[GB][LDP], d2s[GB][LDP];
    __shared__ float es[GB][MLP_OUT];
    __shared__ double ls[GB];
    const int s = threadIdx.x;
    const int local = blockIdx.x * GB + s;
    const bool live = local < count;
    float xh[MLP_IN], th[MLP_OUT];
    if (live) {
        const int64_t r = idx ? static_cast<int64_t>(idx[first + local]) : first + local;
        const float* row = rows + r * ROW;
#pragma unroll
        for (int j = 0; j < MLP_IN; ++j) xh[j] = (row[j] - norm[j]) / norm[MLP_IN + j];
#pragma unroll
        for (int o = 0; o < MLP_OUT; ++o)
            th[o] = (row[MLP_IN + o] - norm[2 * MLP_IN + o]) / norm[2 * MLP_IN + MLP_OUT + o];
    } else {
#pragma unroll
        for (int j = 0; j < MLP_IN; ++j) xh[j] = 0.f;
#pragma unroll
        for (int o = 0; o < MLP_OUT; ++o) th[o] = 0.f;
    }
#pragma unroll
    for (int j = 0; j < MLP_IN; ++j) xs[s][j] = xh[j];
    // forward: a1, a2 into this sample's LDS rows; z1 parked in the delta-1 row
    float a1[MLP_W];
#pragma unroll
    for (int i = 0; i < MLP_W; ++i) {
        float acc = P[B1_OFF + i];
#pragma unroll
        for (int j = 0; j < MLP_IN; ++j) acc += P[W1_OFF + i * MLP_IN + j] * xh[j];
        d1s[s][i] = acc;
        a1[i] = gelu(acc);
        a1s[s][i] = a1[i];
    }
    float z2[MLP_W];
    float out[MLP_OUT];
#pragma unroll
    for (int o = 0; o < MLP_OUT; ++o) out[o] = P[B3_OFF + o];
#pragma unroll
    for (int i = 0; i < MLP_W; ++i) {
        float acc = P[B2_OFF + i];
#pragma unroll
        for (int j = 0; j < MLP_W; ++j) acc += P[W2_OFF + i * MLP_W + j] * a1[j];
        z2[i] = acc;
        const float a2 = gelu(acc);
        a2s[s][i] = a2;
#pragma unroll
        for (int o = 0; o < MLP_OUT; ++o) out[o] += P[W3_OFF + o * MLP_W + i] * a2;
    }
    float e[MLP_OUT];
    double l2 = 0.0;
#pragma unroll
    for (int o = 0; o < MLP_OUT; ++o) {
        const float d = out[o] - th[o];
        e[o] = live ? d * inv_n : 0.f;
        l2 += live ? static_cast<double>(d) * d : 0.0;
        es[s][o] = e[o];
    }
    // delta2 = W3^T e .* gelu'(z2);  delta1 = W2^T delta2 .* gelu'(z1)
    float d2[MLP_W];
#pragma unroll
    for (int i = 0; i < MLP_W; ++i) {
        float acc = 0.f;
#pragma unroll
        for (int o = 0; o < MLP_OUT; ++o) acc += P[W3_OFF + o * MLP_W + i] * e[o];
        d2[i] = acc * gelu_d(z2[i]);
        d2s[s][i] = d2[i];
    }
    for (int j = 0; j < MLP_W; ++j) {
        float acc = 0.f;
#pragma unroll
        for (int i = 0; i < MLP_W; ++i) acc += P[W2_OFF + i * MLP_W + j] * d2[i];
        d1s[s][j] = acc * gelu_d(d1s[s][j]);
    }
    ls[s] = l2;
    __syncthreads();
    // parameter gradients: sums over the workgroup's samples in sample order
    float* gp = partial + static_cast<size_t>(blockIdx.x) * MLP_P;
    for (int p = s; p < MLP_P; p += GB) {
        float acc = 0.f;
        if (p < B1_OFF) {
            const int i = p / MLP_IN, j = p % MLP_IN;
            for (int t = 0; t < GB; ++t) acc += d1s[t][i] * xs[t][j];
        } else if (p < W2_OFF) {
            const int i = p - B1_OFF;
            for (int t = 0; t < GB; ++t) acc += d1s[t][i];
        } else if (p < B2_OFF) {
            const int i = (p - W2_OFF) / MLP_W, j = (p - W2_OFF) % MLP_W;
            for (int t = 0; t < GB; ++t) acc += d2s[t][i] * a1s[t][j];
        } else if (p < W3_OFF) {
            const int i = p - B2_OFF;
            for (int t = 0; t < GB; ++t) acc += d2s[t][i];
        } else if (p < B3_OFF) {
            const int o = (p - W3_OFF) / MLP_W, j = (p - W3_OFF) % MLP_W;
            for (int t = 0; t < GB; ++t) acc += es[t][o] * a2s[t][j];
        } else {
            const int o = p - B3_OFF;
            for (int t = 0; t < GB; ++t) acc += es[t][o];
        }
        gp[p] = acc;
    }
    if (s == 0) {
        double L = 0.0;
        for (int t = 0; t < GB; ++t) L += ls[t];
        partial_loss[blockIdx.x] = L;
    }
}

// Sum the partials in block order; optionally take one Adam step.
__global__ void k_mlp_reduce(const float* partial, const double* partial_loss, int nblk,
                             float* grad, float* P, float* m, float* v, float lr, float b1c,
                             float b2c, int adam, double* loss_acc) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0 && loss_acc) {
        double L = 0.0;
        for (int b = 0; b < nblk; ++b) L += partial_loss[b];
        *loss_acc += L;
    }
    if (p >= MLP_P) return;
    float g = 0.f;
    for (int b = 0; b < nblk; ++b) g += partial[static_cast<size_t>(b) * MLP_P + p];
    if (grad) grad[p] = g;
    if (adam) {
        const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
        const float mm = beta1 * m[p] + (1.f - beta1) * g;
        const float vv = beta2 * v[p] + (1.f - beta2) * g * g;
        m[p] = mm;
        v[p] = vv;
        P[p] -= lr * (mm / b1c) / (sqrtf(vv / b2c) + eps);
    }
}

// Column sums and sums of squares of the sample rows (FP64), per chunk.
__global__ void k_colstats(const float* rows, int64_t nrows, int64_t chunk, double* out) {
    __shared__ double s1[256], s2[256];
    const int col = blockIdx.x, ch = blockIdx.y;
    const int64_t lo = ch * chunk, hi = std::min(nrows, lo + chunk);
    double a = 0.0, b = 0.0;
    for (int64_t r = lo + threadIdx.x; r < hi; r += blockDim.x) {
        const double v = rows[r * ROW + col];
        a += v;
        b += v * v;
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) { s1[threadIdx.x] += s1[threadIdx.x + w]; s2[threadIdx.x] += s2[threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[(static_cast<size_t>(ch) * ROW + col) * 2] = s1[0];
        out[(static_cast<size_t>(ch) * ROW + col) * 2 + 1] = s2[0];
    }
}

void launch_mlp_infer(Ctx& c) {
    if (!c.model || c.K == 0) return;
    hipLaunchKernelGGL(k_mlp_infer, dim3((c.K + MLP_NODES_WG - 1) / MLP_NODES_WG), dim3(MLP_WAVES * 64), 0,
                       c.stream, c.model->params, c.model->norm, c.mdf_feat, c.mdf_pred, c.K, c.mdf,
                       c.mdf_tol);
    hipLaunchKernelGGL(k_mdf_smooth, dim3((c.K + 255) / 256), dim3(256), 0, c.stream, c.mdf_pred,
                       c.spread, c.off, c.owner, c.x, c.y, c.K, c.ny, c.mdf, c.mdf_tol);
}

static int ensure_partial(Mlp& m, int nblk) {
    if (m.partial_blocks >= nblk) return SF_OK;
    if (m.partial) hipFree(m.partial);
    m.partial = nullptr;
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&m.partial),
                     static_cast<size_t>(nblk) * (MLP_P * sizeof(float) + sizeof(double))));
    m.partial_blocks = nblk;
    return SF_OK;
}

static double* partial_loss(Mlp& m) {
    return reinterpret_cast<double*>(m.partial + static_cast<size_t>(m.partial_blocks) * MLP_P);
}

// Gradient (and optionally an Adam step) over count rows starting at first.
static int grad_pass(Mlp& m, const float* rows, const int* idx, int64_t first, int count,
                     float* grad, float lr, bool adam, double* loss_acc) {
    const int nblk = (count + GB - 1) / GB;
    int rc = ensure_partial(m, nblk);
    if (rc != SF_OK) return rc;
    hipLaunchKernelGGL(k_mlp_grad, dim3(nblk), dim3(GB), 0, m.stream, m.params, m.norm, rows, idx,
                       first, count, 1.0f / static_cast<float>(count), m.partial, partial_loss(m));
    float b1c = 1.f, b2c = 1.f;
    if (adam) {
        m.adam_t += 1;
        b1c = 1.f - std::pow(0.9f, static_cast<float>(m.adam_t));
        b2c = 1.f - std::pow(0.999f, static_cast<float>(m.adam_t));
    }
    hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_P + 255) / 256), dim3(256), 0, m.stream, m.partial,
                       partial_loss(m), nblk, grad, m.params, m.adam_m, m.adam_v, lr, b1c, b2c,
                       adam ? 1 : 0, loss_acc);
    SF_HIP(hipGetLastError());
    return SF_OK;
}

// Column statistics of the rows -> normalisation (mean, standard deviation;
// unit scale where a column is constant).
static int fit_norm(Mlp& m, const float* rows, int64_t nrows) {
    const int64_t chunk = 65536;
    const int nch = static_cast<int>((nrows + chunk - 1) / chunk);
    double* d = nullptr;
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&d), static_cast<size_t>(nch) * ROW * 2 * sizeof(double)));
    hipLaunchKernelGGL(k_colstats, dim3(ROW, nch), dim3(256), 0, m.stream, rows, nrows, chunk, d);
    std::vector<double> h(static_cast<size_t>(nch) * ROW * 2);
    hipError_t e = hipStreamSynchronize(m.stream);
    if (e == hipSuccess) e = hipMemcpy(h.data(), d, h.size() * sizeof(double), hipMemcpyDeviceToHost);
    hipFree(d);
    SF_HIP(e);
    float nrm[MLP_NORM];
    for (int col = 0; col < ROW; ++col) {
        double a = 0.0, b = 0.0;
        for (int ch = 0; ch < nch; ++ch) {
            a += h[(static_cast<size_t>(ch) * ROW + col) * 2];
            b += h[(static_cast<size_t>(ch) * ROW + col) * 2 + 1];
        }
        const double mean = a / static_cast<double>(nrows);
        const double var = std::max(0.0, b / static_cast<double>(nrows) - mean * mean);
        double sd = std::sqrt(var);
        if (!(sd > 1e-30 && sd > 1e-7 * std::abs(mean))) sd = std::max(std::abs(mean), 1e-30);
        if (col < MLP_IN) { nrm[col] = static_cast<float>(mean); nrm[MLP_IN + col] = static_cast<float>(sd); }
        else {
            const int o = col - MLP_IN;
            nrm[2 * MLP_IN + o] = static_cast<float>(mean);
            nrm[2 * MLP_IN + MLP_OUT + o] = static_cast<float>(sd);
        }
    }
    SF_HIP(hipMemcpy(m.norm, nrm, sizeof(nrm), hipMemcpyHostToDevice));
    m.norm_set = true;
    return SF_OK;
}

}  // namespace sf

using namespace sf;

extern "C" {

int sf_mlp_create(sf_mlp** out, int device, uint64_t seed) {
    if (!out) return fail(SF_E_ARG, "null pointer");
    *out = nullptr;
    int ndev = 0;
    SF_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(SF_E_ARG, "no such HIP device");
    SF_HIP(hipSetDevice(device));
    auto* h = new sf_mlp();
    Mlp& m = h->m;
    m.device = device;
    bool ok = hipStreamCreateWithFlags(&m.stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&m.params), MLP_P * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&m.norm), MLP_NORM * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&m.adam_m), MLP_P * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&m.adam_v), MLP_P * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&m.acc), sizeof(double)) == hipSuccess;
    if (!ok) { sf_mlp_destroy(h); return fail(SF_E_HIP, "network allocation failed"); }
    // Glorot-uniform weights, zero biases, identity normalisation
    std::vector<float> P(MLP_P, 0.f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int fan_out, int fan_in) {
        const double a = std::sqrt(6.0 / static_cast<double>(fan_in + fan_out));
        std::uniform_real_distribution<double> U(-a, a);
        for (int i = 0; i < fan_out * fan_in; ++i) P[off + i] = static_cast<float>(U(rng));
    };
    fill(W1_OFF, MLP_W, MLP_IN);
    fill(W2_OFF, MLP_W, MLP_W);
    fill(W3_OFF, MLP_OUT, MLP_W);
    float nrm[MLP_NORM];
    for (int j = 0; j < MLP_NORM; ++j) nrm[j] = (j >= MLP_IN && j < 2 * MLP_IN) || j >= 2 * MLP_IN + MLP_OUT ? 1.f : 0.f;
    std::vector<float> zeros(MLP_P, 0.f);
    ok = hipMemcpy(m.params, P.data(), MLP_P * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(m.norm, nrm, sizeof(nrm), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(m.adam_m, zeros.data(), MLP_P * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(m.adam_v, zeros.data(), MLP_P * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) { sf_mlp_destroy(h); return fail(SF_E_HIP, "network upload failed"); }
    *out = h;
    return SF_OK;
}

int sf_mlp_destroy(sf_mlp* h) {
    if (!h) return SF_OK;
    Mlp& m = h->m;
    hipSetDevice(m.device);
    if (m.stream) hipStreamSynchronize(m.stream);
    void* ptrs[] = {m.params, m.norm, m.adam_m, m.adam_v, m.partial, m.acc};
    for (void* p : ptrs) if (p) hipFree(p);
    if (m.stream) hipStreamDestroy(m.stream);
    delete h;
    return SF_OK;
}

int sf_mlp_num_params(int* n_params, int* n_norm) {
    if (n_params) *n_params = MLP_P;
    if (n_norm) *n_norm = MLP_NORM;
    return SF_OK;
}

int sf_mlp_get_params(sf_mlp* h, float* params, float* norm) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    Mlp& m = h->m;
    SF_HIP(hipSetDevice(m.device));
    SF_HIP(hipStreamSynchronize(m.stream));
    if (params) SF_HIP(hipMemcpy(params, m.params, MLP_P * sizeof(float), hipMemcpyDeviceToHost));
    if (norm) SF_HIP(hipMemcpy(norm, m.norm, MLP_NORM * sizeof(float), hipMemcpyDeviceToHost));
    return SF_OK;
}

int sf_mlp_set_params(sf_mlp* h, const float* params, const float* norm) {
    if (!h) return fail(SF_E_ARG, "null pointer");
    Mlp& m = h->m;
    SF_HIP(hipSetDevice(m.device));
    SF_HIP(hipStreamSynchronize(m.stream));
    if (norm) {
        for (int j = MLP_IN; j < 2 * MLP_IN; ++j)
            if (!(norm[j] > 0.f)) return fail(SF_E_ARG, "input scales must be positive");
        for (int j = 2 * MLP_IN + MLP_OUT; j < MLP_NORM; ++j)
            if (!(norm[j] > 0.f)) return fail(SF_E_ARG, "output scales must be positive");
        SF_HIP(hipMemcpy(m.norm, norm, MLP_NORM * sizeof(float), hipMemcpyHostToDevice));
        m.norm_set = true;
    }
    if (params) {
        SF_HIP(hipMemcpy(m.params, params, MLP_P * sizeof(float), hipMemcpyHostToDevice));
        std::vector<float> zeros(MLP_P, 0.f);
        SF_HIP(hipMemcpy(m.adam_m, zeros.data(), MLP_P * sizeof(float), hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(m.adam_v, zeros.data(), MLP_P * sizeof(float), hipMemcpyHostToDevice));
        m.adam_t = 0;
    }
    return SF_OK;
}

int sf_mlp_predict(sf_mlp* h, int n, const float* x, float* y) {
    if (!h || !x || !y || n < 0) return fail(SF_E_ARG, "bad arguments");
    if (n == 0) return SF_OK;
    Mlp& m = h->m;
    SF_HIP(hipSetDevice(m.device));
    float *dx = nullptr, *dy = nullptr;
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&dx), static_cast<size_t>(n) * MLP_IN * sizeof(float)));
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&dy), static_cast<size_t>(n) * MLP_OUT * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(dx, x, static_cast<size_t>(n) * MLP_IN * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_mlp_predict, dim3((n + MLP_NODES_WG - 1) / MLP_NODES_WG), dim3(MLP_WAVES * 64),
                           0, m.stream, m.params, m.norm, dx, dy, n);
        e = hipStreamSynchronize(m.stream);
    }
    if (e == hipSuccess) e = hipMemcpy(y, dy, static_cast<size_t>(n) * MLP_OUT * sizeof(float), hipMemcpyDeviceToHost);
    hipFree(dx); hipFree(dy);
    SF_HIP(e);
    return SF_OK;
}

int sf_mlp_loss_grad(sf_mlp* h, int n, const float* x, const float* t, double* loss, float* grad) {
    if (!h || !x || !t || n <= 0) return fail(SF_E_ARG, "bad arguments");
    Mlp& m = h->m;
    SF_HIP(hipSetDevice(m.device));
    std::vector<float> rows(static_cast<size_t>(n) * ROW);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < MLP_IN; ++j) rows[static_cast<size_t>(i) * ROW + j] = x[static_cast<size_t>(i) * MLP_IN + j];
        for (int o = 0; o < MLP_OUT; ++o) rows[static_cast<size_t>(i) * ROW + MLP_IN + o] = t[static_cast<size_t>(i) * MLP_OUT + o];
    }
    float *drows = nullptr, *dgrad = nullptr;
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&drows), rows.size() * sizeof(float)));
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&dgrad), MLP_P * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(drows, rows.data(), rows.size() * sizeof(float), hipMemcpyHostToDevice);
    const double zero = 0.0;
    if (e == hipSuccess) e = hipMemcpy(m.acc, &zero, sizeof(double), hipMemcpyHostToDevice);
    int rc = SF_OK;
    if (e == hipSuccess) rc = grad_pass(m, drows, nullptr, 0, n, dgrad, 0.f, false, m.acc);
    if (e == hipSuccess && rc == SF_OK) e = hipStreamSynchronize(m.stream);
    double L = 0.0;
    if (e == hipSuccess && rc == SF_OK && grad) e = hipMemcpy(grad, dgrad, MLP_P * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess && rc == SF_OK) e = hipMemcpy(&L, m.acc, sizeof(double), hipMemcpyDeviceToHost);
    hipFree(drows); hipFree(dgrad);
    if (rc != SF_OK) return rc;
    SF_HIP(e);
    if (loss) *loss = 0.5 * L / static_cast<double>(n);
    return SF_OK;
}

int sf_mlp_train(sf_mlp* h, sf_ctx* ctx, int epochs, int batch, float lr, uint64_t seed,
                 double* epoch_loss) {
    if (!h || !ctx) return fail(SF_E_ARG, "null pointer");
    if (epochs < 0 || batch < 1 || !(lr > 0.f)) return fail(SF_E_ARG, "bad training arguments");
    Mlp& m = h->m;
    Ctx& c = ctx->c;
    if (c.device != m.device) return fail(SF_E_ARG, "network and context are on different devices");
    if (!c.ds_buf) return fail(SF_E_STATE, "no sample set (call sf_mdf_collect and step first)");
    SF_HIP(hipSetDevice(m.device));
    launch_mdf_fold(c);
    SF_HIP(hipStreamSynchronize(c.stream));
    MdfDev md{};
    SF_HIP(hipMemcpy(&md, c.mdf, sizeof(MdfDev), hipMemcpyDeviceToHost));
    const int64_t nrows = md.ds_rows;
    if (nrows < 1) return fail(SF_E_STATE, "the sample set is empty");
    if (nrows > 0x7fffffff) return fail(SF_E_ARG, "too many samples");
    if (!m.norm_set) {
        int rc = fit_norm(m, c.ds_buf, nrows);
        if (rc != SF_OK) return rc;
    }
    int* didx = nullptr;
    SF_HIP(hipMalloc(reinterpret_cast<void**>(&didx), nrows * sizeof(int)));
    std::vector<int> idx(nrows);
    std::mt19937_64 rng(seed);
    int rc = SF_OK;
    for (int ep = 0; ep < epochs && rc == SF_OK; ++ep) {
        std::iota(idx.begin(), idx.end(), 0);
        std::shuffle(idx.begin(), idx.end(), rng);
        hipError_t e = hipMemcpyAsync(didx, idx.data(), nrows * sizeof(int), hipMemcpyHostToDevice, m.stream);
        const double zero = 0.0;
        if (e == hipSuccess) e = hipMemcpyAsync(m.acc, &zero, sizeof(double), hipMemcpyHostToDevice, m.stream);
        if (e != hipSuccess) { rc = fail(SF_E_HIP, hipGetErrorString(e)); break; }
        for (int64_t b = 0; b < nrows && rc == SF_OK; b += batch) {
            const int cnt = static_cast<int>(std::min<int64_t>(batch, nrows - b));
            rc = grad_pass(m, c.ds_buf, didx, b, cnt, nullptr, lr * 1.0f, true, m.acc);
        }
        double L = 0.0;
        if (rc == SF_OK) {
            e = hipStreamSynchronize(m.stream);
            if (e == hipSuccess) e = hipMemcpy(&L, m.acc, sizeof(double), hipMemcpyDeviceToHost);
            if (e != hipSuccess) rc = fail(SF_E_HIP, hipGetErrorString(e));
        }
        // the accumulated squared error over the epoch, as a mean loss
        if (epoch_loss) epoch_loss[ep] = 0.5 * L / static_cast<double>(nrows);
    }
    hipFree(didx);
    return rc;
}

}  // extern "C"
