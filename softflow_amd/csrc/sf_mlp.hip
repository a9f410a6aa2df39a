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

// The network on the matrix cores, one wavefront per 16 nodes.  Lane
// l = 16 q + j works on node j of the wave.  Each layer is a product
// [64 units x K] x [K x 16 nodes] done as four 16x16 output tiles of
// v_mfma_f32_16x16x4f32 (K in steps of 4).  Operand layout of that
// instruction (wave64): lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; accumulator r of lane l holds D[4 (l / 16) + r][l % 16].
// After a layer, lane (q, j) therefore holds units 16 t + 4 q + r (tile t,
// accumulator r) of node j, which is the B operand the next layer needs at
// K step 4 t + r when that layer's K order is permuted the same way: the
// activations never leave the registers.  The weights are read in that
// permuted order from a lane-major copy (k_mlp_pack), one 16-byte load per
// four values, so a wavefront needs neither LDS nor a barrier, and a
// launch has one workgroup per 16 nodes to spread over the chip.  The sums
// run in the matrix cores' order; the training forward (k_mlp_grad) runs
// them in index order, so the two agree to FP32 rounding.
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MLP_NODES = 16;            // nodes per wavefront (and workgroup)
constexpr int K1_STEPS = 5;              // layer-1 K steps: 17 inputs padded to 20
// per-lane slots of the packed copy
constexpr int SL_W1 = 0;                              // [tile 4][step 5]
constexpr int SL_W2 = SL_W1 + 4 * K1_STEPS;           // [out tile 4][in tile 4][acc 4]
constexpr int SL_W3 = SL_W2 + 64;                     // [output 2][tile 4][acc 4]
constexpr int SL_B1 = SL_W3 + 32;                     // [tile 4][acc 4]
constexpr int SL_B2 = SL_B1 + 16;
constexpr int SL_MEAN = SL_B2 + 16;                   // input means, [step 5]
constexpr int SL_SCALE = SL_MEAN + K1_STEPS;          // input scales, [step 5]
constexpr int MLP_SLOTS = (SL_SCALE + K1_STEPS + 3) / 4 * 4;
static_assert(MLP_SLOTS == MLP_PACKED / 64, "packed copy size");

// Slot v of lane l: the weight, bias or input normalisation that lane feeds
// to the matrix cores or its own arithmetic (zero-weight, unit-scale padding
// past input 17).
__device__ float mlp_slot(const float* P, const float* norm, int v, int l) {
    const int q = l >> 4, j = l & 15;
    if (v < SL_W2) {
        const int t = v / K1_STEPS, f = 4 * (v % K1_STEPS) + q;
        return f < MLP_IN ? P[W1_OFF + (16 * t + j) * MLP_IN + f] : 0.f;
    }
    if (v < SL_W3) {
        const int u = v - SL_W2, to = u >> 4, ti = (u >> 2) & 3, r = u & 3;
        return P[W2_OFF + (16 * to + j) * MLP_W + 16 * ti + 4 * q + r];
    }
    if (v < SL_B1) {
        const int u = v - SL_W3, o = u >> 4, t = (u >> 2) & 3, r = u & 3;
        return P[W3_OFF + o * MLP_W + 16 * t + 4 * q + r];
    }
    if (v < SL_B2) return P[B1_OFF + 16 * ((v - SL_B1) >> 2) + 4 * q + ((v - SL_B1) & 3)];
    if (v < SL_MEAN) return P[B2_OFF + 16 * ((v - SL_B2) >> 2) + 4 * q + ((v - SL_B2) & 3)];
    if (v < SL_SCALE) {
        const int f = 4 * (v - SL_MEAN) + q;
        return f < MLP_IN ? norm[f] : 0.f;
    }
    if (v < SL_SCALE + K1_STEPS) {
        const int f = 4 * (v - SL_SCALE) + q;
        return f < MLP_IN ? norm[MLP_IN + f] : 1.f;
    }
    return 0.f;
}

// The lane-major copy: packed[(v / 4) * 256 + 4 l + v % 4] = slot v of lane l.
__global__ void k_mlp_pack(const float* __restrict__ P, const float* __restrict__ norm, float* packed) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < MLP_PACKED; i += gridDim.x * blockDim.x)
        packed[i] = mlp_slot(P, norm, (i >> 8) * 4 + (i & 3), (i >> 2) & 63);
}

struct MlpOut {
    float y0, y1;      // the two outputs of node j before denormalisation
    bool ok;           // all inputs of node j finite
};

// This lane's slots of the packed copy, into registers.
__device__ __forceinline__ void mlp_weights(const float* __restrict__ packed, float (&w)[MLP_SLOTS]) {
    const int l = threadIdx.x & 63;
    const f32x4* pk = reinterpret_cast<const f32x4*>(packed);
#pragma unroll
    for (int u = 0; u < MLP_SLOTS / 4; ++u) {
        const f32x4 v = pk[u * 64 + l];
        w[4 * u] = v[0]; w[4 * u + 1] = v[1]; w[4 * u + 2] = v[2]; w[4 * u + 3] = v[3];
    }
}

// The network on nodes k0 .. k0+15 of x (rows of MLP_IN, clipped at n; x may
// point into LDS), with this lane's weight slots w.  Every lane returns the
// result of its node j = l % 16.
__device__ __forceinline__ MlpOut mlp_tile(const float (&w)[MLP_SLOTS], const float* __restrict__ P,
                                           const float* x, int k0, int n) {
    const int l = threadIdx.x & 63, q = l >> 4, j = l & 15;
    const int node = k0 + j;
    float xh[K1_STEPS];
    int bad = 0;
#pragma unroll
    for (int s = 0; s < K1_STEPS; ++s) {
        const int f = 4 * s + q;
        const float v = node < n && f < MLP_IN ? x[static_cast<size_t>(node) * MLP_IN + f] : 0.f;
        bad |= !isfinite(v);
        xh[s] = (v - w[SL_MEAN + s]) / w[SL_SCALE + s];
    }
    bad |= __shfl_xor(bad, 16);
    bad |= __shfl_xor(bad, 32);
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < K1_STEPS; ++s)
#pragma unroll
        for (int t = 0; t < 4; ++t)
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(w[SL_W1 + t * K1_STEPS + s], xh[s], acc[t], 0, 0, 0);
    float h[16];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) h[4 * t + r] = gelu(acc[t][r] + w[SL_B1 + 4 * t + r]);
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 16; ++s)
#pragma unroll
        for (int t = 0; t < 4; ++t)
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(w[SL_W2 + 16 * t + s], h[s], acc[t], 0, 0, 0);
    // the 64 -> 2 layer: this lane's 16 units, then the sum over the four q
    float o0 = 0.f, o1 = 0.f;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float g = gelu(acc[t][r] + w[SL_B2 + 4 * t + r]);
            o0 += w[SL_W3 + 4 * t + r] * g;
            o1 += w[SL_W3 + 16 + 4 * t + r] * g;
        }
    o0 += __shfl_xor(o0, 16);
    o1 += __shfl_xor(o1, 16);
    o0 += __shfl_xor(o0, 32);
    o1 += __shfl_xor(o1, 32);
    return MlpOut{P[B3_OFF] + o0, P[B3_OFF + 1] + o1, bad == 0};
}

__device__ MlpOut mlp_wave(const float* __restrict__ packed, const float* __restrict__ P, const float* x,
                           int k0, int n) {
    float w[MLP_SLOTS];
    mlp_weights(packed, w);
    return mlp_tile(w, P, x, k0, n);
}

// The gains of one node from the network's outputs (the in-step inference): held to
// g_par in [0, 4] and g_perp in [-2, 2], (1, 0) where an input or an output is
// not finite.
__device__ __forceinline__ void mlp_gains(const MlpOut& r, const float* __restrict__ norm, double& g0, double& g1) {
    g0 = 1.0;
    g1 = 0.0;
    if (r.ok) {
        const float gp = norm[2 * MLP_IN] + norm[2 * MLP_IN + MLP_OUT] * r.y0;
        const float gq = norm[2 * MLP_IN + 1] + norm[2 * MLP_IN + MLP_OUT + 1] * r.y1;
        if (isfinite(gp) && isfinite(gq)) {
            g0 = fminf(fmaxf(gp, 0.f), 4.f);
            g1 = fminf(fmaxf(gq, -2.f), 2.f);
        }
    }
}

// The network alone on n inputs, on the matrix cores: y = out_mean + out_scale * net(x^).
__global__ __launch_bounds__(64) void k_mlp_predict(const float* __restrict__ packed, const float* __restrict__ P,
                                                    const float* __restrict__ norm, const float* x, float* y, int n) {
    const int k0 = blockIdx.x * MLP_NODES;
    const MlpOut r = mlp_wave(packed, P, x, k0, n);
    const int l = threadIdx.x, k = k0 + (l & 15);
    if (l < MLP_NODES && k < n) {
        y[static_cast<size_t>(k) * MLP_OUT] = norm[2 * MLP_IN] + norm[2 * MLP_IN + MLP_OUT] * r.y0;
        y[static_cast<size_t>(k) * MLP_OUT + 1] = norm[2 * MLP_IN + 1] + norm[2 * MLP_IN + MLP_OUT + 1] * r.y1;
    }
}

// The gains, low-passed along each membrane with the binomial weights
// (1 4 6 4 1)/16, scale the first plain correction dF1 (smooth, since both
// velocities it compares are interpolated).  Node-by-node prediction errors
// would otherwise alternate between neighbours: that zig-zag barely moves the
// interpolated velocity, so the residual cannot see it, yet it pushes the
// fluid.
// Node k's smoothed gains scale its record's dF1; gain[2 m] holds node m's
// gains for m in [base, base + count) (the node's capsule lies inside).
__device__ __forceinline__ void mdf_smooth_node(const double* gain, int base, NodeSpread* rec, const int* off,
                                                const int* owner, const double* x, const double* y, int ny, int k) {
    const int cap = owner[k];
    const int o = off[cap], n = off[cap + 1] - o, i = k - o;
    double gp = gain[2 * (k - base)], gq = gain[2 * (k - base) + 1];
    if (n >= 5) {
        const int j[5] = {o + (i - 2 + n) % n, o + (i - 1 + n) % n, k, o + (i + 1) % n, o + (i + 2) % n};
        const double w[5] = {1.0, 4.0, 6.0, 4.0, 1.0};
        gp = 0.0;
        gq = 0.0;
        for (int t = 0; t < 5; ++t) {
            gp += w[t] * gain[2 * (j[t] - base)];
            gq += w[t] * gain[2 * (j[t] - base) + 1];
        }
        gp *= 0.0625;
        gq *= 0.0625;
    }
    const NodeSpread r = rec[k];    // dF1
    // the whole record, valid where the correction is finite
    rec[k] = node_record(x[k], y[k], gp * r.fx - gq * r.fy, gp * r.fy + gq * r.fx, r.ds, k, ny);
}

// The learned warm start in one launch, one workgroup per group of whole
// capsules (at most WARM_NODES nodes, or one larger capsule): every node's
// inputs (mdf_node_features, also kept for the sample rows) and dF1 into
// f1; then, unless the residual is already within tolerance, the network on
// 16-node tiles of the group (one wavefront per tile, the weights loaded into
// registers once per wavefront), the gains held in LDS, and the smoothed
// correction into each node's record.  Everything is the three-kernel path's
// arithmetic as the network alone (k_mlp_predict): each node's network output
// depends on its own inputs only, whichever tile holds it.
constexpr int WARM_NODES = 64;           // nodes a group of several capsules may hold
constexpr int WARM_MAX = 256;            // nodes of the largest capsule
struct WarmArgs {
    FeatArgs fa;
    const int* grp;                      // first capsule of each group, ngroups + 1 entries
    float* feat;
    double* f1;
    NodeSpread* rec;
    const float *packed, *P, *norm;
    const MdfDev* md;
    double tol;
    int ny;
};

__global__ __launch_bounds__(256) void k_mdf_warm(WarmArgs a) {
    __shared__ float sx[WARM_MAX * MLP_IN];
    __shared__ double sg[2 * WARM_MAX];
    const int g = blockIdx.x;
    const int base = a.fa.off[a.grp[g]], n = a.fa.off[a.grp[g + 1]] - base;
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const int k = base + t;
        a.f1[2 * k] = a.rec[k].fx;
        a.f1[2 * k + 1] = a.rec[k].fy;
        float f[MLP_IN];
        mdf_node_features(a.fa, k, f);
#pragma unroll
        for (int j = 0; j < MLP_IN; ++j) {
            sx[t * MLP_IN + j] = f[j];
            a.feat[static_cast<size_t>(k) * MLP_IN + j] = f[j];
        }
    }
    if (a.tol > 0.0 && mdf_res(a.md, 1) <= a.tol) return;    // the same for the whole grid
    __syncthreads();
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (wv * MLP_NODES < n) {
        float w[MLP_SLOTS];
        mlp_weights(a.packed, w);
        for (int t0 = wv * MLP_NODES; t0 < n; t0 += 4 * MLP_NODES) {
            const MlpOut r = mlp_tile(w, a.P, sx, t0, n);
            if (l < MLP_NODES && t0 + l < n) mlp_gains(r, a.norm, sg[2 * (t0 + l)], sg[2 * (t0 + l) + 1]);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += blockDim.x)
        mdf_smooth_node(sg, base, a.rec, a.fa.off, a.fa.owner, a.fa.x, a.fa.y, a.ny, base + t);
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

// The first correction's warm start (k_mdf_warm), in place of launch_mdf_features.
void launch_mdf_warm(Ctx& c) {
    if (!c.model || c.K == 0) return;
    const WarmArgs a{feat_args(c), c.warm_grp, c.mdf_feat, c.mdf_f1, c.spread, c.model->packed, c.model->params,
                     c.model->norm, c.mdf, c.mdf_tol, c.ny};
    hipLaunchKernelGGL(k_mdf_warm, dim3(c.warm_ngroups), dim3(256), 0, c.stream, a);
}

// Capsule groups of the warm start: consecutive capsules while their nodes
// fit WARM_NODES (a larger capsule alone).
std::vector<int> warm_groups(const int32_t* off, int ncaps) {
    std::vector<int> grp{0};
    int first = 0;
    for (int i = 1; i <= ncaps; ++i) {
        if (i == ncaps || off[i + 1] - off[first] > WARM_NODES) {
            grp.push_back(i);
            first = i;
        }
    }
    return grp;
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

// Refresh the lane-major copy inference reads; every change of the weights or
// the normalisation ends here, before the device is used on another stream.
static int repack(Mlp& m) {
    hipLaunchKernelGGL(k_mlp_pack, dim3(MLP_PACKED / 256), dim3(256), 0, m.stream, m.params, m.norm, m.packed);
    SF_HIP(hipGetLastError());
    SF_HIP(hipStreamSynchronize(m.stream));
    return SF_OK;
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
    return repack(m);
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
    ok = ok && hipMalloc(reinterpret_cast<void**>(&m.packed), MLP_PACKED * sizeof(float)) == hipSuccess;
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
    if (!ok || repack(m) != SF_OK) { sf_mlp_destroy(h); return fail(SF_E_HIP, "network upload failed"); }
    *out = h;
    return SF_OK;
}

int sf_mlp_destroy(sf_mlp* h) {
    if (!h) return SF_OK;
    Mlp& m = h->m;
    hipSetDevice(m.device);
    if (m.stream) hipStreamSynchronize(m.stream);
    void* ptrs[] = {m.params, m.norm, m.packed, m.adam_m, m.adam_v, m.partial, m.acc};
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
    return repack(m);
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
        hipLaunchKernelGGL(k_mlp_predict, dim3((n + MLP_NODES - 1) / MLP_NODES), dim3(64), 0, m.stream,
                           m.packed, m.params, m.norm, dx, dy, n);
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
    const int rp = repack(m);
    return rc != SF_OK ? rc : rp;
}

}  // extern "C"
