// Lattice kernels: fused D2Q9 collide + push-stream with in-kernel halfway
// bounce-back and periodic wrap, the advection-diffusion step, and the
// on-demand macroscopic / reduction kernels.
//
// Layout in HBM: populations in SoA planes f[q][y*nx+x] (q-major, x fastest),
// double-buffered.  FP64 mode stores f itself and evaluates every expression
// in the reference's order (bit parity); FP32 mode stores f - w_q and
// collides in FP32.  The raw spread force is an interleaved (Fx,Fy) pair per
// cell; body force and the |F| cap are applied in-kernel.
#include <type_traits>

#include "sf_internal.h"

namespace sf {

// ---------------------------------------------------------------------------
// FP64 collision operators in the reference's operation order.
// BGK + Guo: lbm/bgk_collision.cpp:56-85
__device__ __forceinline__ void col_bgk(double* f, double om, double r, double vx, double vy,
                                        double fx, double fy) {
    const double om1 = 1.0 - om;
    const double guo = 1.0 - 0.5 * om;
    const double uu = vx * vx + vy * vy;
    const double uu_term = -0.5 * 3.0 * uu;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const double ex = static_cast<double>(dcx(q));
        const double ey = static_cast<double>(dcy(q));
        const double eu = ex * vx + ey * vy;
        double fe = dw(q) * r * (1.0 + eu * 3.0 + 0.5 * eu * eu * 9.0 + uu_term);
        double tx = (ex - vx) * 3.0 + (eu * 9.0) * ex;
        double ty = (ey - vy) * 3.0 + (eu * 9.0) * ey;
        double Fi = guo * dw(q) * (tx * fx + ty * fy);
        f[q] = om1 * f[q] + om * fe + Fi;
    }
}

// MRT + Guo in moment space: lbm/mrt_collision.cpp:65-138
__device__ __forceinline__ void col_mrt(double* f, double sn, double s1, double s2, double s4,
                                        double r, double vx, double vy, double fx, double fy) {
    const double s6 = s4;
    double f0 = f[0], f1 = f[1], f2 = f[2], f3 = f[3], f4 = f[4];
    double f5 = f[5], f6 = f[6], f7 = f[7], f8 = f[8];
    double m0 = f0 + f1 + f2 + f3 + f4 + f5 + f6 + f7 + f8;
    double m1 = -4.0 * f0 - f1 - f2 - f3 - f4 + 2.0 * f5 + 2.0 * f6 + 2.0 * f7 + 2.0 * f8;
    double m2 = 4.0 * f0 - 2.0 * f1 - 2.0 * f2 - 2.0 * f3 - 2.0 * f4 + f5 + f6 + f7 + f8;
    double m3 = f1 - f3 + f5 - f6 - f7 + f8;
    double m4 = -2.0 * f1 + 2.0 * f3 + f5 - f6 - f7 + f8;
    double m5 = f2 - f4 + f5 + f6 - f7 - f8;
    double m6 = -2.0 * f2 + 2.0 * f4 + f5 + f6 - f7 - f8;
    double m7 = f1 - f2 + f3 - f4;
    double m8 = f5 - f6 + f7 - f8;
    double uu = vx * vx + vy * vy;
    double e1 = -2.0 * r + 3.0 * r * uu;
    double e2 = r - 3.0 * r * uu;
    double e4 = -r * vx;
    double e6 = -r * vy;
    double e7 = r * (vx * vx - vy * vy);
    double e8 = r * vx * vy;
    double uF = vx * fx + vy * fy;
    double S1 = 6.0 * uF, S2 = -6.0 * uF, S3 = fx, S4 = -fx, S5 = fy, S6 = -fy;
    double S7 = 2.0 * (vx * fx - vy * fy);
    double S8 = vx * fy + vy * fx;
    m1 = m1 - s1 * (m1 - e1) + (1.0 - s1 * 0.5) * S1;
    m2 = m2 - s2 * (m2 - e2) + (1.0 - s2 * 0.5) * S2;
    m3 = m3 + S3;
    m4 = m4 - s4 * (m4 - e4) + (1.0 - s4 * 0.5) * S4;
    m5 = m5 + S5;
    m6 = m6 - s6 * (m6 - e6) + (1.0 - s6 * 0.5) * S6;
    m7 = m7 - sn * (m7 - e7) + (1.0 - sn * 0.5) * S7;
    m8 = m8 - sn * (m8 - e8) + (1.0 - sn * 0.5) * S8;
    f[0] = m0 / 9.0 - m1 / 9.0 + m2 / 9.0;
    f[1] = m0 / 9.0 - m1 / 36.0 - m2 / 18.0 + m3 / 6.0 - m4 / 6.0 + m7 / 4.0;
    f[2] = m0 / 9.0 - m1 / 36.0 - m2 / 18.0 + m5 / 6.0 - m6 / 6.0 - m7 / 4.0;
    f[3] = m0 / 9.0 - m1 / 36.0 - m2 / 18.0 - m3 / 6.0 + m4 / 6.0 + m7 / 4.0;
    f[4] = m0 / 9.0 - m1 / 36.0 - m2 / 18.0 - m5 / 6.0 + m6 / 6.0 - m7 / 4.0;
    f[5] = m0 / 9.0 + m1 / 18.0 + m2 / 36.0 + m3 / 6.0 + m4 / 12.0 + m5 / 6.0 + m6 / 12.0 + m8 / 4.0;
    f[6] = m0 / 9.0 + m1 / 18.0 + m2 / 36.0 - m3 / 6.0 - m4 / 12.0 + m5 / 6.0 + m6 / 12.0 - m8 / 4.0;
    f[7] = m0 / 9.0 + m1 / 18.0 + m2 / 36.0 - m3 / 6.0 - m4 / 12.0 - m5 / 6.0 - m6 / 12.0 + m8 / 4.0;
    f[8] = m0 / 9.0 + m1 / 18.0 + m2 / 36.0 + m3 / 6.0 + m4 / 12.0 - m5 / 6.0 - m6 / 12.0 - m8 / 4.0;
}

// Regularized BGK keeping the first-order non-equilibrium moment:
// lbm/regularized_bgk.cpp:35-91; equilibrium as lbm/lattice.h:35-39.
__device__ __forceinline__ double feq_ref(int q, double rho, double ux, double uy) {
    constexpr double cs2 = 1.0 / 3.0;
    constexpr double cs4 = cs2 * cs2;
    double eu = static_cast<double>(dcx(q)) * ux + static_cast<double>(dcy(q)) * uy;
    double uu = ux * ux + uy * uy;
    return dw(q) * rho * (1.0 + eu / cs2 + (eu * eu) / (2.0 * cs4) - uu / (2.0 * cs2));
}
__device__ __forceinline__ void col_reg(double* f, double omega, double rho, double ux, double uy,
                                        double Fx, double Fy) {
    double fe[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) fe[q] = feq_ref(q, rho, ux, uy);
    double jx = 0.0, jy = 0.0, pxx = 0.0, pxy = 0.0, pyy = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        double fn = f[q] - fe[q];
        jx += fn * dcx(q);
        jy += fn * dcy(q);
        pxx += fn * dcx(q) * dcx(q);
        pxy += fn * dcx(q) * dcy(q);
        pyy += fn * dcy(q) * dcy(q);
    }
    constexpr double cs2 = 1.0 / 3.0;
    constexpr double cs4_2 = 2.0 * cs2 * cs2;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int cx = dcx(q), cy = dcy(q);
        double Qxx = cx * cx - cs2;
        double Qxy = cx * cy;
        double Qyy = cy * cy - cs2;
        double reg = dw(q) * ((cx * jx + cy * jy) / cs2 + (Qxx * pxx + 2.0 * Qxy * pxy + Qyy * pyy) / cs4_2);
        double eu = cx * ux + cy * uy;
        double Fi = (1.0 - 0.5 * omega) * dw(q) *
                    ((cx - ux + eu * cx / cs2) * Fx / cs2 + (cy - uy + eu * cy / cs2) * Fy / cs2);
        f[q] = fe[q] + (1.0 - omega) * reg + Fi;
    }
}

// ---------------------------------------------------------------------------
// FP32 collision operators on deviations d_q = f_q - w_q.  The moments are
// formed from the deviations (rho = 1 + sum d), so the small Guo increments
// are not lost against w_q ~ 0.1.
__device__ __forceinline__ void dcol_bgk(float* d, float om, float drho, float rho, float ux,
                                         float uy, float fx, float fy) {
    const float guo = 1.0f - 0.5f * om;
    const float uu15 = 1.5f * (ux * ux + uy * uy);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const float ex = static_cast<float>(dcx(q)), ey = static_cast<float>(dcy(q));
        const float w = static_cast<float>(dw(q));
        const float eu = ex * ux + ey * uy;
        const float deq = w * (drho + rho * (3.0f * eu + 4.5f * eu * eu - uu15));
        const float src = guo * w * ((3.0f * (ex - ux) + 9.0f * eu * ex) * fx +
                                     (3.0f * (ey - uy) + 9.0f * eu * ey) * fy);
        d[q] = d[q] - om * (d[q] - deq) + src;
    }
}

__device__ __forceinline__ void dcol_mrt(float* d, float sn, float s1, float s2, float s4,
                                         float drho, float rho, float ux, float uy, float fx,
                                         float fy) {
    // moments of the deviations; the weights contribute m1 = -2, m2 = 1, which
    // cancel in m - m_eq and in M^-1 (their image is w itself).
    const float d0 = d[0], d1 = d[1], d2 = d[2], d3 = d[3], d4 = d[4];
    const float d5 = d[5], d6 = d[6], d7 = d[7], d8 = d[8];
    const float a14 = d1 + d2 + d3 + d4, a58 = d5 + d6 + d7 + d8;
    float m1 = -4.0f * d0 - a14 + 2.0f * a58;
    float m2 = 4.0f * d0 - 2.0f * a14 + a58;
    float m3 = (d1 - d3) + (d5 - d6) + (d8 - d7);
    float m4 = -2.0f * (d1 - d3) + (d5 - d6) + (d8 - d7);
    float m5 = (d2 - d4) + (d5 + d6) - (d7 + d8);
    float m6 = -2.0f * (d2 - d4) + (d5 + d6) - (d7 + d8);
    float m7 = (d1 + d3) - (d2 + d4);
    float m8 = (d5 + d7) - (d6 + d8);
    const float uu = ux * ux + uy * uy;
    const float e1 = -2.0f * drho + 3.0f * rho * uu;
    const float e2 = drho - 3.0f * rho * uu;
    const float e4 = -rho * ux, e6 = -rho * uy;
    const float e7 = rho * (ux * ux - uy * uy), e8 = rho * ux * uy;
    const float uF = ux * fx + uy * fy;
    const float s6 = s4;
    m1 = m1 - s1 * (m1 - e1) + (1.0f - 0.5f * s1) * (6.0f * uF);
    m2 = m2 - s2 * (m2 - e2) - (1.0f - 0.5f * s2) * (6.0f * uF);
    m3 = m3 + fx;
    m4 = m4 - s4 * (m4 - e4) - (1.0f - 0.5f * s4) * fx;
    m5 = m5 + fy;
    m6 = m6 - s6 * (m6 - e6) - (1.0f - 0.5f * s6) * fy;
    m7 = m7 - sn * (m7 - e7) + (1.0f - 0.5f * sn) * (2.0f * (ux * fx - uy * fy));
    m8 = m8 - sn * (m8 - e8) + (1.0f - 0.5f * sn) * (ux * fy + uy * fx);
    const float m0 = drho;
    const float b = m0 * (1.0f / 9.0f);
    const float c14 = b - m1 * (1.0f / 36.0f) - m2 * (1.0f / 18.0f);
    const float c58 = b + m1 * (1.0f / 18.0f) + m2 * (1.0f / 36.0f);
    d[0] = b - m1 * (1.0f / 9.0f) + m2 * (1.0f / 9.0f);
    const float hx = m3 * (1.0f / 6.0f) - m4 * (1.0f / 6.0f);
    const float hy = m5 * (1.0f / 6.0f) - m6 * (1.0f / 6.0f);
    const float q7 = m7 * 0.25f;
    d[1] = c14 + hx + q7;
    d[2] = c14 + hy - q7;
    d[3] = c14 - hx + q7;
    d[4] = c14 - hy - q7;
    const float gx = m3 * (1.0f / 6.0f) + m4 * (1.0f / 12.0f);
    const float gy = m5 * (1.0f / 6.0f) + m6 * (1.0f / 12.0f);
    const float q8 = m8 * 0.25f;
    d[5] = c58 + gx + gy + q8;
    d[6] = c58 - gx + gy - q8;
    d[7] = c58 - gx - gy + q8;
    d[8] = c58 + gx - gy - q8;
}

__device__ __forceinline__ void dcol_reg(float* d, float om, float drho, float rho, float ux,
                                         float uy, float fx, float fy) {
    float deq[9];
    const float uu15 = 1.5f * (ux * ux + uy * uy);
    float jx = 0.f, jy = 0.f, pxx = 0.f, pxy = 0.f, pyy = 0.f;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const float ex = static_cast<float>(dcx(q)), ey = static_cast<float>(dcy(q));
        const float eu = ex * ux + ey * uy;
        deq[q] = static_cast<float>(dw(q)) * (drho + rho * (3.0f * eu + 4.5f * eu * eu - uu15));
        const float fn = d[q] - deq[q];
        jx += fn * ex; jy += fn * ey;
        pxx += fn * ex * ex; pxy += fn * ex * ey; pyy += fn * ey * ey;
    }
    const float guo = 1.0f - 0.5f * om;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const float ex = static_cast<float>(dcx(q)), ey = static_cast<float>(dcy(q));
        const float w = static_cast<float>(dw(q));
        const float Qxx = ex * ex - (1.0f / 3.0f), Qyy = ey * ey - (1.0f / 3.0f), Qxy = ex * ey;
        const float reg = w * (3.0f * (ex * jx + ey * jy) + 4.5f * (Qxx * pxx + 2.0f * Qxy * pxy + Qyy * pyy));
        const float eu = ex * ux + ey * uy;
        const float src = guo * w * ((3.0f * (ex - ux) + 9.0f * eu * ex) * fx +
                                     (3.0f * (ey - uy) + 9.0f * eu * ey) * fy);
        d[q] = deq[q] + (1.0f - om) * reg + src;
    }
}

// ---------------------------------------------------------------------------
// Streaming accesses: every population is read once and written once per
// step, so the loads and stores bypass cache retention (nontemporal).
template <typename T>
__device__ __forceinline__ T ld_nt(const T* p) { return __builtin_nontemporal_load(p); }
template <typename T>
__device__ __forceinline__ void st_nt(T* p, T v) { __builtin_nontemporal_store(v, p); }

struct LbmArgs {
    const void* fsrc;
    void* fdst;
    const void* Fcur;        // raw spread force of this step
    const void* Fshift;      // raw force feeding the collision velocity
    const uint8_t* flags;
    // with generic flags: per cell, bit q set when the neighbour in direction q
    // (both axes wrapped) is SOLID, and the cell's own type in bits 9-11
    const uint16_t* nbm;
    const double* srho;      // stored macro (SHIFT_STORED)
    const double* su;
    int* counters;
    int nx, ny, N;
    int periodic_x, periodic_y, walls_x, walls_y;
    int shift;
    int has_body;
    double bfx, bfy, fmax, fmax2;
    double om, se, seps, sq;
    // scalar transport riding along (the step before this one's)
    const void* gsrc[4];
    void* gdst[4];
    const double* Cst[4];
    int nsp, from_store, scalar_periodic_y;
    double oms[4];
    // launch geometry: 1-D grid over nbx x nby blocks of blockDim cells; with
    // remap, workgroup b (dispatched to XCD b % 8) takes tile (b % 8) * per + b / 8,
    // so each XCD streams one contiguous band of rows.  persist > 0: a grid of
    // 8 * persist workgroups; the persist workgroups of an XCD walk its band
    // with stride persist.
    int nbx, nblk, per_xcd, remap, persist;
    // batched collide-stream: this member takes two cells per lane (k_lbm_pair's body);
    // flat: its workgroups take consecutive runs of the row-major cell index
    // (across row ends) instead of rectangles of bx x by cells
    int pair, flat;
    // moving walls (lbm/moving_wall.cpp): x velocities of the solid rows y = 0 and y = ny-1
    int moving;
    double wall_bot, wall_top;
    // Zou-He inlet / outlet (lbm/zou_he_boundary.cpp) and the scalar's inlet concentration
    double inlet_ux, outlet_rho;
    double cin[4];
};

// FLUID cells take the body force and collide (bgk_collision.cpp:43,
// simulation.cpp:541-609); INLET and OUTLET cells only stream.
__device__ __forceinline__ bool is_fluid_type(const LbmArgs& a, int x, int y, bool use_flags) {
    if (use_flags) return a.flags[y * a.nx + x] == SF_FLUID;
    return !((a.walls_y && (y == 0 || y == a.ny - 1)) || (a.walls_x && (x == 0 || x == a.nx - 1)));
}

__device__ __forceinline__ bool block_origin(const LbmArgs& a, int b, int& x, int& y) {
    const int t = a.remap ? (b % 8) * a.per_xcd + b / 8 : b;
    if (t >= a.nblk) return false;
    x = (t % a.nbx) * blockDim.x + threadIdx.x;
    y = (t / a.nbx) * blockDim.y + threadIdx.y;
    return true;
}

__device__ __forceinline__ bool is_solid(const LbmArgs& a, int x, int y, bool use_flags) {
    if (use_flags) return a.flags[y * a.nx + x] == SF_SOLID;
    return (a.walls_y && (y == 0 || y == a.ny - 1)) || (a.walls_x && (x == 0 || x == a.nx - 1));
}

// The collide-stream's view of the cell types: one 16-bit mask per cell
// instead of a flag load per neighbour.  A neighbour is looked up only where
// the push consults it (in bounds, or wrapped by a periodic axis), where the
// wrapped position is the neighbour's own.
__device__ __forceinline__ int mask_type(unsigned m) { return static_cast<int>((m >> 9) & 7u); }
__device__ __forceinline__ bool mask_solid_nb(unsigned m, int q) { return (m >> q) & 1u; }

template <typename T>
__device__ __forceinline__ void load_force(const void* F, int n, T& fx, T& fy) {
    const T* p = static_cast<const T*>(F) + 2 * static_cast<size_t>(n);
    fx = p[0];
    fy = p[1];
}

// Force of the collision: raw spread + body force (FLUID cells) then the |F|
// cap (simulation.cpp:541-609).  Returns 1 if the cap fired.
template <typename T>
__device__ __forceinline__ int force_finish(const LbmArgs& a, bool fluid, T& fx, T& fy) {
    if (a.has_body && fluid) {
        fx += static_cast<T>(a.bfx);
        fy += static_cast<T>(a.bfy);
    }
    T m2 = fx * fx + fy * fy;
    if (m2 > static_cast<T>(a.fmax2)) {
        T sc = static_cast<T>(a.fmax) / sqrt(m2);
        fx *= sc;
        fy *= sc;
        return 1;
    }
    return 0;
}

template <typename T>
__device__ __forceinline__ int collision_force(const LbmArgs& a, const void* F, int n, bool fluid,
                                               T& fx, T& fy) {
    load_force<T>(F, n, fx, fy);
    return force_finish<T>(a, fluid, fx, fy);
}

// One cell of the advection-diffusion step (advection_diffusion.cpp:111-223):
// BGK with first-order equilibrium from C and the post-LBM velocity, then a
// push with half-way bounce-back; x is always periodic for the scalar.
template <bool FP32, bool FLAGS>
__device__ __forceinline__ void adr_cell(const LbmArgs& a, int x, int y, int n, double ux, double uy,
                                         unsigned m) {
    using T = typename std::conditional<FP32, float, double>::type;
    const size_t N = static_cast<size_t>(a.N);
    for (int s = 0; s < a.nsp; ++s) {
        const T* gs = static_cast<const T*>(a.gsrc[s]);
        T* gd = static_cast<T*>(a.gdst[s]);
        T g[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) g[q] = ld_nt(gs + q * N + n);
        double C;
        if (a.from_store) {
            C = a.Cst[s][n];
        } else {
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < 9; ++q) acc += static_cast<double>(g[q]);
            C = acc;
        }
        const double oms = a.oms[s];
        // INLET cells relax towards the inlet concentration (advection_diffusion.cpp:111-150)
        const bool inlet = FLAGS && mask_type(m) == SF_INLET;
        if (inlet) C = a.cin[s];
        if (!FP32 || inlet) {
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                const double eu = static_cast<double>(dcx(q)) * ux + static_cast<double>(dcy(q)) * uy;
                const double ge = dw(q) * C * (1.0 + eu * 3.0);
                g[q] = static_cast<T>(static_cast<double>(g[q]) - oms * (static_cast<double>(g[q]) - ge));
            }
        } else {
            // FP32: the eight moving populations relax in FP32; the rest
            // population takes C minus their sum (FP64), which is its exact-
            // arithmetic value.  Evaluated directly, the FP32 weights (which do
            // not sum to one) would drift the scalar mass by ~1e-8 per step.
            const float Cf = static_cast<float>(C), omf = static_cast<float>(oms);
            double moving = 0.0;
#pragma unroll
            for (int q = 1; q < 9; ++q) {
                const float eu = static_cast<float>(dcx(q) * ux + dcy(q) * uy);
                const float ge = static_cast<float>(dw(q)) * Cf * (1.0f + 3.0f * eu);
                g[q] = g[q] - omf * (g[q] - ge);
                moving += static_cast<double>(g[q]);
            }
            g[0] = static_cast<float>(C - moving);
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            int xd = x + dcx(q);
            xd = xd < 0 ? xd + a.nx : (xd >= a.nx ? xd - a.nx : xd);
            int yd = y + dcy(q);
            bool bounce;
            const bool nb_solid = FLAGS ? mask_solid_nb(m, q) : false;
            if (a.scalar_periodic_y) {
                yd = yd < 0 ? yd + a.ny : (yd >= a.ny ? yd - a.ny : yd);
                bounce = FLAGS ? nb_solid : is_solid(a, xd, yd, false);
            } else {
                bounce = (yd < 0 || yd >= a.ny) || (FLAGS ? nb_solid : is_solid(a, xd, yd, false));
            }
            if (bounce) st_nt(gd + dopp(q) * N + n, g[q]);
            else st_nt(gd + q * N + (yd * a.nx + xd), g[q]);
        }
    }
}

template <int OP, bool FP32, bool FLAGS, bool ADRF>
__device__ __forceinline__ int lbm_cell(const LbmArgs& a, int x, int y) {
    int capped = 0;
    if (x < a.nx && y < a.ny) {
        const int n = y * a.nx + x;
        unsigned m = 0;
        if constexpr (FLAGS) m = a.nbm[n];
        const bool solid = FLAGS ? mask_type(m) == SF_SOLID : is_solid(a, x, y, false);
        const bool fluid = !solid;
        const bool collide = FLAGS ? mask_type(m) == SF_FLUID : fluid;
        using T = typename std::conditional<FP32, float, double>::type;
        T Fx, Fy;
        capped = collision_force<T>(a, a.Fcur, n, collide, Fx, Fy);
        if (fluid) {
            const T* src = static_cast<const T*>(a.fsrc);
            T* dst = static_cast<T*>(a.fdst);
            const size_t N = static_cast<size_t>(a.N);
            T f[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) f[q] = ld_nt(src + q * N + n);
            double rho_wall = 0.0;    // the density a moving wall reflects with
            double adr_ux = 0.0, adr_uy = 0.0;   // the scalar step's velocity (after the push)
            if constexpr (!FP32) {
                double r = 0.0, vx = 0.0, vy = 0.0;
#pragma unroll
                for (int q = 0; q < 9; ++q) {
                    double fi = f[q];
                    r += fi;
                    vx += fi * static_cast<double>(dcx(q));
                    vy += fi * static_cast<double>(dcy(q));
                }
                double sx = 0.0, sy = 0.0;
                if (a.shift == SHIFT_PREV) collision_force<double>(a, a.Fshift, n, collide, sx, sy);
                else if (a.shift == SHIFT_PARTIAL) load_force<double>(a.Fshift, n, sx, sy);
                double ux, uy;
                if (a.shift == SHIFT_STORED) {
                    r = a.srho[n];
                    ux = a.su[2 * n];
                    uy = a.su[2 * n + 1];
                } else if (r > 1e-15) {
                    ux = (vx + 0.5 * sx) / r;
                    uy = (vy + 0.5 * sy) / r;
                } else {
                    ux = 0.0;
                    uy = 0.0;
                }
                rho_wall = r;
                adr_ux = ux;
                adr_uy = uy;
                if (!collide) {
                } else if (OP == SF_REGULARIZED) col_reg(f, a.om, r, ux, uy, Fx, Fy);
                else if (OP == SF_MRT) col_mrt(f, a.om, a.se, a.seps, a.sq, r, ux, uy, Fx, Fy);
                else col_bgk(f, a.om, r, ux, uy, Fx, Fy);
            } else {
                float drho = ((f[0] + (f[1] + f[3])) + (f[2] + f[4])) + ((f[5] + f[7]) + (f[6] + f[8]));
                float jx = (f[1] - f[3]) + (f[5] - f[6]) + (f[8] - f[7]);
                float jy = (f[2] - f[4]) + (f[5] + f[6]) - (f[7] + f[8]);
                float sx = 0.f, sy = 0.f;
                if (a.shift == SHIFT_PREV) collision_force<float>(a, a.Fshift, n, collide, sx, sy);
                else if (a.shift == SHIFT_PARTIAL) load_force<float>(a.Fshift, n, sx, sy);
                float rho = 1.0f + drho, ux, uy;
                if (a.shift == SHIFT_STORED) {
                    rho = static_cast<float>(a.srho[n]);
                    drho = static_cast<float>(a.srho[n] - 1.0);
                    ux = static_cast<float>(a.su[2 * n]);
                    uy = static_cast<float>(a.su[2 * n + 1]);
                } else {
                    const float inv = 1.0f / rho;
                    ux = (jx + 0.5f * sx) * inv;
                    uy = (jy + 0.5f * sy) * inv;
                }
                rho_wall = static_cast<double>(rho);
                adr_ux = static_cast<double>(ux);
                adr_uy = static_cast<double>(uy);
                const float om = static_cast<float>(a.om);
                if (!collide) {
                } else if (OP == SF_REGULARIZED) dcol_reg(f, om, drho, rho, ux, uy, Fx, Fy);
                else if (OP == SF_MRT)
                    dcol_mrt(f, om, static_cast<float>(a.se), static_cast<float>(a.seps),
                             static_cast<float>(a.sq), drho, rho, ux, uy, Fx, Fy);
                else dcol_bgk(f, om, drho, rho, ux, uy, Fx, Fy);
            }
            // push with halfway bounce-back (lbm_solver.cpp:212-235, bounce_back.cpp:12-40)
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                int xd = x + dcx(q), yd = y + dcy(q);
                bool oob = false;
                if (a.periodic_x) xd = xd < 0 ? xd + a.nx : (xd >= a.nx ? xd - a.nx : xd);
                else oob |= (xd < 0 || xd >= a.nx);
                if (a.periodic_y) yd = yd < 0 ? yd + a.ny : (yd >= a.ny ? yd - a.ny : yd);
                else oob |= (yd < 0 || yd >= a.ny);
                if (!oob) {
                    if (FLAGS ? mask_solid_nb(m, q) : is_solid(a, xd, yd, false)) {
                        T v = f[q];
                        if (a.moving && (yd == 0 || yd == a.ny - 1)) {
                            // a moving wall row (moving_wall.cpp:20-38): minus
                            // 2 w_q rho (c_q . U) / cs^2, in the reference's order
                            const double U = yd == 0 ? a.wall_bot : a.wall_top;
                            if (fabs(U) > 1e-15) {
                                const double eu = static_cast<double>(dcx(q)) * U;
                                v = static_cast<T>(static_cast<double>(f[q]) -
                                                   2.0 * dw(q) * rho_wall * eu / (1.0 / 3.0));
                            }
                        }
                        st_nt(dst + dopp(q) * N + n, v);
                    } else {
                        st_nt(dst + q * N + (yd * a.nx + xd), f[q]);
                    }
                }
                if (q != 0) {
                    const int xs = x - dcx(q), ys = y - dcy(q);
                    const bool soob = (!a.periodic_x && (xs < 0 || xs >= a.nx)) ||
                                      (!a.periodic_y && (ys < 0 || ys >= a.ny));
                    if (soob) st_nt(dst + q * N + n, f[q]);
                }
            }
            // the scalar step after the populations have left, so the two sets
            // are never held at once
            if constexpr (ADRF) adr_cell<FP32, FLAGS>(a, x, y, n, adr_ux, adr_uy, m);
        }
    }
    return capped;
}

template <int OP, bool FP32, bool FLAGS, bool ADRF>
__global__ __launch_bounds__(256) void k_lbm(LbmArgs a) {
    int capped = 0;
    if (a.persist == 0) {
        int x, y;
        if (!block_origin(a, blockIdx.x, x, y)) return;
        capped = lbm_cell<OP, FP32, FLAGS, ADRF>(a, x, y);
    } else {
        const int b = blockIdx.x, xcd = b % 8, l = b / 8;
        const int t_end = min((xcd + 1) * a.per_xcd, a.nblk);
        for (int t = xcd * a.per_xcd + l; t < t_end; t += a.persist) {
            const int x = (t % a.nbx) * blockDim.x + threadIdx.x;
            const int y = (t / a.nbx) * blockDim.y + threadIdx.y;
            capped += lbm_cell<OP, FP32, FLAGS, ADRF>(a, x, y);
        }
    }
    // count capped cells: one atomic per wavefront
    for (int off = 32; off > 0; off >>= 1) capped += __shfl_down(capped, off);
    if ((threadIdx.x + threadIdx.y * blockDim.x) % warpSize == 0 && capped)
        atomicAdd(&a.counters[0], capped);
}

// lanes_x: lanes across a row (nx, or nx / 2 for the paired collide-stream)
static void set_geometry(const Ctx& c, LbmArgs& a, int bx, int by, int remap, int persist, dim3& g,
                         dim3& b, int lanes_x = 0) {
    b = dim3(bx, by);
    a.nbx = ((lanes_x > 0 ? lanes_x : c.nx) + bx - 1) / bx;
    const int nby = (c.ny + by - 1) / by;
    a.nblk = a.nbx * nby;
    a.per_xcd = (a.nblk + 7) / 8;
    a.remap = remap;
    a.persist = persist > 0 && persist < a.per_xcd ? persist : 0;
    g = dim3(a.persist ? 8 * a.persist : (remap ? 8 * a.per_xcd : a.nblk));
}


// ---------------------------------------------------------------------------
// Paired collide-stream for FP32 channels: each lane takes the cells (x, x+1),
// x even, of a lattice that is periodic in x with no x walls, no generic cell
// types and no moving walls.  There every solid, bounce-back and out-of-bounds
// decision depends on the row alone, so the two cells share it: their loads,
// the unshifted pushes, the bounce-backs and the self-copies move as aligned
// 8-byte pairs, and the pushes along x as 8-byte pairs one cell over (two
// 4-byte stores where one of them wraps).  The arithmetic of each cell is lbm_cell's FP32 branch and
// adr_cell's FP32 relaxation, expression for expression, so results are
// bit-identical to one cell per lane.  With 4-byte lanes a wave moves half the
// bytes per memory instruction, which holds a streaming kernel well below the
// bandwidth the same traffic reaches with 8-byte lanes.
typedef float pair_f2 __attribute__((ext_vector_type(2)));

// the same pair at a 4-byte-aligned address (the pushes along x, one 8-byte
// store where neither cell's push wraps; the code object runs in the
// runtime's unaligned access mode)
typedef float pair_f2u __attribute__((ext_vector_type(2), aligned(4)));
__device__ __forceinline__ void st_pair_u(float* p, float v0, float v1) {
    pair_f2u v;
    v.x = v0;
    v.y = v1;
    __builtin_nontemporal_store(v, reinterpret_cast<pair_f2u*>(p));
}

__device__ __forceinline__ pair_f2 ld_pair(const float* p) {
    return __builtin_nontemporal_load(reinterpret_cast<const pair_f2*>(p));
}
__device__ __forceinline__ void st_pair(float* p, float v0, float v1) {
    pair_f2 v;
    v.x = v0;
    v.y = v1;
    __builtin_nontemporal_store(v, reinterpret_cast<pair_f2*>(p));
}

// The raw forces of buffer F at cells n and n + 1 (n even: one 16-byte load).
__device__ __forceinline__ float4 pair_force_raw(const void* F, int n) {
    return *reinterpret_cast<const float4*>(static_cast<const float*>(F) + 2 * static_cast<size_t>(n));
}

// The pair's collision forces (collision_force of each cell; c0/c1: the cells
// take the body force) and the forces of their collision velocities (the
// shift: the previous step's collision force, or the raw partial force), each
// buffer's pair in one 16-byte load.  Returns the number of capped collision
// forces.
__device__ __forceinline__ int pair_forces(const LbmArgs& a, int n, bool c0, bool c1, float (&F)[4],
                                           float (&S)[4]) {
    const float4 fc = pair_force_raw(a.Fcur, n);
    F[0] = fc.x; F[1] = fc.y; F[2] = fc.z; F[3] = fc.w;
    int capped = force_finish<float>(a, c0, F[0], F[1]);
    capped += force_finish<float>(a, c1, F[2], F[3]);
    S[0] = 0.f; S[1] = 0.f; S[2] = 0.f; S[3] = 0.f;
    if (a.shift == SHIFT_PREV || a.shift == SHIFT_PARTIAL) {
        const float4 fs = pair_force_raw(a.Fshift, n);
        S[0] = fs.x; S[1] = fs.y; S[2] = fs.z; S[3] = fs.w;
        if (a.shift == SHIFT_PREV) {
            force_finish<float>(a, c0, S[0], S[1]);
            force_finish<float>(a, c1, S[2], S[3]);
        }
    }
    return capped;
}

// lbm_cell's FP32 macroscopic state and collision of one cell; (sx, sy) is
// the force of its collision velocity (pair_forces)
template <int OP>
__device__ __forceinline__ void pair_collide(const LbmArgs& a, float (&f)[9], int n, bool collide, float Fx,
                                             float Fy, float sx, float sy, float& ux, float& uy) {
    float drho = ((f[0] + (f[1] + f[3])) + (f[2] + f[4])) + ((f[5] + f[7]) + (f[6] + f[8]));
    float jx = (f[1] - f[3]) + (f[5] - f[6]) + (f[8] - f[7]);
    float jy = (f[2] - f[4]) + (f[5] + f[6]) - (f[7] + f[8]);
    float rho = 1.0f + drho;
    if (a.shift == SHIFT_STORED) {
        rho = static_cast<float>(a.srho[n]);
        drho = static_cast<float>(a.srho[n] - 1.0);
        ux = static_cast<float>(a.su[2 * n]);
        uy = static_cast<float>(a.su[2 * n + 1]);
    } else {
        const float inv = 1.0f / rho;
        ux = (jx + 0.5f * sx) * inv;
        uy = (jy + 0.5f * sy) * inv;
    }
    const float om = static_cast<float>(a.om);
    if (!collide) {
    } else if (OP == SF_REGULARIZED) dcol_reg(f, om, drho, rho, ux, uy, Fx, Fy);
    else if (OP == SF_MRT)
        dcol_mrt(f, om, static_cast<float>(a.se), static_cast<float>(a.seps), static_cast<float>(a.sq), drho,
                 rho, ux, uy, Fx, Fy);
    else dcol_bgk(f, om, drho, rho, ux, uy, Fx, Fy);
}

// adr_cell's FP32 relaxation of one cell (INLET cells: its FP64 form towards
// the inlet concentration)
__device__ __forceinline__ void pair_relax(const LbmArgs& a, int s, float (&g)[9], int n, double ux, double uy,
                                           bool inlet = false) {
    double C;
    if (a.from_store) {
        C = a.Cst[s][n];
    } else {
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < 9; ++q) acc += static_cast<double>(g[q]);
        C = acc;
    }
    if (inlet) {
        C = a.cin[s];
        const double oms = a.oms[s];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const double eu = static_cast<double>(dcx(q)) * ux + static_cast<double>(dcy(q)) * uy;
            const double ge = dw(q) * C * (1.0 + eu * 3.0);
            g[q] = static_cast<float>(static_cast<double>(g[q]) - oms * (static_cast<double>(g[q]) - ge));
        }
        return;
    }
    const float Cf = static_cast<float>(C), omf = static_cast<float>(a.oms[s]);
    double moving = 0.0;
#pragma unroll
    for (int q = 1; q < 9; ++q) {
        const float eu = static_cast<float>(dcx(q) * ux + dcy(q) * uy);
        const float ge = static_cast<float>(dw(q)) * Cf * (1.0f + 3.0f * eu);
        g[q] = g[q] - omf * (g[q] - ge);
        moving += static_cast<double>(g[q]);
    }
    g[0] = static_cast<float>(C - moving);
}

__device__ __forceinline__ int wrap_x(int x, int nx) { return x < 0 ? x + nx : (x >= nx ? x - nx : x); }

// One cell's store of population q: the push to (x + cx, yd), or the bounce
// into its own opposite slot (generic cell types, where the pair's cells may
// differ).
__device__ __forceinline__ void pair_store_one(float* d, size_t N, int n, int x, int nx, int q, int yd, bool bounce,
                                               float v) {
    if (bounce) st_nt(d + dopp(q) * N + n, v);
    else st_nt(d + q * N + (static_cast<size_t>(yd) * nx + wrap_x(x + dcx(q), nx)), v);
}

// The pair's stores of population q when both cells push (or both bounce):
// 8-byte pairs as in the channel form.
__device__ __forceinline__ void pair_store_both(float* d, size_t N, int n, int x, int nx, int q, int yd, bool bounce,
                                                float v0, float v1) {
    if (bounce) {
        st_pair(d + dopp(q) * N + n, v0, v1);
    } else if (dcx(q) == 0) {
        st_pair(d + q * N + (static_cast<size_t>(yd) * nx + x), v0, v1);
    } else {
        const size_t row = static_cast<size_t>(yd) * nx;
        const int xa = x + dcx(q);
        if (xa >= 0 && xa + 1 < nx) {
            st_pair_u(d + q * N + row + xa, v0, v1);
        } else {
            st_nt(d + q * N + row + wrap_x(x + dcx(q), nx), v0);
            st_nt(d + q * N + row + wrap_x(x + 1 + dcx(q), nx), v1);
        }
    }
}


// Generic cell types: each cell of the pair takes its own type and neighbour
// mask (lbm_cell with FLAGS); where both cells are fluid and decide a
// direction alike, the two stores still move as one pair.
template <int OP, bool ADRF>
__device__ __forceinline__ int lbm_pair_types(const LbmArgs& a, int x, int y, int n) {
    const uint32_t mm = *reinterpret_cast<const uint32_t*>(a.nbm + n);   // n even: 4-byte aligned
    const unsigned m0 = mm & 0xffffu, m1 = mm >> 16;
    const bool fl0 = mask_type(m0) != SF_SOLID, fl1 = mask_type(m1) != SF_SOLID;
    const bool co0 = mask_type(m0) == SF_FLUID, co1 = mask_type(m1) == SF_FLUID;
    float F[4], S[4];
    const int capped = pair_forces(a, n, co0, co1, F, S);
    if (!fl0 && !fl1) return capped;
    const float* src = static_cast<const float*>(a.fsrc);
    float* dst = static_cast<float*>(a.fdst);
    const size_t N = static_cast<size_t>(a.N);
    float f0[9], f1[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const pair_f2 v = ld_pair(src + q * N + n);
        f0[q] = v.x;
        f1[q] = v.y;
    }
    float ux0 = 0.f, uy0 = 0.f, ux1 = 0.f, uy1 = 0.f;
    if (fl0) pair_collide<OP>(a, f0, n, co0, F[0], F[1], S[0], S[1], ux0, uy0);
    if (fl1) pair_collide<OP>(a, f1, n + 1, co1, F[2], F[3], S[2], S[3], ux1, uy1);
    const bool both = fl0 && fl1;
    // the populations leave before the scalar's are loaded, so the two sets are
    // never held at once
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        int yd = y + dcy(q);
        bool oob = false;
        if (a.periodic_y) yd = yd < 0 ? yd + a.ny : (yd >= a.ny ? yd - a.ny : yd);
        else oob = (yd < 0 || yd >= a.ny);
        if (!oob) {
            const bool b0 = mask_solid_nb(m0, q), b1 = mask_solid_nb(m1, q);
            if (both && b0 == b1) {
                pair_store_both(dst, N, n, x, a.nx, q, yd, b0, f0[q], f1[q]);
            } else {
                if (fl0) pair_store_one(dst, N, n, x, a.nx, q, yd, b0, f0[q]);
                if (fl1) pair_store_one(dst, N, n + 1, x + 1, a.nx, q, yd, b1, f1[q]);
            }
        }
        if (q != 0) {
            const int ys = y - dcy(q);
            if (!a.periodic_y && (ys < 0 || ys >= a.ny)) {
                if (both) st_pair(dst + q * N + n, f0[q], f1[q]);
                else if (fl0) st_nt(dst + q * N + n, f0[q]);
                else st_nt(dst + q * N + n + 1, f1[q]);
            }
        }
    }
    if constexpr (ADRF) {
        for (int s = 0; s < a.nsp; ++s) {
            const float* gs = static_cast<const float*>(a.gsrc[s]);
            float* gd = static_cast<float*>(a.gdst[s]);
            float g0[9], g1[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                const pair_f2 v = ld_pair(gs + q * N + n);
                g0[q] = v.x;
                g1[q] = v.y;
            }
            if (fl0) pair_relax(a, s, g0, n, static_cast<double>(ux0), static_cast<double>(uy0),
                                mask_type(m0) == SF_INLET);
            if (fl1) pair_relax(a, s, g1, n + 1, static_cast<double>(ux1), static_cast<double>(uy1),
                                mask_type(m1) == SF_INLET);
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                int yd = y + dcy(q);
                bool yout = false;
                if (a.scalar_periodic_y) yd = yd < 0 ? yd + a.ny : (yd >= a.ny ? yd - a.ny : yd);
                else yout = yd < 0 || yd >= a.ny;
                const bool b0 = yout || mask_solid_nb(m0, q), b1 = yout || mask_solid_nb(m1, q);
                if (both && b0 == b1) {
                    pair_store_both(gd, N, n, x, a.nx, q, yd, b0, g0[q], g1[q]);
                } else {
                    if (fl0) pair_store_one(gd, N, n, x, a.nx, q, yd, b0, g0[q]);
                    if (fl1) pair_store_one(gd, N, n + 1, x + 1, a.nx, q, yd, b1, g1[q]);
                }
            }
        }
    }
    return capped;
}

template <int OP, bool ADRF, bool FLAGS>
__device__ __forceinline__ int lbm_pair(const LbmArgs& a, int x, int y) {
    if (x >= a.nx || y >= a.ny) return 0;     // nx is even: x + 1 < nx as well
    const int n = y * a.nx + x;
    if constexpr (FLAGS) return lbm_pair_types<OP, ADRF>(a, x, y, n);
    const bool fluid = !is_solid(a, x, y, false);
    const float* src = static_cast<const float*>(a.fsrc);
    float* dst = static_cast<float*>(a.fdst);
    const size_t N = static_cast<size_t>(a.N);
    float f0[9], f1[9];
    if (fluid) {
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const pair_f2 v = ld_pair(src + q * N + n);
            f0[q] = v.x;
            f1[q] = v.y;
        }
    }
    float F[4], S[4];
    const int capped = pair_forces(a, n, fluid, fluid, F, S);
    if (!fluid) return capped;
    float ux0, uy0, ux1, uy1;
    pair_collide<OP>(a, f0, n, true, F[0], F[1], S[0], S[1], ux0, uy0);
    pair_collide<OP>(a, f1, n + 1, true, F[2], F[3], S[2], S[3], ux1, uy1);
    const int xp[2] = {x, x + 1};
    // push with halfway bounce-back (lbm_solver.cpp:212-235, bounce_back.cpp:12-40); the
    // populations leave before the scalar's are loaded, so the two sets are
    // never held at once
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        int yd = y + dcy(q);
        bool oob = false;
        if (a.periodic_y) yd = yd < 0 ? yd + a.ny : (yd >= a.ny ? yd - a.ny : yd);
        else oob = (yd < 0 || yd >= a.ny);
        if (!oob) {
            if (is_solid(a, 0, yd, false)) {
                st_pair(dst + dopp(q) * N + n, f0[q], f1[q]);
            } else if (dcx(q) == 0) {
                st_pair(dst + q * N + (static_cast<size_t>(yd) * a.nx + x), f0[q], f1[q]);
            } else {
                const size_t row = static_cast<size_t>(yd) * a.nx;
                const int xa = xp[0] + dcx(q);
                if (xa >= 0 && xa + 1 < a.nx) {
                    st_pair_u(dst + q * N + row + xa, f0[q], f1[q]);
                } else {
                    st_nt(dst + q * N + row + wrap_x(xp[0] + dcx(q), a.nx), f0[q]);
                    st_nt(dst + q * N + row + wrap_x(xp[1] + dcx(q), a.nx), f1[q]);
                }
            }
        }
        if (q != 0) {
            const int ys = y - dcy(q);
            if (!a.periodic_y && (ys < 0 || ys >= a.ny)) st_pair(dst + q * N + n, f0[q], f1[q]);
        }
    }
    if constexpr (ADRF) {
        for (int s = 0; s < a.nsp; ++s) {
            const float* gs = static_cast<const float*>(a.gsrc[s]);
            float* gd = static_cast<float*>(a.gdst[s]);
            float g0[9], g1[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                const pair_f2 v = ld_pair(gs + q * N + n);
                g0[q] = v.x;
                g1[q] = v.y;
            }
            pair_relax(a, s, g0, n, static_cast<double>(ux0), static_cast<double>(uy0));
            pair_relax(a, s, g1, n + 1, static_cast<double>(ux1), static_cast<double>(uy1));
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                int yd = y + dcy(q);
                bool bounce;
                if (a.scalar_periodic_y) {
                    yd = yd < 0 ? yd + a.ny : (yd >= a.ny ? yd - a.ny : yd);
                    bounce = is_solid(a, 0, yd, false);
                } else {
                    bounce = (yd < 0 || yd >= a.ny) || is_solid(a, 0, yd, false);
                }
                if (bounce) {
                    st_pair(gd + dopp(q) * N + n, g0[q], g1[q]);
                } else if (dcx(q) == 0) {
                    st_pair(gd + q * N + (static_cast<size_t>(yd) * a.nx + x), g0[q], g1[q]);
                } else {
                    const size_t row = static_cast<size_t>(yd) * a.nx;
                    const int xa = xp[0] + dcx(q);
                    if (xa >= 0 && xa + 1 < a.nx) {
                        st_pair_u(gd + q * N + row + xa, g0[q], g1[q]);
                    } else {
                        st_nt(gd + q * N + row + wrap_x(xp[0] + dcx(q), a.nx), g0[q]);
                        st_nt(gd + q * N + row + wrap_x(xp[1] + dcx(q), a.nx), g1[q]);
                    }
                }
            }
        }
    }
    return capped;
}

template <int OP, bool ADRF, bool FLAGS>
__device__ __forceinline__ void lbm_pair_grid(const LbmArgs& a) {
    int capped = 0;
    if (a.persist == 0) {
        const int b = blockIdx.x;
        const int t = a.remap ? (b % 8) * a.per_xcd + b / 8 : b;
        if (t >= a.nblk) return;
        const int x = 2 * ((t % a.nbx) * blockDim.x + threadIdx.x);
        const int y = (t / a.nbx) * blockDim.y + threadIdx.y;
        capped = lbm_pair<OP, ADRF, FLAGS>(a, x, y);
    } else {
        const int b = blockIdx.x, xcd = b % 8, l = b / 8;
        const int t_end = min((xcd + 1) * a.per_xcd, a.nblk);
        for (int t = xcd * a.per_xcd + l; t < t_end; t += a.persist) {
            const int x = 2 * ((t % a.nbx) * blockDim.x + threadIdx.x);
            const int y = (t / a.nbx) * blockDim.y + threadIdx.y;
            capped += lbm_pair<OP, ADRF, FLAGS>(a, x, y);
        }
    }
    for (int off = 32; off > 0; off >>= 1) capped += __shfl_down(capped, off);
    if ((threadIdx.x + threadIdx.y * blockDim.x) % warpSize == 0 && capped)
        atomicAdd(&a.counters[0], capped);
}

// Channels: at least four waves per SIMD.  With the populations stored before
// the scalar's are loaded, the MRT form with the scalar holds 88 VGPRs (five
// waves, no spill); held to six waves it spills and runs slower.
template <int OP, bool ADRF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_lbm_pair(LbmArgs a) {
    lbm_pair_grid<OP, ADRF, false>(a);
}

// Generic cell types: not held to four waves.  With that bound the MRT form
// with the scalar spills 96 bytes per lane and runs slower, so the allocator
// keeps its own count (154 VGPRs, three waves per SIMD).
template <int OP, bool ADRF>
__global__ __launch_bounds__(256) void k_lbm_pair_types(LbmArgs a) {
    lbm_pair_grid<OP, ADRF, true>(a);
}

// The paired form applies: FP32, x periodic without x walls, no moving walls,
// nx even (generic cell types take the per-cell form, lbm_pair_types).
bool lbm_pair_ok(const Ctx& c) {
    const bool moving = fabs(c.p.bottom_wall_velocity) > 1e-15 || fabs(c.p.top_wall_velocity) > 1e-15;
    return c.fp32 && c.periodic_x && !c.walls_x && !moving && c.nx % 2 == 0;
}

template <int OP, bool FP32>
static void lbm_dispatch_flags(Ctx& c, const LbmArgs& a, dim3 g, dim3 b, bool adr) {
    if (adr) {
        if (c.generic_flags) hipLaunchKernelGGL((k_lbm<OP, FP32, true, true>), g, b, 0, c.stream, a);
        else hipLaunchKernelGGL((k_lbm<OP, FP32, false, true>), g, b, 0, c.stream, a);
    } else {
        if (c.generic_flags) hipLaunchKernelGGL((k_lbm<OP, FP32, true, false>), g, b, 0, c.stream, a);
        else hipLaunchKernelGGL((k_lbm<OP, FP32, false, false>), g, b, 0, c.stream, a);
    }
}

static void scalar_args(Ctx& c, LbmArgs& a) {
    a.nsp = c.nsp;
    a.from_store = c.c_from_store ? 1 : 0;
    a.scalar_periodic_y = c.p.scalar_periodic_y;
    for (int s = 0; s < c.nsp; ++s) {
        a.gsrc[s] = c.g[s][c.gcur];
        a.gdst[s] = c.g[s][1 - c.gcur];
        a.Cst[s] = c.Cstore[s];
        a.oms[s] = 1.0 / (3.0 * c.p.diffusivity[s] + 0.5);
        a.cin[s] = c.p.inlet_concentration[s];
    }
}

static LbmArgs lbm_args(Ctx& c, int shift, bool fuse_adr) {
    LbmArgs a{};
    a.fsrc = c.f[c.fcur];
    a.fdst = c.f[1 - c.fcur];
    a.Fcur = c.F[c.Fcur];
    a.Fshift = shift == SHIFT_PARTIAL ? c.F[c.Fpart] : c.F[c.Fprev];
    a.flags = c.flags;
    a.nbm = c.nbm;
    a.srho = c.stored_rho;
    a.su = c.stored_u;
    a.counters = c.counters;
    a.nx = c.nx; a.ny = c.ny; a.N = c.N;
    a.periodic_x = c.periodic_x; a.periodic_y = c.periodic_y;
    a.walls_x = c.walls_x; a.walls_y = c.walls_y;
    a.shift = shift;
    a.has_body = (c.p.body_force_x != 0.0 || c.p.body_force_y != 0.0);
    a.bfx = c.p.body_force_x; a.bfy = c.p.body_force_y;
    a.fmax = c.p.max_lattice_force; a.fmax2 = a.fmax * a.fmax;
    a.om = 1.0 / c.p.tau;
    a.se = c.p.mrt_se; a.seps = c.p.mrt_s_eps; a.sq = c.p.mrt_sq;
    a.wall_bot = c.p.bottom_wall_velocity;
    a.wall_top = c.p.top_wall_velocity;
    a.moving = (fabs(a.wall_bot) > 1e-15 || fabs(a.wall_top) > 1e-15) ? 1 : 0;
    a.inlet_ux = c.p.inlet_velocity;
    a.outlet_rho = c.p.outlet_density;
    if (fuse_adr) scalar_args(c, a);
    return a;
}

void launch_lbm_cfg(Ctx& c, int shift, bool fuse_adr, int bx, int by, int remap, int persist, int pair) {
    LbmArgs a = lbm_args(c, shift, fuse_adr);
    dim3 b, g;
    const int op = c.p.collision_model;
    // pair: 1 channels, 2 also generic cell types (lbm_pair_types)
    if (pair && lbm_pair_ok(c) && (!c.generic_flags || pair == 2)) {
        set_geometry(c, a, bx, by, remap, persist, g, b, c.nx / 2);
        auto go = [&](auto kern) { hipLaunchKernelGGL(kern, g, b, 0, c.stream, a); };
        auto by_op = [&](auto adr, auto flags) {
            constexpr bool AD = decltype(adr)::value, FL = decltype(flags)::value;
            if constexpr (FL) {
                if (op == SF_REGULARIZED) go(k_lbm_pair_types<SF_REGULARIZED, AD>);
                else if (op == SF_MRT) go(k_lbm_pair_types<SF_MRT, AD>);
                else go(k_lbm_pair_types<SF_BGK, AD>);
            } else {
                if (op == SF_REGULARIZED) go(k_lbm_pair<SF_REGULARIZED, AD>);
                else if (op == SF_MRT) go(k_lbm_pair<SF_MRT, AD>);
                else go(k_lbm_pair<SF_BGK, AD>);
            }
        };
        using T_ = std::true_type;
        using F_ = std::false_type;
        if (fuse_adr) {
            if (c.generic_flags) by_op(T_{}, T_{});
            else by_op(T_{}, F_{});
        } else {
            if (c.generic_flags) by_op(F_{}, T_{});
            else by_op(F_{}, F_{});
        }
        return;
    }
    set_geometry(c, a, bx, by, remap, persist, g, b);
    if (c.fp32) {
        if (op == SF_REGULARIZED) lbm_dispatch_flags<SF_REGULARIZED, true>(c, a, g, b, fuse_adr);
        else if (op == SF_MRT) lbm_dispatch_flags<SF_MRT, true>(c, a, g, b, fuse_adr);
        else lbm_dispatch_flags<SF_BGK, true>(c, a, g, b, fuse_adr);
    } else {
        if (op == SF_REGULARIZED) lbm_dispatch_flags<SF_REGULARIZED, false>(c, a, g, b, fuse_adr);
        else if (op == SF_MRT) lbm_dispatch_flags<SF_MRT, false>(c, a, g, b, fuse_adr);
        else lbm_dispatch_flags<SF_BGK, false>(c, a, g, b, fuse_adr);
    }
}

// ZouHeBoundary::apply (lbm/zou_he_boundary.cpp:18-75) on the streamed
// populations: a parabolic velocity inlet on the INLET cells of column 0 and a
// fixed-density outlet on the OUTLET cells of column nx-1, each rebuilding the
// three populations that enter from outside (one thread per row, FP64).
template <bool FP32>
__global__ void k_zou_he(LbmArgs a) {
    using T = typename std::conditional<FP32, float, double>::type;
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= a.ny) return;
    T* fd = static_cast<T*>(a.fdst);
    const size_t N = static_cast<size_t>(a.N);
    // full populations (FP32 storage holds f - w)
    auto get = [&](int x, int q) -> double {
        const double v = static_cast<double>(fd[q * N + static_cast<size_t>(y) * a.nx + x]);
        return FP32 ? v + dw(q) : v;
    };
    auto put = [&](int x, int q, double v) {
        fd[q * N + static_cast<size_t>(y) * a.nx + x] = static_cast<T>(FP32 ? v - dw(q) : v);
    };
    if (a.flags[static_cast<size_t>(y) * a.nx] == SF_INLET) {
        const int x = 0;
        const double y_center = static_cast<double>(a.ny) * 0.5;
        const double H = static_cast<double>(a.ny - 2);
        const double y_norm = (static_cast<double>(y) - y_center) / (H * 0.5);
        const double ux = a.inlet_ux * (1.0 - y_norm * y_norm);
        const double uy = 0.0;
        const double f0 = get(x, 0), f2 = get(x, 2), f3 = get(x, 3), f4 = get(x, 4);
        const double f6 = get(x, 6), f7 = get(x, 7);
        const double rho = (f0 + f2 + f4 + 2.0 * (f3 + f6 + f7)) / (1.0 - ux);
        put(x, 1, f3 + (2.0 / 3.0) * rho * ux);
        put(x, 5, f7 + 0.5 * (f4 - f2) + (1.0 / 6.0) * rho * ux + 0.5 * rho * uy);
        put(x, 8, f6 - 0.5 * (f4 - f2) + (1.0 / 6.0) * rho * ux - 0.5 * rho * uy);
    }
    if (a.flags[static_cast<size_t>(y) * a.nx + a.nx - 1] == SF_OUTLET) {
        const int x = a.nx - 1;
        const double rho = a.outlet_rho;
        const double f0 = get(x, 0), f1 = get(x, 1), f2 = get(x, 2), f4 = get(x, 4);
        const double f5 = get(x, 5), f8 = get(x, 8);
        const double ux = -1.0 + (f0 + f2 + f4 + 2.0 * (f1 + f5 + f8)) / rho;
        const double uy = 0.0;
        put(x, 3, f1 - (2.0 / 3.0) * rho * ux);
        put(x, 7, f5 - 0.5 * (f4 - f2) - (1.0 / 6.0) * rho * ux - 0.5 * rho * uy);
        put(x, 6, f8 + 0.5 * (f4 - f2) - (1.0 / 6.0) * rho * ux + 0.5 * rho * uy);
    }
}

// InterpolatedBounceBack::apply (lbm/interpolated_bounce_back.cpp:55-96) on the
// streamed populations.  A group is one cell's links in direction order, the
// order of the reference's link list within a cell.  Groups [0, par) read and
// write only their own cell's slots and slots no link writes: one thread each.
// The rest (cells that read a slot another cell's link writes: fluid runs two
// cells long between solids) run on one thread in the reference's cell order.
// FP32 storage holds f - w: the update is affine with weights summing to one
// and w is equal for opposite directions, so it applies to f - w unchanged.
template <typename T>
__device__ __forceinline__ void ibb_group(T* f, size_t N, int cell, int l0, int l1, const int* lq,
                                          const int* lup, const double* lqf) {
    for (int l = l0; l < l1; ++l) {
        const int q = lq[l], o = dopp(q);
        const double qf = lqf[l];
        const double fq = static_cast<double>(f[q * N + cell]);
        double r;
        if (qf < 0.5) {
            const int u = lup[l];
            if (u >= 0) r = 2.0 * qf * fq + (1.0 - 2.0 * qf) * static_cast<double>(f[q * N + u]);
            else r = fq;
        } else {
            const double inv2q = 1.0 / (2.0 * qf);
            r = inv2q * fq + (1.0 - inv2q) * static_cast<double>(f[o * N + cell]);
        }
        f[o * N + cell] = static_cast<T>(r);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_ibb(T* f, int N, int g0, int g1, int serial, const int* gcell,
                                             const int* goff, const int* lq, const int* lup,
                                             const double* lqf) {
    if (serial) {
        if (blockIdx.x != 0 || threadIdx.x != 0) return;
        for (int g = g0; g < g1; ++g)
            ibb_group<T>(f, static_cast<size_t>(N), gcell[g], goff[g], goff[g + 1], lq, lup, lqf);
        return;
    }
    const int g = g0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= g1) return;
    ibb_group<T>(f, static_cast<size_t>(N), gcell[g], goff[g], goff[g + 1], lq, lup, lqf);
}

void launch_ibb(Ctx& c) {
    if (c.ibb_groups == 0) return;
    void* f = c.f[1 - c.fcur];
    auto run = [&](auto* fp) {
        using T = std::remove_pointer_t<decltype(fp)>;
        if (c.ibb_par > 0)
            hipLaunchKernelGGL(k_ibb<T>, dim3((c.ibb_par + 255) / 256), dim3(256), 0, c.stream, fp, c.N, 0,
                               c.ibb_par, 0, c.ibb_cell, c.ibb_off, c.ibb_q, c.ibb_up, c.ibb_qf);
        if (c.ibb_groups > c.ibb_par)
            hipLaunchKernelGGL(k_ibb<T>, dim3(1), dim3(64), 0, c.stream, fp, c.N, c.ibb_par, c.ibb_groups, 1,
                               c.ibb_cell, c.ibb_off, c.ibb_q, c.ibb_up, c.ibb_qf);
    };
    if (c.fp32) run(static_cast<float*>(f));
    else run(static_cast<double*>(f));
}

void launch_lbm(Ctx& c, int shift, bool fuse_adr) {
    launch_lbm_cfg(c, shift, fuse_adr, c.lbm_bx, c.lbm_by, c.lbm_remap, c.lbm_persist, c.lbm_pair);
    launch_ibb(c);
    if (c.inlet_outlet) {
        LbmArgs a = lbm_args(c, shift, false);
        a.fdst = c.f[1 - c.fcur];
        const dim3 g((c.ny + 63) / 64), b(64);
        if (c.fp32) hipLaunchKernelGGL(k_zou_he<true>, g, b, 0, c.stream, a);
        else hipLaunchKernelGGL(k_zou_he<false>, g, b, 0, c.stream, a);
    }
    c.fcur = 1 - c.fcur;
    if (fuse_adr) {
        c.gcur = 1 - c.gcur;
        c.c_from_store = false;
    }
}

// Batched collide-stream over the members of an ensemble: workgroup b runs
// workgroup map[b].y of member map[b].x's own linear grid (the map itself
// orders the workgroups across XCDs).  ADRF: the members' lagged scalar steps
// ride along (a member without a scalar has nsp = 0 and runs none).
template <int OP, bool FP32, bool ADRF>
__global__ __launch_bounds__(256) void kb_lbm(const LbmArgs* packs, const int2* map) {
    const int2 w = map[blockIdx.x];
    if (w.x < 0) return;
    const LbmArgs a = packs[w.x];
    int capped;
    // flat: lane l of workgroup w.y takes cell (pair) w.y * threads + l of the
    // row-major index; lanes past the lattice get x = nx and do nothing
    const int lane = threadIdx.y * blockDim.x + threadIdx.x;
    const int threads = blockDim.x * blockDim.y;
    if (FP32 && a.pair) {
        if (w.y >= a.nblk) return;
        int x, y;
        if (a.flat) {
            const int n = 2 * (w.y * threads + lane);      // nx even: a pair stays in its row
            x = n < a.N ? n % a.nx : a.nx;
            y = n < a.N ? n / a.nx : 0;
        } else {
            x = 2 * ((w.y % a.nbx) * blockDim.x + threadIdx.x);
            y = (w.y / a.nbx) * blockDim.y + threadIdx.y;
        }
        capped = lbm_pair<OP, ADRF, false>(a, x, y);
    } else {
        int x, y;
        if (a.flat) {
            if (w.y >= a.nblk) return;
            const int n = w.y * threads + lane;
            x = n < a.N ? n % a.nx : a.nx;
            y = n < a.N ? n / a.nx : 0;
        } else if (!block_origin(a, w.y, x, y)) {
            return;
        }
        capped = lbm_cell<OP, FP32, false, ADRF>(a, x, y);
    }
    for (int off = 32; off > 0; off >>= 1) capped += __shfl_down(capped, off);
    if ((threadIdx.x + threadIdx.y * blockDim.x) % warpSize == 0 && capped)
        atomicAdd(&a.counters[0], capped);
}

// The steady step's launch (previous-step shift force; with a scalar, the
// lagged scalar step of the previous step fused in, as one_step fuses it) with
// a linear grid of bx x by workgroups; swaps the population (and scalar)
// buffers as launch_lbm does.
void batch_items_lbm(Ctx& c, int bx, int by, int pair, int flat, int shift, BatchItem* it) {
    const bool fuse = c.nsp > 0;
    LbmArgs a = lbm_args(c, shift, fuse);
    dim3 g, b;
    a.pair = pair && lbm_pair_ok(c) && !c.generic_flags ? 1 : 0;
    set_geometry(c, a, bx, by, 0, 0, g, b, a.pair ? c.nx / 2 : 0);
    a.flat = flat ? 1 : 0;
    if (a.flat) {
        const int per_block = bx * by * (a.pair ? 2 : 1);
        a.nblk = (c.N + per_block - 1) / per_block;
    }
    batch_put(it[0], a.nblk, a);
    c.fcur = 1 - c.fcur;
    if (fuse) {
        c.gcur = 1 - c.gcur;
        c.c_from_store = false;
        c.adr_pending = true;       // this step's scalar step runs in the next collide-stream
    }
}

template <bool FP32, bool ADRF>
static void batch_lbm_op(const BatchVariant& v, const LbmArgs* packs, const int2* map, int nblocks,
                         hipStream_t s) {
    const dim3 g(nblocks), b(v.bx, v.by);
    if (v.op == SF_REGULARIZED) hipLaunchKernelGGL((kb_lbm<SF_REGULARIZED, FP32, ADRF>), g, b, 0, s, packs, map);
    else if (v.op == SF_MRT) hipLaunchKernelGGL((kb_lbm<SF_MRT, FP32, ADRF>), g, b, 0, s, packs, map);
    else hipLaunchKernelGGL((kb_lbm<SF_BGK, FP32, ADRF>), g, b, 0, s, packs, map);
}

void batch_launch_lbm(const BatchVariant& v, const void* packs, const int2* map, int nblocks,
                      hipStream_t s) {
    if (nblocks <= 0) return;
    const LbmArgs* p = static_cast<const LbmArgs*>(packs);
    if (v.fp32) {
        if (v.adr) batch_lbm_op<true, true>(v, p, map, nblocks, s);
        else batch_lbm_op<true, false>(v, p, map, nblocks, s);
    } else {
        if (v.adr) batch_lbm_op<false, true>(v, p, map, nblocks, s);
        else batch_lbm_op<false, false>(v, p, map, nblocks, s);
    }
}

// ---------------------------------------------------------------------------
// Velocity of a cell after the LBM step (computeMacroscopic with the step's
// collision force): used by the scalar step and the macroscopic views.
template <bool FP32>
__device__ __forceinline__ void cell_macro(const void* fv, size_t N, int n, const void* F,
                                           const LbmArgs& a, bool zero_force, bool body, double& r,
                                           double& ux, double& uy) {
    if constexpr (!FP32) {
        const double* f = static_cast<const double*>(fv);
        double rr = 0.0, vx = 0.0, vy = 0.0;
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            double fi = f[q * N + n];
            rr += fi;
            vx += fi * static_cast<double>(dcx(q));
            vy += fi * static_cast<double>(dcy(q));
        }
        double sx = 0.0, sy = 0.0;
        if (!zero_force) collision_force<double>(a, F, n, body, sx, sy);
        r = rr;
        if (rr > 1e-15) {
            ux = (vx + 0.5 * sx) / rr;
            uy = (vy + 0.5 * sy) / rr;
        } else {
            ux = 0.0;
            uy = 0.0;
        }
    } else {
        const float* f = static_cast<const float*>(fv);
        float d[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) d[q] = f[q * N + n];
        float drho = ((d[0] + (d[1] + d[3])) + (d[2] + d[4])) + ((d[5] + d[7]) + (d[6] + d[8]));
        float jx = (d[1] - d[3]) + (d[5] - d[6]) + (d[8] - d[7]);
        float jy = (d[2] - d[4]) + (d[5] + d[6]) - (d[7] + d[8]);
        float sx = 0.f, sy = 0.f;
        if (!zero_force) collision_force<float>(a, F, n, body, sx, sy);
        float rho = 1.0f + drho;
        r = rho;
        ux = (jx + 0.5f * sx) / rho;
        uy = (jy + 0.5f * sy) / rho;
    }
}

// Stand-alone advection-diffusion step (advection_diffusion.cpp:84-223): the
// velocity is recomputed from the populations after the LBM step and that
// step's collision force.  SOLID cells are never written: their g stays at its
// initial value in both buffers, as the reference's copy does.
template <bool FP32, bool FLAGS>
__global__ __launch_bounds__(256) void k_adr(LbmArgs a, const void* f) {
    int x, y;
    if (!block_origin(a, blockIdx.x, x, y)) return;
    if (x >= a.nx || y >= a.ny) return;
    const int n = y * a.nx + x;
    if (is_solid(a, x, y, FLAGS)) return;
    double r, ux, uy;
    cell_macro<FP32>(f, static_cast<size_t>(a.N), n, a.Fcur, a, false, is_fluid_type(a, x, y, FLAGS), r,
                     ux, uy);
    adr_cell<FP32, FLAGS>(a, x, y, n, ux, uy, FLAGS ? static_cast<unsigned>(a.nbm[n]) : 0u);
}

static LbmArgs view_args(Ctx& c) {
    LbmArgs a{};
    a.flags = c.flags;
    a.nbm = c.nbm;
    a.nx = c.nx; a.ny = c.ny; a.N = c.N;
    a.periodic_x = c.periodic_x; a.periodic_y = c.periodic_y;
    a.walls_x = c.walls_x; a.walls_y = c.walls_y;
    a.has_body = (c.p.body_force_x != 0.0 || c.p.body_force_y != 0.0);
    a.bfx = c.p.body_force_x; a.bfy = c.p.body_force_y;
    a.fmax = c.p.max_lattice_force; a.fmax2 = a.fmax * a.fmax;
    a.Fcur = c.F[c.Fcur];
    return a;
}

void launch_adr(Ctx& c, int Fidx) {
    LbmArgs a = view_args(c);
    a.Fcur = c.F[Fidx];
    scalar_args(c, a);
    const void* f = c.f[c.fcur];
    dim3 b, g;
    set_geometry(c, a, c.lbm_bx, c.lbm_by, c.lbm_remap, 0, g, b);
    if (c.fp32) {
        if (c.generic_flags) hipLaunchKernelGGL((k_adr<true, true>), g, b, 0, c.stream, a, f);
        else hipLaunchKernelGGL((k_adr<true, false>), g, b, 0, c.stream, a, f);
    } else {
        if (c.generic_flags) hipLaunchKernelGGL((k_adr<false, true>), g, b, 0, c.stream, a, f);
        else hipLaunchKernelGGL((k_adr<false, false>), g, b, 0, c.stream, a, f);
    }
    c.gcur = 1 - c.gcur;
    c.c_from_store = false;
}

// ---------------------------------------------------------------------------
// Macroscopic view: rho/ux/uy/Fx/Fy exactly as LatticeField holds them after
// the last step.  mode: 0 = force of the last collision (after a step),
// 1 = zero force (fresh initialisation), 2 = stored macro arrays.
template <bool FP32, bool FLAGS>
__device__ __forceinline__ void macro_force(const LbmArgs& a, int mode, int n, double& Fx, double& Fy) {
    const int x = n % a.nx, y = n / a.nx;
    const bool solid = is_solid(a, x, y, FLAGS);
    if (mode == 1) {
        Fx = 0.0; Fy = 0.0;
    } else if (FP32) {
        float fx, fy;
        collision_force<float>(a, a.Fcur, n, !solid && is_fluid_type(a, x, y, FLAGS), fx, fy);
        Fx = fx; Fy = fy;
    } else {
        collision_force<double>(a, a.Fcur, n, !solid && is_fluid_type(a, x, y, FLAGS), Fx, Fy);
    }
}

template <bool FP32, bool FLAGS>
__device__ __forceinline__ void macro_state(const LbmArgs& a, const void* f, int mode, int n, double& r,
                                            double& u, double& v) {
    const int x = n % a.nx, y = n / a.nx;
    if (mode == 2) {
        r = a.srho[n]; u = a.su[2 * n]; v = a.su[2 * n + 1];
        return;
    }
    if (is_solid(a, x, y, FLAGS)) {
        r = 0.0; u = 0.0; v = 0.0;
        return;
    }
    cell_macro<FP32>(f, static_cast<size_t>(a.N), n, a.Fcur, a, mode == 1, is_fluid_type(a, x, y, FLAGS), r,
                     u, v);
}

template <bool FP32, bool FLAGS>
__global__ void k_macro(LbmArgs a, const void* f, int mode, double* rho, double* ux, double* uy,
                        double* Fx, double* Fy) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= a.N) return;
    if (Fx) {
        double fx, fy;
        macro_force<FP32, FLAGS>(a, mode, n, fx, fy);
        Fx[n] = fx; Fy[n] = fy;
    }
    if (!rho) return;
    double r, u, v;
    macro_state<FP32, FLAGS>(a, f, mode, n, r, u, v);
    rho[n] = r; ux[n] = u; uy[n] = v;
}

// rho, ux, uy at listed cells, the values k_macro gives there
template <bool FP32, bool FLAGS>
__global__ void k_macro_cells(LbmArgs a, const void* f, int mode, const int* cells, int count, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    double r, u, v;
    macro_state<FP32, FLAGS>(a, f, mode, cells[i], r, u, v);
    out[3 * i] = r; out[3 * i + 1] = u; out[3 * i + 2] = v;
}

void launch_macro_cells(Ctx& c, const int* cells, int count, double* out) {
    if (count <= 0) return;
    LbmArgs a = view_args(c);
    a.srho = c.stored_rho;
    a.su = c.stored_u;
    const int mode = c.shift_next == SHIFT_STORED ? 2 : (c.step_count == 0 ? 1 : 0);
    dim3 b(64), g((count + 63) / 64);
    const void* f = c.f[c.fcur];
    if (c.fp32) {
        if (c.generic_flags) hipLaunchKernelGGL((k_macro_cells<true, true>), g, b, 0, c.stream, a, f, mode, cells, count, out);
        else hipLaunchKernelGGL((k_macro_cells<true, false>), g, b, 0, c.stream, a, f, mode, cells, count, out);
    } else {
        if (c.generic_flags) hipLaunchKernelGGL((k_macro_cells<false, true>), g, b, 0, c.stream, a, f, mode, cells, count, out);
        else hipLaunchKernelGGL((k_macro_cells<false, false>), g, b, 0, c.stream, a, f, mode, cells, count, out);
    }
}

// Sum of n values: 256 fixed partials, then one thread adds them in order
// (deterministic).  ws holds 256 doubles.
__global__ void k_sum_partial(const double* v, int64_t n, double* ws) {
    __shared__ double s[256];
    const int t = threadIdx.x;
    double acc = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + t; i < n; i += 256 * 256) acc += v[i];
    s[t] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) s[t] += s[t + w];
        __syncthreads();
    }
    if (t == 0) ws[blockIdx.x] = s[0];
}

__global__ void k_sum_final(const double* ws, double* out) {
    if (threadIdx.x != 0) return;
    double acc = 0.0;
    for (int b = 0; b < 256; ++b) acc += ws[b];
    *out = acc;
}

void launch_sum(Ctx& c, const double* v, int64_t n, double* ws, double* out) {
    hipLaunchKernelGGL(k_sum_partial, dim3(256), dim3(256), 0, c.stream, v, n, ws);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(64), 0, c.stream, ws, out);
}

void launch_macro(Ctx& c, double* rho, double* ux, double* uy, double* Fx, double* Fy) {
    LbmArgs a = view_args(c);
    a.srho = c.stored_rho;
    a.su = c.stored_u;
    const int mode = c.shift_next == SHIFT_STORED ? 2 : (c.step_count == 0 ? 1 : 0);
    dim3 b(256), g((c.N + 255) / 256);
    const void* f = c.f[c.fcur];
    if (c.fp32) {
        if (c.generic_flags) hipLaunchKernelGGL((k_macro<true, true>), g, b, 0, c.stream, a, f, mode, rho, ux, uy, Fx, Fy);
        else hipLaunchKernelGGL((k_macro<true, false>), g, b, 0, c.stream, a, f, mode, rho, ux, uy, Fx, Fy);
    } else {
        if (c.generic_flags) hipLaunchKernelGGL((k_macro<false, true>), g, b, 0, c.stream, a, f, mode, rho, ux, uy, Fx, Fy);
        else hipLaunchKernelGGL((k_macro<false, false>), g, b, 0, c.stream, a, f, mode, rho, ux, uy, Fx, Fy);
    }
}

// FP64 reductions over FLUID cells: per-block partials in a fixed order, then
// one block folds the partials (deterministic for a given launch shape).
__device__ __forceinline__ void stats_add(double r, double u, double v, double (&acc)[5]) {
    if (!isfinite(r) || !isfinite(u) || !isfinite(v)) { acc[4] += 1.0; return; }
    acc[0] += r; acc[1] += r * u; acc[2] += r * v;
    acc[3] = fmax(acc[3], u * u + v * v);
}

__device__ __forceinline__ void stats_block_store(const double (&acc)[5], double* part) {
    __shared__ double s[5][256];
    const int t = threadIdx.x;
    for (int k = 0; k < 5; ++k) s[k][t] = acc[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            s[0][t] += s[0][t + w]; s[1][t] += s[1][t + w]; s[2][t] += s[2][t + w];
            s[3][t] = fmax(s[3][t], s[3][t + w]); s[4][t] += s[4][t + w];
        }
        __syncthreads();
    }
    if (t == 0)
        for (int k = 0; k < 5; ++k) part[blockIdx.x * 5 + k] = s[k][0];
}

__global__ __launch_bounds__(256) void k_stats_partial(const double* rho, const double* ux, const double* uy,
                                                       const uint8_t* flags, LbmArgs a, double* part) {
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < a.N; n += gridDim.x * blockDim.x) {
        const int x = n % a.nx, y = n / a.nx;
        if (is_solid(a, x, y, flags != nullptr)) continue;
        stats_add(rho[n], ux[n], uy[n], acc);
    }
    stats_block_store(acc, part);
}

// The same partials with each cell's rho and u computed in place, as k_macro
// computes them, instead of read from arrays: the per-cell values and the
// summation order are those of k_stats_partial over k_macro's output, so the
// results are identical, and no N-sized buffer is needed.
template <bool FP32, bool FLAGS>
__global__ __launch_bounds__(256) void k_stats_fused(LbmArgs a, const void* f, int mode, double* part) {
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < a.N; n += gridDim.x * blockDim.x) {
        const int x = n % a.nx, y = n / a.nx;
        if (is_solid(a, x, y, FLAGS)) continue;
        double r, u, v;
        macro_state<FP32, FLAGS>(a, f, mode, n, r, u, v);
        stats_add(r, u, v, acc);
    }
    stats_block_store(acc, part);
}

__global__ void k_stats_final(const double* part, int nb, double* out) {
    if (threadIdx.x != 0) return;
    double m = 0, px = 0, py = 0, um = 0, bad = 0;
    for (int b = 0; b < nb; ++b) {
        m += part[b * 5]; px += part[b * 5 + 1]; py += part[b * 5 + 2];
        um = fmax(um, part[b * 5 + 3]); bad += part[b * 5 + 4];
    }
    out[0] = m; out[1] = px; out[2] = py; out[3] = sqrt(um); out[4] = bad;
}

// Global statistics over the non-solid cells (io/global_stats_writer.cpp:55-80,
// simulation.cpp:820-907), FP64: count, sum rho, min rho, max rho, max |u|^2,
// sum |u|, sum rho |u|^2 / 2, non-finite cells, the density farthest from
// rho0, sum rho ux, sum rho uy.  Non-finite cells are counted and left out of
// the rest.
constexpr int GSTATS = 11;

__device__ __forceinline__ void gstats_combine(double* a, const double* b, double rho0) {
    a[0] += b[0]; a[1] += b[1];
    a[2] = fmin(a[2], b[2]); a[3] = fmax(a[3], b[3]); a[4] = fmax(a[4], b[4]);
    a[5] += b[5]; a[6] += b[6]; a[7] += b[7];
    if (fabs(b[8] - rho0) > fabs(a[8] - rho0)) a[8] = b[8];
    a[9] += b[9]; a[10] += b[10];
}

__global__ void k_gstats_partial(const double* rho, const double* ux, const double* uy,
                                 const uint8_t* flags, LbmArgs a, double rho0, double* part) {
    __shared__ double s[GSTATS][256];
    const int t = threadIdx.x;
    double v[GSTATS] = {0.0, 0.0, 1e300, -1e300, 0.0, 0.0, 0.0, 0.0, rho0, 0.0, 0.0};
    for (int n = blockIdx.x * blockDim.x + t; n < a.N; n += gridDim.x * blockDim.x) {
        const int x = n % a.nx, y = n / a.nx;
        if (is_solid(a, x, y, flags != nullptr)) continue;
        const double r = rho[n], u = ux[n], w = uy[n];
        if (!isfinite(r) || !isfinite(u) || !isfinite(w)) { v[7] += 1.0; continue; }
        const double v2 = u * u + w * w;
        v[0] += 1.0; v[1] += r;
        v[2] = fmin(v[2], r); v[3] = fmax(v[3], r); v[4] = fmax(v[4], v2);
        v[5] += sqrt(v2); v[6] += 0.5 * r * v2;
        if (fabs(r - rho0) > fabs(v[8] - rho0)) v[8] = r;
        v[9] += r * u; v[10] += r * w;
    }
    for (int k = 0; k < GSTATS; ++k) s[k][t] = v[k];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) {
            double a0[GSTATS], b0[GSTATS];
            for (int k = 0; k < GSTATS; ++k) { a0[k] = s[k][t]; b0[k] = s[k][t + h]; }
            gstats_combine(a0, b0, rho0);
            for (int k = 0; k < GSTATS; ++k) s[k][t] = a0[k];
        }
        __syncthreads();
    }
    if (t == 0)
        for (int k = 0; k < GSTATS; ++k) part[blockIdx.x * GSTATS + k] = s[k][0];
}

__global__ void k_gstats_final(const double* part, int nb, double rho0, double* out) {
    if (threadIdx.x != 0) return;
    double a0[GSTATS];
    for (int k = 0; k < GSTATS; ++k) a0[k] = part[k];
    for (int b = 1; b < nb; ++b) gstats_combine(a0, part + b * GSTATS, rho0);
    for (int k = 0; k < GSTATS; ++k) out[k] = a0[k];
}

void launch_global_stats(Ctx& c, double rho0, double* dev_out) {
    // dev_out: GSTATS results (padded to 16), then 3N + GSTATS * nb doubles of scratch
    double* rho = dev_out + 16;
    double* ux = rho + c.N;
    double* uy = ux + c.N;
    double* part = uy + c.N;
    launch_macro(c, rho, ux, uy, nullptr, nullptr);
    LbmArgs a = view_args(c);
    const int nb = 256;
    hipLaunchKernelGGL(k_gstats_partial, dim3(nb), dim3(256), 0, c.stream, rho, ux, uy,
                       c.generic_flags ? c.flags : nullptr, a, rho0, part);
    hipLaunchKernelGGL(k_gstats_final, dim3(1), dim3(64), 0, c.stream, part, nb, rho0, dev_out);
}

void launch_field_stats(Ctx& c, double* dev_out) {
    // dev_out holds 5 results followed by scratch: 3N + 5*nb doubles
    double* rho = dev_out + 8;
    double* ux = rho + c.N;
    double* uy = ux + c.N;
    double* part = uy + c.N;
    launch_macro(c, rho, ux, uy, nullptr, nullptr);
    LbmArgs a = view_args(c);
    const int nb = STATS_BLOCKS;
    hipLaunchKernelGGL(k_stats_partial, dim3(nb), dim3(256), 0, c.stream, rho, ux, uy,
                       c.generic_flags ? c.flags : nullptr, a, part);
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(64), 0, c.stream, part, nb, dev_out);
}

void launch_field_stats_fused(Ctx& c, double* part, double* dev_out) {
    // part: 5 * STATS_BLOCKS doubles; dev_out: 5 doubles
    LbmArgs a = view_args(c);
    a.srho = c.stored_rho;
    a.su = c.stored_u;
    const int mode = c.shift_next == SHIFT_STORED ? 2 : (c.step_count == 0 ? 1 : 0);
    const void* f = c.f[c.fcur];
    const dim3 g(STATS_BLOCKS), b(256);
    if (c.fp32) {
        if (c.generic_flags) hipLaunchKernelGGL((k_stats_fused<true, true>), g, b, 0, c.stream, a, f, mode, part);
        else hipLaunchKernelGGL((k_stats_fused<true, false>), g, b, 0, c.stream, a, f, mode, part);
    } else {
        if (c.generic_flags) hipLaunchKernelGGL((k_stats_fused<false, true>), g, b, 0, c.stream, a, f, mode, part);
        else hipLaunchKernelGGL((k_stats_fused<false, false>), g, b, 0, c.stream, a, f, mode, part);
    }
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(64), 0, c.stream, part, STATS_BLOCKS, dev_out);
}

}  // namespace sf
