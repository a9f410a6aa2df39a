// CPU restatement of SoftFlow's per-timestep hot path (the oracle).
//
// TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load this; the product (softflow_amd/) never does.
//
// Every floating-point expression keeps the reference's operation order so
// that, built with -O2 -ffp-contract=off (no -ffast-math, no -march=native),
// this code reproduces the reference engine's bits at one OpenMP thread.
// Citations are to ramc77/SoftFlow; each function names the file:line it
// restates.  The data layout is this build's own (flat SoA node arrays with
// CSR capsule offsets) rather than the reference's per-capsule vectors.
#include "sfo.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <random>
#include <vector>

namespace {

constexpr double kPI = 3.14159265358979323846;            // core/types.h:19

// D2Q9 tables: src/lbm/lattice.h:9-31
constexpr int    Q = 9;
constexpr double W[Q] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                         1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
constexpr int    CX[Q] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int    CY[Q] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr int    OPP[Q] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
constexpr double CS2 = 1.0 / 3.0;
constexpr double CS4 = CS2 * CS2;

enum Cell : int { FLUID = 0, SOLID = 1, INLET = 2, OUTLET = 3, EMPTY = 4 };

// lattice.h:35-39
inline double feq(int q, double rho, double ux, double uy) {
    double eu = static_cast<double>(CX[q]) * ux + static_cast<double>(CY[q]) * uy;
    double uu = ux * ux + uy * uy;
    return W[q] * rho * (1.0 + eu / CS2 + (eu * eu) / (2.0 * CS4) - uu / (2.0 * CS2));
}

// ---------------------------------------------------------------------------
// Peskin 4-point kernel as a 1025-entry table: coupling/delta_function.h:15-59
double phi_exact(double r) {
    double ar = std::abs(r);
    if (ar <= 1.0) return 0.125 * (3.0 - 2.0 * ar + std::sqrt(1.0 + 4.0 * ar - 4.0 * ar * ar));
    if (ar <= 2.0) return 0.125 * (5.0 - 2.0 * ar - std::sqrt(-7.0 + 12.0 * ar - 4.0 * ar * ar));
    return 0.0;
}
struct PhiLUT {
    double v[1025];
    PhiLUT() {
        for (int i = 0; i <= 1024; ++i) v[i] = phi_exact(2.0 * static_cast<double>(i) / 1024);
    }
};
const PhiLUT& lut() { static const PhiLUT t; return t; }
inline double phi(double r) {
    double ar = std::abs(r);
    if (ar >= 2.0) return 0.0;
    double idx = ar * (1024 * 0.5);
    int i = static_cast<int>(idx);
    if (i >= 1024) return 0.0;
    double frac = idx - static_cast<double>(i);
    const double* t = lut().v;
    return t[i] * (1.0 - frac) + t[i + 1] * frac;
}
inline double delta2(double rx, double ry) { return phi(rx) * phi(ry); }

// ---------------------------------------------------------------------------
// Collision kernels on one cell.  f[q] in/out.
// BGK + Guo: lbm/bgk_collision.cpp:15-86
inline void collide_bgk(double* f, double om, double r, double vx, double vy,
                        double fx, double fy) {
    const double om1 = 1.0 - om;
    const double guo = 1.0 - 0.5 * om;
    const double uu = vx * vx + vy * vy;
    const double uu_term = -0.5 * 3.0 * uu;
    for (int q = 0; q < Q; ++q) {
        const double ex = static_cast<double>(CX[q]);
        const double ey = static_cast<double>(CY[q]);
        const double eu = ex * vx + ey * vy;
        double fe = W[q] * r * (1.0 + eu * 3.0 + 0.5 * eu * eu * 9.0 + uu_term);
        double tx = (ex - vx) * 3.0 + (eu * 9.0) * ex;
        double ty = (ey - vy) * 3.0 + (eu * 9.0) * ey;
        double Fi = guo * W[q] * (tx * fx + ty * fy);
        f[q] = om1 * f[q] + om * fe + Fi;
    }
}

// MRT (Lallemand-Luo moment basis) + Guo in moment space:
// lbm/mrt_collision.cpp:31-140
inline void collide_mrt(double* f, double sn, double s1, double s2, double s4,
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

// Regularized BGK keeping the first-order non-equilibrium moment + Guo:
// lbm/regularized_bgk.cpp:15-93
inline void collide_reg(double* f, double omega, double rho, double ux,
                        double uy, double Fx, double Fy) {
    double fe[Q];
    for (int q = 0; q < Q; ++q) fe[q] = feq(q, rho, ux, uy);
    double jx = 0.0, jy = 0.0, pxx = 0.0, pxy = 0.0, pyy = 0.0;
    for (int q = 0; q < Q; ++q) {
        double fn = f[q] - fe[q];
        jx += fn * CX[q];
        jy += fn * CY[q];
        pxx += fn * CX[q] * CX[q];
        pxy += fn * CX[q] * CY[q];
        pyy += fn * CY[q] * CY[q];
    }
    constexpr double cs2 = 1.0 / 3.0;
    constexpr double cs4_2 = 2.0 * cs2 * cs2;
    for (int q = 0; q < Q; ++q) {
        double Qxx = CX[q] * CX[q] - cs2;
        double Qxy = CX[q] * CY[q];
        double Qyy = CY[q] * CY[q] - cs2;
        double reg = W[q] * ((CX[q] * jx + CY[q] * jy) / cs2 +
                             (Qxx * pxx + 2.0 * Qxy * pxy + Qyy * pyy) / cs4_2);
        double eu = CX[q] * ux + CY[q] * uy;
        double Fi = (1.0 - 0.5 * omega) * W[q] *
                    ((CX[q] - ux + eu * CX[q] / cs2) * Fx / cs2 +
                     (CY[q] - uy + eu * CY[q] / cs2) * Fy / cs2);
        f[q] = fe[q] + (1.0 - omega) * reg + Fi;
    }
}

struct V2 { double x, y; };

// ---------------------------------------------------------------------------
struct Caps {
    // per capsule
    std::vector<int> off{0};         // CSR node offsets
    std::vector<int> type, model, helf, rigid;
    std::vector<double> k_stretch, G_s, C_sk, k_bend, kappa0, k_area, k_per, gvisc, eta;
    std::vector<double> wlc_ratio, wlc_kbt, wlc_kpow;
    std::vector<double> rest_area, rest_per, density;
    std::vector<int> periodic_nx, has_prev;
    // per node
    std::vector<double> x, y, vx, vy, fx, fy;
    std::vector<double> L0, th0, ds0, Lprev;
    int n() const { return static_cast<int>(off.size()) - 1; }

};

// An obstacle: geometry/circle_obstacle.h, rect_obstacle.h,
// polygon_obstacle.h/.cpp.  Each query keeps the reference's operations.
struct OObstacle {
    int kind;                       // 0 circle, 1 rectangle, 2 polygon
    double a[4];                    // circle cx, cy, r; rectangle x0, y0, x1, y1
    std::vector<V2> poly;

    bool contains(double x, double y) const {
        if (kind == 0) {
            double dx = x - a[0], dy = y - a[1];
            return (dx * dx + dy * dy) <= a[2] * a[2];
        }
        if (kind == 1) return x >= a[0] && x <= a[2] && y >= a[1] && y <= a[3];
        bool in = false;                                   // polygon_obstacle.cpp:10-24
        const int n = static_cast<int>(poly.size());
        for (int i = 0, j = n - 1; i < n; j = i++) {
            const V2 &pi = poly[i], &pj = poly[j];
            if (((pi.y > y) != (pj.y > y)) && (x < (pj.x - pi.x) * (y - pi.y) / (pj.y - pi.y) + pi.x))
                in = !in;
        }
        return in;
    }
    V2 nearest(double x, double y) const {
        if (kind == 0) {
            double dx = x - a[0], dy = y - a[1];
            double d = std::sqrt(dx * dx + dy * dy);
            if (d < 1e-15) return V2{a[0] + a[2], a[1]};
            return V2{a[0] + a[2] * dx / d, a[1] + a[2] * dy / d};
        }
        if (kind == 1) {                                   // rect_obstacle.h:27-41
            double cx = std::clamp(x, a[0], a[2]), cy = std::clamp(y, a[1], a[3]);
            if (cx == x && cy == y) {
                double dl = x - a[0], dr = a[2] - x, db = y - a[1], dt = a[3] - y;
                double m = std::min({dl, dr, db, dt});
                if (m == dl) return V2{a[0], y};
                if (m == dr) return V2{a[2], y};
                if (m == db) return V2{x, a[1]};
                return V2{x, a[3]};
            }
            return V2{cx, cy};
        }
        double best = std::numeric_limits<double>::max();  // polygon_obstacle.cpp:26-54
        V2 out{0.0, 0.0};
        const int n = static_cast<int>(poly.size());
        for (int i = 0; i < n; ++i) {
            const V2 v0 = poly[i], v1 = poly[(i + 1) % n];
            V2 e{v1.x - v0.x, v1.y - v0.y};
            double len2 = e.x * e.x + e.y * e.y;
            V2 q = v0;
            if (!(len2 < 1e-30)) {
                double t = ((x - v0.x) * e.x + (y - v0.y) * e.y) / len2;
                t = std::max(0.0, std::min(1.0, t));
                q = V2{v0.x + t * e.x, v0.y + t * e.y};
            }
            double ddx = q.x - x, ddy = q.y - y;
            double d2 = ddx * ddx + ddy * ddy;
            if (d2 < best) { best = d2; out = q; }
        }
        return out;
    }
    double signed_distance(double x, double y) const {
        if (kind == 0) {
            double dx = x - a[0], dy = y - a[1];
            return std::sqrt(dx * dx + dy * dy) - a[2];
        }
        if (kind == 1) {                                   // rect_obstacle.h:18-25
            double dx = std::max({a[0] - x, x - a[2], 0.0});
            double dy = std::max({a[1] - y, y - a[3], 0.0});
            double out = std::sqrt(dx * dx + dy * dy);
            if (out > 1e-15) return out;
            return -std::min({x - a[0], a[2] - x, y - a[1], a[3] - y});
        }
        V2 q = nearest(x, y);
        double d = std::sqrt((q.x - x) * (q.x - x) + (q.y - y) * (q.y - y));
        return contains(x, y) ? -d : d;
    }
    V2 normal(double x, double y) const {
        if (kind == 0) {
            double dx = x - a[0], dy = y - a[1];
            double d = std::sqrt(dx * dx + dy * dy);
            if (d < 1e-15) return V2{1.0, 0.0};
            return V2{dx / d, dy / d};
        }
        V2 q = nearest(x, y);
        double dx = x - q.x, dy = y - q.y;
        double d = std::sqrt(dx * dx + dy * dy);
        if (kind == 1) {                                   // rect_obstacle.h:43-61
            if (d < 1e-15) {
                double dl = std::abs(x - a[0]), dr = std::abs(x - a[2]);
                double db = std::abs(y - a[1]), dt = std::abs(y - a[3]);
                double m = std::min({dl, dr, db, dt});
                if (m == dl) return V2{-1.0, 0.0};
                if (m == dr) return V2{1.0, 0.0};
                if (m == db) return V2{0.0, -1.0};
                return V2{0.0, 1.0};
            }
            return V2{dx / d, dy / d};
        }
        if (d < 1e-15) return V2{0.0, 1.0};                // polygon_obstacle.cpp:62-77
        V2 nn{dx / d, dy / d};
        if (contains(x, y)) { nn.x = -nn.x; nn.y = -nn.y; }
        return nn;
    }
};

struct IbbLink { int x, y, q; double qf; };

}  // namespace

struct sfo_sim {
    sfo_params p;
    std::vector<OObstacle> obs;
    std::vector<IbbLink> ibb;
    int nx, ny, N;
    bool periodic_x, periodic_y;
    std::vector<double> f, ftmp, rho, ux, uy, Fx, Fy;
    std::vector<int> flags;
    std::vector<int> nb;          // pull-source table, q-major
    std::vector<int> solid;
    Caps caps;
    int capped = 0;
    long step_count = 0;
    // scalar transport (node-major g[n*Q+q], advection_diffusion.h:149-161)
    int nsp = 0;
    std::vector<std::vector<double>> g, gtmp, C;
    std::vector<double> tau_s, c_inlet;
    // MDF control of the device engine (no reference counterpart): residual
    // mode stops once max_k |V_k - U(X_k)| <= tol; statistics per step
    double mdf_tol = 0.0;
    int mdf_max = 0, mdf_probe = 0;
    double st_applied = 0.0, st_final = 0.0, st_max = 0.0, st_steps = 0.0;
};

namespace {

// minimum image in x used by the capsule geometry: membrane/capsule.h:140-148
inline V2 mimg(const Caps& c, int cap, double ax, double ay, double bx, double by) {
    V2 d{ax - bx, ay - by};
    if (c.periodic_nx[cap] > 0) {
        double Lx = static_cast<double>(c.periodic_nx[cap]);
        if (d.x > 0.5 * Lx) d.x -= Lx;
        if (d.x < -0.5 * Lx) d.x += Lx;
    }
    return d;
}
inline double norm(V2 v) { return std::sqrt(v.x * v.x + v.y * v.y); }
inline double cross(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }
inline double dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }

// Capsule::area  membrane/capsule.cpp:175-195 (signed sum before abs)
double signed_area_sum(const Caps& c, int cap) {
    int o = c.off[cap], n = c.off[cap + 1] - o;
    double A = 0.0;
    if (c.periodic_nx[cap] > 0 && n > 0) {
        std::vector<V2> u(n);
        u[0] = V2{c.x[o], c.y[o]};
        for (int i = 1; i < n; ++i) {
            V2 d = mimg(c, cap, c.x[o + i], c.y[o + i], c.x[o + i - 1], c.y[o + i - 1]);
            u[i] = V2{u[i - 1].x + d.x, u[i - 1].y + d.y};
        }
        for (int i = 0; i < n; ++i) A += cross(u[i], u[(i + 1) % n]);
        return A;
    }
    for (int i = 0; i < n; ++i) {
        int j = (i + 1) % n;
        A += cross(V2{c.x[o + i], c.y[o + i]}, V2{c.x[o + j], c.y[o + j]});
    }
    return A;
}
double cap_area(const Caps& c, int cap) { return 0.5 * std::abs(signed_area_sum(c, cap)); }

// Capsule::perimeter  membrane/capsule.cpp:197-204
double cap_perimeter(const Caps& c, int cap) {
    int o = c.off[cap], n = c.off[cap + 1] - o;
    double P = 0.0;
    for (int i = 0; i < n; ++i) {
        int j = (i + 1) % n;
        P += norm(mimg(c, cap, c.x[o + j], c.y[o + j], c.x[o + i], c.y[o + i]));
    }
    return P;
}

// Capsule::centroid  membrane/capsule.cpp:159-173
V2 cap_centroid(const Caps& c, int cap) {
    int o = c.off[cap], n = c.off[cap + 1] - o;
    if (c.periodic_nx[cap] > 0 && n > 0) {
        V2 ref{c.x[o], c.y[o]};
        V2 s = ref;
        for (int i = 1; i < n; ++i) {
            V2 d = mimg(c, cap, c.x[o + i], c.y[o + i], ref.x, ref.y);
            s.x += ref.x + d.x;
            s.y += ref.y + d.y;
        }
        return V2{s.x / static_cast<double>(n), s.y / static_cast<double>(n)};
    }
    V2 s{0.0, 0.0};
    for (int i = 0; i < n; ++i) { s.x += c.x[o + i]; s.y += c.y[o + i]; }
    return V2{s.x / static_cast<double>(n), s.y / static_cast<double>(n)};
}

// Capsule::secondMoments / deformationIndex / inclinationAngle
// membrane/capsule.cpp:210-265
void cap_moments(const Caps& c, int cap, double& mxx, double& myy, double& mxy) {
    int o = c.off[cap], n = c.off[cap + 1] - o;
    std::vector<V2> p(n);
    p[0] = V2{c.x[o], c.y[o]};
    for (int i = 1; i < n; ++i) {
        V2 d = mimg(c, cap, c.x[o + i], c.y[o + i], c.x[o + i - 1], c.y[o + i - 1]);
        p[i] = V2{p[i - 1].x + d.x, p[i - 1].y + d.y};
    }
    double A = 0.0, cx = 0.0, cy = 0.0;
    for (int i = 0; i < n; ++i) {
        V2 a = p[i], b = p[(i + 1) % n];
        double cr = a.x * b.y - b.x * a.y;
        A += cr; cx += (a.x + b.x) * cr; cy += (a.y + b.y) * cr;
    }
    A *= 0.5;
    mxx = myy = mxy = 0.0;
    if (std::abs(A) < 1e-15) return;
    cx /= (6.0 * A);
    cy /= (6.0 * A);
    for (int i = 0; i < n; ++i) {
        double x0 = p[i].x - cx, y0 = p[i].y - cy;
        V2 q = p[(i + 1) % n];
        double x1 = q.x - cx, y1 = q.y - cy;
        double cr = x0 * y1 - x1 * y0;
        mxx += (x0 * x0 + x0 * x1 + x1 * x1) * cr;
        myy += (y0 * y0 + y0 * y1 + y1 * y1) * cr;
        mxy += (x0 * y1 + 2.0 * x0 * y0 + 2.0 * x1 * y1 + x1 * y0) * cr;
    }
    double sg = (A > 0.0) ? 1.0 : -1.0;
    mxx *= sg / 12.0; myy *= sg / 12.0; mxy *= sg / 24.0;
}
double cap_taylor_d(const Caps& c, int cap) {
    int n = c.off[cap + 1] - c.off[cap];
    if (n < 3) return 0.0;
    double mxx, myy, mxy;
    cap_moments(c, cap, mxx, myy, mxy);
    double tr = 0.5 * (mxx + myy);
    double dd = std::sqrt(0.25 * (mxx - myy) * (mxx - myy) + mxy * mxy);
    double J1 = tr + dd, J2 = std::max(tr - dd, 0.0);
    double s1 = std::sqrt(J1), s2 = std::sqrt(J2);
    return (s1 + s2 > 1e-15) ? (s1 - s2) / (s1 + s2) : 0.0;
}

// Capsule::Capsule (rest geometry): membrane/capsule.cpp:7-141
void add_capsule(Caps& c, double cx, double cy, double radius, int num_nodes,
                 const sfo_membrane& m, int type) {
    int n = num_nodes > 0 ? num_nodes
                          : std::max(static_cast<int>(std::ceil(2.0 * kPI * radius / 0.75)), 12);
    int cap = c.n();
    int o = c.off.back();
    c.off.push_back(o + n);
    c.type.push_back(type); c.model.push_back(m.model); c.helf.push_back(m.use_helfrich);
    c.rigid.push_back(m.is_rigid);
    c.k_stretch.push_back(m.k_stretch); c.G_s.push_back(m.G_s); c.C_sk.push_back(m.C_skalak);
    c.k_bend.push_back(m.k_bend); c.kappa0.push_back(m.kappa_0); c.k_area.push_back(m.k_area);
    c.k_per.push_back(m.k_perimeter); c.gvisc.push_back(m.gamma_visc); c.eta.push_back(m.eta_membrane);
    c.density.push_back(m.density);
    c.periodic_nx.push_back(0); c.has_prev.push_back(0);
    c.x.resize(o + n); c.y.resize(o + n);
    c.vx.resize(o + n, 0.0); c.vy.resize(o + n, 0.0);
    c.fx.resize(o + n, 0.0); c.fy.resize(o + n, 0.0);
    c.L0.resize(o + n); c.th0.resize(o + n, 0.0); c.ds0.resize(o + n, 0.0); c.Lprev.resize(o + n);

    auto shape_pt = [&](double th) -> V2 {
        if (m.shape == 1 || m.shape == 3)
            return V2{radius * std::cos(th), radius * m.aspect_ratio * std::sin(th)};
        if (m.shape == 2) {
            double r = radius * (1.0 - m.indent_depth * std::sin(th) * std::sin(th));
            return V2{r * std::cos(th), r * std::sin(th)};
        }
        return V2{radius * std::cos(th), radius * std::sin(th)};
    };
    if (m.shape == 0) {
        for (int i = 0; i < n; ++i) {
            V2 r = shape_pt(2.0 * kPI * i / n);
            c.x[o + i] = cx + r.x; c.y[o + i] = cy + r.y;
        }
    } else {
        const int M = 256 * n;
        std::vector<V2> pts(M + 1);
        std::vector<double> cum(M + 1, 0.0);
        for (int k = 0; k <= M; ++k) {
            pts[k] = shape_pt(2.0 * kPI * k / M);
            if (k > 0) cum[k] = cum[k - 1] + norm(V2{pts[k].x - pts[k - 1].x, pts[k].y - pts[k - 1].y});
        }
        const double total = cum[M];
        int k = 0;
        for (int i = 0; i < n; ++i) {
            const double target = total * i / n;
            while (k < M - 1 && cum[k + 1] < target) ++k;
            const double seg = cum[k + 1] - cum[k];
            const double w = (seg > 1e-15) ? (target - cum[k]) / seg : 0.0;
            V2 d{pts[k + 1].x - pts[k].x, pts[k + 1].y - pts[k].y};
            V2 r{pts[k].x + d.x * w, pts[k].y + d.y * w};
            c.x[o + i] = cx + r.x; c.y[o + i] = cy + r.y;
        }
    }
    for (int i = 0; i < n; ++i) {
        int j = (i + 1) % n;
        c.L0[o + i] = norm(V2{c.x[o + j] - c.x[o + i], c.y[o + j] - c.y[o + i]});
        c.Lprev[o + i] = c.L0[o + i];
    }
    c.rest_area.push_back(cap_area(c, cap));
    c.rest_per.push_back(cap_perimeter(c, cap));
    double kbt = m.wlc_kBT_p, kpow = m.wlc_k_pow;
    if (m.model == 3 && n > 0) {
        double L0 = c.L0[o];
        double Lmax = m.wlc_L_max_ratio * L0;
        double x0 = L0 / Lmax;
        if (kbt <= 0.0) {
            double denom = 1.0 / (2.0 * (1.0 - x0) * (1.0 - x0) * (1.0 - x0)) + 1.0;
            kbt = m.G_s * Lmax / (L0 * denom);
        }
        if (kpow <= 0.0) {
            double Fr = kbt * (1.0 / (4.0 * (1.0 - x0) * (1.0 - x0)) - 0.25 + x0);
            kpow = Fr * L0;
        }
    }
    c.wlc_ratio.push_back(m.wlc_L_max_ratio); c.wlc_kbt.push_back(kbt); c.wlc_kpow.push_back(kpow);
    if (m.use_helfrich) {
        for (int i = 0; i < n; ++i) {
            int ip = (i - 1 + n) % n, in = (i + 1) % n;
            V2 ep = mimg(c, cap, c.x[o + i], c.y[o + i], c.x[o + ip], c.y[o + ip]);
            V2 en = mimg(c, cap, c.x[o + in], c.y[o + in], c.x[o + i], c.y[o + i]);
            c.ds0[o + i] = 0.5 * (norm(ep) + norm(en));
            c.th0[o + i] = (m.kappa_0 != 0.0) ? m.kappa_0 * c.ds0[o + i]
                                              : std::atan2(cross(ep, en), dot(ep, en));
        }
    }
}

// Capsule::computeMembraneForces and helpers: membrane/capsule.cpp:149-157, 320-671
void membrane_forces(Caps& c, int cap) {
    if (c.rigid[cap]) return;
    const int o = c.off[cap], n = c.off[cap + 1] - o;
    double* X = c.x.data() + o;
    double* Y = c.y.data() + o;
    double* FX = c.fx.data() + o;
    double* FY = c.fy.data() + o;
    auto nxt = [n](int i) { return (i + 1) % n; };
    auto prv = [n](int i) { return (i - 1 + n) % n; };
    // stretching  (:320-470)
    for (int i = 0; i < n; ++i) {
        int j = nxt(i);
        V2 d = mimg(c, cap, X[j], Y[j], X[i], Y[i]);
        double L = norm(d);
        if (L < 1e-15) continue;
        double L0 = c.L0[o + i];
        double Fm;
        switch (c.model[cap]) {
        case 1: {
            double lam = L / L0;
            double li3 = 1.0 / (lam * lam * lam);
            Fm = c.G_s[cap] * (lam - li3);
            break; }
        case 2: {
            double lam = L / L0;
            double lam2 = lam * lam;
            Fm = c.G_s[cap] * lam * (lam2 - 1.0) * (1.0 + c.C_sk[cap]);
            break; }
        case 3: {
            double Lmax = c.wlc_ratio[cap] * L0;
            double xx = L / Lmax;
            if (xx > 0.99) xx = 0.99;
            double Fw = c.wlc_kbt[cap] * (1.0 / (4.0 * (1.0 - xx) * (1.0 - xx)) - 0.25 + xx);
            double Fp = -c.wlc_kpow[cap] / L;
            Fm = Fw + Fp;
            break; }
        default:
            Fm = c.k_stretch[cap] * (L / L0 - 1.0);
        }
        V2 dir{d.x / L, d.y / L};
        V2 F{dir.x * Fm, dir.y * Fm};
        FX[i] += F.x; FY[i] += F.y;
        FX[j] -= F.x; FY[j] -= F.y;
    }
    // bending (:476-541)
    if (c.helf[cap]) {
        for (int i = 0; i < n; ++i) {
            int ip = prv(i), in = nxt(i);
            V2 ep = mimg(c, cap, X[i], Y[i], X[ip], Y[ip]);
            V2 en = mimg(c, cap, X[in], Y[in], X[i], Y[i]);
            double lp2 = dot(ep, ep), ln2 = dot(en, en);
            if (lp2 < 1e-30 || ln2 < 1e-30) continue;
            double dth = std::atan2(cross(ep, en), dot(ep, en)) - c.th0[o + i];
            double cc = -c.k_bend[cap] * dth / c.ds0[o + i];
            V2 gn{-en.y / ln2, en.x / ln2};
            V2 gp{-ep.y / lp2, ep.x / lp2};
            FX[in] += gn.x * cc; FY[in] += gn.y * cc;
            FX[ip] += gp.x * cc; FY[ip] += gp.y * cc;
            V2 s{gn.x + gp.x, gn.y + gp.y};
            FX[i] -= s.x * cc; FY[i] -= s.y * cc;
        }
    } else {
        for (int i = 0; i < n; ++i) {
            V2 dp = mimg(c, cap, X[prv(i)], Y[prv(i)], X[i], Y[i]);
            V2 dn = mimg(c, cap, X[nxt(i)], Y[nxt(i)], X[i], Y[i]);
            V2 lap{dp.x + dn.x, dp.y + dn.y};
            FX[i] += lap.x * c.k_bend[cap]; FY[i] += lap.y * c.k_bend[cap];
        }
    }
    // area penalty (:547-586)
    {
        double A = cap_area(c, cap);
        double A0 = c.rest_area[cap];
        if (!(A0 < 1e-15)) {
            double coeff = -c.k_area[cap] * (A - A0) / A0;
            double sa = signed_area_sum(c, cap);
            double sign = (sa >= 0.0) ? 1.0 : -1.0;
            for (int i = 0; i < n; ++i) {
                V2 ep = mimg(c, cap, X[i], Y[i], X[prv(i)], Y[prv(i)]);
                V2 en = mimg(c, cap, X[nxt(i)], Y[nxt(i)], X[i], Y[i]);
                V2 np{ep.y, -ep.x}, nn{en.y, -en.x};
                V2 out{(np.x + nn.x) * 0.5, (np.y + nn.y) * 0.5};
                double len = norm(out);
                if (len < 1e-15) continue;
                out = V2{out.x / len * sign, out.y / len * sign};
                FX[i] += out.x * coeff; FY[i] += out.y * coeff;
            }
        }
    }
    // perimeter penalty (:588-604)
    {
        double P = cap_perimeter(c, cap);
        double P0 = c.rest_per[cap];
        if (!(P0 < 1e-15)) {
            double coeff = -c.k_per[cap] * (P - P0) / P0;
            for (int i = 0; i < n; ++i) {
                int j = nxt(i);
                V2 d = mimg(c, cap, X[j], Y[j], X[i], Y[i]);
                double L = norm(d);
                if (L < 1e-15) continue;
                V2 t{d.x / L, d.y / L};
                FX[i] -= t.x * coeff; FY[i] -= t.y * coeff;
                FX[j] += t.x * coeff; FY[j] += t.y * coeff;
            }
        }
    }
    // nodal damping (:606-610)
    for (int i = 0; i < n; ++i) {
        FX[i] -= c.vx[o + i] * c.gvisc[cap];
        FY[i] -= c.vy[o + i] * c.gvisc[cap];
    }
    // Kelvin-Voigt (:612-671)
    if (c.eta[cap] > 0.0) {
        if (!c.has_prev[cap]) {
            for (int i = 0; i < n; ++i) {
                int j = nxt(i);
                c.Lprev[o + i] = norm(mimg(c, cap, X[j], Y[j], X[i], Y[i]));
            }
            c.has_prev[cap] = 1;
        } else {
            for (int i = 0; i < n; ++i) {
                int j = nxt(i);
                V2 d = mimg(c, cap, X[j], Y[j], X[i], Y[i]);
                double L = norm(d);
                if (L < 1e-15) continue;
                double sr = (L - c.Lprev[o + i]) / c.L0[o + i];
                double Fv = c.eta[cap] * sr;
                V2 dir{d.x / L, d.y / L};
                V2 F{dir.x * Fv, dir.y * Fv};
                FX[i] += F.x; FY[i] += F.y;
                FX[j] -= F.x; FY[j] -= F.y;
                c.Lprev[o + i] = L;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Repulsion: membrane/repulsion.cpp:6-178 and membrane/cell_list.h:36-146
struct Repulsion {
    double eps, sig, rcut, gn, mu;
    int pw;
    double Lx;
    double mix(double dx) const {
        if (Lx > 0.0) {
            if (dx > 0.5 * Lx) dx -= Lx;
            if (dx < -0.5 * Lx) dx += Lx;
        }
        return dx;
    }
    V2 pair(V2 pa, V2 va, V2 pb, V2 vb) const {
        V2 d{pa.x - pb.x, pa.y - pb.y};
        d.x = mix(d.x);
        double r = norm(d);
        if (r < 1e-15 || r >= rcut) return V2{0.0, 0.0};
        V2 nn{d.x / r, d.y / r};
        double Fr = eps * std::pow(sig / r, pw);
        double Fn = Fr;
        V2 Ft{0.0, 0.0};
        if (gn > 0.0 || mu > 0.0) {
            V2 vr{va.x - vb.x, va.y - vb.y};
            double vn = vr.x * nn.x + vr.y * nn.y;
            Fn = Fr - gn * vn;
            if (Fn < 0.0) Fn = 0.0;
            if (mu > 0.0) {
                V2 vt{vr.x - nn.x * vn, vr.y - nn.y * vn};
                double vtm = norm(vt);
                if (vtm > 1e-12) {
                    double s = -mu * Fn;
                    Ft = V2{vt.x / vtm * s, vt.y / vtm * s};
                }
            }
        }
        return V2{nn.x * Fn + Ft.x, nn.y * Fn + Ft.y};
    }
};

void repulsion_all(Caps& c, const Repulsion& R, double ybot, double ytop) {
    const int ncap = c.n();
    if (ncap == 0) return;
    const int K = c.off[ncap];
    auto wall = [&](int cap) {
        for (int k = c.off[cap]; k < c.off[cap + 1]; ++k) {
            V2 p{c.x[k], c.y[k]}, v{c.vx[k], c.vy[k]};
            V2 a = R.pair(p, v, V2{p.x, ybot}, V2{0.0, 0.0});
            c.fx[k] += a.x; c.fy[k] += a.y;
            V2 b = R.pair(p, v, V2{p.x, ytop}, V2{0.0, 0.0});
            c.fx[k] += b.x; c.fy[k] += b.y;
        }
    };
    if (K <= 100) {
        // one-sided all-pairs (the OpenMP build's path): repulsion.cpp:103-112
        for (int i = 0; i < ncap; ++i) {
            for (int j = 0; j < ncap; ++j) {
                if (i == j) continue;
                for (int a = c.off[i]; a < c.off[i + 1]; ++a) {
                    for (int b = c.off[j]; b < c.off[j + 1]; ++b) {
                        V2 F = R.pair(V2{c.x[a], c.y[a]}, V2{c.vx[a], c.vy[a]},
                                      V2{c.x[b], c.y[b]}, V2{c.vx[b], c.vy[b]});
                        c.fx[a] += F.x; c.fy[a] += F.y;
                    }
                }
            }
            wall(i);
        }
        return;
    }
    // cell list
    double xext, xorg;
    if (R.Lx > 0.0) { xext = R.Lx; xorg = 0.0; }
    else {
        double xmin = 1e30, xmax = -1e30;
        for (int k = 0; k < K; ++k) { if (c.x[k] < xmin) xmin = c.x[k]; if (c.x[k] > xmax) xmax = c.x[k]; }
        xorg = xmin - R.rcut;
        xext = (xmax - xmin) + 2.0 * R.rcut;
    }
    int ncx = std::max(1, static_cast<int>(std::floor(xext / R.rcut)));
    int ncy = std::max(1, static_cast<int>(std::floor((ytop - ybot) / R.rcut)));
    double dx = xext / static_cast<double>(ncx);
    double dy = (ytop - ybot) / static_cast<double>(ncy);
    auto cell_of = [&](double px0, double py) {
        if (std::isnan(px0) || std::isnan(py) || std::isinf(px0) || std::isinf(py)) return 0;
        double px = px0 - xorg;
        if (R.Lx > 0.0) { px = std::fmod(px, R.Lx); if (px < 0.0) px += R.Lx; }
        int ix = static_cast<int>(std::floor(px / dx));
        int iy = static_cast<int>(std::floor((py - ybot) / dy));
        if (ix < 0) ix = 0;
        if (ix >= ncx) ix = ncx - 1;
        if (iy < 0) iy = 0;
        if (iy >= ncy) iy = ncy - 1;
        return iy * ncx + ix;
    };
    std::vector<std::vector<int>> cells(static_cast<size_t>(ncx) * ncy);
    std::vector<int> owner(K);
    for (int cap = 0; cap < ncap; ++cap)
        for (int k = c.off[cap]; k < c.off[cap + 1]; ++k) {
            owner[k] = cap;
            int ci = cell_of(c.x[k], c.y[k]);
            if (ci >= 0 && ci < ncx * ncy) cells[ci].push_back(k);
        }
    for (int cap = 0; cap < ncap; ++cap) {
        for (int a = c.off[cap]; a < c.off[cap + 1]; ++a) {
            int ci = cell_of(c.x[a], c.y[a]);
            int cx0 = ci % ncx, cy0 = ci / ncx;
            for (int ddy = -1; ddy <= 1; ++ddy) {
                int yy = cy0 + ddy;
                if (yy < 0 || yy >= ncy) continue;
                for (int ddx = -1; ddx <= 1; ++ddx) {
                    int xx = cx0 + ddx;
                    if (R.Lx > 0.0) { if (xx < 0) xx += ncx; if (xx >= ncx) xx -= ncx; }
                    else if (xx < 0 || xx >= ncx) continue;
                    for (int b : cells[yy * ncx + xx]) {
                        if (owner[b] == cap) continue;
                        V2 F = R.pair(V2{c.x[a], c.y[a]}, V2{c.vx[a], c.vy[a]},
                                      V2{c.x[b], c.y[b]}, V2{c.vx[b], c.vy[b]});
                        c.fx[a] += F.x; c.fy[a] += F.y;
                    }
                }
            }
        }
        wall(cap);
    }
}

// ---------------------------------------------------------------------------
// IBM: coupling/immersed_boundary.cpp
double node_ds(const sfo_sim& s, int cap, int k) {
    const Caps& c = s.caps;
    int o = c.off[cap], n = c.off[cap + 1] - o;
    int kp = (k - 1 + n) % n, kn = (k + 1) % n;
    double Lx = static_cast<double>(s.nx);
    auto mds = [Lx](double ax, double ay, double bx, double by) {   // :16-23
        V2 d{ax - bx, ay - by};
        if (Lx > 0.0) {
            if (d.x > 0.5 * Lx) d.x -= Lx;
            if (d.x < -0.5 * Lx) d.x += Lx;
        }
        return norm(d);
    };
    double dp = mds(c.x[o + k], c.y[o + k], c.x[o + kp], c.y[o + kp]);
    double dn = mds(c.x[o + kn], c.y[o + kn], c.x[o + k], c.y[o + k]);
    return 0.5 * (dp + dn);
}

// spreadNodeForce(Local): :102-172
void spread_node(sfo_sim& s, double px, double py, double fx, double fy, double ds) {
    if (std::isnan(px) || std::isnan(py) || std::isinf(px) || std::isinf(py)) return;
    if (std::isnan(fx) || std::isnan(fy)) return;
    int ix0 = static_cast<int>(std::floor(px));
    int iy0 = static_cast<int>(std::floor(py));
    if (iy0 < -2 || iy0 >= s.ny + 2) return;
    for (int iy = iy0 - 1; iy <= iy0 + 2; ++iy) {
        for (int ix = ix0 - 1; ix <= ix0 + 2; ++ix) {
            int jx = ix;
            if (jx < 0) jx += s.nx;
            if (jx >= s.nx) jx -= s.nx;
            int jy = iy;
            if (jy < 0 || jy >= s.ny) continue;
            double w = delta2(px - static_cast<double>(ix), py - static_cast<double>(iy));
            if (w > 1e-15) {
                int n = jy * s.nx + jx;
                s.Fx[n] += fx * w * ds;
                s.Fy[n] += fy * w * ds;
            }
        }
    }
}

// interpolateAtPoint: :191-229
V2 interp_at(const sfo_sim& s, double px, double py) {
    if (std::isnan(px) || std::isnan(py) || std::isinf(px) || std::isinf(py)) return V2{0.0, 0.0};
    int ix0 = static_cast<int>(std::floor(px));
    int iy0 = static_cast<int>(std::floor(py));
    if (iy0 < -2 || iy0 >= s.ny + 2) return V2{0.0, 0.0};
    V2 u{0.0, 0.0};
    for (int iy = iy0 - 1; iy <= iy0 + 2; ++iy) {
        for (int ix = ix0 - 1; ix <= ix0 + 2; ++ix) {
            int jx = ix;
            if (jx < 0) jx += s.nx;
            if (jx >= s.nx) jx -= s.nx;
            int jy = iy;
            if (jy < 0 || jy >= s.ny) continue;
            double w = delta2(px - static_cast<double>(ix), py - static_cast<double>(iy));
            if (w > 1e-15) {
                int n = jy * s.nx + jx;
                u.x += s.ux[n] * w;
                u.y += s.uy[n] * w;
            }
        }
    }
    return u;
}

void spread_all(sfo_sim& s) {                                   // :25-100
    const Caps& c = s.caps;
    for (int cap = 0; cap < c.n(); ++cap) {
        int o = c.off[cap], n = c.off[cap + 1] - o;
        for (int k = 0; k < n; ++k)
            spread_node(s, c.x[o + k], c.y[o + k], c.fx[o + k], c.fy[o + k], node_ds(s, cap, k));
    }
}

// LatticeField::computeMacroscopic: lbm/lattice_field.cpp:28-73
void macroscopic(sfo_sim& s) {
    const int N = s.N;
    for (int n = 0; n < N; ++n) {
        if (s.flags[n] == SOLID || s.flags[n] == EMPTY) {
            s.rho[n] = 0.0; s.ux[n] = 0.0; s.uy[n] = 0.0;
            continue;
        }
        double r = 0.0, vx = 0.0, vy = 0.0;
        for (int q = 0; q < Q; ++q) {
            double fi = s.f[q * N + n];
            r += fi;
            vx += fi * static_cast<double>(CX[q]);
            vy += fi * static_cast<double>(CY[q]);
        }
        s.rho[n] = r;
        if (r > 1e-15) {
            s.ux[n] = (vx + 0.5 * s.Fx[n]) / r;
            s.uy[n] = (vy + 0.5 * s.Fy[n]) / r;
        } else {
            s.ux[n] = 0.0; s.uy[n] = 0.0;
        }
    }
}

// multiDirectForcing: :231-286, with ncorr = iterations - 1 corrections.
// Residual mode (tol > 0, no reference counterpart): before each correction
// the no-slip residual r = max_k |V_k - U(X_k)| is taken from the same
// interpolation, and the loop stops once r <= tol.
void mdf(sfo_sim& s, int ncorr) {
    spread_all(s);
    if (ncorr <= 0) return;
    Caps& c = s.caps;
    const double tol = s.mdf_tol;
    const int K = static_cast<int>(c.x.size());
    std::vector<double> dFx(K), dFy(K);
    bool conv = false;
    double r_final = 0.0;
    int applied = 0;
    auto residual_pass = [&](bool keep) {
        macroscopic(s);
        double r = 0.0;
        for (int cap = 0; cap < c.n(); ++cap) {
            int o = c.off[cap], n = c.off[cap + 1] - o;
            for (int k = 0; k < n; ++k) {
                double px = c.x[o + k], py = c.y[o + k];
                V2 ui = interp_at(s, px, py);
                V2 du{c.vx[o + k] - ui.x, c.vy[o + k] - ui.y};
                double rl = 1.0;
                int ix = static_cast<int>(px + 0.5);
                int iy = static_cast<int>(py + 0.5);
                if (ix < 0) ix += s.nx;
                if (ix >= s.nx) ix -= s.nx;
                if (iy >= 0 && iy < s.ny) rl = s.rho[iy * s.nx + ix];
                double two_rho = 2.0 * rl;
                if (keep) { dFx[o + k] = du.x * two_rho; dFy[o + k] = du.y * two_rho; }
                double rk = std::sqrt(du.x * du.x + du.y * du.y);
                r = (rk > r || rk != rk) ? rk : r;
            }
        }
        return r;
    };
    for (int it = 1; it <= ncorr; ++it) {
        double r = residual_pass(true);
        r_final = r;
        if (tol > 0.0 && r <= tol) { conv = true; break; }
        ++applied;
        for (int cap = 0; cap < c.n(); ++cap) {
            int o = c.off[cap], n = c.off[cap + 1] - o;
            for (int k = 0; k < n; ++k)
                spread_node(s, c.x[o + k], c.y[o + k], dFx[o + k], dFy[o + k], node_ds(s, cap, k));
        }
    }
    if (!conv && s.mdf_probe) {
        // diagnostic only: the fields the collision reads stay as they were
        std::vector<double> r0 = s.rho, u0 = s.ux, v0 = s.uy;
        r_final = residual_pass(false);
        s.rho.swap(r0); s.ux.swap(u0); s.uy.swap(v0);
    }
    s.st_applied += applied;
    s.st_final += r_final;
    s.st_max = (r_final > s.st_max || r_final != r_final) ? r_final : s.st_max;
    s.st_steps += 1.0;
}

// ---------------------------------------------------------------------------
// LBM step: lbm/lbm_solver.cpp:172-235 and lbm/bounce_back.cpp:12-40
void build_neighbors(sfo_sim& s) {                             // :72-104
    const int nx = s.nx, ny = s.ny, N = s.N;
    s.nb.assign(static_cast<size_t>(Q) * N, 0);
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            int n = y * nx + x;
            for (int q = 0; q < Q; ++q) {
                int xs = x - CX[q], ys = y - CY[q];
                if (s.periodic_x) xs = (xs + nx) % nx;
                if (s.periodic_y) ys = (ys + ny) % ny;
                bool oobx = !s.periodic_x && (xs < 0 || xs >= nx);
                bool ooby = !s.periodic_y && (ys < 0 || ys >= ny);
                s.nb[q * N + n] = (oobx || ooby) ? n : ys * nx + xs;
            }
        }
}

// MovingWall::apply (lbm/moving_wall.cpp:14-63), after the bounce-back: for a
// wall row moving with velocity U along x, each link from a non-solid cell
// into a solid cell of that row returns f_q - 2 w_q rho (c_q . U) / cs^2, with
// rho the fluid cell's density from the last macroscopic pass.
static void moving_wall_row(sfo_sim& s, int yw, double U) {
    const int nx = s.nx, ny = s.ny, N = s.N;
    const double cs2 = 1.0 / 3.0;
    for (int x = 0; x < nx; ++x) {
        if (s.flags[yw * nx + x] != SOLID) continue;
        for (int q = 1; q < Q; ++q) {
            int xn = x - CX[q];
            const int yn = yw - CY[q];
            if (s.periodic_x) xn = (xn + nx) % nx;
            if (xn < 0 || xn >= nx || yn < 0 || yn >= ny) continue;
            const int m = yn * nx + xn;
            if (s.flags[m] == SOLID) continue;
            const double eu = CX[q] * U;
            s.f[OPP[q] * N + m] = s.f[q * N + yw * nx + x] - 2.0 * W[q] * s.rho[m] * eu / cs2;
        }
    }
}

static void moving_walls(sfo_sim& s) {
    const double top = s.p.top_wall_velocity, bot = s.p.bottom_wall_velocity;
    if (!(std::abs(top) > 1e-15 || std::abs(bot) > 1e-15)) return;
    if (std::abs(bot) > 1e-15) moving_wall_row(s, 0, bot);
    if (std::abs(top) > 1e-15) moving_wall_row(s, s.ny - 1, top);
}

// ZouHeBoundary::apply over the inlet and outlet node lists
// (lbm/zou_he_boundary.cpp:18-75): a parabolic velocity inlet on the INLET
// cells of column 0 and a fixed-density outlet on the OUTLET cells of column
// nx-1, each rebuilding the three populations that stream in from outside.
static void zou_he(sfo_sim& s) {
    const int nx = s.nx, ny = s.ny, N = s.N;
    auto F = [&](int x, int y, int q) -> double& { return s.f[static_cast<size_t>(q) * N + y * nx + x]; };
    for (int y = 0; y < ny; ++y) {
        if (s.flags[y * nx] != INLET) continue;
        const int x = 0;
        const double y_center = static_cast<double>(ny) * 0.5;
        const double H = static_cast<double>(ny - 2);
        const double y_norm = (static_cast<double>(y) - y_center) / (H * 0.5);
        const double ux = s.p.inlet_velocity * (1.0 - y_norm * y_norm);
        const double uy = 0.0;
        const double f0 = F(x, y, 0), f2 = F(x, y, 2), f3 = F(x, y, 3), f4 = F(x, y, 4);
        const double f6 = F(x, y, 6), f7 = F(x, y, 7);
        const double rho = (f0 + f2 + f4 + 2.0 * (f3 + f6 + f7)) / (1.0 - ux);
        F(x, y, 1) = f3 + (2.0 / 3.0) * rho * ux;
        F(x, y, 5) = f7 + 0.5 * (f4 - f2) + (1.0 / 6.0) * rho * ux + 0.5 * rho * uy;
        F(x, y, 8) = f6 - 0.5 * (f4 - f2) + (1.0 / 6.0) * rho * ux - 0.5 * rho * uy;
    }
    for (int y = 0; y < ny; ++y) {
        const int x = nx - 1;
        if (s.flags[y * nx + x] != OUTLET) continue;
        const double rho = s.p.outlet_density;
        const double f0 = F(x, y, 0), f1 = F(x, y, 1), f2 = F(x, y, 2), f4 = F(x, y, 4);
        const double f5 = F(x, y, 5), f8 = F(x, y, 8);
        const double ux = -1.0 + (f0 + f2 + f4 + 2.0 * (f1 + f5 + f8)) / rho;
        const double uy = 0.0;
        F(x, y, 3) = f1 - (2.0 / 3.0) * rho * ux;
        F(x, y, 7) = f5 - 0.5 * (f4 - f2) - (1.0 / 6.0) * rho * ux - 0.5 * rho * uy;
        F(x, y, 6) = f8 + 0.5 * (f4 - f2) - (1.0 / 6.0) * rho * ux + 0.5 * rho * uy;
    }
}

// InterpolatedBounceBack::initialize (lbm/interpolated_bounce_back.cpp:7-53)
void ibb_build(sfo_sim& s) {
    s.ibb.clear();
    for (int y = 0; y < s.ny; ++y)
        for (int x = 0; x < s.nx; ++x) {
            if (s.flags[y * s.nx + x] == SOLID) continue;
            for (int q = 1; q < Q; ++q) {
                int xn = x + CX[q], yn = y + CY[q];
                if (xn < 0 || xn >= s.nx || yn < 0 || yn >= s.ny) continue;
                if (s.flags[yn * s.nx + xn] != SOLID) continue;
                double best = 1.0;
                for (const OObstacle& o : s.obs) {
                    double lo = 0.0, hi = 1.0;
                    for (int it = 0; it < 20; ++it) {
                        double mid = 0.5 * (lo + hi);
                        if (o.contains(static_cast<double>(x) + mid * static_cast<double>(CX[q]),
                                       static_cast<double>(y) + mid * static_cast<double>(CY[q])))
                            hi = mid;
                        else
                            lo = mid;
                    }
                    if (hi < best) best = hi;
                }
                s.ibb.push_back(IbbLink{x, y, q, best});
            }
        }
}

// InterpolatedBounceBack::apply (lbm/interpolated_bounce_back.cpp:55-96), links
// in list order
void ibb_apply(sfo_sim& s) {
    const int N = s.N, nx = s.nx, ny = s.ny;
    auto F = [&](int x, int y, int q) -> double& { return s.f[static_cast<size_t>(q) * N + y * nx + x]; };
    for (const IbbLink& L : s.ibb) {
        const int o = OPP[L.q];
        if (L.qf < 0.5) {
            int xu = L.x - CX[L.q], yu = L.y - CY[L.q];
            if (s.periodic_x) xu = ((xu % nx) + nx) % nx;
            if (xu >= 0 && xu < nx && yu >= 0 && yu < ny)
                F(L.x, L.y, o) = 2.0 * L.qf * F(L.x, L.y, L.q) + (1.0 - 2.0 * L.qf) * F(xu, yu, L.q);
            else
                F(L.x, L.y, o) = F(L.x, L.y, L.q);
        } else {
            double inv2q = 1.0 / (2.0 * L.qf);
            F(L.x, L.y, o) = inv2q * F(L.x, L.y, L.q) + (1.0 - inv2q) * F(L.x, L.y, o);
        }
    }
}

// RepulsionForce::computeObstacleRepulsion (membrane/repulsion.cpp:180-236)
void obstacle_repulsion(Caps& c, const sfo_params& p, const OObstacle& ob) {
    const int K = c.off[c.n()];
    for (int k = 0; k < K; ++k) {
        double sd = ob.signed_distance(c.x[k], c.y[k]);
        if (sd >= p.rep_r_cut) continue;
        double r = std::abs(sd);
        if (r < 1e-15) r = 1e-15;
        V2 nn = ob.normal(c.x[k], c.y[k]);
        double Frep = p.rep_epsilon * std::pow(p.rep_sigma / r, p.rep_power);
        if (sd < 0.0) {
            c.fx[k] += nn.x * (Frep * 10.0);
            c.fy[k] += nn.y * (Frep * 10.0);
            continue;
        }
        double Fn = Frep;
        V2 Ft{0.0, 0.0};
        if (p.rep_damping_normal > 0.0 || p.rep_friction_coeff > 0.0) {
            double vn = c.vx[k] * nn.x + c.vy[k] * nn.y;
            Fn = Frep - p.rep_damping_normal * vn;
            if (Fn < 0.0) Fn = 0.0;
            if (p.rep_friction_coeff > 0.0) {
                V2 vt{c.vx[k] - nn.x * vn, c.vy[k] - nn.y * vn};
                double m = norm(vt);
                if (m > 1e-12) Ft = V2{vt.x / m * (-p.rep_friction_coeff * Fn), vt.y / m * (-p.rep_friction_coeff * Fn)};
            }
        }
        c.fx[k] += nn.x * Fn + Ft.x;
        c.fy[k] += nn.y * Fn + Ft.y;
    }
}

void lbm_step(sfo_sim& s) {
    const int N = s.N;
    const sfo_params& p = s.p;
    const bool reg = p.collision_model == 2;
    const bool mrt = !reg && (p.use_mrt || p.collision_model == 1);
    const double om = 1.0 / p.tau;
    double fl[Q];
    for (int n = 0; n < N; ++n) {
        if (s.flags[n] != FLUID) continue;
        for (int q = 0; q < Q; ++q) fl[q] = s.f[q * N + n];
        if (reg) collide_reg(fl, om, s.rho[n], s.ux[n], s.uy[n], s.Fx[n], s.Fy[n]);
        else if (mrt) collide_mrt(fl, om, p.mrt_se, p.mrt_s_eps, p.mrt_sq, s.rho[n], s.ux[n],
                                  s.uy[n], s.Fx[n], s.Fy[n]);
        else collide_bgk(fl, om, s.rho[n], s.ux[n], s.uy[n], s.Fx[n], s.Fy[n]);
        for (int q = 0; q < Q; ++q) s.f[q * N + n] = fl[q];
    }
    for (int q = 0; q < Q; ++q)
        for (int n = 0; n < N; ++n) s.ftmp[q * N + n] = s.f[q * N + s.nb[q * N + n]];
    s.f.swap(s.ftmp);
    for (int sn : s.solid) {
        int x = sn % s.nx, y = sn / s.nx;
        for (int q = 1; q < Q; ++q) {
            int xn = x - CX[q], yn = y - CY[q];
            if (s.periodic_x) xn = (xn + s.nx) % s.nx;
            if (s.periodic_y) yn = (yn + s.ny) % s.ny;
            if (xn < 0 || xn >= s.nx || yn < 0 || yn >= s.ny) continue;
            int ct = s.flags[yn * s.nx + xn];
            if (ct == SOLID || ct == EMPTY) continue;
            s.f[OPP[q] * N + yn * s.nx + xn] = s.f[q * N + sn];
        }
    }
    moving_walls(s);
    if (p.use_interpolated_bb) ibb_apply(s);                       // lbm_solver.cpp:193-196
    if (p.boundary_type == 0) zou_he(s);
    macroscopic(s);
}

// ---------------------------------------------------------------------------
// ADR: lbm/advection_diffusion.cpp:84-223
void adr_step(sfo_sim& s) {
    const int nx = s.nx, ny = s.ny, N = s.N;
    for (int sp = 0; sp < s.nsp; ++sp) {
        const double oms = 1.0 / s.tau_s[sp];
        const double Cin = s.c_inlet[sp];
        const double* src = s.g[sp].data();
        double* dst = s.gtmp[sp].data();
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                int n = y * nx + x, qb = n * Q;
                int ct = s.flags[n];
                if (ct == SOLID) { for (int q = 0; q < Q; ++q) dst[qb + q] = src[qb + q]; continue; }
                const double Cc = (ct == INLET) ? Cin : s.C[sp][n];
                const double u = s.ux[n], v = s.uy[n];
                for (int q = 0; q < Q; ++q) {
                    const double eu = static_cast<double>(CX[q]) * u + static_cast<double>(CY[q]) * v;
                    const double ge = W[q] * Cc * (1.0 + eu * 3.0);
                    dst[qb + q] = src[qb + q] - oms * (src[qb + q] - ge);
                }
            }
        const double* t = s.gtmp[sp].data();
        double* g = s.g[sp].data();
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                int n = y * nx + x, qb = n * Q;
                if (s.flags[n] == SOLID) continue;
                for (int q = 0; q < Q; ++q) {
                    int xs = x - CX[q], ys = y - CY[q];
                    xs = ((xs % nx) + nx) % nx;
                    if (s.p.scalar_periodic_y) ys = ((ys % ny) + ny) % ny;
                    if (ys < 0 || ys >= ny) { g[qb + q] = t[qb + OPP[q]]; continue; }
                    int nu = ys * nx + xs;
                    g[qb + q] = (s.flags[nu] == SOLID) ? t[qb + OPP[q]] : t[nu * Q + q];
                }
            }
        for (int n = 0; n < N; ++n) {
            double Cs = 0.0;
            for (int q = 0; q < Q; ++q) Cs += g[n * Q + q];
            s.C[sp][n] = Cs;
        }
    }
}

// Capsule::moveNodes: membrane/capsule.cpp:267-297
void move_capsule(Caps& c, int cap, double dt) {
    int o = c.off[cap], n = c.off[cap + 1] - o;
    if (c.rigid[cap]) {
        V2 ctr = cap_centroid(c, cap);
        V2 vcm{0.0, 0.0};
        for (int i = 0; i < n; ++i) { vcm.x += c.vx[o + i]; vcm.y += c.vy[o + i]; }
        vcm.x *= (1.0 / n); vcm.y *= (1.0 / n);
        double num = 0.0, den = 0.0;
        for (int i = 0; i < n; ++i) {
            V2 r{c.x[o + i] - ctr.x, c.y[o + i] - ctr.y};
            num += r.x * c.vy[o + i] - r.y * c.vx[o + i];
            den += r.x * r.x + r.y * r.y;
        }
        double omg = (den > 1e-10) ? num / den : 0.0;
        for (int i = 0; i < n; ++i) {
            V2 r{c.x[o + i] - ctr.x, c.y[o + i] - ctr.y};
            c.vx[o + i] = vcm.x - omg * r.y;
            c.vy[o + i] = vcm.y + omg * r.x;
        }
    }
    for (int i = 0; i < n; ++i) {
        c.x[o + i] += c.vx[o + i] * dt;
        c.y[o + i] += c.vy[o + i] * dt;
    }
    if (c.periodic_nx[cap] > 0) {
        double Lx = static_cast<double>(c.periodic_nx[cap]);
        for (int i = 0; i < n; ++i) {
            while (c.x[o + i] < 0.0) c.x[o + i] += Lx;
            while (c.x[o + i] >= Lx) c.x[o + i] -= Lx;
        }
    }
}

}  // namespace

// ===========================================================================
extern "C" {

void sfo_default_params(sfo_params* p) {                       // core/parameters.h
    std::memset(p, 0, sizeof(*p));
    p->nx = 200; p->ny = 50; p->dt = 1.0; p->ibm_iterations = 1;
    p->tau = 0.8; p->rho0 = 1.0; p->boundary_type = 0;
    p->inlet_velocity = 0.02; p->outlet_density = 1.0;          /* parameters.h:44-45 */
    p->apply_gravity_to_fluid = 1; p->apply_gravity_to_capsules = 1;
    p->mrt_se = 1.4; p->mrt_s_eps = 1.4; p->mrt_sq = 1.4;
    p->max_lattice_force = 0.01;
    p->rep_epsilon = 0.05; p->rep_sigma = 1.0; p->rep_r_cut = 3.5; p->rep_power = 4;
    p->n_species = 1; p->diffusivity[0] = 0.01;
}

void sfo_default_membrane(sfo_membrane* m) {
    std::memset(m, 0, sizeof(*m));
    m->model = 0; m->k_stretch = 0.1; m->G_s = 0.1; m->C_skalak = 1.0;
    m->wlc_L_max_ratio = 2.2; m->k_bend = 0.005; m->k_area = 0.5; m->k_perimeter = 0.1;
    m->gamma_visc = 0.01; m->eta_membrane = 0.05; m->viscosity_ratio = 1.0; m->density = 1.0;
    m->aspect_ratio = 1.0; m->indent_depth = 0.4;
}

sfo_sim* sfo_create(const sfo_params* p) {
    sfo_sim* s = new sfo_sim();
    s->p = *p;
    s->nx = p->nx; s->ny = p->ny; s->N = p->nx * p->ny;
    s->periodic_x = p->boundary_type == 1;
    s->periodic_y = p->periodic_y != 0;
    size_t N = static_cast<size_t>(s->N);
    s->f.assign(Q * N, 0.0); s->ftmp.assign(Q * N, 0.0);
    s->rho.assign(N, 0.0); s->ux.assign(N, 0.0); s->uy.assign(N, 0.0);
    s->Fx.assign(N, 0.0); s->Fy.assign(N, 0.0);
    s->flags.assign(N, FLUID);
    build_neighbors(*s);
    if (p->scalar_enabled) {
        s->nsp = p->n_species;
        s->g.assign(s->nsp, std::vector<double>(Q * N, 0.0));
        s->gtmp.assign(s->nsp, std::vector<double>(Q * N, 0.0));
        s->C.assign(s->nsp, std::vector<double>(N, 0.0));
        s->tau_s.resize(s->nsp); s->c_inlet.resize(s->nsp);
        for (int k = 0; k < s->nsp; ++k) {
            s->tau_s[k] = 3.0 * p->diffusivity[k] + 0.5;
            s->c_inlet[k] = p->inlet_concentration[k];
        }
    }
    return s;
}

void sfo_destroy(sfo_sim* s) { delete s; }

void sfo_add_capsule(sfo_sim* s, double cx, double cy, double radius, int num_nodes,
                     const sfo_membrane* m, int type) {
    add_capsule(s->caps, cx, cy, radius, num_nodes, *m, type);
    if (s->periodic_x) s->caps.periodic_nx.back() = s->nx;
}

// Simulation::addCapsuleRandom: engine/simulation.cpp:220-274
int sfo_add_capsule_random(sfo_sim* s, int count, double x0, double y0, double x1,
                           double y1, double rmin, double rmax, int num_nodes,
                           const sfo_membrane* m, int type, int seed, double min_gap,
                           int max_attempts) {
    std::mt19937 rng(static_cast<unsigned>(seed));
    std::uniform_real_distribution<double> dx(x0, x1), dy(y0, y1), dr(rmin, rmax);
    int placed = 0, attempts = 0;
    int total = (max_attempts > 0) ? max_attempts : count * 500;
    const double wall_bot = 0.0, wall_top = static_cast<double>(s->ny - 1);
    while (placed < count && attempts < total) {
        double cx = dx(rng);
        double cy = dy(rng);
        double r = dr(rng);
        bool overlap = false;
        for (int c = 0; c < s->caps.n(); ++c) {
            V2 oc = cap_centroid(s->caps, c);
            double orr = std::sqrt(s->caps.rest_area[c] / kPI);
            double dist = norm(V2{cx - oc.x, cy - oc.y});
            if (dist < r + orr + min_gap) { overlap = true; break; }
        }
        if (cy - r < wall_bot + min_gap || cy + r > wall_top - min_gap) overlap = true;
        for (const OObstacle& ob : s->obs)
            if (ob.signed_distance(cx, cy) < r + min_gap) { overlap = true; break; }
        if (!overlap) {
            int nn = num_nodes > 0 ? num_nodes
                                   : std::max(static_cast<int>(std::ceil(2.0 * kPI * r / 0.75)), 12);
            sfo_add_capsule(s, cx, cy, r, nn, m, type);
            ++placed;
        }
        ++attempts;
    }
    return placed;
}

// Simulation::initialize + LBMSolver::initialize: simulation.cpp:429-468,
// lbm_solver.cpp:130-170, channel_builder.cpp:6-36
void sfo_add_obstacle(sfo_sim* s, int kind, const double* c, int n) {
    OObstacle o{};
    o.kind = kind;
    if (kind == 0) { o.a[0] = c[0]; o.a[1] = c[1]; o.a[2] = c[2]; }
    else if (kind == 1) {                                          // rect_obstacle.h:10-12
        o.a[0] = std::min(c[0], c[2]); o.a[1] = std::min(c[1], c[3]);
        o.a[2] = std::max(c[0], c[2]); o.a[3] = std::max(c[1], c[3]);
    } else {
        for (int i = 0; i + 1 < n; i += 2) o.poly.push_back(V2{c[i], c[i + 1]});
    }
    s->obs.push_back(o);
}

double sfo_obstacle_distance(const sfo_sim* s, int i, double x, double y, double* normal) {
    const OObstacle& o = s->obs[i];
    if (normal) { V2 nn = o.normal(x, y); normal[0] = nn.x; normal[1] = nn.y; }
    return o.signed_distance(x, y);
}

int sfo_ibb_links(const sfo_sim* s, int* xyq, double* qf) {
    const int n = static_cast<int>(s->ibb.size());
    for (int i = 0; i < n; ++i) {
        if (xyq) { xyq[3 * i] = s->ibb[i].x; xyq[3 * i + 1] = s->ibb[i].y; xyq[3 * i + 2] = s->ibb[i].q; }
        if (qf) qf[i] = s->ibb[i].qf;
    }
    return n;
}

void sfo_initialize(sfo_sim* s) {
    const int nx = s->nx, ny = s->ny;
    for (int y = 0; y < ny; ++y)                                   // channel_builder.cpp:27-33
        for (int x = 0; x < nx; ++x)
            for (const OObstacle& o : s->obs)
                if (o.contains(static_cast<double>(x), static_cast<double>(y))) {
                    s->flags[y * nx + x] = SOLID;
                    break;
                }
    for (int x = 0; x < nx; ++x) { s->flags[x] = SOLID; s->flags[(ny - 1) * nx + x] = SOLID; }
    if (s->p.boundary_type == 0) {
        for (int y = 1; y < ny - 1; ++y) { s->flags[y * nx] = INLET; s->flags[y * nx + nx - 1] = OUTLET; }
    } else if (s->p.boundary_type == 2) {
        for (int y = 0; y < ny; ++y) { s->flags[y * nx] = SOLID; s->flags[y * nx + nx - 1] = SOLID; }
    }
    if (s->periodic_y)
        for (int x = 0; x < nx; ++x) { s->flags[x] = FLUID; s->flags[(ny - 1) * nx + x] = FLUID; }
    const int N = s->N;
    for (int n = 0; n < N; ++n) {
        s->rho[n] = s->p.rho0; s->ux[n] = 0.0; s->uy[n] = 0.0;
        for (int q = 0; q < Q; ++q) s->f[q * N + n] = feq(q, s->p.rho0, 0.0, 0.0);
    }
    std::fill(s->Fx.begin(), s->Fx.end(), 0.0);
    std::fill(s->Fy.begin(), s->Fy.end(), 0.0);
    macroscopic(*s);
    s->solid.clear();
    for (int n = 0; n < N; ++n) if (s->flags[n] == SOLID) s->solid.push_back(n);
    if (s->p.use_interpolated_bb) ibb_build(*s);                   // simulation.cpp:440-442
    if (s->nsp) sfo_scalar_initialize(s, 0.0);
    if (s->periodic_x) for (auto& v : s->caps.periodic_nx) v = s->nx;
    s->step_count = 0;
}

void sfo_scalar_initialize(sfo_sim* s, double c0) {          // advection_diffusion.cpp:46-57
    for (int k = 0; k < s->nsp; ++k)
        for (int n = 0; n < s->N; ++n) {
            s->C[k][n] = c0;
            for (int q = 0; q < Q; ++q) s->g[k][n * Q + q] = W[q] * c0;
        }
}

void sfo_set_scalar_point(sfo_sim* s, int x, int y, double c, int sp) {   // :72-80
    if (x < 0 || x >= s->nx || y < 0 || y >= s->ny || sp < 0 || sp >= s->nsp) return;
    int n = y * s->nx + x;
    s->C[sp][n] = c;
    for (int q = 0; q < Q; ++q) s->g[sp][n * Q + q] = W[q] * c;
}

// Simulation::step: engine/simulation.cpp:470-787 (in-scope stages only)
void sfo_step(sfo_sim* s, int nsteps) {
    const sfo_params& p = s->p;
    Repulsion R{p.rep_epsilon, p.rep_sigma, p.rep_r_cut, p.rep_damping_normal,
                p.rep_friction_coeff, p.rep_power,
                p.boundary_type == 1 ? static_cast<double>(p.nx) : 0.0};
    for (int it = 0; it < nsteps; ++it) {
        std::fill(s->Fx.begin(), s->Fx.end(), 0.0);                 // :474
        std::fill(s->Fy.begin(), s->Fy.end(), 0.0);
        Caps& c = s->caps;
        std::fill(c.fx.begin(), c.fx.end(), 0.0);                   // :477
        std::fill(c.fy.begin(), c.fy.end(), 0.0);
        for (int cap = 0; cap < c.n(); ++cap) membrane_forces(c, cap);   // :481
        repulsion_all(c, R, 0.5, static_cast<double>(p.ny) - 1.5);   // :486-497
        for (const OObstacle& ob : s->obs) obstacle_repulsion(c, p, ob);
        if (p.apply_gravity_to_capsules && (p.gravity_x != 0.0 || p.gravity_y != 0.0)) {
            for (int cap = 0; cap < c.n(); ++cap) {                 // :516-527, capsule.cpp:673-690
                double drho = c.density[cap] - p.rho0;
                if (std::abs(drho) < 1e-12) continue;
                int o = c.off[cap], n = c.off[cap + 1] - o;
                double A = cap_area(c, cap);
                double inv_n = 1.0 / static_cast<double>(n);
                double gx = drho * A * p.gravity_x * inv_n, gy = drho * A * p.gravity_y * inv_n;
                for (int i = 0; i < n; ++i) { c.fx[o + i] += gx; c.fy[o + i] += gy; }
            }
        }
        const int ncorr = s->mdf_tol > 0.0 ? s->mdf_max : p.ibm_iterations - 1;
        if (ncorr > 0) mdf(*s, ncorr);                              // :531-538
        else spread_all(*s);
        const bool hb = (p.body_force_x != 0.0 || p.body_force_y != 0.0);   // :541-572
        const bool hg = p.apply_gravity_to_fluid && (p.gravity_x != 0.0 || p.gravity_y != 0.0);
        if (hb || hg) {
            for (int n = 0; n < s->N; ++n) {
                if (s->flags[n] != FLUID) continue;
                if (hb) { s->Fx[n] += p.body_force_x; s->Fy[n] += p.body_force_y; }
                if (hg) { double r = s->rho[n]; s->Fx[n] += r * p.gravity_x; s->Fy[n] += r * p.gravity_y; }
            }
        }
        s->capped = sfo_cap_forces(s->Fx.data(), s->Fy.data(), s->N, p.max_lattice_force);   // :588-609
        lbm_step(*s);                                               // :627-629
        if (s->nsp) adr_step(*s);                                   // :637-654
        for (int cap = 0; cap < c.n(); ++cap) {                     // :657-659
            for (int k = c.off[cap]; k < c.off[cap + 1]; ++k) {
                V2 u = interp_at(*s, c.x[k], c.y[k]);
                c.vx[k] = u.x; c.vy[k] = u.y;
            }
        }
        for (int cap = 0; cap < c.n(); ++cap) move_capsule(c, cap, p.dt);   // :662
        if (!p.periodic_y)                                          // :665-675
            sfo_clamp_nodes_y(c.y.data(), static_cast<int>(c.y.size()), p.ny);
        if (p.capsule_periodic_x && p.boundary_type != 1)           // :687-700
            sfo_recirculate_x(c.x.data(), c.y.data(), c.off.data(), c.n(), p.nx);
        s->step_count++;
    }
}

// Simulation::step 6b, simulation.cpp:588-609: every cell (any type) whose
// |F|^2 exceeds max^2 is scaled onto |F| = max; the count is reported.
int sfo_cap_forces(double* Fx, double* Fy, int N, double max_force) {
    const double Fm = max_force, Fm2 = Fm * Fm;
    int capped = 0;
    for (int n = 0; n < N; ++n) {
        double fx = Fx[n], fy = Fy[n];
        double m2 = fx * fx + fy * fy;
        if (m2 > Fm2) {
            double sc = Fm / std::sqrt(m2);
            Fx[n] *= sc; Fy[n] *= sc;
            ++capped;
        }
    }
    return capped;
}

// Simulation::step 10, simulation.cpp:665-675 (walls in y only).
void sfo_clamp_nodes_y(double* y, int K, int ny) {
    const double ymin = 1.5, ymax = static_cast<double>(ny) - 1.5;
    for (int k = 0; k < K; ++k) {
        if (y[k] < ymin) y[k] = ymin;
        if (y[k] > ymax) y[k] = ymax;
    }
}

// Simulation::step 10a, simulation.cpp:687-700.  Over a non-periodic fluid the
// capsules carry no periodic length, so Capsule::centroid is the plain mean
// of the nodes (membrane/capsule.cpp:168-172).
int sfo_recirculate_x(double* x, const double* y, const int* off, int ncaps, int nx) {
    (void)y;
    const double Lx = static_cast<double>(nx);
    int moved = 0;
    for (int c = 0; c < ncaps; ++c) {
        const int o = off[c], n = off[c + 1] - o;
        double sx = 0.0;
        for (int i = 0; i < n; ++i) sx += x[o + i];
        const double cx = sx / static_cast<double>(n);
        double shift = 0.0;
        if (cx >= Lx) shift = -Lx;
        else if (cx < 0.0) shift = Lx;
        if (shift != 0.0) {
            for (int i = 0; i < n; ++i) x[o + i] += shift;
            ++moved;
        }
    }
    return moved;
}

int sfo_num_capsules(const sfo_sim* s) { return s->caps.n(); }
int sfo_total_nodes(const sfo_sim* s) { return s->caps.off.back(); }
void sfo_capsule_offsets(const sfo_sim* s, int* off) {
    for (size_t i = 0; i < s->caps.off.size(); ++i) off[i] = s->caps.off[i];
}
void sfo_get_nodes(const sfo_sim* s, double* x, double* y, double* vx, double* vy,
                   double* fx, double* fy) {
    const Caps& c = s->caps;
    size_t K = c.x.size();
    if (x) std::memcpy(x, c.x.data(), K * 8);
    if (y) std::memcpy(y, c.y.data(), K * 8);
    if (vx) std::memcpy(vx, c.vx.data(), K * 8);
    if (vy) std::memcpy(vy, c.vy.data(), K * 8);
    if (fx) std::memcpy(fx, c.fx.data(), K * 8);
    if (fy) std::memcpy(fy, c.fy.data(), K * 8);
}
void sfo_set_nodes(sfo_sim* s, const double* x, const double* y, const double* vx,
                   const double* vy) {
    Caps& c = s->caps;
    size_t K = c.x.size();
    if (x) std::memcpy(c.x.data(), x, K * 8);
    if (y) std::memcpy(c.y.data(), y, K * 8);
    if (vx) std::memcpy(c.vx.data(), vx, K * 8);
    if (vy) std::memcpy(c.vy.data(), vy, K * 8);
}
void sfo_get_field(const sfo_sim* s, double* rho, double* ux, double* uy, double* fx,
                   double* fy) {
    size_t N = s->N;
    if (rho) std::memcpy(rho, s->rho.data(), N * 8);
    if (ux) std::memcpy(ux, s->ux.data(), N * 8);
    if (uy) std::memcpy(uy, s->uy.data(), N * 8);
    if (fx) std::memcpy(fx, s->Fx.data(), N * 8);
    if (fy) std::memcpy(fy, s->Fy.data(), N * 8);
}
void sfo_get_f(const sfo_sim* s, double* f) { std::memcpy(f, s->f.data(), s->f.size() * 8); }
void sfo_set_f(sfo_sim* s, const double* f) { std::memcpy(s->f.data(), f, s->f.size() * 8); }
void sfo_set_macro(sfo_sim* s, const double* rho, const double* ux, const double* uy) {
    size_t N = s->N;
    if (rho) std::memcpy(s->rho.data(), rho, N * 8);
    if (ux) std::memcpy(s->ux.data(), ux, N * 8);
    if (uy) std::memcpy(s->uy.data(), uy, N * 8);
}
void sfo_get_flags(const sfo_sim* s, int* flags) { std::memcpy(flags, s->flags.data(), s->N * 4); }
// LatticeField::setCellType for every cell, then LBMSolver::rebuildBoundaryNodeLists
// (lbm_solver.h:47, lbm_solver.cpp:109-128)
void sfo_set_flags(sfo_sim* s, const int* flags) {
    std::memcpy(s->flags.data(), flags, s->N * 4);
    s->solid.clear();
    for (int n = 0; n < s->N; ++n) if (s->flags[n] == SOLID) s->solid.push_back(n);
}
void sfo_get_scalar(const sfo_sim* s, double* C, int sp) {
    if (sp >= 0 && sp < s->nsp) std::memcpy(C, s->C[sp].data(), s->N * 8);
}
void sfo_set_concentration(sfo_sim* s, const double* C, int sp) {
    if (sp >= 0 && sp < s->nsp) std::memcpy(s->C[sp].data(), C, s->N * 8);
}
int sfo_last_capped(const sfo_sim* s) { return s->capped; }

double sfo_max_speed(const sfo_sim* s) {                        // simulation.cpp:892-907
    double m = 0.0;
    for (int n = 0; n < s->N; ++n) {
        if (s->flags[n] == SOLID) continue;
        double sp = s->ux[n] * s->ux[n] + s->uy[n] * s->uy[n];
        if (sp > m) m = sp;
    }
    return std::sqrt(m);
}

// what: 0 area, 1 perimeter, 2 centroid x, 3 centroid y, 4 Taylor D,
//       5 rest area, 6 rest perimeter, 7 inclination
double sfo_capsule_geometry(const sfo_sim* s, int c, int what) {
    const Caps& cc = s->caps;
    switch (what) {
    case 0: return cap_area(cc, c);
    case 1: return cap_perimeter(cc, c);
    case 2: return cap_centroid(cc, c).x;
    case 3: return cap_centroid(cc, c).y;
    case 4: return cap_taylor_d(cc, c);
    case 5: return cc.rest_area[c];
    case 6: return cc.rest_per[c];
    case 7: {
        double mxx, myy, mxy;
        if (cc.off[c + 1] - cc.off[c] < 3) return 0.0;
        cap_moments(cc, c, mxx, myy, mxy);
        return 0.5 * std::atan2(2.0 * mxy, mxx - myy);
    }
    }
    return 0.0;
}

void sfo_rest_lengths(const sfo_sim* s, double* out) {
    std::memcpy(out, s->caps.L0.data(), s->caps.L0.size() * 8);
}

double sfo_phi(double r) { return phi(r); }
double sfo_phi_exact(double r) { return phi_exact(r); }

void sfo_collide_cell(int model, double tau, double se, double seps, double sq, double* f9,
                      double rho, double ux, double uy, double Fx, double Fy) {
    double om = 1.0 / tau;
    if (model == 2) collide_reg(f9, om, rho, ux, uy, Fx, Fy);
    else if (model == 1) collide_mrt(f9, om, se, seps, sq, rho, ux, uy, Fx, Fy);
    else collide_bgk(f9, om, rho, ux, uy, Fx, Fy);
}

void sfo_pair_force(const sfo_sim* s, const double* pa, const double* va, const double* pb,
                    const double* vb, double* out) {
    const sfo_params& p = s->p;
    Repulsion R{p.rep_epsilon, p.rep_sigma, p.rep_r_cut, p.rep_damping_normal,
                p.rep_friction_coeff, p.rep_power,
                p.boundary_type == 1 ? static_cast<double>(p.nx) : 0.0};
    V2 F = R.pair(V2{pa[0], pa[1]}, V2{va[0], va[1]}, V2{pb[0], pb[1]}, V2{vb[0], vb[1]});
    out[0] = F.x;
    out[1] = F.y;
}

void sfo_membrane_forces_only(sfo_sim* s) {
    Caps& c = s->caps;
    std::fill(c.fx.begin(), c.fx.end(), 0.0);
    std::fill(c.fy.begin(), c.fy.end(), 0.0);
    for (int cap = 0; cap < c.n(); ++cap) membrane_forces(c, cap);
}


void sfo_set_mdf_control(sfo_sim* s, double tol, int max_corrections, int probe) {
    s->mdf_tol = tol > 0.0 ? tol : 0.0;
    s->mdf_max = tol > 0.0 ? max_corrections : 0;
    s->mdf_probe = probe ? 1 : 0;
}

void sfo_mdf_stats(sfo_sim* s, double* out4) {
    const double n = s->st_steps;
    out4[0] = n > 0 ? s->st_applied / n : 0.0;
    out4[1] = n > 0 ? s->st_final / n : 0.0;
    out4[2] = s->st_max;
    out4[3] = n;
    s->st_applied = s->st_final = s->st_max = s->st_steps = 0.0;
}

// The warm-start network of the device engine (no reference counterpart),
// restated in FP64: x^ = (x - in_mean) / in_scale; a1 = GELU(W1 x^ + b1);
// a2 = GELU(W2 a1 + b2); out = W3 a2 + b3; y = out_mean + out_scale * out.
// GELU(z) = z Phi(z) (exact, with erf).  Parameter block W1[w][17], b1[w],
// W2[w][w], b2[w], W3[2][w], b3[2]; norm = in_mean[17], in_scale[17],
// out_mean[2], out_scale[2].
namespace {
constexpr int NI = 17, NO = 2;
inline double gelu_(double z) { return 0.5 * z * (1.0 + std::erf(z / std::sqrt(2.0))); }
inline double gelu_d_(double z) {
    return 0.5 * (1.0 + std::erf(z / std::sqrt(2.0))) + z * std::exp(-0.5 * z * z) / std::sqrt(2.0 * M_PI);
}
struct MlpView {
    int w;
    const double *W1, *b1, *W2, *b2, *W3, *b3;
    explicit MlpView(const double* P, int width) : w(width) {
        W1 = P; b1 = W1 + w * NI; W2 = b1 + w; b2 = W2 + w * w; W3 = b2 + w; b3 = W3 + NO * w;
    }
};
}  // namespace

int sfo_mlp_num_params(int width) { return width * NI + width + width * width + width + NO * width + NO; }

void sfo_mlp_predict(const double* P, const double* norm, int width, int n, const double* x, double* y) {
    MlpView v(P, width);
    std::vector<double> a1(width), a2(width);
    for (int s = 0; s < n; ++s) {
        double xh[NI];
        for (int j = 0; j < NI; ++j) xh[j] = (x[s * NI + j] - norm[j]) / norm[NI + j];
        for (int i = 0; i < width; ++i) {
            double z = v.b1[i];
            for (int j = 0; j < NI; ++j) z += v.W1[i * NI + j] * xh[j];
            a1[i] = gelu_(z);
        }
        for (int i = 0; i < width; ++i) {
            double z = v.b2[i];
            for (int j = 0; j < width; ++j) z += v.W2[i * width + j] * a1[j];
            a2[i] = gelu_(z);
        }
        for (int o = 0; o < NO; ++o) {
            double z = v.b3[o];
            for (int j = 0; j < width; ++j) z += v.W3[o * width + j] * a2[j];
            y[s * NO + o] = norm[2 * NI + o] + norm[2 * NI + NO + o] * z;
        }
    }
}

// L = 1/(2n) sum_s |out_s - t^_s|^2 with t^ = (t - out_mean) / out_scale, and dL/dP.
double sfo_mlp_loss_grad(const double* P, const double* norm, int width, int n, const double* x,
                         const double* t, double* grad) {
    MlpView v(P, width);
    const int np = sfo_mlp_num_params(width);
    std::fill(grad, grad + np, 0.0);
    double* gW1 = grad; double* gb1 = gW1 + width * NI; double* gW2 = gb1 + width;
    double* gb2 = gW2 + width * width; double* gW3 = gb2 + width; double* gb3 = gW3 + NO * width;
    std::vector<double> z1(width), a1(width), z2(width), a2(width), d1(width), d2(width);
    double L = 0.0;
    for (int s = 0; s < n; ++s) {
        double xh[NI], e[NO];
        for (int j = 0; j < NI; ++j) xh[j] = (x[s * NI + j] - norm[j]) / norm[NI + j];
        for (int i = 0; i < width; ++i) {
            double z = v.b1[i];
            for (int j = 0; j < NI; ++j) z += v.W1[i * NI + j] * xh[j];
            z1[i] = z; a1[i] = gelu_(z);
        }
        for (int i = 0; i < width; ++i) {
            double z = v.b2[i];
            for (int j = 0; j < width; ++j) z += v.W2[i * width + j] * a1[j];
            z2[i] = z; a2[i] = gelu_(z);
        }
        for (int o = 0; o < NO; ++o) {
            double z = v.b3[o];
            for (int j = 0; j < width; ++j) z += v.W3[o * width + j] * a2[j];
            const double th = (t[s * NO + o] - norm[2 * NI + o]) / norm[2 * NI + NO + o];
            const double d = z - th;
            L += d * d;
            e[o] = d / n;
        }
        for (int i = 0; i < width; ++i) {
            double acc = 0.0;
            for (int o = 0; o < NO; ++o) acc += v.W3[o * width + i] * e[o];
            d2[i] = acc * gelu_d_(z2[i]);
        }
        for (int j = 0; j < width; ++j) {
            double acc = 0.0;
            for (int i = 0; i < width; ++i) acc += v.W2[i * width + j] * d2[i];
            d1[j] = acc * gelu_d_(z1[j]);
        }
        for (int o = 0; o < NO; ++o) {
            for (int j = 0; j < width; ++j) gW3[o * width + j] += e[o] * a2[j];
            gb3[o] += e[o];
        }
        for (int i = 0; i < width; ++i) {
            for (int j = 0; j < width; ++j) gW2[i * width + j] += d2[i] * a1[j];
            gb2[i] += d2[i];
            for (int j = 0; j < NI; ++j) gW1[i * NI + j] += d1[i] * xh[j];
            gb1[i] += d1[i];
        }
    }
    return 0.5 * L / n;
}

}  // extern "C"
