// Capsule membrane forces (one wavefront per capsule) and capsule contact
// forces (cell list built by atomic binning + per-cell sort).
//
// Each node's force is the same ordered sequence of += / -= the reference
// applies in its per-capsule loops (membrane/capsule.cpp:149-157, 320-671):
// lanes compute per-edge and per-angle terms into LDS, then every node applies
// the terms that touch it in the reference's loop order.  Ring reductions
// (area, perimeter, unwrapped coordinates) are sequential FP64 sums.
#include "sf_internal.h"
#include "sf_math.h"

namespace sf {

// nodes per capsule held in LDS: the kernel is instantiated for capsules of
// up to 64 nodes (small LDS, so every capsule's workgroup can be resident at
// once) and up to 256

struct MemArgs {
    double *x, *y, *vx, *vy, *fx, *fy;
    const double *L0, *th0, *ds0;
    double* Lprev;
    int* has_prev;
    const int* off;
    const sf_capsule_desc* desc;
    int periodic_nx;
    int* counters;    // counters[0] (capped cells of this step) is cleared here
    // when set: each node's arc length element for the spread (half the
    // distances to its ring neighbours, minimum image over ibm_nx)
    double* arc;
    int ibm_nx;
};

__device__ __forceinline__ void mimg(double ax, double ay, double bx, double by, double Lx,
                                     double& dx, double& dy) {
    dx = ax - bx;
    dy = ay - by;
    if (Lx > 0.0) {
        if (dx > 0.5 * Lx) dx -= Lx;
        if (dx < -0.5 * Lx) dx += Lx;
    }
}

// Sequential FP64 sum init + v[0] + v[1] + ... + v[n-1] (one lane; the loads
// of a chunk go out before its additions).
__device__ __forceinline__ double serial_sum(const double* v, int n, double init) {
    double acc = init;
    int i = 0;
    for (; i + 8 <= n; i += 8) {
        double t[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) t[u] = v[i + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += t[u];
    }
    for (; i < n; ++i) acc += v[i];
    return acc;
}

template <int MAXN>
__device__ __forceinline__ void membrane_body(MemArgs a, int bx) {
    __shared__ double X[MAXN], Y[MAXN], FX[MAXN], FY[MAXN];
    __shared__ double Ex[MAXN], Ey[MAXN], Gx[MAXN], Gy[MAXN], Cc[MAXN], Ln[MAXN];
    // per-node inputs, all read from memory in the first pass
    __shared__ double VX[MAXN], VY[MAXN], RL[MAXN], RTH[MAXN], RDS[MAXN], LP[MAXN];
    __shared__ int skip[MAXN];
    __shared__ double red[2];
    const int cap = bx;
    const int lane = threadIdx.x;
    if (cap == 0 && lane == 0) a.counters[0] = 0;
    const int o = a.off[cap], n = a.off[cap + 1] - o;
    const sf_capsule_desc d = a.desc[cap];
    const double Lx = static_cast<double>(a.periodic_nx);
    for (int i = lane; i < n; i += 64) {
        X[i] = a.x[o + i];
        Y[i] = a.y[o + i];
        VX[i] = a.vx[o + i];
        VY[i] = a.vy[o + i];
        RL[i] = a.L0[o + i];
        RTH[i] = a.th0[o + i];
        RDS[i] = a.ds0[o + i];
        LP[i] = a.Lprev[o + i];
        FX[i] = 0.0;
        FY[i] = 0.0;
    }
    __syncthreads();
    if (a.arc) {
        const double Li = static_cast<double>(a.ibm_nx);
        for (int i = lane; i < n; i += 64) {
            const int ip = (i - 1 + n) % n, in = (i + 1) % n;
            const double dp = min_image_ds(X[i], Y[i], X[ip], Y[ip], Li);
            const double dn = min_image_ds(X[in], Y[in], X[i], Y[i], Li);
            a.arc[o + i] = 0.5 * (dp + dn);
        }
    }
    if (d.is_rigid) {
        for (int i = lane; i < n; i += 64) { a.fx[o + i] = 0.0; a.fy[o + i] = 0.0; }
        return;
    }
    auto nxt = [n](int i) { return (i + 1) % n; };
    auto prv = [n](int i) { return (i - 1 + n) % n; };
    // edge forces applied as +F to node i and -F to node i+1, in edge order
    auto apply_edges = [&](bool plus_first) {
        for (int j = lane; j < n; j += 64) {
            if (j >= 1) {
                const int e = j - 1;
                if (!skip[e]) { if (plus_first) { FX[j] -= Ex[e]; FY[j] -= Ey[e]; } else { FX[j] += Ex[e]; FY[j] += Ey[e]; } }
                if (!skip[j]) { if (plus_first) { FX[j] += Ex[j]; FY[j] += Ey[j]; } else { FX[j] -= Ex[j]; FY[j] -= Ey[j]; } }
            } else {
                if (!skip[0]) { if (plus_first) { FX[0] += Ex[0]; FY[0] += Ey[0]; } else { FX[0] -= Ex[0]; FY[0] -= Ey[0]; } }
                const int e = n - 1;
                if (!skip[e]) { if (plus_first) { FX[0] -= Ex[e]; FY[0] -= Ey[e]; } else { FX[0] += Ex[e]; FY[0] += Ey[e]; } }
            }
        }
    };

    // -- stretching: capsule.cpp:320-470 --------------------------------
    for (int i = lane; i < n; i += 64) {
        const int j = nxt(i);
        double dx, dy;
        mimg(X[j], Y[j], X[i], Y[i], Lx, dx, dy);
        const double L = sqrt(dx * dx + dy * dy);
        skip[i] = L < 1e-15;
        if (skip[i]) continue;
        const double L0 = RL[i];
        double Fm;
        if (d.model == 1) {
            double lam = L / L0;
            double li3 = 1.0 / (lam * lam * lam);
            Fm = d.G_s * (lam - li3);
        } else if (d.model == 2) {
            double lam = L / L0;
            double lam2 = lam * lam;
            Fm = d.G_s * lam * (lam2 - 1.0) * (1.0 + d.C_skalak);
        } else if (d.model == 3) {
            double Lmax = d.wlc_L_max_ratio * L0;
            double xx = L / Lmax;
            if (xx > 0.99) xx = 0.99;
            double Fw = d.wlc_kBT_p * (1.0 / (4.0 * (1.0 - xx) * (1.0 - xx)) - 0.25 + xx);
            double Fp = -d.wlc_k_pow / L;
            Fm = Fw + Fp;
        } else {
            Fm = d.k_stretch * (L / L0 - 1.0);
        }
        const double ux = dx / L, uy = dy / L;
        Ex[i] = ux * Fm;
        Ey[i] = uy * Fm;
    }
    __syncthreads();
    apply_edges(true);
    __syncthreads();

    // -- bending: capsule.cpp:476-541 ------------------------------------
    if (d.use_helfrich) {
        for (int i = lane; i < n; i += 64) {
            const int ip = prv(i), in = nxt(i);
            double epx, epy, enx, eny;
            mimg(X[i], Y[i], X[ip], Y[ip], Lx, epx, epy);
            mimg(X[in], Y[in], X[i], Y[i], Lx, enx, eny);
            const double lp2 = epx * epx + epy * epy, ln2 = enx * enx + eny * eny;
            skip[i] = (lp2 < 1e-30 || ln2 < 1e-30);
            if (skip[i]) continue;
            const double dth = atan2(epx * eny - epy * enx, epx * enx + epy * eny) - RTH[i];
            Cc[i] = -d.k_bend * dth / RDS[i];
            Ex[i] = -eny / ln2;   // d theta / d x_{i+1}
            Ey[i] = enx / ln2;
            Gx[i] = -epy / lp2;   // d theta / d x_{i-1}
            Gy[i] = epx / lp2;
        }
        __syncthreads();
        for (int j = lane; j < n; j += 64) {
            int cand[3] = {prv(j), j, nxt(j)};
            // ascending, de-duplicated angle indices touching node j
            for (int u = 0; u < 3; ++u)
                for (int v = u + 1; v < 3; ++v)
                    if (cand[v] < cand[u]) { int t = cand[u]; cand[u] = cand[v]; cand[v] = t; }
            for (int u = 0; u < 3; ++u) {
                const int i = cand[u];
                if (u > 0 && cand[u - 1] == i) continue;
                if (skip[i]) continue;
                const double c = Cc[i];
                if (nxt(i) == j) { FX[j] += Ex[i] * c; FY[j] += Ey[i] * c; }
                if (prv(i) == j) { FX[j] += Gx[i] * c; FY[j] += Gy[i] * c; }
                if (i == j) {
                    const double sx = Ex[i] + Gx[i], sy = Ey[i] + Gy[i];
                    FX[j] -= sx * c; FY[j] -= sy * c;
                }
            }
        }
    } else {
        for (int i = lane; i < n; i += 64) {
            double px, py, qx, qy;
            mimg(X[prv(i)], Y[prv(i)], X[i], Y[i], Lx, px, py);
            mimg(X[nxt(i)], Y[nxt(i)], X[i], Y[i], Lx, qx, qy);
            const double lx = px + qx, ly = py + qy;
            FX[i] += lx * d.k_bend;
            FY[i] += ly * d.k_bend;
        }
    }
    __syncthreads();

    // -- area penalty: capsule.cpp:547-586 (area() and the signed sum) ----
    // Periodic x: the ring unwrapped from node 0 (capsule.cpp:177-186) by the
    // sequential sum of the minimum-image steps, then the shoelace terms; one
    // lane adds them in node order.  Ex/Ey hold the steps, Gx/Gy the unwrapped
    // coordinates, Cc the terms.
    if (Lx > 0.0) {
        for (int i = lane; i < n; i += 64)
            if (i >= 1) mimg(X[i], Y[i], X[i - 1], Y[i - 1], Lx, Ex[i], Ey[i]);
        __syncthreads();
        if (lane == 0 && n > 0) {
            double ux = X[0], uy = Y[0];
            Gx[0] = ux; Gy[0] = uy;
            int i = 1;
            for (; i + 8 <= n; i += 8) {
                double tx[8], ty[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) { tx[u] = Ex[i + u]; ty[u] = Ey[i + u]; }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    ux = ux + tx[u]; uy = uy + ty[u];
                    Gx[i + u] = ux; Gy[i + u] = uy;
                }
            }
            for (; i < n; ++i) {
                ux = ux + Ex[i]; uy = uy + Ey[i];
                Gx[i] = ux; Gy[i] = uy;
            }
        }
        __syncthreads();
        for (int i = lane; i < n; i += 64) {
            const int j = (i + 1) % n;
            Cc[i] = Gx[i] * Gy[j] - Gy[i] * Gx[j];
        }
    } else {
        for (int i = lane; i < n; i += 64) {
            const int j = nxt(i);
            Cc[i] = X[i] * Y[j] - Y[i] * X[j];
        }
    }
    __syncthreads();
    if (lane == 0) red[0] = serial_sum(Cc, n, 0.0);
    __syncthreads();
    {
        const double A0 = d.rest_area;
        if (!(A0 < 1e-15)) {
            const double sa = red[0];
            const double area = 0.5 * fabs(sa);
            const double coeff = -d.k_area * (area - A0) / A0;
            const double sign = (sa >= 0.0) ? 1.0 : -1.0;
            for (int i = lane; i < n; i += 64) {
                double epx, epy, enx, eny;
                mimg(X[i], Y[i], X[prv(i)], Y[prv(i)], Lx, epx, epy);
                mimg(X[nxt(i)], Y[nxt(i)], X[i], Y[i], Lx, enx, eny);
                double ox = (epy + eny) * 0.5, oy = (-epx + -enx) * 0.5;
                const double len = sqrt(ox * ox + oy * oy);
                if (len < 1e-15) continue;
                ox = ox / len * sign;
                oy = oy / len * sign;
                FX[i] += ox * coeff;
                FY[i] += oy * coeff;
            }
        }
    }
    __syncthreads();

    // -- perimeter penalty: capsule.cpp:588-604 --------------------------
    for (int i = lane; i < n; i += 64) {
        const int j = nxt(i);
        double dx, dy;
        mimg(X[j], Y[j], X[i], Y[i], Lx, dx, dy);
        const double L = sqrt(dx * dx + dy * dy);
        Ln[i] = L;
        skip[i] = L < 1e-15;
        if (!skip[i]) { Ex[i] = dx / L; Ey[i] = dy / L; }
    }
    __syncthreads();
    if (lane == 0) red[1] = serial_sum(Ln, n, 0.0);
    __syncthreads();
    {
        const double P0 = d.rest_perimeter;
        if (!(P0 < 1e-15)) {
            const double coeff = -d.k_perimeter * (red[1] - P0) / P0;
            for (int i = lane; i < n; i += 64)
                if (!skip[i]) { Ex[i] = Ex[i] * coeff; Ey[i] = Ey[i] * coeff; }
            __syncthreads();
            apply_edges(false);
        }
    }
    __syncthreads();

    // -- nodal damping: capsule.cpp:606-610 ------------------------------
    for (int i = lane; i < n; i += 64) {
        FX[i] -= VX[i] * d.gamma_visc;
        FY[i] -= VY[i] * d.gamma_visc;
    }
    __syncthreads();

    // -- Kelvin-Voigt membrane viscosity: capsule.cpp:612-671 ------------
    if (d.eta_membrane > 0.0) {
        const bool first = a.has_prev[cap] == 0;
        for (int i = lane; i < n; i += 64) {
            const int j = nxt(i);
            double dx, dy;
            mimg(X[j], Y[j], X[i], Y[i], Lx, dx, dy);
            const double L = sqrt(dx * dx + dy * dy);
            if (first) { a.Lprev[o + i] = L; skip[i] = 1; continue; }
            skip[i] = L < 1e-15;
            if (skip[i]) continue;
            const double sr = (L - LP[i]) / RL[i];
            const double Fv = d.eta_membrane * sr;
            const double ux = dx / L, uy = dy / L;
            Ex[i] = ux * Fv;
            Ey[i] = uy * Fv;
            a.Lprev[o + i] = L;
        }
        __syncthreads();
        if (!first) apply_edges(true);
        __syncthreads();
        if (first && lane == 0) a.has_prev[cap] = 1;
    }
    for (int i = lane; i < n; i += 64) {
        a.fx[o + i] = FX[i];
        a.fy[o + i] = FY[i];
    }
}

template <int MAXN>
__global__ __launch_bounds__(64) void k_membrane(MemArgs a) {
    membrane_body<MAXN>(a, blockIdx.x);
}

static MemArgs mem_args(const Ctx& c) {
    MemArgs a{};
    a.x = c.x; a.y = c.y; a.vx = c.vx; a.vy = c.vy; a.fx = c.fx; a.fy = c.fy;
    a.L0 = c.L0; a.th0 = c.th0; a.ds0 = c.ds0; a.Lprev = c.Lprev; a.has_prev = c.has_prev;
    a.off = c.off; a.desc = c.desc;
    a.periodic_nx = c.periodic_x ? c.nx : 0;
    a.counters = c.counters;
    a.arc = c.arc;
    a.ibm_nx = c.nx;
    return a;
}

void launch_membrane(Ctx& c) {
    if (c.ncaps == 0) return;
    const MemArgs a = mem_args(c);
    if (c.max_nodes <= 64)
        hipLaunchKernelGGL(k_membrane<64>, dim3(c.ncaps), dim3(64), 0, c.stream, a);
    else
        hipLaunchKernelGGL(k_membrane<256>, dim3(c.ncaps), dim3(64), 0, c.stream, a);
}

// ---------------------------------------------------------------------------
// Contact forces: membrane/repulsion.cpp:6-178, membrane/cell_list.h:36-146.
struct RepArgs {
    const double *x, *y, *vx, *vy;
    double *fx, *fy;
    const int *off, *owner;
    int K, ncaps;
    double eps, sig, rcut, gn, mu, Lx;
    int pw;
    double ybot, ytop;
    // cell list: counts and per-entry node data in node-index order
    const int* cnt;
    const double4* pay;
    const int* own;
    double rcut2_hi;   // > rcut^2 by a margin: beyond it no pair can be within rcut
    int ncx, ncy;
    double xorg, cdx, cdy;
    // when set: distance from each node to the nearest node of another
    // capsule among those examined (capped at 2 r_cut)
    float* dmin;
    // obstacles, after the wall images (simulation.cpp:495-499)
    const Obstacle* obs;
    int nobs;
};

__device__ __forceinline__ void obstacle_forces(const RepArgs& a, double px, double py, double vx,
                                                double vy, double& fx, double& fy) {
    for (int o = 0; o < a.nobs; ++o) {
        const Obstacle& ob = a.obs[o];
        if (obstacle_far(ob, px, py, a.rcut)) continue;
        obstacle_contact(ob, px, py, vx, vy, a.eps, a.sig, a.pw, a.rcut, a.gn, a.mu, fx, fy);
    }
}

__device__ __forceinline__ double pair_dist2(const RepArgs& a, double px, double py, double bx,
                                             double by) {
    double dx = bx - px;
    if (a.Lx > 0.0) {
        if (dx > 0.5 * a.Lx) dx -= a.Lx;
        if (dx < -0.5 * a.Lx) dx += a.Lx;
    }
    const double dy = by - py;
    return dx * dx + dy * dy;
}

__device__ __forceinline__ void pair_force(const RepArgs& a, double pax, double pay, double vax,
                                           double vay, double pbx, double pby, double vbx,
                                           double vby, double& Fx, double& Fy) {
    double dx = pax - pbx, dy = pay - pby;
    if (a.Lx > 0.0) {
        if (dx > 0.5 * a.Lx) dx -= a.Lx;
        if (dx < -0.5 * a.Lx) dx += a.Lx;
    }
    const double d2 = dx * dx + dy * dy;
    if (d2 > a.rcut2_hi) { Fx = 0.0; Fy = 0.0; return; }   // r >= rcut for certain
    const double r = sqrt(d2);
    if (r < 1e-15 || r >= a.rcut) { Fx = 0.0; Fy = 0.0; return; }
    const double nx = dx / r, ny = dy / r;
    const double Fr = a.eps * pow_int_cr(a.sig / r, a.pw);
    double Fn = Fr, tx = 0.0, ty = 0.0;
    if (a.gn > 0.0 || a.mu > 0.0) {
        const double rx = vax - vbx, ry = vay - vby;
        const double vn = rx * nx + ry * ny;
        Fn = Fr - a.gn * vn;
        if (Fn < 0.0) Fn = 0.0;
        if (a.mu > 0.0) {
            const double vtx = rx - nx * vn, vty = ry - ny * vn;
            const double vtm = sqrt(vtx * vtx + vty * vty);
            if (vtm > 1e-12) {
                const double s = -a.mu * Fn;
                tx = vtx / vtm * s;
                ty = vty / vtm * s;
            }
        }
    }
    Fx = nx * Fn + tx;
    Fy = ny * Fn + ty;
}

__device__ __forceinline__ int cell_index(const RepArgs& a, double px0, double py) {
    return contact_cell(px0, py, a.xorg, a.Lx, a.cdx, a.cdy, a.ybot, a.ncx, a.ncy);
}

// Every lane of the wavefront takes part in the slot claims (wave_claim):
// consecutive nodes mostly share a cell.
__device__ __forceinline__ void cells_fill_body(RepArgs a, int* cnt, int* list, int* counters, int bx) {
    const int k = bx * blockDim.x + threadIdx.x;
    const int ci = k < a.K ? cell_index(a, a.x[k], a.y[k]) : -1;
    contact_cell_claim(cnt, list, counters, ci, k);
}

__global__ void k_cells_fill(RepArgs a, int* cnt, int* list, int* counters) {
    cells_fill_body(a, cnt, list, counters, blockIdx.x);
}

// Sort each cell's list by node index (the reference's insertion order), lay
// the listed nodes' data out in that order, and clear the other step's count
// array for the next step.
__device__ __forceinline__ void cells_sort_body(int ncells, int nclear, const int* cnt, int* list, int* other,
                             const double* x, const double* y, const double* vx, const double* vy,
                             const int* owner, double4* pay, int* own, int bx) {
    const int ci = bx * blockDim.x + threadIdx.x;
    if (ci < nclear) other[ci] = 0;
    if (ci >= ncells) return;
    int m = cnt[ci];
    if (m > CELL_MAXC) m = CELL_MAXC;
    const size_t base = static_cast<size_t>(ci) * CELL_MAXC;
    int* l = list + base;
    for (int i = 1; i < m; ++i) {
        const int v = l[i];
        int j = i - 1;
        while (j >= 0 && l[j] > v) { l[j + 1] = l[j]; --j; }
        l[j + 1] = v;
    }
    for (int i = 0; i < m; ++i) {
        const int k = l[i];
        pay[base + i] = make_double4(x[k], y[k], vx[k], vy[k]);
        own[base + i] = owner[k];
    }
}

__global__ void k_cells_sort(int ncells, int nclear, const int* cnt, int* list, int* other,
                             const double* x, const double* y, const double* vx, const double* vy,
                             const int* owner, double4* pay, int* own) {
    cells_sort_body(ncells, nclear, cnt, list, other, x, y, vx, vy, owner, pay, own, blockIdx.x);
}

template <bool CELLS>
__device__ __forceinline__ void repulsion_body(RepArgs a, int bx) {
    const int k = bx * blockDim.x + threadIdx.x;
    if (k >= a.K) return;
    const int cap = a.owner[k];
    const double px = a.x[k], py = a.y[k], vx = a.vx[k], vy = a.vy[k];
    double fx = a.fx[k], fy = a.fy[k];
    double Fx, Fy;
    double d2min = 4.0 * a.rcut * a.rcut;
    if (!CELLS) {
        for (int j = 0; j < a.ncaps; ++j) {
            if (j == cap) continue;
            for (int b = a.off[j]; b < a.off[j + 1]; ++b) {
                pair_force(a, px, py, vx, vy, a.x[b], a.y[b], a.vx[b], a.vy[b], Fx, Fy);
                fx += Fx; fy += Fy;
                if (a.dmin) d2min = fmin(d2min, pair_dist2(a, px, py, a.x[b], a.y[b]));
            }
        }
    } else {
        const int ci = cell_index(a, px, py);
        const int cx0 = ci % a.ncx, cy0 = ci / a.ncx;
        for (int dyc = -1; dyc <= 1; ++dyc) {
            const int yy = cy0 + dyc;
            if (yy < 0 || yy >= a.ncy) continue;
            for (int dxc = -1; dxc <= 1; ++dxc) {
                int xx = cx0 + dxc;
                if (a.Lx > 0.0) { if (xx < 0) xx += a.ncx; if (xx >= a.ncx) xx -= a.ncx; }
                else if (xx < 0 || xx >= a.ncx) continue;
                const int cell = yy * a.ncx + xx;
                int m = a.cnt[cell];
                if (m > CELL_MAXC) m = CELL_MAXC;
                const size_t base = static_cast<size_t>(cell) * CELL_MAXC;
                for (int t = 0; t < m; ++t) {
                    if (a.own[base + t] == cap) continue;
                    const double4 b = a.pay[base + t];
                    pair_force(a, px, py, vx, vy, b.x, b.y, b.z, b.w, Fx, Fy);
                    fx += Fx; fy += Fy;
                    if (a.dmin) d2min = fmin(d2min, pair_dist2(a, px, py, b.x, b.y));
                }
            }
        }
    }
    pair_force(a, px, py, vx, vy, px, a.ybot, 0.0, 0.0, Fx, Fy);
    fx += Fx; fy += Fy;
    pair_force(a, px, py, vx, vy, px, a.ytop, 0.0, 0.0, Fx, Fy);
    fx += Fx; fy += Fy;
    obstacle_forces(a, px, py, vx, vy, fx, fy);
    a.fx[k] = fx;
    a.fy[k] = fy;
    if (a.dmin) a.dmin[k] = static_cast<float>(sqrt(d2min));
}

template <bool CELLS>
__global__ void k_repulsion(RepArgs a) {
    repulsion_body<CELLS>(a, blockIdx.x);
}

// Cell-list contact with REP_LANES lanes per node: lane u < 9 scans neighbour cell u
// (the reference's dy-outer, dx-inner order), lanes 9 and 10 take the two
// wall images, all in parallel; every non-zero pair force is kept in LDS with
// its place in that order, and the node's lane 0 adds them to the membrane
// force in order -- the sequence of additions of the serial loop, minus its
// exact zeros.  A lane with more than REP_SLOTS non-zero pairs makes lane 0
// redo the node serially.
constexpr int REP_LANES = 12;      // lanes per node: 9 cells, 2 wall images, one spare
constexpr int REP_NODES = 256 / REP_LANES;   // nodes per 256-thread workgroup (21; 4 threads idle)
constexpr int REP_SLOTS = 6;       // non-zero pair forces a lane keeps

// Obstacles: the lanes of a node screen the obstacles' holding discs (staged
// in LDS) in parallel and mark the ones within reach; the node's lane 0 then
// adds their terms in list order.  Up to REP_OBS_LDS obstacles; beyond that
// lane 0 screens them all itself.
constexpr int REP_OBS_LDS = 256;

// Lane 0 of a node: its membrane force plus the pair forces the lanes kept, in
// the reference's order (or the serial loop after an overflow), then the
// obstacles; stores the node's force and nearest distance.
__device__ __forceinline__ void repulsion_lanes_sum(const RepArgs& a, int k, int g, int cap, double px, double py,
                                                    double vx, double vy, bool obs_lds,
                                                    double (*sfx)[11][REP_SLOTS], double (*sfy)[11][REP_SLOTS],
                                                    int (*scnt)[REP_LANES], float (*sdmin)[REP_LANES],
                                                    unsigned long long (*snear)[REP_OBS_LDS / 64], double fx,
                                                    double fy) {
    double Fx, Fy;
    bool overflow = false;
    for (int u = 0; u < 11; ++u) overflow |= scnt[g][u] > REP_SLOTS;
    if (!overflow) {
        for (int u = 0; u < 11; ++u)
            for (int j = 0; j < scnt[g][u]; ++j) { fx += sfx[g][u][j]; fy += sfy[g][u][j]; }
    } else {
        // the serial loop of the reference (cells, then the wall images)
        const int ci = cell_index(a, px, py);
        const int cx0 = ci % a.ncx, cy0 = ci / a.ncx;
        for (int dyc = -1; dyc <= 1; ++dyc) {
            const int yy = cy0 + dyc;
            if (yy < 0 || yy >= a.ncy) continue;
            for (int dxc = -1; dxc <= 1; ++dxc) {
                int xx = cx0 + dxc;
                if (a.Lx > 0.0) { if (xx < 0) xx += a.ncx; if (xx >= a.ncx) xx -= a.ncx; }
                else if (xx < 0 || xx >= a.ncx) continue;
                const int cell = yy * a.ncx + xx;
                int m = a.cnt[cell];
                if (m > CELL_MAXC) m = CELL_MAXC;
                const size_t base = static_cast<size_t>(cell) * CELL_MAXC;
                for (int t = 0; t < m; ++t) {
                    if (a.own[base + t] == cap) continue;
                    const double4 b = a.pay[base + t];
                    pair_force(a, px, py, vx, vy, b.x, b.y, b.z, b.w, Fx, Fy);
                    fx += Fx; fy += Fy;
                }
            }
        }
        pair_force(a, px, py, vx, vy, px, a.ybot, 0.0, 0.0, Fx, Fy);
        fx += Fx; fy += Fy;
        pair_force(a, px, py, vx, vy, px, a.ytop, 0.0, 0.0, Fx, Fy);
        fx += Fx; fy += Fy;
    }
    if (obs_lds && snear) {
        for (int w = 0; w < REP_OBS_LDS / 64; ++w) {
            unsigned long long bits = snear[g][w];
            while (bits) {
                const int o = w * 64 + __ffsll(static_cast<long long>(bits)) - 1;
                bits &= bits - 1ull;
                obstacle_contact(a.obs[o], px, py, vx, vy, a.eps, a.sig, a.pw, a.rcut, a.gn, a.mu, fx, fy);
            }
        }
    } else {
        obstacle_forces(a, px, py, vx, vy, fx, fy);
    }
    a.fx[k] = fx;
    a.fy[k] = fy;
    if (a.dmin) {
        float dm = sdmin[g][0];
        for (int u = 1; u < 9; ++u) dm = fminf(dm, sdmin[g][u]);
        a.dmin[k] = dm;
    }
}

// OBS: the lattice may hold obstacles (their holding discs are staged in LDS,
// which the form without them leaves free for more resident workgroups).
template <bool OBS>
__device__ __forceinline__ void repulsion_lanes_body(RepArgs a, int bx) {
    __shared__ double sfx[REP_NODES][11][REP_SLOTS], sfy[REP_NODES][11][REP_SLOTS];
    __shared__ int scnt[REP_NODES][REP_LANES];
    __shared__ float sdmin[REP_NODES][REP_LANES];
    __shared__ double sbc[OBS ? REP_OBS_LDS : 1][3];
    __shared__ unsigned long long snear_lds[OBS ? REP_NODES : 1][REP_OBS_LDS / 64];
    unsigned long long (*snear)[REP_OBS_LDS / 64] = OBS ? snear_lds : nullptr;
    const int g = static_cast<int>(threadIdx.x) / REP_LANES, l = static_cast<int>(threadIdx.x) % REP_LANES;
    const bool grp = g < REP_NODES;          // the workgroup's last threads hold no node
    const int k = bx * REP_NODES + g;
    const bool live = grp && k < a.K;
    int cap = 0;
    double px = 0.0, py = 0.0, vx = 0.0, vy = 0.0;
    double fx0 = 0.0, fy0 = 0.0;   // the membrane force, read by the node's lane 0 up front
    if (live) { cap = a.owner[k]; px = a.x[k]; py = a.y[k]; vx = a.vx[k]; vy = a.vy[k]; }
    if (live && l == 0) { fx0 = a.fx[k]; fy0 = a.fy[k]; }
    const bool obs_lds = OBS && a.nobs > 0 && a.nobs <= REP_OBS_LDS;
    if (obs_lds) {
        for (int i = threadIdx.x; i < a.nobs; i += blockDim.x) {
            sbc[i][0] = a.obs[i].bc[0];
            sbc[i][1] = a.obs[i].bc[1];
            sbc[i][2] = a.obs[i].bc[2];
        }
        if (grp && l < REP_OBS_LDS / 64) snear[g][l] = 0ull;
        __syncthreads();
        if (live)
            for (int o = l; o < a.nobs; o += REP_LANES) {
                const double dx = px - sbc[o][0], dy = py - sbc[o][1];
                const double lim = sbc[o][2] + a.rcut;
                if (!(dx * dx + dy * dy > lim * lim * (1.0 + 1e-9)))   // obstacle_far, negated
                    atomicOr(&snear[g][o >> 6], 1ull << (o & 63));
            }
    }
    int nz = 0;
    double d2min = 4.0 * a.rcut * a.rcut;
    double Fx, Fy;
    auto keep = [&](double fx, double fy) {
        if (fx == 0.0 && fy == 0.0) return;
        if (nz < REP_SLOTS) { sfx[g][l][nz] = fx; sfy[g][l][nz] = fy; }
        ++nz;
    };
    if (live && l < 9) {
        const int ci = cell_index(a, px, py);
        const int cx0 = ci % a.ncx, cy0 = ci / a.ncx;
        const int yy = cy0 + l / 3 - 1;
        int xx = cx0 + l % 3 - 1;
        bool ok = yy >= 0 && yy < a.ncy;
        if (a.Lx > 0.0) { if (xx < 0) xx += a.ncx; if (xx >= a.ncx) xx -= a.ncx; }
        else ok = ok && xx >= 0 && xx < a.ncx;
        if (ok) {
            const int cell = yy * a.ncx + xx;
            int m = a.cnt[cell];
            if (m > CELL_MAXC) m = CELL_MAXC;
            const size_t base = static_cast<size_t>(cell) * CELL_MAXC;
            for (int t = 0; t < m; ++t) {
                if (a.own[base + t] == cap) continue;
                const double4 b = a.pay[base + t];
                pair_force(a, px, py, vx, vy, b.x, b.y, b.z, b.w, Fx, Fy);
                keep(Fx, Fy);
                if (a.dmin) d2min = fmin(d2min, pair_dist2(a, px, py, b.x, b.y));
            }
        }
    } else if (live && l < 11) {
        pair_force(a, px, py, vx, vy, px, l == 9 ? a.ybot : a.ytop, 0.0, 0.0, Fx, Fy);
        keep(Fx, Fy);
    }
    if (grp) {
        scnt[g][l] = nz;
        sdmin[g][l] = static_cast<float>(sqrt(d2min));
    }
    __syncthreads();
    if (live && l == 0)
        repulsion_lanes_sum(a, k, g, cap, px, py, vx, vy, obs_lds, sfx, sfy, scnt, sdmin, snear, fx0, fy0);
}

template <bool OBS>
__global__ __launch_bounds__(256) void k_repulsion_lanes(RepArgs a) {
    repulsion_lanes_body<OBS>(a, blockIdx.x);
}

// x extent of the cell grid when x is not periodic (cell_list.h:50-66):
// bounding box of the nodes, computed on the device by one workgroup.
struct CellGeom { double xorg, cdx; int ncx, pad; };
__global__ void k_cell_geom(const double* x, int K, double rcut, int max_ncx, CellGeom* out,
                            int* counters) {
    __shared__ double mn[256], mx[256];
    const int t = threadIdx.x;
    double a = 1e30, b = -1e30;
    for (int k = t; k < K; k += 256) { a = fmin(a, x[k]); b = fmax(b, x[k]); }
    mn[t] = a; mx[t] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) { mn[t] = fmin(mn[t], mn[t + w]); mx[t] = fmax(mx[t], mx[t + w]); }
        __syncthreads();
    }
    if (t == 0) {
        const double xorg = mn[0] - rcut;
        const double ext = (mx[0] - mn[0]) + 2.0 * rcut;
        int ncx = static_cast<int>(floor(ext / rcut));
        if (ncx < 1) ncx = 1;
        if (ncx > max_ncx) { atomicOr(&counters[2], 2); ncx = max_ncx; }
        out->xorg = xorg;
        out->cdx = ext / static_cast<double>(ncx);
        out->ncx = ncx;
    }
}

static RepArgs rep_args(const Ctx& c) {
    RepArgs a{};
    a.x = c.x; a.y = c.y; a.vx = c.vx; a.vy = c.vy; a.fx = c.fx; a.fy = c.fy;
    a.off = c.off; a.owner = c.owner; a.K = c.K; a.ncaps = c.ncaps;
    a.eps = c.p.rep_epsilon; a.sig = c.p.rep_sigma; a.rcut = c.p.rep_r_cut;
    a.gn = c.p.rep_damping_normal; a.mu = c.p.rep_friction_coeff; a.pw = c.p.rep_power;
    a.Lx = c.periodic_x ? static_cast<double>(c.nx) : 0.0;
    a.ybot = 0.5; a.ytop = static_cast<double>(c.ny) - 1.5;
    a.rcut2_hi = a.rcut * a.rcut * (1.0 + 1e-12);
    a.dmin = c.mdf_hist_on ? c.mdf_dmin : nullptr;
    a.obs = c.obs;
    a.nobs = c.nobs;
    return a;
}

// contacts are summed over all pairs up to this many nodes, over the cell list above
constexpr int REP_ALLPAIRS_MAX = 100;

bool contact_uses_cells(const Ctx& c) {
    return c.ncaps > 0 && c.K > REP_ALLPAIRS_MAX && c.periodic_x;
}

// The contact cell list of the current positions: filled here, or already by
// the previous step's closing interpolation (cells_prefilled).
void launch_contact_fill(Ctx& c) {
    RepArgs a = rep_args(c);
    a.ncy = c.ncy; a.cdy = c.cdy;
    a.ncx = c.ncx; a.xorg = 0.0; a.cdx = c.cdx;
    const int nmax = c.ncx * c.ncy;
    int* cnt = c.cell_cnt + static_cast<size_t>(c.cell_cur) * nmax;
    const int bs = 128, nb = (c.K + bs - 1) / bs;
    if (!c.cells_prefilled)
        hipLaunchKernelGGL(k_cells_fill, dim3(nb), dim3(bs), 0, c.stream, a, cnt, c.cell_list, c.counters);
    c.cells_prefilled = false;
    c.cell_built = cnt;
    c.cell_other = c.cell_cnt + static_cast<size_t>(1 - c.cell_cur) * nmax;
    c.cell_cur = 1 - c.cell_cur;
}

// The fill the step's closing interpolation does for the next step: the
// count array launch_contact_fill would fill (cleared by this step's sort), at
// the moved positions.  None without the cell list, or where the positions
// still change after the interpolation (rigid capsules move in their own
// kernel).
CellFill contact_fill_ahead(Ctx& c, bool move) {
    if (!move || !contact_uses_cells(c) || !c.cell_cnt) return CellFill{};
    const int nmax = c.ncx * c.ncy;
    c.cells_prefilled = true;
    return CellFill{c.cell_cnt + static_cast<size_t>(c.cell_cur) * nmax, c.cell_list, c.counters, 0.0,
                    static_cast<double>(c.nx), c.cdx, c.cdy, 0.5, c.ncx, c.ncy};
}

// Positions changed from the host: forget a fill done ahead.
void drop_contact_prefill(Ctx& c) {
    if (!c.cells_prefilled) return;
    if (c.cell_cnt) launch_fill_int(c, c.cell_cnt + static_cast<size_t>(c.cell_cur) * c.ncx * c.ncy,
                                    c.ncx * c.ncy, 0);
    c.cells_prefilled = false;
}

// Membrane forces (workgroups [0, ncaps)) and the contact cell sort (the rest,
// one thread per cell) in one launch of 64-thread workgroups.
template <int MAXN>
__global__ __launch_bounds__(64) void k_membrane_sort(MemArgs a, int ncaps, int ncells, int nclear,
                                                      const int* cnt, int* list, int* other,
                                                      const double* x, const double* y,
                                                      const double* vx, const double* vy,
                                                      const int* owner, double4* pay, int* own) {
    if (static_cast<int>(blockIdx.x) < ncaps) membrane_body<MAXN>(a, blockIdx.x);
    else cells_sort_body(ncells, nclear, cnt, list, other, x, y, vx, vy, owner, pay, own,
                         blockIdx.x - ncaps);
}

void launch_membrane_sort(Ctx& c) {
    const MemArgs a = mem_args(c);
    const int nmax = c.ncx * c.ncy;
    const dim3 g(c.ncaps + (nmax + 63) / 64);
    if (c.max_nodes <= 64)
        hipLaunchKernelGGL(k_membrane_sort<64>, g, dim3(64), 0, c.stream, a, c.ncaps, nmax, nmax,
                           c.cell_built, c.cell_list, c.cell_other, c.x, c.y, c.vx, c.vy, c.owner,
                           c.cell_pay, c.cell_own);
    else
        hipLaunchKernelGGL(k_membrane_sort<256>, g, dim3(64), 0, c.stream, a, c.ncaps, nmax, nmax,
                           c.cell_built, c.cell_list, c.cell_other, c.x, c.y, c.vx, c.vy, c.owner,
                           c.cell_pay, c.cell_own);
}

void launch_repulsion(Ctx& c) {
    if (c.ncaps == 0) return;
    RepArgs a = rep_args(c);
    const int bs = 128, nb = (c.K + bs - 1) / bs;
    if (c.K <= REP_ALLPAIRS_MAX) {
        hipLaunchKernelGGL((k_repulsion<false>), dim3(nb), dim3(bs), 0, c.stream, a);
        return;
    }
    a.ncy = c.ncy;
    a.cdy = c.cdy;
    if (c.periodic_x) {
        a.ncx = c.ncx; a.xorg = 0.0; a.cdx = c.cdx;
    } else {
        // the grid changes with the node bounding box; the kernels read it
        // from a device record written by k_cell_geom
        CellGeom* g = reinterpret_cast<CellGeom*>(c.counters + 8);
        hipLaunchKernelGGL(k_cell_geom, dim3(1), dim3(256), 0, c.stream, c.x, c.K, a.rcut, c.ncx, g, c.counters);
        // non-periodic grids are resolved on the host after a sync
        CellGeom h;
        hipMemcpyAsync(&h, g, sizeof(h), hipMemcpyDeviceToHost, c.stream);
        hipStreamSynchronize(c.stream);
        a.ncx = h.ncx; a.xorg = h.xorg; a.cdx = h.cdx;
    }
    int* cnt = c.cell_built;
    c.cell_built = nullptr;
    if (!cnt) {
        const int ncells = a.ncx * a.ncy;
        const int nmax = c.ncx * c.ncy;
        cnt = c.cell_cnt + static_cast<size_t>(c.cell_cur) * nmax;
        int* other = c.cell_cnt + static_cast<size_t>(1 - c.cell_cur) * nmax;
        hipLaunchKernelGGL(k_cells_fill, dim3(nb), dim3(bs), 0, c.stream, a, cnt, c.cell_list, c.counters);
        hipLaunchKernelGGL(k_cells_sort, dim3((nmax + 255) / 256), dim3(256), 0, c.stream, ncells, nmax, cnt,
                           c.cell_list, other, c.x, c.y, c.vx, c.vy, c.owner, c.cell_pay, c.cell_own);
        c.cell_cur = 1 - c.cell_cur;
    }
    a.cnt = cnt;
    a.pay = c.cell_pay;
    a.own = c.cell_own;
    auto kern = c.nobs > 0 ? k_repulsion_lanes<true> : k_repulsion_lanes<false>;
    hipLaunchKernelGGL(kern, dim3((c.K + REP_NODES - 1) / REP_NODES), dim3(256), 0,
                       c.stream, a);
}


// ---------------------------------------------------------------------------
// Batched launches over the members of an ensemble: each workgroup looks up
// (member, own workgroup index) in the map and runs the single-context body
// on that member's arguments.
struct CellsFillPack {
    RepArgs a;
    int* cnt;
    int* list;
    int* counters;
};
struct CellsSortPack {
    int ncells, nclear;
    const int* cnt;
    int* list;
    int* other;
    const double *x, *y, *vx, *vy;
    const int* owner;
    double4* pay;
    int* own;
};
struct ContactPack {
    RepArgs a;
    int allpairs;    // all-pairs sum (one thread per node) instead of the cell list
};
constexpr int BATCH_FILL_THREADS = 128;
constexpr int BATCH_ALLPAIRS_THREADS = 256;

__global__ __launch_bounds__(BATCH_FILL_THREADS) void kb_cells_fill(const CellsFillPack* packs,
                                                                   const int2* map) {
    const int2 w = map[blockIdx.x];
    if (w.x < 0) return;
    const CellsFillPack p = packs[w.x];
    cells_fill_body(p.a, p.cnt, p.list, p.counters, w.y);
}

template <int MAXN>
__global__ __launch_bounds__(64) void kb_membrane_sort(const MemArgs* mp, const CellsSortPack* sp,
                                                       const int2* map, int M) {
    const int2 w = map[blockIdx.x];
    if (w.x < 0) return;
    if (w.x < M) {
        membrane_body<MAXN>(mp[w.x], w.y);
    } else {
        const CellsSortPack p = sp[w.x - M];
        cells_sort_body(p.ncells, p.nclear, p.cnt, p.list, p.other, p.x, p.y, p.vx, p.vy, p.owner,
                        p.pay, p.own, w.y);
    }
}

__global__ __launch_bounds__(256) void kb_contact(const ContactPack* packs, const int2* map) {
    const int2 w = map[blockIdx.x];
    if (w.x < 0) return;
    const ContactPack p = packs[w.x];
    if (p.allpairs) repulsion_body<false>(p.a, w.y);
    else repulsion_lanes_body<true>(p.a, w.y);
}

void batch_items_membrane(Ctx& c, BatchItem* it) {
    batch_put(it[0], c.ncaps, mem_args(c));
}

// The same arguments and cell-count buffer alternation as launch_repulsion
// (x-periodic channels: the cell grid is fixed).
void batch_items_contact(Ctx& c, BatchItem* it) {
    RepArgs a = rep_args(c);
    if (c.K <= REP_ALLPAIRS_MAX) {
        batch_put(it[0], 0, CellsFillPack{});
        batch_put(it[1], 0, CellsSortPack{});
        batch_put(it[2], (c.K + BATCH_ALLPAIRS_THREADS - 1) / BATCH_ALLPAIRS_THREADS, ContactPack{a, 1});
        return;
    }
    a.ncy = c.ncy; a.cdy = c.cdy;
    a.ncx = c.ncx; a.xorg = 0.0; a.cdx = c.cdx;
    const int nmax = c.ncx * c.ncy;
    int* cnt = c.cell_cnt + static_cast<size_t>(c.cell_cur) * nmax;
    int* other = c.cell_cnt + static_cast<size_t>(1 - c.cell_cur) * nmax;
    batch_put(it[0], (c.K + BATCH_FILL_THREADS - 1) / BATCH_FILL_THREADS,
              CellsFillPack{a, cnt, c.cell_list, c.counters});
    batch_put(it[1], (nmax + 63) / 64,           // 64-thread workgroups, beside the membrane ones
              CellsSortPack{a.ncx * a.ncy, nmax, cnt, c.cell_list, other, c.x, c.y, c.vx, c.vy,
                            c.owner, c.cell_pay, c.cell_own});
    c.cell_cur = 1 - c.cell_cur;
    a.cnt = cnt;
    a.pay = c.cell_pay;
    a.own = c.cell_own;
    batch_put(it[2], (c.K + REP_NODES - 1) / REP_NODES, ContactPack{a, 0});
}

void batch_launch_membrane_sort(const BatchVariant& v, const void* mem_packs, const void* sort_packs,
                                const int2* map, int nblocks, int M, hipStream_t s) {
    if (nblocks <= 0) return;
    const MemArgs* mp = static_cast<const MemArgs*>(mem_packs);
    const CellsSortPack* sp = static_cast<const CellsSortPack*>(sort_packs);
    if (v.maxn <= 64) hipLaunchKernelGGL(kb_membrane_sort<64>, dim3(nblocks), dim3(64), 0, s, mp, sp, map, M);
    else hipLaunchKernelGGL(kb_membrane_sort<256>, dim3(nblocks), dim3(64), 0, s, mp, sp, map, M);
}

void batch_launch_membrane(int kind, const BatchVariant& v, const void* packs, const int2* map,
                           int nblocks, hipStream_t s) {
    if (nblocks <= 0) return;
    const dim3 g(nblocks);
    switch (kind) {
    case BK_CELLS_FILL:
        hipLaunchKernelGGL(kb_cells_fill, g, dim3(BATCH_FILL_THREADS), 0, s,
                           static_cast<const CellsFillPack*>(packs), map);
        break;
    case BK_CONTACT:
        hipLaunchKernelGGL(kb_contact, g, dim3(256), 0, s, static_cast<const ContactPack*>(packs), map);
        break;
    default:
        break;
    }
}

}  // namespace sf
