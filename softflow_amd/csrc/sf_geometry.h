// Obstacle geometry shared by the host (cell flags, interpolated bounce-back
// links, capsule placement) and the device (obstacle contact forces).
//
// The three shapes of geometry/{circle,rect,polygon}_obstacle.h, with their
// contains / signedDistance / nearestPoint / normalAt evaluated in the same
// IEEE double operations as the reference (no FMA contraction: the library
// and the engine are compiled with -ffp-contract=off), so that cell flags,
// link distances and contact forces come out bit for bit.
#pragma once
#include "sf_math.h"

#include <cmath>

namespace sf {

enum ObstacleKind { OBS_CIRCLE = 0, OBS_RECT = 1, OBS_POLYGON = 2 };

// One obstacle.  Circle: p = (cx, cy, r).  Rectangle: p = (x0, y0, x1, y1)
// with x0 <= x1 and y0 <= y1 (the constructor's min/max, rect_obstacle.h:10-12).
// Polygon: nv vertices at v[0..nv) of the vertex array the caller passes.
struct Obstacle {
    int kind, nv;
    double p[4];
    const double* v;   // polygon vertices, (x, y) pairs
    double bc[3];      // a disc (x, y, radius) holding the whole obstacle
};

struct ONearest { double x, y; };

SF_HD inline double omin(double a, double b) { return b < a ? b : a; }   // std::min
SF_HD inline double omax(double a, double b) { return a < b ? b : a; }   // std::max

// A node farther than r_cut from the disc that holds the obstacle is farther
// than r_cut from the obstacle: its contact term is zero and is skipped (the
// margin covers the rounding of the exact distance).
SF_HD inline bool obstacle_far(const Obstacle& o, double px, double py, double rcut) {
    const double dx = px - o.bc[0], dy = py - o.bc[1];
    const double lim = o.bc[2] + rcut;
    return dx * dx + dy * dy > lim * lim * (1.0 + 1e-9);
}

// The holding disc of an obstacle, a little larger than the tightest one.
inline void obstacle_bound(Obstacle& o, const double* verts) {
    double cx, cy, r;
    if (o.kind == OBS_CIRCLE) {
        cx = o.p[0]; cy = o.p[1]; r = o.p[2];
    } else if (o.kind == OBS_RECT) {
        cx = 0.5 * (o.p[0] + o.p[2]); cy = 0.5 * (o.p[1] + o.p[3]);
        r = 0.5 * std::sqrt((o.p[2] - o.p[0]) * (o.p[2] - o.p[0]) + (o.p[3] - o.p[1]) * (o.p[3] - o.p[1]));
    } else {
        double x0 = verts[0], x1 = verts[0], y0 = verts[1], y1 = verts[1];
        for (int i = 1; i < o.nv; ++i) {
            x0 = omin(x0, verts[2 * i]); x1 = omax(x1, verts[2 * i]);
            y0 = omin(y0, verts[2 * i + 1]); y1 = omax(y1, verts[2 * i + 1]);
        }
        cx = 0.5 * (x0 + x1); cy = 0.5 * (y0 + y1);
        r = 0.0;
        for (int i = 0; i < o.nv; ++i) {
            const double dx = verts[2 * i] - cx, dy = verts[2 * i + 1] - cy;
            r = omax(r, std::sqrt(dx * dx + dy * dy));
        }
    }
    o.bc[0] = cx; o.bc[1] = cy;
    o.bc[2] = std::fabs(r) * (1.0 + 1e-12) + 1e-12;
}

// circle_obstacle.h:11-38
SF_HD inline bool circle_contains(const Obstacle& o, double x, double y) {
    const double dx = x - o.p[0], dy = y - o.p[1];
    return (dx * dx + dy * dy) <= o.p[2] * o.p[2];
}

// rect_obstacle.h:14-62
SF_HD inline bool rect_contains(const Obstacle& o, double x, double y) {
    return x >= o.p[0] && x <= o.p[2] && y >= o.p[1] && y <= o.p[3];
}

SF_HD inline ONearest rect_nearest(const Obstacle& o, double x, double y) {
    const double x0 = o.p[0], y0 = o.p[1], x1 = o.p[2], y1 = o.p[3];
    // std::clamp(v, lo, hi): lo if v < lo, hi if hi < v, else v
    const double cx = x < x0 ? x0 : (x1 < x ? x1 : x);
    const double cy = y < y0 ? y0 : (y1 < y ? y1 : y);
    if (cx == x && cy == y) {
        const double dl = x - x0, dr = x1 - x, db = y - y0, dt = y1 - y;
        const double dmin = omin(omin(omin(dl, dr), db), dt);
        if (dmin == dl) return {x0, y};
        if (dmin == dr) return {x1, y};
        if (dmin == db) return {x, y0};
        return {x, y1};
    }
    return {cx, cy};
}

// polygon_obstacle.cpp:10-24: ray casting
SF_HD inline bool poly_contains(const Obstacle& o, double x, double y) {
    bool inside = false;
    const int n = o.nv;
    for (int i = 0, j = n - 1; i < n; j = i++) {
        const double xi = o.v[2 * i], yi = o.v[2 * i + 1];
        const double xj = o.v[2 * j], yj = o.v[2 * j + 1];
        if (((yi > y) != (yj > y)) && (x < (xj - xi) * (y - yi) / (yj - yi) + xi)) inside = !inside;
    }
    return inside;
}

// polygon_obstacle.cpp:26-54: the nearest point over the edges, first minimum kept
SF_HD inline ONearest poly_nearest(const Obstacle& o, double x, double y) {
    double best = 1.7976931348623157e308;
    ONearest c{0.0, 0.0};
    const int n = o.nv;
    for (int i = 0; i < n; ++i) {
        const int j = (i + 1) % n;
        const double v0x = o.v[2 * i], v0y = o.v[2 * i + 1];
        const double ex = o.v[2 * j] - v0x, ey = o.v[2 * j + 1] - v0y;
        const double len2 = ex * ex + ey * ey;
        ONearest q{v0x, v0y};
        if (!(len2 < 1e-30)) {
            double t = ((x - v0x) * ex + (y - v0y) * ey) / len2;
            t = omax(0.0, omin(1.0, t));
            q = {v0x + t * ex, v0y + t * ey};
        }
        const double dx = q.x - x, dy = q.y - y;
        const double d2 = dx * dx + dy * dy;
        if (d2 < best) { best = d2; c = q; }
    }
    return c;
}

SF_HD inline bool obstacle_contains(const Obstacle& o, double x, double y) {
    if (o.kind == OBS_CIRCLE) return circle_contains(o, x, y);
    if (o.kind == OBS_RECT) return rect_contains(o, x, y);
    return poly_contains(o, x, y);
}

SF_HD inline double obstacle_signed_distance(const Obstacle& o, double x, double y) {
    if (o.kind == OBS_CIRCLE) {
        const double dx = x - o.p[0], dy = y - o.p[1];
        return std::sqrt(dx * dx + dy * dy) - o.p[2];
    }
    if (o.kind == OBS_RECT) {
        const double x0 = o.p[0], y0 = o.p[1], x1 = o.p[2], y1 = o.p[3];
        const double dx = omax(omax(x0 - x, x - x1), 0.0);
        const double dy = omax(omax(y0 - y, y - y1), 0.0);
        const double outside = std::sqrt(dx * dx + dy * dy);
        if (outside > 1e-15) return outside;
        const double dmin = omin(omin(omin(x - x0, x1 - x), y - y0), y1 - y);
        return -dmin;
    }
    const ONearest q = poly_nearest(o, x, y);
    const double dist = std::sqrt((q.x - x) * (q.x - x) + (q.y - y) * (q.y - y));
    return poly_contains(o, x, y) ? -dist : dist;
}

SF_HD inline ONearest obstacle_nearest_point(const Obstacle& o, double x, double y) {
    if (o.kind == OBS_CIRCLE) {
        const double dx = x - o.p[0], dy = y - o.p[1];
        const double d = std::sqrt(dx * dx + dy * dy);
        if (d < 1e-15) return {o.p[0] + o.p[2], o.p[1]};
        return {o.p[0] + o.p[2] * dx / d, o.p[1] + o.p[2] * dy / d};
    }
    if (o.kind == OBS_RECT) return rect_nearest(o, x, y);
    return poly_nearest(o, x, y);
}

SF_HD inline ONearest obstacle_normal(const Obstacle& o, double x, double y) {
    if (o.kind == OBS_CIRCLE) {
        const double dx = x - o.p[0], dy = y - o.p[1];
        const double d = std::sqrt(dx * dx + dy * dy);
        if (d < 1e-15) return {1.0, 0.0};
        return {dx / d, dy / d};
    }
    if (o.kind == OBS_RECT) {
        const ONearest q = rect_nearest(o, x, y);
        const double dx = x - q.x, dy = y - q.y;
        const double d = std::sqrt(dx * dx + dy * dy);
        if (d < 1e-15) {
            const double dl = std::fabs(x - o.p[0]), dr = std::fabs(x - o.p[2]);
            const double db = std::fabs(y - o.p[1]), dt = std::fabs(y - o.p[3]);
            const double dmin = omin(omin(omin(dl, dr), db), dt);
            if (dmin == dl) return {-1.0, 0.0};
            if (dmin == dr) return {1.0, 0.0};
            if (dmin == db) return {0.0, -1.0};
            return {0.0, 1.0};
        }
        return {dx / d, dy / d};
    }
    const ONearest q = poly_nearest(o, x, y);
    const double dx = x - q.x, dy = y - q.y;
    const double len = std::sqrt(dx * dx + dy * dy);
    if (len < 1e-15) return {0.0, 1.0};
    ONearest nrm{dx / len, dy / len};
    if (poly_contains(o, x, y)) { nrm.x = -nrm.x; nrm.y = -nrm.y; }
    return nrm;
}

// RepulsionForce::computeObstacleRepulsion for one node and one obstacle
// (membrane/repulsion.cpp:180-236): adds the obstacle's contact force to
// (fx, fy) in the reference's operation order.
SF_HD inline void obstacle_contact(const Obstacle& o, double px, double py, double vx, double vy,
                                   double eps, double sig, int pw, double rcut, double gn, double mu,
                                   double& fx, double& fy) {
    const double sd = obstacle_signed_distance(o, px, py);
    if (sd >= rcut) return;
    double r = std::fabs(sd);
    if (r < 1e-15) r = 1e-15;
    const ONearest nrm = obstacle_normal(o, px, py);
    const double Frep = eps * pow_int_cr(sig / r, pw);
    if (sd < 0.0) {
        const double s = Frep * 10.0;
        fx += nrm.x * s;
        fy += nrm.y * s;
        return;
    }
    double Fn = Frep, tx = 0.0, ty = 0.0;
    if (gn > 0.0 || mu > 0.0) {
        const double vn = vx * nrm.x + vy * nrm.y;
        Fn = Frep - gn * vn;
        if (Fn < 0.0) Fn = 0.0;
        if (mu > 0.0) {
            const double vtx = vx - nrm.x * vn, vty = vy - nrm.y * vn;
            const double vtm = std::sqrt(vtx * vtx + vty * vty);
            if (vtm > 1e-12) {
                const double s = -mu * Fn;
                tx = vtx / vtm * s;
                ty = vty / vtm * s;
            }
        }
    }
    fx += nrm.x * Fn + tx;
    fy += nrm.y * Fn + ty;
}

}  // namespace sf
