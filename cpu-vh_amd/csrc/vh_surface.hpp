// Freezeout surface elements of one lattice cell, for the host C ABI
// (vh_surface.cpp) and the device's element kernels (vh_kernels.hip): one
// source, so both build every element bit for bit alike (compiled with
// -ffp-contract=off; sqrt and division correctly rounded on both).
//
// The constant-e surface inside a square (2D), a cube (3D, the (tau, x, y)
// cell of a 2+1D run) or a hypercube (4D, the (tau, x, y, eta) cell of a
// 3+1D run) as the surface finder the reference links, Cornelius 1.3
// (P. Huovinen and H. Petersen, Eur. Phys. J. A 48 (2012) 171;
// freezeout/cornelius-c++-1.3/cornelius.cpp), builds it, so that the
// elements — centroid and normal vector — are the ones the reference writes
// to surface.dat:
//   * on each square the crossing points of the edges, taken in the order
//     (edge u at v=0, edge v at u=0, edge v at u=du, edge u at v=dv), become
//     segments pairwise; a corner exactly at the level counts as a crossing
//     1e-9 of the edge from it when its neighbour is below; with four
//     crossings the mean of the corners decides the pairing and the low side
//     (cornelius.cpp:1177-1337);
//   * each segment carries a point on the low side (the mean of the corners
//     below the level, or the ambiguous case's choice); a segment's normal is
//     (-dv, du) turned towards that point (:259-296, :91-106);
//   * a cube collects the segments of its six faces (faces at fixed tau,
//     then x, then y, each at 0 then at the cell's width; :1502-1541); when
//     no face is ambiguous and there are not exactly six segments they form
//     one polygon, otherwise they are chained end to start into polygons
//     (:1552-1646);
//   * a polygon's centroid is the area-weighted mean of the triangles
//     (segment, mean of the corners), or the mean itself for a triangle; its
//     normal is the sum of the triangles' (segment, centroid) vector areas,
//     each turned towards its segment's low-side point (:478-624);
//   * a hypercube collects the polygons of its eight cubes; they form one
//     polyhedron unless a face was ambiguous (or 24 segments from two corners
//     on one side), else polygons join the polyhedron one of whose segments
//     starts where one of theirs starts or ends; the polyhedron's centroid
//     and normal come from the tetrahedra (segment, polygon centroid, mean)
//     (:1759-1967, :652-873).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace vh {
namespace surf {

#define VH_HD __host__ __device__ inline

// A list of at most N elements held in place: no allocation per cell.  The
// capacities used below are the most one cell can produce (two segments per
// square, six faces per cube, polygons of at least three segments, eight
// cubes per hypercube), so push_back never meets a full list.
template <class T, int N>
struct Small {
    T v[N];
    int n = 0;
    VH_HD int size() const { return n; }
    VH_HD bool empty() const { return n == 0; }
    VH_HD void clear() { n = 0; }
    VH_HD void push_back(const T &x) {
        if (n < N) v[n++] = x;
    }
    VH_HD T &emplace_back() {
        v[n < N ? n : N - 1] = T();
        if (n < N) ++n;
        return v[n - 1];
    }
    VH_HD T &back() { return v[n - 1]; }
    VH_HD const T &back() const { return v[n - 1]; }
    VH_HD T &operator[](int i) { return v[i]; }
    VH_HD const T &operator[](int i) const { return v[i]; }
};

// Points live in the four coordinates (tau, x, y, eta) of the largest cell;
// a square or cube of a lower-dimensional cell is embedded with its unused
// leading coordinates fixed at 0, as the reference embeds them.
struct Seg {
    double a[4], b[4];  // end points ("start", "end")
    double out[4];      // a point on the low side
};
constexpr int MAX_SEGS = 12;  // per cube: six faces, two segments each
using SegList = Small<Seg, MAX_SEGS>;
using Poly = Small<int, MAX_SEGS>;            // segment indices, in chain order
using PolyList = Small<Poly, MAX_SEGS / 3>;  // per cube
constexpr int MAX_POLYS = 8 * (MAX_SEGS / 3);  // per hypercube

// element construction results
enum { SURF_OK = 0, SURF_STATE = 1, SURF_CAPACITY = 2 };

VH_HD void swap_d(double &a, double &b) {
    const double t = a;
    a = b;
    b = t;
}

// Contour segments of one square.  p[iu][iv] are the corner values at
// (iu du, iv dv); put(u, v, dst) places a square-local point.  Returns
// whether the square is ambiguous (four crossings).
template <class PUT>
VH_HD bool square_segments(const double p[2][2], double E0, double du, double dv, SegList &segs, PUT put) {
    int above = 0;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            if (p[i][j] >= E0) ++above;
    if (above == 0 || above == 4) return false;
    double cut[4][2];
    int nc = 0;
    // one edge from corner value pa (parameter 0) to pb (parameter L)
    auto edge = [&](double pa, double pb, double L, double fixed, bool along_u) {
        double tpar;
        if ((pa - E0) * (pb - E0) < 0) tpar = (pa - E0) / (pa - pb) * L;
        else if (pa == E0 && pb < E0) tpar = 1e-9 * L;
        else if (pb == E0 && pa < E0) tpar = (1.0 - 1e-9) * L;
        else return;
        cut[nc][0] = along_u ? tpar : fixed;
        cut[nc][1] = along_u ? fixed : tpar;
        ++nc;
    };
    edge(p[0][0], p[1][0], du, 0.0, true);   // v = 0
    edge(p[0][0], p[0][1], dv, 0.0, false);  // u = 0
    edge(p[1][0], p[1][1], dv, du, false);   // u = du
    edge(p[0][1], p[1][1], du, dv, true);    // v = dv
    if (nc != 2 && nc != 4) return false;    // cannot happen for finite corner values
    double out[2][2] = {{0, 0}, {0, 0}};
    bool ambiguous = false;
    if (nc == 4) {
        ambiguous = true;
        double mid = 0;
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) mid += 0.25 * p[i][j];
        // default pairing (0,1),(2,3); when the centre and corner (0,0) lie on
        // the same side the pairing is (0,2),(1,3)
        if ((p[0][0] < E0 && mid < E0) || (p[0][0] > E0 && mid > E0))
            for (int k = 0; k < 2; ++k) swap_d(cut[1][k], cut[2][k]);
        if (mid - E0 < 0) {
            out[0][0] = out[1][0] = 0.5 * du;
            out[0][1] = out[1][1] = 0.5 * dv;
        } else if (p[0][0] - E0 < 0) {
            out[0][0] = 0; out[0][1] = 0;
            out[1][0] = du; out[1][1] = dv;
        } else {
            out[0][0] = du; out[0][1] = 0;
            out[1][0] = 0; out[1][1] = dv;
        }
    } else {
        int nout = 0;
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                if (p[i][j] < E0) {
                    out[0][0] += i * du;
                    out[0][1] += j * dv;
                    ++nout;
                }
        if (nout > 0) {
            out[0][0] = out[0][0] / double(nout);
            out[0][1] = out[0][1] / double(nout);
        }
    }
    for (int k = 0; k < nc; k += 2) {
        Seg s;
        put(cut[k][0], cut[k][1], s.a);
        put(cut[k + 1][0], cut[k + 1][1], s.b);
        put(out[k / 2][0], out[k / 2][1], s.out);
        segs.push_back(s);
    }
    return ambiguous;
}

// Normal of a segment in its square's plane (free coordinates iu < iv):
// (-(b_v - a_v), b_u - a_u), turned towards the low side.
VH_HD void segment_element(const Seg &s, int iu, int iv, double *centroid, double *normal) {
    for (int k = 0; k < 4; ++k) {
        centroid[k] = (s.a[k] + s.b[k]) / 2.0;
        normal[k] = 0;
    }
    normal[iu] = -(s.b[iv] - s.a[iv]);
    normal[iv] = s.b[iu] - s.a[iu];
    double dot = 0;
    for (int k = 0; k < 4; ++k) dot += (s.out[k] - centroid[k]) * normal[k];
    if (dot < 0)
        for (int k = 0; k < 4; ++k) normal[k] = -normal[k];
}

VH_HD bool same_point(const double *p, const double *q) {
    double d = 0;
    for (int k = 0; k < 4; ++k) d += fabs(p[k] - q[k]);
    return d < 1e-10;
}

// The polygons of one cube: coordinate cf fixed at cv, free coordinates
// f[0] < f[1] < f[2], c[i][j][k] the value at (i, j, k) along them.
struct CubeSurface {
    int f[3];
    SegList segs;
    PolyList polys;
    bool ambiguous = false;
};

VH_HD void cube_surface(const double c[2][2][2], double E0, const double *d, int cf, double cv, CubeSurface &cs) {
    for (int k = 0, n = 0; k < 4; ++k)
        if (k != cf) cs.f[n++] = k;
    // faces: fixed free coordinate f[a] at 0 then at its width, a = 0, 1, 2
    for (int a = 0; a < 3; ++a) {
        const int u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
        const int iu = cs.f[u], iv = cs.f[v], ia = cs.f[a];
        for (int j = 0; j < 2; ++j) {
            double p[2][2];
            for (int i1 = 0; i1 < 2; ++i1)
                for (int i2 = 0; i2 < 2; ++i2)
                    p[i1][i2] = a == 0 ? c[j][i1][i2] : (a == 1 ? c[i1][j][i2] : c[i1][i2][j]);
            const double fixed = j * d[ia];
            cs.ambiguous |= square_segments(p, E0, d[iu], d[iv], cs.segs, [&](double x, double y, double *dst) {
                dst[cf] = cv;
                dst[ia] = fixed;
                dst[iu] = x;
                dst[iv] = y;
            });
        }
    }
    const int n = cs.segs.size();
    if (n == 0) return;
    if (!cs.ambiguous && n != 6) {
        cs.polys.emplace_back();
        for (int i = 0; i < n; ++i) cs.polys.back().push_back(i);
        return;
    }
    // chain segments end to start; a segment joined by its end is flipped
    cs.ambiguous = true;
    char used[MAX_SEGS] = {};
    int nused = 0;
    while (nused < n) {
        if (n - nused < 3) {
            cs.polys.clear();
            return;
        }
        cs.polys.emplace_back();
        Poly &poly = cs.polys.back();
        for (int i = 0; i < n; ++i) {
            if (used[i]) continue;
            Seg &s = cs.segs[i];
            bool take = poly.empty();
            if (!take) {
                const double *last = cs.segs[poly.back()].b;
                if (same_point(s.a, last)) {
                    take = true;
                } else if (same_point(s.b, last)) {
                    for (int k = 0; k < 4; ++k) swap_d(s.a[k], s.b[k]);
                    take = true;
                }
            }
            if (take) {
                poly.push_back(i);
                used[i] = 1;
                ++nused;
                i = 0;  // rescan from the second segment, as the reference's loop does
            }
        }
    }
}

// Centroid of one polygon: the mean of its corners for a triangle, else the
// area-weighted mean of the triangles (segment, mean).
VH_HD void polygon_centroid(const CubeSurface &cs, const Poly &poly, double *centroid) {
    double mean[4] = {0, 0, 0, 0};
    for (int q = 0; q < poly.size(); ++q) {
        const Seg &s = cs.segs[poly[q]];
        for (int k = 0; k < 4; ++k) mean[k] += s.a[k] + s.b[k];
    }
    for (int k = 0; k < 4; ++k) mean[k] = mean[k] / double(2.0 * poly.size());
    if (poly.size() == 3) {
        for (int k = 0; k < 4; ++k) centroid[k] = mean[k];
        return;
    }
    const int x1 = cs.f[0], x2 = cs.f[1], x3 = cs.f[2];
    double up[4] = {0, 0, 0, 0}, down = 0;
    for (int q = 0; q < poly.size(); ++q) {
        const Seg &s = cs.segs[poly[q]];
        double cm[4], a[4], b[4];
        for (int k = 0; k < 4; ++k) {
            cm[k] = (s.a[k] + s.b[k] + mean[k]) / 3.0;
            a[k] = s.a[k] - mean[k];
            b[k] = s.b[k] - mean[k];
        }
        const double c0 = a[x2] * b[x3] - a[x3] * b[x2], c1 = a[x1] * b[x3] - a[x3] * b[x1],
                     c2 = a[x2] * b[x1] - a[x1] * b[x2];
        const double area = 0.5 * sqrt(c0 * c0 + c1 * c1 + c2 * c2);
        for (int k = 0; k < 4; ++k) up[k] += cm[k] * area;
        down += area;
    }
    for (int k = 0; k < 4; ++k) centroid[k] = up[k] / down;
}

// Normal of one polygon: the sum of the (segment, centroid) triangles' vector
// areas, each turned towards its segment's low-side point.
VH_HD void polygon_normal(const CubeSurface &cs, const Poly &poly, const double *centroid, double *normal) {
    const int x1 = cs.f[0], x2 = cs.f[1], x3 = cs.f[2];
    for (int k = 0; k < 4; ++k) normal[k] = 0;
    for (int q = 0; q < poly.size(); ++q) {
        const Seg &s = cs.segs[poly[q]];
        double a[4], b[4], n[4] = {0, 0, 0, 0};
        for (int k = 0; k < 4; ++k) {
            a[k] = s.a[k] - centroid[k];
            b[k] = s.b[k] - centroid[k];
        }
        n[x1] = 0.5 * (a[x2] * b[x3] - a[x3] * b[x2]);
        n[x2] = -0.5 * (a[x1] * b[x3] - a[x3] * b[x1]);
        n[x3] = 0.5 * (a[x1] * b[x2] - a[x2] * b[x1]);
        double dot = 0;
        for (int k = 0; k < 4; ++k) dot += (s.out[k] - centroid[k]) * n[k];
        if (dot < 0)
            for (int k = 0; k < 4; ++k) n[k] = -n[k];
        for (int k = 0; k < 4; ++k) normal[k] += n[k];
    }
}

// The 4D vector orthogonal to a, b, c whose norm is the volume of the
// tetrahedron they span: n_i = (-1)^i det(a, b, c without column i) / 6.
VH_HD void tetra_normal(const double *a, const double *b, const double *c, double *n) {
    const double bc01 = b[0] * c[1] - b[1] * c[0], bc02 = b[0] * c[2] - b[2] * c[0], bc03 = b[0] * c[3] - b[3] * c[0];
    const double bc12 = b[1] * c[2] - b[2] * c[1], bc13 = b[1] * c[3] - b[3] * c[1], bc23 = b[2] * c[3] - b[3] * c[2];
    n[0] = 1.0 / 6.0 * (a[1] * bc23 - a[2] * bc13 + a[3] * bc12);
    n[1] = -1.0 / 6.0 * (a[0] * bc23 - a[2] * bc03 + a[3] * bc02);
    n[2] = 1.0 / 6.0 * (a[0] * bc13 - a[1] * bc03 + a[3] * bc01);
    n[3] = -1.0 / 6.0 * (a[0] * bc12 - a[1] * bc02 + a[2] * bc01);
}

// Elements of one square: corners[iu][iv] as p[2][2] row-major, d = (du, dv);
// 4 doubles per element (centroid u v, normal u v).
VH_HD int square_elements(const double *corners, double E0, const double *d, double *elements, int cap, int *n_el) {
    const double p[2][2] = {{corners[0], corners[1]}, {corners[2], corners[3]}};
    SegList segs;
    square_segments(p, E0, d[0], d[1], segs, [](double u, double v, double *dst) {
        dst[0] = 0;
        dst[1] = 0;
        dst[2] = u;
        dst[3] = v;
    });
    *n_el = segs.size();
    if (segs.size() > cap || (!elements && !segs.empty())) return SURF_CAPACITY;
    for (int e = 0; e < segs.size(); ++e) {
        double c[4], n[4];
        segment_element(segs[e], 2, 3, c, n);
        for (int k = 0; k < 2; ++k) {
            elements[4 * e + k] = c[2 + k];
            elements[4 * e + 2 + k] = n[2 + k];
        }
    }
    return SURF_OK;
}

// Elements of one cube: corners[4 i + 2 j + k] at (i, j, k) along (tau, x, y),
// d = (dtau, dx, dy); 6 doubles per element (centroid, normal).  With
// elements NULL and cap 0 only *n_el is set (SURF_OK).
VH_HD int cube_elements(const double *corners, double E0, const double *d, double *elements, int cap, int *n_el) {
    double c[2][2][2];
    int above = 0;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                c[i][j][k] = corners[4 * i + 2 * j + k];
                if (c[i][j][k] >= E0) ++above;
            }
    *n_el = 0;
    if (above == 0 || above == 8) return SURF_OK;
    const double d4[4] = {1, d[0], d[1], d[2]};
    CubeSurface cs;
    cube_surface(c, E0, d4, 0, 0.0, cs);
    if (cs.polys.empty()) return cs.segs.empty() ? SURF_OK : SURF_STATE;
    *n_el = cs.polys.size();
    if (!elements && cap == 0) return SURF_OK;
    if (cs.polys.size() > cap || !elements) return SURF_CAPACITY;
    for (int e = 0; e < cs.polys.size(); ++e) {
        double cen[4], nor[4];
        polygon_centroid(cs, cs.polys[e], cen);
        polygon_normal(cs, cs.polys[e], cen, nor);
        for (int k = 0; k < 3; ++k) {
            elements[6 * e + k] = cen[1 + k];
            elements[6 * e + 3 + k] = nor[1 + k];
        }
    }
    return SURF_OK;
}

// Elements of one hypercube: corners[8 i + 4 j + 2 k + l] at (i, j, k, l)
// along (tau, x, y, eta), d = (dtau, dx, dy, deta); 8 doubles per element
// (centroid, normal).  With elements NULL and cap 0 only *n_el is set.
VH_HD int hypercube_elements(const double *corners, double E0, const double *d, double *elements, int cap,
                             int *n_el) {
    double h[2][2][2][2];
    int above = 0, below = 0;
    for (int i = 0; i < 16; ++i) {
        h[i >> 3][(i >> 2) & 1][(i >> 1) & 1][i & 1] = corners[i];
        if (corners[i] >= E0) ++above;
        if (corners[i] < E0) ++below;
    }
    *n_el = 0;
    if (above == 0 || above == 16) return SURF_OK;
    // the eight cubes: coordinate i fixed at 0 then at its width
    CubeSurface cubes[8];
    bool ambiguous = false;
    int nlines = 0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 2; ++j) {
            double cu[2][2][2];
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b)
                    for (int e = 0; e < 2; ++e)
                        cu[a][b][e] = i == 0 ? h[j][a][b][e]
                                             : (i == 1 ? h[a][j][b][e] : (i == 2 ? h[a][b][j][e] : h[a][b][e][j]));
            CubeSurface &cs = cubes[2 * i + j];
            cube_surface(cu, E0, d, i, j * d[i], cs);
            if (cs.polys.empty() && !cs.segs.empty()) return SURF_STATE;
            ambiguous |= cs.ambiguous;
            nlines += cs.segs.size();
        }
    // with no ambiguous face, two corners on one side and 24 segments can
    // still be two separate pieces (the reference's one extra test)
    if (!ambiguous) {
        const int pts = below > 8 ? 16 - below : below;
        if (nlines == 24 && pts == 2) ambiguous = true;
    }
    // polygons in cube order: (cube, polygon) and the polygon's centroid
    int pcube[MAX_POLYS], ppoly[MAX_POLYS];
    int np = 0;
    for (int q = 0; q < 8; ++q)
        for (int k = 0; k < cubes[q].polys.size(); ++k) {
            pcube[np] = q;
            ppoly[np] = k;
            ++np;
        }
    // polyhedra: hstart[e] .. hstart[e+1] index horder, the polygons in the
    // order they joined
    int horder[MAX_POLYS], hstart[MAX_POLYS + 1];
    int nh = 0;
    if (!ambiguous) {
        for (int i = 0; i < np; ++i) horder[i] = i;
        hstart[0] = 0;
        hstart[1] = np;
        nh = 1;
    } else {
        // a polygon joins a polyhedron when one of its segments starts or ends
        // where a segment of a polygon already in it starts
        auto touches = [&](int p, int q) {
            const CubeSurface &cp = cubes[pcube[p]], &cq = cubes[pcube[q]];
            const Poly &sp = cp.polys[ppoly[p]], &sq = cq.polys[ppoly[q]];
            for (int j = 0; j < sp.size(); ++j)
                for (int k = 0; k < sq.size(); ++k) {
                    const Seg &l1 = cp.segs[sp[j]], &l2 = cq.segs[sq[k]];
                    if (same_point(l1.a, l2.a) || same_point(l1.b, l2.a)) return true;
                }
            return false;
        };
        char used[MAX_POLYS] = {};
        int nused = 0, pos = 0;
        while (nused < np) {
            hstart[nh++] = pos;
            const int first = pos;
            for (int i = 0; i < np; ++i) {
                if (used[i]) continue;
                bool take = pos == first;
                for (int k = first; !take && k < pos; ++k) take = touches(i, horder[k]);
                if (take) {
                    horder[pos++] = i;
                    used[i] = 1;
                    ++nused;
                    i = 0;  // rescan from the second polygon, as the reference's loop does
                }
            }
        }
        hstart[nh] = pos;
    }
    *n_el = nh;
    if (!elements && cap == 0) return SURF_OK;
    if (nh > cap || !elements) return SURF_CAPACITY;
    double pcen[MAX_POLYS][4];
    for (int i = 0; i < np; ++i) polygon_centroid(cubes[pcube[i]], cubes[pcube[i]].polys[ppoly[i]], pcen[i]);
    for (int e = 0; e < nh; ++e) {
        // centroid: volume-weighted mean of the tetrahedra (segment, polygon
        // centroid, mean of all corners)
        int ntet = 0;
        double mean[4] = {0, 0, 0, 0};
        for (int r = hstart[e]; r < hstart[e + 1]; ++r) {
            const int pi = horder[r];
            const CubeSurface &cs = cubes[pcube[pi]];
            const Poly &poly = cs.polys[ppoly[pi]];
            for (int q = 0; q < poly.size(); ++q) {
                const Seg &s = cs.segs[poly[q]];
                for (int k = 0; k < 4; ++k) mean[k] += s.a[k] + s.b[k];
                ++ntet;
            }
        }
        for (int k = 0; k < 4; ++k) mean[k] = mean[k] / double(2.0 * ntet);
        double up[4] = {0, 0, 0, 0}, down = 0;
        for (int r = hstart[e]; r < hstart[e + 1]; ++r) {
            const int pi = horder[r];
            const CubeSurface &cs = cubes[pcube[pi]];
            const Poly &poly = cs.polys[ppoly[pi]];
            const double *cent = pcen[pi];
            for (int q = 0; q < poly.size(); ++q) {
                const Seg &s = cs.segs[poly[q]];
                double cm[4], a[4], b[4], c[4], n[4];
                for (int k = 0; k < 4; ++k) {
                    cm[k] = (s.a[k] + s.b[k] + cent[k] + mean[k]) / 4.0;
                    a[k] = s.a[k] - mean[k];
                    b[k] = s.b[k] - mean[k];
                    c[k] = cent[k] - mean[k];
                }
                tetra_normal(a, b, c, n);
                double vol = 0;
                for (int k = 0; k < 4; ++k) vol += n[k] * n[k];
                vol = sqrt(vol);
                for (int k = 0; k < 4; ++k) up[k] += cm[k] * vol;
                down += vol;
            }
        }
        double cen[4], nor[4] = {0, 0, 0, 0};
        for (int k = 0; k < 4; ++k) cen[k] = up[k] / down;
        // normal: the tetrahedra (segment, polygon centroid, centroid), each
        // turned towards its segment's low-side point
        for (int r = hstart[e]; r < hstart[e + 1]; ++r) {
            const int pi = horder[r];
            const CubeSurface &cs = cubes[pcube[pi]];
            const Poly &poly = cs.polys[ppoly[pi]];
            const double *cent = pcen[pi];
            for (int q = 0; q < poly.size(); ++q) {
                const Seg &s = cs.segs[poly[q]];
                double a[4], b[4], c[4], n[4];
                for (int k = 0; k < 4; ++k) {
                    a[k] = s.a[k] - cen[k];
                    b[k] = s.b[k] - cen[k];
                    c[k] = cent[k] - cen[k];
                }
                tetra_normal(a, b, c, n);
                double dot = 0;
                for (int k = 0; k < 4; ++k) dot += (s.out[k] - cen[k]) * n[k];
                if (dot < 0)
                    for (int k = 0; k < 4; ++k) n[k] = -n[k];
                for (int k = 0; k < 4; ++k) nor[k] += n[k];
            }
        }
        for (int k = 0; k < 4; ++k) {
            elements[8 * e + k] = cen[k];
            elements[8 * e + 4 + k] = nor[k];
        }
    }
    return SURF_OK;
}

#undef VH_HD

}  // namespace surf
}  // namespace vh
