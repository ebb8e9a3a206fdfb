/*
 * ples_oracle.c -- CPU FP64 restatement of the PLES pressure-Poisson path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Never part of the product path;
 * see ples_oracle.h for the layout conventions and the cited reference
 * routines.  Written from the reference's algorithm; loop orders and the
 * order of floating-point sums follow the cited Fortran so that the
 * restatement reproduces the reference trajectory on one global domain.
 */
#include "ples_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define GIDX(d, i, j, k) \
    ((size_t)(i) + (size_t)((d)->nx + 2) * ((size_t)(j) + (size_t)((d)->ny + 2) * (size_t)(k)))
#define NIDX(f, i, j, k) \
    ((size_t)(i) + (size_t)((f)->nx + 1) * ((size_t)(j) + (size_t)((f)->ny + 1) * (size_t)(k)))

static size_t gsize(const orc_dims* d) {
    return (size_t)(d->nx + 2) * (size_t)(d->ny + 2) * (size_t)(d->nz + 2);
}

/* ------------------------------------------------------------------ */
/* case generation                                                     */
/* ------------------------------------------------------------------ */

void orc_make_nodes(int nx, int ny, int nz, double Lx, double Ly, double Lz, double gamma,
                    double amp, double* x, double* y, double* z) {
    const double pi = 3.14159265358979323846;
    const double tg = gamma > 0.0 ? tanh(gamma) : 1.0;
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                size_t n = (size_t)i + (size_t)(nx + 1) * ((size_t)j + (size_t)(ny + 1) * (size_t)k);
                double xi = (double)i * Lx / (double)nx;
                double zk = (double)k * Lz / (double)nz;
                double eta = (double)j / (double)ny;
                double s = gamma > 0.0 ? 0.5 * (1.0 + tanh(gamma * (2.0 * eta - 1.0)) / tg) : eta;
                double zb = amp * sin(2.0 * pi * xi / (Lx / 4.0)) * cos(2.0 * pi * zk / (Lz / 4.0));
                x[n] = xi;
                y[n] = zb + (Ly - zb) * s;
                z[n] = zk;
            }
}

static uint64_t splitmix64(uint64_t v) {
    v += 0x9E3779B97F4A7C15ull;
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
    return v ^ (v >> 31);
}

double orc_gauss(uint64_t seed, uint32_t field, int64_t ctr) {
    uint64_t key = splitmix64(seed ^ ((uint64_t)field << 56));
    uint64_t h1 = splitmix64(key ^ (2ull * (uint64_t)ctr));
    uint64_t h2 = splitmix64(key ^ (2ull * (uint64_t)ctr + 1ull));
    double u1 = ((double)(h1 >> 11) + 1.0) * 0x1.0p-53; /* (0,1] */
    double u2 = (double)(h2 >> 11) * 0x1.0p-53;         /* [0,1) */
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
}

void orc_make_velocity(const orc_dims* d, uint64_t seed, double* u, double* v, double* w) {
    double* f[3] = {u, v, w};
    for (int c = 0; c < 3; ++c) {
        double* a = f[c];
        memset(a, 0, gsize(d) * sizeof(double));
        for (int k = 1; k <= d->nz; ++k)
            for (int j = 1; j <= d->ny; ++j)
                for (int i = 1; i <= d->nx; ++i) {
                    size_t id = GIDX(d, i, j, k);
                    a[id] = orc_gauss(seed, (uint32_t)c, (int64_t)id);
                }
        /* periodic ghost layers: copies of the opposite interior layer */
        if (d->ip == 0)
            for (int k = 1; k <= d->nz; ++k)
                for (int j = 1; j <= d->ny; ++j) {
                    a[GIDX(d, 0, j, k)] = a[GIDX(d, d->nx, j, k)];
                    a[GIDX(d, d->nx + 1, j, k)] = a[GIDX(d, 1, j, k)];
                }
        if (d->jp == 0)
            for (int k = 1; k <= d->nz; ++k)
                for (int i = 0; i <= d->nx + 1; ++i) {
                    a[GIDX(d, i, 0, k)] = a[GIDX(d, i, d->ny, k)];
                    a[GIDX(d, i, d->ny + 1, k)] = a[GIDX(d, i, 1, k)];
                }
        if (d->kp == 0)
            for (int j = 0; j <= d->ny + 1; ++j)
                for (int i = 0; i <= d->nx + 1; ++i) {
                    a[GIDX(d, i, j, 0)] = a[GIDX(d, i, j, d->nz)];
                    a[GIDX(d, i, j, d->nz + 1)] = a[GIDX(d, i, j, 1)];
                }
    }
}

/* ------------------------------------------------------------------ */
/* geometry                                                            */
/* ------------------------------------------------------------------ */

static int floordiv(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }

void orc_node(const orc_dims* f, const double* x, const double* y, const double* z, int i,
              int j, int k, double o[3]) {
    /* ghost-node extension of cellep (fsrc/cellep.f03:29-98), which fills
     * the ghost node layers in every direction the periodic way: shift along
     * the direction by (node(n) - node(0)) per period, copy the other two
     * coordinates. */
    if (k < 0 || k > f->nz) {
        int m = floordiv(k, f->nz);
        double a[3], b0[3], b1[3];
        orc_node(f, x, y, z, i, j, k - m * f->nz, a);
        orc_node(f, x, y, z, i, j, 0, b0);
        orc_node(f, x, y, z, i, j, f->nz, b1);
        o[0] = a[0];
        o[1] = a[1];
        o[2] = a[2] + (double)m * (b1[2] - b0[2]);
        return;
    }
    if (j < 0 || j > f->ny) {
        int m = floordiv(j, f->ny);
        double a[3], b0[3], b1[3];
        orc_node(f, x, y, z, i, j - m * f->ny, k, a);
        orc_node(f, x, y, z, i, 0, k, b0);
        orc_node(f, x, y, z, i, f->ny, k, b1);
        o[0] = a[0];
        o[1] = a[1] + (double)m * (b1[1] - b0[1]);
        o[2] = a[2];
        return;
    }
    if (i < 0 || i > f->nx) {
        int m = floordiv(i, f->nx);
        double a[3], b0[3], b1[3];
        orc_node(f, x, y, z, i - m * f->nx, j, k, a);
        orc_node(f, x, y, z, 0, j, k, b0);
        orc_node(f, x, y, z, f->nx, j, k, b1);
        o[0] = a[0] + (double)m * (b1[0] - b0[0]);
        o[1] = a[1];
        o[2] = a[2];
        return;
    }
    size_t n = NIDX(f, i, j, k);
    o[0] = x[n];
    o[1] = y[n];
    o[2] = z[n];
}

typedef struct {
    const orc_dims* f;
    const double *x, *y, *z;
} nodes_t;

static void nd(const nodes_t* N, int i, int j, int k, double o[3]) {
    orc_node(N->f, N->x, N->y, N->z, i, j, k, o);
}

/* average of four nodes, component-wise, in the given order */
static void avg4(const nodes_t* N, const int a[4][3], double o[3]) {
    double p0[3], p1[3], p2[3], p3[3];
    nd(N, a[0][0], a[0][1], a[0][2], p0);
    nd(N, a[1][0], a[1][1], a[1][2], p1);
    nd(N, a[2][0], a[2][1], a[2][2], p2);
    nd(N, a[3][0], a[3][1], a[3][2], p3);
    for (int c = 0; c < 3; ++c) o[c] = .25 * (p0[c] + p1[c] + p2[c] + p3[c]);
}

/* half-difference of two node pairs: .5*(A+B) - .5*(C+D) */
static void hdiff(const nodes_t* N, int a0, int a1, int a2, int b0, int b1, int b2, int c0, int c1,
                  int c2, int d0, int d1, int d2, double o[3]) {
    double A[3], B[3], C[3], D[3];
    nd(N, a0, a1, a2, A);
    nd(N, b0, b1, b2, B);
    nd(N, c0, c1, c2, C);
    nd(N, d0, d1, d2, D);
    for (int c = 0; c < 3; ++c) o[c] = .5 * (A[c] + B[c]) - .5 * (C[c] + D[c]);
}

/* vectors of the gprima family: cs = eta x zeta, et = zeta x csi,
 * zt = csi x eta; giac = csi . (eta x zeta)  (fsrc/gprima.f03:22-36) */
typedef struct {
    double cs[3], et[3], zt[3], giac;
} frame_t;

static void frame(const double c[3], const double e[3], const double t[3], frame_t* F) {
    /* c = (xcsi,ycsi,zcsi), e = (xeta,...), t = (xzet,...) */
    double xcsi = c[0], ycsi = c[1], zcsi = c[2];
    double xeta = e[0], yeta = e[1], zeta = e[2];
    double xzet = t[0], yzet = t[1], zzet = t[2];
    F->giac = xcsi * (yeta * zzet - yzet * zeta) - xeta * (ycsi * zzet - yzet * zcsi) +
              xzet * (ycsi * zeta - yeta * zcsi);
    F->cs[0] = yeta * zzet - yzet * zeta;
    F->cs[1] = xzet * zeta - xeta * zzet;
    F->cs[2] = xeta * yzet - xzet * yeta;
    F->et[0] = yzet * zcsi - ycsi * zzet;
    F->et[1] = xcsi * zzet - xzet * zcsi;
    F->et[2] = xzet * ycsi - xcsi * yzet;
    F->zt[0] = ycsi * zeta - yeta * zcsi;
    F->zt[1] = xeta * zcsi - xcsi * zeta;
    F->zt[2] = xcsi * yeta - xeta * ycsi;
}

static double dot3(const double a[3], const double b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static double sq3(const double a[3]) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }

void orc_metrics(const orc_dims* f, const double* x, const double* y, const double* z, int st,
                 double* const g[9], double* const csv[9], double* giac) {
    nodes_t N = {f, x, y, z};
    orc_dims lv = {f->nx / st, f->ny / st, f->nz / st, f->ip, f->jp, f->kp};
    const orc_dims* d = &lv;
    const double pot = (double)st;
    const size_t n = gsize(d);
    for (int m = 0; m < 9; ++m) memset(g[m], 0, n * sizeof(double));
    if (csv)
        for (int m = 0; m < 9; ++m) memset(csv[m], 0, n * sizeof(double));

    /* ---- csi = const faces: fsrc/metrica.f03:85-197, multigrid.f03:2520-2673 */
    for (int k = 1; k <= d->nz; ++k)
        for (int j = 1; j <= d->ny; ++j)
            for (int i = 0; i <= d->nx; ++i) {
                int j2 = j * st, j1 = (j - 1) * st, k2 = k * st, k1 = (k - 1) * st, ic = i * st;
                double c[3], e[3], t[3];
#define FACEX(a, o)                                                        \
    do {                                                                   \
        const int q_[4][3] = {{a, j2, k2}, {a, j2, k1}, {a, j1, k2}, {a, j1, k1}}; \
        avg4(&N, q_, o);                                                   \
    } while (0)
                if (i == 0 && d->ip == 1) {
                    double x0[3], x1[3], x2[3];
                    FACEX(0, x0);
                    FACEX(st, x1);
                    FACEX(2 * st, x2);
                    for (int q = 0; q < 3; ++q) c[q] = .5 * (-3. * x0[q] + 4. * x1[q] - x2[q]);
                } else if (i == d->nx && d->ip == 1) {
                    double x0[3], x1[3], x2[3];
                    FACEX(f->nx, x0);
                    FACEX((i - 1) * st, x1);
                    FACEX((i - 2) * st, x2);
                    for (int q = 0; q < 3; ++q) c[q] = .5 * (3. * x0[q] - 4. * x1[q] + x2[q]);
                } else {
                    double xd[3], xs[3];
                    FACEX((i + 1) * st, xd);
                    FACEX((i - 1) * st, xs);
                    for (int q = 0; q < 3; ++q) c[q] = .5 * (xd[q] - xs[q]);
                }
#undef FACEX
                hdiff(&N, ic, j2, k2, ic, j2, k1, ic, j1, k2, ic, j1, k1, e);
                hdiff(&N, ic, j2, k2, ic, j1, k2, ic, j2, k1, ic, j1, k1, t);
                frame_t F;
                frame(c, e, t, &F);
                size_t id = GIDX(d, i, j, k);
                g[0][id] = (sq3(F.cs) / F.giac) / pot;
                g[1][id] = (dot3(F.cs, F.et) / F.giac) / pot;
                g[2][id] = (dot3(F.cs, F.zt) / F.giac) / pot;
                if (csv) {
                    csv[0][id] = F.cs[0];
                    csv[1][id] = F.cs[1];
                    csv[2][id] = F.cs[2];
                }
            }

    /* ---- eta = const faces: fsrc/metrica.f03:203-294, multigrid.f03:2679-2827 */
    for (int k = 1; k <= d->nz; ++k)
        for (int j = 0; j <= d->ny; ++j)
            for (int i = 1; i <= d->nx; ++i) {
                int i2 = i * st, i1 = (i - 1) * st, k2 = k * st, k1 = (k - 1) * st, jc = j * st;
                double c[3], e[3], t[3];
#define FACEY(b, o)                                                        \
    do {                                                                   \
        const int q_[4][3] = {{i2, b, k2}, {i2, b, k1}, {i1, b, k1}, {i1, b, k2}}; \
        avg4(&N, q_, o);                                                   \
    } while (0)
                if (j == 0 && d->jp == 1) {
                    double x0[3], x1[3], x2[3];
                    FACEY(0, x0);
                    FACEY(st, x1);
                    FACEY(2 * st, x2);
                    for (int q = 0; q < 3; ++q) e[q] = .5 * (-3. * x0[q] + 4. * x1[q] - x2[q]);
                } else if (j == d->ny && d->jp == 1) {
                    double x0[3], x1[3], x2[3];
                    FACEY(f->ny, x0);
                    FACEY((j - 1) * st, x1);
                    FACEY((j - 2) * st, x2);
                    for (int q = 0; q < 3; ++q) e[q] = .5 * (3. * x0[q] - 4. * x1[q] + x2[q]);
                } else {
                    double xp[3], xm[3];
                    FACEY((j + 1) * st, xp);
                    FACEY((j - 1) * st, xm);
                    for (int q = 0; q < 3; ++q) e[q] = .5 * (xp[q] - xm[q]);
                }
#undef FACEY
                hdiff(&N, i2, jc, k2, i1, jc, k2, i2, jc, k1, i1, jc, k1, t);
                hdiff(&N, i2, jc, k2, i2, jc, k1, i1, jc, k2, i1, jc, k1, c);
                frame_t F;
                frame(c, e, t, &F);
                size_t id = GIDX(d, i, j, k);
                g[3][id] = (dot3(F.et, F.cs) / F.giac) / pot;
                g[4][id] = (sq3(F.et) / F.giac) / pot;
                g[5][id] = (dot3(F.et, F.zt) / F.giac) / pot;
                if (csv) {
                    csv[3][id] = F.et[0];
                    csv[4][id] = F.et[1];
                    csv[5][id] = F.et[2];
                }
            }

    /* ---- zeta = const faces: fsrc/metrica.f03:316-419, multigrid.f03:2834-2988 */
    for (int k = 0; k <= d->nz; ++k)
        for (int j = 1; j <= d->ny; ++j)
            for (int i = 1; i <= d->nx; ++i) {
                int i2 = i * st, i1 = (i - 1) * st, j2 = j * st, j1 = (j - 1) * st, kc = k * st;
                double c[3], e[3], t[3];
#define FACEZ(cc, o)                                                       \
    do {                                                                   \
        const int q_[4][3] = {{i2, j2, cc}, {i2, j1, cc}, {i1, j1, cc}, {i1, j2, cc}}; \
        avg4(&N, q_, o);                                                   \
    } while (0)
                if (k == 0 && d->kp == 1) {
                    double x0[3], x1[3], x2[3];
                    FACEZ(0, x0);
                    FACEZ(st, x1);
                    FACEZ(2 * st, x2);
                    for (int q = 0; q < 3; ++q) t[q] = .5 * (-3. * x0[q] + 4. * x1[q] - x2[q]);
                } else if (k == d->nz && d->kp == 1) {
                    double x0[3], x1[3], x2[3];
                    FACEZ(f->nz, x0);
                    FACEZ((k - 1) * st, x1);
                    FACEZ((k - 2) * st, x2);
                    for (int q = 0; q < 3; ++q) t[q] = .5 * (3. * x0[q] - 4. * x1[q] + x2[q]);
                } else {
                    double xa[3], xb[3];
                    FACEZ((k + 1) * st, xa);
                    FACEZ((k - 1) * st, xb);
                    for (int q = 0; q < 3; ++q) t[q] = .5 * (xa[q] - xb[q]);
                }
#undef FACEZ
                hdiff(&N, i2, j2, kc, i2, j1, kc, i1, j2, kc, i1, j1, kc, c);
                hdiff(&N, i2, j2, kc, i1, j2, kc, i2, j1, kc, i1, j1, kc, e);
                frame_t F;
                frame(c, e, t, &F);
                size_t id = GIDX(d, i, j, k);
                g[6][id] = (dot3(F.zt, F.cs) / F.giac) / pot;
                g[7][id] = (dot3(F.zt, F.et) / F.giac) / pot;
                g[8][id] = (sq3(F.zt) / F.giac) / pot;
                if (csv) {
                    csv[6][id] = F.zt[0];
                    csv[7][id] = F.zt[1];
                    csv[8][id] = F.zt[2];
                }
            }

    /* ---- centroid Jacobian: fsrc/metrica.f03:432-483 */
    if (giac) {
        memset(giac, 0, n * sizeof(double));
        for (int k = 1; k <= d->nz; ++k)
            for (int j = 1; j <= d->ny; ++j)
                for (int i = 1; i <= d->nx; ++i) {
                    double P[2][2][2][3];
                    for (int a = 0; a < 2; ++a)
                        for (int b = 0; b < 2; ++b)
                            for (int cc = 0; cc < 2; ++cc)
                                nd(&N, (i - 1 + a) * st, (j - 1 + b) * st, (k - 1 + cc) * st,
                                   P[a][b][cc]);
                    double c[3], e[3], t[3];
                    for (int q = 0; q < 3; ++q) {
                        c[q] = .25 * (P[1][1][1][q] + P[1][1][0][q] + P[1][0][0][q] + P[1][0][1][q]) -
                               .25 * (P[0][1][1][q] + P[0][1][0][q] + P[0][0][0][q] + P[0][0][1][q]);
                        e[q] = .25 * (P[1][1][1][q] + P[1][1][0][q] + P[0][1][0][q] + P[0][1][1][q]) -
                               .25 * (P[1][0][1][q] + P[1][0][0][q] + P[0][0][0][q] + P[0][0][1][q]);
                        t[q] = .25 * (P[1][1][1][q] + P[1][0][1][q] + P[0][0][1][q] + P[0][1][1][q]) -
                               .25 * (P[1][1][0][q] + P[1][0][0][q] + P[0][0][0][q] + P[0][1][0][q]);
                    }
                    frame_t F;
                    frame(c, e, t, &F);
                    giac[GIDX(d, i, j, k)] = F.giac;
                }
    }
}

/* ------------------------------------------------------------------ */
/* multigrid operators                                                 */
/* ------------------------------------------------------------------ */

#define P(i, j, k) p[GIDX(d, i, j, k)]
#define G(m, i, j, k) g[m][GIDX(d, i, j, k)]

/* Edge and corner ghosts of mul_boun_sndrcv (fsrc/multigrid.f03:1358-1742,
 * jp = 1 only): every ghost that is neither a face ghost nor interior is a
 * copy of one other value -- its periodic image along the periodic
 * directions, or its neighbour towards the interior along a wall direction.
 * The MPI exchanges of rank 0 <-> nproc-1 there become copies between the
 * planes k = 0, 1, nz, nz+1 of the one global array; the edge strips of
 * fsrc/multigrid.f03:1750-1931 are slab-internal halos and change nothing. */
void orc_edge_fill(const orc_dims* d, double* p) {
    const int nx = d->nx, ny = d->ny, nz = d->nz;
    /* edges and corners: fsrc/multigrid.f03:1358-1742 (jp = 1 only) */
    if (d->jp == 1) {
        if (d->kp == 0) {
            /* 1) corner lines along x */
            for (int i = 1; i <= nx; ++i) {
                P(i, ny + 1, 0) = P(i, ny + 1, nz);
                P(i, 0, 0) = P(i, 0, nz);
                P(i, ny + 1, nz + 1) = P(i, ny + 1, 1);
                P(i, 0, nz + 1) = P(i, 0, 1);
            }
            /* 2) vertical frame on planes k = 0, nz+1 */
            if (d->ip == 0) {
                for (int j = 0; j <= ny + 1; ++j) {
                    P(0, j, 0) = P(nx, j, nz);
                    P(nx + 1, j, 0) = P(1, j, nz);
                    P(0, j, nz + 1) = P(nx, j, 1);
                    P(nx + 1, j, nz + 1) = P(1, j, 1);
                }
            } else {
                for (int j = 0; j <= ny + 1; ++j) {
                    P(nx + 1, j, 0) = P(nx + 1, j, nz);
                    P(0, j, 0) = P(0, j, nz);
                    P(nx + 1, j, nz + 1) = P(nx + 1, j, 1);
                    P(0, j, nz + 1) = P(0, j, 1);
                }
            }
            /* 3) horizontal frame */
            if (d->ip == 0) {
                for (int k = 1; k <= nz; ++k) {
                    P(0, ny + 1, k) = P(nx, ny + 1, k);
                    P(0, 0, k) = P(nx, 0, k);
                    P(nx + 1, ny + 1, k) = P(1, ny + 1, k);
                    P(nx + 1, 0, k) = P(1, 0, k);
                }
            } else {
                for (int k = 1; k <= nz; ++k) {
                    P(0, ny + 1, k) = P(0, ny, k);
                    P(0, 0, k) = P(0, 1, k);
                    P(nx + 1, ny + 1, k) = P(nx + 1, ny, k);
                    P(nx + 1, 0, k) = P(nx + 1, 1, k);
                }
            }
        } else {
            /* 1) corner lines along x */
            for (int i = 1; i <= nx; ++i) {
                P(i, 0, 0) = P(i, 1, 0);
                P(i, ny + 1, 0) = P(i, ny, 0);
            }
            for (int i = 1; i <= nx; ++i) {
                P(i, 0, nz + 1) = P(i, 1, nz + 1);
                P(i, ny + 1, nz + 1) = P(i, ny, nz + 1);
            }
            /* 2) vertical frame */
            if (d->ip == 0) {
                for (int j = 0; j <= ny + 1; ++j) {
                    P(0, j, 0) = P(nx, j, 0);
                    P(nx + 1, j, 0) = P(1, j, 0);
                }
                for (int j = 0; j <= ny + 1; ++j) {
                    P(0, j, nz + 1) = P(nx, j, nz + 1);
                    P(nx + 1, j, nz + 1) = P(1, j, nz + 1);
                }
            } else {
                for (int j = 0; j <= ny + 1; ++j) {
                    P(0, j, 0) = P(1, j, 0);
                    P(nx + 1, j, 0) = P(nx, j, 0);
                }
                for (int j = 0; j <= ny + 1; ++j) {
                    P(0, j, nz + 1) = P(1, j, nz + 1);
                    P(nx + 1, j, nz + 1) = P(nx, j, nz + 1);
                }
            }
            /* 3) horizontal frame */
            if (d->ip == 0) {
                for (int k = 1; k <= nz; ++k) {
                    P(0, ny + 1, k) = P(nx, ny + 1, k);
                    P(0, 0, k) = P(nx, 0, k);
                    P(nx + 1, ny + 1, k) = P(1, ny + 1, k);
                    P(nx + 1, 0, k) = P(1, 0, k);
                }
            } else {
                for (int k = 1; k <= nz; ++k) {
                    P(0, ny + 1, k) = P(1, ny + 1, k);
                    P(0, 0, k) = P(1, 0, k);
                    P(nx + 1, ny + 1, k) = P(nx, ny + 1, k);
                    P(nx + 1, 0, k) = P(nx, 0, k);
                }
            }
        }
    }
}

void orc_ghost_fill(const orc_dims* d, double* p, const double* const g[9],
                    const double* const cs[6], double an, double dt, int npass) {
    const int nx = d->nx, ny = d->ny, nz = d->nz;
    const double inv_dt = 1. / dt, coef = 0.25;
    const int have_cs = (cs != NULL) && an != 0.0;
#define CS(m, a, b, lda) (have_cs ? cs[m][(size_t)((a)-1) + (size_t)(lda) * (size_t)((b)-1)] : 0.0)
    for (int pass = 0; pass < npass; ++pass) {
        /* left/right: fsrc/multigrid.f03:1113-1158 */
        if (d->ip == 1) {
            for (int k = 1; k <= nz; ++k)
                for (int j = 1; j <= ny; ++j) {
                    double inv = 1. / G(0, 0, j, k);
                    P(0, j, k) = P(1, j, k) - an * CS(0, j, k, ny) * inv_dt * inv +
                                 coef * inv *
                                     (G(1, 0, j, k) * (P(1, j + 1, k) + P(0, j + 1, k) - P(1, j - 1, k) -
                                                       P(0, j - 1, k)) +
                                      G(2, 0, j, k) * (P(1, j, k + 1) + P(0, j, k + 1) - P(1, j, k - 1) -
                                                       P(0, j, k - 1)));
                    inv = 1. / G(0, nx, j, k);
                    P(nx + 1, j, k) =
                        P(nx, j, k) + an * CS(1, j, k, ny) * inv_dt * inv -
                        coef * inv *
                            (G(1, nx, j, k) * (P(nx + 1, j + 1, k) + P(nx, j + 1, k) -
                                               P(nx + 1, j - 1, k) - P(nx, j - 1, k)) +
                             G(2, nx, j, k) * (P(nx + 1, j, k + 1) + P(nx, j, k + 1) -
                                               P(nx + 1, j, k - 1) - P(nx, j, k - 1)));
                }
        } else {
            for (int k = 1; k <= nz; ++k)
                for (int j = 1; j <= ny; ++j) {
                    P(0, j, k) = P(nx, j, k);
                    P(nx + 1, j, k) = P(1, j, k);
                }
        }
        /* bottom/upper: fsrc/multigrid.f03:1161-1278 (freesurface = 0) */
        if (d->jp == 1) {
            for (int k = 1; k <= nz; ++k)
                for (int i = 1; i <= nx; ++i) {
                    double inv = 1. / G(4, i, 0, k);
                    P(i, 0, k) = P(i, 1, k) - an * CS(2, i, k, nx) * inv_dt * inv +
                                 coef * inv *
                                     (G(3, i, 0, k) * (P(i + 1, 1, k) + P(i + 1, 0, k) - P(i - 1, 1, k) -
                                                       P(i - 1, 0, k)) +
                                      G(5, i, 0, k) * (P(i, 1, k + 1) + P(i, 0, k + 1) - P(i, 1, k - 1) -
                                                       P(i, 0, k - 1)));
                    inv = 1. / G(4, i, ny, k);
                    P(i, ny + 1, k) =
                        P(i, ny, k) + an * CS(3, i, k, nx) * inv_dt * inv -
                        coef * inv *
                            (G(3, i, ny, k) * (P(i + 1, ny + 1, k) + P(i + 1, ny, k) -
                                               P(i - 1, ny + 1, k) - P(i - 1, ny, k)) +
                             G(5, i, ny, k) * (P(i, ny + 1, k + 1) + P(i, ny, k + 1) -
                                               P(i, ny + 1, k - 1) - P(i, ny, k - 1)));
                }
        } else {
            for (int k = 1; k <= nz; ++k)
                for (int i = 1; i <= nx; ++i) {
                    P(i, 0, k) = P(i, ny, k);
                    P(i, ny + 1, k) = P(i, 1, k);
                }
        }
        /* back/front: fsrc/multigrid.f03:1281-1352 */
        if (d->kp == 1) {
            for (int i = 1; i <= nx; ++i)
                for (int j = 1; j <= ny; ++j) {
                    double inv = 1. / G(8, i, j, 0);
                    P(i, j, 0) = P(i, j, 1) - an * CS(4, i, j, nx) * inv_dt * inv +
                                 coef * inv *
                                     (G(6, i, j, 0) * (P(i + 1, j, 1) + P(i + 1, j, 0) - P(i - 1, j, 1) -
                                                       P(i - 1, j, 0)) +
                                      G(7, i, j, 0) * (P(i, j + 1, 1) + P(i, j + 1, 0) - P(i, j - 1, 1) -
                                                       P(i, j - 1, 0)));
                }
            for (int i = 1; i <= nx; ++i)
                for (int j = 1; j <= ny; ++j) {
                    double inv = 1. / G(8, i, j, nz);
                    P(i, j, nz + 1) =
                        P(i, j, nz) + an * CS(5, i, j, nx) * inv_dt * inv -
                        coef * inv *
                            (G(6, i, j, nz) * (P(i + 1, j, nz + 1) + P(i + 1, j, nz) -
                                               P(i - 1, j, nz + 1) - P(i - 1, j, nz)) +
                             G(7, i, j, nz) * (P(i, j + 1, nz + 1) + P(i, j + 1, nz) -
                                               P(i, j - 1, nz + 1) - P(i, j - 1, nz)));
                }
        } else {
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    P(i, j, 0) = P(i, j, nz);
                    P(i, j, nz + 1) = P(i, j, 1);
                }
        }
        orc_edge_fill(d, p);
    }
#undef CS
}

/* six face fluxes (times s) of cell (i,j,k): fsrc/resid_sndrcv.f03:85-142 */
#define FLUXES                                                                                  \
    double res_dx = (G(0, i, j, k) * (P(i + 1, j, k) - P(i, j, k)) +                              \
                     coef * (G(1, i, j, k) * (P(i + 1, j + 1, k) + P(i, j + 1, k) - P(i + 1, j - 1, k) - \
                                              P(i, j - 1, k)) +                                  \
                             G(2, i, j, k) * (P(i + 1, j, k + 1) + P(i, j, k + 1) - P(i + 1, j, k - 1) - \
                                              P(i, j, k - 1)))) *                                \
                    inv_pot;                                                                    \
    double res_sn = (G(0, i - 1, j, k) * (P(i, j, k) - P(i - 1, j, k)) +                          \
                     coef * (G(1, i - 1, j, k) * (P(i, j + 1, k) + P(i - 1, j + 1, k) - P(i, j - 1, k) - \
                                                  P(i - 1, j - 1, k)) +                          \
                             G(2, i - 1, j, k) * (P(i, j, k + 1) + P(i - 1, j, k + 1) - P(i, j, k - 1) - \
                                                  P(i - 1, j, k - 1)))) *                        \
                    inv_pot;                                                                    \
    double res_sop = (G(4, i, j, k) * (P(i, j + 1, k) - P(i, j, k)) +                             \
                      coef * (G(3, i, j, k) * (P(i + 1, j + 1, k) + P(i + 1, j, k) - P(i - 1, j + 1, k) - \
                                               P(i - 1, j, k)) +                                 \
                              G(5, i, j, k) * (P(i, j + 1, k + 1) + P(i, j, k + 1) - P(i, j + 1, k - 1) - \
                                               P(i, j, k - 1)))) *                               \
                     inv_pot;                                                                   \
    double res_sot = (G(4, i, j - 1, k) * (P(i, j, k) - P(i, j - 1, k)) +                         \
                      coef * (G(3, i, j - 1, k) * (P(i + 1, j, k) + P(i + 1, j - 1, k) - P(i - 1, j, k) - \
                                                   P(i - 1, j - 1, k)) +                         \
                              G(5, i, j - 1, k) * (P(i, j, k + 1) + P(i, j - 1, k + 1) - P(i, j, k - 1) - \
                                                   P(i, j - 1, k - 1)))) *                       \
                     inv_pot;                                                                   \
    double res_av = (G(8, i, j, k) * (P(i, j, k + 1) - P(i, j, k)) +                              \
                     coef * (G(6, i, j, k) * (P(i + 1, j, k + 1) + P(i + 1, j, k) - P(i - 1, j, k + 1) - \
                                              P(i - 1, j, k)) +                                  \
                             G(7, i, j, k) * (P(i, j + 1, k + 1) + P(i, j + 1, k) - P(i, j - 1, k + 1) - \
                                              P(i, j - 1, k)))) *                                \
                    inv_pot;                                                                    \
    double res_ind = (G(8, i, j, k - 1) * (P(i, j, k) - P(i, j, k - 1)) +                         \
                      coef * (G(6, i, j, k - 1) * (P(i + 1, j, k) + P(i + 1, j, k - 1) - P(i - 1, j, k) - \
                                                   P(i - 1, j, k - 1)) +                         \
                              G(7, i, j, k - 1) * (P(i, j + 1, k) + P(i, j + 1, k - 1) - P(i, j - 1, k) - \
                                                   P(i, j - 1, k - 1)))) *                       \
                     inv_pot;

void orc_sor_sweep(const orc_dims* d, double* p, const double* rh, const double* const g[9],
                   double pot, double omega) {
    const int nx = d->nx, ny = d->ny, nz = d->nz;
    const double inv_pot = 1. / pot, ppot2 = 1. / pot / pot, coef = .25;
    /* periodic-k plane exchange before the sweep: fsrc/multigrid.f03:784-812 */
    if (d->kp == 0) {
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) P(i, j, 0) = P(i, j, nz);
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) P(i, j, nz + 1) = P(i, j, 1);
    }
    /* lexicographic sweep: fsrc/multigrid.f03:815-919 */
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j) {
            if (d->ip == 0) {
                P(0, j, k) = P(nx, j, k);
                P(nx + 1, j, k) = P(1, j, k);
            }
            for (int i = 1; i <= nx; ++i) {
                if (d->jp == 0) {
                    P(i, 0, k) = P(i, ny, k);
                    P(i, ny + 1, k) = P(i, 1, k);
                }
                FLUXES
                double den = G(0, i, j, k) * ppot2 + G(0, i - 1, j, k) * ppot2 + G(4, i, j, k) * ppot2 +
                             G(4, i, j - 1, k) * ppot2 + G(8, i, j, k) * ppot2 +
                             G(8, i, j, k - 1) * ppot2;
                double resi = (res_dx - res_sn + res_sop - res_sot + res_av - res_ind) * inv_pot -
                              rh[GIDX(d, i, j, k)];
                P(i, j, k) = P(i, j, k) + omega * resi / den;
            }
        }
}

void orc_resid(const orc_dims* d, const double* p, const double* rh, const double* const g[9],
               double pot, double* r) {
    const double inv_pot = 1. / pot, coef = .25;
    for (int k = 1; k <= d->nz; ++k)
        for (int j = 1; j <= d->ny; ++j)
            for (int i = 1; i <= d->nx; ++i) {
                FLUXES
                r[GIDX(d, i, j, k)] =
                    (res_dx - res_sn + res_sop - res_sot + res_av - res_ind) * inv_pot - rh[GIDX(d, i, j, k)];
            }
}

/* orc_resid with its planes split over nthr OpenMP threads (bitwise the same) */
static void orc_resid_mt(const orc_dims* d, const double* p, const double* rh, const double* const g[9],
                         double pot, double* r, int nthr) {
    const double inv_pot = 1. / pot, coef = .25;
#pragma omp parallel for num_threads(nthr) if (nthr > 1) schedule(static)
    for (int k = 1; k <= d->nz; ++k)
        for (int j = 1; j <= d->ny; ++j)
            for (int i = 1; i <= d->nx; ++i) {
                FLUXES
                r[GIDX(d, i, j, k)] =
                    (res_dx - res_sn + res_sop - res_sot + res_av - res_ind) * inv_pot - rh[GIDX(d, i, j, k)];
            }
}
#undef FLUXES

void orc_restrict(const orc_dims* dc, const double* rf, double* rc) {
    orc_dims fd = {dc->nx * 2, dc->ny * 2, dc->nz * 2, dc->ip, dc->jp, dc->kp};
    const orc_dims* f = &fd;
    const double inv_8 = 1. / 8.;
    for (int k = 1; k <= dc->nz; ++k)
        for (int j = 1; j <= dc->ny; ++j)
            for (int i = 1; i <= dc->nx; ++i) {
                int id = 2 * i, jd = 2 * j, kd = 2 * k;
                double s = rf[GIDX(f, id, jd, kd)] + rf[GIDX(f, id - 1, jd, kd)] + rf[GIDX(f, id, jd - 1, kd)] +
                           rf[GIDX(f, id - 1, jd - 1, kd)] + rf[GIDX(f, id, jd, kd - 1)] +
                           rf[GIDX(f, id - 1, jd, kd - 1)] + rf[GIDX(f, id, jd - 1, kd - 1)] +
                           rf[GIDX(f, id - 1, jd - 1, kd - 1)];
                rc[GIDX(dc, i, j, k)] = -s * inv_8;
            }
}

void orc_prolong(const orc_dims* dc, const double* ec, double* pf) {
    orc_dims fd = {dc->nx * 2, dc->ny * 2, dc->nz * 2, dc->ip, dc->jp, dc->kp};
    const orc_dims* f = &fd;
    const double inv_64 = 1. / 64.;
    for (int k = 0; k <= dc->nz; ++k)
        for (int j = 0; j <= dc->ny; ++j)
            for (int i = 0; i <= dc->nx; ++i) {
                const double c[8] = {ec[GIDX(dc, i, j, k)],         ec[GIDX(dc, i + 1, j, k)],
                                     ec[GIDX(dc, i, j + 1, k)],     ec[GIDX(dc, i, j, k + 1)],
                                     ec[GIDX(dc, i + 1, j + 1, k)], ec[GIDX(dc, i + 1, j, k + 1)],
                                     ec[GIDX(dc, i, j + 1, k + 1)], ec[GIDX(dc, i + 1, j + 1, k + 1)]};
                /* coarse offsets of c[] in the order a1..a8 */
                static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                              {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
                /* fine children (di,dj,dk) in the order of fsrc/prolong.f03:81-257 */
                static const int order[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                                {1, 1, 0}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
                for (int q = 0; q < 8; ++q) {
                    const int di = order[q][0], dj = order[q][1], dk = order[q][2];
                    double w[8];
                    for (int a = 0; a < 8; ++a)
                        w[a] = (off[a][0] == di ? 3. : 1.) * (off[a][1] == dj ? 3. : 1.) *
                               (off[a][2] == dk ? 3. : 1.);
                    double s = w[0] * c[0] + w[1] * c[1] + w[2] * c[2] + w[3] * c[3] + w[4] * c[4] +
                               w[5] * c[5] + w[6] * c[6] + w[7] * c[7];
                    size_t id = GIDX(f, 2 * i + di, 2 * j + dj, 2 * k + dk);
                    pf[id] = pf[id] + s * inv_64;
                }
            }
}

/* ---- line relaxation in eta (islor = 1) ------------------------------ */

void orc_tridag(int n, double* aa, double* bb, double* cc, double* ff) {
    /* fsrc/tridag.f03:11-30, rows 0..n-1 for the Fortran's 1..n */
    bb[0] = 1. / bb[0];
    aa[0] = ff[0] * bb[0];
    for (int k = 1; k < n; ++k) {
        const int k1 = k - 1;
        cc[k1] = cc[k1] * bb[k1];
        bb[k] = bb[k] - aa[k] * cc[k1];
        bb[k] = 1. / bb[k];
        aa[k] = (ff[k] - aa[k] * aa[k1]) * bb[k];
    }
    ff[n - 1] = aa[n - 1];
    for (int k = n - 2; k >= 0; --k) ff[k] = aa[k] - cc[k] * ff[k + 1];
}

/* wall flags of sett (fsrc/sett.f03:53-108): 0 on a non-periodic boundary
 * face of the cell, 1 elsewhere.  Global semantics: both zeta walls are
 * flagged (the reference flags i_in on rank 0 and i_av on rank nproc-1,
 * fsrc/sett.f03:95-108; a single rank would flag only i_in). */
typedef struct {
    int dx, sn, sp, st, av, in;
} flags_t;

static flags_t cell_flags(const orc_dims* d, int i, int j, int k) {
    flags_t f = {1, 1, 1, 1, 1, 1};
    if (i == 1) f.sn = 1 - d->ip;
    if (i == d->nx) f.dx = 1 - d->ip;
    if (j == 1) f.st = 1 - d->jp;
    if (j == d->ny) f.sp = 1 - d->jp;
    if (k == 1) f.in = 1 - d->kp;
    if (k == d->nz) f.av = 1 - d->kp;
    return f;
}

/* the line coefficients of column (i, k) (rows 0..ny+1) of sett,
 * fsrc/sett.f03:119-143; d is the level (k, nz global) */
static void sett_column(const orc_dims* d, const double* const g[9], double pot, int i, int k, int kl,
                        double* aa, double* bb, double* cc) {
    const double ppot2 = 1. / pot / pot;
    const int ny = d->ny;
    /* metrics of the level are indexed by the local plane kl (slab offset) */
    const orc_dims dl = {d->nx, d->ny, 1, d->ip, d->jp, d->kp};
#define GL(m, a, b, c) g[m][GIDX(&dl, a, b, c)]
    for (int j = 1; j <= ny; ++j) {
        const flags_t f = cell_flags(d, i, j, k);
        aa[j] = ppot2 * GL(4, i, j - 1, kl);
        bb[j] = -ppot2 * (GL(4, i, j, kl) + GL(4, i, j - 1, kl) + GL(0, i, j, kl) * f.dx + GL(0, i - 1, j, kl) * f.sn +
                          GL(8, i, j, kl) * f.av + GL(8, i, j, kl - 1) * f.in);
        cc[j] = ppot2 * GL(4, i, j, kl);
    }
    /* side bottom (3): aa = 0; side upper (4): cc = 0 */
    aa[0] = 0.;
    bb[0] = GL(4, i, 0, kl) * ppot2;
    cc[0] = -GL(4, i, 0, kl) * ppot2;
    aa[ny + 1] = -GL(4, i, ny, kl) * ppot2;
    bb[ny + 1] = GL(4, i, ny, kl) * ppot2;
    cc[ny + 1] = 0.;
#undef GL
}

void orc_sett(const orc_dims* d, const double* const g[9], double pot, double* aa, double* bb, double* cc) {
    const int ny = d->ny;
    const size_t n = gsize(d);
    memset(aa, 0, n * sizeof(double));
    memset(bb, 0, n * sizeof(double));
    memset(cc, 0, n * sizeof(double));
    double* col = (double*)malloc((size_t)(ny + 2) * 3 * sizeof(double));
    for (int k = 1; k <= d->nz; ++k)
        for (int i = 1; i <= d->nx; ++i) {
            sett_column(d, g, pot, i, k, k, col, col + ny + 2, col + 2 * (ny + 2));
            for (int j = 0; j <= ny + 1; ++j) {
                const size_t id = GIDX(d, i, j, k);
                aa[id] = col[j];
                bb[id] = col[ny + 2 + j];
                cc[id] = col[2 * (ny + 2) + j];
            }
        }
    free(col);
}

/* the line solves of one SLOR sweep over the planes of d (a k-slab of the
 * level dglob: global plane = local + kofs; p, rh and g offset to the slab;
 * the wall flags are the level's) */
static void slor_planes(const orc_dims* d, const orc_dims* dglob, int kofs, double* p, const double* rh,
                        const double* const g[9], const double* const cs[6], double an, double dt,
                        double pot, double omega) {
    const int nx = d->nx, ny = d->ny, nz = d->nz;
    const orc_dims dg = *dglob;
    const double ppot2 = 1. / pot / pot, q = .25;
    const int have_cs = (cs != NULL) && an != 0.0;
#define CSV(m, a, b, lda) (have_cs ? cs[m][(size_t)((a)-1) + (size_t)(lda) * (size_t)((b)-1)] : 0.0)
    double* aa = (double*)malloc((size_t)(ny + 2) * 4 * sizeof(double));
    double *bb = aa + (ny + 2), *cc = bb + (ny + 2), *ff = cc + (ny + 2);
    for (int k = 1; k <= nz; ++k) {
        if (d->ip == 0)
            for (int j = 1; j <= ny; ++j) {
                P(0, j, k) = P(nx, j, k);
                P(nx + 1, j, k) = P(1, j, k);
            }
        for (int i = 1; i <= nx; ++i) {
            for (int j = 1; j <= ny; ++j) { /* 2128-2183 */
                const flags_t f = cell_flags(&dg, i, j, k + kofs);
                double s =
                    -G(0, i, j, k) * P(i + 1, j, k) * f.dx - (1. - f.dx) * an * CSV(1, j, k + kofs, ny) / dt -
                    G(0, i - 1, j, k) * P(i - 1, j, k) * f.sn + (1. - f.sn) * an * CSV(0, j, k + kofs, ny) / dt -
                    q * G(1, i, j, k) * f.dx * (P(i + 1, j + 1, k) + P(i, j + 1, k) - P(i + 1, j - 1, k) - P(i, j - 1, k)) -
                    q * G(2, i, j, k) * f.dx * (P(i + 1, j, k + 1) + P(i, j, k + 1) - P(i + 1, j, k - 1) - P(i, j, k - 1)) +
                    q * G(1, i - 1, j, k) * f.sn *
                        (P(i, j + 1, k) + P(i - 1, j + 1, k) - P(i, j - 1, k) - P(i - 1, j - 1, k)) +
                    q * G(2, i - 1, j, k) * f.sn *
                        (P(i, j, k + 1) + P(i - 1, j, k + 1) - P(i, j, k - 1) - P(i - 1, j, k - 1)) -
                    G(8, i, j, k) * P(i, j, k + 1) * f.av - (1. - f.av) * an * CSV(5, i, j, nx) / dt -
                    G(8, i, j, k - 1) * P(i, j, k - 1) * f.in + (1. - f.in) * an * CSV(4, i, j, nx) / dt -
                    q * G(6, i, j, k) * f.av * (P(i + 1, j, k + 1) + P(i + 1, j, k) - P(i - 1, j, k + 1) - P(i - 1, j, k)) -
                    q * G(7, i, j, k) * f.av * (P(i, j + 1, k + 1) + P(i, j + 1, k) - P(i, j - 1, k + 1) - P(i, j - 1, k)) +
                    q * G(6, i, j, k - 1) * f.in *
                        (P(i + 1, j, k) + P(i + 1, j, k - 1) - P(i - 1, j, k) - P(i - 1, j, k - 1)) +
                    q * G(7, i, j, k - 1) * f.in *
                        (P(i, j + 1, k) + P(i, j + 1, k - 1) - P(i, j - 1, k) - P(i, j - 1, k - 1)) -
                    q * G(3, i, j, k) * (P(i + 1, j + 1, k) + P(i + 1, j, k) - P(i - 1, j + 1, k) - P(i - 1, j, k)) -
                    q * G(5, i, j, k) * (P(i, j + 1, k + 1) + P(i, j, k + 1) - P(i, j + 1, k - 1) - P(i, j, k - 1)) +
                    q * G(3, i, j - 1, k) * (P(i + 1, j, k) + P(i + 1, j - 1, k) - P(i - 1, j, k) - P(i - 1, j - 1, k)) +
                    q * G(5, i, j - 1, k) * (P(i, j, k + 1) + P(i, j - 1, k + 1) - P(i, j, k - 1) - P(i, j - 1, k - 1));
                ff[j] = s * ppot2 + rh[GIDX(d, i, j, k)];
            }
            /* wall rows of the ghosts, 2185-2205 */
            ff[0] = (-an * CSV(2, i, k + kofs, nx) / dt +
                     q * G(3, i, 0, k) * (P(i + 1, 1, k) + P(i + 1, 0, k) - P(i - 1, 1, k) - P(i - 1, 0, k)) +
                     q * G(5, i, 0, k) * (P(i, 1, k + 1) + P(i, 0, k + 1) - P(i, 1, k - 1) - P(i, 0, k - 1))) *
                    ppot2;
            ff[ny + 1] = (an * CSV(3, i, k + kofs, nx) / dt -
                          q * G(3, i, ny, k) *
                              (P(i + 1, ny + 1, k) + P(i + 1, ny, k) - P(i - 1, ny + 1, k) - P(i - 1, ny, k)) -
                          q * G(5, i, ny, k) *
                              (P(i, ny + 1, k + 1) + P(i, ny, k + 1) - P(i, ny + 1, k - 1) - P(i, ny, k - 1))) *
                         ppot2;
            sett_column(&dg, g, pot, i, k + kofs, k, aa, bb, cc);
            orc_tridag(ny + 2, aa, bb, cc, ff);
            for (int j = 0; j <= ny + 1; ++j) P(i, j, k) = P(i, j, k) + omega * (ff[j] - P(i, j, k));
        }
    }
    free(aa);
#undef CSV
}

void orc_slor_sweep(const orc_dims* d, double* p, const double* rh, const double* const g[9],
                    const double* const cs[6], double an, double dt, double pot, double omega, int nbc) {
    /* one pseudo-time iteration (kk) of solut_sndrcv_slor,
     * fsrc/multigrid.f03:2050-2199 */
    const int nx = d->nx, ny = d->ny, nz = d->nz;
    orc_ghost_fill(d, p, g, cs, an, dt, nbc); /* mul_boun_sndrcv (2050-2057) */
    if (d->kp == 0) {                         /* periodic-k planes (2068-2094) */
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) P(i, j, 0) = P(i, j, nz);
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) P(i, j, nz + 1) = P(i, j, 1);
    }
    slor_planes(d, d, 0, p, rh, g, cs, an, dt, pot, omega);
}
#undef P
#undef G

/* ------------------------------------------------------------------ */
/* V-cycle driver                                                      */
/* ------------------------------------------------------------------ */

/* Level hierarchy of multi: per-level dims, metrics (metrica / mul_met,
 * computed once as the reference does at init, fsrc/stratiParticle.f03:767-769)
 * and the work arrays pr, rhs, residual of fsrc/multigrid.f03:144-154. */
struct orc_hier {
    orc_dims d;
    int L;
    orc_dims* dl;
    double*** gl;
    double **pr, **rv, **rn;
    double** slab; /* k-slab buffers (nslab > 1) */
    int nslab_alloc;
};

orc_hier* orc_hier_create(const orc_dims* d, int nlevel, const double* x, const double* y, const double* z,
                          const double* const g1[9]) {
    const int L = nlevel;
    orc_hier* H = (orc_hier*)calloc(1, sizeof(orc_hier));
    H->d = *d;
    H->L = L;
    H->dl = (orc_dims*)calloc((size_t)L, sizeof(orc_dims));
    H->gl = (double***)calloc((size_t)L, sizeof(double**));
    H->pr = (double**)calloc((size_t)L, sizeof(double*));
    H->rv = (double**)calloc((size_t)L, sizeof(double*));
    H->rn = (double**)calloc((size_t)L, sizeof(double*));
    for (int n = 0; n < L; ++n) {
        const int st = 1 << n;
        H->dl[n] = (orc_dims){d->nx / st, d->ny / st, d->nz / st, d->ip, d->jp, d->kp};
        const size_t sz = gsize(&H->dl[n]);
        H->gl[n] = (double**)calloc(9, sizeof(double*));
        for (int m = 0; m < 9; ++m) H->gl[n][m] = (double*)calloc(sz, sizeof(double));
        H->pr[n] = (double*)calloc(sz, sizeof(double));
        H->rv[n] = (double*)calloc(sz, sizeof(double));
        H->rn[n] = (double*)calloc(sz, sizeof(double));
        if (n == 0 && g1) {
            for (int m = 0; m < 9; ++m) memcpy(H->gl[0][m], g1[m], sz * sizeof(double));
        } else {
            orc_metrics(d, x, y, z, st, H->gl[n], NULL, NULL);
        }
    }
    return H;
}

void orc_hier_destroy(orc_hier* H) {
    if (!H) return;
    for (int n = 0; n < H->L; ++n) {
        for (int m = 0; m < 9; ++m) free(H->gl[n][m]);
        free(H->gl[n]);
        free(H->pr[n]);
        free(H->rv[n]);
        free(H->rn[n]);
    }
    for (int s = 0; s < H->nslab_alloc; ++s) free(H->slab[s]);
    free(H->slab);
    free(H->dl);
    free(H->gl);
    free(H->pr);
    free(H->rv);
    free(H->rn);
    free(H);
}

/* One smoother sweep of level d on ns k-slabs, the reference's MPI slab
 * semantics (fsrc/multigrid.f03:929-966): every slab relaxes its own planes
 * in the reference order, reading the planes next to it (its halo) as they
 * were when the sweep started; the halos are exchanged after the sweep.
 * The slabs run on ns OpenMP threads.  The ghost fill before the sweep is
 * global.  Results do not depend on the thread count, only on ns. */
static void sweep_slabs(orc_hier* H, const orc_dims* d, double* p, const double* rh, const double* const g[9],
                        const double* const cs[6], double an, double dt, double pot, double omega, int islor,
                        int ns) {
    const int nz = d->nz, nzs = nz / ns;
    const size_t sz = (size_t)(d->nx + 2) * (size_t)(d->ny + 2);
    if (d->kp == 0) { /* periodic-k planes before the sweep (fsrc/multigrid.f03:784-812) */
        for (int j = 1; j <= d->ny; ++j)
            for (int i = 1; i <= d->nx; ++i) {
                p[GIDX(d, i, j, 0)] = p[GIDX(d, i, j, nz)];
                p[GIDX(d, i, j, nz + 1)] = p[GIDX(d, i, j, 1)];
            }
    }
#pragma omp parallel for num_threads(ns) schedule(static, 1)
    for (int s = 0; s < ns; ++s)
        memcpy(H->slab[s], p + sz * (size_t)(s * nzs), (size_t)(nzs + 2) * sz * sizeof(double));
#pragma omp parallel for num_threads(ns) schedule(static, 1)
    for (int s = 0; s < ns; ++s) {
        const size_t off = sz * (size_t)(s * nzs);
        const orc_dims ds = {d->nx, d->ny, nzs, d->ip, d->jp, 1};
        const double* gs[9];
        for (int m = 0; m < 9; ++m) gs[m] = g[m] + off;
        if (islor)
            slor_planes(&ds, d, s * nzs, H->slab[s], rh + off, gs, cs, an, dt, pot, omega);
        else
            orc_sor_sweep(&ds, H->slab[s], rh + off, gs, pot, omega);
    }
#pragma omp parallel for num_threads(ns) schedule(static, 1)
    for (int s = 0; s < ns; ++s)
        memcpy(p + sz * (size_t)(s * nzs + 1), H->slab[s] + sz, (size_t)nzs * sz * sizeof(double));
}

static void smooth_level(orc_hier* H, int n, const double* rh, const double* const cs[6], double dt,
                         double omega, int kss, int nbc, int islor, int ns) {
    const orc_dims* dn = &H->dl[n];
    const double pot = (double)(1 << n), an = n == 0 ? 1. : 0.;
    const double* const* g = (const double* const*)H->gl[n];
    for (int s = 0; s < kss; ++s) {
        if (ns > 1) {
            orc_ghost_fill(dn, H->pr[n], g, cs, an, dt, nbc);
            sweep_slabs(H, dn, H->pr[n], rh, g, cs, an, dt, pot, omega, islor, ns);
        } else if (islor) {
            orc_slor_sweep(dn, H->pr[n], rh, g, cs, an, dt, pot, omega, nbc);
        } else {
            orc_ghost_fill(dn, H->pr[n], g, cs, an, dt, nbc);
            orc_sor_sweep(dn, H->pr[n], rh, g, pot, omega);
        }
    }
}

int orc_hier_multi(orc_hier* H, double* fi, const double* rhs, const double* const cs[6], double dt,
                   double omega, double eps, int ficycle, int kss1, int kssc, int nbc, int islor, int nslab,
                   double* resmax_out, double* res_hist) {
    const int L = H->L;
    const orc_dims* d = &H->d;
    const int ns = nslab > 1 ? nslab : 1;
    if (ns > 1) {
        if (d->nz % (ns << (L - 1)) != 0) return -1;
        if (H->nslab_alloc != ns) { /* buffers hold nz / ns + 2 planes */
            for (int s = 0; s < H->nslab_alloc; ++s) free(H->slab[s]);
            free(H->slab);
            const size_t sz = (size_t)(d->nx + 2) * (size_t)(d->ny + 2);
            H->slab = (double**)calloc((size_t)ns, sizeof(double*));
            for (int s = 0; s < ns; ++s) H->slab[s] = (double*)malloc((size_t)(d->nz / ns + 2) * sz * sizeof(double));
            H->nslab_alloc = ns;
        }
    }
    const size_t sz0 = gsize(d);
    double resmax = 1.;
    int ktime = 0;
    /* fsrc/multigrid.f03:174-637 */
    while (ktime < ficycle && resmax >= eps) {
        ++ktime;
        memcpy(H->pr[0], fi, sz0 * sizeof(double));
        for (int n = 1; n < L; ++n) memset(H->pr[n], 0, gsize(&H->dl[n]) * sizeof(double));
        /* down leg */
        for (int n = 0; n < L; ++n) {
            const double pot = (double)(1 << n);
            const double* rh = (n == 0) ? rhs : H->rv[n];
            smooth_level(H, n, rh, cs, dt, omega, n == 0 ? kss1 : kssc, nbc, islor, ns);
            if (n == 0 || n < L - 1) {
                orc_resid_mt(&H->dl[n], H->pr[n], rh, (const double* const*)H->gl[n], pot, H->rn[n], ns);
                if (n < L - 1) orc_restrict(&H->dl[n + 1], H->rn[n], H->rv[n + 1]);
            }
        }
        /* up leg */
        for (int n = L - 2; n >= 0; --n) {
            const double* rh = (n == 0) ? rhs : H->rv[n];
            orc_prolong(&H->dl[n + 1], H->pr[n + 1], H->pr[n]);
            smooth_level(H, n, rh, cs, dt, omega, n == 0 ? kss1 : kssc, nbc, islor, ns);
        }
        if (L > 1) orc_resid_mt(&H->dl[0], H->pr[0], rhs, (const double* const*)H->gl[0], 1., H->rn[0], ns);
        memcpy(fi, H->pr[0], sz0 * sizeof(double));
        resmax = orc_absmax(d, H->rn[0]);
        if (res_hist) res_hist[ktime - 1] = resmax;
    }
    if (resmax_out) *resmax_out = resmax;
    return ktime;
}

int orc_multi_ex(const orc_dims* d, int nlevel, const double* x, const double* y, const double* z,
                 const double* const g1[9], double* fi, const double* rhs, const double* const cs[6], double dt,
                 double omega, double eps, int ficycle, int kss1, int kssc, int nbc, int islor,
                 double* resmax_out, double* res_hist) {
    orc_hier* H = orc_hier_create(d, nlevel, x, y, z, g1);
    const int n = orc_hier_multi(H, fi, rhs, cs, dt, omega, eps, ficycle, kss1, kssc, nbc, islor, 1,
                                 resmax_out, res_hist);
    orc_hier_destroy(H);
    return n;
}

int orc_multi(const orc_dims* d, int nlevel, const double* x, const double* y, const double* z,
              const double* const g1[9], double* fi, const double* rhs, const double* const cs[6],
              double dt, double omega, double eps, int ficycle, int kss1, int kssc, int nbc,
              double* resmax_out, double* res_hist) {
    return orc_multi_ex(d, nlevel, x, y, z, g1, fi, rhs, cs, dt, omega, eps, ficycle, kss1, kssc, nbc, 0,
                        resmax_out, res_hist);
}

/* ------------------------------------------------------------------ */
/* projection step                                                     */
/* ------------------------------------------------------------------ */

void orc_contra(const orc_dims* d, const double* const csv[9], const double* u, const double* v,
                const double* w, double* uc, double* vc, double* wc) {
    const size_t n = gsize(d);
    memset(uc, 0, n * sizeof(double));
    memset(vc, 0, n * sizeof(double));
    memset(wc, 0, n * sizeof(double));
#define A(arr, i, j, k) arr[GIDX(d, i, j, k)]
    /* fsrc/contra.f03:73-87 (interior x-faces i = ip..nx-ip) */
    for (int k = 1; k <= d->nz; ++k)
        for (int j = 1; j <= d->ny; ++j)
            for (int i = d->ip; i <= d->nx - d->ip; ++i) {
                double ui = .5 * (A(u, i, j, k) + A(u, i + 1, j, k));
                double vi = .5 * (A(v, i, j, k) + A(v, i + 1, j, k));
                double wi = .5 * (A(w, i, j, k) + A(w, i + 1, j, k));
                A(uc, i, j, k) = ui * A(csv[0], i, j, k) + vi * A(csv[1], i, j, k) + wi * A(csv[2], i, j, k);
            }
    /* fsrc/contra.f03:132-146 */
    for (int k = 1; k <= d->nz; ++k)
        for (int j = d->jp; j <= d->ny - d->jp; ++j)
            for (int i = 1; i <= d->nx; ++i) {
                double ui = .5 * (A(u, i, j, k) + A(u, i, j + 1, k));
                double vi = .5 * (A(v, i, j, k) + A(v, i, j + 1, k));
                double wi = .5 * (A(w, i, j, k) + A(w, i, j + 1, k));
                A(vc, i, j, k) = ui * A(csv[3], i, j, k) + vi * A(csv[4], i, j, k) + wi * A(csv[5], i, j, k);
            }
    /* fsrc/contra.f03:204-220 */
    for (int k = d->kp; k <= d->nz - d->kp; ++k)
        for (int j = 1; j <= d->ny; ++j)
            for (int i = 1; i <= d->nx; ++i) {
                double ui = .5 * (A(u, i, j, k) + A(u, i, j, k + 1));
                double vi = .5 * (A(v, i, j, k) + A(v, i, j, k + 1));
                double wi = .5 * (A(w, i, j, k) + A(w, i, j, k + 1));
                A(wc, i, j, k) = ui * A(csv[6], i, j, k) + vi * A(csv[7], i, j, k) + wi * A(csv[8], i, j, k);
            }
}

void orc_diver(const orc_dims* d, const double* uc, const double* vc, const double* wc, double dt,
               double* rhs) {
    /* fsrc/diver.f03:20-28 */
    for (int k = 1; k <= d->nz; ++k)
        for (int j = 1; j <= d->ny; ++j)
            for (int i = 1; i <= d->nx; ++i)
                A(rhs, i, j, k) = (A(uc, i, j, k) - A(uc, i - 1, j, k) + A(vc, i, j, k) - A(vc, i, j - 1, k) +
                                   A(wc, i, j, k) - A(wc, i, j, k - 1)) /
                                  dt;
}

double orc_project(const orc_dims* d, double* rhs, const double* const cs[6], double dt) {
    double s = 0.0;
    for (int k = 1; k <= d->nz; ++k)
        for (int j = 1; j <= d->ny; ++j)
            for (int i = 1; i <= d->nx; ++i) s += A(rhs, i, j, k);
    double bsum = 0.0;
    if (cs) {
        if (d->ip == 1)
            for (int k = 0; k < d->nz; ++k)
                for (int j = 0; j < d->ny; ++j)
                    bsum += cs[1][(size_t)j + (size_t)d->ny * k] - cs[0][(size_t)j + (size_t)d->ny * k];
        if (d->jp == 1)
            for (int k = 0; k < d->nz; ++k)
                for (int i = 0; i < d->nx; ++i)
                    bsum += cs[3][(size_t)i + (size_t)d->nx * k] - cs[2][(size_t)i + (size_t)d->nx * k];
        if (d->kp == 1)
            for (int j = 0; j < d->ny; ++j)
                for (int i = 0; i < d->nx; ++i)
                    bsum += cs[5][(size_t)i + (size_t)d->nx * j] - cs[4][(size_t)i + (size_t)d->nx * j];
    }
    double mean = (s - bsum / dt) / ((double)d->nx * d->ny * d->nz);
    for (int k = 1; k <= d->nz; ++k)
        for (int j = 1; j <= d->ny; ++j)
            for (int i = 1; i <= d->nx; ++i) A(rhs, i, j, k) -= mean;
    return mean;
}

void orc_gradie(const orc_dims* d, const double* fi, const double* const csv[9], const double* giac,
                const double* const g[9], double dt, double* gra1, double* gra2, double* gra3,
                double* cgra1, double* cgra2, double* cgra3) {
    const double coef_dt = -0.5 * dt;
    const size_t n = gsize(d);
    memset(gra1, 0, n * sizeof(double));
    memset(gra2, 0, n * sizeof(double));
    memset(gra3, 0, n * sizeof(double));
    memset(cgra1, 0, n * sizeof(double));
    memset(cgra2, 0, n * sizeof(double));
    memset(cgra3, 0, n * sizeof(double));
    /* fsrc/gradie.f03:28-66 */
    for (int k = 1; k <= d->nz; ++k)
        for (int j = 1; j <= d->ny; ++j)
            for (int i = 1; i <= d->nx; ++i) {
                double coef = coef_dt / A(giac, i, j, k);
                double* out[3] = {gra1, gra2, gra3};
                for (int c = 0; c < 3; ++c) {
                    const double* cx = csv[c];
                    const double* ex = csv[3 + c];
                    const double* zx = csv[6 + c];
                    double s = A(cx, i, j, k) * (A(fi, i, j, k) + A(fi, i + 1, j, k)) -
                               A(cx, i - 1, j, k) * (A(fi, i, j, k) + A(fi, i - 1, j, k)) +
                               A(ex, i, j, k) * (A(fi, i, j, k) + A(fi, i, j + 1, k)) -
                               A(ex, i, j - 1, k) * (A(fi, i, j, k) + A(fi, i, j - 1, k)) +
                               A(zx, i, j, k) * (A(fi, i, j, k) + A(fi, i, j, k + 1)) -
                               A(zx, i, j, k - 1) * (A(fi, i, j, k) + A(fi, i, j, k - 1));
                    A(out[c], i, j, k) = coef * s;
                }
            }
    /* fsrc/gradie.f03:69-117 */
#define F(i, j, k) A(fi, i, j, k)
    for (int k = 1; k <= d->nz; ++k)
        for (int j = 1; j <= d->ny; ++j)
            for (int i = 0; i <= d->nx; ++i)
                A(cgra1, i, j, k) =
                    A(g[0], i, j, k) * (F(i + 1, j, k) - F(i, j, k)) +
                    .25 * A(g[1], i, j, k) * (F(i + 1, j + 1, k) + F(i, j + 1, k) - F(i + 1, j - 1, k) - F(i, j - 1, k)) +
                    .25 * A(g[2], i, j, k) * (F(i + 1, j, k + 1) + F(i, j, k + 1) - F(i + 1, j, k - 1) - F(i, j, k - 1));
    for (int k = 1; k <= d->nz; ++k)
        for (int j = 0; j <= d->ny; ++j)
            for (int i = 1; i <= d->nx; ++i)
                A(cgra2, i, j, k) =
                    A(g[4], i, j, k) * (F(i, j + 1, k) - F(i, j, k)) +
                    .25 * A(g[3], i, j, k) * (F(i + 1, j + 1, k) + F(i + 1, j, k) - F(i - 1, j + 1, k) - F(i - 1, j, k)) +
                    .25 * A(g[5], i, j, k) * (F(i, j + 1, k + 1) + F(i, j, k + 1) - F(i, j + 1, k - 1) - F(i, j, k - 1));
    for (int k = 0; k <= d->nz; ++k)
        for (int j = 1; j <= d->ny; ++j)
            for (int i = 1; i <= d->nx; ++i)
                A(cgra3, i, j, k) =
                    A(g[8], i, j, k) * (F(i, j, k + 1) - F(i, j, k)) +
                    .25 * A(g[6], i, j, k) * (F(i + 1, j, k + 1) + F(i + 1, j, k) - F(i - 1, j, k + 1) - F(i - 1, j, k)) +
                    .25 * A(g[7], i, j, k) * (F(i, j + 1, k + 1) + F(i, j + 1, k) - F(i, j - 1, k + 1) - F(i, j - 1, k));
#undef F
}

/* ------------------------------------------------------------------ */
/* time-loop neighbours of the solve: update, contra, vel_up, divergence */
/* ------------------------------------------------------------------ */

/* Parabolic extrapolation to the wall face from the first three cells. */
static double parab(double a1, double a2, double a3) { return 1.875 * a1 - 1.25 * a2 + .375 * a3; }

void orc_update(const orc_dims* d, double* const vel[3], const double* const del[3]) {
    /* fsrc/update.f03:29-38 interior increment */
    for (int c = 0; c < 3; ++c) {
        double* u = vel[c];
        const double* du = del[c];
        for (int k = 1; k <= d->nz; ++k)
            for (int j = 1; j <= d->ny; ++j)
                for (int i = 1; i <= d->nx; ++i) A(u, i, j, k) += A(du, i, j, k);
    }
    for (int c = 0; c < 3; ++c) {
        double* u = vel[c];
        const int nx = d->nx, ny = d->ny, nz = d->nz;
        /* fsrc/update.f03:43-91: xi sides, walls (ip=1) or periodic copies */
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j) {
                if (d->ip == 1) {
                    A(u, 0, j, k) = parab(A(u, 1, j, k), A(u, 2, j, k), A(u, 3, j, k));
                    A(u, nx + 1, j, k) = parab(A(u, nx, j, k), A(u, nx - 1, j, k), A(u, nx - 2, j, k));
                } else {
                    A(u, 0, j, k) = A(u, nx, j, k);
                    A(u, nx + 1, j, k) = A(u, 1, j, k);
                }
            }
        /* fsrc/update.f03:93-139: eta sides */
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i) {
                if (d->jp == 1) {
                    A(u, i, 0, k) = parab(A(u, i, 1, k), A(u, i, 2, k), A(u, i, 3, k));
                    A(u, i, ny + 1, k) = parab(A(u, i, ny, k), A(u, i, ny - 1, k), A(u, i, ny - 2, k));
                } else {
                    A(u, i, 0, k) = A(u, i, ny, k);
                    A(u, i, ny + 1, k) = A(u, i, 1, k);
                }
            }
        /* fsrc/update.f03:141-235: zeta sides (the periodic plane exchange
         * copies i = 1..nx, j = 1..ny) */
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                if (d->kp == 1) {
                    A(u, i, j, 0) = parab(A(u, i, j, 1), A(u, i, j, 2), A(u, i, j, 3));
                    A(u, i, j, nz + 1) = parab(A(u, i, j, nz), A(u, i, j, nz - 1), A(u, i, j, nz - 2));
                } else {
                    A(u, i, j, 0) = A(u, i, j, nz);
                    A(u, i, j, nz + 1) = A(u, i, j, 1);
                }
            }
    }
}

/* wall face flux from the ghost-cell velocity and the face area vector */
static double wall_flux(const orc_dims* d, const double* const vel[3], const double* const cv[3], int gi,
                        int gj, int gk, int fi, int fj, int fk) {
    return A(vel[0], gi, gj, gk) * A(cv[0], fi, fj, fk) + A(vel[1], gi, gj, gk) * A(cv[1], fi, fj, fk) +
           A(vel[2], gi, gj, gk) * A(cv[2], fi, fj, fk);
}

void orc_contra_step(const orc_dims* d, const double* const csv[9], const double* const vel[3],
                     double* const fc[3], double* const cs[6], int potenziale) {
    const int nx = d->nx, ny = d->ny, nz = d->nz;
    const double* cx[3] = {csv[0], csv[1], csv[2]};
    const double* ex[3] = {csv[3], csv[4], csv[5]};
    const double* zx[3] = {csv[6], csv[7], csv[8]};
    double *uc = fc[0], *vc = fc[1], *wc = fc[2];
    /* fsrc/contra.f03:35-65: xi walls; cs1/cs2 (ny,nz) */
    if (d->ip == 1)
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j) {
                const size_t o = (size_t)(j - 1) + (size_t)ny * (k - 1);
                double f0 = wall_flux(d, vel, cx, 0, j, k, 0, j, k);
                cs[0][o] = potenziale ? -A(uc, 0, j, k) : f0 - A(uc, 0, j, k);
                A(uc, 0, j, k) = f0;
                double f1 = wall_flux(d, vel, cx, nx + 1, j, k, nx, j, k);
                cs[1][o] = potenziale ? -A(uc, nx, j, k) : f1 - A(uc, nx, j, k);
                A(uc, nx, j, k) = f1;
            }
    /* fsrc/contra.f03:70-87 */
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = d->ip; i <= nx - d->ip; ++i)
                A(uc, i, j, k) = .5 * (A(vel[0], i, j, k) + A(vel[0], i + 1, j, k)) * A(cx[0], i, j, k) +
                                 .5 * (A(vel[1], i, j, k) + A(vel[1], i + 1, j, k)) * A(cx[1], i, j, k) +
                                 .5 * (A(vel[2], i, j, k) + A(vel[2], i + 1, j, k)) * A(cx[2], i, j, k);
    /* fsrc/contra.f03:95-127: eta walls; cs3/cs4 (nx,nz) */
    if (d->jp == 1)
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i) {
                const size_t o = (size_t)(i - 1) + (size_t)nx * (k - 1);
                double f0 = wall_flux(d, vel, ex, i, 0, k, i, 0, k);
                cs[2][o] = potenziale ? -A(vc, i, 0, k) : f0 - A(vc, i, 0, k);
                A(vc, i, 0, k) = f0;
                double f1 = wall_flux(d, vel, ex, i, ny + 1, k, i, ny, k);
                cs[3][o] = potenziale ? -A(vc, i, ny, k) : f1 - A(vc, i, ny, k);
                A(vc, i, ny, k) = f1;
            }
    /* fsrc/contra.f03:132-146 */
    for (int k = 1; k <= nz; ++k)
        for (int j = d->jp; j <= ny - d->jp; ++j)
            for (int i = 1; i <= nx; ++i)
                A(vc, i, j, k) = .5 * (A(vel[0], i, j, k) + A(vel[0], i, j + 1, k)) * A(ex[0], i, j, k) +
                                 .5 * (A(vel[1], i, j, k) + A(vel[1], i, j + 1, k)) * A(ex[1], i, j, k) +
                                 .5 * (A(vel[2], i, j, k) + A(vel[2], i, j + 1, k)) * A(ex[2], i, j, k);
    /* fsrc/contra.f03:152-194: zeta walls on the first / last slab; cs5/cs6 (nx,ny) */
    if (d->kp == 1)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const size_t o = (size_t)(i - 1) + (size_t)nx * (j - 1);
                double f0 = wall_flux(d, vel, zx, i, j, 0, i, j, 0);
                cs[4][o] = potenziale ? -A(wc, i, j, 0) : f0 - A(wc, i, j, 0);
                A(wc, i, j, 0) = f0;
                double f1 = wall_flux(d, vel, zx, i, j, nz + 1, i, j, nz);
                cs[5][o] = potenziale ? -A(wc, i, j, nz) : f1 - A(wc, i, j, nz);
                A(wc, i, j, nz) = f1;
            }
    /* fsrc/contra.f03:198-221 (k = kp..nz-kp over the slabs) */
    for (int k = d->kp; k <= nz - d->kp; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                A(wc, i, j, k) = .5 * (A(vel[0], i, j, k) + A(vel[0], i, j, k + 1)) * A(zx[0], i, j, k) +
                                 .5 * (A(vel[1], i, j, k) + A(vel[1], i, j, k + 1)) * A(zx[1], i, j, k) +
                                 .5 * (A(vel[2], i, j, k) + A(vel[2], i, j, k + 1)) * A(zx[2], i, j, k);
}

void orc_vel_up(const orc_dims* d, double* const vel[3], double* const fc[3], const double* const gra[3],
                const double* const cgra[3], double dt, int potenziale) {
    const int nx = d->nx, ny = d->ny, nz = d->nz;
    /* fsrc/vel_up.f03:39-80: cell velocities */
    for (int c = 0; c < 3; ++c)
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i)
                    A(vel[c], i, j, k) = (potenziale ? 0.0 : A(vel[c], i, j, k)) + A(gra[c], i, j, k);
    /* fsrc/vel_up.f03:85-101: periodic xi copies, j = 0..ny+1 */
    if (d->ip == 0)
        for (int c = 0; c < 3; ++c)
            for (int k = 1; k <= nz; ++k)
                for (int j = 0; j <= ny + 1; ++j) {
                    A(vel[c], 0, j, k) = A(vel[c], nx, j, k);
                    A(vel[c], nx + 1, j, k) = A(vel[c], 1, j, k);
                }
    /* fsrc/vel_up.f03:133-186: periodic zeta planes, i = 1..nx, j = 0..ny+1
     * (eta is never periodic for this solver) */
    if (d->kp == 0)
        for (int c = 0; c < 3; ++c)
            for (int j = 0; j <= ny + 1; ++j)
                for (int i = 1; i <= nx; ++i) {
                    A(vel[c], i, j, 0) = A(vel[c], i, j, nz);
                    A(vel[c], i, j, nz + 1) = A(vel[c], i, j, 1);
                }
    /* fsrc/vel_up.f03:191-289: contravariant fluxes on the non-wall faces */
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = d->ip; i <= nx - d->ip; ++i)
                A(fc[0], i, j, k) = (potenziale ? 0.0 : A(fc[0], i, j, k)) - dt * A(cgra[0], i, j, k);
    for (int k = 1; k <= nz; ++k)
        for (int j = d->jp; j <= ny - d->jp; ++j)
            for (int i = 1; i <= nx; ++i)
                A(fc[1], i, j, k) = (potenziale ? 0.0 : A(fc[1], i, j, k)) - dt * A(cgra[1], i, j, k);
    for (int k = d->kp; k <= nz - d->kp; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i)
                A(fc[2], i, j, k) = (potenziale ? 0.0 : A(fc[2], i, j, k)) - dt * A(cgra[2], i, j, k);
}

void orc_check_divergence(const orc_dims* d, const double* const fc[3], const double* giac, double out[4]) {
    /* fsrc/check_divergence.f03:131-190 (bodyforce = 0): max |div| and
     * max |div/giac|, first over i=2..nx-1, j=2..ny-1, k=2..nz-1 ("inside
     * field"), then over all cells */
    double m[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 1; k <= d->nz; ++k)
        for (int j = 1; j <= d->ny; ++j)
            for (int i = 1; i <= d->nx; ++i) {
                const double div = A(fc[0], i, j, k) - A(fc[0], i - 1, j, k) + A(fc[1], i, j, k) -
                                   A(fc[1], i, j - 1, k) + A(fc[2], i, j, k) - A(fc[2], i, j, k - 1);
                const double a = fabs(div), ag = fabs(div / A(giac, i, j, k));
                const int inside = i >= 2 && i <= d->nx - 1 && j >= 2 && j <= d->ny - 1 && k != 1 && k != d->nz;
                if (inside) {
                    if (a > m[0]) m[0] = a;
                    if (ag > m[1]) m[1] = ag;
                }
                if (a > m[2]) m[2] = a;
                if (ag > m[3]) m[3] = ag;
            }
    for (int q = 0; q < 4; ++q) out[q] = m[q];
}

/* ------------------------------------------------------------------ */
/* norms                                                               */
/* ------------------------------------------------------------------ */

double orc_absmax(const orc_dims* d, const double* r) {
    double m = 0.0;
    for (int k = 1; k <= d->nz; ++k)
        for (int j = 1; j <= d->ny; ++j)
            for (int i = 1; i <= d->nx; ++i) {
                double a = fabs(A(r, i, j, k));
                if (a > m) m = a;
            }
    return m;
}

double orc_l2(const orc_dims* d, const double* r) {
    double s = 0.0;
    for (int k = 1; k <= d->nz; ++k)
        for (int j = 1; j <= d->ny; ++j)
            for (int i = 1; i <= d->nx; ++i) s += A(r, i, j, k) * A(r, i, j, k);
    return sqrt(s);
}

double orc_mean(const orc_dims* d, const double* p) {
    double s = 0.0;
    for (int k = 1; k <= d->nz; ++k)
        for (int j = 1; j <= d->ny; ++j)
            for (int i = 1; i <= d->nx; ++i) s += A(p, i, j, k);
    return s / ((double)d->nx * d->ny * d->nz);
}
#undef A
