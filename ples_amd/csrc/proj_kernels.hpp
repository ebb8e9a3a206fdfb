// proj_kernels.hpp -- gfx950 kernels of the time-loop steps around the
// pressure solve: update (predictor increment + wall ghosts), contra
// (contravariant face fluxes and the wall flux changes cs1..cs6), vel_up
// (pressure correction of cell velocities and face fluxes) and the
// divergence check.  All arrays use the ghosted layout of mg_kernels.hpp;
// face fluxes are stored at the index of their low cell (xi-face i at
// idx(i,j,k)).  Every kernel is one elementwise pass over a (64 x 4) tile
// grid of the ghosted box, one zeta plane per grid row, so a wavefront reads
// and writes whole i-rows.
#pragma once
#include "mg_kernels.hpp"

namespace plesmg {

struct Vel3 {
    double* c[3];
};
struct CVel3 {
    const double* c[3];
};
struct CsOut {
    double* c[6];
};

// ghosted (i, j, k) of this thread in a (nx+2, ny+2, nz+2) box launch
__device__ inline bool box_cell(const Geo& G, int& i, int& j, int& k) {
    i = blockIdx.x * blockDim.x + threadIdx.x;
    j = blockIdx.y * blockDim.y + threadIdx.y;
    k = blockIdx.z;
    return i <= G.nx + 1 && j <= G.ny + 1 && k <= G.nz + 1;
}

__device__ inline bool interior(const Geo& G, int i, int j, int k) {
    return i >= 1 && i <= G.nx && j >= 1 && j <= G.ny && k >= 1 && k <= G.nz;
}

// fsrc/update.f03:29-38: vel += del in the interior
__global__ void __launch_bounds__(256) k_update_cells(Geo G, Vel3 V, CVel3 D) {
    int i, j, k;
    if (!box_cell(G, i, j, k) || !interior(G, i, j, k)) return;
    const long c = G.idx(i, j, k);
#pragma unroll
    for (int q = 0; q < 3; ++q) V.c[q][c] += D.c[q][c];
}

// parabola through the first three cells evaluated at the wall face
__device__ inline double parab(double a1, double a2, double a3) {
    return 1.875 * a1 - 1.25 * a2 + .375 * a3;
}

// fsrc/update.f03:41-235: ghost faces (edges excluded) of each direction from
// the interior -- parabolic wall extrapolation or periodic copies.  Every
// written position reads interior cells only, so the three directions run in
// one pass.
__global__ void __launch_bounds__(256) k_update_ghosts(Geo G, Vel3 V) {
    int i, j, k;
    if (!box_cell(G, i, j, k)) return;
    const bool ii = i >= 1 && i <= G.nx, jj = j >= 1 && j <= G.ny, kk = k >= 1 && k <= G.nz;
    long s;   // stride towards the interior
    int wall; // 1: extrapolate, 0: copy
    long src; // periodic source
    if ((i == 0 || i == G.nx + 1) && jj && kk) {
        s = (i == 0) ? 1 : -1;
        wall = G.ip;
        src = G.idx(i == 0 ? G.nx : 1, j, k);
    } else if ((j == 0 || j == G.ny + 1) && ii && kk) {
        s = (j == 0) ? G.sy : -G.sy;
        wall = G.jp;
        src = G.idx(i, j == 0 ? G.ny : 1, k);
    } else if ((k == 0 || k == G.nz + 1) && ii && jj) {
        s = (k == 0) ? G.sz : -G.sz;
        wall = G.kp;
        src = G.idx(i, j, k == 0 ? G.nz : 1);
    } else {
        return;
    }
    const long c = G.idx(i, j, k);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        double* u = V.c[q];
        u[c] = wall ? parab(u[c + s], u[c + 2 * s], u[c + 3 * s]) : u[src];
    }
}

// fsrc/contra.f03:35-221: contravariant face fluxes.  Non-wall faces take
// the two-cell average of the velocity dotted with the face vector; wall
// faces the ghost-cell velocity, and their flux change new - old (-old when
// potenziale = 1) goes to cs (reference shapes, first index fastest).
__global__ void __launch_bounds__(256) k_contra(Geo G, CsvP V, CVel3 U, Vel3 F, CsOut cs, int potenziale) {
    int i, j, k;
    if (!box_cell(G, i, j, k)) return;
    const long c = G.idx(i, j, k);
    const bool ii = i >= 1 && i <= G.nx, jj = j >= 1 && j <= G.ny, kk = k >= 1 && k <= G.nz;
    const int n[3] = {G.nx, G.ny, G.nz};
    const int idx3[3] = {i, j, k};
    const long st[3] = {1, G.sy, G.sz};
    const int wallf[3] = {G.ip, G.jp, G.kp};
    const bool other[3] = {jj && kk, ii && kk, ii && jj};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int f = idx3[d];
        if (f > n[d] || !other[d]) continue;
        const double* vx = V.v[3 * d];
        const double* vy = V.v[3 * d + 1];
        const double* vz = V.v[3 * d + 2];
        if (wallf[d] && (f == 0 || f == n[d])) {
            const long g = (f == 0) ? c : c + st[d];  // ghost cell of this wall face
            const double nf = U.c[0][g] * vx[c] + U.c[1][g] * vy[c] + U.c[2][g] * vz[c];
            const double old = F.c[d][c];
            long o;
            if (d == 0) o = (long)(j - 1) + (long)G.ny * (k - 1);
            else if (d == 1) o = (long)(i - 1) + (long)G.nx * (k - 1);
            else o = (long)(i - 1) + (long)G.nx * (j - 1);
            cs.c[2 * d + (f == 0 ? 0 : 1)][o] = potenziale ? -old : nf - old;
            F.c[d][c] = nf;
        } else {
            const long b = c + st[d];
            F.c[d][c] = .5 * (U.c[0][c] + U.c[0][b]) * vx[c] + .5 * (U.c[1][c] + U.c[1][b]) * vy[c] +
                        .5 * (U.c[2][c] + U.c[2][b]) * vz[c];
        }
    }
}

// fsrc/vel_up.f03:39-80, 188-289 (freesurface = 0): cell velocities += gra,
// non-wall face fluxes -= dt * cgra (potenziale = 1: replaced)
__global__ void __launch_bounds__(256) k_vel_up(Geo G, Vel3 U, Vel3 F, CVel3 gra, CVel3 cgra, double dt,
                                                int potenziale) {
    int i, j, k;
    if (!box_cell(G, i, j, k)) return;
    const long c = G.idx(i, j, k);
    const double keep = potenziale ? 0.0 : 1.0;
    if (interior(G, i, j, k)) {
#pragma unroll
        for (int q = 0; q < 3; ++q) U.c[q][c] = keep * U.c[q][c] + gra.c[q][c];
    }
    const bool ii = i >= 1 && i <= G.nx, jj = j >= 1 && j <= G.ny, kk = k >= 1 && k <= G.nz;
    if (i >= G.ip && i <= G.nx - G.ip && jj && kk) F.c[0][c] = keep * F.c[0][c] - dt * cgra.c[0][c];
    if (j >= G.jp && j <= G.ny - G.jp && ii && kk) F.c[1][c] = keep * F.c[1][c] - dt * cgra.c[1][c];
    if (k >= G.kp && k <= G.nz - G.kp && ii && jj) F.c[2][c] = keep * F.c[2][c] - dt * cgra.c[2][c];
}

// fsrc/vel_up.f03:85-186: periodic xi copies (j = 0..ny+1, interior k) and
// periodic zeta planes (interior i, j = 0..ny+1); both read interior i / k
// positions only, which this pass does not write
__global__ void __launch_bounds__(256) k_vel_up_images(Geo G, Vel3 U) {
    int i, j, k;
    if (!box_cell(G, i, j, k)) return;
    const bool ii = i >= 1 && i <= G.nx, kk = k >= 1 && k <= G.nz;
    long src;
    if (G.ip == 0 && (i == 0 || i == G.nx + 1) && kk) src = G.idx(i == 0 ? G.nx : 1, j, k);
    else if (G.kp == 0 && (k == 0 || k == G.nz + 1) && ii) src = G.idx(i, j, k == 0 ? G.nz : 1);
    else return;
    const long c = G.idx(i, j, k);
#pragma unroll
    for (int q = 0; q < 3; ++q) U.c[q][c] = U.c[q][src];
}

// fsrc/check_divergence.f03:131-190 (bodyforce = 0): per-block maxima of
// |div| and |div/giac| inside the field (i=2..nx-1, j=2..ny-1, k=2..nz-1) and
// over all cells; column-per-thread launch for block_reduce_store
__global__ void __launch_bounds__(256) k_divmax(Geo G, CVel3 F, const double* __restrict__ giac,
                                                double* partials, int nblocks) {
    int i, k;
    const bool act = column_of(G, i, k);
    double m[4] = {0, 0, 0, 0};
    if (act) {
        const bool ik = i >= 2 && i <= G.nx - 1 && k != 1 && k != G.nz;
        for (int j = 1; j <= G.ny; ++j) {
            const long c = G.idx(i, j, k);
            const double div = F.c[0][c] - F.c[0][c - 1] + F.c[1][c] - F.c[1][c - G.sy] + F.c[2][c] -
                               F.c[2][c - G.sz];
            const double a = fabs(div), ag = fabs(div / giac[c]);
            if (ik && j >= 2 && j <= G.ny - 1) {
                m[0] = fmax(m[0], a);
                m[1] = fmax(m[1], ag);
            }
            m[2] = fmax(m[2], a);
            m[3] = fmax(m[3], ag);
        }
    }
    const bool mx[4] = {true, true, true, true};
    block_reduce_store<4>(m, partials, nblocks, mx);
}

__global__ void k_fin_max(const double* __restrict__ part, int nb, int nv, double* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int v = 0; v < nv; ++v) {
        double s = 0;
        for (int q = 0; q < nb; ++q) s = fmax(s, part[(long)v * nb + q]);
        out[v] = s;
    }
}

}  // namespace plesmg
