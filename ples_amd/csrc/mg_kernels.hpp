// mg_kernels.hpp -- gfx950 kernels of the PLES pressure multigrid.
//
// Layout: every level stores cell arrays "ghosted", (nx+2)(ny+2)(nz+2),
// i fastest; face metrics use the same layout (face f of a cell row stored
// at cell index f).  One thread owns one eta-column (i,k) and streams up the
// column in j: the 19-point stencil then needs only nine column loads per
// step (rolling windows in registers), all loads of a wavefront are
// contiguous rows in i, and the eta-line solve of the smoother runs in the
// same thread without any cross-thread traffic.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace plesmg {

struct Geo {
    int nx, ny, nz;
    int ip, jp, kp;  // 1 = walls, 0 = periodic
    long sy, sz;     // strides of j and k
    long n;          // total ghosted cells
    __host__ __device__ long idx(int i, int j, int k) const {
        return (long)i + sy * (long)j + sz * (long)k;
    }
};

template <class TM>
struct Met {
    const TM* g[9];  // g11 g12 g13 | g21 g22 g23 | g31 g32 g33
};

// 16-bit smoother metrics of a level whose g13 and g31 vanish, packed per
// cell in 4-byte pairs so a row of the smoother issues fewer, wider metric
// loads than seven 2-byte ones (more bytes in flight per register):
// a[c] = (g11, g12), z[c] = (g32, g33), b[c] = (g21, g23), g22 apart.
using H2 = _Float16 __attribute__((ext_vector_type(2)));
struct MetPk {
    const H2* a;
    const H2* z;
    const H2* b;
    const _Float16* g22;
};

// metric q (0..8 as in Met) of cell c
template <class TM>
__host__ __device__ inline TM mg(const Met<TM>& M, int q, long c) {
    return M.g[q][c];
}
__device__ inline _Float16 mg(const MetPk& M, int q, long c) {
    switch (q) {
        case 0: return M.a[c].x;
        case 1: return M.a[c].y;
        case 3: return M.b[c].x;
        case 4: return M.g22[c];
        case 5: return M.b[c].y;
        case 7: return M.z[c].x;
        case 8: return M.z[c].y;
        default: return (_Float16)0;  // g13, g31
    }
}

// metric views read as 16-bit values by the smoother
template <class MV>
struct MVHalf {
    static constexpr bool value = false;
};
template <>
struct MVHalf<Met<_Float16>> {
    static constexpr bool value = true;
};
template <>
struct MVHalf<MetPk> {
    static constexpr bool value = true;
};

// ------------------------------------------------------------------------
// block reductions (wavefront 64, deterministic tree order)
// ------------------------------------------------------------------------
template <int NV>
__device__ inline void block_reduce_store(double (&v)[NV], double* partials, int nblocks_stride,
                                          const bool (&is_max)[NV]) {
    __shared__ double red[NV][16];
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    const int lane = tid & 63, wid = tid >> 6;
    const int nw = (blockDim.x * blockDim.y + 63) >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double a = v[q];
        for (int off = 32; off > 0; off >>= 1) {
            double b = __shfl_down(a, off, 64);
            a = is_max[q] ? fmax(a, b) : a + b;
        }
        if (lane == 0) red[q][wid] = a;
    }
    __syncthreads();
    if (tid == 0) {
        const long bid = blockIdx.x + (long)gridDim.x * blockIdx.y;
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double a = red[q][0];
            for (int w = 1; w < nw; ++w) a = is_max[q] ? fmax(a, red[q][w]) : a + red[q][w];
            partials[(long)q * nblocks_stride + bid] = a;
        }
    }
}

// ------------------------------------------------------------------------
// column-streaming 19-point stencil
// ------------------------------------------------------------------------
// Rolling state of the nine columns around (i,k) at rows j-1, j, j+1.
template <class T>
struct Win {
    T pm, p0, pp;        // (i  ,k  )
    T xmm, xm0, xmp;     // (i-1,k  )
    T xpm, xp0, xpp;     // (i+1,k  )
    T zmm, zm0, zmp;     // (i  ,k-1)
    T zpm, zp0, zpp;     // (i  ,k+1)
    T dmm, dpm, dmp, dpp;// (i-1,k-1) (i+1,k-1) (i-1,k+1) (i+1,k+1) at row j
};

template <class T, class P>
__host__ __device__ inline void win_init(Win<T>& w, const P* __restrict__ p, long c0, long sy, long sz) {
    // c0 = idx(i,0,k): rows 0 (j-1 of j=1) and 1 (j)
    w.pm = (T)p[c0];
    w.p0 = (T)p[c0 + sy];
    w.xmm = (T)p[c0 - 1];
    w.xm0 = (T)p[c0 - 1 + sy];
    w.xpm = (T)p[c0 + 1];
    w.xp0 = (T)p[c0 + 1 + sy];
    w.zmm = (T)p[c0 - sz];
    w.zm0 = (T)p[c0 - sz + sy];
    w.zpm = (T)p[c0 + sz];
    w.zp0 = (T)p[c0 + sz + sy];
}

// load row j+1 of the five face columns and row j of the four diagonals;
// c = idx(i,j,k).  X0: the diagonals only enter through g13 and g31, which
// vanish on the level, so they are not loaded.
template <bool X0 = false, class T, class P>
__host__ __device__ inline void win_load(Win<T>& w, const P* __restrict__ p, long c, long sy, long sz) {
    w.pp = (T)p[c + sy];
    w.xmp = (T)p[c - 1 + sy];
    w.xpp = (T)p[c + 1 + sy];
    w.zmp = (T)p[c - sz + sy];
    w.zpp = (T)p[c + sz + sy];
    if constexpr (!X0) {
        w.dmm = (T)p[c - 1 - sz];
        w.dpm = (T)p[c + 1 - sz];
        w.dmp = (T)p[c - 1 + sz];
        w.dpp = (T)p[c + 1 + sz];
    }
}

template <class T>
__host__ __device__ inline void win_shift(Win<T>& w) {
    w.pm = w.p0; w.p0 = w.pp;
    w.xmm = w.xm0; w.xm0 = w.xmp;
    w.xpm = w.xp0; w.xp0 = w.xpp;
    w.zmm = w.zm0; w.zm0 = w.zmp;
    w.zpm = w.zp0; w.zp0 = w.zpp;
}

// eta-face metrics of row j-1 carried between steps
template <class T>
struct G2 {
    T g21, g22, g23;
};

// s * sum of face-flux differences at (i,j,k), the reference expression
// of fsrc/resid_sndrcv.f03:85-147 (without the "- rh").  X0: the level's
// xi-zeta cross metrics g13, g31 are identically zero (orthogonal xi and
// zeta grid lines, as on terrain-following grids over x-z separable meshes),
// so they are not loaded.
// Face fluxes of cell (i,j,k) from the window and the metrics of its six
// faces (g: the cell's own high faces, h: the low xi and zeta faces, stored
// at c-1 and c-sz).  X0 drops the g13/g31 terms (and their diagonal values).
template <bool X0, class T>
__host__ __device__ inline T flux_xi_hi(const Win<T>& w, T g11, T g12, T g13, T inv_pot) {
    const T q = (T)0.25;
    if constexpr (X0) return (g11 * (w.xp0 - w.p0) + q * (g12 * (w.xpp + w.pp - w.xpm - w.pm))) * inv_pot;
    else
        return (g11 * (w.xp0 - w.p0) +
                q * (g12 * (w.xpp + w.pp - w.xpm - w.pm) + g13 * (w.dpp + w.zp0 - w.dpm - w.zm0))) *
               inv_pot;
}
template <bool X0, class T>
__host__ __device__ inline T flux_xi_lo(const Win<T>& w, T h11, T h12, T h13, T inv_pot) {
    const T q = (T)0.25;
    if constexpr (X0) return (h11 * (w.p0 - w.xm0) + q * (h12 * (w.pp + w.xmp - w.pm - w.xmm))) * inv_pot;
    else
        return (h11 * (w.p0 - w.xm0) +
                q * (h12 * (w.pp + w.xmp - w.pm - w.xmm) + h13 * (w.zp0 + w.dmp - w.zm0 - w.dmm))) *
               inv_pot;
}
template <class T>
__host__ __device__ inline T flux_eta_hi(const Win<T>& w, T g21, T g22, T g23, T inv_pot) {
    const T q = (T)0.25;
    return (g22 * (w.pp - w.p0) +
            q * (g21 * (w.xpp + w.xp0 - w.xmp - w.xm0) + g23 * (w.zpp + w.zp0 - w.zmp - w.zm0))) *
           inv_pot;
}
template <class T>
__host__ __device__ inline T flux_eta_lo(const Win<T>& w, T g21, T g22, T g23, T inv_pot) {
    const T q = (T)0.25;
    return (g22 * (w.p0 - w.pm) +
            q * (g21 * (w.xp0 + w.xpm - w.xm0 - w.xmm) + g23 * (w.zp0 + w.zpm - w.zm0 - w.zmm))) *
           inv_pot;
}
template <bool X0, class T>
__host__ __device__ inline T flux_zeta_hi(const Win<T>& w, T g31, T g32, T g33, T inv_pot) {
    const T q = (T)0.25;
    if constexpr (X0) return (g33 * (w.zp0 - w.p0) + q * (g32 * (w.zpp + w.pp - w.zpm - w.pm))) * inv_pot;
    else
        return (g33 * (w.zp0 - w.p0) +
                q * (g31 * (w.dpp + w.xp0 - w.dmp - w.xm0) + g32 * (w.zpp + w.pp - w.zpm - w.pm))) *
               inv_pot;
}
template <bool X0, class T>
__host__ __device__ inline T flux_zeta_lo(const Win<T>& w, T h31, T h32, T h33, T inv_pot) {
    const T q = (T)0.25;
    if constexpr (X0) return (h33 * (w.p0 - w.zm0) + q * (h32 * (w.pp + w.zmp - w.pm - w.zmm))) * inv_pot;
    else
        return (h33 * (w.p0 - w.zm0) +
                q * (h31 * (w.xp0 + w.dpm - w.xm0 - w.dmm) + h32 * (w.pp + w.zmp - w.pm - w.zmm))) *
               inv_pot;
}

template <bool X0 = false, class T, class MV>
__host__ __device__ inline T stencil(const Win<T>& w, const MV& M, long c, long sz, const G2<T>& gm,
                            const G2<T>& g0, T inv_pot) {
    const T g11 = (T)mg(M, 0, c), g12 = (T)mg(M, 1, c), g13 = X0 ? (T)0 : (T)mg(M, 2, c);
    const T h11 = (T)mg(M, 0, c - 1), h12 = (T)mg(M, 1, c - 1), h13 = X0 ? (T)0 : (T)mg(M, 2, c - 1);
    const T g31 = X0 ? (T)0 : (T)mg(M, 6, c), g32 = (T)mg(M, 7, c), g33 = (T)mg(M, 8, c);
    const T h31 = X0 ? (T)0 : (T)mg(M, 6, c - sz), h32 = (T)mg(M, 7, c - sz), h33 = (T)mg(M, 8, c - sz);
    const T res_dx = flux_xi_hi<X0>(w, g11, g12, g13, inv_pot);
    const T res_sn = flux_xi_lo<X0>(w, h11, h12, h13, inv_pot);
    const T res_sop = flux_eta_hi(w, g0.g21, g0.g22, g0.g23, inv_pot);
    const T res_sot = flux_eta_lo(w, gm.g21, gm.g22, gm.g23, inv_pot);
    const T res_av = flux_zeta_hi<X0>(w, g31, g32, g33, inv_pot);
    const T res_ind = flux_zeta_lo<X0>(w, h31, h32, h33, inv_pot);
    return (res_dx - res_sn + res_sop - res_sot + res_av - res_ind) * inv_pot;
}

template <class T, class MV>
__host__ __device__ inline G2<T> load_g2(const MV& M, long c) {
    return G2<T>{(T)mg(M, 3, c), (T)mg(M, 4, c), (T)mg(M, 5, c)};
}

__device__ inline bool column_of(const Geo& G, int& i, int& k) {
    i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    k = blockIdx.y * blockDim.y + threadIdx.y + 1;
    return i <= G.nx && k <= G.nz;
}

// interior cell of a one-thread-per-cell launch (grid nx/bx x ny/by x nz,
// rows along the block's x: coalesced)
__device__ inline bool cell_of(const Geo& G, long& c) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > G.nx || j > G.ny) return false;
    c = G.idx(i, j, (int)blockIdx.z + 1);
    return true;
}

// write value v of row j of column (i,k) also to its periodic images
template <class T>
__host__ __device__ inline void write_images(const Geo& G, T* __restrict__ p, int i, int j, int k, T v) {
    int ia[3], ka[3];
    int ni = 0, nk = 0;
    ia[ni++] = i;
    ka[nk++] = k;
    if (G.ip == 0) {
        if (i == G.nx) ia[ni++] = 0;
        if (i == 1) ia[ni++] = G.nx + 1;
    }
    if (G.kp == 0) {
        if (k == G.nz) ka[nk++] = 0;
        if (k == 1) ka[nk++] = G.nz + 1;
    }
    if (ni == 1 && nk == 1) return;
    for (int a = 0; a < ni; ++a)
        for (int b = 0; b < nk; ++b)
            if (a != 0 || b != 0) p[G.idx(ia[a], j, ka[b])] = v;
}

// ------------------------------------------------------------------------
// Column segments: the eta-columns are split into P equal segments, one
// thread per (column, segment); a block holds bx columns x P segments of one
// zeta plane k.  P = 1 gives one thread per column.
// ------------------------------------------------------------------------
struct Seg {
    int i, k, s, j0, j1;
};

// Blocks are dispatched round-robin over the 8 XCDs (each with its own L2):
// linear block b runs on XCD b % 8.  Remap so that XCD x works through the
// contiguous tile range [x T/8, (x+1) T/8) in order, i.e. neighbouring zeta
// planes (which share the k-1/k+1 stencil data) meet in the same L2.
__device__ inline void xcd_tile(int& bx, int& bk) {
    const int gx = gridDim.x, T = gridDim.x * gridDim.y;
    const int b = blockIdx.x + gx * blockIdx.y;
    int t = b;
    if ((T & 7) == 0) t = (b & 7) * (T >> 3) + (b >> 3);
    bx = t % gx;
    bk = t / gx;
}

__device__ inline bool seg_of(const Geo& G, int P, Seg& q) {
    int bxi, bk;
    xcd_tile(bxi, bk);
    q.i = bxi * blockDim.x + threadIdx.x + 1;
    q.k = bk + 1;
    q.s = threadIdx.y;
    const int m = G.ny / P;
    q.j0 = q.s * m + 1;
    q.j1 = q.j0 + m - 1;
    return q.i <= G.nx;
}

// Factors of the segmented eta-line solve, all metric-only (per level):
// invm, cpr: Thomas factorisation of the segment block M_s (the coupling to
// the neighbouring segments removed), v = M_s^-1 (a_j0 e_first) and
// w = M_s^-1 (c_j1 e_last), the responses to the neighbouring segments' end
// values.
struct LineF {
    const float* invm;
    const float* cpr;
    const float* v;
    const float* w;
    const _Float16* v16;  // 16-bit copies of v, w: read by the smoother with 16-bit metrics
    const _Float16* w16;
    const H2* vw16;       // the same pairs packed (v, w), read with packed metrics
};

// spike value of row c: the 16-bit copy for the 16-bit-metric smoother
template <bool HALF>
__device__ inline float spike(const float* a, const _Float16* a16, long c) {
    if constexpr (HALF) return (float)a16[c];
    else return a[c];
}

// ------------------------------------------------------------------------
// eta-line Jacobi smoother:  p_new = p + omega * M_line^{-1} (b - A p)
// M_line = tridiag(-s^2 g22(j-1), den_j, -s^2 g22(j)) over the whole column,
// den_j the s^2-scaled sum of the normal metrics of the non-wall faces.
// Partitioned solve: each segment solves its block locally (forward pass
// fused with the 19-point residual, storing e_j = p_j + omega d'_j + c'_j
// p_{j+1}; backward pass q_j = e_j - c'_j q_{j+1} = p_j + local solution),
// the 2P end values of a column are coupled through a block-tridiagonal 2x2
// system solved in LDS by every thread of the column, and a last pass adds
// the responses v, w to the neighbours' end values.  The column then writes
// its eta-wall ghosts (homogeneous flux condition, neighbours lagged) and
// its periodic images.  zero_in: the input iterate is identically zero.
// ------------------------------------------------------------------------
template <int P, int BX, class T>
__device__ inline void reduced_solve(T (&red)[8][P][BX], int tx, int s, T& uLm, T& uFp) {
    // unknowns x_t = (F_t, L_t): x_t + A_t x_{t-1} + C_t x_{t+1} = Y_t,
    // A_t = [[0, vF_t],[0, vL_t]], C_t = [[wF_t, 0],[wL_t, 0]].
    // Block Thomas by the column's first thread, in place in LDS:
    // red[0..1][t] (Y) -> G_t -> x_t, red[2..3][t] (vF, vL) and red[6..7][t]
    // -> D_t^{-1}; red[4..5][t] keep (wF, wL).
    if (s == 0) {
#pragma unroll 1
        for (int t = 0; t < P; ++t) {
            const T yF = red[0][t][tx], yL = red[1][t][tx];
            const T vF = red[2][t][tx], vL = red[3][t][tx];
            T d00 = 1, d01 = 0, d10 = 0, d11 = 1, g0 = yF, g1 = yL;
            if (t > 0) {
                const T i10 = red[6][t - 1][tx], i11 = red[7][t - 1][tx];
                const T m00 = vF * i10, m01 = vF * i11, m10 = vL * i10, m11 = vL * i11;
                const T wFp = red[4][t - 1][tx], wLp = red[5][t - 1][tx];
                const T gp0 = red[0][t - 1][tx], gp1 = red[1][t - 1][tx];
                d00 = 1 - (m00 * wFp + m01 * wLp);
                d10 = -(m10 * wFp + m11 * wLp);
                g0 = yF - (m00 * gp0 + m01 * gp1);
                g1 = yL - (m10 * gp0 + m11 * gp1);
            }
            const T det = d00 * d11 - d01 * d10;
            red[0][t][tx] = g0;
            red[1][t][tx] = g1;
            red[2][t][tx] = d11 / det;   // Dinv00
            red[3][t][tx] = -d01 / det;  // Dinv01
            red[6][t][tx] = -d10 / det;  // Dinv10
            red[7][t][tx] = d00 / det;   // Dinv11
        }
        T xF = 0;
#pragma unroll 1
        for (int t = P - 1; t >= 0; --t) {
            T g0 = red[0][t][tx], g1 = red[1][t][tx];
            if (t < P - 1) {
                g0 -= red[4][t][tx] * xF;
                g1 -= red[5][t][tx] * xF;
            }
            const T nF = red[2][t][tx] * g0 + red[3][t][tx] * g1;
            const T nL = red[6][t][tx] * g0 + red[7][t][tx] * g1;
            red[0][t][tx] = nF;
            red[1][t][tx] = nL;
            xF = nF;
        }
    }
    __syncthreads();
    uLm = (s > 0) ? red[1][s - 1][tx] : (T)0;
    uFp = (s < P - 1) ? red[0][s + 1][tx] : (T)0;
}

// lanes of a segmented-smoother block along xi for P segments
template <int P>
struct SegBX {
    // 256/P rounded down to a power of two, clamped to [32, 128]
    static constexpr int raw = 256 / P;
    static constexpr int pow2 = raw >= 128 ? 128 : (raw >= 64 ? 64 : 32);
    static constexpr int value = pow2;
};

// eta-wall ghosts of column (i,k) after its new values pn1 = p(i,1,k),
// pnn = p(i,ny,k): fsrc/multigrid.f03:1166-1186 with an = 0, neighbouring
// columns with their input (lagged) values; writes the ghosts' images too.
template <class T, class MV>
__device__ inline void yline_ghosts(const Geo& G, const MV& M, const T* __restrict__ pin,
                                    T* __restrict__ pout, int i, int k, bool lo_side, bool hi_side, T pn1,
                                    T pnn, int zero_in) {
    const T qq = (T)0.25;
    const long sy = G.sy, sz = G.sz;
    if (lo_side) {
        T lo = pn1;
        const long c0 = G.idx(i, 0, k);
        if (!zero_in) {
            const T inv = (T)1 / (T)mg(M, 4, c0);
            lo = pn1 + qq * inv *
                           ((T)mg(M, 3, c0) * ((T)pin[c0 + 1 + sy] + (T)pin[c0 + 1] - (T)pin[c0 - 1 + sy] -
                                             (T)pin[c0 - 1]) +
                            (T)mg(M, 5, c0) * ((T)pin[c0 + sy + sz] + (T)pin[c0 + sz] - (T)pin[c0 + sy - sz] -
                                             (T)pin[c0 - sz]));
        }
        pout[c0] = lo;
        write_images(G, pout, i, 0, k, lo);
    }
    if (hi_side) {
        T hi = pnn;
        const long ct = G.idx(i, G.ny, k);
        if (!zero_in) {
            const T invt = (T)1 / (T)mg(M, 4, ct);
            hi = pnn - qq * invt *
                           ((T)mg(M, 3, ct) * ((T)pin[ct + 1 + sy] + (T)pin[ct + 1] - (T)pin[ct - 1 + sy] -
                                             (T)pin[ct - 1]) +
                            (T)mg(M, 5, ct) * ((T)pin[ct + sy + sz] + (T)pin[ct + sz] - (T)pin[ct + sy - sz] -
                                             (T)pin[ct - sz]));
        }
        pout[ct + sy] = hi;
        write_images(G, pout, i, G.ny + 1, k, hi);
    }
}

// one whole column (a single segment), the local solve staged through pout
template <class T, class TM, class TR>
__device__ inline void yline_column(const Geo& G, const Met<TM>& M, const LineF& F, const T* __restrict__ pin,
                                    T* __restrict__ pout, const TR* __restrict__ b, T omega, T inv_pot,
                                    int zero_in, int i, int k) {
    const long sy = G.sy, sz = G.sz;
    const T s2 = inv_pot * inv_pot;
    const long cb = G.idx(i, 0, k);
    T dprev = 0;
    if (zero_in) {
        T g22m = (T)M.g[4][cb];
        for (int j = 1; j <= G.ny; ++j) {
            const long c = cb + sy * j;
            const T a = (j > 1) ? -s2 * g22m : (T)0;
            const T d = (-(T)b[c] - a * dprev) * (T)F.invm[c];
            pout[c] = omega * d;
            dprev = d;
            g22m = (T)M.g[4][c];
        }
    } else {
        Win<T> w;
        win_init(w, pin, cb, sy, sz);
        G2<T> gm = load_g2<T>(M, cb);
        for (int j = 1; j <= G.ny; ++j) {
            const long c = cb + sy * j;
            win_load(w, pin, c, sy, sz);
            const G2<T> g0 = load_g2<T>(M, c);
            const T r = stencil(w, M, c, sz, gm, g0, inv_pot) - (T)b[c];
            const T a = (j > 1) ? -s2 * gm.g22 : (T)0;
            const T d = (r - a * dprev) * (T)F.invm[c];
            pout[c] = w.p0 + omega * d + (T)F.cpr[c] * w.pp;
            dprev = d;
            gm = g0;
            win_shift(w);
        }
    }
    T qn = 0, pnn = 0;
    for (int j = G.ny; j >= 1; --j) {
        const long c = cb + sy * j;
        qn = pout[c] - (T)F.cpr[c] * qn;
        pout[c] = qn;
        write_images(G, pout, i, j, k, qn);
        if (j == G.ny) pnn = qn;
    }
    yline_ghosts(G, M, pin, pout, i, k, true, true, qn, pnn, zero_in);
}

// Coarsest level: nu sweeps in one launch of one workgroup (the level fits
// in a block's columns), ping-pong p0 -> p1 -> p0 ... with a block barrier
// between sweeps.  The first sweep starts from zero.
template <class T, class TM>
__device__ __forceinline__ void d_coarse(Geo G, Met<TM> M, LineF F, T* __restrict__ p0,
                                         T* __restrict__ p1, const T* __restrict__ b, T omega,
                                         T inv_pot, int nu) {
    const int ncol = G.nx * G.nz;
    for (int s = 0; s < nu; ++s) {
        const T* pin = (s & 1) ? p1 : p0;
        T* pout = (s & 1) ? p0 : p1;
        for (int t = threadIdx.x; t < ncol; t += blockDim.x)
            yline_column(G, M, F, pin, pout, b, omega, inv_pot, s == 0 ? 1 : 0, t % G.nx + 1, t / G.nx + 1);
        __syncthreads();
    }
}
template <class T, class TM>
__global__ void __launch_bounds__(1024) k_coarse(Geo G, Met<TM> M, LineF F, T* __restrict__ p0,
                                                 T* __restrict__ p1, const T* __restrict__ b, T omega,
                                                 T inv_pot, int nu) {
    d_coarse<T, TM>(G, M, F, p0, p1, b, omega, inv_pot, nu);
}


// Coarsest level held in LDS: the iterate pair, rhs, metrics and line
// factors are staged into the workgroup's LDS once, the nu sweeps run on LDS
// (no global-memory latency per row), the result is written back.  Layout of
// the dynamic LDS: 12 double arrays of n = ghosted cells (p0, p1, b, g11..g33),
// then 2 float arrays (invm, cpr).
template <class T>
__device__ __forceinline__ void d_coarse_lds(Geo G, Met<T> M, LineF F, T* __restrict__ pout_g,
                                             const T* __restrict__ b_g, T omega, T inv_pot, int nu) {
    extern __shared__ double smem[];
    const long n = G.n;
    T* base = reinterpret_cast<T*>(smem);
    T* p0 = base;
    T* p1 = base + n;
    T* bb = base + 2 * n;
    Met<T> ML;
    for (int q = 0; q < 9; ++q) ML.g[q] = base + (3 + q) * n;
    float* fl = reinterpret_cast<float*>(base + 12 * n);
    LineF FL{fl, fl + n, nullptr, nullptr, nullptr, nullptr};
    for (long t = threadIdx.x; t < n; t += blockDim.x) {
        p0[t] = 0;
        p1[t] = 0;
        bb[t] = b_g[t];
        for (int q = 0; q < 9; ++q) const_cast<T*>(ML.g[q])[t] = M.g[q][t];
        fl[t] = F.invm[t];
        fl[n + t] = F.cpr[t];
    }
    __syncthreads();
    const int ncol = G.nx * G.nz;
    for (int s = 0; s < nu; ++s) {
        const T* pin = (s & 1) ? p1 : p0;
        T* pout = (s & 1) ? p0 : p1;
        for (int t = threadIdx.x; t < ncol; t += blockDim.x)
            yline_column(G, ML, FL, pin, pout, bb, omega, inv_pot, s == 0 ? 1 : 0, t % G.nx + 1, t / G.nx + 1);
        __syncthreads();
    }
    const T* res = (nu & 1) ? p1 : p0;
    for (long t = threadIdx.x; t < n; t += blockDim.x) pout_g[t] = res[t];
}
template <class T>
__global__ void __launch_bounds__(256) k_coarse_lds(Geo G, Met<T> M, LineF F, T* __restrict__ pout_g,
                                                    const T* __restrict__ b_g, T omega, T inv_pot, int nu) {
    d_coarse_lds<T>(G, M, F, pout_g, b_g, omega, inv_pot, nu);
}


// Segmented smoother with the segment held in registers (M rows per segment
// known at compile time): forward pass fused with the 19-point residual,
// local backward substitution in registers, reduced 2x2 block system of the
// column in LDS, final pass adds the neighbour responses and writes once.
// FP32 variants whose coefficients fit 32 KB of LDS factor the segment's
// tridiagonal block on the fly from the metrics the stencil loads anyway
// (diagonal = s^2 * sum of the non-wall normal face metrics, off-diagonals
// -s^2 g22), keeping c'_j in LDS for the backward pass; the others read the
// precomputed factors (invm, cpr).
template <int P, int MR, class T>
struct YlineOtf {
    static constexpr int BX = SegBX<P>::value;
    static constexpr bool value = sizeof(T) == 4 && MR * P * BX * 4 <= 32768;
};

// minimum waves per SIMD of the FP32 smoother with FP32 / 16-bit metrics
// (register budget per lane: 512 / waves)
#ifndef PLES_YLINE_WPE_F32
#define PLES_YLINE_WPE_F32 4
#endif
#ifndef PLES_YLINE_WPE_F16
#define PLES_YLINE_WPE_F16 4
#endif
// minimum waves per SIMD of a smoother variant: the FP32 ones as above; the
// FP64 X0 one with 16 segments of 16 rows (the 512x256x512 fine level) 4,
// which holds it at 128 VGPRs (at 130 it ran at 3 waves per SIMD and 20%
// slower); the other FP64 ones unconstrained (pinning them to 4 spills)
template <class T, class MV, int P = 1, int MR = 0, bool X0 = false>
struct YlineWpe {
    static constexpr int value = sizeof(T) == 4 ? (MVHalf<MV>::value ? PLES_YLINE_WPE_F16 : PLES_YLINE_WPE_F32)
                                                : ((P == 16 && MR == 16 && X0) ? 4 : 1);
};
// Metrics of one cell row as the smoother reads them: the cell's high faces
// (g), its low xi and zeta faces (h: stored at c-1, k: stored at c-sz), kept
// in their storage type until used.  One row is loaded ahead of its use.
template <class MV, bool X0>
struct MRow;
template <class TM, bool X0>
struct MRow<Met<TM>, X0> {
    TM g11, g12, g13, h11, h12, h13, g21, g22, g23, g31, g32, g33, k31, k32, k33;
    __device__ void load(const Met<TM>& M, long c, long sz) {
        g11 = M.g[0][c]; g12 = M.g[1][c]; h11 = M.g[0][c - 1]; h12 = M.g[1][c - 1];
        g21 = M.g[3][c]; g22 = M.g[4][c]; g23 = M.g[5][c];
        g32 = M.g[7][c]; g33 = M.g[8][c]; k32 = M.g[7][c - sz]; k33 = M.g[8][c - sz];
        if constexpr (!X0) {
            g13 = M.g[2][c]; h13 = M.g[2][c - 1]; g31 = M.g[6][c]; k31 = M.g[6][c - sz];
        } else {
            g13 = h13 = g31 = k31 = (TM)0;
        }
    }
};
template <bool X0>
struct MRow<MetPk, X0> {
    H2 a, am, z, zm, b;
    _Float16 g22v;
    __device__ void load(const MetPk& M, long c, long sz) {
        a = M.a[c]; am = M.a[c - 1]; z = M.z[c]; zm = M.z[c - sz]; b = M.b[c]; g22v = M.g22[c];
    }
};
template <class T, class TM, bool X0>
__device__ inline void mrow_get(const MRow<Met<TM>, X0>& r, T (&o)[15]) {
    o[0] = (T)r.g11; o[1] = (T)r.g12; o[2] = (T)r.g13; o[3] = (T)r.h11; o[4] = (T)r.h12; o[5] = (T)r.h13;
    o[6] = (T)r.g21; o[7] = (T)r.g22; o[8] = (T)r.g23; o[9] = (T)r.g31; o[10] = (T)r.g32; o[11] = (T)r.g33;
    o[12] = (T)r.k31; o[13] = (T)r.k32; o[14] = (T)r.k33;
}
template <class T, bool X0>
__device__ inline void mrow_get(const MRow<MetPk, X0>& r, T (&o)[15]) {
    o[0] = (T)r.a.x; o[1] = (T)r.a.y; o[2] = (T)0; o[3] = (T)r.am.x; o[4] = (T)r.am.y; o[5] = (T)0;
    o[6] = (T)r.b.x; o[7] = (T)r.g22v; o[8] = (T)r.b.y; o[9] = (T)0; o[10] = (T)r.z.x; o[11] = (T)r.z.y;
    o[12] = (T)0; o[13] = (T)r.zm.x; o[14] = (T)r.zm.y;
}

// everything the forward pass reads for one row: iterate row j+1 of the five
// face columns (and row j of the diagonals without X0), rhs, metrics and, for
// the variants with stored factors, invm and c'
template <class T, class MV, class TR, bool X0, bool OTF>
struct RowIn {
    T pp, xmp, xpp, zmp, zpp, dmm, dpm, dmp, dpp;
    TR rhs;
    float invm;
    MRow<MV, X0> m;
    __device__ void load(const T* __restrict__ p, const TR* __restrict__ b, const MV& M, const LineF& F, long c,
                         long sy, long sz) {
        pp = p[c + sy]; xmp = p[c - 1 + sy]; xpp = p[c + 1 + sy]; zmp = p[c - sz + sy]; zpp = p[c + sz + sy];
        if constexpr (!X0) {
            dmm = p[c - 1 - sz]; dpm = p[c + 1 - sz]; dmp = p[c - 1 + sz]; dpp = p[c + 1 + sz];
        }
        rhs = b[c];
        if constexpr (!OTF) invm = F.invm[c];
        m.load(M, c, sz);
    }
};

// s * the face-flux differences of the next row of a column (the stencil
// of fsrc/resid_sndrcv.f03:85-147 without the "- rh"): the window is advanced
// to it with the loaded row, Flo holds the flux through its lower eta face
// (the upper-face flux of the row before) and is advanced; m gets the row's
// metrics.
template <bool X0, class T, class MV, class TR, bool OTF>
__device__ inline T row_apply(Win<T>& w, const RowIn<T, MV, TR, X0, OTF>& cur, T& Flo, T inv_pot, T (&m)[15]) {
    w.pp = cur.pp; w.xmp = cur.xmp; w.xpp = cur.xpp; w.zmp = cur.zmp; w.zpp = cur.zpp;
    if constexpr (!X0) {
        w.dmm = cur.dmm; w.dpm = cur.dpm; w.dmp = cur.dmp; w.dpp = cur.dpp;
    }
    mrow_get<T>(cur.m, m);
    const T Fhi = flux_eta_hi(w, m[6], m[7], m[8], inv_pot);
    const T ap = (flux_xi_hi<X0>(w, m[0], m[1], m[2], inv_pot) - flux_xi_lo<X0>(w, m[3], m[4], m[5], inv_pot) +
                  Fhi - Flo + flux_zeta_hi<X0>(w, m[9], m[10], m[11], inv_pot) -
                  flux_zeta_lo<X0>(w, m[12], m[13], m[14], inv_pot)) *
                 inv_pot;
    Flo = Fhi;
    return ap;
}

// reciprocal of a line pivot: the hardware reciprocal in FP32 (1 ulp: the
// smoother is a preconditioner), a division in FP64
template <class T>
__device__ inline T pivot_inv(T x) {
    if constexpr (sizeof(T) == 4) return __builtin_amdgcn_rcpf(x);
    else return (T)1 / x;
}

// spike pair (v_j, w_j) of row c in the type the variant reads
template <class MV>
struct SpikeRow {
    float v, w;
    __device__ void load(const LineF& F, long c) { v = F.v[c]; w = F.w[c]; }
    __device__ float vv() const { return v; }
    __device__ float ww() const { return w; }
};
template <>
struct SpikeRow<Met<_Float16>> {
    _Float16 v, w;
    __device__ void load(const LineF& F, long c) { v = F.v16[c]; w = F.w16[c]; }
    __device__ float vv() const { return (float)v; }
    __device__ float ww() const { return (float)w; }
};
template <>
struct SpikeRow<MetPk> {
    H2 vw;
    __device__ void load(const LineF& F, long c) { vw = F.vw16[c]; }
    __device__ float vv() const { return (float)vw.x; }
    __device__ float ww() const { return (float)vw.y; }
};

// loads of the next row issued before a row is computed: where the two rows'
// registers fit the variant's budget without spilling (FP32 with packed
// 16-bit metrics, up to 8 segments)
template <int P, int MR, class T, class MV>
struct YlinePf {
    static constexpr bool value = sizeof(T) == 4 && std::is_same<MV, MetPk>::value && P <= 8;
};

// MV: the metric view, Met<T>, Met<_Float16> or the packed MetPk (X0 levels).
// The forward pass keeps the loads of the next row in flight while a row is
// computed, carries the eta flux through a row's upper face into the next
// row (it is that row's lower-face flux), and the last pass loads all the
// segment's spikes before its first store, so no load of that pass waits on
// a store.
template <int P, int MR, class T, class MV, class TR, bool X0 = false>
__device__ __forceinline__ void d_yline(Geo G, MV M, LineF F, const T* __restrict__ pin, T* __restrict__ pout,
                                        const TR* __restrict__ b, T omega, T inv_pot, int zero_in) {
    constexpr int BX = SegBX<P>::value;
    constexpr bool OTF = YlineOtf<P, MR, T>::value;
    constexpr bool PF = YlinePf<P, MR, T, MV>::value;
    __shared__ T red[8][P][BX];
    __shared__ T cps[OTF ? MR : 1][P][BX];
    Seg q;
    const bool act = seg_of(G, P, q);
    constexpr int mr = MR;
    const int tx = threadIdx.x, s = q.s;
    const long sy = G.sy, sz = G.sz;
    const T s2 = inv_pot * inv_pot;
    const long cb = act ? G.idx(q.i, q.j0 - 1, q.k) : 0;  // row j0-1
    // wall faces are not part of the line operator's diagonal
    const T mxm = (G.ip == 1 && q.i == 1) ? (T)0 : (T)1, mxp = (G.ip == 1 && q.i == G.nx) ? (T)0 : (T)1;
    const T mzm = (G.kp == 1 && q.k == 1) ? (T)0 : (T)1, mzp = (G.kp == 1 && q.k == G.nz) ? (T)0 : (T)1;
    const T myf = (q.j0 == 1) ? (T)0 : (T)1, myl = (q.j1 == G.ny) ? (T)0 : (T)1;
    T e[MR];
    T yF = 0, yL = 0, vF = 0, vL = 0, wF = 0, wL = 0;
    T elast = 0;  // final value of the segment's last row
    if (act) {
        T dprev = 0, pfirst = 0, plast = 0, cprev = 0;
        // factor of row jj: 1/m and c'; g22m = g22(j-1), g22p = g22(j)
        auto factor = [&](int jj, T g11, T h11, T g33, T h33, T g22m, T g22p, T& inv) {
            const T den = s2 * (mxp * g11 + mxm * h11 + mzp * g33 + mzm * h33 +
                                ((jj == mr - 1) ? myl : (T)1) * g22p + ((jj == 0) ? myf : (T)1) * g22m);
            const T a = (jj > 0) ? -s2 * g22m : (T)0;
            const T cc = (jj < mr - 1) ? -s2 * g22p : (T)0;
            inv = pivot_inv(den - a * cprev);
            const T cp = cc * inv;
            cps[jj][s][tx] = cp;
            cprev = cp;
            return a;
        };
        if (zero_in) {
            T g22m = (T)mg(M, 4, cb);
#pragma unroll
            for (int jj = 0; jj < MR; ++jj) {
                const long c = cb + sy * (jj + 1);
                const T g22p = (T)mg(M, 4, c);
                T a, inv;
                if constexpr (OTF) {
                    a = factor(jj, (T)mg(M, 0, c), (T)mg(M, 0, c - 1), (T)mg(M, 8, c), (T)mg(M, 8, c - sz), g22m,
                               g22p, inv);
                } else {
                    a = (jj > 0) ? -s2 * g22m : (T)0;
                    inv = (T)F.invm[c];
                }
                const T d = (-(T)b[c] - a * dprev) * inv;
                e[jj] = omega * d;
                dprev = d;
                g22m = g22p;
            }
        } else {
            Win<T> w;
            win_init(w, pin, cb, sy, sz);
            pfirst = w.p0;
            // eta flux through the segment's lower face (row j0-1's eta-face metrics)
            T g22m = (T)mg(M, 4, cb);
            T Flo = flux_eta_lo(w, (T)mg(M, 3, cb), g22m, (T)mg(M, 5, cb), inv_pot);
            RowIn<T, MV, TR, X0, OTF> nx;
            if constexpr (PF) nx.load(pin, b, M, F, cb + sy, sy, sz);
#pragma unroll
            for (int jj = 0; jj < MR; ++jj) {
                const long c = cb + sy * (jj + 1);
                RowIn<T, MV, TR, X0, OTF> cur;
                if constexpr (PF) {
                    cur = nx;
                    if (jj + 1 < MR) nx.load(pin, b, M, F, c + sy, sy, sz);
                } else {
                    cur.load(pin, b, M, F, c, sy, sz);
                }
                T m[15];
                const T r = row_apply(w, cur, Flo, inv_pot, m) - (T)cur.rhs;
                T a, inv, cp;
                if constexpr (OTF) {
                    a = factor(jj, m[0], m[3], m[11], m[14], g22m, m[7], inv);
                    cp = cprev;
                } else {
                    a = (jj > 0) ? -s2 * g22m : (T)0;
                    inv = (T)cur.invm;
                    cp = (T)F.cpr[c];
                }
                const T d = (r - a * dprev) * inv;
                e[jj] = w.p0 + omega * d + cp * w.pp;
                dprev = d;
                plast = w.p0;
                g22m = m[7];
                win_shift(w);
            }
        }
        T qn = 0;
#pragma unroll
        for (int jj = MR - 1; jj >= 0; --jj) {
            T cp;
            if constexpr (OTF) cp = cps[jj][s][tx];
            else cp = (T)F.cpr[cb + sy * (jj + 1)];  // re-read: cache hit, saves registers
            qn = e[jj] - cp * qn;
            e[jj] = qn;
            if (jj == mr - 1) elast = qn;
        }
        yF = e[0] - pfirst;
        yL = elast - plast;
        if (P > 1) {
            SpikeRow<MV> sF, sL;
            sF.load(F, cb + sy);
            sL.load(F, cb + sy * mr);
            vF = (T)sF.vv();
            wF = (T)sF.ww();
            vL = (T)sL.vv();
            wL = (T)sL.ww();
        }
    }
    T uLm = 0, uFp = 0;
    if (P > 1) {
        red[0][s][tx] = yF;
        red[1][s][tx] = yL;
        red[2][s][tx] = vF;
        red[3][s][tx] = vL;
        red[4][s][tx] = wF;
        red[5][s][tx] = wL;
        __syncthreads();
        reduced_solve<P, BX, T>(red, tx, s, uLm, uFp);
    }
    if (!act) return;
    const int i = q.i, k = q.k;
    SpikeRow<MV> sv[MR];
    if (P > 1) {
#pragma unroll
        for (int jj = 0; jj < MR; ++jj) sv[jj].load(F, cb + sy * (jj + 1));
    }
    // periodic images of the column: (ia, k), (i, ka), (ia, ka) on the same row
    const int ia = (G.ip == 0) ? ((i == G.nx) ? 0 : (i == 1 ? G.nx + 1 : -1)) : -1;
    const int ka = (G.kp == 0) ? ((k == G.nz) ? 0 : (k == 1 ? G.nz + 1 : -1)) : -1;
    const long ox = (long)(ia - i), oz = (long)(ka - k) * sz;
    T pfin = 0;
#pragma unroll
    for (int jj = 0; jj < MR; ++jj) {
        const long c = cb + sy * (jj + 1);
        T pn = e[jj];
        if (P > 1) pn -= uLm * (T)sv[jj].vv() + uFp * (T)sv[jj].ww();
        e[jj] = pn;
        pout[c] = pn;
        if (ia >= 0) pout[c + ox] = pn;
        if (ka >= 0) pout[c + oz] = pn;
        if (ia >= 0 && ka >= 0) pout[c + ox + oz] = pn;
        if (jj == mr - 1) pfin = pn;
    }
    yline_ghosts(G, M, pin, pout, i, k, q.j0 == 1, q.j1 == G.ny, e[0], pfin, zero_in);
}
template <int P, int MR, class T, class MV, class TR, bool X0 = false>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(YlineWpe<T, MV, P, MR, X0>::value)))
k_yline(Geo G, MV M, LineF F, const T* __restrict__ pin, T* __restrict__ pout, const TR* __restrict__ b, T omega,
        T inv_pot, int zero_in) {
    d_yline<P, MR, T, MV, TR, X0>(G, M, F, pin, pout, b, omega, inv_pot, zero_in);
}


// Segmented smoother for segment lengths without a register variant: the
// local solve is staged through pout (e, then q), same algorithm.
template <int P, class T, class TM, class TR>
__device__ __forceinline__ void d_yline_staged(Geo G, Met<TM> M, LineF F, const T* __restrict__ pin,
                                               T* __restrict__ pout, const TR* __restrict__ b, T omega,
                                               T inv_pot, int zero_in) {
    constexpr int BX = SegBX<P>::value;
    __shared__ T red[8][P][BX];
    Seg q;
    const bool act = seg_of(G, P, q);
    const int tx = threadIdx.x, s = q.s;
    const long sy = G.sy, sz = G.sz;
    const T s2 = inv_pot * inv_pot;
    const long cb = act ? G.idx(q.i, q.j0 - 1, q.k) : 0;
    T yF = 0, yL = 0, vF = 0, vL = 0, wF = 0, wL = 0;
    if (act) {
        T dprev = 0, pfirst = 0, plast = 0;
        if (zero_in) {
            T g22m = (T)M.g[4][cb];
            for (int j = q.j0; j <= q.j1; ++j) {
                const long c = cb + sy * (j - q.j0 + 1);
                const T a = (j > q.j0) ? -s2 * g22m : (T)0;
                const T d = (-(T)b[c] - a * dprev) * (T)F.invm[c];
                pout[c] = omega * d;
                dprev = d;
                g22m = (T)M.g[4][c];
            }
        } else {
            Win<T> w;
            win_init(w, pin, cb, sy, sz);
            G2<T> gm = load_g2<T>(M, cb);
            pfirst = w.p0;
            for (int j = q.j0; j <= q.j1; ++j) {
                const long c = cb + sy * (j - q.j0 + 1);
                win_load(w, pin, c, sy, sz);
                const G2<T> g0 = load_g2<T>(M, c);
                const T r = stencil(w, M, c, sz, gm, g0, inv_pot) - (T)b[c];
                const T a = (j > q.j0) ? -s2 * gm.g22 : (T)0;
                const T d = (r - a * dprev) * (T)F.invm[c];
                pout[c] = w.p0 + omega * d + (T)F.cpr[c] * w.pp;
                dprev = d;
                plast = w.p0;
                gm = g0;
                win_shift(w);
            }
        }
        T qn = 0;
        for (int j = q.j1; j >= q.j0; --j) {
            const long c = cb + sy * (j - q.j0 + 1);
            qn = pout[c] - (T)F.cpr[c] * qn;
            pout[c] = qn;
            if (j == q.j1) yL = qn - plast;
        }
        yF = qn - pfirst;
        if (P > 1) {
            const long cF = cb + sy, cL = cb + sy * (q.j1 - q.j0 + 1);
            vF = (T)spike<sizeof(TM) == 2>(F.v, F.v16, cF);
            vL = (T)spike<sizeof(TM) == 2>(F.v, F.v16, cL);
            wF = (T)spike<sizeof(TM) == 2>(F.w, F.w16, cF);
            wL = (T)spike<sizeof(TM) == 2>(F.w, F.w16, cL);
        }
    }
    T uLm = 0, uFp = 0;
    if (P > 1) {
        red[0][s][tx] = yF;
        red[1][s][tx] = yL;
        red[2][s][tx] = vF;
        red[3][s][tx] = vL;
        red[4][s][tx] = wF;
        red[5][s][tx] = wL;
        __syncthreads();
        reduced_solve<P, BX, T>(red, tx, s, uLm, uFp);
    }
    if (!act) return;
    const int i = q.i, k = q.k;
    T pn1 = 0, pnn = 0;
    for (int j = q.j0; j <= q.j1; ++j) {
        const long c = cb + sy * (j - q.j0 + 1);
        T pn = pout[c];
        if (P > 1)
            pn -= uLm * (T)spike<sizeof(TM) == 2>(F.v, F.v16, c) + uFp * (T)spike<sizeof(TM) == 2>(F.w, F.w16, c);
        pout[c] = pn;
        write_images(G, pout, i, j, k, pn);
        if (j == q.j0) pn1 = pn;
        if (j == q.j1) pnn = pn;
    }
    yline_ghosts(G, M, pin, pout, i, k, q.j0 == 1, q.j1 == G.ny, pn1, pnn, zero_in);
}
template <int P, class T, class TM, class TR>
__global__ void __launch_bounds__(512) k_yline_staged(Geo G, Met<TM> M, LineF F, const T* __restrict__ pin,
                                                      T* __restrict__ pout, const TR* __restrict__ b, T omega,
                                                      T inv_pot, int zero_in) {
    d_yline_staged<P, T, TM, TR>(G, M, F, pin, pout, b, omega, inv_pot, zero_in);
}


// ------------------------------------------------------------------------
// Residual family (segmented columns).
//   MODE 0: r = A p - b                     (V-cycle residual)
//   MODE 1: r = b - A p, partial sums of r.r, max|r|, b.b
//   MODE 2: w = A p, partial sums of w.w, w.wp, r.w, r.wp
// ------------------------------------------------------------------------
template <int MODE, class T, class TM, class TR, class TO, class TV = double, bool X0 = false>
__device__ __forceinline__ void d_resid(Geo G, Met<TM> M, const T* __restrict__ p,
                                        const TR* __restrict__ b, TO* __restrict__ out,
                                        T inv_pot, const TV* __restrict__ rvec,
                                        const TV* __restrict__ wp, double* partials,
                                        int nblocks, int P) {
    Seg q;
    const bool act = seg_of(G, P, q);
    double acc[4] = {0, 0, 0, 0};
    if constexpr (MODE == 2 && std::is_same<T, TM>::value && std::is_same<T, TO>::value &&
                  std::is_same<T, TV>::value) {
        // the Krylov operator w = A p: the next row (iterate, metrics, r and
        // the previous direction's w) is loaded while a row is computed, the
        // eta flux carried between rows
        using RI = RowIn<T, Met<TM>, TV, X0, true>;  // its rhs field carries r
        const LineF F{};
        if (act) {
            const long sy = G.sy, sz = G.sz;
            const long cb = G.idx(q.i, q.j0 - 1, q.k);
            const int m = q.j1 - q.j0 + 1;
            Win<T> w;
            win_init(w, p, cb, sy, sz);
            T Flo = flux_eta_lo(w, (T)M.g[3][cb], (T)M.g[4][cb], (T)M.g[5][cb], inv_pot);
            RI nx;
            nx.load(p, rvec, M, F, cb + sy, sy, sz);
            TV wpn = wp ? wp[cb + sy] : (TV)0;
            for (int r = 0; r < m; ++r) {
                const long c = cb + sy * (r + 1);
                const RI cur = nx;
                const TV wpc = wpn;
                if (r + 1 < m) {
                    nx.load(p, rvec, M, F, c + sy, sy, sz);
                    if (wp) wpn = wp[c + sy];
                }
                T mm[15];
                const TO wo = (TO)row_apply(w, cur, Flo, inv_pot, mm);
                out[c] = wo;
                const double wv = (double)wo, rv = (double)cur.rhs;
                acc[0] += wv * wv;
                if (wp) {
                    acc[1] += wv * (double)wpc;
                    acc[3] += rv * (double)wpc;
                }
                acc[2] += rv * wv;
                win_shift(w);
            }
        }
        double v[4] = {acc[0], acc[1], acc[2], acc[3]};
        const bool mx[4] = {false, false, false, false};
        block_reduce_store<4>(v, partials, nblocks, mx);
        return;
    }
    if (act) {
        const long sy = G.sy, sz = G.sz;
        const long cb = G.idx(q.i, q.j0 - 1, q.k);
        Win<T> w;
        win_init(w, p, cb, sy, sz);
        G2<T> gm = load_g2<T>(M, cb);
        for (int j = q.j0; j <= q.j1; ++j) {
            const long c = cb + sy * (j - q.j0 + 1);
            win_load<X0>(w, p, c, sy, sz);
            const G2<T> g0 = load_g2<T>(M, c);
            const T ap = stencil<X0>(w, M, c, sz, gm, g0, inv_pot);
            if (MODE == 0) {
                out[c] = (TO)(ap - (T)b[c]);
            } else if (MODE == 1) {
                const double bb = (double)b[c];
                const double r = bb - (double)ap;
                out[c] = (TO)r;
                acc[0] += r * r;
                acc[1] = fmax(acc[1], fabs(r));
                acc[2] += bb * bb;
            } else {
                const TO wo = (TO)ap;
                out[c] = wo;
                const double wv = (double)wo;  // the dots see the stored w
                const double rv = (double)rvec[c];
                acc[0] += wv * wv;
                if (wp) {
                    const double wpv = (double)wp[c];
                    acc[1] += wv * wpv;
                    acc[3] += rv * wpv;
                }
                acc[2] += rv * wv;
            }
            gm = g0;
            win_shift(w);
        }
    }
    if (MODE == 1) {
        double v[3] = {acc[0], acc[1], acc[2]};
        const bool mx[3] = {false, true, false};
        block_reduce_store<3>(v, partials, nblocks, mx);
    } else if (MODE == 2) {
        double v[4] = {acc[0], acc[1], acc[2], acc[3]};
        const bool mx[4] = {false, false, false, false};
        block_reduce_store<4>(v, partials, nblocks, mx);
    }
}
template <int MODE, class T, class TM, class TR, class TO, class TV = double, bool X0 = false>
__global__ void __launch_bounds__(512) k_resid(Geo G, Met<TM> M, const T* __restrict__ p,
                                               const TR* __restrict__ b, TO* __restrict__ out,
                                               T inv_pot, const TV* __restrict__ rvec,
                                               const TV* __restrict__ wp, double* partials,
                                               int nblocks, int P) {
    d_resid<MODE, T, TM, TR, TO, TV, X0>(G, M, p, b, out, inv_pot, rvec, wp, partials, nblocks, P);
}


// V-cycle residual of a level restricted straight into the next coarser
// level's rhs, without storing the fine residual: rc(I,J,K) = -(1/8) * the
// sum of the 2x2x2 fine residuals r = A p - b (k_resid MODE 0 then
// k_restrict).  Each thread walks the two zeta planes 2K-1, 2K of its column
// together, one row pair at a time (two stencil windows), so the 2x2 (j, k)
// part of a child block stays in registers; the i pair is the neighbouring
// lane (bx even); fx = 1 (a coarse level with the fine xi count): -(1/4) of
// the four (j, k) children of the thread's own column.  Segments have an even
// number of rows and start at an odd row.  Block: bx lanes along i x P
// segments; XCD-aware block order.
// Rows of the fused residual-restriction with the next row of each plane
// loaded while a row is computed and the eta flux carried between rows (the
// FP64 kernel, whose loads-at-use form is scheduled by the compiler with one
// load in flight: 1.7x slower; and the FP32 one with the packed 16-bit
// metrics of the smoother, a preconditioner-side residual whose metric
// rounding, like the smoother's, only perturbs the coarse correction).
template <class T, class MV, class TR, class TC, bool X0>
__device__ __forceinline__ void resid_restrict_rows(const Geo& G, const MV& M, const T* __restrict__ p,
                                                    const TR* __restrict__ b, const Geo& Gc, TC* __restrict__ rc,
                                                    T inv_pot, int m, int i, int j0, int kc, int tx, bool act,
                                                    long cb0, long cb1, int fx) {
    using RI = RowIn<T, MV, TR, X0, true>;
    const LineF F{};
    const long sy = G.sy, sz = G.sz;
    Win<T> w0, w1;
    T F0 = 0, F1 = 0;
    RI n0, n1;
    if (act) {
        win_init(w0, p, cb0, sy, sz);
        win_init(w1, p, cb1, sy, sz);
        F0 = flux_eta_lo(w0, (T)mg(M, 3, cb0), (T)mg(M, 4, cb0), (T)mg(M, 5, cb0), inv_pot);
        F1 = flux_eta_lo(w1, (T)mg(M, 3, cb1), (T)mg(M, 4, cb1), (T)mg(M, 5, cb1), inv_pot);
        n0.load(p, b, M, F, cb0 + sy, sy, sz);
        n1.load(p, b, M, F, cb1 + sy, sy, sz);
    }
    for (int rp = 0; rp < m / 2; ++rp) {
        T v = 0;
        if (act) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int row = 2 * rp + h;
                const RI c0 = n0, c1 = n1;
                if (row + 1 < m) {
                    n0.load(p, b, M, F, cb0 + sy * (row + 2), sy, sz);
                    n1.load(p, b, M, F, cb1 + sy * (row + 2), sy, sz);
                }
                T mm[15];
                v += row_apply(w0, c0, F0, inv_pot, mm) - (T)c0.rhs;
                v += row_apply(w1, c1, F1, inv_pot, mm) - (T)c1.rhs;
                win_shift(w0);
                win_shift(w1);
            }
        }
        if (fx == 1) {
            if (act) rc[Gc.idx(i, (j0 + 1) / 2 + rp, kc)] = (TC)(-v * (T)0.25);
            continue;
        }
        v += __shfl_down(v, 1, 64);
        if (act && !(tx & 1)) rc[Gc.idx((i + 1) >> 1, (j0 + 1) / 2 + rp, kc)] = (TC)(-v * (T)0.125);
    }
}

template <class T, class TM, class TR, class TC, bool X0 = false>
__device__ __forceinline__ void d_resid_restrict(Geo G, Met<TM> M, const T* __restrict__ p,
                                                 const TR* __restrict__ b, Geo Gc, TC* __restrict__ rc,
                                                 T inv_pot, int P, int fx) {
    int bxi, bk;
    xcd_tile(bxi, bk);
    const int bx = blockDim.x, tx = threadIdx.x, s = threadIdx.y;
    const int m = G.ny / P;
    const int i = bxi * bx + tx + 1, kc = bk + 1, j0 = s * m + 1;
    const bool act = i <= G.nx;
    const long sy = G.sy, sz = G.sz;
    const long cb0 = act ? G.idx(i, j0 - 1, 2 * kc - 1) : 0, cb1 = cb0 + sz;
    if constexpr (sizeof(T) == 8) {
        resid_restrict_rows<T, Met<TM>, TR, TC, X0>(G, M, p, b, Gc, rc, inv_pot, m, i, j0, kc, tx, act, cb0, cb1,
                                                      fx);
        return;
    }
    Win<T> w0, w1;
    G2<T> gm0{}, gm1{};
    if (act) {
        win_init(w0, p, cb0, sy, sz);
        win_init(w1, p, cb1, sy, sz);
        gm0 = load_g2<T>(M, cb0);
        gm1 = load_g2<T>(M, cb1);
    }
    for (int rp = 0; rp < m / 2; ++rp) {
        T v = 0;
        if (act) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const long c0 = cb0 + sy * (2 * rp + h + 1), c1 = c0 + sz;
                win_load<X0>(w0, p, c0, sy, sz);
                win_load<X0>(w1, p, c1, sy, sz);
                const G2<T> g00 = load_g2<T>(M, c0), g01 = load_g2<T>(M, c1);
                v += stencil<X0>(w0, M, c0, sz, gm0, g00, inv_pot) - (T)b[c0];
                v += stencil<X0>(w1, M, c1, sz, gm1, g01, inv_pot) - (T)b[c1];
                gm0 = g00;
                gm1 = g01;
                win_shift(w0);
                win_shift(w1);
            }
        }
        if (fx == 1) {  // coarse level with the fine xi count: four children
            if (act) rc[Gc.idx(i, (j0 + 1) / 2 + rp, kc)] = (TC)(-v * (T)0.25);
            continue;
        }
        v += __shfl_down(v, 1, 64);
        if (act && !(tx & 1)) rc[Gc.idx((i + 1) >> 1, (j0 + 1) / 2 + rp, kc)] = (TC)(-v * (T)0.125);
    }
}
template <class T, class TM, class TR, class TC, bool X0 = false>
__global__ void __launch_bounds__(512) k_resid_restrict(Geo G, Met<TM> M, const T* __restrict__ p,
                                                        const TR* __restrict__ b, Geo Gc, TC* __restrict__ rc,
                                                        T inv_pot, int P, int fx) {
    d_resid_restrict<T, TM, TR, TC, X0>(G, M, p, b, Gc, rc, inv_pot, P, fx);
}
// FP32 fused residual-restriction reading the packed 16-bit metrics (levels
// without g13/g31 whose smoother reads them): same launch shape and XCD order
template <class TR, class TC>
__device__ __forceinline__ void d_resid_restrict_pk(Geo G, MetPk M, const float* __restrict__ p,
                                                    const TR* __restrict__ b, Geo Gc, TC* __restrict__ rc,
                                                    float inv_pot, int P, int fx) {
    int bxi, bk;
    xcd_tile(bxi, bk);
    const int bx = blockDim.x, tx = threadIdx.x, s = threadIdx.y;
    const int m = G.ny / P;
    const int i = bxi * bx + tx + 1, kc = bk + 1, j0 = s * m + 1;
    const bool act = i <= G.nx;
    const long cb0 = act ? G.idx(i, j0 - 1, 2 * kc - 1) : 0, cb1 = cb0 + G.sz;
    resid_restrict_rows<float, MetPk, TR, TC, true>(G, M, p, b, Gc, rc, inv_pot, m, i, j0, kc, tx, act, cb0, cb1,
                                                     fx);
}
template <class TR, class TC>
__global__ void __launch_bounds__(512) k_resid_restrict_pk(Geo G, MetPk M, const float* __restrict__ p,
                                                           const TR* __restrict__ b, Geo Gc, TC* __restrict__ rc,
                                                           float inv_pot, int P, int fx) {
    d_resid_restrict_pk<TR, TC>(G, M, p, b, Gc, rc, inv_pot, P, fx);
}


// ------------------------------------------------------------------------
// Outer-loop scalar logic on the device (no host round trip per iteration).
// ------------------------------------------------------------------------
// S_DONE: set by k_fin_update once the stop test (S_MODE 0: r.r < S_THR,
// 1: max|r| < S_THR) holds; later refinement steps then leave x and r alone,
// so the host may enqueue steps ahead of reading the residual.
// Device-driven iterative refinement (batched mixed solves) keeps its state
// in the outer set: S_RN (norm the inner rhs is scaled by), S_STALL and
// S_MOLD (refinements that did not halve the outer measure; the measure
// before), S_EPS, S_BN (|b|), S_MAXC, S_ABS (stop mode), S_TOTC (V-cycles
// applied), S_REF / S_NOREF (this step ends a refinement / it does not).
enum { S_RR = 0, S_RMAX, S_BB, S_WW, S_WWP, S_RW, S_RWP, S_WPWP, S_ALPHA, S_BETA, S_HAVEP, S_RR0,
       S_SUM, S_SUM2, S_DONE, S_THR, S_MODE, S_CYC, S_RN, S_STALL, S_MOLD, S_EPS, S_BN, S_MAXC, S_ABS,
       S_TOTC, S_REF, S_NOREF, S_COUNT = 32 };

// sum (or max) of nb partials per quantity, one block of 256 threads,
// fixed order: strided per-thread sums then a tree
template <int NV>
__device__ inline void fin_reduce(const double* __restrict__ part, int nb, const bool (&mx)[NV],
                                  double (&out)[NV]) {
    __shared__ double sh[NV][256];
    const int t = threadIdx.x;
    // four strided accumulators per quantity, so every thread has 4 NV loads
    // in flight instead of one dependent load at a time; fixed order
    double a[NV][4];
#pragma unroll
    for (int q = 0; q < NV; ++q)
#pragma unroll
        for (int k = 0; k < 4; ++k) a[q][k] = 0.0;
    int u = t;
    for (; u + 768 < nb; u += 1024)
#pragma unroll
        for (int q = 0; q < NV; ++q)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const double v = part[(long)q * nb + u + 256 * k];
                a[q][k] = mx[q] ? fmax(a[q][k], v) : a[q][k] + v;
            }
    for (; u < nb; u += 256)
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            const double v = part[(long)q * nb + u];
            a[q][0] = mx[q] ? fmax(a[q][0], v) : a[q][0] + v;
        }
#pragma unroll
    for (int q = 0; q < NV; ++q)
        sh[q][t] = mx[q] ? fmax(fmax(a[q][0], a[q][1]), fmax(a[q][2], a[q][3]))
                         : (a[q][0] + a[q][1]) + (a[q][2] + a[q][3]);
    __syncthreads();
    for (int wdt = 128; wdt > 0; wdt >>= 1) {
        if (t < wdt)
#pragma unroll
            for (int q = 0; q < NV; ++q) sh[q][t] = mx[q] ? fmax(sh[q][t], sh[q][t + wdt]) : sh[q][t] + sh[q][t + wdt];
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q] = sh[q][0];
}

// finalize the partial sums written by k_resid MODE 1 into S[0..2]
__global__ void __launch_bounds__(256) k_fin_resid(const double* __restrict__ part, int nb, double* S) {
    const bool mx[3] = {false, true, false};
    double o[3];
    fin_reduce<3>(part, nb, mx, o);
    if (threadIdx.x == 0) {
        S[S_RR] = o[0];
        S[S_RMAX] = o[1];
        S[S_BB] = o[2];
    }
}

// finalize k_resid MODE 2 sums and compute the GCR step (beta, alpha)
__device__ __forceinline__ void d_fin_apply(const double* __restrict__ part, int nb, double* S, int krylov) {
    const bool mx[4] = {false, false, false, false};
    double o[4];
    fin_reduce<4>(part, nb, mx, o);
    if (threadIdx.x != 0) return;
    if (S[S_DONE] != 0.0) {
        S[S_ALPHA] = 0.0;
        S[S_BETA] = 0.0;
        return;
    }
    const double ww = o[0], wwp = o[1], rw = o[2], rwp = o[3];
    double beta = 0.0;
    if (krylov && S[S_HAVEP] != 0.0 && S[S_WPWP] > 0.0) beta = wwp / S[S_WPWP];
    const double wwn = ww - 2.0 * beta * wwp + beta * beta * S[S_WPWP];
    const double rwn = rw - beta * rwp;
    const double alpha = krylov ? ((wwn > 0.0) ? rwn / wwn : 0.0) : 1.0;
    S[S_WW] = ww;
    S[S_WWP] = wwp;
    S[S_RW] = rw;
    S[S_RWP] = rwp;
    S[S_BETA] = beta;
    S[S_ALPHA] = alpha;
    S[S_WPWP] = wwn;
    S[S_HAVEP] = 1.0;
}
__global__ void __launch_bounds__(256) k_fin_apply(const double* __restrict__ part, int nb, double* S, int krylov) {
    d_fin_apply(part, nb, S, krylov);
}


// w <- w - beta wp ; z <- z - beta zp ; x += alpha z ; r -= alpha w ;
// partial sums of r.r and max|r| (interior cells, segmented columns)
template <class T>
__device__ __forceinline__ void d_gcr_update(Geo G, T* __restrict__ x, T* __restrict__ r,
                                             T* __restrict__ z, T* __restrict__ w,
                                             const T* __restrict__ zp,
                                             const T* __restrict__ wp, const double* S,
                                             double* partials, int nblocks, int P) {
    if (S[S_DONE] != 0.0) return;  // converged: uniform exit, partials unused
    Seg q;
    const bool act = seg_of(G, P, q);
    double acc[2] = {0, 0};
    if (act) {
        const double alpha = S[S_ALPHA], beta = S[S_BETA];
        const long cb = G.idx(q.i, q.j0 - 1, q.k);
        for (int j = q.j0; j <= q.j1; ++j) {
            const long c = cb + G.sy * (j - q.j0 + 1);
            T zv = z[c], wv = w[c];
            if (beta != 0.0) {
                zv = (T)(zv - beta * zp[c]);
                wv = (T)(wv - beta * wp[c]);
                z[c] = zv;
                w[c] = wv;
            }
            x[c] = (T)(x[c] + alpha * zv);
            const T rv = (T)(r[c] - alpha * wv);
            r[c] = rv;
            acc[0] += (double)rv * (double)rv;
            acc[1] = fmax(acc[1], fabs((double)rv));
        }
    }
    double v[2] = {acc[0], acc[1]};
    const bool mx[2] = {false, true};
    block_reduce_store<2>(v, partials, nblocks, mx);
}
template <class T>
__global__ void __launch_bounds__(512) k_gcr_update(Geo G, T* __restrict__ x, T* __restrict__ r,
                                                    T* __restrict__ z, T* __restrict__ w,
                                                    const T* __restrict__ zp,
                                                    const T* __restrict__ wp, const double* S,
                                                    double* partials, int nblocks, int P) {
    d_gcr_update<T>(G, x, r, z, w, zp, wp, S, partials, nblocks, P);
}


__device__ __forceinline__ void d_fin_update(const double* __restrict__ part, int nb, double* S) {
    if (S[S_DONE] != 0.0) return;  // uniform: r.r and max|r| stay those of the converged step
    const bool mx[2] = {false, true};
    double o[2];
    fin_reduce<2>(part, nb, mx, o);
    if (threadIdx.x == 0) {
        S[S_RR] = o[0];
        S[S_RMAX] = o[1];
        const double m = (S[S_MODE] != 0.0) ? o[1] : o[0];
        S[S_CYC] += 1.0;  // steps that changed x since the stop test was armed
        if (m < S[S_THR]) S[S_DONE] = 1.0;
    }
}
__global__ void __launch_bounds__(256) k_fin_update(const double* __restrict__ part, int nb, double* S) {
    d_fin_update(part, nb, S);
}


// copy the scalars of n solves into one buffer (one read-back per batch)
struct SPtrs {
    const double* const* s;
};
__global__ void k_gather_S(const double* const* __restrict__ S, int n, double* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * S_COUNT) return;
    out[t] = S[t / S_COUNT][t % S_COUNT];
}

// ------------------------------------------------------------------------
// Transfers
// ------------------------------------------------------------------------
// rc = -(1/8) * sum of the eight children (fsrc/restrict_sndrcv.f03:32-53);
// a coarse level that keeps the xi (fx = 1) or the xi and eta counts (fx = fy
// = 1) averages its four or two children the same way
template <class TF, class TC>
__device__ __forceinline__ void d_restrict(Geo Gc, Geo Gf, const TF* __restrict__ rf,
                                           TC* __restrict__ rc, int fx, int fy, int gx) {
    // grid x = gx blocks per coarse row x planes (the plane index folded into x)
    const int i = (blockIdx.x % gx) * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.x / gx + 1;
    if (i > Gc.nx || j > Gc.ny) return;
    const long sy = Gf.sy, sz = Gf.sz;
    if (fx == 1 && fy == 1) {
        const long f = Gf.idx(i, j, 2 * k);
        rc[Gc.idx(i, j, k)] = (TC)(-(rf[f] + rf[f - sz]) * (TF)0.5);
        return;
    }
    if (fx == 1) {
        const long f = Gf.idx(i, 2 * j, 2 * k);
        const TF s = rf[f] + rf[f - sy] + rf[f - sz] + rf[f - sy - sz];
        rc[Gc.idx(i, j, k)] = (TC)(-s * (TF)0.25);
        return;
    }
    const long f = Gf.idx(2 * i, 2 * j, 2 * k);
    TF s = rf[f] + rf[f - 1] + rf[f - sy] + rf[f - 1 - sy] + rf[f - sz] + rf[f - 1 - sz] +
           rf[f - sy - sz] + rf[f - 1 - sy - sz];
    rc[Gc.idx(i, j, k)] = (TC)(-s * (TF)0.125);
}
template <class TF, class TC>
__global__ void __launch_bounds__(256) k_restrict(Geo Gc, Geo Gf, const TF* __restrict__ rf,
                                                  TC* __restrict__ rc, int fx, int fy, int gx) {
    d_restrict<TF, TC>(Gc, Gf, rf, rc, fx, fy, gx);
}


// pf += trilinear cell-centred interpolation of ec over the ghost-inclusive
// fine range (fsrc/prolong.f03:64-261): weight 3 per axis for the coarse cell
// on the fine child's side, 1 for the other, over 64.  Fine cells 2I, 2I+1
// (and likewise in j, k) take their values from the same coarse cells I, I+1,
// so one thread per (I, J, K), I in 0..ncx (ghosts included), loads that
// 2x2x2 coarse block once and updates its 8 fine children, two consecutive i
// at a time (the fine strides are even, so the pair is aligned).
template <class TC, class TF>
__device__ __forceinline__ void d_prolong_blk(Geo Gc, Geo Gf, const TC* __restrict__ ec,
                                              TF* __restrict__ pf, int gx) {
    // grid x = gx blocks per coarse row x planes (the plane index folded into x)
    const int I = (blockIdx.x % gx) * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    const int Kc = blockIdx.x / gx;
    if (I > Gc.nx || J > Gc.ny) return;
    const long c = Gc.idx(I, J, Kc);
    const long sy = Gc.sy, sz = Gc.sz;
    TF e[2][2][2];  // [k][j][i] offsets
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int bb = 0; bb < 2; ++bb)
#pragma unroll
            for (int d = 0; d < 2; ++d) e[a][bb][d] = (TF)ec[c + d + sy * bb + sz * a];
    using V2 = TF __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int dk = 0; dk < 2; ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj) {
            // weights 3 for the coarse cell on the fine child's side, 1 for the other
            const TF wk0 = dk ? (TF)1 : (TF)3, wk1 = dk ? (TF)3 : (TF)1;
            const TF wj0 = dj ? (TF)1 : (TF)3, wj1 = dj ? (TF)3 : (TF)1;
            TF s2[2];
#pragma unroll
            for (int d = 0; d < 2; ++d)
                s2[d] = wk0 * (wj0 * e[0][0][d] + wj1 * e[0][1][d]) + wk1 * (wj0 * e[1][0][d] + wj1 * e[1][1][d]);
            // fine i = 2I: 3 * (coarse I) + 1 * (I+1); fine 2I+1: 1 * I + 3 * (I+1)
            V2* f = (V2*)(pf + Gf.idx(2 * I, 2 * J + dj, 2 * Kc + dk));
            V2 v = *f;
            v.x += ((TF)3 * s2[0] + s2[1]) * (TF)(1.0 / 64.0);
            v.y += (s2[0] + (TF)3 * s2[1]) * (TF)(1.0 / 64.0);
            *f = v;
        }
}
template <class TC, class TF>
__global__ void __launch_bounds__(256) k_prolong_blk(Geo Gc, Geo Gf, const TC* __restrict__ ec,
                                                     TF* __restrict__ pf, int gx) {
    d_prolong_blk<TC, TF>(Gc, Gf, ec, pf, gx);
}


// pf += bilinear (eta, zeta) interpolation of ec for a coarse level with the
// fine xi count: the prolongation above without its xi factor (weights 3 and
// 1 per axis over 16, the fine xi index taking the coarse one's value).  One
// thread per coarse (i, J, K), i in 0..nx+1, J in 0..ncy, K in 0..ncz (ghosts
// included), updating its four fine children.
template <class TC, class TF>
__device__ __forceinline__ void d_prolong_yz(Geo Gc, Geo Gf, const TC* __restrict__ ec,
                                             TF* __restrict__ pf, int gx) {
    // grid x = gx blocks per coarse row x planes (the plane index folded into x)
    const int I = (blockIdx.x % gx) * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    const int Kc = blockIdx.x / gx;
    if (I > Gc.nx + 1 || J > Gc.ny) return;
    const long c = Gc.idx(I, J, Kc);
    const long sy = Gc.sy, sz = Gc.sz;
    const TF e00 = (TF)ec[c], e01 = (TF)ec[c + sy], e10 = (TF)ec[c + sz], e11 = (TF)ec[c + sy + sz];  // [k][j]
#pragma unroll
    for (int dk = 0; dk < 2; ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj) {
            const TF wk0 = dk ? (TF)1 : (TF)3, wk1 = dk ? (TF)3 : (TF)1;
            const TF wj0 = dj ? (TF)1 : (TF)3, wj1 = dj ? (TF)3 : (TF)1;
            pf[Gf.idx(I, 2 * J + dj, 2 * Kc + dk)] +=
                (wk0 * (wj0 * e00 + wj1 * e01) + wk1 * (wj0 * e10 + wj1 * e11)) * (TF)(1.0 / 16.0);
        }
}
template <class TC, class TF>
__global__ void __launch_bounds__(256) k_prolong_yz(Geo Gc, Geo Gf, const TC* __restrict__ ec,
                                                    TF* __restrict__ pf, int gx) {
    d_prolong_yz<TC, TF>(Gc, Gf, ec, pf, gx);
}


// pf += linear zeta interpolation of ec for a coarse level with the fine xi
// and eta counts (weights 3 and 1 over 4).  One thread per coarse (i, j, K),
// i in 0..nx+1, j in 0..ny+1, K in 0..ncz (ghosts included), updating its two
// fine children (i, j, 2K) and (i, j, 2K+1).
template <class TC, class TF>
__device__ __forceinline__ void d_prolong_z(Geo Gc, Geo Gf, const TC* __restrict__ ec,
                                            TF* __restrict__ pf, int gx) {
    // grid x = gx blocks per coarse row x planes (the plane index folded into x)
    const int I = (blockIdx.x % gx) * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    const int Kc = blockIdx.x / gx;
    if (I > Gc.nx + 1 || J > Gc.ny + 1) return;
    const long c = Gc.idx(I, J, Kc);
    const TF e0 = (TF)ec[c], e1 = (TF)ec[c + Gc.sz];
    pf[Gf.idx(I, J, 2 * Kc)] += ((TF)3 * e0 + e1) * (TF)0.25;
    pf[Gf.idx(I, J, 2 * Kc + 1)] += (e0 + (TF)3 * e1) * (TF)0.25;
}
template <class TC, class TF>
__global__ void __launch_bounds__(256) k_prolong_z(Geo Gc, Geo Gf, const TC* __restrict__ ec,
                                                   TF* __restrict__ pf, int gx) {
    d_prolong_z<TC, TF>(Gc, Gf, ec, pf, gx);
}


// ------------------------------------------------------------------------
// Exact ghost fill (outer FP64 operator, walls in xi/zeta): face relations
// of fsrc/multigrid.f03:1113-1317 solved by red/black passes over each wall
// face, neighbours along periodic directions read through the wrap, then
// every non-face ghost (periodic images, edges, corners) filled from its
// source by the edge rules of fsrc/multigrid.f03:1358-1742.
// ------------------------------------------------------------------------
struct BCs {
    const double* cs[6];  // cs1..cs6 (may be null)
    double an_over_dt;    // an / dt
};

__host__ __device__ inline int wrapi(const Geo& G, int i) {
    if (G.ip == 0) {
        if (i == 0) return G.nx;
        if (i == G.nx + 1) return 1;
    }
    return i;
}
__host__ __device__ inline int wrapk(const Geo& G, int k) {
    if (G.kp == 0) {
        if (k == 0) return G.nz;
        if (k == G.nz + 1) return 1;
    }
    return k;
}

// one launch per face set (0 eta, 1 xi, 2 zeta) and colour, so that no
// ghost read in a launch is written by the same launch
template <class T, class TM>
__host__ __device__ inline void halo_face_point(const Geo& G, const Met<TM>& M, T* __restrict__ p, long t, int color,
                                       const BCs& bc, int set, int images) {
    if (set == 1) t += 2L * G.nx * G.nz;
    if (set == 2) t += 2L * G.nx * G.nz + (G.ip == 1 ? 2L * G.ny * G.nz : 0L);
    const double an = bc.an_over_dt;
    const T q = (T)0.25;
#define PV(a, b, c) ((T)p[G.idx(wrapi(G, a), b, wrapk(G, c))])
    // eta faces (jp = 1 required)
    {
        const long nf = 2L * G.nx * G.nz;
        if (t < nf) {
            const int side = (int)(t / ((long)G.nx * G.nz));
            const long rem = t % ((long)G.nx * G.nz);
            const int i = (int)(rem % G.nx) + 1, k = (int)(rem / G.nx) + 1;
            if (((i + k) & 1) != color) return;
            if (side == 0) {
                const long cb = G.idx(i, 0, k);
                const T inv = (T)1 / (T)M.g[4][cb];
                const T csv = bc.cs[2] ? (T)bc.cs[2][(i - 1) + (long)G.nx * (k - 1)] : (T)0;
                const T val = PV(i, 1, k) - (T)an * csv * inv +
                        q * inv *
                            ((T)M.g[3][cb] * (PV(i + 1, 1, k) + PV(i + 1, 0, k) - PV(i - 1, 1, k) - PV(i - 1, 0, k)) +
                             (T)M.g[5][cb] * (PV(i, 1, k + 1) + PV(i, 0, k + 1) - PV(i, 1, k - 1) - PV(i, 0, k - 1)));
                p[cb] = val;
                if (images) write_images(G, p, i, 0, k, val);
            } else {
                const int ny = G.ny;
                const long ct = G.idx(i, ny, k);
                const T inv = (T)1 / (T)M.g[4][ct];
                const T csv = bc.cs[3] ? (T)bc.cs[3][(i - 1) + (long)G.nx * (k - 1)] : (T)0;
                const T val = PV(i, ny, k) + (T)an * csv * inv -
                               q * inv *
                                   ((T)M.g[3][ct] * (PV(i + 1, ny + 1, k) + PV(i + 1, ny, k) - PV(i - 1, ny + 1, k) -
                                                     PV(i - 1, ny, k)) +
                                    (T)M.g[5][ct] * (PV(i, ny + 1, k + 1) + PV(i, ny, k + 1) - PV(i, ny + 1, k - 1) -
                                                     PV(i, ny, k - 1)));
                p[ct + G.sy] = val;
                if (images) write_images(G, p, i, ny + 1, k, val);
            }
            return;
        }
        if (set == 0) return;
        t -= nf;
    }
    if (G.ip == 1) {
        const long nf = 2L * G.ny * G.nz;
        if (t < nf) {
            const int side = (int)(t / ((long)G.ny * G.nz));
            const long rem = t % ((long)G.ny * G.nz);
            const int j = (int)(rem % G.ny) + 1, k = (int)(rem / G.ny) + 1;
            if (((j + k) & 1) != color) return;
            if (side == 0) {
                const long cb = G.idx(0, j, k);
                const T inv = (T)1 / (T)M.g[0][cb];
                const T csv = bc.cs[0] ? (T)bc.cs[0][(j - 1) + (long)G.ny * (k - 1)] : (T)0;
                const T val = PV(1, j, k) - (T)an * csv * inv +
                        q * inv *
                            ((T)M.g[1][cb] * (PV(1, j + 1, k) + PV(0, j + 1, k) - PV(1, j - 1, k) - PV(0, j - 1, k)) +
                             (T)M.g[2][cb] * (PV(1, j, k + 1) + PV(0, j, k + 1) - PV(1, j, k - 1) - PV(0, j, k - 1)));
                p[cb] = val;
            } else {
                const int nx = G.nx;
                const long ct = G.idx(nx, j, k);
                const T inv = (T)1 / (T)M.g[0][ct];
                const T csv = bc.cs[1] ? (T)bc.cs[1][(j - 1) + (long)G.ny * (k - 1)] : (T)0;
                const T val = PV(nx, j, k) + (T)an * csv * inv -
                            q * inv *
                                ((T)M.g[1][ct] * (PV(nx + 1, j + 1, k) + PV(nx, j + 1, k) - PV(nx + 1, j - 1, k) -
                                                  PV(nx, j - 1, k)) +
                                 (T)M.g[2][ct] * (PV(nx + 1, j, k + 1) + PV(nx, j, k + 1) - PV(nx + 1, j, k - 1) -
                                                  PV(nx, j, k - 1)));
                p[ct + 1] = val;
            }
            return;
        }
        if (set == 1) return;
        t -= nf;
    }
    if (G.kp == 1) {
        const long nf = 2L * G.nx * G.ny;
        if (t < nf) {
            const int side = (int)(t / ((long)G.nx * G.ny));
            const long rem = t % ((long)G.nx * G.ny);
            const int i = (int)(rem % G.nx) + 1, j = (int)(rem / G.nx) + 1;
            if (((i + j) & 1) != color) return;
            if (side == 0) {
                const long cb = G.idx(i, j, 0);
                const T inv = (T)1 / (T)M.g[8][cb];
                const T csv = bc.cs[4] ? (T)bc.cs[4][(i - 1) + (long)G.nx * (j - 1)] : (T)0;
                const T val = PV(i, j, 1) - (T)an * csv * inv +
                        q * inv *
                            ((T)M.g[6][cb] * (PV(i + 1, j, 1) + PV(i + 1, j, 0) - PV(i - 1, j, 1) - PV(i - 1, j, 0)) +
                             (T)M.g[7][cb] * (PV(i, j + 1, 1) + PV(i, j + 1, 0) - PV(i, j - 1, 1) - PV(i, j - 1, 0)));
                p[cb] = val;
            } else {
                const int nz = G.nz;
                const long ct = G.idx(i, j, nz);
                const T inv = (T)1 / (T)M.g[8][ct];
                const T csv = bc.cs[5] ? (T)bc.cs[5][(i - 1) + (long)G.nx * (j - 1)] : (T)0;
                const T val = PV(i, j, nz) + (T)an * csv * inv -
                               q * inv *
                                   ((T)M.g[6][ct] * (PV(i + 1, j, nz + 1) + PV(i + 1, j, nz) - PV(i - 1, j, nz + 1) -
                                                     PV(i - 1, j, nz)) +
                                    (T)M.g[7][ct] * (PV(i, j + 1, nz + 1) + PV(i, j + 1, nz) - PV(i, j - 1, nz + 1) -
                                                     PV(i, j - 1, nz)));
                p[ct + G.sz] = val;
            }
        }
    }
#undef PV
}

// Source of a non-face ghost position: periodic wrap first, then the
// wall-wall edge rules (fsrc/multigrid.f03:1640-1738 for kp = 1,
// 1358-1635 for kp = 0).  Returns false for interior / face positions.
__host__ __device__ inline bool ghost_source(const Geo& G, int i, int j, int k, int& si, int& sj, int& sk) {
    const bool gi = (i == 0 || i == G.nx + 1), gj = (j == 0 || j == G.ny + 1),
               gk = (k == 0 || k == G.nz + 1);
    const int ng = (int)gi + (int)gj + (int)gk;
    if (ng == 0) return false;
    // a single ghost index in a wall direction: face ghost, not a copy
    if (ng == 1 && ((gi && G.ip == 1) || (gj && G.jp == 1) || (gk && G.kp == 1))) return false;
    si = i; sj = j; sk = k;
    if (gi && G.ip == 0) si = (i == 0) ? G.nx : 1;
    if (gk && G.kp == 0) sk = (k == 0) ? G.nz : 1;
    if (gj && G.jp == 0) sj = (j == 0) ? G.ny : 1;
    for (int guard = 0; guard < 3; ++guard) {
        const bool hi = (si == 0 || si == G.nx + 1), hj = (sj == 0 || sj == G.ny + 1),
                   hk = (sk == 0 || sk == G.nz + 1);
        const int n2 = (int)hi + (int)hj + (int)hk;
        if (n2 <= 1) break;
        if (hi && hj && !hk) {
            // xi-eta edge: kp = 0 copies along eta from the xi-ghost, kp = 1 along xi
            if (G.kp == 0) sj = (sj == 0) ? 1 : G.ny;
            else si = (si == 0) ? 1 : G.nx;
        } else if (hj && hk && !hi) {
            sj = (sj == 0) ? 1 : G.ny;  // eta-zeta edge: copy along eta from the zeta-ghost
        } else if (hi && hk && !hj) {
            si = (si == 0) ? 1 : G.nx;  // xi-zeta edge: copy along xi from the zeta-ghost
        } else {
            si = (si == 0) ? 1 : G.nx;  // corner: along xi first
        }
    }
    return true;
}

template <class T>
__host__ __device__ inline void halo_edge_point(const Geo& G, T* __restrict__ p, long t) {
    // every ghost position: enumerate the six ghost planes of the box, each
    // position counted once by giving planes i-ghost > j-ghost > k-ghost
    const long NX = G.nx + 2, NY = G.ny + 2, NZ = G.nz + 2;
    int i, j, k;
    const long ni = 2 * NY * NZ;  // i in {0, nx+1}
    const long nj = 2 * NX * NZ;  // j in {0, ny+1}, i interior
    const long nk = 2 * NX * NY;  // k in {0, nz+1}, i,j interior
    if (t < ni) {
        i = (t / (NY * NZ)) ? G.nx + 1 : 0;
        const long r = t % (NY * NZ);
        j = (int)(r % NY);
        k = (int)(r / NY);
    } else if ((t -= ni) < nj) {
        j = (t / (NX * NZ)) ? G.ny + 1 : 0;
        const long r = t % (NX * NZ);
        i = (int)(r % NX);
        k = (int)(r / NX);
        if (i == 0 || i == G.nx + 1) return;
    } else if ((t -= nj) < nk) {
        k = (t / (NX * NY)) ? G.nz + 1 : 0;
        const long r = t % (NX * NY);
        i = (int)(r % NX);
        j = (int)(r / NX);
        if (i == 0 || i == G.nx + 1 || j == 0 || j == G.ny + 1) return;
    } else {
        return;
    }
    int si, sj, sk;
    if (!ghost_source(G, i, j, k, si, sj, sk)) return;
    p[G.idx(i, j, k)] = p[G.idx(si, sj, sk)];
}

// Eta-wall ghosts of a level periodic in xi and zeta (no other walls).  The
// face relation of fsrc/multigrid.f03:1161-1264 couples a wall's ghosts only
// along the wall plane: g = f + cA (g(i+1) - g(i-1)) + cB (g(k+1) - g(k-1)),
// f the part from the interior row and the wall flux.  One block owns a
// TX x TK tile of one wall and carries a halo of 2*passes points each side
// (periodic wrap), runs `passes` red/black passes on it in LDS and writes its
// tile with the images: the same passes as the launch-per-colour fill, in one
// launch.  The starting ghost values come from a snapshot of the ghost rows
// (k_ghost_rows), so the result does not depend on the order in which
// blocks run.  Dynamic LDS: 4 (TX + 4 passes)(TK + 4 passes) values of T.
// Copy of the two eta-wall ghost rows (snap[side][k][i], ghosted i and k):
// the starting guess of k_eta_ghosts, taken before the launch so no block
// reads a ghost that another block of that launch writes.
template <class T>
__device__ __forceinline__ void d_ghost_rows(Geo G, const T* __restrict__ p, T* __restrict__ snap) {
    const long NX = G.nx + 2, NZ = G.nz + 2;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * NX * NZ) return;
    const int side = (int)(t / (NX * NZ));
    const long r = t % (NX * NZ);
    snap[t] = p[G.idx((int)(r % NX), side ? G.ny + 1 : 0, (int)(r / NX))];
}
template <class T>
__global__ void __launch_bounds__(256) k_ghost_rows(Geo G, const T* __restrict__ p, T* __restrict__ snap) {
    d_ghost_rows<T>(G, p, snap);
}


template <class T, class TM, int TX, int TK>
__device__ __forceinline__ void d_eta_ghosts(Geo G, Met<TM> M, T* __restrict__ p, const T* __restrict__ snap,
                                             BCs bc, int passes) {
    extern __shared__ __align__(16) unsigned char eta_lds[];
    const int Hw = 2 * passes, WX = TX + 2 * Hw, WK = TK + 2 * Hw, NW = WX * WK;
    T* sf = (T*)eta_lds;
    T* sa = sf + NW;
    T* sb = sa + NW;
    T* sg = sb + NW;
    const int tx = (G.nx + TX - 1) / TX, tk = (G.nz + TK - 1) / TK;
    const int side = blockIdx.x / (tx * tk), rem = blockIdx.x % (tx * tk);
    const int ti = (rem % tx) * TX, tkk = (rem / tx) * TK;  // tile origin, 0-based
    const int jr = side == 0 ? 1 : G.ny;                    // interior row at the wall
    const int jg = side == 0 ? 0 : G.ny + 1;                // ghost row
    const int jm = side == 0 ? 0 : G.ny;                    // row of the wall-face metrics
    const T sgn = side == 0 ? (T)1 : (T)-1;
    const T q = (T)0.25;
    const double an = bc.an_over_dt;
    const double* cs = bc.cs[2 + side];
    auto wrap = [](int a, int n) { return ((a % n) + n) % n + 1; };  // 0-based, any sign -> 1..n
    for (int t = threadIdx.x; t < NW; t += blockDim.x) {
        const int i = wrap(ti + (t % WX) - Hw, G.nx), k = wrap(tkk + (t / WX) - Hw, G.nz);
        const long cm = G.idx(i, jm, k);
        const T inv = (T)1 / (T)M.g[4][cm];
        const T cA = sgn * q * inv * (T)M.g[3][cm], cB = sgn * q * inv * (T)M.g[5][cm];
        const T csv = cs ? (T)cs[(i - 1) + (long)G.nx * (k - 1)] : (T)0;
        const int ip1 = i == G.nx ? 1 : i + 1, im1 = i == 1 ? G.nx : i - 1;
        const int kp1 = k == G.nz ? 1 : k + 1, km1 = k == 1 ? G.nz : k - 1;
        sf[t] = (T)p[G.idx(i, jr, k)] - sgn * (T)an * csv * inv +
                cA * ((T)p[G.idx(ip1, jr, k)] - (T)p[G.idx(im1, jr, k)]) +
                cB * ((T)p[G.idx(i, jr, kp1)] - (T)p[G.idx(i, jr, km1)]);
        sa[t] = cA;
        sb[t] = cB;
        sg[t] = snap[(long)side * (G.nx + 2) * (G.nz + 2) + i + (long)(G.nx + 2) * k];
    }
    __syncthreads();
    // each colour phase walks only that colour's points (WX is even: half a row each)
    const int hx = WX / 2;
    for (int ps = 0; ps < passes; ++ps)
        for (int color = 0; color < 2; ++color) {
            for (int u = threadIdx.x; u < hx * WK; u += blockDim.x) {
                const int wk = u / hx, wi = 2 * (u % hx) + ((wk + color) & 1);
                if (wi == 0 || wi == WX - 1 || wk == 0 || wk == WK - 1) continue;
                const int t = wi + WX * wk;
                sg[t] = sf[t] + sa[t] * (sg[t + 1] - sg[t - 1]) + sb[t] * (sg[t + WX] - sg[t - WX]);
            }
            __syncthreads();
        }
    for (int t = threadIdx.x; t < TX * TK; t += blockDim.x) {
        const int i = ti + (t % TX) + 1, k = tkk + (t / TX) + 1;
        if (i > G.nx || k > G.nz) continue;
        const T v = sg[(t % TX + Hw) + WX * (t / TX + Hw)];
        p[G.idx(i, jg, k)] = v;
        write_images(G, p, i, jg, k, v);
    }
}
template <class T, class TM, int TX, int TK>
__global__ void __launch_bounds__(1024) k_eta_ghosts(Geo G, Met<TM> M, T* __restrict__ p, const T* __restrict__ snap,
                                                     BCs bc, int passes) {
    d_eta_ghosts<T, TM, TX, TK>(G, M, p, snap, bc, passes);
}


template <class T, class TM>
__device__ __forceinline__ void d_halo_face(Geo G, Met<TM> M, T* __restrict__ p, int color, BCs bc, int set, int images) {
    halo_face_point(G, M, p, (long)blockIdx.x * blockDim.x + threadIdx.x, color, bc, set, images);
}
template <class T, class TM>
__global__ void k_halo_face(Geo G, Met<TM> M, T* __restrict__ p, int color, BCs bc, int set, int images) {
    d_halo_face<T, TM>(G, M, p, color, bc, set, images);
}


template <class T>
__device__ __forceinline__ void d_halo_edges(Geo G, T* __restrict__ p) {
    halo_edge_point(G, p, (long)blockIdx.x * blockDim.x + threadIdx.x);
}
template <class T>
__global__ void k_halo_edges(Geo G, T* __restrict__ p) {
    d_halo_edges<T>(G, p);
}


// Exact coarsest-level solve: p (every ghosted position) = K b, K the dense
// map built at set-up (rows: ghosted positions, columns: interior cells, i
// fastest, each row padded with zeros to kp columns, a multiple of 256), b
// the interior of the ghosted rhs, gathered through gi (the ghosted index of
// each interior cell).  The block stages b's interior in LDS (the dense solve
// is for levels of at most kDenseCols cells); each wavefront takes one row
// and issues all its 16-byte loads of the row before the first product (an
// FP64 row of 2048 entries is 16 loads per lane), then sums in a fixed order
// and a fixed shuffle tree: deterministic.  An FP32 map (the mixed solver's:
// the V-cycle it preconditions runs in FP32) halves the bytes; sums are FP64.
constexpr int kDenseCols = 2048, kDenseWaves = 8;
template <class TK>
__device__ __forceinline__ void d_coarse_dense(Geo G, const TK* __restrict__ K, const int* __restrict__ gi,
                                               int N, int kp, const double* __restrict__ b,
                                               double* __restrict__ p) {
    __shared__ double bs[kDenseCols];
    constexpr int VW = 16 / sizeof(TK);          // map entries per 16-byte load
    constexpr int NL = kDenseCols / (64 * VW);   // loads per lane of the longest row
    using V = typename std::conditional<sizeof(TK) == 4, float4, double2>::type;
    const int lane = threadIdx.x & 63;
    const long row = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const bool act = row < G.n;  // uniform over the wavefront
    // the row's loads first: they are in flight while the block stages b
    V kv[NL];
    if (act) {
        const V* kr = reinterpret_cast<const V*>(K + row * (long)kp) + lane;
#pragma unroll
        for (int u = 0; u < NL; ++u)
            if (u * 64 * VW < kp) kv[u] = kr[u * 64];
    }
    for (int c = threadIdx.x; c < kp; c += blockDim.x) bs[c] = (c < N) ? b[gi[c]] : 0.0;
    __syncthreads();
    if (!act) return;
    double a = 0.0;
#pragma unroll
    for (int u = 0; u < NL; ++u) {
        if (u * 64 * VW >= kp) continue;  // uniform; no break, so the loop unrolls and kv stays in registers
        const int c = (u * 64 + lane) * VW;
        if constexpr (VW == 4) {
            a += ((double)kv[u].x * bs[c] + (double)kv[u].y * bs[c + 1]) +
                 ((double)kv[u].z * bs[c + 2] + (double)kv[u].w * bs[c + 3]);
        } else {
            a += (double)kv[u].x * bs[c] + (double)kv[u].y * bs[c + 1];
        }
    }
    for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
    if (lane == 0) p[row] = a;
}
template <class TK>
__global__ void __launch_bounds__(64 * kDenseWaves) k_coarse_dense(Geo G, const TK* __restrict__ K,
                                                                   const int* __restrict__ gi, int N, int kp,
                                                                   const double* __restrict__ b,
                                                                   double* __restrict__ p) {
    d_coarse_dense<TK>(G, K, gi, N, kp, b, p);
}

// ghosted index of each interior cell (i fastest): the dense solve's gather
__global__ void __launch_bounds__(256) k_interior_index(Geo G, int* __restrict__ gi) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long)G.nx * G.ny * G.nz) return;
    const int i = (int)(c % G.nx) + 1, j = (int)((c / G.nx) % G.ny) + 1, k = (int)(c / ((long)G.nx * G.ny)) + 1;
    gi[c] = (int)G.idx(i, j, k);
}

// the set-up's FP64 map (rows of N entries) in the solve's layout: rows
// padded with zeros to kp entries, converted to TK
template <class TK>
__global__ void __launch_bounds__(256) k_pad_map(const double* __restrict__ src, long rows, int N, int kp,
                                                 TK* __restrict__ dst) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows * kp) return;
    const long r = t / kp;
    const int c = (int)(t % kp);
    dst[t] = (TK)(c < N ? src[r * N + c] : 0.0);
}

// ------------------------------------------------------------------------
// Gauss-Jordan inverse with partial pivoting of an nb x nb matrix B, held
// with the identity as the row-major nb x 2nb array M = [B | I]; after the
// nb column steps the right half is B^-1 (set-up of the exact coarse solve).
// Column step col: one block picks the pivot row (largest |M[r][col]|, r >=
// col, the smallest such r on ties), swaps it into row col, scales that row
// and records every row's elimination factor; then the other rows are
// updated in the columns right of col (the columns left of it are already
// unit columns, and column col itself is not read again).
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) k_gj_pivot(double* __restrict__ M, long nb, long col,
                                                   double* __restrict__ f, int* __restrict__ status) {
    __shared__ double bv[1024];
    __shared__ long bi[1024];
    const long w = 2 * nb;
    const int t = threadIdx.x;
    double best = -1.0;
    long bidx = col;
    for (long r = col + t; r < nb; r += blockDim.x) {
        const double a = fabs(M[r * w + col]);
        if (a > best) {
            best = a;
            bidx = r;
        }
    }
    bv[t] = best;
    bi[t] = bidx;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if (t < h && (bv[t + h] > bv[t] || (bv[t + h] == bv[t] && bi[t + h] < bi[t]))) {
            bv[t] = bv[t + h];
            bi[t] = bi[t + h];
        }
        __syncthreads();
    }
    const long p = bi[0];
    const double pv = M[p * w + col];
    if (!(fabs(pv) > 0.0)) {  // singular (every thread sees the same value)
        if (t == 0) *status = 1;
        return;
    }
    __syncthreads();  // every thread has read the pivot before the rows change
    const double inv = 1.0 / pv;
    for (long c = t; c < w; c += blockDim.x) {
        const double a = M[p * w + c], b = M[col * w + c];
        M[col * w + c] = a * inv;
        if (p != col) M[p * w + c] = b;
    }
    __syncthreads();
    for (long r = t; r < nb; r += blockDim.x) f[r] = (r == col) ? 0.0 : M[r * w + col];
}
// grid: (columns right of col, 256 per block) x rows
__global__ void __launch_bounds__(256) k_gj_elim(double* __restrict__ M, long nb, long col,
                                                 const double* __restrict__ f) {
    const long w = 2 * nb;
    const long c = col + 1 + (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long r = blockIdx.y;
    if (c >= w) return;
    const double fr = f[r];
    if (fr == 0.0) return;
    M[r * w + c] -= fr * M[col * w + c];
}

// The exact coarse solve's map K_full = E K (ghosted rows x interior
// columns): E's column m is the ghosted image of the unit vector e_m (stored
// column-major, ng per column), K the inverse's top-left N x N block, the
// right half of the Gauss-Jordan array M (row m at M + m w + nb).  An interior
// row g is the inverse's row of its cell; a ghost row is the combination of
// the inverse's rows its image entries weight.  One thread per (row g,
// column c); a block row shares g, so E's entries are block-uniform.
__global__ void __launch_bounds__(256) k_dense_map(Geo G, const double* __restrict__ E, const double* __restrict__ M,
                                                   long nb, long N, double* __restrict__ KF) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long g = blockIdx.y;
    if (c >= N) return;
    const long w = 2 * nb;
    const long NX = G.nx + 2, NY = G.ny + 2;
    const int i = (int)(g % NX), j = (int)((g / NX) % NY), k = (int)(g / (NX * NY));
    double acc = 0.0;
    if (i >= 1 && i <= G.nx && j >= 1 && j <= G.ny && k >= 1 && k <= G.nz) {
        const long m = (long)(i - 1) + (long)G.nx * ((j - 1) + (long)G.ny * (k - 1));
        acc = M[m * w + nb + c];
    } else {
        const long ng = G.n;
        for (long m = 0; m < N; ++m) {
            const double e = E[m * ng + g];
            if (e != 0.0) acc += e * M[m * w + nb + c];
        }
    }
    KF[g * N + c] = acc;
}

// ------------------------------------------------------------------------
// Smoother factors of the segmented eta-line solve (see LineF), computed in
// FP64 and stored in FP32; the v and w solves use the stored factors so the
// partitioned solve is consistent with the smoother's local solves.
// ------------------------------------------------------------------------
template <class TM>
__global__ void k_line_factors(Geo G, Met<TM> M, double inv_pot, int P, float* __restrict__ invm,
                               float* __restrict__ cpr, float* __restrict__ v, float* __restrict__ w) {
    Seg q;
    if (!seg_of(G, P, q)) return;
    const int i = q.i, k = q.k;
    const double s2 = inv_pot * inv_pot;
    auto den_at = [&](int j) {
        const long c = G.idx(i, j, k);
        const double g11p = (G.ip == 1 && i == G.nx) ? 0.0 : (double)M.g[0][c];
        const double g11m = (G.ip == 1 && i == 1) ? 0.0 : (double)M.g[0][c - 1];
        const double g22p = (j == G.ny) ? 0.0 : (double)M.g[4][c];
        const double g22m = (j == 1) ? 0.0 : (double)M.g[4][c - G.sy];
        const double g33p = (G.kp == 1 && k == G.nz) ? 0.0 : (double)M.g[8][c];
        const double g33m = (G.kp == 1 && k == 1) ? 0.0 : (double)M.g[8][c - G.sz];
        return s2 * (g11p + g11m + g22p + g22m + g33p + g33m);
    };
    // couplings of the whole column: a_j to j-1, c_j to j+1 (zero at the walls)
    auto a_at = [&](int j) { return (j == 1) ? 0.0 : -s2 * (double)M.g[4][G.idx(i, j - 1, k)]; };
    auto c_at = [&](int j) { return (j == G.ny) ? 0.0 : -s2 * (double)M.g[4][G.idx(i, j, k)]; };
    double cprev = 0.0;
    for (int j = q.j0; j <= q.j1; ++j) {
        const long c = G.idx(i, j, k);
        const double a = (j > q.j0) ? a_at(j) : 0.0;
        const double cc = (j < q.j1) ? c_at(j) : 0.0;
        const double m = den_at(j) - a * cprev;
        const double cp = cc / m;
        invm[c] = (float)(1.0 / m);
        cpr[c] = (float)cp;
        cprev = (double)(float)cp;
    }
    // v = M_s^-1 (a_j0 e_first), w = M_s^-1 (c_j1 e_last), forward then back;
    // a single segment has no neighbours (and no v, w arrays)
    if (P == 1) return;
    const double aj0 = a_at(q.j0), cj1 = c_at(q.j1);
    double dv = 0.0, dw = 0.0;
    for (int j = q.j0; j <= q.j1; ++j) {
        const long c = G.idx(i, j, k);
        const double a = (j > q.j0) ? a_at(j) : 0.0;
        dv = ((j == q.j0 ? aj0 : 0.0) - a * dv) * (double)invm[c];
        dw = ((j == q.j1 ? cj1 : 0.0) - a * dw) * (double)invm[c];
        v[c] = (float)dv;
        w[c] = (float)dw;
    }
    double vn = 0.0, wn = 0.0;
    for (int j = q.j1; j >= q.j0; --j) {
        const long c = G.idx(i, j, k);
        vn = (double)v[c] - (double)cpr[c] * vn;
        wn = (double)w[c] - (double)cpr[c] * wn;
        v[c] = (float)vn;
        w[c] = (float)wn;
    }
}

// ------------------------------------------------------------------------
// Geometry: face metrics J^-1 g^{ij} / pot of the level with node stride st
// (metrica, fsrc/metrica.f03:85-419, st = 1; mul_met,
// fsrc/multigrid.f03:2520-2988).  Node coordinates come from the finest
// node arrays with the periodic extension of cellep (fsrc/cellep.f03:29-98)
// evaluated on the fly.  At coarse wall faces where the reference's
// second-order one-sided derivative inverts the face Jacobian (strongly
// stretched walls with few coarse cells), the first-order one-sided
// derivative is used instead.
// ------------------------------------------------------------------------
struct Nodes {
    const double *x, *y, *z;
    int nx, ny, nz, ip, jp, kp;  // finest level
};

__device__ inline int fdiv(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }

__device__ inline void node_at(const Nodes& N, int i, int j, int k, double o[3]) {
    // periodic-style extension in every direction (as cellep fills the ghost
    // node layers), shift by the node span per period
    const int mk = (k < 0 || k > N.nz) ? fdiv(k, N.nz) : 0;
    const int mj = (j < 0 || j > N.ny) ? fdiv(j, N.ny) : 0;
    const int mi = (i < 0 || i > N.nx) ? fdiv(i, N.nx) : 0;
    const int ii = i - mi * N.nx, jj = j - mj * N.ny, kk = k - mk * N.nz;
    auto at = [&](int a, int b, int c, int comp) -> double {
        const long n = (long)a + (long)(N.nx + 1) * ((long)b + (long)(N.ny + 1) * c);
        return comp == 0 ? N.x[n] : (comp == 1 ? N.y[n] : N.z[n]);
    };
    o[0] = at(ii, jj, kk, 0);
    o[1] = at(ii, jj, kk, 1);
    o[2] = at(ii, jj, kk, 2);
    if (mi) o[0] += mi * (at(N.nx, jj, kk, 0) - at(0, jj, kk, 0));
    if (mj) o[1] += mj * (at(ii, N.ny, kk, 1) - at(ii, 0, kk, 1));
    if (mk) o[2] += mk * (at(ii, jj, N.nz, 2) - at(ii, jj, 0, 2));
}

struct Frame {
    double cs[3], et[3], zt[3], giac;
};

__device__ inline Frame make_frame(const double c[3], const double e[3], const double t[3]) {
    Frame F;
    const double xcsi = c[0], ycsi = c[1], zcsi = c[2];
    const double xeta = e[0], yeta = e[1], zeta = e[2];
    const double xzet = t[0], yzet = t[1], zzet = t[2];
    F.giac = xcsi * (yeta * zzet - yzet * zeta) - xeta * (ycsi * zzet - yzet * zcsi) +
             xzet * (ycsi * zeta - yeta * zcsi);
    F.cs[0] = yeta * zzet - yzet * zeta;
    F.cs[1] = xzet * zeta - xeta * zzet;
    F.cs[2] = xeta * yzet - xzet * yeta;
    F.et[0] = yzet * zcsi - ycsi * zzet;
    F.et[1] = xcsi * zzet - xzet * zcsi;
    F.et[2] = xzet * ycsi - xcsi * yzet;
    F.zt[0] = ycsi * zeta - yeta * zcsi;
    F.zt[1] = xeta * zcsi - xcsi * zeta;
    F.zt[2] = xcsi * yeta - xeta * ycsi;
    return F;
}

__device__ inline double dot3(const double* a, const double* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ inline double sq3(const double* a) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }

// four-node average in the given order
__device__ inline void avg4(const Nodes& N, const int q[4][3], double o[3]) {
    double a[3], b[3], c[3], d[3];
    node_at(N, q[0][0], q[0][1], q[0][2], a);
    node_at(N, q[1][0], q[1][1], q[1][2], b);
    node_at(N, q[2][0], q[2][1], q[2][2], c);
    node_at(N, q[3][0], q[3][1], q[3][2], d);
    for (int m = 0; m < 3; ++m) o[m] = .25 * (a[m] + b[m] + c[m] + d[m]);
}

__device__ inline void hdiff(const Nodes& N, int a0, int a1, int a2, int b0, int b1, int b2, int c0,
                             int c1, int c2, int d0, int d1, int d2, double o[3]) {
    double A[3], B[3], C[3], D[3];
    node_at(N, a0, a1, a2, A);
    node_at(N, b0, b1, b2, B);
    node_at(N, c0, c1, c2, C);
    node_at(N, d0, d1, d2, D);
    for (int m = 0; m < 3; ++m) o[m] = .5 * (A[m] + B[m]) - .5 * (C[m] + D[m]);
}

// one-sided derivative at a wall face: second order as the reference, with
// the first-order fallback when the second-order one points the wrong way
__device__ inline void wall_deriv(const double x0[3], const double x1[3], const double x2[3],
                                  double sgn, int st, double o[3]) {
    double s2[3], s1[3];
    for (int m = 0; m < 3; ++m) {
        s2[m] = sgn * .5 * (-3. * x0[m] + 4. * x1[m] - x2[m]);
        s1[m] = sgn * (x1[m] - x0[m]);
    }
    if (st > 1 && dot3(s2, s1) <= 0.25 * sq3(s1)) {
        for (int m = 0; m < 3; ++m) o[m] = s1[m];
    } else {
        for (int m = 0; m < 3; ++m) o[m] = s2[m];
    }
}

// dir 0: xi faces (i = 0..nx), 1: eta faces (j = 0..ny), 2: zeta faces.
// sx, sy, sz: node strides of the level (2^(n-1) each for the reference's
// level n; a coarse level below one that kept the xi or the xi and eta counts
// has half the stride there); coarse: the level is not the finest (the wall
// derivative fallback).  The metrics are divided by div = sx sy / sz: with the
// operator's 1/sz^2 this scales the coarse-index flux sum by the number of fine
// cells per coarse cell (sx sy sz), the reference's 1/pot when all are pot.
template <class TO>
__global__ void k_metrics(Nodes N, Geo G, int sx, int sy, int sz, int coarse, int dir, TO* g0, TO* g1, TO* g2,
                          double* v0, double* v1, double* v2) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y * blockDim.y + threadIdx.y;
    const int cidx = blockIdx.z;
    int i, j, k;
    if (dir == 0) {
        i = a; j = b + 1; k = cidx + 1;
        if (i > G.nx || j > G.ny) return;
    } else if (dir == 1) {
        i = a + 1; j = b; k = cidx + 1;
        if (i > G.nx || j > G.ny) return;
    } else {
        i = a + 1; j = b + 1; k = cidx;
        if (i > G.nx || j > G.ny) return;
    }
    const double pot = (double)sx * (double)sy / (double)sz;
    const int st = coarse ? 2 : 1;  // wall_deriv: first-order fallback on coarse levels
    double c[3], e[3], t[3];
    if (dir == 0) {
        const int j2 = j * sy, j1 = (j - 1) * sy, k2 = k * sz, k1 = (k - 1) * sz, ic = i * sx;
        auto face = [&](int aa, double o[3]) {
            const int q[4][3] = {{aa, j2, k2}, {aa, j2, k1}, {aa, j1, k2}, {aa, j1, k1}};
            avg4(N, q, o);
        };
        if (i == 0 && G.ip == 1) {
            double x0[3], x1[3], x2[3];
            face(0, x0); face(sx, x1); face(2 * sx, x2);
            wall_deriv(x0, x1, x2, 1.0, st, c);
        } else if (i == G.nx && G.ip == 1) {
            double x0[3], x1[3], x2[3];
            face(N.nx, x0); face((i - 1) * sx, x1); face((i - 2) * sx, x2);
            wall_deriv(x0, x1, x2, -1.0, st, c);
        } else {
            double xd[3], xs[3];
            face((i + 1) * sx, xd); face((i - 1) * sx, xs);
            for (int m = 0; m < 3; ++m) c[m] = .5 * (xd[m] - xs[m]);
        }
        hdiff(N, ic, j2, k2, ic, j2, k1, ic, j1, k2, ic, j1, k1, e);
        hdiff(N, ic, j2, k2, ic, j1, k2, ic, j2, k1, ic, j1, k1, t);
        const Frame F = make_frame(c, e, t);
        const long id = G.idx(i, j, k);
        g0[id] = (TO)((sq3(F.cs) / F.giac) / pot);
        g1[id] = (TO)((dot3(F.cs, F.et) / F.giac) / pot);
        g2[id] = (TO)((dot3(F.cs, F.zt) / F.giac) / pot);
        if (v0) { v0[id] = F.cs[0]; v1[id] = F.cs[1]; v2[id] = F.cs[2]; }
    } else if (dir == 1) {
        const int i2 = i * sx, i1 = (i - 1) * sx, k2 = k * sz, k1 = (k - 1) * sz, jc = j * sy;
        auto face = [&](int bb, double o[3]) {
            const int q[4][3] = {{i2, bb, k2}, {i2, bb, k1}, {i1, bb, k1}, {i1, bb, k2}};
            avg4(N, q, o);
        };
        if (j == 0 && G.jp == 1) {
            double x0[3], x1[3], x2[3];
            face(0, x0); face(sy, x1); face(2 * sy, x2);
            wall_deriv(x0, x1, x2, 1.0, st, e);
        } else if (j == G.ny && G.jp == 1) {
            double x0[3], x1[3], x2[3];
            face(N.ny, x0); face((j - 1) * sy, x1); face((j - 2) * sy, x2);
            wall_deriv(x0, x1, x2, -1.0, st, e);
        } else {
            double xp[3], xm[3];
            face((j + 1) * sy, xp); face((j - 1) * sy, xm);
            for (int m = 0; m < 3; ++m) e[m] = .5 * (xp[m] - xm[m]);
        }
        hdiff(N, i2, jc, k2, i1, jc, k2, i2, jc, k1, i1, jc, k1, t);
        hdiff(N, i2, jc, k2, i2, jc, k1, i1, jc, k2, i1, jc, k1, c);
        const Frame F = make_frame(c, e, t);
        const long id = G.idx(i, j, k);
        g0[id] = (TO)((dot3(F.et, F.cs) / F.giac) / pot);
        g1[id] = (TO)((sq3(F.et) / F.giac) / pot);
        g2[id] = (TO)((dot3(F.et, F.zt) / F.giac) / pot);
        if (v0) { v0[id] = F.et[0]; v1[id] = F.et[1]; v2[id] = F.et[2]; }
    } else {
        const int i2 = i * sx, i1 = (i - 1) * sx, j2 = j * sy, j1 = (j - 1) * sy, kc = k * sz;
        auto face = [&](int cc, double o[3]) {
            const int q[4][3] = {{i2, j2, cc}, {i2, j1, cc}, {i1, j1, cc}, {i1, j2, cc}};
            avg4(N, q, o);
        };
        if (k == 0 && G.kp == 1) {
            double x0[3], x1[3], x2[3];
            face(0, x0); face(sz, x1); face(2 * sz, x2);
            wall_deriv(x0, x1, x2, 1.0, st, t);
        } else if (k == G.nz && G.kp == 1) {
            double x0[3], x1[3], x2[3];
            face(N.nz, x0); face((k - 1) * sz, x1); face((k - 2) * sz, x2);
            wall_deriv(x0, x1, x2, -1.0, st, t);
        } else {
            double xa[3], xb[3];
            face((k + 1) * sz, xa); face((k - 1) * sz, xb);
            for (int m = 0; m < 3; ++m) t[m] = .5 * (xa[m] - xb[m]);
        }
        hdiff(N, i2, j2, kc, i2, j1, kc, i1, j2, kc, i1, j1, kc, c);
        hdiff(N, i2, j2, kc, i1, j2, kc, i2, j1, kc, i1, j1, kc, e);
        const Frame F = make_frame(c, e, t);
        const long id = G.idx(i, j, k);
        g0[id] = (TO)((dot3(F.zt, F.cs) / F.giac) / pot);
        g1[id] = (TO)((dot3(F.zt, F.et) / F.giac) / pot);
        g2[id] = (TO)((sq3(F.zt) / F.giac) / pot);
        if (v0) { v0[id] = F.zt[0]; v1[id] = F.zt[1]; v2[id] = F.zt[2]; }
    }
}

// centroid Jacobian giac (fsrc/metrica.f03:432-483), finest level
__global__ void k_giac(Nodes N, Geo G, double* giac) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > G.nx || j > G.ny) return;
    double P[2][2][2][3];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) node_at(N, i - 1 + a, j - 1 + b, k - 1 + c, P[a][b][c]);
    double c[3], e[3], t[3];
    for (int q = 0; q < 3; ++q) {
        c[q] = .25 * (P[1][1][1][q] + P[1][1][0][q] + P[1][0][0][q] + P[1][0][1][q]) -
               .25 * (P[0][1][1][q] + P[0][1][0][q] + P[0][0][0][q] + P[0][0][1][q]);
        e[q] = .25 * (P[1][1][1][q] + P[1][1][0][q] + P[0][1][0][q] + P[0][1][1][q]) -
               .25 * (P[1][0][1][q] + P[1][0][0][q] + P[0][0][0][q] + P[0][0][1][q]);
        t[q] = .25 * (P[1][1][1][q] + P[1][0][1][q] + P[0][0][1][q] + P[0][1][1][q]) -
               .25 * (P[1][1][0][q] + P[1][0][0][q] + P[0][0][0][q] + P[0][1][0][q]);
    }
    giac[G.idx(i, j, k)] = make_frame(c, e, t).giac;
}

// ------------------------------------------------------------------------
// Synthetic case: counter-based N(0,1) cell velocity, contra fluxes, diver
// ------------------------------------------------------------------------
__device__ inline uint64_t splitmix64(uint64_t v) {
    v += 0x9E3779B97F4A7C15ull;
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
    return v ^ (v >> 31);
}

__device__ inline double gauss(uint64_t seed, uint32_t field, int64_t ctr) {
    const uint64_t key = splitmix64(seed ^ ((uint64_t)field << 56));
    const uint64_t h1 = splitmix64(key ^ (2ull * (uint64_t)ctr));
    const uint64_t h2 = splitmix64(key ^ (2ull * (uint64_t)ctr + 1ull));
    const double u1 = ((double)(h1 >> 11) + 1.0) * 0x1.0p-53;
    const double u2 = (double)(h2 >> 11) * 0x1.0p-53;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
}

// velocity component c of cell (i,j,k) incl. periodic ghost layers (ghost
// cells in wall directions are never read: wall fluxes are zero)
__device__ inline double vel(const Geo& G, uint64_t seed, int c, int i, int j, int k) {
    if (G.ip == 0) { if (i == 0) i = G.nx; else if (i == G.nx + 1) i = 1; }
    if (G.jp == 0) { if (j == 0) j = G.ny; else if (j == G.ny + 1) j = 1; }
    if (G.kp == 0) { if (k == 0) k = G.nz; else if (k == G.nz + 1) k = 1; }
    return gauss(seed, (uint32_t)c, (int64_t)G.idx(i, j, k));
}

// contravariant flux of face f (dir) from the seeded velocity
// (fsrc/contra.f03:73-87,132-146,204-220), zero on wall faces
__device__ inline double face_flux(const Geo& G, const double* const* csv, uint64_t seed, int dir,
                                   int i, int j, int k) {
    int i2 = i, j2 = j, k2 = k;
    if (dir == 0) {
        if (G.ip == 1 && (i == 0 || i == G.nx)) return 0.0;
        i2 = i + 1;
    } else if (dir == 1) {
        if (G.jp == 1 && (j == 0 || j == G.ny)) return 0.0;
        j2 = j + 1;
    } else {
        if (G.kp == 1 && (k == 0 || k == G.nz)) return 0.0;
        k2 = k + 1;
    }
    const double ui = .5 * (vel(G, seed, 0, i, j, k) + vel(G, seed, 0, i2, j2, k2));
    const double vi = .5 * (vel(G, seed, 1, i, j, k) + vel(G, seed, 1, i2, j2, k2));
    const double wi = .5 * (vel(G, seed, 2, i, j, k) + vel(G, seed, 2, i2, j2, k2));
    const long id = G.idx(i, j, k);
    return ui * csv[3 * dir][id] + vi * csv[3 * dir + 1][id] + wi * csv[3 * dir + 2][id];
}

struct CsvP {
    const double* v[9];
};

// rhs = div(flux)/dt (fsrc/diver.f03:20-28) of the seeded case; partial sums
__global__ void __launch_bounds__(256) k_case_rhs(Geo G, CsvP V, uint64_t seed, double dt,
                                                  double* __restrict__ rhs, double* partials,
                                                  int nblocks) {
    int i, k;
    const bool act = column_of(G, i, k);
    double acc[1] = {0};
    if (act) {
        for (int j = 1; j <= G.ny; ++j) {
            const double d = face_flux(G, V.v, seed, 0, i, j, k) - face_flux(G, V.v, seed, 0, i - 1, j, k) +
                             face_flux(G, V.v, seed, 1, i, j, k) - face_flux(G, V.v, seed, 1, i, j - 1, k) +
                             face_flux(G, V.v, seed, 2, i, j, k) - face_flux(G, V.v, seed, 2, i, j, k - 1);
            const double r = d / dt;
            rhs[G.idx(i, j, k)] = r;
            acc[0] += r;
        }
    }
    const bool mx[1] = {false};
    block_reduce_store<1>(acc, partials, nblocks, mx);
}

// diver from caller fluxes in the ghosted face layout
__global__ void __launch_bounds__(256) k_diver(Geo G, const double* __restrict__ uc,
                                               const double* __restrict__ vc,
                                               const double* __restrict__ wc, double dt,
                                               double* __restrict__ rhs) {
    int i, k;
    if (!column_of(G, i, k)) return;
    for (int j = 1; j <= G.ny; ++j) {
        const long c = G.idx(i, j, k);
        rhs[c] = (uc[c] - uc[c - 1] + vc[c] - vc[c - G.sy] + wc[c] - wc[c - G.sz]) / dt;
    }
}

// column sums of the interior (for the compatibility projection)
__global__ void __launch_bounds__(256) k_sum(Geo G, const double* __restrict__ a, double* partials,
                                             int nblocks) {
    int i, k;
    const bool act = column_of(G, i, k);
    double acc[2] = {0, 0};
    if (act) {
        // four rows per trip, four partial sums each (the loads of a trip are
        // independent); fixed order
        const long c0 = G.idx(i, 1, k), sy = G.sy;
        double s[4] = {0, 0, 0, 0}, q[4] = {0, 0, 0, 0};
        int j = 0;
#pragma unroll 2
        for (; j + 4 <= G.ny; j += 4)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double v = a[c0 + sy * (j + u)];
                s[u] += v;
                q[u] += v * v;
            }
        for (; j < G.ny; ++j) {
            const double v = a[c0 + sy * j];
            s[0] += v;
            q[0] += v * v;
        }
        acc[0] = (s[0] + s[1]) + (s[2] + s[3]);
        acc[1] = (q[0] + q[1]) + (q[2] + q[3]);
    }
    const bool mx[2] = {false, false};
    block_reduce_store<2>(acc, partials, nblocks, mx);
}

// sums of the partials of k_sum / k_case_rhs into S[S_SUM], S[S_SUM+1]
// (one block of 256 threads, fixed order)
__global__ void __launch_bounds__(256) k_fin_sum(const double* __restrict__ part, int nb, double* S) {
    const bool mx[2] = {false, false};
    double o[2];
    fin_reduce<2>(part, nb, mx, o);
    if (threadIdx.x == 0) {
        S[S_SUM] = o[0];
        S[S_SUM + 1] = o[1];
    }
}

// a[interior] -= (S[S_SUM] - bflux) / ncell
__global__ void __launch_bounds__(256) k_sub_mean(Geo G, double* __restrict__ a, const double* S,
                                                  double bflux, double ncell) {
    long c;
    if (!cell_of(G, c)) return;
    a[c] -= (S[S_SUM] - bflux) / ncell;
}

// interior perturbation p += rel * std * N(0,1)
__global__ void __launch_bounds__(256) k_perturb(Geo G, double* __restrict__ a, const double* S,
                                                 uint64_t seed, double rel, double ncell) {
    long c;
    if (!cell_of(G, c)) return;
    const double mean = S[S_SUM] / ncell;
    const double var = fmax(S[S_SUM + 1] / ncell - mean * mean, 0.0);
    a[c] += rel * sqrt(var) * gauss(seed, 7u, (int64_t)c);
}

// interior = N(0,1) per (seed, cell) (a test field)
__global__ void __launch_bounds__(256) k_fill_gauss(Geo G, double* __restrict__ a, uint64_t seed) {
    long c;
    if (!cell_of(G, c)) return;
    a[c] = gauss(seed, 9u, (int64_t)c);
}

// per-block maxima of |a - b| over n values (grid-stride)
__global__ void __launch_bounds__(256) k_absdiff_max(long n, const double* __restrict__ a,
                                                     const double* __restrict__ b, double* partials) {
    double v[1] = {0.0};
    const bool mx[1] = {true};
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        v[0] = fmax(v[0], fabs(a[t] - b[t]));
    block_reduce_store<1>(v, partials, gridDim.x, mx);
}

// r32 = s * r64 over the interior (the FP32 inner problem's rhs)
__global__ void __launch_bounds__(256) k_scale_to32(Geo G, const double* __restrict__ r, double s,
                                                    float* __restrict__ r32) {
    long c;
    if (!cell_of(G, c)) return;
    r32[c] = (float)(s * r[c]);
}

// x += s * e32 over the interior (FP64 iterative-refinement update)
__global__ void __launch_bounds__(256) k_add_back(Geo G, double* __restrict__ x, double s,
                                                  const float* __restrict__ e32) {
    long c;
    if (!cell_of(G, c)) return;
    x[c] += s * (double)e32[c];
}

// ------------------------------------------------------------------------
// Device-driven iterative refinement: the refinement loop of the mixed
// solver (inner FP32 GCR to a relative tolerance, x += |r| e, FP64 residual,
// next tolerance) with every decision taken on the device, so a batch of
// solves advances by graph-replayed steps without host round trips.  Each
// step is one inner step followed by this tail, whose kernels do nothing
// unless the step ends a refinement (S_REF).
// ------------------------------------------------------------------------
// Does this step end a refinement?  The inner solve met its tolerance, or it
// ran 40 steps, or the cycle budget is spent.
__device__ __forceinline__ void d_ir_check(double* S, const double* S2) {
    if (threadIdx.x != 0) return;
    double ref = 0.0;
    if (S[S_DONE] == 0.0) {
        const double n = S2[S_CYC];
        const double lim = fmin(40.0, S[S_MAXC] - S[S_TOTC]);
        if (S2[S_DONE] != 0.0 || n >= lim) ref = 1.0;
    }
    S[S_REF] = ref;
    S[S_NOREF] = 1.0 - ref;
}
__global__ void k_ir_check(double* S, const double* S2) { d_ir_check(S, S2); }

// x += |r| e32 at the end of a refinement
// interior cell c of a 1-D launch over the ghosted array
__device__ inline bool flat_interior(const Geo& G, long& c) {
    c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= G.n) return false;
    const long NX = G.nx + 2, NY = G.ny + 2;
    const int i = (int)(c % NX), j = (int)((c / NX) % NY), k = (int)(c / (NX * NY));
    return i >= 1 && i <= G.nx && j >= 1 && j <= G.ny && k >= 1 && k <= G.nz;
}

__device__ __forceinline__ void d_add_back_if(Geo G, double* __restrict__ x, const double* __restrict__ S,
                                              const float* __restrict__ e32) {
    if (S[S_REF] == 0.0) return;
    long c;
    if (!flat_interior(G, c)) return;
    x[c] += S[S_RN] * (double)e32[c];
}
__global__ void __launch_bounds__(256) k_add_back_if(Geo G, double* __restrict__ x, const double* __restrict__ S,
                                                     const float* __restrict__ e32) {
    d_add_back_if(G, x, S, e32);
}

// the new FP64 residual's norms (k_resid MODE 1 partials), the stop and
// stall tests, and the next refinement's scale and inner tolerance (S_DONE:
// 1 converged, 2 not converged, 3 non-finite); a finished solve clears S_REF
__device__ __forceinline__ void d_fin_ir(const double* __restrict__ part, int nb, double* S, double* S2) {
    if (S[S_REF] == 0.0) return;  // uniform
    const bool mx[3] = {false, true, false};
    double o[3];
    fin_reduce<3>(part, nb, mx, o);
    if (threadIdx.x != 0) return;
    const double rr = o[0], rmax = o[1];
    S[S_RR] = rr;
    S[S_RMAX] = rmax;
    S[S_TOTC] += S2[S_CYC];
    const bool absm = S[S_ABS] != 0.0;
    const double m = absm ? rmax : sqrt(rr) / S[S_BN];
    double done = 0.0;
    if (!isfinite(rr)) {
        done = 3.0;
    } else if (m < S[S_EPS]) {
        done = 1.0;
    } else {
        if (m > 0.5 * S[S_MOLD]) {
            S[S_STALL] += 1.0;
            if (S[S_STALL] >= 3.0) done = 2.0;  // FP32 cannot reduce it further
        } else {
            S[S_STALL] = 0.0;
        }
        S[S_MOLD] = m;
        if (done == 0.0 && S[S_TOTC] >= S[S_MAXC]) done = 2.0;
    }
    const double rn = sqrt(rr);
    if (done == 0.0 && !(rn > 0.0)) done = 1.0;
    if (done != 0.0) {
        S[S_DONE] = done;
        S[S_REF] = 0.0;
        S[S_NOREF] = 1.0;
        return;
    }
    S[S_RN] = rn;
    const double need = absm ? S[S_EPS] / fmax(rmax, 1e-300) : S[S_EPS] * S[S_BN] / rn;
    const double tol = fmin(0.1, fmax(1e-5, 0.5 * need));
    S2[S_DONE] = 0.0;
    S2[S_THR] = tol * tol;
    S2[S_MODE] = 0.0;
    S2[S_CYC] = 0.0;
    S2[S_HAVEP] = 0.0;
}
__global__ void __launch_bounds__(256) k_fin_ir(const double* __restrict__ part, int nb, double* S, double* S2) {
    d_fin_ir(part, nb, S, S2);
}

// next refinement's inner rhs r / |r| in FP32 and inner iterate e = 0
__device__ __forceinline__ void d_scale32_if(Geo G, const double* __restrict__ r, const double* __restrict__ S,
                                             float* __restrict__ r32, float* __restrict__ e32) {
    if (S[S_REF] == 0.0 || S[S_DONE] != 0.0) return;
    long c;
    if (!flat_interior(G, c)) return;
    r32[c] = (float)((1.0 / S[S_RN]) * r[c]);
    e32[c] = 0.0f;
}
__global__ void __launch_bounds__(256) k_scale32_if(Geo G, const double* __restrict__ r,
                                                    const double* __restrict__ S, float* __restrict__ r32,
                                                    float* __restrict__ e32) {
    d_scale32_if(G, r, S, r32, e32);
}

// Bandwidth probes.  Each lane keeps four independent 16-byte loads in
// flight (four float4 one block-width apart, so every load instruction of a
// wavefront reads 1 KB contiguous), then stores them (copy) or folds them
// into a register written once per lane (read).
__global__ void __launch_bounds__(256) k_copy16(long n, const float4* __restrict__ a, float4* __restrict__ b) {
    const long bw = blockDim.x;
    const long stride = (long)gridDim.x * 4 * bw;
    for (long t = (long)blockIdx.x * 4 * bw + threadIdx.x; t < n; t += stride) {
        float4 v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (t + q * bw < n) v[q] = a[t + q * bw];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (t + q * bw < n) b[t + q * bw] = v[q];
    }
}

__global__ void __launch_bounds__(256) k_read16(long n, const float4* __restrict__ a, float* __restrict__ out) {
    const long bw = blockDim.x;
    const long stride = (long)gridDim.x * 4 * bw;
    float acc = 0.f;
    for (long t = (long)blockIdx.x * 4 * bw + threadIdx.x; t < n; t += stride) {
        float4 v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = (t + q * bw < n) ? a[t + q * bw] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int q = 0; q < 4; ++q) acc += (v[q].x + v[q].y) + (v[q].z + v[q].w);
    }
    out[(long)blockIdx.x * bw + threadIdx.x] = acc;
}

// Eleven FP32 streams read and one written with 4-byte accesses, one element
// of each per lane and trip: the array count and access width of a row of the
// fine smoother (iterate at three zeta planes, rhs, packed metrics, spikes,
// output), without its arithmetic.  Bandwidth probe of that access pattern.
struct Streams11 {
    const float* a[11];
};
__global__ void __launch_bounds__(256) k_streams12(long n, Streams11 S, float* __restrict__ out) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
        float v[11];
#pragma unroll
        for (int q = 0; q < 11; ++q) v[q] = S.a[q][t];
        float acc = 0.f;
#pragma unroll
        for (int q = 0; q < 11; ++q) acc += v[q];
        out[t] = acc;
    }
}

// elementwise helpers over the full ghosted array
template <class TI, class TO>
__device__ __forceinline__ void d_convert(long n, const TI* __restrict__ a, TO* __restrict__ b) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long stride = (long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) b[t] = (TO)a[t];
}
template <class TI, class TO>
__global__ void k_convert(long n, const TI* __restrict__ a, TO* __restrict__ b) {
    d_convert<TI, TO>(n, a, b);
}


// b = 16-bit copy of a * scale (round to nearest)
__global__ void k_to_half(long n, const float* __restrict__ a, _Float16* __restrict__ b, float scale) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long stride = (long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) b[t] = (_Float16)(a[t] * scale);
}

// packed 16-bit smoother metrics (MetPk) of an FP32 level, scaled like k_to_half
__global__ void k_pack_met16(long n, const float* __restrict__ g11, const float* __restrict__ g12,
                             const float* __restrict__ g32, const float* __restrict__ g33,
                             const float* __restrict__ g21, const float* __restrict__ g23, float scale,
                             H2* __restrict__ a, H2* __restrict__ zz, H2* __restrict__ b) {
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        H2 v;
        v.x = (_Float16)(g11[t] * scale);
        v.y = (_Float16)(g12[t] * scale);
        a[t] = v;
        H2 vz;
        vz.x = (_Float16)(g32[t] * scale);
        vz.y = (_Float16)(g33[t] * scale);
        zz[t] = vz;
        H2 u;
        u.x = (_Float16)(g21[t] * scale);
        u.y = (_Float16)(g23[t] * scale);
        b[t] = u;
    }
}

// packed 16-bit spikes (v, w)
__global__ void k_pack_vw(long n, const float* __restrict__ v, const float* __restrict__ w, H2* __restrict__ vw) {
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        H2 u;
        u.x = (_Float16)v[t];
        u.y = (_Float16)w[t];
        vw[t] = u;
    }
}

// per-block maxima of |a| over n values (grid-stride), one partial per block
template <class TI>
__global__ void __launch_bounds__(256) k_absmax(long n, const TI* __restrict__ a, double* partials) {
    double v[1] = {0.0};
    const bool mx[1] = {true};
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long stride = (long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) v[0] = fmax(v[0], fabs((double)a[t]));
    block_reduce_store<1>(v, partials, gridDim.x, mx);
}

// gradie (fsrc/gradie.f03:28-117)
__global__ void __launch_bounds__(256) k_gradie(Geo G, const double* __restrict__ fi, CsvP V,
                                                const double* __restrict__ giac, Met<double> M,
                                                double dt, double* gra1, double* gra2, double* gra3,
                                                double* cg1, double* cg2, double* cg3) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > G.nx || j > G.ny || k > G.nz) return;
    const long c = G.idx(i, j, k);
    const long sy = G.sy, sz = G.sz;
    const bool in = i >= 1 && j >= 1 && k >= 1;
    if (in) {
        const double coef = (-0.5 * dt) / giac[c];
        double* out[3] = {gra1, gra2, gra3};
        for (int q = 0; q < 3; ++q) {
            const double* cx = V.v[q];
            const double* ex = V.v[3 + q];
            const double* zx = V.v[6 + q];
            const double s = cx[c] * (fi[c] + fi[c + 1]) - cx[c - 1] * (fi[c] + fi[c - 1]) +
                             ex[c] * (fi[c] + fi[c + sy]) - ex[c - sy] * (fi[c] + fi[c - sy]) +
                             zx[c] * (fi[c] + fi[c + sz]) - zx[c - sz] * (fi[c] + fi[c - sz]);
            out[q][c] = coef * s;
        }
    }
    if (j >= 1 && k >= 1)
        cg1[c] = M.g[0][c] * (fi[c + 1] - fi[c]) +
                 .25 * M.g[1][c] * (fi[c + 1 + sy] + fi[c + sy] - fi[c + 1 - sy] - fi[c - sy]) +
                 .25 * M.g[2][c] * (fi[c + 1 + sz] + fi[c + sz] - fi[c + 1 - sz] - fi[c - sz]);
    if (i >= 1 && k >= 1)
        cg2[c] = M.g[4][c] * (fi[c + sy] - fi[c]) +
                 .25 * M.g[3][c] * (fi[c + 1 + sy] + fi[c + 1] - fi[c - 1 + sy] - fi[c - 1]) +
                 .25 * M.g[5][c] * (fi[c + sy + sz] + fi[c + sz] - fi[c + sy - sz] - fi[c - sz]);
    if (i >= 1 && j >= 1)
        cg3[c] = M.g[8][c] * (fi[c + sz] - fi[c]) +
                 .25 * M.g[6][c] * (fi[c + 1 + sz] + fi[c + 1] - fi[c - 1 + sz] - fi[c - 1]) +
                 .25 * M.g[7][c] * (fi[c + sy + sz] + fi[c + sy] - fi[c - sy + sz] - fi[c - sy]);
}

// reference-shaped array <-> ghosted array: the box [i0..i1]x[j0..j1]x[k0..k1]
// of the ghosted layout stored compactly (first index fastest).
template <class T>
__global__ void k_pack(Geo G, int i0, int i1, int j0, int j1, int k0, const T* __restrict__ compact,
                       T* __restrict__ ghosted) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y * blockDim.y + threadIdx.y;
    const int c = blockIdx.z;
    const int ni = i1 - i0 + 1, nj = j1 - j0 + 1;
    if (a >= ni || b >= nj) return;
    ghosted[G.idx(i0 + a, j0 + b, k0 + c)] = compact[(long)a + (long)ni * ((long)b + (long)nj * c)];
}

template <class T>
__global__ void k_unpack(Geo G, int i0, int i1, int j0, int j1, int k0, const T* __restrict__ ghosted,
                         T* __restrict__ compact) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y * blockDim.y + threadIdx.y;
    const int c = blockIdx.z;
    const int ni = i1 - i0 + 1, nj = j1 - j0 + 1;
    if (a >= ni || b >= nj) return;
    compact[(long)a + (long)ni * ((long)b + (long)nj * c)] = ghosted[G.idx(i0 + a, j0 + b, k0 + c)];
}

}  // namespace plesmg
