// ples_mg.hip -- host side of the MI355X pressure multigrid: level
// hierarchy, V-cycle, FP64 refinement loop (GCR-accelerated), C ABI.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <array>
#include <map>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <memory>
#include <new>
#include <mutex>
#include <thread>
#include <chrono>
#include <deque>
#include <unordered_map>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/ples_mg.h"
#include "group_launch.hpp"
#include "mg_kernels.hpp"
#include "proj_kernels.hpp"

using namespace plesmg;

namespace {

struct Level {
    Geo G{};
    bool f32 = false;  // V-cycle precision of this level
    double pot = 1.0;  // node stride in eta and zeta (2^n); the operator's 1/s
    int sx = 1, sy = 1;  // node strides in xi and eta: pot, or pot / 2 below a level that kept the count
    double* g64[9] = {};
    float* g32[9] = {};
    _Float16* g16[9] = {};  // smoother copy of g32 * 4^m (FP32 levels, smoother_fp16)
    double h16s = 1.0;      // 2^-m: 1/pot of the smoother reading g16
    bool use16 = false;     // the smoother reads g16
    bool x0 = false;        // g13 and g31 identically zero: stencil kernels skip them
    void* kd = nullptr;     // coarsest level: dense exact solve map (ghosted x interior, rows padded to kdp)
    bool kd32 = false;      // the map is FP32 (mixed solver), else FP64
    int kdp = 0;            // padded row length of the map
    int* kdi = nullptr;     // ghosted index of each interior cell (the map's columns)
    _Float16* v16 = nullptr;  // 16-bit copies of lv_v, lv_w (levels with g16)
    _Float16* w16 = nullptr;
    H2* pk_a = nullptr;       // packed 16-bit smoother metrics (levels with g16 and x0)
    H2* pk_z = nullptr;
    H2* pk_b = nullptr;
    H2* vw16 = nullptr;       // packed 16-bit spikes (levels with pk_a)
    float* invm = nullptr;
    float* cpr = nullptr;
    float* lv_v = nullptr;  // segmented line-solve responses
    float* lv_w = nullptr;
    int P = 1;              // eta segments per column (smoother, outer residual)
    int rP = 0, rbx = 0;    // V-cycle residual launch: segments x lanes (0 = as the smoother)
    void* p[2] = {nullptr, nullptr};
    void* b = nullptr;
    void* r = nullptr;
    int cur = 0;
    long ncell() const { return (long)G.nx * G.ny * G.nz; }
};

struct Rec {
    int fam;
    double bytes;
    hipEvent_t a, b;
};

// one recorded kernel launch of a problem: the grouped launcher of its device
// body, the launch shape and the problem's table entry (stop flag and
// arguments, bytes of a GEntry)
struct KRec {
    hipError_t (*go)(const void* tab, int n, dim3 g, dim3 b, size_t shm, hipStream_t s);
    dim3 g, b;
    size_t shm;
    std::vector<unsigned char> e;
};

}  // namespace

struct ples_mg_handle {
    int dev = 0;
    hipStream_t s = nullptr;
    ples_mg_opts o{};
    int jx = 0, jy = 0, jz = 0, ip = 0, jp = 1, kp = 0, L = 1;
    int Lv = 1;  // levels the V-cycle visits: with coarse_exact it stops at the
                 // first coarse level of at most kCoarseDenseMax cells, solved exactly
    std::vector<Level> lv;
    double* nodes[3] = {};
    double* x = nullptr;  // resident iterate (FP64, ghosted)
    double* bvec = nullptr;
    double* rvec = nullptr;
    double* Z[2] = {};
    double* W[2] = {};
    double* cs[6] = {};
    bool have_cs = false;
    double cs_sum = 0.0;  // sum of the wall fluxes cs (outflow - inflow) of the resident cs
    double dt = 1.0;
    double omega_now = 0.8;               // damping of the sweep being launched
    int halo_passes32 = 6;                // wall-ghost passes of the FP32 inner direction
    // (Z, W) pair of the next FP64 / FP32 refinement step.  Steps alternate the
    // pair across solves and refinements too: a step leaves the levels' ping-pong
    // selectors where the other pair's captured graph starts (the dense coarse
    // solve flips its level once per step), so restarting at pair 0 after an odd
    // number of steps would drop and recapture both step graphs.  A restart of
    // the GCR history (no previous direction) reads the other pair only through
    // terms multiplied by zero, so the pair does not change the results.
    int qnext[2] = {0, 0};
    void* snap = nullptr;                 // fine-level eta-wall ghost rows (k_eta_ghosts start)
    size_t snap_bytes = 0;
    double cheb_lo = 0.0, cheb_hi = 0.0;  // Chebyshev sweep schedule interval (0 = constant omega)
    double* S = nullptr;   // device scalars
    double* Sh = nullptr;  // pinned host mirror
    double* part = nullptr;
    int part_n = 0;
    size_t bytes = 0;
    std::vector<std::pair<void*, size_t>> allocs;
    int pooled = -1;  // buffers from the stream-ordered pool (1) or hipMalloc (0); -1 before the first
    unsigned timing = 0;  // bit f: time kernel family f
    int timing_every = 1;  // with timing on: every n-th refinement step is launched eagerly and timed
    long timing_step = 0;  // refinement steps enqueued since timing was set
    std::vector<Rec> recs;
    std::vector<KRec>* rec = nullptr;  // set: kernel launches are recorded, not enqueued
    // timing of a single-dispatch scope: the next launch records these events
    // at its own start and end (per-dispatch timestamps, no gap of their own)
    hipEvent_t kt_a = nullptr, kt_b = nullptr;
    bool kt_fired = false;
    // slot of S whose nonzero value makes a grouped launch skip this problem
    // (S_DONE; S_NOREF for the refinement tail of device-driven refinement)
    int skip = S_DONE;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    hipEvent_t sev[4] = {};          // read-back ring of the refinement loop
    hipGraphExec_t gexec[2] = {};    // captured refinement step per (Z, W) pair
    std::vector<int> gin[2], gout[2];
    double last_removed = 0.0;
    int inner = 1;                   // V-cycles per refinement step (resolved from opts)
    int lds_max = 65536;             // LDS bytes per workgroup for the coarsest-level solver
    // FP32 inner FGCR of the mixed solver (level-0 FP32 vectors, own scalars)
    bool ir32 = false;
    float* e32 = nullptr;            // inner correction
    float* rr32 = nullptr;           // inner residual (the V-cycle's level-0 rhs)
    float* Z32[2] = {};
    float* W32[2] = {};
    double* S2 = nullptr;
    const float* rhs32 = nullptr;    // set while a V-cycle runs on rr32
    hipGraphExec_t gexec32[2] = {};
    std::vector<int> gin32[2], gout32[2];
    // projection-step state (allocated on first use): level-1 face vectors
    // and Jacobian, cell velocities, face fluxes, increments, gradients
    bool pj_ready = false;
    double* pj_csv[9] = {};
    double* pj_giac = nullptr;
    double* pj_vel[3] = {};
    double* pj_fc[3] = {};
    double* pj_del[3] = {};
    double* pj_gra[3] = {};
    double* pj_cgra[3] = {};
    double* pj_dmax = nullptr;
};

namespace {

using H = ples_mg_handle;

int hip_status(hipError_t e) {
    if (e == hipSuccess) return PLES_MG_OK;
    if (e == hipErrorOutOfMemory) return PLES_MG_ERR_OOM;
    fprintf(stderr, "ples_mg: HIP error %s\n", hipGetErrorString(e));
    return PLES_MG_ERR_HIP;
}

#define CK(x)                                  \
    do {                                       \
        hipError_t e_ = (x);                   \
        if (e_ != hipSuccess) return hip_status(e_); \
    } while (0)
#define CKS(x)                          \
    do {                                \
        int st_ = (x);                  \
        if (st_ != PLES_MG_OK) return st_; \
    } while (0)

// Device buffers of a handle come from the device's stream-ordered memory
// pool, allocated and freed in order on the handle's stream: a cold start
// (about a hundred buffers) and an eviction's frees cost pool bookkeeping
// instead of a driver allocation and a device-wide synchronising free each.
// The pool keeps freed memory for the next handles (release threshold at the
// maximum), so the process holds its peak; a device whose runtime lacks the
// pool falls back to hipMalloc / hipFree.
std::mutex g_pool_mu;
std::vector<int> g_pool_state;  // per device: 0 unknown, 1 pool, 2 plain allocations

bool use_pool(int dev) {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    if ((int)g_pool_state.size() <= dev) g_pool_state.resize((size_t)dev + 1, 0);
    if (g_pool_state[(size_t)dev] == 0) {
        hipMemPool_t pool = nullptr;
        uint64_t thr = UINT64_MAX;
        const bool ok = hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess && pool &&
                        hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr) == hipSuccess;
        if (!ok) (void)hipGetLastError();
        g_pool_state[(size_t)dev] = ok ? 1 : 2;
    }
    return g_pool_state[(size_t)dev] == 1;
}

int dalloc(H* h, void** p, size_t nbytes) {
    *p = nullptr;
    if (nbytes == 0) return PLES_MG_OK;
    if (h->pooled < 0) h->pooled = use_pool(h->dev) ? 1 : 0;
    hipError_t e = h->pooled ? hipMallocAsync(p, nbytes, h->s) : hipMalloc(p, nbytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return PLES_MG_ERR_OOM;
    }
    e = hipMemsetAsync(*p, 0, nbytes, h->s);
    if (e != hipSuccess) return hip_status(e);
    h->allocs.push_back({*p, nbytes});
    h->bytes += nbytes;
    return PLES_MG_OK;
}

void dfree(H* h, void* p) {
    if (!p) return;
    for (auto it = h->allocs.begin(); it != h->allocs.end(); ++it)
        if (it->first == p) {
            h->bytes -= it->second;
            h->allocs.erase(it);
            break;
        }
    if (h->pooled > 0) (void)hipFreeAsync(p, h->s);
    else (void)hipFree(p);
}

Geo make_geo(int nx, int ny, int nz, int ip, int jp, int kp) {
    Geo G;
    G.nx = nx; G.ny = ny; G.nz = nz;
    G.ip = ip; G.jp = jp; G.kp = kp;
    G.sy = nx + 2;
    G.sz = (long)(nx + 2) * (ny + 2);
    G.n = G.sz * (nz + 2);
    return G;
}

int pow2ceil(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// column launch: block (bx, by) over (i, k)
struct ColLaunch {
    dim3 grid, block;
    int nblocks;
};
// one thread per interior cell (kernels using cell_of)
ColLaunch cell_launch(const Geo& G) {
    ColLaunch c;
    c.block = dim3(64, 4, 1);
    c.grid = dim3((unsigned)((G.nx + 63) / 64), (unsigned)((G.ny + 3) / 4), (unsigned)G.nz);
    c.nblocks = (int)(c.grid.x * c.grid.y * c.grid.z);
    return c;
}

ColLaunch col_launch(const Geo& G) {
    ColLaunch c;
    const int bx = std::min(64, pow2ceil(G.nx));
    const int by = std::max(1, std::min(256 / bx, pow2ceil(G.nz)));
    c.block = dim3(bx, by, 1);
    c.grid = dim3((G.nx + bx - 1) / bx, (G.nz + by - 1) / by, 1);
    c.nblocks = (int)(c.grid.x * c.grid.y);
    return c;
}

// lanes per block of a segmented launch: 256/P rounded down to a power of
// two, clamped to [32, 128] (SegBX on the device)
int seg_bx(int P) {
    const int raw = 256 / P;
    return raw >= 128 ? 128 : (raw >= 64 ? 64 : 32);
}

// segmented-column launch: block (bx, P) over (i, segment), one zeta plane per block row
ColLaunch seg_launch(const Geo& G, int P) {
    ColLaunch c;
    const int bxmax = seg_bx(P);  // = SegBX<P>
    const int bx = std::min(bxmax, pow2ceil(G.nx));
    c.block = dim3(bx, P, 1);
    c.grid = dim3((G.nx + bx - 1) / bx, G.nz, 1);
    c.nblocks = (int)(c.grid.x * c.grid.y);
    return c;
}

// eta segments per column, fixed rule: among the counts P with a register
// smoother variant (ny / P in {4, 8, 16, 32}, P in 1..16 dividing ny), the
// fewest that still give the level ncol * P >= 2^16 threads (long register
// rows where the grid fills the device), else 8 rows per segment where
// possible (the small, latency-bound levels); the autotuner's choice on the
// benchmark shapes.  Row counts without a register variant: choose_segments.
int choose_segments(int ny);
int fixed_segments(int ny, long ncol) {
    std::vector<int> cand;
    for (int P : {1, 2, 3, 4, 6, 8, 12, 16}) {
        const int m = (ny % P == 0) ? ny / P : 0;
        if (m == 4 || m == 8 || m == 16 || m == 32) cand.push_back(P);
    }
    if (cand.empty()) return choose_segments(ny);
    for (int P : cand)
        if (ncol * P >= (1L << 16)) return P;
    for (int P : cand)
        if (ny / P == 8) return P;
    return cand.back();
}

// eta segments per column: the largest power of two up to 16 that divides ny
// and leaves at least 16 rows per segment
int choose_segments(int ny) {
    int P = 1;
    while (P < 16 && ny % (2 * P) == 0 && ny / (2 * P) >= 16) P *= 2;
    // a segment length without a register variant: try three-way splits,
    // which turn the 3*2^k row counts (48, 96, ...) into 4..32
    const int m = ny / P;
    if (m != 4 && m != 8 && m != 16 && m != 32)
        for (int Q : {3, 6, 12})
            if (ny % Q == 0) {
                const int mq = ny / Q;
                if (mq == 4 || mq == 8 || mq == 16 || mq == 32) return Q;
            }
    return P;
}

LineF line_f(const Level& L) { return LineF{L.invm, L.cpr, L.lv_v, L.lv_w, L.v16, L.w16, L.vw16}; }

// per-launch event timing
// Kernel time of a scope: events recorded on the stream before and after it,
// or, for a scope of one kernel launch (one = true), the launch's own start
// and end timestamps (the kernel's duration, without the gaps that event
// records between launches add).
struct KT {
    H* h;
    int fam;
    double bytes;
    hipEvent_t a = nullptr, b = nullptr;
    bool on, one;
    KT(H* h_, int fam_, double bytes_, bool one_ = false)
        : h(h_), fam(fam_), bytes(bytes_), on(((h_->timing >> fam_) & 1u) != 0), one(one_) {
        if (on) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            if (one) {
                h->kt_a = a;
                h->kt_b = b;
                h->kt_fired = false;
            } else {
                (void)hipEventRecord(a, h->s);
            }
        }
    }
    ~KT() {
        if (!on) return;
        if (one) {
            h->kt_a = h->kt_b = nullptr;
            if (!h->kt_fired) {  // nothing was launched (recorded for a grouped launch instead)
                (void)hipEventDestroy(a);
                (void)hipEventDestroy(b);
                return;
            }
        } else {
            (void)hipEventRecord(b, h->s);
        }
        h->recs.push_back(Rec{fam, bytes, a, b});
    }
};

// Launch of device body F (kernel kern, its single-problem wrapper) with
// arguments a: enqueued on the handle's stream, or recorded for a grouped
// launch while h->rec is set (2-D grids only: blockIdx.z selects the problem)
template <auto F, int LB, int WPE, class K, class... A>
void klaunch(H* h, K kern, dim3 g, dim3 b, size_t shm, A... a) {
    if (h->rec) {
        const GEntry<A...> e{h->S + h->skip, make_pack(a...)};
        KRec r{&launch_group<F, LB, WPE, A...>, g, b, shm, {}};
        r.e.assign((const unsigned char*)&e, (const unsigned char*)&e + sizeof e);
        if (g.z != 1) r.go = nullptr;  // refused when the batch groups the records
        h->rec->push_back(std::move(r));
        return;
    }
    if (h->kt_a && !h->kt_fired) {
        hipExtLaunchKernelGGL(kern, g, b, (std::uint32_t)shm, h->s, h->kt_a, h->kt_b, 0u, a...);
        h->kt_fired = true;
        return;
    }
    hipLaunchKernelGGL(kern, g, b, shm, h->s, a...);
}
// KLAUNCH(h, name, TPL(template args), LB, WPE, grid, block, shm, args...):
// d_name / k_name of the kernel headers
#define TPL(...) <__VA_ARGS__>
#define KLAUNCH(h, NAME, TA, LB, WPE, g, b, shm, ...) \
    klaunch<&d_##NAME TA, LB, WPE>(h, &k_##NAME TA, g, b, shm, __VA_ARGS__)
#define KLAUNCH0(h, NAME, LB, WPE, g, b, shm, ...) klaunch<&d_##NAME, LB, WPE>(h, &k_##NAME, g, b, shm, __VA_ARGS__)

// metric arrays the level's stencil kernels read
double nmet(const Level& L) { return L.x0 ? 7.0 : 9.0; }

template <class T>
Met<T> met(const Level& L);
template <>
Met<float> met<float>(const Level& L) {
    Met<float> m;
    for (int q = 0; q < 9; ++q) m.g[q] = L.g32[q];
    return m;
}
template <>
Met<double> met<double>(const Level& L) {
    Met<double> m;
    for (int q = 0; q < 9; ++q) m.g[q] = L.g64[q];
    return m;
}

// ---------------------------------------------------------------- halo
long face_set_size(const Geo& G, int set) {
    if (set == 0) return 2L * G.nx * G.nz;
    if (set == 1) return G.ip == 1 ? 2L * G.ny * G.nz : 0L;
    return G.kp == 1 ? 2L * G.nx * G.ny : 0L;
}

// Eta-wall ghosts of a level periodic in xi and zeta by one k_eta_ghosts
// launch (the passes of every tile in LDS); false where that does not apply
// (xi or zeta walls, or a tile halo beyond 64 KB of LDS).
template <class T, class TM>
bool eta_ghosts(H* h, const Geo& G, const Met<TM>& m, T* p, const BCs& bc, int passes) {
    if (G.ip != 0 || G.kp != 0 || G.jp != 1 || passes < 1) return false;
    constexpr int TX = sizeof(T) == 4 ? 64 : 32, TK = 8;  // the FP32 tile's halo fits 64 KB of LDS at 64 wide
    const int hw = 2 * passes;
    const size_t lds = (size_t)(TX + 2 * hw) * (TK + 2 * hw) * 4 * sizeof(T);
    const long nrow = 2L * (G.nx + 2) * (G.nz + 2);
    if (lds > 65536 || (size_t)nrow * sizeof(T) > h->snap_bytes) return false;
    KLAUNCH(h, ghost_rows, TPL(T), 256, 1, dim3((unsigned)((nrow + 255) / 256)), dim3(256), 0, G, (const T*)p,
            (T*)h->snap);
    const int nb = 2 * ((G.nx + TX - 1) / TX) * ((G.nz + TK - 1) / TK);
    KLAUNCH(h, eta_ghosts, TPL(T, TM, TX, TK), 1024, 1, dim3(nb), dim3(1024), lds, G, m, p, (const T*)h->snap, bc,
            passes);
    return true;
}

// interior_images: the periodic images of the interior are already valid
// (a smoother output), only the wall ghosts and their images are needed
template <class T>
int halo(H* h, const Level& L, T* p, double an_over_dt, bool with_cs, int passes, bool interior_images = false) {
    const Geo& G = L.G;
    BCs bc;
    for (int q = 0; q < 6; ++q) bc.cs[q] = (with_cs && h->have_cs) ? h->cs[q] : nullptr;
    bc.an_over_dt = an_over_dt;
    {
        const long NX = G.nx + 2, NY = G.ny + 2, NZ = G.nz + 2;
        const long ne = 2 * NY * NZ + 2 * NX * NZ + 2 * NX * NY;
        KT kt(h, PLES_MG_K_HALO, (double)(2L * G.nx * G.nz * 8 + (interior_images ? 0 : ne * 2)) * sizeof(T));
        if (eta_ghosts<T, T>(h, G, met<T>(L), p, bc, passes)) {
            if (!interior_images)
                KLAUNCH(h, halo_edges, TPL(T), 1024, 1, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, G, p);
            CK(hipGetLastError());
            return PLES_MG_OK;
        }
    }
    long nf = 2L * G.nx * G.nz;
    if (G.ip == 1) nf += 2L * G.ny * G.nz;
    if (G.kp == 1) nf += 2L * G.nx * G.ny;
    const long NX = G.nx + 2, NY = G.ny + 2, NZ = G.nz + 2;
    const long ne = 2 * NY * NZ + 2 * NX * NZ + 2 * NX * NY;
    const double bytes = (double)(nf * 6 + ne * 2) * passes * sizeof(T);
    KT kt(h, PLES_MG_K_HALO, bytes);
    // faces (red, black), then the edges that the next pass reads
    for (int ps = 0; ps < passes; ++ps) {
        for (int set = 0; set < 3; ++set) {
            const long nset = face_set_size(G, set);
            if (nset == 0) continue;
            for (int color = 0; color < 2; ++color)
                KLAUNCH(h, halo_face, TPL(T, T), 1024, 1, dim3((unsigned)((nset + 255) / 256)), dim3(256), 0, G,
                        met<T>(L), p, color, bc, set, 0);
        }
        KLAUNCH(h, halo_edges, TPL(T), 1024, 1, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, G, p);
    }
    CK(hipGetLastError());
    return PLES_MG_OK;
}

// exact ghost fill of an FP64 fine-level vector with the outer metrics.
// fast: periodic xi and zeta with valid interior images -- only the eta-wall
// ghosts are iterated (neighbours read through the periodic wrap) and each
// pass writes their images, no edge launch.
int halo64(H* h, double* p, bool with_bc, bool fast = false, int passes_override = 0) {
    const Level& L0 = h->lv[0];
    Met<double> m = met<double>(L0);
    const Geo& G = L0.G;
    BCs bc;
    for (int q = 0; q < 6; ++q) bc.cs[q] = (with_bc && h->have_cs) ? h->cs[q] : nullptr;
    bc.an_over_dt = with_bc ? 1.0 / h->dt : 0.0;
    long nf = 2L * G.nx * G.nz;
    if (G.ip == 1) nf += 2L * G.ny * G.nz;
    if (G.kp == 1) nf += 2L * G.nx * G.ny;
    const long NX = G.nx + 2, NY = G.ny + 2, NZ = G.nz + 2;
    const long ne = 2 * NY * NZ + 2 * NX * NZ + 2 * NX * NY;
    fast = fast && G.ip == 0 && G.kp == 0;
    // every pass counts: the outer GCR needs the ghosts at their fixed point
    // to machine precision, or the attainable residual stalls
    const int passes = passes_override > 0 ? passes_override : std::max(1, h->o.halo_passes);
    {
        KT kt(h, PLES_MG_K_HALO, (double)(2L * G.nx * G.nz * 8 + (fast ? 0 : ne * 2)) * 8.0);
        if (eta_ghosts<double, double>(h, G, m, p, bc, passes)) {
            if (!fast)
                KLAUNCH(h, halo_edges, TPL(double), 1024, 1, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, G, p);
            CK(hipGetLastError());
            return PLES_MG_OK;
        }
    }
    KT kt(h, PLES_MG_K_HALO, (double)(nf * 6 + (fast ? 0 : ne * 2)) * passes * 8.0);
    for (int ps = 0; ps < passes; ++ps) {
        for (int set = 0; set < 3; ++set) {
            const long nset = face_set_size(G, set);
            if (nset == 0) continue;
            for (int color = 0; color < 2; ++color)
                KLAUNCH(h, halo_face, TPL(double, double), 1024, 1, dim3((unsigned)((nset + 255) / 256)), dim3(256),
                        0, G, m, p, color, bc, set, fast ? 1 : 0);
        }
        if (!fast)
            KLAUNCH(h, halo_edges, TPL(double), 1024, 1, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, G, p);
    }
    CK(hipGetLastError());
    return PLES_MG_OK;
}

// ---------------------------------------------------------------- V-cycle pieces
// register variant of a segment length: 4, 8, 16 or 32 rows; other
// lengths are staged through memory (measured faster than the register
// variant with guarded row loops, whose larger register arrays lower the
// occupancy)
int yline_mr(int m) { return (m == 4 || m == 8 || m == 16 || m == 32) ? m : 0; }

// one smoother sweep with metrics m (MR = 0: the staged variant)
template <int P, int MR, class T, class TM, class TR>
void yline_go(H* h, Level& L, const Met<TM>& m, double inv_pot, const TR* rhs, int zin) {
    const ColLaunch cl = seg_launch(L.G, P);
    if constexpr (MR > 0) {
        auto go = [&](auto x0) {
            constexpr bool X0 = decltype(x0)::value;
            KLAUNCH(h, yline, TPL(P, MR, T, Met<TM>, TR, X0), 512, (YlineWpe<T, Met<TM>, P, MR, X0>::value), cl.grid, cl.block,
                    0, L.G, m, line_f(L), (const T*)L.p[L.cur], (T*)L.p[1 - L.cur], rhs, (T)h->omega_now,
                    (T)inv_pot, zin);
        };
        // the variant without the g13, g31 loads: measured faster for the 16-bit-metric
        // and FP64 smoothers, slower for the FP32-metric one (kept with the loads); the
        // 16-bit one reads the packed metrics and spikes
        if constexpr (sizeof(TM) == 2) {
            if (L.x0 && L.pk_a && (P == 1 || L.vw16)) {
                const MetPk mp{L.pk_a, L.pk_z, L.pk_b, L.g16[4]};
                KLAUNCH(h, yline, TPL(P, MR, T, MetPk, TR, true), 512, (YlineWpe<T, MetPk, P, MR, true>::value), cl.grid,
                        cl.block, 0, L.G, mp, line_f(L), (const T*)L.p[L.cur], (T*)L.p[1 - L.cur], rhs,
                        (T)h->omega_now, (T)inv_pot, zin);
            } else if (L.x0) {
                go(std::true_type{});
            } else {
                go(std::false_type{});
            }
        } else if constexpr (sizeof(T) == 8) {
            if (L.x0) go(std::true_type{});
            else go(std::false_type{});
        } else {
            go(std::false_type{});
        }
    } else
        KLAUNCH(h, yline_staged, TPL(P, T, TM, TR), 512, 1, cl.grid, cl.block, 0, L.G, m, line_f(L),
                (const T*)L.p[L.cur], (T*)L.p[1 - L.cur], rhs, (T)h->omega_now, (T)inv_pot, zin);
}

template <int P, int MR, class T, class TR>
void launch_yline_mr(H* h, Level& L, const TR* rhs, int zin) {
    if constexpr (sizeof(T) == 4) {
        if (L.use16 && L.g16[0]) {  // 16-bit metrics scaled by 4^m: 1/pot scaled by 2^-m
            Met<_Float16> m;
            for (int q = 0; q < 9; ++q) m.g[q] = L.g16[q];
            yline_go<P, MR, T, _Float16, TR>(h, L, m, L.h16s / L.pot, rhs, zin);
            return;
        }
    }
    yline_go<P, MR, T, T, TR>(h, L, met<T>(L), 1.0 / L.pot, rhs, zin);
}

template <int P, class T, class TR>
void launch_yline(H* h, Level& L, const TR* rhs, int zin) {
    switch (yline_mr(L.G.ny / P)) {
        case 4: launch_yline_mr<P, 4, T, TR>(h, L, rhs, zin); break;
        case 8: launch_yline_mr<P, 8, T, TR>(h, L, rhs, zin); break;
        case 16: launch_yline_mr<P, 16, T, TR>(h, L, rhs, zin); break;
        case 32: launch_yline_mr<P, 32, T, TR>(h, L, rhs, zin); break;
        default: launch_yline_mr<P, 0, T, TR>(h, L, rhs, zin); break;
    }
}

// segment lengths with a register-resident smoother variant
bool reg_rows(int m) { return yline_mr(m) != 0; }

template <class T, class TR>
int smooth(H* h, Level& L, const TR* rhs, int nu, bool zero_start, bool fine) {
    const double sm = (L.use16 && L.g16[0]) ? 2.0 : sizeof(T), sp = sizeof(T), sr = sizeof(TR);
    const double N = (double)L.ncell();
    const double seg = L.P > 1 ? ((L.use16 && L.g16[0]) ? 4.0 : 8.0) : 0.0;  // v, w reads of the last pass
    for (int s = 0; s < nu; ++s) {
        const int zin = (zero_start && s == 0) ? 1 : 0;
        // damping of sweep s: the constant omega, or with the Chebyshev schedule
        // 1 / (the roots of the degree-nu Chebyshev polynomial on [lo, hi]),
        // so the nu sweeps together damp that part of the spectrum of
        // M_line^-1 A by 1 / T_nu((hi + lo) / (hi - lo)) (1/9 for nu = 3)
        h->omega_now = h->o.omega;
        if (h->cheb_hi > 0.0 && nu > 1) {
            const double c = 0.5 * (h->cheb_hi + h->cheb_lo), d = 0.5 * (h->cheb_hi - h->cheb_lo);
            h->omega_now = 1.0 / (c + d * std::cos(M_PI * (2.0 * s + 1.0) / (2.0 * nu)));
        }
        // forward: rhs, metrics, iterate, invm, cpr; last pass: (v, w), write.  The staged variant
        // adds the e write/read, the cpr re-read and the q write/read.  FP32 register variants
        // factor on the fly (no invm, cpr; the zero-start sweep reads g11, g33 instead).
        const int MR = yline_mr(L.G.ny / L.P);
        const int bxs = seg_bx(L.P);
        const bool otf = sizeof(T) == 4 && MR > 0 && MR * L.P * bxs * 4 <= 32768;
        const double stage = MR > 0 ? 0.0 : (4 * sp + 4);
        const double fac = otf ? 0.0 : 8.0;
        const double bytes = zin ? N * (sr + sm + (otf ? 2 * sm : 8.0) + sp + seg + stage)
                                 : N * (((MR > 0 && (sm == 2.0 || sizeof(T) == 8)) ? nmet(L) : 9.0) * sm + sp + sr +
                                        fac + sp + seg + stage);
        {
            KT kt(h, fine ? PLES_MG_K_SMOOTH_FINE : PLES_MG_K_SMOOTH, bytes, true);
            switch (L.P) {
                case 16: launch_yline<16, T, TR>(h, L, rhs, zin); break;
                case 12: launch_yline<12, T, TR>(h, L, rhs, zin); break;
                case 8: launch_yline<8, T, TR>(h, L, rhs, zin); break;
                case 6: launch_yline<6, T, TR>(h, L, rhs, zin); break;
                case 4: launch_yline<4, T, TR>(h, L, rhs, zin); break;
                case 3: launch_yline<3, T, TR>(h, L, rhs, zin); break;
                case 2: launch_yline<2, T, TR>(h, L, rhs, zin); break;
                default: launch_yline<1, T, TR>(h, L, rhs, zin); break;
            }
        }
        L.cur ^= 1;
        if (L.G.ip == 1 || L.G.kp == 1) CKS(halo<T>(h, L, (T*)L.p[L.cur], 0.0, false, 1));
    }
    CK(hipGetLastError());
    return PLES_MG_OK;
}

template <class T, class TR>
int vresid(H* h, Level& L, const TR* rhs, bool fine) {
    ColLaunch cl = seg_launch(L.G, L.P);
    int P = L.P;
    if (L.rP > 0) {  // tuned launch shape (the residual of a cell does not depend on it)
        P = L.rP;
        cl.block = dim3(L.rbx, P, 1);
        cl.grid = dim3((L.G.nx + L.rbx - 1) / L.rbx, L.G.nz, 1);
    }
    const double N = (double)L.ncell();
    KT kt(h, fine ? PLES_MG_K_RESID_FINE : PLES_MG_K_RESID, N * (nmet(L) * sizeof(T) + 2 * sizeof(T) + sizeof(TR)), true);
    auto go = [&](auto x0) {
        constexpr bool X0 = decltype(x0)::value;
        KLAUNCH(h, resid, TPL(0, T, T, TR, T, double, X0), 512, 1, cl.grid, cl.block, 0, L.G, met<T>(L),
                (const T*)L.p[L.cur], rhs, (T*)L.r, (T)(1.0 / L.pot), (const double*)nullptr, (const double*)nullptr,
                (double*)nullptr, 0, P);
    };
    if (L.x0) go(std::true_type{});
    else go(std::false_type{});
    CK(hipGetLastError());
    return PLES_MG_OK;
}

template <class TF, class TC>
int restrict_to(H* h, Level& F, Level& Cc) {
    dim3 block(std::min(64, pow2ceil(Cc.G.nx)), 4);
    const int gx = (Cc.G.nx + block.x - 1) / block.x;
    dim3 grid(gx * Cc.G.nz, (Cc.G.ny + 3) / 4);
    KT kt(h, PLES_MG_K_TRANSFER, (double)F.ncell() * sizeof(TF) + (double)Cc.ncell() * sizeof(TC), true);
    KLAUNCH(h, restrict, TPL(TF, TC), 256, 1, grid, block, 0, Cc.G, F.G, (const TF*)F.r, (TC*)Cc.b,
            F.G.nx / Cc.G.nx, F.G.ny / Cc.G.ny, gx);
    CK(hipGetLastError());
    return PLES_MG_OK;
}

// V-cycle residual of level F restricted straight into Cc.b (one launch, no
// stored fine residual); returns 1 where the launch shape does not apply (odd
// segment rows, planes or lanes) and the caller runs the two kernels instead
template <class T, class TR, class TC>
int resid_restrict(H* h, Level& F, Level& Cc, const TR* rhs, bool fine) {
    const int P = F.rP > 0 ? F.rP : F.P;
    int bx = F.rbx > 0 ? F.rbx : seg_bx(P);
    const int m = F.G.ny / P;
    const int fx = F.G.nx / Cc.G.nx;  // 1: the coarse level keeps the xi count
    if (F.G.ny == Cc.G.ny) return 1;  // eta count kept too: residual, then the two-child restriction
    if (m < 2 || (m & 1) || (F.G.nz & 1) || (fx == 2 && (F.G.nx & 1)) || F.G.ny % P) return 1;
    while (bx * P > 512 && bx > 2) bx >>= 1;
    if ((bx * P) % 64 != 0 || (fx == 2 && (bx & 1))) return 1;
    const dim3 block((unsigned)bx, (unsigned)P, 1), grid((unsigned)((F.G.nx + bx - 1) / bx), (unsigned)(F.G.nz / 2));
    const double N = (double)F.ncell();
    if constexpr (sizeof(T) == 4) {
        if (F.x0 && F.pk_a) {  // the smoother's packed 16-bit metrics: 14 B per cell instead of 28
            KT kt(h, fine ? PLES_MG_K_RESID_FINE : PLES_MG_K_RESID,
                  N * (14.0 + sizeof(T) + sizeof(TR)) + (double)Cc.ncell() * sizeof(TC), true);
            const MetPk mp{F.pk_a, F.pk_z, F.pk_b, F.g16[4]};
            KLAUNCH(h, resid_restrict_pk, TPL(TR, TC), 512, 1, grid, block, 0, F.G, mp, (const float*)F.p[F.cur],
                    rhs, Cc.G, (TC*)Cc.b, (float)(F.h16s / F.pot), P, fx);  // metrics scaled by 2^2m
            CK(hipGetLastError());
            return PLES_MG_OK;
        }
    }
    KT kt(h, fine ? PLES_MG_K_RESID_FINE : PLES_MG_K_RESID,
          N * (nmet(F) * sizeof(T) + sizeof(T) + sizeof(TR)) + (double)Cc.ncell() * sizeof(TC), true);
    auto go = [&](auto x0) {
        constexpr bool X0 = decltype(x0)::value;
        KLAUNCH(h, resid_restrict, TPL(T, T, TR, TC, X0), 512, 1, grid, block, 0, F.G, met<T>(F),
                (const T*)F.p[F.cur], rhs, Cc.G, (TC*)Cc.b, (T)(1.0 / F.pot), P, fx);
    };
    if (F.x0) go(std::true_type{});
    else go(std::false_type{});
    CK(hipGetLastError());
    return PLES_MG_OK;
}

template <class TC, class TF>
int prolong_to(H* h, Level& Cc, Level& F) {
    if (Cc.G.ny == F.G.ny) {  // coarsened in zeta only: 2 fine children per coarse cell
        dim3 block(std::min(64, pow2ceil(Cc.G.nx + 2)), 4);
        const int gx = (Cc.G.nx + 2 + block.x - 1) / block.x;
        dim3 grid(gx * (Cc.G.nz + 1), (Cc.G.ny + 2 + 3) / 4);
        KT kt(h, PLES_MG_K_TRANSFER, 2.0 * F.ncell() * sizeof(TF) + (double)Cc.ncell() * sizeof(TC), true);
        KLAUNCH(h, prolong_z, TPL(TC, TF), 256, 1, grid, block, 0, Cc.G, F.G, (const TC*)Cc.p[Cc.cur],
                (TF*)F.p[F.cur], gx);
        CK(hipGetLastError());
        return PLES_MG_OK;
    }
    if (Cc.G.nx == F.G.nx) {  // coarsened in eta and zeta only: 4 fine children per coarse cell
        dim3 block(std::min(64, pow2ceil(Cc.G.nx + 2)), 4);
        const int gx = (Cc.G.nx + 2 + block.x - 1) / block.x;
        dim3 grid(gx * (Cc.G.nz + 1), (Cc.G.ny + 1 + 3) / 4);
        KT kt(h, PLES_MG_K_TRANSFER, 2.0 * F.ncell() * sizeof(TF) + (double)Cc.ncell() * sizeof(TC), true);
        KLAUNCH(h, prolong_yz, TPL(TC, TF), 256, 1, grid, block, 0, Cc.G, F.G, (const TC*)Cc.p[Cc.cur],
                (TF*)F.p[F.cur], gx);
        CK(hipGetLastError());
        return PLES_MG_OK;
    }
    // one thread per coarse block (I, J, K) in 0..nc (ghosts included): 8 fine children each
    dim3 block(std::min(64, pow2ceil(Cc.G.nx + 1)), 4);
    const int gx = (Cc.G.nx + 1 + block.x - 1) / block.x;
    dim3 grid(gx * (Cc.G.nz + 1), (Cc.G.ny + 1 + 3) / 4);
    KT kt(h, PLES_MG_K_TRANSFER, 2.0 * F.ncell() * sizeof(TF) + (double)Cc.ncell() * sizeof(TC), true);
    KLAUNCH(h, prolong_blk, TPL(TC, TF), 256, 1, grid, block, 0, Cc.G, F.G, (const TC*)Cc.p[Cc.cur],
            (TF*)F.p[F.cur], gx);
    CK(hipGetLastError());
    return PLES_MG_OK;
}

int lv_smooth(H* h, int n, const double* r0, int nu, bool zero) {
    Level& L = h->lv[n];
    if (n == 0 && h->rhs32) return smooth<float, float>(h, L, h->rhs32, nu, zero, true);
    if (n == 0) {
        return L.f32 ? smooth<float, double>(h, L, r0, nu, zero, true)
                     : smooth<double, double>(h, L, r0, nu, zero, true);
    }
    return L.f32 ? smooth<float, float>(h, L, (const float*)L.b, nu, zero, false)
                 : smooth<double, double>(h, L, (const double*)L.b, nu, zero, false);
}

int lv_resid(H* h, int n, const double* r0) {
    Level& L = h->lv[n];
    if (n == 0 && h->rhs32) return vresid<float, float>(h, L, h->rhs32, true);
    if (n == 0) return L.f32 ? vresid<float, double>(h, L, r0, true) : vresid<double, double>(h, L, r0, true);
    return L.f32 ? vresid<float, float>(h, L, (const float*)L.b, false)
                 : vresid<double, double>(h, L, (const double*)L.b, false);
}

int lv_restrict(H* h, int n);

// residual of level n restricted to level n+1's rhs: fused where it applies
int lv_resid_restrict(H* h, int n, const double* r0) {
    Level &F = h->lv[n], &Cc = h->lv[n + 1];
    int st = 1;
    if (n == 0 && h->rhs32)
        st = Cc.f32 ? resid_restrict<float, float, float>(h, F, Cc, h->rhs32, true)
                    : resid_restrict<float, float, double>(h, F, Cc, h->rhs32, true);
    else if (n == 0 && F.f32)
        st = Cc.f32 ? resid_restrict<float, double, float>(h, F, Cc, r0, true)
                    : resid_restrict<float, double, double>(h, F, Cc, r0, true);
    else if (n == 0)
        st = resid_restrict<double, double, double>(h, F, Cc, r0, true);
    else if (F.f32)
        st = Cc.f32 ? resid_restrict<float, float, float>(h, F, Cc, (const float*)F.b, false)
                    : resid_restrict<float, float, double>(h, F, Cc, (const float*)F.b, false);
    else
        st = resid_restrict<double, double, double>(h, F, Cc, (const double*)F.b, false);
    if (st != 1) return st;
    CKS(lv_resid(h, n, r0));
    return lv_restrict(h, n);
}

int lv_restrict(H* h, int n) {
    Level &F = h->lv[n], &Cc = h->lv[n + 1];
    if (F.f32 && Cc.f32) return restrict_to<float, float>(h, F, Cc);
    if (F.f32 && !Cc.f32) return restrict_to<float, double>(h, F, Cc);
    if (!F.f32 && Cc.f32) return restrict_to<double, float>(h, F, Cc);
    return restrict_to<double, double>(h, F, Cc);
}

int lv_prolong(H* h, int n) {  // n+1 -> n
    Level &F = h->lv[n], &Cc = h->lv[n + 1];
    if (F.f32 && Cc.f32) return prolong_to<float, float>(h, Cc, F);
    if (F.f32 && !Cc.f32) return prolong_to<double, float>(h, Cc, F);
    if (!F.f32 && Cc.f32) return prolong_to<float, double>(h, Cc, F);
    return prolong_to<double, double>(h, Cc, F);
}

// coarsest level in one launch of one workgroup when its columns fit a block
// and it has no xi/zeta walls (whose ghosts need the separate halo launches)
// coarse levels up to this many cells get the exact dense solve: the
// V-cycle stops at the first such level (a level of 2048 cells, e.g.
// 16x8x16, costs one 53 MB matrix-vector product instead of its sweeps,
// transfers and the levels below it).  Levels with xi or zeta walls keep 512:
// the set-up's ghost fixed point of each unit vector converges slowly there.
constexpr long kCoarseDenseMax = 2048, kCoarseDenseMaxWalls = 512;
static_assert(kCoarseDenseMax <= kDenseCols, "the dense solve stages at most kDenseCols rhs values");
long dense_max(int ip, int kp) { return (ip == 0 && kp == 0) ? kCoarseDenseMax : kCoarseDenseMaxWalls; }
// interval of the Chebyshev sweep schedule on the spectrum of M_line^-1 A
// (measured on the terrain workloads: one V-cycle fewer than omega = 0.8 at
// 256x128x256 in both precisions and at 512x256x512; the cycle counts are
// sensitive to the lower end, [0.3, 2.0] costs five more)
constexpr double kChebLo = 0.48, kChebHi = 2.4;

// zeta-to-xi coupling ratio (sum g33 / sum g11 on the fine level) from which
// the first coarse level halves zeta only, or eta and zeta (opts.semicoarsen)
constexpr double kSemiRatio = 2.0;

bool coarse_one_block(const Level& L) {
    return !L.f32 && L.G.ip == 0 && L.G.kp == 0 && (long)L.G.nx * L.G.nz <= 1024;
}

size_t coarse_lds_bytes(const Level& L) { return (size_t)L.G.n * (12 * sizeof(double) + 2 * sizeof(float)); }

int coarse_solve(H* h, Level& L, const double* rhs, int nu) {
    if (L.kd) {  // exact: one matrix-vector product gives the ghosted solution
        KT kt(h, PLES_MG_K_SMOOTH, (double)L.G.n * L.kdp * (L.kd32 ? 4.0 : 8.0));
        const dim3 grid((unsigned)((L.G.n + kDenseWaves - 1) / kDenseWaves)), blk(64 * kDenseWaves);
        if (L.kd32)
            KLAUNCH(h, coarse_dense, TPL(float), 64 * kDenseWaves, 1, grid, blk, 0, L.G, (const float*)L.kd,
                    (const int*)L.kdi, (int)L.ncell(), L.kdp, rhs, (double*)L.p[1 - L.cur]);
        else
            KLAUNCH(h, coarse_dense, TPL(double), 64 * kDenseWaves, 1, grid, blk, 0, L.G, (const double*)L.kd,
                    (const int*)L.kdi, (int)L.ncell(), L.kdp, rhs, (double*)L.p[1 - L.cur]);
        CK(hipGetLastError());
        L.cur ^= 1;
        return PLES_MG_OK;
    }
    const int ncol = L.G.nx * L.G.nz;
    KT kt(h, PLES_MG_K_SMOOTH, (double)L.ncell() * (9 * 8 + 8 + 8 + 8 + 8));
    const size_t lds = coarse_lds_bytes(L);
    if (lds <= (size_t)h->lds_max) {
        // result lands in p[1 - cur]
        KLAUNCH(h, coarse_lds, TPL(double), 256, 1, dim3(1), dim3(256), lds, L.G, met<double>(L), line_f(L),
                (double*)L.p[1 - L.cur], rhs, (double)h->o.omega, 1.0 / L.pot, nu);
        CK(hipGetLastError());
        L.cur ^= 1;
        return PLES_MG_OK;
    }
    const int threads = std::min(1024, ((ncol + 63) / 64) * 64);
    KLAUNCH(h, coarse, TPL(double, double), 1024, 1, dim3(1), dim3(threads), 0, L.G, met<double>(L), line_f(L),
            (double*)L.p[L.cur], (double*)L.p[1 - L.cur], rhs, (double)h->o.omega, 1.0 / L.pot, nu);
    CK(hipGetLastError());
    if (nu & 1) L.cur ^= 1;
    return PLES_MG_OK;
}

// One V-cycle for A z = r0 on the fine iterate of level 0: from a zero
// initial guess (fine_zero) or continuing from the current one; the coarse
// levels always solve for corrections from zero.
int vcycle(H* h, const double* r0, bool fine_zero = true) {
    const int L = h->Lv;
    for (int n = 0; n < L; ++n) {
        const int nu = (n == L - 1) ? h->o.nu_coarse : h->o.nu_pre;
        Level& lvn = h->lv[n];
        if (n == L - 1 && (lvn.kd || coarse_one_block(lvn))) {
            CKS(coarse_solve(h, lvn, n == 0 ? r0 : (const double*)lvn.b, std::max(1, nu)));
            continue;
        }
        CKS(lv_smooth(h, n, r0, std::max(1, nu), n > 0 || fine_zero));
        if (n < L - 1) CKS(lv_resid_restrict(h, n, r0));
    }
    for (int n = L - 2; n >= 0; --n) {
        CKS(lv_prolong(h, n));
        if (h->o.nu_post > 0) CKS(lv_smooth(h, n, r0, h->o.nu_post, false));
    }
    return PLES_MG_OK;
}

// ---------------------------------------------------------------- outer loop
int read_S(H* h) {
    CK(hipMemcpyAsync(h->Sh, h->S, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, h->s));
    CK(hipStreamSynchronize(h->s));
    return PLES_MG_OK;
}

int set_S(H* h, int idx, double v) {
    h->Sh[idx] = v;
    CK(hipMemcpyAsync(h->S + idx, h->Sh + idx, sizeof(double), hipMemcpyHostToDevice, h->s));
    return PLES_MG_OK;
}

// r = b - A x (exact ghosts of x with the boundary fluxes), sums into S
int outer_residual(H* h, double* x, double* r) {
    CKS(halo64(h, x, true));
    const Level& L0 = h->lv[0];
    const ColLaunch cl = seg_launch(L0.G, L0.P);
    {
        KT kt(h, PLES_MG_K_OUTER_APPLY, (double)L0.ncell() * (8 * nmet(L0) + 24));
        auto go = [&](auto kern) {
            hipLaunchKernelGGL(kern, cl.grid, cl.block, 0, h->s, L0.G, met<double>(L0), (const double*)x,
                               (const double*)h->bvec, r, 1.0, (const double*)nullptr, (const double*)nullptr,
                               h->part, cl.nblocks, L0.P);
        };
        if (L0.x0) go(k_resid<1, double, double, double, double, double, true>);
        else go(k_resid<1, double, double, double, double, double, false>);
    }
    hipLaunchKernelGGL(k_fin_resid, dim3(1), dim3(256), 0, h->s, (const double*)h->part, cl.nblocks, h->S);
    CK(hipGetLastError());
    return PLES_MG_OK;
}

// one refinement step: z = V(r) (FP32 or FP64 V-cycle), w = A z, GCR(2)
// orthogonalisation against the previous (z, w), x += alpha z, r -= alpha w,
// r.r and max|r| into S.  Scalars stay on the device; a converged solve
// (S_DONE) leaves x and r unchanged.  q selects the (Z, W) ping-pong pair.
int step_body(H* h, int q) {
    Level& L0 = h->lv[0];
    const ColLaunch cl = seg_launch(L0.G, L0.P);
    const double N0 = (double)L0.ncell();
    // z = M r: inner_cycles V-cycles on A z = r (the first from zero)
    for (int c = 0; c < h->inner; ++c) CKS(vcycle(h, h->rvec, c == 0));
    double* Zq = h->Z[q];
    double* Wq = h->W[q];
    double* Zp = h->Z[1 - q];
    double* Wp = h->W[1 - q];
    {
        const long n = L0.G.n;
        const int nb = (int)std::min<long>((n + 255) / 256, 8192);
        if (L0.f32)
            KLAUNCH(h, convert, TPL(float, double), 1024, 1, dim3(nb), dim3(256), 0, n, (const float*)L0.p[L0.cur],
                    Zq);
        else
            KLAUNCH(h, convert, TPL(double, double), 1024, 1, dim3(nb), dim3(256), 0, n,
                    (const double*)L0.p[L0.cur], Zq);
    }
    CKS(halo64(h, Zq, false, true));
    const bool kry = h->o.krylov != 0;
    {
        KT kt(h, PLES_MG_K_OUTER_APPLY, N0 * (8 * nmet(L0) + 24 + (kry ? 8 : 0)));
        auto go = [&](auto x0) {
            constexpr bool X0 = decltype(x0)::value;
            KLAUNCH(h, resid, TPL(2, double, double, double, double, double, X0), 512, 1, cl.grid, cl.block, 0,
                    L0.G, met<double>(L0), (const double*)Zq, (const double*)nullptr, Wq, 1.0,
                    (const double*)h->rvec, (const double*)(kry ? Wp : nullptr), h->part, cl.nblocks, L0.P);
        };
        if (L0.x0) go(std::true_type{});
        else go(std::false_type{});
    }
    KLAUNCH0(h, fin_apply, 256, 1, dim3(1), dim3(256), 0, (const double*)h->part, cl.nblocks, h->S, h->o.krylov);
    {
        KT kt(h, PLES_MG_K_OUTER_UPDATE, N0 * (48 + (kry ? 32 + 16 : 0)));
        KLAUNCH(h, gcr_update, TPL(double), 512, 1, cl.grid, cl.block, 0, L0.G, h->x, h->rvec, Zq, Wq,
                (const double*)Zp, (const double*)Wp, (const double*)h->S, h->part, cl.nblocks, L0.P);
    }
    KLAUNCH0(h, fin_update, 256, 1, dim3(1), dim3(256), 0, (const double*)h->part, cl.nblocks, h->S);
    CK(hipGetLastError());
    return PLES_MG_OK;
}

// host-side state a step depends on (ping-pong selectors of every level)
std::vector<int> step_state(const H* h) {
    std::vector<int> v;
    for (const Level& L : h->lv) v.push_back(L.cur);
    return v;
}
void set_step_state(H* h, const std::vector<int>& v) {
    for (size_t n = 0; n < h->lv.size(); ++n) h->lv[n].cur = v[n];
}

void drop_graphs(H* h) {
    for (int q = 0; q < 2; ++q) {
        if (h->gexec[q]) (void)hipGraphExecDestroy(h->gexec[q]);
        h->gexec[q] = nullptr;
        h->gin[q].clear();
        h->gout[q].clear();
    }
}

// capture step_body(q) into a graph on the handle's stream (not executed)
int capture_step(H* h, int q) {
    h->gin[q] = step_state(h);
    CK(hipStreamBeginCapture(h->s, hipStreamCaptureModeThreadLocal));
    const int st = step_body(h, q);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(h->s, &g);
    if (st != PLES_MG_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    CK(e);
    const hipError_t ei = hipGraphInstantiate(&h->gexec[q], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    CK(ei);
    h->gout[q] = step_state(h);
    return PLES_MG_OK;
}

// whether the next refinement step is launched eagerly with its kernels timed
// (kernel timing on, and the step one of every timing_every); the other steps
// replay their graphs, captured with timing suspended
bool timed_step(H* h) {
    if (h->timing == 0) return false;
    const bool t = h->timing_every <= 1 || h->timing_step % h->timing_every == 0;
    ++h->timing_step;
    return t;
}

// enqueue one refinement step: replay of the captured graph of pair q when
// graphs are on and the step is not a timed one, else eager launches
int enqueue_step(H* h, int q) {
    if (!h->o.graphs || timed_step(h)) return step_body(h, q);
    if (h->gexec[q] && step_state(h) != h->gin[q]) drop_graphs(h);  // state moved: recapture
    if (!h->gexec[q]) {
        const std::vector<int> before = step_state(h);
        const unsigned tm = h->timing;
        h->timing = 0;
        const int st = capture_step(h, q);
        h->timing = tm;
        CKS(st);
        set_step_state(h, before);
    }
    CK(hipGraphLaunch(h->gexec[q], h->s));
    set_step_state(h, h->gout[q]);
    return PLES_MG_OK;
}

// stop test parameters of the scalar set Sd on the device; clears S_DONE, the
// step count and the GCR history.  mode 1: max|r| < thr, 0: r.r < thr.
int arm_stop(H* h, double* Sd, double thr, int mode) {
    double* st = h->Sh + (Sd == h->S ? 6 : 7) * S_COUNT;  // pinned staging per scalar set
    st[0] = 0.0;                    // S_DONE
    st[1] = thr;                    // S_THR
    st[2] = mode ? 1.0 : 0.0;       // S_MODE
    st[3] = 0.0;                    // S_CYC
    st[4] = 0.0;                    // S_HAVEP
    CK(hipMemcpyAsync(Sd + S_DONE, st, 4 * sizeof(double), hipMemcpyHostToDevice, h->s));
    CK(hipMemcpyAsync(Sd + S_HAVEP, st + 4, sizeof(double), hipMemcpyHostToDevice, h->s));
    return PLES_MG_OK;
}

// One step of the FP32 inner FGCR(2) of the mixed solver on A32 e = r32:
// z = V(rr32) (FP32 V-cycle), its exact homogeneous ghosts, w = A32 z with
// the dots, the GCR update of e and rr32.  Scalars in S2, FP64 sums.
int step32_body(H* h, int q) {
    Level& L0 = h->lv[0];
    const ColLaunch cl = seg_launch(L0.G, L0.P);
    float* Zq = h->Z32[q];
    float* Wq = h->W32[q];
    float* Zp = h->Z32[1 - q];
    float* Wp = h->W32[1 - q];
    // One V-cycle from zero flips the fine ping-pong pair max(1, nu_pre) +
    // nu_post times: the buffer its last sweep writes is known, so Zq stands in
    // for it and receives z directly (no copy); Zp is not touched.
    const bool direct = h->inner == 1 && h->Lv >= 2;
    const int fin = L0.cur ^ ((std::max(1, h->o.nu_pre) + std::max(0, h->o.nu_post)) & 1);
    void* held = L0.p[fin];
    if (direct) L0.p[fin] = Zq;
    h->rhs32 = h->rr32;
    int st = PLES_MG_OK;
    for (int c = 0; c < h->inner && st == PLES_MG_OK; ++c) st = vcycle(h, nullptr, c == 0);
    h->rhs32 = nullptr;
    if (direct) {
        const bool landed = L0.cur == fin;
        L0.p[fin] = held;
        if (!landed) return PLES_MG_ERR_ARG;  // the sweep count did not match (internal)
    }
    CKS(st);
    if (!direct) {
        const long n = L0.G.n;
        const int nb = (int)std::min<long>((n + 255) / 256, 8192);
        KLAUNCH(h, convert, TPL(float, float), 1024, 1, dim3(nb), dim3(256), 0, n, (const float*)L0.p[L0.cur], Zq);
    }
    CKS(halo<float>(h, L0, Zq, 0.0, false, std::max(1, h->halo_passes32), true));
    const double N0 = (double)L0.ncell();
    const bool kry = h->o.krylov != 0;
    {
        KT kt(h, PLES_MG_K_OUTER_APPLY, N0 * (4 * nmet(L0) + 12 + (kry ? 4 : 0)));
        auto go = [&](auto x0) {
            constexpr bool X0 = decltype(x0)::value;
            KLAUNCH(h, resid, TPL(2, float, float, float, float, float, X0), 512, 1, cl.grid, cl.block, 0, L0.G,
                    met<float>(L0), (const float*)Zq, (const float*)nullptr, Wq, 1.0f, (const float*)h->rr32,
                    (const float*)(kry ? Wp : nullptr), h->part, cl.nblocks, L0.P);
        };
        if (L0.x0) go(std::true_type{});
        else go(std::false_type{});
    }
    KLAUNCH0(h, fin_apply, 256, 1, dim3(1), dim3(256), 0, (const double*)h->part, cl.nblocks, h->S2, h->o.krylov);
    {
        KT kt(h, PLES_MG_K_OUTER_UPDATE, N0 * (24 + (kry ? 16 + 8 : 0)));
        KLAUNCH(h, gcr_update, TPL(float), 512, 1, cl.grid, cl.block, 0, L0.G, h->e32, h->rr32, Zq, Wq,
                (const float*)Zp, (const float*)Wp, (const double*)h->S2, h->part, cl.nblocks, L0.P);
    }
    KLAUNCH0(h, fin_update, 256, 1, dim3(1), dim3(256), 0, (const double*)h->part, cl.nblocks, h->S2);
    CK(hipGetLastError());
    return PLES_MG_OK;
}

// Tail of a step of device-driven iterative refinement (after step32_body):
// where the inner solve is done (or out of steps) it adds x += |r| e, takes
// the FP64 residual, tests convergence and stalling and arms the next
// refinement, all on the device; in a grouped launch the tail's field
// kernels skip a problem whose step does not end a refinement (S_NOREF), and
// every kernel of it tests S_REF itself.  1-D grids (groupable).
int ir_tail(H* h) {
    Level& L0 = h->lv[0];
    const ColLaunch cl = seg_launch(L0.G, L0.P);
    const dim3 g1((unsigned)((L0.G.n + 255) / 256)), b1(256);
    const int sk = h->skip;
    KLAUNCH0(h, ir_check, 64, 1, dim3(1), dim3(64), 0, h->S, (const double*)h->S2);
    h->skip = S_NOREF;
    KLAUNCH0(h, add_back_if, 256, 1, g1, b1, 0, L0.G, h->x, (const double*)h->S, (const float*)h->e32);
    int st = halo64(h, h->x, true);
    if (st == PLES_MG_OK) {
        auto go = [&](auto x0) {
            constexpr bool X0 = decltype(x0)::value;
            KLAUNCH(h, resid, TPL(1, double, double, double, double, double, X0), 512, 1, cl.grid, cl.block, 0, L0.G,
                    met<double>(L0), (const double*)h->x, (const double*)h->bvec, h->rvec, 1.0,
                    (const double*)nullptr, (const double*)nullptr, h->part, cl.nblocks, L0.P);
        };
        if (L0.x0) go(std::true_type{});
        else go(std::false_type{});
        KLAUNCH0(h, fin_ir, 256, 1, dim3(1), dim3(256), 0, (const double*)h->part, cl.nblocks, h->S, h->S2);
        KLAUNCH0(h, scale32_if, 256, 1, g1, b1, 0, L0.G, (const double*)h->rvec, (const double*)h->S, h->rr32,
                 h->e32);
    }
    h->skip = sk;
    CKS(st);
    CK(hipGetLastError());
    return PLES_MG_OK;
}

// one step of device-driven iterative refinement (batched mixed solves)
int step_ir_body(H* h, int q) {
    CKS(step32_body(h, q));
    return ir_tail(h);
}

void drop_graphs32(H* h) {
    for (int q = 0; q < 2; ++q) {
        if (h->gexec32[q]) (void)hipGraphExecDestroy(h->gexec32[q]);
        h->gexec32[q] = nullptr;
        h->gin32[q].clear();
        h->gout32[q].clear();
    }
}

int enqueue_step32(H* h, int q) {
    if (!h->o.graphs || timed_step(h)) return step32_body(h, q);
    if (h->gexec32[q] && step_state(h) != h->gin32[q]) drop_graphs32(h);
    if (!h->gexec32[q]) {
        const std::vector<int> before = step_state(h);
        h->gin32[q] = before;
        const unsigned tm = h->timing;
        h->timing = 0;  // no timing events inside the capture
        CK(hipStreamBeginCapture(h->s, hipStreamCaptureModeThreadLocal));
        const int st = step32_body(h, q);
        h->timing = tm;
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(h->s, &g);
        if (st != PLES_MG_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        CK(e);
        const hipError_t ei = hipGraphInstantiate(&h->gexec32[q], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        CK(ei);
        h->gout32[q] = step_state(h);
        set_step_state(h, before);
    }
    CK(hipGraphLaunch(h->gexec32[q], h->s));
    set_step_state(h, h->gout32[q]);
    return PLES_MG_OK;
}

// Run refinement steps (kind 0: FP64 GCR step, 1: FP32 inner step, scalars
// in Sd) until the device stop test holds or maxsteps steps ran.  Steps are
// enqueued ahead of the residual read-back (a ring of pinned snapshots of Sd,
// one event each) while the observed contraction predicts that the step in
// flight will not converge yet, so the device never idles on the host
// between steps; near convergence the host waits for each step.  A step
// enqueued past convergence changes nothing (S_DONE).  meas(rr, rmax) is the
// host's measure of the stop test, eps its threshold.
struct StepRun {
    int applied = 0;
    bool done = false;
    bool nonfinite = false;
    double rr = 0.0, rmax = 0.0;
};

template <class Meas>
int run_steps(H* h, int kind, double* Sd, double start, double eps, int maxsteps, Meas meas, StepRun& out) {
    int q = h->qnext[kind], enq = 0, seen = 0;
    double last = start, rho = 0.5;
    const int depth = 2;
    out = StepRun{};
    while (true) {
        while (enq < maxsteps && enq - seen < depth) {
            if (enq > seen && last * std::pow(rho, (double)(enq - seen)) < 4.0 * eps) break;
            CKS(kind == 0 ? enqueue_step(h, q) : enqueue_step32(h, q));
            q = 1 - q;
            double* slot = h->Sh + 2 * S_COUNT + (enq % 4) * S_COUNT;
            CK(hipMemcpyAsync(slot, Sd, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, h->s));
            CK(hipEventRecord(h->sev[enq % 4], h->s));
            ++enq;
        }
        if (seen == enq) break;
        CK(hipEventSynchronize(h->sev[seen % 4]));
        const double* slot = h->Sh + 2 * S_COUNT + (seen % 4) * S_COUNT;
        ++seen;
        out.rr = slot[S_RR];
        out.rmax = slot[S_RMAX];
        if (!std::isfinite(out.rr)) {
            out.nonfinite = true;
            out.applied = seen;
            break;
        }
        if (slot[S_DONE] != 0.0) {
            out.done = true;
            out.applied = (int)slot[S_CYC];
            break;
        }
        const double m = meas(out.rr, out.rmax);
        if (last > 0.0 && m > 0.0) rho = std::min(0.95, std::max(1e-3, m / last));
        last = m;
        out.applied = seen;
    }
    h->qnext[kind] = q;
    // steps still in flight past convergence are no-ops; the callers' next
    // work (the true residual and its read-back) follows them on the stream
    return PLES_MG_OK;
}

// FP64 multigrid-preconditioned GCR(2) (both precisions' V-cycles).
int solve_gcr64(H* h, double eps, int maxcyc, int& cycles, double& rr, double& rmax, double bnorm) {
    const bool absmode = h->o.stop_mode == PLES_MG_STOP_ABS_MAX;
    auto measure = [&](double r2, double rm) { return absmode ? rm : std::sqrt(r2) / bnorm; };
    const double thr = absmode ? eps : (eps * bnorm) * (eps * bnorm);
    int status = PLES_MG_NOT_CONVERGED;
    bool conv = measure(rr, rmax) < eps;
    if (conv) return PLES_MG_OK;
    CKS(arm_stop(h, h->S, thr, absmode));
    while (!conv && cycles < maxcyc) {
        StepRun run;
        CKS(run_steps(h, 0, h->S, measure(rr, rmax), eps, maxcyc - cycles, measure, run));
        cycles += run.applied;
        rr = run.rr;
        rmax = run.rmax;
        if (run.nonfinite) return PLES_MG_ERR_NONFINITE;
        if (!run.done) break;
        // confirm with the true residual; restart the recurrence if it drifted
        CKS(outer_residual(h, h->x, h->rvec));
        CKS(read_S(h));
        rr = h->Sh[S_RR];
        rmax = h->Sh[S_RMAX];
        if (measure(rr, rmax) < eps) {
            conv = true;
            status = PLES_MG_OK;
        } else {
            CKS(arm_stop(h, h->S, thr, absmode));
        }
    }
    return status;
}

// Mixed precision: FP64 iterative refinement around an FP32 FGCR(2)
// preconditioned by FP32 V-cycles.  Each refinement takes the FP64 residual
// r = b - A x (FP64 metrics), solves A32 e = r/|r| in FP32 to a relative
// tolerance the outer test needs (at most 1e-5: FP32 precision), and adds
// x += |r| e in FP64.  The exact coarse solve keeps FP64 vectors (its dense
// map is FP32 in this solver).
int solve_ir32(H* h, double eps, int maxcyc, int& cycles, double& rr, double& rmax, double bnorm) {
    const bool absmode = h->o.stop_mode == PLES_MG_STOP_ABS_MAX;
    auto measure = [&](double r2, double rm) { return absmode ? rm : std::sqrt(r2) / bnorm; };
    const Level& L0 = h->lv[0];
    const ColLaunch ce = cell_launch(L0.G);
    int stall = 0;
    while (cycles < maxcyc) {
        const double m = measure(rr, rmax);
        if (m < eps) return PLES_MG_OK;
        const double rn = std::sqrt(rr);
        if (!(rn > 0.0)) return PLES_MG_OK;
        hipLaunchKernelGGL(k_scale_to32, ce.grid, ce.block, 0, h->s, L0.G, (const double*)h->rvec, 1.0 / rn, h->rr32);
        CK(hipMemsetAsync(h->e32, 0, (size_t)L0.G.n * sizeof(float), h->s));
        CK(hipGetLastError());
        // relative reduction the outer test still needs (L2 terms), FP32-bounded
        const double need = absmode ? eps / std::max(rmax, 1e-300) : eps * bnorm / rn;
        const double tol = std::min(0.1, std::max(1e-5, 0.5 * need));
        CKS(arm_stop(h, h->S2, tol * tol, 0));
        auto rel = [&](double r2, double) { return std::sqrt(r2); };
        StepRun run;
        CKS(run_steps(h, 1, h->S2, 1.0, tol, std::min(40, maxcyc - cycles), rel, run));
        cycles += run.applied;
        if (run.nonfinite) return PLES_MG_ERR_NONFINITE;
        hipLaunchKernelGGL(k_add_back, ce.grid, ce.block, 0, h->s, L0.G, h->x, rn, (const float*)h->e32);
        CK(hipGetLastError());
        CKS(outer_residual(h, h->x, h->rvec));
        CKS(read_S(h));
        const double m_old = m;
        rr = h->Sh[S_RR];
        rmax = h->Sh[S_RMAX];
        if (!std::isfinite(rr)) return PLES_MG_ERR_NONFINITE;
        if (measure(rr, rmax) > 0.5 * m_old) {
            if (++stall >= 3) break;  // FP32 cannot reduce it further
        } else {
            stall = 0;
        }
    }
    return measure(rr, rmax) < eps ? PLES_MG_OK : PLES_MG_NOT_CONVERGED;
}

int solve_core(H* h, double eps, int maxcyc, ples_mg_result* res) {
    CK(hipEventRecord(h->t0, h->s));
    CKS(outer_residual(h, h->x, h->rvec));
    CKS(read_S(h));
    const double bb = h->Sh[S_BB];
    const double bnorm = std::sqrt(bb) > 0 ? std::sqrt(bb) : 1.0;
    double rr = h->Sh[S_RR], rmax = h->Sh[S_RMAX];
    if (!std::isfinite(rr)) return PLES_MG_ERR_NONFINITE;
    int cycles = 0;
    int status = h->ir32 ? solve_ir32(h, eps, maxcyc, cycles, rr, rmax, bnorm)
                         : solve_gcr64(h, eps, maxcyc, cycles, rr, rmax, bnorm);
    if (status != PLES_MG_OK && status != PLES_MG_NOT_CONVERGED && status != PLES_MG_ERR_NONFINITE) return status;
    // leave the iterate with exact ghosts (the reference returns fi with ghosts)
    CKS(halo64(h, h->x, true));
    CK(hipEventRecord(h->t1, h->s));
    CK(hipStreamSynchronize(h->s));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, h->t0, h->t1));
    if (res) {
        res->status = status;
        res->cycles = cycles;
        res->rel_res = std::sqrt(rr) / bnorm;
        res->abs_res = rmax;
        res->rhs_mean_removed = h->last_removed;
        res->seconds = ms * 1e-3;
    }
    return status;
}

// ---------------------------------------------------------------- packing
template <class T>
int pack_box(H* h, const Geo& G, int i0, int i1, int j0, int j1, int k0, int k1, const T* host, T* dghost,
             T* staging) {
    const long ni = i1 - i0 + 1, nj = j1 - j0 + 1, nk = k1 - k0 + 1;
    CK(hipMemcpyAsync(staging, host, sizeof(T) * ni * nj * nk, hipMemcpyHostToDevice, h->s));
    dim3 block(64, 4);
    dim3 grid((unsigned)((ni + 63) / 64), (unsigned)((nj + 3) / 4), (unsigned)nk);
    hipLaunchKernelGGL((k_pack<T>), grid, block, 0, h->s, G, i0, i1, j0, j1, k0, (const T*)staging, dghost);
    CK(hipGetLastError());
    return PLES_MG_OK;
}

template <class T>
int unpack_box(H* h, const Geo& G, int i0, int i1, int j0, int j1, int k0, int k1, const T* dghost, T* host,
               T* staging) {
    const long ni = i1 - i0 + 1, nj = j1 - j0 + 1, nk = k1 - k0 + 1;
    dim3 block(64, 4);
    dim3 grid((unsigned)((ni + 63) / 64), (unsigned)((nj + 3) / 4), (unsigned)nk);
    hipLaunchKernelGGL((k_unpack<T>), grid, block, 0, h->s, G, i0, i1, j0, j1, k0, dghost, staging);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(host, staging, sizeof(T) * ni * nj * nk, hipMemcpyDeviceToHost, h->s));
    CK(hipStreamSynchronize(h->s));
    return PLES_MG_OK;
}

// metric face boxes of the reference shapes (dir of metric m = m/3)
void face_box(const Geo& G, int dir, int& i0, int& i1, int& j0, int& j1, int& k0, int& k1) {
    i0 = 1; i1 = G.nx; j0 = 1; j1 = G.ny; k0 = 1; k1 = G.nz;
    if (dir == 0) i0 = 0;
    if (dir == 1) j0 = 0;
    if (dir == 2) k0 = 0;
}

Nodes nodes_of(const H* h) {
    Nodes N;
    N.x = h->nodes[0]; N.y = h->nodes[1]; N.z = h->nodes[2];
    N.nx = h->jx; N.ny = h->jy; N.nz = h->jz;
    N.ip = h->ip; N.jp = h->jp; N.kp = h->kp;
    return N;
}

// FP64 metrics of level n into g[9] (ghosted, zero elsewhere); optional face vectors
int compute_metrics(H* h, int n, double* const g[9], double* const v[9]) {
    const Geo& G = h->lv[n].G;
    const int st = 1 << n, sx = h->lv[n].sx, sy = h->lv[n].sy;
    const Nodes N = nodes_of(h);
    for (int dir = 0; dir < 3; ++dir) {
        const int na = (dir == 0) ? G.nx + 1 : G.nx;
        const int nbv = (dir == 1) ? G.ny + 1 : G.ny;
        const int nc = (dir == 2) ? G.nz + 1 : G.nz;
        dim3 block(std::min(64, pow2ceil(na)), 4);
        dim3 grid((na + block.x - 1) / block.x, (nbv + 3) / 4, nc);
        hipLaunchKernelGGL((k_metrics<double>), grid, block, 0, h->s, N, G, sx, sy, st, n > 0 ? 1 : 0, dir, g[3 * dir],
                           g[3 * dir + 1], g[3 * dir + 2], v ? v[3 * dir] : nullptr, v ? v[3 * dir + 1] : nullptr,
                           v ? v[3 * dir + 2] : nullptr);
    }
    CK(hipGetLastError());
    return PLES_MG_OK;
}

int level_factors(H* h, Level& L) {
    if (L.P > 1 && (L.lv_v == nullptr || L.lv_w == nullptr)) return PLES_MG_ERR_ARG;
    if (L.G.ny % L.P != 0 || L.P > 16) return PLES_MG_ERR_ARG;
    const ColLaunch cl = seg_launch(L.G, L.P);
    if (L.f32)
        hipLaunchKernelGGL((k_line_factors<float>), cl.grid, cl.block, 0, h->s, L.G, met<float>(L), 1.0 / L.pot,
                           L.P, L.invm, L.cpr, L.lv_v, L.lv_w);
    else
        hipLaunchKernelGGL((k_line_factors<double>), cl.grid, cl.block, 0, h->s, L.G, met<double>(L),
                           1.0 / L.pot, L.P, L.invm, L.cpr, L.lv_v, L.lv_w);
    CK(hipGetLastError());
    if (L.g16[0] && L.P > 1) {  // the 16-bit-metric smoother reads 16-bit spikes
        const long n = L.G.n;
        if (!L.v16) CKS(dalloc(h, (void**)&L.v16, (size_t)n * 2));
        if (!L.w16) CKS(dalloc(h, (void**)&L.w16, (size_t)n * 2));
        const int nb = (int)std::min<long>((n + 255) / 256, 8192);
        hipLaunchKernelGGL(k_to_half, dim3(nb), dim3(256), 0, h->s, n, (const float*)L.lv_v, L.v16, 1.0f);
        hipLaunchKernelGGL(k_to_half, dim3(nb), dim3(256), 0, h->s, n, (const float*)L.lv_w, L.w16, 1.0f);
        if (L.pk_a) {
            if (!L.vw16) CKS(dalloc(h, (void**)&L.vw16, (size_t)n * sizeof(H2)));
            hipLaunchKernelGGL(k_pack_vw, dim3(nb), dim3(256), 0, h->s, n, (const float*)L.lv_v,
                               (const float*)L.lv_w, L.vw16);
        }
        CK(hipGetLastError());
    }
    return PLES_MG_OK;
}

// ---------------------------------------------------------------- autotuner
// Launch shapes of the level kernels, chosen by timing the candidates on the
// level's own arrays: the smoother's eta segments per column P (1..16,
// dividing ny, >= 4 rows each; the lanes per block follow from P) and the
// V-cycle residual's segments x lanes.  Results are kept per level shape and
// precision for the life of the process, so every handle of a shape runs the
// same kernels (the smoother's rounding depends on P; the residual's does not).
struct TuneVal {
    int P, rP, rbx, use16;
};
// key: nx ny nz ip kp, FP32 level, fine level, 16-bit smoother metrics present
using TuneKey = std::array<int, 8>;
std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tune;

bool valid_segments(int P) {
    return P == 1 || P == 2 || P == 3 || P == 4 || P == 6 || P == 8 || P == 12 || P == 16;
}

// a segment count the kernels and the autotuner accept for ny rows: a valid
// count dividing ny into segments of at least four rows
bool seg_ok(int ny, int P) { return valid_segments(P) && ny % P == 0 && ny / P >= 4; }

std::vector<int> seg_candidates(int ny) {
    std::vector<int> v;
    for (int P : {1, 2, 3, 4, 6, 8, 12, 16})
        if (seg_ok(ny, P)) v.push_back(P);
    return v;
}

template <class F>
int time_launches(H* h, int reps, double& ms, F&& f) {
    hipEvent_t a, b;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    int st = f();  // warm-up
    if (st == PLES_MG_OK) {
        (void)hipEventRecord(a, h->s);
        for (int r = 0; r < reps && st == PLES_MG_OK; ++r) st = f();
        (void)hipEventRecord(b, h->s);
        (void)hipEventSynchronize(b);
        float e = 0.f;
        (void)hipEventElapsedTime(&e, a, b);
        ms = e / reps;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return st;
}

int lv_smooth(H* h, int n, const double* r0, int nu, bool zero);
int lv_resid(H* h, int n, const double* r0);
int lv_resid_restrict(H* h, int n, const double* r0);
bool coarse_one_block(const Level& L);

// time_missing: time the candidates of a shape the table does not hold (else
// the level keeps the fixed rule's shape)
int tune_level(H* h, int n, bool time_missing = true) {
    Level& L = h->lv[n];
    if (n >= h->Lv) return PLES_MG_OK;  // below the V-cycle's coarsest level: not visited
    if (n == h->Lv - 1 && (L.kd || coarse_one_block(L))) return PLES_MG_OK;  // one-launch coarse solver
    const bool half = L.g16[0] != nullptr;
    const TuneKey key{L.G.nx, L.G.ny, L.G.nz, L.G.ip, L.G.kp, L.f32 ? 1 : 0, n == 0 ? 1 : 0, half ? 1 : 0};
    std::lock_guard<std::mutex> lock(g_tune_mu);
    const size_t cells = (size_t)L.G.n;
    auto ensure_vw = [&]() -> int {
        if (!L.lv_v) CKS(dalloc(h, (void**)&L.lv_v, cells * 4));
        if (!L.lv_w) CKS(dalloc(h, (void**)&L.lv_w, cells * 4));
        return PLES_MG_OK;
    };
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
        if (it->second.P > 1) CKS(ensure_vw());
        L.P = it->second.P;
        L.rP = it->second.rP;
        L.rbx = it->second.rbx;
        L.use16 = half && it->second.use16 != 0;
        return level_factors(h, L);
    }
    if (!time_missing) return PLES_MG_OK;
    const double* r0 = h->rvec;  // level-0 rhs (any content: timing only)
    // the fixed rule's shape is the incumbent; another shape must be at least
    // 3% faster to replace it, which keeps choices stable between runs
    const int P0 = L.P;
    TuneVal best{P0, 0, 0, L.use16 ? 1 : 0};
    double tbest = 1e30;
    std::vector<int> cands = seg_candidates(L.G.ny);
    std::stable_partition(cands.begin(), cands.end(), [&](int P) { return P == P0; });
    for (int P : cands) {
        if (P > 1) CKS(ensure_vw());
        L.P = P;
        CKS(level_factors(h, L));
        double ms = 0;
        CKS(time_launches(h, 3, ms, [&] { return lv_smooth(h, n, r0, 2, false); }));
        if (ms < 0.97 * tbest || tbest == 1e30) {
            tbest = ms;
            best.P = P;
        }
    }
    L.P = best.P;
    CKS(level_factors(h, L));
    if (half) {  // smoother metrics: the fixed rule's choice is the incumbent
        tbest = 1e30;
        const bool u0 = L.use16;
        for (bool u : {u0, !u0}) {
            L.use16 = u;
            double ms = 0;
            CKS(time_launches(h, 3, ms, [&] { return lv_smooth(h, n, r0, 2, false); }));
            if (ms < 0.97 * tbest || tbest == 1e30) {
                tbest = ms;
                best.use16 = u ? 1 : 0;
            }
        }
        L.use16 = best.use16 != 0;
    }
    tbest = 1e30;
    const ColLaunch c0 = seg_launch(L.G, L.P);
    std::vector<std::pair<int, int>> rc{{L.P, (int)c0.block.x}};
    for (int P : seg_candidates(L.G.ny))
        for (int bx = 32; bx <= 256; bx *= 2)
            if (bx * P <= 512 && bx <= 2 * L.G.nx && !(P == L.P && bx == (int)c0.block.x)) rc.push_back({P, bx});
    for (auto [P, bx] : rc) {
        L.rP = P;
        L.rbx = bx;
        double ms = 0;
        // the V-cycle runs the residual fused with the restriction below the coarsest level
        CKS(time_launches(h, 3, ms, [&] { return n < h->Lv - 1 ? lv_resid_restrict(h, n, r0) : lv_resid(h, n, r0); }));
        if (ms < 0.97 * tbest || tbest == 1e30) {
            tbest = ms;
            best.rP = P;
            best.rbx = bx;
        }
    }
    L.rP = best.rP;
    L.rbx = best.rbx;
    g_tune[key] = best;
    return PLES_MG_OK;
}

int check_grid(const ples_mg_grid* g) {
    if (!g || !g->x || !g->y || !g->z) return PLES_MG_ERR_ARG;
    if (g->jx < 1 || g->jy < 1 || g->jz < 1 || g->nlevel < 1 || g->nlevel > 16) return PLES_MG_ERR_BAD_DIMS;
    if (g->jp != 1) return PLES_MG_ERR_UNSUPPORTED;  // the reference supports eta walls only
    if ((g->ip != 0 && g->ip != 1) || (g->kp != 0 && g->kp != 1)) return PLES_MG_ERR_ARG;
    const int st = 1 << (g->nlevel - 1);
    if (g->jx % st || g->jy % st || g->jz % st) return PLES_MG_ERR_BAD_DIMS;
    const int cx = g->jx / st, cy = g->jy / st, cz = g->jz / st;
    if (cy < 2 || (g->ip == 1 && cx < 2) || (g->kp == 1 && cz < 2)) return PLES_MG_ERR_BAD_DIMS;
    return PLES_MG_OK;
}

// ---------------------------------------------------------------- exact coarse solve
// The coarsest level's operator as a dense matrix, built on the host with the
// device's own stencil and ghost rules (compiled for both): column m is the
// 19-point operator applied to the unit vector e_m whose wall ghosts sit at
// the fixed point of the homogeneous face relations and edge rules (the
// coarse-level boundary condition).  The operator is singular (constants);
// the flux form sums to zero over the box, so 1 is also its left null vector
// and the system bordered with the zero-mean constraint, [[A, 1], [1^T, 0]],
// is regular.  Its inverse's N x N block K maps a rhs b to the zero-mean
// solution of A p = b - mean(b).  Stored: E K, E the map from the interior to
// the ghosted vector (ghosts by the same fixed point), so one matrix-vector
// product gives the whole ghosted solution.
// Process-wide cache of dense coarsest-level maps: the map depends only on
// the level's shape, node stride and metrics, so a handle re-created for the
// same grid (a served case admitted again after its eviction, a batch of
// equal problems) takes the stored map instead of rebuilding it on the host
// (tens to hundreds of milliseconds).  Keyed by the shape and stride; a hit
// also compares the metric bytes, so the map used is always the one these
// metrics give.  Least recently used entries leave first; the limit holds the
// maps of a few dozen served cases with 2048-cell coarse levels (53 MB each).
// Pinned host buffer (falls back to pageable memory if pinning fails): the
// cache's stored maps (a hit is one DMA copy to the device, the buffer shared
// with the handles copying from it while the cache may drop it) and the
// server's node copies (uploaded at each cold start).
struct PinnedBuf {
    double* p = nullptr;
    size_t n = 0;
    bool pinned = false;
    explicit PinnedBuf(size_t n_) : n(n_) {
        if (hipHostMalloc((void**)&p, n * sizeof(double), hipHostMallocDefault) == hipSuccess) {
            pinned = true;
        } else {
            (void)hipGetLastError();
            p = new (std::nothrow) double[n];  // nullptr when out of memory: callers check
        }
    }
    ~PinnedBuf() {
        if (pinned) (void)hipHostFree(p);
        else delete[] p;
    }
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
};
struct DenseEntry {
    int nx, ny, nz, ip, jp, kp;
    double pot;
    std::vector<double> met;  // the nine metric arrays, concatenated
    std::shared_ptr<const PinnedBuf> map;
};
std::mutex g_dense_mu;
// least recently used first; never destroyed (its pinned buffers are not
// released during the runtime's own teardown at exit)
std::deque<DenseEntry>& g_dense = *new std::deque<DenseEntry>;
size_t g_dense_bytes = 0;
constexpr size_t kDenseCacheMax = size_t(4) << 30;

std::shared_ptr<const PinnedBuf> dense_lookup(const Geo& G, double pot, const std::vector<double>& met) {
    std::lock_guard<std::mutex> lk(g_dense_mu);
    for (auto it = g_dense.begin(); it != g_dense.end(); ++it) {
        const DenseEntry& e = *it;
        if (e.nx == G.nx && e.ny == G.ny && e.nz == G.nz && e.ip == G.ip && e.jp == G.jp && e.kp == G.kp &&
            e.pot == pot && e.met.size() == met.size() &&
            std::memcmp(e.met.data(), met.data(), met.size() * sizeof(double)) == 0) {
            std::shared_ptr<const PinnedBuf> out = e.map;
            DenseEntry hit = std::move(*it);  // most recently used: to the back
            g_dense.erase(it);
            g_dense.push_back(std::move(hit));
            return out;
        }
    }
    return nullptr;
}

void dense_store(const Geo& G, double pot, std::vector<double> met, std::shared_ptr<const PinnedBuf> map) {
    const size_t b = (met.size() + map->n) * sizeof(double);
    if (b > kDenseCacheMax) return;
    std::lock_guard<std::mutex> lk(g_dense_mu);
    while (!g_dense.empty() && g_dense_bytes + b > kDenseCacheMax) {
        g_dense_bytes -= (g_dense.front().met.size() + g_dense.front().map->n) * sizeof(double);
        g_dense.pop_front();
    }
    g_dense.push_back(DenseEntry{G.nx, G.ny, G.nz, G.ip, G.jp, G.kp, pot, std::move(met), std::move(map)});
    g_dense_bytes += b;
}

// the solve's copy of a level's FP64 dense map (device, rows of N entries):
// rows padded to a multiple of 256 entries, FP32 for the mixed solver
int install_map(H* h, Level& L, const double* kf) {
    const long ng = L.G.n;
    const int N = (int)L.ncell();
    L.kdp = (N + 255) / 256 * 256;
    L.kd32 = h->o.precision == PLES_MG_MIXED;
    CKS(dalloc(h, &L.kd, (size_t)ng * L.kdp * (L.kd32 ? 4 : 8)));
    CKS(dalloc(h, (void**)&L.kdi, (size_t)N * sizeof(int)));
    hipLaunchKernelGGL(k_interior_index, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, h->s, L.G, L.kdi);
    const dim3 grid((unsigned)((ng * L.kdp + 255) / 256));
    if (L.kd32)
        hipLaunchKernelGGL(k_pad_map<float>, grid, dim3(256), 0, h->s, kf, ng, N, L.kdp, (float*)L.kd);
    else
        hipLaunchKernelGGL(k_pad_map<double>, grid, dim3(256), 0, h->s, kf, ng, N, L.kdp, (double*)L.kd);
    CK(hipGetLastError());
    return PLES_MG_OK;
}

int build_coarse_dense(H* h, Level& L) {
    const Geo& G = L.G;
    const long N = L.ncell(), ng = G.n;
    std::vector<std::vector<double>> gm(9, std::vector<double>((size_t)ng));
    for (int q = 0; q < 9; ++q)
        CK(hipMemcpyAsync(gm[q].data(), L.g64[q], (size_t)ng * 8, hipMemcpyDeviceToHost, h->s));
    CK(hipStreamSynchronize(h->s));
    std::vector<double> metcat;
    metcat.reserve((size_t)(9 * ng));
    for (int q = 0; q < 9; ++q) metcat.insert(metcat.end(), gm[q].begin(), gm[q].end());
    if (std::shared_ptr<const PinnedBuf> cached = dense_lookup(G, L.pot, metcat)) {
        double* kf = nullptr;
        CKS(dalloc(h, (void**)&kf, (size_t)(ng * N) * 8));
        CK(hipMemcpyAsync(kf, cached->p, (size_t)(ng * N) * 8, hipMemcpyHostToDevice, h->s));
        CKS(install_map(h, L, kf));
        CK(hipStreamSynchronize(h->s));
        dfree(h, kf);
        return PLES_MG_OK;
    }
    Met<double> M;
    for (int q = 0; q < 9; ++q) M.g[q] = gm[q].data();
    BCs bc0;
    for (int q = 0; q < 6; ++q) bc0.cs[q] = nullptr;
    bc0.an_over_dt = 0.0;
    const long NX = G.nx + 2, NY = G.ny + 2, NZ = G.nz + 2;
    const long ne = 2 * NY * NZ + 2 * NX * NZ + 2 * NX * NY;
    auto cell = [&](long c) { return G.idx((int)(c % G.nx) + 1, (int)((c / G.nx) % G.ny) + 1, (int)(c / ((long)G.nx * G.ny)) + 1); };
    std::vector<double> A((size_t)(N * N), 0.0);   // column-major
    std::vector<double> E((size_t)(ng * N), 0.0);  // column-major: ghosted image of e_m
    const double inv_pot = 1.0 / L.pot;
    // columns m are independent: each worker takes a contiguous range with
    // its own work vectors (column m writes only column m of A and of E)
    auto columns = [&](long m0, long m1) {
        std::vector<double> v((size_t)ng), prev((size_t)ng);
        std::vector<long> seen((size_t)N, -1), rows((size_t)N);
        for (long m = m0; m < m1; ++m) {
            std::fill(v.begin(), v.end(), 0.0);
            v[(size_t)cell(m)] = 1.0;
            for (int pass = 0; pass < 400; ++pass) {
                prev = v;
                for (int set = 0; set < 3; ++set) {
                    const long nset = face_set_size(G, set);
                    for (int color = 0; color < 2; ++color)
                        for (long t = 0; t < nset; ++t) halo_face_point(G, M, v.data(), t, color, bc0, set, 0);
                }
                for (long t = 0; t < ne; ++t) halo_edge_point(G, v.data(), t);
                double d = 0.0;
                for (long t = 0; t < ng; ++t) d = std::max(d, std::fabs(v[(size_t)t] - prev[(size_t)t]));
                if (d <= 1e-17) break;  // unit-scale vector: converged to rounding
            }
            std::copy(v.begin(), v.end(), E.begin() + (size_t)(m * ng));
            // the 19-point stencil of a row reads only cells within one cell of
            // it: rows next to a nonzero entry of the ghosted image (the unit
            // cell, its periodic images, the wall ghosts it sets) are the
            // column's nonzeros
            long nrows = 0;
            for (long gc = 0; gc < ng; ++gc) {
                if (v[(size_t)gc] == 0.0) continue;
                const int gi = (int)(gc % NX), gj = (int)((gc / NX) % NY), gk = (int)(gc / (NX * NY));
                for (int dk = -1; dk <= 1; ++dk)
                    for (int dj = -1; dj <= 1; ++dj)
                        for (int di = -1; di <= 1; ++di) {
                            const int i = gi + di, j = gj + dj, k = gk + dk;
                            if (i < 1 || i > G.nx || j < 1 || j > G.ny || k < 1 || k > G.nz) continue;
                            const long r = (long)(i - 1) + (long)G.nx * ((j - 1) + (long)G.ny * (k - 1));
                            if (seen[(size_t)r] != m) {
                                seen[(size_t)r] = m;
                                rows[(size_t)nrows++] = r;
                            }
                        }
            }
            for (long q = 0; q < nrows; ++q) {
                const long r = rows[(size_t)q];
                const long c = cell(r);
                Win<double> w;
                win_init(w, v.data(), c - G.sy, G.sy, G.sz);
                win_load(w, v.data(), c, G.sy, G.sz);
                const G2<double> g0 = load_g2<double>(M, c), gmv = load_g2<double>(M, c - G.sy);
                A[(size_t)(r + N * m)] = stencil(w, M, c, G.sz, gmv, g0, inv_pot);
            }
        }
    };
    {
        const long nt = std::max<long>(1, std::min<long>({(long)std::thread::hardware_concurrency(), 16, N / 64}));
        std::vector<std::thread> pool;
        long t = 1;
        for (; t < nt; ++t) {
            try {
                pool.emplace_back(columns, N * t / nt, N * (t + 1) / nt);
            } catch (...) {  // no more threads: this one takes the remaining ranges
                break;
            }
        }
        columns(0, N / nt);
        if (t < nt) columns(N * t / nt, N);
        for (std::thread& th : pool) th.join();
    }
    // Gauss-Jordan inverse of the bordered matrix (partial pivoting) on the
    // device: [B | I] is nb x 2nb, the right half becomes B^-1; then the map
    // K_full = E K (ghosted rows x interior columns) from it on the device
    const long nb = N + 1, w = 2 * nb;
    auto KF = std::make_shared<PinnedBuf>((size_t)(ng * N));
    if (!KF->p) return PLES_MG_ERR_OOM;
    double* kf = nullptr;  // the FP64 map on the device
    {
        std::vector<double> Mh((size_t)(nb * w), 0.0);
        for (long r = 0; r < N; ++r) {
            for (long c = 0; c < N; ++c) Mh[(size_t)(r * w + c)] = A[(size_t)(r + N * c)];
            Mh[(size_t)(r * w + N)] = 1.0;
            Mh[(size_t)(N * w + r)] = 1.0;
        }
        for (long r = 0; r < nb; ++r) Mh[(size_t)(r * w + nb + r)] = 1.0;
        double* Md = nullptr;
        double* fd = nullptr;
        double* Ed = nullptr;
        int* std_ = nullptr;
        CKS(dalloc(h, (void**)&Md, (size_t)(nb * w) * 8));
        CKS(dalloc(h, (void**)&fd, (size_t)nb * 8));
        CKS(dalloc(h, (void**)&Ed, (size_t)(ng * N) * 8));
        CKS(dalloc(h, (void**)&std_, sizeof(int)));
        CKS(dalloc(h, (void**)&kf, (size_t)(ng * N) * 8));
        CK(hipMemcpyAsync(Md, Mh.data(), (size_t)(nb * w) * 8, hipMemcpyHostToDevice, h->s));
        CK(hipMemcpyAsync(Ed, E.data(), (size_t)(ng * N) * 8, hipMemcpyHostToDevice, h->s));
        CK(hipMemsetAsync(std_, 0, sizeof(int), h->s));
        for (long col = 0; col < nb; ++col) {
            hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(1024), 0, h->s, Md, nb, col, fd, std_);
            const long nc = w - col - 1;
            if (nc > 0)
                hipLaunchKernelGGL(k_gj_elim, dim3((unsigned)((nc + 255) / 256), (unsigned)nb), dim3(256), 0, h->s, Md,
                                   nb, col, (const double*)fd);
        }
        hipLaunchKernelGGL(k_dense_map, dim3((unsigned)((N + 255) / 256), (unsigned)ng), dim3(256), 0, h->s, G,
                           (const double*)Ed, (const double*)Md, nb, N, kf);
        CK(hipGetLastError());
        int sing = 0;
        CK(hipMemcpyAsync(KF->p, kf, (size_t)(ng * N) * 8, hipMemcpyDeviceToHost, h->s));
        CK(hipMemcpyAsync(&sing, std_, sizeof(int), hipMemcpyDeviceToHost, h->s));
        CK(hipStreamSynchronize(h->s));
        dfree(h, Md);
        dfree(h, fd);
        dfree(h, Ed);
        dfree(h, std_);
        if (sing) {
            dfree(h, kf);
            return PLES_MG_ERR_NONFINITE;
        }
    }
    for (size_t t = 0; t < KF->n; ++t)
        if (!std::isfinite(KF->p[t])) {
            dfree(h, kf);
            return PLES_MG_ERR_NONFINITE;
        }
    CKS(install_map(h, L, kf));
    CK(hipStreamSynchronize(h->s));
    dfree(h, kf);
    dense_store(G, L.pot, std::move(metcat), std::move(KF));
    return PLES_MG_OK;
}

// Whether the xi-zeta cross metrics g13 and g31 are identically zero on a
// level (orthogonal xi and zeta grid lines everywhere, e.g. terrain-following
// grids over x-z separable meshes): then the stencil kernels are launched in
// their variant that does not load them; the operator is the same.
int find_x0(H* h, Level& L) {
    const long n = L.G.n;
    const int nb = (int)std::min<long>((n + 255) / 256, 1024);
    double* part = nullptr;
    CKS(dalloc(h, (void**)&part, (size_t)nb * 2 * sizeof(double)));
    hipLaunchKernelGGL(k_absmax<double>, dim3(nb), dim3(256), 0, h->s, n, (const double*)L.g64[2], part);
    hipLaunchKernelGGL(k_absmax<double>, dim3(nb), dim3(256), 0, h->s, n, (const double*)L.g64[6], part + nb);
    CK(hipGetLastError());
    std::vector<double> hp((size_t)nb * 2);
    CK(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, h->s));
    CK(hipStreamSynchronize(h->s));
    dfree(h, part);
    double mx = 0.0;
    for (double v : hp) mx = std::max(mx, v);
    L.x0 = mx == 0.0;
    return PLES_MG_OK;
}

// interior sum of metric m of level L (FP64 metrics), column sums on the device
int metric_sum(H* h, const Level& L, int m, double& out) {
    const ColLaunch cl = col_launch(L.G);
    double* part = nullptr;
    CKS(dalloc(h, (void**)&part, (size_t)cl.nblocks * 2 * sizeof(double)));
    hipLaunchKernelGGL(k_sum, cl.grid, cl.block, 0, h->s, L.G, (const double*)L.g64[m], part, cl.nblocks);
    CK(hipGetLastError());
    std::vector<double> hp((size_t)cl.nblocks);
    CK(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, h->s));
    CK(hipStreamSynchronize(h->s));
    dfree(h, part);
    out = 0.0;
    for (double v : hp) out += v;
    return PLES_MG_OK;
}

// Wall-ghost passes with xi or zeta walls (automatic count).  The red/black
// passes contract slowly where the tangential cross metrics are large
// against the normal one (strongly warped walled grids), and a fixed count
// then leaves the outer residual short of its fixed point.  At create, one
// pass at a time is applied to the ghosts of a Gaussian test field; the
// largest ratio of successive ghost changes over the later passes estimates
// the contraction rho, and the counts are those that take the first change
// down to 1e-15 (FP64 fills) and 1e-8 (the FP32 inner direction), within
// [24, 400] and [6, 400].
int auto_halo_passes(H* h) {
    const Level& L0 = h->lv[0];
    const Geo& G = L0.G;
    double* a = h->Z[0];
    double* b = h->Z[1];
    const long n = G.n;
    const int nb = (int)std::min<long>((n + 255) / 256, 1024);
    double* part = nullptr;
    CKS(dalloc(h, (void**)&part, (size_t)nb * sizeof(double)));
    CK(hipMemsetAsync(a, 0, (size_t)n * 8, h->s));
    const ColLaunch ce = cell_launch(G);
    hipLaunchKernelGGL(k_fill_gauss, ce.grid, ce.block, 0, h->s, G, a, (uint64_t)0x5eedULL);
    CK(hipGetLastError());
    std::vector<double> hp((size_t)nb);
    double d1 = 0.0, d4 = 0.0, dprev = 0.0, rho = 0.0;
    int used = 0;
    constexpr int kMaxProbe = 48;
    for (int ps = 1; ps <= kMaxProbe; ++ps) {
        CK(hipMemcpyAsync(b, a, (size_t)n * 8, hipMemcpyDeviceToDevice, h->s));
        CKS(halo64(h, a, false, false, 1));
        hipLaunchKernelGGL(k_absdiff_max, dim3(nb), dim3(256), 0, h->s, n, (const double*)a, (const double*)b, part);
        CK(hipGetLastError());
        CK(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, h->s));
        CK(hipStreamSynchronize(h->s));
        double d = 0.0;
        for (double v : hp) d = std::max(d, v);
        if (!std::isfinite(d)) {
            dfree(h, part);
            return PLES_MG_ERR_NONFINITE;
        }
        if (ps == 1) d1 = d;
        else if (ps >= 4 && dprev > 0.0) rho = std::max(rho, d / dprev);
        if (ps == 4) d4 = d;
        dprev = d;
        used = ps;
        if (d <= 1e-15 * std::max(d1, 1e-300)) break;
    }
    dfree(h, part);
    CK(hipMemsetAsync(a, 0, (size_t)n * 8, h->s));
    CK(hipMemsetAsync(b, 0, (size_t)n * 8, h->s));
    // changes still growing after all probe passes: the face relations have no
    // reachable fixed point on this grid (the reference's ghost iteration,
    // fsrc/multigrid.f03:1109-1317, diverges there too); refuse the grid here
    // rather than return non-finite pressures from the solve
    if (used == kMaxProbe && dprev > d4) return PLES_MG_ERR_UNSUPPORTED;
    auto count = [&](double target, int lo) {
        if (rho <= 0.0) return std::max(lo, used);
        if (rho >= 0.999) return 400;
        const int k = (int)std::ceil(std::log(target) / std::log(rho)) + 1;
        return std::min(400, std::max(lo, k));
    };
    h->o.halo_passes = count(1e-15, 24);
    h->halo_passes32 = count(1e-8, 6);
    return PLES_MG_OK;
}

// 16-bit copy of an FP32 level's metrics for its smoother, scaled by 4^m so
// that the largest normal metric lands in [2^8, 2^10): far from the FP16
// overflow, and values down to 2^-24 of it stay normal numbers.  The scale
// is exact and is taken back out through 1/pot (2^-m).
int make_half_metrics(H* h, Level& L) {
    const long cells = L.G.n;
    const int nb = (int)std::min<long>((cells + 255) / 256, 1024);
    double* part = nullptr;
    CKS(dalloc(h, (void**)&part, (size_t)nb * 3 * sizeof(double)));
    const int diag[3] = {0, 4, 8};
    for (int d = 0; d < 3; ++d)
        hipLaunchKernelGGL(k_absmax<float>, dim3(nb), dim3(256), 0, h->s, cells, (const float*)L.g32[diag[d]],
                           part + d * nb);
    CK(hipGetLastError());
    std::vector<double> hp((size_t)nb * 3);
    CK(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, h->s));
    CK(hipStreamSynchronize(h->s));
    dfree(h, part);
    double mx = 0.0;
    for (double v : hp) mx = std::max(mx, v);
    if (!(mx > 0.0) || !std::isfinite(mx)) return PLES_MG_ERR_NONFINITE;
    const int e = (int)std::floor(std::log2(mx));
    const int m = (int)std::floor((9.0 - (double)e) / 2.0);  // e + 2m in {8, 9}
    const float scale = std::ldexp(1.0f, 2 * m);
    L.h16s = std::ldexp(1.0, -m);
    for (int q = 0; q < 9; ++q) {
        CKS(dalloc(h, (void**)&L.g16[q], (size_t)cells * 2));
        hipLaunchKernelGGL(k_to_half, dim3(nb), dim3(256), 0, h->s, cells, (const float*)L.g32[q], L.g16[q], scale);
    }
    if (L.x0) {  // g13 = g31 = 0: the register smoother reads the packed form
        CKS(dalloc(h, (void**)&L.pk_a, (size_t)cells * sizeof(H2)));
        CKS(dalloc(h, (void**)&L.pk_z, (size_t)cells * sizeof(H2)));
        CKS(dalloc(h, (void**)&L.pk_b, (size_t)cells * sizeof(H2)));
        hipLaunchKernelGGL(k_pack_met16, dim3(nb), dim3(256), 0, h->s, cells, (const float*)L.g32[0],
                           (const float*)L.g32[1], (const float*)L.g32[7], (const float*)L.g32[8],
                           (const float*)L.g32[3], (const float*)L.g32[5], scale, L.pk_a, L.pk_z, L.pk_b);
    }
    CK(hipGetLastError());
    return PLES_MG_OK;
}

// An upper bound where the first coarse level may keep the xi (and eta)
// count (opts.semicoarsen: decided at create from the metrics).
size_t estimate(const ples_mg_grid* g, const ples_mg_opts* o) {
    size_t tot = 0;
    const bool mixed = o && o->precision == PLES_MG_MIXED;
    const bool semi = o && o->semicoarsen;
    // levels the V-cycle visits (create's rule; with semicoarsening the level
    // shapes depend on the grid, so both hierarchies' smallest first level counts)
    int lv = g->nlevel;
    if (o && o->coarse_exact)
        for (int n = 1; n < g->nlevel && lv == g->nlevel; ++n) {
            const int st = 1 << n, sx = semi ? st / 2 : st, sy = (semi && o->semicoarsen == 1) ? st / 2 : st;
            const long full = (long)(g->jx / st) * (g->jy / st) * (g->jz / st);
            const long semic = (long)(g->jx / sx) * (g->jy / sy) * (g->jz / st);
            if (std::min(full, semic) <= dense_max(g->ip, g->kp)) lv = n + 1;
        }
    for (int n = 0; n < g->nlevel; ++n) {
        const int st = 1 << n, sx = (semi && n > 0) ? st / 2 : st;
        const int sy = (semi && n > 0 && o->semicoarsen == 1) ? st / 2 : st;
        const size_t cells = (size_t)(g->jx / sx + 2) * (g->jy / sy + 2) * (g->jz / st + 2);
        const bool f32 = mixed && n < lv - 1;
        const size_t sp = f32 ? 4 : 8;
        tot += cells * 9 * sp;                  // V-cycle metrics
        if (f32 && o->smoother_fp16) tot += cells * (9 * 2 + 4 + 12 + 4);  // smoother's 16-bit metrics, spikes, packed
        if (n == 0 && f32) tot += cells * 9 * 8;  // FP64 outer metrics
        tot += cells * 16;                      // invm, cpr, v, w
        tot += cells * sp * 2;                  // ping-pong
        if (n > 0) tot += cells * sp;           // rhs
        if (n < g->nlevel - 1) tot += cells * sp;  // residual
        if (n == lv - 1 && o && o->coarse_exact) {  // dense map of either hierarchy, the larger
            size_t dense = 0;
            for (int f : {sx, st}) {
                const int fy = f == st ? st : sy;
                const size_t ni = (size_t)(g->jx / f) * (g->jy / fy) * (g->jz / st);
                const size_t cf = (size_t)(g->jx / f + 2) * (g->jy / fy + 2) * (g->jz / st + 2);
                if ((long)ni <= dense_max(g->ip, g->kp)) dense = std::max(dense, cf * ni * 8);
            }
            tot += dense;
        }
    }
    const size_t c0 = (size_t)(g->jx + 2) * (g->jy + 2) * (g->jz + 2);
    tot += c0 * 8 * 7;                                             // x, b, r, Z[2], W[2]
    if (mixed && g->nlevel >= 2) tot += c0 * 4 * 6;                // FP32 inner e, r, Z[2], W[2]
    tot += (size_t)(g->jx + 1) * (g->jy + 1) * (g->jz + 1) * 24;   // nodes
    return tot;
}

}  // namespace

// ---------------------------------------------------------------- projection step
dim3 box_grid(const Geo& G) { return dim3((unsigned)((G.nx + 2 + 63) / 64), (unsigned)((G.ny + 2 + 3) / 4), (unsigned)(G.nz + 2)); }
const dim3 kBoxBlock(64, 4);

// wall directions need three interior cells for the parabolic extrapolation
bool proj_dims_ok(const Geo& G) {
    return G.ny >= 3 && (G.ip == 0 || G.nx >= 3) && (G.kp == 0 || G.nz >= 3);
}

// ====================================================================== C ABI
extern "C" {

void ples_mg_default_opts(ples_mg_opts* o) {
    if (!o) return;
    o->precision = PLES_MG_FP64;
    o->nu_pre = 3;
    o->nu_post = 3;
    o->nu_coarse = 16;
    o->omega = 0.8;
    o->krylov = 1;
    o->stop_mode = PLES_MG_STOP_REL_L2;
    o->project_rhs = 1;
    o->halo_passes = 0;  // auto
    o->device = 0;
    o->graphs = 1;
    o->inner_cycles = 0;  // auto
    o->autotune = 0;
    o->fp32_krylov = 1;
    o->smoother_fp16 = 1;
    o->coarse_exact = 1;
    o->sweep_schedule = 1;
    o->semicoarsen = 0;
}

const char* ples_mg_status_string(int s) {
    switch (s) {
        case PLES_MG_OK: return "ok";
        case PLES_MG_NOT_CONVERGED: return "not converged";
        case PLES_MG_ERR_NONFINITE: return "non-finite value";
        case PLES_MG_ERR_OOM: return "out of device memory";
        case PLES_MG_ERR_BAD_DIMS: return "bad dimensions";
        case PLES_MG_ERR_HIP: return "HIP error";
        case PLES_MG_ERR_ARG: return "bad argument";
        case PLES_MG_ERR_UNSUPPORTED: return "unsupported configuration";
        case PLES_MG_ERR_BUSY: return "not admitted";
        default: return "unknown";
    }
}

size_t ples_mg_estimate_bytes(const ples_mg_grid* g, const ples_mg_opts* o) {
    if (check_grid(g) != PLES_MG_OK) return 0;
    return estimate(g, o);
}

int ples_mg_destroy(ples_mg_handle* h) {
    if (!h) return PLES_MG_OK;
    (void)hipSetDevice(h->dev);
    if (h->s) (void)hipStreamSynchronize(h->s);
    for (auto& r : h->recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    drop_graphs(h);  // before the buffers they reference
    drop_graphs32(h);
    for (auto& a : h->allocs) {
        if (h->pooled > 0) (void)hipFreeAsync(a.first, h->s);
        else (void)hipFree(a.first);
    }
    if (h->s && h->pooled > 0) (void)hipStreamSynchronize(h->s);  // the frees, before the stream goes
    for (int q = 0; q < 4; ++q)
        if (h->sev[q]) (void)hipEventDestroy(h->sev[q]);
    if (h->Sh) (void)hipHostFree(h->Sh);
    if (h->t0) (void)hipEventDestroy(h->t0);
    if (h->t1) (void)hipEventDestroy(h->t1);
    if (h->s) (void)hipStreamDestroy(h->s);
    delete h;
    return PLES_MG_OK;
}

static int create_impl(const ples_mg_grid* g, const ples_mg_opts* oin, H* h) {
    ples_mg_opts o;
    ples_mg_default_opts(&o);
    if (oin) o = *oin;
    if (o.precision != PLES_MG_FP64 && o.precision != PLES_MG_MIXED) return PLES_MG_ERR_ARG;
    if (o.nu_pre < 0 || o.nu_post < 0 || o.nu_coarse < 0 || !(o.omega > 0.0 && o.omega < 2.0))
        return PLES_MG_ERR_ARG;
    const bool auto_passes = o.halo_passes <= 0;
    if (o.halo_passes <= 0) o.halo_passes = (g->ip == 1 || g->kp == 1) ? 24 : 6;
    if (o.inner_cycles < 0 || o.inner_cycles > 8) return PLES_MG_ERR_ARG;
    if (o.semicoarsen < 0 || o.semicoarsen > 2) return PLES_MG_ERR_ARG;
    h->o = o;
    h->inner = o.inner_cycles > 0 ? o.inner_cycles : 1;
    h->dev = o.device;
    CK(hipSetDevice(h->dev));
    CK(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
    // workgroup LDS the coarsest-level solver may use (up to 160 KB per
    // workgroup on this GPU; 64 KB without the opt-in)
    {
        int v = 0;
        h->lds_max = 64 * 1024;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, h->dev) == hipSuccess &&
            v > h->lds_max &&
            hipFuncSetAttribute(reinterpret_cast<const void*>(&k_coarse_lds<double>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, v) == hipSuccess)
            h->lds_max = v;
        (void)hipGetLastError();
    }
    CK(hipEventCreate(&h->t0));
    CK(hipEventCreate(&h->t1));
    // pinned S mirror: [0, S_COUNT) scalars, then 4 divergence maxima, then
    // the 4-slot read-back ring of the refinement loop, then the arming
    // staging of the two scalar sets
    CK(hipHostMalloc((void**)&h->Sh, 8 * S_COUNT * sizeof(double), hipHostMallocDefault));
    std::memset(h->Sh, 0, 8 * S_COUNT * sizeof(double));
    for (int q = 0; q < 4; ++q) CK(hipEventCreateWithFlags(&h->sev[q], hipEventDisableTiming));
    h->jx = g->jx; h->jy = g->jy; h->jz = g->jz;
    h->ip = g->ip; h->jp = g->jp; h->kp = g->kp;
    h->L = g->nlevel;
    h->Lv = h->L;
    const bool mixed = o.precision == PLES_MG_MIXED;
    // nodes
    const size_t nn = (size_t)(g->jx + 1) * (g->jy + 1) * (g->jz + 1);
    const double* src[3] = {g->x, g->y, g->z};
    for (int c = 0; c < 3; ++c) {
        CKS(dalloc(h, (void**)&h->nodes[c], nn * 8));
        CK(hipMemcpyAsync(h->nodes[c], src[c], nn * 8, hipMemcpyHostToDevice, h->s));
    }
    h->lv.resize(h->L);
    if (o.sweep_schedule == 1) {
        h->cheb_hi = kChebHi;
        h->cheb_lo = kChebLo;
    }
    // first coarse level coarsened in zeta only (semicoarsen 1) or in eta and
    // zeta (2) when the fine zeta coupling dominates the xi coupling: the
    // eta-line smoother treats xi and zeta explicitly, and the error it leaves
    // oscillatory in xi (and, where the eta spacing is wide, in eta) is not
    // representable on a level with half those counts
    bool semi = false;
    for (int n = 0; n < h->L; ++n) {
        Level& L = h->lv[n];
        const int st = 1 << n;
        L.sx = (semi && n > 0) ? st / 2 : st;
        L.sy = (semi && n > 0 && o.semicoarsen == 1) ? st / 2 : st;
        L.G = make_geo(g->jx / L.sx, g->jy / L.sy, g->jz / st, g->ip, g->jp, g->kp);
        L.pot = (double)st;
        // the first coarse level small enough for the dense solve ends the V-cycle
        if (h->Lv == h->L && n >= 1 && o.coarse_exact && L.ncell() <= dense_max(g->ip, g->kp)) h->Lv = n + 1;
        L.f32 = mixed && n < h->Lv - 1;
        const size_t cells = (size_t)L.G.n;
        for (int q = 0; q < 9; ++q) CKS(dalloc(h, (void**)&L.g64[q], cells * 8));
        CKS(compute_metrics(h, n, L.g64, nullptr));
        if (n == 0 && g->g[0]) {
            // caller-provided level-1 metrics in the reference shapes
            double* staging = nullptr;
            CKS(dalloc(h, (void**)&staging, cells * 8));
            for (int q = 0; q < 9; ++q) {
                if (!g->g[q]) return PLES_MG_ERR_ARG;
                int i0, i1, j0, j1, k0, k1;
                face_box(L.G, q / 3, i0, i1, j0, j1, k0, k1);
                CKS(pack_box<double>(h, L.G, i0, i1, j0, j1, k0, k1, g->g[q], L.g64[q], staging));
            }
            CK(hipStreamSynchronize(h->s));
            dfree(h, staging);
        }
        CKS(find_x0(h, L));
        if (n == 0 && o.semicoarsen && h->L >= 2) {
            double s11 = 0.0, s33 = 0.0;
            CKS(metric_sum(h, L, 0, s11));
            CKS(metric_sum(h, L, 8, s33));
            semi = s33 >= kSemiRatio * s11;
        }
        if (L.f32) {
            for (int q = 0; q < 9; ++q) {
                CKS(dalloc(h, (void**)&L.g32[q], cells * 4));
                const int nb = (int)std::min<size_t>((cells + 255) / 256, 8192);
                hipLaunchKernelGGL((k_convert<double, float>), dim3(nb), dim3(256), 0, h->s, (long)cells,
                                   (const double*)L.g64[q], L.g32[q]);
            }
            CK(hipGetLastError());
            if (o.smoother_fp16) {
                CKS(make_half_metrics(h, L));
                // fixed rule: the fine level, and the coarser levels (from 512
                // cells) whose segments fit a register variant of the smoother,
                // which reads the packed form where g13 = g31 = 0 (the autotuner
                // chose these on the benchmark shapes; the staged variant and the
                // smallest levels are faster with FP32 metrics); the autotuner
                // times both
                L.use16 = n == 0 || (L.ncell() >= 512 &&
                                     yline_mr(L.G.ny / fixed_segments(L.G.ny, (long)L.G.nx * L.G.nz)) != 0);
            }
            if (n > 0) {
                CK(hipStreamSynchronize(h->s));
                for (int q = 0; q < 9; ++q) {
                    dfree(h, L.g64[q]);
                    L.g64[q] = nullptr;
                }
            }
        }
        L.P = fixed_segments(L.G.ny, (long)L.G.nx * L.G.nz);
        {  // V-cycle residual, fixed rule: twice the smoother's segments (up to 16,
           // an even row count each), 64 lanes up to 8 segments, else 32 (the
           // autotuner's usual choice on the benchmark shapes)
            const int r2 = std::min(2 * L.P, 16);
            L.rP = (L.G.ny % r2 == 0 && L.G.ny / r2 >= 2 && (L.G.ny / r2) % 2 == 0) ? r2 : L.P;
            L.rbx = L.rP <= 8 ? 64 : 32;
        }
        CKS(dalloc(h, (void**)&L.invm, cells * 4));
        CKS(dalloc(h, (void**)&L.cpr, cells * 4));
        if (L.P > 1) {
            CKS(dalloc(h, (void**)&L.lv_v, cells * 4));
            CKS(dalloc(h, (void**)&L.lv_w, cells * 4));
        }
        CKS(level_factors(h, L));
        if (n == h->Lv - 1 && o.coarse_exact && L.ncell() <= dense_max(g->ip, g->kp)) CKS(build_coarse_dense(h, L));
        const size_t sp = L.f32 ? 4 : 8;
        CKS(dalloc(h, &L.p[0], cells * sp));
        CKS(dalloc(h, &L.p[1], cells * sp));
        if (n > 0) CKS(dalloc(h, &L.b, cells * sp));
        if (n < h->L - 1) CKS(dalloc(h, &L.r, cells * sp));
    }
    const size_t c0 = (size_t)h->lv[0].G.n;
    h->snap_bytes = 2 * (size_t)(h->lv[0].G.nx + 2) * (h->lv[0].G.nz + 2) * sizeof(double);
    CKS(dalloc(h, &h->snap, h->snap_bytes));
    CK(hipGetLastError());
    CKS(dalloc(h, (void**)&h->x, c0 * 8));
    CKS(dalloc(h, (void**)&h->bvec, c0 * 8));
    CKS(dalloc(h, (void**)&h->rvec, c0 * 8));
    for (int q = 0; q < 2; ++q) {
        CKS(dalloc(h, (void**)&h->Z[q], c0 * 8));
        CKS(dalloc(h, (void**)&h->W[q], c0 * 8));
    }
    h->ir32 = mixed && h->Lv >= 2 && o.fp32_krylov != 0;
    if (h->ir32) {
        CKS(dalloc(h, (void**)&h->e32, c0 * 4));
        CKS(dalloc(h, (void**)&h->rr32, c0 * 4));
        for (int q = 0; q < 2; ++q) {
            CKS(dalloc(h, (void**)&h->Z32[q], c0 * 4));
            CKS(dalloc(h, (void**)&h->W32[q], c0 * 4));
        }
        CKS(dalloc(h, (void**)&h->S2, S_COUNT * sizeof(double)));
    }
    const size_t csz[6] = {(size_t)g->jy * g->jz, (size_t)g->jy * g->jz, (size_t)g->jx * g->jz,
                           (size_t)g->jx * g->jz, (size_t)g->jx * g->jy, (size_t)g->jx * g->jy};
    for (int q = 0; q < 6; ++q) CKS(dalloc(h, (void**)&h->cs[q], csz[q] * 8));
    CKS(dalloc(h, (void**)&h->S, S_COUNT * sizeof(double)));
    // partial sums of the widest launch any segment count can give (32 lanes)
    h->part_n = std::max(col_launch(h->lv[0].G).nblocks, ((h->lv[0].G.nx + 31) / 32) * h->lv[0].G.nz);
    h->part_n = std::max(h->part_n, 1);
    CKS(dalloc(h, (void**)&h->part, (size_t)h->part_n * 4 * sizeof(double)));
    h->halo_passes32 = h->o.halo_passes;
    if (auto_passes && (g->ip == 1 || g->kp == 1)) CKS(auto_halo_passes(h));
    if (o.autotune >= 0)  // loaded table entries (and timing with autotune > 0)
        for (int n = 0; n < h->L; ++n) CKS(tune_level(h, n, o.autotune > 0));
    CK(hipStreamSynchronize(h->s));
    return PLES_MG_OK;
}

int ples_mg_create(const ples_mg_grid* g, const ples_mg_opts* o, ples_mg_handle** out) {
    if (!out) return PLES_MG_ERR_ARG;
    *out = nullptr;
    CKS(check_grid(g));
    H* h = new H();
    const int st = create_impl(g, o, h);
    if (st != PLES_MG_OK) {
        ples_mg_destroy(h);
        return st;
    }
    *out = h;
    return PLES_MG_OK;
}

void* ples_mg_stream(ples_mg_handle* h) { return h ? (void*)h->s : nullptr; }
size_t ples_mg_device_bytes(const ples_mg_handle* h) { return h ? h->bytes : 0; }

// ---------------------------------------------------------------- rhs helpers
// FP64 compatibility projection of the resident rhs on the handle's stream:
// S[S_SUM] = sum(b), then b -= (S[S_SUM] - bflux) / N when on (no host sync)
static int project_rhs_async(H* h, double bflux, bool on) {
    const Level& L0 = h->lv[0];
    const ColLaunch cl = col_launch(L0.G);
    hipLaunchKernelGGL(k_sum, cl.grid, cl.block, 0, h->s, L0.G, (const double*)h->bvec, h->part, cl.nblocks);
    hipLaunchKernelGGL(k_fin_sum, dim3(1), dim3(256), 0, h->s, (const double*)h->part, cl.nblocks, h->S);
    if (on) {
        const ColLaunch ce = cell_launch(L0.G);
        hipLaunchKernelGGL(k_sub_mean, ce.grid, ce.block, 0, h->s, L0.G, h->bvec, (const double*)h->S, bflux,
                           (double)L0.ncell());
    }
    CK(hipGetLastError());
    return PLES_MG_OK;
}

// the value project_rhs_async took off, once h->Sh holds S
static double removed_mean(const H* h, double bflux, bool on) {
    return on ? (h->Sh[S_SUM] - bflux) / (double)h->lv[0].ncell() : 0.0;
}

static int project_rhs(H* h, double bflux, bool force = false) {
    const bool on = force || h->o.project_rhs;
    CKS(project_rhs_async(h, bflux, on));
    CKS(read_S(h));
    h->last_removed = removed_mean(h, bflux, on);
    return PLES_MG_OK;
}

static double boundary_flux(const H* h, const double* cs[6], double dt) {
    double s = 0.0;
    if (h->ip == 1 && cs[0] && cs[1])
        for (long q = 0; q < (long)h->jy * h->jz; ++q) s += cs[1][q] - cs[0][q];
    if (h->jp == 1 && cs[2] && cs[3])
        for (long q = 0; q < (long)h->jx * h->jz; ++q) s += cs[3][q] - cs[2][q];
    if (h->kp == 1 && cs[4] && cs[5])
        for (long q = 0; q < (long)h->jx * h->jy; ++q) s += cs[5][q] - cs[4][q];
    return s / dt;
}

static int upload_cs(H* h, const double* cs[6]) {
    const size_t csz[6] = {(size_t)h->jy * h->jz, (size_t)h->jy * h->jz, (size_t)h->jx * h->jz,
                           (size_t)h->jx * h->jz, (size_t)h->jx * h->jy, (size_t)h->jx * h->jy};
    bool any = false;
    for (int q = 0; q < 6; ++q) {
        if (cs[q]) {
            any = true;
            CK(hipMemcpyAsync(h->cs[q], cs[q], csz[q] * 8, hipMemcpyHostToDevice, h->s));
        } else {
            CK(hipMemsetAsync(h->cs[q], 0, csz[q] * 8, h->s));
        }
    }
    h->have_cs = any;
    h->cs_sum = boundary_flux(h, cs, 1.0);
    return PLES_MG_OK;
}

// cs_sum of the resident (device) wall fluxes, e.g. after contra computed them
static int cs_sum_device(H* h) {
    std::vector<double> hcs[6];
    const double* csp[6] = {};
    const size_t csz[6] = {(size_t)h->jy * h->jz, (size_t)h->jy * h->jz, (size_t)h->jx * h->jz,
                           (size_t)h->jx * h->jz, (size_t)h->jx * h->jy, (size_t)h->jx * h->jy};
    const int wall[3] = {h->ip, h->jp, h->kp};
    for (int q = 0; q < 6; ++q) {
        if (!wall[q / 2]) continue;
        hcs[q].resize(csz[q]);
        CK(hipMemcpyAsync(hcs[q].data(), h->cs[q], csz[q] * 8, hipMemcpyDeviceToHost, h->s));
        csp[q] = hcs[q].data();
    }
    CK(hipStreamSynchronize(h->s));
    h->cs_sum = boundary_flux(h, csp, 1.0);
    return PLES_MG_OK;
}

static int upload_fi(H* h, const double* fi, int deepl) {
    const Geo& G = h->lv[0].G;
    // fi(0:jx+1,0:jy+1,1-deepl:...) -> planes k = 0..jz+1 start at offset (deepl-1) planes
    CK(hipMemcpyAsync(h->x, fi + (size_t)(deepl - 1) * G.sz, (size_t)G.n * 8, hipMemcpyHostToDevice, h->s));
    return PLES_MG_OK;
}

static int download_fi(H* h, double* fi, int deepl) {
    const Geo& G = h->lv[0].G;
    CK(hipMemcpyAsync(fi + (size_t)(deepl - 1) * G.sz, h->x, (size_t)G.n * 8, hipMemcpyDeviceToHost, h->s));
    CK(hipStreamSynchronize(h->s));
    return PLES_MG_OK;
}

int ples_mg_solve(ples_mg_handle* h, double* fi, int deepl, int deepr, const double* rhs, const double* cs1,
                  const double* cs2, const double* cs3, const double* cs4, const double* cs5, const double* cs6,
                  double dt, double eps, int ficycle, ples_mg_result* res) {
    if (!h || !fi || !rhs || deepl < 1 || deepr < 1 || !(dt > 0.0) || !(eps > 0.0) || ficycle < 0)
        return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    const Geo& G = h->lv[0].G;
    const double* cs[6] = {cs1, cs2, cs3, cs4, cs5, cs6};
    h->dt = dt;
    CKS(upload_cs(h, cs));
    CKS(upload_fi(h, fi, deepl));
    CKS(pack_box<double>(h, G, 1, G.nx, 1, G.ny, 1, G.nz, rhs, h->bvec, h->W[1]));
    CKS(project_rhs(h, h->cs_sum / dt));
    const int st = solve_core(h, eps, ficycle, res);
    if (st != PLES_MG_OK && st != PLES_MG_NOT_CONVERGED && st != PLES_MG_ERR_NONFINITE) return st;
    CKS(download_fi(h, fi, deepl));
    return st;
}

int ples_mg_residual(ples_mg_handle* h, const double* fi, int deepl, int deepr, const double* rhs,
                     const double* cs1, const double* cs2, const double* cs3, const double* cs4, const double* cs5,
                     const double* cs6, double dt, double* r) {
    if (!h || !fi || !rhs || !r || deepl < 1 || deepr < 1 || !(dt > 0.0)) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    const Level& L0 = h->lv[0];
    const Geo& G = L0.G;
    const double* cs[6] = {cs1, cs2, cs3, cs4, cs5, cs6};
    h->dt = dt;
    CKS(upload_cs(h, cs));
    double* p = h->Z[0];
    double* b = h->Z[1];
    double* out = h->W[0];
    CK(hipMemcpyAsync(p, fi + (size_t)(deepl - 1) * G.sz, (size_t)G.n * 8, hipMemcpyHostToDevice, h->s));
    CKS(pack_box<double>(h, G, 1, G.nx, 1, G.ny, 1, G.nz, rhs, b, h->W[1]));
    CKS(halo64(h, p, true));
    const ColLaunch cl = seg_launch(G, L0.P);
    hipLaunchKernelGGL((k_resid<0, double, double, double, double>), cl.grid, cl.block, 0, h->s, G,
                       met<double>(L0), (const double*)p, (const double*)b, out, 1.0, (const double*)nullptr,
                       (const double*)nullptr, (double*)nullptr, 0, L0.P);
    CK(hipGetLastError());
    return unpack_box<double>(h, G, 1, G.nx, 1, G.ny, 1, G.nz, out, r, h->W[1]);
}

int ples_mg_metrica(ples_mg_handle* h, double* const g[9], double* const csv[9], double* giac) {
    if (!h) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    const Geo& G = h->lv[0].G;
    const size_t cells = (size_t)G.n;
    std::vector<double*> tmp;
    double* gg[9];
    double* vv[9];
    for (int q = 0; q < 9; ++q) {
        CKS(dalloc(h, (void**)&gg[q], cells * 8));
        CKS(dalloc(h, (void**)&vv[q], cells * 8));
        tmp.push_back(gg[q]);
        tmp.push_back(vv[q]);
    }
    double* gi = nullptr;
    CKS(dalloc(h, (void**)&gi, cells * 8));
    tmp.push_back(gi);
    CKS(compute_metrics(h, 0, gg, vv));
    {
        dim3 block(64, 4);
        dim3 grid((G.nx + 63) / 64, (G.ny + 3) / 4, G.nz);
        hipLaunchKernelGGL(k_giac, grid, block, 0, h->s, nodes_of(h), G, gi);
        CK(hipGetLastError());
    }
    int st = PLES_MG_OK;
    for (int q = 0; q < 9 && st == PLES_MG_OK; ++q) {
        int i0, i1, j0, j1, k0, k1;
        face_box(G, q / 3, i0, i1, j0, j1, k0, k1);
        if (g && g[q]) st = unpack_box<double>(h, G, i0, i1, j0, j1, k0, k1, gg[q], g[q], h->W[1]);
        if (st == PLES_MG_OK && csv && csv[q])
            st = unpack_box<double>(h, G, i0, i1, j0, j1, k0, k1, vv[q], csv[q], h->W[1]);
    }
    if (st == PLES_MG_OK && giac) st = unpack_box<double>(h, G, 1, G.nx, 1, G.ny, 1, G.nz, gi, giac, h->W[1]);
    (void)hipStreamSynchronize(h->s);
    for (double* p : tmp) {
        dfree(h, p);
    }
    return st;
}

int ples_mg_level_metrics(ples_mg_handle* h, int level, double* const g[9]) {
    if (!h || !g || level < 1 || level > h->L) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    const int n = level - 1;
    const Geo& G = h->lv[n].G;
    const size_t cells = (size_t)G.n;
    double* gg[9];
    for (int q = 0; q < 9; ++q) CKS(dalloc(h, (void**)&gg[q], cells * 8));
    int st = compute_metrics(h, n, gg, nullptr);
    if (st == PLES_MG_OK && n == 0 && h->lv[0].g64[0]) {
        for (int q = 0; q < 9; ++q)
            (void)hipMemcpyAsync(gg[q], h->lv[0].g64[q], cells * 8, hipMemcpyDeviceToDevice, h->s);
    }
    for (int q = 0; q < 9 && st == PLES_MG_OK; ++q) {
        if (!g[q]) continue;
        if (hipMemcpyAsync(g[q], gg[q], cells * 8, hipMemcpyDeviceToHost, h->s) != hipSuccess) st = PLES_MG_ERR_HIP;
    }
    (void)hipStreamSynchronize(h->s);
    for (int q = 0; q < 9; ++q) {
        dfree(h, gg[q]);
    }
    return st;
}

// host double <-> device level array of element type T (ghosted level layout)
static int level_upload(H* h, void* dst, bool f32, const double* src, long n) {
    if (!f32) {
        CK(hipMemcpyAsync(dst, src, (size_t)n * 8, hipMemcpyHostToDevice, h->s));
        CK(hipStreamSynchronize(h->s));
        return PLES_MG_OK;
    }
    std::vector<float> t((size_t)n);
    for (long q = 0; q < n; ++q) t[(size_t)q] = (float)src[q];
    CK(hipMemcpyAsync(dst, t.data(), (size_t)n * 4, hipMemcpyHostToDevice, h->s));
    CK(hipStreamSynchronize(h->s));
    return PLES_MG_OK;
}

static int level_download(H* h, double* dst, bool f32, const void* src, long n) {
    if (!f32) {
        CK(hipMemcpyAsync(dst, src, (size_t)n * 8, hipMemcpyDeviceToHost, h->s));
        CK(hipStreamSynchronize(h->s));
        return PLES_MG_OK;
    }
    std::vector<float> t((size_t)n);
    CK(hipMemcpyAsync(t.data(), src, (size_t)n * 4, hipMemcpyDeviceToHost, h->s));
    CK(hipStreamSynchronize(h->s));
    for (long q = 0; q < n; ++q) dst[q] = (double)t[(size_t)q];
    return PLES_MG_OK;
}

int ples_mg_level_op(ples_mg_handle* h, int level, int op, int P, double omega, const double* in0,
                     const double* in1, double* out) {
    if (!h || !in0 || !out || level < 1 || level > h->L) return PLES_MG_ERR_ARG;
    const int n = level - 1;
    Level& L = h->lv[n];
    const bool coarse_op = op == PLES_MG_OP_RESTRICT || op == PLES_MG_OP_RESID_RESTRICT || op == PLES_MG_OP_PROLONG;
    if (coarse_op && n == h->L - 1) return PLES_MG_ERR_ARG;
    if ((op == PLES_MG_OP_SMOOTH || op == PLES_MG_OP_RESID || op == PLES_MG_OP_RESID_RESTRICT ||
         op == PLES_MG_OP_PROLONG) && !in1)
        return PLES_MG_ERR_ARG;
    if (op == PLES_MG_OP_RESID && !L.r) return PLES_MG_ERR_ARG;
    if (op == PLES_MG_OP_SMOOTH && (P < 0 || (P > 0 && !seg_ok(L.G.ny, P)) || !(omega > 0.0 && omega < 2.0)))
        return PLES_MG_ERR_ARG;
    if (op < PLES_MG_OP_SMOOTH || op > PLES_MG_OP_PROLONG) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    CK(hipStreamSynchronize(h->s));
    const long nf = L.G.n;
    // level-0 rhs of the smoother and residual: the FP64 scratch vector rvec
    // (the V-cycle reads the fine rhs as r0); coarser levels keep theirs in b
    const double* r0 = h->rvec;
    auto upload_rhs = [&]() -> int {
        if (n == 0) return level_upload(h, h->rvec, false, in1, nf);
        return level_upload(h, L.b, L.f32, in1, nf);
    };
    const float* rhs32 = h->rhs32;
    h->rhs32 = nullptr;
    int st = PLES_MG_OK;
    switch (op) {
        case PLES_MG_OP_SMOOTH: {
            CKS(level_upload(h, L.p[L.cur], L.f32, in0, nf));
            CKS(upload_rhs());
            const int P0 = L.P;
            const double om0 = h->o.omega;
            if (P > 0 && P != L.P) {
                if (P > 1 && !L.lv_v) CKS(dalloc(h, (void**)&L.lv_v, (size_t)nf * 4));
                if (P > 1 && !L.lv_w) CKS(dalloc(h, (void**)&L.lv_w, (size_t)nf * 4));
                L.P = P;
                CKS(level_factors(h, L));
            }
            h->o.omega = omega;  // one sweep: the constant damping
            st = lv_smooth(h, n, r0, 1, false);
            h->o.omega = om0;
            if (st == PLES_MG_OK) st = level_download(h, out, L.f32, L.p[L.cur], nf);
            if (L.P != P0) {
                L.P = P0;
                CKS(level_factors(h, L));
            }
            break;
        }
        case PLES_MG_OP_RESID:
            CKS(level_upload(h, L.p[L.cur], L.f32, in0, nf));
            CKS(upload_rhs());
            st = lv_resid(h, n, r0);
            if (st == PLES_MG_OK) st = level_download(h, out, L.f32, L.r, nf);
            break;
        case PLES_MG_OP_RESTRICT: {
            Level& C = h->lv[n + 1];
            CKS(level_upload(h, L.r, L.f32, in0, nf));
            CK(hipMemsetAsync(C.b, 0, (size_t)C.G.n * (C.f32 ? 4 : 8), h->s));
            st = lv_restrict(h, n);
            if (st == PLES_MG_OK) st = level_download(h, out, C.f32, C.b, C.G.n);
            break;
        }
        case PLES_MG_OP_RESID_RESTRICT: {
            Level& C = h->lv[n + 1];
            CKS(level_upload(h, L.p[L.cur], L.f32, in0, nf));
            CKS(upload_rhs());
            CK(hipMemsetAsync(C.b, 0, (size_t)C.G.n * (C.f32 ? 4 : 8), h->s));
            st = lv_resid_restrict(h, n, r0);
            if (st == PLES_MG_OK) st = level_download(h, out, C.f32, C.b, C.G.n);
            break;
        }
        case PLES_MG_OP_PROLONG: {
            Level& C = h->lv[n + 1];
            CKS(level_upload(h, C.p[C.cur], C.f32, in0, C.G.n));
            CKS(level_upload(h, L.p[L.cur], L.f32, in1, nf));
            st = lv_prolong(h, n);
            if (st == PLES_MG_OK) st = level_download(h, out, L.f32, L.p[L.cur], nf);
            break;
        }
        default: st = PLES_MG_ERR_ARG;
    }
    h->rhs32 = rhs32;
    CK(hipStreamSynchronize(h->s));
    return st;
}

int ples_mg_diver(ples_mg_handle* h, const double* uc, const double* vc, const double* wc, double dt, double* rhs) {
    if (!h || !uc || !vc || !wc || !rhs || !(dt > 0.0)) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    const Geo& G = h->lv[0].G;
    double* du = h->Z[0];
    double* dv = h->Z[1];
    double* dw = h->W[0];
    double* out = h->rvec;
    CKS(pack_box<double>(h, G, 0, G.nx, 1, G.ny, 1, G.nz, uc, du, h->W[1]));
    CK(hipStreamSynchronize(h->s));
    CKS(pack_box<double>(h, G, 1, G.nx, 0, G.ny, 1, G.nz, vc, dv, h->W[1]));
    CK(hipStreamSynchronize(h->s));
    CKS(pack_box<double>(h, G, 1, G.nx, 1, G.ny, 0, G.nz, wc, dw, h->W[1]));
    const ColLaunch cl = col_launch(G);
    hipLaunchKernelGGL(k_diver, cl.grid, cl.block, 0, h->s, G, (const double*)du, (const double*)dv,
                       (const double*)dw, dt, out);
    CK(hipGetLastError());
    return unpack_box<double>(h, G, 1, G.nx, 1, G.ny, 1, G.nz, out, rhs, h->W[1]);
}

int ples_mg_gradie(ples_mg_handle* h, const double* fi, int deepl, int deepr, double dt, double* gra1,
                   double* gra2, double* gra3, double* cgra1, double* cgra2, double* cgra3) {
    if (!h || !fi || deepl < 1 || deepr < 1) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    const Level& L0 = h->lv[0];
    const Geo& G = L0.G;
    const size_t cells = (size_t)G.n;
    std::vector<double*> tmp;
    double* gg[9];
    double* vv[9];
    double* gi = nullptr;
    double* outs[6];
    for (int q = 0; q < 9; ++q) {
        CKS(dalloc(h, (void**)&gg[q], cells * 8));
        CKS(dalloc(h, (void**)&vv[q], cells * 8));
        tmp.push_back(gg[q]);
        tmp.push_back(vv[q]);
    }
    CKS(dalloc(h, (void**)&gi, cells * 8));
    tmp.push_back(gi);
    for (int q = 0; q < 6; ++q) {
        CKS(dalloc(h, (void**)&outs[q], cells * 8));
        tmp.push_back(outs[q]);
    }
    CKS(compute_metrics(h, 0, gg, vv));
    {
        dim3 block(64, 4);
        dim3 grid((G.nx + 63) / 64, (G.ny + 3) / 4, G.nz);
        hipLaunchKernelGGL(k_giac, grid, block, 0, h->s, nodes_of(h), G, gi);
    }
    // gradie works with the metrics the solver uses on level 1
    Met<double> m = met<double>(L0);
    double* p = h->Z[0];
    CK(hipMemcpyAsync(p, fi + (size_t)(deepl - 1) * G.sz, cells * 8, hipMemcpyHostToDevice, h->s));
    CsvP V;
    for (int q = 0; q < 9; ++q) V.v[q] = vv[q];
    {
        dim3 block(64, 4);
        dim3 grid((G.nx + 1 + 63) / 64, (G.ny + 1 + 3) / 4, G.nz + 1);
        hipLaunchKernelGGL(k_gradie, grid, block, 0, h->s, G, (const double*)p, V, (const double*)gi, m, dt,
                           outs[0], outs[1], outs[2], outs[3], outs[4], outs[5]);
        CK(hipGetLastError());
    }
    int st = PLES_MG_OK;
    double* dst[6] = {gra1, gra2, gra3, cgra1, cgra2, cgra3};
    for (int q = 0; q < 6 && st == PLES_MG_OK; ++q) {
        if (!dst[q]) continue;
        int i0 = 1, i1 = G.nx, j0 = 1, j1 = G.ny, k0 = 1, k1 = G.nz;
        if (q == 3) i0 = 0;
        if (q == 4) j0 = 0;
        if (q == 5) k0 = 0;
        st = unpack_box<double>(h, G, i0, i1, j0, j1, k0, k1, outs[q], dst[q], h->W[1]);
    }
    (void)hipStreamSynchronize(h->s);
    for (double* t : tmp) {
        dfree(h, t);
    }
    return st;
}

int ples_mg_case_rhs(ples_mg_handle* h, uint64_t seed, double dt) {
    if (!h || !(dt > 0.0)) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    const Geo& G = h->lv[0].G;
    const size_t cells = (size_t)G.n;
    double* gg[9];
    double* vv[9];
    for (int q = 0; q < 9; ++q) {
        CKS(dalloc(h, (void**)&gg[q], cells * 8));
        CKS(dalloc(h, (void**)&vv[q], cells * 8));
    }
    int st = compute_metrics(h, 0, gg, vv);
    CsvP V;
    for (int q = 0; q < 9; ++q) V.v[q] = vv[q];
    const ColLaunch cl = col_launch(G);
    if (st == PLES_MG_OK) {
        hipLaunchKernelGGL(k_case_rhs, cl.grid, cl.block, 0, h->s, G, V, seed, dt, h->bvec, h->part, cl.nblocks);
        if (hipGetLastError() != hipSuccess) st = PLES_MG_ERR_HIP;
    }
    h->dt = dt;
    const double* none[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    if (st == PLES_MG_OK) st = upload_cs(h, none);
    if (st == PLES_MG_OK) st = project_rhs(h, 0.0);
    (void)hipStreamSynchronize(h->s);
    for (int q = 0; q < 9; ++q) {
        dfree(h, gg[q]);
        dfree(h, vv[q]);
    }
    return st;
}

int ples_mg_zero_fi(ples_mg_handle* h) {
    if (!h) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    CK(hipMemsetAsync(h->x, 0, (size_t)h->lv[0].G.n * 8, h->s));
    CK(hipStreamSynchronize(h->s));
    return PLES_MG_OK;
}

int ples_mg_perturb_fi(ples_mg_handle* h, uint64_t seed, double rel) {
    if (!h) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    const Level& L0 = h->lv[0];
    const ColLaunch cl = col_launch(L0.G);
    hipLaunchKernelGGL(k_sum, cl.grid, cl.block, 0, h->s, L0.G, (const double*)h->x, h->part, cl.nblocks);
    hipLaunchKernelGGL(k_fin_sum, dim3(1), dim3(256), 0, h->s, (const double*)h->part, cl.nblocks, h->S);
    const ColLaunch ce = cell_launch(L0.G);
    hipLaunchKernelGGL(k_perturb, ce.grid, ce.block, 0, h->s, L0.G, h->x, (const double*)h->S, seed, rel,
                       (double)L0.ncell());
    CK(hipGetLastError());
    CK(hipStreamSynchronize(h->s));
    return PLES_MG_OK;
}

int ples_mg_solve_resident(ples_mg_handle* h, double eps, int maxcycles, ples_mg_result* res) {
    if (!h || !(eps > 0.0) || maxcycles < 0) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    // the FP64 compatibility projection of the resident rhs, every solve
    if (h->o.project_rhs) CKS(project_rhs(h, h->cs_sum / h->dt));
    return solve_core(h, eps, maxcycles, res);
}

int ples_mg_project_rhs(ples_mg_handle* h, double* rhs, const double* cs1, const double* cs2, const double* cs3,
                        const double* cs4, const double* cs5, const double* cs6, double dt, double* removed) {
    if (!h || !rhs || !(dt > 0.0)) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    const Geo& G = h->lv[0].G;
    const double* cs[6] = {cs1, cs2, cs3, cs4, cs5, cs6};
    h->dt = dt;
    CKS(upload_cs(h, cs));
    CKS(pack_box<double>(h, G, 1, G.nx, 1, G.ny, 1, G.nz, rhs, h->bvec, h->W[1]));
    CKS(project_rhs(h, h->cs_sum / dt, true));
    if (removed) *removed = h->last_removed;
    return unpack_box<double>(h, G, 1, G.nx, 1, G.ny, 1, G.nz, h->bvec, rhs, h->W[1]);
}

int ples_mg_get_fi(ples_mg_handle* h, double* fi) {
    if (!h || !fi) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    return download_fi(h, fi, 1);
}

int ples_mg_set_fi(ples_mg_handle* h, const double* fi) {
    if (!h || !fi) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    CKS(upload_fi(h, fi, 1));
    CK(hipStreamSynchronize(h->s));
    return PLES_MG_OK;
}

int ples_mg_get_rhs(ples_mg_handle* h, double* rhs) {
    if (!h || !rhs) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    const Geo& G = h->lv[0].G;
    return unpack_box<double>(h, G, 1, G.nx, 1, G.ny, 1, G.nz, h->bvec, rhs, h->W[1]);
}

int ples_mg_set_rhs(ples_mg_handle* h, const double* rhs) {
    if (!h || !rhs) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    const Geo& G = h->lv[0].G;
    CKS(pack_box<double>(h, G, 1, G.nx, 1, G.ny, 1, G.nz, rhs, h->bvec, h->W[1]));
    CK(hipStreamSynchronize(h->s));
    return PLES_MG_OK;
}

int ples_mg_set_timing(ples_mg_handle* h, int mask) {
    if (!h) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    CK(hipStreamSynchronize(h->s));
    for (auto& r : h->recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    h->recs.clear();
    h->timing = (unsigned)mask;
    h->timing_step = 0;
    return PLES_MG_OK;
}

int ples_mg_set_timing_every(ples_mg_handle* h, int every) {
    if (!h || every < 1) return PLES_MG_ERR_ARG;
    h->timing_every = every;
    h->timing_step = 0;
    return PLES_MG_OK;
}

int ples_mg_kernel_stats(ples_mg_handle* h, int fam, long* launches, double* ms, double* bytes) {
    if (!h || fam < 0 || fam >= PLES_MG_K_COUNT) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    CK(hipStreamSynchronize(h->s));
    long n = 0;
    double t = 0.0, by = 0.0;
    for (auto& r : h->recs) {
        const bool match = r.fam == fam || (fam == PLES_MG_K_SMOOTH && r.fam == PLES_MG_K_SMOOTH_FINE) ||
                           (fam == PLES_MG_K_RESID && r.fam == PLES_MG_K_RESID_FINE);
        if (!match) continue;
        float e = 0.f;
        if (hipEventElapsedTime(&e, r.a, r.b) == hipSuccess) t += e;
        by += r.bytes;
        ++n;
    }
    if (launches) *launches = n;
    if (ms) *ms = t;
    if (bytes) *bytes = by;
    return PLES_MG_OK;
}


// ---------------------------------------------------------------- projection step
static int ensure_proj(H* h) {
    if (h->pj_ready) return PLES_MG_OK;
    const Geo& G = h->lv[0].G;
    if (!proj_dims_ok(G)) return PLES_MG_ERR_BAD_DIMS;
    const size_t cells = (size_t)G.n;
    for (int q = 0; q < 9; ++q) CKS(dalloc(h, (void**)&h->pj_csv[q], cells * 8));
    CKS(dalloc(h, (void**)&h->pj_giac, cells * 8));
    for (int q = 0; q < 3; ++q) {
        CKS(dalloc(h, (void**)&h->pj_vel[q], cells * 8));
        CKS(dalloc(h, (void**)&h->pj_fc[q], cells * 8));
        CKS(dalloc(h, (void**)&h->pj_del[q], cells * 8));
        CKS(dalloc(h, (void**)&h->pj_gra[q], cells * 8));
        CKS(dalloc(h, (void**)&h->pj_cgra[q], cells * 8));
    }
    CKS(dalloc(h, (void**)&h->pj_dmax, 4 * sizeof(double)));
    double* gg[9] = {};
    int st = PLES_MG_OK;
    for (int q = 0; q < 9 && st == PLES_MG_OK; ++q) st = dalloc(h, (void**)&gg[q], cells * 8);
    if (st == PLES_MG_OK) st = compute_metrics(h, 0, gg, h->pj_csv);
    if (st == PLES_MG_OK) {
        dim3 block(64, 4);
        dim3 grid((G.nx + 63) / 64, (G.ny + 3) / 4, G.nz);
        hipLaunchKernelGGL(k_giac, grid, block, 0, h->s, nodes_of(h), G, h->pj_giac);
        if (hipGetLastError() != hipSuccess) st = PLES_MG_ERR_HIP;
    }
    (void)hipStreamSynchronize(h->s);
    for (int q = 0; q < 9; ++q)
        if (gg[q]) dfree(h, gg[q]);
    if (st == PLES_MG_OK) h->pj_ready = true;
    return st;
}

// cell arrays shaped like fi: (0:jx+1, 0:jy+1, 1-deepl:jz+deepr), planes 0..jz+1 used
static int upload_cells(H* h, double* dst, const double* host, int deepl) {
    const Geo& G = h->lv[0].G;
    CK(hipMemcpyAsync(dst, host + (size_t)(deepl - 1) * G.sz, (size_t)G.n * 8, hipMemcpyHostToDevice, h->s));
    return PLES_MG_OK;
}
static int download_cells(H* h, double* host, const double* src, int deepl) {
    const Geo& G = h->lv[0].G;
    CK(hipMemcpyAsync(host + (size_t)(deepl - 1) * G.sz, src, (size_t)G.n * 8, hipMemcpyDeviceToHost, h->s));
    return PLES_MG_OK;
}

// face flux boxes in the reference shapes: uc (0:jx,1:jy,1:jz), vc (1:jx,0:jy,1:jz), wc (1:jx,1:jy,0:jz)
static int upload_faces(H* h, double* const dst[3], const double* const fc[3]) {
    const Geo& G = h->lv[0].G;
    for (int d = 0; d < 3; ++d) {
        int i0, i1, j0, j1, k0, k1;
        face_box(G, d, i0, i1, j0, j1, k0, k1);
        CK(hipMemsetAsync(dst[d], 0, (size_t)G.n * 8, h->s));
        CKS(pack_box<double>(h, G, i0, i1, j0, j1, k0, k1, fc[d], dst[d], h->W[1]));
        CK(hipStreamSynchronize(h->s));  // the staging buffer is reused
    }
    return PLES_MG_OK;
}
static int download_faces(H* h, double* const fc[3], double* const src[3]) {
    const Geo& G = h->lv[0].G;
    for (int d = 0; d < 3; ++d) {
        int i0, i1, j0, j1, k0, k1;
        face_box(G, d, i0, i1, j0, j1, k0, k1);
        CKS(unpack_box<double>(h, G, i0, i1, j0, j1, k0, k1, src[d], fc[d], h->W[1]));
    }
    return PLES_MG_OK;
}

static int launch_update(H* h) {
    const Geo& G = h->lv[0].G;
    Vel3 V{{h->pj_vel[0], h->pj_vel[1], h->pj_vel[2]}};
    CVel3 D{{h->pj_del[0], h->pj_del[1], h->pj_del[2]}};
    hipLaunchKernelGGL(k_update_cells, box_grid(G), kBoxBlock, 0, h->s, G, V, D);
    hipLaunchKernelGGL(k_update_ghosts, box_grid(G), kBoxBlock, 0, h->s, G, V);
    CK(hipGetLastError());
    return PLES_MG_OK;
}

// contra into the resident face fluxes; cs into the solver's boundary fluxes
static int launch_contra(H* h, int potenziale) {
    const Geo& G = h->lv[0].G;
    CsvP V;
    for (int q = 0; q < 9; ++q) V.v[q] = h->pj_csv[q];
    CVel3 U{{h->pj_vel[0], h->pj_vel[1], h->pj_vel[2]}};
    Vel3 F{{h->pj_fc[0], h->pj_fc[1], h->pj_fc[2]}};
    CsOut cs{{h->cs[0], h->cs[1], h->cs[2], h->cs[3], h->cs[4], h->cs[5]}};
    hipLaunchKernelGGL(k_contra, box_grid(G), kBoxBlock, 0, h->s, G, V, U, F, cs, potenziale);
    CK(hipGetLastError());
    h->have_cs = true;
    return PLES_MG_OK;
}

static int launch_gradie(H* h, const double* fi, double dt) {
    const Level& L0 = h->lv[0];
    const Geo& G = L0.G;
    CsvP V;
    for (int q = 0; q < 9; ++q) V.v[q] = h->pj_csv[q];
    dim3 block(64, 4);
    dim3 grid((G.nx + 1 + 63) / 64, (G.ny + 1 + 3) / 4, G.nz + 1);
    hipLaunchKernelGGL(k_gradie, grid, block, 0, h->s, G, fi, V, (const double*)h->pj_giac, met<double>(L0), dt,
                       h->pj_gra[0], h->pj_gra[1], h->pj_gra[2], h->pj_cgra[0], h->pj_cgra[1], h->pj_cgra[2]);
    CK(hipGetLastError());
    return PLES_MG_OK;
}

static int launch_vel_up(H* h, double dt, int potenziale) {
    const Geo& G = h->lv[0].G;
    Vel3 U{{h->pj_vel[0], h->pj_vel[1], h->pj_vel[2]}};
    Vel3 F{{h->pj_fc[0], h->pj_fc[1], h->pj_fc[2]}};
    CVel3 gr{{h->pj_gra[0], h->pj_gra[1], h->pj_gra[2]}};
    CVel3 cg{{h->pj_cgra[0], h->pj_cgra[1], h->pj_cgra[2]}};
    hipLaunchKernelGGL(k_vel_up, box_grid(G), kBoxBlock, 0, h->s, G, U, F, gr, cg, dt, potenziale);
    hipLaunchKernelGGL(k_vel_up_images, box_grid(G), kBoxBlock, 0, h->s, G, U);
    CK(hipGetLastError());
    return PLES_MG_OK;
}

static int launch_divmax(H* h, double out[4]) {
    const Geo& G = h->lv[0].G;
    const ColLaunch cl = col_launch(G);
    CVel3 F{{h->pj_fc[0], h->pj_fc[1], h->pj_fc[2]}};
    hipLaunchKernelGGL(k_divmax, cl.grid, cl.block, 0, h->s, G, F, (const double*)h->pj_giac, h->part, cl.nblocks);
    hipLaunchKernelGGL(k_fin_max, dim3(1), dim3(1), 0, h->s, (const double*)h->part, cl.nblocks, 4, h->pj_dmax);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(h->Sh + S_COUNT, h->pj_dmax, 4 * sizeof(double), hipMemcpyDeviceToHost, h->s));
    CK(hipStreamSynchronize(h->s));
    for (int q = 0; q < 4; ++q) out[q] = h->Sh[S_COUNT + q];
    return PLES_MG_OK;
}

static int download_cs(H* h, double* const cs[6]) {
    const size_t csz[6] = {(size_t)h->jy * h->jz, (size_t)h->jy * h->jz, (size_t)h->jx * h->jz,
                           (size_t)h->jx * h->jz, (size_t)h->jx * h->jy, (size_t)h->jx * h->jy};
    const int wall[3] = {h->ip, h->jp, h->kp};
    for (int q = 0; q < 6; ++q)
        if (cs && cs[q] && wall[q / 2])
            CK(hipMemcpyAsync(cs[q], h->cs[q], csz[q] * 8, hipMemcpyDeviceToHost, h->s));
    CK(hipStreamSynchronize(h->s));
    return PLES_MG_OK;
}

int ples_mg_update(ples_mg_handle* h, double* u, double* v, double* w, int deepl, int deepr, const double* delu,
                   const double* delv, const double* delw) {
    if (!h || !u || !v || !w || !delu || !delv || !delw || deepl < 1 || deepr < 1) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    CKS(ensure_proj(h));
    const Geo& G = h->lv[0].G;
    double* vel[3] = {u, v, w};
    const double* del[3] = {delu, delv, delw};
    for (int q = 0; q < 3; ++q) {
        CKS(upload_cells(h, h->pj_vel[q], vel[q], deepl));
        CK(hipMemcpyAsync(h->pj_del[q], del[q], (size_t)G.n * 8, hipMemcpyHostToDevice, h->s));
    }
    CKS(launch_update(h));
    for (int q = 0; q < 3; ++q) CKS(download_cells(h, vel[q], h->pj_vel[q], deepl));
    CK(hipStreamSynchronize(h->s));
    return PLES_MG_OK;
}

int ples_mg_contra(ples_mg_handle* h, const double* u, const double* v, const double* w, int deepl, int deepr,
                   double* uc, double* vc, double* wc, double* const cs[6], int potenziale) {
    if (!h || !u || !v || !w || !uc || !vc || !wc || deepl < 1 || deepr < 1) return PLES_MG_ERR_ARG;
    if (potenziale != 0 && potenziale != 1) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    CKS(ensure_proj(h));
    const double* vel[3] = {u, v, w};
    for (int q = 0; q < 3; ++q) CKS(upload_cells(h, h->pj_vel[q], vel[q], deepl));
    const double* fcin[3] = {uc, vc, wc};
    CKS(upload_faces(h, h->pj_fc, fcin));
    CKS(launch_contra(h, potenziale));
    CKS(cs_sum_device(h));
    double* fcout[3] = {uc, vc, wc};
    CKS(download_faces(h, fcout, h->pj_fc));
    return download_cs(h, cs);
}

int ples_mg_vel_up(ples_mg_handle* h, double* u, double* v, double* w, int deepl, int deepr, double* uc,
                   double* vc, double* wc, const double* const gra[3], const double* const cgra[3], double dt,
                   int potenziale) {
    if (!h || !u || !v || !w || !uc || !vc || !wc || !gra || !cgra || deepl < 1 || deepr < 1 || !(dt > 0.0))
        return PLES_MG_ERR_ARG;
    if (potenziale != 0 && potenziale != 1) return PLES_MG_ERR_ARG;
    for (int q = 0; q < 3; ++q)
        if (!gra[q] || !cgra[q]) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    CKS(ensure_proj(h));
    const Geo& G = h->lv[0].G;
    double* vel[3] = {u, v, w};
    for (int q = 0; q < 3; ++q) {
        CKS(upload_cells(h, h->pj_vel[q], vel[q], deepl));
        CK(hipMemsetAsync(h->pj_gra[q], 0, (size_t)G.n * 8, h->s));
        CKS(pack_box<double>(h, G, 1, G.nx, 1, G.ny, 1, G.nz, gra[q], h->pj_gra[q], h->W[1]));
        CK(hipStreamSynchronize(h->s));
    }
    CKS(upload_faces(h, h->pj_cgra, cgra));
    const double* fcin[3] = {uc, vc, wc};
    CKS(upload_faces(h, h->pj_fc, fcin));
    CKS(launch_vel_up(h, dt, potenziale));
    for (int q = 0; q < 3; ++q) CKS(download_cells(h, vel[q], h->pj_vel[q], deepl));
    double* fcout[3] = {uc, vc, wc};
    return download_faces(h, fcout, h->pj_fc);
}

int ples_mg_check_divergence(ples_mg_handle* h, const double* uc, const double* vc, const double* wc,
                             double out[4]) {
    if (!h || !uc || !vc || !wc || !out) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    CKS(ensure_proj(h));
    const double* fcin[3] = {uc, vc, wc};
    CKS(upload_faces(h, h->pj_fc, fcin));
    return launch_divmax(h, out);
}

int ples_mg_projection_step(ples_mg_handle* h, double* u, double* v, double* w, int deepl, int deepr,
                            const double* delu, const double* delv, const double* delw, double* uc, double* vc,
                            double* wc, double* const cs[6], double* fi, double dt, double eps, int ficycle,
                            int potenziale, ples_mg_result* res, double div[4]) {
    if (!h || !u || !v || !w || !delu || !delv || !delw || !uc || !vc || !wc || !fi || deepl < 1 || deepr < 1 ||
        !(dt > 0.0) || !(eps > 0.0) || ficycle < 0)
        return PLES_MG_ERR_ARG;
    if (potenziale != 0 && potenziale != 1) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(h->dev));
    CKS(ensure_proj(h));
    const Level& L0 = h->lv[0];
    const Geo& G = L0.G;
    double* vel[3] = {u, v, w};
    const double* del[3] = {delu, delv, delw};
    for (int q = 0; q < 3; ++q) {
        CKS(upload_cells(h, h->pj_vel[q], vel[q], deepl));
        CK(hipMemcpyAsync(h->pj_del[q], del[q], (size_t)G.n * 8, hipMemcpyHostToDevice, h->s));
    }
    const double* fcin[3] = {uc, vc, wc};
    CKS(upload_faces(h, h->pj_fc, fcin));
    CKS(upload_fi(h, fi, deepl));
    // update -> contra -> diver (fsrc/stratiParticle.f03:2678-2845)
    CKS(launch_update(h));
    CKS(launch_contra(h, potenziale));
    {
        const ColLaunch cl = col_launch(G);
        hipLaunchKernelGGL(k_diver, cl.grid, cl.block, 0, h->s, G, (const double*)h->pj_fc[0],
                           (const double*)h->pj_fc[1], (const double*)h->pj_fc[2], dt, h->bvec);
        CK(hipGetLastError());
    }
    // boundary flux of the compatibility projection from the new cs
    CKS(cs_sum_device(h));
    h->dt = dt;
    CKS(project_rhs(h, h->cs_sum / dt));
    // multi (fsrc/stratiParticle.f03:2881-2887)
    const int st = solve_core(h, eps, ficycle, res);
    if (st != PLES_MG_OK && st != PLES_MG_NOT_CONVERGED) return st;
    // gradie -> vel_up -> check_divergence (fsrc/stratiParticle.f03:2932-2937, 3526)
    CKS(launch_gradie(h, h->x, dt));
    CKS(launch_vel_up(h, dt, potenziale));
    double dm[4];
    CKS(launch_divmax(h, dm));
    if (div)
        for (int q = 0; q < 4; ++q) div[q] = dm[q];
    for (int q = 0; q < 3; ++q) CKS(download_cells(h, vel[q], h->pj_vel[q], deepl));
    CKS(download_fi(h, fi, deepl));
    double* fcout[3] = {uc, vc, wc};
    CKS(download_faces(h, fcout, h->pj_fc));
    CKS(download_cs(h, cs));
    return st;
}


// ---------------------------------------------------------------- batched solves
}  // extern "C"

struct ples_mg_batch {
    int dev = 0;
    std::vector<H*> hs;
    hipStream_t ms = nullptr;  // batch stream: graph launches, gathers, read-backs
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    int K = 2;                              // refinement steps per launch
    std::vector<int> q;                     // (Z, W) pair of each problem's next step
    hipGraphExec_t gexec[2] = {};           // K steps of every problem, by start state
    std::vector<std::vector<int>> gin[2], gout[2];
    const double** dS = nullptr;            // device array of the problems' S
    double* Sall = nullptr;                 // gathered S (device)
    double* ring = nullptr;                 // pinned: 2 slots x n x S_COUNT
    hipEvent_t rev[2] = {};
    size_t bytes = 0;
    // grouped launches: the problems of one family (identical launch
    // sequences: same shape and options) share one launch per kernel of the
    // K steps; tables of per-problem arguments per graph slot
    bool grouped = true;
    struct Fam {
        std::vector<int> mem;       // problem indices
        std::vector<KRec> recs;     // launch shapes (entries unused)
        std::vector<size_t> off;    // table offset of each launch
    };
    std::vector<Fam> fam[2];
    void* gtab[2] = {};
    // mixed handles with the FP32 inner Krylov: device-driven iterative
    // refinement steps (inner FP32 GCR step + refinement tail), else FP64 GCR
    // steps around the FP32 V-cycle
    bool ir = false;
};

namespace {

using B = ples_mg_batch;

std::vector<std::vector<int>> batch_state(const B* b);
void set_batch_state(B* b, const std::vector<std::vector<int>>& v);

void batch_drop_graph(B* b) {
    if (b->ms && (b->gtab[0] || b->gtab[1])) (void)hipStreamSynchronize(b->ms);  // tables in use by a replay
    for (int g = 0; g < 2; ++g) {
        if (b->gexec[g]) (void)hipGraphExecDestroy(b->gexec[g]);
        b->gexec[g] = nullptr;
        b->gin[g].clear();
        b->gout[g].clear();
        b->fam[g].clear();
        if (b->gtab[g]) (void)hipFree(b->gtab[g]);
        b->gtab[g] = nullptr;
    }
}

bool same_launch(const KRec& a, const KRec& c) {
    return a.go == c.go && a.g.x == c.g.x && a.g.y == c.g.y && a.b.x == c.b.x && a.b.y == c.b.y &&
           a.b.z == c.b.z && a.shm == c.shm && a.e.size() == c.e.size();
}

// Record K steps of every problem (the launches they would enqueue), group
// the problems whose launch sequences match into families, and upload one
// table per (family, launch) of the problems' entries into slot g.  Every
// problem's host state advances by K steps, as by a capture.  Returns 1
// (nothing uploaded, states as on entry) where a launch cannot be grouped.
int batch_group_prepare(B* b, int slot) {
    const int n = (int)b->hs.size();
    const auto before = batch_state(b);
    std::vector<std::vector<KRec>> rec(n);
    int st = PLES_MG_OK;
    for (int i = 0; i < n && st == PLES_MG_OK; ++i) {
        H* h = b->hs[i];
        h->rec = &rec[i];
        for (int k = 0; k < b->K && st == PLES_MG_OK; ++k) {
            st = b->ir ? step_ir_body(h, b->q[i]) : step_body(h, b->q[i]);
            b->q[i] ^= 1;
        }
        h->rec = nullptr;
    }
    bool ok = st == PLES_MG_OK;
    for (int i = 0; i < n && ok; ++i)
        for (const KRec& r : rec[i]) ok = ok && r.go != nullptr;
    if (!ok) {
        set_batch_state(b, before);
        return st != PLES_MG_OK ? st : 1;
    }
    std::vector<B::Fam>& fams = b->fam[slot];
    fams.clear();
    for (int i = 0; i < n; ++i) {
        B::Fam* f = nullptr;
        for (B::Fam& c : fams) {
            const auto& r0 = rec[c.mem[0]];
            if (r0.size() != rec[i].size()) continue;
            bool eq = true;
            for (size_t r = 0; r < r0.size() && eq; ++r) eq = same_launch(r0[r], rec[i][r]);
            if (eq) {
                f = &c;
                break;
            }
        }
        if (!f) {
            fams.push_back(B::Fam{});
            f = &fams.back();
        }
        f->mem.push_back(i);
    }
    size_t tot = 0;
    for (B::Fam& f : fams) {
        f.recs = rec[f.mem[0]];
        f.off.resize(f.recs.size());
        for (size_t r = 0; r < f.recs.size(); ++r) {
            f.off[r] = tot;
            tot += (f.recs[r].e.size() * f.mem.size() + 255) / 256 * 256;
        }
    }
    std::vector<unsigned char> host(std::max<size_t>(tot, 256));
    for (const B::Fam& f : fams)
        for (size_t r = 0; r < f.recs.size(); ++r) {
            const size_t es = f.recs[r].e.size();
            for (size_t m = 0; m < f.mem.size(); ++m)
                std::memcpy(host.data() + f.off[r] + m * es, rec[f.mem[m]][r].e.data(), es);
        }
    if (b->gtab[slot]) (void)hipFree(b->gtab[slot]);
    b->gtab[slot] = nullptr;
    CK(hipMalloc(&b->gtab[slot], host.size()));
    CK(hipMemcpy(b->gtab[slot], host.data(), host.size(), hipMemcpyHostToDevice));
    return PLES_MG_OK;
}

// the grouped K steps of slot g: each family on the stream of its first
// problem (forked from and joined back into the batch stream), one launch
// per recorded kernel over the family's problems, then the gather of all S
int batch_group_body(B* b, int slot) {
    const int n = (int)b->hs.size();
    CK(hipEventRecord(b->fork, b->ms));
    for (const B::Fam& f : b->fam[slot]) {
        const int lead = f.mem[0];
        hipStream_t s = b->hs[lead]->s;
        CK(hipStreamWaitEvent(s, b->fork, 0));
        for (size_t r = 0; r < f.recs.size(); ++r) {
            const KRec& k = f.recs[r];
            CK(k.go((const unsigned char*)b->gtab[slot] + f.off[r], (int)f.mem.size(), k.g, k.b, k.shm, s));
        }
        CK(hipEventRecord(b->join[lead], s));
        CK(hipStreamWaitEvent(b->ms, b->join[lead], 0));
    }
    const int tot = n * S_COUNT;
    hipLaunchKernelGGL(k_gather_S, dim3((tot + 255) / 256), dim3(256), 0, b->ms, (const double* const*)b->dS, n,
                       b->Sall);
    CK(hipGetLastError());
    return PLES_MG_OK;
}

// K steps of every problem on its own stream, forked from and joined back
// into the batch stream, then the gather of all S into Sall
int batch_chunk_body(B* b) {
    const int n = (int)b->hs.size();
    CK(hipEventRecord(b->fork, b->ms));
    for (int i = 0; i < n; ++i) {
        H* h = b->hs[i];
        CK(hipStreamWaitEvent(h->s, b->fork, 0));
        for (int k = 0; k < b->K; ++k) {
            CKS(b->ir ? step_ir_body(h, b->q[i]) : step_body(h, b->q[i]));
            b->q[i] ^= 1;
        }
        CK(hipEventRecord(b->join[i], h->s));
        CK(hipStreamWaitEvent(b->ms, b->join[i], 0));
    }
    const int tot = n * S_COUNT;
    hipLaunchKernelGGL(k_gather_S, dim3((tot + 255) / 256), dim3(256), 0, b->ms, (const double* const*)b->dS, n,
                       b->Sall);
    CK(hipGetLastError());
    return PLES_MG_OK;
}

std::vector<std::vector<int>> batch_state(const B* b) {
    std::vector<std::vector<int>> v;
    for (size_t i = 0; i < b->hs.size(); ++i) {
        v.push_back(step_state(b->hs[i]));
        v.back().push_back(b->q[i]);
    }
    return v;
}

void set_batch_state(B* b, const std::vector<std::vector<int>>& v) {
    for (size_t i = 0; i < b->hs.size(); ++i) {
        set_step_state(b->hs[i], v[i]);
        b->q[i] = v[i].back();
    }
}

// one launch of K steps for the whole batch: replay of the captured graph
// whose start state (every level's ping-pong, every (Z, W) selector) is the
// current one -- two graphs cover odd K and odd smoothing counts -- else
// capture it; eager launches when graphs are off
int batch_enqueue_chunk(B* b, bool use_graph) {
    if (!use_graph) return batch_chunk_body(b);
    const auto now = batch_state(b);
    int g = -1;
    for (int t = 0; t < 2; ++t)
        if (b->gexec[t] && b->gin[t] == now) g = t;
    if (g < 0) {
        g = !b->gexec[0] ? 0 : (!b->gexec[1] ? 1 : -1);
        if (g < 0) {
            batch_drop_graph(b);
            g = 0;
        }
        // grouped launches: recorded and tabled before the capture (which
        // then leaves the states where the recording advanced them)
        int grp = 1;
        if (b->grouped) {
            grp = batch_group_prepare(b, g);
            if (grp != 1) CKS(grp);
        }
        CK(hipStreamBeginCapture(b->ms, hipStreamCaptureModeGlobal));
        const int st = grp == 0 ? batch_group_body(b, g) : batch_chunk_body(b);
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(b->ms, &gr);
        if (st != PLES_MG_OK) {
            if (gr) (void)hipGraphDestroy(gr);
            return st;
        }
        CK(e);
        const hipError_t ei = hipGraphInstantiate(&b->gexec[g], gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        CK(ei);
        b->gin[g] = now;
        b->gout[g] = batch_state(b);
        set_batch_state(b, now);
    }
    CK(hipGraphLaunch(b->gexec[g], b->ms));
    set_batch_state(b, b->gout[g]);
    return PLES_MG_OK;
}

// Batched mixed solves by device-driven iterative refinement: the host arms
// each problem's first refinement (FP64 residual, FP32 rhs r / |r|, inner
// tolerance: the single solve's first refinement), then replays graphs of K
// steps of the whole batch (each an inner FP32 GCR step and the refinement
// tail) and only reads the stop codes back.  Per problem the steps, decisions
// and arithmetic are those of the single handle's refinement loop.
int batch_solve_ir(B* b, double eps, int maxcycles, int K, ples_mg_result* res) {
    const int n = (int)b->hs.size();
    b->q.assign(n, 0);
    bool use_graph = true;
    for (H* h : b->hs) use_graph = use_graph && h->o.graphs && h->timing == 0;
    hipEvent_t t0, t1;
    CK(hipEventCreate(&t0));
    CK(hipEventCreate(&t1));
    CK(hipEventRecord(t0, b->ms));
    std::vector<double> bnorm(n), rr(n), rmax(n);
    std::vector<int> done(n, 0), cycles(n, 0), status(n, PLES_MG_NOT_CONVERGED);
    for (int i = 0; i < n; ++i) {
        H* h = b->hs[i];
        CK(hipStreamWaitEvent(h->s, t0, 0));
        if (h->o.project_rhs) CKS(project_rhs_async(h, h->cs_sum / h->dt, true));
        CKS(outer_residual(h, h->x, h->rvec));
        CK(hipMemcpyAsync(h->Sh, h->S, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, h->s));
    }
    for (int i = 0; i < n; ++i) {
        H* h = b->hs[i];
        CK(hipStreamSynchronize(h->s));
        if (h->o.project_rhs) h->last_removed = removed_mean(h, h->cs_sum / h->dt, true);
        const bool absmode = h->o.stop_mode == PLES_MG_STOP_ABS_MAX;
        bnorm[i] = std::sqrt(h->Sh[S_BB]) > 0 ? std::sqrt(h->Sh[S_BB]) : 1.0;
        rr[i] = h->Sh[S_RR];
        rmax[i] = h->Sh[S_RMAX];
        const double m = absmode ? rmax[i] : std::sqrt(rr[i]) / bnorm[i];
        const double rn = std::sqrt(rr[i]);
        if (!std::isfinite(rr[i])) {
            status[i] = PLES_MG_ERR_NONFINITE;
            done[i] = 1;
        } else if (m < eps || !(rn > 0.0)) {
            status[i] = PLES_MG_OK;
            done[i] = 1;
        }
        const Level& L0 = h->lv[0];
        if (!done[i]) {
            const ColLaunch ce = cell_launch(L0.G);
            hipLaunchKernelGGL(k_scale_to32, ce.grid, ce.block, 0, h->s, L0.G, (const double*)h->rvec, 1.0 / rn,
                               h->rr32);
            CK(hipMemsetAsync(h->e32, 0, (size_t)L0.G.n * sizeof(float), h->s));
            CK(hipGetLastError());
            const double need = absmode ? eps / std::max(rmax[i], 1e-300) : eps * bnorm[i] / rn;
            const double tol = std::min(0.1, std::max(1e-5, 0.5 * need));
            CKS(arm_stop(h, h->S2, tol * tol, 0));
        }
        // the refinement fields of the outer set (staging after arm_stop's)
        double* st = h->Sh + 6 * S_COUNT + 8;
        st[0] = rn;                             // S_RN
        st[1] = 0.0;                            // S_STALL
        st[2] = m;                              // S_MOLD
        st[3] = eps;                            // S_EPS
        st[4] = bnorm[i];                       // S_BN
        st[5] = (double)maxcycles;              // S_MAXC
        st[6] = absmode ? 1.0 : 0.0;            // S_ABS
        st[7] = 0.0;                            // S_TOTC
        st[8] = 0.0;                            // S_REF
        st[9] = 1.0;                            // S_NOREF
        st[10] = done[i] ? 1.0 : 0.0;           // S_DONE
        CK(hipMemcpyAsync(h->S + S_RN, st, 10 * sizeof(double), hipMemcpyHostToDevice, h->s));
        CK(hipMemcpyAsync(h->S + S_DONE, st + 10, sizeof(double), hipMemcpyHostToDevice, h->s));
        CK(hipEventRecord(b->join[i], h->s));
        CK(hipStreamWaitEvent(b->ms, b->join[i], 0));
    }
    auto all_done = [&]() {
        for (int i = 0; i < n; ++i)
            if (!done[i]) return false;
        return true;
    };
    // every solve stops within maxcycles + 1 steps (the device caps the inner
    // steps at the budget)
    const int max_launch = (maxcycles + 1 + K - 1) / K + 1;
    int enq = 0, seen = 0;
    while (!all_done()) {
        while (enq - seen < 2 && enq < max_launch) {
            CKS(batch_enqueue_chunk(b, use_graph));
            double* slot = b->ring + (size_t)(enq % 2) * n * S_COUNT;
            CK(hipMemcpyAsync(slot, b->Sall, (size_t)n * S_COUNT * sizeof(double), hipMemcpyDeviceToHost, b->ms));
            CK(hipEventRecord(b->rev[enq % 2], b->ms));
            ++enq;
        }
        if (seen == enq) break;
        CK(hipEventSynchronize(b->rev[seen % 2]));
        const double* slot = b->ring + (size_t)(seen % 2) * n * S_COUNT;
        ++seen;
        for (int i = 0; i < n; ++i) {
            if (done[i]) continue;
            const double* Si = slot + (size_t)i * S_COUNT;
            if (Si[S_DONE] == 0.0) continue;
            done[i] = 1;
            status[i] = Si[S_DONE] == 1.0 ? PLES_MG_OK : (Si[S_DONE] == 3.0 ? PLES_MG_ERR_NONFINITE : PLES_MG_NOT_CONVERGED);
            cycles[i] = (int)Si[S_TOTC];
            rr[i] = Si[S_RR];
            rmax[i] = Si[S_RMAX];
        }
    }
    CK(hipStreamSynchronize(b->ms));
    for (int i = 0; i < n; ++i) {
        H* h = b->hs[i];
        if (!done[i] && seen > 0) {  // out of launches (not reached: the device caps every solve)
            const double* Si = b->ring + (size_t)((seen - 1) % 2) * n * S_COUNT + (size_t)i * S_COUNT;
            cycles[i] = (int)Si[S_TOTC];
            rr[i] = Si[S_RR];
            rmax[i] = Si[S_RMAX];
        }
        CKS(halo64(h, h->x, true));
        CK(hipEventRecord(b->join[i], h->s));
        CK(hipStreamWaitEvent(b->ms, b->join[i], 0));
    }
    CK(hipEventRecord(t1, b->ms));
    CK(hipEventSynchronize(t1));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    int worst = PLES_MG_OK;
    for (int i = 0; i < n; ++i) {
        H* h = b->hs[i];
        if (res) {
            res[i].status = status[i];
            res[i].cycles = cycles[i];
            res[i].rel_res = std::sqrt(rr[i]) / bnorm[i];
            res[i].abs_res = rmax[i];
            res[i].rhs_mean_removed = h->last_removed;
            res[i].seconds = ms * 1e-3;
        }
        if (status[i] != PLES_MG_OK && worst == PLES_MG_OK) worst = status[i];
    }
    return worst;
}

int batch_set_done(B* b, int i, double v) {
    H* h = b->hs[i];
    h->Sh[S_DONE] = v;
    CK(hipMemcpyAsync(h->S + S_DONE, h->Sh + S_DONE, sizeof(double), hipMemcpyHostToDevice, b->ms));
    return PLES_MG_OK;
}

}  // namespace

extern "C" {

int ples_mg_batch_create(int n, const ples_mg_grid* grids, const ples_mg_opts* opts, ples_mg_batch** out) {
    if (!out || n < 1 || !grids) return PLES_MG_ERR_ARG;
    *out = nullptr;
    B* b = new B();
    int st = PLES_MG_OK;
    for (int i = 0; i < n && st == PLES_MG_OK; ++i) {
        ples_mg_handle* h = nullptr;
        st = ples_mg_create(&grids[i], opts, &h);
        if (st == PLES_MG_OK) {
            b->hs.push_back(h);
            b->bytes += h->bytes;
        }
    }
    auto fail = [&](int code) {
        ples_mg_batch_destroy(b);
        return code;
    };
    if (st != PLES_MG_OK) return fail(st);
    b->dev = b->hs[0]->dev;
    b->ir = true;
    for (H* h : b->hs) b->ir = b->ir && h->ir32;
    if (hipSetDevice(b->dev) != hipSuccess) return fail(PLES_MG_ERR_HIP);
    if (hipStreamCreateWithFlags(&b->ms, hipStreamNonBlocking) != hipSuccess) return fail(PLES_MG_ERR_HIP);
    if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) return fail(PLES_MG_ERR_HIP);
    b->join.resize(n);
    for (int i = 0; i < n; ++i)
        if (hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) != hipSuccess) return fail(PLES_MG_ERR_HIP);
    for (int q = 0; q < 2; ++q)
        if (hipEventCreateWithFlags(&b->rev[q], hipEventDisableTiming) != hipSuccess) return fail(PLES_MG_ERR_HIP);
    std::vector<const double*> hp(n);
    for (int i = 0; i < n; ++i) hp[i] = b->hs[i]->S;
    if (hipMalloc((void**)&b->dS, n * sizeof(double*)) != hipSuccess) return fail(PLES_MG_ERR_OOM);
    if (hipMalloc((void**)&b->Sall, (size_t)n * S_COUNT * sizeof(double)) != hipSuccess) return fail(PLES_MG_ERR_OOM);
    if (hipMemcpy(b->dS, hp.data(), n * sizeof(double*), hipMemcpyHostToDevice) != hipSuccess)
        return fail(PLES_MG_ERR_HIP);
    if (hipHostMalloc((void**)&b->ring, 2 * (size_t)n * S_COUNT * sizeof(double), hipHostMallocDefault) != hipSuccess)
        return fail(PLES_MG_ERR_OOM);
    b->bytes += n * sizeof(double*) + (size_t)n * S_COUNT * sizeof(double);
    *out = b;
    return PLES_MG_OK;
}

int ples_mg_batch_destroy(ples_mg_batch* b) {
    if (!b) return PLES_MG_OK;
    if (!b->hs.empty()) (void)hipSetDevice(b->hs[0]->dev);
    if (b->ms) (void)hipStreamSynchronize(b->ms);
    batch_drop_graph(b);
    for (H* h : b->hs) ples_mg_destroy(h);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (hipEvent_t e : b->join)
        if (e) (void)hipEventDestroy(e);
    for (int q = 0; q < 2; ++q)
        if (b->rev[q]) (void)hipEventDestroy(b->rev[q]);
    if (b->dS) (void)hipFree(b->dS);
    if (b->Sall) (void)hipFree(b->Sall);
    if (b->ring) (void)hipHostFree(b->ring);
    if (b->ms) (void)hipStreamDestroy(b->ms);
    delete b;
    return PLES_MG_OK;
}

int ples_mg_batch_size(const ples_mg_batch* b) { return b ? (int)b->hs.size() : 0; }

int ples_mg_batch_set_grouped(ples_mg_batch* b, int on) {
    if (!b || (on != 0 && on != 1)) return PLES_MG_ERR_ARG;
    if (b->grouped != (on == 1)) {
        (void)hipSetDevice(b->dev);
        batch_drop_graph(b);
        b->grouped = on == 1;
    }
    return PLES_MG_OK;
}

int ples_mg_batch_families(const ples_mg_batch* b) {
    if (!b) return PLES_MG_ERR_ARG;
    for (int g = 0; g < 2; ++g)
        if (!b->fam[g].empty()) return (int)b->fam[g].size();
    return 0;
}

ples_mg_handle* ples_mg_batch_handle(ples_mg_batch* b, int i) {
    if (!b || i < 0 || i >= (int)b->hs.size()) return nullptr;
    return b->hs[i];
}

size_t ples_mg_batch_device_bytes(const ples_mg_batch* b) {
    if (!b) return 0;
    size_t t = 0;
    for (const H* h : b->hs) t += h->bytes;
    return t + b->hs.size() * (sizeof(double*) + S_COUNT * sizeof(double));
}

int ples_mg_batch_solve(ples_mg_batch* b, double eps, int maxcycles, int steps_per_launch, ples_mg_result* res) {
    if (!b || !(eps > 0.0) || maxcycles < 0 || steps_per_launch < 0 || steps_per_launch > 16) return PLES_MG_ERR_ARG;
    const int n = (int)b->hs.size();
    CK(hipSetDevice(b->dev));
    const int K = steps_per_launch > 0 ? steps_per_launch : 2;
    if (K != b->K) {
        batch_drop_graph(b);
        b->K = K;
    }
    if (b->ir) return batch_solve_ir(b, eps, maxcycles, K, res);
    b->q.assign(n, 0);
    bool use_graph = true;
    for (H* h : b->hs) use_graph = use_graph && h->o.graphs && h->timing == 0;
    hipEvent_t t0, t1;
    CK(hipEventCreate(&t0));
    CK(hipEventCreate(&t1));
    CK(hipEventRecord(t0, b->ms));
    // initial residuals, every problem on its own stream
    std::vector<double> bnorm(n), thr(n), last(n), rho(n, 0.5);
    std::vector<int> done(n, 0), cycles(n, 0), status(n, PLES_MG_NOT_CONVERGED);
    std::vector<double> rr(n), rmax(n);
    for (int i = 0; i < n; ++i) {
        H* h = b->hs[i];
        CK(hipStreamWaitEvent(h->s, t0, 0));
        // the FP64 compatibility projection of every rhs, as ples_mg_solve_resident
        if (h->o.project_rhs) CKS(project_rhs_async(h, h->cs_sum / h->dt, true));
        CKS(outer_residual(h, h->x, h->rvec));
        CK(hipMemcpyAsync(h->Sh, h->S, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, h->s));
    }
    for (int i = 0; i < n; ++i) {
        H* h = b->hs[i];
        CK(hipStreamSynchronize(h->s));
        if (h->o.project_rhs) h->last_removed = removed_mean(h, h->cs_sum / h->dt, true);
        const bool absmode = h->o.stop_mode == PLES_MG_STOP_ABS_MAX;
        bnorm[i] = std::sqrt(h->Sh[S_BB]) > 0 ? std::sqrt(h->Sh[S_BB]) : 1.0;
        thr[i] = absmode ? eps : (eps * bnorm[i]) * (eps * bnorm[i]);
        rr[i] = h->Sh[S_RR];
        rmax[i] = h->Sh[S_RMAX];
        if (!std::isfinite(rr[i])) {
            status[i] = PLES_MG_ERR_NONFINITE;
            done[i] = 1;
        }
        last[i] = absmode ? rmax[i] : std::sqrt(rr[i]) / bnorm[i];
        CKS(arm_stop(h, h->S, thr[i], h->o.stop_mode == PLES_MG_STOP_ABS_MAX));
        if (!done[i] && last[i] < eps) {
            status[i] = PLES_MG_OK;
            done[i] = 1;
        }
        if (done[i]) {
            h->Sh[S_DONE] = 1.0;
            CK(hipMemcpyAsync(h->S + S_DONE, h->Sh + S_DONE, sizeof(double), hipMemcpyHostToDevice, h->s));
        }
        CK(hipEventRecord(b->join[i], h->s));
        CK(hipStreamWaitEvent(b->ms, b->join[i], 0));
    }
    auto measure = [&](int i, double r2, double rm) {
        return b->hs[i]->o.stop_mode == PLES_MG_STOP_ABS_MAX ? rm : std::sqrt(r2) / bnorm[i];
    };
    auto all_done = [&]() {
        for (int i = 0; i < n; ++i)
            if (!done[i]) return false;
        return true;
    };
    int enq = 0, seen = 0;     // launches (of K steps) enqueued / read back
    int base = 0;              // steps applied before the current arming
    while (!all_done()) {
        while (enq - seen < 2 && base + (enq + 1) * K <= std::max(maxcycles, K)) {
            if (enq > seen) {
                // stop running ahead once every open problem is predicted to
                // converge inside the launch in flight
                bool any_left = false;
                for (int i = 0; i < n; ++i)
                    if (!done[i] && last[i] * std::pow(rho[i], (double)((enq - seen) * K)) >= 4.0 * eps) any_left = true;
                if (!any_left) break;
            }
            CKS(batch_enqueue_chunk(b, use_graph));
            double* slot = b->ring + (size_t)(enq % 2) * n * S_COUNT;
            CK(hipMemcpyAsync(slot, b->Sall, (size_t)n * S_COUNT * sizeof(double), hipMemcpyDeviceToHost, b->ms));
            CK(hipEventRecord(b->rev[enq % 2], b->ms));
            ++enq;
        }
        if (seen == enq) break;  // cycle cap
        CK(hipEventSynchronize(b->rev[seen % 2]));
        const double* slot = b->ring + (size_t)(seen % 2) * n * S_COUNT;
        ++seen;
        for (int i = 0; i < n; ++i) {
            if (done[i]) continue;
            const double* Si = slot + (size_t)i * S_COUNT;
            const double m = measure(i, Si[S_RR], Si[S_RMAX]);
            if (!std::isfinite(Si[S_RR])) {
                status[i] = PLES_MG_ERR_NONFINITE;
                done[i] = 1;
                CKS(batch_set_done(b, i, 1.0));
                continue;
            }
            if (Si[S_DONE] != 0.0) {
                done[i] = 2;  // converged on the device; confirmed below
                cycles[i] = base + (int)Si[S_CYC];
                continue;
            }
            if (last[i] > 0 && m > 0) rho[i] = std::min(0.95, std::max(1e-3, std::pow(m / last[i], 1.0 / K)));
            last[i] = m;
            if (base + (int)Si[S_CYC] >= maxcycles) {
                cycles[i] = base + (int)Si[S_CYC];
                done[i] = 1;  // cap reached: freeze
                CKS(batch_set_done(b, i, 1.0));
            }
        }
        if (!all_done()) continue;
        // every problem finished: confirm the device stop test with true residuals
        CK(hipStreamSynchronize(b->ms));
        for (int i = 0; i < n; ++i) {
            if (done[i] != 2) continue;
            H* h = b->hs[i];
            CKS(outer_residual(h, h->x, h->rvec));
            CK(hipMemcpyAsync(h->Sh, h->S, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, h->s));
        }
        bool restart = false;
        for (int i = 0; i < n; ++i) {
            if (done[i] != 2) continue;
            H* h = b->hs[i];
            CK(hipStreamSynchronize(h->s));
            rr[i] = h->Sh[S_RR];
            rmax[i] = h->Sh[S_RMAX];
            if (measure(i, rr[i], rmax[i]) < eps) {
                status[i] = PLES_MG_OK;
                done[i] = 1;
            } else if (cycles[i] < maxcycles) {
                // the recurrence drifted from the true residual: restart it
                CKS(arm_stop(h, h->S, thr[i], h->o.stop_mode == PLES_MG_STOP_ABS_MAX));
                done[i] = 0;
                last[i] = measure(i, rr[i], rmax[i]);
                restart = true;
            } else {
                done[i] = 1;
            }
        }
        if (restart) {
            // re-armed problems count their steps from zero again
            int mn = maxcycles;
            for (int i = 0; i < n; ++i)
                if (!done[i]) mn = std::min(mn, cycles[i]);
            base = mn;
            enq = seen = 0;
            for (int i = 0; i < n; ++i) {
                CK(hipEventRecord(b->join[i], b->hs[i]->s));
                CK(hipStreamWaitEvent(b->ms, b->join[i], 0));
            }
        }
    }
    CK(hipStreamSynchronize(b->ms));
    // final state: ghosts of every iterate, last residual norms of unconfirmed ones
    for (int i = 0; i < n; ++i) {
        H* h = b->hs[i];
        if (status[i] != PLES_MG_OK && seen > 0) {
            const double* Si = b->ring + (size_t)((seen - 1) % 2) * n * S_COUNT + (size_t)i * S_COUNT;
            rr[i] = Si[S_RR];
            rmax[i] = Si[S_RMAX];
            if (cycles[i] == 0) cycles[i] = base + (int)Si[S_CYC];
        }
        CKS(halo64(h, h->x, true));
        CK(hipEventRecord(b->join[i], h->s));
        CK(hipStreamWaitEvent(b->ms, b->join[i], 0));
    }
    CK(hipEventRecord(t1, b->ms));
    CK(hipEventSynchronize(t1));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    int worst = PLES_MG_OK;
    for (int i = 0; i < n; ++i) {
        if (res) {
            res[i].status = status[i];
            res[i].cycles = cycles[i];
            res[i].rel_res = std::sqrt(rr[i]) / bnorm[i];
            res[i].abs_res = rmax[i];
            res[i].rhs_mean_removed = b->hs[i]->last_removed;
            res[i].seconds = ms * 1e-3;
        }
        if (status[i] != PLES_MG_OK && worst == PLES_MG_OK) worst = status[i];
    }
    return worst;
}


int ples_mg_level_shape(const ples_mg_handle* h, int level, int* nx, int* ny, int* nz) {
    if (!h || level < 1 || level > h->L) return PLES_MG_ERR_ARG;
    const Geo& G = h->lv[level - 1].G;
    if (nx) *nx = G.nx;
    if (ny) *ny = G.ny;
    if (nz) *nz = G.nz;
    return PLES_MG_OK;
}

int ples_mg_tiles(const ples_mg_handle* h, int level, int* P, int* rP, int* rbx) {
    if (!h || level < 1 || level > h->L) return PLES_MG_ERR_ARG;
    const Level& L = h->lv[level - 1];
    const ColLaunch cl = seg_launch(L.G, L.P);
    if (P) *P = L.P;
    if (rP) *rP = L.rP > 0 ? L.rP : L.P;
    if (rbx) *rbx = L.rP > 0 ? L.rbx : (int)cl.block.x;
    return PLES_MG_OK;
}

int ples_mg_cycle_levels(const ples_mg_handle* h, int* levels) {
    if (!h || !levels) return PLES_MG_ERR_ARG;
    *levels = h->Lv;
    return PLES_MG_OK;
}


// ---------------------------------------------------------------- serving loop
}  // extern "C"

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// a handle being created for a cold start on the server's worker thread
struct Creation {
    std::atomic<bool> done{false};
    ples_mg_handle* h = nullptr;
    int status = PLES_MG_OK;
};

struct SCase {
    uint64_t id = 0;
    int jx = 0, jy = 0, jz = 0, ip = 0, kp = 0, nlevel = 1;
    std::shared_ptr<PinnedBuf> x, y, z;  // host copy of the nodes (pinned)
    uint64_t rhs_seed = 0;
    double dt = 1.0;
    H* h = nullptr;               // resident handle (nullptr = cold)
    std::shared_ptr<Creation> creating;  // cold start in flight (its est already counted in used)
    size_t est = 0;               // device bytes when resident
    bool busy = false;
    bool solved = false;          // the resident iterate holds a previous solution
    uint64_t last_use = 0;
    std::vector<double> seed_fi;  // pressure the next admission starts from (host, ples_mg_set_fi layout)
};

enum { RQ_QUEUED = 0, RQ_INIT, RQ_RUN, RQ_CONFIRM, RQ_DONE };

struct SReq {
    int64_t ticket = 0;
    uint64_t case_id = 0;
    uint64_t seed = 0;
    double rel = 0.0, eps = 1e-10;
    int maxcyc = 200;
    int state = RQ_QUEUED;
    bool warm = false;
    double t_submit = 0, t_admit = 0, t_done = 0;
    double bnorm = 1, thr = 0, last = 0, rho = 0.5;
    int enq = 0, seen = 0, cycles = 0, base = 0;
    hipEvent_t ev = nullptr;       // init / confirmation read-back
    ples_mg_result res{};
};

}  // namespace

struct ples_mg_server {
    ples_mg_server_opts so{};
    ples_mg_opts o{};
    std::unordered_map<uint64_t, SCase> cases;
    std::deque<int64_t> queue;                    // waiting tickets, arrival order
    std::unordered_map<int64_t, SReq> reqs;
    std::vector<int64_t> active;
    int64_t next_ticket = 1;
    uint64_t clock = 0;
    size_t used = 0;
    long completed = 0, evictions = 0, cold_starts = 0;
    // worker thread: creates the handles of cold starts and destroys evicted
    // ones, so neither the create (tens of ms) nor the frees of an eviction
    // (hipFree of the large buffers: up to 0.7 ms each) stall the requests
    // the serving loop keeps running
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv;       // jobs queued / stop
    std::condition_variable landed;   // a job finished
    std::deque<std::function<void()>> jobs;   // creations: run first
    std::deque<std::function<void()>> frees;  // destructions of evicted handles
    bool stop = false;
};

namespace {

using SV = ples_mg_server;

ples_mg_grid case_grid(const SCase& c) {
    ples_mg_grid g{};
    g.jx = c.jx; g.jy = c.jy; g.jz = c.jz;
    g.ip = c.ip; g.jp = 1; g.kp = c.kp;
    g.nlevel = c.nlevel;
    g.x = c.x->p; g.y = c.y->p; g.z = c.z->p;
    return g;
}

void post(SV* sv, std::function<void()> job, bool is_free = false) {
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        (is_free ? sv->frees : sv->jobs).push_back(std::move(job));
    }
    sv->cv.notify_one();
}

void worker_loop(SV* sv) {
    for (;;) {
        std::function<void()> job;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->cv.wait(lk, [&] { return sv->stop || !sv->jobs.empty() || !sv->frees.empty(); });
            // a cold start waiting is served before any queued destruction
            std::deque<std::function<void()>>& q = !sv->jobs.empty() ? sv->jobs : sv->frees;
            if (q.empty()) return;  // stop requested and nothing left
            job = std::move(q.front());
            q.pop_front();
        }
        job();
        {  // a waiter checks the jobs' flags under mu: take it once so it cannot miss this notification
            std::lock_guard<std::mutex> lk(sv->mu);
        }
        sv->landed.notify_all();
    }
}

// an idle case's handle leaves: its bytes are free for admission at once,
// the handle is destroyed on the worker thread (its stream is idle)
void evict(SV* sv, SCase& c) {
    if (!c.h) return;
    sv->used -= c.est;
    H* h = c.h;
    const int dev = sv->so.device;
    post(
        sv,
        [h, dev] {
            (void)hipSetDevice(dev);
            ples_mg_destroy(h);
        },
        true);
    c.h = nullptr;
    c.solved = false;
    ++sv->evictions;
}

int resident_count(const SV* sv) {
    int n = 0;
    for (const auto& kv : sv->cases) n += kv.second.h != nullptr || kv.second.creating != nullptr;
    return n;
}

// make room for `need` bytes (and a resident slot) by evicting idle cases,
// least recently used first; false if that cannot be done now
bool make_room(SV* sv, size_t need, bool need_slot) {
    for (;;) {
        const bool mem_ok = sv->used + need <= sv->so.hbm_budget;
        const bool slot_ok = !need_slot || resident_count(sv) < sv->so.warm_cases;
        if (mem_ok && slot_ok) return true;
        SCase* lru = nullptr;
        for (auto& kv : sv->cases) {
            SCase& c = kv.second;
            if (c.h && !c.busy && (!lru || c.last_use < lru->last_use)) lru = &c;
        }
        if (!lru) return false;
        evict(sv, *lru);
    }
}

// admission: the case must be idle, resident or creatable within the budget
int try_admit(SV* sv, SReq& r) {
    SCase& c = sv->cases.at(r.case_id);
    if (c.busy) return 0;
    if ((int)sv->active.size() >= sv->so.max_active) return 0;
    if (!c.h) {
        if (c.creating) {  // cold start in flight on the worker
            if (!c.creating->done.load(std::memory_order_acquire)) return 0;
            const std::shared_ptr<Creation> job = std::move(c.creating);
            if (job->status != PLES_MG_OK) {
                sv->used -= c.est;
                return job->status == PLES_MG_ERR_OOM ? 0 : job->status;  // out of memory: retried later
            }
            c.h = job->h;
            if (c.h->bytes > c.est) {
                sv->used += c.h->bytes - c.est;
                c.est = c.h->bytes;
            }
            ++sv->cold_starts;
        } else {
            // reserve the bytes and the slot, create (and generate the rhs) on the worker
            if (!make_room(sv, c.est, true)) return 0;
            sv->used += c.est;
            auto job = std::make_shared<Creation>();
            c.creating = job;
            ples_mg_opts o = sv->o;
            o.device = sv->so.device;
            const SCase* cp = &c;  // cases are never removed: the node arrays outlive the job
            post(sv, [job, cp, o] {
                (void)hipSetDevice(o.device);
                ples_mg_grid g = case_grid(*cp);
                ples_mg_handle* h = nullptr;
                int st = ples_mg_create(&g, &o, &h);
                if (st == PLES_MG_OK) {
                    st = ples_mg_case_rhs(h, cp->rhs_seed, cp->dt);
                    if (st == PLES_MG_OK && hipStreamSynchronize(h->s) != hipSuccess) st = PLES_MG_ERR_HIP;
                    if (st != PLES_MG_OK) {
                        ples_mg_destroy(h);
                        h = nullptr;
                    }
                }
                job->h = h;
                job->status = st;
                job->done.store(true, std::memory_order_release);
            });
            return 0;
        }
    }
    if (!c.seed_fi.empty()) {  // pressure given by the caller (e.g. read from a restart file)
        CKS(ples_mg_set_fi(c.h, c.seed_fi.data()));
        std::vector<double>().swap(c.seed_fi);
        c.solved = true;
    }
    H* h = c.h;
    c.busy = true;
    c.last_use = ++sv->clock;
    r.warm = c.solved;
    r.t_admit = now_s();
    // warm: perturb the previous solution; cold: start from zero
    if (r.warm) {
        const Level& L0 = h->lv[0];
        const ColLaunch cl = col_launch(L0.G);
        hipLaunchKernelGGL(k_sum, cl.grid, cl.block, 0, h->s, L0.G, (const double*)h->x, h->part, cl.nblocks);
        hipLaunchKernelGGL(k_fin_sum, dim3(1), dim3(256), 0, h->s, (const double*)h->part, cl.nblocks, h->S);
        const ColLaunch ce = cell_launch(L0.G);
        hipLaunchKernelGGL(k_perturb, ce.grid, ce.block, 0, h->s, L0.G, h->x, (const double*)h->S, r.seed, r.rel,
                           (double)L0.ncell());
    } else {
        CK(hipMemsetAsync(h->x, 0, (size_t)h->lv[0].G.n * 8, h->s));
    }
    CK(hipGetLastError());
    CKS(outer_residual(h, h->x, h->rvec));
    CK(hipMemcpyAsync(h->Sh, h->S, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, h->s));
    if (!r.ev) CK(hipEventCreateWithFlags(&r.ev, hipEventDisableTiming));
    CK(hipEventRecord(r.ev, h->s));
    r.state = RQ_INIT;
    sv->active.push_back(r.ticket);
    return 1;
}

void finish(SV* sv, SReq& r, int status, double rr, double rmax) {
    SCase& c = sv->cases.at(r.case_id);
    H* h = c.h;
    (void)halo64(h, h->x, true);
    (void)hipStreamSynchronize(h->s);
    r.t_done = now_s();
    r.state = RQ_DONE;
    r.res.status = status;
    r.res.cycles = r.cycles;
    r.res.rel_res = std::sqrt(rr) / r.bnorm;
    r.res.abs_res = rmax;
    r.res.rhs_mean_removed = h->last_removed;
    r.res.seconds = r.t_done - r.t_submit;
    c.busy = false;
    c.solved = c.solved || status == PLES_MG_OK;
    c.last_use = ++sv->clock;
    if (r.ev) (void)hipEventDestroy(r.ev);
    r.ev = nullptr;
    ++sv->completed;
}

// advance one running request without blocking: enqueue refinement steps
// (graph replays on its handle's stream) up to two in flight, consume the
// read-backs that have landed, confirm convergence with the true residual
int advance(SV* sv, SReq& r) {
    SCase& c = sv->cases.at(r.case_id);
    H* h = c.h;
    const bool absmode = h->o.stop_mode == PLES_MG_STOP_ABS_MAX;
    auto measure = [&](double rr, double rm) { return absmode ? rm : std::sqrt(rr) / r.bnorm; };
    if (r.state == RQ_INIT || r.state == RQ_CONFIRM) {
        const hipError_t q = hipEventQuery(r.ev);
        if (q == hipErrorNotReady) return 0;
        CK(q);
        const double rr = h->Sh[S_RR], rmax = h->Sh[S_RMAX];
        if (r.state == RQ_INIT) {
            r.bnorm = std::sqrt(h->Sh[S_BB]) > 0 ? std::sqrt(h->Sh[S_BB]) : 1.0;
            r.thr = absmode ? r.eps : (r.eps * r.bnorm) * (r.eps * r.bnorm);
            r.last = measure(rr, rmax);
        }
        if (!std::isfinite(rr)) {
            finish(sv, r, PLES_MG_ERR_NONFINITE, rr, rmax);
            return 1;
        }
        if (measure(rr, rmax) < r.eps) {
            finish(sv, r, PLES_MG_OK, rr, rmax);
            return 1;
        }
        if (r.cycles >= r.maxcyc) {
            finish(sv, r, PLES_MG_NOT_CONVERGED, rr, rmax);
            return 1;
        }
        CKS(arm_stop(h, h->S, r.thr, absmode));
        r.base = r.cycles;
        r.enq = r.seen = 0;
        r.last = measure(rr, rmax);
        r.state = RQ_RUN;
    }
    int progressed = 0;
    // consume landed read-backs
    while (r.seen < r.enq) {
        const hipError_t q = hipEventQuery(h->sev[r.seen % 4]);
        if (q == hipErrorNotReady) break;
        CK(q);
        const double* slot = h->Sh + 2 * S_COUNT + (r.seen % 4) * S_COUNT;
        ++r.seen;
        progressed = 1;
        const double rr = slot[S_RR], rmax = slot[S_RMAX];
        if (!std::isfinite(rr)) {
            CK(hipStreamSynchronize(h->s));
            finish(sv, r, PLES_MG_ERR_NONFINITE, rr, rmax);
            return 1;
        }
        r.cycles = r.base + (int)slot[S_CYC];
        if (slot[S_DONE] != 0.0 || r.cycles >= r.maxcyc) {
            // confirm with the true residual after the steps in flight
            CKS(outer_residual(h, h->x, h->rvec));
            CK(hipMemcpyAsync(h->Sh, h->S, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, h->s));
            CK(hipEventRecord(r.ev, h->s));
            r.state = RQ_CONFIRM;
            return 1;
        }
        const double m = measure(rr, rmax);
        if (r.last > 0 && m > 0) r.rho = std::min(0.95, std::max(1e-3, m / r.last));
        r.last = m;
    }
    // keep up to two steps in flight unless the one in flight should converge
    while (r.enq - r.seen < 2 && r.base + r.enq < r.maxcyc) {
        if (r.enq > r.seen && r.last * std::pow(r.rho, (double)(r.enq - r.seen)) < 4.0 * r.eps) break;
        CKS(enqueue_step(h, h->qnext[0]));
        h->qnext[0] ^= 1;
        double* slot = h->Sh + 2 * S_COUNT + (r.enq % 4) * S_COUNT;
        CK(hipMemcpyAsync(slot, h->S, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, h->s));
        CK(hipEventRecord(h->sev[r.enq % 4], h->s));
        ++r.enq;
        progressed = 1;
    }
    return progressed;
}

}  // namespace

extern "C" {

void ples_mg_server_default_opts(ples_mg_server_opts* so) {
    if (!so) return;
    so->hbm_budget = 0;
    so->max_active = 8;
    so->warm_cases = 16;
    so->device = 0;
}

int ples_mg_server_create(const ples_mg_server_opts* so, const ples_mg_opts* opts, ples_mg_server** out) {
    if (!out) return PLES_MG_ERR_ARG;
    *out = nullptr;
    SV* sv = new SV();
    ples_mg_server_default_opts(&sv->so);
    if (so) sv->so = *so;
    ples_mg_default_opts(&sv->o);
    if (opts) sv->o = *opts;
    if (sv->so.max_active < 1 || sv->so.warm_cases < 1) {
        delete sv;
        return PLES_MG_ERR_ARG;
    }
    if (hipSetDevice(sv->so.device) != hipSuccess) {
        delete sv;
        return PLES_MG_ERR_HIP;
    }
    if (sv->so.hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            delete sv;
            return PLES_MG_ERR_HIP;
        }
        sv->so.hbm_budget = (size_t)(0.8 * (double)fr);
    }
    sv->worker = std::thread(worker_loop, sv);
    *out = sv;
    return PLES_MG_OK;
}

int ples_mg_server_destroy(ples_mg_server* sv) {
    if (!sv) return PLES_MG_OK;
    {  // finish the queued creations and destructions, then stop the worker
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv.notify_all();
    if (sv->worker.joinable()) sv->worker.join();
    (void)hipSetDevice(sv->so.device);
    for (auto& kv : sv->cases) {
        if (kv.second.h) ples_mg_destroy(kv.second.h);
        if (kv.second.creating && kv.second.creating->h) ples_mg_destroy(kv.second.creating->h);
    }
    for (auto& kv : sv->reqs)
        if (kv.second.ev) (void)hipEventDestroy(kv.second.ev);
    delete sv;
    return PLES_MG_OK;
}

int ples_mg_server_add_case(ples_mg_server* sv, uint64_t case_id, const ples_mg_grid* g, uint64_t rhs_seed,
                            double dt) {
    if (!sv || !g || !(dt > 0.0) || sv->cases.count(case_id)) return PLES_MG_ERR_ARG;
    CKS(check_grid(g));
    SCase c;
    c.id = case_id;
    c.jx = g->jx; c.jy = g->jy; c.jz = g->jz;
    c.ip = g->ip; c.kp = g->kp;
    c.nlevel = g->nlevel;
    const size_t nn = (size_t)(g->jx + 1) * (g->jy + 1) * (g->jz + 1);
    std::shared_ptr<PinnedBuf>* dst[3] = {&c.x, &c.y, &c.z};
    const double* src[3] = {g->x, g->y, g->z};
    for (int q = 0; q < 3; ++q) {
        *dst[q] = std::make_shared<PinnedBuf>(nn);
        if (!(*dst[q])->p) return PLES_MG_ERR_OOM;
        std::memcpy((*dst[q])->p, src[q], nn * sizeof(double));
    }
    c.rhs_seed = rhs_seed;
    c.dt = dt;
    c.est = estimate(g, &sv->o) + ((size_t)64 << 20);  // + allocator slack and temporaries
    if (c.est > sv->so.hbm_budget) return PLES_MG_ERR_BUSY;  // can never be admitted
    sv->cases.emplace(case_id, std::move(c));
    return PLES_MG_OK;
}

int ples_mg_server_submit(ples_mg_server* sv, uint64_t case_id, uint64_t perturb_seed, double perturb_rel,
                          double eps, int maxcycles, int64_t* ticket) {
    if (!sv || !ticket || !(eps > 0.0) || maxcycles < 1 || !sv->cases.count(case_id)) return PLES_MG_ERR_ARG;
    SReq r;
    r.ticket = sv->next_ticket++;
    r.case_id = case_id;
    r.seed = perturb_seed;
    r.rel = perturb_rel;
    r.eps = eps;
    r.maxcyc = maxcycles;
    r.t_submit = now_s();
    sv->reqs.emplace(r.ticket, r);
    sv->queue.push_back(r.ticket);
    *ticket = r.ticket;
    return PLES_MG_OK;
}

int ples_mg_server_step(ples_mg_server* sv, int* n_done) {
    if (!sv) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(sv->so.device));
    int done = 0;
    // admit in arrival order; a request whose case is busy waits without
    // blocking the ones behind it, except later requests of the same case: a
    // case's time steps run in the order they were submitted (a cold start
    // landing on the worker during this scan must not let a later step of
    // the case start before an earlier one that was passed over)
    std::vector<uint64_t> passed;
    for (auto it = sv->queue.begin(); it != sv->queue.end();) {
        if ((int)sv->active.size() >= sv->so.max_active) break;
        SReq& r = sv->reqs.at(*it);
        if (std::find(passed.begin(), passed.end(), r.case_id) != passed.end()) {
            ++it;
            continue;
        }
        const int a = try_admit(sv, r);
        if (a < 0 || a > 1) return a;  // error status
        if (a == 1) {
            it = sv->queue.erase(it);
        } else {
            passed.push_back(r.case_id);
            ++it;
        }
    }
    for (size_t k = 0; k < sv->active.size();) {
        SReq& r = sv->reqs.at(sv->active[k]);
        const int st = advance(sv, r);
        if (st < 0 || st > 1) return st;
        if (r.state == RQ_DONE) {
            ++done;
            sv->active.erase(sv->active.begin() + k);
        } else {
            ++k;
        }
    }
    if (n_done) *n_done = done;
    return PLES_MG_OK;
}

int ples_mg_server_wait(ples_mg_server* sv) {
    if (!sv) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(sv->so.device));
    // block on the oldest outstanding read-back of any running request, else
    // on the next cold start the worker finishes
    for (int64_t t : sv->active) {
        SReq& r = sv->reqs.at(t);
        H* h = sv->cases.at(r.case_id).h;
        if ((r.state == RQ_INIT || r.state == RQ_CONFIRM) && r.ev) {
            CK(hipEventSynchronize(r.ev));
            return PLES_MG_OK;
        }
        if (r.state == RQ_RUN && r.seen < r.enq) {
            CK(hipEventSynchronize(h->sev[r.seen % 4]));
            return PLES_MG_OK;
        }
    }
    auto creation_landed = [&] {
        bool any = false;
        for (const auto& kv : sv->cases)
            if (kv.second.creating) {
                any = true;
                if (kv.second.creating->done.load(std::memory_order_acquire)) return true;
            }
        return !any;
    };
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->landed.wait(lk, creation_landed);
    return PLES_MG_OK;
}

int ples_mg_server_poll(ples_mg_server* sv, int64_t ticket, ples_mg_result* res, int* warm) {
    if (!sv || !sv->reqs.count(ticket)) return PLES_MG_ERR_ARG;
    const SReq& r = sv->reqs.at(ticket);
    if (r.state != RQ_DONE) return PLES_MG_ERR_BUSY;
    if (res) *res = r.res;
    if (warm) *warm = r.warm ? 1 : 0;
    return PLES_MG_OK;
}

int ples_mg_server_pending(const ples_mg_server* sv) {
    return sv ? (int)(sv->queue.size() + sv->active.size()) : 0;
}

int ples_mg_server_get_fi(ples_mg_server* sv, uint64_t case_id, double* fi) {
    if (!sv || !fi || !sv->cases.count(case_id)) return PLES_MG_ERR_ARG;
    SCase& c = sv->cases.at(case_id);
    if (!c.h || c.busy) return PLES_MG_ERR_BUSY;
    return ples_mg_get_fi(c.h, fi);
}

int ples_mg_server_set_fi(ples_mg_server* sv, uint64_t case_id, const double* fi) {
    if (!sv || !fi || !sv->cases.count(case_id)) return PLES_MG_ERR_ARG;
    SCase& c = sv->cases.at(case_id);
    if (c.busy) return PLES_MG_ERR_BUSY;
    if (c.h) {
        CKS(ples_mg_set_fi(c.h, fi));
        c.solved = true;
        return PLES_MG_OK;
    }
    const size_t n = (size_t)(c.jx + 2) * (c.jy + 2) * (c.jz + 2);
    c.seed_fi.assign(fi, fi + n);
    return PLES_MG_OK;
}

int ples_mg_server_stats(const ples_mg_server* sv, size_t* used_bytes, int* resident, long* completed,
                         long* evictions, long* cold_starts) {
    if (!sv) return PLES_MG_ERR_ARG;
    if (used_bytes) *used_bytes = sv->used;
    if (resident) *resident = resident_count(sv);
    if (completed) *completed = sv->completed;
    if (evictions) *evictions = sv->evictions;
    if (cold_starts) *cold_starts = sv->cold_starts;
    return PLES_MG_OK;
}


// Copy bandwidth of this device: a 16-byte-per-lane copy of `bytes` bytes,
// reps launches timed with HIP events.  The measured streaming ceiling that
// the stencil kernels' achieved bandwidth is read against.
// copy (read + write bytes) or read-only probe of `bytes` bytes, reps launches
// twelve-stream probe: bytes = read + written bytes of one launch
static int probe_streams12(int device, size_t bytes, int reps, double* gbs) {
    if (!gbs || bytes < 48 * 4096 || reps < 1) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(device));
    const size_t n = bytes / 48;
    float* buf = nullptr;
    CK(hipMalloc((void**)&buf, 12 * n * sizeof(float)));
    Streams11 S;
    for (int q = 0; q < 11; ++q) S.a[q] = buf + (size_t)q * n;
    float* out = buf + 11 * n;
    hipStream_t s;
    hipEvent_t e0, e1;
    int st = PLES_MG_OK;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) st = PLES_MG_ERR_HIP;
    if (st == PLES_MG_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        st = PLES_MG_ERR_HIP;
    if (st == PLES_MG_OK) {
        (void)hipMemsetAsync(buf, 0, 12 * n * sizeof(float), s);
        const int nb = 8192, bt = 256;
        auto launch = [&]() { hipLaunchKernelGGL(k_streams12, dim3(nb), dim3(bt), 0, s, (long)n, S, out); };
        launch();
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps; ++r) launch();
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess) st = PLES_MG_ERR_HIP;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        *gbs = 48.0 * (double)n * reps / (ms * 1e-3) / 1e9;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
    }
    (void)hipFree(buf);
    return st;
}

static int probe_bandwidth(int device, size_t bytes, int reps, bool read_only, double* gbs) {
    if (!gbs || bytes < 4096 || reps < 1) return PLES_MG_ERR_ARG;
    CK(hipSetDevice(device));
    const size_t n4 = bytes / 16;
    const int nb = 8192, bt = 256;
    float4 *a = nullptr, *b = nullptr;
    float* sink = nullptr;
    CK(hipMalloc((void**)&a, n4 * 16));
    const size_t second = read_only ? (size_t)nb * bt * sizeof(float) : n4 * 16;
    if (hipMalloc(read_only ? (void**)&sink : (void**)&b, second) != hipSuccess) {
        (void)hipFree(a);
        return PLES_MG_ERR_OOM;
    }
    hipStream_t s;
    hipEvent_t e0, e1;
    int st = PLES_MG_OK;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) st = PLES_MG_ERR_HIP;
    if (st == PLES_MG_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        st = PLES_MG_ERR_HIP;
    if (st == PLES_MG_OK) {
        (void)hipMemsetAsync(a, 0, n4 * 16, s);
        auto launch = [&]() {
            if (read_only)
                hipLaunchKernelGGL(k_read16, dim3(nb), dim3(bt), 0, s, (long)n4, (const float4*)a, sink);
            else
                hipLaunchKernelGGL(k_copy16, dim3(nb), dim3(bt), 0, s, (long)n4, (const float4*)a, b);
        };
        launch();
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps; ++r) launch();
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess) st = PLES_MG_ERR_HIP;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        *gbs = (read_only ? 1.0 : 2.0) * (double)(n4 * 16) * reps / (ms * 1e-3) / 1e9;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
    }
    (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (sink) (void)hipFree(sink);
    return st;
}

}  // extern "C"

// ---------------------------------------------------------------- no-allocation projection
// The compatibility projection of a device-resident rhs for callers that
// cannot allocate (graph capture, tight serving loops): the block partial
// sums and the removed value live in this file-scope device array, one copy
// per GPU, shared by every call on that GPU (calls that overlap on one GPU
// race on it).  Host code never passes the array's name: it takes the
// device address from hipGetSymbolAddress once per device.
namespace {
constexpr int kNaBlocks = 1024;  // reduction grid: one partial per block
constexpr int kNaThreads = 256;
// 32-bit cell indices in the kernels (a grid-stride index stays below 2^31)
constexpr long kNaMaxCells = 0x7fffffffL - (long)kNaBlocks * kNaThreads;
static __device__ __align__(256) double scratch[kNaBlocks + 8];

std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_ptr;

// device address of `scratch` on the current device (looked up once per device)
int scratch_address(int dev, double** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) {
        *out = it->second;
        return PLES_MG_OK;
    }
    double* p = nullptr;
    if (hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
        (void)hipGetLastError();
        return PLES_MG_ERR_HIP;
    }
    g_scratch_ptr[dev] = p;
    *out = p;
    return PLES_MG_OK;
}

struct NaIn {
    const double* a[7];  // rhs, cs1..cs6 (NULL: absent)
    int n[7];            // their lengths
    double w[7];         // weights: 1 for the rhs, -/+ 1/dt for the inflow/outflow fluxes
};

// part[b] = sum over the arrays of w * (the elements block b visits), fixed order
__global__ void __launch_bounds__(kNaThreads) k_na_sum(NaIn in, double* __restrict__ part) {
    double s = 0.0;
    const int stride = gridDim.x * blockDim.x;
    for (int q = 0; q < 7; ++q) {
        if (!in.a[q]) continue;
        double t = 0.0;
        for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < in.n[q]; c += stride) t += in.a[q][c];
        s += in.w[q] * t;
    }
    __shared__ double red[kNaThreads / 64];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = 0.0;
        for (int w = 0; w < kNaThreads / 64; ++w) b += red[w];
        part[blockIdx.x] = b;
    }
}

// mean = sum(part) / ncell into part[kNaBlocks] (and *removed when given)
__global__ void __launch_bounds__(kNaThreads) k_na_fin(double* __restrict__ part, int nb, double inv_n,
                                                       double* __restrict__ removed) {
    double s = 0.0;
    for (int q = threadIdx.x; q < nb; q += blockDim.x) s += part[q];
    __shared__ double red[kNaThreads / 64];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = 0.0;
        for (int w = 0; w < kNaThreads / 64; ++w) b += red[w];
        part[kNaBlocks] = b * inv_n;
        if (removed) *removed = b * inv_n;
    }
}

__global__ void __launch_bounds__(kNaThreads) k_na_sub(int n, double* __restrict__ a, const double* __restrict__ mean) {
    const double m = mean[0];
    const int stride = gridDim.x * blockDim.x;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += stride) a[c] -= m;
}
}  // namespace

extern "C" {

int ples_mg_project_rhs_noalloc(double* rhs, int jx, int jy, int jz, const double* cs1, const double* cs2,
                                const double* cs3, const double* cs4, const double* cs5, const double* cs6,
                                double dt, double* removed, void* stream) {
    if (!rhs || jx < 1 || jy < 1 || jz < 1 || !(dt > 0.0)) return PLES_MG_ERR_ARG;
    const long n = (long)jx * jy * jz;
    if (n > kNaMaxCells || (long)jy * jz > kNaMaxCells || (long)jx * jz > kNaMaxCells ||
        (long)jx * jy > kNaMaxCells)
        return PLES_MG_ERR_ARG;
    int dev = 0;
    CK(hipGetDevice(&dev));
    double* part = nullptr;
    CKS(scratch_address(dev, &part));
    const hipStream_t s = (hipStream_t)stream;
    NaIn in{};
    const double* cs[6] = {cs1, cs2, cs3, cs4, cs5, cs6};
    const int len[6] = {jy * jz, jy * jz, jx * jz, jx * jz, jx * jy, jx * jy};
    in.a[0] = rhs;
    in.n[0] = (int)n;
    in.w[0] = 1.0;
    for (int q = 0; q < 6; ++q) {  // sum(rhs) - sum(outflow - inflow) / dt
        in.a[q + 1] = cs[q];
        in.n[q + 1] = len[q];
        in.w[q + 1] = (q & 1) ? -1.0 / dt : 1.0 / dt;
    }
    const int nb = (int)std::min<long>(kNaBlocks, (n + 8L * kNaThreads - 1) / (8L * kNaThreads));
    hipLaunchKernelGGL(k_na_sum, dim3(nb), dim3(kNaThreads), 0, s, in, part);
    hipLaunchKernelGGL(k_na_fin, dim3(1), dim3(kNaThreads), 0, s, part, nb, 1.0 / (double)n, removed);
    hipLaunchKernelGGL(k_na_sub, dim3(nb), dim3(kNaThreads), 0, s, (int)n, rhs, (const double*)(part + kNaBlocks));
    CK(hipGetLastError());
    return PLES_MG_OK;
}

int ples_mg_copy_bandwidth(int device, size_t bytes, int reps, double* gbs) {
    return probe_bandwidth(device, bytes, reps, false, gbs);
}

int ples_mg_streams_bandwidth(int device, size_t bytes, int reps, double* gbs) {
    return probe_streams12(device, bytes, reps, gbs);
}

int ples_mg_read_bandwidth(int device, size_t bytes, int reps, double* gbs) {
    return probe_bandwidth(device, bytes, reps, true, gbs);
}


// ---- tuning table: one line per level shape
//   nx ny nz ip kp f32 fine half P rP rbx use16
int ples_mg_tuning_load(const char* path) {
    if (!path) return PLES_MG_ERR_ARG;
    FILE* f = std::fopen(path, "r");
    if (!f) return PLES_MG_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_tune_mu);
    char line[256];
    int n = 0;
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        TuneKey k{};
        TuneVal v{};
        // nx ny nz ip kp f32 fine half  P rP rbx use16; lines of ten fields
        // (no half column, no use16) are read as half = use16 = 0
        const int got = std::sscanf(line, "%d %d %d %d %d %d %d %d %d %d %d %d", &k[0], &k[1], &k[2], &k[3], &k[4],
                                    &k[5], &k[6], &k[7], &v.P, &v.rP, &v.rbx, &v.use16);
        if (got == 10) {
            k[7] = 0;
            v.use16 = 0;
            if (std::sscanf(line, "%*d %*d %*d %*d %*d %*d %*d %d %d %d", &v.P, &v.rP, &v.rbx) != 3) continue;
        } else if (got != 12) {
            continue;
        }
        if ((v.use16 != 0 && v.use16 != 1) || (k[7] != 0 && k[7] != 1) || (v.use16 && !k[7])) continue;
        const bool okP = seg_ok(k[1], v.P);
        const bool okR = seg_ok(k[1], v.rP) && v.rbx >= 1 && (v.rbx & (v.rbx - 1)) == 0 && v.rbx <= 256 &&
                         v.rbx * v.rP <= 512;
        if (!okP || !okR) continue;
        g_tune[k] = v;
        ++n;
    }
    std::fclose(f);
    return n > 0 ? PLES_MG_OK : PLES_MG_ERR_ARG;
}

int ples_mg_tuning_clear(void) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune.clear();
    return PLES_MG_OK;
}

int ples_mg_tuning_save(const char* path) {
    if (!path) return PLES_MG_ERR_ARG;
    FILE* f = std::fopen(path, "w");
    if (!f) return PLES_MG_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_tune_mu);
    std::fprintf(f, "# nx ny nz ip kp f32 fine half  P rP rbx use16  (launch shapes chosen by the autotuner)\n");
    for (const auto& kv : g_tune) {
        const auto& k = kv.first;
        std::fprintf(f, "%d %d %d %d %d %d %d %d %d %d %d %d\n", k[0], k[1], k[2], k[3], k[4], k[5], k[6], k[7],
                     kv.second.P, kv.second.rP, kv.second.rbx, kv.second.use16);
    }
    std::fclose(f);
    return PLES_MG_OK;
}

}  // extern "C"
