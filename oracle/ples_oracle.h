/*
 * ples_oracle.h -- CPU FP64 restatement of the PLES pressure-Poisson path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared
 * against; it is never linked into, loaded by, or called from the product
 * library (ples_amd/).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.
 *
 * Every function cites the reference routine it restates (paths relative to
 * the PLES source tree).  The reference runs on MPI k-slabs; this
 * restatement holds the whole domain in one array, so every MPI halo,
 * periodic and edge-strip exchange of the reference becomes a plain copy
 * between planes of that array ("global" semantics: what the reference
 * computes with >= 2 ranks, where both periodic partners exist).
 *
 * Layout conventions (shared with the HIP library):
 *   - cell arrays are "ghosted": index (i,j,k) with i in 0..nx+1,
 *     j in 0..ny+1, k in 0..nz+1, i fastest:
 *         GIDX(i,j,k) = i + (nx+2)*(j + (ny+2)*k)
 *     rhs / residual arrays use the same layout (only the interior is read).
 *   - face metrics use the same ghosted layout: the x-face i (0..nx) of cell
 *     row (j,k) is stored at GIDX(i,j,k); y-face j at GIDX(i,j,k);
 *     z-face k at GIDX(i,j,k).  Order of the 9 metric pointers:
 *         g11 g12 g13 | g21 g22 g23 | g31 g32 g33
 *   - node coordinates are (nx+1)(ny+1)(nz+1), i fastest:
 *         NIDX(i,j,k) = i + (nx+1)*(j + (ny+1)*k)
 *   - boundary flux arrays follow the reference shapes
 *     (fsrc/iniz.f03:63-68): cs1,cs2 (ny,nz); cs3,cs4 (nx,nz); cs5,cs6 (nx,ny),
 *     first index fastest.
 *   - periodic flags follow the reference (Aboundary.in): ip/jp/kp = 1 means
 *     walls (non-periodic), 0 means periodic.
 *
 * What pins each part (tests/test_oracle*.py):
 *   - metrics: exact on affine node maps at every level (constant metric
 *     tensor with cross terms), and on node maps quadratic in each index with
 *     walls in every direction at every level, where the central and the
 *     second-order one-sided wall derivatives are exact at the face centres
 *     (a first-order wall derivative fails it);
 *   - operator: exact on quadratic pressures, annihilates affine ones;
 *   - wall ghosts: an affine pressure with its exact wall flux is reproduced;
 *   - edge/corner ghosts: a field that satisfies every boundary condition is
 *     kept, and fields constant along chosen directions fix the direction of
 *     every copy;
 *   - transfers: restriction of constants and the sum rule, prolongation of
 *     constants and linear fields;
 *   - tridag against a banded solve; sett's rows against the operator on a
 *     column; SLOR: the first relaxed column solves its equations, the exact
 *     solution is a fixed point on warped walled grids with fluxes;
 *   - orc_multi: the exact FFT/tridiagonal channel solution and a dense solve
 *     on a warped grid; the SLOR V-cycle reaches the SOR V-cycle's solution;
 *   - case rhs = divergence of the face fluxes; gradie exact on an affine
 *     pressure; projection-step routines (tests/test_oracle_projection.py).
 * Parity unpinned: the per-cycle residual trajectory of the reference's
 * smoothers (only a run of the reference itself would fix it).
 */
#ifndef PLES_ORACLE_H
#define PLES_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_dims {
    int nx, ny, nz; /* cells on this level */
    int ip, jp, kp; /* 1 = wall, 0 = periodic */
} orc_dims;

/* ---- case generation (test inputs) ---------------------------------- */

/* Node coordinates of the channel / terrain-warped grids:
 *   x_i = i*Lx/nx, z_k = k*Lz/nz, s(eta) = 0.5*(1 + tanh(gamma*(2 eta-1))/tanh gamma)
 *   y = zb(x,z) + (Ly - zb(x,z)) * s(eta_j),
 *   zb = amp * sin(2 pi x/(Lx/4)) * cos(2 pi z/(Lz/4))   (amp = 0 -> channel)
 * gamma <= 0 gives a uniform y spacing. */
void orc_make_nodes(int nx, int ny, int nz, double Lx, double Ly, double Lz,
                    double gamma, double amp, double* x, double* y, double* z);

/* Counter-based N(0,1) sample (SplitMix64 -> Box-Muller, cosine branch). */
double orc_gauss(uint64_t seed, uint32_t field, int64_t ctr);

/* Cell-centred velocity u,v,w (ghosted layout) ~ N(0,1) per cell, periodic
 * ghost layers copied from the opposite interior layer. */
void orc_make_velocity(const orc_dims* d, uint64_t seed, double* u, double* v, double* w);

/* ---- geometry: metrica / mul_met / gprima family -------------------- */

/* Node coordinate with the periodic extension of cellep
 * (fsrc/cellep.f03:29-98): out-of-range indices in periodic directions
 * wrap and shift the coordinate along that direction by the period. */
void orc_node(const orc_dims* fine, const double* x, const double* y, const double* z,
              int i, int j, int k, double out[3]);

/* Face metrics J^-1 g^{ij} of the level with node stride st (1 = finest),
 * divided by pot = st, as metrica (fsrc/metrica.f03:85-419, st = 1) and
 * mul_met (fsrc/multigrid.f03:2520-2988, st = 2^(n-1)).  `fine` holds the
 * finest-level dims; the level dims are fine/st.  g[9] are outputs in the
 * ghosted layout of the level.  If csv != NULL it receives the 9 face area
 * vectors csx,csy,csz | etx,ety,etz | ztx,zty,ztz (st must be 1); if
 * giac != NULL it receives the centroid Jacobian (fsrc/metrica.f03:432-483). */
void orc_metrics(const orc_dims* fine, const double* x, const double* y, const double* z,
                 int st, double* const g[9], double* const csv[9], double* giac);

/* ---- multigrid operators ------------------------------------------- */

/* Ghost fill of mul_boun_sndrcv (fsrc/multigrid.f03:1109-1933): npass
 * passes (the reference uses 5) of face ghosts (walls: face flux = an*cs/dt,
 * periodic: copies) followed by the edge rules.  cs may be NULL (all zero)
 * and is only read when an != 0. */
void orc_ghost_fill(const orc_dims* d, double* p, const double* const g[9],
                    const double* const cs[6], double an, double dt, int npass);

/* One lexicographic in-place SOR sweep (fsrc/multigrid.f03:784-919) incl.
 * the per-row periodic ghost refresh; pot = 2^(n-1). */
void orc_sor_sweep(const orc_dims* d, double* p, const double* rh,
                   const double* const g[9], double pot, double omega);

/* Residual r = s^2 * sum(face flux differences) - rh, s = 1/pot
 * (fsrc/resid_sndrcv.f03:80-151).  r in ghosted layout, interior written. */
void orc_resid(const orc_dims* d, const double* p, const double* rh,
               const double* const g[9], double pot, double* r);

/* Restriction rc = -(1/8) * sum of the 8 children (fsrc/restrict_sndrcv.f03:32-53). */
void orc_restrict(const orc_dims* dc, const double* rf, double* rc);

/* Trilinear cell-centred prolongation-and-add over the ghost-inclusive
 * range (fsrc/prolong.f03:64-261). */
void orc_prolong(const orc_dims* dc, const double* ec, double* pf);

/* The reference V-cycle driver multi (fsrc/multigrid.f03:163-637) with L
 * levels (the reference caps L at 4).  fi (ghosted, in/out) is the warm
 * start and result; rhs ghosted; cs may be NULL.  g1 may be NULL, then the
 * level-1 metrics are computed from the nodes.  kss1/kssc are the sweeps
 * per visit (reference: 5 / 10), nbc the ghost passes per sweep (5).
 * Stops when max|r| < eps or after ficycle cycles.  res_hist (may be NULL)
 * receives max|r| after each cycle (up to ficycle entries).
 * Returns the number of cycles. */
int orc_multi(const orc_dims* d, int nlevel, const double* x, const double* y,
              const double* z, const double* const g1[9], double* fi,
              const double* rhs, const double* const cs[6], double dt, double omega,
              double eps, int ficycle, int kss1, int kssc, int nbc,
              double* resmax_out, double* res_hist);

/* The same with the smoother of the reference's islor switch
 * (fsrc/multigrid.f03:223-563): islor = 0 the lexicographic point SOR,
 * islor = 1 the eta-line SOR (orc_slor_sweep). */
int orc_multi_ex(const orc_dims* d, int nlevel, const double* x, const double* y,
                 const double* z, const double* const g1[9], double* fi,
                 const double* rhs, const double* const cs[6], double dt, double omega,
                 double eps, int ficycle, int kss1, int kssc, int nbc, int islor,
                 double* resmax_out, double* res_hist);

/* Level hierarchy of multi built once (metrics of every level, as the
 * reference's init does with metrica / mul_met, and the work arrays), then
 * any number of solves.  nslab > 1 runs every smoother sweep on nslab
 * k-slabs with the reference's MPI slab semantics (each slab relaxes its own
 * planes with its halo planes as they were at the start of the sweep,
 * fsrc/multigrid.f03:929-966), one OpenMP thread per slab when built with
 * OpenMP; requires nz % (nslab * 2^(nlevel-1)) == 0 (else returns -1).
 * nslab <= 1 is orc_multi_ex. */
typedef struct orc_hier orc_hier;
orc_hier* orc_hier_create(const orc_dims* d, int nlevel, const double* x, const double* y,
                          const double* z, const double* const g1[9]);
void orc_hier_destroy(orc_hier* H);
int orc_hier_multi(orc_hier* H, double* fi, const double* rhs, const double* const cs[6],
                   double dt, double omega, double eps, int ficycle, int kss1, int kssc,
                   int nbc, int islor, int nslab, double* resmax_out, double* res_hist);

/* ---- line relaxation (islor = 1) ------------------------------------ */

/* Thomas algorithm of tridag (fsrc/tridag.f03:11-30) on rows 0..n-1:
 * aa[r] x[r-1] + bb[r] x[r] + cc[r] x[r+1] = ff[r]; aa, bb, cc are
 * overwritten as the reference does, the solution replaces ff. */
void orc_tridag(int n, double* aa, double* bb, double* cc, double* ff);

/* Line coefficients of sett (fsrc/sett.f03:119-143) of a level with 1/s =
 * pot, ghosted layout, j = 0..ny+1 of every column (i,k): interior rows
 * aa = s^2 g22(j-1), bb = -s^2 (g22(j) + g22(j-1) + the xi/zeta normal
 * metrics of the non-wall faces), cc = s^2 g22(j); the wall rows
 * 0 and ny+1 are the flux relations of the eta-wall ghosts. */
void orc_sett(const orc_dims* d, const double* const g[9], double pot, double* aa,
              double* bb, double* cc);

/* One sweep of solut_sndrcv_slor (fsrc/multigrid.f03:2050-2199): the ghost
 * fill (nbc passes, an, cs, dt as in orc_ghost_fill), the periodic-k plane
 * exchange, then for every plane k and column i (in that order) the
 * tridiagonal eta-line system of rows 0..ny+1 with the explicit terms at
 * their current values, solved by orc_tridag, and p += omega (line solution
 * - p) on the whole column including its wall ghosts. */
void orc_slor_sweep(const orc_dims* d, double* p, const double* rh, const double* const g[9],
                    const double* const cs[6], double an, double dt, double pot, double omega,
                    int nbc);

/* Edge and corner ghosts only (the last step of every orc_ghost_fill pass,
 * fsrc/multigrid.f03:1358-1742): each is a copy of its periodic image or of
 * its neighbour towards the interior across a wall. */
void orc_edge_fill(const orc_dims* d, double* p);

/* ---- projection step ------------------------------------------------ */

/* Contravariant face fluxes of contra (fsrc/contra.f03:73-87,132-146,204-220)
 * from cell velocities, with wall-face fluxes set to zero.  uc/vc/wc are in
 * the ghosted face layout (x-face i at GIDX(i,j,k), ...). */
void orc_contra(const orc_dims* d, const double* const csv[9], const double* u,
                const double* v, const double* w, double* uc, double* vc, double* wc);

/* rhs = (div of face fluxes)/dt (fsrc/diver.f03:20-28). */
void orc_diver(const orc_dims* d, const double* uc, const double* vc, const double* wc,
               double dt, double* rhs);

/* Compatibility projection: rhs -= (sum(rhs) - boundary flux sum/dt)/N.
 * Returns the removed mean.  (The reference only prints the integral,
 * fsrc/stratiParticle.f03:2847-2866.) */
double orc_project(const orc_dims* d, double* rhs, const double* const cs[6], double dt);

/* Centroid and contravariant pressure gradients of gradie
 * (fsrc/gradie.f03:28-117).  gra1..3 ghosted cell layout (interior);
 * cgra1..3 ghosted face layout (x-faces i=0..nx, y-faces j=0..ny,
 * z-faces k=0..nz). */
void orc_gradie(const orc_dims* d, const double* fi, const double* const csv[9],
                const double* giac, const double* const g[9], double dt, double* gra1,
                double* gra2, double* gra3, double* cgra1, double* cgra2, double* cgra3);

/* ---- time-loop neighbours of the solve ------------------------------ */

/* update (fsrc/update.f03:29-235): vel[c] += del[c] in the interior, then
 * the ghost faces (not edges) of each direction: parabolic extrapolation
 * 1.875/-1.25/0.375 from the first three cells at walls, periodic copies
 * otherwise.  vel, del ghosted cell layout. */
void orc_update(const orc_dims* d, double* const vel[3], const double* const del[3]);

/* contra (fsrc/contra.f03:35-221): face fluxes fc = (uc, vc, wc) (ghosted
 * face layout, in/out).  Wall faces take the ghost-cell velocity dotted
 * with the face vector and return cs = new - old wall flux (potenziale = 1:
 * cs = -old) in cs1..cs6 (reference shapes; only wall directions written);
 * other faces take the average of the two cells. */
void orc_contra_step(const orc_dims* d, const double* const csv[9], const double* const vel[3],
                     double* const fc[3], double* const cs[6], int potenziale);

/* vel_up (fsrc/vel_up.f03:39-289, freesurface = 0): vel += gra (= gra if
 * potenziale), periodic xi / zeta ghost copies, fc -= dt*cgra on the
 * non-wall faces (fc = -dt*cgra if potenziale). */
void orc_vel_up(const orc_dims* d, double* const vel[3], double* const fc[3], const double* const gra[3],
                const double* const cgra[3], double dt, int potenziale);

/* check_divergence (fsrc/check_divergence.f03:131-190, bodyforce = 0):
 * out = {max|div| inside, max|div/giac| inside, max|div| all, max|div/giac| all}. */
void orc_check_divergence(const orc_dims* d, const double* const fc[3], const double* giac, double out[4]);

/* ---- norms ------------------------------------------------------------ */
double orc_absmax(const orc_dims* d, const double* r);
double orc_l2(const orc_dims* d, const double* r);
double orc_mean(const orc_dims* d, const double* p);

#ifdef __cplusplus
}
#endif
#endif
