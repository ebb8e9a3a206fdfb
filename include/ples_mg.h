/*
 * ples_mg.h -- C ABI of the MI355X pressure-Poisson multigrid for PLES.
 *
 * Drop-in replacement for the pressure solve of the PLES projection step:
 *     call multi(eps, ficycle, nlevel, jxc, jyc, jzc, ...)
 * (fsrc/stratiParticle.f03:2881-2887, signature fsrc/multigrid.f03:42-48),
 * together with the metric construction it depends on (metrica,
 * fsrc/metrica.f03:2-501; mul_met, fsrc/multigrid.f03:2293-3052) and the
 * projection-step kernels around it (diver, fsrc/diver.f03:2-46; gradie,
 * fsrc/gradie.f03:2-120).  All entry points are plain C: pointers, sizes and
 * status codes; no exceptions cross the boundary.
 *
 * Array conventions (host side, as the Fortran caller holds them, first
 * index fastest, single MPI rank = whole domain):
 *   fi   (0:jx+1, 0:jy+1, 1-deepl:jz+deepr)   pressure incl. ghosts (in/out)
 *   rhs  (1:jx, 1:jy, 1:jz)
 *   cs1,cs2 (jy,jz); cs3,cs4 (jx,jz); cs5,cs6 (jx,jy)   boundary fluxes
 *   g11,g12,g13 (0:jx, 1:jy, 1:jz)   xi-face metrics   (myarrays_metri3)
 *   g21,g22,g23 (1:jx, 0:jy, 1:jz)   eta-face metrics
 *   g31,g32,g33 (1:jx, 1:jy, 0:jz)   zeta-face metrics
 *   x,y,z (0:jx, 0:jy, 0:jz)          grid nodes (the interior part of the
 *                                      reference's (-8:n1+8,...) arrays)
 * Periodicity flags follow Aboundary.in: ip/jp/kp = 1 walls, 0 periodic.
 * Walls in eta (jp = 1) are required, as in the reference
 * (fsrc/multigrid.f03:1744-1746).
 */
#ifndef PLES_MG_H
#define PLES_MG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference has none: fsrc/multigrid.f03:174) ---- */
#define PLES_MG_OK 0
#define PLES_MG_NOT_CONVERGED 1 /* iteration cap reached before the tolerance */
#define PLES_MG_ERR_NONFINITE 2 /* NaN/Inf in the input or the iteration */
#define PLES_MG_ERR_OOM 3       /* device allocation failed / over budget */
#define PLES_MG_ERR_BAD_DIMS 4  /* dims not divisible by 2^(nlevel-1), too few cells */
#define PLES_MG_ERR_HIP 5       /* HIP runtime error */
#define PLES_MG_ERR_ARG 6       /* NULL or out-of-range argument */
#define PLES_MG_ERR_UNSUPPORTED 7
#define PLES_MG_ERR_BUSY 8      /* serving: request not admitted */

/* ---- options ---- */
#define PLES_MG_FP64 0  /* FP64 V-cycles */
#define PLES_MG_MIXED 1 /* FP32 V-cycles inside FP64 refinement */

#define PLES_MG_STOP_REL_L2 0  /* ||b - A p||_2 / ||b||_2 < eps */
#define PLES_MG_STOP_ABS_MAX 1 /* max|b - A p| < eps  (the reference's test,
                                  fsrc/multigrid.f03:598-631) */

typedef struct ples_mg_grid {
    int jx, jy, jz;        /* fine-level cells (reference n1, n2, n3) */
    int ip, jp, kp;        /* 1 = walls, 0 = periodic */
    int nlevel;            /* multigrid levels, >= 1 (the reference caps it at 4) */
    const double* x;       /* nodes, required */
    const double* y;
    const double* z;
    const double* g[9];    /* optional level-1 metrics g11..g33 in the reference
                              shapes; all NULL -> built from the nodes on device */
} ples_mg_grid;

typedef struct ples_mg_opts {
    int precision;   /* PLES_MG_FP64 or PLES_MG_MIXED */
    int nu_pre;      /* smoothing sweeps per level on the way down (default 3) */
    int nu_post;     /* ... on the way up (default 3) */
    int nu_coarse;   /* sweeps on the coarsest level (default 16) */
    double omega;    /* damping of the eta-line Jacobi smoother (default 0.8):
                        every sweep with sweep_schedule = 0, the coarsest-level
                        sweeps otherwise */
    int krylov;      /* 1 = GCR acceleration of the refinement loop (default), 0 = plain */
    int stop_mode;   /* PLES_MG_STOP_REL_L2 (default) or PLES_MG_STOP_ABS_MAX */
    int project_rhs; /* 1 = FP64 compatibility projection of the rhs (default) */
    int halo_passes; /* red/black passes of the exact wall-ghost solve (0 = auto:
                        6 with eta walls only; with xi/zeta walls measured at
                        create from the contraction of the passes on a test
                        field, 24 to 400; create fails with
                        PLES_MG_ERR_UNSUPPORTED where the passes diverge) */
    int device;      /* HIP device ordinal (default 0) */
    int graphs;      /* 1 = replay each refinement step from a captured HIP graph
                        (default), 0 = launch its kernels one by one; steps are
                        always launched one by one while kernel timing is on */
    int inner_cycles; /* V-cycles per refinement step, the first from zero, each
                        further one continuing from the last (0 = auto = 1) */
    int autotune;    /* launch shapes of every level (smoother segments per
                        column and 16-bit metrics, residual segments x lanes):
                        0 (default) = the process-wide table's shape where it
                        holds the level's (ples_mg_tuning_load), else the fixed
                        rule; 1 = as 0, and a shape the table lacks is timed at
                        create (the candidates on the level's own arrays, the
                        fastest kept and added to the table, so equal shapes get
                        equal kernels); -1 = the fixed rule only */
    int fp32_krylov; /* PLES_MG_MIXED only: 1 = FP64 iterative refinement around
                        an FP32 GCR(2) preconditioned by the FP32 V-cycle (default);
                        0 = the FP32 V-cycle directly inside the FP64 GCR(2).
                        Batched solves follow the option (the refinement loop
                        driven on the device); served solves use the latter. */
    int smoother_fp16; /* PLES_MG_MIXED only: 1 = the smoother of the FP32 levels
                        reads a 16-bit copy of the level's metrics (scaled by a
                        power of four, folded into 1/pot), and so does the
                        V-cycle's fused residual-restriction on levels without
                        g13/g31; the FP32 operator of the Krylov loop and the
                        FP64 outer residual keep their metrics, so only the
                        preconditioner changes, not the solution.  Used on the
                        fine level (the autotuner may choose per level for the
                        smoother).  0 = FP32 metrics everywhere.  Default 1 */
    int coarse_exact; /* 1 = the V-cycle descends to the first coarse level
                        of at most 2048 cells (512 on grids with xi or zeta
                        walls; else to the last level) and solves that level
                        exactly when it is that small: a dense map built at create from the level's
                        operator with its wall ghosts eliminated (inverted on
                        the device), applied as one matrix-vector product
                        (the map in FP32 for PLES_MG_MIXED, FP64 vectors);
                        levels below it are built but not visited (default).
                        0 = all nlevel levels, nu_coarse sweeps on the last */
    int sweep_schedule; /* damping of the nu sweeps of a visit: 1 = 1 / (the
                        roots of the degree-nu Chebyshev polynomial on
                        [0.48, 2.4]), which damps that part of the smoother's
                        spectrum by 1 / T_nu(1.5) together (default); 0 = omega
                        for every sweep */
    int semicoarsen; /* first coarse level when the fine level's zeta coupling
                        is at least twice its xi coupling (sum g33 >= 2 sum g11,
                        e.g. dz <= dx / 1.41): 1 = halves zeta only (keeps the
                        xi and eta counts), 2 = halves eta and zeta (keeps xi);
                        the levels below halve all three.  0 = every level
                        halves all three counts, the reference's hierarchy
                        (fsrc/multigrid.f03:2293-3052) (default).  The
                        converged pressure is the same. */
} ples_mg_opts;

typedef struct ples_mg_result {
    int status;             /* PLES_MG_* */
    int cycles;             /* V-cycles applied */
    double rel_res;         /* ||b - A p||_2 / ||b||_2, FP64 */
    double abs_res;         /* max |b - A p| */
    double rhs_mean_removed;/* mean taken off the rhs by the projection */
    double seconds;         /* device time of the solve */
} ples_mg_result;

typedef struct ples_mg_handle ples_mg_handle;

void ples_mg_default_opts(ples_mg_opts* opts);
const char* ples_mg_status_string(int status);

/* Build the level hierarchy on the device: level-1 metrics (metrica) and
 * coarse metrics (mul_met) from the nodes, smoother factors, work vectors.
 * Replaces the init-time metrica -> mul_met -> wall sequence
 * (fsrc/stratiParticle.f03:767-769). */
int ples_mg_create(const ples_mg_grid* grid, const ples_mg_opts* opts, ples_mg_handle** out);
int ples_mg_destroy(ples_mg_handle* h);

/* Drop-in solve (replaces `call multi(...)`, fsrc/multigrid.f03:42-653).
 * fi: in = warm start, out = solution incl. ghost layers.  fi has
 * (jx+2)*(jy+2)*(jz+deepl+deepr) entries; only k = 0..jz+1 are touched.
 * rhs as `rhs(1:jx,1:jy,1:jz)`.  cs1..cs6 may be NULL (zero fluxes).
 * eps / ficycle as in the reference; the stop test is opts.stop_mode. */
int ples_mg_solve(ples_mg_handle* h, double* fi, int deepl, int deepr, const double* rhs,
                  const double* cs1, const double* cs2, const double* cs3, const double* cs4,
                  const double* cs5, const double* cs6, double dt, double eps, int ficycle,
                  ples_mg_result* res);

/* Residual r = s^2 * sum(face fluxes) - rhs of the fine level
 * (fsrc/resid_sndrcv.f03:80-151) after the exact ghost fill of fi with the
 * boundary fluxes.  r as `r(1:jx,1:jy,1:jz)`.  fi is not modified. */
int ples_mg_residual(ples_mg_handle* h, const double* fi, int deepl, int deepr,
                     const double* rhs, const double* cs1, const double* cs2,
                     const double* cs3, const double* cs4, const double* cs5,
                     const double* cs6, double dt, double* r);

/* Zero-mean (compatibility) projection of a rhs, in place on host arrays:
 * rhs -= (sum(rhs) - (sum of the wall fluxes cs, outflow - inflow) / dt) / N,
 * in FP64 on the device; *removed (may be NULL) receives the value taken off.
 * The reference has no such step (it only prints the rhs integral,
 * fsrc/stratiParticle.f03:2847-2866); on a Neumann / periodic problem an
 * inconsistent rhs has no solution.  ples_mg_solve, ples_mg_solve_resident
 * and ples_mg_projection_step apply it by themselves (opts.project_rhs).
 * The projected rhs and cs become the handle's resident ones. */
int ples_mg_project_rhs(ples_mg_handle* h, double* rhs, const double* cs1, const double* cs2,
                        const double* cs3, const double* cs4, const double* cs5, const double* cs6,
                        double dt, double* removed);

/* ples_mg_project_rhs on device arrays, allocating nothing.  No counterpart
 * in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  rhs: device array (1:jx,1:jy,1:jz),
 * projected in place; cs1..cs6: device arrays in the shapes above, NULL for
 * periodic directions (no flux); removed (may be NULL): a device double that
 * receives the value taken off; stream: hipStream_t (NULL = the null stream).
 * The work is enqueued on the stream and not waited for.  Scratch: a
 * file-scope device array of 1032 doubles (8 KB) per GPU, held until the
 * process ends and shared by every call on that GPU: calls that overlap on
 * one GPU race.  At most 2^31 - 2^18 cells (and per cs array), else
 * PLES_MG_ERR_ARG. */
int ples_mg_project_rhs_noalloc(double* rhs, int jx, int jy, int jz, const double* cs1,
                                const double* cs2, const double* cs3, const double* cs4,
                                const double* cs5, const double* cs6, double dt, double* removed,
                                void* stream);

/* Level-1 metrics as metrica builds them (fsrc/metrica.f03:85-483), in the
 * reference shapes: g[9] (see above), face vectors csv[9] = csx,csy,csz
 * (0:jx,1:jy,1:jz), etx,ety,etz (1:jx,0:jy,1:jz), ztx,zty,ztz
 * (1:jx,1:jy,0:jz), giac (1:jx,1:jy,1:jz).  Any output may be NULL. */
int ples_mg_metrica(ples_mg_handle* h, double* const g[9], double* const csv[9], double* giac);

/* Metrics of multigrid level n (1-based, n >= 1) in the ghosted device layout
 * copied to the host: each g[m] receives (nx+2)(ny+2)(nz+2) values, the
 * level's counts from ples_mg_level_shape, face index f stored at cell index
 * f.  For parity tests. */
int ples_mg_level_metrics(ples_mg_handle* h, int level, double* const g[9]);

/* One operation of multigrid level n (1-based) on host arrays in the
 * ghosted layout of that level ((nx+2)(ny+2)(nz+2) values, i fastest;
 * counts from ples_mg_level_shape), run by the kernels the V-cycle runs, in
 * the level's precision (inputs of an FP32 level are rounded to FP32).  For
 * kernel-level parity tests:
 *   PLES_MG_OP_SMOOTH: one eta-line smoother sweep, the relaxation of
 *     solut_sndrcv_sor / solut_sndrcv_slor (fsrc/multigrid.f03:656-1010,
 *     1939-2290): in0 = iterate (its ghosts as given), in1 = rhs, out = the
 *     iterate after the sweep (p + omega M_line^-1 (b - A p) in the interior,
 *     eta-wall ghosts and periodic images rewritten); P = eta segments per
 *     column (0: the level's own), omega = the damping;
 *   PLES_MG_OP_RESID: r = A p - b (fsrc/resid_sndrcv.f03:80-151), levels
 *     below the coarsest: in0 = p, in1 = b, out = r;
 *   PLES_MG_OP_RESTRICT: rc = -(1/8) sum of the 8 children
 *     (fsrc/restrict_sndrcv.f03:32-53): in0 = fine residual (level n), out =
 *     coarse rhs (level n+1);
 *   PLES_MG_OP_RESID_RESTRICT: the two fused as the V-cycle runs them: in0 =
 *     p, in1 = b (level n), out = coarse rhs (level n+1);
 *   PLES_MG_OP_PROLONG: pf += trilinear prolongation of ec over the
 *     ghost-inclusive range (fsrc/prolong.f03:64-261): in0 = ec (level n+1,
 *     ghosts included), in1 = pf (level n), out = the updated pf.
 * The operations overwrite the handle's level work arrays (not the resident
 * iterate or rhs of level 1). */
#define PLES_MG_OP_SMOOTH 0
#define PLES_MG_OP_RESID 1
#define PLES_MG_OP_RESTRICT 2
#define PLES_MG_OP_RESID_RESTRICT 3
#define PLES_MG_OP_PROLONG 4
int ples_mg_level_op(ples_mg_handle* h, int level, int op, int P, double omega, const double* in0,
                     const double* in1, double* out);

/* Cell counts of multigrid level n (1-based): jz/2^(n-1), and jx/2^(n-1),
 * jy/2^(n-1), or jx/2^(n-2) (and jy/2^(n-2)) below a first coarse level that
 * kept those counts (opts.semicoarsen); the reference's jxc/jyc/jzc
 * (fsrc/iniz_metrica.f03:22-47) for a full-coarsening hierarchy.  Outputs may
 * be NULL. */
int ples_mg_level_shape(const ples_mg_handle* h, int level, int* nx, int* ny, int* nz);

/* rhs = (div of contravariant fluxes)/dt (fsrc/diver.f03:20-28) with
 * uc (0:jx,1:jy,1:jz), vc (1:jx,0:jy,1:jz), wc (1:jx,1:jy,0:jz). */
int ples_mg_diver(ples_mg_handle* h, const double* uc, const double* vc, const double* wc,
                  double dt, double* rhs);

/* Pressure gradients of gradie (fsrc/gradie.f03:28-117) from fi (ghosted as
 * in ples_mg_solve): gra1..3 (1:jx,1:jy,1:jz), cgra1 (0:jx,1:jy,1:jz),
 * cgra2 (1:jx,0:jy,1:jz), cgra3 (1:jx,1:jy,0:jz).  Outputs may be NULL. */
int ples_mg_gradie(ples_mg_handle* h, const double* fi, int deepl, int deepr, double dt,
                   double* gra1, double* gra2, double* gra3, double* cgra1, double* cgra2,
                   double* cgra3);

/* ---- the time-loop steps around the solve ----
 * Cell velocities u, v, w are shaped like fi (0:jx+1, 0:jy+1, 1-deepl:jz+deepr)
 * (fsrc/iniz.f03:95-97); face fluxes uc (0:jx,1:jy,1:jz), vc (1:jx,0:jy,1:jz),
 * wc (1:jx,1:jy,0:jz) (fsrc/iniz.f03:103-105); cs1..cs6 as above.
 * potenziale = 1 is the reference's potential-flow switch.  Walls need at
 * least three cells across (PLES_MG_ERR_BAD_DIMS otherwise). */

/* update (fsrc/update.f03:2-289): u,v,w += delu,delv,delw in the interior,
 * then every ghost face (not edge) from the interior: parabolic extrapolation
 * 1.875/-1.25/0.375 at walls, periodic copies otherwise.  delu.. are
 * (0:jx+1, 0:jy+1, 0:jz+1) (fsrc/stratiParticle.f03:315); only the interior
 * is read. */
int ples_mg_update(ples_mg_handle* h, double* u, double* v, double* w, int deepl, int deepr,
                   const double* delu, const double* delv, const double* delw);

/* contra (fsrc/contra.f03:2-241): intermediate contravariant fluxes uc,vc,wc
 * (in/out) from u,v,w, and on wall faces the flux change cs = new - old
 * (potenziale = 1: cs = -old) written to cs[0..5] = cs1..cs6 for the wall
 * directions (entries may be NULL).  The handle keeps cs as the boundary
 * fluxes of its next resident solve. */
int ples_mg_contra(ples_mg_handle* h, const double* u, const double* v, const double* w, int deepl,
                   int deepr, double* uc, double* vc, double* wc, double* const cs[6], int potenziale);

/* vel_up (fsrc/vel_up.f03:2-313, freesurface = 0): u,v,w += gra1..3
 * (1:jx,1:jy,1:jz), periodic ghost copies, uc,vc,wc -= dt*cgra1..3 on the
 * non-wall faces (cgra in the uc,vc,wc shapes, as ples_mg_gradie returns
 * them). */
int ples_mg_vel_up(ples_mg_handle* h, double* u, double* v, double* w, int deepl, int deepr,
                   double* uc, double* vc, double* wc, const double* const gra[3],
                   const double* const cgra[3], double dt, int potenziale);

/* check_divergence (fsrc/check_divergence.f03:131-190, bodyforce = 0):
 * out = {max|div| inside the field (i=2..jx-1, j=2..jy-1, k=2..jz-1),
 *        max|div/giac| inside, max|div| all cells, max|div/giac| all}. */
int ples_mg_check_divergence(ples_mg_handle* h, const double* uc, const double* vc,
                             const double* wc, double out[4]);

/* The projection step of the time loop in one call, on the device:
 * update -> contra -> diver -> (compatibility projection) -> multi ->
 * gradie -> vel_up -> check_divergence
 * (fsrc/stratiParticle.f03:2678-2937, 3526).  All arrays in/out as in the
 * single-step entry points; fi is the warm start and the result; cs may be
 * NULL; div (may be NULL) receives the check_divergence maxima. */
int ples_mg_projection_step(ples_mg_handle* h, double* u, double* v, double* w, int deepl, int deepr,
                            const double* delu, const double* delv, const double* delw,
                            double* uc, double* vc, double* wc, double* const cs[6], double* fi,
                            double dt, double eps, int ficycle, int potenziale,
                            ples_mg_result* res, double div[4]);

/* ---- device-resident use (inputs stay in HBM between calls) ---- */

/* Synthetic divergence rhs on the device: N(0,1) cell velocity from a
 * counter-based generator (seed), contra-style face fluxes with zero wall
 * fluxes, diver, zero-mean projection.  Also resets nothing else. */
int ples_mg_case_rhs(ples_mg_handle* h, uint64_t seed, double dt);
/* Resident iterate: set to zero / perturb by rel*std(p)*N(0,1) (seed). */
int ples_mg_zero_fi(ples_mg_handle* h);
int ples_mg_perturb_fi(ples_mg_handle* h, uint64_t seed, double rel);
/* Solve with the resident rhs and resident iterate as warm start. */
int ples_mg_solve_resident(ples_mg_handle* h, double eps, int maxcycles, ples_mg_result* res);
/* Copy the resident iterate / rhs to or from host arrays (layouts as in
 * ples_mg_solve, deepl = deepr = 1 for fi). */
int ples_mg_get_fi(ples_mg_handle* h, double* fi);
int ples_mg_set_fi(ples_mg_handle* h, const double* fi);
int ples_mg_get_rhs(ples_mg_handle* h, double* rhs);
int ples_mg_set_rhs(ples_mg_handle* h, const double* rhs);

/* ---- batched solves ----
 * n independent problems of any sizes (the grids may differ; one opts for
 * all), each with its own handle and HIP stream.  ples_mg_batch_solve runs
 * the refinement loops of all of them together: one HIP graph holds
 * steps_per_launch refinement steps (V-cycle + FP64 update) of every problem,
 * forked onto the problems' streams so their kernels overlap on the device,
 * and one launch of it advances the whole batch.  Grouped (the default):
 * problems whose refinement steps issue the same launch sequence (same shape
 * and options) form a family, and each kernel of the family's steps is ONE
 * launch over all its problems (the problem is the grid's z index, its
 * arguments come from a device table); families run side by side.  Each
 * problem stops on its own test (device-side; a converged problem's blocks
 * exit at once), the host reads all residuals back with one copy per launch.
 * Results are bitwise those of solving each problem alone.  There is no
 * reference counterpart: PLES solves one pressure field per time step. */
typedef struct ples_mg_batch ples_mg_batch;
int ples_mg_batch_create(int n, const ples_mg_grid* grids, const ples_mg_opts* opts,
                         ples_mg_batch** out);
int ples_mg_batch_destroy(ples_mg_batch* b);
int ples_mg_batch_size(const ples_mg_batch* b);
/* The handle of problem i (owned by the batch): rhs and iterate set-up and
 * read-back use the device-resident entry points above. */
ples_mg_handle* ples_mg_batch_handle(ples_mg_batch* b, int i);
size_t ples_mg_batch_device_bytes(const ples_mg_batch* b);
/* Solve every problem from its resident iterate to eps (opts.stop_mode);
 * res[i] per problem (seconds = the whole batch).  steps_per_launch 0 = 2.
 * Returns PLES_MG_OK when all converged, else the first failing status. */
int ples_mg_batch_solve(ples_mg_batch* b, double eps, int maxcycles, int steps_per_launch,
                        ples_mg_result* res);
/* Grouped launches on (1, the default) or off (0: every problem's kernels
 * launched on its own stream).  The number of families of the last grouped
 * graph (0 before the first grouped solve) by ples_mg_batch_families. */
int ples_mg_batch_set_grouped(ples_mg_batch* b, int on);
int ples_mg_batch_families(const ples_mg_batch* b);

/* ---- serving loop ----
 * A solve server for many cases (independent simulations, each a grid and a
 * seeded divergence rhs).  Each request is one time step of a case: it
 * perturbs the case's previous pressure, kept resident on the device as the
 * warm start (relative std perturb_rel, seeded), or starts from zero when the
 * case is cold, and solves to its own tolerance.  Requests are admitted while
 * fewer than max_active run and the case is, or can be made, resident within
 * hbm_budget device bytes and warm_cases resident cases (idle cases are
 * evicted least recently used first, losing their warm start).  Requests that
 * arrive while others run join at the next ples_mg_server_step; every running
 * request advances on its own stream (graph-replayed refinement steps), so
 * their kernels overlap on the device, and none waits for another to finish.
 * A cold case's handle is created (and its rhs generated) on the server's
 * worker thread, its bytes and slot reserved at once; the request is admitted
 * at the first ples_mg_server_step after the creation lands, and the requests
 * behind it are admitted meanwhile.  An evicted handle is destroyed on the
 * same thread: its bytes count as free at eviction, so device memory in use
 * can exceed hbm_budget by the evictions not yet freed.  ples_mg_server_wait
 * also returns when a creation lands.
 * No reference counterpart: PLES solves one field per time step in-line. */
typedef struct ples_mg_server ples_mg_server;
typedef struct ples_mg_server_opts {
    size_t hbm_budget; /* device bytes for resident cases (0 = 80% of free memory) */
    int max_active;    /* requests running at once (default 8) */
    int warm_cases;    /* resident cases kept (default 16) */
    int device;
} ples_mg_server_opts;
void ples_mg_server_default_opts(ples_mg_server_opts* so);
int ples_mg_server_create(const ples_mg_server_opts* so, const ples_mg_opts* opts,
                          ples_mg_server** out);
int ples_mg_server_destroy(ples_mg_server* sv);
/* Register a case (nodes are copied; no device memory until admitted).
 * PLES_MG_ERR_BUSY if the case alone exceeds the budget. */
int ples_mg_server_add_case(ples_mg_server* sv, uint64_t case_id, const ples_mg_grid* grid,
                            uint64_t rhs_seed, double dt);
int ples_mg_server_submit(ples_mg_server* sv, uint64_t case_id, uint64_t perturb_seed,
                          double perturb_rel, double eps, int maxcycles, int64_t* ticket);
/* Admit waiting requests and advance every running one without blocking;
 * *n_done = requests completed in this call. */
int ples_mg_server_step(ples_mg_server* sv, int* n_done);
/* Block until some running request's next device result is available. */
int ples_mg_server_wait(ples_mg_server* sv);
/* PLES_MG_OK and the result (seconds = submit-to-completion wall time,
 * *warm = 1 if it started from the case's previous solution) once done,
 * PLES_MG_ERR_BUSY while pending. */
int ples_mg_server_poll(ples_mg_server* sv, int64_t ticket, ples_mg_result* res, int* warm);
int ples_mg_server_pending(const ples_mg_server* sv);
/* The case's resident pressure (layout of ples_mg_get_fi); ERR_BUSY if cold or running. */
int ples_mg_server_get_fi(ples_mg_server* sv, uint64_t case_id, double* fi);
/* Start the case's next request from this pressure (layout of ples_mg_set_fi),
 * as a warm start: a field PLES wrote to its restart file (fsrc/restart.f03:33-98
 * reads it back into fi) resumes a served simulation where it stopped.  Copied
 * to the device now if the case is resident, else kept on the host until the
 * case is admitted; ERR_BUSY while the case runs. */
int ples_mg_server_set_fi(ples_mg_server* sv, uint64_t case_id, const double* fi);
int ples_mg_server_stats(const ples_mg_server* sv, size_t* used_bytes, int* resident,
                         long* completed, long* evictions, long* cold_starts);

/* ---- restart file of the flow field ----
 * The layout PLES reads at restart (old_res_form, fsrc/restart.f03:33-98):
 * header nscal / ti / bbx dbbx, then one record u v w fi rhov(1..nscal) per
 * cell over (0:jx+1,0:jy+1,0:jz+1), then uc (0:jx,1:jy,1:jz), vc
 * (1:jx,0:jy,1:jz), wc (1:jx,1:jy,0:jz) one value per record, data in
 * Fortran format 10e18.10.  Arrays in the shapes above (rhov(isc,i,j,k),
 * isc fastest); read outputs may be NULL.  Host-only. */
int ples_mg_restart_write(const char* path, int jx, int jy, int jz, int nscal, double ti, double bbx,
                          double dbbx, const double* u, const double* v, const double* w,
                          const double* fi, const double* rhov, const double* uc, const double* vc,
                          const double* wc);
int ples_mg_restart_read(const char* path, int jx, int jy, int jz, int nscal, double* ti, double* bbx,
                         double* dbbx, double* u, double* v, double* w, double* fi, double* rhov,
                         double* uc, double* vc, double* wc);

/* Autotuner table (process-wide): load entries from a text file (one level
 * shape per line: nx ny nz ip kp f32 fine half P rP rbx use16 -- half: the
 * level has the 16-bit smoother metrics, use16: its smoother reads them;
 * lines of the older form nx ny nz ip kp f32 fine P rP rbx are read with
 * half = use16 = 0), so handles created afterwards (opts.autotune >= 0) take
 * those shapes without timing; save the current table. */
int ples_mg_tuning_load(const char* path);
int ples_mg_tuning_save(const char* path);
/* Empty the table: handles created afterwards take the fixed rule's shapes
 * (or time theirs with opts.autotune = 1). */
int ples_mg_tuning_clear(void);

/* HBM bandwidth probes (no counterpart in the reference: the roofline's
 * measured ceiling).  Copy: GB/s of read + write bytes of a copy of `bytes`
 * bytes repeated reps times; read: GB/s of a read-only sweep.  Each lane
 * keeps four independent 16-byte loads in flight. */
int ples_mg_copy_bandwidth(int device, size_t bytes, int reps, double* gbs);
int ples_mg_read_bandwidth(int device, size_t bytes, int reps, double* gbs);
/* Twelve FP32 streams, eleven read and one written, 4-byte accesses (the
 * array count and width of a fine-smoother row): GB/s of the `bytes` read +
 * written per launch. */
int ples_mg_streams_bandwidth(int device, size_t bytes, int reps, double* gbs);

/* Launch shape of level `level` (1-based): smoother segments per column P,
 * V-cycle residual segments rP x lanes rbx (outputs may be NULL). */
int ples_mg_tiles(const ples_mg_handle* h, int level, int* P, int* rP, int* rbx);

/* Number of levels the V-cycle visits: nlevel, or fewer with coarse_exact,
 * where the cycle stops at the first coarse level small enough for the exact
 * dense solve (the reference's multi visits all nlevel levels,
 * fsrc/multigrid.f03:216-563). */
int ples_mg_cycle_levels(const ples_mg_handle* h, int* levels);

/* HIP stream the handle launches on (hipStream_t as void*). */
void* ples_mg_stream(ples_mg_handle* h);
/* Device bytes held by the handle. */
size_t ples_mg_device_bytes(const ples_mg_handle* h);
/* Device bytes ples_mg_create would allocate for this grid and options. */
size_t ples_mg_estimate_bytes(const ples_mg_grid* grid, const ples_mg_opts* opts);

/* ---- per-kernel timing (HIP events on the handle's stream) ---- */
#define PLES_MG_K_SMOOTH 0   /* eta-line smoother, all levels */
#define PLES_MG_K_SMOOTH_FINE 1 /* eta-line smoother, finest level only */
#define PLES_MG_K_RESID 2    /* V-cycle residual, all levels */
#define PLES_MG_K_RESID_FINE 3
#define PLES_MG_K_OUTER_APPLY 4 /* FP64 operator + dots of the refinement loop */
#define PLES_MG_K_OUTER_UPDATE 5
#define PLES_MG_K_TRANSFER 6 /* restriction + prolongation */
#define PLES_MG_K_HALO 7
#define PLES_MG_K_COUNT 8
#define PLES_MG_TIME_ALL 0xFF
/* Event timing of every launch of the families whose bit is set in mask
 * (bit f = family f; PLES_MG_TIME_ALL, 0 = off); resets the stats. */
int ples_mg_set_timing(ples_mg_handle* h, int mask);
/* With timing on, time one refinement step in every `every` (launched
 * eagerly with the events); the other steps replay their captured graphs, so
 * the timed run keeps the graphs' launch rate.  1 = every step (default).
 * No counterpart in the reference: the per-kernel measurement of bench-style
 * callers that time kernels inside a throughput run. */
int ples_mg_set_timing_every(ples_mg_handle* h, int every);
/* Launch count, summed device milliseconds and algorithmic bytes per kernel
 * family since timing was enabled. */
int ples_mg_kernel_stats(ples_mg_handle* h, int family, long* launches, double* ms,
                         double* bytes);

#ifdef __cplusplus
}
#endif
#endif
