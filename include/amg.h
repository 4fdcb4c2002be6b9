/*
 * amg.h -- C ABI of the MI355X (gfx950) pressure-projection Poisson solver.
 *
 * Drop-in replacement for the reference's multigrid pressure solver mgMK
 * (archive libmg.a, Makefile:56,124-145,205-206), which the Fortran DNS/LES
 * drivers call only through four external subroutines:
 *
 *   mg_read_params(infile, stat)          mgMK/mg_read_params.F90:1
 *   MG_init(stat)                         mgMK/mg_init.F90:1
 *   mg_solver(phif, rhsf, iter, relmax)   mgMK/mg_solver.F90:1
 *   mg_end(ierr)                          mgMK/mg_end.F90:1
 *
 * and the projection step around it in dns/:
 *
 *   Psource(U,V,W,SRC,dt,stat)            dns/NSsolver.F90:2167
 *   Pgrad(p,U,V,W,dt)                     dns/NSsolver.F90:2023
 *   divergence(u,v,w,tmp,plane_write,ierr) dns/Routines.F90:54
 *
 * Every field crosses this boundary in the reference's own layout: a FP64
 * Fortran array phi(sx-1:ex+1, sy-1:ey+1, sz-1:ez+1), i.e. a contiguous
 * block of (nx+2)*(ny+2)*(nz+2) doubles, x fastest, one ghost layer.
 * Functions named *_host take host pointers (the Fortran arrays), functions
 * named *_device take device pointers (HBM-resident fields).
 *
 * All functions return AMG_OK (0) or a negative AMG_E* code; amg_strerror()
 * names it.  No function of this header allocates on the caller's behalf
 * except amg_create / amg_batch_create / amg_server_create.
 */
#ifndef AMG_H
#define AMG_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define AMG_OK            0
#define AMG_EINVAL       -1   /* bad argument / NULL pointer                        */
#define AMG_EBADSIZE     -2   /* grid sizes violate the level constraints          */
#define AMG_EBADBC       -3   /* boundary code outside {0,1,2} (MG_init stat=-1)   */
#define AMG_ENOMEM       -4   /* device allocation failed / HBM budget exceeded    */
#define AMG_ENOTCONV     -5   /* maxcy cycles without reaching the tolerance       */
#define AMG_ENAN         -6   /* NaN/Inf in the right-hand side or the iterates    */
#define AMG_EHIP         -7   /* HIP runtime error (message via amg_last_error)    */
#define AMG_EUNSUPPORTED -8   /* feature outside this build (e.g. 3-D density)     */
#define AMG_EPARSE       -9   /* INI file missing / key missing / bad value        */
#define AMG_ENOGPU      -10   /* no HIP device visible                             */
#define AMG_EBUSY       -11   /* server queue full / request id unknown yet        */

/* ---- boundary codes: the reference's tbc values (mgMK/mg_read_params.F90:74-80) */
#define AMG_BC_PERIODIC  0
#define AMG_BC_NEUMANN   1
#define AMG_BC_DIRICHLET 2

/* ---- smoothers: the reference's [MGsolver] Smoother key (mgMK/mg_cycle.F90:53-73) */
#define AMG_SMOOTHER_RBPGS 0  /* point red-black Gauss-Seidel (mgMK/mg_relax.F90:85)   */
#define AMG_SMOOTHER_LJ    2  /* z-line red-black Gauss-Seidel (mgMK/mg_relax.F90:423)  */

/* ---- precision of the V-cycle ---- */
#define AMG_FP64  0           /* plain FP64 V-cycles on p                               */
#define AMG_MIXED 1           /* FP32 V-cycle inside an FP64 defect-correction loop     */

/* ---- stop rules ---- */
#define AMG_STOP_RESIDUAL 0   /* ||b - A p||_inf / ||b||_inf <= tol                     */
#define AMG_STOP_RELMAX   1   /* max|p_new-p_old|/max|p_new| <= tol (mg_residual.F90:60-94) */

#define AMG_MAX_LEVELS 16

/*
 * Solver controls.  Mirrors the control fields of the reference's module
 * mgVARS (include/mg_modules.F90:14-54) read by mg_read_params
 * (mgMK/mg_read_params.F90:44-214).  Face order of tbc/vbc is the
 * reference's: [0]=X2max [1]=X3min [2]=X2min [3]=X3max [4]=X1min [5]=X1max.
 */
typedef struct amg_config {
  int    nlevels[3];            /* nxlevels, nylevels, nzlevels; all 0 = automatic   */
  int    smoother;              /* AMG_SMOOTHER_*                                     */
  int    maxcy;                 /* maximum V-cycles                                   */
  double tol;                   /* tolerance of stop_rule                             */
  int    stop_rule;             /* AMG_STOP_*                                         */
  int    pre_relax[AMG_MAX_LEVELS];  /* PreRelax, coarsest level first (reference order) */
  int    post_relax[AMG_MAX_LEVELS]; /* PstRelax, coarsest level first                   */
  int    tbc[6];                /* AMG_BC_* per face                                  */
  double vbc[6];                /* Neumann flux / Dirichlet value per face            */
  int    precision;             /* AMG_FP64 or AMG_MIXED                              */
  int    report_relmax;         /* FP64 mode: also compute relmax (costs a copy of p) */
  int    verbose;               /* print one line per V-cycle to stderr               */
  int    strict;                /* accepted for compatibility; one GPU has no halos   */
  int    use_graph;             /* capture the V-cycle into a hipGraph (default 1)    */
  int    coarse_direct;         /* 1 (default): solve the first level below the finest
                                   with at most 2^18 cells exactly (separable solve in
                                   FP64) when z is not periodic and the level is above
                                   the coarse tail; the levels below it are not cycled.
                                   0: cycle every level down to the coarsest           */
  int    pair_cycles;           /* 1 (default): mixed precision with the residual stop
                                   rule may close two FP32 V-cycles with one FP64
                                   correction when the solve's own convergence rate
                                   says one more cycle cannot reach tol (the device
                                   decides after every correction); 0: one correction
                                   per V-cycle                                          */
} amg_config;

/*
 * Grid and background density, as the reference's modules Grid and Flow
 * hold them (include/modules.F90:17-55): per direction the arrays
 * xe, xc, dxe, dxc of length n+2 (ghost-inclusive, loadGrid conventions,
 * utils/loadGrid.F90:79-90), and rho(z) of length nz+2.  rho == NULL means
 * rho = 1 (the Boussinesq case).  The reference's Flow::r is 3-D but is
 * horizontally uniform (dns/IC.F90:1085-1092); the Fortran shim passes
 * one column of it.
 */
typedef struct amg_grid {
  int nx, ny, nz;
  const double *xe, *xc, *dxe, *dxc;
  const double *ye, *yc, *dye, *dyc;
  const double *ze, *zc, *dze, *dzc;
  const double *rho;
} amg_grid;

/* Per-solve report (all optional outputs). */
typedef struct amg_stats {
  int    iters;                 /* V-cycles run (maxcy+1 when not converged, as a Fortran DO) */
  int    converged;
  double relres;                /* final ||b - A p||_inf / ||b||_inf (FP64)          */
  double relmax;                /* final reference-style relmax                        */
  double res0;                  /* initial relative residual                          */
  double bnorm;                 /* ||b||_inf after boundary folding / compat projection */
  double ms_total;              /* device time of the solve (HIP events)               */
  double ms_cycles;             /* device time of the V-cycles only                    */
  double bytes_cycles;          /* algorithmic bytes of the V-cycles run              */
  int    nlevels;
  int    level_n[AMG_MAX_LEVELS][3];
} amg_stats;

typedef struct amg_solver_s* amg_handle;

/* ---- configuration -------------------------------------------------------- */

/* Defaults: automatic levels, LJ smoother, V(2,2), tol 1e-10 on the relative
 * residual, maxcy 100, mixed precision, periodic x/y, Neumann z. */
void amg_config_default(amg_config* cfg);

/* Reads [MGsolver] and [BCs] of a reference-format INI file
 * (mgMK/mg_read_params.F90:44-214; keys Smoother, Strict, Verbose,
 * nxlevels, nylevels, nzlevels, maxcy, tolmax, PreRelax, PstRelax, X1/X2/X3
 * and [X?min]/[X?max] PTYPE/PVALUE).  tolmax sets tol with stop_rule
 * AMG_STOP_RELMAX, as in the reference; an extra key 'tolres' selects the
 * residual rule.  Replaces mg_read_params(infile, stat). */
int amg_read_params(const char* infile, amg_config* cfg);

/* ---- single solver -------------------------------------------------------- */

/* Validates sizes (mgMK/mg_check_size.F90:43-151), builds the level
 * hierarchy and 1-D coefficient tables, allocates device memory.
 * Replaces MG_init(stat) (mgMK/mg_init.F90:1-351). */
int amg_create(const amg_config* cfg, const amg_grid* grid, amg_handle* out);

/* Frees everything.  Replaces mg_end(ierr) (mgMK/mg_end.F90:1). */
int amg_destroy(amg_handle h);

/* Solve with host (Fortran) arrays: phif = initial guess in, solution out
 * with filled ghost layers (mgMK/mg_solver.F90:139); rhsf is read only
 * (the reference clobbers it with the fine residual; callers never reuse
 * it, dns/NSsolver.F90:278,293).  iter and relmax as mg_solver returns
 * them.  Returns AMG_ENOTCONV (after printing the reference's WARNING line)
 * when maxcy cycles do not converge.  Replaces mg_solver(phif,rhsf,iter,relmax). */
int amg_solve_host(amg_handle h, double* phif, const double* rhsf,
                   int* iter, double* relmax, amg_stats* st);

/* Same with device-resident ghosted arrays; runs on the handle's stream and
 * returns after the solve completes. */
int amg_solve_device(amg_handle h, double* d_phif, const double* d_rhsf,
                     int* iter, double* relmax, amg_stats* st);

/* The HIP stream (hipStream_t) the handle launches on. */
void* amg_stream(amg_handle h);

/* Number of levels and their sizes (finest first). */
int amg_levels(amg_handle h, int* nlevels, int sizes[][3]);

/* The level solved directly by the separable solve (amg_config.coarse_direct),
 * or -1 when the V-cycle goes down to the coarsest level. */
int amg_direct_level(amg_handle h);

/* ---- projection step (FP64, device arrays, reference layout) ---------------
 * Psource: SRC = D(rho u*)/dt on interior cells, ghosts zeroed
 *          (dns/NSsolver.F90:2204-2219).
 * Pgrad:   u -= dt G(p/rho) on every interior face index, then the periodic
 *          velocity ghost wrap (dns/NSsolver.F90:2057-2082).
 * Divergence: div field (optional output), volume-weighted mean |div|,
 *          max |div| and its 1-based (i,j,k) (dns/Routines.F90:123-173). */
int amg_psource_device(amg_handle h, const double* U, const double* V, const double* W,
                       double* SRC, double dt);
int amg_pgrad_device(amg_handle h, const double* p, double* U, double* V, double* W, double dt);
int amg_divergence_device(amg_handle h, const double* U, const double* V, const double* W,
                          double* div_or_null, double* avg, double* maxabs, int loc[3]);

/* Allocation-free divergence report: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * kernels and results as amg_divergence_device, which allocates its block
 * partials on its first call; this one keeps them in a static 256 KiB
 * device array per GPU instead and allocates nothing.  Limit: grids of at
 * most 33 554 432 cells (AMG_EBADSIZE beyond).  Calls that overlap on one
 * GPU share that array and race. */
int amg_divergence_device_noalloc(amg_handle h, const double* U, const double* V,
                                  const double* W, double* div_or_null, double* avg,
                                  double* maxabs, int loc[3]);

/* ---- the time stepper's boundary handling and vertical viscous step -------
 * Boundary conditions of the flow fields, the reference's module boundC
 * (include/modules.F90:82-85): TB(1:3,1:2,1:12) (type) and VB(1:3,1:2,1:12)
 * (value) per direction, face and variable class, in Fortran order
 * (index (d-1) + 3*(f-1) + 6*(v-1)), so a Fortran caller passes its arrays
 * as they are.  Types: 1 Dirichlet, 2 Neumann, 3 periodic, 4 Neumann on the
 * ghost only, 6 wall (mean = value, z), 7 wall (zero, z), 8 linear
 * extrapolation (x). */
typedef struct amg_flowbc {
  int    tb[72];
  double vb[72];
} amg_flowbc;

/* ghost(var, vtype, stat) (dns/BC.F90:1-187, bound_update :334-584), serial:
 * every ghost takes its periodic image, then each face's rule runs in the
 * order X1min, X1max, X2min, X2max, X3min, X3max.  vtype is the reference's
 * type name ("u", "utemp", "p", "rho", "cfluc", ...); an unknown name is
 * AMG_EINVAL (the reference prints and goes on). */
int amg_ghost_device(amg_handle h, double* var, const char* vtype, const amg_flowbc* bc);

/* Pgrad(p,U,V,W,dt) complete (dns/NSsolver.F90:2023-2085): the velocity
 * correction of amg_pgrad_device followed by ghost(U,'utemp'),
 * ghost(V,'vtemp'), ghost(W,'wtemp') with the given flow BCs. */
int amg_pgrad_ghost_device(amg_handle h, const double* p, double* U, double* V, double* W,
                           double dt, const amg_flowbc* bc);

/* var_advance_CN_PL(veltemp1,veltemp2,f1,f2,f3,rkdt,var)
 * (dns/NSsolver.F90:3054-3465), the DNS branch on one GPU: Crank-Nicolson in
 * z for variable var (1 u, 2 v, 3 w, 5 rho, 8 scal1) of
 *   veltemp2 <- f1*veltemp2 + f2*veltemp1*f3 + rkdt/2 (L veltemp2 + L veltemp2_new)
 * with the wall rows of bc, a Thomas solve per interior column
 * (mgMK/mg_relax.F90:541-595) and ghost(veltemp2, var's type).  p_or_null
 * non-NULL is the reference's pcorrect (predicted pressure gradient).
 * rRe, rPr, rSc are 1/Re, 1/Pr, 1/Sc (module Parameters).  Wall types the
 * reference stops on, and systems with a zero pivot, are AMG_EINVAL. */
int amg_cn_advance_device(amg_handle h, const double* veltemp1, double* veltemp2, double f1,
                          double f2, double f3, double rkdt, int var, const double* p_or_null,
                          const amg_flowbc* bc, double rRe, double rPr, double rSc);

/* ---- batched uneven instances (one launch per level step) ---------------- */
typedef struct amg_batch_s* amg_batch;

/* All instances share cfg's smoother, sweep counts, precision, tol and
 * maxcy; grids (and so level counts) may differ. */
int amg_batch_create(int n, const amg_config* cfg, const amg_grid* grids, amg_batch* out);
int amg_batch_destroy(amg_batch b);
/* d_phi[i], d_rhs[i]: device ghosted arrays of instance i.  Each instance
 * stops when it converges; the result of each equals its solve alone. */
int amg_batch_solve_device(amg_batch b, double* const* d_phi, const double* const* d_rhs,
                           int* iters, double* relres, amg_stats* total);

/* ---- serving: continuous batching of concurrent requests ------------------ */
typedef struct amg_server_s* amg_server;

/* hbm_budget_bytes: device bytes the server may hold for resident
 * instances (hierarchies + warm-start pressure); max_slots: resident
 * instances; max_batch: instances advanced per round. */
int amg_server_create(const amg_config* cfg, double hbm_budget_bytes, int max_slots,
                      int max_batch, amg_server* out);
int amg_server_destroy(amg_server s);
/* Registers an ensemble member's grid (its hierarchy stays resident while
 * the budget allows); returns a member id >= 0 or an error. */
int amg_server_add_member(amg_server s, const amg_grid* grid);
/* Enqueue a solve of member m with device RHS (ghosted, FP64).  The
 * member's previous pressure is the warm start; the solution is written
 * to d_phi_out (ghosted) on completion.  Returns a request id >= 0. */
int amg_server_submit(amg_server s, int member, const double* d_rhs, double* d_phi_out);
/* Non-blocking completion check: 1 done, 0 pending, <0 error.  On done
 * fills the request's iterations, relres, and submit->complete latency. */
int amg_server_poll(amg_server s, int req, int* iters, double* relres, double* latency_ms);
/* Block until every submitted request completed. */
int amg_server_drain(amg_server s);

/* ---- measurement and test hooks (no reference counterpart) ------------------ */
/* Runs one kernel of the hierarchy 'reps' times back to back on the handle's
 * stream without synchronising; 'precision' must be the handle's.
 *   which 0: red smoother half-sweep on 'level'     4: black half-sweep
 *         1: residual + restriction level -> level+1
 *         2: prolongation-add level+1 -> level
 *         3: FP64 finest-level kernel (FP64: residual check; mixed: correct+defect)
 *         5: coarsest-level solve ('level' must be the coarsest)
 *         6: red half-sweep with a zero initial guess
 *         7/8: red/black z-line half-sweep of the coarse tail (short columns)
 *         10/11: red/black z-line half-sweep in the short form (FP32)
 *         12/13: red/black z-line half-sweep in the lean form
 *         14: separable direct solve u = A^-1 f ('level' must be amg_direct_level)
 *         15: the V-cycle's prolongation-add level+1 -> level, which leaves the
 *             cells that the level's first post-smoothing half-sweep (red)
 *             rewrites without reading them
 *         16: the V-cycle's residual + restriction level -> level+1, which
 *             takes as zero the residuals that the level's last
 *             pre-smoothing half-sweep (black) solved (marching restriction
 *             only; otherwise as 1)
 * Returns the kernel's algorithmic bytes per launch.  The caller times the
 * stream (amg_stream) with its own events. */
int amg_run_kernel(amg_handle h, int which, int level, int precision, int reps, double* bytes);

/* Copies field 'field' (0 = u, 1 = f) of level 'level' between the device
 * and buf (nz*ny*nx doubles, x fastest, no ghosts): dir 0 device -> buf,
 * dir 1 buf -> device (converted to the level's precision).  Level 0 of a
 * mixed handle is the FP32 correction level; level -1 is the FP64 finest
 * pressure/RHS (p, b) of either precision.  Synchronous. */
int amg_level_field(amg_handle h, int level, int field, int dir, double* buf);

const char* amg_strerror(int code);
const char* amg_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* AMG_H */
