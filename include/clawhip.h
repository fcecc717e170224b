/*
 * clawhip.h -- C ABI of libclawhip.so, the MI355X (gfx950) implementation of
 * PyClaw's 2D time step.  Plain pointers and sizes only; every array argument
 * is a device pointer; every call is stream-ordered on the stream passed in
 * and returns an int status (CLAW_OK or a CLAW_E* code, never aborts).
 *
 * Replaces the reference's f2py module ABI:
 *   classic2.step2   -- src/pyclaw/classic/step2.f90:2-5   (called at classic/solver.py:537-539)
 *   classic2.step2ds -- src/pyclaw/classic/step2ds.f90:2-5 (called at classic/solver.py:525-531)
 *   sharpclaw2.flux2 -- src/pyclaw/sharpclaw/flux2.f90:2   (called at sharpclaw/solver.py:577)
 * and the Riemann-solver plugin (the rpn2/rpt2 procedure pointers passed at
 * classic/solver.py:518-519, with scalar parameters in the /cparam/ common block
 * set at state.py:162-182) by a Riemann-solver id plus parameters.
 *
 * Array layout ("planar padded"): component m of cell (i, j), 0-based interior
 * indices, lives at  a[m*plane + (j+ng)*ldq + (i+ng)],  x fastest.  ng ghost
 * layers are stored but are only read for sides whose boundary code is
 * CLAW_BC_STORED; the other codes synthesise ghost values in the kernel.
 * Kernels write interior cells only.
 */
#ifndef CLAWHIP_H
#define CLAWHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CLAW_OK = 0,
    CLAW_EINVAL = 1,       /* invalid argument or unsupported option */
    CLAW_EDEVICE = 2,      /* HIP runtime error */
    CLAW_ENOMEM = 3,       /* device allocation failed */
    CLAW_ETOOLARGE = 4     /* size beyond a documented maximum */
};

/* Riemann solvers (clawpack.riemann names). */
enum {
    CLAW_RP_ACOUSTICS = 0,     /* rp2_acoustics:    p[0..3] = rho, bulk, cc, zz     */
    CLAW_RP_VC_ACOUSTICS = 1,  /* rp2_vc_acoustics: aux = (rho, c) per cell        */
    CLAW_RP_EULER = 2,         /* rp2_euler_5wave:  p[0..1] = gamma, gamma1; efix  */
    CLAW_RP_SHALLOW = 3        /* rp2_shallow_roe_with_efix: p[0] = grav           */
};

/* Boundary codes (pyclaw.BC, solver.py:12-18); STORED reads the ghost cells as
   they are in the array (the caller filled them, e.g. for user/custom BCs). */
enum { CLAW_BC_STORED = 0, CLAW_BC_EXTRAP = 1, CLAW_BC_PERIODIC = 2, CLAW_BC_WALL = 3 };

/* Arithmetic of a step.  CLAW_F64: double precision with every multiply and
   add rounded separately, as the reference's CPU build (bit-identical
   results).  CLAW_F32: single precision.  CLAW_F64_FMA: double precision with
   multiply-adds fused (one rounding) and one division per shared
   denominator, for the classic and SharpClaw steps (not the batched runs or
   the adjoint).  Classic steps agree with CLAW_F64 to 1e-13 relative per
   step (tested); SharpClaw steps agree to rounding on rough data, while on
   smooth data WENO5's weights amplify any rounding difference (one-ulp input
   noise moves the reference arithmetic itself by ~1e-9 per step). */
enum { CLAW_F64 = 0, CLAW_F32 = 1, CLAW_F64_FMA = 2 };

typedef struct {
    int32_t mx, my;      /* interior cells                                     */
    int32_t ng;          /* ghost layers stored in q/aux (>= 2 classic, 3 SharpClaw) */
    int32_t pad_;
    int64_t ldq;         /* row pitch in elements (>= mx + 2 ng)               */
    int64_t plane;       /* component pitch in elements (>= ldq * (my + 2 ng)) */
    double dx, dy;       /* cell sizes (Dimension.delta, geometry.py:425-427)  */
} ClawGrid;

typedef struct {
    int32_t order;        /* 1 Godunov, 2 with limited corrections (method(2))  */
    int32_t transverse;   /* 0 none, 1 trans_inc, 2 trans_cor (method(3))       */
    int32_t split;        /* 0 unsplit step2, 1 dimensional splitting step2ds   */
    int32_t fwave;        /* must be 0: no f-wave solver is provided            */
    int32_t num_waves;
    int32_t mthlim[8];    /* limiter per wave family (philim ids; 0 = none)     */
} ClawMethod;

typedef struct {
    int32_t kind;         /* CLAW_RP_*                                          */
    int32_t efix;         /* Euler: entropy fix on (reference default)          */
    double p[8];          /* solver parameters, see CLAW_RP_*                   */
} ClawRiemann;

typedef struct {
    int32_t lower[2], upper[2];          /* q: per dimension (x, y)            */
    int32_t aux_lower[2], aux_upper[2];  /* aux                                */
} ClawBC;

/* Context: owns scratch buffers (split intermediate state), the words through
   which a step kernel combines its wavefronts' Courant numbers (allocated and
   zeroed on the device by claw_ctx_create; every launch leaves them at zero)
   and launch choices.  One context per device; calls on one context are
   stream-ordered and must not overlap on different host threads, nor run
   concurrently on different streams.  A context created where its device is
   absent serves argument checks only: its launches return CLAW_EDEVICE. */
int claw_ctx_create(int device, void **ctx);
int claw_ctx_destroy(void *ctx);

/* No counterpart in the reference.  The unsplit classic kernels address each
   component plane through a 32-bit per-lane offset from the plane's base when
   the plane is under 4 GiB, and through 64-bit offsets otherwise; on = 1 makes
   this context use the 64-bit kernels for every grid (the results are the
   same; for tests of that path on small grids).  on = 0 restores the default. */
int claw_ctx_set_wide_offsets(void *ctx, int on);

/* One classic step (unsplit or split) from qold into qnew (distinct buffers).
   Replaces ClawSolver2D.step_hyperbolic's step2/step2ds calls
   (src/pyclaw/classic/solver.py:498-547, step2.f90, step2ds.f90).
   dt_dev: device double, the step size.  cfl_dev: device double, receives the
   Courant number of the step (max over the same interfaces as the reference);
   it is overwritten.  precision: CLAW_F64 (reference arithmetic), CLAW_F64_FMA
   or CLAW_F32.
   aux may be NULL when the solver uses no aux fields. */
int claw2d_classic_step(void *ctx, const ClawGrid *grid, const ClawMethod *method,
                        const ClawRiemann *rp, const ClawBC *bc, int precision,
                        const void *qold, void *qnew, const void *aux,
                        const double *dt_dev, double *cfl_dev, void *stream);

/* Same, with flags.  CLAW_STEP_ACCUMULATE_CFL: do not reset *cfl_dev; the
   step's Courant number is max-combined into it (one launch per unsplit step,
   which lets a caller time the step kernel alone or run several steps and
   read one maximum). */
enum { CLAW_STEP_ACCUMULATE_CFL = 1 };
int claw2d_classic_step_ex(void *ctx, const ClawGrid *grid, const ClawMethod *method,
                           const ClawRiemann *rp, const ClawBC *bc, int precision,
                           const void *qold, void *qnew, const void *aux,
                           const double *dt_dev, double *cfl_dev, int flags, void *stream);

/* claw2d_classic_step without any device allocation: no counterpart in the
   reference.  For callers that cannot allocate device memory (code under HIP
   graph capture, tight serving loops, harnesses that forbid hipMalloc).
   Arguments and results as claw2d_classic_step.  The split step's
   intermediate state, which claw2d_classic_step keeps in memory the context
   allocates, lives here in a fixed device array of the library
   (CLAW_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends): a split
   step needs num_eqn * grid->plane * sizeof(real) bytes of it and returns
   CLAW_ETOOLARGE beyond that; unsplit steps use none.  The current device
   must be the context's device.  Every call on one GPU shares that GPU's
   array: calls that may overlap on one GPU (different streams, threads)
   race and must be serialised by the caller. */
#define CLAW_NOALLOC_SCRATCH_BYTES (32u << 20)
int claw2d_classic_step_noalloc(void *ctx, const ClawGrid *grid, const ClawMethod *method,
                                const ClawRiemann *rp, const ClawBC *bc, int precision,
                                const void *qold, void *qnew, const void *aux,
                                const double *dt_dev, double *cfl_dev, void *stream);

/* Autotuner of the unsplit step's launch shape: times the fused kernel for a
   set of row-band heights (the rows one wavefront marches through; more
   bands mean more wavefronts in flight, fewer mean fewer halo rows; the
   candidates include the height that makes one band per wavefront slot of
   the device), each with the kernel compiled for one and for two
   wavefronts per SIMD and with blocks of one or four wavefronts, and the
   paired fill decomposition where it tiles the grid (band 0 below), and
   keeps the fastest for this context and problem shape (Riemann solver,
   order, transverse mode, precision, mx, my); later classic steps of that
   shape use it.  With CLAW_F64 and CLAW_F32 results do not depend on the
   launch shape (every cell is computed the same way and the Courant maximum
   is order-free), so tuning never changes answers; with CLAW_F64_FMA the
   march runs two rows per loop trip and, for the light solvers, a band's
   first row in a copy of its own, and the compiler's contractions can round
   these copies differently, so answers can move in the last bits with the
   band starts (Euler and acoustics: equal for every shape whose bands start
   on even rows, the paired fill decomposition included).  Runs real steps
   from qold into qnew and
   max-combines into *cfl_dev.  Blocks until the timing launches finish.
   *band_out (optional) receives the chosen height, negated when the
   two-waves-per-SIMD build won; 0 for the paired fill decomposition and
   for the split step (not tuned).
   No counterpart in the reference (its step has no launch shape). */
int claw2d_classic_autotune(void *ctx, const ClawGrid *grid, const ClawMethod *method,
                            const ClawRiemann *rp, const ClawBC *bc, int precision,
                            const void *qold, void *qnew, const void *aux,
                            const double *dt_dev, double *cfl_dev, void *stream, int *band_out);

/* Sets the launch shape the autotuner would store: row-band height `band`
   (>= 1) and the build for `waves_per_simd` (1 or 2) wavefronts per SIMD,
   for classic steps of this problem shape on this context.  For callers
   that tune once and reuse the shape.  No counterpart in the reference. */
int claw2d_classic_set_launch(void *ctx, const ClawGrid *grid, const ClawMethod *method,
                              const ClawRiemann *rp, int precision, int band, int waves_per_simd);

/* Same, also setting the wavefronts per block (1, 2 or 4; the shape set by
   claw2d_classic_set_launch uses 4).  Band 0 with waves_per_simd 2 and
   waves_per_block 1 selects the paired fill decomposition: one-wavefront
   blocks, one per wavefront slot of the device; the two wavefronts that
   share a SIMD split their rows about 2 : 1, the larger part to the one
   dispatched first, which the SIMD's issue arbitration favours; every
   piece lies within one column strip.  Where that cannot tile the grid
   (fewer pairs than strips, or units under a row) the step runs the
   default bands.  The default shape for FP64 grids of at least 16384
   strip-rows (strips of 60 columns).  No counterpart in the reference. */
int claw2d_classic_set_launch_ex(void *ctx, const ClawGrid *grid, const ClawMethod *method,
                                 const ClawRiemann *rp, int precision, int band, int waves_per_simd,
                                 int waves_per_block);

/* The launch shape classic steps of this problem shape use on this context
   (tuned, set, or the default; band 0 is the paired fill decomposition).
   No counterpart in the reference. */
int claw2d_classic_get_launch(void *ctx, const ClawGrid *grid, const ClawMethod *method,
                              const ClawRiemann *rp, int precision, int *band, int *waves_per_simd,
                              int *waves_per_block);

/* Number of kernel launches one classic step issues (for timing). */
int claw2d_classic_launches(const ClawMethod *method);

/* ---- SharpClaw (WENO5 component-wise, lim_type=2, char_decomp=0) ---------- */

/* dq = dt * dq/dt over the interior, with ng >= 3 ghost layers.  Replaces
   sharpclaw2.flux2 (src/pyclaw/sharpclaw/flux2.f90:2-101, called from
   SharpClawSolver2D.dq_hyperbolic, sharpclaw/solver.py:535-582).  dq uses the
   same padded layout as q; its ghost cells are not written.  *cfl_dev is
   overwritten with the Courant number of the evaluation. */
int claw2d_sharpclaw_dq(void *ctx, const ClawGrid *grid, const ClawRiemann *rp, const ClawBC *bc,
                        int precision, const void *q, void *dq, const void *aux,
                        const double *dt_dev, double *cfl_dev, void *stream);

/* One Runge-Kutta stage: evaluates dq on qin and writes the stage
   combination (sharpclaw/solver.py:201-251) into out:
     0 out = dq                    1 out = qin + dq   (Euler; SSP33 stage 1)
     2 out = qin + dq/6.           3 t = qin + dq/6.; s2 = q0/25. + (9./25)*t; out = 15.*s2 - 5.*t
     4 out = (s2 + 0.6*qin) + 0.1*dq
     5 out = 0.75*q0 + 0.25*(qin + dq)        6 out = 1./3.*q0 + 2./3.*(qin + dq)
   q0 is the step's initial state (modes 3, 5, 6); s2 is written by mode 3
   and read by mode 4.  *cfl_dev is overwritten. */
int claw2d_sharpclaw_stage(void *ctx, const ClawGrid *grid, const ClawRiemann *rp, const ClawBC *bc,
                           int precision, int mode, const void *qin, const void *q0, void *s2,
                           void *out, const void *aux, const double *dt_dev, double *cfl_dev,
                           void *stream);

/* A whole SSP104 step (10 stages, sharpclaw/solver.py:221-251) from q into
   qnew, with s1a/s1b/s2 as stage registers (all five buffers distinct).
   cfl10_dev receives the 10 stages' Courant numbers; the caller applies the
   reference's rule (the first stage above cfl_max aborts the step, else the
   last stage's value is the step's).  Ghost codes CLAW_BC_STORED are read
   from each stage register, so callers with user boundary functions fill
   them per stage with claw2d_sharpclaw_stage instead. */
int claw2d_sharpclaw_ssp104_step(void *ctx, const ClawGrid *grid, const ClawRiemann *rp,
                                 const ClawBC *bc, int precision, const void *q, void *qnew,
                                 void *s1a, void *s1b, void *s2, const void *aux,
                                 const double *dt_dev, double *cfl10_dev, void *stream);

/* Autotuner of the SharpClaw stage kernel's row-band height (the same idea as
   claw2d_classic_autotune): times dq evaluations of q into dq for heights
   4..64 and the fill decomposition (band height 0, below), each with the
   kernel built for the default occupancy and for three wavefronts per SIMD,
   and keeps the fastest for this context and shape; every later SharpClaw
   call of that shape (dq, stage, SSP104 step) uses it.  Results do not
   depend on it.  Overwrites dq and *cfl_dev.  *band_out (optional) receives
   the height, negated when the three-wave build won.  No counterpart in the
   reference. */
int claw2d_sharpclaw_autotune(void *ctx, const ClawGrid *grid, const ClawRiemann *rp, const ClawBC *bc,
                              int precision, const void *q, void *dq, const void *aux,
                              const double *dt_dev, double *cfl_dev, void *stream, int *band_out);

/* Sets the SharpClaw launch shape directly: row-band height `band` (>= 1) and
   the build for the default occupancy (waves_per_simd 1) or for three
   wavefronts per SIMD (3).  band = 0 is the fill decomposition: one
   wavefront per slot of the build (CUs x the wavefronts per CU that the
   runtime's occupancy query gives for the kernel), each taking an equal
   share of the grid's rows strip by strip (the default for large grids, with
   the default build in one-wavefront blocks).  No counterpart in the
   reference. */
int claw2d_sharpclaw_set_launch(void *ctx, const ClawGrid *grid, const ClawRiemann *rp, int precision,
                                int band, int waves_per_simd);

/* Same, also setting the wavefronts per block (1, 2 or 4; the shape set by
   claw2d_sharpclaw_set_launch uses 4).  No counterpart in the reference. */
int claw2d_sharpclaw_set_launch_ex(void *ctx, const ClawGrid *grid, const ClawRiemann *rp, int precision,
                                   int band, int waves_per_simd, int waves_per_block);

/* The launch shape SharpClaw stages of this problem shape use on this
   context (tuned, set, or the default).  No counterpart in the reference. */
int claw2d_sharpclaw_get_launch(void *ctx, const ClawGrid *grid, const ClawRiemann *rp, int precision,
                                int *band, int *waves_per_simd, int *waves_per_block);

/* ---- Batched runs of many independent grids ---------------------------------
   No counterpart in the reference (which runs one grid per process); this is
   the batching of the build's plan: instances of different sizes advance
   together, one persistent work-queue launch per step, with the reference's
   per-grid accept/reject and dt rule (solver.py:651-687) applied on the device.
   All instances share the method, Riemann solver and boundary codes. */
typedef struct {
    int32_t mx, my;
    int64_t q_offset[2];    /* element offsets in q_pool of the instance's two state buffers */
    int64_t aux_offset;     /* element offset in aux_pool, or -1 without aux                */
    int64_t ldq, plane;     /* planar padded layout of this instance (see above)           */
    double dx, dy;
} ClawBatchInstance;

/* Builds the work list (instances sorted largest first, one item per
   60-column strip x 32-row band; blocks of four wavefronts) and device
   descriptors. */
int claw2d_batch_plan_create(void *ctx, const ClawBatchInstance *inst, int n, int ng,
                             const ClawMethod *method, const ClawRiemann *rp, const ClawBC *bc,
                             int precision, void **plan);
int claw2d_batch_plan_destroy(void *plan);
int claw2d_batch_work_items(void *plan);

/* Tile shape of the plan's work items: rows per band (>= 1) and wavefronts
   per block of the persistent launch (1, 2 or 4).  Rebuilds the work list
   (waits for the device first) and drops a captured step graph.  Results
   do not depend on the shape.  No counterpart in the reference. */
int claw2d_batch_set_shape(void *plan, int rows, int waves_per_block);

/* Autotuner of the batched step: times one step for band heights
   {8,16,32,64,128} x {1,2,4} wavefronts per block and keeps the fastest.
   The timed steps write the instances' other buffers (parity is not
   flipped), so the state, dt, t and the counters are unchanged; blocks
   until done.  rows_out / waves_per_block_out (optional) receive the
   choice.  No counterpart in the reference. */
int claw2d_batch_autotune(void *plan, void *q_pool, const void *aux_pool, const double *dt,
                          const int *parity, void *stream, int *rows_out, int *waves_per_block_out);

/* nsteps unsplit steps of every instance.  Per-instance device arrays (length
   n): dt (in/out), t (in/out), cflmax (in/out), parity (0: state in buffer 0,
   1: buffer 1; flipped on each accepted step), naccepted, nrejected (counts).
   use_graph != 0 captures the whole sequence into a hipGraph (cached for
   repeated calls with the same arguments) and launches it on `stream`. */
int claw2d_batch_steps(void *plan, void *q_pool, const void *aux_pool, int nsteps,
                       double cfl_max, double cfl_desired, double dt_max, double *dt, double *t,
                       double *cflmax, int32_t *parity, int32_t *naccepted, int32_t *nrejected,
                       int use_graph, void *stream);

/* The same per-step bookkeeping for n grids stepped by other entry points:
   reads cfl[k] (then zeroes it), accepts if cfl <= cfl_max (t += dt, parity
   flips, naccepted++), else nrejected++; dt = min(dt_max, dt*cfl_desired/cfl). */
int claw2d_step_update(int n, double cfl_max, double cfl_desired, double dt_max, double *dt,
                       double *cfl, double *t, double *cflmax, int32_t *parity,
                       int32_t *naccepted, int32_t *nrejected, void *stream);

/* ---- evolve_to_time's step loop on the device -------------------------------
   Replaces the host loop of src/pyclaw/solver.py:629-696 (dt clipped to tend,
   accept if cfl <= cfl_max else restore q, dt = min(dt_max,
   dt*cfl_desired/cfl), t += dt, status counters) for runs with no per-step
   host work, so that many steps are queued between host synchronisations.
   The state is double-buffered: iteration k reads cur = states[k%2] and
   cfl[k%2] and writes next = states[(k+1)%2] and cfl[(k+1)%2] = 0; its step
   reads dt from &cur->dt and accumulates into cfl[k%2].  Iterations after the
   loop's end (t >= tend, or max_steps iterations) leave every field as is. */
typedef struct {
    double cfl_max, cfl_desired, dt_max, tend, tstart;
    int32_t max_steps;
    int32_t ncfl;   /* Courant numbers per step (0 or 1; 10 for the SSP104 stages, of which the
                       first above cfl_max decides, else the last: sharpclaw/solver.py:221-251) */
    int32_t dt_variable;   /* 0: fixed dt, t = tstart + n*dt, and a step above cfl_max ends the
                              loop with failed = 1, its result kept (solver.py:655-677) */
    int32_t pad;
} ClawLoopRule;

typedef struct {
    double dt;           /* the next step's dt (clipped to tend) */
    double t;            /* solution time */
    double cflmax, dtmin, dtmax;   /* status['cflmax'/'dtmin'/'dtmax'] */
    double cfl_last;     /* Courant number of the last iteration */
    int32_t naccepted;   /* status['numsteps'] */
    int32_t niter;       /* iterations taken (accepted + rejected) */
    int32_t done;        /* 1 once t >= tend or niter == max_steps (or failed) */
    int32_t failed;      /* fixed dt only: the last step's Courant number was above cfl_max */
} ClawLoopState;

/* Gauges recorded by the loop entry points below (solver.py:713-742 with the
   default compute_gauge_values, FP64): after each accepted iteration,
   vals[row*n + k] = q[offsets[k]] of the new state and times[row] = t, with
   row = naccepted - 1 - base (rows outside [0, cap) are not written). */
typedef struct {
    const long long *offsets;
    double *vals;
    double *times;
    int32_t n, base, cap, pad;
} ClawLoopGauges;

/* One iteration's bookkeeping after the step qold -> qnew: qnew := qold
   (nbytes, the whole padded buffer) when the step is rejected or the loop has
   ended, so the state is always in qnew afterwards.  cfl_cur and cfl_next
   hold ncfl values each. */
int claw2d_loop_advance(const ClawLoopRule *rule, const ClawLoopState *cur, ClawLoopState *next,
                        const double *cfl_cur, double *cfl_next, const void *qold, void *qnew,
                        long long nbytes, void *stream);

/* niter iterations of that loop for the classic step, launched from here:
   iteration i steps q[(k0+i)%2] -> q[(k0+i+1)%2] (q = {qa, qb}) with
   claw2d_classic_step_ex(..., CLAW_STEP_ACCUMULATE_CFL) reading dt from
   states[(k0+i)%2] and accumulating into cfl[(k0+i)%2], then
   claw2d_loop_advance, then (gauges != NULL, FP64 only) the gauge record.
   states: two ClawLoopState, cfl: two doubles, device memory.  No host
   synchronisation. */
int claw2d_classic_loop(void *ctx, const ClawGrid *grid, const ClawMethod *meth, const ClawRiemann *rp,
                        const ClawBC *bc, int precision, void *qa, void *qb, const void *aux,
                        const ClawLoopRule *rule, ClawLoopState *states, double *cfl, long long nbytes, int k0,
                        int niter, const ClawLoopGauges *gauges, void *stream);

/* The same for SharpClaw: whole SSP104 steps (claw2d_sharpclaw_ssp104_step,
   stage registers s1a/s1b/s2) with rule->ncfl = 10 and cfl holding 2 x 10
   doubles (iteration i uses cfl + 10*((k0+i)%2)). */
int claw2d_sharpclaw_loop(void *ctx, const ClawGrid *grid, const ClawRiemann *rp, const ClawBC *bc,
                          int precision, void *qa, void *qb, void *s1a, void *s1b, void *s2, const void *aux,
                          const ClawLoopRule *rule, ClawLoopState *states, double *cfl, long long nbytes, int k0,
                          int niter, const ClawLoopGauges *gauges, void *stream);

/* ---- Discrete adjoint of the classic step (variable-coefficient acoustics) ---
   No counterpart in the reference, which has no adjoint.  The forward map is
   claw2d_classic_step with CLAW_RP_VC_ACOUSTICS (aux = (rho, c), c > 0 in
   every cell), unsplit, order 1 or 2 with any limiter, transverse 0/1/2, at a
   fixed dt, with boundary codes extrap, periodic or wall for q and aux
   (stored ghost cells are not differentiated: CLAW_EINVAL). */

/* Bytes of device workspace claw2d_vc_adjoint_step needs for this grid. */
long long claw2d_vc_adjoint_workspace_bytes(const ClawGrid *grid);

/* lam_out = (d q_{n+1} / d q_n)^T lam_in, and gaux += (d q_{n+1} / d aux)^T lam_in,
   with q = q_n (the state the forward step started from).  lam_in is read at
   interior cells only; lam_out is written at interior cells; gaux is
   accumulated at interior cells (component 0: rho, 1: c).  lam_in != lam_out.
   Deterministic (fixed summation order, no atomics). */
int claw2d_vc_adjoint_step(const ClawGrid *grid, const ClawMethod *method, const ClawBC *bc,
                           const double *q, const double *aux, const double *lam_in, double *lam_out,
                           double *gaux, double dt, void *workspace, long long workspace_bytes,
                           void *stream);

/* The same step followed by the receiver-misfit source of the previous step:
   lam_out[offsets[r]] += traces[r] - observed[r] for r < nrec (component 0
   offsets into the padded layout; receivers must be distinct interior cells). */
int claw2d_vc_adjoint_step_misfit(const ClawGrid *grid, const ClawMethod *method, const ClawBC *bc,
                                  const double *q, const double *aux, const double *lam_in,
                                  double *lam_out, double *gaux, double dt, void *workspace,
                                  long long workspace_bytes, const long long *offsets, int nrec,
                                  const double *traces, const double *observed, void *stream);

/* out[k] = a[offsets[k]] (receiver traces). */
int claw2d_gather(const double *a, const long long *offsets, int n, double *out, void *stream);

/* a[offsets[k]] += traces[k] - observed[k] (distinct offsets). */
int claw2d_misfit_source(double *a, const long long *offsets, int n, const double *traces,
                         const double *observed, void *stream);

/* ---- Source terms as device operations --------------------------------------
   Replace the host source routines of the Euler shock-bubble application,
   apps/euler_2d/shockbubble.py:89-124 (step_Euler_radial, set as
   solver.step_source, classic/solver.py:131-149) and :127-153
   (dq_Euler_radial, set as solver.dq_src, sharpclaw/solver.py:279-280):
   geometric source terms of the Euler equations with radial symmetry, radius
   = aux component 0.  Interior cells only, q in the padded planar layout of
   grid; num_eqn >= 4 (components 4.. are a passive tracer); gamma1 = gamma - 1.
   FP64 results equal the reference's bit for bit. */

/* One 2-stage Runge-Kutta step of length dt, in place on q. */
int claw2d_euler_radial_step(const ClawGrid *grid, int precision, int num_eqn, void *q, const void *aux,
                             double dt, double gamma1, void *stream);

/* dq = dt * source(q) (components >= 4 zero); dq has q's layout and is a
   different buffer. */
int claw2d_euler_radial_dq(const ClawGrid *grid, int precision, int num_eqn, const void *q, const void *aux,
                           void *dq, double dt, double gamma1, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* CLAWHIP_H */
