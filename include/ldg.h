/*
 * ldg.h -- C ABI of the MI355X LDG right-hand-side library (libldg.so).
 *
 * The reference (NoseKnowsAll/Local-DG) has no plugin or FFI layer: its
 * hot path is reached through the Solver class (solver.h:110-113), the
 * array.h column-major layout and the driver.cpp options.  Each entry point
 * below names the reference interface it stands in for.  Plain pointers
 * and sizes only; no C++ types cross this boundary.
 *
 * Layout contract (array.h:17-430, solver.cpp:62, 319-325): a state u is
 * u(iN, iS, iK) column-major, i.e. u[iN + dofs*(iS + 5*iK)], with
 * iN = ix + n*iy + n*n*iz on ascending Chebyshev-Lobatto nodes (n = p+1),
 * iS in (rho, rho u, rho v, rho w, rho E), iK = ex + nx*ey + nx*ny*ez.
 * A gradient Du is Dus(iN, iS, iK, l), i.e. Du[iN + dofs*(iS + 5*(iK + K*l))].
 *
 * Errors: every int-returning call returns LDG_OK (0) or a negative code;
 * ldg_last_error() gives the message.  No C++ exception crosses the ABI.
 * Threading: calls on one handle must be serialised by the caller; separate
 * handles are independent.  Every call is complete when it returns.
 */
#ifndef LDG_H
#define LDG_H

#ifdef __cplusplus
extern "C" {
#endif

#define LDG_OK 0
#define LDG_EINVAL -1       /* bad argument or configuration */
#define LDG_ENOMEM -2       /* device allocation failed */
#define LDG_EDEVICE -3      /* HIP runtime error, or no device */
#define LDG_ENONFINITE -4   /* the state or a diagnostic became non-finite */

#define LDG_FP64 0          /* plain FP64 path */
#define LDG_MIXED 1         /* FP32 fluxes/quadrature, FP64 state and RK */

typedef struct ldg_handle ldg_handle;

/* Mesh, physics and precision of one solver instance.
   Mesh: Mesh(nx, ny, nz, botLeft, topRight, mpi) (mesh.cpp:51);
   physics: the constants set in Solver::Solver (solver.cpp:30-43);
   order p: Solver(p, ...) (solver.cpp:7). */
typedef struct ldg_config {
  int p;                  /* polynomial order, 1..5 */
  int nx, ny, nz;         /* elements per direction, each >= 3 (mesh.cpp:102-107) */
  double xmin[3];         /* botLeft */
  double xmax[3];         /* topRight */
  double L;               /* length scale L (solver.cpp:37) */
  double Re, Pr, gamma, M0, rho0, V0, R;
  int precision;          /* LDG_FP64 or LDG_MIXED */
} ldg_config;

/* Fills the reference defaults: p=2, 10^3 elements on [-pi L, pi L]^3,
   Re=1600, Pr=0.71, gamma=1.4, M0=0.1, rho0=V0=L=1, R=8.314, FP64
   (driver.cpp:26-39, solver.cpp:30-36). */
int ldg_config_default(ldg_config *cfg);

/* Solver::Solver(p, dtSnaps, tf, L, mesh) (solver.cpp:7-71): builds the
   element operators and allocates the device state. */
int ldg_create(const ldg_config *cfg, ldg_handle **out);

/* Releases device memory (Solver's destructor). */
void ldg_destroy(ldg_handle *h);

/* Last error message of the handle (or, when h == NULL, of the last failed
   call without a handle: ldg_create, ldg_ensemble_create, ldg_rhs_noalloc). */
const char *ldg_last_error(const ldg_handle *h);

/* dofs = (p+1)^3 and the element count K (Solver::dofs, mesh.nElements). */
int ldg_sizes(const ldg_handle *h, int *dofs, int *nelem);

/* Taylor-Green state at the nodes on the device
   (Solver::initialCondition, solver.cpp:194-236). */
int ldg_initial_condition(ldg_handle *h);

/* Copy a host state in (out) in the layout above (the reference's
   Solver::u member, solver.cpp:62). */
int ldg_set_state(ldg_handle *h, const double *u);
int ldg_get_state(ldg_handle *h, double *u);

/* Device address of the resident state (for callers that keep data in HBM). */
int ldg_state_device_ptr(ldg_handle *h, double **dptr);

/* The reference time-step rule (solver.cpp:47): cfl*min(dx,dy,dz)/(V0 (2p+1));
   the reference uses cfl = 0.01. */
double ldg_time_step(const ldg_handle *h, double cfl);

/* One right-hand-side evaluation k = rhs(u) (Solver::rk4Rhs,
   solver.cpp:486-535).  u, k are host arrays (dofs,5,K); Du, if not NULL,
   receives the auxiliary gradient Dus (dofs,5,K,3). */
int ldg_rhs(ldg_handle *h, const double *u, double *k, double *Du);

/* Device scratch of ldg_rhs_noalloc per GPU, in bytes. */
#define LDG_NOALLOC_SCRATCH_BYTES (64u << 20)

/* k = rhs(u) exactly as ldg_rhs computes it, for callers that cannot
   allocate device memory: code under HIP graph capture, tight serving
   loops, harnesses that forbid hipMalloc.  No counterpart in the reference
   (its rk4Rhs, solver.cpp:486-535, allocates scratch on every call).
   u and k are DEVICE arrays (dofs,5,K) in the layout above; the mesh,
   physics and precision come from cfg; there is no handle.  The operators,
   instance tables and face fluxes live in one fixed device array of
   LDG_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends), so the
   mesh must satisfy 8192 + 30 w K (p+1)^2 <= LDG_NOALLOC_SCRATCH_BYTES with
   w = 8 (LDG_FP64) or 4 (LDG_MIXED) bytes -- e.g. up to 22^3 elements at
   p = 4 in FP64 -- and a larger one returns LDG_EINVAL.  Runs on `stream`
   (a hipStream_t; NULL: the default stream) and is complete when it
   returns; calls no hipMalloc/hipFree.  All calls on one GPU share its
   array: calls that overlap in time on one GPU (several threads or
   streams) race, and must be serialised by the caller. */
int ldg_rhs_noalloc(const ldg_config *cfg, const double *u, double *k, void *stream);

/* nsteps RK4 steps of size dt on the resident state (the step loop of
   Solver::dgTimeStep, solver.cpp:347-442).  If ke != NULL it receives the
   reference kinetic energy after each step (kes(iStep), solver.cpp:437). */
int ldg_step(ldg_handle *h, double dt, int nsteps, double *ke);

/* Diagnostics of the resident state: out[0] = reference kinetic energy
   (Solver::computeKE, solver.cpp:1007-1034); out[1] = kinetic energy and
   out[2] = enstrophy by Gauss quadrature of the interpolated state and LDG
   gradient (no reference counterpart).  There is no separate kinetic-energy
   or enstrophy entry point: the reference's per-step KE record (kes,
   solver.cpp:437) is ldg_step's ke output, and this call covers the rest. */
int ldg_diagnostics(ldg_handle *h, double out[3]);

/* Average device time per launch of each kernel over nsteps RK4 steps of
   the resident state, measured with HIP events on the handle's stream:
   ms[0] = gradient kernel, ms[1] = right-hand-side kernel.  The state is
   advanced.  No reference counterpart (the reference only prints wall
   time, solver.cpp:344-372). */
int ldg_kernel_times(ldg_handle *h, double dt, int nsteps, double ms[2]);

/* ------------------------------------------------------------------------
 * Ensembles: many independent solver instances advanced by the same kernel
 * launches.  No reference counterpart: the reference runs one case per
 * process (driver.cpp:20-60).  Instances of one order share one launch of
 * each kernel; the domain, Pr, gamma, rho0, R, L and precision come from the
 * base configuration, and each instance has its own order, mesh, Reynolds
 * and Mach numbers, velocity scale and a phase shift of the Taylor-Green
 * field.  The single-instance calls above act on instance 0; ldg_rhs,
 * ldg_step and ldg_kernel_times need a handle of one instance.
 * ------------------------------------------------------------------------ */
typedef struct ldg_instance {
  int p;                  /* polynomial order, 1..5 */
  int nx, ny, nz;         /* elements per direction, each >= 3 */
  double Re, M0, V0;
  double phase[3];        /* initial state = Taylor-Green evaluated at x + phase */
} ldg_instance;

/* n instances with parameters drawn from seed: meshes 8^3, 12^3, 16^3 or
   24^3, p in 2..4, Re log-uniform in [400, 3200], M0 uniform in
   [0.05, 0.30], V0 = 1 + 0.05 N(0,1), phases uniform in [0, 2 pi). */
int ldg_ensemble_sample(unsigned long long seed, int n, ldg_instance *out);

/* Handle of n instances (state zero until set or initialised). */
int ldg_ensemble_create(const ldg_config *base, const ldg_instance *inst, int n, ldg_handle **out);

int ldg_instance_count(const ldg_handle *h, int *n);
int ldg_instance_sizes(const ldg_handle *h, int i, int *p, int *dofs, int *nelem);

/* ldg_time_step's rule for instance i. */
int ldg_instance_time_step(const ldg_handle *h, int i, double cfl, double *dt);

/* State of instance i, layout as for ldg_set_state with that instance's
   dofs and K. */
int ldg_instance_set_state(ldg_handle *h, int i, const double *u);
int ldg_instance_get_state(ldg_handle *h, int i, double *u);

/* ldg_diagnostics for instance i. */
int ldg_instance_diagnostics(ldg_handle *h, int i, double out[3]);

/* nsteps RK4 steps of every instance, instance i with step dt[i].  If ke is
   not NULL, ke[s * n + i] receives instance i's reference kinetic energy
   after step s. */
int ldg_ensemble_step(ldg_handle *h, const double *dt, int nsteps, double *ke);

/* enable != 0: ldg_step / ldg_ensemble_step replay a HIP graph of two RK4
   steps instead of launching the kernels one by one (an odd last step is
   launched eagerly).  Results are identical. */
int ldg_set_graph(ldg_handle *h, int enable);

/* Times the candidate elements-per-workgroup tilings of each kernel on the
   resident state and keeps the fastest per order.  The state is not
   changed.  If chosen is not NULL it receives, per order group in order of
   first appearance, (p, elements per workgroup of the gradient kernel, of
   the right-hand-side kernel): 3 ints per group, at most 5 groups. */
int ldg_autotune(ldg_handle *h, int *chosen);

/* ------------------------------------------------------------------------
 * Learned closure for under-resolved runs.  No reference counterpart: the
 * reference has no closure.  A per-element MLP maps one element's state and
 * LDG gradient (20 fields at the (p+1)^3 nodes, 540 values at p = 2) to a
 * correction of its right-hand side (5 fields, 135 values):
 * in -> hidden -> hidden -> out with SiLU, FP32 weights entering BF16
 * matrix-core products with FP32 accumulation.  It is trained on
 * high-order runs of this solver: target = P rhs_fine(u) - rhs_coarse(P u)
 * with P the element-wise L2 projection to the coarse order.
 * ------------------------------------------------------------------------ */
typedef struct ldg_closure ldg_closure;

/* Closure for order p with `hidden` units per hidden layer and room for
   `capacity` training samples; weights drawn from seed. */
int ldg_closure_create(int p, int hidden, long capacity, unsigned long long seed, ldg_closure **out);
void ldg_closure_destroy(ldg_closure *c);
const char *ldg_closure_last_error(const ldg_closure *c);
int ldg_closure_sizes(const ldg_closure *c, int *nin, int *nout, int *hidden, long *nparam);

/* Parameters (nparam floats: W1, b1, W2, b2, W3, b3, each W row-major
   out x in) and normalisation (2 (nin + nout) floats: input mean, input
   1/std, output mean, output std).  Either pointer may be NULL.  Setting
   parameters resets the optimiser state. */
int ldg_closure_get(ldg_closure *c, float *params, float *norm);
int ldg_closure_set(ldg_closure *c, const float *params, const float *norm);

/* Network on B normalised rows (host arrays): Y (B x nout) = net(X);
   ldg_closure_gradient: the gradient of mean((net(X) - T)^2) w.r.t. the
   parameters into grad (nparam) and the loss. */
int ldg_closure_forward(ldg_closure *c, const float *X, int B, float *Y);
int ldg_closure_gradient(ldg_closure *c, const float *X, const float *T, int B, float *grad, double *loss);

/* L2 projection of instance i's state of `fine` onto instance j of
   `coarse` (same mesh, coarse order <= fine order), on the device. */
int ldg_project_state(ldg_handle *fine, int i, ldg_handle *coarse, int j);

/* Training samples from the resident states of `fine`: instance i of
   `coarse` (the closure's order, same mesh) receives the projected state;
   a fraction `keep` of the elements (chosen from seed) is appended as
   (features, target) pairs.  Overwrites coarse's state. */
int ldg_closure_collect(ldg_closure *c, ldg_handle *fine, ldg_handle *coarse, double keep, unsigned long long seed);
/* Append n samples given as host arrays (raw, not normalised). */
int ldg_closure_add_samples(ldg_closure *c, const float *X, const float *T, long n);
int ldg_closure_samples(const ldg_closure *c, long *n);
int ldg_closure_get_samples(ldg_closure *c, long first, long n, float *X, float *T);

/* Loss weighting, chosen before the normalisation is fitted: mode 0
   (default) weighs every sample alike; mode 1 weighs each collection (one
   ldg_closure_collect or ldg_closure_add_samples call: one snapshot) by the
   inverse of its mean normalised target energy, so that snapshots with small
   corrections count as much as those with large ones.  Weights average 1;
   ldg_closure_get_weights reads them after the fit. */
int ldg_closure_set_weighting(ldg_closure *c, int mode);
int ldg_closure_get_weights(ldg_closure *c, long first, long n, float *w);

/* Per-feature mean/std of the collected samples become the normalisation;
   the dataset is normalised in place (once) and the loss weights are set. */
int ldg_closure_fit_normalization(ldg_closure *c);

/* Adam on minibatches of a seeded random (1 - val_fraction) of the samples;
   losses[2 e] / losses[2 e + 1] = training / hold-out weighted MSE of epoch
   e (normalised units). */
int ldg_closure_train(ldg_closure *c, int epochs, int batch, double lr, double val_fraction, unsigned long long seed,
                      double *losses);

/* Attach c (NULL detaches) to every instance of h with the closure's
   order: each RK stage (and ldg_rhs) adds the correction element by
   element, unless the RMS of the normalised correction exceeds `bound`
   (that element keeps the plain right-hand side; bound 0 switches every
   element off) or, with flags & LDG_CLOSURE_DISSIPATIVE, unless the
   correction would raise the element's nodal kinetic energy, or, with
   flags & LDG_CLOSURE_ENERGIZING, unless it would lower it (not both).
   With flags & LDG_CLOSURE_PER_STEP, ldg_step evaluates the correction
   and the gate once per RK4 step, from the step's initial state in stage
   1, and adds them in all four stages (ldg_rhs always evaluates them).  c
   must outlive the attachment.  ldg_closure_gated: elements switched off in
   the last evaluated stage, of `total`. */
#define LDG_CLOSURE_DISSIPATIVE 1
#define LDG_CLOSURE_ENERGIZING 2
#define LDG_CLOSURE_PER_STEP 4
int ldg_set_closure(ldg_handle *h, const ldg_closure *c, double bound, int flags);
int ldg_closure_gated(ldg_handle *h, long *off, long *total);

#ifdef __cplusplus
}
#endif

#endif /* LDG_H */
