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

/* Last error message of the handle (or of the last failed create when
   h == NULL). */
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

/* nsteps RK4 steps of size dt on the resident state (the step loop of
   Solver::dgTimeStep, solver.cpp:347-442).  If ke != NULL it receives the
   reference kinetic energy after each step (kes(iStep), solver.cpp:437). */
int ldg_step(ldg_handle *h, double dt, int nsteps, double *ke);

/* Diagnostics of the resident state: out[0] = reference kinetic energy
   (Solver::computeKE, solver.cpp:1007-1034); out[1] = kinetic energy and
   out[2] = enstrophy by Gauss quadrature of the interpolated state and LDG
   gradient (no reference counterpart). */
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

#ifdef __cplusplus
}
#endif

#endif /* LDG_H */
