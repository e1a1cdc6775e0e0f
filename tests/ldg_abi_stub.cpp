// ldg_abi_stub.cpp -- a CPU test double of the libldg.so entry points that
// the C++ host (local-dg_amd/host) calls, so that the host logic (driver
// options, wall-clock limit, KE history, checkpoints) runs under the address
// and undefined-behaviour sanitizers on a machine without a GPU.  No
// numerics: a "step" adds 1 to every state entry and records KE = step
// count, which makes every host-side bookkeeping error visible.
#include <cmath>
#include <cstring>
#include <vector>

#include "ldg.h"

struct ldg_handle {
  ldg_config cfg;
  int dofs = 0, K = 0;
  long steps = 0;
  std::vector<double> u;
};

extern "C" {

int ldg_config_default(ldg_config *c) {
  std::memset(c, 0, sizeof(*c));
  c->p = 2; c->nx = c->ny = c->nz = 10;
  for (int d = 0; d < 3; ++d) { c->xmin[d] = -M_PI; c->xmax[d] = M_PI; }
  c->L = 1; c->Re = 1600; c->Pr = 0.71; c->gamma = 1.4; c->M0 = 0.1; c->rho0 = 1; c->V0 = 1; c->R = 8.314;
  return LDG_OK;
}

int ldg_create(const ldg_config *c, ldg_handle **out) {
  if (!c || !out || c->p < 1 || c->nx < 3) return LDG_EINVAL;
  ldg_handle *h = new ldg_handle;
  h->cfg = *c;
  h->dofs = (c->p + 1) * (c->p + 1) * (c->p + 1);
  h->K = c->nx * c->ny * c->nz;
  h->u.assign((size_t)h->dofs * 5 * h->K, 0.0);
  *out = h;
  return LDG_OK;
}

void ldg_destroy(ldg_handle *h) { delete h; }
const char *ldg_last_error(const ldg_handle *) { return "stub"; }

int ldg_sizes(const ldg_handle *h, int *dofs, int *nelem) {
  *dofs = h->dofs;
  *nelem = h->K;
  return LDG_OK;
}

double ldg_time_step(const ldg_handle *h, double cfl) {
  const double dx = (h->cfg.xmax[0] - h->cfg.xmin[0]) / h->cfg.nx;
  return cfl * dx / (h->cfg.V0 * (2 * h->cfg.p + 1));
}

int ldg_initial_condition(ldg_handle *h) {
  for (size_t i = 0; i < h->u.size(); ++i) h->u[i] = (double)(i % 7);
  return LDG_OK;
}

int ldg_set_state(ldg_handle *h, const double *u) {
  std::memcpy(h->u.data(), u, h->u.size() * sizeof(double));
  return LDG_OK;
}

int ldg_get_state(ldg_handle *h, double *u) {
  std::memcpy(u, h->u.data(), h->u.size() * sizeof(double));
  return LDG_OK;
}

int ldg_step(ldg_handle *h, double dt, int nsteps, double *ke) {
  if (!(dt > 0) || nsteps < 0) return LDG_EINVAL;
  for (int s = 0; s < nsteps; ++s) {
    for (double &v : h->u) v += 1.0;
    ++h->steps;
    if (ke) ke[s] = (double)h->steps;   // writes exactly nsteps entries
  }
  return LDG_OK;
}

}  // extern "C"
