/*
 * C/OpenMP restatement of the reference Wilson operator and CG, for the CPU
 * baseline timing.  TEST INFRASTRUCTURE ONLY (checked against the NumPy
 * oracle in tests/test_oracle_c.py); never linked into the product.
 *
 * Follows jaxqft/fermions/wilson.py:228-256 (sign = -1: D, +1: D^dag),
 * the diagonal m + 4r (wilson.py:134-143), the boundary phase on the
 * x_mu = L_mu - 1 link slice (su3_wilson_nf2.py:726-759) and the
 * jax.scipy.sparse.linalg.cg iteration used at su3_wilson_nf2.py:1243-1269.
 * Arrays are the reference layouts: psi (L0,L1,L2,L3,4,nc) complex128,
 * U BMXYIJ (4,L0,L1,L2,L3,nc,nc) complex128 (U(1): (4,L0..L3)).
 */
#include <complex.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* gamma tables of gamma.py:14-46 in 4D: (g_mu psi)_s = ph[mu][s] psi[perm[mu][s]] */
static const int PERM[4][4] = {{3, 2, 1, 0}, {3, 2, 1, 0}, {2, 3, 0, 1}, {0, 1, 2, 3}};
static const cplx PH[4][4] = {{1, 1, 1, 1}, {-I, I, -I, I}, {-I, -I, I, I}, {1, 1, -1, -1}};

static void apply_one(const cplx* U, const cplx* psi, cplx* out, const int* L, int nc, double mass,
                      const cplx* bc, int sign) {
  const long V = (long)L[0] * L[1] * L[2] * L[3];
  const long str[4] = {(long)L[1] * L[2] * L[3], (long)L[2] * L[3], L[3], 1};
  const int ns_nc = 4 * nc;
#pragma omp parallel for schedule(static)
  for (long s = 0; s < V; ++s) {
    int x[4];
    long rem = s;
    for (int mu = 3; mu >= 0; --mu) {
      x[mu] = (int)(rem % L[mu]);
      rem /= L[mu];
    }
    cplx acc[12];
    for (int c = 0; c < ns_nc; ++c) acc[c] = 0;
    for (int mu = 0; mu < 4; ++mu) {
      long sf = (x[mu] == L[mu] - 1) ? s - (long)(L[mu] - 1) * str[mu] : s + str[mu];
      long sb = (x[mu] == 0) ? s + (long)(L[mu] - 1) * str[mu] : s - str[mu];
      cplx pf = (x[mu] == L[mu] - 1) ? bc[mu] : 1.0;
      cplx pb = (x[mu] == 0) ? conj(bc[mu]) : 1.0;
      const cplx* Uf = U + ((long)mu * V + s) * nc * nc;
      const cplx* Ub = U + ((long)mu * V + sb) * nc * nc;
      const cplx* qf = psi + sf * ns_nc;
      const cplx* qb = psi + sb * ns_nc;
      cplx uf[12], ub[12];
      /* colour multiply on all four spin rows: U psi(x+mu), U(x-mu)^dag psi(x-mu) */
      for (int sp = 0; sp < 4; ++sp)
        for (int a = 0; a < nc; ++a) {
          cplx f = 0, b = 0;
          for (int bb = 0; bb < nc; ++bb) {
            f += Uf[a * nc + bb] * qf[sp * nc + bb];
            b += conj(Ub[bb * nc + a]) * qb[sp * nc + bb];
          }
          uf[sp * nc + a] = pf * f;
          ub[sp * nc + a] = pb * b;
        }
      /* (r + sign g) on the forward hop, (r - sign g) on the backward hop, r = 1 */
      for (int sp = 0; sp < 4; ++sp)
        for (int a = 0; a < nc; ++a) {
          cplx gf = PH[mu][sp] * uf[PERM[mu][sp] * nc + a];
          cplx gb = PH[mu][sp] * ub[PERM[mu][sp] * nc + a];
          cplx fwd = uf[sp * nc + a] + (double)sign * gf;
          cplx bwd = ub[sp * nc + a] - (double)sign * gb;
          acc[sp * nc + a] -= 0.5 * (fwd + bwd);
        }
    }
    const double m0 = mass + 4.0;
    for (int c = 0; c < ns_nc; ++c) out[s * ns_nc + c] = m0 * psi[s * ns_nc + c] + acc[c];
  }
}

void wo_apply(const double* U, const double* psi, double* out, const int* L, int nc, double mass,
              const double* bc_re, const double* bc_im, int sign) {
  cplx bc[4];
  for (int i = 0; i < 4; ++i) bc[i] = bc_re[i] + I * bc_im[i];
  apply_one((const cplx*)U, (const cplx*)psi, (cplx*)out, L, nc, mass, bc, sign);
}

static double vdot_re(const cplx* a, const cplx* b, long n) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (long i = 0; i < n; ++i) s += creal(conj(a[i]) * b[i]);
  return s;
}

/* Runs CG on D^dag D x = b (x0 = 0) for at most maxiter iterations or until
 * |r|^2 <= tol^2 |b|^2.  Returns the number of iterations. */
int wo_cg(const double* Ud, const double* bd, double* xd, const int* L, int nc, double mass, const double* bc_re,
          const double* bc_im, double tol, int maxiter) {
  const cplx* U = (const cplx*)Ud;
  const cplx* b = (const cplx*)bd;
  cplx* x = (cplx*)xd;
  cplx bc[4];
  for (int i = 0; i < 4; ++i) bc[i] = bc_re[i] + I * bc_im[i];
  const long n = (long)L[0] * L[1] * L[2] * L[3] * 4 * nc;
  cplx* r = malloc(sizeof(cplx) * n);
  cplx* p = malloc(sizeof(cplx) * n);
  cplx* t = malloc(sizeof(cplx) * n);
  cplx* Ap = malloc(sizeof(cplx) * n);
  memset(x, 0, sizeof(cplx) * n);
  memcpy(r, b, sizeof(cplx) * n);
  memcpy(p, b, sizeof(cplx) * n);
  double rr = vdot_re(r, r, n);
  const double stop = tol * tol * vdot_re(b, b, n);
  int k = 0;
  while (rr > stop && k < maxiter) {
    apply_one(U, p, t, L, nc, mass, bc, -1);
    apply_one(U, t, Ap, L, nc, mass, bc, +1);
    double alpha = rr / vdot_re(p, Ap, n);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
    }
    double rr_new = vdot_re(r, r, n);
    double beta = rr_new / rr;
    rr = rr_new;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
    ++k;
  }
  free(r);
  free(p);
  free(t);
  free(Ap);
  return k;
}
