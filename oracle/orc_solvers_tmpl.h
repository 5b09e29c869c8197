/*
 * orc_solvers_tmpl.h -- oracle Krylov solvers, instantiated twice by
 * orc_solvers.c (real double and double _Complex) through the macros
 *   T, SFX, DOT, NRM, DIFF, MUL, ADD, SUB, RDIV, TDIV, TRDIV, FROMR, NEG,
 *   RMUL, RSUB, ZERO, CONVABS, ABSV.
 * TEST INFRASTRUCTURE ONLY (see orc.h).
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)

static T* CAT(valloc_, SFX)(int n)
{
  T* v = (T*)malloc(sizeof(T) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; i++) v[i] = ZERO;
  return v;
}

/* CG: generic_cg.cpp:132-231 (f64), :278-382 (c128).  The complex version
 * keeps alpha complex (rsq / <p,Ap>) and beta complex with zero imaginary part. */
orc_info CAT(orc_cg_, SFX)(T* phi, T* phi0, int size, int max_iter, double eps, MVT mv, void* extra,
                          orc_verb* verb, double* hist)
{
  orc_info inf = {0.0, 0, 0, 0, "CG"};
  T* r = CAT(valloc_, SFX)(size);
  T* p = CAT(valloc_, SFX)(size);
  T* Ap = CAT(valloc_, SFX)(size);
  const double bsqrt = sqrt(NRM(phi0, size));
  mv(p, phi, extra); inf.ops_count++;
  for (int i = 0; i < size; i++) r[i] = SUB(phi0[i], p[i]);
  for (int i = 0; i < size; i++) p[i] = r[i];
  for (int i = 0; i < size; i++) Ap[i] = ZERO;
  mv(Ap, p, extra); inf.ops_count++;
  double rsq = NRM(r, size), rsqNew = 0.0;
  int k;
  for (k = 0; k < max_iter; k++) {
    const T alpha = RDIV(rsq, DOT(p, Ap, size));
    for (int i = 0; i < size; i++) {
      phi[i] = ADD(phi[i], MUL(alpha, p[i]));
      r[i] = SUB(r[i], MUL(alpha, Ap[i]));
    }
    rsqNew = NRM(r, size);
    if (hist) hist[k] = sqrt(rsqNew) / bsqrt;
    orc_print_resid(verb, "CG", k + 1, inf.ops_count, sqrt(rsqNew) / bsqrt);
    if (sqrt(rsqNew) < eps * bsqrt || k == max_iter - 1) break;
    const T beta = FROMR(rsqNew / rsq);
    rsq = rsqNew;
    for (int i = 0; i < size; i++) p[i] = ADD(r[i], MUL(beta, p[i]));
    mv(Ap, p, extra); inf.ops_count++;
  }
  inf.success = (k == max_iter - 1) ? 0 : 1;
  k++;
  mv(Ap, phi, extra); inf.ops_count++;
  const double truersq = DIFF(Ap, phi0, size);
  free(r); free(p); free(Ap);
  orc_print_summary(verb, "CG", inf.success, k, inf.ops_count, sqrt(truersq) / bsqrt);
  inf.resSq = truersq;
  inf.iter = k;
  return inf;
}

/* Restart driver shared by CG(n), BiCGStab(n), GCR(n): generic_cg.cpp:235-275,
 * generic_bicgstab.cpp:162-201, generic_gcr.cpp:160-199 (and the c128 copies). */
typedef orc_info (*CAT(inner_fn_, SFX))(T*, T*, int, int, double, MVT, void*, orc_verb*, double*);
static orc_info CAT(restart_drive_, SFX)(CAT(inner_fn_, SFX) fn, const char* base, T* phi, T* phi0, int size,
                                         int max_iter, double eps, int restart_freq, MVT mv, void* extra,
                                         orc_verb* verb)
{
  const double bsqrt = sqrt(NRM(phi0, size));
  orc_verb inner;
  orc_verb_restart(&inner, verb);
  char name[32];
  snprintf(name, sizeof(name), "%s(%d)", base, restart_freq);
  int iter = 0, ops = 0;
  orc_info inf;
  do {
    inf = fn(phi, phi0, size, restart_freq, eps, mv, extra, verb ? &inner : NULL, NULL);
    if (verb) { verb->len = inner.len; }
    iter += inf.iter;
    ops += inf.ops_count;
    orc_print_restart(verb, name, iter, ops, sqrt(inf.resSq) / bsqrt);
    if (verb) { inner.len = verb->len; }
  } while (iter < max_iter && inf.success == 0 && sqrt(inf.resSq) / bsqrt > eps);
  inf.iter = iter;
  inf.ops_count = ops;
  orc_print_summary(verb, name, inf.success, iter, ops, sqrt(inf.resSq) / bsqrt);
  snprintf(inf.name, sizeof(inf.name), "%s", name);
  inf.success = (sqrt(inf.resSq) / bsqrt > eps) ? 0 : 1;
  return inf;
}

orc_info CAT(orc_cg_restart_, SFX)(T* phi, T* phi0, int size, int max_iter, double eps, int restart_freq, MVT mv,
                                   void* extra, orc_verb* verb)
{
  return CAT(restart_drive_, SFX)(CAT(orc_cg_, SFX), "CG", phi, phi0, size, max_iter, eps, restart_freq, mv, extra, verb);
}

/* BiCGStab: generic_bicgstab.cpp:22-158 (f64), :205-341 (c128). */
orc_info CAT(orc_bicgstab_, SFX)(T* phi, T* phi0, int size, int max_iter, double eps, MVT mv, void* extra,
                                orc_verb* verb, double* hist)
{
  orc_info inf = {0.0, 0, 0, 0, "BiCGStab"};
  T* r = CAT(valloc_, SFX)(size);
  T* r0 = CAT(valloc_, SFX)(size);
  T* p = CAT(valloc_, SFX)(size);
  T* Ap = CAT(valloc_, SFX)(size);
  T* s = CAT(valloc_, SFX)(size);
  T* As = CAT(valloc_, SFX)(size);
  const double bsqrt = sqrt(NRM(phi0, size));
  mv(Ap, phi, extra); inf.ops_count++;
  for (int i = 0; i < size; i++) r[i] = SUB(phi0[i], Ap[i]);
  for (int i = 0; i < size; i++) r0[i] = r[i];
  for (int i = 0; i < size; i++) p[i] = r[i];
  T rho = DOT(r0, r, size);
  mv(Ap, p, extra); inf.ops_count++;
  int k;
  for (k = 0; k < max_iter; k++) {
    const T alpha = TDIV(rho, DOT(r0, Ap, size));
    for (int i = 0; i < size; i++) s[i] = SUB(r[i], MUL(alpha, Ap[i]));
    mv(As, s, extra); inf.ops_count++;
    const T omega = TDIV(DOT(As, s, size), DOT(As, As, size));
    for (int i = 0; i < size; i++) phi[i] = ADD(ADD(phi[i], MUL(alpha, p[i])), MUL(omega, s[i]));
    for (int i = 0; i < size; i++) r[i] = SUB(s[i], MUL(omega, As[i]));
    const double rsq = NRM(r, size);
    if (hist) hist[k] = sqrt(rsq) / bsqrt;
    orc_print_resid(verb, "BiCGStab", k + 1, inf.ops_count, sqrt(rsq) / bsqrt);
    if (sqrt(rsq) < eps * bsqrt || k == max_iter - 1) break;
    const T rhoNew = DOT(r0, r, size);
    const T beta = MUL(TDIV(rhoNew, rho), TDIV(alpha, omega));
    rho = rhoNew;
    for (int i = 0; i < size; i++) p[i] = ADD(r[i], MUL(beta, SUB(p[i], MUL(omega, Ap[i]))));
    for (int i = 0; i < size; i++) Ap[i] = ZERO;
    mv(Ap, p, extra); inf.ops_count++;
  }
  inf.success = (k == max_iter - 1) ? 0 : 1;
  k++;
  mv(Ap, phi, extra); inf.ops_count++;
  const double truersq = DIFF(Ap, phi0, size);
  free(r); free(r0); free(p); free(Ap); free(s); free(As);
  orc_print_summary(verb, "BiCGStab", inf.success, k, inf.ops_count, sqrt(truersq) / bsqrt);
  inf.resSq = truersq;
  inf.iter = k;
  return inf;
}

orc_info CAT(orc_bicgstab_restart_, SFX)(T* phi, T* phi0, int size, int max_iter, double eps, int restart_freq,
                                         MVT mv, void* extra, orc_verb* verb)
{
  return CAT(restart_drive_, SFX)(CAT(orc_bicgstab_, SFX), "BiCGStab", phi, phi0, size, max_iter, eps, restart_freq,
                                  mv, extra, verb);
}

/* GCR (Saad 6.9): generic_gcr.cpp:27-156 (f64), :202-333 (c128).  Every
 * direction (p, Ap) is kept; classical Gram-Schmidt against A r. */
orc_info CAT(orc_gcr_, SFX)(T* phi, T* phi0, int size, int max_iter, double eps, MVT mv, void* extra,
                           orc_verb* verb, double* hist)
{
  orc_info inf = {0.0, 0, 0, 0, "GCR"};
  T* x = CAT(valloc_, SFX)(size);
  T* r = CAT(valloc_, SFX)(size);
  T* Ar = CAT(valloc_, SFX)(size);
  T* p = CAT(valloc_, SFX)(size);
  T* Ap = CAT(valloc_, SFX)(size);
  int cap = 16, nstore = 0;
  T** ps = (T**)malloc(sizeof(T*) * (size_t)cap);
  T** Aps = (T**)malloc(sizeof(T*) * (size_t)cap);
  for (int i = 0; i < size; i++) x[i] = phi[i];
  const double bsqrt = sqrt(NRM(phi0, size));
  mv(p, x, extra); inf.ops_count++;
  for (int i = 0; i < size; i++) r[i] = SUB(phi0[i], p[i]);
  for (int i = 0; i < size; i++) p[i] = r[i];
  mv(Ap, p, extra); inf.ops_count++;
  int k;
  for (k = 0; k < max_iter; k++) {
    if (nstore == cap) {
      cap *= 2;
      ps = (T**)realloc(ps, sizeof(T*) * (size_t)cap);
      Aps = (T**)realloc(Aps, sizeof(T*) * (size_t)cap);
    }
    ps[nstore] = p;
    Aps[nstore] = Ap;
    nstore++;
    const T alpha = TRDIV(DOT(Ap, r, size), NRM(Ap, size));
    for (int i = 0; i < size; i++) {
      x[i] = ADD(x[i], MUL(alpha, p[i]));
      r[i] = SUB(r[i], MUL(alpha, Ap[i]));
    }
    const double rsq = NRM(r, size);
    if (hist) hist[k] = sqrt(rsq) / bsqrt;
    orc_print_resid(verb, "GCR", k + 1, inf.ops_count, sqrt(rsq) / bsqrt);
    if (sqrt(rsq) < eps * bsqrt || k == max_iter - 1) break;
    for (int i = 0; i < size; i++) Ar[i] = ZERO;
    mv(Ar, r, extra); inf.ops_count++;
    p = CAT(valloc_, SFX)(size);
    Ap = CAT(valloc_, SFX)(size);
    for (int i = 0; i < size; i++) { p[i] = r[i]; Ap[i] = Ar[i]; }
    for (int j = 0; j <= k; j++) {
      const T beta = TRDIV(NEG(DOT(Aps[j], Ar, size)), NRM(Aps[j], size));
      for (int i = 0; i < size; i++) {
        p[i] = ADD(p[i], MUL(beta, ps[j][i]));
        Ap[i] = ADD(Ap[i], MUL(beta, Aps[j][i]));
      }
    }
  }
  inf.success = (k == max_iter - 1) ? 0 : 1;
  k++;
  for (int i = 0; i < size; i++) p[i] = ZERO;
  mv(p, x, extra); inf.ops_count++;
  const double truersq = DIFF(p, phi0, size);
  for (int i = 0; i < size; i++) phi[i] = x[i];
  free(x); free(r); free(Ar);
  /* p and Ap were pushed into storage on every iteration; when the loop
   * never ran they are still the first buffers */
  if (nstore == 0) { free(p); free(Ap); }
  for (int j = 0; j < nstore; j++) { free(ps[j]); free(Aps[j]); }
  free(ps); free(Aps);
  orc_print_summary(verb, "GCR", inf.success, k, inf.ops_count, sqrt(truersq) / bsqrt);
  inf.resSq = truersq;
  inf.iter = k;
  return inf;
}

orc_info CAT(orc_gcr_restart_, SFX)(T* phi, T* phi0, int size, int max_iter, double eps, int restart_freq, MVT mv,
                                    void* extra, orc_verb* verb)
{
  return CAT(restart_drive_, SFX)(CAT(orc_gcr_, SFX), "GCR", phi, phi0, size, max_iter, eps, restart_freq, mv, extra,
                                  verb);
}

#define SWAPT(a, b) do { T t_ = (a); (a) = (b); (b) = t_; } while (0)
#define SWAPP(a, b) do { T* t_ = (a); (a) = (b); (b) = t_; } while (0)

/* Multishift CG (Jegerlehner, hep-lat/9612014): generic_cg_m.cpp:23-310 (f64),
 * :312-600 (c128).  Zero initial guess; converged shifts are swapped to the
 * tail of the active range every resid_freq_check iterations, permuting the
 * caller's phi[] and shifts[] until the end of the call. */
orc_info CAT(orc_cg_m_, SFX)(T** phi, T* phi0, int n_shift, int size, int resid_freq_check, int max_iter, double eps,
                            double* shifts, MVT mv, void* extra, int worst_first, orc_verb* verb, double* resSqmrhs)
{
  orc_info inf = {0.0, 0, 0, 0, "CG-M"};
  int n_rem = n_shift;
  T* alpha_s = (T*)malloc(sizeof(T) * (size_t)n_shift);
  T* beta_s = (T*)malloc(sizeof(T) * (size_t)n_shift);
  T* zeta_s = (T*)malloc(sizeof(T) * (size_t)n_shift);
  T* zeta_p = (T*)malloc(sizeof(T) * (size_t)n_shift);
  T** p_s = (T**)malloc(sizeof(T*) * (size_t)n_shift);
  int* mapping = (int*)malloc(sizeof(int) * (size_t)n_shift);
  for (int n = 0; n < n_shift; n++) p_s[n] = CAT(valloc_, SFX)(size);
  T* r = CAT(valloc_, SFX)(size);
  T* p = CAT(valloc_, SFX)(size);
  T* Ap = CAT(valloc_, SFX)(size);
  for (int n = 0; n < n_shift; n++) mapping[n] = n;
  double rsq = 0.0, rsqNew = 0.0;
  for (int n = 0; n < n_shift; n++) {
    beta_s[n] = zeta_s[n] = zeta_p[n] = FROMR(1.0);
    alpha_s[n] = FROMR(0.0);
  }
  T beta = FROMR(1.0), alpha = FROMR(0.0), beta_prev;
  const double bsqrt = sqrt(NRM(phi0, size));
  for (int n = 0; n < n_shift; n++)
    for (int i = 0; i < size; i++) { p_s[n][i] = phi0[i]; phi[n][i] = ZERO; }
  for (int i = 0; i < size; i++) { p[i] = phi0[i]; r[i] = phi0[i]; }
  mv(Ap, p, extra); inf.ops_count++;
  rsq = NRM(r, size);
  int k;
  for (k = 0; k < max_iter; k++) {
    beta_prev = beta;
    beta = RDIV(-rsq, DOT(p, Ap, size));
    for (int n = 0; n < n_rem; n++) {
      const T tmp = zeta_s[n];
      const T num = MUL(MUL(zeta_s[n], zeta_p[n]), beta_prev);
      const T den = ADD(MUL(MUL(beta, alpha), SUB(zeta_p[n], zeta_s[n])),
                        MUL(MUL(zeta_p[n], beta_prev), RSUB(1.0, RMUL(shifts[n], beta))));
      zeta_s[n] = TDIV(num, den);
      zeta_p[n] = tmp;
      beta_s[n] = TDIV(MUL(beta, zeta_s[n]), zeta_p[n]);
      for (int i = 0; i < size; i++) phi[n][i] = SUB(phi[n][i], MUL(beta_s[n], p_s[n][i]));
    }
    for (int i = 0; i < size; i++) r[i] = ADD(r[i], MUL(beta, Ap[i]));
    rsqNew = NRM(r, size);
    orc_print_resid(verb, "CG-M", k + 1, inf.ops_count, sqrt(rsqNew) / bsqrt);
    if (k % resid_freq_check == 0) {
      for (int n = 0; n < n_rem; n++) {
        if (CONVABS(zeta_s[n]) * sqrt(rsqNew) < eps * bsqrt) {
          n_rem--;
          if (n_rem != n) {
            int ti = mapping[n_rem]; mapping[n_rem] = mapping[n]; mapping[n] = ti;
            SWAPP(phi[n_rem], phi[n]);
            SWAPP(p_s[n_rem], p_s[n]);
            SWAPT(alpha_s[n_rem], alpha_s[n]);
            SWAPT(beta_s[n_rem], beta_s[n]);
            SWAPT(zeta_s[n_rem], zeta_s[n]);
            SWAPT(zeta_p[n_rem], zeta_p[n]);
            double td = shifts[n_rem]; shifts[n_rem] = shifts[n]; shifts[n] = td;
            n--;
          }
        }
      }
    }
    if ((worst_first && ABSV(zeta_s[0]) * sqrt(rsqNew) < eps * bsqrt) || n_rem == 0 || k == max_iter - 1) break;
    alpha = FROMR(rsqNew / rsq);
    rsq = rsqNew;
    for (int n = 0; n < n_rem; n++) {
      alpha_s[n] = TDIV(MUL(MUL(alpha, zeta_s[n]), beta_s[n]), MUL(zeta_p[n], beta));
      for (int i = 0; i < size; i++) p_s[n][i] = ADD(MUL(zeta_s[n], r[i]), MUL(alpha_s[n], p_s[n][i]));
    }
    for (int i = 0; i < size; i++) p[i] = ADD(r[i], MUL(alpha, p[i]));
    mv(Ap, p, extra); inf.ops_count++;
  }
  inf.success = (k == max_iter - 1) ? 0 : 1;
  k++;
  for (int n = 0; n < n_shift; n++) {
    if (mapping[n] != n) {
      for (int m = n + 1; m < n_shift; m++) {
        if (mapping[m] == n) {
          SWAPP(phi[m], phi[n]);
          double td = shifts[m]; shifts[m] = shifts[n]; shifts[n] = td;
          mapping[m] = mapping[n];
          mapping[n] = n;
          n--;
          break;
        }
      }
    }
  }
  double* relres = (double*)malloc(sizeof(double) * (size_t)n_shift);
  for (int n = 0; n < n_shift; n++) {
    for (int i = 0; i < size; i++) Ap[i] = ZERO;
    mv(Ap, phi[n], extra); inf.ops_count++;
    for (int i = 0; i < size; i++) Ap[i] = ADD(Ap[i], RMUL(shifts[n], phi[n][i]));
    const double rs = DIFF(Ap, phi0, size);
    if (resSqmrhs) resSqmrhs[n] = rs;
    relres[n] = sqrt(rs) / bsqrt;
  }
  free(r); free(p); free(Ap);
  for (int n = 0; n < n_shift; n++) free(p_s[n]);
  free(p_s); free(alpha_s); free(beta_s); free(zeta_s); free(zeta_p); free(mapping);
  orc_print_summary_multi(verb, inf.success, k, inf.ops_count, relres, n_shift);
  free(relres);
  inf.resSq = 0.0; /* the reference never assigns it (generic_cg_m.cpp:306) */
  inf.iter = k;
  return inf;
}

#undef SWAPT
#undef SWAPP
#undef CAT
#undef CAT2
