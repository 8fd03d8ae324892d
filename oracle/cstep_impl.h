/* Body of the C oracle stepper, included twice by cstep.c with REAL set to
 * double and to float.  TEST INFRASTRUCTURE ONLY.
 *
 * Restates rfx's per-step update in the reference's sub-step order
 * (rfx/simulation.py:1256-1681): H update (rfx/core/yee.py:265-291, graded
 * :429-451), CPML-H (rfx/boundaries/cpml.py:804-1065), E update
 * (yee.py:379-404, graded :470-492), CPML-E (cpml.py:522-801), outer-face PEC
 * (rfx/boundaries/pec.py:13-48), PEC mask (pec.py:179-208), then the record
 * operations in the lane's order.  Expressions follow oracle/fdtd_np.py term
 * for term so the two agree to rounding. */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUFFIX)

typedef struct {
  REAL *f[6];        /* ex ey ez hx hy hz */
  REAL *ca, *cb, *ch, *ce_p, *ch_p;
  REAL *ca3[3], *cb3[3];  /* per component (subpixel smoothing), else ca, cb */
  REAL *psi[3][2][2]; /* [axis][side][term]: H psi */
  REAL *psie[3][2][2];/* E psi */
} FN(state_t);

static inline REAL FN(at)(const REAL *a, const oracle_desc *d, long i, long j, long k) {
  if (i < 0 || j < 0 || k < 0 || i >= d->nx || j >= d->ny || k >= d->nz) return (REAL)0;
  return a[(i * d->ny + j) * d->nz + k];
}

static inline REAL FN(dfwd)(const REAL *a, const oracle_desc *d, long i, long j, long k, int ax) {
  REAL c = a[(i * d->ny + j) * d->nz + k];
  REAL nb = FN(at)(a, d, i + (ax == 0), j + (ax == 1), k + (ax == 2));
  if (!d->graded) return (nb - c) / (REAL)d->dx;
  long q = ax == 0 ? i : ax == 1 ? j : k;
  return (nb - c) * (REAL)d->inv_h[ax][q];
}

static inline REAL FN(dbwd)(const REAL *a, const oracle_desc *d, long i, long j, long k, int ax) {
  REAL c = a[(i * d->ny + j) * d->nz + k];
  REAL nb = FN(at)(a, d, i - (ax == 0), j - (ax == 1), k - (ax == 2));
  if (!d->graded) return (c - nb) / (REAL)d->dx;
  long q = ax == 0 ? i : ax == 1 ? j : k;
  return (c - nb) * (REAL)d->inv_e[ax][q];
}

/* CPML term table (field, source field, sign) per axis: cpml.py:875-1043 (H)
 * and :608-777 (E). Field ids: 0..2 = ex..ez, 3..5 = hx..hz. */
static const int FN(h_terms)[3][2][2] = {{{4, 2}, {5, 1}}, {{3, 2}, {5, 0}}, {{3, 1}, {4, 0}}};
static const REAL FN(h_sign)[3][2] = {{1, -1}, {-1, 1}, {1, -1}};
static const int FN(e_terms)[3][2][2] = {{{1, 5}, {2, 4}}, {{0, 5}, {2, 3}}, {{0, 4}, {1, 3}}};
static const REAL FN(e_sign)[3][2] = {{-1, 1}, {1, -1}, {-1, 1}};

static void FN(cpml)(FN(state_t) *s, const oracle_desc *d, int is_h) {
  long N[3] = {d->nx, d->ny, d->nz};
  long n = d->cpml_n;
  for (int ax = 0; ax < 3; ++ax) {
    if (!(d->cpml_axes & (1 << ax))) continue;
    for (int side = 0; side < 2; ++side) {
      long lo = side == 0 ? 0 : N[ax] - n;
      const float *pb = d->prof_b[ax][side], *pc = d->prof_c[ax][side], *pk = d->prof_k[ax][side];
      REAL dd = (REAL)d->face_d[ax][side];
      for (int t = 0; t < 2; ++t) {
        int fld = is_h ? FN(h_terms)[ax][t][0] : FN(e_terms)[ax][t][0];
        int src = is_h ? FN(h_terms)[ax][t][1] : FN(e_terms)[ax][t][1];
        REAL sign = is_h ? FN(h_sign)[ax][t] : FN(e_sign)[ax][t];
        REAL *psi = is_h ? s->psi[ax][side][t] : s->psie[ax][side][t];
        REAL *coef = is_h ? s->ch_p : s->ce_p;
        long e0[3] = {0, 0, 0}, e1[3] = {N[0], N[1], N[2]};
        e0[ax] = lo;
        e1[ax] = lo + n;
        long sy = e1[1] - e0[1], sz = e1[2] - e0[2];
#pragma omp parallel for schedule(static)
        for (long i = e0[0]; i < e1[0]; ++i)
          for (long j = e0[1]; j < e1[1]; ++j)
            for (long k = e0[2]; k < e1[2]; ++k) {
              long l = ax == 0 ? i - lo : ax == 1 ? j - lo : k - lo;
              long c = (i * d->ny + j) * d->nz + k;
              long p = ((i - e0[0]) * sy + (j - e0[1])) * sz + (k - e0[2]);
              const REAL *a = s->f[src];
              REAL g;
              if (is_h) {
                REAL nb = FN(at)(a, d, i + (ax == 0), j + (ax == 1), k + (ax == 2));
                g = (nb - a[c]) / dd;
              } else {
                REAL nb = FN(at)(a, d, i - (ax == 0), j - (ax == 1), k - (ax == 2));
                g = (a[c] - nb) / dd;
              }
              REAL b = (REAL)pb[l], cc_ = (REAL)pc[l], kk = (REAL)pk[l];
              REAL ps = b * psi[p] + cc_ * g;
              psi[p] = ps;
              REAL cf = coef[c];
              REAL u = s->f[fld][c];
              u = u + sign * cf * ps;
              u = u + sign * cf * ((REAL)1 / kk - (REAL)1) * g;
              s->f[fld][c] = u;
            }
      }
    }
  }
}

static void FN(step_fields)(FN(state_t) *s, const oracle_desc *d) {
  long nx = d->nx, ny = d->ny, nz = d->nz;
  REAL *ex = s->f[0], *ey = s->f[1], *ez = s->f[2], *hx = s->f[3], *hy = s->f[4], *hz = s->f[5];
#pragma omp parallel for schedule(static)
  for (long i = 0; i < nx; ++i)
    for (long j = 0; j < ny; ++j)
      for (long k = 0; k < nz; ++k) {
        long c = (i * ny + j) * nz + k;
        REAL cx = FN(dfwd)(ez, d, i, j, k, 1) - FN(dfwd)(ey, d, i, j, k, 2);
        REAL cy = FN(dfwd)(ex, d, i, j, k, 2) - FN(dfwd)(ez, d, i, j, k, 0);
        REAL cz = FN(dfwd)(ey, d, i, j, k, 0) - FN(dfwd)(ex, d, i, j, k, 1);
        hx[c] = hx[c] - s->ch[c] * cx;
        hy[c] = hy[c] - s->ch[c] * cy;
        hz[c] = hz[c] - s->ch[c] * cz;
      }
  if (d->cpml_n > 0) FN(cpml)(s, d, 1);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < nx; ++i)
    for (long j = 0; j < ny; ++j)
      for (long k = 0; k < nz; ++k) {
        long c = (i * ny + j) * nz + k;
        REAL cx = FN(dbwd)(hz, d, i, j, k, 1) - FN(dbwd)(hy, d, i, j, k, 2);
        REAL cy = FN(dbwd)(hx, d, i, j, k, 2) - FN(dbwd)(hz, d, i, j, k, 0);
        REAL cz = FN(dbwd)(hy, d, i, j, k, 0) - FN(dbwd)(hx, d, i, j, k, 1);
        ex[c] = s->ca3[0][c] * ex[c] + s->cb3[0][c] * cx;
        ey[c] = s->ca3[1][c] * ey[c] + s->cb3[1][c] * cy;
        ez[c] = s->ca3[2][c] * ez[c] + s->cb3[2][c] * cz;
      }
  if (d->cpml_n > 0) FN(cpml)(s, d, 0);
  /* pec.py:13-48 */
  if (d->pec_axes & 1)
    for (long j = 0; j < ny; ++j)
      for (long k = 0; k < nz; ++k) {
        long a = (0 * ny + j) * nz + k, b = ((nx - 1) * ny + j) * nz + k;
        ey[a] = ey[b] = ez[a] = ez[b] = 0;
      }
  if (d->pec_axes & 2)
    for (long i = 0; i < nx; ++i)
      for (long k = 0; k < nz; ++k) {
        long a = (i * ny + 0) * nz + k, b = (i * ny + ny - 1) * nz + k;
        ex[a] = ex[b] = ez[a] = ez[b] = 0;
      }
  if (d->pec_axes & 4)
    for (long i = 0; i < nx; ++i)
      for (long j = 0; j < ny; ++j) {
        long a = (i * ny + j) * nz + 0, b = (i * ny + j) * nz + nz - 1;
        ex[a] = ex[b] = ey[a] = ey[b] = 0;
        ez[b] = 0;
      }
  if (d->pec_bits) {
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nx * ny * nz; ++c) {
      unsigned char m = d->pec_bits[c];
      if (m & 1) ex[c] = ex[c] * (REAL)0;
      if (m & 2) ey[c] = ey[c] * (REAL)0;
      if (m & 4) ez[c] = ez[c] * (REAL)0;
    }
  }
}

static REAL FN(port_i)(FN(state_t) *s, const oracle_desc *d, long cell, int comp, const double *loop) {
  long nz = d->nz, ny = d->ny;
  long i = cell / (ny * nz), j = (cell / nz) % ny, k = cell % nz;
  REAL *hx = s->f[3], *hy = s->f[4], *hz = s->f[5];
  if (comp == 2)
    return (FN(at)(hy, d, i, j, k) - FN(at)(hy, d, i - 1, j, k)) * (REAL)loop[1] -
           (FN(at)(hx, d, i, j, k) - FN(at)(hx, d, i, j - 1, k)) * (REAL)loop[0];
  if (comp == 0)
    return (FN(at)(hz, d, i, j, k) - FN(at)(hz, d, i, j - 1, k)) * (REAL)loop[2] -
           (FN(at)(hy, d, i, j, k) - FN(at)(hy, d, i, j, k - 1)) * (REAL)loop[1];
  return (FN(at)(hx, d, i, j, k) - FN(at)(hx, d, i, j, k - 1)) * (REAL)loop[0] -
         (FN(at)(hz, d, i, j, k) - FN(at)(hz, d, i - 1, j, k)) * (REAL)loop[2];
}

static void FN(planes)(FN(state_t) *s, const oracle_desc *d, double t) {
  long N[3] = {d->nx, d->ny, d->nz};
  for (int q = 0; q < d->n_plane; ++q) {
    int ax = d->plane_axis[q];
    long idx = d->plane_index[q];
    const REAL *fld = s->f[d->plane_comp[q]];
    int a1 = ax == 0 ? 1 : 0, a2 = ax == 2 ? 1 : 2;
    long n1 = N[a1], n2 = N[a2];
    int nf = d->plane_nf[q];
    double *acc = d->plane_acc[q]; /* [f][n1][n2][re,im] */
    for (int f = 0; f < nf; ++f) {
      double w = -2.0 * M_PI * (double)d->plane_freqs[q][f] * t;
      double pr = cos(w) * d->dt, pim = sin(w) * d->dt;
#pragma omp parallel for schedule(static)
      for (long u = 0; u < n1; ++u)
        for (long v = 0; v < n2; ++v) {
          long ii[3];
          ii[ax] = idx;
          ii[a1] = u;
          ii[a2] = v;
          double x = (double)fld[(ii[0] * N[1] + ii[1]) * N[2] + ii[2]];
          long o = ((long)f * n1 * n2 + u * n2 + v) * 2;
          acc[o] += pr * x;
          acc[o + 1] += pim * x;
        }
    }
  }
}

static int FN(run)(const oracle_desc *d, oracle_out *out) {
  long ncell = (long)d->nx * d->ny * d->nz;
  FN(state_t) s;
  memset(&s, 0, sizeof(s));
  for (int c = 0; c < 6; ++c) {
    s.f[c] = (REAL *)calloc(ncell, sizeof(REAL));
    if (!s.f[c]) return -1;
  }
  REAL **co[5] = {&s.ca, &s.cb, &s.ch, &s.ce_p, &s.ch_p};
  for (int c = 0; c < 5; ++c) {
    *co[c] = (REAL *)malloc(ncell * sizeof(REAL));
    if (!*co[c]) return -1;
  }
  REAL dt = (REAL)d->dt;
  for (long c = 0; c < ncell; ++c) {
    REAL eps = (REAL)d->eps_r[c] * (REAL)ORACLE_EPS0;
    REAL sig = (REAL)d->sigma[c];
    REAL mu = (REAL)d->mu_r[c] * (REAL)ORACLE_MU0;
    REAL L = sig * dt / ((REAL)2 * eps);
    s.ca[c] = ((REAL)1 - L) / ((REAL)1 + L);
    s.cb[c] = (dt / eps) / ((REAL)1 + L);
    s.ch[c] = dt / mu;
    s.ce_p[c] = dt / ((REAL)d->eps_r[c] * (REAL)ORACLE_EPS0);
    s.ch_p[c] = dt / ((REAL)d->mu_r[c] * (REAL)ORACLE_MU0);
  }
  /* rfx/core/yee.py:824-901 (update_e_aniso): per-component eps with the
     cell's isotropic sigma; the CPML correction keeps eps_r (ce_p) */
  for (int q = 0; q < 3; ++q) {
    if (!d->aniso_eps[q]) {
      s.ca3[q] = s.ca;
      s.cb3[q] = s.cb;
      continue;
    }
    s.ca3[q] = (REAL *)malloc(ncell * sizeof(REAL));
    s.cb3[q] = (REAL *)malloc(ncell * sizeof(REAL));
    if (!s.ca3[q] || !s.cb3[q]) return -1;
    for (long c = 0; c < ncell; ++c) {
      REAL eps = (REAL)d->aniso_eps[q][c] * (REAL)ORACLE_EPS0;
      REAL L = (REAL)d->sigma[c] * dt / ((REAL)2 * eps);
      s.ca3[q][c] = ((REAL)1 - L) / ((REAL)1 + L);
      s.cb3[q][c] = (dt / eps) / ((REAL)1 + L);
    }
  }
  long N[3] = {d->nx, d->ny, d->nz};
  if (d->cpml_n > 0)
    for (int ax = 0; ax < 3; ++ax)
      for (int side = 0; side < 2; ++side)
        for (int t = 0; t < 2; ++t) {
          long sz = d->cpml_n;
          for (int b = 0; b < 3; ++b)
            if (b != ax) sz *= N[b];
          s.psi[ax][side][t] = (REAL *)calloc(sz, sizeof(REAL));
          s.psie[ax][side][t] = (REAL *)calloc(sz, sizeof(REAL));
        }
  if (d->init_fields)
    for (int c = 0; c < 6; ++c)
      for (long q = 0; q < ncell; ++q) s.f[c][q] = (REAL)d->init_fields[c][q];
  for (long n = 0; n < d->n_steps; ++n) {
    FN(step_fields)(&s, d);
    if (d->record_order == 0) {
      for (int p = 0; p < d->n_port; ++p) {
        out->port_v[p * d->n_steps + n] = -(double)s.f[d->port_comp[p]][d->port_cell[p]] * d->port_dpar[p];
        out->port_i[p * d->n_steps + n] = (double)FN(port_i)(&s, d, d->port_cell[p], d->port_comp[p], d->port_loop + 3 * p);
      }
      for (int q = 0; q < d->n_src; ++q)
        s.f[d->src_comp[q]][d->src_cell[q]] += (REAL)d->src_wave[n * d->n_src + q];
      for (int p = 0; p < d->n_probe; ++p)
        out->ts[n * d->n_probe + p] = (double)s.f[d->probe_comp[p]][d->probe_cell[p]];
      FN(planes)(&s, d, (double)(n + 1) * d->dt);
    } else {
      for (int q = 0; q < d->n_src; ++q)
        s.f[d->src_comp[q]][d->src_cell[q]] += (REAL)d->src_wave[n * d->n_src + q];
      for (int p = 0; p < d->n_port; ++p) {
        out->port_v[p * d->n_steps + n] = -(double)s.f[d->port_comp[p]][d->port_cell[p]] * d->port_dpar[p];
        out->port_i[p * d->n_steps + n] = (double)FN(port_i)(&s, d, d->port_cell[p], d->port_comp[p], d->port_loop + 3 * p);
      }
      FN(planes)(&s, d, (double)n * d->dt);
      for (int p = 0; p < d->n_probe; ++p)
        out->ts[n * d->n_probe + p] = (double)s.f[d->probe_comp[p]][d->probe_cell[p]];
    }
  }
  if (out->fields)
    for (int c = 0; c < 6; ++c)
      for (long q = 0; q < ncell; ++q) out->fields[c][q] = (double)s.f[c][q];
  for (int c = 0; c < 6; ++c) free(s.f[c]);
  for (int q = 0; q < 3; ++q)
    if (d->aniso_eps[q]) {
      free(s.ca3[q]);
      free(s.cb3[q]);
    }
  for (int c = 0; c < 5; ++c) free(*co[c]);
  if (d->cpml_n > 0)
    for (int ax = 0; ax < 3; ++ax)
      for (int side = 0; side < 2; ++side)
        for (int t = 0; t < 2; ++t) {
          free(s.psi[ax][side][t]);
          free(s.psie[ax][side][t]);
        }
  return 0;
}

#undef FN
#undef CAT
#undef CAT2
