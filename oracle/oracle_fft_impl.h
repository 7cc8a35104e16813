/*
 * oracle_fft_impl.h — fast mixed-radix FFT body, included by oracle.c once per
 * precision with ORA_REAL / ORA_SFX defined.  TEST INFRASTRUCTURE (checker and
 * CPU baseline only).
 *
 * Algorithm: Stockham autosort, stage s with radix R and prior product p:
 *   for i < N/R: k = i mod p, v[r] = x[i + r N/R] * w_{pR}^{r k},
 *                V = DFT_R(v), y[(i-k)R + k + r p] = V[r].
 * Prime factors above 61 go through Bluestein with a 2^a 3^b 5^c convolution.
 * The inverse is conj(FFT(conj(x))), unnormalised.
 */

typedef struct ORA_SFX(fplan_s) {
  int64_t n;
  int nst;
  int radix[64];
  int64_t twoff[64];
  ORA_REAL *tw;      /* stage twiddles, complex interleaved */
  ORA_REAL *roots[64]; /* per-stage w_R^m tables for generic radices */
  int bluestein;
  int64_t m;
  struct ORA_SFX(fplan_s) *sub;
  ORA_REAL *chirp; /* n complex, exp(-i pi k^2 / n) */
  ORA_REAL *bhat;  /* m complex, FFT_m(b) / m */
} ORA_SFX(fplan);

static void ORA_SFX(fplan_free)(ORA_SFX(fplan) *pl) {
  if (!pl) return;
  free(pl->tw);
  for (int s = 0; s < pl->nst; ++s) free(pl->roots[s]);
  if (pl->sub) ORA_SFX(fplan_free)(pl->sub);
  free(pl->chirp);
  free(pl->bhat);
  free(pl);
}

static int64_t ORA_SFX(smooth235_at_least)(int64_t v) {
  int64_t best = 1;
  while (best < v) best <<= 1;
  for (int64_t a = 1; a < 2 * v; a *= 2)
    for (int64_t b = a; b < 2 * v; b *= 3)
      for (int64_t c = b; c < 2 * v; c *= 5)
        if (c >= v && c < best) best = c;
  return best;
}

static void ORA_SFX(fplan_exec_row)(const ORA_SFX(fplan) *pl, const ORA_REAL *x, ORA_REAL *y, ORA_REAL *scratch);

static ORA_SFX(fplan) *ORA_SFX(fplan_make)(int64_t n) {
  ORA_SFX(fplan) *pl = (ORA_SFX(fplan) *)calloc(1, sizeof(ORA_SFX(fplan)));
  pl->n = n;
  int64_t rest = n;
  int nst = 0;
  while (rest % 4 == 0) { pl->radix[nst++] = 4; rest /= 4; }
  while (rest % 2 == 0) { pl->radix[nst++] = 2; rest /= 2; }
  for (int64_t f = 3; f <= 61; f += 2)
    while (rest % f == 0) { pl->radix[nst++] = (int)f; rest /= f; }
  if (rest > 1) {
    /* a large prime factor remains: Bluestein over the whole length */
    pl->bluestein = 1;
    pl->nst = 0;
    int64_t m = ORA_SFX(smooth235_at_least)(2 * n - 1);
    pl->m = m;
    pl->sub = ORA_SFX(fplan_make)(m);
    pl->chirp = (ORA_REAL *)malloc(sizeof(ORA_REAL) * 2 * (size_t)n);
    ORA_REAL *b = (ORA_REAL *)calloc(2 * (size_t)m, sizeof(ORA_REAL));
    for (int64_t k = 0; k < n; ++k) {
      int64_t r = (int64_t)(((__int128)k * k) % (2 * n));
      long double c, s;
      unit_root_ld(r, 2 * n, &c, &s);
      pl->chirp[2 * k] = (ORA_REAL)c;
      pl->chirp[2 * k + 1] = (ORA_REAL)(-s);
      b[2 * k] = (ORA_REAL)c;
      b[2 * k + 1] = (ORA_REAL)s;
      if (k) {
        b[2 * (m - k)] = (ORA_REAL)c;
        b[2 * (m - k) + 1] = (ORA_REAL)s;
      }
    }
    pl->bhat = (ORA_REAL *)malloc(sizeof(ORA_REAL) * 2 * (size_t)m);
    ORA_REAL *scratch = (ORA_REAL *)malloc(sizeof(ORA_REAL) * 4 * (size_t)m);
    ORA_SFX(fplan_exec_row)(pl->sub, b, pl->bhat, scratch);
    for (int64_t j = 0; j < 2 * m; ++j) pl->bhat[j] /= (ORA_REAL)m;
    free(scratch);
    free(b);
    return pl;
  }
  pl->nst = nst;
  int64_t total = 0, p = 1;
  for (int s = 0; s < nst; ++s) {
    pl->twoff[s] = total;
    total += p * pl->radix[s];
    p *= pl->radix[s];
  }
  pl->tw = (ORA_REAL *)malloc(sizeof(ORA_REAL) * 2 * (size_t)(total > 0 ? total : 1));
  p = 1;
  for (int s = 0; s < nst; ++s) {
    int R = pl->radix[s];
    for (int64_t k = 0; k < p; ++k)
      for (int r = 0; r < R; ++r) {
        long double c, sn;
        unit_root_ld((int64_t)r * k, p * R, &c, &sn);
        pl->tw[2 * (pl->twoff[s] + k * R + r)] = (ORA_REAL)c;
        pl->tw[2 * (pl->twoff[s] + k * R + r) + 1] = (ORA_REAL)(-sn);
      }
    if (R > 5 && R != 4) {
      pl->roots[s] = (ORA_REAL *)malloc(sizeof(ORA_REAL) * 2 * (size_t)R);
      for (int q = 0; q < R; ++q) {
        long double c, sn;
        unit_root_ld(q, R, &c, &sn);
        pl->roots[s][2 * q] = (ORA_REAL)c;
        pl->roots[s][2 * q + 1] = (ORA_REAL)(-sn);
      }
    }
    p *= R;
  }
  return pl;
}

/* Forward radix-R DFT of v (complex interleaved, in place). */
static void ORA_SFX(dft_small)(ORA_REAL *v, int R, const ORA_REAL *roots, ORA_REAL *tmp) {
  if (R == 2) {
    ORA_REAL ar = v[0], ai = v[1], br = v[2], bi = v[3];
    v[0] = ar + br; v[1] = ai + bi; v[2] = ar - br; v[3] = ai - bi;
    return;
  }
  if (R == 4) {
    ORA_REAL s0r = v[0] + v[4], s0i = v[1] + v[5], d0r = v[0] - v[4], d0i = v[1] - v[5];
    ORA_REAL s1r = v[2] + v[6], s1i = v[3] + v[7], d1r = v[2] - v[6], d1i = v[3] - v[7];
    v[0] = s0r + s1r; v[1] = s0i + s1i;
    v[4] = s0r - s1r; v[5] = s0i - s1i;
    /* X1 = d0 - i d1, X3 = d0 + i d1 */
    v[2] = d0r + d1i; v[3] = d0i - d1r;
    v[6] = d0r - d1i; v[7] = d0i + d1r;
    return;
  }
  if (R == 3) {
    const ORA_REAL c1 = (ORA_REAL)-0.5L, s1 = (ORA_REAL)0.866025403784438646763723170752936183L;
    ORA_REAL sr = v[2] + v[4], si = v[3] + v[5], dr = v[2] - v[4], di = v[3] - v[5];
    ORA_REAL x0r = v[0], x0i = v[1];
    v[0] = x0r + sr; v[1] = x0i + si;
    ORA_REAL tr = x0r + c1 * sr, ti = x0i + c1 * si;
    /* X1 = t - i s1 d, X2 = t + i s1 d */
    v[2] = tr + s1 * di; v[3] = ti - s1 * dr;
    v[4] = tr - s1 * di; v[5] = ti + s1 * dr;
    return;
  }
  if (R == 5) {
    const ORA_REAL c1 = (ORA_REAL)0.309016994374947424102293417182819059L;
    const ORA_REAL c2 = (ORA_REAL)-0.809016994374947424102293417182819059L;
    const ORA_REAL s1 = (ORA_REAL)0.951056516295153572116439333379382143L;
    const ORA_REAL s2 = (ORA_REAL)0.587785252292473129168705954639072769L;
    ORA_REAL a1r = v[2] + v[8], a1i = v[3] + v[9], b1r = v[2] - v[8], b1i = v[3] - v[9];
    ORA_REAL a2r = v[4] + v[6], a2i = v[5] + v[7], b2r = v[4] - v[6], b2i = v[5] - v[7];
    ORA_REAL x0r = v[0], x0i = v[1];
    v[0] = x0r + a1r + a2r; v[1] = x0i + a1i + a2i;
    ORA_REAL t1r = x0r + c1 * a1r + c2 * a2r, t1i = x0i + c1 * a1i + c2 * a2i;
    ORA_REAL t2r = x0r + c2 * a1r + c1 * a2r, t2i = x0i + c2 * a1i + c1 * a2i;
    ORA_REAL u1r = s1 * b1r + s2 * b2r, u1i = s1 * b1i + s2 * b2i;
    ORA_REAL u2r = s2 * b1r - s1 * b2r, u2i = s2 * b1i - s1 * b2i;
    /* X1 = t1 - i u1, X4 = t1 + i u1, X2 = t2 - i u2, X3 = t2 + i u2 */
    v[2] = t1r + u1i; v[3] = t1i - u1r;
    v[8] = t1r - u1i; v[9] = t1i + u1r;
    v[4] = t2r + u2i; v[5] = t2i - u2r;
    v[6] = t2r - u2i; v[7] = t2i + u2r;
    return;
  }
  for (int q = 0; q < R; ++q) {
    ORA_REAL sr = 0, si = 0;
    int idx = 0;
    for (int r = 0; r < R; ++r) {
      ORA_REAL wr = roots[2 * idx], wi = roots[2 * idx + 1];
      sr += v[2 * r] * wr - v[2 * r + 1] * wi;
      si += v[2 * r] * wi + v[2 * r + 1] * wr;
      idx += q;
      if (idx >= R) idx -= R;
    }
    tmp[2 * q] = sr;
    tmp[2 * q + 1] = si;
  }
  memcpy(v, tmp, sizeof(ORA_REAL) * 2 * (size_t)R);
}

/* Forward transform of one row; scratch holds >= 2n complex (4m for Bluestein). */
static void ORA_SFX(fplan_exec_row)(const ORA_SFX(fplan) *pl, const ORA_REAL *x, ORA_REAL *y, ORA_REAL *scratch) {
  const int64_t n = pl->n;
  if (pl->bluestein) {
    const int64_t m = pl->m;
    ORA_REAL *a = scratch, *sub_scratch = scratch + 2 * m, *ah = y; /* y too short: use a separate area */
    ORA_REAL *tmp = (ORA_REAL *)malloc(sizeof(ORA_REAL) * 2 * (size_t)m);
    ORA_REAL *sub2 = (ORA_REAL *)malloc(sizeof(ORA_REAL) * 4 * (size_t)m);
    (void)ah;
    for (int64_t j = 0; j < m; ++j) {
      if (j < n) {
        ORA_REAL xr = x[2 * j], xi = x[2 * j + 1], cr = pl->chirp[2 * j], ci = pl->chirp[2 * j + 1];
        a[2 * j] = xr * cr - xi * ci;
        a[2 * j + 1] = xr * ci + xi * cr;
      } else {
        a[2 * j] = 0;
        a[2 * j + 1] = 0;
      }
    }
    ORA_SFX(fplan_exec_row)(pl->sub, a, tmp, sub2);
    /* pointwise with bhat, conj for the inverse via forward transform */
    for (int64_t j = 0; j < m; ++j) {
      ORA_REAL ar = tmp[2 * j], ai = tmp[2 * j + 1], br = pl->bhat[2 * j], bi = pl->bhat[2 * j + 1];
      a[2 * j] = ar * br - ai * bi;
      a[2 * j + 1] = -(ar * bi + ai * br);
    }
    ORA_SFX(fplan_exec_row)(pl->sub, a, tmp, sub2);
    for (int64_t k = 0; k < n; ++k) {
      ORA_REAL wr = tmp[2 * k], wi = -tmp[2 * k + 1], cr = pl->chirp[2 * k], ci = pl->chirp[2 * k + 1];
      y[2 * k] = wr * cr - wi * ci;
      y[2 * k + 1] = wr * ci + wi * cr;
    }
    (void)sub_scratch;
    free(tmp);
    free(sub2);
    return;
  }
  if (pl->nst == 0) { /* n == 1 */
    y[0] = x[0];
    y[1] = x[1];
    return;
  }
  ORA_REAL *bufs[2] = {scratch, scratch + 2 * n};
  const ORA_REAL *src = x;
  ORA_REAL v[2 * 64], tmp[2 * 64];
  int64_t p = 1;
  for (int s = 0; s < pl->nst; ++s) {
    const int R = pl->radix[s];
    const int64_t T = n / R;
    ORA_REAL *dst = (s == pl->nst - 1) ? y : bufs[s & 1];
    const ORA_REAL *tw = pl->tw + 2 * pl->twoff[s];
    for (int64_t i = 0; i < T; ++i) {
      const int64_t k = i % p;
      for (int r = 0; r < R; ++r) {
        ORA_REAL xr = src[2 * (i + r * T)], xi = src[2 * (i + r * T) + 1];
        if (p > 1 && r > 0) {
          ORA_REAL wr = tw[2 * (k * R + r)], wi = tw[2 * (k * R + r) + 1];
          v[2 * r] = xr * wr - xi * wi;
          v[2 * r + 1] = xr * wi + xi * wr;
        } else {
          v[2 * r] = xr;
          v[2 * r + 1] = xi;
        }
      }
      ORA_SFX(dft_small)(v, R, pl->roots[s], tmp);
      const int64_t j = (i - k) * R + k;
      for (int r = 0; r < R; ++r) {
        dst[2 * (j + r * p)] = v[2 * r];
        dst[2 * (j + r * p) + 1] = v[2 * r + 1];
      }
    }
    src = dst;
    p *= R;
  }
}

typedef struct {
  const ORA_SFX(fplan) *pl;
  const ORA_REAL *in;
  ORA_REAL *out;
  int sign;
} ORA_SFX(c2c_ctx);

static void ORA_SFX(c2c_rows)(void *vctx, int64_t begin, int64_t end) {
  ORA_SFX(c2c_ctx) *c = (ORA_SFX(c2c_ctx) *)vctx;
  const int64_t n = c->pl->n;
  const int64_t scratch_len = c->pl->bluestein ? 4 * c->pl->m : 4 * n;
  ORA_REAL *scratch = (ORA_REAL *)malloc(sizeof(ORA_REAL) * (size_t)scratch_len);
  ORA_REAL *row = (ORA_REAL *)malloc(sizeof(ORA_REAL) * 2 * (size_t)n);
  ORA_REAL *res = (ORA_REAL *)malloc(sizeof(ORA_REAL) * 2 * (size_t)n);
  for (int64_t b = begin; b < end; ++b) {
    const ORA_REAL *x = c->in + 2 * b * n;
    for (int64_t j = 0; j < n; ++j) {
      row[2 * j] = x[2 * j];
      row[2 * j + 1] = c->sign > 0 ? -x[2 * j + 1] : x[2 * j + 1];
    }
    ORA_SFX(fplan_exec_row)(c->pl, row, res, scratch);
    ORA_REAL *y = c->out + 2 * b * n;
    for (int64_t j = 0; j < n; ++j) {
      y[2 * j] = res[2 * j];
      y[2 * j + 1] = c->sign > 0 ? -res[2 * j + 1] : res[2 * j + 1];
    }
  }
  free(scratch);
  free(row);
  free(res);
}

int ORA_SFX(ora_fft_c2c)(const ORA_REAL *in, ORA_REAL *out, int64_t n, int64_t batch, int sign, int nthreads) {
  if (n <= 0 || batch <= 0) return 1;
  ORA_SFX(fplan) *pl = ORA_SFX(fplan_make)(n);
  ORA_SFX(c2c_ctx) ctx = {pl, in, out, sign};
  parallel_for(batch, nthreads, ORA_SFX(c2c_rows), &ctx);
  ORA_SFX(fplan_free)(pl);
  return 0;
}

/* R2C by promotion to a full complex transform, then the first n/2+1 bins
 * (the reference's expand -> C2C -> pack, src/exec/raw_nodes.cpp:419-456). */
typedef struct {
  const ORA_SFX(fplan) *pl;
  const ORA_REAL *in;
  ORA_REAL *out;
  int64_t idist, odist;
} ORA_SFX(real_ctx);

static void ORA_SFX(r2c_rows)(void *vctx, int64_t begin, int64_t end) {
  ORA_SFX(real_ctx) *c = (ORA_SFX(real_ctx) *)vctx;
  const int64_t n = c->pl->n, h = n / 2 + 1;
  const int64_t scratch_len = c->pl->bluestein ? 4 * c->pl->m : 4 * n;
  ORA_REAL *scratch = (ORA_REAL *)malloc(sizeof(ORA_REAL) * (size_t)scratch_len);
  ORA_REAL *row = (ORA_REAL *)malloc(sizeof(ORA_REAL) * 2 * (size_t)n);
  ORA_REAL *res = (ORA_REAL *)malloc(sizeof(ORA_REAL) * 2 * (size_t)n);
  for (int64_t b = begin; b < end; ++b) {
    const ORA_REAL *x = c->in + b * c->idist;
    for (int64_t j = 0; j < n; ++j) {
      row[2 * j] = x[j];
      row[2 * j + 1] = 0;
    }
    ORA_SFX(fplan_exec_row)(c->pl, row, res, scratch);
    ORA_REAL *y = c->out + 2 * b * c->odist;
    memcpy(y, res, sizeof(ORA_REAL) * 2 * (size_t)h);
  }
  free(scratch);
  free(row);
  free(res);
}

/* C2R: Hermitian expansion with Im(X[0]) and Im(X[n/2]) (even n) dropped
 * (python/flagfft_codegen/kernels.py:1159-1192), inverse, real part
 * (kernels.py:1195-1221). */
static void ORA_SFX(c2r_rows)(void *vctx, int64_t begin, int64_t end) {
  ORA_SFX(real_ctx) *c = (ORA_SFX(real_ctx) *)vctx;
  const int64_t n = c->pl->n, h = n / 2 + 1;
  const int64_t scratch_len = c->pl->bluestein ? 4 * c->pl->m : 4 * n;
  ORA_REAL *scratch = (ORA_REAL *)malloc(sizeof(ORA_REAL) * (size_t)scratch_len);
  ORA_REAL *row = (ORA_REAL *)malloc(sizeof(ORA_REAL) * 2 * (size_t)n);
  ORA_REAL *res = (ORA_REAL *)malloc(sizeof(ORA_REAL) * 2 * (size_t)n);
  for (int64_t b = begin; b < end; ++b) {
    const ORA_REAL *x = c->in + 2 * b * c->idist;
    for (int64_t k = 0; k < n; ++k) {
      ORA_REAL xr, xi;
      if (k < h) {
        xr = x[2 * k];
        xi = x[2 * k + 1];
      } else {
        xr = x[2 * (n - k)];
        xi = -x[2 * (n - k) + 1];
      }
      if (k == 0 || (n % 2 == 0 && k == n / 2)) xi = 0;
      /* inverse via conj trick: conj input */
      row[2 * k] = xr;
      row[2 * k + 1] = -xi;
    }
    ORA_SFX(fplan_exec_row)(c->pl, row, res, scratch);
    ORA_REAL *y = c->out + b * c->odist;
    for (int64_t j = 0; j < n; ++j) y[j] = res[2 * j];
  }
  free(scratch);
  free(row);
  free(res);
}

int ORA_SFX(ora_fft_r2c)(const ORA_REAL *in, ORA_REAL *out, int64_t n, int64_t batch, int64_t idist, int64_t odist,
                         int nthreads) {
  if (n <= 0 || batch <= 0) return 1;
  ORA_SFX(fplan) *pl = ORA_SFX(fplan_make)(n);
  ORA_SFX(real_ctx) ctx = {pl, in, out, idist, odist};
  parallel_for(batch, nthreads, ORA_SFX(r2c_rows), &ctx);
  ORA_SFX(fplan_free)(pl);
  return 0;
}

int ORA_SFX(ora_fft_c2r)(const ORA_REAL *in, ORA_REAL *out, int64_t n, int64_t batch, int64_t idist, int64_t odist,
                         int nthreads) {
  if (n <= 0 || batch <= 0) return 1;
  ORA_SFX(fplan) *pl = ORA_SFX(fplan_make)(n);
  ORA_SFX(real_ctx) ctx = {pl, in, out, idist, odist};
  parallel_for(batch, nthreads, ORA_SFX(c2r_rows), &ctx);
  ORA_SFX(fplan_free)(pl);
  return 0;
}
