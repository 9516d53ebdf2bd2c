/* Type-generic core of the CPU oracle, included once per element type with
 *   T    the real type (float, double, long double)
 *   SUF  the name suffix (_f32, _f64, _ld)
 *   TW   the reference's pick_width(T) on an AVX-512 host (kernels.jl:11-22)
 * ORACLE ONLY: test infrastructure, never a product path.
 */
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define F(name) CAT(name, SUF)

/* ------------------------------------------------------------------ chain */

typedef struct {
  int n;
  int nst;
  int rad[40];
  T* twr[40];
  T* twi[40];
} F(Chain);

/* make_stage (plan.jl:66-95), access-order layout: entry (r-1)*p + j - 1 is
 * w^{p*j}, w = exp(D*2*pi*i/n), angle from an exact integer numerator */
static void F(make_twiddles)(int r, int n, int D, T** twr, T** twi) {
  int m = n / r;
  *twr = (T*)malloc(sizeof(T) * (size_t)((r - 1) * m + 1));
  *twi = (T*)malloc(sizeof(T) * (size_t)((r - 1) * m + 1));
  for (int p = 0; p < m; ++p)
    for (int j = 1; j < r; ++j) {
      long double c, s;
      expi_exact((int64_t)p * j, n, D, &c, &s);
      (*twr)[(r - 1) * p + j - 1] = (T)c;
      (*twi)[(r - 1) * p + j - 1] = (T)s;
    }
}

static int F(chain_init)(F(Chain) * ch, int64_t n, int D) {
  int rad[64];
  int nst = build_chain(n, TW, rad);
  if (nst < 0) return -1;
  ch->n = (int)n;
  ch->nst = nst;
  int ncur = (int)n;
  for (int i = 0; i < nst; ++i) {
    ch->rad[i] = rad[i];
    F(make_twiddles)(rad[i], ncur, D, &ch->twr[i], &ch->twi[i]);
    ncur /= rad[i];
  }
  return 0;
}

static void F(chain_free)(F(Chain) * ch) {
  for (int i = 0; i < ch->nst; ++i) {
    free(ch->twr[i]);
    free(ch->twi[i]);
  }
  ch->nst = 0;
}

/* One DIF Stockham stage (stockham.jl:26-130, 351-428): n = current length,
 * s = stride (lanes), m = n/r butterflies; for p < m, q < s
 *   a_j = x[s*p + j*m*s + q],  y[r*s*p + k*s + q] = w^{p*k} * DFT_r(a)_k     */
static void F(stage)(T* yre, T* yim, const T* xre, const T* xim, int n, int64_t s, int r, const T* twr,
                     const T* twi, int D) {
  int m = n / r;
  int64_t ms = (int64_t)m * s;
  int h = (r - 1) / 2;
  T cc[64], ss[64];
  if (r % 2 == 1) { /* bfp_consts (kernels.jl:158-163) */
    for (int k = 1; k <= h; ++k)
      for (int j = 1; j <= h; ++j) {
        long double c, sn;
        expi_exact((int64_t)j * k, r, 1, &c, &sn);
        cc[(k - 1) * h + j - 1] = (T)c;
        ss[(k - 1) * h + j - 1] = (T)(D * sn);
      }
  }
  const T c8 = (T)0.70710678118654752440084436210484903928L;
  for (int p = 0; p < m; ++p) {
    int64_t xb = s * p, yb = (int64_t)r * s * p;
    for (int64_t q = 0; q < s; ++q) {
      T ar[16], ai[16], yr[16], yi[16];
      for (int j = 0; j < r; ++j) {
        ar[j] = xre[xb + j * ms + q];
        ai[j] = xim[xb + j * ms + q];
      }
      if (r == 2) { /* bf2 (kernels.jl:93-98) */
        yr[0] = ar[0] + ar[1];
        yi[0] = ai[0] + ai[1];
        yr[1] = ar[0] - ar[1];
        yi[1] = ai[0] - ai[1];
      } else if (r == 4) {
        F(bf4_core)(ar[0], ai[0], ar[1], ai[1], ar[2], ai[2], ar[3], ai[3], D, yr, yi);
      } else if (r == 8) { /* bf8_core (kernels.jl:124-144) */
        T er[4], ei[4], or_[4], oi[4];
        F(bf4_core)(ar[0], ai[0], ar[2], ai[2], ar[4], ai[4], ar[6], ai[6], D, er, ei);
        F(bf4_core)(ar[1], ai[1], ar[3], ai[3], ar[5], ai[5], ar[7], ai[7], D, or_, oi);
        T t1r, t1i, t2r, t2i, t3r, t3i;
        if (D == -1) {
          t1r = c8 * (or_[1] + oi[1]);
          t1i = c8 * (oi[1] - or_[1]);
          t2r = oi[2];
          t2i = -or_[2];
          t3r = c8 * (oi[3] - or_[3]);
          t3i = -c8 * (or_[3] + oi[3]);
        } else {
          t1r = c8 * (or_[1] - oi[1]);
          t1i = c8 * (oi[1] + or_[1]);
          t2r = -oi[2];
          t2i = or_[2];
          t3r = -c8 * (or_[3] + oi[3]);
          t3i = c8 * (or_[3] - oi[3]);
        }
        yr[0] = er[0] + or_[0]; yi[0] = ei[0] + oi[0];
        yr[1] = er[1] + t1r;    yi[1] = ei[1] + t1i;
        yr[2] = er[2] + t2r;    yi[2] = ei[2] + t2i;
        yr[3] = er[3] + t3r;    yi[3] = ei[3] + t3i;
        yr[4] = er[0] - or_[0]; yi[4] = ei[0] - oi[0];
        yr[5] = er[1] - t1r;    yi[5] = ei[1] - t1i;
        yr[6] = er[2] - t2r;    yi[6] = ei[2] - t2i;
        yr[7] = er[3] - t3r;    yi[7] = ei[3] - t3i;
      } else { /* bfp (kernels.jl:174-228) */
        T tr[8], ti[8], dr[8], di[8];
        T y0r = ar[0], y0i = ai[0];
        for (int j = 1; j <= h; ++j) {
          tr[j] = ar[j] + ar[r - j];
          ti[j] = ai[j] + ai[r - j];
          dr[j] = ar[j] - ar[r - j];
          di[j] = ai[j] - ai[r - j];
          y0r += tr[j];
          y0i += ti[j];
        }
        yr[0] = y0r;
        yi[0] = y0i;
        for (int k = 1; k <= h; ++k) {
          T ur = ar[0], ui = ai[0], vr = 0, vi = 0;
          for (int j = h; j >= 1; --j) {
            ur += cc[(k - 1) * h + j - 1] * tr[j];
            ui += cc[(k - 1) * h + j - 1] * ti[j];
            vr += ss[(k - 1) * h + j - 1] * dr[j];
            vi += ss[(k - 1) * h + j - 1] * di[j];
          }
          yr[k] = ur - vi;
          yi[k] = ui + vr;
          yr[r - k] = ur + vi;
          yi[r - k] = ui - vr;
        }
      }
      yre[yb + q] = yr[0];
      yim[yb + q] = yi[0];
      for (int k = 1; k < r; ++k) {
        T wr = twr[(r - 1) * p + k - 1], wi = twi[(r - 1) * p + k - 1];
        yre[yb + k * s + q] = wr * yr[k] - wi * yi[k];
        yim[yb + k * s + q] = wr * yi[k] + wi * yr[k];
      }
    }
  }
}

/* run_chain! (plan.jl:311-321): ping-pong, stride s*B with B inner lanes;
 * returns 0 if the result is in (are, aim), 1 if in (bre, bim) */
static int F(chain_run)(const F(Chain) * ch, T* are, T* aim, T* bre, T* bim, int64_t B, int D) {
  int ncur = ch->n;
  int64_t s = 1;
  T *cr = are, *ci = aim, *orr = bre, *oi = bim;
  for (int i = 0; i < ch->nst; ++i) {
    F(stage)(orr, oi, cr, ci, ncur, s * B, ch->rad[i], ch->twr[i], ch->twi[i], D);
    T* t = cr; cr = orr; orr = t;
    t = ci; ci = oi; oi = t;
    s *= ch->rad[i];
    ncur /= ch->rad[i];
  }
  return cr == are ? 0 : 1;
}

/* FFTPlan call (plan.jl:359-396): deinterleave, chain, scaled interleave */
static int F(fft_chain)(int64_t n, int D, T scale, const T* x, T* y) {
  F(Chain) ch;
  if (F(chain_init)(&ch, n, D)) return -1;
  T* buf = (T*)malloc(sizeof(T) * 4 * (size_t)n);
  T *are = buf, *aim = buf + n, *bre = buf + 2 * n, *bim = buf + 3 * n;
  for (int64_t j = 0; j < n; ++j) {
    are[j] = x[2 * j];
    aim[j] = x[2 * j + 1];
  }
  int which = F(chain_run)(&ch, are, aim, bre, bim, 1, D);
  T* rr = which ? bre : are;
  T* ri = which ? bim : aim;
  for (int64_t j = 0; j < n; ++j) {
    y[2 * j] = rr[j] * scale;
    y[2 * j + 1] = ri[j] * scale;
  }
  free(buf);
  F(chain_free)(&ch);
  return 0;
}

/* -------------------------------------------------------------- four-step */

/* FourStepPlan (fourstep.jl:272-495), one task: j = j1 + N1*j2, k = k2 + N2*k1.
 * Pass A: strips of S columns, chain over j2 with S lanes, twiddle
 * w^{j1*k2} from the two-level table hi[e/K]*lo[e%K] (K = isqrt(n-1)+1,
 * fourstep.jl:297-308), store z[j1 + N1*k2].  Pass B: groups of G columns
 * k2 of z transposed into lanes, chain over j1, scaled natural-order rows.
 * The reference's further lane factor b1*b2 and the nested inner split are
 * register/cache details and are not restated.
 * nt > 1 runs strips and groups on nt threads (tforeach_chunks). */
typedef struct {
  int64_t n, N1, N2, S, K;
  int D;
  T scale;
  const T* x;
  T* y;
  T *zre, *zim;
  T *hir, *hii, *lor, *loi;
  F(Chain) chA, chB;
} F(FsCtx);

static void F(fs_pass_a)(F(FsCtx) * c, int64_t slo, int64_t shi) {
  int64_t N1 = c->N1, N2 = c->N2, S = c->S;
  T* buf = (T*)malloc(sizeof(T) * 4 * (size_t)(S * N2));
  T *are = buf, *aim = buf + S * N2, *bre = buf + 2 * S * N2, *bim = buf + 3 * S * N2;
  for (int64_t st = slo; st < shi; ++st) {
    int64_t c0 = st * S;
    int64_t lim = N1 - c0 < S ? N1 - c0 : S;
    for (int64_t j2 = 0; j2 < N2; ++j2)
      for (int64_t l = 0; l < S; ++l) {
        int64_t e = c0 + l + N1 * j2;
        are[S * j2 + l] = l < lim ? c->x[2 * e] : 0;
        aim[S * j2 + l] = l < lim ? c->x[2 * e + 1] : 0;
      }
    int which = F(chain_run)(&c->chA, are, aim, bre, bim, S, c->D);
    T* rr = which ? bre : are;
    T* ri = which ? bim : aim;
    for (int64_t k2 = 0; k2 < N2; ++k2)
      for (int64_t l = 0; l < lim; ++l) {
        int64_t j1 = c0 + l;
        int64_t e = j1 * k2;
        int64_t hi = e / c->K, lo = e % c->K;
        T tr = c->hir[hi] * c->lor[lo] - c->hii[hi] * c->loi[lo];
        T ti = c->hir[hi] * c->loi[lo] + c->hii[hi] * c->lor[lo];
        T vr = rr[S * k2 + l], vi = ri[S * k2 + l];
        c->zre[j1 + N1 * k2] = vr * tr - vi * ti;
        c->zim[j1 + N1 * k2] = vr * ti + vi * tr;
      }
  }
  free(buf);
}

static void F(fs_pass_b)(F(FsCtx) * c, int64_t glo, int64_t ghi) {
  int64_t N1 = c->N1, N2 = c->N2, G = c->S;
  T* buf = (T*)malloc(sizeof(T) * 4 * (size_t)(G * N1));
  T *are = buf, *aim = buf + G * N1, *bre = buf + 2 * G * N1, *bim = buf + 3 * G * N1;
  for (int64_t gr = glo; gr < ghi; ++gr) {
    int64_t c2 = gr * G;
    int64_t lim = N2 - c2 < G ? N2 - c2 : G;
    for (int64_t g = 0; g < G; ++g)
      for (int64_t j1 = 0; j1 < N1; ++j1) {
        are[G * j1 + g] = g < lim ? c->zre[j1 + N1 * (c2 + g)] : 0;
        aim[G * j1 + g] = g < lim ? c->zim[j1 + N1 * (c2 + g)] : 0;
      }
    int which = F(chain_run)(&c->chB, are, aim, bre, bim, G, c->D);
    T* rr = which ? bre : are;
    T* ri = which ? bim : aim;
    for (int64_t k1 = 0; k1 < N1; ++k1)
      for (int64_t g = 0; g < lim; ++g) {
        int64_t e = c2 + g + N2 * k1;
        c->y[2 * e] = rr[G * k1 + g] * c->scale;
        c->y[2 * e + 1] = ri[G * k1 + g] * c->scale;
      }
  }
  free(buf);
}

typedef struct {
  F(FsCtx) * c;
  int pass;
  int64_t lo, hi;
} F(FsJob);

static void* F(fs_worker)(void* arg) {
  F(FsJob)* j = (F(FsJob)*)arg;
  if (j->pass == 0) F(fs_pass_a)(j->c, j->lo, j->hi);
  else F(fs_pass_b)(j->c, j->lo, j->hi);
  return NULL;
}

static void F(fs_run_pass)(F(FsCtx) * c, int pass, int64_t count, int nt) {
  if (nt <= 1) {
    if (pass == 0) F(fs_pass_a)(c, 0, count);
    else F(fs_pass_b)(c, 0, count);
    return;
  }
  pthread_t th[256];
  F(FsJob) jobs[256];
  if (nt > 256) nt = 256;
  int used = 0;
  for (int i = 0; i < nt; ++i) {
    int64_t lo, hi;
    chunk_range(count, nt, i, 1, &lo, &hi);
    if (lo >= hi) continue;
    jobs[used].c = c;
    jobs[used].pass = pass;
    jobs[used].lo = lo;
    jobs[used].hi = hi;
    pthread_create(&th[used], NULL, F(fs_worker), &jobs[used]);
    ++used;
  }
  for (int i = 0; i < used; ++i) pthread_join(th[i], NULL);
}

static int F(fourstep)(int64_t n, int64_t N1, int64_t N2, int64_t S, int D, T scale, const T* x, T* y, int nt) {
  if (N1 * N2 != n || S < 1) return -1;
  F(FsCtx) c;
  memset(&c, 0, sizeof(c));
  c.n = n;
  c.N1 = N1;
  c.N2 = N2;
  c.S = S;
  c.D = D;
  c.scale = scale;
  c.x = x;
  c.y = y;
  if (F(chain_init)(&c.chA, N2, D) || F(chain_init)(&c.chB, N1, D)) return -1;
  int64_t K = isqrt64(n - 1) + 1, nh = (n + K - 1) / K;
  c.K = K;
  c.hir = (T*)malloc(sizeof(T) * (size_t)nh);
  c.hii = (T*)malloc(sizeof(T) * (size_t)nh);
  c.lor = (T*)malloc(sizeof(T) * (size_t)K);
  c.loi = (T*)malloc(sizeof(T) * (size_t)K);
  for (int64_t i = 0; i < nh; ++i) {
    long double cc, ss;
    expi_exact(i * K, n, D, &cc, &ss);
    c.hir[i] = (T)cc;
    c.hii[i] = (T)ss;
  }
  for (int64_t i = 0; i < K; ++i) {
    long double cc, ss;
    expi_exact(i, n, D, &cc, &ss);
    c.lor[i] = (T)cc;
    c.loi[i] = (T)ss;
  }
  c.zre = (T*)malloc(sizeof(T) * (size_t)n);
  c.zim = (T*)malloc(sizeof(T) * (size_t)n);
  F(fs_run_pass)(&c, 0, (N1 + S - 1) / S, nt);
  F(fs_run_pass)(&c, 1, (N2 + S - 1) / S, nt);
  free(c.zre);
  free(c.zim);
  free(c.hir);
  free(c.hii);
  free(c.lor);
  free(c.loi);
  F(chain_free)(&c.chA);
  F(chain_free)(&c.chB);
  return 0;
}

/* -------------------------------------------------------------- resample */

/* ResamplePlan (resample.jl:175-276) with table weights (:232-245) */
static int F(resample)(int64_t n, int D, T scale, const T* x, T* y, long double tol) {
  int64_t m = nextsmooth(2 * n);
  double ratio = (double)n / (double)m;
  if (ratio < 0.5) ratio = 0.5;
  int w0 = (int)ceil(log(1.0 / (double)tol) / (M_PI * sqrt(1.0 - ratio)));
  int w = 4 * ((w0 + 3) / 4);
  int h = w / 2;
  double beta = 0.97 * M_PI * w * (1.0 - (double)n / (2.0 * m));
  int nq = 2 * w + 20;
  double zq[128], wq[128];
  gauss_legendre(nq, zq, wq);
  int64_t hp = (n + 1) / 2;
  T* corr = (T*)malloc(sizeof(T) * (size_t)n);
  for (int64_t j = 0; j < n; ++j) {
    int64_t jp = j < hp ? j : j - n;
    double nu = (double)jp / (double)m;
    double phi = 0.0;
    for (int q = 0; q < nq; ++q)
      if (zq[q] > 0) phi += wq[q] * es_kernel(zq[q], beta) * cos(M_PI * nu * w * zq[q]);
    corr[j] = (T)((double)scale / (w * phi));
  }
  F(Chain) ch;
  if (F(chain_init)(&ch, m, D)) {
    free(corr);
    return -1;
  }
  T* buf = (T*)malloc(sizeof(T) * 4 * (size_t)m);
  T *are = buf, *aim = buf + m, *bre = buf + 2 * m, *bim = buf + 3 * m;
  for (int64_t j = 0; j < m; ++j) are[j] = aim[j] = 0;
  for (int64_t j = 0; j < hp; ++j) {
    are[j] = x[2 * j] * corr[j];
    aim[j] = x[2 * j + 1] * corr[j];
  }
  for (int64_t j = hp; j < n; ++j) {
    int64_t o = m - n + j;
    are[o] = x[2 * j] * corr[j];
    aim[o] = x[2 * j + 1] * corr[j];
  }
  int which = F(chain_run)(&ch, are, aim, bre, bim, 1, D);
  T* gr = which ? bre : are;
  T* gi = which ? bim : aim;
  for (int64_t k = 0; k < n; ++k) {
    int64_t km = k * m, q = km / n, r = km % n;
    int64_t l0 = q - h + 1;
    double delta = (double)r / (double)n;
    T re = 0, im = 0;
    for (int i = 0; i < w; ++i) {
      T wt = (T)es_kernel((delta + h - 1 - i) / h, beta);
      int64_t l = ((l0 + i) % m + m) % m;
      re += wt * gr[l];
      im += wt * gi[l];
    }
    y[2 * k] = re;
    y[2 * k + 1] = im;
  }
  free(buf);
  free(corr);
  F(chain_free)(&ch);
  return 0;
}

/* 1-D routing for the oracle: a smooth length runs its stage chain, any
 * other length the resampling path (plan.jl:192-206 without the CPU
 * speed heuristics, which do not change the result beyond rounding) */
int F(oracle_fft)(int64_t n, int D, T scale, const T* x, T* y) {
  if (n < 1) return -1;
  if (is_smooth13(n)) return F(fft_chain)(n, D, scale, x, y);
  return F(resample)(n, D, scale, x, y, F(resample_tol)());
}

int F(oracle_fft_batch)(int64_t n, int64_t batch, int D, T scale, const T* x, T* y) {
  for (int64_t b = 0; b < batch; ++b)
    if (F(oracle_fft)(n, D, scale, x + 2 * n * b, y + 2 * n * b)) return -1;
  return 0;
}

int F(oracle_fourstep)(int64_t n, int64_t N1, int64_t N2, int D, T scale, const T* x, T* y, int nt) {
  return F(fourstep)(n, N1, N2, TW, D, scale, x, y, nt);
}

int F(oracle_resample)(int64_t n, int D, T scale, const T* x, T* y) {
  return F(resample)(n, D, scale, x, y, F(resample_tol)());
}

/* ----------------------------------------------------------------- HvdH */

/* Length-n DFT by Gaussian resampling (hvdh.jl:1-19, after Harvey & van der
 * Hoeven) onto the smooth t of hvdh_t, Delta = t - n, rho = t/n:
 *   v[Delta*k mod t] = (S u)_k = (1/alpha) sum_j exp(-b (k n/t - j)^2) u_(j mod n),
 *                      b = pi/alpha^2, 2h taps (hvdh.jl:360-470);
 *   Y = F_t v, unnormalized, sign D;
 *   (n/t) Y[q] = sum_l exp(-a (rho*l - q)^2) x_l,  a = pi*alpha^2,
 *                x_l = X_(Delta*l mod n)     (Poisson summation of the spread)
 * on the rows q = [rho*l] nearest each unknown; the other rows are
 * identities.  Indexed by q the system is banded (entries exp(-a(d+beta)^2),
 * |d| <= B) and is solved by LU without pivoting in chunks of L rows with H
 * halo rows per side (hvdh.jl:474-686); the inverse decays, so each chunk's
 * core rows are exact to rounding.  Parameters as hvdh.jl:154-190: alpha =
 * 1.0 (F32) / 1.45 (F64), tol = eps/4, the AVX-512 chunk counts K = 16 / 8
 * and the row target 256.  Weights by the reference's factorisations,
 * exp(-b(delta-o)^2) = E G^o exp(-b o^2) and exp(-a(d+beta)^2) =
 * A' g^d exp(-a d^2), two exponentials per output / per column. */

typedef struct {
  int64_t n, t, delta, H, L, G, K;
  int B, h;
  long double a, b, alpha, tol;
} F(Hvdh);

/* e = 2t*l' - 2n*q + n for the first unknown l' with rho*l' >= q - 1/2
 * (hvdh.jl:147-151); row q holds unknown l' iff e < 2n */
static int64_t F(hvdh_e)(int64_t n, int64_t t, int64_t q, int64_t* lp) {
  const int64_t num = 2 * n * q - n, den = 2 * t;
  int64_t l = num >= 0 ? (num + den - 1) / den : -((-num) / den);
  *lp = l;
  return 2 * t * l - 2 * n * q + n;
}

/* banded LU without pivoting, in place: A is m x m with half band w, stored
 * dense row-major; solves A x = rhs */
static int F(band_solve)(long double* A, long double* x, int64_t m, int64_t w) {
  for (int64_t k = 0; k < m; ++k) {
    const long double piv = A[k * m + k];
    if (piv == 0) return -1;
    const int64_t hi = k + w < m - 1 ? k + w : m - 1;
    for (int64_t i = k + 1; i <= hi; ++i) {
      const long double f = A[i * m + k] / piv;
      if (f == 0) continue;
      for (int64_t j = k; j <= hi; ++j) A[i * m + j] -= f * A[k * m + j];
      x[i] -= f * x[k];
    }
  }
  for (int64_t k = m - 1; k >= 0; --k) {
    long double acc = x[k];
    const int64_t hi = k + w < m - 1 ? k + w : m - 1;
    for (int64_t j = k + 1; j <= hi; ++j) acc -= A[k * m + j] * x[j];
    x[k] = acc / A[k * m + k];
  }
  return 0;
}

/* halo rows per side (hvdh.jl:118-145): rows of the inverse of local
 * (2P+1)-windows around 18 probe unknowns decay below tol within H columns */
static int64_t F(hvdh_halo)(int64_t n, int64_t t, long double a, int B, long double tol) {
  const long double rho = (long double)t / n;
  /* c = (t - n)^-1 mod n */
  int64_t c = 1;
  {
    int64_t r0 = n, r1 = (t - n) % n, s0 = 0, s1 = 1;
    while (r1) {
      const int64_t qq = r0 / r1, r2 = r0 - qq * r1, s2 = s0 - qq * s1;
      r0 = r1;
      r1 = r2;
      s0 = s1;
      s1 = s2;
    }
    c = ((s0 % n) + n) % n;
  }
  int64_t probes[18];
  int np = 0;
  const int64_t cnt = n < 16 ? n : 16;
  for (int64_t k = 0; k < cnt; ++k)
    probes[np++] = cnt > 1 ? llrintl((long double)k * (n - 1) / (cnt - 1)) : 0;
  probes[np++] = (int64_t)(((__int128)(n / 2) * c) % n);
  probes[np++] = (int64_t)(((__int128)(n / 2 + 1) * c) % n);
  for (int64_t P = 64; P <= 8192; P *= 2) {
    const int64_t m = 2 * P + 1;
    long double* M = (long double*)malloc(sizeof(long double) * (size_t)(m * m));
    long double* r = (long double*)malloc(sizeof(long double) * (size_t)m);
    int64_t H = 0;
    for (int pi = 0; pi < np; ++pi) {
      const int64_t l0 = probes[pi];
      /* M^T, entry (j, i) = M[i, j] = exp(-a (rho l_j - fld(2t l_i + n, 2n))^2) */
      memset(M, 0, sizeof(long double) * (size_t)(m * m));
      for (int64_t i = 0; i < m; ++i) {
        const int64_t li = l0 - P + i;
        const int64_t num = 2 * t * li + n, den = 2 * n;
        const int64_t qi = num >= 0 ? num / den : -((-num + den - 1) / den);
        for (int64_t j = i - B - 1; j <= i + B + 1; ++j) {
          if (j < 0 || j >= m) continue;
          const long double d = rho * (long double)(l0 - P + j) - (long double)qi;
          M[j * m + i] = expl(-a * d * d);
        }
        r[i] = 0;
      }
      r[P] = 1;
      if (F(band_solve)(M, r, m, B + 1)) {
        free(M);
        free(r);
        return -1;
      }
      long double mx = 0;
      for (int64_t j = 0; j < m; ++j) mx = fabsl(r[j]) > mx ? fabsl(r[j]) : mx;
      for (int64_t j = 0; j < m; ++j) {
        const int64_t dj = j - P < 0 ? P - j : j - P;
        if (fabsl(r[j]) > tol * mx && dj > H) H = dj;
      }
    }
    free(M);
    free(r);
    if (H < P - 2 * B) return (int64_t)ceill(rho * H) + B;
  }
  return -1;
}

static int F(hvdh_init)(F(Hvdh) * p, int64_t n) {
  p->n = n;
  p->t = hvdh_t(n);
  if (p->t < 0) return -1;
  p->delta = p->t - n;
  const int f64 = sizeof(T) >= 8;
  p->alpha = f64 ? 1.45L : 1.0L;
  p->a = (long double)M_PI * p->alpha * p->alpha;
  p->b = (long double)M_PI / (p->alpha * p->alpha);
  p->tol = f64 ? ldexpl(1.0L, -54) : ldexpl(1.0L, -25);
  p->h = (int)ceill(p->alpha * sqrtl(logl(1 / p->tol) / (long double)M_PI));
  p->B = (int)floorl(sqrtl(logl(1 / p->tol) / p->a) + 0.5L);
  p->H = F(hvdh_halo)(n, p->t, p->a, p->B, p->tol);
  if (p->H < 0) return -1;
  p->K = f64 ? 8 : 16;
  const int64_t Lt = 256;
  p->G = (p->t + p->K * Lt - 1) / (p->K * Lt);
  if (p->G < 1) p->G = 1;
  p->L = (p->t + p->K * p->G - 1) / (p->K * p->G);
  return 0;
}

/* halo < 0: the probed halo; halo >= 0 overrides it (tests plant a
 * truncated halo there) */
static int F(hvdh)(int64_t n, int D, T scale, const T* x, T* y, int64_t halo) {
  F(Hvdh) p;
  if (F(hvdh_init)(&p, n)) return -1;
  if (halo >= 0) p.H = halo;
  const int64_t t = p.t, dl = p.delta;
  const int h = p.h, B = p.B;
  /* spread, in transform order */
  T* v = (T*)calloc(2 * (size_t)t, sizeof(T));
  T* Y = (T*)malloc(sizeof(T) * 2 * (size_t)t);
  T* kc = (T*)malloc(sizeof(T) * (size_t)(2 * h + 1));
  for (int o = -h; o <= h; ++o) kc[o + h] = (T)expl(-p.b * o * o);
  const T bT = (T)p.b, ia = (T)(1 / p.alpha);
  for (int64_t k = 0; k < t; ++k) {
    const int64_t kn_mod = (int64_t)(((__int128)k * n) % t);
    int64_t s = (int64_t)(((__int128)k * n) / t);
    T dlt;
    if (kn_mod == 0) {
      s -= 1;
      dlt = (T)1;
    } else {
      dlt = (T)((long double)kn_mod / t);
    }
    const T E = (T)exp((double)(-bT * dlt * dlt)) * ia;
    const T Gp = (T)exp((double)(2 * bT * dlt)), Gm = (T)1 / Gp;
    T re = 0, im = 0;
    T P = E;
    for (int o = 1; o <= h; ++o) { /* taps right of s: weight E G^o exp(-b o^2) */
      P *= Gp;
      const T W = P * kc[o + h];
      const int64_t j = ((s + o) % n + n) % n;
      re += W * x[2 * j];
      im += W * x[2 * j + 1];
    }
    P = E;
    for (int o = 0; o >= 1 - h; --o) {
      if (o < 0) P *= Gm;
      const T W = P * kc[o + h];
      const int64_t j = ((s + o) % n + n) % n;
      re += W * x[2 * j];
      im += W * x[2 * j + 1];
    }
    const int64_t q = (int64_t)(((__int128)dl * k) % t);
    v[2 * q] = re;
    v[2 * q + 1] = im;
  }
  free(kc);
  if (F(fft_chain)(t, D, (T)1, v, Y)) {
    free(v);
    free(Y);
    return -1;
  }
  free(v);
  /* chunk solves: band storage, row i holds columns i-B..i+B; LU without
   * pivoting in the plan's precision, one factorisation for the real and
   * imaginary right-hand sides */
  const int64_t R = p.L + 2 * p.H, W = 2 * B + 1;
  T* A = (T*)malloc(sizeof(T) * (size_t)(R * W));
  T* br = (T*)malloc(sizeof(T) * (size_t)R);
  T* bi = (T*)malloc(sizeof(T) * (size_t)R);
  T* beta = (T*)malloc(sizeof(T) * (size_t)R);
  T* Ap = (T*)malloc(sizeof(T) * (size_t)R);
  T* gp = (T*)malloc(sizeof(T) * (size_t)R);
  T* ch = (T*)malloc(sizeof(T) * (size_t)(B + 1));
  int64_t* out = (int64_t*)malloc(sizeof(int64_t) * (size_t)R);
  for (int d = 0; d <= B; ++d) ch[d] = (T)expl(-p.a * d * d);
  const T nt_ratio = (T)((long double)n / t), aT = (T)p.a;
  int rc = 0;
  for (int64_t c = 0; c < p.G * p.K && !rc; ++c) {
    if (c * p.L >= t) break;
    const int64_t q0 = c * p.L - p.H;
    for (int64_t i = 0; i < R; ++i) {
      int64_t lp;
      const int64_t e = F(hvdh_e)(n, t, q0 + i, &lp);
      out[i] = e < 2 * n ? (int64_t)((((__int128)dl * lp) % n + n) % n) : -1;
      beta[i] = (T)((long double)(e - n) / (2.0L * n));
      /* column state: A' = exp(-a beta^2), g = exp(-2 a beta) */
      Ap[i] = out[i] < 0 ? (T)0 : (T)exp((double)(-aT * beta[i] * beta[i]));
      gp[i] = (T)exp((double)(-2 * aT * beta[i]));
    }
    for (int64_t i = 0; i < R; ++i) {
      T* row = A + i * W;
      const int64_t qm = ((q0 + i) % t + t) % t;
      for (int64_t m = 0; m < W; ++m) row[m] = 0;
      if (out[i] < 0) {
        row[B] = 1;
        br[i] = bi[i] = 0;
        continue;
      }
      for (int m = -B; m <= B; ++m) {
        const int64_t j = i + m;
        if (j < 0 || j >= R || out[j] < 0) continue;
        /* exp(-a (m + beta_j)^2) = A'_j g_j^m exp(-a m^2) */
        T gm = 1;
        const T gb = m >= 0 ? gp[j] : (T)1 / gp[j];
        for (int r = 0; r < (m >= 0 ? m : -m); ++r) gm *= gb;
        row[m + B] = Ap[j] * gm * ch[m >= 0 ? m : -m];
      }
      br[i] = nt_ratio * Y[2 * qm];
      bi[i] = nt_ratio * Y[2 * qm + 1];
    }
    for (int64_t k = 0; k < R && !rc; ++k) {
      const T piv = A[k * W + B];
      if (piv == 0) {
        rc = -1;
        break;
      }
      for (int64_t r = 1; r <= B && k + r < R; ++r) {
        T* ri = A + (k + r) * W;
        const T f = ri[B - r] / piv;
        if (f == 0) continue;
        for (int64_t j = 0; j <= B; ++j) ri[B - r + j] -= f * A[k * W + B + j];
        br[k + r] -= f * br[k];
        bi[k + r] -= f * bi[k];
      }
    }
    for (int64_t k = R - 1; k >= 0 && !rc; --k) {
      T sr = br[k], si = bi[k];
      for (int64_t j = 1; j <= B && k + j < R; ++j) {
        sr -= A[k * W + B + j] * br[k + j];
        si -= A[k * W + B + j] * bi[k + j];
      }
      br[k] = sr / A[k * W + B];
      bi[k] = si / A[k * W + B];
    }
    for (int64_t i = p.H; i < p.H + p.L && !rc; ++i) {
      if (q0 + i >= t || out[i] < 0) continue;
      y[2 * out[i]] = br[i] * scale;
      y[2 * out[i] + 1] = bi[i] * scale;
    }
  }
  free(A);
  free(br);
  free(bi);
  free(beta);
  free(Ap);
  free(gp);
  free(ch);
  free(out);
  free(Y);
  return rc;
}

int F(oracle_hvdh)(int64_t n, int D, T scale, const T* x, T* y) { return F(hvdh)(n, D, scale, x, y, -1); }
int F(oracle_hvdh_halo)(int64_t n, int D, T scale, const T* x, T* y, int64_t halo) {
  return F(hvdh)(n, D, scale, x, y, halo);
}

/* ---------------------------------------------------------------- split */

/* SplitPlan (composite.jl:103-179, call :181-200): n = s*p with s the
 * 7-smooth part of n (composite.jl:45-52) and p odd.  Decimation in time:
 * the s inputs x_r[j] = x[r + s*j] (the gather, :214-220) through length-p
 * HvdH unnormalized, A[k1 + p*r] = w_n^(r*k1) Y_r[k1] (the twiddle,
 * :222-262), then the length-s chain over r with p inner lanes, whose
 * output X[k1 + p*k2] is natural order, scaled (:195-198).  The reference
 * takes this route for even n with s < 16 (:138-142); any n with a smooth
 * part and an odd cofactor is accepted here. */
static int F(split)(int64_t n, int D, T scale, const T* x, T* y) {
  static const int fs[4] = {2, 3, 5, 7};
  int64_t s = 1;
  for (int i = 0; i < 4; ++i)
    while (n % (s * fs[i]) == 0) s *= fs[i];
  const int64_t p = n / s;
  if (s < 2 || p < 3 || p % 2 == 0) return -1;
  F(Chain) ch;
  if (F(chain_init)(&ch, s, D)) return -1;
  T* sub = (T*)malloc(sizeof(T) * 4 * (size_t)p);
  T* buf = (T*)malloc(sizeof(T) * 4 * (size_t)n);
  T *xs = sub, *ys = sub + 2 * p;
  T *are = buf, *aim = buf + n, *bre = buf + 2 * n, *bim = buf + 3 * n;
  int rc = 0;
  for (int64_t r = 0; r < s && rc == 0; ++r) {
    for (int64_t j = 0; j < p; ++j) {
      xs[2 * j] = x[2 * (r + s * j)];
      xs[2 * j + 1] = x[2 * (r + s * j) + 1];
    }
    rc = F(hvdh)(p, D, (T)1, xs, ys, -1);
    for (int64_t k1 = 0; k1 < p && rc == 0; ++k1) {
      long double c, sn;
      expi_exact(r * k1, n, D, &c, &sn);
      const T wr = (T)c, wi = (T)sn, zr = ys[2 * k1], zi = ys[2 * k1 + 1];
      are[k1 + p * r] = zr * wr - zi * wi;
      aim[k1 + p * r] = zr * wi + zi * wr;
    }
  }
  if (rc == 0) {
    const int which = F(chain_run)(&ch, are, aim, bre, bim, p, D);
    const T* rr = which ? bre : are;
    const T* ri = which ? bim : aim;
    for (int64_t k = 0; k < n; ++k) {
      y[2 * k] = rr[k] * scale;
      y[2 * k + 1] = ri[k] * scale;
    }
  }
  free(sub);
  free(buf);
  F(chain_free)(&ch);
  return rc;
}

int F(oracle_split)(int64_t n, int D, T scale, const T* x, T* y) { return F(split)(n, D, scale, x, y); }

/* ------------------------------------------------------------------ real */

/* rfft (real.jl:80-142): even n packs to n/2 complex + twist, odd n full */
int F(oracle_rfft)(int64_t n, const T* x, T* y) {
  if (n < 1) return -1;
  if (n % 2) {
    T* xc = (T*)calloc(2 * (size_t)n, sizeof(T));
    T* yc = (T*)malloc(sizeof(T) * 2 * (size_t)n);
    for (int64_t j = 0; j < n; ++j) xc[2 * j] = x[j];
    int rc = F(oracle_fft)(n, -1, (T)1, xc, yc);
    memcpy(y, yc, sizeof(T) * 2 * (size_t)(n / 2 + 1));
    free(xc);
    free(yc);
    return rc;
  }
  int64_t h = n / 2;
  T* Z = (T*)malloc(sizeof(T) * 2 * (size_t)h);
  if (F(oracle_fft)(h, -1, (T)1, x, Z)) {
    free(Z);
    return -1;
  }
  y[0] = Z[0] + Z[1];
  y[1] = 0;
  y[2 * h] = Z[0] - Z[1];
  y[2 * h + 1] = 0;
  for (int64_t k = 1; k < h; ++k) {
    long double c, s;
    expi_exact(k, n, -1, &c, &s);
    T wr = (T)c, wi = (T)s;
    T ar = Z[2 * k], ai = Z[2 * k + 1];
    T br = Z[2 * (h - k)], bi = -Z[2 * (h - k) + 1];
    T xer = (ar + br) * (T)0.5, xei = (ai + bi) * (T)0.5;
    T xor_ = (ar - br) * (T)0.5, xoi = (ai - bi) * (T)0.5;
    /* xe + w*(-i*xo), -i*xo = (xoi, -xor) */
    T tr = xoi, ti = -xor_;
    y[2 * k] = xer + (wr * tr - wi * ti);
    y[2 * k + 1] = xei + (wr * ti + wi * tr);
  }
  free(Z);
  return 0;
}

/* irfft / brfft (real.jl:144-211) */
int F(oracle_irfft)(int64_t n, const T* X, T* y, int normalized) {
  if (n < 1) return -1;
  if (n % 2) {
    int64_t h = n / 2;
    T* buf = (T*)malloc(sizeof(T) * 2 * (size_t)n);
    T* out = (T*)malloc(sizeof(T) * 2 * (size_t)n);
    for (int64_t k = 0; k <= h; ++k) {
      buf[2 * k] = X[2 * k];
      buf[2 * k + 1] = X[2 * k + 1];
    }
    for (int64_t k = 1; k <= h; ++k) {
      buf[2 * (n - k)] = X[2 * k];
      buf[2 * (n - k) + 1] = -X[2 * k + 1];
    }
    int rc = F(oracle_fft)(n, 1, normalized ? (T)(1.0L / n) : (T)1, buf, out);
    for (int64_t j = 0; j < n; ++j) y[j] = out[2 * j];
    free(buf);
    free(out);
    return rc;
  }
  int64_t h = n / 2;
  T pre = normalized ? (T)0.5 : (T)1;
  T* Z = (T*)malloc(sizeof(T) * 2 * (size_t)h);
  for (int64_t k = 0; k < h; ++k) {
    long double c, s;
    expi_exact(k, n, 1, &c, &s);
    T wr = (T)c, wi = (T)s;
    T ar = X[2 * k], ai = X[2 * k + 1];
    T br = X[2 * (h - k)], bi = -X[2 * (h - k) + 1];
    T xer = (ar + br) * pre, xei = (ai + bi) * pre;
    T xdr = (ar - br) * pre, xdi = (ai - bi) * pre;
    T tr = wr * xdr - wi * xdi, ti = wr * xdi + wi * xdr;
    Z[2 * k] = xer - ti;
    Z[2 * k + 1] = xei + tr;
  }
  int rc = F(oracle_fft)(h, 1, normalized ? (T)(1.0L / h) : (T)1, Z, y);
  free(Z);
  return rc;
}

/* ------------------------------------------------------------------- N-d */

/* NdPlan (ndplan.jl:57-118, 270-327): one pass per dim of length > 1, each
 * pencil transformed by the 1-D route; column-major, dims[0] fastest;
 * the scale is applied on the last pass */
int F(oracle_ndfft)(int rank, const int64_t* dims, int D, T scale, const T* x, T* y) {
  int64_t len = 1;
  for (int i = 0; i < rank; ++i) len *= dims[i];
  memcpy(y, x, sizeof(T) * 2 * (size_t)len);
  int last = -1;
  for (int i = 0; i < rank; ++i)
    if (dims[i] > 1) last = i;
  if (last < 0) {
    for (int64_t j = 0; j < 2 * len; ++j) y[j] *= scale;
    return 0;
  }
  for (int ax = 0; ax < rank; ++ax) {
    int64_t nd = dims[ax];
    if (nd == 1) continue;
    int64_t Bi = 1, Bo = 1;
    for (int i = 0; i < ax; ++i) Bi *= dims[i];
    for (int i = ax + 1; i < rank; ++i) Bo *= dims[i];
    T* pin = (T*)malloc(sizeof(T) * 2 * (size_t)nd);
    T* pout = (T*)malloc(sizeof(T) * 2 * (size_t)nd);
    T sc = ax == last ? scale : (T)1;
    for (int64_t bo = 0; bo < Bo; ++bo)
      for (int64_t bi = 0; bi < Bi; ++bi) {
        int64_t base = bi + bo * Bi * nd;
        for (int64_t i = 0; i < nd; ++i) {
          pin[2 * i] = y[2 * (base + i * Bi)];
          pin[2 * i + 1] = y[2 * (base + i * Bi) + 1];
        }
        if (F(oracle_fft)(nd, D, sc, pin, pout)) {
          free(pin);
          free(pout);
          return -1;
        }
        for (int64_t i = 0; i < nd; ++i) {
          y[2 * (base + i * Bi)] = pout[2 * i];
          y[2 * (base + i * Bi) + 1] = pout[2 * i + 1];
        }
      }
    free(pin);
    free(pout);
  }
  return 0;
}

/* ------------------------------------------------------- naive reference */

/* X_k = sum_j x_j exp(D*2*pi*i*jk/n), exact integer angle reduction */
int F(dft_naive)(int64_t n, int D, const T* x, T* y) {
  for (int64_t k = 0; k < n; ++k) {
    long double sr = 0, si = 0;
    for (int64_t j = 0; j < n; ++j) {
      long double c, s;
      expi_exact((j * k) % n, n, D, &c, &s);
      sr += c * x[2 * j] - s * x[2 * j + 1];
      si += c * x[2 * j + 1] + s * x[2 * j];
    }
    y[2 * k] = (T)sr;
    y[2 * k + 1] = (T)si;
  }
  return 0;
}

/* ------------------------------------------------- threaded CPU baseline */

/* the reference's loop of 1-D plans over a batch, spread over nt threads in
 * contiguous chunks (threads.jl:15-38); one chain (and its twiddle tables)
 * per thread, reused across that thread's transforms */
typedef struct {
  int64_t n, lo, hi;
  int D;
  T scale;
  const T* x;
  T* y;
} F(BatchJob);

static void* F(batch_worker)(void* arg) {
  F(BatchJob)* j = (F(BatchJob)*)arg;
  F(Chain) ch;
  if (F(chain_init)(&ch, j->n, j->D)) return NULL;
  int64_t n = j->n;
  T* buf = (T*)malloc(sizeof(T) * 4 * (size_t)n);
  T *are = buf, *aim = buf + n, *bre = buf + 2 * n, *bim = buf + 3 * n;
  for (int64_t b = j->lo; b < j->hi; ++b) {
    const T* x = j->x + 2 * n * b;
    T* y = j->y + 2 * n * b;
    for (int64_t i = 0; i < n; ++i) {
      are[i] = x[2 * i];
      aim[i] = x[2 * i + 1];
    }
    int which = F(chain_run)(&ch, are, aim, bre, bim, 1, j->D);
    T* rr = which ? bre : are;
    T* ri = which ? bim : aim;
    for (int64_t i = 0; i < n; ++i) {
      y[2 * i] = rr[i] * j->scale;
      y[2 * i + 1] = ri[i] * j->scale;
    }
  }
  free(buf);
  F(chain_free)(&ch);
  return NULL;
}

int F(baseline_fft_batch)(int64_t n, int64_t batch, int D, T scale, const T* x, T* y, int nt) {
  if (!is_smooth13(n)) return -1;
  if (nt < 1) nt = 1;
  if (nt > 256) nt = 256;
  pthread_t th[256];
  F(BatchJob) jobs[256];
  int used = 0;
  for (int i = 0; i < nt; ++i) {
    int64_t lo, hi;
    chunk_range(batch, nt, i, 1, &lo, &hi);
    if (lo >= hi) continue;
    F(BatchJob) j = {n, lo, hi, D, scale, x, y};
    jobs[used] = j;
    pthread_create(&th[used], NULL, F(batch_worker), &jobs[used]);
    ++used;
  }
  for (int i = 0; i < used; ++i) pthread_join(th[i], NULL);
  return 0;
}

/* the reference's N-d plan on nt threads (ndplan.jl:57-118, 270-327): axis
 * by axis from `first`, the pencils of an axis in contiguous chunks over
 * the threads (threads.jl:15-38), one chain per thread; in place on y, the
 * scale on the last transformed axis.  13-smooth dims only. */
typedef struct {
  int64_t nd, Bi, lo, hi;
  int D;
  T scale;
  T* y;
  int err;
} F(NdJob);

static void* F(nd_worker)(void* arg) {
  F(NdJob)* j = (F(NdJob)*)arg;
  F(Chain) ch;
  if (F(chain_init)(&ch, j->nd, j->D)) {
    j->err = 1;
    return NULL;
  }
  const int64_t nd = j->nd, Bi = j->Bi;
  T* buf = (T*)malloc(sizeof(T) * 4 * (size_t)nd);
  T *are = buf, *aim = buf + nd, *bre = buf + 2 * nd, *bim = buf + 3 * nd;
  for (int64_t p = j->lo; p < j->hi; ++p) {
    const int64_t bi = p % Bi, bo = p / Bi;
    T* base = j->y + 2 * (bi + bo * Bi * nd);
    for (int64_t i = 0; i < nd; ++i) {
      are[i] = base[2 * i * Bi];
      aim[i] = base[2 * i * Bi + 1];
    }
    int which = F(chain_run)(&ch, are, aim, bre, bim, 1, j->D);
    T* rr = which ? bre : are;
    T* ri = which ? bim : aim;
    for (int64_t i = 0; i < nd; ++i) {
      base[2 * i * Bi] = rr[i] * j->scale;
      base[2 * i * Bi + 1] = ri[i] * j->scale;
    }
  }
  free(buf);
  F(chain_free)(&ch);
  return NULL;
}

int F(baseline_ndfft)(int rank, const int64_t* dims, int first, int D, T scale, T* y, int nt) {
  if (nt < 1) nt = 1;
  if (nt > 256) nt = 256;
  int last = -1;
  for (int i = first; i < rank; ++i) {
    if (!is_smooth13(dims[i])) return -1;
    if (dims[i] > 1) last = i;
  }
  for (int ax = first; ax < rank; ++ax) {
    const int64_t nd = dims[ax];
    if (nd == 1) continue;
    int64_t Bi = 1, Bo = 1;
    for (int i = 0; i < ax; ++i) Bi *= dims[i];
    for (int i = ax + 1; i < rank; ++i) Bo *= dims[i];
    pthread_t th[256];
    F(NdJob) jobs[256];
    int used = 0;
    for (int i = 0; i < nt; ++i) {
      int64_t lo, hi;
      chunk_range(Bi * Bo, nt, i, 1, &lo, &hi);
      if (lo >= hi) continue;
      F(NdJob) j = {nd, Bi, lo, hi, D, ax == last ? scale : (T)1, y, 0};
      jobs[used] = j;
      pthread_create(&th[used], NULL, F(nd_worker), &jobs[used]);
      ++used;
    }
    int err = 0;
    for (int i = 0; i < used; ++i) {
      pthread_join(th[i], NULL);
      err |= jobs[i].err;
    }
    if (err) return -1;
  }
  return 0;
}

/* rfft / irfft rows of an even length n on nt threads (real.jl:80-131,
 * 144-193 per row): one half-length chain and twist table per thread */
typedef struct {
  int64_t n, lo, hi;
  int inverse, normalized, err;
  const T* x;
  T* y;
} F(RealJob);

static void* F(real_worker)(void* arg) {
  F(RealJob)* j = (F(RealJob)*)arg;
  const int64_t n = j->n, h = n / 2;
  const int D = j->inverse ? 1 : -1;
  F(Chain) ch;
  if (F(chain_init)(&ch, h, D)) {
    j->err = 1;
    return NULL;
  }
  T* buf = (T*)malloc(sizeof(T) * 6 * (size_t)h);
  T *are = buf, *aim = buf + h, *bre = buf + 2 * h, *bim = buf + 3 * h, *wr = buf + 4 * h, *wi = buf + 5 * h;
  for (int64_t k = 0; k < h; ++k) {
    long double c, s;
    expi_exact(k, n, D, &c, &s);
    wr[k] = (T)c;
    wi[k] = (T)s;
  }
  for (int64_t b = j->lo; b < j->hi; ++b) {
    if (!j->inverse) {
      const T* x = j->x + n * b;
      T* y = j->y + 2 * (h + 1) * b;
      for (int64_t i = 0; i < h; ++i) {
        are[i] = x[2 * i];
        aim[i] = x[2 * i + 1];
      }
      int which = F(chain_run)(&ch, are, aim, bre, bim, 1, D);
      T* zr = which ? bre : are;
      T* zi = which ? bim : aim;
      y[0] = zr[0] + zi[0];
      y[1] = 0;
      y[2 * h] = zr[0] - zi[0];
      y[2 * h + 1] = 0;
      for (int64_t k = 1; k < h; ++k) {
        T ar = zr[k], ai = zi[k], br = zr[h - k], bi = -zi[h - k];
        T xer = (ar + br) * (T)0.5, xei = (ai + bi) * (T)0.5;
        T tr = (ai - bi) * (T)0.5, ti = -(ar - br) * (T)0.5;
        y[2 * k] = xer + (wr[k] * tr - wi[k] * ti);
        y[2 * k + 1] = xei + (wr[k] * ti + wi[k] * tr);
      }
    } else {
      const T* X = j->x + 2 * (h + 1) * b;
      T* y = j->y + n * b;
      const T pre = j->normalized ? (T)0.5 : (T)1;
      for (int64_t k = 0; k < h; ++k) {
        T ar = X[2 * k], ai = X[2 * k + 1], br = X[2 * (h - k)], bi = -X[2 * (h - k) + 1];
        T xer = (ar + br) * pre, xei = (ai + bi) * pre;
        T xdr = (ar - br) * pre, xdi = (ai - bi) * pre;
        T tr = wr[k] * xdr - wi[k] * xdi, ti = wr[k] * xdi + wi[k] * xdr;
        are[k] = xer - ti;
        aim[k] = xei + tr;
      }
      int which = F(chain_run)(&ch, are, aim, bre, bim, 1, D);
      T* zr = which ? bre : are;
      T* zi = which ? bim : aim;
      const T sc = j->normalized ? (T)(1.0L / h) : (T)1;
      for (int64_t i = 0; i < h; ++i) {
        y[2 * i] = zr[i] * sc;
        y[2 * i + 1] = zi[i] * sc;
      }
    }
  }
  free(buf);
  F(chain_free)(&ch);
  return NULL;
}

int F(baseline_real_batch)(int64_t n, int64_t batch, int inverse, int normalized, const T* x, T* y, int nt) {
  if (n < 2 || n % 2 || !is_smooth13(n / 2)) return -1;
  if (nt < 1) nt = 1;
  if (nt > 256) nt = 256;
  pthread_t th[256];
  F(RealJob) jobs[256];
  int used = 0;
  for (int i = 0; i < nt; ++i) {
    int64_t lo, hi;
    chunk_range(batch, nt, i, 1, &lo, &hi);
    if (lo >= hi) continue;
    F(RealJob) j = {n, lo, hi, inverse, normalized, 0, x, y};
    jobs[used] = j;
    pthread_create(&th[used], NULL, F(real_worker), &jobs[used]);
    ++used;
  }
  int err = 0;
  for (int i = 0; i < used; ++i) {
    pthread_join(th[i], NULL);
    err |= jobs[i].err;
  }
  return err ? -1 : 0;
}

#undef F
#undef CAT
#undef CAT2
