/*
 * canon_fft_impl.h -- canonical power-of-two FFT, instantiated for double and float.
 * TEST INFRASTRUCTURE ONLY (part of the oracle; see sfft_oracle.h).
 *
 * Included twice by sfft_oracle.c with
 *   CF_REAL  (double | float), CF_CPLX (orc_c64 | orc_c32), CF_FMA (fma | fmaf),
 *   CF_SFX   (64 | 32).
 *
 * The arithmetic below is the definition the HIP kernels reproduce bit for bit:
 *  - passes: log2N is split into p = ceil(log2N / 11) passes, the first (log2N mod p)
 *    passes one bit larger (orc_fft_plan);
 *  - pass i (radix R = 2^r, Lprev = product of earlier radices, Lcur = Lprev*R) maps
 *      a[v] = in[(q + (N/Lcur) v) Lprev + t]                       v < R
 *      a[v] *= W_Lcur^(t v)          (passes after the first, v >= 1)
 *      A = subfft_R(a)
 *      out[q Lcur + t + Lprev k] = A[k]                            k < R
 *    for q < N/Lcur, t < Lprev (self-sorting Stockham with a big radix);
 *  - W_T^e for log2 T <= 11 is the table value tw(e, T); above, with
 *    s = min(ceil(log2 T / 2), 8) (a low table of at most 256 entries),
 *    W_T^e = cmul(tw((e >> s) << s, T), tw(e & (2^s - 1), T));
 *  - subfft_R: radix-4 Stockham stages (then one radix-2 stage if r is odd); stage with
 *    span ns and radix rho maps c[u] = a[j + u R/rho], multiplies c[u] (u >= 1, ns > 1) by
 *    tw((j mod ns) u R/(ns rho), R), applies the butterfly, and stores
 *    a[(j/ns) ns rho + (j mod ns) + u ns] = c[u];
 *  - cmul(a, w) = (fma(a.x, w.x, -(a.y w.y)), fma(a.x, w.y, a.y w.x));
 *  - radix-4 forward: t0=a0+a2, t1=a0-a2, t2=a1+a3, t3=(a1-a3)*(-i);
 *    y0=t0+t2, y1=t1+t3, y2=t0-t2, y3=t1-t3 (inverse uses +i);
 *  - tw(e, T) = (cos 2 pi e/T, -+ sin 2 pi e/T) from orc_twiddle64, rounded to CF_REAL.
 */

#define CF_CAT2(a, b) a##b
#define CF_CAT(a, b) CF_CAT2(a, b)
#define CF_FN(name) CF_CAT(name, CF_SFX)

static inline CF_CPLX CF_FN(cf_cmul)(CF_CPLX a, CF_CPLX w) {
  CF_CPLX r;
  r.x = CF_FMA(a.x, w.x, -(a.y * w.y));
  r.y = CF_FMA(a.x, w.y, a.y * w.x);
  return r;
}

static inline CF_CPLX CF_FN(cf_tw)(uint64_t e, uint64_t T, int inverse) {
  orc_c64 d;
  orc_twiddle64(e, T, inverse, &d);
  CF_CPLX r;
  r.x = (CF_REAL)d.x;
  r.y = (CF_REAL)d.y;
  return r;
}

/* in-place sub-FFT of size R = 2^r on a (tmp has R entries) */
static void CF_FN(cf_subfft)(CF_CPLX *a, int r, int inverse, CF_CPLX *tmp, const CF_CPLX *twR) {
  const int R = 1 << r;
  int ns = 1;
  int rem = r;
  while (rem > 0) {
    const int rho = (rem >= 2) ? 4 : 2;
    memcpy(tmp, a, sizeof(CF_CPLX) * (size_t)R);
    const int stride = R / rho;
    const int tstep = R / (ns * rho);
    for (int j = 0; j < stride; ++j) {
      CF_CPLX c[4];
      for (int u = 0; u < rho; ++u) c[u] = tmp[j + u * stride];
      if (ns > 1) {
        const int jm = j % ns;
        for (int u = 1; u < rho; ++u) c[u] = CF_FN(cf_cmul)(c[u], twR[(size_t)jm * u * tstep]);
      }
      CF_CPLX y[4];
      if (rho == 4) {
        CF_CPLX t0, t1, t2, t3, d;
        t0.x = c[0].x + c[2].x; t0.y = c[0].y + c[2].y;
        t1.x = c[0].x - c[2].x; t1.y = c[0].y - c[2].y;
        t2.x = c[1].x + c[3].x; t2.y = c[1].y + c[3].y;
        d.x = c[1].x - c[3].x;  d.y = c[1].y - c[3].y;
        if (!inverse) { t3.x = d.y; t3.y = -d.x; } else { t3.x = -d.y; t3.y = d.x; }
        y[0].x = t0.x + t2.x; y[0].y = t0.y + t2.y;
        y[1].x = t1.x + t3.x; y[1].y = t1.y + t3.y;
        y[2].x = t0.x - t2.x; y[2].y = t0.y - t2.y;
        y[3].x = t1.x - t3.x; y[3].y = t1.y - t3.y;
      } else {
        y[0].x = c[0].x + c[1].x; y[0].y = c[0].y + c[1].y;
        y[1].x = c[0].x - c[1].x; y[1].y = c[0].y - c[1].y;
      }
      const int base = (j / ns) * ns * rho + (j % ns);
      for (int u = 0; u < rho; ++u) a[base + u * ns] = y[u];
    }
    ns *= rho;
    rem -= (rho == 4) ? 2 : 1;
  }
}

void CF_FN(orc_fft)(CF_CPLX *data, int log2N, int inverse, CF_CPLX *scratch) {
  if (log2N <= 0) return;
  int logs[8];
  const int np = orc_fft_plan(log2N, logs);
  const size_t N = (size_t)1 << log2N;
  CF_CPLX *in = data, *out = scratch;
  size_t Lprev = 1;
  CF_CPLX *a = (CF_CPLX *)malloc(sizeof(CF_CPLX) * 4096);
  CF_CPLX *tmp = (CF_CPLX *)malloc(sizeof(CF_CPLX) * 4096);
  CF_CPLX *twR = (CF_CPLX *)malloc(sizeof(CF_CPLX) * 4096);
  int logLprev = 0;
  for (int pi = 0; pi < np; ++pi) {
    const int r = logs[pi];
    const size_t R = (size_t)1 << r;
    const size_t Lcur = Lprev * R;
    const int logLcur = logLprev + r;
    const size_t nq = N / Lcur;
    for (size_t e = 0; e < R; ++e) twR[e] = CF_FN(cf_tw)(e, R, inverse);
    /* outer twiddle tables: direct below 2^12, two-level above (values from cf_tw) */
    CF_CPLX *tdir = NULL, *thi = NULL, *tlo = NULL;
    int s = 0;
    if (pi > 0) {
      if (logLcur <= 11) {
        tdir = (CF_CPLX *)malloc(sizeof(CF_CPLX) * Lcur);
        for (size_t e = 0; e < Lcur; ++e) tdir[e] = CF_FN(cf_tw)(e, Lcur, inverse);
      } else {
        s = (logLcur + 1) / 2 < 8 ? (logLcur + 1) / 2 : 8;
        const size_t nlo = (size_t)1 << s, nhi = Lcur >> s;
        tlo = (CF_CPLX *)malloc(sizeof(CF_CPLX) * nlo);
        thi = (CF_CPLX *)malloc(sizeof(CF_CPLX) * nhi);
        for (size_t e = 0; e < nlo; ++e) tlo[e] = CF_FN(cf_tw)(e, Lcur, inverse);
        for (size_t e = 0; e < nhi; ++e) thi[e] = CF_FN(cf_tw)(e << s, Lcur, inverse);
      }
    }
    for (size_t q = 0; q < nq; ++q) {
      for (size_t t = 0; t < Lprev; ++t) {
        for (size_t v = 0; v < R; ++v) a[v] = in[(q + nq * v) * Lprev + t];
        if (pi > 0) {
          for (size_t v = 1; v < R; ++v) {
            const uint64_t e = (uint64_t)t * v;
            CF_CPLX w;
            if (tdir) w = tdir[e];
            else w = CF_FN(cf_cmul)(thi[e >> s], tlo[e & (((uint64_t)1 << s) - 1)]);
            a[v] = CF_FN(cf_cmul)(a[v], w);
          }
        }
        CF_FN(cf_subfft)(a, r, inverse, tmp, twR);
        for (size_t k = 0; k < R; ++k) out[q * Lcur + t + Lprev * k] = a[k];
      }
    }
    free(tdir); free(thi); free(tlo);
    CF_CPLX *sw = in; in = out; out = sw;
    Lprev = Lcur;
    logLprev = logLcur;
  }
  if (in != data) memcpy(data, in, sizeof(CF_CPLX) * N);
  free(a); free(tmp); free(twR);
}

#undef CF_FN
#undef CF_CAT
#undef CF_CAT2
