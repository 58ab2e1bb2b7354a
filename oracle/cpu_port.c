/*
 * C restatement of the reference's CPU path for the PRNU hot path
 * (TEST INFRASTRUCTURE ONLY: used by tests/ and by the CPU-baseline leg of
 * bench.py, never by the product library).
 *
 * Reference paths are relative to
 * src/main/java/nl/esciencecenter/common_source_identification/.
 *
 * Built by oracle/Makefile into oracle/_build/libprnu_cpu_port.so.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* NormalizedCrossCorrelation.applyCPU (kernels/compare/NormalizedCrossCorrelation.java:127-133):
 * sum_xy accumulates the FP32 product x[i]*y[i] in FP64, and
 * NCC = sum_xy / sqrt(sumsq_x * sumsq_y). */
double port_ncc_cpu(const float *x, double sumsq_x, const float *y, double sumsq_y, int64_t n) {
    double sum_xy = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        float p = x[i] * y[i];
        sum_xy += (double)p;
    }
    return sum_xy / sqrt(sumsq_x * sumsq_y);
}

/* Sum of FP32 squares in FP64 (the sumsq argument the caller of applyCPU passes). */
double port_sumsq(const float *x, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        float p = x[i] * x[i];
        s += (double)p;
    }
    return s;
}

/* A list of pairs over a pattern table, parallel over pairs with `threads`
 * OpenMP threads (0 = all).  patterns[k] points at pattern k (n floats);
 * pairs holds (i, j) index pairs; sumsq[k] is port_sumsq(patterns[k]). */
void port_ncc_pairs(const float *const *patterns, const double *sumsq, const int32_t *pairs, int64_t n_pairs,
                    int64_t n, double *out, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t p = 0; p < n_pairs; ++p) {
        const int32_t i = pairs[2 * p], j = pairs[2 * p + 1];
        out[p] = port_ncc_cpu(patterns[i], sumsq[i], patterns[j], sumsq[j], n);
    }
}

void port_sumsq_many(const float *const *patterns, int64_t count, int64_t n, double *out, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t k = 0; k < count; ++k) out[k] = port_sumsq(patterns[k], n);
}

int port_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ======================================================================
 * Extraction chain and PCE on the CPU, for timing the reference's CPU path
 * (the applyCPU methods of kernels/filter/).  FP32 data as the Java code; the FFT is
 * this file's own (the reference uses JTransforms 2.4, which is not in its
 * tree): a self-sorting (Stockham) mixed-radix FFT, radices 4, 2, 3, 5,
 * twiddles computed in FP64.  PCE has no CPU code in the reference; its
 * restatement follows kernels/compare/peaktocorrelationenergy.cu:102-368.
 * ====================================================================== */

static const double kPi = 3.14159265358979323846;

typedef struct {
    float re, im;
} cpx;

typedef struct {
    int n, nstage;
    int radix[40];
    cpx *root; /* exp(-2 pi i t / n), t < n */
} fft_plan;

static int fft_plan_init(fft_plan *p, int n) {
    p->n = n;
    p->nstage = 0;
    int m = n;
    while (m % 4 == 0) { p->radix[p->nstage++] = 4; m /= 4; }
    while (m % 2 == 0) { p->radix[p->nstage++] = 2; m /= 2; }
    while (m % 3 == 0) { p->radix[p->nstage++] = 3; m /= 3; }
    while (m % 5 == 0) { p->radix[p->nstage++] = 5; m /= 5; }
    if (m != 1) return -1;
    p->root = (cpx *)malloc(sizeof(cpx) * (size_t)n);
    if (!p->root) return -1;
    for (int t = 0; t < n; ++t) {
        const double a = -2.0 * kPi * (double)t / (double)n;
        p->root[t].re = (float)cos(a);
        p->root[t].im = (float)sin(a);
    }
    return 0;
}

static void fft_plan_free(fft_plan *p) { free(p->root); }

static inline cpx cmul(cpx a, cpx b) {
    cpx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}

/* In-place forward (dir = -1) or unnormalised inverse (dir = +1) transform
 * of x (n values) with work buffer y.  Stage s of radix R with span Ns (the
 * product of the earlier radices): for j < n/R, the R inputs x[j + r n/R]
 * are twiddled by w^(r (j mod Ns) n/(Ns R)), transformed by an R-point DFT
 * and written to y[(j/Ns) Ns R + j mod Ns + r Ns]. */
static void fft_run(const fft_plan *p, cpx *x, cpx *y, int dir) {
    const int n = p->n;
    cpx *a = x, *b = y;
    int Ns = 1;
    const double c3 = -0.5, s3 = sqrt(3.0) / 2.0;
    const double c51 = cos(2.0 * kPi / 5.0), c52 = cos(4.0 * kPi / 5.0);
    const double s51 = sin(2.0 * kPi / 5.0), s52 = sin(4.0 * kPi / 5.0);
    const float sg = dir < 0 ? 1.0f : -1.0f; /* sign of the sines of exp(-2 pi i / R) */
    for (int st = 0; st < p->nstage; ++st) {
        const int R = p->radix[st], m = n / R, step = n / (Ns * R);
        for (int j = 0; j < m; ++j) {
            const int jm = j % Ns;
            cpx v[5] = {{0.0f, 0.0f}};
            for (int r = 0; r < R; ++r) {
                cpx w = p->root[(size_t)r * jm * step];
                if (dir > 0) w.im = -w.im;
                v[r] = r ? cmul(a[j + r * m], w) : a[j + r * m];
            }
            cpx o[5];
            if (R == 2) {
                o[0].re = v[0].re + v[1].re; o[0].im = v[0].im + v[1].im;
                o[1].re = v[0].re - v[1].re; o[1].im = v[0].im - v[1].im;
            } else if (R == 4) {
                const cpx t0 = {v[0].re + v[2].re, v[0].im + v[2].im}, t1 = {v[0].re - v[2].re, v[0].im - v[2].im};
                const cpx t2 = {v[1].re + v[3].re, v[1].im + v[3].im}, t3 = {v[1].re - v[3].re, v[1].im - v[3].im};
                o[0].re = t0.re + t2.re; o[0].im = t0.im + t2.im;
                o[2].re = t0.re - t2.re; o[2].im = t0.im - t2.im;
                /* -i * t3 (forward) or +i * t3 (inverse) */
                o[1].re = t1.re + sg * t3.im; o[1].im = t1.im - sg * t3.re;
                o[3].re = t1.re - sg * t3.im; o[3].im = t1.im + sg * t3.re;
            } else if (R == 3) {
                const float sr = v[1].re + v[2].re, si = v[1].im + v[2].im;
                const float dr = v[1].re - v[2].re, di = v[1].im - v[2].im;
                const float mr = v[0].re + (float)c3 * sr, mi = v[0].im + (float)c3 * si;
                o[0].re = v[0].re + sr; o[0].im = v[0].im + si;
                o[1].re = mr + sg * (float)s3 * di; o[1].im = mi - sg * (float)s3 * dr;
                o[2].re = mr - sg * (float)s3 * di; o[2].im = mi + sg * (float)s3 * dr;
            } else { /* R == 5 */
                const float s1r = v[1].re + v[4].re, s1i = v[1].im + v[4].im;
                const float d1r = v[1].re - v[4].re, d1i = v[1].im - v[4].im;
                const float s2r = v[2].re + v[3].re, s2i = v[2].im + v[3].im;
                const float d2r = v[2].re - v[3].re, d2i = v[2].im - v[3].im;
                const float a1r = v[0].re + (float)c51 * s1r + (float)c52 * s2r;
                const float a1i = v[0].im + (float)c51 * s1i + (float)c52 * s2i;
                const float a2r = v[0].re + (float)c52 * s1r + (float)c51 * s2r;
                const float a2i = v[0].im + (float)c52 * s1i + (float)c51 * s2i;
                const float b1r = (float)s51 * d1r + (float)s52 * d2r, b1i = (float)s51 * d1i + (float)s52 * d2i;
                const float b2r = (float)s52 * d1r - (float)s51 * d2r, b2i = (float)s52 * d1i - (float)s51 * d2i;
                o[0].re = v[0].re + s1r + s2r; o[0].im = v[0].im + s1i + s2i;
                o[1].re = a1r + sg * b1i; o[1].im = a1i - sg * b1r;
                o[4].re = a1r - sg * b1i; o[4].im = a1i + sg * b1r;
                o[2].re = a2r + sg * b2i; o[2].im = a2i - sg * b2r;
                o[3].re = a2r - sg * b2i; o[3].im = a2i + sg * b2r;
            }
            const int base = (j / Ns) * Ns * R + jm;
            for (int r = 0; r < R; ++r) b[base + r * Ns] = o[r];
        }
        cpx *t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    if (a != x) memcpy(x, a, sizeof(cpx) * (size_t)n);
}

/* Forward or inverse 2-D transform of an h x w row-major complex array:
 * rows, then columns through a gathered buffer (as WienerFilter.applyCPU
 * transforms its columns).  Single thread; work: 2 * max(h, w) values. */
static void fft2(const fft_plan *ph, const fft_plan *pw, cpx *a, int h, int w, int dir, cpx *work) {
    cpx *col = work, *tmp = work + (h > w ? h : w);
    for (int y = 0; y < h; ++y) fft_run(pw, a + (size_t)y * w, tmp, dir);
    for (int x = 0; x < w; ++x) {
        for (int y = 0; y < h; ++y) col[y] = a[(size_t)y * w + x];
        fft_run(ph, col, tmp, dir);
        for (int y = 0; y < h; ++y) a[(size_t)y * w + x] = col[y];
    }
}

/* Forward complex FFT of n values (tests). */
int port_fft(cpx *x, int n, int inverse) {
    fft_plan p;
    if (fft_plan_init(&p, n)) return -1;
    cpx *work = (cpx *)malloc(sizeof(cpx) * (size_t)n);
    if (!work) { fft_plan_free(&p); return -1; }
    fft_run(&p, x, work, inverse ? 1 : -1);
    free(work);
    fft_plan_free(&p);
    return 0;
}

/* GrayscaleFilter.applyCPU (kernels/filter/GrayscaleFilter.java:87-100):
 * bytes 0, 1, 2 of a pixel are taken as b, g, r: 0.299 r + 0.587 g + 0.114 b.
 * Gray input (fmt 1) is one byte fed as r = g = b. */
static void gray_cpu(const uint8_t *img, int fmt, int h, int w, float *out) {
    const int64_t n = (int64_t)h * w;
    for (int64_t i = 0; i < n; ++i) {
        if (fmt == 1) {
            const float v = (float)img[i];
            out[i] = 0.299f * v + 0.587f * v + 0.114f * v;
        } else {
            const float b = (float)img[3 * i], g = (float)img[3 * i + 1], r = (float)img[3 * i + 2];
            out[i] = 0.299f * r + 0.587f * g + 0.114f * b;
        }
    }
}

/* FastNoiseFilter.applyCPU (kernels/filter/FastNoiseFilter.java:130-204):
 * gradients (forward differences at the borders, centred inside),
 * normalised by 1 / (EPS + |grad|) with EPS = 1, then the same differences
 * of the normalised gradients summed into the image (its divergence). */
static void fastnoise_cpu(float *p, int h, int w, float *dx, float *dy) {
    for (int y = 0; y < h; ++y) {
        const float *s = p + (size_t)y * w;
        float *d = dx + (size_t)y * w;
        d[0] = s[1] - s[0];
        d[w - 1] = s[w - 1] - s[w - 2];
        for (int i = 1; i < w - 1; ++i) d[i] = 0.5f * (s[i + 1] - s[i - 1]);
    }
    for (int x = 0; x < w; ++x) {
        dy[x] = p[(size_t)w + x] - p[x];
        dy[(size_t)(h - 1) * w + x] = p[(size_t)(h - 1) * w + x] - p[(size_t)(h - 2) * w + x];
    }
    for (int y = 1; y < h - 1; ++y)
        for (int x = 0; x < w; ++x) dy[(size_t)y * w + x] = 0.5f * (p[(size_t)(y + 1) * w + x] - p[(size_t)(y - 1) * w + x]);
    const int64_t n = (int64_t)h * w;
    for (int64_t i = 0; i < n; ++i) {
        const float gx = dx[i], gy = dy[i];
        const float norm = sqrtf(gx * gx + gy * gy);
        const float sc = 1.0f / (1.0f + norm);
        dx[i] = gx * sc;
        dy[i] = gy * sc;
    }
    for (int y = 0; y < h; ++y) {
        const float *s = dx + (size_t)y * w;
        float *d = p + (size_t)y * w;
        d[0] = s[1] - s[0];
        d[w - 1] = s[w - 1] - s[w - 2];
        for (int i = 1; i < w - 1; ++i) d[i] = 0.5f * (s[i + 1] - s[i - 1]);
    }
    for (int x = 0; x < w; ++x) {
        p[x] += dy[(size_t)w + x] - dy[x];
        p[(size_t)(h - 1) * w + x] += dy[(size_t)(h - 1) * w + x] - dy[(size_t)(h - 2) * w + x];
    }
    for (int y = 1; y < h - 1; ++y)
        for (int x = 0; x < w; ++x)
            p[(size_t)y * w + x] += 0.5f * (dy[(size_t)(y + 1) * w + x] - dy[(size_t)(y - 1) * w + x]);
}

/* ZeroMeanTotalFilter.applyCPU (kernels/filter/ZeroMeanTotalFilter.java:108-163):
 * every row, then every column, loses the mean of its even and of its odd
 * elements (FP32 sums). */
static void zeromean_cpu(float *p, int h, int w) {
    for (int y = 0; y < h; ++y) {
        float *r = p + (size_t)y * w, se = 0.0f, so = 0.0f;
        for (int i = 0; i < w - 1; i += 2) { se += r[i]; so += r[i + 1]; }
        if (w & 1) se += r[w - 1];
        const float me = se / (float)((w + 1) >> 1), mo = so / (float)(w >> 1);
        for (int i = 0; i < w - 1; i += 2) { r[i] -= me; r[i + 1] -= mo; }
        if (w & 1) r[w - 1] -= me;
    }
    for (int x = 0; x < w; ++x) {
        float se = 0.0f, so = 0.0f;
        for (int i = 0; i < h - 1; i += 2) { se += p[(size_t)i * w + x]; so += p[(size_t)(i + 1) * w + x]; }
        if (h & 1) se += p[(size_t)(h - 1) * w + x];
        const float me = se / (float)((h + 1) >> 1), mo = so / (float)(h >> 1);
        for (int i = 0; i < h - 1; i += 2) { p[(size_t)i * w + x] -= me; p[(size_t)(i + 1) * w + x] -= mo; }
        if (h & 1) p[(size_t)(h - 1) * w + x] -= me;
    }
}

/* WienerFilter.applyCPU (kernels/filter/WienerFilter.java:256-422): FFT2 of
 * the pattern; P = |F|^2 into a zero-bordered array; for each filter size
 * f in {3, 5, 7, 9} the f x f box sums of P / f^2 by running column sums
 * (FP32) and the minimum over f; var = n / (n - 1) * sum x^2 (FP64 sum);
 * F *= var / max(var, estimate); x = Re(IFFT2(F)) / n. */
static void wiener_cpu(float *p, int h, int w, const fft_plan *ph, const fft_plan *pw, cpx *F, float *est,
                       float *mag, float *colsum, cpx *work) {
    enum { kB = 4 };
    const int64_t n = (int64_t)h * w;
    const int pw2 = w + 2 * kB;
    double sumsq = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        F[i].re = p[i];
        F[i].im = 0.0f;
        sumsq += (double)(p[i] * p[i]);
    }
    fft2(ph, pw, F, h, w, -1, work);
    memset(mag, 0, sizeof(float) * (size_t)(h + 2 * kB + 1) * pw2);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const cpx f = F[(size_t)y * w + x];
            mag[(size_t)(y + kB) * pw2 + kB + x] = f.re * f.re + f.im * f.im;
        }
    for (int64_t i = 0; i < n; ++i) est[i] = 3.402823466e38f;
    static const int sizes[4] = {3, 5, 7, 9};
    for (int q = 0; q < 4; ++q) {
        const int fs = sizes[q], off = kB - (fs - 1) / 2, cols = pw2 - 2 * off;
        const float sc = 1.0f / (float)(fs * fs);
        memset(colsum, 0, sizeof(float) * (size_t)(pw2 + 1));
        for (int y = 0; y < fs; ++y) {
            const float *row = mag + (size_t)(y + off) * pw2;
            for (int x = 0; x < cols; ++x) colsum[x + 1] += row[off + x];
        }
        for (int x = 1; x <= cols; ++x) colsum[x] *= sc;
        for (int y = 0; y < h; ++y) {
            float s = 0.0f;
            for (int x = 1; x < fs; ++x) s += colsum[x];
            float *e = est + (size_t)y * w;
            for (int x = 0; x < w; ++x) {
                s += colsum[x + fs] - colsum[x];
                if (s < e[x]) e[x] = s;
            }
            const float *top = mag + (size_t)(y + off) * pw2, *bot = mag + (size_t)(y + off + fs) * pw2;
            for (int x = 0; x < cols; ++x) colsum[x + 1] += sc * (bot[off + x] - top[off + x]);
        }
    }
    const float var = (float)((sumsq * (double)n) / (double)(n - 1));
    for (int64_t i = 0; i < n; ++i) {
        const float s = var / (var > est[i] ? var : est[i]);
        F[i].re *= s;
        F[i].im *= s;
    }
    fft2(ph, pw, F, h, w, 1, work);
    const float inv = 1.0f / (float)n;
    for (int64_t i = 0; i < n; ++i) p[i] = F[i].re * inv;
}

typedef struct {
    fft_plan ph, pw;
    cpx *F, *work;
    float *a, *b, *est, *mag, *colsum;
} chain_ws;

static int chain_ws_init(chain_ws *s, int h, int w) {
    memset(s, 0, sizeof(*s));
    if (fft_plan_init(&s->ph, h) || fft_plan_init(&s->pw, w)) return -1;
    const size_t n = (size_t)h * w, mx = (size_t)(h > w ? h : w);
    s->F = (cpx *)malloc(sizeof(cpx) * n);
    s->work = (cpx *)malloc(sizeof(cpx) * 2 * mx);
    s->a = (float *)malloc(sizeof(float) * n);
    s->b = (float *)malloc(sizeof(float) * n);
    s->est = (float *)malloc(sizeof(float) * n);
    s->mag = (float *)malloc(sizeof(float) * (size_t)(h + 9) * (size_t)(w + 8));
    s->colsum = (float *)malloc(sizeof(float) * (size_t)(w + 9));
    return (s->F && s->work && s->a && s->b && s->est && s->mag && s->colsum) ? 0 : -1;
}

static void chain_ws_free(chain_ws *s) {
    fft_plan_free(&s->ph);
    fft_plan_free(&s->pw);
    free(s->F); free(s->work); free(s->a); free(s->b); free(s->est); free(s->mag); free(s->colsum);
}

/* PRNUFilter.applyCPU chain (kernels/filter/PRNUFilter.java:124-138) of one
 * image into `pattern` (h * w floats); with `spectrum` (not NULL) also the
 * R2C half spectrum h x (w/2+1) (SpectralFilter.applyCPU,
 * kernels/filter/SpectralFilter.java:56-75). */
static void chain_one(chain_ws *s, const uint8_t *img, int fmt, int h, int w, float *pattern, float *spectrum) {
    gray_cpu(img, fmt, h, w, s->a);
    fastnoise_cpu(s->a, h, w, s->b, s->est);
    zeromean_cpu(s->a, h, w);
    wiener_cpu(s->a, h, w, &s->ph, &s->pw, s->F, s->est, s->mag, s->colsum, s->work);
    memcpy(pattern, s->a, sizeof(float) * (size_t)h * w);
    if (spectrum) {
        const int64_t n = (int64_t)h * w;
        const int hw = w / 2 + 1;
        for (int64_t i = 0; i < n; ++i) { s->F[i].re = s->a[i]; s->F[i].im = 0.0f; }
        fft2(&s->ph, &s->pw, s->F, h, w, -1, s->work);
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < hw; ++x) {
                spectrum[2 * ((size_t)y * hw + x)] = s->F[(size_t)y * w + x].re;
                spectrum[2 * ((size_t)y * hw + x) + 1] = s->F[(size_t)y * w + x].im;
            }
    }
}

/* `count` images (count * h * w * (fmt ? 1 : 3) bytes) to patterns (and
 * spectra when not NULL), one image per thread.  Returns 0, -1 on a size
 * the FFT does not take (factors other than 2, 3, 5) or no memory. */
int port_prnu_batch(const uint8_t *imgs, int count, int h, int w, int fmt, float *patterns, float *spectra,
                    int threads) {
    if (h < 2 || w < 2) return -1;
    const size_t in_stride = (size_t)h * w * (fmt == 1 ? 1 : 3), hw = (size_t)h * (w / 2 + 1) * 2;
    int err = 0;
#ifdef _OPENMP
    if (threads < 1) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads) reduction(| : err)
#endif
    {
        chain_ws s;
        if (chain_ws_init(&s, h, w)) {
            err = 1;
        } else {
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
            for (int k = 0; k < count; ++k)
                chain_one(&s, imgs + k * in_stride, fmt, h, w, patterns + (size_t)k * h * w,
                          spectra ? spectra + k * hw : NULL);
        }
        chain_ws_free(&s);
    }
    return err ? -1 : 0;
}

/* PCE of two R2C half spectra (h x (w/2+1) complex, FP32 interleaved), as
 * kernels/compare/peaktocorrelationenergy.cu:102-368 computes it.
 * variant 0 (reference): the full cross spectrum takes conj(X) Y for
 * j <= w/2 and X conj(Y) of column w - j on the same row for j > w/2; the
 * plane is the unnormalised inverse FFT; peak location argmax |c|, energy of
 * c^2 outside the 9 x 9 window (rows and columns strictly within 5 of the
 * peak, clipped) over n - 121 in FP64; the peak value is the plane's last
 * element, squared in FP32.
 * variant 1 (textbook): the Hermitian full spectrum conj(X) Y, peak at argmax
 * |c|, energy outside the 11 x 11 window with cyclic wrap over the rest,
 * PCE = sign(c_p) c_p^2 / energy. */
static double pce_one(const float *xs, const float *ys, int h, int w, int variant, const fft_plan *ph,
                      const fft_plan *pw, cpx *C, cpx *work) {
    const int hw = w / 2 + 1;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            cpx a, b, c;
            if (x < hw) {
                const size_t i = 2 * ((size_t)y * hw + x);
                a.re = xs[i]; a.im = -xs[i + 1];  /* conj(X) */
                b.re = ys[i]; b.im = ys[i + 1];
            } else if (variant == 0) {
                const size_t i = 2 * ((size_t)y * hw + (w - x));
                a.re = xs[i]; a.im = xs[i + 1];   /* X conj(Y), same row */
                b.re = ys[i]; b.im = -ys[i + 1];
            } else {
                const int yy = (h - y) % h;       /* Hermitian: X[y][x] = conj(X[h-y][w-x]) */
                const size_t i = 2 * ((size_t)yy * hw + (w - x));
                a.re = xs[i]; a.im = xs[i + 1];
                b.re = ys[i]; b.im = -ys[i + 1];
            }
            c = cmul(a, b);
            C[(size_t)y * w + x] = c;
        }
    fft2(ph, pw, C, h, w, 1, work);
    const int64_t n = (int64_t)h * w;
    int64_t pk = 0;
    float best = -1.0f;
    for (int64_t i = 0; i < n; ++i) {
        const float v = fabsf(C[i].re);
        if (v > best) { best = v; pk = i; }
    }
    const int py = (int)(pk / w), px = (int)(pk % w);
    double energy = 0.0;
    if (variant == 0) {
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                if (y > py - 5 && y < py + 5 && x > px - 5 && x < px + 5) continue;
                const double v = C[(size_t)y * w + x].re;
                energy += v * v;
            }
        energy /= (double)(n - 121);
        const float peak = C[n - 1].re;
        return (double)(peak * peak) / energy;
    }
    int64_t excluded = 0;
    for (int y = 0; y < h; ++y) {
        int dy = abs(y - py);
        if (h - dy < dy) dy = h - dy;
        for (int x = 0; x < w; ++x) {
            int dx = abs(x - px);
            if (w - dx < dx) dx = w - dx;
            if (dy <= 5 && dx <= 5) { ++excluded; continue; }
            const double v = C[(size_t)y * w + x].re;
            energy += v * v;
        }
    }
    energy /= (double)(n - excluded);
    const double cp = C[pk].re;
    return (cp < 0 ? -1.0 : 1.0) * cp * cp / energy;
}

/* PCE of `n_pairs` (i, j) pairs of spectra (count x h x (w/2+1) complex),
 * one pair per thread. */
int port_pce_pairs(const float *spectra, int h, int w, const int32_t *pairs, int64_t n_pairs, int variant,
                   double *out, int threads) {
    const size_t hw = (size_t)h * (w / 2 + 1) * 2;
    int err = 0;
#ifdef _OPENMP
    if (threads < 1) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads) reduction(| : err)
#endif
    {
        fft_plan ph, pw;
        cpx *C = (cpx *)malloc(sizeof(cpx) * (size_t)h * w);
        cpx *work = (cpx *)malloc(sizeof(cpx) * 2 * (size_t)(h > w ? h : w));
        const int bad = fft_plan_init(&ph, h) | fft_plan_init(&pw, w);
        if (bad || !C || !work) {
            err = 1;
        } else {
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
            for (int64_t q = 0; q < n_pairs; ++q)
                out[q] = pce_one(spectra + pairs[2 * q] * hw, spectra + pairs[2 * q + 1] * hw, h, w, variant, &ph,
                                 &pw, C, work);
            fft_plan_free(&ph);
            fft_plan_free(&pw);
        }
        free(C);
        free(work);
    }
    return err ? -1 : 0;
}
