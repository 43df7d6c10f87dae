/*
 * FP64 CPU restatement of ISP_MW's BM3D filter (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker and the timed CPU baseline.  The
 * product library never links or calls this code.
 *
 * The reference itself cannot be compiled here: it needs FFTW3 (absent) and
 * MSVC-only headers and constructs.  This file restates its algorithm in C, in
 * the reference's loop order and in double precision (FLType = double,
 * include/Type.h:26-30).  The FFTW plans of source/BM3D.cpp:13-24 are replaced
 * by the transforms FFTW documents for REDFT10 / REDFT01, computed here with a
 * radix-2 fast DCT for power-of-two lengths and dense cosine sums otherwise.
 *
 * Deviations from the reference, each forced or documented:
 *  - BM3D::process_Plane_FL passes a 0x0 intermediate plane to basic.Kernel
 *    (source/BM3D.cpp:466-468), which the reference then overruns; here the
 *    intermediate is sized like src (the evident intent).
 *  - When more than GroupSize candidates pass the threshold the reference calls
 *    std::partial_sort (include/Block.h:690-695), whose order among equal keys
 *    is unspecified; here equal keys keep scan order (the stable outcome).
 *  - A plane that the 3-plane Kernel does not filter (sigma <= 0) is left
 *    unwritten by the reference (source/BM3D.cpp:200-217); here it is copied
 *    from src, which is what the two-step Frame path then sees.
 */
#include "bm3d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXN 64

static int g_threads = 1;

int orc_set_threads(int n)
{
    int old = g_threads;
    g_threads = n < 1 ? 1 : n;
    return old;
}

/* ------------------------------------------------------------------------ */
/* 1D transforms.  M_n[k][j] = cos(pi (2j+1) k / (2n)).
 * dct2_sum: y = M x            (REDFT10 = 2 * dct2_sum)
 * dct3_sum: x = M^T y          (REDFT01(X) = dct3_sum(diag(1,2,2,..) X))      */

static double g_inv2cos[ORC_MAXN + 1][ORC_MAXN / 2]; /* 1 / (2 cos(pi(2j+1)/(2n))) */
static double g_dense[ORC_MAXN + 1][ORC_MAXN * ORC_MAXN];
static int g_tables_ready = 0;

static void init_tables(void)
{
    if (g_tables_ready) return;
    for (int n = 1; n <= ORC_MAXN; ++n) {
        for (int j = 0; j < n / 2; ++j)
            g_inv2cos[n][j] = (double)(1.0L / (2.0L * cosl(3.14159265358979323846264338327950288L * (2 * j + 1) / (2.0L * n))));
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                g_dense[n][k * n + j] = (double)cosl(3.14159265358979323846264338327950288L * (long double)((2 * j + 1) * k) / (2.0L * n));
    }
    g_tables_ready = 1;
}

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* Lee's radix-2 factorisation of the DCT-II sum, applied along the row index
 * of an n x w array (every one of the w columns transformed at once). */
static void lee_dct2(double *x, int n, int w, double *tmp)
{
    if (n == 1) return;
    int h = n / 2;
    double *a = tmp, *b = tmp + (size_t)h * w;
    for (int j = 0; j < h; ++j) {
        const double *p = x + (size_t)j * w, *q = x + (size_t)(n - 1 - j) * w;
        double *pa = a + (size_t)j * w, *pb = b + (size_t)j * w;
        const double s = g_inv2cos[n][j];
        for (int c = 0; c < w; ++c) {
            pa[c] = p[c] + q[c];
            pb[c] = (p[c] - q[c]) * s;
        }
    }
    lee_dct2(a, h, w, tmp + (size_t)n * w);
    lee_dct2(b, h, w, tmp + (size_t)n * w);
    for (int k = 0; k < h; ++k) memcpy(x + (size_t)(2 * k) * w, a + (size_t)k * w, sizeof(double) * (size_t)w);
    for (int k = 0; k < h - 1; ++k) {
        double *o = x + (size_t)(2 * k + 1) * w;
        const double *p = b + (size_t)k * w, *q = b + (size_t)(k + 1) * w;
        for (int c = 0; c < w; ++c) o[c] = p[c] + q[c];
    }
    memcpy(x + (size_t)(n - 1) * w, b + (size_t)(h - 1) * w, sizeof(double) * (size_t)w);
}

/* Transpose of lee_dct2. */
static void lee_dct3(double *y, int n, int w, double *tmp)
{
    if (n == 1) return;
    int h = n / 2;
    double *v = tmp, *u = tmp + (size_t)h * w;
    for (int k = 0; k < h; ++k) {
        memcpy(v + (size_t)k * w, y + (size_t)(2 * k) * w, sizeof(double) * (size_t)w);
        double *pu = u + (size_t)k * w;
        const double *p = y + (size_t)(2 * k + 1) * w;
        if (k > 0) {
            const double *q = y + (size_t)(2 * k - 1) * w;
            for (int c = 0; c < w; ++c) pu[c] = p[c] + q[c];
        } else {
            for (int c = 0; c < w; ++c) pu[c] = p[c] + 0.0;
        }
    }
    lee_dct3(v, h, w, tmp + (size_t)n * w);
    lee_dct3(u, h, w, tmp + (size_t)n * w);
    for (int j = 0; j < h; ++j) {
        double *o1 = y + (size_t)j * w, *o2 = y + (size_t)(n - 1 - j) * w;
        const double *pv = v + (size_t)j * w, *pu = u + (size_t)j * w;
        const double s = g_inv2cos[n][j];
        for (int c = 0; c < w; ++c) {
            double t = pu[c] * s;
            o1[c] = pv[c] + t;
            o2[c] = pv[c] - t;
        }
    }
}

/* REDFT10 (inverse = 0) or REDFT01 (inverse = 1) along the row index of an
 * n x w array.  tmp holds at least 2*n*w doubles. */
static void redft_rows(double *x, int n, int w, int inverse, double *tmp)
{
    if (inverse)
        for (size_t k = (size_t)w; k < (size_t)n * w; ++k) x[k] *= 2.0;
    if (is_pow2(n)) {
        if (inverse) lee_dct3(x, n, w, tmp);
        else lee_dct2(x, n, w, tmp);
    } else {
        const double *m = g_dense[n];
        double *out = tmp;
        for (int a = 0; a < n; ++a) {
            double *o = out + (size_t)a * w;
            for (int c = 0; c < w; ++c) o[c] = 0.0;
            for (int k = 0; k < n; ++k) {
                const double coef = inverse ? m[k * n + a] : m[a * n + k];
                const double *p = x + (size_t)k * w;
                for (int c = 0; c < w; ++c) o[c] += coef * p[c];
            }
        }
        memcpy(x, out, sizeof(double) * (size_t)n * w);
    }
    if (!inverse)
        for (size_t k = 0; k < (size_t)n * w; ++k) x[k] *= 2.0;
}

static void transpose(double *dst, const double *src, int rows, int cols)
{
    for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) dst[(size_t)c * rows + r] = src[(size_t)r * cols + c];
}

void orc_dct3d(double *data, int n0, int n1, int n2, int inverse)
{
    init_tables();
    enum { CAP = ORC_MAXN * 64 };
    double stackbuf[2 * CAP + CAP];
    double *tmp = stackbuf, *tr = stackbuf + 2 * CAP;
    double *heap = NULL;
    size_t slice = (size_t)n1 * n2, total = (size_t)n0 * slice;
    if (total > CAP || n0 > ORC_MAXN || n1 > ORC_MAXN || n2 > ORC_MAXN) {
        heap = (double *)malloc(sizeof(double) * 3 * (total > slice ? total : slice) * 2);
        tmp = heap;
        tr = heap + 4 * total;
    }
    /* axis 0: the n0 rows of width n1*n2 */
    redft_rows(data, n0, (int)slice, inverse, tmp);
    for (int a = 0; a < n0; ++a) {
        double *s = data + (size_t)a * slice;
        /* axis 1: n1 rows of width n2 */
        redft_rows(s, n1, n2, inverse, tmp);
        /* axis 2: via the transposed slice */
        transpose(tr, s, n1, n2);
        redft_rows(tr, n2, n1, inverse, tmp);
        transpose(s, tr, n2, n1);
    }
    free(heap);
}

/* ------------------------------------------------------------------------ */
/* Parameters (include/BM3D.h:237-259; source/BM3D.cpp:9-73).              */

static void opp_norms(double *normY, double *normU, double *normV)
{
    /* ColorMatrix_RGB2YUV_Parameter, OPP branch (include/Specification.h:391-407) */
    double Yr = (double)(1.0L / 3.0L), Yg = (double)(1.0L / 3.0L), Yb = (double)(1.0L / 3.0L);
    double Ur = (double)(1.0L / 2.0L), Ug = 0.0, Ub = (double)(-1.0L / 2.0L);
    double Vr = (double)(1.0L / 4.0L), Vg = (double)(-1.0L / 2.0L), Vb = (double)(1.0L / 4.0L);
    *normY = sqrt(Yr * Yr + Yg * Yg + Yb * Yb);
    *normU = sqrt(Ur * Ur + Ug * Ug + Ub * Ub);
    *normV = sqrt(Vr * Vr + Vg * Vg + Vb * Vb);
}

typedef struct {
    orc_step_params p;   /* with thMSE scaled by normY */
    double sigma[3];     /* sigma[c]/255*norm_c */
    int wiener;
    double finalAMP[ORC_MAXN];
    double wienerSigmaSqr[3][ORC_MAXN];
    double *thr[3][ORC_MAXN]; /* thrTable[c][n-1], n*B*B entries */
} filter_data;

static void filter_data_init(filter_data *f, const orc_step_params *p, int wiener)
{
    double norm[3];
    opp_norms(&norm[0], &norm[1], &norm[2]);
    memset(f, 0, sizeof(*f));
    f->p = *p;
    f->p.thMSE = p->thMSE * norm[0];
    f->wiener = wiener;
    int G = p->group_size, B = p->block_size;
    for (int c = 0; c < 3; ++c) f->sigma[c] = p->sigma[c] / (double)255 * norm[c];
    for (int i = 1; i <= G && i <= ORC_MAXN; ++i) {
        f->finalAMP[i - 1] = 2 * i * 2 * B * 2 * B;
        double forwardAMP = sqrt(f->finalAMP[i - 1]);
        for (int c = 0; c < 3; ++c) {
            if (!(p->sigma[c] > 0)) continue;
            double sigma = f->sigma[c];
            if (wiener) {
                f->wienerSigmaSqr[c][i - 1] = sigma * forwardAMP * sigma * forwardAMP;
            } else {
                double thrBase = sigma * p->lambda * forwardAMP;
                double thr[4] = {thrBase, thrBase * sqrt(2.0), thrBase * 2.0, thrBase * sqrt(8.0)};
                double *t = (double *)malloc(sizeof(double) * (size_t)i * B * B);
                double *d = t;
                for (int z = 0; z < i; ++z)
                    for (int y = 0; y < B; ++y)
                        for (int x = 0; x < B; ++x, ++d) *d = thr[(x == 0) + (y == 0) + (z == 0)];
                f->thr[c][i - 1] = t;
            }
        }
    }
}

static void filter_data_free(filter_data *f)
{
    for (int c = 0; c < 3; ++c)
        for (int i = 0; i < ORC_MAXN; ++i) free(f->thr[c][i]);
}

/* ------------------------------------------------------------------------ */
/* Grid and search window.                                                  */

int orc_grid(int n, int block_size, int block_step, int *out, int cap)
{
    if (block_step <= 0) return -1;
    int last = n - block_size, cnt = 0;
    for (int j = 0;; j += block_step) {
        if (j >= last + block_step) break;
        else if (j > last) j = last;
        if (cnt < cap && out) out[cnt] = j;
        ++cnt;
    }
    return cnt;
}

int orc_search_boundary(int pos, int plane_boundary, int search_range, int search_step)
{
    int sb;
    search_range = search_range / search_step * search_step;
    if (pos == plane_boundary) {
        sb = plane_boundary;
    } else if (pos > plane_boundary) {
        sb = pos - search_range;
        while (sb < plane_boundary) sb += search_step;
    } else {
        sb = pos + search_range;
        while (sb > plane_boundary) sb -= search_step;
    }
    return sb;
}

/* ------------------------------------------------------------------------ */
/* Block matching (include/Block.h:578-711).                                */

typedef struct {
    double key;
    int seq;
    int y, x;
    int z;   /* frame of a multi-frame candidate (0 otherwise) */
} cand;

static int cand_cmp(const void *a, const void *b)
{
    const cand *p = (const cand *)a, *q = (const cand *)b;
    if (p->key < q->key) return -1;
    if (p->key > q->key) return 1;
    return (p->seq > q->seq) - (p->seq < q->seq);
}

/* Returns N; writes N positions and keys.  buf holds at least window size + 1. */
static int match_one(const filter_data *f, const double *ref, int w, int h, int stride,
                     int j, int i, cand *buf, int32_t *pos_out, double *key_out)
{
    const orc_step_params *p = &f->p;
    const int B = p->block_size;
    int n = 0;
    if (p->group_size == 1 || p->thMSE <= 0) {
        pos_out[0] = (j << 16) | i;
        if (key_out) key_out[0] = 0.0;
        return 1;
    }
    double MSE2SSE = (double)(B * B) * 1.0 * 1.0 / (double)(255 * 255);
    double distMul = 1.0 / MSE2SSE;
    double thSSE = p->thMSE * MSE2SSE;

    int step = p->bm_step;
    int range = p->bm_range / step * step;
    int l = orc_search_boundary(i, 0, range, step);
    int r = orc_search_boundary(i, w - B, range, step);
    int t = orc_search_boundary(j, 0, range, step);
    int b = orc_search_boundary(j, h - B, range, step);

    const double *rb = ref + (size_t)j * stride + i;
    buf[n].key = 0.0; buf[n].seq = -1; buf[n].y = j; buf[n].x = i; ++n;
    int seq = 0;
    for (int y = t; y <= b; y += step) {
        /* Four horizontally adjacent candidates at a time: each keeps its own
         * accumulator, summed in the reference's row-major order, so every
         * distance is bit-identical to the one-at-a-time loop. */
        for (int x0 = l; x0 <= r; x0 += 4 * step) {
            int nc = (r - x0) / step + 1;
            if (nc > 4) nc = 4;
            double dist[4] = {0.0, 0.0, 0.0, 0.0};
            const double *cb = ref + (size_t)y * stride + x0;
            if (nc == 4) {
                for (int u = 0; u < B; ++u) {
                    const double *rr = rb + (size_t)u * stride, *cc = cb + (size_t)u * stride;
                    for (int v = 0; v < B; ++v) {
                        double t0 = rr[v] - cc[v], t1 = rr[v] - cc[v + step];
                        double t2 = rr[v] - cc[v + 2 * step], t3 = rr[v] - cc[v + 3 * step];
                        dist[0] += t0 * t0;
                        dist[1] += t1 * t1;
                        dist[2] += t2 * t2;
                        dist[3] += t3 * t3;
                    }
                }
            } else {
                for (int c = 0; c < nc; ++c)
                    for (int u = 0; u < B; ++u) {
                        const double *rr = rb + (size_t)u * stride, *cc = cb + (size_t)u * stride + c * step;
                        for (int v = 0; v < B; ++v) {
                            double temp = rr[v] - cc[v];
                            dist[c] += temp * temp;
                        }
                    }
            }
            for (int c = 0; c < nc; ++c) {
                int x = x0 + c * step;
                if (y == j && x == i) continue;
                if (dist[c] <= thSSE && dist[c] != 0) {
                    buf[n].key = dist[c] * distMul; buf[n].seq = seq; buf[n].y = y; buf[n].x = x; ++n;
                }
                ++seq;
            }
        }
    }
    qsort(buf, (size_t)n, sizeof(cand), cand_cmp);
    int G = p->group_size;
    if (G > 0 && n > G) n = G;
    for (int z = 0; z < n; ++z) {
        pos_out[z] = (buf[z].y << 16) | buf[z].x;
        if (key_out) key_out[z] = buf[z].key;
    }
    return n;
}

static int window_cap(const orc_step_params *p)
{
    int step = p->bm_step > 0 ? p->bm_step : 1;
    int side = 2 * (p->bm_range / step) + 1;
    return side * side + 1;
}

int orc_match_all(const orc_step_params *p, const double *ref, int w, int h, int stride,
                  int *counts, int32_t *pos, double *keys)
{
    if (p->block_size <= 0 || w < p->block_size || h < p->block_size || p->bm_step <= 0 ||
        p->group_size <= 0 || p->group_size > ORC_MAXN)
        return -1;
    filter_data f;
    filter_data_init(&f, p, 0);
    int ngx = orc_grid(w, p->block_size, p->block_step, NULL, 0);
    int ngy = orc_grid(h, p->block_size, p->block_step, NULL, 0);
    if (ngx < 0 || ngy < 0) { filter_data_free(&f); return -1; }
    int *gx = (int *)malloc(sizeof(int) * (size_t)ngx), *gy = (int *)malloc(sizeof(int) * (size_t)ngy);
    orc_grid(w, p->block_size, p->block_step, gx, ngx);
    orc_grid(h, p->block_size, p->block_step, gy, ngy);
    int G = p->group_size;
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads)
    for (int a = 0; a < ngy; ++a) {
        cand *buf = (cand *)malloc(sizeof(cand) * (size_t)window_cap(p));
        for (int c = 0; c < ngx; ++c) {
            size_t g = (size_t)a * ngx + c;
            counts[g] = match_one(&f, ref, w, h, stride, gy[a], gx[c], buf, pos + g * G,
                                  keys ? keys + g * G : NULL);
        }
        free(buf);
    }
    free(gx); free(gy);
    filter_data_free(&f);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Collaborative filtering (source/BM3D.cpp:339-385 and 411-457).           */

static const int g_zero_frames[ORC_MAXN] = {0};

/* Member z lies in frame fz[z] of the planes srcs[] (Pos3 gather, BlockGroup::From with a
 * plane vector, include/Block.h:1141-1158). */
static void gather(double *grp, const double *const *srcs, const int *fz, int stride, const int32_t *pos, int N,
                   int B)
{
    double *d = grp;
    for (int z = 0; z < N; ++z) {
        int y = pos[z] >> 16, x = pos[z] & 0xFFFF;
        const double *s = srcs[fz[z]] + (size_t)y * stride + x;
        for (int u = 0; u < B; ++u)
            for (int v = 0; v < B; ++v) *d++ = s[(size_t)u * stride + v];
    }
}

/* ------------------------------------------------------------------------ */
/* Opt-in variant modes of the filter, which the reference does not          */
/* have (it uses a DCT along the group and a uniform weight per block):       */
/*   Haar: the group axis is transformed by the orthonormal Haar matrix, and  */
/*         a group is truncated to the largest power of two of its members;  */
/*   Kaiser: each filtered block is aggregated with a Kaiser window of shape  */
/*         beta (w * k[u] k[v] in num, w * k[u] k[v] in den).                 */
/* Both run in the orthonormal formulation (a uniform lambda*sigma threshold; */
/* Wiener noise sigma^2 / 2^zeta, zeta counting the zero indices, the DC row  */
/* of the Haar matrix being index 0), which the reference's unnormalised      */
/* tables reduce to in the DCT mode.                                          */

static int g_haar = 0;
static double g_kaiser_beta = 0.0;

int orc_set_variant(int haar, double kaiser_beta)
{
    g_haar = haar != 0;
    g_kaiser_beta = kaiser_beta;
    return 0;
}

/* Orthonormal Haar matrix of a power-of-two n, row-major:
 * H_1 = [1], H_2m = [H_m (x) [1 1]; I_m (x) [1 -1]] / sqrt(2). */
static void haar_matrix(double *h, int n)
{
    h[0] = 1.0;
    for (int m = 1; m < n; m *= 2) {
        double t[ORC_MAXN * ORC_MAXN];
        const double r = (double)(1.0L / sqrtl(2.0L));
        for (int i = 0; i < 2 * m * 2 * m; ++i) t[i] = 0.0;
        for (int r0 = 0; r0 < m; ++r0)
            for (int c = 0; c < m; ++c) {
                t[r0 * 2 * m + 2 * c] = h[r0 * m + c] * r;
                t[r0 * 2 * m + 2 * c + 1] = h[r0 * m + c] * r;
            }
        for (int r0 = 0; r0 < m; ++r0) {
            t[(m + r0) * 2 * m + 2 * r0] = r;
            t[(m + r0) * 2 * m + 2 * r0 + 1] = -r;
        }
        for (int i = 0; i < 2 * m * 2 * m; ++i) h[i] = t[i];
    }
}

/* Orthonormal DCT-II matrix of length n, row-major */
static void dct_matrix(double *c, int n)
{
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            c[k * n + j] = (double)((k == 0 ? sqrtl(1.0L / n) : sqrtl(2.0L / n)) *
                                    cosl(3.14159265358979323846264338327950288L * (2 * j + 1) * k / (2.0L * n)));
}

/* Modified Bessel function of the first kind, order 0 (its power series) */
static double bessel_i0(double x)
{
    double sum = 1.0, term = 1.0;
    for (int m = 1; m < 200; ++m) {
        term *= (x / (2.0 * m)) * (x / (2.0 * m));
        sum += term;
        if (term < 1e-18 * sum) break;
    }
    return sum;
}

/* Kaiser window of length n and shape beta: I0(beta sqrt(1 - t^2)) / I0(beta),
 * t = (2i - (n - 1)) / (n - 1); 1 for n = 1 */
static void kaiser_window(double *k, int n, double beta)
{
    for (int i = 0; i < n; ++i) {
        if (n == 1) {
            k[i] = 1.0;
            continue;
        }
        const double t = (2.0 * i - (n - 1)) / (double)(n - 1);
        k[i] = bessel_i0(beta * sqrt(fmax(0.0, 1.0 - t * t))) / bessel_i0(beta);
    }
}

/* y = M x along one axis of g[n0][n1][n2] (axis 0, 1 or 2), M n x n row-major;
 * transposed (x = M^T y) when inverse. */
static void axis_product(double *g, int n0, int n1, int n2, int axis, const double *M, int inverse)
{
    const int n = axis == 0 ? n0 : (axis == 1 ? n1 : n2);
    const size_t stride = axis == 0 ? (size_t)n1 * n2 : (axis == 1 ? (size_t)n2 : 1);
    const int lines0 = axis == 0 ? 1 : n0, lines1 = axis == 2 ? n1 : (axis == 1 ? n2 : n1 * n2);
    double in[ORC_MAXN];
    for (int a = 0; a < lines0; ++a)
        for (int b = 0; b < lines1; ++b) {
            double *p;
            if (axis == 0) p = g + b;
            else if (axis == 1) p = g + (size_t)a * n1 * n2 + b;
            else p = g + (size_t)a * n1 * n2 + (size_t)b * n2;
            for (int i = 0; i < n; ++i) in[i] = p[(size_t)i * stride];
            for (int k = 0; k < n; ++k) {
                double s = 0.0;
                for (int i = 0; i < n; ++i) s += (inverse ? M[i * n + k] : M[k * n + i]) * in[i];
                p[(size_t)k * stride] = s;
            }
        }
}

static void variant_transform(double *g, int N, int B, int inverse)
{
    double cz[ORC_MAXN * ORC_MAXN], cb[ORC_MAXN * ORC_MAXN];
    if (g_haar) haar_matrix(cz, N);
    else dct_matrix(cz, N);
    dct_matrix(cb, B);
    if (!inverse) {
        axis_product(g, N, B, B, 2, cb, 0);
        axis_product(g, N, B, B, 1, cb, 0);
        axis_product(g, N, B, B, 0, cz, 0);
    } else {
        axis_product(g, N, B, B, 0, cz, 1);
        axis_product(g, N, B, B, 1, cb, 1);
        axis_product(g, N, B, B, 2, cb, 1);
    }
}

static void collaborative_filter_variant(const filter_data *f, int c, double *const *num, double *const *den,
                                         int row0, int stride, const double *const *src, const double *const *ref,
                                         const int32_t *pos, const int *fz, int N, double *grp, double *gref)
{
    const int B = f->p.block_size;
    if (g_haar)
        while (N & (N - 1)) N &= N - 1;  /* the largest power of two of the members */
    const int cnt = N * B * B;
    const double sigma = f->sigma[c];
    gather(grp, src, fz, stride, pos, N, B);
    variant_transform(grp, N, B, 0);
    double w;
    if (!f->wiener) {
        const double thr = sigma * f->p.lambda;
        int retained = 0;
        for (int k = 0; k < cnt; ++k) {
            if (fabs(grp[k]) <= thr) grp[k] = 0;
            else ++retained;
        }
        w = retained < 1 ? 1.0 : 1.0 / (double)retained;
    } else {
        gather(gref, ref, fz, stride, pos, N, B);
        variant_transform(gref, N, B, 0);
        double L2 = 0;
        for (int k = 0; k < cnt; ++k) {
            const int z = k / (B * B), u = (k / B) % B, v = k % B;
            const int zeta = (z == 0) + (u == 0) + (v == 0);
            const double yy = gref[k] * gref[k];
            const double wc = yy / (yy + sigma * sigma / (double)(1 << zeta));
            grp[k] *= wc;
            L2 += wc * wc;
        }
        w = L2 <= 0 ? 1.0 : 1.0 / L2;
    }
    variant_transform(grp, N, B, 1);
    double kw[ORC_MAXN];
    kaiser_window(kw, B, g_kaiser_beta);
    const double *s = grp;
    for (int z = 0; z < N; ++z) {
        int y = (pos[z] >> 16) - row0, x = pos[z] & 0xFFFF;
        double *dn = num[fz[z]] + (size_t)y * stride + x;
        double *dd = den[fz[z]] + (size_t)y * stride + x;
        for (int u = 0; u < B; ++u)
            for (int v = 0; v < B; ++v) {
                const double ww = w * kw[u] * kw[v];
                dn[(size_t)u * stride + v] += *s++ * ww;
                dd[(size_t)u * stride + v] += ww;
            }
    }
}

/* Accumulates into num/den of each member's frame, whose row 0 is image row
 * `row0`.  A single-frame group has every fz[z] = 0. */
static void collaborative_filter_pos3(const filter_data *f, int c, double *const *num, double *const *den, int row0,
                                      int stride, const double *const *src, const double *const *ref,
                                      const int32_t *pos, const int *fz, int N, double *grp, double *gref)
{
    if (g_haar || g_kaiser_beta != 0.0) {
        collaborative_filter_variant(f, c, num, den, row0, stride, src, ref, pos, fz, N, grp, gref);
        return;
    }
    const int B = f->p.block_size;
    const int cnt = N * B * B;
    gather(grp, src, fz, stride, pos, N, B);
    orc_dct3d(grp, N, B, B, 0);
    double denWeight;
    if (!f->wiener) {
        const double *thr = f->thr[c][N - 1];
        int retained = 0;
        for (int k = 0; k < cnt; ++k) {
            if (fabs(grp[k]) <= thr[k]) grp[k] = 0;
            else ++retained;
        }
        denWeight = retained < 1 ? 1.0 : 1.0 / (double)retained;
    } else {
        gather(gref, ref, fz, stride, pos, N, B);
        orc_dct3d(gref, N, B, B, 0);
        const double sigmaSquare = f->wienerSigmaSqr[c][N - 1];
        double L2 = 0;
        for (int k = 0; k < cnt; ++k) {
            double ySquare = gref[k] * gref[k];
            double wc = ySquare / (ySquare + sigmaSquare);
            grp[k] *= wc;
            L2 += wc * wc;
        }
        denWeight = L2 <= 0 ? 1.0 : 1.0 / L2;
    }
    orc_dct3d(grp, N, B, B, 1);
    double numWeight = denWeight / f->finalAMP[N - 1];
    const double *s = grp;
    for (int z = 0; z < N; ++z) {
        int y = (pos[z] >> 16) - row0, x = pos[z] & 0xFFFF;
        double *dn = num[fz[z]] + (size_t)y * stride + x;
        for (int u = 0; u < B; ++u)
            for (int v = 0; v < B; ++v) dn[(size_t)u * stride + v] += *s++ * numWeight;
    }
    for (int z = 0; z < N; ++z) {
        int y = (pos[z] >> 16) - row0, x = pos[z] & 0xFFFF;
        double *dd = den[fz[z]] + (size_t)y * stride + x;
        for (int u = 0; u < B; ++u)
            for (int v = 0; v < B; ++v) dd[(size_t)u * stride + v] += denWeight;
    }
}

/* Accumulates into num/den, whose row 0 is image row `row0`. */
static void collaborative_filter(const filter_data *f, int c, double *num, double *den, int row0,
                                 int stride, const double *src, const double *ref,
                                 const int32_t *pos, int N, double *grp, double *gref)
{
    double *nm[1] = {num}, *dn[1] = {den};
    const double *sr[1] = {src}, *rf[1] = {ref};
    collaborative_filter_pos3(f, c, nm, dn, row0, stride, sr, rf, pos, g_zero_frames, N, grp, gref);
}

static int degenerate(const filter_data *f, int nplanes)
{
    const orc_step_params *p = &f->p;
    int nosigma = nplanes == 1 ? !(p->sigma[0] > 0)
                               : (!(p->sigma[0] > 0) && !(p->sigma[1] > 0) && !(p->sigma[2] > 0));
    return nosigma || p->group_size == 0 || p->block_size <= 0 || p->bm_range <= 0 ||
           p->bm_range < p->bm_step || p->thMSE <= 0;
}

int orc_kernel(const orc_step_params *p, int wiener, int nplanes,
               double *const *dst, const double *const *src, const double *const *ref,
               int w, int h)
{
    if (nplanes != 1 && nplanes != 3) return -1;
    init_tables();
    filter_data f;
    filter_data_init(&f, p, wiener);
    size_t npx = (size_t)w * h;
    if (degenerate(&f, nplanes)) {
        for (int c = 0; c < nplanes; ++c)
            if (dst[c] != src[c]) memcpy(dst[c], src[c], sizeof(double) * npx);
        filter_data_free(&f);
        return 1;
    }
    const int B = p->block_size, G = p->group_size;
    if (w < B || h < B || p->block_step <= 0 || p->bm_step <= 0 || G < 0 || G > ORC_MAXN) {
        filter_data_free(&f);
        return -2;
    }
    int active[3] = {0, 0, 0};
    for (int c = 0; c < nplanes; ++c) active[c] = p->sigma[c] > 0;

    int ngx = orc_grid(w, B, p->block_step, NULL, 0);
    int ngy = orc_grid(h, B, p->block_step, NULL, 0);
    int *gx = (int *)malloc(sizeof(int) * (size_t)ngx), *gy = (int *)malloc(sizeof(int) * (size_t)ngy);
    orc_grid(w, B, p->block_step, gx, ngx);
    orc_grid(h, B, p->block_step, gy, ngy);

    double *num[3] = {0, 0, 0}, *den[3] = {0, 0, 0};
    for (int c = 0; c < nplanes; ++c)
        if (active[c]) {
            num[c] = (double *)calloc(npx, sizeof(double));
            den[c] = (double *)calloc(npx, sizeof(double));
        }

    int range = p->bm_range / p->bm_step * p->bm_step;
    int nthreads = g_threads;
    if (nthreads > ngy) nthreads = ngy;
    int Gcap = G > 0 ? G : ORC_MAXN;

    if (nthreads <= 1) {
        /* The reference's single loop: accumulate straight into num/den. */
        cand *buf = (cand *)malloc(sizeof(cand) * (size_t)window_cap(p));
        int32_t pos[ORC_MAXN];
        double *grp = (double *)malloc(sizeof(double) * (size_t)Gcap * B * B);
        double *gref = (double *)malloc(sizeof(double) * (size_t)Gcap * B * B);
        for (int a = 0; a < ngy; ++a)
            for (int b = 0; b < ngx; ++b) {
                int N = match_one(&f, ref[0], w, h, w, gy[a], gx[b], buf, pos, NULL);
                if (G > 0 && N > G) N = G;
                for (int c = 0; c < nplanes; ++c)
                    if (active[c])
                        collaborative_filter(&f, c, num[c], den[c], 0, w, src[c], ref[c], pos, N, grp, gref);
            }
        free(buf); free(grp); free(gref);
    } else {
        /* Row bands of the grid, each with private num/den rows covering the
         * bands' reach; summed afterwards in band order. */
        int per = (ngy + nthreads - 1) / nthreads;
#pragma omp parallel for schedule(static, 1) num_threads(nthreads)
        for (int t = 0; t < nthreads; ++t) {
            int a0 = t * per, a1 = a0 + per < ngy ? a0 + per : ngy;
            if (a0 >= a1) continue;
            int r0 = gy[a0] - range; if (r0 < 0) r0 = 0;
            int r1 = gy[a1 - 1] + range + B; if (r1 > h) r1 = h;
            size_t bpx = (size_t)(r1 - r0) * w;
            double *bn[3] = {0, 0, 0}, *bd[3] = {0, 0, 0};
            for (int c = 0; c < nplanes; ++c)
                if (active[c]) { bn[c] = (double *)calloc(bpx, sizeof(double)); bd[c] = (double *)calloc(bpx, sizeof(double)); }
            cand *buf = (cand *)malloc(sizeof(cand) * (size_t)window_cap(p));
            int32_t pos[ORC_MAXN];
            double *grp = (double *)malloc(sizeof(double) * (size_t)Gcap * B * B);
            double *gref = (double *)malloc(sizeof(double) * (size_t)Gcap * B * B);
            for (int a = a0; a < a1; ++a)
                for (int b = 0; b < ngx; ++b) {
                    int N = match_one(&f, ref[0], w, h, w, gy[a], gx[b], buf, pos, NULL);
                    if (G > 0 && N > G) N = G;
                    for (int c = 0; c < nplanes; ++c)
                        if (active[c])
                            collaborative_filter(&f, c, bn[c], bd[c], r0, w, src[c], ref[c], pos, N, grp, gref);
                }
            free(buf); free(grp); free(gref);
#pragma omp critical
            {
                for (int c = 0; c < nplanes; ++c)
                    if (active[c]) {
                        double *dn = num[c] + (size_t)r0 * w, *dd = den[c] + (size_t)r0 * w;
                        for (size_t k = 0; k < bpx; ++k) { dn[k] += bn[c][k]; dd[k] += bd[c][k]; }
                    }
            }
            for (int c = 0; c < nplanes; ++c) { free(bn[c]); free(bd[c]); }
        }
    }

    for (int c = 0; c < nplanes; ++c) {
        if (active[c]) {
            for (size_t k = 0; k < npx; ++k) dst[c][k] = num[c][k] / den[c][k];
            free(num[c]); free(den[c]);
        } else if (dst[c] != src[c]) {
            memcpy(dst[c], src[c], sizeof(double) * npx);
        }
    }
    free(gx); free(gy);
    filter_data_free(&f);
    return 0;
}

int orc_bm3d_plane(const orc_step_params *basic, const orc_step_params *final_,
                   double *dst, const double *src, const double *ref, int w, int h)
{
    size_t npx = (size_t)w * h;
    double *tmp = (double *)malloc(sizeof(double) * npx);
    const double *s1[1] = {src}, *r1[1] = {ref ? ref : src};
    double *d1[1] = {tmp};
    int rc = orc_kernel(basic, 0, 1, d1, s1, r1, w, h);
    if (rc >= 0) {
        const double *r2[1] = {tmp};
        double *d2[1] = {dst};
        rc = orc_kernel(final_, 1, 1, d2, s1, r2, w, h);
    }
    free(tmp);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* OPP colour (include/Conversion.hpp:547-578, 418-432, 739-762) for float  */
/* planes with Floor 0 / Ceil 1 (chroma Floor -0.5 / Neutral 0 / Ceil 0.5). */

void orc_rgb2opp(double *y, double *u, double *v, const double *r, const double *g,
                 const double *b, long n)
{
    double gainY = 1.0 / (1.0 * 3.0), offsetY = -0.0 * 3.0 * gainY + 0.0;
    double gainU = 1.0 / (1.0 * 2.0), gainV = 1.0 / (1.0 * 4.0), offsetC = 0.0;
    for (long k = 0; k < n; ++k) {
        if (y) y[k] = (r[k] + g[k] + b[k]) * gainY + offsetY;
        if (u) u[k] = (r[k] - b[k]) * gainU + offsetC;
        if (v) v[k] = (r[k] - g[k] * 2.0 + b[k]) * gainV + offsetC;
    }
}

static double clip01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }

void orc_opp2rgb(double *r, double *g, double *b, const double *y, const double *u,
                 const double *v, long n, int clip)
{
    double Ry = 1.0, Ru = 1.0, Rv = (double)(2.0L / 3.0L);
    double Gy = 1.0, Gv = (double)(-4.0L / 3.0L);
    double By = 1.0, Bu = -1.0, Bv = (double)(2.0L / 3.0L);
    double offsetR = -0.0 * Ry - 0.0 * (Ru + Rv) + 0.0;
    double offsetG = -0.0 * Gy - 0.0 * (0.0 + Gv) + 0.0;
    double offsetB = -0.0 * By - 0.0 * (Bu + Bv) + 0.0;
    for (long k = 0; k < n; ++k) {
        double tr = Ry * y[k] + Ru * u[k] + Rv * v[k] + offsetR;
        double tg = Gy * y[k] + Gv * v[k] + offsetG;
        double tb = By * y[k] + Bu * u[k] + Bv * v[k] + offsetB;
        r[k] = clip ? clip01(tr) : tr;
        g[k] = clip ? clip01(tg) : tg;
        b[k] = clip ? clip01(tb) : tb;
    }
}

static int wrapper_degenerate(const orc_step_params *p)
{
    /* BM3D_Base::process_Frame (source/BM3D.cpp:284-290) */
    return (!(p->sigma[0] > 0) && !(p->sigma[1] > 0) && !(p->sigma[2] > 0)) || p->group_size == 1 ||
           p->block_size <= 0 || p->bm_range <= 0 || p->bm_range < p->bm_step || p->thMSE <= 0;
}

int orc_bm3d_rgb(const orc_step_params *basic, const orc_step_params *final_, int mode,
                 double *const *dst, const double *const *src, const double *const *ref,
                 int w, int h)
{
    long npx = (long)w * h;
    if (mode != 1 && mode != 2 && mode != 3) return -1;
    const orc_step_params *single = mode == 1 ? basic : final_;
    if (mode != 3 && wrapper_degenerate(single)) {
        for (int c = 0; c < 3; ++c) memcpy(dst[c], src[c], sizeof(double) * (size_t)npx);
        return 1;
    }
    double *buf = (double *)malloc(sizeof(double) * (size_t)npx * 9);
    double *s[3] = {buf, buf + npx, buf + 2 * npx};
    double *t[3] = {buf + 3 * npx, buf + 4 * npx, buf + 5 * npx};
    double *r[3] = {buf + 6 * npx, buf + 7 * npx, buf + 8 * npx};
    orc_rgb2opp(s[0], s[1], s[2], src[0], src[1], src[2], npx);
    int have_ref = ref && ref[0];
    int rc;
    const double *sc[3] = {s[0], s[1], s[2]};
    if (mode == 3) {
        const double *rb[3] = {s[0], s[1], s[2]};
        if (have_ref) {
            orc_rgb2opp(r[0], NULL, NULL, ref[0], ref[1], ref[2], npx);
            rb[0] = r[0];
        }
        rc = orc_kernel(basic, 0, 3, t, sc, rb, w, h);
        if (rc >= 0) {
            const double *tc[3] = {t[0], t[1], t[2]};
            rc = orc_kernel(final_, 1, 3, s, sc, tc, w, h);
        }
    } else {
        const double *rb[3] = {s[0], s[1], s[2]};
        if (have_ref) {
            if (mode == 1) orc_rgb2opp(r[0], NULL, NULL, ref[0], ref[1], ref[2], npx);
            else orc_rgb2opp(r[0], r[1], r[2], ref[0], ref[1], ref[2], npx);
            rb[0] = r[0]; rb[1] = r[1]; rb[2] = r[2];
        }
        rc = orc_kernel(single, mode == 2, 3, s, sc, rb, w, h);
    }
    if (rc >= 0) orc_opp2rgb(dst[0], dst[1], dst[2], s[0], s[1], s[2], npx, 1);
    free(buf);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Integer planes and frames (Plane / Frame of include/Image_Type.h:51-403). */

/* isChroma / isPCChroma on integer quantisation (include/Image_Type.hpp:144-167). */
static int q_chroma(int floor_, int neutral) { return floor_ < neutral; }
static int q_pc_chroma(int floor_, int neutral, int ceil_) { return floor_ < neutral && (floor_ + ceil_) % 2 == 1; }

/* RangeConvert int -> float plane (include/Conversion.hpp:134-256, general
 * branch): luma to [0,1] (dst Floor 0, Neutral 0, Ceil 1), chroma to
 * [-0.5,0.5] (dst Neutral 0).  No clip, except that a PC-range chroma source
 * is always clipped (:241-242). */
void orc_range_to_float(double *dst, const int32_t *src, long n, int floor_, int neutral, int ceil_)
{
    int chroma = q_chroma(floor_, neutral);
    double sRange = (double)(ceil_ - floor_);
    double gain = 1.0 / sRange;
    double offset = -(double)neutral * gain + 0.0;
    int clip = q_pc_chroma(floor_, neutral, ceil_);
    double lo = chroma ? -0.5 : 0.0, hi = chroma ? 0.5 : 1.0;
    for (long k = 0; k < n; ++k) {
        double v = (double)src[k] * gain + offset;
        dst[k] = clip ? (v < lo ? lo : (v > hi ? hi : v)) : v;
    }
}

/* RangeConvert float -> int plane with clip (include/Conversion.hpp:241-256):
 * v * range + Neutral + 0.5 (0.499999 for PC-range chroma), clipped to
 * [Floor, Ceil], truncated. */
void orc_range_to_int(int32_t *dst, const double *src, long n, int floor_, int neutral, int ceil_)
{
    double dRange = (double)(ceil_ - floor_);
    double gain = dRange / 1.0;
    double offset = -0.0 * gain + (double)neutral + (q_pc_chroma(floor_, neutral, ceil_) ? 0.499999 : 0.5);
    for (long k = 0; k < n; ++k) {
        double v = src[k] * gain + offset;
        v = v < floor_ ? floor_ : (v > ceil_ ? ceil_ : v);
        dst[k] = (int32_t)v;
    }
}

static int plane_wrapper_degenerate(const orc_step_params *p)
{
    /* BM3D_Base::process_Plane (source/BM3D.cpp:250-255) */
    double normY, normU, normV;
    opp_norms(&normY, &normU, &normV);
    return !(p->sigma[0] > 0) || p->group_size == 1 || p->block_size <= 0 || p->bm_range <= 0 ||
           p->bm_range < p->bm_step || p->thMSE * normY <= 0;
}

int orc_bm3d_plane_int(const orc_step_params *basic, const orc_step_params *final_, int mode,
                       int32_t *dst, const int32_t *src, const int32_t *ref, int w, int h,
                       int floor_, int neutral, int ceil_)
{
    long npx = (long)w * h;
    const orc_step_params *single = mode == 1 ? basic : final_;
    if (mode != 3 && plane_wrapper_degenerate(single)) {
        memcpy(dst, src, sizeof(int32_t) * (size_t)npx);
        return 1;
    }
    double *s = (double *)malloc(sizeof(double) * (size_t)npx * 3);
    double *r = s + npx, *d = s + 2 * npx;
    orc_range_to_float(s, src, npx, floor_, neutral, ceil_);
    const double *rr = s;
    if (ref && ref != src) {
        orc_range_to_float(r, ref, npx, floor_, neutral, ceil_);
        rr = r;
    }
    int rc;
    if (mode == 3) {
        rc = orc_bm3d_plane(basic, final_, d, s, rr, w, h);
    } else {
        const double *s1[1] = {s}, *r1[1] = {rr};
        double *d1[1] = {d};
        rc = orc_kernel(single, mode == 2, 1, d1, s1, r1, w, h);
    }
    if (rc >= 0) orc_range_to_int(dst, d, npx, floor_, neutral, ceil_);
    free(s);
    return rc;
}

/* Frame (int RGB, Floor/Ceil) two-step or single-step BM3D:
 * MatrixConvert_RGB2YUV OPP from int (include/Conversion.hpp:547-578),
 * filtering, MatrixConvert_YUV2RGB OPP to int with +0.5 and clip
 * (include/Conversion.hpp:739-762). */
int orc_bm3d_rgb_int(const orc_step_params *basic, const orc_step_params *final_, int mode,
                     int32_t *const *dst, const int32_t *const *src, const int32_t *const *ref,
                     int w, int h, int floor_, int ceil_)
{
    long npx = (long)w * h;
    if (mode != 1 && mode != 2 && mode != 3) return -1;
    const orc_step_params *single = mode == 1 ? basic : final_;
    if (mode != 3 && wrapper_degenerate(single)) {
        for (int c = 0; c < 3; ++c) memcpy(dst[c], src[c], sizeof(int32_t) * (size_t)npx);
        return 1;
    }
    double sRange = (double)(ceil_ - floor_);
    double gainY = 1.0 / (sRange * 3.0), offsetY = -(double)floor_ * 3.0 * gainY + 0.0;
    double gainU = 1.0 / (sRange * 2.0), gainV = 1.0 / (sRange * 4.0), offsetC = 0.0;
    double *buf = (double *)malloc(sizeof(double) * (size_t)npx * 9);
    double *s[3] = {buf, buf + npx, buf + 2 * npx};
    double *t[3] = {buf + 3 * npx, buf + 4 * npx, buf + 5 * npx};
    double *r[3] = {buf + 6 * npx, buf + 7 * npx, buf + 8 * npx};
    for (long k = 0; k < npx; ++k) {
        double R = src[0][k], G = src[1][k], B = src[2][k];
        s[0][k] = (R + G + B) * gainY + offsetY;
        s[1][k] = (R - B) * gainU + offsetC;
        s[2][k] = (R - G * 2.0 + B) * gainV + offsetC;
    }
    int have_ref = ref && ref[0] && (ref[0] != src[0] || ref[1] != src[1] || ref[2] != src[2]);
    if (have_ref)
        for (long k = 0; k < npx; ++k) {
            double R = ref[0][k], G = ref[1][k], B = ref[2][k];
            r[0][k] = (R + G + B) * gainY + offsetY;
            r[1][k] = (R - B) * gainU + offsetC;
            r[2][k] = (R - G * 2.0 + B) * gainV + offsetC;
        }
    const double *sc[3] = {s[0], s[1], s[2]};
    int rc;
    if (mode == 3) {
        const double *rb[3] = {have_ref ? r[0] : s[0], s[1], s[2]};
        rc = orc_kernel(basic, 0, 3, t, sc, rb, w, h);
        if (rc >= 0) {
            const double *tc[3] = {t[0], t[1], t[2]};
            rc = orc_kernel(final_, 1, 3, s, sc, tc, w, h);
        }
    } else {
        const double *rb[3] = {have_ref ? r[0] : s[0], have_ref ? r[1] : s[1], have_ref ? r[2] : s[2]};
        rc = orc_kernel(single, mode == 2, 3, s, sc, rb, w, h);
    }
    if (rc >= 0) {
        double dRange = sRange, gY = dRange / 1.0, gC = dRange / 1.0;
        double Ry = 1.0 * gY, Ru = 1.0 * gC, Rv = (double)(2.0L / 3.0L) * gC;
        double Gy = 1.0 * gY, Gv = (double)(-4.0L / 3.0L) * gC;
        double By = 1.0 * gY, Bu = -1.0 * gC, Bv = (double)(2.0L / 3.0L) * gC;
        double offR = -0.0 * Ry - 0.0 * (Ru + Rv) + (double)floor_ + 0.5;
        double offG = -0.0 * Gy - 0.0 * (0.0 + Gv) + (double)floor_ + 0.5;
        double offB = -0.0 * By - 0.0 * (Bu + Bv) + (double)floor_ + 0.5;
        for (long k = 0; k < npx; ++k) {
            double Y = s[0][k], U = s[1][k], V = s[2][k];
            double v[3] = {Ry * Y + Ru * U + Rv * V + offR, Gy * Y + Gv * V + offG, By * Y + Bu * U + Bv * V + offB};
            for (int c = 0; c < 3; ++c) {
                double x = v[c] < floor_ ? floor_ : (v[c] > ceil_ ? ceil_ : v[c]);
                dst[c][k] = (int32_t)x;
            }
        }
    }
    free(buf);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* NLMeans (include/NLMeans.h:11-64, source/NLMeans.cpp:6-375): the same     */
/* block matching (thMSE not scaled by normY), a weighted average of each    */
/* group with weights exp(-key / strength^2), optionally corrected by a      */
/* stationarity test, and uniform aggregation.                               */

static int nlm_degenerate(const orc_nlm_params *p)
{
    /* NLMeans::process_Plane_FL / process_Frame (source/NLMeans.cpp:152-158, 281-287) */
    return p->strength <= 0 || p->group_size == 1 || p->block_size <= 0 || p->bm_range <= 0 ||
           p->bm_range < p->bm_step || p->thMSE <= 0;
}

/* WeightedAverage / WeightedAverage_Correction (source/NLMeans.cpp:6-147) of
 * one group on one plane (sample range `range`), into blk[B*B]. */
static void nlm_average(const orc_nlm_params *p, const double *src, int stride, double range, int j, int i,
                        const int32_t *pos, const double *keys, int N, double *blk)
{
    const int B = p->block_size;
    int G = N;
    if (p->group_size > 0 && G > p->group_size) G = p->group_size;
    if (G < 2) {
        for (int y = 0; y < B; ++y)
            for (int x = 0; x < B; ++x) blk[y * B + x] = src[(size_t)(j + y) * stride + i + x];
        return;
    }
    double *s1 = (double *)calloc((size_t)B * B * 2, sizeof(double)), *s2 = s1 + B * B;
    const double exponentMul = -1 / (p->strength * p->strength);
    double wsum = 0;
    for (int k = 0; k < G; ++k) {
        const double w = exp(keys[k] * exponentMul);
        wsum += w;
        const int py = pos[k] >> 16, px = pos[k] & 0xFFFF;
        for (int y = 0; y < B; ++y)
            for (int x = 0; x < B; ++x) {
                const double v = src[(size_t)(py + y) * stride + px + x];
                if (p->correction) {
                    const double t = v * w;
                    s1[y * B + x] += t;
                    s2[y * B + x] += v * t;
                } else {
                    s1[y * B + x] += v * w;
                }
            }
    }
    const double rec = 1.0 / wsum;
    if (p->correction) {
        const double sigma = p->sigma * range / 255, VarN = sigma * sigma;
        for (int y = 0; y < B; ++y)
            for (int x = 0; x < B; ++x) {
                const double X = src[(size_t)(j + y) * stride + i + x];
                const double EX = s1[y * B + x] * rec, VarX = s2[y * B + x] * rec - EX * EX;
                blk[y * B + x] = VarX > VarN ? X - (VarN / VarX) * (X - EX) : EX;
            }
    } else {
        for (int q = 0; q < B * B; ++q) blk[q] = s1[q] * rec;
    }
    free(s1);
}

/* The grid loop shared by the plane and frame forms: matching on ref (range
 * 1), nplanes source planes of sample range `range`, num / den sums.  Groups
 * and their block estimates are formed in parallel over grid rows; the sums
 * are then taken in the reference's raster order, so the result does not
 * depend on the thread count. */
static int nlm_sums(const orc_nlm_params *p, int nplanes, const double *const *src, const double *ref, double range,
                    int w, int h, double *const *num, double *den)
{
    const int B = p->block_size;
    if (w < B || h < B || p->bm_step <= 0 || p->block_step <= 0 || p->group_size < 0 || p->group_size > 4096) return -1;
    filter_data f;
    memset(&f, 0, sizeof(f));
    f.p.block_size = B;
    f.p.block_step = p->block_step;
    f.p.group_size = p->group_size;
    f.p.bm_range = p->bm_range;
    f.p.bm_step = p->bm_step;
    f.p.thMSE = p->thMSE;
    const int cap = window_cap(&f.p);
    /* reference grid (source/NLMeans.cpp:179-201) */
    const int right = w - B, bottom = h - B;
    int *gy = (int *)malloc(sizeof(int) * (size_t)(h + 2)), *gx = (int *)malloc(sizeof(int) * (size_t)(w + 2));
    int ngy = 0, ngx = 0;
    for (int j = 0;; j += p->block_step) {
        if (j >= bottom + p->block_step) break;
        if (j > bottom) j = bottom;
        gy[ngy++] = j;
    }
    for (int i = 0;; i += p->block_step) {
        if (i >= right + p->block_step) break;
        if (i > right) i = right;
        gx[ngx++] = i;
    }
    const size_t bsz = (size_t)B * B;
    double *blocks = (double *)malloc(sizeof(double) * bsz * (size_t)nplanes * (size_t)ngx * (size_t)ngy);
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads)
    for (int a = 0; a < ngy; ++a) {
        cand *buf = (cand *)malloc(sizeof(cand) * (size_t)cap);
        int32_t *pos = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
        double *keys = (double *)malloc(sizeof(double) * (size_t)cap);
        for (int b = 0; b < ngx; ++b) {
            const int n = match_one(&f, ref, w, h, w, gy[a], gx[b], buf, pos, keys);
            for (int c = 0; c < nplanes; ++c)
                nlm_average(p, src[c], w, range, gy[a], gx[b], pos, keys, n,
                            blocks + (((size_t)a * ngx + b) * nplanes + c) * bsz);
        }
        free(buf);
        free(pos);
        free(keys);
    }
    for (int a = 0; a < ngy; ++a)
        for (int b = 0; b < ngx; ++b) {
            const int j = gy[a], i = gx[b];
            for (int c = 0; c < nplanes; ++c) {
                const double *blk = blocks + (((size_t)a * ngx + b) * nplanes + c) * bsz;
                for (int y = 0; y < B; ++y)
                    for (int x = 0; x < B; ++x) num[c][(size_t)(j + y) * w + i + x] += blk[y * B + x];
            }
            for (int y = 0; y < B; ++y)
                for (int x = 0; x < B; ++x) den[(size_t)(j + y) * w + i + x] += 1.0;
        }
    free(blocks);
    free(gy);
    free(gx);
    return 0;
}

int orc_nlm_plane(const orc_nlm_params *p, double *dst, const double *src, const double *ref, int w, int h)
{
    const long npx = (long)w * h;
    if (nlm_degenerate(p)) {
        memcpy(dst, src, sizeof(double) * (size_t)npx);
        return 1;
    }
    double *num = (double *)calloc((size_t)npx * 2, sizeof(double)), *den = num + npx;
    double *nums[1] = {num};
    const double *srcs[1] = {src};
    const int rc = nlm_sums(p, 1, srcs, ref ? ref : src, 1.0, w, h, nums, den);
    if (rc == 0)
        for (long k = 0; k < npx; ++k) dst[k] = num[k] / den[k];
    free(num);
    return rc;
}

int orc_nlm_rgb_int(const orc_nlm_params *p, int32_t *const *dst, const int32_t *const *src,
                    const int32_t *const *ref, int w, int h, int floor_, int ceil_)
{
    const long npx = (long)w * h;
    if (nlm_degenerate(p)) {
        for (int c = 0; c < 3; ++c) memcpy(dst[c], src[c], sizeof(int32_t) * (size_t)npx);
        return 1;
    }
    /* ConvertToY of the guide, OPP (include/Conversion.hpp:418-432) */
    const double sRange = (double)(ceil_ - floor_), gainY = 1.0 / (sRange * 3.0);
    const double offsetY = -(double)floor_ * 3.0 * gainY + 0.0;
    const int32_t *const *g = (ref && ref[0]) ? ref : src;
    double *buf = (double *)calloc((size_t)npx * 8, sizeof(double));
    double *y = buf, *s[3] = {buf + npx, buf + 2 * npx, buf + 3 * npx};
    double *num[3] = {buf + 4 * npx, buf + 5 * npx, buf + 6 * npx}, *den = buf + 7 * npx;
    for (long k = 0; k < npx; ++k) {
        const double t = ((double)g[0][k] + (double)g[1][k] + (double)g[2][k]) * gainY + offsetY;
        y[k] = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);  /* clip = true (:487, :502) */
        for (int c = 0; c < 3; ++c) s[c][k] = (double)src[c][k];
    }
    const double *sc[3] = {s[0], s[1], s[2]};
    const int rc = nlm_sums(p, 3, sc, y, sRange, w, h, num, den);
    if (rc == 0)
        for (int c = 0; c < 3; ++c)
            for (long k = 0; k < npx; ++k) dst[c][k] = (int32_t)(num[c][k] / den[k] + 0.5);
    free(buf);
    return rc;
}

/* NLMeans::process_Plane (source/NLMeans.cpp:232-257): Plane_FL src_data(src)
 * converts with clip (include/Image_Type.cpp Plane_FL(const Plane&) ->
 * RangeConvert(dst, src, clip = true), include/Conversion.hpp:277), then the
 * float filter and RangeConvert back with +0.5 rounding.  The output clip of
 * orc_range_to_int changes nothing here: outputs are convex combinations of
 * clipped samples. */
int orc_nlm_plane_int(const orc_nlm_params *p, int32_t *dst, const int32_t *src, const int32_t *ref, int w, int h,
                      int floor_, int neutral, int ceil_)
{
    const long npx = (long)w * h;
    if (nlm_degenerate(p)) {
        memcpy(dst, src, sizeof(int32_t) * (size_t)npx);
        return 1;
    }
    double *buf = (double *)malloc(sizeof(double) * (size_t)npx * 3), *s = buf, *r = buf + npx, *d = buf + 2 * npx;
    const int chroma = q_chroma(floor_, neutral);
    const double lo = chroma ? -0.5 : 0.0, hi = chroma ? 0.5 : 1.0;
    const int32_t *in[2] = {src, ref && ref != src ? ref : NULL};
    double *out[2] = {s, r};
    for (int j = 0; j < 2; ++j) {
        if (!in[j]) continue;
        orc_range_to_float(out[j], in[j], npx, floor_, neutral, ceil_);
        for (long k = 0; k < npx; ++k) out[j][k] = out[j][k] < lo ? lo : (out[j][k] > hi ? hi : out[j][k]);
    }
    const int rc = orc_nlm_plane(p, d, s, in[1] ? r : s, w, h);
    if (rc == 0) orc_range_to_int(dst, d, npx, floor_, neutral, ceil_);
    free(buf);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Multi-frame grouping.  The reference carries Pos3 group codes (frame, y, x)
 * and gathers / aggregates over a vector of planes (BlockGroup,
 * include/Block.h:843-846, 927-936, 1141-1208) but never forms such groups.
 * This restatement defines them as the direct extension of
 * BlockMatchingMulti (include/Block.h:578-711) to a stack of frames:
 *   - the reference block is frame t's; candidates are every position of the
 *     clipped search window in each frame z with |z - t| <= radius, visited
 *     frame by frame in the order t, t-1, t+1, t-2, t+2, ... and row-major
 *     within a frame (the scan order that breaks ties);
 *   - the acceptance test is the reference's (SSE <= thSSE and SSE != 0),
 *     the position (t, j, i) itself is the group's first member;
 *   - members are filtered as one group by the step's collaborative filter
 *     and each is aggregated into its own frame.
 * With one frame or radius 0 every group and every output equals the
 * single-frame filter's. */

static int stack_frame_order(int t, int frames, int radius, int *order)
{
    int n = 0;
    order[n++] = t;
    for (int d = 1; d <= radius; ++d) {
        if (t - d >= 0) order[n++] = t - d;
        if (t + d < frames) order[n++] = t + d;
    }
    return n;
}

/* Group of reference block (t, j, i) over the guide planes guide[frame]. */
static int match_stack(const filter_data *f, const double *const *guide, int frames, int radius, int t, int w, int h,
                       int j, int i, cand *buf, int32_t *pos_out, int *fz_out)
{
    const orc_step_params *p = &f->p;
    const int B = p->block_size;
    pos_out[0] = (j << 16) | i;
    fz_out[0] = t;
    if (p->group_size == 1 || p->thMSE <= 0) return 1;
    const double MSE2SSE = (double)(B * B) * 1.0 * 1.0 / (double)(255 * 255);
    const double distMul = 1.0 / MSE2SSE;
    const double thSSE = p->thMSE * MSE2SSE;
    const int step = p->bm_step;
    const int range = p->bm_range / step * step;
    const int l = orc_search_boundary(i, 0, range, step), r = orc_search_boundary(i, w - B, range, step);
    const int tp = orc_search_boundary(j, 0, range, step), b = orc_search_boundary(j, h - B, range, step);
    const double *rb = guide[t] + (size_t)j * w + i;
    int order[2 * 64 + 1];
    const int nf = stack_frame_order(t, frames, radius, order);
    int n = 0, seq = 0;
    buf[n].key = 0.0; buf[n].seq = -1; buf[n].y = j; buf[n].x = i; buf[n].z = t; ++n;
    for (int q = 0; q < nf; ++q) {
        const int z = order[q];
        const double *g = guide[z];
        for (int y = tp; y <= b; y += step)
            for (int x = l; x <= r; x += step) {
                if (z == t && y == j && x == i) continue;
                const double *cb = g + (size_t)y * w + x;
                double dist = 0.0;
                for (int u = 0; u < B; ++u)
                    for (int v = 0; v < B; ++v) {
                        const double d = rb[(size_t)u * w + v] - cb[(size_t)u * w + v];
                        dist += d * d;
                    }
                if (dist <= thSSE && dist != 0) {
                    buf[n].key = dist * distMul; buf[n].seq = seq; buf[n].y = y; buf[n].x = x; buf[n].z = z; ++n;
                }
                ++seq;
            }
    }
    qsort(buf, (size_t)n, sizeof(cand), cand_cmp);
    const int G = p->group_size;
    if (G > 0 && n > G) n = G;
    for (int k = 0; k < n; ++k) {
        pos_out[k] = (buf[k].y << 16) | buf[k].x;
        fz_out[k] = buf[k].z;
    }
    return n;
}

int orc_match_stack(const orc_step_params *p, const double *const *guide, int frames, int radius, int t, int w,
                    int h, int *counts, int32_t *pos, int *fz)
{
    if (p->block_size <= 0 || w < p->block_size || h < p->block_size || p->bm_step <= 0 || p->group_size <= 0 ||
        p->group_size > ORC_MAXN || frames < 1 || radius < 0 || radius > 64 || t < 0 || t >= frames)
        return -1;
    filter_data f;
    filter_data_init(&f, p, 0);
    const int B = p->block_size, G = p->group_size;
    const int ngx = orc_grid(w, B, p->block_step, NULL, 0), ngy = orc_grid(h, B, p->block_step, NULL, 0);
    int *gx = (int *)malloc(sizeof(int) * (size_t)ngx), *gy = (int *)malloc(sizeof(int) * (size_t)ngy);
    orc_grid(w, B, p->block_step, gx, ngx);
    orc_grid(h, B, p->block_step, gy, ngy);
    cand *buf = (cand *)malloc(sizeof(cand) * (size_t)window_cap(p) * (size_t)(2 * radius + 1));
    for (int a = 0; a < ngy; ++a)
        for (int c = 0; c < ngx; ++c) {
            const size_t g = (size_t)a * ngx + c;
            counts[g] = match_stack(&f, guide, frames, radius, t, w, h, gy[a], gx[c], buf, pos + g * G, fz + g * G);
        }
    free(buf);
    free(gx);
    free(gy);
    filter_data_free(&f);
    return 0;
}

/* One BM3D step over a stack: dst/src/ref are [frame * nplanes + plane];
 * matching on ref[frame * nplanes] (the luma guide), Wiener guides ref[...]. */
int orc_kernel_stack(const orc_step_params *p, int wiener, int nplanes, int frames, int radius,
                     double *const *dst, const double *const *src, const double *const *ref, int w, int h)
{
    if ((nplanes != 1 && nplanes != 3) || frames < 1 || radius < 0 || radius > 64) return -1;
    init_tables();
    filter_data f;
    filter_data_init(&f, p, wiener);
    const size_t npx = (size_t)w * h;
    const int nall = frames * nplanes;
    if (degenerate(&f, nplanes)) {
        for (int k = 0; k < nall; ++k)
            if (dst[k] != src[k]) memcpy(dst[k], src[k], sizeof(double) * npx);
        filter_data_free(&f);
        return 1;
    }
    const int B = p->block_size, G = p->group_size;
    if (w < B || h < B || p->block_step <= 0 || p->bm_step <= 0 || G < 0 || G > ORC_MAXN) {
        filter_data_free(&f);
        return -2;
    }
    int active[3] = {0, 0, 0};
    for (int c = 0; c < nplanes; ++c) active[c] = p->sigma[c] > 0;
    const int ngx = orc_grid(w, B, p->block_step, NULL, 0), ngy = orc_grid(h, B, p->block_step, NULL, 0);
    int *gx = (int *)malloc(sizeof(int) * (size_t)ngx), *gy = (int *)malloc(sizeof(int) * (size_t)ngy);
    orc_grid(w, B, p->block_step, gx, ngx);
    orc_grid(h, B, p->block_step, gy, ngy);
    /* planes by frame: num/den/src/ref of plane c */
    double **num = (double **)calloc((size_t)nall * 2, sizeof(double *)), **den = num + nall;
    const double **sp = (const double **)malloc(sizeof(double *) * (size_t)frames * 3);
    const double **rp = sp + frames, **gp = sp + 2 * frames;
    for (int k = 0; k < nall; ++k)
        if (active[k % nplanes]) {
            num[k] = (double *)calloc(npx, sizeof(double));
            den[k] = (double *)calloc(npx, sizeof(double));
        }
    for (int z = 0; z < frames; ++z) gp[z] = ref[z * nplanes];
    const int Gcap = G > 0 ? G : ORC_MAXN;
    cand *buf = (cand *)malloc(sizeof(cand) * (size_t)window_cap(p) * (size_t)(2 * radius + 1));
    int32_t pos[ORC_MAXN];
    int fz[ORC_MAXN];
    double *grp = (double *)malloc(sizeof(double) * (size_t)Gcap * B * B);
    double *gref = (double *)malloc(sizeof(double) * (size_t)Gcap * B * B);
    double **nm = (double **)malloc(sizeof(double *) * (size_t)frames * 2), **dn = nm + frames;
    /* frames in order, each frame's grid in the reference's raster order */
    for (int t = 0; t < frames; ++t)
        for (int a = 0; a < ngy; ++a)
            for (int b = 0; b < ngx; ++b) {
                int N = match_stack(&f, gp, frames, radius, t, w, h, gy[a], gx[b], buf, pos, fz);
                if (G > 0 && N > G) N = G;
                for (int c = 0; c < nplanes; ++c) {
                    if (!active[c]) continue;
                    for (int z = 0; z < frames; ++z) {
                        sp[z] = src[z * nplanes + c];
                        rp[z] = ref[z * nplanes + c];
                        nm[z] = num[z * nplanes + c];
                        dn[z] = den[z * nplanes + c];
                    }
                    collaborative_filter_pos3(&f, c, nm, dn, 0, w, sp, rp, pos, fz, N, grp, gref);
                }
            }
    for (int k = 0; k < nall; ++k) {
        if (active[k % nplanes]) {
            for (size_t q = 0; q < npx; ++q) dst[k][q] = num[k][q] / den[k][q];
            free(num[k]);
            free(den[k]);
        } else if (dst[k] != src[k]) {
            memcpy(dst[k], src[k], sizeof(double) * npx);
        }
    }
    free(num); free(sp); free(buf); free(grp); free(gref); free(nm); free(gx); free(gy);
    filter_data_free(&f);
    return 0;
}

/* Two-step (mode 3) or single-step (1 basic, 2 final) BM3D of a stack of grey
 * planes (nplanes 1) or float RGB frames (nplanes 3, the Frame semantics:
 * OPP without clip, matching on Y, OPP back with clip).  ref (or NULL): as
 * for the single-frame forms, the step-1 matching guide (modes 1 and 3) or
 * the Wiener guide (mode 2). */
int orc_bm3d_stack(const orc_step_params *basic, const orc_step_params *final_, int mode, int nplanes, int frames,
                   int radius, double *const *dst, const double *const *src, const double *const *ref, int w, int h)
{
    if ((mode != 1 && mode != 2 && mode != 3) || (nplanes != 1 && nplanes != 3) || frames < 1) return -1;
    const size_t npx = (size_t)w * h;
    const int nall = frames * nplanes;
    double *buf = (double *)malloc(sizeof(double) * npx * (size_t)nall * 3);
    double **pl = (double **)malloc(sizeof(double *) * (size_t)nall * 3);
    double **s = pl, **t = pl + nall, **r = pl + 2 * nall;
    for (int k = 0; k < nall; ++k) {
        s[k] = buf + (size_t)k * npx;
        t[k] = buf + (size_t)(nall + k) * npx;
        r[k] = buf + (size_t)(2 * nall + k) * npx;
    }
    const int have_ref = ref && ref[0];
    for (int z = 0; z < frames; ++z) {
        if (nplanes == 3) {
            const double *const *sz = src + z * 3;
            orc_rgb2opp(s[z * 3], s[z * 3 + 1], s[z * 3 + 2], sz[0], sz[1], sz[2], (long)npx);
            if (have_ref) {
                const double *const *rz = ref + z * 3;
                orc_rgb2opp(r[z * 3], r[z * 3 + 1], r[z * 3 + 2], rz[0], rz[1], rz[2], (long)npx);
            }
        } else {
            memcpy(s[z], src[z], sizeof(double) * npx);
            if (have_ref) memcpy(r[z], ref[z], sizeof(double) * npx);
        }
    }
    const double *const *sc = (const double *const *)s;
    int rc;
    if (mode == 3) {
        /* step 1 matches on the ref's luma (or src's); its planes filter src */
        const double **rb = (const double **)malloc(sizeof(double *) * (size_t)nall);
        for (int k = 0; k < nall; ++k) rb[k] = (have_ref && k % nplanes == 0) ? r[k] : s[k];
        rc = orc_kernel_stack(basic, 0, nplanes, frames, radius, t, sc, rb, w, h);
        free(rb);
        if (rc >= 0) rc = orc_kernel_stack(final_, 1, nplanes, frames, radius, s, sc, (const double *const *)t, w, h);
    } else {
        const double **rb = (const double **)malloc(sizeof(double *) * (size_t)nall);
        for (int k = 0; k < nall; ++k) rb[k] = have_ref && (mode == 2 || k % nplanes == 0) ? r[k] : s[k];
        rc = orc_kernel_stack(mode == 1 ? basic : final_, mode == 2, nplanes, frames, radius, s, sc, rb, w, h);
        free(rb);
    }
    for (int z = 0; z < frames && rc >= 0; ++z) {
        if (nplanes == 3) orc_opp2rgb(dst[z * 3], dst[z * 3 + 1], dst[z * 3 + 2], s[z * 3], s[z * 3 + 1],
                                      s[z * 3 + 2], (long)npx, 1);
        else memcpy(dst[z], s[z], sizeof(double) * npx);
    }
    free(buf);
    free(pl);
    return rc;
}
