/*
 * eplace_oracle.c -- FP64 CPU restatement of the ePlace hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product.  Built with -ffp-contract=off so that the FP32 bin-index routine
 * performs exactly the IEEE operations written here.
 *
 * Reference paths below are relative to the searsm8/AIEplace tree.
 */
#include "eplace_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

void eo_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int eo_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* WA primitives                                                      */
/* ------------------------------------------------------------------ */

/* computeTerm.py:6-13 -- a = 1 + a/2^prec, then prec squarings. */
double eo_fast_exp(double a, int prec)
{
    double r = 1.0 + a / ldexp(1.0, prec);
    for (int i = 0; i < prec; ++i) r *= r;
    return r;
}

static double wa_exp(double a, int use_fast)
{
    return use_fast ? eo_fast_exp(a, 16) : exp(a);
}

/* computeTerm.py:15-43 (a_plus, a_minus, b_term, c_term);
 * AIEplace.cpp:612-647 (compute_a_terms_CPU / compute_bc_terms_CPU). */
void eo_wa_terms(const double *x, int d, double gamma, int use_fast,
                 double *a_plus, double *a_minus, double *bc)
{
    double xmax = x[0], xmin = x[0];
    for (int i = 1; i < d; ++i) {
        if (x[i] > xmax) xmax = x[i];
        if (x[i] < xmin) xmin = x[i];
    }
    double bp = 0, cp = 0, bm = 0, cm = 0;
    for (int i = 0; i < d; ++i) {
        double ap = wa_exp((x[i] - xmax) / gamma, use_fast);
        double am = wa_exp((xmin - x[i]) / gamma, use_fast);
        if (a_plus) a_plus[i] = ap;
        if (a_minus) a_minus[i] = am;
        bp += ap;
        cp += x[i] * ap;
        bm += am;
        cm += x[i] * am;
    }
    bc[0] = bp;
    bc[1] = cp;
    bc[2] = bm;
    bc[3] = cm;
}

/* computeTerm.py:50-56; AIEplace.cpp:660-668:
 *   ((1 + x/g) b+ - c+/g) a+ / b+^2  -  ((1 - x/g) b- + c-/g) a- / b-^2 */
static double wa_partials_ws(const double *x, int d, double gamma, int use_fast,
                             double *partials, double *scratch /* 2d */)
{
    double bc[4];
    double *ap = scratch, *am = scratch + d;
    eo_wa_terms(x, d, gamma, use_fast, ap, am, bc);
    for (int i = 0; i < d; ++i) {
        double t1 = ((1.0 + x[i] / gamma) * bc[0] - bc[1] / gamma) * ap[i] / (bc[0] * bc[0]);
        double t2 = ((1.0 - x[i] / gamma) * bc[2] + bc[3] / gamma) * am[i] / (bc[2] * bc[2]);
        partials[i] = t1 - t2;
    }
    return bc[1] / bc[0] - bc[3] / bc[2];
}

void eo_wa_partials(const double *x, int d, double gamma, int use_fast, double *partials)
{
    double *s = (double *)malloc(sizeof(double) * (size_t)d * 2);
    wa_partials_ws(x, d, gamma, use_fast, partials, s);
    free(s);
}

/* computeTerm.py:45-47 */
double eo_wa_value(const double *x, int d, double gamma, int use_fast)
{
    double bc[4];
    eo_wa_terms(x, d, gamma, use_fast, NULL, NULL, bc);
    return bc[1] / bc[0] - bc[3] / bc[2];
}

/* ------------------------------------------------------------------ */
/* netlist-level helpers                                              */
/* ------------------------------------------------------------------ */

static inline int n_place(const eo_design *d) { return d->n_mov + d->n_fill; }

static inline void cell_pos(const eo_design *d, const double *x, const double *y, int c,
                            double *px, double *py)
{
    int np = n_place(d);
    if (c < np) {
        *px = x[c];
        *py = y[c];
    } else {
        *px = d->fix_x[c - np];
        *py = d->fix_y[c - np];
    }
}

/* AIEplace.cpp:574-673 over every net of degree >= 2, pins with offsets.
 * Per-pin partials are formed net-parallel, then summed per cell in pin order
 * (deterministic for any thread count). */
void eo_wa_grad(const eo_design *d, const double *x, const double *y, double gamma,
                double *gx, double *gy, double *wa)
{
    int np = n_place(d);
    int P = d->net_start[d->n_nets];
    double *pgx = (double *)malloc(sizeof(double) * (size_t)(P > 0 ? P : 1));
    double *pgy = (double *)malloc(sizeof(double) * (size_t)(P > 0 ? P : 1));
    double wa_sum = 0.0;
#pragma omp parallel reduction(+ : wa_sum)
    {
        int cap = 64;
        double *buf = (double *)malloc(sizeof(double) * (size_t)cap * 5);
#pragma omp for schedule(dynamic, 1024)
        for (int n = 0; n < d->n_nets; ++n) {
            int s = d->net_start[n], e = d->net_start[n + 1];
            int deg = e - s;
            if (deg < 2) {
                for (int p = s; p < e; ++p) pgx[p] = pgy[p] = 0.0;
                continue;
            }
            if (deg > cap) {
                cap = deg;
                free(buf);
                buf = (double *)malloc(sizeof(double) * (size_t)cap * 5);
            }
            double *bx = buf, *by = buf + cap, *sc = buf + 2 * cap;
            for (int k = 0; k < deg; ++k) {
                int p = s + k;
                double cx, cy;
                cell_pos(d, x, y, d->pin_cell[p], &cx, &cy);
                bx[k] = cx + d->pin_dx[p];
                by[k] = cy + d->pin_dy[p];
            }
            wa_sum += wa_partials_ws(bx, deg, gamma, 0, pgx + s, sc);
            wa_sum += wa_partials_ws(by, deg, gamma, 0, pgy + s, sc);
        }
        free(buf);
    }
    memset(gx, 0, sizeof(double) * (size_t)np);
    memset(gy, 0, sizeof(double) * (size_t)np);
    for (int p = 0; p < P; ++p) {
        int c = d->pin_cell[p];
        if (c < d->n_mov) {
            gx[c] += pgx[p];
            gy[c] += pgy[p];
        }
    }
    free(pgx);
    free(pgy);
    if (wa) *wa = wa_sum;
}

/* Net.cpp:136-155, DataBase.cpp:143-149: sum over nets of (max-min)x + (max-min)y. */
double eo_hpwl(const eo_design *d, const double *x, const double *y)
{
    double total = 0.0;
    for (int n = 0; n < d->n_nets; ++n) {
        int s = d->net_start[n], e = d->net_start[n + 1];
        if (e - s < 1) continue;
        double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
        for (int p = s; p < e; ++p) {
            double cx, cy;
            cell_pos(d, x, y, d->pin_cell[p], &cx, &cy);
            double qx = cx + d->pin_dx[p], qy = cy + d->pin_dy[p];
            if (qx < xmin) xmin = qx;
            if (qx > xmax) xmax = qx;
            if (qy < ymin) ymin = qy;
            if (qy > ymax) ymax = qy;
        }
        total += (xmax - xmin) + (ymax - ymin);
    }
    return total;
}

/* ------------------------------------------------------------------ */
/* density                                                            */
/* ------------------------------------------------------------------ */

/* Grid.cpp:40-61 index rule (floor of offset / bin size) in IEEE FP32. */
void eo_cell_bins_f32(int n, const float *x, const float *y, const float *w, const float *h,
                      float xl, float yl, float bw, float bh, float sqrt2_bw, float sqrt2_bh,
                      int Mx, int My, int stretch, int32_t *out)
{
    for (int c = 0; c < n; ++c) {
        float x0, x1, y0, y1;
        if (stretch) {
            float cx = x[c] + 0.5f * w[c];
            float cy = y[c] + 0.5f * h[c];
            float sx = fmaxf(w[c], sqrt2_bw);
            float sy = fmaxf(h[c], sqrt2_bh);
            x0 = cx - 0.5f * sx;
            x1 = cx + 0.5f * sx;
            y0 = cy - 0.5f * sy;
            y1 = cy + 0.5f * sy;
        } else {
            x0 = x[c];
            x1 = x[c] + w[c];
            y0 = y[c];
            y1 = y[c] + h[c];
        }
        int i0 = (int)floorf((x0 - xl) / bw);
        int i1 = (int)floorf((x1 - xl) / bw);
        int j0 = (int)floorf((y0 - yl) / bh);
        int j1 = (int)floorf((y1 - yl) / bh);
        if (i0 < 0) i0 = 0;
        if (j0 < 0) j0 = 0;
        if (i1 > Mx - 1) i1 = Mx - 1;
        if (j1 > My - 1) j1 = My - 1;
        out[4 * c + 0] = i0;
        out[4 * c + 1] = i1;
        out[4 * c + 2] = j0;
        out[4 * c + 3] = j1;
    }
}

/* Bin.h:78-121: overlap = (min(right) - max(left)) * (min(top) - max(bottom)),
 * accumulated only where positive; scaled by `scale`. */
static void deposit_rect(const eo_design *d, double x0, double x1, double y0, double y1,
                         double scale, double *map)
{
    int Mx = d->bins_x, My = d->bins_y;
    double bw = (d->xh - d->xl) / Mx, bh = (d->yh - d->yl) / My;
    int i0 = (int)floor((x0 - d->xl) / bw), i1 = (int)floor((x1 - d->xl) / bw);
    int j0 = (int)floor((y0 - d->yl) / bh), j1 = (int)floor((y1 - d->yl) / bh);
    if (i0 < 0) i0 = 0;
    if (j0 < 0) j0 = 0;
    if (i1 > Mx - 1) i1 = Mx - 1;
    if (j1 > My - 1) j1 = My - 1;
    for (int i = i0; i <= i1; ++i) {
        double bx0 = d->xl + i * bw, bx1 = d->xl + (i + 1) * bw;
        double ox = fmin(bx1, x1) - fmax(bx0, x0);
        if (ox <= 0) continue;
        for (int j = j0; j <= j1; ++j) {
            double by0 = d->yl + j * bh, by1 = d->yl + (j + 1) * bh;
            double oy = fmin(by1, y1) - fmax(by0, y0);
            if (oy <= 0) continue;
            map[(size_t)i * My + j] += scale * ox * oy;
        }
    }
}

static void stretched_rect(const eo_design *d, double x, double y, int c, double *x0,
                           double *x1, double *y0, double *y1, double *ratio)
{
    double bw = (d->xh - d->xl) / d->bins_x, bh = (d->yh - d->yl) / d->bins_y;
    double w = d->w[c], h = d->h[c];
    double sx = fmax(w, M_SQRT2 * bw), sy = fmax(h, M_SQRT2 * bh);
    double cx = x + 0.5 * w, cy = y + 0.5 * h;
    *x0 = cx - 0.5 * sx;
    *x1 = cx + 0.5 * sx;
    *y0 = cy - 0.5 * sy;
    *y1 = cy + 0.5 * sy;
    *ratio = (w * h) / (sx * sy);
}

void eo_density_exact(const eo_design *d, const double *x, const double *y, int c0, int c1,
                      double *map)
{
    for (int c = c0; c < c1; ++c)
        deposit_rect(d, x[c], x[c] + d->w[c], y[c], y[c] + d->h[c], 1.0, map);
}

void eo_density_stretched(const eo_design *d, const double *x, const double *y, int c0, int c1,
                          double *map)
{
    for (int c = c0; c < c1; ++c) {
        double x0, x1, y0, y1, r;
        stretched_rect(d, x[c], y[c], c, &x0, &x1, &y0, &y1, &r);
        deposit_rect(d, x0, x1, y0, y1, r, map);
    }
}

void eo_density_fixed(const eo_design *d, double *map)
{
    int np = n_place(d);
    for (int k = 0; k < d->n_fix; ++k) {
        int c = np + k;
        deposit_rect(d, d->fix_x[k], d->fix_x[k] + d->w[c], d->fix_y[k], d->fix_y[k] + d->h[c],
                     1.0, map);
    }
}

/* Bin.h:123-126 + Grid.cpp:144-151, normalised by movable area and with rho_t. */
double eo_overflow(const eo_design *d, const double *map_mov, const double *map_fix,
                   double rho_t)
{
    int Mx = d->bins_x, My = d->bins_y;
    double bw = (d->xh - d->xl) / Mx, bh = (d->yh - d->yl) / My;
    double cap = rho_t * bw * bh;
    double sum = 0.0;
    for (size_t b = 0; b < (size_t)Mx * My; ++b) {
        double v = map_mov[b] + rho_t * map_fix[b] - cap;
        if (v > 0) sum += v;
    }
    double area = 0.0;
    for (int c = 0; c < d->n_mov; ++c) area += d->w[c] * d->h[c];
    return area > 0 ? sum / area : 0.0;
}

/* ------------------------------------------------------------------ */
/* transforms                                                         */
/* ------------------------------------------------------------------ */

/* DCT.cpp:24-40 */
void eo_dct(const double *x, int n, double *y)
{
    for (int k = 0; k < n; ++k) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += x[i] * cos(M_PI / n * (i + 0.5) * k);
        y[k] = s;
    }
}

/* DCT.cpp:45-59 */
void eo_idct(const double *x, int n, double *y)
{
    for (int k = 0; k < n; ++k) {
        double s = 0.0;
        for (int i = 1; i < n; ++i) s += x[i] * cos(M_PI / n * (k + 0.5) * i);
        y[k] = 0.5 * x[0] + s;
    }
}

/* DCT.cpp:64-82: reverse x_1..x_{N-1}, keep x_0, IDCT, negate odd outputs. */
void eo_idxst(const double *x, int n, double *y)
{
    double *t = (double *)malloc(sizeof(double) * (size_t)n);
    t[0] = x[0];
    for (int i = 1; i < n; ++i) t[i] = x[n - i];
    eo_idct(t, n, y);
    for (int k = 1; k < n; k += 2) y[k] = -y[k];
    free(t);
}

/* customDCT.py:44-55 */
void eo_idst(const double *x, int n, double *y)
{
    for (int k = 0; k < n; ++k) {
        double s = 0.0;
        for (int i = 1; i < n; ++i) s += x[i] * sin(M_PI * i * (k + 0.5) / n);
        y[k] = s;
    }
}

typedef void (*tr1d)(const double *, int, double *);
static tr1d kind_fn(int kind)
{
    switch (kind) {
    case 0: return eo_dct;
    case 1: return eo_idct;
    case 2: return eo_idxst;
    default: return eo_idst;
    }
}

/* customDCT.py:79-158: rows (second index) first, then columns. */
void eo_transform2d(const double *in, int nx, int ny, int col_kind, int row_kind, double *out)
{
    tr1d fr = kind_fn(row_kind), fc = kind_fn(col_kind);
    double *tmp = (double *)malloc(sizeof(double) * (size_t)nx * ny);
    for (int i = 0; i < nx; ++i) fr(in + (size_t)i * ny, ny, tmp + (size_t)i * ny);
    double *col = (double *)malloc(sizeof(double) * (size_t)nx * 2);
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) col[i] = tmp[(size_t)i * ny + j];
        fc(col, nx, col + nx);
        for (int i = 0; i < nx; ++i) out[(size_t)i * ny + j] = col[nx + i];
    }
    free(col);
    free(tmp);
}

/* C[m x n] = A[m x k] * B[k x n], row-major, double. */
static void dgemm(int m, int n, int k, const double *A, const double *B, double *C)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < m; ++i) {
        double *c = C + (size_t)i * n;
        for (int j = 0; j < n; ++j) c[j] = 0.0;
        for (int p = 0; p < k; ++p) {
            double a = A[(size_t)i * k + p];
            if (a == 0.0) continue;
            const double *b = B + (size_t)p * n;
            for (int j = 0; j < n; ++j) c[j] += a * b[j];
        }
    }
}

/* cos/sin table T[u][i] = cos|sin(pi u (i+1/2)/M) */
static double *trig_table(int M, int sine, int transpose)
{
    double *t = (double *)malloc(sizeof(double) * (size_t)M * M);
    for (int u = 0; u < M; ++u)
        for (int i = 0; i < M; ++i) {
            double a = M_PI * u * (i + 0.5) / M;
            double v = sine ? sin(a) : cos(a);
            if (transpose)
                t[(size_t)i * M + u] = v;
            else
                t[(size_t)u * M + i] = v;
        }
    return t;
}

/* DCT.cpp:24-82: DCT_naive accumulates in double, IDCT_naive in float, one
 * cos() per term; IDXST_naive reverses x_1..x_{N-1}, applies IDCT_naive and
 * negates the odd outputs. */
void eo_rows_naive(const float *in, int rows, int n, int kind, float *out)
{
#pragma omp parallel for schedule(static)
    for (int r = 0; r < rows; ++r) {
        const float *x = in + (size_t)r * n;
        float *y = out + (size_t)r * n;
        if (kind == 0) {
            for (int k = 0; k < n; ++k) {
                double sum = 0;
                for (int j = 0; j < n; ++j) sum += x[j] * cos(M_PI / n * (j + .5) * k);
                y[k] = (float)sum;
            }
            continue;
        }
        float tmp[16384];
        const float *src = x;
        if (kind == 2) {
            tmp[0] = x[0];
            for (int j = 1; j < n; ++j) tmp[j] = x[n - j];
            src = tmp;
        }
        for (int k = 0; k < n; ++k) {
            float sum = 0;
            for (int j = 1; j < n; ++j) sum += src[j] * (float)cos(M_PI / n * (k + .5) * j);
            y[k] = 0.5f * src[0] + sum;
            if (kind == 2 && (k & 1)) y[k] = -y[k];
        }
    }
}

/* fft.cpp:130-191 spectral step, shared by both field routines. */
static void spectral(double *a, int Mx, int My, double wy_scale, double *ph, double *ehx,
                     double *ehy)
{
    double s = 4.0 / Mx / My;
    for (int u = 0; u < Mx; ++u) {
        double wx = M_PI * u / Mx;
        for (int v = 0; v < My; ++v) {
            double wy = M_PI * v / My * wy_scale;
            size_t b = (size_t)u * My + v;
            double val = a[b] * s;
            if (u == 0) val *= 0.5;
            if (v == 0) val *= 0.5;
            double p = 0.0;
            if (u != 0 || v != 0) p = val / (wx * wx + wy * wy);
            ph[b] = p;
            ehx[b] = p * wx;
            ehy[b] = p * wy;
        }
    }
}

void eo_field(const double *rho, int Mx, int My, double wy_scale, double *phi, double *ex,
              double *ey)
{
    size_t n = (size_t)Mx * My;
    double *CyT = trig_table(My, 0, 1);  /* [j][v] = cos(pi v (j+1/2)/My) */
    double *Cy = trig_table(My, 0, 0);   /* [v][j]                       */
    double *Sy = trig_table(My, 1, 0);   /* [v][j] = sin(...)             */
    double *Cx = trig_table(Mx, 0, 0);   /* [u][i]                       */
    double *CxT = trig_table(Mx, 0, 1);  /* [i][u]                       */
    double *SxT = trig_table(Mx, 1, 1);  /* [i][u] = sin(pi u (i+1/2)/Mx) */
    double *T = (double *)malloc(sizeof(double) * n);
    double *A = (double *)malloc(sizeof(double) * n);
    double *ph = (double *)malloc(sizeof(double) * n);
    double *hx = (double *)malloc(sizeof(double) * n);
    double *hy = (double *)malloc(sizeof(double) * n);

    dgemm(Mx, My, My, rho, CyT, T); /* T[i][v] = sum_j rho[i][j] cos_v(j) */
    dgemm(Mx, My, Mx, Cx, T, A);    /* A[u][v] = sum_i cos_u(i) T[i][v]   */
    spectral(A, Mx, My, wy_scale, ph, hx, hy);
    if (phi) {
        dgemm(Mx, My, My, ph, Cy, T);
        dgemm(Mx, My, Mx, CxT, T, phi);
    }
    dgemm(Mx, My, My, hx, Cy, T);
    dgemm(Mx, My, Mx, SxT, T, ex);
    dgemm(Mx, My, My, hy, Sy, T);
    dgemm(Mx, My, Mx, CxT, T, ey);

    free(CyT); free(Cy); free(Sy); free(Cx); free(CxT); free(SxT);
    free(T); free(A); free(ph); free(hx); free(hy);
}

void eo_field_direct(const double *rho, int Mx, int My, double wy_scale, double *phi,
                     double *ex, double *ey)
{
    size_t n = (size_t)Mx * My;
    double *A = (double *)calloc(n, sizeof(double));
    double *ph = (double *)malloc(sizeof(double) * n);
    double *hx = (double *)malloc(sizeof(double) * n);
    double *hy = (double *)malloc(sizeof(double) * n);
    for (int u = 0; u < Mx; ++u)
        for (int v = 0; v < My; ++v) {
            double s = 0.0;
            for (int i = 0; i < Mx; ++i)
                for (int j = 0; j < My; ++j)
                    s += rho[(size_t)i * My + j] * cos(M_PI * u * (i + 0.5) / Mx) *
                         cos(M_PI * v * (j + 0.5) / My);
            A[(size_t)u * My + v] = s;
        }
    spectral(A, Mx, My, wy_scale, ph, hx, hy);
    for (int i = 0; i < Mx; ++i)
        for (int j = 0; j < My; ++j) {
            double sp = 0, sx = 0, sy = 0;
            for (int u = 0; u < Mx; ++u)
                for (int v = 0; v < My; ++v) {
                    size_t b = (size_t)u * My + v;
                    double cu = cos(M_PI * u * (i + 0.5) / Mx), su = sin(M_PI * u * (i + 0.5) / Mx);
                    double cv = cos(M_PI * v * (j + 0.5) / My), sv = sin(M_PI * v * (j + 0.5) / My);
                    sp += ph[b] * cu * cv;
                    sx += hx[b] * su * cv;
                    sy += hy[b] * cu * sv;
                }
            size_t o = (size_t)i * My + j;
            if (phi) phi[o] = sp;
            ex[o] = sx;
            ey[o] = sy;
        }
    free(A); free(ph); free(hx); free(hy);
}

/* AIEplace.cpp:1004-1016: force = sum over overlapped bins of overlap * E_b
 * (smoothed rectangle, scaled by the same area ratio as the deposit). */
void eo_density_force(const eo_design *d, const double *x, const double *y, const double *ex,
                      const double *ey, double *fx, double *fy)
{
    int Mx = d->bins_x, My = d->bins_y;
    double bw = (d->xh - d->xl) / Mx, bh = (d->yh - d->yl) / My;
    int np = n_place(d);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < np; ++c) {
        double x0, x1, y0, y1, r;
        stretched_rect(d, x[c], y[c], c, &x0, &x1, &y0, &y1, &r);
        int i0 = (int)floor((x0 - d->xl) / bw), i1 = (int)floor((x1 - d->xl) / bw);
        int j0 = (int)floor((y0 - d->yl) / bh), j1 = (int)floor((y1 - d->yl) / bh);
        if (i0 < 0) i0 = 0;
        if (j0 < 0) j0 = 0;
        if (i1 > Mx - 1) i1 = Mx - 1;
        if (j1 > My - 1) j1 = My - 1;
        double sx = 0, sy = 0;
        for (int i = i0; i <= i1; ++i) {
            double ox = fmin(d->xl + (i + 1) * bw, x1) - fmax(d->xl + i * bw, x0);
            if (ox <= 0) continue;
            for (int j = j0; j <= j1; ++j) {
                double oy = fmin(d->yl + (j + 1) * bh, y1) - fmax(d->yl + j * bh, y0);
                if (oy <= 0) continue;
                size_t b = (size_t)i * My + j;
                sx += ox * oy * ex[b];
                sy += ox * oy * ey[b];
            }
        }
        fx[c] = r * sx;
        fy[c] = r * sy;
    }
}

/* ------------------------------------------------------------------ */
/* full gradient and the Nesterov placer                              */
/* ------------------------------------------------------------------ */

typedef struct ws {
    size_t nb;
    double *map_mov, *map_fix, *map_field, *rho, *ex, *ey;
    double *gwx, *gwy, *fx, *fy;
    int32_t *npins;
} ws_t;

static void ws_init(const eo_design *d, ws_t *w)
{
    int np = n_place(d);
    w->nb = (size_t)d->bins_x * d->bins_y;
    w->map_mov = (double *)calloc(w->nb, sizeof(double));
    w->map_fix = (double *)calloc(w->nb, sizeof(double));
    w->map_field = (double *)calloc(w->nb, sizeof(double));
    w->rho = (double *)calloc(w->nb, sizeof(double));
    w->ex = (double *)calloc(w->nb, sizeof(double));
    w->ey = (double *)calloc(w->nb, sizeof(double));
    w->gwx = (double *)calloc((size_t)np, sizeof(double));
    w->gwy = (double *)calloc((size_t)np, sizeof(double));
    w->fx = (double *)calloc((size_t)np, sizeof(double));
    w->fy = (double *)calloc((size_t)np, sizeof(double));
    w->npins = (int32_t *)calloc((size_t)np, sizeof(int32_t));
    eo_density_fixed(d, w->map_fix);
    for (int n = 0; n < d->n_nets; ++n)
        for (int p = d->net_start[n]; p < d->net_start[n + 1]; ++p) {
            int c = d->pin_cell[p];
            if (c < d->n_mov) w->npins[c]++;
        }
}

static void ws_free(ws_t *w)
{
    free(w->map_mov); free(w->map_fix); free(w->map_field); free(w->rho);
    free(w->ex); free(w->ey); free(w->gwx); free(w->gwy); free(w->fx); free(w->fy);
    free(w->npins);
}

/* density pass only: fills map_mov (exact, non-filler) and returns tau */
static double density_tau(const eo_design *d, const double *x, const double *y, double rho_t,
                          ws_t *w)
{
    memset(w->map_mov, 0, sizeof(double) * w->nb);
    eo_density_exact(d, x, y, 0, d->n_mov, w->map_mov);
    return eo_overflow(d, w->map_mov, w->map_fix, rho_t);
}

/* WL gradient, density map, field, force; returns tau.  Components in ws. */
static double eval_components(const eo_design *d, const double *x, const double *y, double gamma,
                              double rho_t, ws_t *w, double *hpwl)
{
    int Mx = d->bins_x, My = d->bins_y;
    double bw = (d->xh - d->xl) / Mx, bh = (d->yh - d->yl) / My;
    double tau = density_tau(d, x, y, rho_t, w);
    memset(w->map_field, 0, sizeof(double) * w->nb);
    eo_density_stretched(d, x, y, 0, n_place(d), w->map_field);
    double inv_area = 1.0 / (bw * bh);
    for (size_t b = 0; b < w->nb; ++b) w->rho[b] = (w->map_field[b] + rho_t * w->map_fix[b]) * inv_area;
    eo_field(w->rho, Mx, My, bw / bh, NULL, w->ex, w->ey);
    eo_density_force(d, x, y, w->ex, w->ey, w->fx, w->fy);
    eo_wa_grad(d, x, y, gamma, w->gwx, w->gwy, NULL);
    if (hpwl) *hpwl = eo_hpwl(d, x, y);
    return tau;
}

static void assemble(const eo_design *d, const ws_t *w, double lambda, double *gx, double *gy)
{
    int np = n_place(d);
    for (int c = 0; c < np; ++c) {
        double pre = (double)w->npins[c] + lambda * d->w[c] * d->h[c];
        if (pre < 1.0) pre = 1.0;
        gx[c] = (w->gwx[c] - lambda * w->fx[c]) / pre;
        gy[c] = (w->gwy[c] - lambda * w->fy[c]) / pre;
    }
}

double eo_gradient(const eo_design *d, const double *x, const double *y, double rho_t,
                   double gamma, double lambda, double *gx, double *gy, double *hpwl,
                   double *gwx, double *gwy, double *fx, double *fy, double *ex, double *ey)
{
    ws_t w;
    ws_init(d, &w);
    double tau = eval_components(d, x, y, gamma, rho_t, &w, hpwl);
    assemble(d, &w, lambda, gx, gy);
    int np = n_place(d);
    if (gwx) memcpy(gwx, w.gwx, sizeof(double) * (size_t)np);
    if (gwy) memcpy(gwy, w.gwy, sizeof(double) * (size_t)np);
    if (fx) memcpy(fx, w.fx, sizeof(double) * (size_t)np);
    if (fy) memcpy(fy, w.fy, sizeof(double) * (size_t)np);
    if (ex) memcpy(ex, w.ex, sizeof(double) * w.nb);
    if (ey) memcpy(ey, w.ey, sizeof(double) * w.nb);
    ws_free(&w);
    return tau;
}

static double gamma_of(const eo_params *p, double bw, double tau)
{
    double g = p->gamma_bins * bw;
    if (p->gamma_schedule) g *= pow(10.0, (tau - 0.1) * 20.0 / 9.0 - 1.0);
    return g;
}

static void clamp_pos(const eo_design *d, double *x, double *y)
{
    int np = n_place(d);
    for (int c = 0; c < np; ++c) {
        double xmax = d->xh - d->w[c], ymax = d->yh - d->h[c];
        if (x[c] < d->xl) x[c] = d->xl;
        if (x[c] > xmax) x[c] = xmax;
        if (y[c] < d->yl) y[c] = d->yl;
        if (y[c] > ymax) y[c] = ymax;
    }
}

static double dist2(int n, const double *ax, const double *ay, const double *bx, const double *by)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        double dx = ax[i] - bx[i], dy = ay[i] - by[i];
        s += dx * dx + dy * dy;
    }
    return s;
}

int eo_place(const eo_design *d, const eo_params *p, double *x, double *y, eo_iter_log *log,
             int log_cap, double *final_hpwl, double *final_overflow, int64_t *total_evals)
{
    int np = n_place(d);
    size_t sz = sizeof(double) * (size_t)np;
    double bw = (d->xh - d->xl) / d->bins_x;
    double rho_t = p->target_density;
    ws_t w;
    ws_init(d, &w);

    double *ux = (double *)malloc(sz), *uy = (double *)malloc(sz);
    double *vx = (double *)malloc(sz), *vy = (double *)malloc(sz);
    double *gx = (double *)malloc(sz), *gy = (double *)malloc(sz);
    double *nux = (double *)malloc(sz), *nuy = (double *)malloc(sz);
    double *nvx = (double *)malloc(sz), *nvy = (double *)malloc(sz);
    double *ngx = (double *)malloc(sz), *ngy = (double *)malloc(sz);
    int64_t evals = 0;

    clamp_pos(d, x, y);
    memcpy(vx, x, sz);
    memcpy(vy, y, sz);
    memcpy(ux, x, sz);
    memcpy(uy, y, sz);

    /* iteration-0 parameters */
    double tau = density_tau(d, vx, vy, rho_t, &w);
    double gamma = gamma_of(p, bw, tau);
    double hpwl;
    tau = eval_components(d, vx, vy, gamma, rho_t, &w, &hpwl);
    evals++;
    double sw = 0, sf = 0;
    for (int c = 0; c < np; ++c) {
        sw += fabs(w.gwx[c]) + fabs(w.gwy[c]);
        sf += fabs(w.fx[c]) + fabs(w.fy[c]);
    }
    double lambda = sf > 0 ? p->lambda_eta * sw / sf : p->lambda_eta;
    assemble(d, &w, lambda, gx, gy);

    /* alpha_0 from a probe point v_-1 = v_0 - eps * g_0 */
    double gmax = 0;
    for (int c = 0; c < np; ++c) {
        if (fabs(gx[c]) > gmax) gmax = fabs(gx[c]);
        if (fabs(gy[c]) > gmax) gmax = fabs(gy[c]);
    }
    double eps = gmax > 0 ? p->init_disp_bins * bw / gmax : 0.0;
    for (int c = 0; c < np; ++c) {
        nvx[c] = vx[c] - eps * gx[c];
        nvy[c] = vy[c] - eps * gy[c];
    }
    clamp_pos(d, nvx, nvy);
    eval_components(d, nvx, nvy, gamma, rho_t, &w, NULL);
    evals++;
    assemble(d, &w, lambda, ngx, ngy);
    double num = sqrt(dist2(np, vx, vy, nvx, nvy)), den = sqrt(dist2(np, gx, gy, ngx, ngy));
    double alpha = den > 0 ? num / den : eps;
    double ak = 1.0;
    double prev_hpwl = hpwl;
    int it = 0;
    double cur_tau = tau, cur_hpwl = hpwl;

    for (it = 0; it < p->max_iters; ++it) {
        double ak1 = 0.5 * (1.0 + sqrt(4.0 * ak * ak + 1.0));
        double coef = (ak - 1.0) / ak1;
        double alpha_hat = alpha, alpha_used = alpha;
        int trials = 0;
        for (int t = 0; t < p->max_backtrack; ++t) {
            for (int c = 0; c < np; ++c) {
                nux[c] = vx[c] - alpha * gx[c];
                nuy[c] = vy[c] - alpha * gy[c];
            }
            clamp_pos(d, nux, nuy);
            for (int c = 0; c < np; ++c) {
                nvx[c] = nux[c] + coef * (nux[c] - ux[c]);
                nvy[c] = nuy[c] + coef * (nuy[c] - uy[c]);
            }
            clamp_pos(d, nvx, nvy);
            cur_tau = eval_components(d, nvx, nvy, gamma, rho_t, &w, &cur_hpwl);
            evals++;
            trials++;
            assemble(d, &w, lambda, ngx, ngy);
            double nn = sqrt(dist2(np, nvx, nvy, vx, vy));
            double dd = sqrt(dist2(np, ngx, ngy, gx, gy));
            alpha_hat = dd > 0 ? nn / dd : alpha;
            alpha_used = alpha;
            if (alpha_hat >= 0.95 * alpha || t == p->max_backtrack - 1) break;
            alpha = alpha_hat;
        }
        /* accept */
        memcpy(ux, nux, sz); memcpy(uy, nuy, sz);
        memcpy(vx, nvx, sz); memcpy(vy, nvy, sz);
        memcpy(gx, ngx, sz); memcpy(gy, ngy, sz);
        alpha = alpha_hat;
        ak = ak1;
        if (log && it < log_cap) {
            log[it].iter = it;
            log[it].evals = trials;
            log[it].hpwl = cur_hpwl;
            log[it].overflow = cur_tau;
            log[it].lambda = lambda;
            log[it].gamma = gamma;
            log[it].alpha = alpha;
            log[it].alpha_used = alpha_used;
            log[it].coef = coef;
        }
        /* lambda / gamma schedules (RePlAce-style mu, DREAMplace gamma) */
        double dh = cur_hpwl - prev_hpwl;
        double mu = dh < 0 ? p->mu_max : pow(p->mu_max, 1.0 - dh / p->hpwl_ref);
        if (mu < p->mu_min) mu = p->mu_min;
        if (mu > p->mu_max) mu = p->mu_max;
        lambda *= mu;
        prev_hpwl = cur_hpwl;
        gamma = gamma_of(p, bw, cur_tau);
        if (cur_tau <= p->stop_overflow) {
            it++;
            break;
        }
    }

    memcpy(x, vx, sz);
    memcpy(y, vy, sz);
    if (final_hpwl) *final_hpwl = cur_hpwl;
    if (final_overflow) *final_overflow = cur_tau;
    if (total_evals) *total_evals = evals;
    free(ux); free(uy); free(vx); free(vy); free(gx); free(gy);
    free(nux); free(nuy); free(nvx); free(nvy); free(ngx); free(ngy);
    ws_free(&w);
    return it;
}

/* ------------------------------------------------------------------ */
/* step-length MLP (FP64 reference for the device training kernels)   */
/* ------------------------------------------------------------------ */

#define MI 16
#define MH 64
#define OFF_W1 34
#define OFF_B1 (OFF_W1 + MH * MI)
#define OFF_W2 (OFF_B1 + MH)
#define OFF_B2 (OFF_W2 + MH * MH)
#define OFF_W3 (OFF_B2 + MH)
#define OFF_B3 (OFF_W3 + MH)
#define MLP_N (OFF_B3 + 1)

int eo_mlp_nweights(void) { return MLP_N; }

static double gelu(double z) { return 0.5 * z * (1.0 + erf(z / M_SQRT2)); }
static double gelu_d(double z)
{
    return 0.5 * (1.0 + erf(z / M_SQRT2)) + z * exp(-0.5 * z * z) / sqrt(2.0 * M_PI);
}

/* one sample: normalised input -> z1, a1, z2, a2, normalised output */
static double mlp_sample(const double *w, const double *xin, double *xn, double *z1, double *a1,
                         double *z2, double *a2)
{
    for (int i = 0; i < MI; ++i) xn[i] = (xin[i] - w[i]) / w[16 + i];
    for (int j = 0; j < MH; ++j) {
        double s = w[OFF_B1 + j];
        for (int i = 0; i < MI; ++i) s += w[OFF_W1 + j * MI + i] * xn[i];
        z1[j] = s;
        a1[j] = gelu(s);
    }
    for (int j = 0; j < MH; ++j) {
        double s = w[OFF_B2 + j];
        for (int k = 0; k < MH; ++k) s += w[OFF_W2 + j * MH + k] * a1[k];
        z2[j] = s;
        a2[j] = gelu(s);
    }
    double o = w[OFF_B3];
    for (int k = 0; k < MH; ++k) o += w[OFF_W3 + k] * a2[k];
    return o;
}

void eo_mlp_forward(const double *w, const double *x, int n, double *out)
{
    double xn[MI], z1[MH], a1[MH], z2[MH], a2[MH];
    for (int s = 0; s < n; ++s)
        out[s] = mlp_sample(w, x + (size_t)s * MI, xn, z1, a1, z2, a2) * w[33] + w[32];
}

double eo_mlp_grad(const double *w, const double *x, const double *y, int n, double *grad)
{
    double xn[MI], z1[MH], a1[MH], z2[MH], a2[MH], d2[MH], d1[MH];
    memset(grad, 0, sizeof(double) * MLP_N);
    double loss = 0.0;
    for (int s = 0; s < n; ++s) {
        double o = mlp_sample(w, x + (size_t)s * MI, xn, z1, a1, z2, a2);
        double t = (y[s] - w[32]) / w[33];
        double e = o - t;
        loss += e * e;
        double dout = 2.0 * e / n;
        grad[OFF_B3] += dout;
        for (int k = 0; k < MH; ++k) {
            grad[OFF_W3 + k] += dout * a2[k];
            d2[k] = dout * w[OFF_W3 + k] * gelu_d(z2[k]);
        }
        for (int j = 0; j < MH; ++j) {
            grad[OFF_B2 + j] += d2[j];
            for (int k = 0; k < MH; ++k) grad[OFF_W2 + j * MH + k] += d2[j] * a1[k];
        }
        for (int k = 0; k < MH; ++k) {
            double s2 = 0.0;
            for (int j = 0; j < MH; ++j) s2 += d2[j] * w[OFF_W2 + j * MH + k];
            d1[k] = s2 * gelu_d(z1[k]);
        }
        for (int j = 0; j < MH; ++j) {
            grad[OFF_B1 + j] += d1[j];
            for (int i = 0; i < MI; ++i) grad[OFF_W1 + j * MI + i] += d1[j] * xn[i];
        }
    }
    return loss / n;
}
