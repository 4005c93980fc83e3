/*
 * gempic_oracle2d.c -- CPU restatement of GEMPIC.jl's 2D particle-mesh
 * coupling and 2D Maxwell/Poisson solvers.
 *
 * TEST INFRASTRUCTURE (see gempic_oracle.h): the checker, never the product.
 *
 * Each function follows the cited reference function with its rounding order
 * (compiled with -ffp-contract=off). The reference's FFTW transforms are
 * restated as direct DFT sums over long-double twiddle tables (the GPU uses
 * the same sums and tables); FFTW rounds differently, which the reference's
 * own tests bound at 1e-4 for these solvers (test/test_maxwell_2d_fem.jl).
 *
 * Layout: 2D arrays are nx*ny doubles, column-major with x fastest, as the
 * reference's vec(reshape(..., nx, ny)); node (i, j) (0-based) at i + j*nx.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "gempic_oracle.h"

static int64_t imod(int64_t a, int64_t n)
{
    int64_t r = a % n;
    return r < 0 ? r + n : r;
}

/* ------------------------------------------------------------------------ */
/* particle_mesh_coupling_2d.jl                                              */
/* ------------------------------------------------------------------------ */

/* ParticleMeshCoupling2D constructor (particle_mesh_coupling_2d.jl:22-45). */
int or_pmc2d_init(or_pmc2d *p, double xmin, double xmax, int64_t nx, double ymin, double ymax,
                  int64_t ny, int degree, int smoothing)
{
    if (degree < 0 || degree > OR_MAXDEG || nx < 1 || ny < 1) return -1;
    memset(p, 0, sizeof(*p));
    p->nx = nx;
    p->ny = ny;
    p->xmin = xmin;
    p->ymin = ymin;
    p->dx = (xmax - xmin) / (double)nx;     /* TwoDGrid (mesh.jl:38-39) */
    p->dy = (ymax - ymin) / (double)ny;
    p->degree = degree;
    p->n_span = degree + 1;
    if (smoothing == OR_COLLOCATION) p->scaling = 1.0 / (p->dx * p->dy);
    else if (smoothing == OR_GALERKIN) p->scaling = 1.0;
    else return -1;
    or_splinepp_tables(degree, p->pp, p->pp_fp);
    return 0;
}

/* compute_shape_factor (:54-65): ceil() cell index, offset in (0, 1]; the
 * 2D uniform_bsplines_eval_basis! (low_level_bsplines.jl:82-110) runs the
 * x and y recurrences with the same operations as the 1D one. */
static void shape_factor(const or_pmc2d *p, double xp, double yp, int64_t *ix, int64_t *iy,
                         double *vx, double *vy)
{
    xp = (xp - p->xmin) / p->dx;
    yp = (yp - p->ymin) / p->dy;
    int64_t ip = (int64_t)ceil(xp);
    int64_t jp = (int64_t)ceil(yp);
    double dxp = xp - (double)(ip - 1);
    double dyp = yp - (double)(jp - 1);
    or_bspline_basis(p->degree, dxp, vx);
    or_bspline_basis(p->degree, dyp, vy);
    *ix = ip - p->degree;
    *iy = jp - p->degree;
}

/* index_1dto2d_column_major (:76-82), 0-based result */
static int64_t idx2d(const or_pmc2d *p, int64_t i1, int64_t i2)
{
    return imod(i1, p->nx) + imod(i2, p->ny) * p->nx;
}

/* add_charge! 2D (:94-105): rho[.] += wp * scaling * Sx * Sy */
void or_pmc2d_add_charge(double *rho, const or_pmc2d *p, double x, double y, double wp)
{
    int64_t ix, iy;
    double vx[OR_MAXDEG + 1], vy[OR_MAXDEG + 1];
    shape_factor(p, x, y, &ix, &iy, vx, vy);
    for (int i1 = 1; i1 <= p->n_span; ++i1) {
        int64_t a = ix + i1 - 2;
        for (int i2 = 1; i2 <= p->n_span; ++i2) {
            int64_t b = iy + i2 - 2;
            rho[idx2d(p, a, b)] += wp * p->scaling * vx[i1 - 1] * vy[i2 - 1];
        }
    }
}

/* horner_1d of column `index` (0-based) of a (deg+1) x m column-major table */
static double horner_col(int degree, const double *pp, int64_t ld, double x, int64_t index)
{
    double res = pp[0 + index * ld];
    for (int i = 1; i <= degree; ++i) res = res * x + pp[i + index * ld];
    return res;
}

/* add_charge_pp! 2D (:127-148): floor() + 1 cell index, horner_m_2d! weights
 * (splinepp.jl:295-300) */
void or_pmc2d_add_charge_pp(double *rho, const or_pmc2d *p, double x, double y, double wp)
{
    double xp = (x - p->xmin) / p->dx;
    double yp = (y - p->ymin) / p->dy;
    int64_t ip = (int64_t)floor(xp) + 1;
    int64_t jp = (int64_t)floor(yp) + 1;
    double dxp = xp - (double)(ip - 1);
    double dyp = yp - (double)(jp - 1);
    ip = ip - p->degree;
    jp = jp - p->degree;
    const int d = p->degree, dp1 = d + 1;
    double vx[OR_MAXDEG + 1], vy[OR_MAXDEG + 1];
    for (int i = 0; i < dp1; ++i) {
        vx[i] = horner_col(d, p->pp, dp1, dxp, i);
        vy[i] = horner_col(d, p->pp, dp1, dyp, i);
    }
    for (int i1 = 1; i1 <= p->n_span; ++i1) {
        int64_t a = ip + i1 - 2;
        for (int i2 = 1; i2 <= p->n_span; ++i2) {
            int64_t b = jp + i2 - 2;
            rho[idx2d(p, a, b)] += wp * p->scaling * vx[i1 - 1] * vy[i2 - 1];
        }
    }
}

/* evaluate 2D (:189-203) */
double or_pmc2d_evaluate(const or_pmc2d *p, double x, double y, const double *field)
{
    int64_t ix, iy;
    double vx[OR_MAXDEG + 1], vy[OR_MAXDEG + 1];
    shape_factor(p, x, y, &ix, &iy, vx, vy);
    double value = 0.0;
    for (int i1 = 1; i1 <= p->n_span; ++i1) {
        int64_t a = ix + i1 - 2;
        for (int i2 = 1; i2 <= p->n_span; ++i2) {
            int64_t b = iy + i2 - 2;
            value += field[idx2d(p, a, b)] * vx[i1 - 1] * vy[i2 - 1];
        }
    }
    return value;
}

/* evaluate_multiple (:214-231): two fields, c = Sx * Sy shared */
void or_pmc2d_evaluate_multiple(const or_pmc2d *p, double x, double y, const double *f1,
                                const double *f2, double *out2)
{
    int64_t ix, iy;
    double vx[OR_MAXDEG + 1], vy[OR_MAXDEG + 1];
    shape_factor(p, x, y, &ix, &iy, vx, vy);
    double v1 = 0.0, v2 = 0.0;
    for (int i1 = 1; i1 <= p->n_span; ++i1) {
        int64_t a = ix + i1 - 2;
        for (int i2 = 1; i2 <= p->n_span; ++i2) {
            int64_t b = iy + i2 - 2;
            int64_t k = idx2d(p, a, b);
            double c = vx[i1 - 1] * vy[i2 - 1];
            v1 += f1[k] * c;
            v2 += f2[k] * c;
        }
    }
    out2[0] = v1;
    out2[1] = v2;
}

/* The reference's loop over particles (one call per particle, in order) for
 * n particles: add_charge! and evaluate / evaluate_multiple. */
void or_pmc2d_add_charge_n(double *rho, const or_pmc2d *p, int64_t n, const double *x,
                           const double *y, const double *wp)
{
    for (int64_t i = 0; i < n; ++i) or_pmc2d_add_charge(rho, p, x[i], y[i], wp[i]);
}

void or_pmc2d_evaluate_n(const or_pmc2d *p, int64_t n, const double *x, const double *y,
                         const double *field, double *out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = or_pmc2d_evaluate(p, x[i], y[i], field);
}

void or_pmc2d_evaluate_multiple_n(const or_pmc2d *p, int64_t n, const double *x,
                                  const double *y, const double *f1, const double *f2,
                                  double *out1, double *out2)
{
    double v[2];
    for (int64_t i = 0; i < n; ++i) {
        or_pmc2d_evaluate_multiple(p, x[i], y[i], f1, f2, v);
        out1[i] = v[0];
        out2[i] = v[1];
    }
}

/* b_to_pp_1d_cell! (splinepp.jl:310-316): pp[j] = sum_i b[i] * coeffs[j, i],
 * accumulated from zero with i outer */
static void b_to_pp_1d_cell(double *pp, int64_t stride, const double *pc, int dp1,
                            const double *b)
{
    for (int j = 0; j < dp1; ++j) pp[j * stride] = 0.0;
    for (int i = 0; i < dp1; ++i)
        for (int j = 0; j < dp1; ++j) pp[j * stride] += b[i] * pc[j + i * dp1];
}

/* b_to_pp_2d! (splinepp.jl:329-399): pp is ((d+1)^2) x (nx*ny) column-major */
void or_pmc2d_b_to_pp(const or_pmc2d *p, const double *b, double *pp)
{
    const int64_t n1 = p->nx, n2 = p->ny;
    const int d = p->degree, dp1 = d + 1, dp2 = d + 1;
    const int64_t ld = (int64_t)dp1 * dp2;
    double b1[OR_MAXDEG + 1], b2[OR_MAXDEG + 1], p2[OR_MAXDEG + 1];
    for (int64_t j = 1; j <= n2; ++j)
        for (int64_t i = 1; i <= n1; ++i) {
            double *col = pp + (i - 1 + n1 * (j - 1)) * ld;
            for (int l = 0; l < dp2; ++l) {
                for (int k = 0; k < dp1; ++k) {
                    /* every branch of the reference reads these same nodes */
                    int64_t gx = i - dp1 + k;        /* 0-based x node */
                    int64_t gy = j - dp2 + l;        /* 0-based y node */
                    b1[k] = b[imod(gx, n1) + imod(gy, n2) * n1];
                }
                b_to_pp_1d_cell(col + l * dp1, 1, p->pp, dp1, b1);
            }
            for (int k = 0; k < dp1; ++k) {
                for (int l = 0; l < dp2; ++l) b2[l] = col[k + dp1 * l];
                b_to_pp_1d_cell(p2, 1, p->pp, dp2, b2);
                for (int l = 0; l < dp2; ++l) col[k + dp1 * l] = p2[l];
            }
        }
}

/* evaluate_pp 2D (:162-178) with horner_2d (splinepp.jl:412-434); indices
 * floor()+1 are not wrapped (the reference indexes the pp matrix directly).
 * Returns NAN and sets *bounds_error for a position outside the table. */
double or_pmc2d_evaluate_pp(const or_pmc2d *p, double x, double y, const double *pp,
                            int *bounds_error)
{
    double xi = (x - p->xmin) / p->dx;
    double yi = (y - p->ymin) / p->dy;
    int64_t idx = (int64_t)floor(xi) + 1;
    int64_t idy = (int64_t)floor(yi) + 1;
    xi = xi - (double)(idx - 1);
    yi = yi - (double)(idy - 1);
    const int d = p->degree, dp1 = d + 1;
    const int64_t n1 = p->nx;
    if (idx < 1 || idx > n1 || idy < 1 || idy > p->ny) {
        if (bounds_error) *bounds_error = 1;
        return NAN;
    }
    double pp1d[OR_MAXDEG + 1] = {0};
    const int64_t ld = (int64_t)dp1 * dp1;
    const double *col = pp + (idx - 1 + (idy - 1) * n1) * ld;
    for (int i = 0; i <= d; ++i) {
        const double *c = col + i * dp1;
        double res = c[0];
        for (int k = 1; k <= d; ++k) res = res * xi + c[k];
        pp1d[i] = res;
    }
    double res = pp1d[0];
    for (int i = 1; i <= d; ++i) res = res * yi + pp1d[i];
    return res;
}

/* ------------------------------------------------------------------------ */
/* poisson_2d_fem.jl helpers                                                 */
/* ------------------------------------------------------------------------ */

/* spline_fem_mass_line (poisson_2d_fem.jl:87-106) */
void or_spline_fem_mass_line(int degree, double *mass_line)
{
    const int n = degree + 1;
    double x[OR_MAXDEG + 2], w[OR_MAXDEG + 2];
    double val[(OR_MAXDEG + 1) * (OR_MAXDEG + 2)];   /* val[i + j*(deg+1)] */
    or_gausslegendre(n, x, w);
    for (int k = 0; k < n; ++k) {
        x[k] += 1.0;
        x[k] *= 0.5;
        w[k] *= 0.5;
    }
    for (int j = 0; j < n; ++j) or_bspline_basis(degree, x[j], val + j * (degree + 1));
    for (int j = 0; j <= degree; ++j) mass_line[j] = 0.0;
    for (int j = 1; j <= degree + 1; ++j)
        for (int i = j; i <= degree + 1; ++i)
            for (int k = 1; k <= n; ++k)
                mass_line[j - 1] += val[(i - 1) + (k - 1) * (degree + 1)] *
                                    val[(i - j) + (k - 1) * (degree + 1)] * w[k - 1];
}

/* spline_fem_compute_mass_eig (poisson_2d_fem.jl:113-126) */
void or_spline_fem_compute_mass_eig(int64_t n_cells, int degree, const double *mass_line,
                                    double *eig)
{
    const double factor = 6.283185307179586 / (double)n_cells;
    for (int64_t k = 1; k <= n_cells; ++k) {
        double e = mass_line[0];
        for (int j = 1; j <= degree; ++j)
            e += mass_line[j] * 2.0 * cos(factor * (double)(k - 1) * (double)j);
        eig[k - 1] = e;
    }
}

/* ------------------------------------------------------------------------ */
/* Direct DFTs (FFTW sign convention) over long-double twiddles               */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t n;
    double *c, *s;      /* cos/sin(2 pi t / n), t in [0, n) */
} dft_tab;

static void dft_tab_init(dft_tab *t, int64_t n)
{
    t->n = n;
    t->c = (double *)malloc(sizeof(double) * n);
    t->s = (double *)malloc(sizeof(double) * n);
    for (int64_t i = 0; i < n; ++i) {
        long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)i / (long double)n;
        t->c[i] = (double)cosl(a);
        t->s[i] = (double)sinl(a);
    }
}

static void dft_tab_free(dft_tab *t)
{
    free(t->c);
    free(t->s);
}

/* one complex DFT of length n over strided data (re, im at base + k*stride):
 * forward X[k] = sum_j x[j] e^{-2 pi i jk/n}; backward (sign = +1) divides
 * by n, as FFTW's ifft. Sums run over j in order, twiddle index (j*k) mod n. */
static void dft1(const dft_tab *t, const double *re, const double *im, int64_t stride,
                 double *ore, double *oim, int64_t ostride, int backward)
{
    const int64_t n = t->n;
    for (int64_t k = 0; k < n; ++k) {
        double sr = 0.0, si = 0.0;
        int64_t q = 0;
        for (int64_t j = 0; j < n; ++j) {
            const double xr = re[j * stride], xi = im[j * stride];
            const double c = t->c[q], s = t->s[q];
            if (!backward) {
                sr += xr * c + xi * s;
                si += xi * c - xr * s;
            } else {
                sr += xr * c - xi * s;
                si += xi * c + xr * s;
            }
            q += k;
            if (q >= n) q -= n;
        }
        if (backward) {
            sr = sr / (double)n;
            si = si / (double)n;
        }
        ore[k * ostride] = sr;
        oim[k * ostride] = si;
    }
}

/* 2D transform of nx*ny complex data in place: x direction (for each j),
 * then y direction (for each i) */
static void dft2(const dft_tab *tx, const dft_tab *ty, double *re, double *im, int backward,
                 double *wr, double *wi)
{
    const int64_t nx = tx->n, ny = ty->n;
    for (int64_t j = 0; j < ny; ++j) {
        dft1(tx, re + j * nx, im + j * nx, 1, wr, wi, 1, backward);
        memcpy(re + j * nx, wr, sizeof(double) * nx);
        memcpy(im + j * nx, wi, sizeof(double) * nx);
    }
    for (int64_t i = 0; i < nx; ++i) {
        dft1(ty, re + i, im + i, nx, wr, wi, 1, backward);
        for (int64_t j = 0; j < ny; ++j) {
            re[i + j * nx] = wr[j];
            im[i + j * nx] = wi[j];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* maxwell_2d_fem.jl                                                         */
/* ------------------------------------------------------------------------ */
struct or_maxwell2d {
    int64_t nx, ny;
    double dx, dy;
    int s_deg_0, s_deg_1;
    double mass0x[OR_MAXDEG + 1], mass0y[OR_MAXDEG + 1];    /* mass_line_0 .* dx / .* dy */
    double mass1x[OR_MAXDEG + 1], mass1y[OR_MAXDEG + 1];
    double *eig0x, *eig0y, *eig1x, *eig1y;                  /* eig_values_mass_{0,1}_{1,2} */
    /* TwoDPoisson (poisson_2d_fem.jl:23-85) */
    double *p_eig0x, *p_eig0y, *dtm1x, *dtm1y, *d1re, *d1im, *d2re, *d2im;
    dft_tab tx, ty;
    double *wre, *wim, *tre, *tim;                          /* nx*ny scratch */
};

or_maxwell2d *or_maxwell2d_create(double xmin, double xmax, int64_t nx, double ymin, double ymax,
                                  int64_t ny, int degree)
{
    if (degree < 1 || degree > 3 || nx < 2 || ny < 2) return NULL;
    or_maxwell2d *m = (or_maxwell2d *)calloc(1, sizeof(*m));
    m->nx = nx;
    m->ny = ny;
    m->dx = (xmax - xmin) / (double)nx;
    m->dy = (ymax - ymin) / (double)ny;
    m->s_deg_0 = degree;
    m->s_deg_1 = degree - 1;
    double l0[OR_MAXDEG + 1], l1[OR_MAXDEG + 1];
    or_spline_fem_mass_line(degree, l0);
    or_spline_fem_mass_line(degree - 1, l1);
    for (int k = 0; k <= degree; ++k) {
        m->mass0x[k] = l0[k] * m->dx;
        m->mass0y[k] = l0[k] * m->dy;
    }
    for (int k = 0; k < degree; ++k) {
        m->mass1x[k] = l1[k] * m->dx;
        m->mass1y[k] = l1[k] * m->dy;
    }
    m->eig0x = (double *)calloc(nx, sizeof(double));
    m->eig1x = (double *)calloc(nx, sizeof(double));
    m->eig0y = (double *)calloc(ny, sizeof(double));
    m->eig1y = (double *)calloc(ny, sizeof(double));
    or_spline_fem_compute_mass_eig(nx, degree, m->mass0x, m->eig0x);
    or_spline_fem_compute_mass_eig(ny, degree, m->mass0y, m->eig0y);
    or_spline_fem_compute_mass_eig(nx, degree - 1, m->mass1x, m->eig1x);
    or_spline_fem_compute_mass_eig(ny, degree - 1, m->mass1y, m->eig1y);
    /* TwoDPoisson: mass_line_0 * dx (same products as .* dx) */
    m->p_eig0x = (double *)calloc(nx, sizeof(double));
    m->p_eig0y = (double *)calloc(ny, sizeof(double));
    double *p_eig1x = (double *)calloc(nx, sizeof(double));
    double *p_eig1y = (double *)calloc(ny, sizeof(double));
    or_spline_fem_compute_mass_eig(nx, degree, m->mass0x, m->p_eig0x);
    or_spline_fem_compute_mass_eig(ny, degree, m->mass0y, m->p_eig0y);
    or_spline_fem_compute_mass_eig(nx, degree - 1, m->mass1x, p_eig1x);
    or_spline_fem_compute_mass_eig(ny, degree - 1, m->mass1y, p_eig1y);
    m->dtm1x = (double *)calloc(nx, sizeof(double));
    m->dtm1y = (double *)calloc(ny, sizeof(double));
    m->d1re = (double *)calloc(nx, sizeof(double));
    m->d1im = (double *)calloc(nx, sizeof(double));
    m->d2re = (double *)calloc(ny, sizeof(double));
    m->d2im = (double *)calloc(ny, sizeof(double));
    const double twopi = 6.283185307179586;
    for (int64_t j = 2; j <= nx; ++j) {
        double angle = twopi * (double)(j - 1) / (double)nx;
        m->d1re[j - 1] = (1.0 - cos(angle)) / m->dx;
        m->d1im[j - 1] = sin(angle) / m->dx;
        m->dtm1x[j - 1] = 2.0 / (m->dx * m->dx) * (1.0 - cos(angle)) * p_eig1x[j - 1];
    }
    for (int64_t j = 2; j <= ny; ++j) {
        double angle = twopi * (double)(j - 1) / (double)ny;
        m->d2re[j - 1] = (1.0 - cos(angle)) / m->dy;
        m->d2im[j - 1] = sin(angle) / m->dy;
        m->dtm1y[j - 1] = 2.0 / (m->dy * m->dy) * (1.0 - cos(angle)) * p_eig1y[j - 1];
    }
    free(p_eig1x);
    free(p_eig1y);
    dft_tab_init(&m->tx, nx);
    dft_tab_init(&m->ty, ny);
    int64_t nn = nx * ny, mx = nx > ny ? nx : ny;
    m->wre = (double *)calloc(nn, sizeof(double));
    m->wim = (double *)calloc(nn, sizeof(double));
    m->tre = (double *)calloc(mx, sizeof(double));
    m->tim = (double *)calloc(mx, sizeof(double));
    return m;
}

void or_maxwell2d_destroy(or_maxwell2d *m)
{
    if (!m) return;
    free(m->eig0x); free(m->eig0y); free(m->eig1x); free(m->eig1y);
    free(m->p_eig0x); free(m->p_eig0y); free(m->dtm1x); free(m->dtm1y);
    free(m->d1re); free(m->d1im); free(m->d2re); free(m->d2im);
    dft_tab_free(&m->tx);
    dft_tab_free(&m->ty);
    free(m->wre); free(m->wim); free(m->tre); free(m->tim);
    free(m);
}

/* table access for the GPU comparison: which = 0..3 -> eig0x, eig0y, eig1x, eig1y;
 * 4..5 -> mass lines 0 (x, y), 6..7 -> mass lines 1 (x, y) */
const double *or_maxwell2d_table(const or_maxwell2d *m, int which)
{
    switch (which) {
    case 0: return m->eig0x;
    case 1: return m->eig0y;
    case 2: return m->eig1x;
    case 3: return m->eig1y;
    case 4: return m->mass0x;
    case 5: return m->mass0y;
    case 6: return m->mass1x;
    case 7: return m->mass1y;
    default: return NULL;
    }
}

/* TwoDLinearSolverSplineMass.solve (linear_solver_spline_mass_2d.jl:16-32):
 * out = real(ifft2(fft2(rhs) ./ (e1[i] * e2[j] * 1.0))) */
static void kron_solve(or_maxwell2d *m, const double *e1, const double *e2, const double *rhs,
                       double *out)
{
    const int64_t nx = m->nx, ny = m->ny;
    for (int64_t k = 0; k < nx * ny; ++k) {
        m->wre[k] = rhs[k];
        m->wim[k] = 0.0;
    }
    dft2(&m->tx, &m->ty, m->wre, m->wim, 0, m->tre, m->tim);
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            const double d = e1[i] * e2[j] * 1.0;
            m->wre[i + j * nx] = m->wre[i + j * nx] / d;
            m->wim[i + j * nx] = m->wim[i + j * nx] / d;
        }
    dft2(&m->tx, &m->ty, m->wre, m->wim, 1, m->tre, m->tim);
    for (int64_t k = 0; k < nx * ny; ++k) out[k] = m->wre[k];
}

/* the inverse mass matrices of 1-forms and 2-forms (maxwell_2d_fem.jl:56-66) */
static void inv_mass(or_maxwell2d *m, int form, int component, const double *rhs, double *out)
{
    const double *ex, *ey;
    if (form == 1) {
        ex = component == 1 ? m->eig1x : m->eig0x;
        ey = component == 2 ? m->eig1y : m->eig0y;
    } else {
        ex = component == 1 ? m->eig0x : m->eig1x;
        ey = component == 2 ? m->eig0y : m->eig1y;
    }
    kron_solve(m, ex, ey, rhs, out);
}

void or_maxwell2d_solve(or_maxwell2d *m, int form, int component, const double *rhs, double *out)
{
    inv_mass(m, form, component, rhs, out);
}

/* spline_fem_multiply_mass (maxwell_2d_fem.jl:300-336): banded circulant
 * product, out[r] = m[0] in[r] + sum_c m[c] (in[r+c] + in[r-c]), c ascending */
static void multiply_mass_1d(int64_t n, int degree, const double *mass, const double *in,
                             int64_t stride, double *out)
{
    for (int64_t r = 0; r < n; ++r) {
        double o = mass[0] * in[r * stride];
        for (int c = 1; c <= degree; ++c)
            o += mass[c] * (in[imod(r + c, n) * stride] + in[imod(r - c, n) * stride]);
        out[r] = o;
    }
}

/* multiply_mass_2dkron! (:343-363): along x for each column, then along y */
void or_maxwell2d_multiply_mass(or_maxwell2d *m, const double *line1, int deg1,
                                const double *line2, int deg2, const double *in, double *out)
{
    const int64_t nx = m->nx, ny = m->ny;
    double *w = (double *)malloc(sizeof(double) * nx * ny);
    double *t = (double *)malloc(sizeof(double) * (nx > ny ? nx : ny));
    for (int64_t j = 0; j < ny; ++j) {
        multiply_mass_1d(nx, deg1, line1, in + j * nx, 1, t);
        memcpy(w + j * nx, t, sizeof(double) * nx);
    }
    for (int64_t i = 0; i < nx; ++i) {
        multiply_mass_1d(ny, deg2, line2, w + i, nx, t);
        for (int64_t j = 0; j < ny; ++j) out[i + j * nx] = t[j];
    }
    free(w);
    free(t);
}

static void mass_lines_for(or_maxwell2d *m, int form, int component, const double **l1, int *d1,
                           const double **l2, int *d2)
{
    /* inner_product's table (:511-575): 0-form (0,0); 1-form c1 (1,0), c2 (0,1),
     * c3 (0,0); 2-form c1 (0,1), c2 (1,0), c3 (1,1); 3-form (1,1) */
    int a = 0, b = 0;
    if (form == 1) { a = component == 1; b = component == 2; }
    else if (form == 2) { a = component != 1; b = component != 2; }
    else if (form == 3) { a = 1; b = 1; }
    *l1 = a ? m->mass1x : m->mass0x;
    *d1 = a ? m->s_deg_1 : m->s_deg_0;
    *l2 = b ? m->mass1y : m->mass0y;
    *d2 = b ? m->s_deg_1 : m->s_deg_0;
}

/* inner_product (:511-575): sum(c1 .* M c2), summed left to right */
double or_maxwell2d_inner_product(or_maxwell2d *m, const double *c1, const double *c2,
                                  int component, int form)
{
    const int64_t nn = m->nx * m->ny;
    const double *l1, *l2;
    int d1, d2;
    mass_lines_for(m, form, component, &l1, &d1, &l2, &d2);
    double *w = (double *)malloc(sizeof(double) * nn);
    or_maxwell2d_multiply_mass(m, l1, d1, l2, d2, c2, w);
    double r = 0.0;
    for (int64_t k = 0; k < nn; ++k) r += c1[k] * w[k];
    free(w);
    return r;
}

/* compute_e_from_b! 2D (:370-412); e, b: 3 arrays of nx*ny */
void or_maxwell2d_compute_e_from_b(or_maxwell2d *m, double dt, double *const *b, double *const *e)
{
    const int64_t nx = m->nx, ny = m->ny, nn = nx * ny;
    double *w[3], *cb[3];
    for (int c = 0; c < 3; ++c) {
        w[c] = (double *)malloc(sizeof(double) * nn);
        cb[c] = (double *)malloc(sizeof(double) * nn);
    }
    or_maxwell2d_multiply_mass(m, m->mass0x, m->s_deg_0, m->mass1y, m->s_deg_1, b[0], w[0]);
    or_maxwell2d_multiply_mass(m, m->mass1x, m->s_deg_1, m->mass0y, m->s_deg_0, b[1], w[1]);
    or_maxwell2d_multiply_mass(m, m->mass1x, m->s_deg_1, m->mass1y, m->s_deg_1, b[2], w[2]);
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            int64_t k = i + j * nx;
            int64_t k1 = imod(i + 1, nx) + j * nx;       /* next node in x */
            int64_t k2 = i + imod(j + 1, ny) * nx;       /* next node in y */
            cb[0][k] = -(w[2][k] - w[2][k2]) / m->dy;
            cb[1][k] = (w[2][k] - w[2][k1]) / m->dx;
            cb[2][k] = (w[0][k] - w[0][k2]) / m->dy - (w[1][k] - w[1][k1]) / m->dx;
        }
    for (int c = 0; c < 3; ++c) {
        inv_mass(m, 1, c + 1, cb[c], w[c]);
        for (int64_t k = 0; k < nn; ++k) e[c][k] = e[c][k] + dt * w[c][k];
    }
    for (int c = 0; c < 3; ++c) {
        free(w[c]);
        free(cb[c]);
    }
}

/* compute_b_from_e! 2D (:423-444) */
void or_maxwell2d_compute_b_from_e(or_maxwell2d *m, double dt, double *const *e, double *const *b)
{
    const int64_t nx = m->nx, ny = m->ny;
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            int64_t k = i + j * nx;
            int64_t k1 = imod(i - 1, nx) + j * nx;       /* previous node in x */
            int64_t k2 = i + imod(j - 1, ny) * nx;       /* previous node in y */
            b[0][k] += -dt * (e[2][k] - e[2][k2]) / m->dy;
            b[1][k] += dt * (e[2][k] - e[2][k1]) / m->dx;
            b[2][k] += -dt * ((e[1][k] - e[1][k1]) / m->dx - (e[0][k] - e[0][k2]) / m->dy);
        }
}

/* compute_e_from_j! 2D (:455-459): e -= M1^{-1} current (no dx factor) */
void or_maxwell2d_compute_e_from_j(or_maxwell2d *m, const double *current, int component, double *e)
{
    const int64_t nn = m->nx * m->ny;
    double *w = (double *)malloc(sizeof(double) * nn);
    inv_mass(m, 1, component, current, w);
    for (int64_t k = 0; k < nn; ++k) e[k] = e[k] - w[k];
    free(w);
}

/* compute_rho_from_e! (:468-499): rho = -(G^T M1 e) */
void or_maxwell2d_compute_rho_from_e(or_maxwell2d *m, double *const *e, double *rho)
{
    const int64_t nx = m->nx, ny = m->ny, nn = nx * ny;
    double *w1 = (double *)malloc(sizeof(double) * nn);
    double *w2 = (double *)malloc(sizeof(double) * nn);
    or_maxwell2d_multiply_mass(m, m->mass1x, m->s_deg_1, m->mass0y, m->s_deg_0, e[0], w1);
    or_maxwell2d_multiply_mass(m, m->mass0x, m->s_deg_0, m->mass1y, m->s_deg_1, e[1], w2);
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            int64_t k = i + j * nx;
            int64_t k1 = imod(i + 1, nx) + j * nx;
            int64_t k2 = i + imod(j + 1, ny) * nx;
            rho[k] = (w1[k] - w1[k1]) / m->dx + (w2[k] - w2[k2]) / m->dy;
        }
    for (int64_t k = 0; k < nn; ++k) rho[k] *= -1.0;
    free(w1);
    free(w2);
}

/* TwoDPoisson compute_e_from_rho! (poisson_2d_fem.jl:237-262): e[0], e[1] */
void or_maxwell2d_compute_e_from_rho(or_maxwell2d *m, const double *rho, double *const *e)
{
    const int64_t nx = m->nx, ny = m->ny, nn = nx * ny;
    double *sr = (double *)malloc(sizeof(double) * nn), *si = (double *)malloc(sizeof(double) * nn);
    double *xr = (double *)malloc(sizeof(double) * nn), *xi = (double *)malloc(sizeof(double) * nn);
    double *yr = (double *)malloc(sizeof(double) * nn), *yi = (double *)malloc(sizeof(double) * nn);
    for (int64_t k = 0; k < nn; ++k) {
        sr[k] = rho[k];
        si[k] = 0.0;
    }
    dft2(&m->tx, &m->ty, sr, si, 0, m->tre, m->tim);       /* fft2d!: x then y */
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            int64_t k = i + j * nx;
            if (i == 0 && j == 0) {
                sr[k] = 0.0;
                si[k] = 0.0;
            } else {
                double ev = m->dtm1x[i] * m->p_eig0y[j] + m->p_eig0x[i] * m->dtm1y[j];
                sr[k] = sr[k] / ev;
                si[k] = si[k] / ev;
            }
            /* -scratch * eig_d (complex) */
            double ar = -sr[k], ai = -si[k];
            xr[k] = ar * m->d1re[i] - ai * m->d1im[i];
            xi[k] = ar * m->d1im[i] + ai * m->d1re[i];
            yr[k] = ar * m->d2re[j] - ai * m->d2im[j];
            yi[k] = ar * m->d2im[j] + ai * m->d2re[j];
        }
    /* ifft2d!: y direction (for each i) then x direction (for each j) */
    double *pr[2] = {xr, yr}, *pi_[2] = {xi, yi};
    for (int c = 0; c < 2; ++c) {
        for (int64_t i = 0; i < nx; ++i) {
            dft1(&m->ty, pr[c] + i, pi_[c] + i, nx, m->tre, m->tim, 1, 1);
            for (int64_t j = 0; j < ny; ++j) {
                pr[c][i + j * nx] = m->tre[j];
                pi_[c][i + j * nx] = m->tim[j];
            }
        }
        for (int64_t j = 0; j < ny; ++j) {
            dft1(&m->tx, pr[c] + j * nx, pi_[c] + j * nx, 1, m->tre, m->tim, 1, 1);
            for (int64_t i = 0; i < nx; ++i) e[c][i + j * nx] = m->tre[i];
        }
    }
    free(sr); free(si); free(xr); free(xi); free(yr); free(yi);
}

/* compute_fem_rhs! (:212-282) / compute_rhs_from_function (:123-197) for a
 * callback f(x, y, ctx); form/component pick the degrees as the reference. */
void or_maxwell2d_fem_rhs(or_maxwell2d *m, double (*f)(double, double, void *), void *ctx,
                          int component, int form, double *coefs)
{
    int d1, d2;
    const int s0 = m->s_deg_0, s1 = m->s_deg_1;
    if (form == 0) { d1 = s0; d2 = s0; }
    else if (form == 1) { d1 = s0; d2 = s0; if (component == 1) d1 = s1; if (component == 2) d2 = s1; }
    else if (form == 2) { d1 = s1; d2 = s1; if (component == 1) d1 = s0; if (component == 2) d2 = s0; }
    else { d1 = s1; d2 = s1; }
    double x1[OR_MAXDEG + 2], w1[OR_MAXDEG + 2], x2[OR_MAXDEG + 2], w2[OR_MAXDEG + 2];
    double b1[(OR_MAXDEG + 2) * (OR_MAXDEG + 2)], b2[(OR_MAXDEG + 2) * (OR_MAXDEG + 2)];
    or_gausslegendre(d1 + 1, x1, w1);
    or_gausslegendre(d2 + 1, x2, w2);
    for (int k = 0; k <= d1; ++k) { x1[k] += 1.0; x1[k] /= 2.0; w1[k] /= 2.0; }
    for (int k = 0; k <= d2; ++k) { x2[k] += 1.0; x2[k] /= 2.0; w2[k] /= 2.0; }
    /* bspl_d[k, :] = basis at gauss point k: row k, column c at [k + c*(d+1)] */
    for (int k = 0; k <= d1; ++k) {
        double v[OR_MAXDEG + 2];
        or_bspline_basis(d1, x1[k], v);
        for (int c = 0; c <= d1; ++c) b1[k + c * (d1 + 1)] = v[c];
    }
    for (int k = 0; k <= d2; ++k) {
        double v[OR_MAXDEG + 2];
        or_bspline_basis(d2, x2[k], v);
        for (int c = 0; c <= d2; ++c) b2[k + c * (d2 + 1)] = v[c];
    }
    const double dx = m->dx, dy = m->dy;
    int64_t ind = 0;
    for (int64_t i2 = 1; i2 <= m->ny; ++i2)
        for (int64_t i1 = 1; i1 <= m->nx; ++i1) {
            double coef = 0.0;
            for (int j1 = 1; j1 <= d1 + 1; ++j1)
                for (int j2 = 1; j2 <= d2 + 1; ++j2)
                    for (int k1 = 1; k1 <= d1 + 1; ++k1)
                        for (int k2 = 1; k2 <= d2 + 1; ++k2) {
                            double xg = dx * (x1[k1 - 1] + (double)i1 + (double)j1 - 2.0);
                            double yg = dy * (x2[k2 - 1] + (double)i2 + (double)j2 - 2.0);
                            coef += w1[k1 - 1] * w2[k2 - 1] * f(xg, yg, ctx) *
                                    b1[(k1 - 1) + (d1 + 2 - j1 - 1) * (d1 + 1)] *
                                    b2[(k2 - 1) + (d2 + 2 - j2 - 1) * (d2 + 1)];
                        }
            coefs[ind++] = coef * dx * dy;
        }
}

/* l2projection (:288-298): the FEM rhs, then the inverse mass of the form
 * (forms 1 and 2; other forms return the rhs) */
void or_maxwell2d_l2projection(or_maxwell2d *m, double (*f)(double, double, void *), void *ctx,
                               int component, int form, double *out)
{
    const int64_t nn = m->nx * m->ny;
    double *rhs = (double *)malloc(sizeof(double) * nn);
    or_maxwell2d_fem_rhs(m, f, ctx, component, form, rhs);
    if (form == 1 || form == 2) inv_mass(m, form, component, rhs, out);
    else memcpy(out, rhs, sizeof(double) * nn);
    free(rhs);
}
