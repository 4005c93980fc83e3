/*
 * gempic_oracle.c -- CPU restatement of GEMPIC.jl's 1D particle-field path.
 *
 * TEST INFRASTRUCTURE ONLY (see gempic_oracle.h): the checker for the HIP
 * library, never the thing measured or shipped.
 *
 * Compiled with -ffp-contract=off: each a*b+c below is rounded twice, as in
 * the reference's (non-@fastmath, non-@muladd) Julia code.
 */
#include "gempic_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Julia Base semantics                                                      */
/* ------------------------------------------------------------------------ */

/* Base.mod(x::Float64, y::Float64): r = rem(x, y) (exact, = C fmod); a zero r
 * takes the sign of y; a nonzero r whose sign differs from y's gets y added
 * (that sum can round to y itself). Used at hamiltonian_splitting_1d2v.jl:79
 * and hamiltonian_splitting_1d1v.jl:91. */
double or_jl_mod_f64(double x, double y)
{
    double r = fmod(x, y);
    if (r == 0.0) return copysign(0.0, y);
    if ((r > 0.0) != (y > 0.0)) return r + y;
    return r;
}

/* Base.mod(a::Int, n::Int): floored modulus, result in [0, n) for n > 0. */
int64_t or_jl_mod_i64(int64_t a, int64_t n)
{
    int64_t r = a % n;
    if (r != 0 && ((r < 0) != (n < 0))) r += n;
    return r;
}

/* ------------------------------------------------------------------------ */
/* low_level_bsplines.jl                                                     */
/* ------------------------------------------------------------------------ */

/* uniform_bsplines_eval_basis! (low_level_bsplines.jl:63-80): de Boor
 * recurrence for the degree+1 non-zero uniform B-splines at offset in [0,1]. */
void or_bspline_basis(int degree, double offset, double *bspl)
{
    bspl[0] = 1.0;
    for (int j = 1; j <= degree; ++j) {
        double xx = -offset;
        double j_real = (double)j;
        double inv_j = 1.0 / j_real;
        double saved = 0.0;
        for (int r = 0; r < j; ++r) {
            xx = xx + 1.0;
            double temp = bspl[r] * inv_j;
            bspl[r] = saved + xx * temp;
            saved = (j_real - xx) * temp;
        }
        bspl[j] = saved;
    }
}

/* eval_uniform_periodic_spline_curve (low_level_bsplines.jl:121-137). Note
 * the reference sums j = 1..degree only (degree+1 terms would be complete);
 * the last term's value at a knot, B(0)_{degree+1}, is 0, so nothing is lost. */
void or_eval_uniform_periodic_spline_curve(int degree, int64_t n, const double *scoef,
                                           double *sval)
{
    double bspl[OR_MAXDEG + 1];
    or_bspline_basis(degree, 0.0, bspl);
    for (int64_t i = 1; i <= n; ++i) {
        double val = 0.0;
        for (int j = 1; j <= degree; ++j) {
            /* imj = mod1(i - j + n, n) */
            int64_t imj = or_jl_mod_i64(i - j + n - 1, n) + 1;
            val = val + bspl[j - 1] * scoef[imj - 1];
        }
        sval[i - 1] = val;
    }
}

/* ------------------------------------------------------------------------ */
/* Gauss-Legendre rule on [-1, 1]                                            */
/* ------------------------------------------------------------------------ */

/* FastGaussQuadrature.jl's gausslegendre (a dependency absent from the
 * reference tree; its version is unpinned, Project.toml has no Manifest).
 * For n <= 5 it returns the textbook closed forms below; for larger n it runs
 * Newton's method on the three-term Legendre recurrence, restated here with
 * the same closed-form/Newton split. Call sites on the path:
 * particle_mesh_coupling_1d.jl:72 (n = (deg+2)/2), maxwell_1d_fem.jl:195. */
int or_gausslegendre(int n, double *x, double *w)
{
    if (n < 1 || n > 64) return -1;
    if (n == 1) {
        x[0] = 0.0; w[0] = 2.0;
        return 0;
    }
    if (n == 2) {
        x[0] = -1.0 / sqrt(3.0); x[1] = 1.0 / sqrt(3.0);
        w[0] = 1.0; w[1] = 1.0;
        return 0;
    }
    if (n == 3) {
        x[0] = -sqrt(3.0 / 5.0); x[1] = 0.0; x[2] = sqrt(3.0 / 5.0);
        w[0] = 5.0 / 9.0; w[1] = 8.0 / 9.0; w[2] = 5.0 / 9.0;
        return 0;
    }
    if (n == 4) {
        double a = 2.0 / 7.0 * sqrt(6.0 / 5.0);
        x[0] = -sqrt(3.0 / 7.0 + a); x[1] = -sqrt(3.0 / 7.0 - a);
        x[2] = sqrt(3.0 / 7.0 - a);  x[3] = sqrt(3.0 / 7.0 + a);
        w[0] = (18.0 - sqrt(30.0)) / 36.0; w[1] = (18.0 + sqrt(30.0)) / 36.0;
        w[2] = w[1]; w[3] = w[0];
        return 0;
    }
    if (n == 5) {
        double b = 2.0 * sqrt(10.0 / 7.0);
        x[0] = -sqrt(5.0 + b) / 3.0; x[1] = -sqrt(5.0 - b) / 3.0; x[2] = 0.0;
        x[3] = sqrt(5.0 - b) / 3.0;  x[4] = sqrt(5.0 + b) / 3.0;
        w[0] = (322.0 - 13.0 * sqrt(70.0)) / 900.0; w[1] = (322.0 + 13.0 * sqrt(70.0)) / 900.0;
        w[2] = 128.0 / 225.0; w[3] = w[1]; w[4] = w[0];
        return 0;
    }
    /* Newton on P_n with the three-term recurrence, in long double, then rounded. */
    for (int i = 0; i < n; ++i) {
        long double t = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L));
        long double dp = 0.0L;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1.0L, p1 = t;
            for (int k = 2; k <= n; ++k) {
                long double p2 = ((2.0L * k - 1.0L) * t * p1 - (k - 1.0L) * p0) / k;
                p0 = p1; p1 = p2;
            }
            dp = n * (t * p1 - p0) / (t * t - 1.0L);
            long double dt = p1 / dp;
            t -= dt;
            if (fabsl(dt) < 1e-19L) break;
        }
        {
            long double p0 = 1.0L, p1 = t;
            for (int k = 2; k <= n; ++k) {
                long double p2 = ((2.0L * k - 1.0L) * t * p1 - (k - 1.0L) * p0) / k;
                p0 = p1; p1 = p2;
            }
            dp = n * (t * p1 - p0) / (t * t - 1.0L);
        }
        /* nodes ascending */
        x[n - 1 - i] = (double)t;
        w[n - 1 - i] = (double)(2.0L / ((1.0L - t * t) * dp * dp));
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* splinepp.jl                                                               */
/* ------------------------------------------------------------------------ */

/* Exact rational polynomials (coefficients low -> high, common denominator). */
typedef struct { int64_t num[OR_MAXDEG + 2]; int64_t den; int len; } rpoly;

static int64_t gcd64(int64_t a, int64_t b)
{
    if (a < 0) a = -a;
    if (b < 0) b = -b;
    while (b) { int64_t t = a % b; a = b; b = t; }
    return a;
}

/* SplinePP tables (splinepp.jl:30-233). The constants there are the
 * power-form coefficients of the degree+1 pieces of the uniform B-spline on
 * [0,1) (poly_coeffs) and of their primitives divided by x (poly_coeffs_fp).
 * Rather than transcribe them, derive them exactly from the same de Boor
 * recurrence (low_level_bsplines.jl:63-80) in rational arithmetic, then round
 * each coefficient num/den as num * (1.0/den), which is how the reference
 * writes its non-unit entries (e.g. 11 * inv_24); entries the reference
 * writes as 4*inv_6 or 26*inv_120 differ from num*(1/den) of the reduced
 * fraction only by a power-of-two scaling, which rounds identically.
 * Output layout: coef[power_row + piece*(deg+1)], power_row 0 = x^deg, i.e.
 * the same column-major (row = power, column = piece) addressing as
 * poly_coeffs[row, piece] in horner_1d (splinepp.jl:268-274). */
int or_splinepp_tables(int degree, double *poly_coeffs, double *poly_coeffs_fp)
{
    if (degree < 1 || degree > OR_MAXDEG) return -1;
    rpoly b[OR_MAXDEG + 1];
    memset(b, 0, sizeof(b));
    b[0].num[0] = 1; b[0].den = 1; b[0].len = 1;
    for (int j = 1; j <= degree; ++j) {
        rpoly nb[OR_MAXDEG + 1];
        memset(nb, 0, sizeof(nb));
        /* saved = 0; for r: xx = r+1 - x; temp = b[r]/j; nb[r] = saved + xx*temp;
         * saved = (j - xx)*temp = (j - r - 1 + x)*temp */
        rpoly saved; memset(&saved, 0, sizeof(saved)); saved.den = 1; saved.len = 1;
        for (int r = 0; r < j; ++r) {
            rpoly temp = b[r];
            temp.den *= j;
            /* xx*temp with xx = (r+1) - x */
            rpoly a; memset(&a, 0, sizeof(a));
            a.den = temp.den; a.len = temp.len + 1;
            for (int i = 0; i < temp.len; ++i) {
                a.num[i] += (r + 1) * temp.num[i];
                a.num[i + 1] -= temp.num[i];
            }
            /* nb[r] = saved + a, common denominator */
            rpoly s; memset(&s, 0, sizeof(s));
            s.den = saved.den * a.den;
            s.len = a.len > saved.len ? a.len : saved.len;
            for (int i = 0; i < s.len; ++i)
                s.num[i] = saved.num[i] * a.den + a.num[i] * saved.den;
            nb[r] = s;
            /* saved = (j - r - 1 + x) * temp */
            rpoly c; memset(&c, 0, sizeof(c));
            c.den = temp.den; c.len = temp.len + 1;
            for (int i = 0; i < temp.len; ++i) {
                c.num[i] += (j - r - 1) * temp.num[i];
                c.num[i + 1] += temp.num[i];
            }
            saved = c;
        }
        nb[j] = saved;
        for (int r = 0; r <= j; ++r) {
            int64_t g = nb[r].den;
            for (int i = 0; i < nb[r].len; ++i) g = gcd64(g, nb[r].num[i]);
            if (g > 1) {
                nb[r].den /= g;
                for (int i = 0; i < nb[r].len; ++i) nb[r].num[i] /= g;
            }
            b[r] = nb[r];
        }
    }
    int dp1 = degree + 1;
    for (int piece = 0; piece <= degree; ++piece) {
        for (int pw = 0; pw <= degree; ++pw) {
            int64_t num = pw < b[piece].len ? b[piece].num[pw] : 0;
            int64_t den = b[piece].den;
            int64_t g = gcd64(num, den);
            int64_t n1 = g ? num / g : 0, d1 = g ? den / g : 1;
            double v = (n1 == 0) ? 0.0 : (double)n1 * (1.0 / (double)d1);
            poly_coeffs[(degree - pw) + piece * dp1] = v;
            /* primitive / x: coefficient of x^pw is num/(den*(pw+1)) */
            int64_t den2 = den * (pw + 1);
            g = gcd64(num, den2);
            n1 = g ? num / g : 0; d1 = g ? den2 / g : 1;
            v = (n1 == 0) ? 0.0 : (double)n1 * (1.0 / (double)d1);
            poly_coeffs_fp[(degree - pw) + piece * dp1] = v;
        }
    }
    return 0;
}

/* horner_1d (splinepp.jl:268-274): coefficient column `index` (0-based here). */
static double horner_1d(int degree, const double *pp, int64_t ld, double x, int64_t index)
{
    double res = pp[0 + index * ld];
    for (int i = 1; i <= degree; ++i) res = res * x + pp[i + index * ld];
    return res;
}

/* b_to_pp (splinepp.jl:241-261): pp_coeffs is (deg+1) x ncells column-major.
 * Julia's sum over the (deg+1)-vector runs left to right for these lengths. */
void or_b_to_pp(int degree, int64_t ncells, const double *b_coeffs, double *pp_coeffs)
{
    double pc[(OR_MAXDEG + 1) * (OR_MAXDEG + 1)], pf[(OR_MAXDEG + 1) * (OR_MAXDEG + 1)];
    or_splinepp_tables(degree, pc, pf);
    int dp1 = degree + 1;
    double coeffs[OR_MAXDEG + 1];
    for (int64_t i = 1; i <= ncells; ++i) {
        if (i <= degree) {
            /* vcat(b[end-deg+i:end], b[1:i]) */
            int c = 0;
            for (int64_t k = ncells - degree + i; k <= ncells; ++k) coeffs[c++] = b_coeffs[k - 1];
            for (int64_t k = 1; k <= i; ++k) coeffs[c++] = b_coeffs[k - 1];
        } else {
            for (int k = 0; k < dp1; ++k) coeffs[k] = b_coeffs[i - degree - 1 + k];
        }
        for (int j = 0; j < dp1; ++j) {
            double s = coeffs[0] * pc[j + 0 * dp1];
            for (int k = 1; k < dp1; ++k) s = s + coeffs[k] * pc[j + k * dp1];
            pp_coeffs[j + (i - 1) * dp1] = s;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* particle_mesh_coupling_1d.jl                                              */
/* ------------------------------------------------------------------------ */

/* ParticleMeshCoupling1D constructor (particle_mesh_coupling_1d.jl:44-95). */
int or_pmc1d_init(or_pmc1d *p, double xmin, double xmax, int64_t nx, int degree, int smoothing)
{
    memset(p, 0, sizeof(*p));
    if (degree < 0 || degree > OR_MAXDEG || nx < 1) return -1;
    p->n_grid = nx;
    p->xmin = xmin;
    p->Lx = xmax - xmin;
    p->delta_x = (xmax - xmin) / (double)nx;
    p->degree = degree;
    p->n_span = degree + 1;
    if (smoothing == OR_COLLOCATION) p->scaling = 1.0 / p->delta_x;
    else if (smoothing == OR_GALERKIN) p->scaling = 1.0;
    else return -2;   /* ArgumentError in the reference */
    p->n_quad = (degree + 2) / 2;
    or_gausslegendre(p->n_quad, p->quad_x, p->quad_w);
    if (degree >= 1) or_splinepp_tables(degree, p->pp, p->pp_fp);
    return 0;
}

/* add_charge! (particle_mesh_coupling_1d.jl:261-280). */
void or_add_charge(double *rho, const or_pmc1d *p, double x, double q)
{
    double sv[OR_MAXDEG + 1];
    double xi = (x - p->xmin) / p->delta_x;
    int64_t index = (int64_t)xi;       /* trunc(Int, xi) */
    xi = xi - (double)index;
    index = index - p->degree;
    or_bspline_basis(p->degree, xi, sv);
    int64_t nx = p->n_grid;
    for (int i = 1; i <= p->n_span; ++i) {
        int64_t k = or_jl_mod_i64(index + i - 1, nx);   /* mod1(index+i, nx) - 1 */
        rho[k] += q * sv[i - 1] * p->scaling;
    }
}

/* add_charge_pp! (particle_mesh_coupling_1d.jl:106-122). */
void or_add_charge_pp(double *rho, const or_pmc1d *p, double x, double q)
{
    double sv[OR_MAXDEG + 1];
    double xi = (x - p->xmin) / p->delta_x;
    int64_t index = (int64_t)xi + 1;
    xi = xi - (double)(index - 1);
    index = index - p->degree;
    for (int i = 0; i < p->n_span; ++i)
        sv[i] = horner_1d(p->degree, p->pp, p->degree + 1, xi, i);
    for (int i = 1; i <= p->n_span; ++i) {
        int64_t k = or_jl_mod_i64(index + i - 2, p->n_grid);
        rho[k] += q * sv[i - 1] * p->scaling;
    }
}

/* evaluate (particle_mesh_coupling_1d.jl:438-453). */
double or_evaluate(const or_pmc1d *p, double x, const double *field)
{
    double sv[OR_MAXDEG + 1];
    double xi = (x - p->xmin) / p->delta_x;
    int64_t index = (int64_t)xi;
    double dxi = xi - (double)index;
    index = index - p->degree;
    or_bspline_basis(p->degree, dxi, sv);
    double value = 0.0;
    for (int i = 1; i <= p->n_span; ++i) {
        int64_t k = or_jl_mod_i64(index + i - 1, p->n_grid);
        value += field[k] * sv[i - 1];
    }
    return value;
}

/* evaluate_pp (particle_mesh_coupling_1d.jl:242-250): field_pp is the
 * (deg+1) x nx matrix from b_to_pp; the column index is NOT wrapped, as in
 * the reference (positions must lie in [xmin, xmin+Lx)). */
double or_evaluate_pp(const or_pmc1d *p, double x, const double *field_pp)
{
    double xi = (x - p->xmin) / p->delta_x;
    int64_t index = (int64_t)xi + 1;
    xi = xi - (double)(index - 1);
    return horner_1d(p->degree, field_pp, p->degree + 1, xi, index - 1);
}

/* update_jv! with B (particle_mesh_coupling_1d.jl:385-425). */
static double update_jv(double *j, const or_pmc1d *p, double lower, double upper, int64_t index,
                        double q, double qm, double sign, double vi, const double *b)
{
    double sv[OR_MAXDEG + 1], svm[OR_MAXDEG + 1];
    int64_t n_cells = p->n_grid;
    double c1 = 0.5 * (upper - lower);
    double c2 = 0.5 * (upper + lower);
    or_bspline_basis(p->degree, c1 * p->quad_x[0] + c2, sv);
    double s0 = p->quad_w[0] * c1;
    for (int k = 0; k < p->n_span; ++k) sv[k] *= s0;
    for (int jq = 1; jq < p->n_quad; ++jq) {
        or_bspline_basis(p->degree, c1 * p->quad_x[jq] + c2, svm);
        for (int k = 0; k < p->n_span; ++k) sv[k] = sv[k] + svm[k] * p->quad_w[jq] * c1;
    }
    double s1 = sign * p->delta_x;
    for (int k = 0; k < p->n_span; ++k) sv[k] *= s1;
    int ind = 0;
    for (int64_t ig = index - p->degree; ig <= index; ++ig) {
        int64_t im = or_jl_mod_i64(ig, n_cells);
        j[im] += q * sv[ind] * p->scaling;
        if (b) vi = vi - qm * sv[ind] * b[im];
        ind++;
    }
    return vi;
}

/* add_current_update_v! with B (particle_mesh_coupling_1d.jl:296-376):
 * trunc() for both endpoints; segment order old, new, in-between. */
double or_add_current_update_v(double *j, const or_pmc1d *p, double x_old, double x_new,
                               double q, double qm, const double *b, double vi)
{
    double xi = (x_old - p->xmin) / p->delta_x;
    int64_t index_old = (int64_t)xi;
    double r_old = xi - (double)index_old;
    xi = (x_new - p->xmin) / p->delta_x;
    int64_t index_new = (int64_t)xi;
    double r_new = xi - (double)index_new;

    if (index_old == index_new) {
        if (r_old < r_new)
            vi = update_jv(j, p, r_old, r_new, index_old, q, qm, 1.0, vi, b);
        else
            vi = update_jv(j, p, r_new, r_old, index_old, q, qm, -1.0, vi, b);
    } else if (index_old < index_new) {
        vi = update_jv(j, p, r_old, 1.0, index_old, q, qm, 1.0, vi, b);
        vi = update_jv(j, p, 0.0, r_new, index_new, q, qm, 1.0, vi, b);
        for (int64_t ind = index_old + 1; ind <= index_new - 1; ++ind)
            vi = update_jv(j, p, 0.0, 1.0, ind, q, qm, 1.0, vi, b);
    } else {
        vi = update_jv(j, p, r_new, 1.0, index_new, q, qm, -1.0, vi, b);
        vi = update_jv(j, p, 0.0, r_old, index_old, q, qm, -1.0, vi, b);
        for (int64_t ind = index_new + 1; ind <= index_old - 1; ++ind)
            vi = update_jv(j, p, 0.0, 1.0, ind, q, qm, -1.0, vi, b);
    }
    return vi;
}

/* add_current_update_v! without B, the 1d1v case
 * (particle_mesh_coupling_1d.jl:471-529, update_jv! :538-580): trunc() for
 * the old position, floor() for the new one. */
double or_add_current_update_v_nob(double *j, const or_pmc1d *p, double x_old, double x_new,
                                   double q, double qm, double vi)
{
    double xi = (x_old - p->xmin) / p->delta_x;
    int64_t index_old = (int64_t)xi;
    double r_old = xi - (double)index_old;
    xi = (x_new - p->xmin) / p->delta_x;
    int64_t index_new = (int64_t)floor(xi);
    double r_new = xi - (double)index_new;

    if (index_old == index_new) {
        if (r_old < r_new)
            vi = update_jv(j, p, r_old, r_new, index_old, q, qm, 1.0, vi, NULL);
        else
            vi = update_jv(j, p, r_new, r_old, index_old, q, qm, -1.0, vi, NULL);
    } else if (index_old < index_new) {
        vi = update_jv(j, p, r_old, 1.0, index_old, q, qm, 1.0, vi, NULL);
        vi = update_jv(j, p, 0.0, r_new, index_new, q, qm, 1.0, vi, NULL);
        for (int64_t ind = index_old + 1; ind <= index_new - 1; ++ind)
            vi = update_jv(j, p, 0.0, 1.0, ind, q, qm, 1.0, vi, NULL);
    } else {
        vi = update_jv(j, p, r_new, 1.0, index_new, q, qm, -1.0, vi, NULL);
        vi = update_jv(j, p, 0.0, r_old, index_old, q, qm, -1.0, vi, NULL);
        for (int64_t ind = index_new + 1; ind <= index_old - 1; ++ind)
            vi = update_jv(j, p, 0.0, 1.0, ind, q, qm, -1.0, vi, NULL);
    }
    return vi;
}

/* update_jv_pp! (particle_mesh_coupling_1d.jl:198-230): primitive at both
 * bounds via horner_primitive_1d (splinepp.jl:281-285). */
static double update_jv_pp(double *j, const or_pmc1d *p, double lower, double upper,
                           int64_t index, double q, double qm, double vi, const double *b)
{
    double sv[OR_MAXDEG + 1], svm[OR_MAXDEG + 1];
    int dp1 = p->degree + 1;
    for (int i = 0; i < p->n_span; ++i) sv[i] = horner_1d(p->degree, p->pp_fp, dp1, lower, i) * lower;
    for (int i = 0; i < p->n_span; ++i) svm[i] = horner_1d(p->degree, p->pp_fp, dp1, upper, i) * upper;
    for (int i = 0; i < p->n_span; ++i) sv[i] = (svm[i] - sv[i]) * p->delta_x;
    int ind = 0;
    for (int64_t ig = index - p->degree; ig <= index; ++ig) {
        int64_t im = or_jl_mod_i64(ig, p->n_grid);
        j[im] = j[im] + (q * sv[ind] * p->scaling);
        vi = vi - qm * sv[ind] * b[im];
        ind++;
    }
    return vi;
}

/* add_current_update_v_pp! (particle_mesh_coupling_1d.jl:131-190): no sign
 * argument; reversed segments pass (upper, lower) swapped instead. */
double or_add_current_update_v_pp(double *j, const or_pmc1d *p, double x_old, double x_new,
                                  double q, double qm, const double *b, double vi)
{
    double xi = (x_old - p->xmin) / p->delta_x;
    int64_t index_old = (int64_t)xi;
    double r_old = xi - (double)index_old;
    xi = (x_new - p->xmin) / p->delta_x;
    int64_t index_new = (int64_t)xi;
    double r_new = xi - (double)index_new;
    if (index_old == index_new) {
        vi = update_jv_pp(j, p, r_old, r_new, index_old, q, qm, vi, b);
    } else if (index_old < index_new) {
        vi = update_jv_pp(j, p, r_old, 1.0, index_old, q, qm, vi, b);
        vi = update_jv_pp(j, p, 0.0, r_new, index_new, q, qm, vi, b);
        for (int64_t ind = index_old + 1; ind <= index_new - 1; ++ind)
            vi = update_jv_pp(j, p, 0.0, 1.0, ind, q, qm, vi, b);
    } else {
        vi = update_jv_pp(j, p, 1.0, r_new, index_new, q, qm, vi, b);
        vi = update_jv_pp(j, p, r_old, 0.0, index_old, q, qm, vi, b);
        for (int64_t ind = index_new + 1; ind <= index_old - 1; ++ind)
            vi = update_jv_pp(j, p, 1.0, 0.0, ind, q, qm, vi, b);
    }
    return vi;
}

void or_add_charge_n(double *rho, const or_pmc1d *p, int64_t n, const double *x, const double *q)
{
    for (int64_t i = 0; i < n; ++i) or_add_charge(rho, p, x[i], q[i]);
}

/* evaluate(kernel_smoother, field_dofs, x::AbstractArray) (diagnostics.jl:260-268) */
void or_evaluate_n(const or_pmc1d *p, int64_t n, const double *x, const double *field, double *out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = or_evaluate(p, x[i], field);
}

/* ------------------------------------------------------------------------ */
/* maxwell_1d_fem.jl                                                         */
/* ------------------------------------------------------------------------ */

/* Maxwell1DFEM constructor (maxwell_1d_fem.jl:49-177): Eulerian-number mass
 * lines, circulant eigenvalues in FFTW halfcomplex layout (real part of mode
 * k at k, imaginary part at n-k). */
or_maxwell1d *or_maxwell1d_create(double xmin, double xmax, int64_t n, int degree)
{
    if (degree < 1 || degree > 3 || n < 2 || (n % 2) != 0) return NULL;
    or_maxwell1d *m = (or_maxwell1d *)calloc(1, sizeof(*m));
    m->xmin = xmin;
    m->n = n;
    m->Lx = xmax - xmin;
    m->delta_x = m->Lx / (double)n;
    m->s_deg_0 = degree;
    m->s_deg_1 = degree - 1;
    double *mass_0 = m->mass_0, *mass_1 = m->mass_1;
    if (degree == 1) {
        mass_0[0] = 4.0 / 6.0; mass_0[1] = 1.0 / 6.0;
        mass_1[0] = 1.0;
    } else if (degree == 2) {
        mass_0[0] = 66.0 / 120.0; mass_0[1] = 26.0 / 120.0; mass_0[2] = 1.0 / 120.0;
        mass_1[0] = 4.0 / 6.0; mass_1[1] = 1.0 / 6.0;
    } else {
        mass_0[0] = 2416.0 / 5040.0; mass_0[1] = 1191.0 / 5040.0;
        mass_0[2] = 120.0 / 5040.0;  mass_0[3] = 1.0 / 5040.0;
        mass_1[0] = 66.0 / 120.0; mass_1[1] = 26.0 / 120.0; mass_1[2] = 1.0 / 120.0;
    }
    m->eig_mass0 = (double *)calloc(n, sizeof(double));
    m->eig_mass1 = (double *)calloc(n, sizeof(double));
    m->eig_weak_ampere = (double *)calloc(n, sizeof(double));
    m->eig_weak_poisson = (double *)calloc(n, sizeof(double));
    m->work = (double *)calloc(n, sizeof(double));
    m->wsave = (double *)calloc(n, sizeof(double));
    m->eigvals = (double *)calloc(n, sizeof(double));
    m->cos_tab = (double *)calloc(n, sizeof(double));
    m->sin_tab = (double *)calloc(n, sizeof(double));
    for (int64_t t = 0; t < n; ++t) {
        /* exact 2*pi*t/n reduced to one period; the DFT uses (j*k mod n) */
        long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)n;
        m->cos_tab[t] = (double)cosl(ang);
        m->sin_tab[t] = (double)sinl(ang);
    }
    const double pi = 3.141592653589793;
    int s0 = degree;
    m->eig_weak_ampere[0] = 0.0;
    m->eig_weak_poisson[0] = 0.0;
    m->eig_mass0[0] = 1.0;
    m->eig_mass1[0] = 1.0;
    double nd = (double)n;
    for (int64_t k = 1; k <= n / 2 - 1; ++k) {
        double coef0 = mass_0[0];
        double coef1 = mass_1[0];
        for (int j = 1; j <= s0 - 1; ++j) {
            double cos_mode = cos(2.0 * pi * (double)j * (double)k / nd);
            coef0 = coef0 + 2.0 * mass_0[j] * cos_mode;
            coef1 = coef1 + 2.0 * mass_1[j] * cos_mode;
        }
        int j = s0;
        coef0 = coef0 + 2.0 * mass_0[j] * cos(2.0 * pi * (double)j * (double)k / nd);
        m->eig_mass0[k] = coef0;
        m->eig_mass0[n - k] = 0.0;
        m->eig_mass1[k] = coef1;
        m->eig_mass1[n - k] = 0.0;
        double cos_mode = cos(2.0 * pi * (double)k / nd);
        double sin_mode = sin(2.0 * pi * (double)k / nd);
        m->eig_weak_ampere[k] = (coef1 / coef0) * (1.0 - cos_mode);
        m->eig_weak_ampere[n - k] = -(coef1 / coef0) * sin_mode;
        double a = 1.0 - cos_mode;
        m->eig_weak_poisson[k] = 1.0 / (coef1 * (a * a + sin_mode * sin_mode));
        m->eig_weak_poisson[n - k] = 0.0;
    }
    {
        double coef0 = mass_0[0];
        double coef1 = mass_1[0];
        for (int j = 1; j <= s0 - 1; ++j) {
            coef0 = coef0 + 2.0 * mass_0[j] * cos(pi * (double)j);
            coef1 = coef1 + 2.0 * mass_1[j] * cos(pi * (double)j);
        }
        int j = s0;
        coef0 = coef0 + 2.0 * mass_0[j] * cos(pi * (double)j);
        m->eig_mass0[n / 2] = coef0;
        m->eig_mass1[n / 2] = coef1;
        m->eig_weak_ampere[n / 2] = 2.0 * (coef1 / coef0);
        m->eig_weak_poisson[n / 2] = 1.0 / (coef1 * 4.0);
    }
    return m;
}

void or_maxwell1d_destroy(or_maxwell1d *m)
{
    if (!m) return;
    free(m->eig_mass0); free(m->eig_mass1); free(m->eig_weak_ampere); free(m->eig_weak_poisson);
    free(m->work); free(m->wsave); free(m->eigvals); free(m->cos_tab); free(m->sin_tab);
    free(m);
}

void or_maxwell1d_tables(const or_maxwell1d *m, double *e0, double *e1, double *ea, double *ep)
{
    size_t nb = (size_t)m->n * sizeof(double);
    if (e0) memcpy(e0, m->eig_mass0, nb);
    if (e1) memcpy(e1, m->eig_mass1, nb);
    if (ea) memcpy(ea, m->eig_weak_ampere, nb);
    if (ep) memcpy(ep, m->eig_weak_poisson, nb);
}

const double *or_maxwell1d_work(const or_maxwell1d *m) { return m->work; }

/* FFTW R2HC (unnormalised, e^{-i}): out[k] = Re X_k (k <= n/2), out[n-k] = Im X_k.
 * Restated as a direct DFT (FFTW is a dependency absent from the reference
 * tree; only its transform definition is needed). */
void or_r2hc(const or_maxwell1d *m, const double *in, double *out)
{
    int64_t n = m->n;
    for (int64_t k = 0; k <= n / 2; ++k) {
        double re = 0.0, im = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            int64_t t = (j * k) % n;
            re += in[j] * m->cos_tab[t];
            im -= in[j] * m->sin_tab[t];
        }
        out[k] = re;
        if (k > 0 && k < n - k) out[n - k] = im;
    }
}

/* FFTW HC2R (unnormalised inverse of R2HC). */
void or_hc2r(const or_maxwell1d *m, const double *in, double *out)
{
    int64_t n = m->n;
    for (int64_t j = 0; j < n; ++j) {
        double s = in[0];
        for (int64_t k = 1; k < n - k; ++k) {
            int64_t t = (j * k) % n;
            s += 2.0 * (in[k] * m->cos_tab[t] - in[n - k] * m->sin_tab[t]);
        }
        if (n % 2 == 0) s += in[n / 2] * ((j % 2) ? -1.0 : 1.0);
        out[j] = s;
    }
}

/* solve_circulant! (maxwell_1d_fem.jl:222-240): result in m->work. */
void or_solve_circulant(or_maxwell1d *m, const double *eigvals, const double *rhs)
{
    int64_t n = m->n;
    memcpy(m->work, rhs, (size_t)n * sizeof(double));
    or_r2hc(m, m->work, m->wsave);
    m->wsave[0] = m->wsave[0] * eigvals[0];
    for (int64_t k = 2; k <= n / 2; ++k) {
        /* 1-based k and n-k+2 -> 0-based k-1 and n-k+1 */
        double re_p = m->wsave[k - 1] * eigvals[k - 1] - m->wsave[n - k + 1] * eigvals[n - k + 1];
        double im_p = m->wsave[k - 1] * eigvals[n - k + 1] + m->wsave[n - k + 1] * eigvals[k - 1];
        m->wsave[k - 1] = re_p;
        m->wsave[n - k + 1] = im_p;
    }
    m->wsave[n / 2] = m->wsave[n / 2] * eigvals[n / 2];
    or_hc2r(m, m->wsave, m->work);
    for (int64_t i = 0; i < n; ++i) m->work[i] /= (double)n;
}

/* compute_e_from_rho! (maxwell_1d_fem.jl:244-255). */
void or_compute_e_from_rho(double *e, or_maxwell1d *m, const double *rho)
{
    or_solve_circulant(m, m->eig_weak_poisson, rho);
    int64_t n = m->n;
    for (int64_t i = 1; i < n; ++i) e[i] = m->work[i - 1] - m->work[i];
    e[0] = m->work[n - 1] - m->work[0];
}

/* compute_e_from_j! (maxwell_1d_fem.jl:263-289). */
void or_compute_e_from_j(double *e, or_maxwell1d *m, const double *current, int component)
{
    int64_t n = m->n;
    memset(m->eigvals, 0, (size_t)n * sizeof(double));
    const double *src = component == 1 ? m->eig_mass1 : m->eig_mass0;
    for (int64_t i = 0; i <= n / 2; ++i) m->eigvals[i] = 1.0 / src[i];
    or_solve_circulant(m, m->eigvals, current);
    for (int64_t i = 0; i < n; ++i) m->work[i] /= m->delta_x;
    for (int64_t i = 0; i < n; ++i) e[i] -= m->work[i];
}

/* compute_e_from_b! (maxwell_1d_fem.jl:384-396). */
void or_compute_e_from_b(double *e, or_maxwell1d *m, double dt, const double *b)
{
    double coef = dt / m->delta_x;
    or_solve_circulant(m, m->eig_weak_ampere, b);
    for (int64_t i = 0; i < m->n; ++i) e[i] += coef * m->work[i];
}

/* compute_b_from_e! (maxwell_1d_fem.jl:407-420). */
void or_compute_b_from_e(double *b, or_maxwell1d *m, double dt, const double *e)
{
    double coef = dt / m->delta_x;
    int64_t n = m->n;
    for (int64_t i = 1; i < n; ++i) b[i] = b[i] + coef * (e[i - 1] - e[i]);
    b[0] = b[0] + coef * (e[n - 1] - e[0]);
}

static double eval_func(double x, double a_cos, double a_sin, double k, double c)
{
    double v = 0.0;
    if (a_cos != 0.0) v += a_cos * cos(k * x);
    if (a_sin != 0.0) v += a_sin * sin(k * x);
    return v + c;
}

/* compute_rhs_from_function! (maxwell_1d_fem.jl:188-220). */
void or_compute_rhs_from_function(double *coefs, or_maxwell1d *m, double a_cos, double a_sin,
                                  double k, double c, int degree)
{
    double x[8], w[8], bspl[8][8];
    or_gausslegendre(degree + 1, x, w);
    for (int i = 0; i <= degree; ++i) { x[i] = 0.5 * (x[i] + 1.0); w[i] = 0.5 * w[i]; }
    for (int kq = 0; kq <= degree; ++kq) or_bspline_basis(degree, x[kq], bspl[kq]);
    for (int64_t i = 1; i <= m->n; ++i) {
        double coef = 0.0;
        for (int j = 1; j <= degree + 1; ++j)
            for (int kq = 1; kq <= degree + 1; ++kq) {
                /* x[k] + i + j - 2 evaluates left to right in Float64 */
                double xq = ((x[kq - 1] + (double)i) + (double)j) - 2.0;
                coef = coef + w[kq - 1] * eval_func(m->delta_x * xq, a_cos, a_sin, k, c) *
                                  bspl[kq - 1][degree + 2 - j - 1];
            }
        coefs[i - 1] = coef * m->delta_x;
    }
}

/* l2projection! (maxwell_1d_fem.jl:347-374). */
int or_l2projection(double *coefs, or_maxwell1d *m, double a_cos, double a_sin, double k,
                    double c, int degree)
{
    int64_t n = m->n;
    double *eig = (double *)calloc(n, sizeof(double));
    or_compute_rhs_from_function(coefs, m, a_cos, a_sin, k, c, degree);
    if (degree == m->s_deg_0) {
        for (int64_t i = 0; i <= n / 2; ++i) eig[i] = 1.0 / m->eig_mass0[i];
    } else if (degree == m->s_deg_0 - 1) {
        for (int64_t i = 0; i <= n / 2; ++i) eig[i] = 1.0 / m->eig_mass1[i];
    } else {
        free(eig);
        return -1;
    }
    or_solve_circulant(m, eig, coefs);
    free(eig);
    /* the reference divides coefs_dofs (the rhs), not work: the solve result
     * is left in the workspace and coefs_dofs is only rescaled */
    for (int64_t i = 0; i < n; ++i) coefs[i] = coefs[i] / m->delta_x;
    return 0;
}

/* l2norm_squared (maxwell_1d_fem.jl:299-313). */
double or_l2norm_squared(or_maxwell1d *m, const double *coefs, int degree)
{
    if (degree == m->s_deg_0) or_solve_circulant(m, m->eig_mass0, coefs);
    else if (degree == m->s_deg_1) or_solve_circulant(m, m->eig_mass1, coefs);
    double r = 0.0;
    for (int64_t i = 0; i < m->n; ++i) r += coefs[i] * m->work[i];
    return r * m->delta_x;
}

/* inner_product (maxwell_1d_fem.jl:461-475). */
double or_inner_product(or_maxwell1d *m, const double *c1, const double *c2, int degree)
{
    if (degree == m->s_deg_0) or_solve_circulant(m, m->eig_mass0, c2);
    else if (degree == m->s_deg_1) or_solve_circulant(m, m->eig_mass1, c2);
    double r = 0.0;
    for (int64_t i = 0; i < m->n; ++i) r += c1[i] * m->work[i];
    return r * m->delta_x;
}

/* ------------------------------------------------------------------------ */
/* hamiltonian_splitting.jl, _1d2v.jl, _1d1v.jl                             */
/* ------------------------------------------------------------------------ */

struct or_split {
    int dim_v, n_weights, ld;
    int64_t np;
    double charge, mass, common_weight, q_over_m;
    double *a;               /* aliased particle array */
    double *e1, *e2, *b;     /* aliased field arrays */
    double *j1, *j2;         /* own j_dofs[1:2] */
    double *e1m, *e2m, *bm;  /* HamiltonianSplittingBoris: e_dofs_mid, b_dofs_mid */
    double adj_g_length, adj_g_wsum;   /* the last or_strang_adjoint's domain derivatives */
    or_maxwell1d *m;
    or_pmc1d ks0, ks1;
    double Lx, x_min, delta_x;
    int ntasks;              /* n_jobs = nthreads() (hamiltonian_splitting.jl:62) */
    double *tbuf;            /* [ntasks][nx]: each task's buffer = zero(j_dofs) */
};

/* HamiltonianSplitting{D,V} constructor (hamiltonian_splitting.jl:41-86). */
or_split *or_split_create(int64_t nx, double xmin, double xmax, int maxwell_degree, int degree0,
                          int smoothing0, int degree1, int smoothing1, int dim_v,
                          int64_t n_particles, int n_weights, double charge, double mass,
                          double common_weight, double *particles, double *e1, double *e2,
                          double *b)
{
    if (dim_v != 1 && dim_v != 2) return NULL;
    or_split *h = (or_split *)calloc(1, sizeof(*h));
    h->m = or_maxwell1d_create(xmin, xmax, nx, maxwell_degree);
    if (!h->m) { free(h); return NULL; }
    if (or_pmc1d_init(&h->ks0, xmin, xmax, nx, degree0, smoothing0) ||
        or_pmc1d_init(&h->ks1, xmin, xmax, nx, degree1, smoothing1)) {
        or_maxwell1d_destroy(h->m); free(h); return NULL;
    }
    h->dim_v = dim_v;
    h->n_weights = n_weights;
    h->ld = 1 + dim_v + n_weights;
    h->np = n_particles;
    /* ParticleGroup (particle_group.jl:25-45) */
    h->charge = charge;
    h->mass = mass;
    h->common_weight = common_weight == 0.0 ? 1.0 / (double)n_particles : common_weight;
    h->q_over_m = charge / mass;
    h->a = particles;
    h->e1 = e1; h->e2 = e2; h->b = b;
    h->j1 = (double *)calloc(nx, sizeof(double));
    h->j2 = (double *)calloc(nx, sizeof(double));
    h->e1m = (double *)calloc(nx, sizeof(double));
    h->e2m = (double *)calloc(nx, sizeof(double));
    h->bm = (double *)calloc(nx, sizeof(double));
    h->x_min = h->m->xmin;
    h->Lx = h->m->Lx;
    h->delta_x = h->Lx / (double)nx;
    h->ntasks = 1;
    return h;
}

/* The reference's task split (hamiltonian_splitting.jl:61-66): n_jobs =
 * nthreads(), @assert np % n_jobs == 0, chunks = partition(1:np, np / n_jobs).
 * With ntasks > 1 the 1d2v Hp1 and Hp2 deposit each chunk into its own
 * buffer on its own thread and sum the buffers in chunk order,
 * reduce(+, fetch.(tasks)) (hamiltonian_splitting_1d2v.jl:48-90,141-164;
 * Base's reduce over fewer than 1024 vectors is the left fold
 * ((b1 + b2) + b3) + ...), and HE's @threads loop runs on the threads:
 * the reference's own results with nthreads() = ntasks. */
int or_split_set_tasks(or_split *h, int ntasks)
{
    if (ntasks < 1 || h->np % ntasks != 0) return -1;
    free(h->tbuf);
    h->tbuf = NULL;
    h->ntasks = ntasks;
    if (ntasks > 1) {
        h->tbuf = (double *)calloc((size_t)ntasks * (size_t)h->m->n, sizeof(double));
        if (!h->tbuf) { h->ntasks = 1; return -1; }
    }
    return 0;
}

int or_split_set_particles(or_split *h, double *particles, int64_t n)
{
    if (h->ntasks > 1 && n % h->ntasks != 0) return -1;
    h->a = particles;
    h->np = n;
    return 0;
}

void or_split_destroy(or_split *h)
{
    if (!h) return;
    or_maxwell1d_destroy(h->m);
    free(h->j1); free(h->j2);
    free(h->e1m); free(h->e2m); free(h->bm);
    free(h->tbuf);
    free(h);
}

const double *or_split_j(const or_split *h, int component) { return component == 1 ? h->j1 : h->j2; }

#define PX(h, i) ((h)->a[0 + (i) * (h)->ld])
#define PV1(h, i) ((h)->a[1 + (i) * (h)->ld])
#define PV2(h, i) ((h)->a[2 + (i) * (h)->ld])
#define PW(h, i) ((h)->a[1 + (h)->dim_v + (i) * (h)->ld])

/* operatorHB: 1D2V / generic (hamiltonian_splitting_1d2v.jl:234-236) is a
 * pure field update; 1D1V (hamiltonian_splitting_1d1v.jl:113-125) is an
 * E-push without q/m. */
void or_operator_hb(or_split *h, double dt)
{
    if (h->dim_v == 2) {
        or_compute_e_from_b(h->e2, h->m, dt, h->b);
        return;
    }
    for (int64_t i = 0; i < h->np; ++i) {
        double xi = PX(h, i);
        double vi = PV1(h, i);
        double e1 = or_evaluate(&h->ks1, xi, h->e1);
        vi = vi + dt * e1;
        PV1(h, i) = vi;
    }
}

/* operatorHE 1D2V (hamiltonian_splitting_1d2v.jl:191-219). */
void or_operator_he(or_split *h, double dt)
{
    double qm = h->q_over_m;
#pragma omp parallel for num_threads(h->ntasks) schedule(static) if (h->ntasks > 1)
    for (int64_t i = 0; i < h->np; ++i) {
        double v1 = PV1(h, i), v2 = PV2(h, i);
        double xi = PX(h, i);
        double e1 = or_evaluate(&h->ks1, xi, h->e1);
        double e2 = or_evaluate(&h->ks0, xi, h->e2);
        PV1(h, i) = v1 + dt * qm * e1;
        PV2(h, i) = v2 + dt * qm * e2;
    }
    or_compute_b_from_e(h->b, h->m, dt, h->e2);
}

/* h->j = reduce(+, task buffers): the left fold in chunk order */
static void reduce_tasks(const or_split *h, double *j)
{
    int64_t nx = h->m->n;
    for (int64_t k = 0; k < nx; ++k) {
        double v = h->tbuf[k] + h->tbuf[nx + k];
        for (int t = 2; t < h->ntasks; ++t) v = v + h->tbuf[(int64_t)t * nx + k];
        j[k] = v;
    }
}

/* Hp2's particle loop over particles [i0, i1) into j (one task's body) */
static void hp2_chunk(or_split *h, double dt, int64_t i0, int64_t i1, double *j)
{
    double qm = h->q_over_m;
    for (int64_t i = i0; i < i1; ++i) {
        double x1 = PX(h, i), v1 = PV1(h, i), v2 = PV2(h, i);
        double b = or_evaluate(&h->ks1, x1, h->b);
        v1 = v1 + dt * qm * v2 * b;
        PV1(h, i) = v1;
        double w = PW(h, i);
        w = w * h->charge;
        w = w * h->common_weight;
        w = w * v2;
        or_add_charge(j, &h->ks0, x1, w);
    }
}

/* operatorHp2 1D2V (hamiltonian_splitting_1d2v.jl:129-176). */
void or_operator_hp2(or_split *h, double dt)
{
    int64_t nx = h->m->n;
    memset(h->j1, 0, (size_t)nx * sizeof(double));
    memset(h->j2, 0, (size_t)nx * sizeof(double));
    if (h->ntasks == 1) {
        hp2_chunk(h, dt, 0, h->np, h->j2);
    } else {
        int64_t c = h->np / h->ntasks;
        memset(h->tbuf, 0, (size_t)h->ntasks * (size_t)nx * sizeof(double));
#pragma omp parallel for num_threads(h->ntasks) schedule(static, 1)
        for (int t = 0; t < h->ntasks; ++t) hp2_chunk(h, dt, t * c, (t + 1) * c, h->tbuf + t * nx);
        reduce_tasks(h, h->j2);
    }
    for (int64_t k = 0; k < nx; ++k) h->j2[k] = h->j2[k] * dt;
    or_compute_e_from_j(h->e2, h->m, h->j2, 2);
}

/* Hp1's first task loop over particles [i0, i1) into j
 * (hamiltonian_splitting_1d2v.jl:48-86) */
static void hp1_current_chunk(or_split *h, double dt, int64_t i0, int64_t i1, double *j)
{
    double qm = h->q_over_m;
    for (int64_t i = i0; i < i1; ++i) {
        double x_old = PX(h, i), v1_old = PV1(h, i), v2_old = PV2(h, i);
        double x_new = x_old + dt * v1_old;
        double wi = PW(h, i);
        wi = wi * h->charge;
        wi = wi * h->common_weight;
        double v2_new = or_add_current_update_v(j, &h->ks1, x_old, x_new, wi, qm, h->b, v2_old);
        x_new = or_jl_mod_f64(x_new, h->Lx);
        PX(h, i) = x_new;
        PV2(h, i) = v2_new;
    }
}

/* Hp1's second task loop, the rho deposit (hamiltonian_splitting_1d2v.jl:90-106) */
static void hp1_rho_chunk(or_split *h, int64_t i0, int64_t i1, double *j)
{
    for (int64_t i = i0; i < i1; ++i) {
        double x = PX(h, i);
        double w = PW(h, i);
        w = w * h->charge;
        w = w * h->common_weight;
        or_add_charge(j, &h->ks0, x, w);
    }
}

/* operatorHp1: 1D2V (hamiltonian_splitting_1d2v.jl:41-112);
 * 1D1V (hamiltonian_splitting_1d1v.jl:63-98). */
void or_operator_hp1(or_split *h, double dt)
{
    int64_t nx = h->m->n;
    memset(h->j1, 0, (size_t)nx * sizeof(double));
    memset(h->j2, 0, (size_t)nx * sizeof(double));
    double qm = h->q_over_m;
    if (h->dim_v == 1) {
        for (int64_t i = 0; i < h->np; ++i) {
            double x_old = PX(h, i);
            double v_old = PV1(h, i);
            double x_new = x_old + dt * v_old;
            /* get_charge (particle_group.jl:67-69) */
            double wi = h->charge * PW(h, i) * h->common_weight;
            or_add_current_update_v_nob(h->j1, &h->ks1, x_old, x_new, wi, qm, v_old);
            x_new = or_jl_mod_f64(x_new, h->Lx);
            PX(h, i) = x_new;
        }
        or_compute_e_from_j(h->e1, h->m, h->j1, 1);
        return;
    }
    if (h->ntasks == 1) {
        hp1_current_chunk(h, dt, 0, h->np, h->j1);
        hp1_rho_chunk(h, 0, h->np, h->j2);
    } else {
        int64_t c = h->np / h->ntasks;
        memset(h->tbuf, 0, (size_t)h->ntasks * (size_t)nx * sizeof(double));
#pragma omp parallel for num_threads(h->ntasks) schedule(static, 1)
        for (int t = 0; t < h->ntasks; ++t)
            hp1_current_chunk(h, dt, t * c, (t + 1) * c, h->tbuf + t * nx);
        reduce_tasks(h, h->j1);
        memset(h->tbuf, 0, (size_t)h->ntasks * (size_t)nx * sizeof(double));
#pragma omp parallel for num_threads(h->ntasks) schedule(static, 1)
        for (int t = 0; t < h->ntasks; ++t) hp1_rho_chunk(h, t * c, (t + 1) * c, h->tbuf + t * nx);
        reduce_tasks(h, h->j2);
    }
    or_compute_e_from_j(h->e1, h->m, h->j1, 1);
}

/* strang_splitting!: generic (hamiltonian_splitting.jl:98-108) and 1D1V
 * (hamiltonian_splitting_1d1v.jl:11-21). */
void or_strang_splitting(or_split *h, double dt, int64_t nsteps)
{
    for (int64_t s = 0; s < nsteps; ++s) {
        if (h->dim_v == 1) {
            or_operator_hb(h, 0.5 * dt);
            or_operator_hp1(h, dt);
            or_operator_hb(h, 0.5 * dt);
        } else {
            or_operator_hb(h, 0.5 * dt);
            or_operator_he(h, 0.5 * dt);
            or_operator_hp2(h, 0.5 * dt);
            or_operator_hp1(h, 1.0 * dt);
            or_operator_hp2(h, 0.5 * dt);
            or_operator_he(h, 0.5 * dt);
            or_operator_hb(h, 0.5 * dt);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* diagnostics.jl                                                            */
/* ------------------------------------------------------------------------ */

/* solve_poisson! (diagnostics.jl:15-31): rho deposit with ks0, then Poisson. */
void or_solve_poisson(double *efield, or_split *h, double *rho)
{
    memset(rho, 0, (size_t)h->m->n * sizeof(double));
    for (int64_t i = 0; i < h->np; ++i) {
        double wi = h->charge * PW(h, i) * h->common_weight;
        or_add_charge(rho, &h->ks0, PX(h, i), wi);
    }
    or_compute_e_from_rho(efield, h->m, rho);
}

/* write_step! (diagnostics.jl:186-250) with pic_diagnostics_transfer (:45-63),
 * _vvb (:75-92) and _poynting (:106-112). The Poynting scratch is
 * `similar(bfield_dofs)` in the reference, i.e. uninitialised memory that
 * compute_e_from_b! adds into; it is taken as zero here. 1D2V only (the
 * reference reads array rows 2:4). */
void or_write_step(or_split *h, double time, int degree, const double *e1, const double *e2,
                   const double *b, const double *e1_n, const double *e2_n,
                   const double *efield_poisson, double *out)
{
    int64_t n = h->m->n;
    double d0 = 0.0, d1 = 0.0, d2 = 0.0;
    for (int64_t i = 0; i < h->np; ++i) {
        double v1 = PV1(h, i), v2 = PV2(h, i);
        double wi = PW(h, i);
        wi *= h->mass;
        wi *= h->common_weight;
        d0 += (v1 * v1 + v2 * v2) * wi;
        d1 += v1 * wi;
        d2 += v2 * wi;
    }
    double transfer = 0.0;
    for (int64_t i = 0; i < h->np; ++i) {
        double xi = PX(h, i);
        double wi = h->charge * PW(h, i) * h->common_weight;
        double v1 = PV1(h, i), v2 = PV2(h, i);
        double ef1 = or_evaluate(&h->ks1, xi, e1);
        double ef2 = or_evaluate(&h->ks0, xi, e2);
        transfer += (v1 * ef1 + v2 * ef2) * wi;
    }
    double vvb = 0.0;
    for (int64_t i = 0; i < h->np; ++i) {
        double xi = PX(h, i), v1 = PV1(h, i), v2 = PV2(h, i);
        double wi = PW(h, i);
        wi *= h->charge;
        wi *= h->common_weight;
        double bf = or_evaluate(&h->ks1, xi, b);
        vvb += wi * v1 * v2 * bf;
    }
    double *scratch = (double *)calloc(n, sizeof(double));
    or_compute_e_from_b(scratch, h->m, 1.0, b);
    double poynting = or_inner_product(h->m, e2, scratch, degree);
    free(scratch);
    double pe1 = or_inner_product(h->m, e1, e1_n, degree - 1);
    double pe2 = or_inner_product(h->m, e2, e2_n, degree);
    double pb3 = or_l2norm_squared(h->m, b, degree - 1);
    double err = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double d = fabs(e1[i] - efield_poisson[i]);
        if (d > err) err = d;
    }
    out[0] = time; out[1] = d0; out[2] = d1; out[3] = d2;
    out[4] = pe1; out[5] = pe2; out[6] = pb3;
    out[7] = transfer; out[8] = vvb; out[9] = poynting; out[10] = err;
}


/* ------------------------------------------------------------------------ */
/* HamiltonianSplittingBoris (hamiltonian_splitting_boris.jl), on the same   */
/* state as or_split (dim_v = 2): e_dofs and b_dofs aliased, e_dofs_mid and  */
/* b_dofs_mid owned (zero at construction, :50-60), j_dofs = j1, j2.         */
/* ------------------------------------------------------------------------ */

const double *or_boris_mid(const or_split *h, int which)
{
    return which == 0 ? h->e1m : which == 1 ? h->e2m : h->bm;
}

/* push_v_epart! (:189-204): v .+ dt * qm .* [E1_mid(x), E2_mid(x)] */
void or_boris_push_v_epart(or_split *h, double dt)
{
    const double qm = h->q_over_m;
    for (int64_t i = 0; i < h->np; ++i) {
        const double x = PX(h, i);
        const double e1 = or_evaluate(&h->ks1, x, h->e1m);
        const double e2 = or_evaluate(&h->ks0, x, h->e2m);
        const double dtqm = dt * qm;
        PV1(h, i) = PV1(h, i) + dtqm * e1;
        PV2(h, i) = PV2(h, i) + dtqm * e2;
    }
}

/* push_v_bpart! (:211-233): rotation by B_mid(x) */
void or_boris_push_v_bpart(or_split *h, double dt)
{
    const double qmdt = h->q_over_m * 0.5 * dt;
    for (int64_t i = 0; i < h->np; ++i) {
        const double v1 = PV1(h, i), v2 = PV2(h, i);
        double b = or_evaluate(&h->ks1, PX(h, i), h->bm);
        b = qmdt * b;
        double m11 = 1.0 / (1.0 + b * b);
        const double m12 = m11 * b * 2.0;
        m11 = m11 * (1.0 - b * b);
        PV1(h, i) = m11 * v1 + m12 * v2;
        PV2(h, i) = -m12 * v1 + m11 * v2;
    }
}

/* push_x_accumulate_j! (:250-288): j_dofs = midpoint deposits of q v1 (ks1)
 * and q v2 (ks0); x = mod(x + dt v1, Lx) */
void or_boris_push_x_accumulate_j(or_split *h, double dt)
{
    const int64_t nx = h->ks0.n_grid;
    memset(h->j1, 0, sizeof(double) * nx);
    memset(h->j2, 0, sizeof(double) * nx);
    for (int64_t i = 0; i < h->np; ++i) {
        const double x_old = PX(h, i), v1 = PV1(h, i), v2 = PV2(h, i);
        const double x_new = x_old + dt * v1;
        const double wi = h->charge * PW(h, i) * h->common_weight;   /* get_charge */
        const double xm = (x_old + x_new) * 0.5;
        or_add_charge(h->j1, &h->ks1, xm, wi * v1);
        or_add_charge(h->j2, &h->ks0, xm, wi * v2);
        PX(h, i) = or_jl_mod_f64(x_new, h->Lx);
    }
}

/* staggering! (:99-122) */
void or_boris_staggering(or_split *h, double dt)
{
    const int64_t nx = h->ks0.n_grid;
    or_boris_push_x_accumulate_j(h, dt * 0.5);
    for (int64_t k = 0; k < nx; ++k) {
        h->e1m[k] = h->e1[k];
        h->e2m[k] = h->e2[k];
        h->j1[k] = 0.5 * dt * h->j1[k];
        h->j2[k] = 0.5 * dt * h->j2[k];
    }
    or_compute_e_from_j(h->e1m, h->m, h->j1, 1);
    or_compute_e_from_b(h->e2m, h->m, 0.5 * dt, h->b);
    or_compute_e_from_j(h->e2m, h->m, h->j2, 2);
}

/* strang_splitting! (:132-177) */
void or_boris_strang(or_split *h, double dt, int64_t number_steps)
{
    const int64_t nx = h->ks0.n_grid;
    for (int64_t step = 0; step < number_steps; ++step) {
        for (int64_t k = 0; k < nx; ++k) h->bm[k] = h->b[k];
        or_compute_b_from_e(h->b, h->m, dt, h->e2m);
        for (int64_t k = 0; k < nx; ++k) h->bm[k] = (h->bm[k] + h->b[k]) * 0.5;
        or_boris_push_v_epart(h, 0.5 * dt);
        or_boris_push_v_bpart(h, dt);
        or_boris_push_v_epart(h, 0.5 * dt);
        or_boris_push_x_accumulate_j(h, dt);
        for (int64_t k = 0; k < nx; ++k) {
            h->e1[k] = h->e1m[k];
            h->e2[k] = h->e2m[k];
            h->j1[k] = dt * h->j1[k];
            h->j2[k] = dt * h->j2[k];
        }
        or_compute_e_from_j(h->e1m, h->m, h->j1, 1);
        or_compute_e_from_b(h->e2m, h->m, dt, h->b);
        or_compute_e_from_j(h->e2m, h->m, h->j2, 2);
    }
}

void or_boris_set_mid(or_split *h, const double *e1m, const double *e2m, const double *bm)
{
    const int64_t nx = h->ks0.n_grid;
    for (int64_t k = 0; k < nx; ++k) {
        if (e1m) h->e1m[k] = e1m[k];
        if (e2m) h->e2m[k] = e2m[k];
        if (bm) h->bm[k] = bm[k];
    }
}


/* ------------------------------------------------------------------------ */
/* sample!(pg, alpha, k, sigma, mesh) for ParticleGroup{1,1} and {1,2}       */
/* (particle_sampling.jl:236-311) with Sobol.jl's SobolSeq(2).               */
/* ------------------------------------------------------------------------ */

/* Sobol.jl (Project.toml compat "1"; not vendored in the reference) follows
 * the Bratley-Fox / Joe-Kuo construction in Gray-code order and skips the
 * all-zero point: the k-th next!() (k = 1, 2, ...) is the XOR of the
 * direction numbers V_j over the set bits j of gray(k) = k ^ (k >> 1).
 * Dimension 1: V_j = 2^-j (van der Corput); dimension 2: primitive
 * polynomial x + 1, m_1 = 1, m_j = 2 m_{j-1} XOR m_{j-1}, V_j = m_j 2^-j.
 * Values are k/2^32 dyadic rationals, exact in double. */
void or_sobol2(uint64_t k, double *r1, double *r2)
{
    uint64_t g = k ^ (k >> 1);
    uint32_t x1 = 0, x2 = 0, m = 1;
    for (int j = 1; j <= 32; ++j) {
        if (j > 1) m = (m << 1) ^ m;
        if ((g >> (j - 1)) & 1u) {
            x1 ^= 1u << (32 - j);
            x2 ^= m << (32 - j);
        }
    }
    *r1 = ldexp((double)x1, -32);
    *r2 = ldexp((double)x2, -32);
}

/* newton(r, alpha, k) (particle_sampling.jl:236-245): the inverse of the
 * cdf of 1 + alpha cos(k x) on [0, 2 pi / k], in the reference's iteration
 * (two interleaved sequences, stopped when they agree to 1e-12). Returns the
 * number of iterations (limit 200; the reference has none). */
int or_landau_newton(double r, double alpha, double k, double *out)
{
    double x0 = 0.0, x1 = 1.0;
    r *= 6.283185307179586 / k;
    int it = 0;
    while (fabs(x1 - x0) > 1e-12 && it < 200) {
        const double p = x0 + alpha * sin(k * x0) / k;
        const double f = 1.0 + alpha * cos(k * x0);
        const double nx1 = x0 - (p - r) / f;
        x0 = x1;
        x1 = nx1;
        ++it;
    }
    *out = x1;
    return it;
}

/* sample! (:266-311): x from the Sobol point's second coordinate by newton,
 * |v| = sigma sqrt(-2 log((i - 1/2) / N)), angle 2 pi r1; weight = dimx.
 * aos: (1 + dim_v + 1) x n column-major with leading dimension ld. Returns
 * the largest newton iteration count. */
int or_sample_landau(int dim_v, int64_t n, double alpha, double k, double sigma, double dimx,
                     double *aos, int64_t ld)
{
    int itmax = 0;
    for (int64_t i = 1; i <= n; ++i) {
        const double v = sigma * sqrt(-2.0 * log(((double)i - 0.5) / (double)n));
        double r1, r2, x;
        or_sobol2((uint64_t)i, &r1, &r2);
        const double th = r1 * 6.283185307179586;
        const int it = or_landau_newton(r2, alpha, k, &x);
        if (it > itmax) itmax = it;
        double *col = aos + (i - 1) * ld;
        col[0] = x;
        col[1] = v * cos(th);
        if (dim_v == 2) col[2] = v * sin(th);
        col[1 + dim_v] = dimx;
    }
    return itmax;
}

/* ------------------------------------------------------------------------ */
/* Reverse mode (adjoint) of the 1D2V Strang splitting                      */
/* ------------------------------------------------------------------------ */
/*
 * The reference has no derivatives. This is the transpose of the Jacobian of
 * each sub-step of strang_splitting! (hamiltonian_splitting.jl:98-108) as the
 * functions above compute it, applied in reverse order, sub-step by
 * sub-step. It is pinned by central finite differences of the forward
 * functions above (tests/test_oracle_adjoint.py): directional derivatives of
 * the loss along random perturbations of every input, per sub-step and over
 * whole runs.
 *
 * Loss: wloss * sum over the steps s = 1..T of 1/2 l2norm_squared(B_s, p-1)
 * (maxwell_1d_fem.jl:299-313), B_s = b_dofs after step s; with wloss = dt
 * this is the time-integrated magnetic energy of write_step!
 * (diagnostics.jl:236), plus <ct, final state> for given cotangents ct.
 *
 * Derivative of a basis of degree p at offset t: d/dt B_p[r] =
 * B_{p-1}[r-1] - B_{p-1}[r] (entries outside 0..p-1 are zero); positions
 * enter through t = (x - xmin)/dx - cell, so d/dx = (1/dx) d/dt.
 */

void or_bspline_deriv(int degree, double offset, double *d)
{
    double lo[OR_MAXDEG + 1];
    if (degree == 0) { d[0] = 0.0; return; }
    or_bspline_basis(degree - 1, offset, lo);
    for (int r = 0; r <= degree; ++r) {
        double a = r >= 1 ? lo[r - 1] : 0.0;
        double b = r <= degree - 1 ? lo[r] : 0.0;
        d[r] = a - b;
    }
}

/* value and x-derivative of evaluate(p, x, field) at the trunc cell */
static void eval_d(const or_pmc1d *p, double x, const double *field, double *val, double *dval)
{
    double sv[OR_MAXDEG + 1], dv[OR_MAXDEG + 1];
    double xi = (x - p->xmin) / p->delta_x;
    int64_t index = (int64_t)xi;
    double t = xi - (double)index;
    index = index - p->degree;
    or_bspline_basis(p->degree, t, sv);
    or_bspline_deriv(p->degree, t, dv);
    double v = 0.0, d = 0.0;
    for (int i = 1; i <= p->n_span; ++i) {
        int64_t k = or_jl_mod_i64(index + i - 1, p->n_grid);
        v += field[k] * sv[i - 1];
        d += field[k] * dv[i - 1];
    }
    *val = v;
    *dval = d / p->delta_x;
}

/* out[k] += wgt * N_k(x) at the trunc cell (the transpose of evaluate) */
static void scatter(const or_pmc1d *p, double x, double wgt, double *out)
{
    double sv[OR_MAXDEG + 1];
    double xi = (x - p->xmin) / p->delta_x;
    int64_t index = (int64_t)xi;
    double t = xi - (double)index;
    index = index - p->degree;
    or_bspline_basis(p->degree, t, sv);
    for (int i = 1; i <= p->n_span; ++i)
        out[or_jl_mod_i64(index + i - 1, p->n_grid)] += wgt * sv[i - 1];
}

/* y = C^T x for the circulant C of solve_circulant(eig, .): conjugate eigenvalues */
static void solve_circulant_t(or_maxwell1d *m, const double *eig, const double *x, double *y)
{
    int64_t n = m->n;
    double *ec = (double *)malloc((size_t)n * sizeof(double));
    for (int64_t k = 0; k < n; ++k) ec[k] = (k > n / 2) ? -eig[k] : eig[k];
    or_solve_circulant(m, ec, x);
    memcpy(y, m->work, (size_t)n * sizeof(double));
    free(ec);
}

/* lambda_e += coef * D^T lambda_b, (D e)[i] = e[i-1] - e[i] (compute_b_from_e!) */
static void adj_b_from_e(double *le, double coef, const double *lb, int64_t n)
{
    for (int64_t j = 0; j < n; ++j) {
        int64_t jp = j + 1 == n ? 0 : j + 1;
        le[j] += coef * (lb[jp] - lb[j]);
    }
}

typedef struct {
    double *lx, *lv1, *lv2;      /* particle cotangents, one per particle */
    double *le1, *le2, *lb;      /* field cotangents */
    /* explicit derivatives of the loss, state held fixed: in dx (through
     * xi = (x - xmin)/dx and the dt/dx, 1/dx factors), in Lx through the mod
     * of Hp1, and the sum over particles of those in the weights */
    double g_dx, g_mod, g_w;
} or_adj_state;

/* (x - xmin) / dx of a position */
static double xi_of(const or_split *h, double x) { return (x - h->x_min) / h->delta_x; }

/* operatorHB 1d2v: E2 += dt/dx W(B) */
static void adj_hb(or_split *h, double dt, const double *b, or_adj_state *A)
{
    int64_t n = h->m->n;
    double *y = (double *)malloc((size_t)n * sizeof(double));
    solve_circulant_t(h->m, h->m->eig_weak_ampere, A->le2, y);
    double coef = dt / h->m->delta_x;
    /* d/d dx of E2 += (dt/dx) W B: -(1/dx) lambda_E2 . (dt/dx) W B = -(1/dx) coef (W^T lambda_E2) . B */
    double r = 0.0;
    for (int64_t i = 0; i < n; ++i) r += y[i] * b[i];
    A->g_dx += -(coef / h->m->delta_x) * r;
    for (int64_t i = 0; i < n; ++i) A->lb[i] += coef * y[i];
    free(y);
}

/* operatorHE 1d2v; S = particle array and fields before the sub-step */
static void adj_he(or_split *h, double dt, const double *S, const double *e1, const double *e2,
                   or_adj_state *A)
{
    double qm = h->q_over_m;
    int ld = h->ld;
    int64_t n = h->m->n;
    {
        /* d/d dx of B += (dt/dx) D E2: -(1/dx) lambda_B . (dt/dx) D E2 */
        double coef = dt / h->m->delta_x, r = 0.0;
        for (int64_t j = 0; j < n; ++j) r += A->lb[j] * (e2[j == 0 ? n - 1 : j - 1] - e2[j]);
        A->g_dx += -(coef / h->m->delta_x) * r;
    }
    adj_b_from_e(A->le2, dt / h->m->delta_x, A->lb, h->m->n);
    for (int64_t i = 0; i < h->np; ++i) {
        double x = S[i * ld];
        double f1, d1, f2, d2;
        eval_d(&h->ks1, x, e1, &f1, &d1);
        eval_d(&h->ks0, x, e2, &f2, &d2);
        double dlx = dt * qm * (A->lv1[i] * d1 + A->lv2[i] * d2);
        A->g_dx += -xi_of(h, x) * dlx;          /* d f(x)/d dx = -xi f'(x) */
        A->lx[i] += dlx;
        scatter(&h->ks1, x, dt * qm * A->lv1[i], A->le1);
        scatter(&h->ks0, x, dt * qm * A->lv2[i], A->le2);
    }
}

/* operatorHp2 1d2v */
static void adj_hp2(or_split *h, double dt, const double *S, const double *b, or_adj_state *A)
{
    or_maxwell1d *m = h->m;
    int64_t n = m->n;
    int ld = h->ld;
    double qm = h->q_over_m;
    double *mu = (double *)calloc((size_t)n, sizeof(double));
    double *eig = (double *)calloc((size_t)n, sizeof(double));
    for (int64_t i = 0; i <= n / 2; ++i) eig[i] = 1.0 / m->eig_mass0[i];
    /* E2 -= C j2 / dx with j2 = dt * sum q v2 S0(x): dj = -(dt/dx) C^T lambda_E2 */
    solve_circulant_t(m, eig, A->le2, mu);
    for (int64_t k = 0; k < n; ++k) mu[k] = -(dt / m->delta_x) * mu[k];
    for (int64_t i = 0; i < h->np; ++i) {
        double x = S[i * ld], v2 = S[2 + i * ld], w = S[1 + h->dim_v + i * ld];
        double q = w * h->charge * h->common_weight;
        double bv, db, g, dg;
        eval_d(&h->ks1, x, b, &bv, &db);
        eval_d(&h->ks0, x, mu, &g, &dg);
        double lv1 = A->lv1[i];
        A->lv2[i] += dt * qm * bv * lv1 + q * h->ks0.scaling * g;
        double dlx = dt * qm * v2 * lv1 * db + q * h->ks0.scaling * v2 * dg;
        A->lx[i] += dlx;
        /* mu carries the solve's 1/dx: d/d dx of -(dt/dx) M0^-1 J adds -(1/dx) mu . J */
        A->g_dx += -xi_of(h, x) * dlx - q * h->ks0.scaling * v2 * g / m->delta_x;
        A->g_w += h->charge * h->common_weight * h->ks0.scaling * v2 * g;
        scatter(&h->ks1, x, dt * qm * v2 * lv1, A->lb);
    }
    free(mu);
    free(eig);
}

/* operatorHp1 1d2v: x_new = x + dt v1, v2 -= qm int_x^x_new B, j1 += q int S1,
 * x = mod(x_new, Lx), E1 -= C1 j1 / dx */
static void adj_hp1(or_split *h, double dt, const double *S, const double *b, or_adj_state *A)
{
    or_maxwell1d *m = h->m;
    int64_t n = m->n;
    int ld = h->ld;
    double qm = h->q_over_m;
    double *mu = (double *)calloc((size_t)n, sizeof(double));
    double *eig = (double *)calloc((size_t)n, sizeof(double));
    for (int64_t i = 0; i <= n / 2; ++i) eig[i] = 1.0 / m->eig_mass1[i];
    solve_circulant_t(m, eig, A->le1, mu);
    for (int64_t k = 0; k < n; ++k) mu[k] = -mu[k] / m->delta_x;
    or_pmc1d p1 = h->ks1;
    p1.scaling = 1.0;
    for (int64_t i = 0; i < h->np; ++i) {
        double x_old = S[i * ld], v1 = S[1 + i * ld], w = S[1 + h->dim_v + i * ld];
        double q = w * h->charge * h->common_weight;
        double x_new = x_old + dt * v1;
        double bn, dbn, bo, dbo, gn, dgn, go, dgo;
        /* the end points' cells are trunc() cells, as in add_current_update_v! */
        eval_d(&h->ks1, x_new, b, &bn, &dbn);
        eval_d(&h->ks1, x_old, b, &bo, &dbo);
        eval_d(&h->ks1, x_new, mu, &gn, &dgn);
        eval_d(&h->ks1, x_old, mu, &go, &dgo);
        double lv2 = A->lv2[i];
        double lx1 = A->lx[i];
        double lxn = A->lx[i] - qm * lv2 * bn + q * h->ks1.scaling * gn;
        A->lx[i] = lxn + qm * lv2 * bo - q * h->ks1.scaling * go;
        A->lv1[i] += dt * lxn;
        {
            /* int_x^xn f for f = B and f = mu1 (the same trunc-cell pieces):
             * d/d dx = (1/dx) int f - xi_n f(xn) + xi f(x), by parts */
            double *tmp = (double *)calloc((size_t)n, sizeof(double));
            double ib = -or_add_current_update_v(tmp, &p1, x_old, x_new, 0.0, 1.0, b, 0.0);
            double im = -or_add_current_update_v(tmp, &p1, x_old, x_new, 0.0, 1.0, mu, 0.0);
            free(tmp);
            double xo = xi_of(h, x_old), xn = xi_of(h, x_new);
            A->g_dx += -qm * lv2 * (ib / m->delta_x - xn * bn + xo * bo) +
                       q * h->ks1.scaling * (xo * go - xn * gn);
            A->g_w += h->charge * h->common_weight * h->ks1.scaling * im;
            double xm = or_jl_mod_f64(x_new, h->Lx);
            if (xm != x_new) A->g_mod += lx1 * (xm > x_new ? 1.0 : -1.0);
        }
        /* lambda_B[k] += -qm lv2 int S1_k: the same line integral, unscaled */
        or_add_current_update_v(A->lb, &p1, x_old, x_new, -qm * lv2, qm, NULL, 0.0);
    }
    free(mu);
    free(eig);
}

/* loss term of B: wloss * 1/2 l2norm_squared(B) and its gradient wloss dx M1 B */
static double loss_b(or_split *h, double wloss, const double *b, double *lb)
{
    or_maxwell1d *m = h->m;
    double e = or_l2norm_squared(m, b, m->s_deg_1);   /* leaves M1 B in work */
    if (lb)
        for (int64_t i = 0; i < m->n; ++i) lb[i] += wloss * m->delta_x * m->work[i];
    return wloss * 0.5 * e;
}

double or_strang_loss(or_split *h, double dt, int64_t nsteps, double wloss)
{
    double loss = 0.0;
    for (int64_t s = 0; s < nsteps; ++s) {
        or_strang_splitting(h, dt, 1);
        loss += loss_b(h, wloss, h->b, NULL);
    }
    return loss;
}

int or_strang_adjoint(or_split *h, double dt, int64_t nsteps, double wloss, const double *ct_part,
                      const double *ct_e1, const double *ct_e2, const double *ct_b, double *loss,
                      double *g_part, double *g_e1, double *g_e2, double *g_b)
{
    if (h->dim_v != 2 || nsteps < 0) return -1;
    const int64_t n = h->m->n, np = h->np;
    const int ld = h->ld;
    const size_t psz = (size_t)ld * (size_t)np, fsz = 3 * (size_t)n;
    const int nsub = 7;
    /* state before every sub-step of every step: particles, then E1, E2, B */
    double *hist = (double *)malloc(((size_t)nsteps * nsub) * (psz + fsz) * sizeof(double) + 8);
    double *bend = (double *)malloc((size_t)nsteps * n * sizeof(double) + 8);
    if (!hist || !bend) { free(hist); free(bend); return -2; }
    const double hdt = 0.5 * dt;
    double L = 0.0;
    for (int64_t s = 0; s < nsteps; ++s) {
        for (int k = 0; k < nsub; ++k) {
            double *rec = hist + ((size_t)s * nsub + k) * (psz + fsz);
            memcpy(rec, h->a, psz * sizeof(double));
            memcpy(rec + psz, h->e1, n * sizeof(double));
            memcpy(rec + psz + n, h->e2, n * sizeof(double));
            memcpy(rec + psz + 2 * n, h->b, n * sizeof(double));
            switch (k) {
            case 0: or_operator_hb(h, hdt); break;
            case 1: or_operator_he(h, hdt); break;
            case 2: or_operator_hp2(h, hdt); break;
            case 3: or_operator_hp1(h, dt); break;
            case 4: or_operator_hp2(h, hdt); break;
            case 5: or_operator_he(h, hdt); break;
            default: or_operator_hb(h, hdt); break;
            }
        }
        memcpy(bend + (size_t)s * n, h->b, n * sizeof(double));
        L += loss_b(h, wloss, h->b, NULL);
    }
    or_adj_state A;
    A.lx = (double *)calloc((size_t)np, sizeof(double));
    A.lv1 = (double *)calloc((size_t)np, sizeof(double));
    A.lv2 = (double *)calloc((size_t)np, sizeof(double));
    A.le1 = (double *)calloc((size_t)n, sizeof(double));
    A.le2 = (double *)calloc((size_t)n, sizeof(double));
    A.lb = (double *)calloc((size_t)n, sizeof(double));
    A.g_dx = 0.0;
    A.g_mod = 0.0;
    A.g_w = 0.0;
    for (int64_t i = 0; ct_part && i < np; ++i) {
        A.lx[i] = ct_part[i * ld];
        A.lv1[i] = ct_part[1 + i * ld];
        A.lv2[i] = ct_part[2 + i * ld];
    }
    for (int64_t k = 0; k < n; ++k) {
        if (ct_e1) A.le1[k] = ct_e1[k];
        if (ct_e2) A.le2[k] = ct_e2[k];
        if (ct_b) A.lb[k] = ct_b[k];
    }
    for (int64_t s = nsteps - 1; s >= 0; --s) {
        A.g_dx += loss_b(h, wloss, bend + (size_t)s * n, A.lb) / h->m->delta_x;
        for (int k = nsub - 1; k >= 0; --k) {
            const double *rec = hist + ((size_t)s * nsub + k) * (psz + fsz);
            const double *e1 = rec + psz, *e2 = rec + psz + n, *b = rec + psz + 2 * n;
            switch (k) {
            case 0: case 6: adj_hb(h, hdt, b, &A); break;
            case 1: case 5: adj_he(h, hdt, rec, e1, e2, &A); break;
            case 2: case 4: adj_hp2(h, hdt, rec, b, &A); break;
            default: adj_hp1(h, dt, rec, b, &A); break;
            }
        }
    }
    if (loss) *loss = L;
    /* dx = Lx / nx: d/d Lx = (1/nx) d/d dx, plus the periods of the mod */
    h->adj_g_length = A.g_dx / (double)n + A.g_mod;
    h->adj_g_wsum = A.g_w;
    for (int64_t i = 0; g_part && i < np; ++i) {
        for (int r = 0; r < ld; ++r) g_part[r + i * ld] = 0.0;
        g_part[i * ld] = A.lx[i];
        g_part[1 + i * ld] = A.lv1[i];
        g_part[2 + i * ld] = A.lv2[i];
    }
    if (g_e1) memcpy(g_e1, A.le1, n * sizeof(double));
    if (g_e2) memcpy(g_e2, A.le2, n * sizeof(double));
    if (g_b) memcpy(g_b, A.lb, n * sizeof(double));
    free(A.lx); free(A.lv1); free(A.lv2); free(A.le1); free(A.le2); free(A.lb);
    free(hist);
    free(bend);
    return 0;
}

/* Reverse mode of solve_poisson! (diagnostics.jl:15-31) with respect to the
 * positions: E = D C_pois rho(x), rho = sum q S0(x) * scaling (add_charge!),
 * (D phi)[i] = phi[i-1] - phi[i] (compute_e_from_rho!, maxwell_1d_fem.jl:244-255).
 * g_x[i] = d <ct_e, E> / d x_i. */
void or_solve_poisson_adjoint(or_split *h, const double *ct_e, double *g_x)
{
    or_maxwell1d *m = h->m;
    int64_t n = m->n;
    double *dt_ = (double *)calloc((size_t)n, sizeof(double));
    double *nu = (double *)calloc((size_t)n, sizeof(double));
    adj_b_from_e(dt_, 1.0, ct_e, n);
    solve_circulant_t(m, m->eig_weak_poisson, dt_, nu);
    for (int64_t i = 0; i < h->np; ++i) {
        double q = h->charge * PW(h, i) * h->common_weight;
        double g, dg;
        eval_d(&h->ks0, PX(h, i), nu, &g, &dg);
        g_x[i] = q * h->ks0.scaling * dg;
    }
    free(dt_);
    free(nu);
}

/* The explicit derivatives of the last or_strang_adjoint's loss in the domain
 * length (state held fixed; through dx = Lx/nx and the mod) and the sum of
 * its derivatives in the particle weights. */
void or_strang_adjoint_domain(const or_split *h, double *g_length, double *g_wsum)
{
    if (g_length) *g_length = h->adj_g_length;
    if (g_wsum) *g_wsum = h->adj_g_wsum;
}
