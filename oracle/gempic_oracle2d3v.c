/*
 * gempic_oracle2d3v.c -- CPU oracle of the 2D3V Hamiltonian splitting.
 *
 * TEST INFRASTRUCTURE (see gempic_oracle.h): the checker, never the product.
 *
 * The reference includes hamiltonian_splitting_2d3v.jl (src/GEMPIC.jl:33)
 * but the file is empty: there is no 2D3V splitting to restate. This file
 * writes the splitting of the GEMPIC paper (Kraus, Kormann, Morrison and
 * Sonnendruecker, J. Plasma Phys. 83 (2017), the Hamiltonian splitting
 * H = Hp1 + Hp2 + Hp3 + HE + HB) on the reference's own 2D pieces, in the
 * form the reference's 1D2V splitting takes (hamiltonian_splitting_1d2v.jl):
 *
 *   fields (TwoDMaxwell, maxwell_2d_fem.jl): E1 in splines of degree
 *   (p-1, p), E2 (p, p-1), E3 (p, p); B1 (p, p-1), B2 (p-1, p), B3 (p-1, p-1)
 *   (the mass solvers inv_mass_1 / inv_mass_2 of maxwell_2d_fem.jl:56-66);
 *   particles ParticleGroup{2,3}: rows x1, x2, v1, v2, v3, w;
 *   q = charge * w * common_weight, qm = charge / mass.
 *
 *   operatorHp1(dt): x1 += dt v1 along the path at fixed x2, with the exact
 *     line integral of the degree p-1 splines in x1 (split at cell
 *     boundaries, Gauss-Legendre with (p+1)/2 points per piece, as update_jv!
 *     does in 1D, particle_mesh_coupling_1d.jl:385-414):
 *       j1  += q  (dx int S^{p-1}) x S^p(x2)
 *       v2  -= qm sum B3 (dx int S^{p-1}) x S^{p-1}(x2)
 *       v3  += qm sum B2 (dx int S^{p-1}) x S^p(x2)
 *     then x1 = xmin + mod(x1 - xmin, Lx) and E1 -= M1^{-1} j1
 *     (compute_e_from_j!, maxwell_2d_fem.jl:455-459);
 *   operatorHp2(dt): the same along x2 at fixed x1 with j2 (p, p-1),
 *       v1  += qm sum B3 S^{p-1}(x1) x (dy int S^{p-1}),
 *       v3  -= qm sum B1 S^p(x1) x (dy int S^{p-1}), E2 -= M1^{-1} j2;
 *   operatorHp3(dt): no motion in the plane;
 *       v1 -= dt qm v3 B2(x), v2 += dt qm v3 B1(x), j3 = dt sum q v3 S^p S^p,
 *       E3 -= M1^{-1} j3;
 *   operatorHE(dt): v += dt qm E(x), then compute_b_from_e!(B, dt, E);
 *   operatorHB(dt): compute_e_from_b!(E, dt, B);
 *   strang_splitting!: HB HE Hp3 Hp2 Hp1 Hp2 Hp3 HE HB with half steps
 *     around the full Hp1 step, as hamiltonian_splitting.jl:98-108 nests
 *     its 1D2V operators.
 *
 * Point values use the 2D coupling's shape factor (ceil() cell index,
 * particle_mesh_coupling_2d.jl:54-65) with a degree per direction; the line
 * integrals cut the path at integer cell coordinates (floor()).
 *
 * What pins it (tests/test_oracle_2d3v.py): the discrete Gauss law (the
 * deposited charge minus compute_rho_from_e! stays constant to round-off),
 * the second-order energy error of the Strang splitting, each sub-flow's
 * exact invariants (|v| in the plane of the magnetic rotation, the
 * line-integral current equal to the change of the charge deposit), and the
 * reduction to the reference's 1D2V splitting for data that do not depend
 * on x2.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "gempic_oracle.h"

#define OR2_ROWS 6              /* x1, x2, v1, v2, v3, w */
#define OR2_MAXSEG 4096         /* pieces of one line integral */

struct or_split2d {
    int64_t nx, ny, n;
    double xmin, ymin, Lx, Ly, dx, dy;
    int p;
    double charge, mass, cw, qm;
    double *part;                               /* aliased 6 x n, column-major */
    double *e[3], *b[3];                        /* aliased, nx*ny each */
    double *j[3];                               /* owned */
    or_maxwell2d *m;
    double gx[OR_MAXDEG + 1], gw[OR_MAXDEG + 1];     /* Gauss-Legendre on [-1, 1], (p+1)/2 points */
    int nq;
    int64_t *seg_cell;
    double *seg_val;                            /* OR2_MAXSEG x (OR_MAXDEG + 1) */
};

static int64_t imod2(int64_t a, int64_t n)
{
    int64_t r = a % n;
    return r < 0 ? r + n : r;
}

static int64_t node(const or_split2d *h, int64_t i, int64_t j)
{
    return imod2(i, h->nx) + imod2(j, h->ny) * h->nx;
}

/* compute_shape_factor (particle_mesh_coupling_2d.jl:54-65) in one
 * direction with degree `deg`: first node of the deg+1 touched, values */
static int64_t point_basis(double pos, double lo, double d, int deg, double *v)
{
    const double xp = (pos - lo) / d;
    const int64_t ip = (int64_t)ceil(xp);
    or_bspline_basis(deg, xp - (double)(ip - 1), v);
    return ip - 1 - deg;
}

/* The pieces of the path [xa, xb] (cell coordinates) cut at integer
 * coordinates, in order from xa to xb: cell c and
 * val[k] = sign * d * int_lo^up S^deg_k(t) dt over the piece's offsets
 * [lo, up] within the cell (first node c - deg). Returns the count. */
static int line_pieces(or_split2d *h, int deg, double xa, double xb, double d)
{
    const double sign = xb >= xa ? 1.0 : -1.0;
    const double lo_x = xb >= xa ? xa : xb, hi_x = xb >= xa ? xb : xa;
    int64_t c0 = (int64_t)floor(lo_x), c1 = (int64_t)floor(hi_x);
    int ns = 0;
    const int nseg = (int)(c1 - c0 + 1);
    for (int s = 0; s < nseg; ++s) {
        /* pieces in path order: forward from c0 up, backward from c1 down */
        const int64_t c = sign > 0 ? c0 + s : c1 - s;
        double lo = lo_x - (double)c, up = hi_x - (double)c;
        if (lo < 0.0) lo = 0.0;
        if (up > 1.0) up = 1.0;
        if (!(up > lo)) continue;
        if (ns >= OR2_MAXSEG) abort();
        double *val = h->seg_val + (size_t)ns * (OR_MAXDEG + 1);
        const double cc1 = 0.5 * (up - lo), cc2 = 0.5 * (up + lo);
        double sv[OR_MAXDEG + 1];
        for (int k = 0; k <= deg; ++k) val[k] = 0.0;
        for (int q = 0; q < h->nq; ++q) {
            or_bspline_basis(deg, cc1 * h->gx[q] + cc2, sv);
            for (int k = 0; k <= deg; ++k) val[k] = val[k] + sv[k] * h->gw[q] * cc1;
        }
        for (int k = 0; k <= deg; ++k) val[k] = val[k] * (sign * d);
        h->seg_cell[ns] = c;
        ++ns;
    }
    return ns;
}

or_split2d *or_split2d_create(int64_t nx, int64_t ny, double xmin, double xmax, double ymin,
                              double ymax, int degree, int64_t n, double charge, double mass,
                              double common_weight, double *particles, double *const *e,
                              double *const *b)
{
    if (degree < 1 || degree > 3 || nx < 2 * degree + 1 || ny < 2 * degree + 1 || n < 0)
        return NULL;
    or_split2d *h = (or_split2d *)calloc(1, sizeof(*h));
    h->nx = nx;
    h->ny = ny;
    h->n = n;
    h->xmin = xmin;
    h->ymin = ymin;
    h->Lx = xmax - xmin;
    h->Ly = ymax - ymin;
    h->dx = h->Lx / (double)nx;
    h->dy = h->Ly / (double)ny;
    h->p = degree;
    h->charge = charge;
    h->mass = mass;
    h->cw = common_weight == 0.0 ? 1.0 / (double)n : common_weight;
    h->qm = charge / mass;
    h->part = particles;
    for (int c = 0; c < 3; ++c) {
        h->e[c] = e[c];
        h->b[c] = b[c];
        h->j[c] = (double *)calloc((size_t)(nx * ny), sizeof(double));
    }
    h->m = or_maxwell2d_create(xmin, xmax, nx, ymin, ymax, ny, degree);
    h->nq = (degree + 1) / 2;        /* (p-1+2) / 2 points, as n_quad_points of degree p-1 */
    or_gausslegendre(h->nq, h->gx, h->gw);
    h->seg_cell = (int64_t *)malloc(sizeof(int64_t) * OR2_MAXSEG);
    h->seg_val = (double *)malloc(sizeof(double) * OR2_MAXSEG * (OR_MAXDEG + 1));
    if (!h->m) {
        or_split2d_destroy(h);
        return NULL;
    }
    return h;
}

void or_split2d_destroy(or_split2d *h)
{
    if (!h) return;
    for (int c = 0; c < 3; ++c) free(h->j[c]);
    if (h->m) or_maxwell2d_destroy(h->m);
    free(h->seg_cell);
    free(h->seg_val);
    free(h);
}

const double *or_split2d_j(const or_split2d *h, int component) { return h->j[component - 1]; }

static double *prow(or_split2d *h, int64_t i) { return h->part + (size_t)i * OR2_ROWS; }

/* field of degrees (degx, degy) at (x1, x2) (evaluate, particle_mesh_coupling_2d.jl:189-203) */
static double eval2(const or_split2d *h, const double *f, int degx, int degy, double x1, double x2)
{
    double vx[OR_MAXDEG + 1], vy[OR_MAXDEG + 1];
    const int64_t ix = point_basis(x1, h->xmin, h->dx, degx, vx);
    const int64_t iy = point_basis(x2, h->ymin, h->dy, degy, vy);
    double v = 0.0;
    for (int a = 0; a <= degx; ++a)
        for (int c = 0; c <= degy; ++c) v += f[node(h, ix + a, iy + c)] * vx[a] * vy[c];
    return v;
}

void or_split2d_operator_hb(or_split2d *h, double dt)
{
    or_maxwell2d_compute_e_from_b(h->m, dt, h->b, h->e);
}

void or_split2d_operator_he(or_split2d *h, double dt)
{
    const int p = h->p;
    for (int64_t i = 0; i < h->n; ++i) {
        double *r = prow(h, i);
        const double e1 = eval2(h, h->e[0], p - 1, p, r[0], r[1]);
        const double e2 = eval2(h, h->e[1], p, p - 1, r[0], r[1]);
        const double e3 = eval2(h, h->e[2], p, p, r[0], r[1]);
        r[2] = r[2] + dt * h->qm * e1;
        r[3] = r[3] + dt * h->qm * e2;
        r[4] = r[4] + dt * h->qm * e3;
    }
    or_maxwell2d_compute_b_from_e(h->m, dt, h->e, h->b);
}

void or_split2d_operator_hp3(or_split2d *h, double dt)
{
    const int p = h->p;
    const int64_t nn = h->nx * h->ny;
    double *j3 = h->j[2];
    memset(j3, 0, sizeof(double) * (size_t)nn);
    for (int64_t i = 0; i < h->n; ++i) {
        double *r = prow(h, i);
        const double q = (r[5] * h->charge) * h->cw;
        const double b1 = eval2(h, h->b[0], p, p - 1, r[0], r[1]);
        const double b2 = eval2(h, h->b[1], p - 1, p, r[0], r[1]);
        const double v3 = r[4];
        r[2] = r[2] - dt * h->qm * v3 * b2;
        r[3] = r[3] + dt * h->qm * v3 * b1;
        double vx[OR_MAXDEG + 1], vy[OR_MAXDEG + 1];
        const int64_t ix = point_basis(r[0], h->xmin, h->dx, p, vx);
        const int64_t iy = point_basis(r[1], h->ymin, h->dy, p, vy);
        const double wq = q * v3;
        for (int a = 0; a <= p; ++a)
            for (int c = 0; c <= p; ++c) j3[node(h, ix + a, iy + c)] += wq * vx[a] * vy[c];
    }
    for (int64_t k = 0; k < nn; ++k) j3[k] = j3[k] * dt;
    or_maxwell2d_compute_e_from_j(h->m, j3, 3, h->e[2]);
}

void or_split2d_operator_hp1(or_split2d *h, double dt)
{
    const int p = h->p;
    const int64_t nn = h->nx * h->ny;
    double *j1 = h->j[0];
    memset(j1, 0, sizeof(double) * (size_t)nn);
    for (int64_t i = 0; i < h->n; ++i) {
        double *r = prow(h, i);
        const double q = (r[5] * h->charge) * h->cw;
        const double x1n = r[0] + dt * r[2];
        const double xa = (r[0] - h->xmin) / h->dx, xb = (x1n - h->xmin) / h->dx;
        double y0[OR_MAXDEG + 1], y1[OR_MAXDEG + 1];
        const int64_t iy0 = point_basis(r[1], h->ymin, h->dy, p, y0);       /* S^p(x2) */
        const int64_t iy1 = point_basis(r[1], h->ymin, h->dy, p - 1, y1);   /* S^{p-1}(x2) */
        const int ns = line_pieces(h, p - 1, xa, xb, h->dx);
        double v2 = r[3], v3 = r[4];
        for (int s = 0; s < ns; ++s) {
            const int64_t fx = h->seg_cell[s] - (p - 1);
            const double *val = h->seg_val + (size_t)s * (OR_MAXDEG + 1);
            for (int a = 0; a < p; ++a) {
                for (int c = 0; c <= p; ++c) {
                    const int64_t k = node(h, fx + a, iy0 + c);
                    j1[k] += q * val[a] * y0[c];
                    v3 = v3 + h->qm * h->b[1][k] * val[a] * y0[c];
                }
                for (int c = 0; c < p; ++c)
                    v2 = v2 - h->qm * h->b[2][node(h, fx + a, iy1 + c)] * val[a] * y1[c];
            }
        }
        r[3] = v2;
        r[4] = v3;
        r[0] = h->xmin + or_jl_mod_f64(x1n - h->xmin, h->Lx);
    }
    or_maxwell2d_compute_e_from_j(h->m, j1, 1, h->e[0]);
}

void or_split2d_operator_hp2(or_split2d *h, double dt)
{
    const int p = h->p;
    const int64_t nn = h->nx * h->ny;
    double *j2 = h->j[1];
    memset(j2, 0, sizeof(double) * (size_t)nn);
    for (int64_t i = 0; i < h->n; ++i) {
        double *r = prow(h, i);
        const double q = (r[5] * h->charge) * h->cw;
        const double x2n = r[1] + dt * r[3];
        const double ya = (r[1] - h->ymin) / h->dy, yb = (x2n - h->ymin) / h->dy;
        double x0[OR_MAXDEG + 1], x1[OR_MAXDEG + 1];
        const int64_t ix0 = point_basis(r[0], h->xmin, h->dx, p, x0);       /* S^p(x1) */
        const int64_t ix1 = point_basis(r[0], h->xmin, h->dx, p - 1, x1);   /* S^{p-1}(x1) */
        const int ns = line_pieces(h, p - 1, ya, yb, h->dy);
        double v1 = r[2], v3 = r[4];
        for (int s = 0; s < ns; ++s) {
            const int64_t fy = h->seg_cell[s] - (p - 1);
            const double *val = h->seg_val + (size_t)s * (OR_MAXDEG + 1);
            for (int c = 0; c < p; ++c) {
                for (int a = 0; a <= p; ++a) {
                    const int64_t k = node(h, ix0 + a, fy + c);
                    j2[k] += q * x0[a] * val[c];
                    v3 = v3 - h->qm * h->b[0][k] * x0[a] * val[c];
                }
                for (int a = 0; a < p; ++a)
                    v1 = v1 + h->qm * h->b[2][node(h, ix1 + a, fy + c)] * x1[a] * val[c];
            }
        }
        r[2] = v1;
        r[4] = v3;
        r[1] = h->ymin + or_jl_mod_f64(x2n - h->ymin, h->Ly);
    }
    or_maxwell2d_compute_e_from_j(h->m, j2, 2, h->e[1]);
}

void or_split2d_strang(or_split2d *h, double dt, int64_t nsteps)
{
    for (int64_t s = 0; s < nsteps; ++s) {
        or_split2d_operator_hb(h, 0.5 * dt);
        or_split2d_operator_he(h, 0.5 * dt);
        or_split2d_operator_hp3(h, 0.5 * dt);
        or_split2d_operator_hp2(h, 0.5 * dt);
        or_split2d_operator_hp1(h, dt);
        or_split2d_operator_hp2(h, 0.5 * dt);
        or_split2d_operator_hp3(h, 0.5 * dt);
        or_split2d_operator_he(h, 0.5 * dt);
        or_split2d_operator_hb(h, 0.5 * dt);
    }
}

/* rho = sum_i q_i S^p(x1_i) S^p(x2_i) (add_charge!, particle_mesh_coupling_2d.jl:94-105) */
void or_split2d_rho(or_split2d *h, double *rho)
{
    const int p = h->p;
    memset(rho, 0, sizeof(double) * (size_t)(h->nx * h->ny));
    for (int64_t i = 0; i < h->n; ++i) {
        double *r = prow(h, i);
        const double q = (r[5] * h->charge) * h->cw;
        double vx[OR_MAXDEG + 1], vy[OR_MAXDEG + 1];
        const int64_t ix = point_basis(r[0], h->xmin, h->dx, p, vx);
        const int64_t iy = point_basis(r[1], h->ymin, h->dy, p, vy);
        for (int a = 0; a <= p; ++a)
            for (int c = 0; c <= p; ++c) rho[node(h, ix + a, iy + c)] += q * vx[a] * vy[c];
    }
}

/* rho from the fields by the weak Gauss law (compute_rho_from_e!, maxwell_2d_fem.jl:468-499) */
void or_split2d_rho_from_e(or_split2d *h, double *rho)
{
    or_maxwell2d_compute_rho_from_e(h->m, h->e, rho);
}

/* out[0] = 1/2 sum m w cw |v|^2, out[1..3] = 1/2 (E_c, M1 E_c), out[4..6] =
 * 1/2 (B_c, M2 B_c) (inner_product, maxwell_2d_fem.jl:511-575) */
void or_split2d_energy(or_split2d *h, double *out)
{
    double kin = 0.0;
    for (int64_t i = 0; i < h->n; ++i) {
        const double *r = prow(h, i);
        kin += (r[2] * r[2] + r[3] * r[3] + r[4] * r[4]) * ((r[5] * h->mass) * h->cw);
    }
    out[0] = 0.5 * kin;
    for (int c = 0; c < 3; ++c) {
        out[1 + c] = 0.5 * or_maxwell2d_inner_product(h->m, h->e[c], h->e[c], c + 1, 1);
        out[4 + c] = 0.5 * or_maxwell2d_inner_product(h->m, h->b[c], h->b[c], c + 1, 2);
    }
}
