// host_setup.hpp -- set-up computations done once on the host when a handle
// is created: the Maxwell1DFEM eigenvalue tables, the DFT twiddles of the
// circulant solves, the Gauss-Legendre rule of ParticleMeshCoupling1D and
// the SplinePP power-form tables.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace gh {

// Julia Base.mod for Float64 (host copy, for positions on the host side).
inline double jl_mod(double x, double y)
{
    double r = std::fmod(x, y);
    if (r == 0.0) return std::copysign(0.0, y);
    if ((r > 0.0) != (y > 0.0)) return r + y;
    return r;
}

// Gauss-Legendre nodes/weights on [-1,1] for the n <= 5 rules that the path
// uses: particle_mesh_coupling_1d.jl:67,72 takes n = (degree+2) / 2 points of
// FastGaussQuadrature.gausslegendre, whose small-n rules are these closed
// forms.
inline bool gauss_legendre(int n, double *x, double *w)
{
    switch (n) {
    case 1:
        x[0] = 0.0; w[0] = 2.0;
        return true;
    case 2:
        x[0] = -1.0 / std::sqrt(3.0); x[1] = 1.0 / std::sqrt(3.0);
        w[0] = 1.0; w[1] = 1.0;
        return true;
    case 3:
        x[0] = -std::sqrt(3.0 / 5.0); x[1] = 0.0; x[2] = std::sqrt(3.0 / 5.0);
        w[0] = 5.0 / 9.0; w[1] = 8.0 / 9.0; w[2] = 5.0 / 9.0;
        return true;
    case 4: {
        double a = 2.0 / 7.0 * std::sqrt(6.0 / 5.0);
        x[0] = -std::sqrt(3.0 / 7.0 + a); x[1] = -std::sqrt(3.0 / 7.0 - a);
        x[2] = std::sqrt(3.0 / 7.0 - a);  x[3] = std::sqrt(3.0 / 7.0 + a);
        w[0] = (18.0 - std::sqrt(30.0)) / 36.0; w[1] = (18.0 + std::sqrt(30.0)) / 36.0;
        w[2] = w[1]; w[3] = w[0];
        return true;
    }
    default:
        return false;
    }
}

// uniform_bsplines_eval_basis (low_level_bsplines.jl:38-61) on the host: the
// p+1 B-splines of degree p at offset x, with the reference's de Boor
// recurrence and rounding.
inline void bspline_host(int degree, double x, double *b)
{
    b[0] = 1.0;
    for (int j = 1; j <= degree; ++j) {
        double xx = -x;
        const double jr = (double)j;
        const double inv_j = 1.0 / jr;
        double saved = 0.0;
        for (int r = 0; r < j; ++r) {
            xx = xx + 1.0;
            const double temp = b[r] * inv_j;
            b[r] = saved + xx * temp;
            saved = (jr - xx) * temp;
        }
        b[j] = saved;
    }
}

// Eulerian-number mass lines of the degree-p (0-form) and degree-(p-1)
// (1-form) spline mass matrices, maxwell_1d_fem.jl:60-92.
inline bool mass_lines(int degree, double *m0, double *m1)
{
    if (degree == 1) {
        m0[0] = 4.0 / 6.0; m0[1] = 1.0 / 6.0;
        m1[0] = 1.0;
    } else if (degree == 2) {
        m0[0] = 66.0 / 120.0; m0[1] = 26.0 / 120.0; m0[2] = 1.0 / 120.0;
        m1[0] = 4.0 / 6.0; m1[1] = 1.0 / 6.0;
    } else if (degree == 3) {
        m0[0] = 2416.0 / 5040.0; m0[1] = 1191.0 / 5040.0;
        m0[2] = 120.0 / 5040.0; m0[3] = 1.0 / 5040.0;
        m1[0] = 66.0 / 120.0; m1[1] = 26.0 / 120.0; m1[2] = 1.0 / 120.0;
    } else {
        return false;
    }
    return true;
}

struct Maxwell1DTables {
    int64_t n = 0;
    std::vector<double> eig_mass0, eig_mass1, eig_weak_ampere, eig_weak_poisson;
    std::vector<double> inv_mass0, inv_mass1;   // 1/eig in the real slots, 0 elsewhere
    std::vector<double> cos_tab, sin_tab;       // cos/sin(2 pi t / n), t = 0..n-1
    // first column c of the circulant matrix C with C rhs = solve_circulant(eig, rhs)
    // (c = HC2R(eig) / n, in long double) for eig = inv_mass0, inv_mass1,
    // eig_weak_ampere, eig_mass1 and eig_weak_poisson
    std::vector<double> circ_m0, circ_m1, circ_amp, circ_mass1, circ_pois;
};

// c[m] = (1/n) (e[0] + (-1)^m e[n/2] + 2 sum_{k=1}^{n/2-1} (e[k] cos(2 pi k m/n) - e[n-k] sin(2 pi k m/n)))
inline std::vector<double> circulant_column(const std::vector<double> &e, int64_t n)
{
    std::vector<double> c((size_t)n, 0.0);
    const long double tp = 2.0L * 3.14159265358979323846264338327950288L;
    for (int64_t m = 0; m < n; ++m) {
        long double s = (long double)e[0] + ((m & 1) ? -1.0L : 1.0L) * (long double)e[n / 2];
        for (int64_t k = 1; k < n / 2; ++k) {
            const long double ang = tp * (long double)((k * m) % n) / (long double)n;
            s += 2.0L * ((long double)e[k] * cosl(ang) - (long double)e[n - k] * sinl(ang));
        }
        c[m] = (double)(s / (long double)n);
    }
    return c;
}

// Circulant eigenvalues in FFTW halfcomplex layout (real part of mode k at
// k, imaginary at n-k), with the expression order of
// maxwell_1d_fem.jl:108-153; the inverse-mass vectors as built by
// compute_e_from_j! (:267-284).
inline bool maxwell1d_tables(int64_t n, int degree, Maxwell1DTables &t)
{
    if (n < 2 || (n % 2) != 0) return false;
    double m0[4] = {0, 0, 0, 0}, m1[4] = {0, 0, 0, 0};
    if (!mass_lines(degree, m0, m1)) return false;
    t.n = n;
    t.eig_mass0.assign(n, 0.0);
    t.eig_mass1.assign(n, 0.0);
    t.eig_weak_ampere.assign(n, 0.0);
    t.eig_weak_poisson.assign(n, 0.0);
    const double pi = 3.141592653589793;
    const double nd = (double)n;
    const int s0 = degree;
    t.eig_weak_ampere[0] = 0.0;
    t.eig_weak_poisson[0] = 0.0;
    t.eig_mass0[0] = 1.0;
    t.eig_mass1[0] = 1.0;
    for (int64_t k = 1; k <= n / 2 - 1; ++k) {
        double coef0 = m0[0];
        double coef1 = m1[0];
        for (int j = 1; j <= s0 - 1; ++j) {
            double cos_mode = std::cos(2.0 * pi * (double)j * (double)k / nd);
            coef0 = coef0 + 2.0 * m0[j] * cos_mode;
            coef1 = coef1 + 2.0 * m1[j] * cos_mode;
        }
        coef0 = coef0 + 2.0 * m0[s0] * std::cos(2.0 * pi * (double)s0 * (double)k / nd);
        t.eig_mass0[k] = coef0;
        t.eig_mass0[n - k] = 0.0;
        t.eig_mass1[k] = coef1;
        t.eig_mass1[n - k] = 0.0;
        double cos_mode = std::cos(2.0 * pi * (double)k / nd);
        double sin_mode = std::sin(2.0 * pi * (double)k / nd);
        t.eig_weak_ampere[k] = (coef1 / coef0) * (1.0 - cos_mode);
        t.eig_weak_ampere[n - k] = -(coef1 / coef0) * sin_mode;
        double a = 1.0 - cos_mode;
        t.eig_weak_poisson[k] = 1.0 / (coef1 * (a * a + sin_mode * sin_mode));
        t.eig_weak_poisson[n - k] = 0.0;
    }
    {
        double coef0 = m0[0];
        double coef1 = m1[0];
        for (int j = 1; j <= s0 - 1; ++j) {
            coef0 = coef0 + 2.0 * m0[j] * std::cos(pi * (double)j);
            coef1 = coef1 + 2.0 * m1[j] * std::cos(pi * (double)j);
        }
        coef0 = coef0 + 2.0 * m0[s0] * std::cos(pi * (double)s0);
        t.eig_mass0[n / 2] = coef0;
        t.eig_mass1[n / 2] = coef1;
        t.eig_weak_ampere[n / 2] = 2.0 * (coef1 / coef0);
        t.eig_weak_poisson[n / 2] = 1.0 / (coef1 * 4.0);
    }
    t.inv_mass0.assign(n, 0.0);
    t.inv_mass1.assign(n, 0.0);
    for (int64_t i = 0; i <= n / 2; ++i) {
        t.inv_mass0[i] = 1.0 / t.eig_mass0[i];
        t.inv_mass1[i] = 1.0 / t.eig_mass1[i];
    }
    // Twiddles of the transform pair of solve_circulant! (FFTW R2HC/HC2R,
    // :222-240), as exact angles reduced to one period and rounded once.
    t.cos_tab.assign(n, 0.0);
    t.sin_tab.assign(n, 0.0);
    for (int64_t s = 0; s < n; ++s) {
        long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)s /
                          (long double)n;
        t.cos_tab[s] = (double)cosl(ang);
        t.sin_tab[s] = (double)sinl(ang);
    }
    t.circ_m0 = circulant_column(t.inv_mass0, n);
    t.circ_m1 = circulant_column(t.inv_mass1, n);
    t.circ_amp = circulant_column(t.eig_weak_ampere, n);
    t.circ_mass1 = circulant_column(t.eig_mass1, n);
    t.circ_pois = circulant_column(t.eig_weak_poisson, n);
    return true;
}

// Power-form coefficients of the pieces of the uniform B-spline of degree
// `degree` and of their primitives divided by x (SplinePP,
// splinepp.jl:30-233), derived exactly with the de Boor recurrence in
// rational arithmetic and rounded as num * (1.0/den) -- the rounding of the
// reference's `k * inv_m` constants. Layout coef[power_row + piece*(deg+1)],
// power_row 0 = x^degree.
inline bool splinepp_tables(int degree, double *pc, double *pf)
{
    if (degree < 1 || degree > 5) return false;
    struct RP { int64_t num[7]; int64_t den; int len; };
    auto gcd = [](int64_t a, int64_t b) {
        if (a < 0) a = -a;
        if (b < 0) b = -b;
        while (b) { int64_t t = a % b; a = b; b = t; }
        return a;
    };
    RP b[6];
    std::memset(b, 0, sizeof(b));
    b[0].num[0] = 1; b[0].den = 1; b[0].len = 1;
    for (int j = 1; j <= degree; ++j) {
        RP nb[6];
        std::memset(nb, 0, sizeof(nb));
        RP saved; std::memset(&saved, 0, sizeof(saved)); saved.den = 1; saved.len = 1;
        for (int r = 0; r < j; ++r) {
            RP temp = b[r];
            temp.den *= j;
            RP a; std::memset(&a, 0, sizeof(a));
            a.den = temp.den; a.len = temp.len + 1;
            for (int i = 0; i < temp.len; ++i) {
                a.num[i] += (r + 1) * temp.num[i];
                a.num[i + 1] -= temp.num[i];
            }
            RP s; std::memset(&s, 0, sizeof(s));
            s.den = saved.den * a.den;
            s.len = a.len > saved.len ? a.len : saved.len;
            for (int i = 0; i < s.len; ++i) s.num[i] = saved.num[i] * a.den + a.num[i] * saved.den;
            nb[r] = s;
            RP c; std::memset(&c, 0, sizeof(c));
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
            for (int i = 0; i < nb[r].len; ++i) g = gcd(g, nb[r].num[i]);
            if (g > 1) {
                nb[r].den /= g;
                for (int i = 0; i < nb[r].len; ++i) nb[r].num[i] /= g;
            }
            b[r] = nb[r];
        }
    }
    const int dp1 = degree + 1;
    for (int piece = 0; piece <= degree; ++piece) {
        for (int pw = 0; pw <= degree; ++pw) {
            int64_t num = pw < b[piece].len ? b[piece].num[pw] : 0;
            int64_t den = b[piece].den;
            int64_t g = gcd(num, den);
            double v = 0.0;
            if (num != 0) v = (double)(num / g) * (1.0 / (double)(den / g));
            pc[(degree - pw) + piece * dp1] = v;
            int64_t den2 = den * (pw + 1);
            g = gcd(num, den2);
            v = 0.0;
            if (num != 0) v = (double)(num / g) * (1.0 / (double)(den2 / g));
            pf[(degree - pw) + piece * dp1] = v;
        }
    }
    return true;
}

}  // namespace gh
