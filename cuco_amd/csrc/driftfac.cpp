// Cosmological drift and kick factors (driftfac.c:14-126), host side.
//
// The reference fills two 1000-entry tables with GSL qag (GAUSS41, rel 1e-8)
// and interpolates them linearly in ln a.  Here the integrals come from an
// adaptive Simpson rule with Richardson correction to ~1e-13 relative, the
// tables are double, and the interpolation is the reference's
// (driftfac.c:71-94, including the u < 1 branch).
#include <cmath>

#include "cuco_engine.h"

namespace {

struct Cosmo {
    double om, ol, hubble;
    int kind;  // 0: drift 1/(H a^3), 1: kick 1/(H a^2)
};

double integrand(double a, const Cosmo &c)
{
    // driftfac.c:48-52 and :59-63
    double h = c.om / (a * a * a) + (1 - c.om - c.ol) / (a * a) + c.ol;
    h = c.hubble * std::sqrt(h);
    return c.kind == 0 ? 1 / (h * a * a * a) : 1 / (h * a * a);
}

double simpson(double a, double b, double fa, double fm, double fb)
{
    return (b - a) / 6.0 * (fa + 4.0 * fm + fb);
}

double adaptive(double a, double b, double fa, double fm, double fb,
                double whole, const Cosmo &c, double tol, int depth)
{
    const double m = 0.5 * (a + b);
    const double lm = 0.5 * (a + m), rm = 0.5 * (m + b);
    const double flm = integrand(lm, c), frm = integrand(rm, c);
    const double left = simpson(a, m, fa, flm, fm);
    const double right = simpson(m, b, fm, frm, fb);
    const double diff = left + right - whole;
    if (depth > 48 || std::fabs(diff) <= 15.0 * tol)
        return left + right + diff / 15.0;
    return adaptive(a, m, fa, flm, fm, left, c, 0.5 * tol, depth + 1) +
           adaptive(m, b, fm, frm, fb, right, c, 0.5 * tol, depth + 1);
}

double integrate(double a, double b, const Cosmo &c)
{
    const double fa = integrand(a, c), fb = integrand(b, c);
    const double m = 0.5 * (a + b), fm = integrand(m, c);
    const double whole = simpson(a, b, fa, fm, fb);
    return adaptive(a, b, fa, fm, fb, whole, c, 1e-14 * std::fabs(whole), 0);
}

double table_at(const double *t, double lb, double lm, double loga)
{
    // driftfac.c:74-82
    const double u = (loga - lb) / (lm - lb) * CUCO_DRIFT_TABLE_LENGTH;
    int i = (int)u;
    if (i >= CUCO_DRIFT_TABLE_LENGTH) i = CUCO_DRIFT_TABLE_LENGTH - 1;
    if (i < 1) return u * t[0];
    return t[i - 1] + (t[i] - t[i - 1]) * (u - i);
}

}  // namespace

extern "C" {

int cuco_init_drift_table(double a_begin, double a_max, double omega0,
                          double omega_lambda, double hubble, double *drift,
                          double *kick)
{
    if (!(a_begin > 0) || !(a_max > a_begin) || !(hubble > 0) || !drift || !kick)
        return CUCO_ERR_INVALID_ARG;
    const double lb = std::log(a_begin), lm = std::log(a_max);
    // Integrate each table entry from the previous one so the work stays
    // linear; the sum of adjacent integrals equals the integral from a_begin.
    double a_prev = a_begin, d_acc = 0, k_acc = 0;
    const Cosmo cd{omega0, omega_lambda, hubble, 0}, ck{omega0, omega_lambda, hubble, 1};
    for (int i = 0; i < CUCO_DRIFT_TABLE_LENGTH; ++i) {
        // driftfac.c:31: upper limit exp(ln a0 + (ln amax - ln a0)/N*(i+1))
        const double hi = std::exp(lb + ((lm - lb) / CUCO_DRIFT_TABLE_LENGTH) * (i + 1));
        d_acc += integrate(a_prev, hi, cd);
        k_acc += integrate(a_prev, hi, ck);
        drift[i] = d_acc;
        kick[i] = k_acc;
        a_prev = hi;
    }
    return CUCO_OK;
}

double cuco_table_factor(const double *table, double a_begin, double a_max,
                         double loga0, double loga1)
{
    const double lb = std::log(a_begin), lm = std::log(a_max);
    return table_at(table, lb, lm, loga1) - table_at(table, lb, lm, loga0);
}

int cuco_kdk_factors(const double *drift, const double *kick, double a_begin,
                     double a_max, double a_start, double a_end, int32_t nsteps,
                     double *factors)
{
    if (!drift || !kick || !factors || nsteps <= 0 || !(a_start >= a_begin) ||
        !(a_end <= a_max * (1 + 1e-12)) || !(a_end > a_start))
        return CUCO_ERR_INVALID_ARG;
    const double l0 = std::log(a_start), l1 = std::log(a_end);
    const double dl = (l1 - l0) / nsteps;
    for (int s = 0; s < nsteps; ++s) {
        const double la = l0 + dl * s, lb = l0 + dl * (s + 1);
        const double lmid = 0.5 * (la + lb);
        factors[3 * s + 0] = cuco_table_factor(kick, a_begin, a_max, la, lmid);
        factors[3 * s + 1] = cuco_table_factor(drift, a_begin, a_max, la, lb);
        factors[3 * s + 2] = cuco_table_factor(kick, a_begin, a_max, lmid, lb);
    }
    return CUCO_OK;
}

}  // extern "C"
