/*
 * numerov_oracle.c -- CPU oracle for the Numerov-shooting hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path in
 * epseon_backend_amd/csrc is compared against; only tests/, the smoke entry
 * point and the cpu_baseline leg of bench.py load it.  The product path never
 * links, loads or calls it.
 *
 * What it restates.  The reference (epseon_backend) owns the slot for this
 * path -- VibwaAlgorithm<FP>::run, cpp/gpu/include/epseon/gpu/algorithms/
 * vibwa.hpp:605-637 -- but never fills it: there is no shader, no dispatch
 * and no CPU numerics (cpp/cpu/source/libcpu.cpp:8-19 only greets).  The
 * arithmetic below is therefore the written-down FP64 specification both
 * implementations follow; the parameters and their meaning come from the
 * reference records:
 *   - Morse record (D_e, r_e, a, r_min, r_max, point_count):
 *     cpp/gpu/include/epseon/gpu/task_configurator/potential_source.hpp:102-183
 *   - algorithm record (masses, min_distance_to_asymptote, level range, one
 *     output value per level): task_configurator/algorithm_config.hpp:75-199,
 *     level_count at :177, output buffer of level_count elements at :189.
 *
 * Numerics (cm^-1, Angstrom, amu; everything FP64, no contraction, explicit fma):
 *   grid        r_i = r_min + i*h,  h = (r_max - r_min)/(N-1)
 *   beta        = (h*h*mu) / (12*K),  K = hbar^2/(2 amu A^2) in cm^-1
 *   shifted pot ut_i = fma(beta, V_i, -1)            (= beta*V_i - 1)
 *   probe       e = beta*E
 *   Numerov f   f_i = e - ut_i                       (= 1 + beta*(E - V_i))
 *   step        t = fma(-10, f_i, 12); g = f_i*f_{i-1};
 *               w_{i+1} = fma(t, w_i, -(g*w_{i-1}))   w_0 = 0, w_1 = 1
 *     (w_i = y_i * f_i * prod_{j<i} f_j, the division-free homogeneous form of
 *      f_{i+1} y_{i+1} = (12 - 10 f_i) y_i - f_{i-1} y_{i-1})
 *   samples     the sign of w is sampled after every B-th step (B = node
 *               stride, a power of two dividing 32) and once more at i = N-1
 *               when N-2 is not a multiple of B.  The node count is the number
 *               of sign changes between consecutive samples, starting from
 *               w_1 > 0.
 *   rescale     the N-2 steps form super-blocks of 32; after the 17th step of
 *               each full super-block, if max(|w_i|, |w_{i-1}|) >= 2^480 both
 *               are multiplied by 2^-480 (exact) and the rescale counter is
 *               incremented.  (Growth per step is below 16, so 47 steps
 *               between tests cannot overflow.)
 *   trial E     E_j = fma(j+1, dE, E_lo), dE = (E_hi - E_lo)/(M+1), E_lo = min V
 *   walls       the Dirichlet points are the last point left of argmin V and the
 *               first point right of it with V >= V_min + 1/(2 beta) (grid ends
 *               if none); the shot runs on that sub-grid, where
 *               1 + beta (E - V_i) > 1/2 for all E >= V_min.  A potential with
 *               1 + beta (E_hi - V_min) > 3/2 is rejected as unresolved.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EO_K 16.857629191640175        /* cm^-1 amu A^2 (CODATA 2018 h, c, u) */
#define EO_RESCALE_AT 0x1p480
#define EO_RESCALE_BY 0x1p-480
#define EO_MAX_STRIDE 32

/* Threads the OpenMP loops below use (1 without OpenMP). */
int32_t eo_num_threads(void)
{
#ifdef _OPENMP
    return (int32_t)omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ grid */

double eo_grid_step(double r_min, double r_max, int64_t n)
{
    return (r_max - r_min) / (double)(n - 1);
}

double eo_beta(double h, double mu)
{
    return (h * h * mu) / (12.0 * EO_K);
}

void eo_shifted_potential(const double* v, int64_t n, double beta, double* ut)
{
    for (int64_t i = 0; i < n; ++i)
        ut[i] = fma(beta, v[i], -1.0);
}

/* Morse V(r) = D_e (1 - exp(-a (r - r_e)))^2 on the grid r_i = r_min + i h. */
void eo_morse(double de, double re, double a, double r_min, double h, int64_t n, double* v)
{
    for (int64_t i = 0; i < n; ++i) {
        const double r = r_min + (double)i * h;
        const double x = 1.0 - exp(-a * (r - re));
        v[i] = de * (x * x);
    }
}

/* Wall trimming (see header).  Returns 1 when the potential is usable
 * (at least 4 points remain and 1 + beta (e_hi - V_min) <= 3/2). */
int32_t eo_trim(const double* v, int64_t n, double beta, double e_hi, int64_t* first, int64_t* last)
{
    int64_t imin = 0;
    for (int64_t i = 1; i < n; ++i)
        if (v[i] < v[imin])
            imin = i;
    const double vmin = v[imin];
    const double wall = vmin + 0.5 / beta;
    int64_t a = 0, b = n - 1;
    for (int64_t i = imin; i >= 0; --i)
        if (v[i] >= wall) {
            a = i;
            break;
        }
    for (int64_t i = imin; i < n; ++i)
        if (v[i] >= wall) {
            b = i;
            break;
        }
    *first = a;
    *last = b;
    const double f_max = 1.0 + beta * (e_hi - vmin);
    return (f_max <= 1.5 && b - a + 1 >= 4) ? 1 : 0;
}

double eo_trial_energy(int64_t j, double e_lo, double de)
{
    return fma((double)(j + 1), de, e_lo);
}

double eo_trial_spacing(double e_lo, double e_hi, int64_t m)
{
    return (e_hi - e_lo) / (double)(m + 1);
}

/* Node-sampling stride: the largest power of two B <= 32 such that the
 * discrete phase advance over B steps stays below pi/2 at the highest kinetic
 * energy of the window (E_hi - V_min).  Then two sign changes can never fall
 * inside one block, so sampling every B points counts every node. */
int32_t eo_node_stride(double beta, double e_hi, double v_min)
{
    const double f = 1.0 + beta * (e_hi - v_min);
    const double u = 12.0 / f - 10.0;
    if (!(u > -2.0))
        return 1;
    if (u >= 2.0)
        return EO_MAX_STRIDE;
    const double theta = acos(0.5 * u);
    const double bmax = floor(1.5707963267948966 / theta);
    int32_t b = 1;
    while (b * 2 <= EO_MAX_STRIDE && (double)(b * 2) <= bmax)
        b *= 2;
    return b;
}

/* ----------------------------------------------------------------- shoot */

static inline uint32_t eo_signbit(double x)
{
    uint64_t bits;
    memcpy(&bits, &x, sizeof bits);
    return (uint32_t)(bits >> 63);
}

/* One outward shot at probe e = beta*E.  Returns the sampled node count;
 * stores w_{N-1} and the number of 2^-480 rescales. */
int32_t eo_shoot(const double* ut, int64_t n, double e, int32_t stride,
                 double* w_last, int32_t* rescales)
{
    double wm = 0.0, w = 1.0, fm = e - ut[0];
    int32_t nodes = 0, k = 0;
    uint32_t sp = 0;
    const int64_t steps = n - 2;
    for (int64_t q = 0; q < steps; ++q) {
        const int64_t i = q + 1;
        const double f = e - ut[i];
        const double t = fma(-10.0, f, 12.0);
        const double g = f * fm;
        const double wn = fma(t, w, -(g * wm));
        wm = w;
        w = wn;
        fm = f;
        if ((q + 1) % stride == 0) {
            const uint32_t sb = eo_signbit(w);
            nodes += (int32_t)(sb ^ sp);
            sp = sb;
        }
        if ((q & 31) == 16 && q < (steps & ~(int64_t)31) &&
            fmax(fabs(w), fabs(wm)) >= EO_RESCALE_AT) {
            w *= EO_RESCALE_BY;
            wm *= EO_RESCALE_BY;
            ++k;
        }
    }
    if (steps % stride != 0) {
        const uint32_t sb = eo_signbit(w);
        nodes += (int32_t)(sb ^ sp);
    }
    *w_last = w;
    *rescales = k;
    return nodes;
}

/* Same recurrence with the sign compared at every point (stride-free count);
 * used by the tests to show that the sampled count misses no node. */
int32_t eo_shoot_full_count(const double* ut, int64_t n, double e)
{
    double wm = 0.0, w = 1.0, fm = e - ut[0];
    int32_t nodes = 0;
    uint32_t sp = 0;
    for (int64_t i = 1; i <= n - 2; ++i) {
        const double f = e - ut[i];
        const double t = fma(-10.0, f, 12.0);
        const double g = f * fm;
        const double wn = fma(t, w, -(g * wm));
        wm = w;
        w = wn;
        fm = f;
        const uint32_t sb = eo_signbit(w);
        nodes += (int32_t)(sb ^ sp);
        sp = sb;
        if (fmax(fabs(w), fabs(wm)) >= EO_RESCALE_AT) {
            w *= EO_RESCALE_BY;
            wm *= EO_RESCALE_BY;
        }
    }
    return nodes;
}

/* The full trajectory w_0..w_{N-1} without any rescaling (small N only). */
void eo_trajectory(const double* ut, int64_t n, double e, double* w_out)
{
    double wm = 0.0, w = 1.0, fm = e - ut[0];
    w_out[0] = 0.0;
    w_out[1] = 1.0;
    for (int64_t i = 1; i <= n - 2; ++i) {
        const double f = e - ut[i];
        const double t = fma(-10.0, f, 12.0);
        const double g = f * fm;
        const double wn = fma(t, w, -(g * wm));
        wm = w;
        w = wn;
        fm = f;
        w_out[i + 1] = w;
    }
}

/* ------------------------------------------------------------------ scan */

/* Shoot every trial energy E_j, j < m, of the window [e_lo, e_hi]. */
void eo_scan(const double* ut, int64_t n, double beta, double e_lo, double e_hi,
             int64_t m, int32_t stride, int32_t* nodes, double* w_last, int32_t* rescales)
{
    const double de = eo_trial_spacing(e_lo, e_hi, m);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t j = 0; j < m; ++j) {
        const double e = beta * eo_trial_energy(j, e_lo, de);
        nodes[j] = eo_shoot(ut, n, e, stride, &w_last[j], &rescales[j]);
    }
}

/* --------------------------------------------------------------- bracket */

typedef struct eo_bracket {
    double e_lo, e_hi;   /* bracket ends */
    double w_lo, w_hi;   /* far-end values at the ends */
    int32_t k_lo, k_hi;  /* their rescale counts */
    int32_t c_lo, c_hi;  /* node counts at the ends */
    int32_t lo_valid;    /* 0 when e_lo is the window floor (no shot there) */
    int32_t found;       /* 0 when no trial energy has more than v nodes */
} eo_bracket;

/* Level v lies between the first trial energy with more than v nodes and the
 * trial energy before it (or the window floor, below which no level exists
 * because every eigenvalue exceeds min V). */
void eo_bracket_level(const int32_t* nodes, const double* w_last, const int32_t* rescales,
                      int64_t m, double e_lo, double de, int32_t v, eo_bracket* out)
{
    int64_t lo = 0, hi = m;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (nodes[mid] > v)
            hi = mid;
        else
            lo = mid + 1;
    }
    const int64_t j = lo;
    memset(out, 0, sizeof *out);
    if (j == m) {
        out->found = 0;
        return;
    }
    out->found = 1;
    out->e_hi = eo_trial_energy(j, e_lo, de);
    out->c_hi = nodes[j];
    out->w_hi = w_last[j];
    out->k_hi = rescales[j];
    if (j == 0 || nodes[j - 1] > v) {
        out->e_lo = e_lo;
        out->c_lo = 0;
        out->lo_valid = 0;
    } else {
        out->e_lo = eo_trial_energy(j - 1, e_lo, de);
        out->c_lo = nodes[j - 1];
        out->w_lo = w_last[j - 1];
        out->k_lo = rescales[j - 1];
        out->lo_valid = 1;
    }
}

/* ---------------------------------------------------------------- refine */

/* Refine level v inside its bracket.  While the bracket holds more than one
 * level (or one end has no shot) it is halved by node count; once it holds
 * exactly level v the Illinois variant of regula falsi runs on the far-end
 * value F(E) = w_{N-1} * 2^(480 k), with every new point kept at least half a
 * tolerance inside the bracket.  Stops when the width is <= tol =
 * rel_tol*max(|E_lo|,|E_hi|) + abs_tol; the level is the final midpoint.
 * Returns the number of shots (or -1 when max_iter ran out). */
int32_t eo_refine_level(const double* ut, int64_t n, double beta, int32_t stride, int32_t v,
                        const eo_bracket* br, double rel_tol, double abs_tol, int32_t max_iter,
                        double* level)
{
    double el = br->e_lo, eh = br->e_hi;
    double wl = br->w_lo, wh = br->w_hi;
    int32_t kl = br->k_lo, kh = br->k_hi;
    int32_t cl = br->c_lo, ch = br->c_hi;
    int32_t lo_ok = br->lo_valid;
    int32_t side = 0;
    int32_t it = 0;
    for (;;) {
        const double width = eh - el;
        const double tol = fma(rel_tol, fmax(fabs(el), fabs(eh)), abs_tol);
        if (width <= tol)
            break;
        if (it >= max_iter) {
            *level = el + 0.5 * width;
            return -1;
        }
        double em;
        const int32_t isolated = lo_ok && cl == v && ch == v + 1;
        if (isolated) {
            const double r = ldexp(wh / wl, 480 * (kh - kl));
            const double s = 1.0 / (1.0 - r);
            em = fma(s, width, el);
            const double d = 0.5 * tol;
            if (!(em >= el + d))
                em = el + d;
            if (!(em <= eh - d))
                em = eh - d;
        } else {
            em = el + 0.5 * width;
            side = 0;
        }
        double w;
        int32_t k;
        const int32_t c = eo_shoot(ut, n, beta * em, stride, &w, &k);
        ++it;
        if (c <= v) {
            el = em; cl = c; wl = w; kl = k; lo_ok = 1;
            if (isolated && side == -1)
                wh *= 0.5;
            side = isolated ? -1 : 0;
        } else {
            eh = em; ch = c; wh = w; kh = k;
            if (isolated && side == 1)
                wl *= 0.5;
            side = isolated ? 1 : 0;
        }
    }
    *level = el + 0.5 * (eh - el);
    return it;
}

/* Scan + bracket + refine of levels [v_min, v_min + n_levels) for one
 * potential.  levels[] gets NaN for levels that are not below e_hi. */
void eo_solve(const double* ut, int64_t n, double beta, double e_lo, double e_hi, int64_t m,
              int32_t stride, int32_t v_min, int32_t n_levels, double rel_tol, double abs_tol,
              int32_t max_iter, int32_t* nodes, double* w_last, int32_t* rescales, double* levels,
              int32_t* iterations)
{
    eo_scan(ut, n, beta, e_lo, e_hi, m, stride, nodes, w_last, rescales);
    const double de = eo_trial_spacing(e_lo, e_hi, m);
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t l = 0; l < n_levels; ++l) {
        eo_bracket br;
        eo_bracket_level(nodes, w_last, rescales, m, e_lo, de, v_min + l, &br);
        if (!br.found) {
            levels[l] = NAN;
            iterations[l] = 0;
            continue;
        }
        iterations[l] = eo_refine_level(ut, n, beta, stride, v_min + l, &br, rel_tol, abs_tol,
                                        max_iter, &levels[l]);
    }
}
