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
 *   scan table  p_i = fma(beta, V_i, -1) / 12                (scan)
 *   scaled pot  uv_i = beta*V_i                       (refinement, see below)
 *   probe       e = beta*E; scan shots use e' = e / 12
 *   scan step   phi_i = e' - p_i                      (= f_i / 12, f_i = 1 + beta*(E - V_i))
 *               z = phi_{i-1} * v_{i-1}; s = fma(10, v_i, z);
 *               v_{i+1} = fma(-phi_i, s, v_i)         v_0 = 0, v_1 = 1
 *     (v_i = w_i / 12^(i-1) with w the division-free homogeneous Numerov
 *      solution w_{i+1} = fma(t, w_i, -(g*w_{i-1})), t = 12 - 10 f_i,
 *      g = f_i f_{i-1}, w_i = y_i * f_i * prod_{j<i} f_j, of
 *      f_{i+1} y_{i+1} = (12 - 10 f_i) y_i - f_{i-1} y_{i-1}; 4 operations
 *      per step instead of 5)
 *   samples     the sign of w is sampled after every B-th step (B = node
 *               stride, a power of two dividing 32) and once more at i = N-1
 *               when N-2 is not a multiple of B.  The node count is the number
 *               of sign changes between consecutive samples, starting from
 *               w_1 > 0.
 *   rescale     the N-2 steps form super-blocks of 32.  Scan: after the 17th
 *               step of every second full super-block (q % 64 == 48), if
 *               max(|v_i|, |v_{i-1}|) < 2^-240 both are multiplied by 2^480
 *               and the scale-up count r grows (|w| grows by less than 7.31
 *               per step, so v only shrinks, by at most 2^-8.5 per step in the
 *               pair norm: the <= 78 unguarded steps stay above 2^-903).  Far end: the scan reports
 *               v_{N-1} * c and r, with 12^(N-2) = c * 2^I, c = 2^frac(L),
 *               I = floor(L), L = (N-2) log2(12), so w_{N-1} = v_{N-1} c 2^(I - 480 r).
 *               Refinement (difference form): after the 17th step of every
 *               fourth full super-block (step q with q % 128 == 112), if
 *               max(|w_i|, |w_{i-1}|) >= 2^480 both are multiplied by 2^-480
 *               (exact) and its counter grows (|t| < 7 and g < 9/4 bound the
 *               growth per step by 7.31 < 2^2.87, so the at most 142 steps
 *               between tests stay below 2^888).
 *   trial E     E_j = fma(j+1, dE, E_lo), dE = (E_hi - E_lo)/(M+1), E_lo = min V
 *   walls       the Dirichlet points are the last point left of argmin V and the
 *               first point right of it with V >= V_min + 1/(2 beta) (grid ends
 *               if none); the shot runs on that sub-grid, where
 *               1 + beta (E - V_i) > 1/2 for all E >= V_min.  A potential with
 *               1 + beta (E_hi - V_min) > 3/2 is rejected as unresolved.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EO_SCAN_RESCALE_BELOW 0x1p-240
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

/* Set the threads of the OpenMP loops below (timing a single-core run). */
void eo_set_num_threads(int32_t n)
{
#ifdef _OPENMP
    omp_set_num_threads(n < 1 ? 1 : n);
#else
    (void)n;
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

void eo_scan_table(const double* v, int64_t n, double beta, double* pt)
{
    for (int64_t i = 0; i < n; ++i)
        pt[i] = fma(beta, v[i], -1.0) / 12.0;
}

void eo_scaled_potential(const double* v, int64_t n, double beta, double* uv)
{
    for (int64_t i = 0; i < n; ++i)
        uv[i] = beta * v[i];
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

/* Far-end scale of an n-point scan shot: 12^(n-2) = *c * 2^*i. */
void eo_far_scale(int64_t n, double* c, int32_t* i)
{
    const double l2 = (double)(n - 2) * log2(12.0);
    const double fl = floor(l2);
    *i = (int32_t)fl;
    *c = exp2(l2 - fl);
}

/* One outward scan shot at probe e = beta*E over p_i = (beta V_i - 1)/12.
 * Returns the sampled node count; stores the far-end mantissa v_{N-1} * c and
 * the scale-up count r (w_{N-1} = mantissa * 2^(I - 480 r), eo_far_scale). */
int32_t eo_shoot(const double* pt, int64_t n, double e, int32_t stride,
                 double* w_last, int32_t* rescales)
{
    const double ep = e / 12.0;
    double wm = 0.0, w = 1.0, fm = ep - pt[0];
    int32_t nodes = 0, k = 0;
    uint32_t sp = 0;
    const int64_t steps = n - 2;
    for (int64_t q = 0; q < steps; ++q) {
        const double f = ep - pt[q + 1];
        const double z = fm * wm;
        const double sg = fma(10.0, w, z);
        const double wn = fma(-f, sg, w);
        wm = w;
        w = wn;
        fm = f;
        if ((q + 1) % stride == 0) {
            const uint32_t sb = eo_signbit(w);
            nodes += (int32_t)(sb ^ sp);
            sp = sb;
        }
        if ((q & 63) == 48 && q < (steps & ~(int64_t)31) &&
            fmax(fabs(w), fabs(wm)) < EO_SCAN_RESCALE_BELOW) {
            w *= EO_RESCALE_AT;
            wm *= EO_RESCALE_AT;
            ++k;
        }
    }
    if (steps % stride != 0) {
        const uint32_t sb = eo_signbit(w);
        nodes += (int32_t)(sb ^ sp);
    }
    double c;
    int32_t i;
    eo_far_scale(n, &c, &i);
    *w_last = w * c;
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

/* ---------------------------------------------------------------- windows
 *
 * Past its outermost turning points a shot only grows (or, from the left
 * wall, only decays towards the well) at a rate of at least
 * acosh(1 + x/2) per step, x_i = 12 (uv_i - e) > 0 (for a locally constant x
 * the recurrence's dominant root lambda has lambda + 1/lambda >= 2 + x).
 * Moving a Dirichlet point into that region from depth S to depth S' > S
 * changes the eigenvalues only by a relative e^(-2S)-order amount, so shots
 * may stop where the guaranteed depth reaches EO_WINDOW_DEPTH = 15 e-folds
 * (e^-30 ~ 1e-13 of a level spacing).  The depth is summed in 2^-40 fixed
 * point from an integer lower bound g(x) of acosh(1 + x/2) within 9 % of it
 * (eo_growth_term; x < 6, which the wall trim guarantees), so the sum is
 * exact and independent of order.
 *
 * Window of a probe e (= beta E) on the n-point trimmed grid: i_first and
 * i_last are the first and last points with uv_i <= e; hi is the first
 * i > i_last at which sum_{i_last < j <= i} g(x_j) >= 15 (else n - 1), lo the
 * last i < i_first at which sum_{i <= j < i_first} g(x_j) >= 15 (else 0).
 *
 * Energy bands: the window [e_lo, e_hi] splits into EO_BANDS bands with tops
 * E_b = fma(b + 1, (e_hi - e_lo)/EO_BANDS, e_lo); band b's window is the
 * window at E_b (windows only grow with E).  The scan shoots trial energies in
 * groups of 64 (j = 64 g .. 64 g + 63); a group uses the window of the band of
 * its highest energy (the first b with E_b >= E_top, else the last band),
 * and a level is refined in the window of the group of its bracket's upper
 * scan point.  Windows off = every band's window is the whole grid. */
#define EO_BANDS 16
#define EO_WINDOW_DEPTH 15

/* floor(65536 acosh(1 + x/2)/sqrt(x)) at x = 2^k, k = -5..3: the ratio only
 * falls with x, so entry k bounds it on [2^(k-1), 2^k) and entry -5 below. */
static const int64_t eo_growth_ratio[9] = {65450, 65366, 65199, 64871, 64242, 63073, 61029, 57761, 53116};

/* floor(sqrt(n)) for n < 2^53 (the double square root, then corrected). */
static uint64_t eo_isqrt(uint64_t n)
{
    uint64_t r = (uint64_t)sqrt((double)n);
    while (r * r > n)
        --r;
    while ((r + 1) * (r + 1) <= n)
        ++r;
    return r;
}

/* Lower bound of acosh(1 + x/2) in 2^-40 fixed point, x > 0: with
 * x = m 2^k (m in [1/2, 1)), T = the top 16 bits of m's 53-bit significand,
 * x >= T 2^(k-16); sqrt(T 2^(k-16)) 2^40 = sqrt(T 2^r) 2^h (k + 64 = 2 h + r)
 * is taken as isqrt(T 2^(r+32)) 2^(h-16), floored, and scaled by the ratio
 * table.  Integer operations only after frexp, so any two implementations
 * agree.  0 below 2^-65; x >= 8 counts as just below 8. */
static int64_t eo_growth_term(double x)
{
    int k;
    const double m = frexp(x, &k);
    if (!(x > 0.0) || k < -64)
        return 0;
    uint64_t t = (uint64_t)ldexp(m, 16);
    if (k > 3) {
        k = 3;
        t = 0xFFFF;
    }
    const int32_t e = k + 64, h = e >> 1;
    const uint64_t r = eo_isqrt(t << ((e & 1) + 32));
    const uint64_t sq = h >= 16 ? r << (h - 16) : r >> (16 - h);
    const int64_t f = eo_growth_ratio[(k < -5 ? -5 : k) + 5];
    return (int64_t)((sq * (uint64_t)f) >> 16);
}

void eo_window(const double* uv, int64_t n, double e, int64_t* lo, int64_t* hi)
{
    int64_t first = -1, last = -1;
    for (int64_t i = 0; i < n; ++i)
        if (uv[i] <= e) {
            if (first < 0)
                first = i;
            last = i;
        }
    *lo = 0;
    *hi = n - 1;
    if (first < 0)
        return;
    const int64_t depth = (int64_t)EO_WINDOW_DEPTH << 40;
    int64_t acc = 0;
    for (int64_t i = last + 1; i < n; ++i) {
        acc += eo_growth_term(12.0 * (uv[i] - e));
        if (acc >= depth) {
            *hi = i;
            break;
        }
    }
    acc = 0;
    for (int64_t i = first - 1; i >= 0; --i) {
        acc += eo_growth_term(12.0 * (uv[i] - e));
        if (acc >= depth) {
            *lo = i;
            break;
        }
    }
}

double eo_band_top(int32_t b, double e_lo, double e_hi)
{
    return fma((double)(b + 1), (e_hi - e_lo) / EO_BANDS, e_lo);
}

/* win[2 b], win[2 b + 1] = (lo, hi) of band b. */
void eo_band_windows(const double* uv, int64_t n, double beta, double e_lo, double e_hi, int64_t* win)
{
    for (int32_t b = 0; b < EO_BANDS; ++b)
        eo_window(uv, n, beta * eo_band_top(b, e_lo, e_hi), &win[2 * b], &win[2 * b + 1]);
}

void eo_full_windows(int64_t n, int64_t* win)
{
    for (int32_t b = 0; b < EO_BANDS; ++b) {
        win[2 * b] = 0;
        win[2 * b + 1] = n - 1;
    }
}

int32_t eo_group_band(int64_t j, int64_t m, double e_lo, double de, double e_hi)
{
    int64_t jt = (j / 64) * 64 + 63;
    if (jt > m - 1)
        jt = m - 1;
    const double et = eo_trial_energy(jt, e_lo, de);
    for (int32_t b = 0; b < EO_BANDS - 1; ++b)
        if (eo_band_top(b, e_lo, e_hi) >= et)
            return b;
    return EO_BANDS - 1;
}

/* Shoot every trial energy E_j, j < m, of the window [e_lo, e_hi], each in
 * its group's band window win (2 EO_BANDS entries). */
void eo_scan(const double* pt, int64_t n, double beta, double e_lo, double e_hi,
             int64_t m, int32_t stride, const int64_t* win, int32_t* nodes, double* w_last,
             int32_t* rescales)
{
    (void)n;
    const double de = eo_trial_spacing(e_lo, e_hi, m);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t j = 0; j < m; ++j) {
        const int32_t b = eo_group_band(j, m, e_lo, de, e_hi);
        const int64_t lo = win[2 * b], hi = win[2 * b + 1];
        const double e = beta * eo_trial_energy(j, e_lo, de);
        nodes[j] = eo_shoot(pt + lo, hi - lo + 1, e, stride, &w_last[j], &rescales[j]);
    }
}

/* --------------------------------------------------------------- bracket */

/* Interpolation points of a level: at most EO_PTS (energy, F = w 2^k) pairs
 * of its band window, kept sorted by energy. */
#define EO_PTS 12

typedef struct eo_bracket {
    double e_lo, e_hi;   /* bracket ends */
    double w_lo, w_hi;   /* far-end mantissas at the ends */
    int32_t k_lo, k_hi;  /* their binary exponents (F = w 2^k) */
    int32_t c_lo, c_hi;  /* node counts at the ends */
    int32_t lo_valid;    /* 0 when e_lo is the window floor (no shot there) */
    int32_t found;       /* 0 when no trial energy has more than v nodes */
    double rate;         /* b of the model F ~ (c0 + c1 E) e^(b E) (0: none) */
    int32_t band;        /* band of the upper end's scan group: the level's window */
    int32_t lo_foreign;  /* 1 when the lower end's value comes from another band's window */
    int32_t n_pts;       /* scan points of the level's window near the bracket (<= EO_PTS) */
    int32_t i_hi;        /* index of the upper end among them */
    double e_pt[EO_PTS]; /* their energies, ascending */
    double w_pt[EO_PTS]; /* far-end mantissas */
    int32_t k_pt[EO_PTS]; /* binary exponents */
} eo_bracket;

/* ---- deterministic exp / log for the refinement model -----------------
 * Written with +, -, *, /, fma, floor, frexp and ldexp only, so the oracle
 * and the kernels produce the same bits.  Accuracy ~1e-15 relative (the model only
 * steers the probes). */
#define EO_LN2 0.69314718055994530942
/* ln 2 split so that q * EO_LN2_HI is exact for |q| < 2^20 (Cody-Waite). */
#define EO_LN2_HI 6.93147180369123816490e-01
#define EO_LN2_LO 1.90821492927058770002e-10
#define EO_SQRT1_2 0.70710678118654752440

/* e^y = m 2^n with m in [1, 2). */
static void eo_model_exp(double y, double* m, int32_t* n)
{
    const double q = floor(y / EO_LN2);
    const double r = (y - q * EO_LN2_HI) - q * EO_LN2_LO;
    /* Taylor coefficients 1/k!, k = 0..17 (truncation < 2e-19 on [0, ln 2)). */
    static const double c[18] = {1.0, 1.0, 1.0 / 2, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720,
                                 1.0 / 5040, 1.0 / 40320, 1.0 / 362880, 1.0 / 3628800,
                                 1.0 / 39916800, 1.0 / 479001600, 1.0 / 6227020800.0,
                                 1.0 / 87178291200.0, 1.0 / 1307674368000.0,
                                 1.0 / 20922789888000.0, 1.0 / 355687428096000.0};
    double p = c[17];
    for (int i = 16; i >= 0; --i)
        p = fma(p, r, c[i]);
    *m = p;
    *n = (int32_t)q;
}

/* sqrt d for d >= 0 by Newton iterations from a linear first guess. */
static double eo_model_sqrt(double d)
{
    if (!(d > 0.0))
        return 0.0;
    int e = 0;
    double m = frexp(d, &e);
    if (e & 1) {
        m = m * 2.0;
        e = e - 1;
    }
    double y = fma(m, 0.4142135623730950, 0.5857864376269050);
    for (int i = 0; i < 5; ++i)
        y = 0.5 * (y + m / y);
    return ldexp(y, e / 2);
}

/* ln x for x > 0. */
static double eo_model_log(double x)
{
    int e = 0;
    double m = frexp(x, &e);
    if (m < EO_SQRT1_2) {
        m = m * 2.0;
        e = e - 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 1.0 / 25.0;
    for (int i = 11; i >= 0; --i)
        p = fma(p, s2, 1.0 / (double)(2 * i + 1));
    return fma((double)e, EO_LN2, 2.0 * s * p);
}

/* Rate b of the model F(E) = (c0 + c1 (E - E_b)) e^(b (E - E_b)) through the
 * scan points a < b < c (equally spaced by de, F = w 2^k): with
 * x = e^(b de), F_a/F_b x + F_c/F_b / x = 2.  The bracket is [E_b, E_c] (third
 * point below, F changes sign between b and c) or [E_a, E_b] (third point
 * above); in either case exactly one root x is positive.  0 when the points
 * do not allow a fit. */
static double eo_model_rate(double wa, int32_t ka, double wb, int32_t kb, double wc, int32_t kc,
                            int32_t above, double de)
{
    if (wb == 0.0 || ka - kb > 960 || kb - ka > 960 || kc - kb > 960 || kb - kc > 960)
        return 0.0;
    const double ra = ldexp(wa / wb, ka - kb);
    const double rc = ldexp(wc / wb, kc - kb);
    const double disc = 1.0 - ra * rc;
    if (!(disc >= 0.0) || ra == 0.0 || !isfinite(ra) || !isfinite(rc))
        return 0.0;
    const double sq = eo_model_sqrt(disc);
    const double x = above ? (1.0 - sq) / ra : (1.0 + sq) / ra;
    if (!(x > 0.0) || !isfinite(x))
        return 0.0;
    const double lx = eo_model_log(x);
    if (!(fabs(lx) < 600.0))
        return 0.0;
    return lx / de;
}

/* F -> G = F e^(-b (E - e0)) as mantissa/exponent. */
static void eo_model_apply(double b, double e0, double e, double* w, int32_t* k)
{
    double m;
    int32_t n;
    eo_model_exp(-b * (e - e0), &m, &n);
    *w = *w * m;
    *k = *k + n;
}

/* Level v lies between the first trial energy with more than v nodes and the
 * trial energy before it (or the window floor, below which no level exists
 * because every eigenvalue exceeds min V).  win and far_exp hold band b's
 * window (win[2 b], win[2 b + 1]) and far-end exponent.  Scan values from a
 * group whose band window differs from the upper end's are not comparable
 * with the level's shots: such a lower end is flagged (lo_foreign) and a
 * third point of the model must share the window. */
static int32_t eo_same_window(const int64_t* win, int32_t a, int32_t b)
{
    return win[2 * a] == win[2 * b] && win[2 * a + 1] == win[2 * b + 1];
}

void eo_bracket_level(const int32_t* nodes, const double* w_last, const int32_t* rescales,
                      int64_t m, double e_lo, double de, double e_hi, int32_t v,
                      const int64_t* win, const int32_t* far_exp, eo_bracket* out)
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
    const int32_t band = eo_group_band(j, m, e_lo, de, e_hi);
    const int32_t fe = far_exp[band];
    out->found = 1;
    out->band = band;
    out->e_hi = eo_trial_energy(j, e_lo, de);
    out->c_hi = nodes[j];
    out->w_hi = w_last[j];
    out->k_hi = fe - rescales[j] * 480;
    out->rate = 0.0;
    if (j == 0 || nodes[j - 1] > v) {
        out->e_lo = e_lo;
        out->c_lo = 0;
        out->lo_valid = 0;
    } else {
        const int32_t bl = eo_group_band(j - 1, m, e_lo, de, e_hi);
        out->e_lo = eo_trial_energy(j - 1, e_lo, de);
        out->c_lo = nodes[j - 1];
        out->w_lo = w_last[j - 1];
        out->k_lo = far_exp[bl] - rescales[j - 1] * 480;
        out->lo_valid = 1;
        out->lo_foreign = !eo_same_window(win, bl, band);
    }
    /* Interpolation points: scan points j - 6 .. j + 5 whose group shares the
     * level's window, the lower end only when its value is of that window. */
    out->n_pts = 0;
    for (int64_t q = j - EO_PTS / 2; q < j + EO_PTS / 2; ++q) {
        if (q < 0 || q >= m)
            continue;
        if (q == j - 1 && !(out->lo_valid && !out->lo_foreign))
            continue;
        const int32_t bq = eo_group_band(q, m, e_lo, de, e_hi);
        if (q != j && q != j - 1 && !eo_same_window(win, bq, band))
            continue;
        if (q == j)
            out->i_hi = out->n_pts;
        out->e_pt[out->n_pts] = eo_trial_energy(q, e_lo, de);
        out->w_pt[out->n_pts] = w_last[q];
        out->k_pt[out->n_pts] = far_exp[bq] - rescales[q] * 480;
        ++out->n_pts;
    }
    /* Model rate from a third scan point in the same count regime: above
     * (j + 1, same count as j) or below (j - 2, same count as j - 1), all
     * three in the level's band. */
    if (out->lo_valid && !out->lo_foreign && out->c_lo == v && out->c_hi == v + 1) {
        if (j + 1 < m && nodes[j + 1] == out->c_hi &&
            eo_same_window(win, eo_group_band(j + 1, m, e_lo, de, e_hi), band))
            out->rate = eo_model_rate(out->w_lo, out->k_lo, out->w_hi, out->k_hi, w_last[j + 1],
                                      fe - rescales[j + 1] * 480, 1, de);
        else if (j >= 2 && nodes[j - 2] == out->c_lo &&
                 eo_same_window(win, eo_group_band(j - 2, m, e_lo, de, e_hi), band))
            out->rate = eo_model_rate(w_last[j - 2], fe - rescales[j - 2] * 480, out->w_lo, out->k_lo,
                                      out->w_hi, out->k_hi, 0, de);
    }
}

/* --------------------------------------------------- difference form
 *
 * Refinement shots use the same recurrence written on (w_i, d_i = w_i -
 * w_{i-1}) with the small quantities kept at full relative precision:
 *   eps_i = e - uv_i  (uv_i = beta V_i, so eps_i = beta (E - V_i))
 *   sig_i = fma(eps_i, eps_{i-1}, eps_i + eps_{i-1})      (g_i = 1 + sig_i)
 *   d_{i+1} = fma(-10 eps_i, w_i, fma(-sig_i, w_{i-1}, d_i))
 *   w_{i+1} = w_i + d_{i+1}
 * (from w_{i+1} = (2 - 10 eps_i) w_i - (1 + sig_i) w_{i-1}).  With f = 1 + eps
 * held near 1, FP64 quantises E in steps of 2^-52/beta; here the root of
 * w_{N-1}(E) is resolved to ~1e-15 relative.  w_0 = 0, w_1 = d_1 = 1; the
 * sign sampling and the 2^480 rescale test (on max(|w_i|, |w_{i-1}|), scaling
 * w, d and w_{i-1}) sit where they sit in eo_shoot.
 */

int32_t eo_shoot_diff(const double* uv, int64_t n, double e, int32_t stride, double* w_last,
                      int32_t* rescales)
{
    double w = 1.0, d = 1.0, wm = 0.0, em = e - uv[0];
    int32_t nodes = 0, k = 0;
    uint32_t sp = 0;
    const int64_t steps = n - 2;
    for (int64_t q = 0; q < steps; ++q) {
        const double ep = e - uv[q + 1];
        const double e10 = ep * 10.0;
        const double sg = fma(ep, em, ep + em);
        d = fma(-e10, w, fma(-sg, wm, d));
        wm = w;
        w = w + d;
        em = ep;
        if ((q + 1) % stride == 0) {
            const uint32_t sb = eo_signbit(w);
            nodes += (int32_t)(sb ^ sp);
            sp = sb;
        }
        if ((q & 127) == 112 && q < (steps & ~(int64_t)31) && fmax(fabs(w), fabs(wm)) >= EO_RESCALE_AT) {
            w *= EO_RESCALE_BY;
            d *= EO_RESCALE_BY;
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

/* ------------------------------------------------------ segmented shots
 *
 * A difference-form shot split over S segments of the Q = N-2 steps:
 * segment s covers steps [q_s, q_{s+1}), q_s = (s*Q)/S.  Each segment yields
 * the 2x2 matrix 2^E [[m00 m01][m10 m11]] taking (w, d) at the segment start
 * to (w, d) at its end: the two columns start from (w, d) = (1, 0) and (0, 1)
 * with w_prev = w - d, and after every 128th step of the segment and at its
 * end all carried values are divided by 2^x, x the frexp exponent of their
 * largest magnitude (a step grows them by less than 2^2.87 and shrinks the
 * largest by less than 2^4.5, so 128 steps stay in the normal range).  The chain v <- M_s v from v = (w_1, d_1) = (1, 1),
 * renormalised the same way after each product, ends at F = w_{N-1} = v0 2^E.
 * The far-end value comes from the product of all segment matrices, formed
 * as a fixed binary tree (at distance d = 1, 2, 4, ...: M_i <- M_{i+d} M_i for
 * i a multiple of 2d, each product renormalised by the frexp exponent of its
 * largest entry), applied to (w_1, d_1) = (1, 1).  For node counts the start
 * vectors come from the sequential chain v <- M_s v (renormalised after every
 * product), each segment is re-run from its start vector with signs sampled
 * at the global positions of the sequential shot, and the segment counts are
 * joined across segment borders.  S = clamp(Q/32, 1, lanes) for a block of
 * `lanes` threads per level.
 */

typedef struct eo_mat {
    double m00, m01, m10, m11;
    int32_t e;
} eo_mat;

int32_t eo_segment_count(int64_t n, int32_t lanes)
{
    const int64_t q = n - 2;
    int64_t s = q / 32;
    if (s < 1)
        s = 1;
    if (s > lanes)
        s = lanes;
    return (int32_t)s;
}

int64_t eo_segment_start(int64_t n, int32_t segments, int32_t s)
{
    return ((int64_t)s * (n - 2)) / segments;
}

static int32_t eo_max_exponent(const double* x, int32_t k)
{
    double m = 0.0;
    for (int32_t i = 0; i < k; ++i)
        m = fmax(m, fabs(x[i]));
    int e = 0;
    frexp(m, &e);
    return e;
}

static void eo_scale(double* x, int32_t k, int32_t e)
{
    for (int32_t i = 0; i < k; ++i)
        x[i] = ldexp(x[i], -e);
}

void eo_segment_matrix(const double* uv, int64_t q0, int64_t q1, double e, eo_mat* out)
{
    /* x = {wA, dA, wmA, wB, dB, wmB} */
    double x[6] = {1.0, 0.0, 1.0, 0.0, 1.0, -1.0};
    double em = e - uv[q0];
    int32_t ex = 0;
    for (int64_t q = q0; q < q1; ++q) {
        const double ep = e - uv[q + 1];
        const double e10 = ep * 10.0;
        const double sg = fma(ep, em, ep + em);
        for (int c = 0; c < 6; c += 3) {
            x[c + 1] = fma(-e10, x[c], fma(-sg, x[c + 2], x[c + 1]));
            x[c + 2] = x[c];
            x[c] = x[c] + x[c + 1];
        }
        em = ep;
        if (((q - q0) & 127) == 127) {
            const int32_t sh = eo_max_exponent(x, 6);
            eo_scale(x, 6, sh);
            ex += sh;
        }
    }
    double y[4] = {x[0], x[1], x[3], x[4]};
    const int32_t sh = eo_max_exponent(y, 4);
    eo_scale(y, 4, sh);
    out->m00 = y[0];
    out->m10 = y[1];
    out->m01 = y[2];
    out->m11 = y[3];
    out->e = ex + sh;
}

/* Chain the segment matrices from (w_1, d_1) = (1, 1).  starts[2s], [2s+1]
 * receive the normalised (w, d) at the start of segment s (may be NULL). */
void eo_chain(const eo_mat* m, int32_t segments, double* f_mant, int32_t* f_exp, double* starts)
{
    double v[2] = {1.0, 1.0};
    int32_t ex = 0;
    for (int32_t s = 0; s < segments; ++s) {
        if (starts) {
            starts[2 * s] = v[0];
            starts[2 * s + 1] = v[1];
        }
        const double a = fma(m[s].m00, v[0], m[s].m01 * v[1]);
        const double b = fma(m[s].m10, v[0], m[s].m11 * v[1]);
        v[0] = a;
        v[1] = b;
        const int32_t sh = eo_max_exponent(v, 2);
        eo_scale(v, 2, sh);
        ex += m[s].e + sh;
    }
    *f_mant = v[0];
    *f_exp = ex;
}

typedef struct eo_seg_count {
    int32_t internal, first, last, has;
} eo_seg_count;

void eo_segment_count_nodes(const double* uv, int64_t n, int64_t q0, int64_t q1, double e,
                            int32_t stride, double w, double d, eo_seg_count* out)
{
    const int64_t qn = n - 2;
    double x[3] = {w, d, w - d};
    double em = e - uv[q0];
    eo_seg_count c = {0, 0, 0, 0};
    for (int64_t q = q0; q < q1; ++q) {
        const double ep = e - uv[q + 1];
        const double e10 = ep * 10.0;
        const double sg = fma(ep, em, ep + em);
        x[1] = fma(-e10, x[0], fma(-sg, x[2], x[1]));
        x[2] = x[0];
        x[0] = x[0] + x[1];
        em = ep;
        if ((q + 1) % stride == 0 || (q == qn - 1 && qn % stride != 0)) {
            const uint32_t sb = eo_signbit(x[0]);
            if (c.has)
                c.internal += (int32_t)(sb ^ (uint32_t)c.last);
            else
                c.first = (int32_t)sb;
            c.last = (int32_t)sb;
            c.has = 1;
        }
        if (((q - q0) & 127) == 127)
            eo_scale(x, 3, eo_max_exponent(x, 3));
    }
    *out = c;
}

/* C = A B (A the later segment), renormalised. */
static eo_mat eo_matmul(const eo_mat* a, const eo_mat* b)
{
    double c[4];
    c[0] = fma(a->m00, b->m00, a->m01 * b->m10);
    c[1] = fma(a->m00, b->m01, a->m01 * b->m11);
    c[2] = fma(a->m10, b->m00, a->m11 * b->m10);
    c[3] = fma(a->m10, b->m01, a->m11 * b->m11);
    const int32_t sh = eo_max_exponent(c, 4);
    eo_scale(c, 4, sh);
    eo_mat r = {c[0], c[1], c[2], c[3], a->e + b->e + sh};
    return r;
}

/* Far-end value by the binary-tree product of the segment matrices. */
void eo_tree_far_end(eo_mat* m, int32_t segments, double* f_mant, int32_t* f_exp)
{
    for (int32_t d = 1; d < segments; d *= 2)
        for (int32_t i = 0; i + d < segments; i += 2 * d)
            m[i] = eo_matmul(&m[i + d], &m[i]);
    *f_mant = fma(m[0].m00, 1.0, m[0].m01 * 1.0);
    *f_exp = m[0].e;
}

#define EO_MAX_SEGMENTS 1024

/* Far-end value (and, when nodes != NULL, the node count) of a segmented
 * difference-form shot at probe e, `lanes` threads per level. */
void eo_shoot_segmented(const double* uv, int64_t n, double e, int32_t stride, int32_t lanes,
                        double* f_mant, int32_t* f_exp, int32_t* nodes)
{
    const int32_t segs = eo_segment_count(n, lanes > EO_MAX_SEGMENTS ? EO_MAX_SEGMENTS : lanes);
    eo_mat m[EO_MAX_SEGMENTS], t[EO_MAX_SEGMENTS];
    double starts[2 * EO_MAX_SEGMENTS];
    for (int32_t s = 0; s < segs; ++s)
        eo_segment_matrix(uv, eo_segment_start(n, segs, s), eo_segment_start(n, segs, s + 1), e, &m[s]);
    memcpy(t, m, sizeof(eo_mat) * (size_t)segs);
    eo_tree_far_end(t, segs, f_mant, f_exp);
    if (nodes == NULL)
        return;
    double cm;
    int32_t ce;
    eo_chain(m, segs, &cm, &ce, starts);
    int32_t total = 0;
    uint32_t prev = 0;
    for (int32_t s = 0; s < segs; ++s) {
        eo_seg_count c;
        eo_segment_count_nodes(uv, n, eo_segment_start(n, segs, s), eo_segment_start(n, segs, s + 1), e,
                               stride, starts[2 * s], starts[2 * s + 1], &c);
        if (c.has) {
            total += c.internal + (int32_t)((uint32_t)c.first ^ prev);
            prev = (uint32_t)c.last;
        }
    }
    *nodes = total;
}

/* ---------------------------------------------------------------- refine */

/* Interpolation points of one refinement, sorted by energy, the upper end at
 * index ih (the lower end, when it has a value of the level's window, at
 * ih - 1): the bracket's scan points, then every shot. */
typedef struct {
    double e[EO_PTS], w[EO_PTS];
    int32_t k[EO_PTS];
    int32_t n, ih;
} eo_points;

/* Insert a shot at em (inside the bracket: at index ih in energy order).  A
 * full list first drops the outermost point farther from the bracket centre c
 * (the last one on a tie). */
static void eo_points_insert(eo_points* p, double c, double em, double w, int32_t k)
{
    if (p->n == EO_PTS) {
        if (c - p->e[0] > p->e[p->n - 1] - c) {
            for (int32_t i = 1; i < p->n; ++i) {
                p->e[i - 1] = p->e[i];
                p->w[i - 1] = p->w[i];
                p->k[i - 1] = p->k[i];
            }
            --p->ih;
        }
        --p->n;
    }
    for (int32_t i = p->n; i > p->ih; --i) {
        p->e[i] = p->e[i - 1];
        p->w[i] = p->w[i - 1];
        p->k[i] = p->k[i - 1];
    }
    p->e[p->ih] = em;
    p->w[p->ih] = w;
    p->k[p->ih] = k;
    ++p->n;
}

/* Root in (0, 1) of the polynomial through the points in x = (E - el)/width,
 * y = w 2^(k - max k) (Newton divided differences), by at most 8 Newton steps
 * from x0, stopping once a step is at most 2^-30.  0 when there are fewer
 * than 3 points, a derivative vanishes or a step leaves (0, 1). */
static int32_t eo_points_root(const eo_points* p, double el, double width, double x0, double* root)
{
    const int32_t n = p->n;
    if (n < 3)
        return 0;
    const double inv = 1.0 / width;
    int32_t kmax = p->k[0];
    for (int32_t i = 1; i < n; ++i)
        kmax = p->k[i] > kmax ? p->k[i] : kmax;
    double x[EO_PTS], c[EO_PTS];
    for (int32_t i = 0; i < n; ++i) {
        x[i] = (p->e[i] - el) * inv;
        c[i] = ldexp(p->w[i], p->k[i] - kmax);
    }
    for (int32_t j = 1; j < n; ++j)
        for (int32_t i = n - 1; i >= j; --i)
            c[i] = (c[i] - c[i - 1]) / (x[i] - x[i - j]);
    double t = x0;
    for (int32_t it = 0; it < 8; ++it) {
        double f = c[n - 1], d = 0.0;
        for (int32_t i = n - 2; i >= 0; --i) {
            const double u = t - x[i];
            d = fma(d, u, f);
            f = fma(f, u, c[i]);
        }
        if (!(d != 0.0))
            return 0;
        const double dt = f / d;
        t = t - dt;
        if (!(t > 0.0 && t < 1.0))
            return 0;
        if (fabs(dt) <= 0x1p-30)
            break;
    }
    *root = t;
    return 1;
}

/* Refine level v inside its bracket (the bracket ends carry the scan's
 * far-end values).  While the bracket holds more than level v (or its floor
 * has no shot) it is cut with counted shots at the count-interpolated point
 * (where level v would sit if the levels in it were evenly spaced, kept within
 * [1/8, 7/8] of the bracket); once it holds exactly level v
 * the probe is the root of the polynomial through the level's interpolation
 * points (the scan points of its window near the bracket and every shot, at
 * most EO_PTS, see eo_points_root) on F(E) = w_{N-1} 2^k (or on
 * G = F e^(-b (E - E_hi)) when the scan gives a model rate b), started from
 * the regula-falsi point of the ends (Anderson-Bjorck scaling), which is the
 * probe when the root search fails; each probe is kept at least tol/2 inside
 * the bracket, and its count follows from the sign of F (count parity = sign
 * bit of w_{N-1}).  Stops when the
 * width is <= tol = rel_tol*max(|E_lo|,|E_hi|) + abs_tol; the level is the
 * final midpoint.  Shots are difference-form: sequential (segmented == 0) or
 * segmented over a block of `lanes` threads (segmented != 0).  Returns the
 * number of shots, -1 when max_iter ran out. */
int32_t eo_refine_level(const double* uv, int64_t n, double beta, int32_t stride, int32_t v,
                        const eo_bracket* br, double rel_tol, double abs_tol, int32_t max_iter,
                        int32_t segmented, int32_t lanes, double* level)
{
    double el = br->e_lo, eh = br->e_hi;
    double wl = br->w_lo, wh = br->w_hi;
    int32_t kl = br->k_lo, kh = br->k_hi; /* binary exponents */
    int32_t cl = br->c_lo, ch = br->c_hi;
    int32_t lo_ok = br->lo_valid;
    int32_t lo_foreign = br->lo_foreign;
    int32_t side = 0;
    int32_t it = 0;
    /* With a model rate b, regula falsi runs on G = F e^(-b (E - E_hi0)),
     * nearly linear in E where F grows exponentially across the bracket;
     * sign(G) = sign(F), so counts and brackets are unchanged.  Without one
     * (refit set), the first isolated probe is the bracket's midpoint and the
     * rate is fitted to the three equally spaced values it gives. */
    double rate = br->rate;
    const double e0 = br->e_hi;
    int32_t refit = rate == 0.0 ? 1 : 0;
    eo_points pts;
    pts.n = br->n_pts;
    pts.ih = br->i_hi;
    for (int32_t i = 0; i < pts.n; ++i) {
        pts.e[i] = br->e_pt[i];
        pts.w[i] = br->w_pt[i];
        pts.k[i] = br->k_pt[i];
    }
    if (rate != 0.0) {
        eo_model_apply(rate, e0, el, &wl, &kl);
        eo_model_apply(rate, e0, eh, &wh, &kh);
        for (int32_t i = 0; i < pts.n; ++i)
            eo_model_apply(rate, e0, pts.e[i], &pts.w[i], &pts.k[i]);
    }
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
        const int32_t reshoot = isolated && lo_foreign;
        const int32_t mid = isolated && !reshoot && refit;
        if (reshoot) {
            /* The floor's value came from another band's window: shoot it
             * again in this level's window (its count stays the scan's). */
            em = el;
        } else if (mid) {
            em = el + 0.5 * width;
        } else if (isolated) {
            const double r = ldexp(wh / wl, kh - kl);
            const double s = 1.0 / (1.0 - r);
            double x;
            if (eo_points_root(&pts, el, width, fmin(fmax(s, 0.0), 1.0), &x))
                em = fma(x, width, el);
            else
                em = fma(s, width, el);
            const double dd = 0.5 * tol;
            if (!(em >= el + dd))
                em = el + dd;
            if (!(em <= eh - dd))
                em = eh - dd;
        } else {
            /* Several levels in the bracket (or no shot at its floor): probe
             * where level v would sit if the ch - cl levels were evenly
             * spaced, kept within [1/8, 7/8] of the bracket. */
            double fr = ((double)(v - cl) + 0.5) / (double)(ch - cl);
            fr = fmin(fmax(fr, 0.125), 0.875);
            em = fma(fr, width, el);
            side = 0;
        }
        double w;
        int32_t k, c;
        if (segmented) {
            eo_shoot_segmented(uv, n, beta * em, stride, lanes, &w, &k, isolated ? NULL : &c);
        } else {
            c = eo_shoot_diff(uv, n, beta * em, stride, &w, &k);
            k *= 480;
        }
        if (isolated)
            c = reshoot ? v : (((int32_t)eo_signbit(w) == (v & 1)) ? v : v + 1);
        if (mid) {
            /* The rate through (el, em, eh), em above the root when c > v;
             * the ends and the points take the transform from here on. */
            refit = 0;
            const double b = eo_model_rate(wl, kl, w, k, wh, kh, c > v ? 1 : 0, 0.5 * width);
            if (b != 0.0) {
                rate = b;
                eo_model_apply(rate, e0, el, &wl, &kl);
                eo_model_apply(rate, e0, eh, &wh, &kh);
                for (int32_t i = 0; i < pts.n; ++i)
                    eo_model_apply(rate, e0, pts.e[i], &pts.w[i], &pts.k[i]);
            }
        }
        if (rate != 0.0)
            eo_model_apply(rate, e0, em, &w, &k);
        ++it;
        eo_points_insert(&pts, el + 0.5 * width, em, w, k);
        if (c <= v)
            ++pts.ih;
        /* Anderson-Bjorck: when the same end is replaced twice in a row, the
         * kept end's value is scaled by m = 1 - F_new/F_replaced (0.5 when
         * m <= 0 or undefined). */
        if (c <= v) {
            if (isolated && side == -1) {
                const double m = 1.0 - ldexp(w / wl, k - kl);
                wh *= (m > 0.0) ? m : 0.5;
            }
            el = em; cl = c; wl = w; kl = k; lo_ok = 1; lo_foreign = 0;
            side = isolated ? -1 : 0;
        } else {
            if (isolated && side == 1) {
                const double m = 1.0 - ldexp(w / wh, k - kh);
                wl *= (m > 0.0) ? m : 0.5;
            }
            eh = em; ch = c; wh = w; kh = k;
            side = isolated ? 1 : 0;
        }
    }
    *level = el + 0.5 * (eh - el);
    return it;
}

/* Band windows and their far-end exponents: the whole grid in every band
 * when windows == 0. */
void eo_solve_windows(const double* uv, int64_t n, double beta, double e_lo, double e_hi,
                      int32_t windows, int64_t* win, int32_t* far_exp)
{
    if (windows)
        eo_band_windows(uv, n, beta, e_lo, e_hi, win);
    else
        eo_full_windows(n, win);
    for (int32_t b = 0; b < EO_BANDS; ++b) {
        double fc;
        eo_far_scale(win[2 * b + 1] - win[2 * b] + 1, &fc, &far_exp[b]);
    }
}

/* Scan + bracket + refine of levels [v_min, v_min + n_levels) for one
 * potential.  levels[] gets NaN for levels that are not below e_hi. */
void eo_solve(const double* pt, const double* uv, int64_t n, double beta, double e_lo, double e_hi,
              int64_t m, int32_t stride, int32_t v_min, int32_t n_levels, double rel_tol,
              double abs_tol, int32_t max_iter, int32_t segmented, int32_t lanes, int32_t windows,
              int32_t* nodes, double* w_last, int32_t* rescales, double* levels, int32_t* iterations)
{
    int64_t win[2 * EO_BANDS];
    int32_t far_exp[EO_BANDS];
    eo_solve_windows(uv, n, beta, e_lo, e_hi, windows, win, far_exp);
    eo_scan(pt, n, beta, e_lo, e_hi, m, stride, win, nodes, w_last, rescales);
    const double de = eo_trial_spacing(e_lo, e_hi, m);
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t l = 0; l < n_levels; ++l) {
        eo_bracket br;
        eo_bracket_level(nodes, w_last, rescales, m, e_lo, de, e_hi, v_min + l, win, far_exp, &br);
        if (!br.found) {
            levels[l] = NAN;
            iterations[l] = 0;
            continue;
        }
        const int64_t lo = win[2 * br.band], hi = win[2 * br.band + 1];
        iterations[l] = eo_refine_level(uv + lo, hi - lo + 1, beta, stride, v_min + l, &br, rel_tol,
                                        abs_tol, max_iter, segmented, lanes, &levels[l]);
    }
}

/* ---- eigenfunctions and the adjoint ------------------------------------
 *
 * Discrete problem on the trimmed grid (Dirichlet points 0 and n-1), with
 * f_i = beta (E - V_i) = e - uv_i:
 *   (1 + f_{i+1}) y_{i+1} = (2 - 10 f_i) y_i - (1 + f_{i-1}) y_{i-1}.
 * In matrix form A y + B diag(f) y = 0, A = tridiag(1, -2, 1),
 * B = tridiag(1, 10, 1); A and B commute, so diag(V) - B^-1 A / beta is
 * symmetric and dE/dV_i = y_i^2 / sum_j y_j^2 (Hellmann-Feynman for the
 * discrete eigenpair).
 *
 * Shot (difference form on y; f carried at full relative precision):
 *   fn = e - uv_{i+1};  r = 1 / (1 + fn);  s = fma(10, f_i, fn);
 *   c = fma(s, y_i, f_{i-1} y_{i-1});  d = (d - c) r;  y_{i+1} = y_i + d,
 * from (y_0, y_1, d_1) = (0, 1, 1).  Each shot visits its points in blocks of
 * 32 steps; after every full block, if max(|y_i|, |y_{i-1}|) >= 2^240, y, y_m
 * and d are multiplied by 2^-240 and the shot's rescale count k grows.
 *
 * Two shots meet at the matching point m (min V on the trimmed grid, clamped
 * to [1, n-3]): the left one from point 0 visits points 1..m (m steps), the
 * right one runs over the reversed table from point n-1 and visits points
 * n-2..m+1 (n-2-m steps).  Sums of y^2 over the visited points are rescaled
 * with the shot.  The right shot is scaled by the least-squares factor
 *   c = (yl_m yr_m + yl_{m+1} yr_{m+1}) / (yr_m^2 + yr_{m+1}^2)
 * and S = fma(c*c, S_right, S_left), a = 1/sqrt(S h): psi_i = y_i a 2^{240 (k_i - k_left)}
 * on the left, y_i (c a) 2^{240 (k_i - k_right)} on the right; sum psi^2 h = 1.
 */
typedef struct {
    double y, ym, d, f, fm;
    int32_t k;
} eo_yshot;

static void eo_ys_init(eo_yshot* s, double e, double u0, double u1)
{
    s->ym = 0.0;
    s->y = 1.0;
    s->d = 1.0;
    s->fm = e - u0;
    s->f = e - u1;
    s->k = 0;
}

static void eo_ys_step(eo_yshot* s, double u, double e)
{
    const double fn = e - u;
    const double r = 1.0 / (1.0 + fn);
    const double sg = fma(10.0, s->f, fn);
    const double c = fma(sg, s->y, s->fm * s->ym);
    s->d = (s->d - c) * r;
    s->ym = s->y;
    s->y = s->y + s->d;
    s->fm = s->f;
    s->f = fn;
}

static int32_t eo_ys_rescale(eo_yshot* s)
{
    const int big = fmax(fabs(s->y), fabs(s->ym)) >= 0x1p240;
    const int32_t sh = big ? -240 : 0;
    s->y = ldexp(s->y, sh);
    s->ym = ldexp(s->ym, sh);
    s->d = ldexp(s->d, sh);
    s->k += big;
    return sh;
}

int64_t eo_match_point(const double* v, int64_t first, int64_t n)
{
    int64_t imin = 0;
    for (int64_t i = 1; i < n; ++i)
        if (v[first + i] < v[first + imin])
            imin = i;
    if (imin < 1)
        imin = 1;
    if (imin > n - 3)
        imin = n - 3;
    return imin;
}

/* One inward shot over `steps` steps of table u (visiting u's points
 * 1..steps); returns the rescaled sum of y^2 over the visited points.  When
 * psi is given, psi[i0 + dir*j] = y_{j+1} a 2^{240 (k - kf)} for the j-th
 * visited point. */
static double eo_inward(const double* u, int64_t steps, double e, eo_yshot* s, double* psi,
                        int64_t i0, int64_t dir, double a, int32_t kf)
{
    double sum = 0.0;
    double mult = ldexp(a, -240 * kf);
    eo_ys_init(s, e, u[0], u[1]);
    for (int64_t j = 0; j < steps; ++j) {
        if (psi)
            psi[i0 + dir * j] = s->y * mult;
        sum = fma(s->y, s->y, sum);
        eo_ys_step(s, u[2 + j], e);
        if ((j + 1) % 32 == 0) {
            const int32_t sh = eo_ys_rescale(s);
            sum = ldexp(sum, 2 * sh);
            mult = ldexp(a, 240 * (s->k - kf));
        }
    }
    return sum;
}

static double* eo_reversed(const double* uv, int64_t n)
{
    double* r = (double*)malloc((size_t)n * sizeof(double));
    for (int64_t i = 0; i < n; ++i)
        r[i] = uv[n - 1 - i];
    return r;
}

/* Normalisation of the eigenfunction at e = beta E: a[0..1] = (a_left,
 * a_right), k[0..1] = (k_left, k_right). */
void eo_psi_norm(const double* uv, int64_t n, double e, int64_t m, double h, double* a, int32_t* k)
{
    double* ur = eo_reversed(uv, n);
    eo_yshot L, R;
    const double sl = eo_inward(uv, m, e, &L, NULL, 0, 0, 0.0, 0);
    const double sr = eo_inward(ur, n - 2 - m, e, &R, NULL, 0, 0, 0.0, 0);
    free(ur);
    const double yl0 = L.ym, yl1 = L.y;
    const double yr1 = R.ym, yr0 = R.y;
    const double num = fma(yl0, yr0, yl1 * yr1);
    const double den = fma(yr0, yr0, yr1 * yr1);
    const double c = num / den;
    const double s = fma(c * c, sr, sl);
    const double an = 1.0 / sqrt(s * h);
    a[0] = an;
    a[1] = c * an;
    k[0] = L.k;
    k[1] = R.k;
}

/* Normalised eigenfunction on the trimmed grid (psi[0] = psi[n-1] = 0). */
void eo_psi(const double* uv, int64_t n, double e, int64_t m, double h, double* psi)
{
    double a[2];
    int32_t k[2];
    eo_psi_norm(uv, n, e, m, h, a, k);
    double* ur = eo_reversed(uv, n);
    eo_yshot L, R;
    psi[0] = 0.0;
    psi[n - 1] = 0.0;
    (void)eo_inward(uv, m, e, &L, psi, 1, 1, a[0], k[0]);
    (void)eo_inward(ur, n - 2 - m, e, &R, psi, n - 2, -1, a[1], k[1]);
    free(ur);
}

/* Adjoint on the trimmed grid: grad[i] = sum_l g[l] psi_l(i)^2 h over the
 * levels with finite energy (levels in e-units: e_l = beta E_l). */
void eo_level_vjp(const double* uv, int64_t n, const double* e, const double* g, int32_t count,
                  int64_t m, double h, double* grad)
{
    double* psi = (double*)malloc((size_t)n * sizeof(double));
    for (int64_t i = 0; i < n; ++i)
        grad[i] = 0.0;
    for (int32_t l = 0; l < count; ++l) {
        if (!isfinite(e[l]))
            continue;
        eo_psi(uv, n, e[l], m, h, psi);
        for (int64_t i = 0; i < n; ++i)
            grad[i] = fma(g[l] * h * psi[i], psi[i], grad[i]);
    }
    free(psi);
}

/* Test hooks for the model's elementary functions. */
double eo_model_exp_value(double y)
{
    double m;
    int32_t n;
    eo_model_exp(y, &m, &n);
    return ldexp(m, n);
}

double eo_model_log_value(double x) { return eo_model_log(x); }

double eo_model_sqrt_value(double d) { return eo_model_sqrt(d); }
