// records.hpp -- host/device records shared by the kernels and the C ABI.
//
// One PotentialRecord per potential of a batch replaces the per-slot buffer
// set of the reference (staging + 5 device-only + 1 output buffer per
// group_size slot, task_configurator/algorithm_config.hpp:173-198): all
// potentials live in single CSR-style device arrays and the record holds the
// offsets into them.
#pragma once

#include <cstdint>

namespace epseon::hip {

// Rescale thresholds of the recurrences: the refinement shots multiply the
// last values by 2^-480 when |w| reaches 2^480; the scan shots (whose values
// only shrink) multiply them by 2^480 when they fall below 2^-480.  Both are
// exact and counted.  The scan's threshold sits at 2^-240 so that values
// still start from at least 2^-240 when the test runs only every second
// super-block.
constexpr double kRescaleAt = 0x1p480;
constexpr double kRescaleBy = 0x1p-480;
constexpr double kScanRescaleBelow = 0x1p-240;
constexpr int kRescaleLog2 = 480;
constexpr int kMaxStride = 32;
// hbar^2 / (2 amu A^2) in cm^-1 (CODATA 2018 h, c and atomic mass unit).
constexpr double kHbar2Over2Amu = 16.857629191640175;

struct alignas(16) PotentialRecord {
    int64_t v_offset;       // grid point 0 of this potential in the V and u tables
    int64_t ut_offset;      // first point of the trimmed grid (v_offset + first_point)
    int64_t energy_offset;  // first trial energy in the scan arrays
    int64_t level_offset;   // first level in the level arrays
    int64_t first_point;    // index of the left Dirichlet point on the full grid
    int32_t n_points;       // points of the trimmed grid (both Dirichlet points included)
    int32_t n_energies;     // M
    int32_t level_min;      // first level v
    int32_t level_count;    // levels requested
    int32_t stride;         // node-sampling stride B (power of two <= 32)
    int32_t flags;          // EPSEON_POT_*
    double beta;            // h^2 mu / (12 K)
    double e_lo;            // window floor (min V)
    double de;              // trial spacing (e_hi - e_lo)/(M + 1)
    double e_hi;            // window top
    double abs_tol;         // absolute refinement tolerance
    double h;               // grid step
    int64_t user_offset;    // point 0 of this potential in the caller's V layout
    int32_t match;          // matching point of the two-sided eigenfunction shots
                            // (min V, relative to first_point, in [1, n_points - 3])
    int32_t full_points;    // points of the untrimmed grid
    int64_t psi_offset;     // first eigenfunction value (sum over earlier potentials of
                            // level_count * full_points)
    double far_scale;       // 2^frac((n_points - 2) log2 12): scan far-end values are
                            // v_{N-1} * far_scale, with w_{N-1} = that * 2^(far_exp - 480 r)
    int32_t far_exp;        // floor((n_points - 2) log2 12)
    int32_t seg_row;        // first row of this potential's segment table (rows of 64 values)
    int64_t psi_ckpt;       // first eigenfunction-shot checkpoint slot of this potential
    int32_t psi_seg_len;    // steps per segment of the eigenfunction shots (a multiple of 32)
    int32_t psi_segs_left;  // segments of the left shot (match steps)
    int32_t psi_segs_right; // segments of the right shot (n_points - 2 - match steps)
    int32_t psi_pad;
};
static_assert(sizeof(PotentialRecord) == 176, "record layout");

// Segments of the 64-lane segmented refinement of a potential with n points:
// the q = n - 2 steps split into min(64, max(1, q / 32)) segments, segment s
// covering steps [s q / segs, (s + 1) q / segs).
constexpr int32_t seg_count(const int32_t n_points)
{
    const int32_t q = n_points - 2;
    const int32_t s = q / 32;
    return s < 1 ? 1 : (s > 64 ? 64 : s);
}

// Rows of a potential's segment table (row r of column s holds uv at point
// q0_s + r): the longest segment plus the row of its start point and the 32
// rows its walk reads ahead; none for a potential without levels to refine.
constexpr int32_t seg_rows(const int32_t n_points, const int32_t flags, const int32_t level_count)
{
    if (flags != 0 || level_count <= 0)
        return 0;
    const int32_t q = n_points - 2;
    const int32_t segs = seg_count(n_points);
    return (q + segs - 1) / segs + 33;
}

// Segment length of the eigenfunction shots of a potential with n points: a
// multiple of 32 (the rescale block), at least 512 steps, and long enough
// that the n - 2 steps make at most about 256 segments.
constexpr int32_t psi_seg_len(const int32_t n_points)
{
    const int32_t q = n_points - 2;
    const int32_t blocks = (q + 32 * 256 - 1) / (32 * 256);
    return 32 * (blocks < 16 ? 16 : blocks);
}

// State of an eigenfunction shot at the start of one of its segments (after
// j = segment * psi_seg_len steps): y_{j+1}, y_j and their difference as
// mantissas of the binary exponent k (the true values are these times 2^k).
// f and f_prev are recomputed from the table.
struct alignas(32) PsiCheckpoint {
    double y, ym, d;
    int32_t k, pad;
};

// Eigenfunction shot through one segment, from the two basis states (y,
// y_prev, d) = (1, 1, 0) (column a) and (0, -1, 1) (column b): the end state
// of a shot entering with (y, d) is y * a + d * b, times 2^e; s_aa, s_ab,
// s_bb are the Gram sums of the columns' y values over the segment's visited
// points, times 2^(2e).
struct PsiSegMatrix {
    double ay, aym, ad, by, bym, bd;
    double saa, sab, sbb;
    int32_t e, pad;
};

// One wavefront (or workgroup) of the segmented eigenfunction passes: up to
// 64 levels from `first` of `potential`, one segment of one side (0: the
// shot from the left wall, 1: from the right wall).
struct PsiTask {
    int32_t potential;
    int32_t first;
    int32_t side;
    int32_t segment;
};

// Normalisation of one level's eigenfunction, from two inward shots that meet
// at PotentialRecord::match: psi_i = y_i * ldexp(a, k - k_final) with
// (a, k_final) = (a_left, k_left) left of the match and (a_right, k_right)
// right of it, y_i * 2^k being the shot's value at point i (k in bits).
// Zero for levels that did not converge.
struct LevelNorm {
    double a_left, a_right;
    int32_t k_left, k_right;
};

// Rescaling of the eigenfunction shots: 2^240 threshold, checked after every
// 32 steps, keeps y^2 sums far from overflow.
constexpr double kPsiRescaleAt = 0x1p240;
constexpr int kPsiRescaleLog2 = 240;

// Per-potential facts of a V table that loading needs: min V, the first index
// where it occurs, the Dirichlet points of the trimmed grid (last point left of
// imin / first point right of it with V >= min V + 1/(2 beta), else the grid
// ends) and the first non-finite index (n_points if none).
struct PotentialStats {
    double vmin;
    int64_t imin, first, last, bad;
};

// Shot windows.  Past its outermost turning points a shot only grows (from
// the left wall: only decays towards the well) by at least acosh(1 + x/2)
// per step, x_i = 12 (uv_i - e) > 0, bounded below exactly by
// g(x) = 2^(floor((k-1)/2) - 1) for x = m 2^k, m in [1/2, 1).  A window's
// Dirichlet points sit where the sum of g from the outermost allowed point
// (uv_i <= e) first reaches kWindowDepth, or at the grid ends; moving a wall
// from there to the grid end changes the levels by an e^-30-order fraction
// of a spacing.  The sum is kept in 2^-40 fixed point (exact, any order).
// The energy window [e_lo, e_hi] splits into kBands bands with tops
// E_b = fma(b + 1, (e_hi - e_lo) / kBands, e_lo); a band's window is the
// window at its top.  The scan shoots a group of 64 trial energies in the
// window of the band holding the group's highest energy; a level is refined
// in the window of its bracket's upper scan point.
constexpr int kBands = 16;
constexpr int kWindowDepth = 15;

// One band window of a potential: Dirichlet points lo, hi (on the trimmed
// grid) and the far-end scale of its n = hi - lo + 1 point scan shots
// (12^(n-2) = far_scale 2^far_exp); seg_row is the first row of its segment
// table.
struct BandWindow {
    int32_t lo, hi;
    int32_t far_exp;
    int32_t seg_row;
    double far_scale;
};
static_assert(sizeof(BandWindow) == 24, "window layout");

// One segment table to build: the n-point window from point lo of a
// potential's trimmed grid, stored from row seg_row.
struct TableJob {
    int32_t potential;
    int32_t lo, n;
    int32_t seg_row;
};

// Bracket of one level from the scan arrays (written by the bracket pass,
// read by the refinement): ends el < eh with far-end values F = w 2^k (of
// the level's band window) and node counts c, lo_ok = 0 when el is the
// window floor (no shot there), found = 0 when no trial energy has more
// than v nodes, rate = b of the refinement model (0: none), band = the
// upper end's band, lo_foreign = 1 when the lower end's value comes from a
// different band window.
struct Bracket {
    double el, eh, wl, wh, rate;
    int32_t kl, kh, cl, ch, lo_ok, found;
    int32_t band;
    int32_t lo_foreign;
};
static_assert(sizeof(Bracket) == 72, "bracket layout");

// Interpolation points of one level's refinement (written by the bracket
// pass, updated by the refinement): at most kPts (energy, F = w 2^k) pairs of
// the level's band window sorted by energy, the upper bracket end at index
// ih (the lower end, when its value is of that window, at ih - 1).
constexpr int32_t kPts = 12;
struct RefinePoints {
    double e[kPts];
    double w[kPts];
    int32_t k[kPts];
    int32_t n, ih;
};
static_assert(sizeof(RefinePoints) == 248, "points layout");

// Work item of one wavefront: potential index and the first trial energy
// (scan) or first level (refine) of its 64 lanes.
struct WaveTask {
    int32_t potential;
    int32_t first;
};

}  // namespace epseon::hip
