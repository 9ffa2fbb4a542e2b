// eigenfunction_kernels.hip -- gfx950 kernels for the normalised
// eigenfunctions of the refined levels and for the adjoint of the levels with
// respect to the tabulated potential.
//
// The reference has no eigenfunction or gradient code (its only algorithm,
// VibwaAlgorithm<FP>::run, cpp/gpu/include/epseon/gpu/algorithms/vibwa.hpp:
// 605-637, dispatches nothing); these kernels implement what its output
// buffer layout (task_configurator/algorithm_config.hpp:173-198) leaves room
// for.
//
// Discrete problem.  On the trimmed grid (Dirichlet points 0 and n-1) the
// Numerov recurrence with f_i = beta (E - V_i),
//   (1 + f_{i+1}) y_{i+1} = (2 - 10 f_i) y_i - (1 + f_{i-1}) y_{i-1},
// is A y + B diag(f) y = 0 with A = tridiag(1, -2, 1), B = tridiag(1, 10, 1).
// A and B commute, so diag(V) - B^-1 A / beta is symmetric with eigenpairs
// (E_v, y_v): the Hellmann-Feynman derivative of a level is
//   dE_v / dV_i = y_{v,i}^2 / sum_j y_{v,j}^2 = psi_{v,i}^2 h,
// psi normalised to sum_i psi_i^2 h = 1.
//
// Shots.  y is integrated in difference form (f carried at full relative
// precision, as in the refinement):
//   d_{i+1} = (d_i - (f_{i+1} + 10 f_i) y_i - f_{i-1} y_{i-1}) / (1 + f_{i+1}),
//   y_{i+1} = y_i + d_{i+1},
// once from each wall inward, both starting (y_0, y_1) = (0, 1), to the
// matching point m = argmin V (inside every level's classically allowed
// region, so both shots are stable and every lane of a wavefront walks the
// same points).  Each shot rescales by 2^-240 after every full 32 steps in
// which |y| reached 2^240.  The right shot is scaled by the least-squares
// factor c that fits it to the left one at points m and m+1.
//
// Segments.  Each side's shot is cut into segments of psi_seg_len steps.
// The normalisation runs every segment at once from two basis states
// (psi_segment_kernel: the segment's transfer map and the Gram sums of its
// columns), then chains the maps per level from the wall (normalize_kernel),
// which gives each segment's start state (a checkpoint), sum y^2 and the
// match.  The eigenfunction and adjoint passes resume every segment from its
// checkpoint in parallel.
//
// Lanes are levels of one potential; the potential table is wave-uniform and
// read with scalar loads, as in the scan.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "records.hpp"

namespace epseon::hip {

namespace {

__device__ __forceinline__ double opaque(double x)
{
    asm volatile("" : "+v"(x));
    return x;
}

struct YShot {
    double y, ym, d, f, fm;
    int32_t k;

    __device__ __forceinline__ void init(const double e, const double u0, const double u1)
    {
        ym = 0.0;
        y = 1.0;
        d = 1.0;
        fm = e - u0;
        f = e - u1;
        k = 0;
    }

    __device__ __forceinline__ void step(const double u, const double e, const double c10)
    {
        const double fn = e - u;
        const double r = 1.0 / (1.0 + fn);
        const double s = __builtin_fma(c10, f, fn);
        const double c = __builtin_fma(s, y, fm * ym);
        d = (d - c) * r;
        ym = y;
        y = y + d;
        fm = f;
        f = fn;
    }

    // Continue a shot from its checkpoint after j steps: u_prev and u_cur
    // are points j and j + 1 of the side's table.
    __device__ __forceinline__ void resume(const PsiCheckpoint& c, const double e, const double u_prev,
                                           const double u_cur)
    {
        y = c.y;
        ym = c.ym;
        d = c.d;
        k = c.k;
        fm = e - u_prev;
        f = e - u_cur;
    }

    // A level without a refined energy: every value stays zero.
    __device__ __forceinline__ void clear(const double e, const double u_prev, const double u_cur)
    {
        y = ym = d = 0.0;
        k = 0;
        fm = e - u_prev;
        f = e - u_cur;
    }

    // Returns the applied binary shift (0 or -240).
    __device__ __forceinline__ int32_t rescale()
    {
        const bool big = fmax(fabs(y), fabs(ym)) >= kPsiRescaleAt;
        const int32_t sh = big ? -kPsiRescaleLog2 : 0;
        y = __builtin_ldexp(y, sh);
        ym = __builtin_ldexp(ym, sh);
        d = __builtin_ldexp(d, sh);
        k += big ? kPsiRescaleLog2 : 0;
        return sh;
    }
};

__device__ __forceinline__ void load16(double (&buf)[16], const double* __restrict__ p)
{
#pragma unroll
    for (int q = 0; q < 16; ++q)
        buf[q] = p[q];
}

// Visit `steps` grid values p[0..steps): body.step(u, q) per value (q = index
// inside the current 32-step block), body.block_end(count, full) after each
// full block of 32 and after the tail.  Two 16-value buffers, each loaded one
// half-block ahead.  Reads up to 47 values past `steps` (tables are padded).
template <class Body>
__device__ __forceinline__ void psi_walk(const double* __restrict__ p, const int32_t steps, Body& body)
{
    const int32_t nsb = steps >> 5;
    double a[16], b[16];
    load16(a, p);
    for (int32_t sb = 0; sb < nsb; ++sb) {
        load16(b, p + 16);
#pragma unroll
        for (int q = 0; q < 16; ++q)
            body.step(a[q], q);
        load16(a, p + 32);
#pragma unroll
        for (int q = 0; q < 16; ++q)
            body.step(b[q], 16 + q);
        body.block_end(32, true);
        p += 32;
    }
    const int32_t rem = steps - (nsb << 5);
    if (rem > 0) {
        for (int32_t q = 0; q < rem; ++q)
            body.step(p[q], q);
        body.block_end(rem, false);
    }
}

// frexp exponent of m (0 for m = 0).
__device__ __forceinline__ int32_t binary_exponent(const double m)
{
    int e = 0;
    (void)__builtin_frexp(m, &e);
    return e;
}

// Pass 1, per segment: the shot's two basis solutions through the segment,
// started from (y, y_prev, d) = (1, 1, 0) and (0, -1, 1), so that a shot
// entering with (y, d) leaves with y times column a plus d times column b;
// and the Gram sums of their values at the visited points, from which the
// segment's sum of y^2 follows as a quadratic form in (y, d).  All values
// are renormalised by a power of two after every 32 steps and at the end;
// e is the accumulated exponent (the Gram sums carry it twice).
struct MatBody {
    double ay, aym, ad, by, bym, bd;
    double saa, sab, sbb;
    double f, fm, e, c10;
    int32_t ex;
    __device__ __forceinline__ void step(const double u, int)
    {
        saa = __builtin_fma(ay, ay, saa);
        sab = __builtin_fma(ay, by, sab);
        sbb = __builtin_fma(by, by, sbb);
        const double fn = e - u;
        const double r = 1.0 / (1.0 + fn);
        const double sg = __builtin_fma(c10, f, fn);
        const double ca = __builtin_fma(sg, ay, fm * aym);
        const double cb = __builtin_fma(sg, by, fm * bym);
        ad = (ad - ca) * r;
        bd = (bd - cb) * r;
        aym = ay;
        bym = by;
        ay = ay + ad;
        by = by + bd;
        fm = f;
        f = fn;
    }
    __device__ __forceinline__ void renormalize()
    {
        const double mx = fmax(fmax(fmax(fabs(ay), fabs(aym)), fmax(fabs(ad), fabs(by))),
                               fmax(fabs(bym), fabs(bd)));
        const int32_t x = binary_exponent(mx);
        ay = __builtin_ldexp(ay, -x);
        aym = __builtin_ldexp(aym, -x);
        ad = __builtin_ldexp(ad, -x);
        by = __builtin_ldexp(by, -x);
        bym = __builtin_ldexp(bym, -x);
        bd = __builtin_ldexp(bd, -x);
        saa = __builtin_ldexp(saa, -2 * x);
        sab = __builtin_ldexp(sab, -2 * x);
        sbb = __builtin_ldexp(sbb, -2 * x);
        ex += x;
    }
    __device__ __forceinline__ void block_end(int32_t, const bool full)
    {
        if (full)
            renormalize();
    }
};

// Pass 2, eigenfunction output: psi_i = y_i * mult written to row[i0 + dir*j].
struct PsiBody {
    YShot s;
    double e, c10, a, mult;
    int32_t kf, i0, dir, base;
    // Staging: lane l's 32 values of a block go to tile[l], then the wave
    // writes them out two rows per store (32 contiguous values each).
    double (*tile)[33];
    double* rows;  // row of the wave's first level, at the trimmed window's first point
    int64_t stride;
    int32_t nrows;
    __device__ __forceinline__ void step(const double u, const int q)
    {
        tile[threadIdx.x][q] = s.y * mult;
        s.step(u, e, c10);
    }
    __device__ __forceinline__ void block_end(const int32_t count, const bool full)
    {
        __syncthreads();
        const int32_t c = static_cast<int32_t>(threadIdx.x & 31);
        const int32_t h = static_cast<int32_t>(threadIdx.x >> 5);
        for (int32_t r = h; r < nrows; r += 2)
            if (c < count)
                rows[r * stride + i0 + dir * (base + c)] = tile[r][c];
        __syncthreads();
        base += count;
        if (full) {
            (void)s.rescale();
            mult = __builtin_ldexp(a, s.k - kf);
        }
    }
};

// Pass 2, adjoint: lane contribution g h psi_i^2 = G y_i^2 staged in LDS and
// summed over the block's levels per point, 32 points at a time.
template <int W>
struct VjpBody {
    YShot s;
    double e, c10, g0, G;
    int32_t kf, i0, dir, base;
    bool accumulate;
    double (*sc)[32][65];  // [W][32][64 lanes + 1 pad]
    double (*sw)[32];      // [W][32]
    double* out;

    __device__ __forceinline__ void step(const double u, const int q)
    {
        const int32_t wv = static_cast<int32_t>(threadIdx.x >> 6);
        const int32_t ln = static_cast<int32_t>(threadIdx.x & 63);
        sc[wv][q][ln] = (G * s.y) * s.y;
        s.step(u, e, c10);
    }

    __device__ __forceinline__ void block_end(const int32_t count, const bool full)
    {
        const int32_t wv = static_cast<int32_t>(threadIdx.x >> 6);
        const int32_t ln = static_cast<int32_t>(threadIdx.x & 63);
        __syncthreads();
        const int32_t j = ln & 31;
        const int32_t h = (ln >> 5) * 32;
        double acc = 0.0;
#pragma unroll
        for (int t = 0; t < 32; ++t)
            acc += sc[wv][j][h + t];
        acc += __shfl_xor(acc, 32);
        if (ln < 32)
            sw[wv][j] = acc;
        __syncthreads();
        const int32_t t = static_cast<int32_t>(threadIdx.x);
        if (t < count) {
            double tot = sw[0][t];
#pragma unroll
            for (int w = 1; w < W; ++w)
                tot += sw[w][t];
            double* o = out + (i0 + dir * (base + t));
            *o = accumulate ? *o + tot : tot;
        }
        base += count;
        if (full) {
            (void)s.rescale();
            G = __builtin_ldexp(g0, 2 * (s.k - kf));
        }
    }
};

__device__ __forceinline__ bool level_active(const int32_t* __restrict__ status, const int64_t o)
{
    return status[o] == 0;
}

}  // namespace

// Shot segment of a PsiTask: the side's table, the first step j0 and the
// number of steps.
struct PsiSegment {
    const double* p;
    int32_t j0, steps;
};

__device__ __forceinline__ PsiSegment psi_segment(const PotentialRecord& P, const double* __restrict__ uv,
                                                  const double* __restrict__ uvr, const PsiTask& t)
{
    const int32_t total = t.side == 0 ? P.match : P.n_points - 2 - P.match;
    const int32_t j0 = t.segment * P.psi_seg_len;
    return PsiSegment{(t.side == 0 ? uv : uvr) + P.ut_offset, j0, min(P.psi_seg_len, total - j0)};
}

__device__ __forceinline__ int64_t psi_slot(const PotentialRecord& P, const int32_t side, const int32_t segment,
                                            const int32_t l)
{
    return P.psi_ckpt + static_cast<int64_t>((side == 0 ? 0 : P.psi_segs_left) + segment) * P.level_count + l;
}

// Start a segment's shot: from the wall (segment 0), from the normalisation
// pass's checkpoint, or all zero for a level that did not converge.
__device__ __forceinline__ void psi_start(YShot& s, const PotentialRecord& P, const PsiTask& t,
                                          const PsiSegment& g, const PsiCheckpoint* __restrict__ ckpt,
                                          const int32_t l, const bool active, const double e)
{
    if (!active)
        s.clear(e, g.p[g.j0], g.p[g.j0 + 1]);
    else if (t.segment == 0)
        s.init(e, g.p[0], g.p[1]);
    else
        s.resume(ckpt[psi_slot(P, t.side, t.segment, l)], e, g.p[g.j0], g.p[g.j0 + 1]);
}

// Pass 1a: one wavefront per eigenfunction task (potential, 64 levels, side,
// segment): each converged level's segment matrix and Gram sums.
__global__ __launch_bounds__(64) void psi_segment_kernel(
    const double* __restrict__ uv, const double* __restrict__ uvr,
    const PotentialRecord* __restrict__ pots, const PsiTask* __restrict__ tasks,
    const double* __restrict__ levels, const int32_t* __restrict__ status,
    PsiSegMatrix* __restrict__ mats)
{
    const PsiTask task = tasks[blockIdx.x];
    const PotentialRecord& P = pots[task.potential];
    const int32_t l = task.first + static_cast<int32_t>(threadIdx.x);
    if (P.flags != 0 || l >= P.level_count)
        return;
    const int64_t o = P.level_offset + l;
    if (!level_active(status, o))
        return;
    const double e = P.beta * levels[o];
    const PsiSegment g = psi_segment(P, uv, uvr, task);
    MatBody B{};
    B.ay = 1.0;
    B.aym = 1.0;
    B.ad = 0.0;
    B.by = 0.0;
    B.bym = -1.0;
    B.bd = 1.0;
    B.e = e;
    B.c10 = opaque(10.0);
    B.fm = e - g.p[g.j0];
    B.f = e - g.p[g.j0 + 1];
    psi_walk(g.p + 2 + g.j0, g.steps, B);
    B.renormalize();
    mats[psi_slot(P, task.side, task.segment, l)] =
        PsiSegMatrix{B.ay, B.aym, B.ad, B.by, B.bym, B.bd, B.saa, B.sab, B.sbb, B.ex, 0};
}

// Pass 1b: one lane per level (wavefront = up to 64 levels of one potential).
// Chains the segment matrices of each side from its wall, (y, y_prev, d) =
// (1, 0, 1), renormalising by a power of two after each product; writes the
// checkpoint (state and binary exponent) at each later segment start, sums
// y^2 as the segments' quadratic forms, matches the sides at m and writes the
// level's LevelNorm (exponents in bits).
__global__ __launch_bounds__(64) void normalize_kernel(
    const PotentialRecord* __restrict__ pots, const WaveTask* __restrict__ tasks, const int32_t n_tasks,
    const int32_t* __restrict__ status, const PsiSegMatrix* __restrict__ mats,
    PsiCheckpoint* __restrict__ ckpt, LevelNorm* __restrict__ norms)
{
    const int32_t tid = static_cast<int32_t>(blockIdx.x);
    if (tid >= n_tasks)
        return;
    const WaveTask task = tasks[tid];
    const PotentialRecord& P = pots[task.potential];
    const int32_t l = task.first + static_cast<int32_t>(threadIdx.x);
    if (l >= P.level_count)
        return;
    const int64_t o = P.level_offset + l;
    if (P.flags != 0 || !level_active(status, o)) {
        norms[o] = LevelNorm{0.0, 0.0, 0, 0};
        return;
    }
    double ye[2], yme[2], srel[2];
    int32_t ke[2];
    for (int32_t side = 0; side < 2; ++side) {
        const int32_t segs = side == 0 ? P.psi_segs_left : P.psi_segs_right;
        double y = 1.0, ym = 0.0, d = 1.0, sr = 0.0;
        int32_t k = 0;
        for (int32_t sg = 0; sg < segs; ++sg) {
            const int64_t slot = psi_slot(P, side, sg, l);
            if (sg > 0)
                ckpt[slot] = PsiCheckpoint{y, ym, d, k, 0};
            const PsiSegMatrix M = mats[slot];
            const double q = __builtin_fma(y * y, M.saa, __builtin_fma(2.0 * y * d, M.sab, d * d * M.sbb));
            const double yn = __builtin_fma(M.ay, y, M.by * d);
            const double ymn = __builtin_fma(M.aym, y, M.bym * d);
            const double dn = __builtin_fma(M.ad, y, M.bd * d);
            const double mx = fmax(fmax(fabs(yn), fabs(ymn)), fabs(dn));
            const int32_t x = binary_exponent(mx);
            y = __builtin_ldexp(yn, -x);
            ym = __builtin_ldexp(ymn, -x);
            d = __builtin_ldexp(dn, -x);
            // y^2 sum relative to 2^(2k) of the new exponent k + e + x.
            sr = __builtin_ldexp(sr, -2 * (M.e + x)) + __builtin_ldexp(q, -2 * x);
            k += M.e + x;
        }
        ye[side] = y;
        yme[side] = ym;
        srel[side] = sr;
        ke[side] = k;
    }
    // Left: (y_m, y_{m+1}) = (ym, y); right (reversed): (y_{m+1}, y_m) = (ym, y);
    // the right shot scaled by c = cm 2^(k_left - k_right).
    const double yl0 = yme[0], yl1 = ye[0];
    const double yr1 = yme[1], yr0 = ye[1];
    const double num = __builtin_fma(yl0, yr0, yl1 * yr1);
    const double den = __builtin_fma(yr0, yr0, yr1 * yr1);
    const double cm = num / den;
    const double s = __builtin_fma(cm * cm, srel[1], srel[0]);
    const double a = 1.0 / sqrt(s * P.h);
    norms[o] = LevelNorm{a, cm * a, ke[0], ke[1]};
}

// Eigenfunctions: for each task (potential, first level, side, segment), the
// segment's points of rows first.. of that potential (full_points values per
// row at out[psi_offset - base + l * full_points]), re-run from the
// normalisation pass's checkpoints, staged in LDS and stored 32 contiguous
// values at a time; the segment-0 tasks also zero the row outside the open
// trimmed window on their side (the whole row of an unresolved potential,
// whose only task is side 0, segment 0).  Zero for levels that did not
// converge.  One lane per level.
__global__ __launch_bounds__(64) void eigenfunction_kernel(
    const double* __restrict__ uv, const double* __restrict__ uvr,
    const PotentialRecord* __restrict__ pots, const PsiTask* __restrict__ tasks,
    const double* __restrict__ levels, const int32_t* __restrict__ status,
    const LevelNorm* __restrict__ norms, const PsiCheckpoint* __restrict__ ckpt, const int64_t base,
    double* __restrict__ out)
{
    __shared__ double s_tile[64][33];
    const PsiTask task = tasks[blockIdx.x];
    const PotentialRecord& P = pots[task.potential];
    const int32_t l = task.first + static_cast<int32_t>(threadIdx.x);
    const int32_t nrows = min(64, P.level_count - task.first);
    const bool valid = l < P.level_count;
    const int64_t N = P.full_points;
    double* row0 = out + (P.psi_offset - base) + static_cast<int64_t>(task.first) * N;
    double* row = row0 + static_cast<int64_t>(threadIdx.x) * N;
    const int64_t first = P.first_point;
    const int32_t n = P.n_points;
    const bool ok = P.flags == 0;
    if (task.segment == 0 && valid) {
        if (!ok) {
            for (int64_t i = 0; i < N; ++i)
                row[i] = 0.0;
        } else if (task.side == 0) {
            for (int64_t i = 0; i <= first; ++i)
                row[i] = 0.0;
        } else {
            for (int64_t i = first + n - 1; i < N; ++i)
                row[i] = 0.0;
        }
    }
    if (!ok)
        return;
    const int64_t o = P.level_offset + l;
    const bool active = valid && level_active(status, o);
    const LevelNorm nm = active ? norms[o] : LevelNorm{0.0, 0.0, 0, 0};
    const double e = P.beta * (active ? levels[o] : P.e_lo);
    const PsiSegment g = psi_segment(P, uv, uvr, task);

    PsiBody B{};
    B.e = e;
    B.c10 = opaque(10.0);
    B.a = task.side == 0 ? nm.a_left : nm.a_right;
    B.kf = task.side == 0 ? nm.k_left : nm.k_right;
    B.i0 = task.side == 0 ? 1 : n - 2;
    B.dir = task.side == 0 ? 1 : -1;
    B.base = g.j0;
    B.tile = s_tile;
    B.rows = row0 + first;
    B.stride = N;
    B.nrows = nrows;
    psi_start(B.s, P, task, g, ckpt, l, active, e);
    B.mult = __builtin_ldexp(B.a, B.s.k - B.kf);
    psi_walk(g.p + 2 + g.j0, g.steps, B);
}

// Adjoint: grad_v[user_offset + i] = sum_l grad_levels[l] psi_l(i)^2 h over
// the converged levels of the task's potential, zero elsewhere on the
// potential's full grid.  One block of W wavefronts per task (potential,
// side, segment): the segment's points, levels in groups of 64 W, the first
// group storing and later ones adding, so the sum order is fixed.  The
// segment-0 tasks also zero the points outside the open trimmed window on
// their side (all points of an unresolved potential, whose only task is side
// 0, segment 0).
template <int W>
__global__ __launch_bounds__(64 * W) void level_vjp_kernel(
    const double* __restrict__ uv, const double* __restrict__ uvr,
    const PotentialRecord* __restrict__ pots, const PsiTask* __restrict__ tasks,
    const double* __restrict__ levels, const int32_t* __restrict__ status,
    const LevelNorm* __restrict__ norms, const PsiCheckpoint* __restrict__ ckpt,
    const double* __restrict__ grad_levels, double* __restrict__ grad_v)
{
    __shared__ double s_c[W][32][65];
    __shared__ double s_w[W][32];
    const PsiTask task = tasks[blockIdx.x];
    const PotentialRecord& P = pots[task.potential];
    double* out = grad_v + P.user_offset;
    const int32_t N = P.full_points;
    const int32_t n = P.n_points;
    const int32_t first = static_cast<int32_t>(P.first_point);
    const bool ok = P.flags == 0;
    const int32_t t = static_cast<int32_t>(threadIdx.x);
    if (task.segment == 0) {
        if (!ok) {
            for (int32_t i = t; i < N; i += 64 * W)
                out[i] = 0.0;
            return;
        }
        if (task.side == 0)
            for (int32_t i = t; i <= first; i += 64 * W)
                out[i] = 0.0;
        else
            for (int32_t i = first + n - 1 + t; i < N; i += 64 * W)
                out[i] = 0.0;
    }
    if (!ok)
        return;
    const PsiSegment g = psi_segment(P, uv, uvr, task);
    const int32_t i0 = task.side == 0 ? 1 : n - 2;
    const int32_t dir = task.side == 0 ? 1 : -1;
    double* win = out + first;
    if (P.level_count == 0) {
        for (int32_t q = t; q < g.steps; q += 64 * W)
            win[i0 + dir * (g.j0 + q)] = 0.0;
        return;
    }
    const double c10 = opaque(10.0);
    for (int32_t g0 = 0; g0 < P.level_count; g0 += 64 * W) {
        const int32_t l = g0 + t;
        const int64_t o = P.level_offset + l;
        const bool active = l < P.level_count && level_active(status, o);
        const LevelNorm nm = active ? norms[o] : LevelNorm{0.0, 0.0, 0, 0};
        const double gl = active ? grad_levels[o] : 0.0;
        const double e = P.beta * (active ? levels[o] : P.e_lo);
        const double a = task.side == 0 ? nm.a_left : nm.a_right;

        VjpBody<W> B{};
        B.e = e;
        B.c10 = c10;
        B.g0 = gl * P.h * a * a;
        B.kf = task.side == 0 ? nm.k_left : nm.k_right;
        B.i0 = i0;
        B.dir = dir;
        B.base = g.j0;
        B.accumulate = g0 > 0;
        B.sc = s_c;
        B.sw = s_w;
        B.out = win;
        psi_start(B.s, P, task, g, ckpt, l, active, e);
        B.G = __builtin_ldexp(B.g0, 2 * (B.s.k - B.kf));
        psi_walk(g.p + 2 + g.j0, g.steps, B);
    }
}

// ---------------------------------------------------------------- launchers

hipError_t launch_normalize(hipStream_t stream, const double* uv, const double* uvr,
                            const PotentialRecord* pots, const PsiTask* psi_tasks, int32_t n_psi_tasks,
                            const WaveTask* tasks, int32_t n_tasks, const double* levels,
                            const int32_t* status, PsiSegMatrix* mats, PsiCheckpoint* ckpt, LevelNorm* norms)
{
    if (n_psi_tasks > 0) {
        hipLaunchKernelGGL(psi_segment_kernel, dim3(static_cast<unsigned>(n_psi_tasks)), dim3(64), 0, stream,
                           uv, uvr, pots, psi_tasks, levels, status, mats);
        if (const hipError_t e = hipGetLastError(); e != hipSuccess)
            return e;
    }
    if (n_tasks == 0)
        return hipSuccess;
    hipLaunchKernelGGL(normalize_kernel, dim3(static_cast<unsigned>(n_tasks)), dim3(64), 0, stream, pots, tasks,
                       n_tasks, status, mats, ckpt, norms);
    return hipGetLastError();
}

hipError_t launch_eigenfunctions(hipStream_t stream, const double* uv, const double* uvr,
                                 const PotentialRecord* pots, const PsiTask* tasks, int32_t n_tasks,
                                 const double* levels, const int32_t* status, const LevelNorm* norms,
                                 const PsiCheckpoint* ckpt, int64_t base, double* out)
{
    if (n_tasks == 0)
        return hipSuccess;
    hipLaunchKernelGGL(eigenfunction_kernel, dim3(static_cast<unsigned>(n_tasks)), dim3(64), 0, stream,
                       uv, uvr, pots, tasks, levels, status, norms, ckpt, base, out);
    return hipGetLastError();
}

hipError_t launch_level_vjp(hipStream_t stream, int32_t waves, const double* uv, const double* uvr,
                            const PotentialRecord* pots, const PsiTask* tasks, int32_t n_tasks,
                            const double* levels, const int32_t* status, const LevelNorm* norms,
                            const PsiCheckpoint* ckpt, const double* grad_levels, double* grad_v)
{
    if (n_tasks == 0)
        return hipSuccess;
    const dim3 grid(static_cast<unsigned>(n_tasks));
#define EPSEON_VJP(Wv)                                                                               \
    hipLaunchKernelGGL(level_vjp_kernel<Wv>, grid, dim3(64 * (Wv)), 0, stream, uv, uvr, pots, tasks, \
                       levels, status, norms, ckpt, grad_levels, grad_v)
    switch (waves) {
    case 1:
        EPSEON_VJP(1);
        break;
    case 2:
        EPSEON_VJP(2);
        break;
    case 3:
        EPSEON_VJP(3);
        break;
    case 4:
        EPSEON_VJP(4);
        break;
    default:
        return hipErrorInvalidValue;
    }
#undef EPSEON_VJP
    return hipGetLastError();
}

}  // namespace epseon::hip
