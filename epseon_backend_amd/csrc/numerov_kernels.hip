// numerov_kernels.hip -- gfx950 kernels of the Numerov-shooting path.
//
// The reference leaves this dispatch empty (VibwaAlgorithm<FP>::run,
// cpp/gpu/include/epseon/gpu/algorithms/vibwa.hpp:605-637, and no shader
// sources exist).  Design for CDNA4:
//   * one lane = one trial energy (scan) or one level (refine); the FP64
//     recurrence lives in registers;
//   * a wavefront always works on one potential, so the grid value u_i is
//     wave-uniform: it is read with scalar loads into SGPRs and used directly
//     as the SGPR operand of v_add_f64 -- no LDS traffic and no VGPRs for the
//     potential, and every potential point is fetched once per wavefront;
//   * no MFMA: the work is a scalar recurrence, bound by FP64 VALU issue.
// Scan recurrence (division-free homogeneous Numerov scaled by 12 per step,
// 4 FP64 VALU ops/step): with f_i = 1 + beta (E - V_i) and
// w_i = 12^(i-1) v_i, the recurrence
//   f_{i+1} y_{i+1} = (12 - 10 f_i) y_i - f_{i-1} y_{i-1}   (w = y f prod f)
// becomes, with phi_i = f_i / 12 = e' - p_i (p_i = (beta V_i - 1)/12,
// e' = beta E / 12),
//   v_{i+1} = v_i - phi_i (10 v_i + phi_{i-1} v_{i-1}),   v_0 = 0, v_1 = 1,
// i.e. z = phi_{i-1} v_{i-1}; s = fma(10, v_i, z); v_{i+1} = fma(-phi_i, s, v_i).
// |w| grows by less than 7.31 per step, so v only shrinks: after the 17th
// step of every second 32-step super-block, values below 2^-240 are
// multiplied by 2^480.  The sign of v (= sign of w) is sampled once per block of `stride`
// steps (a power of two the host chose so that no two nodes fall in one
// block).  Refinement shots run the difference form on uv_i = beta V_i with
// the overflow test every fourth super-block.  Compiled with
// -ffp-contract=off; every fused operation is an explicit fma, so results are
// bit-identical to the C oracle.
//
// Kernels in this file, in launch order: potential statistics and table
// shifts (load), band windows (each shot stops 15 guaranteed e-folds past its
// band top's outermost turning points), segment tables (load), the scan, the
// bracket pass (one thread per level), and the refinement: lane mode (one
// level per lane over the whole grid) or segmented mode (one block per level,
// each lane a segment's transfer matrix, tree product in LDS, 64-lane blocks
// reading the per-window segment table).  Scan and segmented tasks are binned
// per XCD so that a potential's wavefronts share one L2.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

#include "records.hpp"

namespace epseon::hip {

namespace {

__device__ __forceinline__ uint32_t sign_bit(double x)
{
    return static_cast<uint32_t>(__double_as_longlong(x) >> 63) & 1u;
}

// Opaque loop constant.  10 is not an inline constant of the FP64 VALU; held
// in a VGPR the compiler cannot rebuild it every step.
struct StepConsts {
    double c10;
};

__device__ __forceinline__ StepConsts step_consts()
{
    StepConsts k{10.0};
    asm volatile("" : "+v"(k.c10));
    return k;
}

// Node counter: each sample's sign bit is shifted into `bits` (newest in bit
// 0; one funnel shift per sample), and the sign changes are counted in
// batches.  flush(2^m - 1) adds those among the m newest samples and the
// sample before them (bit m; zero-initialised bits stand for the positive
// start), so m <= 31 samples may wait between flushes.
struct SignCounter {
    int32_t nodes;
    uint32_t bits;
    __device__ __forceinline__ void sample(const double w)
    {
        const uint32_t hi = static_cast<uint32_t>(static_cast<uint64_t>(__double_as_longlong(w)) >> 32);
        bits = __builtin_amdgcn_alignbit(bits, hi, 31u);
    }
    __device__ __forceinline__ void flush(const uint32_t mask)
    {
        nodes += __builtin_popcount((bits ^ (bits >> 1)) & mask);
    }
};

// Scan form on p_i = (beta V_i - 1)/12 with e' = beta E / 12: the 12-scaled
// homogeneous recurrence, 4 FP64 VALU ops per step (w holds v).
struct ScanForm {
    using Energy = double;
    using Counter = SignCounter;
    static constexpr int kRescaleEvery = 2;
    double wm, w, fm;
    __device__ __forceinline__ void init(const double e, const double u0)
    {
        wm = 0.0;
        w = 1.0;
        fm = e - u0;
        rescales = 0;
    }
    __device__ __forceinline__ void step(const double u, const double e, const StepConsts& k)
    {
        const double phi = e - u;
        const double z = fm * wm;
        const double s = __builtin_fma(k.c10, w, z);
        const double wn = __builtin_fma(-phi, s, w);
        wm = w;
        w = wn;
        fm = phi;
    }
    __device__ __forceinline__ void rescale()
    {
        const bool small = fmax(fabs(w), fabs(wm)) < kScanRescaleBelow;
        const int32_t sh = small ? kRescaleLog2 : 0;
        w = __builtin_ldexp(w, sh);
        wm = __builtin_ldexp(wm, sh);
        rescales += small ? 1 : 0;
    }
    __device__ __forceinline__ void sample(SignCounter& c) const { c.sample(w); }
    int32_t rescales;
};

// Difference form on uv_i = beta V_i (refinement): eps_i = e - uv_i carried
// at full relative precision, 7 FP64 VALU ops per step:
//   sig = fma(eps_i, eps_{i-1}, eps_i + eps_{i-1}),
//   d = fma(-10 eps_i, w_i, fma(-sig, w_{i-1}, d)), w += d.
struct DiffForm {
    using Energy = double;
    using Counter = SignCounter;
    static constexpr int kRescaleEvery = 4;
    double w, d, wm, em;
    __device__ __forceinline__ void init(const double e, const double u0)
    {
        w = 1.0;
        d = 1.0;
        wm = 0.0;
        em = e - u0;
        rescales = 0;
    }
    __device__ __forceinline__ void step(const double u, const double e, const StepConsts& k)
    {
        const double ep = e - u;
        const double e10 = ep * k.c10;
        const double sg = __builtin_fma(ep, em, ep + em);
        d = __builtin_fma(-e10, w, __builtin_fma(-sg, wm, d));
        wm = w;
        w = w + d;
        em = ep;
    }
    __device__ __forceinline__ void rescale()
    {
        const bool big = fmax(fabs(w), fabs(wm)) >= kRescaleAt;
        const int32_t sh = big ? -kRescaleLog2 : 0;
        w = __builtin_ldexp(w, sh);
        d = __builtin_ldexp(d, sh);
        wm = __builtin_ldexp(wm, sh);
        rescales += big ? 1 : 0;
    }
    __device__ __forceinline__ void sample(SignCounter& c) const { c.sample(w); }
    int32_t rescales;
};

// Two refinement shots advanced together (two levels per lane): their
// dependency chains interleave, which hides FP64 latency in wavefronts that
// have a SIMD to themselves.  Each member is bit-identical to a lone shot.
struct DiffPair {
    static constexpr int kRescaleEvery = 4;
    struct Energy {
        double a, b;
    };
    struct Counter {
        SignCounter a, b;
        __device__ __forceinline__ void flush(const uint32_t mask)
        {
            a.flush(mask);
            b.flush(mask);
        }
    };
    DiffForm x, y;
    __device__ __forceinline__ void init(const Energy e, const double u0)
    {
        x.init(e.a, u0);
        y.init(e.b, u0);
    }
    __device__ __forceinline__ void step(const double u, const Energy e, const StepConsts& k)
    {
        x.step(u, e.a, k);
        y.step(u, e.b, k);
    }
    __device__ __forceinline__ void rescale()
    {
        x.rescale();
        y.rescale();
    }
    __device__ __forceinline__ void sample(Counter& c) const
    {
        c.a.sample(x.w);
        c.b.sample(y.w);
    }
};

// Scheduling fence: nothing moves across it, so the next buffer's scalar
// loads issue right after the first use of the current buffer.  The wait the
// compiler inserts before that first use then finds only the load issued one
// half super-block earlier still outstanding (scalar loads return out of
// order, so every wait is lgkmcnt(0)).
__device__ __forceinline__ void fence_schedule() { __builtin_amdgcn_sched_barrier(0); }

// Steps between node-count flushes at stride B (at most 16 samples wait).
template <int B>
constexpr int kFlushSteps = B >= 8 ? 128 : 16 * B;

// Step Q of the 128-step loop body: the sign is sampled after every B-th
// step and the counter flushed after every kFlushSteps<B>-th.
template <int B, int Q, class Form>
__device__ __forceinline__ void step_at(const double u, const typename Form::Energy e,
                                        const StepConsts& k, Form& s, typename Form::Counter& c)
{
    s.step(u, e, k);
    if constexpr ((Q + 1) % B == 0)
        s.sample(c);
    if constexpr ((Q + 1) % kFlushSteps<B> == 0)
        c.flush((1u << (kFlushSteps<B> / B)) - 1u);
}

template <int B, int Q0, class Form, int... Is>
__device__ __forceinline__ void steps_1_15(const double (&buf)[16], const typename Form::Energy e,
                                           const StepConsts& k, Form& s, typename Form::Counter& c,
                                           std::integer_sequence<int, Is...>)
{
    (step_at<B, Q0 + Is + 1>(buf[Is + 1], e, k, s, c), ...);
}

// One 32-step super-block (number S of the loop body) fed from the two
// 16-point SGPR buffers: each buffer's scalar loads issue right after the
// first step of the other buffer, so a load has 15 steps to land.  With R
// the rescale test follows the 17th step.
template <int B, bool R, int S, class Form>
__device__ __forceinline__ void super_block(const double*& p, double (&a)[16], double (&b)[16],
                                            const typename Form::Energy e, const StepConsts& k, Form& s,
                                            typename Form::Counter& c)
{
    step_at<B, 32 * S>(a[0], e, k, s, c);
    fence_schedule();
#pragma unroll
    for (int q = 0; q < 16; ++q)
        b[q] = p[16 + q];
    fence_schedule();
    steps_1_15<B, 32 * S>(a, e, k, s, c, std::make_integer_sequence<int, 15>{});
    step_at<B, 32 * S + 16>(b[0], e, k, s, c);
    if constexpr (R)
        s.rescale();
    fence_schedule();
#pragma unroll
    for (int q = 0; q < 16; ++q)
        a[q] = p[32 + q];
    fence_schedule();
    steps_1_15<B, 32 * S + 16>(b, e, k, s, c, std::make_integer_sequence<int, 15>{});
    p += 32;
}

// One outward shot over n points at probe e with node stride B, the grid
// table u being wave-uniform.  The grid is consumed in 32-step super-blocks,
// four per loop iteration so that the body stays one straight-line block;
// the rescale test runs after the 17th step of every second super-block
// (Form::kRescaleEvery = 2) or of every fourth one (4).  The sign is sampled
// after every B-th step and at the last point.  The table is
// padded so the prefetch past the last super-block stays in the allocation.
// The form holds the final state, the counter the node count(s).
template <int B, class Form>
__device__ __forceinline__ void shoot_fixed(const double* __restrict__ u, const int32_t n,
                                            const typename Form::Energy e, Form& s,
                                            typename Form::Counter& c)
{
    static_assert(32 % B == 0, "stride must divide the super-block");
    const StepConsts k = step_consts();
    s.init(e, u[0]);
    c = typename Form::Counter{};
    const int32_t steps = n - 2;
    const int32_t nsb = steps >> 5;
    const double* p = u + 1;
    double a[16], b[16];
#pragma unroll
    for (int q = 0; q < 16; ++q)
        a[q] = p[q];
    constexpr bool odd = Form::kRescaleEvery == 2;
    for (int32_t g = 0; g < (nsb >> 2); ++g) {
        super_block<B, false, 0>(p, a, b, e, k, s, c);
        super_block<B, odd, 1>(p, a, b, e, k, s, c);
        super_block<B, false, 2>(p, a, b, e, k, s, c);
        super_block<B, true, 3>(p, a, b, e, k, s, c);
    }
    const int32_t left = nsb & 3;
    if (left >= 1)
        super_block<B, false, 0>(p, a, b, e, k, s, c);
    if (left >= 2)
        super_block<B, odd, 1>(p, a, b, e, k, s, c);
    if (left >= 3)
        super_block<B, false, 2>(p, a, b, e, k, s, c);
    const int32_t rem = steps - (nsb << 5);
    for (int32_t q = 0; q < rem; ++q) {
        s.step(p[q], e, k);
        if ((q + 1) % B == 0)
            s.sample(c);
    }
    if (rem % B != 0)
        s.sample(c);
    // Samples since the last flush: those of the remaining super-blocks past
    // the last flush point and of the tail (at most 16).
    const int32_t pend = ((32 * left) % kFlushSteps<B>) / B + rem / B + (rem % B != 0 ? 1 : 0);
    c.flush((1u << pend) - 1u);
}

template <class Form>
__device__ __forceinline__ void shoot(const double* __restrict__ u, const int32_t n,
                                      const typename Form::Energy e, const int32_t stride, Form& s,
                                      typename Form::Counter& c)
{
    switch (stride) {
    case 32: return shoot_fixed<32, Form>(u, n, e, s, c);
    case 16: return shoot_fixed<16, Form>(u, n, e, s, c);
    case 8: return shoot_fixed<8, Form>(u, n, e, s, c);
    case 4: return shoot_fixed<4, Form>(u, n, e, s, c);
    case 2: return shoot_fixed<2, Form>(u, n, e, s, c);
    default: return shoot_fixed<1, Form>(u, n, e, s, c);
    }
}

// Single-form shot: node count, far-end value and rescale count.
template <class Form>
__device__ __forceinline__ int32_t shoot(const double* __restrict__ u, const int32_t n,
                                         const double e, const int32_t stride, double& w_last,
                                         int32_t& rescales)
{
    Form s;
    s.rescales = 0;
    SignCounter c;
    shoot<Form>(u, n, e, stride, s, c);
    w_last = s.w;
    rescales = s.rescales;
    return c.nodes;
}

__device__ __forceinline__ int32_t wave_task_index()
{
    const int32_t waves_per_block = static_cast<int32_t>(blockDim.x >> 6);
    const int32_t wave = __builtin_amdgcn_readfirstlane(static_cast<int32_t>(threadIdx.x >> 6));
    return static_cast<int32_t>(blockIdx.x) * waves_per_block + wave;
}

// Task of this wavefront.  With `bins` (9 offsets into `tasks`), the list is
// split into one bin per XCD: workgroups are dispatched round-robin over the
// 8 XCDs, so workgroup b runs on XCD b % 8 and takes the (b / 8)-th group of
// that XCD's bin; all wavefronts of a potential then read its table through
// one L2.  Without bins the list is walked in order.
__device__ __forceinline__ bool fetch_task(const WaveTask* __restrict__ tasks, const int32_t n_tasks,
                                           const int32_t* __restrict__ bins, WaveTask& out)
{
    if (bins == nullptr) {
        const int32_t t = wave_task_index();
        if (t >= n_tasks)
            return false;
        out = tasks[t];
        return true;
    }
    const int32_t waves_per_block = static_cast<int32_t>(blockDim.x >> 6);
    const int32_t wave = __builtin_amdgcn_readfirstlane(static_cast<int32_t>(threadIdx.x >> 6));
    const int32_t x = static_cast<int32_t>(blockIdx.x & 7u);
    const int32_t t = bins[x] + static_cast<int32_t>(blockIdx.x >> 3) * waves_per_block + wave;
    if (t >= bins[x + 1])
        return false;
    out = tasks[t];
    return true;
}

// Task of this workgroup (one task per workgroup), binned per XCD like
// fetch_task when `bins` is given.
__device__ __forceinline__ bool fetch_block_task(const WaveTask* __restrict__ tasks, const int32_t n_tasks,
                                                 const int32_t* __restrict__ bins, WaveTask& out)
{
    int32_t t = static_cast<int32_t>(blockIdx.x);
    if (bins != nullptr) {
        const int32_t x = static_cast<int32_t>(blockIdx.x & 7u);
        t = bins[x] + static_cast<int32_t>(blockIdx.x >> 3);
        if (t >= bins[x + 1])
            return false;
    } else if (t >= n_tasks) {
        return false;
    }
    out = tasks[t];
    return true;
}

}  // namespace

// Statistics of each V table in device memory (one block per potential),
// with the host loader's semantics: first index of the minimum, first
// non-finite index, and the wall-trimmed Dirichlet points.
__global__ __launch_bounds__(256) void potential_stats_kernel(const double* __restrict__ v,
                                                              const int64_t* __restrict__ offsets,
                                                              const int64_t* __restrict__ counts,
                                                              const double* __restrict__ betas,
                                                              PotentialStats* __restrict__ out)
{
    __shared__ double s_v[256];
    __shared__ int64_t s_i[256], s_b[256];
    const int32_t p = static_cast<int32_t>(blockIdx.x);
    const int32_t t = static_cast<int32_t>(threadIdx.x);
    const double* base = v + offsets[p];
    const int64_t n = counts[p];
    double best = __builtin_inf();
    int64_t bi = n, bad = n;
    for (int64_t i = t; i < n; i += 256) {
        const double x = base[i];
        if (!__builtin_isfinite(x)) {
            bad = i < bad ? i : bad;
        } else if (x < best || (x == best && i < bi)) {
            best = x;
            bi = i;
        }
    }
    s_v[t] = best;
    s_i[t] = bi;
    s_b[t] = bad;
    __syncthreads();
    for (int32_t d = 128; d > 0; d >>= 1) {
        if (t < d) {
            const double ov = s_v[t + d];
            const int64_t oi = s_i[t + d];
            if (ov < s_v[t] || (ov == s_v[t] && oi < s_i[t])) {
                s_v[t] = ov;
                s_i[t] = oi;
            }
            s_b[t] = s_b[t + d] < s_b[t] ? s_b[t + d] : s_b[t];
        }
        __syncthreads();
    }
    const double vmin = s_v[0];
    const int64_t imin = s_i[0];
    const int64_t first_bad = s_b[0];
    __syncthreads();
    // Trimming needs a finite table with a minimum.
    int64_t first = 0, last = n - 1;
    if (first_bad == n && imin < n) {
        const double wall = vmin + 0.5 / betas[p];
        int64_t lo = -1, hi = n;
        for (int64_t i = t; i <= imin; i += 256)
            if (base[i] >= wall && i > lo)
                lo = i;
        for (int64_t i = imin + t; i < n; i += 256)
            if (base[i] >= wall && i < hi)
                hi = i;
        s_i[t] = lo;
        s_b[t] = hi;
        __syncthreads();
        for (int32_t d = 128; d > 0; d >>= 1) {
            if (t < d) {
                s_i[t] = s_i[t + d] > s_i[t] ? s_i[t + d] : s_i[t];
                s_b[t] = s_b[t + d] < s_b[t] ? s_b[t + d] : s_b[t];
            }
            __syncthreads();
        }
        if (s_i[0] >= 0)
            first = s_i[0];
        if (s_b[0] < n)
            last = s_b[0];
    }
    if (t == 0)
        out[p] = PotentialStats{vmin, imin, first, last, first_bad};
}

// p_i = fma(beta_p, V_i, -1) / 12 (scan table), uv_i = beta_p V_i (refinement
// table) and the reversed refinement table uvr_i = uv_{n-1-i} (inward shots
// from the right wall) for every point of the trimmed grid of every potential,
// read from V in the caller's layout.
__global__ __launch_bounds__(256) void shift_potential_kernel(const double* __restrict__ v,
                                                              const PotentialRecord* __restrict__ pots,
                                                              double* __restrict__ ut,
                                                              double* __restrict__ uv,
                                                              double* __restrict__ uvr)
{
    const PotentialRecord& P = pots[blockIdx.y];
    const double beta = P.beta;
    const int32_t n = P.n_points;
    const double* src = v + P.user_offset + P.first_point;
    for (int32_t i = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x); i < n;
         i += static_cast<int32_t>(gridDim.x * blockDim.x)) {
        const double x = src[i];
        const double b = beta * x;
        ut[P.ut_offset + i] = __builtin_fma(beta, x, -1.0) / 12.0;
        uv[P.ut_offset + i] = b;
        uvr[P.ut_offset + (n - 1 - i)] = b;
    }
}

namespace {

// Band of the scan group (64 trial energies from a multiple of 64) holding
// trial energy j: the first band whose top is at or above the group's
// highest energy, else the last band.
__device__ __forceinline__ int32_t group_band(const PotentialRecord& P, const int32_t j)
{
    const int32_t jt = min((j / 64) * 64 + 63, P.n_energies - 1);
    const double et = __builtin_fma(static_cast<double>(jt + 1), P.de, P.e_lo);
    const double step = (P.e_hi - P.e_lo) / kBands;
    for (int32_t b = 0; b < kBands - 1; ++b)
        if (__builtin_fma(static_cast<double>(b + 1), step, P.e_lo) >= et)
            return b;
    return kBands - 1;
}

__device__ __forceinline__ bool same_window(const BandWindow* W, const int32_t a, const int32_t b)
{
    return W[a].lo == W[b].lo && W[a].hi == W[b].hi;
}

// floor(sqrt(n)) for n < 2^53 (the double square root, then corrected).
__device__ __forceinline__ uint64_t isqrt_exact(const uint64_t n)
{
    uint64_t r = static_cast<uint64_t>(__builtin_sqrt(static_cast<double>(n)));
    while (r * r > n)
        --r;
    while ((r + 1) * (r + 1) <= n)
        ++r;
    return r;
}

// Lower bound g(x) of the per-step growth acosh(1 + x/2) in 2^-40 fixed
// point, within 9 % of it: with x = m 2^k (m in [1/2, 1)) and T the top 16
// bits of m's significand, x >= T 2^(k-16), and sqrt(T 2^(k-16)) 2^40 =
// sqrt(T 2^r) 2^h (k + 64 = 2 h + r) is taken as isqrt(T 2^(r+32)) 2^(h-16),
// floored, times floor(65536 acosh(1 + y/2)/sqrt(y)) at y = 2^k (the ratio
// falls with y; entry -5 serves below) over 65536.  Integer operations only
// after frexp.  0 below 2^-65; x >= 8 counts as just below 8.
__device__ __forceinline__ int64_t growth_term(const double x)
{
    constexpr int64_t kRatio[9] = {65450, 65366, 65199, 64871, 64242, 63073, 61029, 57761, 53116};
    int k = 0;
    const double m = __builtin_frexp(x, &k);
    if (!(x > 0.0) || k < -64)
        return 0;
    uint64_t t = static_cast<uint64_t>(__builtin_ldexp(m, 16));
    if (k > 3) {
        k = 3;
        t = 0xFFFF;
    }
    const int32_t e = k + 64, h = e >> 1;
    const uint64_t r = isqrt_exact(t << ((e & 1) + 32));
    const uint64_t sq = h >= 16 ? r << (h - 16) : r >> (16 - h);
    int64_t f = kRatio[0];
#pragma unroll
    for (int32_t i = 1; i < 9; ++i)
        if (k == i - 5)
            f = kRatio[i];
    return static_cast<int64_t>((sq * static_cast<uint64_t>(f)) >> 16);
}

__device__ __forceinline__ int64_t wave_inclusive_sum(int64_t x)
{
    const int32_t lane = static_cast<int32_t>(threadIdx.x & 63);
#pragma unroll
    for (int32_t d = 1; d < 64; d *= 2) {
        const int64_t y = __shfl_up(x, d, 64);
        if (lane >= d)
            x += y;
    }
    return x;
}

__device__ __forceinline__ int32_t wave_min(int32_t x)
{
#pragma unroll
    for (int32_t d = 32; d >= 1; d /= 2)
        x = min(x, __shfl_xor(x, d, 64));
    return x;
}

__device__ __forceinline__ int32_t wave_max(int32_t x)
{
#pragma unroll
    for (int32_t d = 32; d >= 1; d /= 2)
        x = max(x, __shfl_xor(x, d, 64));
    return x;
}

}  // namespace

namespace {

// Walk a fringe of the window at e from `start` by `dir` (+1: right, -1:
// left), 64 points per step, until the depth sum reaches kWindowDepth;
// returns the point where it does, or the grid end.  One wavefront.
__device__ __forceinline__ int32_t fringe_end(const double* __restrict__ u, const int32_t n, const double e,
                                              const int32_t start, const int32_t dir)
{
    // Chunks of 64 x kPer points, kPer consecutive ones (in walk order) per
    // lane: each lane sums its terms, one wave prefix sum orders the lanes,
    // and the first lane reaching the depth finds the crossing among its own.
    constexpr int32_t kPer = 8;
    const int32_t lane = static_cast<int32_t>(threadIdx.x & 63);
    const int64_t depth = int64_t(kWindowDepth) << 40;
    int64_t acc = 0;
    for (int32_t base = start; base >= 0 && base < n; base += 64 * kPer * dir) {
        int64_t g[kPer];
        int64_t own = 0;
#pragma unroll
        for (int32_t q = 0; q < kPer; ++q) {
            const int32_t i = base + dir * (kPer * lane + q);
            g[q] = (i >= 0 && i < n) ? growth_term(12.0 * (u[i] - e)) : 0;
            own += g[q];
        }
        const int64_t inc = wave_inclusive_sum(own);
        const uint64_t hit = __ballot(acc + inc >= depth);
        if (hit != 0) {
            const int32_t h = static_cast<int32_t>(__builtin_ctzll(hit));
            int64_t run = acc + inc - own;
            int32_t at = kPer - 1;
            bool found = false;
#pragma unroll
            for (int32_t q = 0; q < kPer; ++q) {
                run += g[q];
                if (!found && run >= depth) {
                    at = q;
                    found = true;
                }
            }
            return base + dir * (kPer * h + __shfl(at, h, 64));
        }
        acc += __shfl(inc, 63, 64);
    }
    return dir > 0 ? n - 1 : 0;
}

}  // namespace

// Band windows: one workgroup of kBandWaves wavefronts per potential.  One
// pass over the table finds the first and last allowed point (uv_i <= e_b)
// of all kBands band tops at once; then the wavefronts walk the 2 kBands
// fringes.  Writes (lo, hi) per band (whole grid for rejected potentials).
constexpr int32_t kBandWaves = 16;
__global__ __launch_bounds__(64 * kBandWaves) void band_window_kernel(const double* __restrict__ uv,
                                                          const PotentialRecord* __restrict__ pots,
                                                          const int32_t n_pots, int2* __restrict__ out)
{
    __shared__ int32_t s_first[kBandWaves][kBands], s_last[kBandWaves][kBands];
    __shared__ int32_t s_win[2][kBands];
    const int32_t p = static_cast<int32_t>(blockIdx.x);
    if (p >= n_pots)
        return;
    const PotentialRecord& P = pots[p];
    const int32_t tid = static_cast<int32_t>(threadIdx.x);
    const int32_t lane = tid & 63, wave = tid >> 6;
    const int32_t n = P.n_points;
    if (P.flags != 0) {
        if (tid < kBands)
            out[p * kBands + tid] = make_int2(0, n - 1);
        return;
    }
    const double* u = uv + P.ut_offset;
    double eb[kBands];
    int32_t first[kBands], last[kBands];
#pragma unroll
    for (int32_t b = 0; b < kBands; ++b) {
        eb[b] = P.beta * __builtin_fma(static_cast<double>(b + 1), (P.e_hi - P.e_lo) / kBands, P.e_lo);
        first[b] = n;
        last[b] = -1;
    }
    // Four loads in flight per thread.
    constexpr int32_t kT = 64 * kBandWaves;
    for (int32_t i0 = tid; i0 < n; i0 += 4 * kT) {
        double x[4];
#pragma unroll
        for (int32_t q = 0; q < 4; ++q)
            x[q] = i0 + q * kT < n ? u[i0 + q * kT] : __builtin_inf();
#pragma unroll
        for (int32_t q = 0; q < 4; ++q)
#pragma unroll
            for (int32_t b = 0; b < kBands; ++b)
                if (x[q] <= eb[b]) {
                    first[b] = min(first[b], i0 + q * kT);
                    last[b] = i0 + q * kT;
                }
    }
#pragma unroll
    for (int32_t b = 0; b < kBands; ++b) {
        first[b] = wave_min(first[b]);
        last[b] = wave_max(last[b]);
    }
    if (lane == 0) {
#pragma unroll
        for (int32_t b = 0; b < kBands; ++b) {
            s_first[wave][b] = first[b];
            s_last[wave][b] = last[b];
        }
    }
    __syncthreads();
    // Fringe walks: task t = 2 b + side, wavefront w takes t = w, w + kBandWaves, ...
    for (int32_t t = wave; t < 2 * kBands; t += kBandWaves) {
        const int32_t b = t >> 1, side = t & 1;
        int32_t f = s_first[0][b], l = s_last[0][b];
#pragma unroll
        for (int32_t w = 1; w < kBandWaves; ++w) {
            f = min(f, s_first[w][b]);
            l = max(l, s_last[w][b]);
        }
        int32_t r = side ? n - 1 : 0;
        if (l >= 0)
            r = side ? fringe_end(u, n, eb[b], l + 1, 1) : fringe_end(u, n, eb[b], f - 1, -1);
        if (lane == 0)
            s_win[side][b] = r;
    }
    __syncthreads();
    if (tid < kBands)
        out[p * kBands + tid] = make_int2(s_win[0][tid], s_win[1][tid]);
}

// Scan: lane = trial energy E_j = fma(j+1, dE, E_lo) shot over the band
// window of its group; outputs node count, far-end mantissa v_{N-1} *
// far_scale and scale-up count r per trial energy (w_{N-1} = mantissa *
// 2^(far_exp - 480 r), N and the scale those of the window).
__global__ __launch_bounds__(256) void numerov_scan_kernel(const double* __restrict__ ut,
                                                           const PotentialRecord* __restrict__ pots,
                                                           const BandWindow* __restrict__ windows,
                                                           const WaveTask* __restrict__ tasks,
                                                           const int32_t n_tasks,
                                                           const int32_t* __restrict__ bins,
                                                           int32_t* __restrict__ nodes,
                                                           double* __restrict__ far_end,
                                                           int32_t* __restrict__ rescales)
{
    WaveTask task;
    if (!fetch_task(tasks, n_tasks, bins, task))
        return;
    const PotentialRecord& P = pots[task.potential];
    const int32_t j = task.first + static_cast<int32_t>(threadIdx.x & 63);
    if (j >= P.n_energies)
        return;
    const BandWindow& W = windows[task.potential * kBands + group_band(P, task.first)];
    const double E = __builtin_fma(static_cast<double>(j + 1), P.de, P.e_lo);
    const double e = (P.beta * E) / 12.0;
    double w;
    int32_t k;
    const int32_t c = shoot<ScanForm>(ut + P.ut_offset + W.lo, W.hi - W.lo + 1, e, P.stride, w, k);
    const int64_t o = P.energy_offset + j;
    nodes[o] = c;
    far_end[o] = w * W.far_scale;
    rescales[o] = k;
}

namespace {

// ---- deterministic exp / log / sqrt of the refinement model -------------
// Same operations as the oracle (+, -, *, /, fma, floor, frexp, ldexp), so
// both produce the same bits.

constexpr double kLn2 = 0.69314718055994530942;
constexpr double kLn2Hi = 6.93147180369123816490e-01;  // q * kLn2Hi exact for |q| < 2^20
constexpr double kLn2Lo = 1.90821492927058770002e-10;
constexpr double kSqrt1_2 = 0.70710678118654752440;

// e^y = m 2^n, m in [1, 2).
__device__ __forceinline__ void model_exp(const double y, double& m, int32_t& n)
{
    const double q = floor(y / kLn2);
    const double r = (y - q * kLn2Hi) - q * kLn2Lo;
    constexpr double c[18] = {1.0, 1.0, 1.0 / 2, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720,
                              1.0 / 5040, 1.0 / 40320, 1.0 / 362880, 1.0 / 3628800,
                              1.0 / 39916800, 1.0 / 479001600, 1.0 / 6227020800.0,
                              1.0 / 87178291200.0, 1.0 / 1307674368000.0,
                              1.0 / 20922789888000.0, 1.0 / 355687428096000.0};
    double p = c[17];
#pragma unroll
    for (int i = 16; i >= 0; --i)
        p = __builtin_fma(p, r, c[i]);
    m = p;
    n = static_cast<int32_t>(q);
}

__device__ __forceinline__ double model_sqrt(const double d)
{
    if (!(d > 0.0))
        return 0.0;
    int e = 0;
    double m = __builtin_frexp(d, &e);
    if (e & 1) {
        m = m * 2.0;
        e = e - 1;
    }
    double y = __builtin_fma(m, 0.4142135623730950, 0.5857864376269050);
#pragma unroll
    for (int i = 0; i < 5; ++i)
        y = 0.5 * (y + m / y);
    return __builtin_ldexp(y, e / 2);
}

__device__ __forceinline__ double model_log(const double x)
{
    int e = 0;
    double m = __builtin_frexp(x, &e);
    if (m < kSqrt1_2) {
        m = m * 2.0;
        e = e - 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 1.0 / 25.0;
#pragma unroll
    for (int i = 11; i >= 0; --i)
        p = __builtin_fma(p, s2, 1.0 / static_cast<double>(2 * i + 1));
    return __builtin_fma(static_cast<double>(e), kLn2, 2.0 * s * p);
}

// Rate b of F(E) = (c0 + c1 (E - E_b)) e^(b (E - E_b)) through equally spaced
// scan points a < b < c: with x = e^(b de), F_a/F_b x + F_c/F_b / x = 2, one
// positive root either way the third point lies.  0 when no fit.  Not
// inlined: its square root and logarithm would raise the register peak of
// the refinement kernels that call it once per level.
__device__ __attribute__((noinline)) double model_rate(const double wa, const int32_t ka, const double wb,
                                             const int32_t kb, const double wc, const int32_t kc,
                                             const bool above, const double de)
{
    if (wb == 0.0 || ka - kb > 960 || kb - ka > 960 || kc - kb > 960 || kb - kc > 960)
        return 0.0;
    const double ra = __builtin_ldexp(wa / wb, ka - kb);
    const double rc = __builtin_ldexp(wc / wb, kc - kb);
    const double disc = 1.0 - ra * rc;
    if (!(disc >= 0.0) || ra == 0.0 || !__builtin_isfinite(ra) || !__builtin_isfinite(rc))
        return 0.0;
    const double sq = model_sqrt(disc);
    const double x = above ? (1.0 - sq) / ra : (1.0 + sq) / ra;
    if (!(x > 0.0) || !__builtin_isfinite(x))
        return 0.0;
    const double lx = model_log(x);
    if (!(fabs(lx) < 600.0))
        return 0.0;
    return lx / de;
}

// F = w 2^k -> G = F e^(-b (e - e0)).
__device__ __forceinline__ void model_apply(const double b, const double e0, const double e, double& w,
                                            int32_t& k)
{
    double m;
    int32_t n;
    model_exp(-b * (e - e0), m, n);
    w = w * m;
    k = k + n;
}

// Bracket of level v from the scan: the first trial energy with more than v
// nodes and the one below it (or the window floor, below which no level
// exists since every eigenvalue exceeds min V).  Exponents are binary.  When
// the bracket holds exactly level v and a third scan point of the same count
// regime exists above or below it, `rate` is the model rate b.  `pts` gets
// the level's interpolation points: scan points j - kPts/2 .. j + kPts/2 - 1
// (j the upper end) whose group shares the level's window, the lower end only
// when its value is of that window.
// W: the potential's kBands band windows (those the scan used).
__device__ __forceinline__ Bracket find_bracket(const PotentialRecord& P, const BandWindow* W,
                                                const int32_t v, const int32_t* __restrict__ nodes,
                                                const double* __restrict__ far_end,
                                                const int32_t* __restrict__ rescales, RefinePoints& pts)
{
    Bracket b{};
    pts.n = 0;
    pts.ih = 0;
    const int32_t m = P.n_energies;
    const int32_t* nd = nodes + P.energy_offset;
    int32_t lo = 0, hi = m;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (nd[mid] > v)
            hi = mid;
        else
            lo = mid + 1;
    }
    const int32_t j = lo;
    if (j == m) {
        b.found = 0;
        return b;
    }
    const int32_t band = group_band(P, j);
    const int32_t fe = W[band].far_exp;
    b.found = 1;
    b.band = band;
    b.eh = __builtin_fma(static_cast<double>(j + 1), P.de, P.e_lo);
    b.ch = nd[j];
    b.wh = far_end[P.energy_offset + j];
    b.kh = fe - rescales[P.energy_offset + j] * kRescaleLog2;
    if (j == 0 || nd[j - 1] > v) {
        b.el = P.e_lo;
        b.cl = 0;
        b.lo_ok = 0;
    } else {
        const int32_t bl = group_band(P, j - 1);
        b.el = __builtin_fma(static_cast<double>(j), P.de, P.e_lo);
        b.cl = nd[j - 1];
        b.wl = far_end[P.energy_offset + j - 1];
        b.kl = W[bl].far_exp - rescales[P.energy_offset + j - 1] * kRescaleLog2;
        b.lo_ok = 1;
        b.lo_foreign = same_window(W, bl, band) ? 0 : 1;
    }
    // Candidates j - kPts/2 + s, compacted in registers (every slot addressed
    // by a constant index), then the record is written whole.
    double pe[kPts], pw[kPts];
    int32_t pk[kPts];
    int32_t np = 0, ih = 0;
#pragma unroll
    for (int32_t s = 0; s < kPts; ++s) {
        pe[s] = 0.0;
        pw[s] = 0.0;
        pk[s] = 0;
    }
#pragma unroll
    for (int32_t s = 0; s < kPts; ++s) {
        const int32_t q = j - kPts / 2 + s;
        bool ok = q >= 0 && q < m;
        int32_t bq = band;
        if (ok && q == j - 1)
            ok = b.lo_ok && !b.lo_foreign;
        else if (ok && q != j) {
            bq = group_band(P, q);
            ok = same_window(W, bq, band);
        }
        if (ok) {
            const double e = __builtin_fma(static_cast<double>(q + 1), P.de, P.e_lo);
            const double w = far_end[P.energy_offset + q];
            const int32_t k = W[bq].far_exp - rescales[P.energy_offset + q] * kRescaleLog2;
#pragma unroll
            for (int32_t t = 0; t < kPts; ++t)
                if (t == np) {
                    pe[t] = e;
                    pw[t] = w;
                    pk[t] = k;
                }
            if (q == j)
                ih = np;
            ++np;
        }
    }
#pragma unroll
    for (int32_t t = 0; t < kPts; ++t) {
        pts.e[t] = pe[t];
        pts.w[t] = pw[t];
        pts.k[t] = pk[t];
    }
    pts.n = np;
    pts.ih = ih;
    b.rate = 0.0;
    if (b.lo_ok && !b.lo_foreign && b.cl == v && b.ch == v + 1) {
        const int64_t o = P.energy_offset;
        if (j + 1 < m && nd[j + 1] == b.ch && same_window(W, group_band(P, j + 1), band))
            b.rate = model_rate(b.wl, b.kl, b.wh, b.kh, far_end[o + j + 1],
                                fe - rescales[o + j + 1] * kRescaleLog2, true, P.de);
        else if (j >= 2 && nd[j - 2] == b.cl && same_window(W, group_band(P, j - 2), band))
            b.rate = model_rate(far_end[o + j - 2], fe - rescales[o + j - 2] * kRescaleLog2, b.wl, b.kl,
                                b.wh, b.kh, false, P.de);
    }
    return b;
}

// Insert a shot at em (inside the bracket: at index ih in energy order).  A
// full list first drops the outermost point farther from the bracket centre c
// (the last one on a tie).  Every slot is addressed by a constant index
// (predicated moves), so a list held in registers stays there.
__device__ __forceinline__ void points_insert(RefinePoints& p, const double c, const double em, const double w,
                                              const int32_t k)
{
    int32_t n = p.n;
    int32_t ih = p.ih;
    if (n == kPts) {
        if (c - p.e[0] > p.e[kPts - 1] - c) {
#pragma unroll
            for (int32_t i = 1; i < kPts; ++i) {
                p.e[i - 1] = p.e[i];
                p.w[i - 1] = p.w[i];
                p.k[i - 1] = p.k[i];
            }
            --ih;
        }
        --n;
    }
    // Slots ih .. n - 1 move up by one (from the top down), slot ih takes the shot.
#pragma unroll
    for (int32_t i = kPts - 1; i >= 1; --i)
        if (i > ih && i <= n) {
            p.e[i] = p.e[i - 1];
            p.w[i] = p.w[i - 1];
            p.k[i] = p.k[i - 1];
        }
#pragma unroll
    for (int32_t i = 0; i < kPts; ++i)
        if (i == ih) {
            p.e[i] = em;
            p.w[i] = w;
            p.k[i] = k;
        }
    p.n = n + 1;
    p.ih = ih;
}

// Interpolation table of a probe: x_i = (E_i - el)/width, and the Newton
// divided differences c_i = y[x_0 .. x_i] of y_i = w_i 2^(k_i - max k).
struct PolyTable {
    double x[kPts], c[kPts];
    int32_t n;
};

// The largest binary exponent of the points.
__device__ __forceinline__ int32_t points_kmax(const RefinePoints& p)
{
    int32_t kmax = p.k[0];
#pragma unroll
    for (int32_t i = 1; i < kPts; ++i)
        if (i < p.n)
            kmax = p.k[i] > kmax ? p.k[i] : kmax;
    return kmax;
}

// The table from the points (inv = 1/width), column by column:
// c_i <- (c_i - c_{i-1}) / (x_i - x_{i-j}) for i = n - 1 down to j.  The
// loops run over all kPts slots (predicated on the count) so that the table
// stays in registers.
__device__ __forceinline__ void poly_table(const RefinePoints& p, const double el, const double inv,
                                           const int32_t kmax, PolyTable& t)
{
    const int32_t n = p.n;
    t.n = n;
#pragma unroll
    for (int32_t i = 0; i < kPts; ++i) {
        t.x[i] = 0.0;
        t.c[i] = 0.0;
        if (i < n) {
            t.x[i] = (p.e[i] - el) * inv;
            t.c[i] = __builtin_ldexp(p.w[i], p.k[i] - kmax);
        }
    }
#pragma unroll
    for (int32_t j = 1; j < kPts; ++j)
#pragma unroll
        for (int32_t i = kPts - 1; i >= j; --i)
            if (i < n)
                t.c[i] = (t.c[i] - t.c[i - 1]) / (t.x[i] - t.x[i - j]);
}

// Root in (0, 1) of the table's polynomial by at most 8 Newton steps from
// x0, stopping once a step is at most 2^-30; false when a derivative
// vanishes or a step leaves (0, 1).
__device__ __forceinline__ bool poly_root(const PolyTable& t, const double x0, double& root)
{
    const int32_t n = t.n;
    double top = t.c[2];
#pragma unroll
    for (int32_t i = 3; i < kPts; ++i)
        top = (i == n - 1) ? t.c[i] : top;
    double r = x0;
    for (int32_t it = 0; it < 8; ++it) {
        double f = top, d = 0.0;
#pragma unroll
        for (int32_t i = kPts - 2; i >= 0; --i) {
            if (i <= n - 2) {
                const double u = r - t.x[i];
                d = __builtin_fma(d, u, f);
                f = __builtin_fma(f, u, t.c[i]);
            }
        }
        if (!(d != 0.0))
            return false;
        const double dt = f / d;
        r = r - dt;
        if (!(r > 0.0 && r < 1.0))
            return false;
        if (fabs(dt) <= 0x1p-30)
            break;
    }
    root = r;
    return true;
}

// A probe's plan: the probe itself (`em`) unless it is the root of the
// interpolating polynomial (`root`), which starts from the regula-falsi
// fraction s and falls back to it.
struct ProbePlan {
    double em, el, width, inv, s, tol;
    int32_t kmax;
    bool root;
};

// Refinement state machine shared by both refine kernels.  While the
// bracket holds more than level v (or its floor has no shot) the probe is the
// count-interpolated point and needs a counted shot; once it holds exactly
// level v the probe is the root of the polynomial through the level's
// interpolation points (points_root), started from the regula-falsi point of
// the ends (Anderson-Bjorck scaling), which is the probe when the root search
// fails; every probe is kept at least tol/2 inside, and the probe's count
// follows from the sign of F.  Every shot joins the points.  Stops when the
// width is <= rel_tol max(|E_lo|, |E_hi|) + abs_tol.
// The model transform of interpolation point i (as of the bracket ends).
__device__ __forceinline__ void point_model(RefinePoints& p, const int32_t i, const double rate, const double e0)
{
    double w = p.w[i];
    int32_t k = p.k[i];
    model_apply(rate, e0, p.e[i], w, k);
    p.w[i] = w;
    p.k[i] = k;
}

// The model transform of every interpolation point.
__device__ __forceinline__ void points_model(RefinePoints& p, const double rate, const double e0)
{
    for (int32_t i = 0; i < p.n; ++i)
        point_model(p, i, rate, e0);
}

struct Refiner {
    Bracket b;
    int32_t v, side, it;
    int32_t isolated, reshoot, mid, refit;
    double rate, e0;

    // Regula falsi runs on G = F e^(-b (E - E_hi0)) when the scan gave a
    // model rate (nearly linear where F grows exponentially across the
    // bracket); sign(G) = sign(F), so counts and brackets are unchanged.
    // Without one (refit set), the first isolated probe is the bracket's
    // midpoint and the rate is fitted to the three equally spaced values.
    __device__ __forceinline__ void start_model()
    {
        rate = b.rate;
        e0 = b.eh;
        refit = rate == 0.0 ? 1 : 0;
        if (rate != 0.0) {
            model_apply(rate, e0, b.el, b.wl, b.kl);
            model_apply(rate, e0, b.eh, b.wh, b.kh);
        }
    }

    __device__ __forceinline__ bool converged(const double rel_tol, const double abs_tol) const
    {
        const double tol = __builtin_fma(rel_tol, fmax(fabs(b.el), fabs(b.eh)), abs_tol);
        return (b.eh - b.el) <= tol;
    }

    // The next probe's plan (sets isolated, reshoot, mid and side).
    __device__ __forceinline__ ProbePlan plan(const double rel_tol, const double abs_tol, const RefinePoints& p)
    {
        ProbePlan q{};
        q.el = b.el;
        q.width = b.eh - b.el;
        q.tol = __builtin_fma(rel_tol, fmax(fabs(b.el), fabs(b.eh)), abs_tol);
        q.root = false;
        isolated = b.lo_ok && b.cl == v && b.ch == v + 1;
        reshoot = isolated && b.lo_foreign;
        mid = isolated && !reshoot && refit;
        if (reshoot) {
            // The floor's value came from another band window: shoot it again
            // in this level's window (its count stays the scan's).
            q.em = b.el;
        } else if (mid) {
            q.em = b.el + 0.5 * q.width;
        } else if (isolated) {
            const double r = __builtin_ldexp(b.wh / b.wl, b.kh - b.kl);
            q.s = 1.0 / (1.0 - r);
            q.em = __builtin_fma(q.s, q.width, b.el);
            if (p.n >= 3) {
                q.root = true;
                q.inv = 1.0 / q.width;
                q.kmax = points_kmax(p);
            }
        } else {
            // Several levels in the bracket (or no shot at its floor): probe
            // where level v would sit if the ch - cl levels were evenly
            // spaced, kept within [1/8, 7/8] of the bracket.
            double fr = (static_cast<double>(v - b.cl) + 0.5) / static_cast<double>(b.ch - b.cl);
            fr = fmin(fmax(fr, 0.125), 0.875);
            q.em = __builtin_fma(fr, q.width, b.el);
            side = 0;
        }
        return q;
    }

    // The probe from its plan and, for a root plan, the polynomial's table;
    // an isolated bracket's probe is kept at least tol/2 inside.
    __device__ __forceinline__ double finish(const ProbePlan& q, const PolyTable* t) const
    {
        double em = q.em;
        if (!isolated || reshoot || mid)
            return em;
        double x;
        if (q.root && poly_root(*t, fmin(fmax(q.s, 0.0), 1.0), x))
            em = __builtin_fma(x, q.width, q.el);
        const double d = 0.5 * q.tol;
        if (!(em >= b.el + d))
            em = b.el + d;
        if (!(em <= b.eh - d))
            em = b.eh - d;
        return em;
    }

    // Plan, table and probe in one lane.
    __device__ __forceinline__ double probe(const double rel_tol, const double abs_tol, const RefinePoints& p)
    {
        const ProbePlan q = plan(rel_tol, abs_tol, p);
        PolyTable t;
        if (q.root)
            poly_table(p, q.el, q.inv, q.kmax, t);
        return finish(q, &t);
    }

    // Take the probe's result; `c` is ignored when the bracket is isolated.
    // True when a midpoint shot gave the level its rate: the caller then
    // applies the transform to every interpolation point (the shot is
    // inserted untransformed, so all of them take it alike).  kMidShot false
    // leaves out the midpoint's rate fit, for callers that never pass one.
    template <bool kMidShot = true>
    __device__ __forceinline__ bool update(const double em, double w, int32_t k, int32_t c, RefinePoints& p)
    {
        if (isolated)
            c = reshoot ? v : ((static_cast<int32_t>(sign_bit(w)) == (v & 1)) ? v : v + 1);
        bool fitted = false;
        if (kMidShot && mid) {
            // The rate through (el, em, eh), em above the root when c > v;
            // the ends take the transform from here on.
            refit = 0;
            const double br = model_rate(b.wl, b.kl, w, k, b.wh, b.kh, c > v, 0.5 * (b.eh - b.el));
            if (br != 0.0) {
                rate = br;
                model_apply(rate, e0, b.el, b.wl, b.kl);
                model_apply(rate, e0, b.eh, b.wh, b.kh);
                fitted = true;
            }
        }
        const double wr = w;
        const int32_t kr = k;
        if (rate != 0.0)
            model_apply(rate, e0, em, w, k);
        ++it;
        points_insert(p, b.el + 0.5 * (b.eh - b.el), em, fitted ? wr : w, fitted ? kr : k);
        if (c <= v)
            ++p.ih;
        // Anderson-Bjorck: when the same end is replaced twice in a row, the
        // kept end's value is scaled by m = 1 - F_new/F_replaced (0.5 when
        // m <= 0 or undefined).
        if (c <= v) {
            if (isolated && side == -1) {
                const double m = 1.0 - __builtin_ldexp(w / b.wl, k - b.kl);
                b.wh *= (m > 0.0) ? m : 0.5;
            }
            b.el = em;
            b.cl = c;
            b.wl = w;
            b.kl = k;
            b.lo_ok = 1;
            b.lo_foreign = 0;
            side = isolated ? -1 : 0;
        } else {
            if (isolated && side == 1) {
                const double m = 1.0 - __builtin_ldexp(w / b.wh, k - b.kh);
                b.wl *= (m > 0.0) ? m : 0.5;
            }
            b.eh = em;
            b.ch = c;
            b.wh = w;
            b.kh = k;
            side = isolated ? 1 : 0;
        }
        return fitted;
    }

    __device__ __forceinline__ double level() const { return b.el + 0.5 * (b.eh - b.el); }
};

__device__ __forceinline__ void write_level(double* levels, int32_t* status, int32_t* iterations,
                                            const int64_t out, const double lv, const int32_t st,
                                            const int32_t it)
{
    levels[out] = lv;
    status[out] = st;
    iterations[out] = it;
}

}  // namespace

// Lane-per-level refinement: lane = level v of the wavefront's potential,
// sequential difference-form shots over the wave-uniform uv table.  With
// K = 2 a lane refines levels first + lane and first + lane + 64 of a
// 128-level task together (one paired shot per iteration; a member that has
// converged rides along on its partner's probe and its result is dropped).
// Every lane walks the same table, so the scan must have used one window
// (the whole grid) in every band.
template <int K>
__global__ __launch_bounds__(256) void numerov_refine_kernel(
    const double* __restrict__ uv, const PotentialRecord* __restrict__ pots,
    const BandWindow* __restrict__ windows, const WaveTask* __restrict__ tasks, const int32_t n_tasks,
    const Bracket* __restrict__ brackets, RefinePoints* __restrict__ points, const double rel_tol,
    const int32_t max_iter, double* __restrict__ levels, int32_t* __restrict__ status,
    int32_t* __restrict__ iterations, int32_t* __restrict__ steps)
{
    static_assert(K == 1 || K == 2, "one or two levels per lane");
    const int32_t tid = wave_task_index();
    if (tid >= n_tasks)
        return;
    const WaveTask task = tasks[tid];
    const PotentialRecord& P = pots[task.potential];
    const BandWindow* W = windows + task.potential * kBands;
    const int32_t lane = static_cast<int32_t>(threadIdx.x & 63);
    Refiner r[K];
    bool live[K];
    int64_t out[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const int32_t l = task.first + lane + 64 * q;
        live[q] = false;
        out[q] = P.level_offset + l;
        r[q] = Refiner{};
        if (l >= P.level_count)
            continue;
        if (P.flags != 0) {
            write_level(levels, status, iterations, out[q], __builtin_nan(""), 3, 0);
            continue;
        }
        r[q].v = P.level_min + l;
        r[q].b = brackets[out[q]];
        steps[out[q]] = W[0].hi - W[0].lo - 1;
        r[q].start_model();
        if (!r[q].b.found) {
            write_level(levels, status, iterations, out[q], __builtin_nan(""), 1, 0);
            continue;
        }
        if (r[q].rate != 0.0)
            points_model(points[out[q]], r[q].rate, r[q].e0);
        live[q] = true;
    }
    const double* u = uv + P.ut_offset + W[0].lo;
    const int32_t n = W[0].hi - W[0].lo + 1;
    for (;;) {
        double em[K];
        bool any = false;
#pragma unroll
        for (int q = 0; q < K; ++q) {
            em[q] = 0.0;
            if (!live[q])
                continue;
            if (r[q].converged(rel_tol, P.abs_tol)) {
                write_level(levels, status, iterations, out[q], r[q].level(), 0, r[q].it);
                live[q] = false;
            } else if (r[q].it >= max_iter) {
                write_level(levels, status, iterations, out[q], r[q].level(), 2, r[q].it);
                live[q] = false;
            } else {
                em[q] = r[q].probe(rel_tol, P.abs_tol, points[out[q]]);
                any = true;
            }
        }
        if (!any)
            break;
        if constexpr (K == 1) {
            double w;
            int32_t k;
            const int32_t c = shoot<DiffForm>(u, n, P.beta * em[0], P.stride, w, k);
            if (r[0].update(em[0], w, k * kRescaleLog2, c, points[out[0]]))
                points_model(points[out[0]], r[0].rate, r[0].e0);
        } else {
            const double ea = live[0] ? em[0] : em[1];
            const double eb = live[1] ? em[1] : em[0];
            DiffPair s;
            DiffPair::Counter c;
            shoot<DiffPair>(u, n, DiffPair::Energy{P.beta * ea, P.beta * eb}, P.stride, s, c);
            if (live[0] && r[0].update(em[0], s.x.w, s.x.rescales * kRescaleLog2, c.a.nodes, points[out[0]]))
                points_model(points[out[0]], r[0].rate, r[0].e0);
            if (live[1] && r[1].update(em[1], s.y.w, s.y.rescales * kRescaleLog2, c.b.nodes, points[out[1]]))
                points_model(points[out[1]], r[1].rate, r[1].e0);
        }
    }
}

namespace {

// ---- segmented difference-form shots (one block per level) -------------

__device__ __forceinline__ int32_t frexp_exponent(const double m)
{
    int e = 0;
    (void)__builtin_frexp(m, &e);
    return e;
}

__device__ __forceinline__ int64_t segment_start(const int32_t q, const int32_t segs, const int32_t s)
{
    return (static_cast<int64_t>(s) * q) / segs;
}

struct SegCount {
    int32_t internal, first, last, has;
};

// Values per VGPR buffer of the segment walks (Body::kChunk): three rotating
// buffers of 8 keep the walk's registers low enough for three wavefronts per
// SIMD (four-level walks run two per SIMD; buffers of 4 values would not
// bring them to three).
template <int C, class Body>
__device__ __forceinline__ void walk_chunk(const double (&buf)[C], const int32_t j, Body& body)
{
#pragma unroll
    for (int k = 0; k < C; ++k)
        body(buf[k], j + k);
}

template <int C, class Body>
__device__ __forceinline__ void walk_tail(const double (&buf)[C], const int32_t j, const int32_t len, Body& body)
{
#pragma unroll
    for (int k = 0; k < C; ++k)
        if (j + k < len)
            body(buf[k], j + k);
}

// Word 3 of a buffer resource for raw 32-bit-element reads on CDNA
// (DATA_FORMAT = 4, i.e. 32-bit elements; everything else zero).
constexpr int32_t kBufferResourceWord3 = 0x00020000;

// Grid values of one segment, value j being the j-th step's point (j
// wave-uniform).  RowSource reads the natural uv table (lane-private
// addresses); ColumnSource reads the lane's column of the segment table
// through a wave-uniform buffer resource (row r at byte off + r * 512: the
// lane's column offset in the vector offset, the row's in the scalar offset),
// so each load instruction of a wavefront covers 512 contiguous bytes and
// needs no address arithmetic on the vector ALU.
struct RowSource {
    const double* p;
    __device__ __forceinline__ double at(const int32_t j) const { return p[j]; }
};

struct ColumnSource {
    __amdgpu_buffer_rsrc_t rs;
    int32_t off;
    __device__ __forceinline__ double row(const int32_t r) const
    {
        typedef int32_t v2i __attribute__((ext_vector_type(2)));
        const v2i x = __builtin_amdgcn_raw_buffer_load_b64(rs, off, r * 512, 0);
        return __builtin_bit_cast(double, x);
    }
    __device__ __forceinline__ double at(const int32_t j) const { return row(j + 1); }
};

template <int C, class Src>
__device__ __forceinline__ void load_chunk(double (&buf)[C], const Src& src, const int32_t j)
{
#pragma unroll
    for (int k = 0; k < C; ++k)
        buf[k] = src.at(j + k);
}

// Visit the `len` values src.at(0 .. len - 1) in order, body(value, index);
// ulen is wave-uniform with ulen <= len <= ulen + 1 (segment lengths differ by
// at most one), so the step index stays wave-uniform and only the last step
// is predicated per lane.  Three C-value VGPR buffers rotate so that every
// load is issued two chunks before its use; vector-memory loads return in
// order, so the waits are partial.  Also reads the 2 C values that follow the
// segment (the tables are padded).
template <int C, class Src, class Body>
__device__ __forceinline__ void walk_segment(const Src& src, const int32_t len, const int32_t ulen, Body& body)
{
    double A[C], B[C], D[C];
    load_chunk(A, src, 0);
    load_chunk(B, src, C);
    int32_t j = 0;
    for (; j + 3 * C <= ulen; j += 3 * C) {
        load_chunk(D, src, j + 2 * C);
        walk_chunk(A, j, body);
        load_chunk(A, src, j + 3 * C);
        walk_chunk(B, j + C, body);
        load_chunk(B, src, j + 4 * C);
        walk_chunk(D, j + 2 * C, body);
    }
    const int32_t rem = ulen - j;
    if (rem >= 2 * C) {
        load_chunk(D, src, j + 2 * C);
        walk_chunk(A, j, body);
        walk_chunk(B, j + C, body);
        walk_tail(D, j + 2 * C, len, body);
    } else if (rem >= C) {
        walk_chunk(A, j, body);
        walk_tail(B, j + C, len, body);
    } else {
        walk_tail(A, j, len, body);
    }
}

// Two difference-form columns through one segment: the segment's transfer
// matrix on (w, d).  All six carried values are renormalised after every
// 128th step: a step grows them by less than 2^2.87 and shrinks the largest
// by less than 2^4.5 (its matrix has determinant f_{i-1}/f_{i+1} > 1/3 and
// norm below 7.4), so 128 steps stay inside the normal range, and scaling by
// a power of two is exact.
struct MatrixBody {
    double e, em;
    double wa, da, pa, wb, db, pb;
    int32_t ex;
    double c10;
    __device__ __forceinline__ void operator()(const double u, const int32_t idx)
    {
        const double ep = e - u;
        const double e10 = ep * c10;
        const double sg = __builtin_fma(ep, em, ep + em);
        da = __builtin_fma(-e10, wa, __builtin_fma(-sg, pa, da));
        db = __builtin_fma(-e10, wb, __builtin_fma(-sg, pb, db));
        pa = wa;
        pb = wb;
        wa = wa + da;
        wb = wb + db;
        em = ep;
        if ((idx & 127) == 127) {
            const double mx = fmax(fmax(fmax(fabs(wa), fabs(da)), fmax(fabs(pa), fabs(wb))),
                                   fmax(fabs(db), fabs(pb)));
            const int32_t sh = frexp_exponent(mx);
            wa = __builtin_ldexp(wa, -sh);
            da = __builtin_ldexp(da, -sh);
            pa = __builtin_ldexp(pa, -sh);
            wb = __builtin_ldexp(wb, -sh);
            db = __builtin_ldexp(db, -sh);
            pb = __builtin_ldexp(pb, -sh);
            ex += sh;
        }
    }
};

// KK levels' transfer matrices through one segment of the same window: the
// bodies share every grid value, so each table load feeds all of them (the
// walk of one or two levels is bound by those loads, not by its FP64 work).
// Each member's arithmetic is that of a lone MatrixBody.
template <int KK>
struct MatrixGroup {
    static constexpr int kChunk = 8;
    MatrixBody m[KK];
    __device__ __forceinline__ void operator()(const double u, const int32_t idx)
    {
#pragma unroll
        for (int q = 0; q < KK; ++q)
            m[q](u, idx);
    }
};

// One difference-form solution through a segment, sampling signs at the
// global sample positions (q + 1) % stride == 0 and the last point (the
// stride is a power of two: mask = stride - 1; last = qn - 1 when qn is not
// a multiple of the stride, else -1).
struct CountBody {
    static constexpr int kChunk = 8;
    double e, em, w, d, wm;
    int32_t q0, mask, last;
    SegCount sc;
    double c10;
    __device__ __forceinline__ void operator()(const double u, const int32_t idx)
    {
        const double ep = e - u;
        const double e10 = ep * c10;
        const double sg = __builtin_fma(ep, em, ep + em);
        d = __builtin_fma(-e10, w, __builtin_fma(-sg, wm, d));
        wm = w;
        w = w + d;
        em = ep;
        const int32_t q = q0 + idx;
        if (((q + 1) & mask) == 0 || q == last) {
            const int32_t sb = static_cast<int32_t>(sign_bit(w));
            if (sc.has)
                sc.internal += sb ^ sc.last;
            else
                sc.first = sb;
            sc.last = sb;
            sc.has = 1;
        }
        if ((idx & 127) == 127) {
            const int32_t sh = frexp_exponent(fmax(fmax(fabs(w), fabs(d)), fabs(wm)));
            w = __builtin_ldexp(w, -sh);
            d = __builtin_ldexp(d, -sh);
            wm = __builtin_ldexp(wm, -sh);
        }
    }
};

}  // namespace

// Segment tables of the 64-lane segmented refinement: for job J, rows
// r < seg_rows(J.n) of 64 columns from row J.seg_row, column s holding the
// window's uv at point min(q0_s + r, n - 1) for the segments s < seg_count(n)
// (zero in the unused columns).  One workgroup per 64-row tile of a job: each
// wavefront reads 64 consecutive points of one segment per load (512
// contiguous bytes), the tile is transposed through LDS, and each wavefront
// then stores 64 consecutive values of one row.
constexpr int32_t kTableTile = 64;
__global__ __launch_bounds__(256) void segment_table_kernel(const double* __restrict__ uv,
                                                            const PotentialRecord* __restrict__ pots,
                                                            const TableJob* __restrict__ jobs,
                                                            const int32_t n_jobs, double* __restrict__ segtab)
{
    __shared__ double tile[64][kTableTile + 1];
    const int32_t lane = static_cast<int32_t>(threadIdx.x & 63);
    const int32_t wave = static_cast<int32_t>(threadIdx.x >> 6);
    for (int32_t jb = static_cast<int32_t>(blockIdx.y); jb < n_jobs; jb += static_cast<int32_t>(gridDim.y)) {
        const TableJob J = jobs[jb];
        const int32_t n = J.n;
        const int32_t rows = seg_rows(n, 0, 1);
        const int32_t qn = n - 2;
        const int32_t segs = seg_count(n);
        const double* u = uv + pots[J.potential].ut_offset + J.lo;
        double* out = segtab + static_cast<int64_t>(J.seg_row) * 64;
        for (int32_t r0 = static_cast<int32_t>(blockIdx.x) * kTableTile; r0 < rows;
             r0 += static_cast<int32_t>(gridDim.x) * kTableTile) {
            // Columns: wavefront w reads segments w, w + 4, ...; lane = row.
            for (int32_t sg = wave; sg < 64; sg += 4) {
                double x = 0.0;
                if (sg < segs) {
                    const int64_t q = segment_start(qn, segs, sg) + r0 + lane;
                    x = u[q < n - 1 ? q : n - 1];
                }
                tile[sg][lane] = x;
            }
            __syncthreads();
            // Rows: wavefront w stores rows r0 + w, r0 + w + 4, ...; lane = column.
            for (int32_t r = wave; r < kTableTile && r0 + r < rows; r += 4)
                out[static_cast<int64_t>(r0 + r) * 64 + lane] = tile[lane][r];
            __syncthreads();
        }
    }
}

namespace {

// Shared memory of the segmented refinement (one shot at a time): the
// segment matrices (entries and binary exponents), the chained start vectors
// and segment node counts of a counted shot, and the joined count.
template <int T>
struct SegShared {
    double m[4][T];
    int32_t e[T];
    double start[2][T];
    SegCount cnt[T];
    int32_t c;
};

// One level's shot geometry in a T-thread block: its band window (Dirichlet
// points lo, hi on the trimmed grid), the q = n - 2 steps split into `segs`
// segments, lane s walking [q0, q1) from the table value u_start.
template <bool TABLE>
struct SegGeom {
    const double* u;
    int32_t lo, hi, seg_row, n, qn, segs, q0, q1, ulen;
    bool active;
    double u_start;
    ColumnSource cs;

    // Lane `lane`'s steps [q0, q1) of the window (once per level):
    // floor(s q / sg) = s (q / sg) + floor(s (q mod sg) / sg), so 32-bit
    // arithmetic gives segment_start's values (s (q mod sg) < 2^16).
    static __device__ __forceinline__ int2 bounds(const BandWindow& B, const int32_t lane, const int32_t T)
    {
        const int32_t q = B.hi - B.lo - 1;
        const int32_t sg = min(T, max(1, q / 32));
        if (lane >= sg)
            return make_int2(0, 0);
        const int32_t base = q / sg, rem = q % sg;
        return make_int2(lane * base + (lane * rem) / sg, (lane + 1) * base + ((lane + 1) * rem) / sg);
    }

    __device__ __forceinline__ void init(const double* uv_pot, const double* segtab, const BandWindow& B,
                                         const int2 qb, const int32_t lane, const int32_t T)
    {
        lo = B.lo;
        hi = B.hi;
        seg_row = B.seg_row;
        u = uv_pot + B.lo;
        n = B.hi - B.lo + 1;
        qn = n - 2;
        segs = min(T, max(1, qn / 32));
        active = lane < segs;
        q0 = qb.x;
        q1 = qb.y;
        // Every segment has qn / segs or qn / segs + 1 steps.
        ulen = __builtin_amdgcn_readfirstlane(qn / segs);
        cs = ColumnSource{};
        if constexpr (TABLE) {
            // With the segment table, lane s reads column s (row r = point
            // q0 + r): consecutive lanes load consecutive addresses.  The
            // resource must be wave-uniform (scalar registers); the
            // readfirstlane calls state what the compiler cannot prove.
            const int32_t rows = seg_rows(n, 0, 1);
            const uint64_t base = reinterpret_cast<uint64_t>(segtab + static_cast<int64_t>(B.seg_row) * 64);
            const uint32_t blo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(base));
            const uint32_t bhi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(base >> 32));
            const int32_t bytes = __builtin_amdgcn_readfirstlane(rows * 512);
            cs.rs = __builtin_amdgcn_make_buffer_rsrc(
                reinterpret_cast<double*>((static_cast<uint64_t>(bhi) << 32) | blo), static_cast<int16_t>(0),
                bytes, kBufferResourceWord3);
            cs.off = lane * 8;
            u_start = cs.row(0);
        } else {
            u_start = u[q0];
        }
    }

    template <class Body>
    __device__ __forceinline__ void walk(Body& body) const
    {
        if constexpr (TABLE)
            walk_segment<Body::kChunk>(cs, q1 - q0, ulen, body);
        else
            walk_segment<Body::kChunk>(RowSource{u + q0 + 1}, q1 - q0, ulen, body);
    }
};

// A lane's segment matrix in registers (entries m00 = w column of (1, 0),
// m01 = w column of (0, 1), m10, m11 the d rows), times 2^e.
struct SegMat {
    double m00, m01, m10, m11;
    int32_t e;
};

// The walk's matrix divided by the frexp exponent of its largest entry.
__device__ __forceinline__ SegMat normalized(const MatrixBody& mb)
{
    const double mx = fmax(fmax(fabs(mb.wa), fabs(mb.da)), fmax(fabs(mb.wb), fabs(mb.db)));
    const int32_t s = frexp_exponent(mx);
    return SegMat{__builtin_ldexp(mb.wa, -s), __builtin_ldexp(mb.wb, -s), __builtin_ldexp(mb.da, -s),
                  __builtin_ldexp(mb.db, -s), mb.ex + s};
}

template <int T>
__device__ __forceinline__ void store_seg(SegShared<T>& sh, const int32_t lane, const SegMat& a)
{
    sh.m[0][lane] = a.m00;
    sh.m[1][lane] = a.m01;
    sh.m[2][lane] = a.m10;
    sh.m[3][lane] = a.m11;
    sh.e[lane] = a.e;
}

// Value of lane + D (lanes whose partner lies outside the wavefront get an
// unused value): data-parallel moves within 16-lane rows for D < 16, a
// swizzle within 32-lane halves for D = 16, lane 32 for D = 32 (used by lane
// 0 only).
template <int D>
__device__ __forceinline__ int32_t lane_plus(const int32_t x)
{
    if constexpr (D < 16)
        return __builtin_amdgcn_mov_dpp(x, 0x100 + D, 0xF, 0xF, false);  // row_shl:D
    else if constexpr (D == 16)
        return __builtin_amdgcn_ds_swizzle(x, 0x401F);  // lane ^ 16 within 32 lanes
    else
        return __builtin_amdgcn_readlane(x, 32);
}

template <int D>
__device__ __forceinline__ double lane_plus(const double x)
{
    const int2 v = __builtin_bit_cast(int2, x);
    return __builtin_bit_cast(double, make_int2(lane_plus<D>(v.x), lane_plus<D>(v.y)));
}

// One level of the binary-tree product in registers: lane i (a multiple of
// 2D with i + D < segs) takes M_{i+D} M_i; the same operations as tree_step.
template <int D>
__device__ __forceinline__ void tree_round(SegMat& a, const int32_t lane, const int32_t segs)
{
    const double b00 = lane_plus<D>(a.m00), b01 = lane_plus<D>(a.m01);
    const double b10 = lane_plus<D>(a.m10), b11 = lane_plus<D>(a.m11);
    const int32_t be = lane_plus<D>(a.e);
    if ((lane & (2 * D - 1)) == 0 && lane + D < segs) {
        const double c00 = __builtin_fma(b00, a.m00, b01 * a.m10);
        const double c01 = __builtin_fma(b00, a.m01, b01 * a.m11);
        const double c10 = __builtin_fma(b10, a.m00, b11 * a.m10);
        const double c11 = __builtin_fma(b10, a.m01, b11 * a.m11);
        const int32_t s = frexp_exponent(fmax(fmax(fabs(c00), fabs(c01)), fmax(fabs(c10), fabs(c11))));
        a.m00 = __builtin_ldexp(c00, -s);
        a.m01 = __builtin_ldexp(c01, -s);
        a.m10 = __builtin_ldexp(c10, -s);
        a.m11 = __builtin_ldexp(c11, -s);
        a.e = be + a.e + s;
    }
}

// The binary-tree products of a wavefront's segment matrices in registers,
// one per level of the shot (d = 1, 2, 4, ... while d < segs; the levels'
// rounds interleaved, so their dependency chains overlap); lane 0 ends with
// each product of all segments.  Every lane of the wavefront takes part (the
// cross-lane moves read all lanes).
template <int D, int KK>
__device__ __forceinline__ void tree_rounds(SegMat (&a)[KK], const int32_t lane, const int32_t segs)
{
    if (segs > D) {
#pragma unroll
        for (int q = 0; q < KK; ++q)
            tree_round<D>(a[q], lane, segs);
    }
}

template <int KK>
__device__ __forceinline__ void trees_in_registers(SegMat (&a)[KK], const int32_t lane, const int32_t segs)
{
    tree_rounds<1>(a, lane, segs);
    tree_rounds<2>(a, lane, segs);
    tree_rounds<4>(a, lane, segs);
    tree_rounds<8>(a, lane, segs);
    tree_rounds<16>(a, lane, segs);
    tree_rounds<32>(a, lane, segs);
}

// Start vector of every segment of a counted shot: the matrices chained from
// (w_1, d_1) = (1, 1), renormalised after each segment (one lane).
template <int T>
__device__ __forceinline__ void chain_starts(SegShared<T>& sh, const int32_t segs)
{
    double v0 = 1.0, v1 = 1.0;
    for (int32_t s = 0; s < segs; ++s) {
        sh.start[0][s] = v0;
        sh.start[1][s] = v1;
        const double a = __builtin_fma(sh.m[0][s], v0, sh.m[1][s] * v1);
        const double b = __builtin_fma(sh.m[2][s], v0, sh.m[3][s] * v1);
        const int32_t x = frexp_exponent(fmax(fabs(a), fabs(b)));
        v0 = __builtin_ldexp(a, -x);
        v1 = __builtin_ldexp(b, -x);
    }
}

// One level of the binary-tree product: M_lane <- M_{lane+d} M_lane.
template <int T>
__device__ __forceinline__ void tree_step(SegShared<T>& sh, const int32_t lane, const int32_t d)
{
    const double a00 = sh.m[0][lane + d], a01 = sh.m[1][lane + d];
    const double a10 = sh.m[2][lane + d], a11 = sh.m[3][lane + d];
    const double b00 = sh.m[0][lane], b01 = sh.m[1][lane];
    const double b10 = sh.m[2][lane], b11 = sh.m[3][lane];
    const double c00 = __builtin_fma(a00, b00, a01 * b10);
    const double c01 = __builtin_fma(a00, b01, a01 * b11);
    const double c10 = __builtin_fma(a10, b00, a11 * b10);
    const double c11 = __builtin_fma(a10, b01, a11 * b11);
    const int32_t s = frexp_exponent(fmax(fmax(fabs(c00), fabs(c01)), fmax(fabs(c10), fabs(c11))));
    sh.m[0][lane] = __builtin_ldexp(c00, -s);
    sh.m[1][lane] = __builtin_ldexp(c01, -s);
    sh.m[2][lane] = __builtin_ldexp(c10, -s);
    sh.m[3][lane] = __builtin_ldexp(c11, -s);
    sh.e[lane] = sh.e[lane + d] + sh.e[lane] + s;
}

// Node count of a counted shot from its segments' counts (one lane).
template <int T>
__device__ __forceinline__ int32_t join_counts(const SegShared<T>& sh, const int32_t segs)
{
    int32_t total = 0, prev = 0;
    for (int32_t s = 0; s < segs; ++s) {
        const SegCount sc = sh.cnt[s];
        if (sc.has) {
            total += sc.internal + (sc.first ^ prev);
            prev = sc.last;
        }
    }
    return total;
}

// One segmented shot of KK levels (1, 2 or 4 sharing the window of g; only a
// lone level's shot may be counted): every lane builds its segment's matrix
// for each energy e[k], then the tree product gives the far end F = fw 2^fk
// of each.  A counted shot stores the matrices in LDS, chains them (lane 0)
// for the segment start vectors, walks every segment again and joins the
// node count across segment borders.  In 64-lane blocks (one wavefront) the
// tree runs in registers.
template <int T, int KK, bool TABLE>
__device__ __forceinline__ void segmented_shot(SegShared<T>& sh, const SegGeom<TABLE>& g, const int32_t lane,
                                               const double (&e)[KK], const bool counted, const int32_t stride,
                                               const StepConsts& k, double (&fw)[KK], int32_t (&fk)[KK],
                                               int32_t& c)
{
    static_assert(KK == 1 || KK == 2 || KK == 4, "1, 2 or 4 levels per shot");
    const int32_t segs = g.segs;
    SegMat a[KK];
#pragma unroll
    for (int q = 0; q < KK; ++q)
        a[q] = SegMat{1.0, 0.0, 0.0, 1.0, 0};
    if (g.active) {
        MatrixGroup<KK> mg;
#pragma unroll
        for (int q = 0; q < KK; ++q)
            mg.m[q] = MatrixBody{e[q], e[q] - g.u_start, 1.0, 0.0, 1.0, 0.0, 1.0, -1.0, 0, k.c10};
        g.walk(mg);
#pragma unroll
        for (int q = 0; q < KK; ++q)
            a[q] = normalized(mg.m[q]);
    }
    constexpr bool kRegisterTree = T == 64;
    if (!kRegisterTree || counted) {
        if (g.active)
            store_seg(sh, lane, a[0]);
        __syncthreads();
    }
    if (counted) {
        if (lane == 0)
            chain_starts(sh, segs);
        __syncthreads();
    }
    if constexpr (kRegisterTree) {
        trees_in_registers(a, lane, segs);
#pragma unroll
        for (int q = 0; q < KK; ++q) {
            // Lane 0's product, to every lane.
            const int2 f = __builtin_bit_cast(int2, __builtin_fma(a[q].m00, 1.0, a[q].m01 * 1.0));
            fw[q] = __builtin_bit_cast(double, make_int2(__builtin_amdgcn_readlane(f.x, 0),
                                                         __builtin_amdgcn_readlane(f.y, 0)));
            fk[q] = __builtin_amdgcn_readlane(a[q].e, 0);
        }
    } else {
        for (int32_t d = 1; d < segs; d *= 2) {
            if ((lane & (2 * d - 1)) == 0 && lane + d < segs)
                tree_step(sh, lane, d);
            __syncthreads();
        }
        fw[0] = __builtin_fma(sh.m[0][0], 1.0, sh.m[1][0] * 1.0);
        fk[0] = sh.e[0];
    }
    c = 0;
    if (!counted)
        return;
    if (g.active) {
        const double w0 = sh.start[0][lane], d0 = sh.start[1][lane];
        CountBody cb{e[0], e[0] - g.u_start, w0, d0, w0 - d0, g.q0, stride - 1,
                     (g.qn & (stride - 1)) != 0 ? g.qn - 1 : -1, {0, 0, 0, 0}, k.c10};
        g.walk(cb);
        sh.cnt[lane] = cb.sc;
    }
    __syncthreads();
    if (lane == 0)
        sh.c = join_counts(sh, segs);
    __syncthreads();
    c = sh.c;
}

}  // namespace

// Refinement state of a segmented block, in LDS: thread q < K runs level
// q's state machine (probe and update) between shots and publishes its probe,
// so the walks carry no refinement state in registers.
template <int K>
struct SegControl {
    Refiner r[K];
    RefinePoints pts[K];
    ProbePlan plan[K];
    PolyTable tab[K];
    double em[K];
    int32_t st[K];
    int32_t live[K];
    int32_t counted[K];
    int32_t root[K];
    int32_t remodel[K];
    // A midpoint shot's result, taken at the start of the next round.
    double held_w[K];
    int32_t held_k[K];
    int32_t held[K];
};

// Value of lane - J within its 16-lane row (J < 16; lanes whose source lies
// outside the row get an unused value).
template <int J>
__device__ __forceinline__ double lane_minus(const double x)
{
    const int2 v = __builtin_bit_cast(int2, x);
    return __builtin_bit_cast(double, make_int2(__builtin_amdgcn_mov_dpp(v.x, 0x110 + J, 0xF, 0xF, false),
                                                __builtin_amdgcn_mov_dpp(v.y, 0x110 + J, 0xF, 0xF, false)));
}

// The interpolation tables of the levels whose probe is a polynomial root,
// built by one 16-lane row per level (lane 16 q + i holds point i of level
// q): the same operations as poly_table, one column per step.  Every lane of
// the wavefront executes the cross-lane moves.
template <int K>
__device__ __forceinline__ void tables_in_lanes(SegControl<K>& ctl, const int32_t lane)
{
    static_assert(kPts <= 16 && K * 16 <= 64, "a row of lanes per level");
    const int32_t q = lane / 16, i = lane % 16;
    const bool mine = q < K && i < kPts && ctl.root[q] != 0;
    double x = 0.0, c = 0.0;
    int32_t n = 0;
    if (mine) {
        const RefinePoints& p = ctl.pts[q];
        const ProbePlan& pl = ctl.plan[q];
        n = p.n;
        if (i < n) {
            x = (p.e[i] - pl.el) * pl.inv;
            c = __builtin_ldexp(p.w[i], p.k[i] - pl.kmax);
        }
    }
#define EPSEON_COLUMN(J)                                    \
    {                                                       \
        const double cp = lane_minus<1>(c);                 \
        const double xp = lane_minus<J>(x);                 \
        if (mine && i >= J && i < n)                        \
            c = (c - cp) / (x - xp);                        \
    }
    EPSEON_COLUMN(1)
    EPSEON_COLUMN(2)
    EPSEON_COLUMN(3)
    EPSEON_COLUMN(4)
    EPSEON_COLUMN(5)
    EPSEON_COLUMN(6)
    EPSEON_COLUMN(7)
    EPSEON_COLUMN(8)
    EPSEON_COLUMN(9)
    EPSEON_COLUMN(10)
    EPSEON_COLUMN(11)
#undef EPSEON_COLUMN
    static_assert(kPts == 12, "one column step per point after the first");
    if (mine) {
        if (i == 0)
            ctl.tab[q].n = n;
        ctl.tab[q].x[i] = x;
        ctl.tab[q].c[i] = c;
    }
}

// Segmented refinement: one T-thread block per level (K = 1) or per group of
// K = 2 or 4 consecutive levels (64-lane blocks).  Lane s owns steps
// [q_s, q_{s+1}) of a level's window (S = clamp(Q/32, 1, T) segments) and
// builds that segment's 2x2 difference-form transfer matrix (columns
// (w, d) = (1, 0) and (0, 1), renormalised every 128 steps and at the end).
// The far-end value is the binary-tree product of the matrices
// (M_i <- M_{i+d} M_i at d = 1, 2, 4, ...) applied to (w_1, d_1) = (1, 1).
// Counted shots chain the matrices sequentially (lane 0) for the segment
// start vectors and re-run every segment.  In each round the live levels of
// a group whose shots use the same window walk it together (up to four per
// walk, three padded with a copy of the first), each grid value loaded once
// for all of them; counted shots and levels of other windows shoot apart.
// Every level's arithmetic is that of a block of its own.
template <int T, bool TABLE, int K>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(K == 4 ? 2 : 3))) void
numerov_refine_segmented_kernel(const double* __restrict__ uv, const double* __restrict__ segtab,
                                const PotentialRecord* __restrict__ pots, const BandWindow* __restrict__ windows,
                                const WaveTask* __restrict__ tasks, const int32_t n_tasks,
                                const int32_t* __restrict__ bins, const Bracket* __restrict__ brackets,
                                RefinePoints* __restrict__ points, const double rel_tol, const int32_t max_iter,
                                double* __restrict__ levels, int32_t* __restrict__ status,
                                int32_t* __restrict__ iterations, int32_t* __restrict__ steps)
{
    static_assert(K == 1 || K == 2 || K == 4, "1, 2 or 4 levels per block");
    static_assert(K == 1 || T == 64, "level groups need 64-lane blocks");
    __shared__ SegShared<T> sh;
    __shared__ SegControl<K> ctl;

    WaveTask task;
    if (!fetch_block_task(tasks, n_tasks, bins, task))
        return;
    const PotentialRecord& P = pots[task.potential];
    const int32_t lane = static_cast<int32_t>(threadIdx.x);
    const BandWindow* W = windows + task.potential * kBands;
    if (lane < K) {
        const int32_t q = lane;
        const int32_t l = task.first + q;
        const int64_t out = P.level_offset + l;
        Refiner& r = ctl.r[q];
        r = Refiner{};
        ctl.live[q] = 0;
        ctl.st[q] = 0;
        ctl.remodel[q] = 0;
        ctl.held[q] = 0;
        ctl.pts[q].n = 0;
        if (l < P.level_count) {
            if (P.flags != 0) {
                write_level(levels, status, iterations, out, __builtin_nan(""), 3, 0);
            } else {
                r.v = P.level_min + l;
                r.b = brackets[out];
                r.start_model();
                if (!r.b.found) {
                    write_level(levels, status, iterations, out, __builtin_nan(""), 1, 0);
                } else {
                    ctl.live[q] = 1;
                    ctl.remodel[q] = r.rate != 0.0 ? 1 : 0;
                    ctl.pts[q] = points[out];
                    steps[out] = W[r.b.band].hi - W[r.b.band].lo - 1;
                }
            }
        }
    }
    __syncthreads();
    const double* uv_pot = uv + P.ut_offset;
    // Each level's window and segment bounds stay in registers between
    // shots; wid[q] is the first level of the group with the same window.
    int2 qb[K];
    int32_t band[K], wid[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
        band[q] = __builtin_amdgcn_readfirstlane(ctl.live[q] ? ctl.r[q].b.band : 0);
        qb[q] = ctl.live[q] ? SegGeom<TABLE>::bounds(W[band[q]], lane, T) : make_int2(0, 0);
        wid[q] = q;
#pragma unroll
        for (int o = q - 1; o >= 0; --o) {
            const BandWindow& a = W[band[o]];
            const BandWindow& b = W[band[q]];
            if (a.lo == b.lo && a.hi == b.hi && a.seg_row == b.seg_row)
                wid[q] = o;
        }
    }
    const StepConsts k = step_consts();

    for (;;) {
        // Midpoint shots of the last round (lane q): their rate fit is kept
        // out of the walks' update phase, where registers are scarce.
        if (lane < K) {
            const int32_t q = lane;
            if (ctl.held[q]) {
                ctl.held[q] = 0;
                ctl.remodel[q] = ctl.r[q].update(ctl.em[q], ctl.held_w[q], ctl.held_k[q], 0, ctl.pts[q]) ? 1 : 0;
            }
        }
        __syncthreads();
        // The model transform of the interpolation points of levels that
        // have just been given a rate (at the start, or by a midpoint shot),
        // one lane per point.
        if (lane < K * kPts) {
            const int32_t q = lane / kPts, i = lane % kPts;
            if (ctl.remodel[q] && i < ctl.pts[q].n)
                point_model(ctl.pts[q], i, ctl.r[q].rate, ctl.r[q].e0);
        }
        __syncthreads();
        // The probes: each level's plan (lane q), the polynomial tables of
        // the root plans (a row of lanes per level), then the probes (lane q).
        if (lane < K) {
            const int32_t q = lane;
            ctl.root[q] = 0;
            ctl.remodel[q] = 0;
            if (ctl.live[q]) {
                Refiner r = ctl.r[q];
                bool done = false;
                if (r.converged(rel_tol, P.abs_tol)) {
                    done = true;
                } else if (r.it >= max_iter) {
                    ctl.st[q] = 2;
                    done = true;
                } else {
                    const ProbePlan pl = r.plan(rel_tol, P.abs_tol, ctl.pts[q]);
                    ctl.plan[q] = pl;
                    ctl.root[q] = pl.root ? 1 : 0;
                    ctl.counted[q] = r.isolated ? 0 : 1;
                    // The fields a plan sets.
                    ctl.r[q].isolated = r.isolated;
                    ctl.r[q].reshoot = r.reshoot;
                    ctl.r[q].mid = r.mid;
                    ctl.r[q].side = r.side;
                }
                if (done) {
                    ctl.live[q] = 0;
                    write_level(levels, status, iterations, P.level_offset + task.first + q, r.level(), ctl.st[q],
                                r.it);
                }
            }
        }
        __syncthreads();
        tables_in_lanes(ctl, lane);
        __syncthreads();
        if (lane < K) {
            const int32_t q = lane;
            if (ctl.live[q]) {
                const Refiner r = ctl.r[q];
                const PolyTable t = ctl.tab[q];
                ctl.em[q] = r.finish(ctl.plan[q], &t);
            }
        }
        __syncthreads();
        // Levels still to shoot this round (bit q), and those whose shot is counted.
        uint32_t pending = 0, cmask = 0;
#pragma unroll
        for (int q = 0; q < K; ++q) {
            pending |= (ctl.live[q] != 0 ? 1u : 0u) << q;
            cmask |= (ctl.counted[q] != 0 ? 1u : 0u) << q;
        }
        pending = __builtin_amdgcn_readfirstlane(pending);
        cmask = __builtin_amdgcn_readfirstlane(cmask) & pending;
        if (pending == 0)
            break;
        while (pending != 0) {
            const int32_t q0 = __builtin_ctz(pending);
            // Level q0's window and bounds (selected, not indexed, so that
            // the arrays stay in registers).
            int32_t band0 = band[0], wid0 = wid[0];
            int2 qb0 = qb[0];
#pragma unroll
            for (int q = 1; q < K; ++q)
                if (q == q0) {
                    band0 = band[q];
                    wid0 = wid[q];
                    qb0 = qb[q];
                }
            // The members (bit q): q0 alone when its shot is counted, else
            // the uncounted pending levels of its window.
            uint32_t mm = 1u << q0;
            if (!((cmask >> q0) & 1u)) {
#pragma unroll
                for (int q = 0; q < K; ++q)
                    if (q > q0 && (((pending & ~cmask) >> q) & 1u) && wid[q] == wid0)
                        mm |= 1u << q;
            }
            pending &= ~mm;
            const int32_t n = __builtin_popcount(mm);
            const int32_t kk = n >= 3 ? 4 : n;
            // Member i's level (the first member again past the last).
            int32_t mem[4];
            {
                uint32_t rest = mm;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    mem[i] = rest != 0 ? __builtin_ctz(rest) : q0;
                    rest &= rest - 1;
                }
            }
            SegGeom<TABLE> g;
            g.init(uv_pot, segtab, W[band0], qb0, lane, T);
            int32_t c = 0;
            double fw[4];
            int32_t fk[4];
            if (kk == 4) {
                if constexpr (K == 4) {
                    const double e[4] = {P.beta * ctl.em[mem[0]], P.beta * ctl.em[mem[1]], P.beta * ctl.em[mem[2]],
                                         P.beta * ctl.em[mem[3]]};
                    double f4[4];
                    int32_t k4[4];
                    segmented_shot<T, 4, TABLE>(sh, g, lane, e, false, P.stride, k, f4, k4, c);
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        fw[i] = f4[i];
                        fk[i] = k4[i];
                    }
                }
            } else if (kk == 2) {
                if constexpr (K >= 2) {
                    const double e[2] = {P.beta * ctl.em[mem[0]], P.beta * ctl.em[mem[1]]};
                    double f2[2];
                    int32_t k2[2];
                    segmented_shot<T, 2, TABLE>(sh, g, lane, e, false, P.stride, k, f2, k2, c);
                    fw[0] = f2[0];
                    fw[1] = f2[1];
                    fk[0] = k2[0];
                    fk[1] = k2[1];
                }
            } else {
                const double e[1] = {P.beta * ctl.em[q0]};
                double f1[1];
                int32_t k1[1];
                segmented_shot<T, 1, TABLE>(sh, g, lane, e, ((cmask >> q0) & 1u) != 0, P.stride, k, f1, k1, c);
                fw[0] = f1[0];
                fk[0] = k1[0];
            }
            // Lane i takes member i's result.
            if (lane < n) {
                double wi = fw[0];
                int32_t ki = fk[0], qi = mem[0];
#pragma unroll
                for (int i = 1; i < 4; ++i)
                    if (lane == i) {
                        wi = fw[i];
                        ki = fk[i];
                        qi = mem[i];
                    }
                if (ctl.r[qi].mid) {
                    ctl.held_w[qi] = wi;
                    ctl.held_k[qi] = ki;
                    ctl.held[qi] = 1;
                } else {
                    Refiner r = ctl.r[qi];
                    RefinePoints pts = ctl.pts[qi];
                    r.update<false>(ctl.em[qi], wi, ki, c, pts);
                    ctl.r[qi] = r;
                    ctl.pts[qi] = pts;
                }
            }
            __syncthreads();
        }
    }
}

// Bracket pass: one thread per level task (potential, level index) finds
// the level's bracket in the scan arrays (binary search over its
// potential's node counts, model rate from three scan points), so the
// refinement's wavefronts start shooting at once.
__global__ __launch_bounds__(256) void bracket_kernel(const PotentialRecord* __restrict__ pots,
                                                      const BandWindow* __restrict__ windows,
                                                      const WaveTask* __restrict__ tasks, const int32_t n_tasks,
                                                      const int32_t* __restrict__ nodes,
                                                      const double* __restrict__ far_end,
                                                      const int32_t* __restrict__ rescales,
                                                      Bracket* __restrict__ brackets,
                                                      RefinePoints* __restrict__ points)
{
    const int32_t t = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= n_tasks)
        return;
    const WaveTask task = tasks[t];
    const PotentialRecord& P = pots[task.potential];
    if (P.flags != 0 || task.first >= P.level_count)
        return;
    const int64_t out = P.level_offset + task.first;
    brackets[out] = find_bracket(P, windows + task.potential * kBands, P.level_min + task.first, nodes, far_end,
                                 rescales, points[out]);
}

// ---------------------------------------------------------------- launchers

hipError_t launch_potential_stats(hipStream_t stream, const double* v, const int64_t* offsets,
                                  const int64_t* counts, const double* betas, int32_t n_pots,
                                  PotentialStats* out)
{
    if (n_pots == 0)
        return hipSuccess;
    hipLaunchKernelGGL(potential_stats_kernel, dim3(static_cast<unsigned>(n_pots)), dim3(256), 0, stream, v,
                       offsets, counts, betas, out);
    return hipGetLastError();
}

hipError_t launch_shift_potential(hipStream_t stream, const double* v, const PotentialRecord* pots,
                                  int32_t n_pots, int32_t max_points, double* ut, double* uv,
                                  double* uvr)
{
    if (n_pots == 0)
        return hipSuccess;
    const int32_t bx = (max_points + 255) / 256;
    dim3 grid(static_cast<unsigned>(bx < 1024 ? bx : 1024), static_cast<unsigned>(n_pots));
    hipLaunchKernelGGL(shift_potential_kernel, grid, dim3(256), 0, stream, v, pots, ut, uv, uvr);
    return hipGetLastError();
}

// bins == nullptr: tasks[0, n_tasks) in order.  Otherwise `bins` (device, 9
// offsets) splits the list per XCD and `max_bin` is the largest bin's size.
hipError_t launch_scan(hipStream_t stream, const double* ut, const PotentialRecord* pots,
                       const BandWindow* windows, const WaveTask* tasks, int32_t n_tasks,
                       const int32_t* bins, int32_t max_bin, int32_t waves_per_block, int32_t* nodes,
                       double* far_end, int32_t* rescales)
{
    if (n_tasks == 0)
        return hipSuccess;
    const int32_t blocks = bins == nullptr ? (n_tasks + waves_per_block - 1) / waves_per_block
                                           : 8 * ((max_bin + waves_per_block - 1) / waves_per_block);
    hipLaunchKernelGGL(numerov_scan_kernel, dim3(static_cast<unsigned>(blocks)),
                       dim3(static_cast<unsigned>(64 * waves_per_block)), 0, stream, ut, pots, windows,
                       tasks, n_tasks, bins, nodes, far_end, rescales);
    return hipGetLastError();
}

hipError_t launch_brackets(hipStream_t stream, const PotentialRecord* pots, const BandWindow* windows,
                           const WaveTask* level_tasks, int32_t n_tasks, const int32_t* nodes,
                           const double* far_end, const int32_t* rescales, Bracket* brackets,
                           RefinePoints* points)
{
    if (n_tasks == 0)
        return hipSuccess;
    hipLaunchKernelGGL(bracket_kernel, dim3(static_cast<unsigned>((n_tasks + 255) / 256)), dim3(256), 0, stream,
                       pots, windows, level_tasks, n_tasks, nodes, far_end, rescales, brackets, points);
    return hipGetLastError();
}

hipError_t launch_refine(hipStream_t stream, int32_t levels_per_lane, const double* uv,
                         const PotentialRecord* pots, const BandWindow* windows, const WaveTask* tasks,
                         int32_t n_tasks, int32_t waves_per_block, const Bracket* brackets, RefinePoints* points,
                         double rel_tol, int32_t max_iter, double* levels, int32_t* status, int32_t* iterations,
                         int32_t* steps)
{
    if (n_tasks == 0)
        return hipSuccess;
    const int32_t blocks = (n_tasks + waves_per_block - 1) / waves_per_block;
    const dim3 grid(static_cast<unsigned>(blocks)), block(static_cast<unsigned>(64 * waves_per_block));
    if (levels_per_lane == 2)
        hipLaunchKernelGGL(numerov_refine_kernel<2>, grid, block, 0, stream, uv, pots, windows, tasks,
                           n_tasks, brackets, points, rel_tol, max_iter, levels, status, iterations, steps);
    else if (levels_per_lane == 1)
        hipLaunchKernelGGL(numerov_refine_kernel<1>, grid, block, 0, stream, uv, pots, windows, tasks,
                           n_tasks, brackets, points, rel_tol, max_iter, levels, status, iterations, steps);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// segtab != nullptr selects the segment tables for 64-lane blocks.
// levels_per_block 2 or 4 (64-lane blocks only) takes tasks of that many
// consecutive levels each.
// bins == nullptr: tasks[0, n_tasks) in order; else `bins` (device, 9
// offsets) splits the list per XCD and `max_bin` is the largest bin's size.
hipError_t launch_refine_segmented(hipStream_t stream, int32_t lanes, int32_t levels_per_block, const double* uv,
                                   const double* segtab, const PotentialRecord* pots,
                                   const BandWindow* windows, const WaveTask* tasks, int32_t n_tasks,
                                   const int32_t* bins, int32_t max_bin, const Bracket* brackets,
                                   RefinePoints* points, double rel_tol, int32_t max_iter, double* levels,
                                   int32_t* status, int32_t* iterations, int32_t* steps)
{
    if (n_tasks == 0)
        return hipSuccess;
    if (levels_per_block != 1 && !((levels_per_block == 2 || levels_per_block == 4) && lanes == 64))
        return hipErrorInvalidValue;
    const dim3 grid(static_cast<unsigned>(bins == nullptr ? n_tasks : 8 * max_bin));
#define EPSEON_SEGMENTED(T, TAB, K)                                                                   \
    hipLaunchKernelGGL((numerov_refine_segmented_kernel<T, TAB, K>), grid, dim3(T), 0, stream, uv,   \
                       segtab, pots, windows, tasks, n_tasks, bins, brackets, points, rel_tol, max_iter, \
                       levels, status, iterations, steps)
    switch (lanes) {
    case 256:
        EPSEON_SEGMENTED(256, false, 1);
        break;
    case 128:
        EPSEON_SEGMENTED(128, false, 1);
        break;
    case 64:
        if (levels_per_block == 4) {
            if (segtab != nullptr)
                EPSEON_SEGMENTED(64, true, 4);
            else
                EPSEON_SEGMENTED(64, false, 4);
        } else if (levels_per_block == 2) {
            if (segtab != nullptr)
                EPSEON_SEGMENTED(64, true, 2);
            else
                EPSEON_SEGMENTED(64, false, 2);
        } else if (segtab != nullptr) {
            EPSEON_SEGMENTED(64, true, 1);
        } else {
            EPSEON_SEGMENTED(64, false, 1);
        }
        break;
    default:
        return hipErrorInvalidValue;
    }
#undef EPSEON_SEGMENTED
    return hipGetLastError();
}

hipError_t launch_segment_tables(hipStream_t stream, const double* uv, const PotentialRecord* pots,
                                 const TableJob* jobs, int32_t n_jobs, int32_t max_rows, double* segtab)
{
    if (n_jobs == 0 || max_rows == 0)
        return hipSuccess;
    const int32_t bx = (max_rows + kTableTile - 1) / kTableTile;
    dim3 grid(static_cast<unsigned>(bx < 1024 ? bx : 1024), static_cast<unsigned>(n_jobs < 65535 ? n_jobs : 65535));
    hipLaunchKernelGGL(segment_table_kernel, grid, dim3(256), 0, stream, uv, pots, jobs, n_jobs, segtab);
    return hipGetLastError();
}

// Refinement summary for the statistics, one wavefront per kSummaryPots
// potentials adding its sums to out (zeroed first): out[0] = sum of
// iterations x steps over levels that shot, out[1] = converged levels,
// out[2] = sum over converged levels of their grid's n - 2 steps.  Integer
// sums, so the result does not depend on the order.
constexpr int32_t kSummaryPots = 16;
__global__ __launch_bounds__(64) void level_summary_kernel(const PotentialRecord* __restrict__ pots,
                                                           const int32_t n_pots, const int32_t* __restrict__ status,
                                                           const int32_t* __restrict__ iterations,
                                                           const int32_t* __restrict__ steps,
                                                           unsigned long long* __restrict__ out)
{
    const int32_t p0 = static_cast<int32_t>(blockIdx.x) * kSummaryPots;
    const int32_t p1 = min(n_pots, p0 + kSummaryPots);
    int64_t sum[3] = {0, 0, 0};
    for (int32_t p = p0; p < p1; ++p) {
        const PotentialRecord& r = pots[p];
        for (int32_t l = static_cast<int32_t>(threadIdx.x); l < r.level_count; l += 64) {
            const int64_t o = r.level_offset + l;
            const int32_t it = iterations[o];
            if (it > 0)
                sum[0] += static_cast<int64_t>(it) * steps[o];
            if (status[o] == 0) {
                sum[1] += 1;
                sum[2] += r.n_points - 2;
            }
        }
    }
#pragma unroll
    for (int32_t q = 0; q < 3; ++q) {
        int64_t x = sum[q];
#pragma unroll
        for (int32_t d = 32; d >= 1; d /= 2)
            x += __shfl_xor(x, d, 64);
        if (threadIdx.x == 0 && x != 0)
            atomicAdd(out + q, static_cast<unsigned long long>(x));
    }
}

__global__ void zero_summary_kernel(unsigned long long* __restrict__ out)
{
    if (threadIdx.x < 3)
        out[threadIdx.x] = 0;
}

hipError_t launch_level_summary(hipStream_t stream, const PotentialRecord* pots, int32_t n_pots,
                                const int32_t* status, const int32_t* iterations, const int32_t* steps,
                                int64_t* out)
{
    unsigned long long* o = reinterpret_cast<unsigned long long*>(out);
    hipLaunchKernelGGL(zero_summary_kernel, dim3(1), dim3(64), 0, stream, o);
    if (n_pots > 0)
        hipLaunchKernelGGL(level_summary_kernel, dim3(static_cast<unsigned>((n_pots + kSummaryPots - 1) / kSummaryPots)),
                           dim3(64), 0, stream, pots, n_pots, status, iterations, steps, o);
    return hipGetLastError();
}

hipError_t launch_band_windows(hipStream_t stream, const double* uv, const PotentialRecord* pots,
                               int32_t n_pots, int2* out)
{
    if (n_pots == 0)
        return hipSuccess;
    hipLaunchKernelGGL(band_window_kernel, dim3(static_cast<unsigned>(n_pots)), dim3(64 * kBandWaves), 0, stream,
                       uv, pots, n_pots, out);
    return hipGetLastError();
}

}  // namespace epseon::hip
