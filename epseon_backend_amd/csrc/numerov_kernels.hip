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
// Recurrence (division-free homogeneous Numerov, 5 FP64 VALU ops/step):
//   f_i = e - u_i, t = fma(-10, f_i, 12), g = f_i f_{i-1},
//   w_{i+1} = fma(t, w_i, -(g w_{i-1})),  w_0 = 0, w_1 = 1,
// with u_i = beta V_i - 1 and e = beta E, so f_i = 1 + beta (E - V_i).
// The sign of w is sampled once per block of `stride` steps (a power of two
// the host chose so that no two nodes fall in one block); the overflow test
// (|w| >= 2^480 -> scale the last two values by 2^-480) runs once per
// 32-step super-block.  Compiled with -ffp-contract=off; every fused operation is an
// explicit fma, so results are bit-identical to the C oracle.
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

struct ShotState {
    double wm, w, fm;
    int32_t nodes, rescales;
    uint32_t last_sign;
};

// Loop constants of t = fma(-10, f, 12).  Neither is an inline constant of
// the FP64 VALU, so the compiler would otherwise rebuild 12.0 in the
// accumulator of a v_fmac every step (two extra v_mov per step); held opaque
// in VGPRs the step is a single three-VGPR v_fma_f64.
struct StepConsts {
    double m10, c12;
};

__device__ __forceinline__ StepConsts step_consts()
{
    StepConsts k{-10.0, 12.0};
    asm volatile("" : "+v"(k.m10), "+v"(k.c12));
    return k;
}

__device__ __forceinline__ void numerov_step(const double u, const double e, const StepConsts& k,
                                             ShotState& s)
{
    const double f = e - u;
    const double t = __builtin_fma(k.m10, f, k.c12);
    const double g = f * s.fm;
    const double wn = __builtin_fma(t, s.w, -(g * s.wm));
    s.wm = s.w;
    s.w = wn;
    s.fm = f;
}

__device__ __forceinline__ void sample_sign(ShotState& s)
{
    const uint32_t sb = sign_bit(s.w);
    s.nodes += static_cast<int32_t>(sb ^ s.last_sign);
    s.last_sign = sb;
}

// Branch-free overflow guard: a branch here would split the super-block and
// let the compiler sink half of the next prefetch behind it.
__device__ __forceinline__ void rescale(ShotState& s)
{
    const bool big = fmax(fabs(s.w), fabs(s.wm)) >= kRescaleAt;
    const int32_t sh = big ? -kRescaleLog2 : 0;
    s.w = __builtin_ldexp(s.w, sh);
    s.wm = __builtin_ldexp(s.wm, sh);
    s.rescales += big ? 1 : 0;
}

// Scheduling fence: nothing moves across it, so the next buffer's scalar
// loads issue right after the first use of the current buffer.  The wait the
// compiler inserts before that first use then finds only the load issued one
// half super-block earlier still outstanding (scalar loads return out of
// order, so every wait is lgkmcnt(0)).
__device__ __forceinline__ void fence_schedule() { __builtin_amdgcn_sched_barrier(0); }

template <int B, int Q>
__device__ __forceinline__ void step_at(const double u, const double e, const StepConsts& k,
                                        ShotState& s)
{
    numerov_step(u, e, k, s);
    if constexpr ((Q + 1) % B == 0)
        sample_sign(s);
}

template <int B, int Q0, int... Is>
__device__ __forceinline__ void steps_1_15(const double (&buf)[16], const double e,
                                           const StepConsts& k, ShotState& s,
                                           std::integer_sequence<int, Is...>)
{
    (step_at<B, Q0 + Is + 1>(buf[Is + 1], e, k, s), ...);
}

// One outward shot over n points at probe e = beta E with node stride B.
// The grid is consumed in super-blocks of 32 steps fed from two 16-point
// SGPR buffers: each buffer's scalar loads issue right after the first step
// of the other buffer, so a load has 15 steps to land.  The rescale test runs
// once per super-block, after its 17th step; the sign is sampled after every
// B-th step.  The table is padded so the prefetch past the last super-block
// stays inside the allocation.
template <int B>
__device__ __forceinline__ int32_t shoot_fixed(const double* __restrict__ u, const int32_t n,
                                               const double e, double& w_last, int32_t& rescales)
{
    static_assert(32 % B == 0, "stride must divide the super-block");
    const StepConsts k = step_consts();
    ShotState s{0.0, 1.0, e - u[0], 0, 0, 0u};
    const int32_t steps = n - 2;
    const int32_t nsb = steps >> 5;
    const double* p = u + 1;
    double a[16], b[16];
#pragma unroll
    for (int q = 0; q < 16; ++q)
        a[q] = p[q];
    for (int32_t sb = 0; sb < nsb; ++sb) {
        step_at<B, 0>(a[0], e, k, s);
        fence_schedule();
#pragma unroll
        for (int q = 0; q < 16; ++q)
            b[q] = p[16 + q];
        fence_schedule();
        steps_1_15<B, 0>(a, e, k, s, std::make_integer_sequence<int, 15>{});
        step_at<B, 16>(b[0], e, k, s);
        rescale(s);
        fence_schedule();
#pragma unroll
        for (int q = 0; q < 16; ++q)
            a[q] = p[32 + q];
        fence_schedule();
        steps_1_15<B, 16>(b, e, k, s, std::make_integer_sequence<int, 15>{});
        p += 32;
    }
    const int32_t rem = steps - (nsb << 5);
    for (int32_t q = 0; q < rem; ++q) {
        numerov_step(p[q], e, k, s);
        if ((q + 1) % B == 0)
            sample_sign(s);
    }
    if (rem % B != 0)
        sample_sign(s);
    w_last = s.w;
    rescales = s.rescales;
    return s.nodes;
}

__device__ __forceinline__ int32_t shoot(const double* __restrict__ u, const int32_t n,
                                         const double e, const int32_t stride, double& w_last,
                                         int32_t& rescales)
{
    switch (stride) {
    case 32: return shoot_fixed<32>(u, n, e, w_last, rescales);
    case 16: return shoot_fixed<16>(u, n, e, w_last, rescales);
    case 8: return shoot_fixed<8>(u, n, e, w_last, rescales);
    case 4: return shoot_fixed<4>(u, n, e, w_last, rescales);
    case 2: return shoot_fixed<2>(u, n, e, w_last, rescales);
    default: return shoot_fixed<1>(u, n, e, w_last, rescales);
    }
}

__device__ __forceinline__ int32_t wave_task_index()
{
    const int32_t waves_per_block = static_cast<int32_t>(blockDim.x >> 6);
    const int32_t wave = __builtin_amdgcn_readfirstlane(static_cast<int32_t>(threadIdx.x >> 6));
    return static_cast<int32_t>(blockIdx.x) * waves_per_block + wave;
}

}  // namespace

// u_i = fma(beta_p, V_i, -1) for every point of every potential.
__global__ __launch_bounds__(256) void shift_potential_kernel(const double* __restrict__ v,
                                                              const PotentialRecord* __restrict__ pots,
                                                              double* __restrict__ ut)
{
    const PotentialRecord& P = pots[blockIdx.y];
    const double beta = P.beta;
    for (int32_t i = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x); i < P.n_points;
         i += static_cast<int32_t>(gridDim.x * blockDim.x))
        ut[P.ut_offset + i] = __builtin_fma(beta, v[P.ut_offset + i], -1.0);
}

// Scan: lane = trial energy E_j = fma(j+1, dE, E_lo); outputs node count,
// far-end value and rescale count per trial energy.
__global__ __launch_bounds__(256) void numerov_scan_kernel(const double* __restrict__ ut,
                                                           const PotentialRecord* __restrict__ pots,
                                                           const WaveTask* __restrict__ tasks,
                                                           const int32_t n_tasks,
                                                           int32_t* __restrict__ nodes,
                                                           double* __restrict__ far_end,
                                                           int32_t* __restrict__ rescales)
{
    const int32_t tid = wave_task_index();
    if (tid >= n_tasks)
        return;
    const WaveTask task = tasks[tid];
    const PotentialRecord& P = pots[task.potential];
    const int32_t j = task.first + static_cast<int32_t>(threadIdx.x & 63);
    if (j >= P.n_energies)
        return;
    const double E = __builtin_fma(static_cast<double>(j + 1), P.de, P.e_lo);
    const double e = P.beta * E;
    double w;
    int32_t k;
    const int32_t c = shoot(ut + P.ut_offset, P.n_points, e, P.stride, w, k);
    const int64_t o = P.energy_offset + j;
    nodes[o] = c;
    far_end[o] = w;
    rescales[o] = k;
}

// Bracket + refine: lane = level v.  The bracket is the first trial energy
// with more than v nodes and the one below it (or the window floor).  While
// the bracket holds more than level v it is halved by node count; then the
// Illinois variant of regula falsi runs on F(E) = w_{N-1} 2^(480 k), each new
// point kept at least tol/2 inside the bracket.  The level is the midpoint of
// the final bracket, whose width is <= rel_tol max(|E_lo|, |E_hi|) + abs_tol.
__global__ __launch_bounds__(256) void numerov_refine_kernel(
    const double* __restrict__ ut, const PotentialRecord* __restrict__ pots,
    const WaveTask* __restrict__ tasks, const int32_t n_tasks, const int32_t* __restrict__ nodes,
    const double* __restrict__ far_end, const int32_t* __restrict__ rescales, const double rel_tol,
    const int32_t max_iter, double* __restrict__ levels, int32_t* __restrict__ status,
    int32_t* __restrict__ iterations)
{
    const int32_t tid = wave_task_index();
    if (tid >= n_tasks)
        return;
    const WaveTask task = tasks[tid];
    const PotentialRecord& P = pots[task.potential];
    const int32_t l = task.first + static_cast<int32_t>(threadIdx.x & 63);
    if (l >= P.level_count)
        return;
    const int64_t out = P.level_offset + l;
    if (P.flags != 0) {
        levels[out] = __builtin_nan("");
        status[out] = 3;  // EPSEON_LEVEL_UNRESOLVED
        iterations[out] = 0;
        return;
    }
    const int32_t v = P.level_min + l;
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
        levels[out] = __builtin_nan("");
        status[out] = 1;  // EPSEON_LEVEL_UNBOUND
        iterations[out] = 0;
        return;
    }
    double eh = __builtin_fma(static_cast<double>(j + 1), P.de, P.e_lo);
    int32_t ch = nd[j];
    double wh = far_end[P.energy_offset + j];
    int32_t kh = rescales[P.energy_offset + j];
    double el, wl = 0.0;
    int32_t cl, kl = 0, lo_ok;
    if (j == 0 || nd[j - 1] > v) {
        el = P.e_lo;
        cl = 0;
        lo_ok = 0;
    } else {
        el = __builtin_fma(static_cast<double>(j), P.de, P.e_lo);
        cl = nd[j - 1];
        wl = far_end[P.energy_offset + j - 1];
        kl = rescales[P.energy_offset + j - 1];
        lo_ok = 1;
    }

    const double* u = ut + P.ut_offset;
    const int32_t n = P.n_points;
    const int32_t stride = P.stride;
    const double beta = P.beta;
    int32_t side = 0;
    int32_t it = 0;
    int32_t st = 0;
    for (;;) {
        const double width = eh - el;
        const double tol = __builtin_fma(rel_tol, fmax(fabs(el), fabs(eh)), P.abs_tol);
        if (width <= tol)
            break;
        if (it >= max_iter) {
            st = 2;  // EPSEON_LEVEL_MAXITER
            break;
        }
        double em;
        const bool isolated = lo_ok && cl == v && ch == v + 1;
        if (isolated) {
            const double r = __builtin_ldexp(wh / wl, kRescaleLog2 * (kh - kl));
            const double s = 1.0 / (1.0 - r);
            em = __builtin_fma(s, width, el);
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
        const int32_t c = shoot(u, n, beta * em, stride, w, k);
        ++it;
        if (c <= v) {
            el = em;
            cl = c;
            wl = w;
            kl = k;
            lo_ok = 1;
            if (isolated && side == -1)
                wh *= 0.5;
            side = isolated ? -1 : 0;
        } else {
            eh = em;
            ch = c;
            wh = w;
            kh = k;
            if (isolated && side == 1)
                wl *= 0.5;
            side = isolated ? 1 : 0;
        }
    }
    levels[out] = el + 0.5 * (eh - el);
    status[out] = st;
    iterations[out] = it;
}

// ---------------------------------------------------------------- launchers

hipError_t launch_shift_potential(hipStream_t stream, const double* v, const PotentialRecord* pots,
                                  int32_t n_pots, int32_t max_points, double* ut)
{
    if (n_pots == 0)
        return hipSuccess;
    const int32_t bx = (max_points + 255) / 256;
    dim3 grid(static_cast<unsigned>(bx < 1024 ? bx : 1024), static_cast<unsigned>(n_pots));
    hipLaunchKernelGGL(shift_potential_kernel, grid, dim3(256), 0, stream, v, pots, ut);
    return hipGetLastError();
}

hipError_t launch_scan(hipStream_t stream, const double* ut, const PotentialRecord* pots,
                       const WaveTask* tasks, int32_t n_tasks, int32_t waves_per_block,
                       int32_t* nodes, double* far_end, int32_t* rescales)
{
    if (n_tasks == 0)
        return hipSuccess;
    const int32_t blocks = (n_tasks + waves_per_block - 1) / waves_per_block;
    hipLaunchKernelGGL(numerov_scan_kernel, dim3(static_cast<unsigned>(blocks)),
                       dim3(static_cast<unsigned>(64 * waves_per_block)), 0, stream, ut, pots,
                       tasks, n_tasks, nodes, far_end, rescales);
    return hipGetLastError();
}

hipError_t launch_refine(hipStream_t stream, const double* ut, const PotentialRecord* pots,
                         const WaveTask* tasks, int32_t n_tasks, int32_t waves_per_block,
                         const int32_t* nodes, const double* far_end, const int32_t* rescales,
                         double rel_tol, int32_t max_iter, double* levels, int32_t* status,
                         int32_t* iterations)
{
    if (n_tasks == 0)
        return hipSuccess;
    const int32_t blocks = (n_tasks + waves_per_block - 1) / waves_per_block;
    hipLaunchKernelGGL(numerov_refine_kernel, dim3(static_cast<unsigned>(blocks)),
                       dim3(static_cast<unsigned>(64 * waves_per_block)), 0, stream, ut, pots,
                       tasks, n_tasks, nodes, far_end, rescales, rel_tol, max_iter, levels, status,
                       iterations);
    return hipGetLastError();
}

}  // namespace epseon::hip
