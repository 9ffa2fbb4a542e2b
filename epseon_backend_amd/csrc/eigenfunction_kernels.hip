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

    // Returns the applied binary shift (0 or -240).
    __device__ __forceinline__ int32_t rescale()
    {
        const bool big = fmax(fabs(y), fabs(ym)) >= kPsiRescaleAt;
        const int32_t sh = big ? -kPsiRescaleLog2 : 0;
        y = __builtin_ldexp(y, sh);
        ym = __builtin_ldexp(ym, sh);
        d = __builtin_ldexp(d, sh);
        k += big ? 1 : 0;
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

// Pass 1: sum of y_i^2 over the visited points (in the shot's current scale).
struct NormBody {
    YShot s;
    double e, c10, sum;
    __device__ __forceinline__ void step(const double u, int)
    {
        sum = __builtin_fma(s.y, s.y, sum);
        s.step(u, e, c10);
    }
    __device__ __forceinline__ void block_end(int32_t, const bool full)
    {
        if (full) {
            const int32_t sh = s.rescale();
            sum = __builtin_ldexp(sum, 2 * sh);
        }
    }
};

// Pass 2, eigenfunction output: psi_i = y_i * mult written to row[i0 + dir*j].
struct PsiBody {
    YShot s;
    double e, c10, a, mult;
    int32_t kf, i0, dir, base;
    double* row;
    __device__ __forceinline__ void step(const double u, const int q)
    {
        row[i0 + dir * (base + q)] = s.y * mult;
        s.step(u, e, c10);
    }
    __device__ __forceinline__ void block_end(const int32_t count, const bool full)
    {
        base += count;
        if (full) {
            (void)s.rescale();
            mult = __builtin_ldexp(a, kPsiRescaleLog2 * (s.k - kf));
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
            G = __builtin_ldexp(g0, 2 * kPsiRescaleLog2 * (s.k - kf));
        }
    }
};

__device__ __forceinline__ bool level_active(const int32_t* __restrict__ status, const int64_t o)
{
    return status[o] == 0;
}

}  // namespace

// Pass 1: one lane per level (wavefront = up to 64 levels of one potential).
// Shoots inward from both walls at the refined E, matches at m and writes the
// level's LevelNorm.
__global__ __launch_bounds__(256) void normalize_kernel(
    const double* __restrict__ uv, const double* __restrict__ uvr,
    const PotentialRecord* __restrict__ pots, const WaveTask* __restrict__ tasks,
    const int32_t n_tasks, const double* __restrict__ levels, const int32_t* __restrict__ status,
    LevelNorm* __restrict__ norms)
{
    const int32_t waves = static_cast<int32_t>(blockDim.x >> 6);
    const int32_t tid = static_cast<int32_t>(blockIdx.x) * waves +
                        __builtin_amdgcn_readfirstlane(static_cast<int32_t>(threadIdx.x >> 6));
    if (tid >= n_tasks)
        return;
    const WaveTask task = tasks[tid];
    const PotentialRecord& P = pots[task.potential];
    const int32_t l = task.first + static_cast<int32_t>(threadIdx.x & 63);
    if (l >= P.level_count)
        return;
    const int64_t o = P.level_offset + l;
    if (P.flags != 0 || !level_active(status, o)) {
        norms[o] = LevelNorm{0.0, 0.0, 0, 0};
        return;
    }
    const double e = P.beta * levels[o];
    const double c10 = opaque(10.0);
    const int32_t n = P.n_points;
    const int32_t m = P.match;
    const double* u = uv + P.ut_offset;
    const double* ur = uvr + P.ut_offset;

    NormBody L{};
    L.e = e;
    L.c10 = c10;
    L.sum = 0.0;
    L.s.init(e, u[0], u[1]);
    psi_walk(u + 2, m, L);

    NormBody R{};
    R.e = e;
    R.c10 = c10;
    R.sum = 0.0;
    R.s.init(e, ur[0], ur[1]);
    psi_walk(ur + 2, n - 2 - m, R);

    // Left: (y_m, y_{m+1}) = (ym, y); right (reversed): (y_{m+1}, y_m) = (ym, y).
    const double yl0 = L.s.ym, yl1 = L.s.y;
    const double yr1 = R.s.ym, yr0 = R.s.y;
    const double num = __builtin_fma(yl0, yr0, yl1 * yr1);
    const double den = __builtin_fma(yr0, yr0, yr1 * yr1);
    const double c = num / den;
    const double s = __builtin_fma(c * c, R.sum, L.sum);
    const double a = 1.0 / sqrt(s * P.h);
    norms[o] = LevelNorm{a, c * a, L.s.k, R.s.k};
}

// Eigenfunctions: for each wavefront task (potential, first level), row l of
// that potential (full_points values at out[psi_offset - base + l * full_points])
// is psi of level l, zero outside the open trimmed window and for levels that
// did not converge.  One lane per level.
__global__ __launch_bounds__(64) void eigenfunction_kernel(
    const double* __restrict__ uv, const double* __restrict__ uvr,
    const PotentialRecord* __restrict__ pots, const WaveTask* __restrict__ tasks,
    const double* __restrict__ levels, const int32_t* __restrict__ status,
    const LevelNorm* __restrict__ norms, const int64_t base, double* __restrict__ out)
{
    const WaveTask task = tasks[blockIdx.x];
    const PotentialRecord& P = pots[task.potential];
    const int32_t l = task.first + static_cast<int32_t>(threadIdx.x);
    if (l >= P.level_count)
        return;
    const int64_t N = P.full_points;
    double* row = out + (P.psi_offset - base) + static_cast<int64_t>(l) * N;
    const int64_t first = P.first_point;
    const int32_t n = P.n_points;
    const bool ok = P.flags == 0;
    const int64_t lo_end = ok ? first + 1 : N;
    for (int64_t i = 0; i < lo_end; ++i)
        row[i] = 0.0;
    if (!ok)
        return;
    for (int64_t i = first + n - 1; i < N; ++i)
        row[i] = 0.0;
    const int64_t o = P.level_offset + l;
    const bool active = level_active(status, o);
    const LevelNorm nm = active ? norms[o] : LevelNorm{0.0, 0.0, 0, 0};
    const double e = P.beta * (active ? levels[o] : P.e_lo);
    const double c10 = opaque(10.0);
    const int32_t m = P.match;
    const double* u = uv + P.ut_offset;
    const double* ur = uvr + P.ut_offset;
    double* win = row + first;

    PsiBody L{};
    L.e = e;
    L.c10 = c10;
    L.a = nm.a_left;
    L.kf = nm.k_left;
    L.mult = __builtin_ldexp(nm.a_left, -kPsiRescaleLog2 * nm.k_left);
    L.i0 = 1;
    L.dir = 1;
    L.base = 0;
    L.row = win;
    L.s.init(e, u[0], u[1]);
    psi_walk(u + 2, m, L);

    PsiBody R{};
    R.e = e;
    R.c10 = c10;
    R.a = nm.a_right;
    R.kf = nm.k_right;
    R.mult = __builtin_ldexp(nm.a_right, -kPsiRescaleLog2 * nm.k_right);
    R.i0 = n - 2;
    R.dir = -1;
    R.base = 0;
    R.row = win;
    R.s.init(e, ur[0], ur[1]);
    psi_walk(ur + 2, n - 2 - m, R);
}

// Adjoint: grad_v[user_offset + i] = sum_l grad_levels[l] psi_l(i)^2 h over
// the converged levels of the block's potential, zero elsewhere on the
// potential's full grid.  One block of W wavefronts per potential; levels in
// groups of 64 W, the first group storing and later ones adding, so the sum
// order is fixed.
template <int W>
__global__ __launch_bounds__(64 * W) void level_vjp_kernel(
    const double* __restrict__ uv, const double* __restrict__ uvr,
    const PotentialRecord* __restrict__ pots, const int32_t* __restrict__ pot_list,
    const double* __restrict__ levels, const int32_t* __restrict__ status,
    const LevelNorm* __restrict__ norms, const double* __restrict__ grad_levels,
    double* __restrict__ grad_v)
{
    __shared__ double s_c[W][32][65];
    __shared__ double s_w[W][32];
    const PotentialRecord& P = pots[pot_list[blockIdx.x]];
    double* out = grad_v + P.user_offset;
    const int32_t N = P.full_points;
    const int32_t n = P.n_points;
    const int32_t first = static_cast<int32_t>(P.first_point);
    const bool ok = P.flags == 0;
    for (int32_t i = static_cast<int32_t>(threadIdx.x); i < N; i += 64 * W)
        if (!ok || i <= first || i >= first + n - 1)
            out[i] = 0.0;
    if (!ok)
        return;
    const double c10 = opaque(10.0);
    const int32_t m = P.match;
    const double* u = uv + P.ut_offset;
    const double* ur = uvr + P.ut_offset;
    double* win = out + first;
    for (int32_t g = 0; g < P.level_count; g += 64 * W) {
        const int32_t l = g + static_cast<int32_t>(threadIdx.x);
        const int64_t o = P.level_offset + l;
        const bool active = l < P.level_count && level_active(status, o);
        const LevelNorm nm = active ? norms[o] : LevelNorm{0.0, 0.0, 0, 0};
        const double gl = active ? grad_levels[o] : 0.0;
        const double e = P.beta * (active ? levels[o] : P.e_lo);

        VjpBody<W> L{};
        L.e = e;
        L.c10 = c10;
        L.g0 = gl * P.h * nm.a_left * nm.a_left;
        L.kf = nm.k_left;
        L.G = __builtin_ldexp(L.g0, -2 * kPsiRescaleLog2 * nm.k_left);
        L.i0 = 1;
        L.dir = 1;
        L.base = 0;
        L.accumulate = g > 0;
        L.sc = s_c;
        L.sw = s_w;
        L.out = win;
        L.s.init(e, u[0], u[1]);
        psi_walk(u + 2, m, L);

        VjpBody<W> R{};
        R.e = e;
        R.c10 = c10;
        R.g0 = gl * P.h * nm.a_right * nm.a_right;
        R.kf = nm.k_right;
        R.G = __builtin_ldexp(R.g0, -2 * kPsiRescaleLog2 * nm.k_right);
        R.i0 = n - 2;
        R.dir = -1;
        R.base = 0;
        R.accumulate = g > 0;
        R.sc = s_c;
        R.sw = s_w;
        R.out = win;
        R.s.init(e, ur[0], ur[1]);
        psi_walk(ur + 2, n - 2 - m, R);
    }
}

// ---------------------------------------------------------------- launchers

hipError_t launch_normalize(hipStream_t stream, const double* uv, const double* uvr,
                            const PotentialRecord* pots, const WaveTask* tasks, int32_t n_tasks,
                            const double* levels, const int32_t* status, LevelNorm* norms)
{
    if (n_tasks == 0)
        return hipSuccess;
    hipLaunchKernelGGL(normalize_kernel, dim3(static_cast<unsigned>(n_tasks)), dim3(64), 0, stream,
                       uv, uvr, pots, tasks, n_tasks, levels, status, norms);
    return hipGetLastError();
}

hipError_t launch_eigenfunctions(hipStream_t stream, const double* uv, const double* uvr,
                                 const PotentialRecord* pots, const WaveTask* tasks, int32_t n_tasks,
                                 const double* levels, const int32_t* status,
                                 const LevelNorm* norms, int64_t base, double* out)
{
    if (n_tasks == 0)
        return hipSuccess;
    hipLaunchKernelGGL(eigenfunction_kernel, dim3(static_cast<unsigned>(n_tasks)), dim3(64), 0, stream,
                       uv, uvr, pots, tasks, levels, status, norms, base, out);
    return hipGetLastError();
}

hipError_t launch_level_vjp(hipStream_t stream, int32_t waves, const double* uv, const double* uvr,
                            const PotentialRecord* pots, const int32_t* pot_list, int32_t n_list,
                            const double* levels, const int32_t* status, const LevelNorm* norms,
                            const double* grad_levels, double* grad_v)
{
    if (n_list == 0)
        return hipSuccess;
    const dim3 grid(static_cast<unsigned>(n_list));
    switch (waves) {
    case 1:
        hipLaunchKernelGGL(level_vjp_kernel<1>, grid, dim3(64), 0, stream, uv, uvr, pots, pot_list,
                           levels, status, norms, grad_levels, grad_v);
        break;
    case 2:
        hipLaunchKernelGGL(level_vjp_kernel<2>, grid, dim3(128), 0, stream, uv, uvr, pots, pot_list,
                           levels, status, norms, grad_levels, grad_v);
        break;
    case 3:
        hipLaunchKernelGGL(level_vjp_kernel<3>, grid, dim3(192), 0, stream, uv, uvr, pots, pot_list,
                           levels, status, norms, grad_levels, grad_v);
        break;
    case 4:
        hipLaunchKernelGGL(level_vjp_kernel<4>, grid, dim3(256), 0, stream, uv, uvr, pots, pot_list,
                           levels, status, norms, grad_levels, grad_v);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace epseon::hip
