// epseon_capi.hip -- host side of libepseon_hip.so: contexts, device arena,
// batch layout and the extern "C" entry points declared in include/epseon_hip.h.
//
// Stands in for what VibwaAlgorithm<FP>::run sets up per task in the
// reference (logical device, VMA allocator, per-slot buffers and descriptor
// sets: cpp/gpu/include/epseon/gpu/algorithms/vibwa.hpp:605-637) and for the
// dispatch that never happens there.  No Vulkan objects, no descriptors: the
// kernels take raw device pointers into one arena plus a table of
// PotentialRecord offsets.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "epseon_hip.h"
#include "records.hpp"

namespace epseon::hip {

static_assert(kBands == EPSEON_BANDS, "band count of the C ABI");

hipError_t launch_shift_potential(hipStream_t, const double*, const PotentialRecord*, int32_t,
                                  int32_t, double*, double*, double*);
hipError_t launch_potential_stats(hipStream_t, const double*, const int64_t*, const int64_t*,
                                  const double*, int32_t, PotentialStats*);
hipError_t launch_scan(hipStream_t, const double*, const PotentialRecord*, const BandWindow*,
                       const WaveTask*, int32_t, const int32_t*, int32_t, int32_t, int32_t*, double*,
                       int32_t*);
hipError_t launch_brackets(hipStream_t, const PotentialRecord*, const BandWindow*, const WaveTask*, int32_t,
                           const int32_t*, const double*, const int32_t*, Bracket*, RefinePoints*);
hipError_t launch_refine(hipStream_t, int32_t, const double*, const PotentialRecord*, const BandWindow*,
                         const WaveTask*, int32_t, int32_t, const Bracket*, RefinePoints*, double, int32_t,
                         double*, int32_t*, int32_t*, int32_t*);
hipError_t launch_refine_segmented(hipStream_t, int32_t, int32_t, const double*, const double*,
                                   const PotentialRecord*, const BandWindow*, const WaveTask*, int32_t,
                                   const int32_t*, int32_t, const Bracket*, RefinePoints*, double, int32_t,
                                   double*, int32_t*, int32_t*, int32_t*);
hipError_t launch_segment_tables(hipStream_t, const double*, const PotentialRecord*, const TableJob*,
                                 int32_t, int32_t, double*);
hipError_t launch_band_windows(hipStream_t, const double*, const PotentialRecord*, int32_t, int2*);
hipError_t launch_level_summary(hipStream_t, const PotentialRecord*, int32_t, const int32_t*, const int32_t*,
                                const int32_t*, int64_t*);
hipError_t launch_normalize(hipStream_t, const double*, const double*, const PotentialRecord*,
                            const PsiTask*, int32_t, const WaveTask*, int32_t, const double*, const int32_t*,
                            PsiSegMatrix*, PsiCheckpoint*, LevelNorm*);
hipError_t launch_eigenfunctions(hipStream_t, const double*, const double*, const PotentialRecord*,
                                 const PsiTask*, int32_t, const double*, const int32_t*,
                                 const LevelNorm*, const PsiCheckpoint*, int64_t, double*);
hipError_t launch_level_vjp(hipStream_t, int32_t, const double*, const double*,
                            const PotentialRecord*, const PsiTask*, int32_t, const double*,
                            const int32_t*, const LevelNorm*, const PsiCheckpoint*, const double*,
                            double*);

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what)
{
    return fail(EPSEON_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define EPSEON_HIP_TRY(expr)                                 \
    do {                                                     \
        const hipError_t epseon_e_ = (expr);                 \
        if (epseon_e_ != hipSuccess)                         \
            return hip_fail(epseon_e_, #expr);               \
    } while (0)

// Device arena: hipMalloc'd blocks of a fixed size carved by a bump pointer
// (256-byte aligned); requests larger than a block get a block of their own.
// reset() keeps the blocks for the next batch.
class Arena {
  public:
    explicit Arena(size_t block_bytes) : block_bytes_(block_bytes) {}
    Arena(const Arena&) = delete;
    Arena& operator=(const Arena&) = delete;
    ~Arena() { release(); }

    hipError_t alloc(size_t bytes, void** out)
    {
        bytes = (bytes + 255) & ~size_t(255);
        if (bytes == 0)
            bytes = 256;
        for (size_t b = cursor_block_; b < blocks_.size(); ++b) {
            Block& blk = blocks_[b];
            if (blk.size - blk.used >= bytes) {
                *out = static_cast<char*>(blk.ptr) + blk.used;
                blk.used += bytes;
                cursor_block_ = b;
                return hipSuccess;
            }
        }
        Block blk{nullptr, std::max(bytes, block_bytes_), 0};
        const hipError_t e = hipMalloc(&blk.ptr, blk.size);
        if (e != hipSuccess)
            return e;
        blk.used = bytes;
        blocks_.push_back(blk);
        cursor_block_ = blocks_.size() - 1;
        *out = blk.ptr;
        return hipSuccess;
    }

    void reset()
    {
        for (Block& b : blocks_)
            b.used = 0;
        cursor_block_ = 0;
    }

    void release()
    {
        for (Block& b : blocks_)
            (void)hipFree(b.ptr);
        blocks_.clear();
        cursor_block_ = 0;
    }

    size_t reserved_bytes() const
    {
        size_t s = 0;
        for (const Block& b : blocks_)
            s += b.size;
        return s;
    }

  private:
    struct Block {
        void* ptr;
        size_t size;
        size_t used;
    };
    size_t block_bytes_;
    std::vector<Block> blocks_;
    size_t cursor_block_ = 0;
};

template <typename T>
hipError_t arena_array(Arena& a, size_t count, T** out)
{
    void* p = nullptr;
    const hipError_t e = a.alloc(count * sizeof(T), &p);
    *out = static_cast<T*>(p);
    return e;
}

// Batches with at least this many levels fill the GPU with one lane per
// level (2048 wavefronts); smaller ones refine with one segmented wavefront
// per level.
constexpr int64_t kLanePerLevelMinLevels = int64_t(64) * 2048;

// Automatic level groups in the segmented refinement need at least this many
// levels per compute unit: pairs (three blocks per SIMD), and groups of four
// (two blocks per SIMD, with several rounds of blocks to balance the XCD bins).
constexpr int64_t kPairMinLevelsPerCu = 24;
constexpr int64_t kQuadMinLevelsPerCu = 96;

// Automatic windows need at least this many scan wavefronts per SIMD.
constexpr int64_t kWindowsMinScanWavesPerSimd = 8;

// Node-sampling stride: the largest power of two B <= 32 with B * theta_max
// <= pi/2, theta_max being the discrete phase advance per step at the
// highest kinetic energy of the window.
int32_t node_stride(double beta, double e_hi, double v_min)
{
    const double f = 1.0 + beta * (e_hi - v_min);
    const double u = 12.0 / f - 10.0;
    if (!(u > -2.0))
        return 1;
    if (u >= 2.0)
        return kMaxStride;
    const double theta = std::acos(0.5 * u);
    const double bmax = std::floor(1.5707963267948966 / theta);
    int32_t b = 1;
    while (b * 2 <= kMaxStride && static_cast<double>(b * 2) <= bmax)
        b *= 2;
    return b;
}

}  // namespace

// Scratch of epseon_level_vjp_noalloc: the normalisation pass's segment maps
// and shot checkpoints, in a file-scope device array instead of the arena.
// Every device the library runs on holds its own copy until the process
// ends.  Host code never uses the array's name (a host-side stand-in): each
// device's address is taken once with hipGetSymbolAddress, with that device
// current, when the first context opens on it, and kept in scratch_ptr.
constexpr size_t kNoAllocScratchBytes = EPSEON_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// Address of `scratch` on the current device `device` (once per device).
hipError_t scratch_setup(int32_t device)
{
    if (device < 0 || device >= kMaxScratchDevices)
        return hipSuccess;  // no scratch: epseon_level_vjp_noalloc reports it
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (g_scratch_ptr[device] != nullptr)
        return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
        return e;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

void* scratch_of(int32_t device)
{
    if (device < 0 || device >= kMaxScratchDevices)
        return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    return g_scratch_ptr[device];
}

}  // namespace epseon::hip

using namespace epseon::hip;

namespace {

// Everything the launches of a captured solve depend on: device pointers
// and task lists change only with a load; the rest is the tuning and the
// refinement arguments passed by value.
struct SolveKey {
    uint64_t load = 0;
    int32_t refine_mode = 0, lanes = 0, waves_per_block = 0, xcd_scan = 0, scan_windows = 0, max_iter = 0;
    int32_t levels_per_block = 1;
    double rel_tol = 0.0;

    bool operator==(const SolveKey& o) const
    {
        return load == o.load && refine_mode == o.refine_mode && lanes == o.lanes &&
               levels_per_block == o.levels_per_block &&
               waves_per_block == o.waves_per_block && xcd_scan == o.xcd_scan && scan_windows == o.scan_windows &&
               max_iter == o.max_iter && rel_tol == o.rel_tol;
    }
};

}  // namespace

struct epseon_ctx {
    int32_t device = 0;
    hipStream_t stream = nullptr;
    // Second stream of the pipelined solve (default priority, like the first).
    hipStream_t stream2 = nullptr;
    // Third and fourth streams: the adjoint's four block-width classes run
    // concurrently (each class is a few long, latency-bound blocks).
    hipStream_t stream3 = nullptr, stream4 = nullptr;
    hipEvent_t ev_join_aux[3] = {};
    hipEvent_t ev[8] = {};
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_solve0 = nullptr, ev_solve1 = nullptr;
    // The segment tables are built on stream2 during the first scan after a
    // load; refinements wait for ev_tables once (tables_pending).
    hipEvent_t ev_tables = nullptr;
    bool tables_pending = false;
    Arena arena;
    int32_t waves_per_block = 1;

    std::vector<epseon_potential_desc> desc;
    std::vector<PotentialRecord> rec;
    int64_t total_points = 0, total_energies = 0, total_levels = 0;
    int32_t max_points = 0;

    double* d_v = nullptr;
    double* d_ut = nullptr;
    double* d_uv = nullptr;
    double* d_uvr = nullptr;
    // Band windows (kBands per potential, in potential order): the
    // whole grid in every band (d_win_full) and the windowed ones
    // (d_win_band, computed on the device at load).  A scan and the
    // refinement that follows it use the same set.
    BandWindow* d_win_full = nullptr;
    BandWindow* d_win_band = nullptr;
    int32_t windows = 2;       // tuning: 1 windowed shots, 0 whole grid, 2 auto
    int compute_units = 0;
    bool scan_windows = true;  // the set the last scan used
    // Segment tables of the 64-lane segmented refinement, one per distinct
    // window of a potential with levels to refine (rows of 64 values): lane
    // s of a level's wavefront reads column s, so its loads are coalesced.
    double* d_segtab_full = nullptr;
    double* d_segtab_band = nullptr;
    bool segtab_ok = false;  // every table fits one 32-bit-addressed buffer
    // Jobs of both table sets (whole-grid first, then the band windows'); a
    // set is built when a refinement first needs it.
    TableJob* d_table_jobs = nullptr;
    int32_t n_jobs_full = 0, n_jobs_band = 0, table_max_rows = 0;
    bool tables_full = false, tables_band = false;
    int32_t* d_steps = nullptr;  // recurrence steps of each level's shots (its window's n - 2)
    Bracket* d_brackets = nullptr;  // each level's bracket from the last scan
    RefinePoints* d_points = nullptr;  // each level's interpolation points (bracket pass, refinement)
    PotentialRecord* d_rec = nullptr;
    WaveTask* d_scan_tasks = nullptr;
    // Scan tasks split into one bin per XCD (whole potentials, balanced by
    // scan work), concatenated; d_xcd_bins holds the 9 bin offsets.
    WaveTask* d_xcd_scan_tasks = nullptr;
    int32_t* d_xcd_bins = nullptr;
    int32_t xcd_max_bin = 0;
    // Segmented-refinement level tasks binned per XCD the same way (weight
    // n_points per level), so the levels of a potential read its tables
    // through one L2.
    WaveTask* d_xcd_level_tasks = nullptr;
    // Windowed scans: the same tasks binned by their windowed work, longest
    // first within each bin.
    WaveTask* d_xcd_scan_tasks_band = nullptr;
    int32_t* d_xcd_bins_band = nullptr;
    int32_t xcd_max_bin_band = 0;
    int32_t* d_xcd_level_bins = nullptr;
    int32_t xcd_max_level_bin = 0;
    // Segmented refinement with G = 2 (index 0) or 4 (index 1) levels per
    // block: tasks (potential, level l) for l = 0, G, 2G, ... of each
    // potential, in task order and binned per XCD like the level tasks.
    WaveTask* d_group_tasks[2] = {nullptr, nullptr};
    WaveTask* d_xcd_group_tasks[2] = {nullptr, nullptr};
    int32_t* d_xcd_group_bins[2] = {nullptr, nullptr};
    int32_t xcd_max_group_bin[2] = {0, 0};
    int32_t n_group_tasks[2] = {0, 0};
    int32_t segmented_levels = 0;  // tuning: levels per segmented block (0: automatic)
    int32_t last_refine_levels = 1;
    bool xcd_scan = true;
    WaveTask* d_refine_tasks = nullptr;
    WaveTask* d_level_tasks = nullptr;
    // Lane-mode refinement tasks: 64 levels (one per lane) for potentials
    // with at most 64 levels, 128 levels (two per lane) for the others.
    WaveTask* d_lane1_tasks = nullptr;
    WaveTask* d_lane2_tasks = nullptr;
    int32_t n_scan_tasks = 0, n_refine_tasks = 0, n_level_tasks = 0;
    // Cumulative task counts and scan work over the potentials in task order
    // (longest grids first): entry k covers the first k potentials.
    std::vector<int32_t> cum_scan_tasks, cum_refine_tasks, cum_level_tasks, cum_lane1_tasks,
        cum_lane2_tasks, cum_group_tasks[2];
    std::vector<double> cum_scan_work;
    double solve_split = 0.0;
    // The single-stream solve's launches recorded as a HIP graph (tuning
    // `graph`), replayed while its key (batch, tuning, tolerance) holds.
    int32_t graph = 1;
    hipGraphExec_t solve_exec = nullptr;
    SolveKey graph_key{};
    SolveKey last_key{};  // key of the last single-stream solve
    bool has_last_key = false;
    uint64_t loads = 0;  // batches loaded so far: part of the graph key
    bool last_solve_graph = false;
    int32_t levels_per_lane = 1;
    int32_t segmented_lanes = 0;  // 0: chosen from the batch size
    bool solve_timed = false;
    double last_split_used = 0.0;
    int32_t refine_mode = EPSEON_REFINE_AUTO;
    int32_t last_refine_mode = EPSEON_REFINE_AUTO;
    int32_t last_refine_lanes = 0;
    int32_t* d_nodes = nullptr;
    double* d_far = nullptr;
    int32_t* d_resc = nullptr;
    double* d_levels = nullptr;
    int32_t* d_status = nullptr;
    int32_t* d_iters = nullptr;
    LevelNorm* d_norms = nullptr;
    double* d_grad_levels = nullptr;
    // Adjoint tasks (potential, side, segment) grouped by block width (1..4
    // wavefronts), and the eigenfunction tasks (potential, 64 levels, side,
    // segment) in potential task order with cumulative counts.
    PsiTask* d_vjp_lists[4] = {};
    int32_t n_vjp[4] = {};
    PsiTask* d_psi_tasks = nullptr;
    std::vector<int32_t> cum_psi_tasks;
    // Eigenfunction-shot checkpoints written by the normalisation pass
    // (carved from the arena on first use).
    PsiCheckpoint* d_ckpt = nullptr;
    PsiSegMatrix* d_psi_mats = nullptr;
    int64_t ckpt_total = 0;
    // Scratch for host-side outputs, carved from the arena on first use.
    double* d_stage = nullptr;
    size_t stage_elems = 0;
    int64_t user_span = 0;
    int64_t psi_total = 0;
    // Position of each potential in task order.
    std::vector<int32_t> order_pos;

    // Host views of the last refine's levels, status and iterations:
    // pointers into the pinned read-back block, set by fetch_levels (valid
    // while host_levels_valid).
    const double* h_levels = nullptr;
    const int32_t* h_status = nullptr;
    const int32_t* h_iters = nullptr;
    bool host_levels_valid = false;
    // The last refine's summary (level_summary_kernel, enqueued after every
    // refinement: refinement point-updates, converged levels, eigenfunction
    // steps), read back with the levels.
    int64_t* d_summary = nullptr;
    int64_t h_summary[3] = {0, 0, 0};
    // Levels, status, iterations and steps live in one device block
    // (d_level_block, level_block_bytes) read back by one copy into pinned
    // host memory.
    char* d_level_block = nullptr;
    size_t level_block_bytes = 0;
    void* h_pinned = nullptr;
    size_t h_pinned_bytes = 0;

    bool loaded = false, scanned = false, refined = false, normalized = false;
    bool scan_timed = false, refine_timed = false, normalize_timed = false, vjp_timed = false;
    double scan_point_updates = 0.0;       // of the last scan
    double scan_pu_full = 0.0, scan_pu_band = 0.0;
    std::vector<BandWindow> h_win_full, h_win_band;

    explicit epseon_ctx(size_t block_bytes) : arena(block_bytes) {}
};

namespace {

int set_device(epseon_ctx* ctx)
{
    EPSEON_HIP_TRY(hipSetDevice(ctx->device));
    return EPSEON_OK;
}

int check_ctx(epseon_ctx* ctx)
{
    if (ctx == nullptr)
        return fail(EPSEON_ERR_INVALID, "null context");
    return set_device(ctx);
}

}  // namespace

// The entry points below have C linkage from their declarations in
// include/epseon_hip.h.

const char* epseon_last_error(void) { return g_last_error.c_str(); }

int epseon_struct_sizes(int64_t* sizes, int32_t count)
{
    if (sizes == nullptr || count < 0)
        return fail(EPSEON_ERR_INVALID, "invalid size buffer");
    const int64_t all[] = {sizeof(epseon_device_info), sizeof(epseon_potential_desc),
                           sizeof(epseon_potential_info), sizeof(epseon_stats),
                           sizeof(epseon_tuning), sizeof(epseon_bracket)};
    constexpr int32_t n_all = static_cast<int32_t>(sizeof all / sizeof all[0]);
    for (int32_t i = 0; i < count && i < n_all; ++i)
        sizes[i] = all[i];
    return EPSEON_OK;
}

int epseon_device_count(int32_t* count)
{
    if (count == nullptr)
        return fail(EPSEON_ERR_INVALID, "null count pointer");
    int n = 0;
    const hipError_t e = hipGetDeviceCount(&n);
    if (e == hipErrorNoDevice) {
        *count = 0;
        return EPSEON_OK;
    }
    if (e != hipSuccess) {
        *count = 0;
        return hip_fail(e, "hipGetDeviceCount");
    }
    *count = n;
    return EPSEON_OK;
}

int epseon_device_info_get(int32_t device, epseon_device_info* info)
{
    if (info == nullptr)
        return fail(EPSEON_ERR_INVALID, "null info pointer");
    int32_t n = 0;
    if (const int rc = epseon_device_count(&n); rc != EPSEON_OK)
        return rc;
    if (device < 0 || device >= n)
        return fail(EPSEON_ERR_DEVICE, "Device not available.");
    hipDeviceProp_t p;
    EPSEON_HIP_TRY(hipGetDeviceProperties(&p, device));
    std::memset(info, 0, sizeof *info);
    std::snprintf(info->name, sizeof info->name, "%s", p.name);
    std::snprintf(info->arch, sizeof info->arch, "%s", p.gcnArchName);
    int rv = 0, dv = 0;
    (void)hipRuntimeGetVersion(&rv);
    (void)hipDriverGetVersion(&dv);
    info->runtime_version = rv;
    info->driver_version = dv;
    info->vendor_id = 0x1002;
    info->pci_device_id = p.pciDeviceID;
    info->compute_units = p.multiProcessorCount;
    info->max_threads_per_block = p.maxThreadsPerBlock;
    for (int i = 0; i < 3; ++i) {
        info->max_block_dim[i] = p.maxThreadsDim[i];
        info->max_grid_dim[i] = p.maxGridSize[i];
    }
    info->shared_mem_per_block = static_cast<int64_t>(p.sharedMemPerBlock);
    info->shared_mem_per_cu = static_cast<int64_t>(p.maxSharedMemoryPerMultiProcessor);
    info->total_global_mem = static_cast<int64_t>(p.totalGlobalMem);
    info->clock_khz = p.clockRate;
    info->warp_size = p.warpSize;
    std::memcpy(info->uuid, p.uuid.bytes, 16);
    return EPSEON_OK;
}

int epseon_ctx_create(int32_t device, uint64_t block_bytes, epseon_ctx** out)
{
    if (out == nullptr)
        return fail(EPSEON_ERR_INVALID, "null output pointer");
    *out = nullptr;
    int32_t n = 0;
    if (const int rc = epseon_device_count(&n); rc != EPSEON_OK)
        return rc;
    if (device < 0 || device >= n)
        return fail(EPSEON_ERR_DEVICE, "Device not available.");
    const size_t bb = block_bytes ? static_cast<size_t>(block_bytes) : (size_t(64) << 20);
    epseon_ctx* ctx = new (std::nothrow) epseon_ctx(bb);
    if (ctx == nullptr)
        return fail(EPSEON_ERR_MEMORY, "host allocation failed");
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&ctx->stream3, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&ctx->stream4, hipStreamNonBlocking);
    for (int i = 0; i < 3 && e == hipSuccess; ++i)
        e = hipEventCreateWithFlags(&ctx->ev_join_aux[i], hipEventDisableTiming);
    for (int i = 0; i < 8 && e == hipSuccess; ++i)
        e = hipEventCreate(&ctx->ev[i]);
    if (e == hipSuccess)
        e = hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess)
        e = hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming);
    if (e == hipSuccess)
        e = hipEventCreate(&ctx->ev_solve0);
    if (e == hipSuccess)
        e = hipEventCreate(&ctx->ev_solve1);
    if (e == hipSuccess)
        e = hipEventCreateWithFlags(&ctx->ev_tables, hipEventDisableTiming);
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&ctx->compute_units, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess)
        e = scratch_setup(device);
    if (e != hipSuccess) {
        const int rc = hip_fail(e, "context creation");
        epseon_ctx_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return EPSEON_OK;
}

int epseon_ctx_destroy(epseon_ctx* ctx)
{
    if (ctx == nullptr)
        return EPSEON_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream != nullptr)
        (void)hipStreamSynchronize(ctx->stream);
    for (hipStream_t st : {ctx->stream2, ctx->stream3, ctx->stream4})
        if (st != nullptr)
            (void)hipStreamSynchronize(st);
    if (ctx->solve_exec != nullptr)
        (void)hipGraphExecDestroy(ctx->solve_exec);
    if (ctx->h_pinned != nullptr)
        (void)hipHostFree(ctx->h_pinned);
    ctx->arena.release();
    for (hipEvent_t& ev : ctx->ev)
        if (ev != nullptr)
            (void)hipEventDestroy(ev);
    for (hipEvent_t ev : {ctx->ev_fork, ctx->ev_join, ctx->ev_solve0, ctx->ev_solve1, ctx->ev_tables,
                          ctx->ev_join_aux[0], ctx->ev_join_aux[1], ctx->ev_join_aux[2]})
        if (ev != nullptr)
            (void)hipEventDestroy(ev);
    if (ctx->stream != nullptr)
        (void)hipStreamDestroy(ctx->stream);
    for (hipStream_t st : {ctx->stream2, ctx->stream3, ctx->stream4})
        if (st != nullptr)
            (void)hipStreamDestroy(st);
    delete ctx;
    return EPSEON_OK;
}

namespace {

int32_t choose_refine_mode(const epseon_ctx* ctx);
int32_t segmented_lanes(const epseon_ctx* ctx);
bool windows_for(const epseon_ctx* ctx, int32_t mode);

// Build the segment tables of the band windows (band) or of the whole
// grids on `stream`, once per load.
hipError_t build_tables(epseon_ctx* ctx, hipStream_t stream, bool band)
{
    bool& built = band ? ctx->tables_band : ctx->tables_full;
    if (built || !ctx->segtab_ok)
        return hipSuccess;
    built = true;
    if (band)
        return launch_segment_tables(stream, ctx->d_uv, ctx->d_rec, ctx->d_table_jobs + ctx->n_jobs_full,
                                     ctx->n_jobs_band, ctx->table_max_rows, ctx->d_segtab_band);
    return launch_segment_tables(stream, ctx->d_uv, ctx->d_rec, ctx->d_table_jobs, ctx->n_jobs_full,
                                 ctx->table_max_rows, ctx->d_segtab_full);
}

// Shared body of epseon_load_potentials (v_host) and
// epseon_load_potentials_device (v_dev).
// Band of the scan group starting at trial energy j0 (as the kernels
// compute it).
int32_t host_group_band(const PotentialRecord& r, int32_t j0)
{
    const int32_t jt = std::min(j0 + 63, r.n_energies - 1);
    const double et = std::fma(static_cast<double>(jt + 1), r.de, r.e_lo);
    const double step = (r.e_hi - r.e_lo) / kBands;
    for (int32_t b = 0; b < kBands - 1; ++b)
        if (std::fma(static_cast<double>(b + 1), step, r.e_lo) >= et)
            return b;
    return kBands - 1;
}

// Per-XCD bins of a task list (potentials in task order, potential k's
// tasks at [cum[k], cum[k + 1]), task t worth work[t]).  Each potential's
// tasks are cut into chunks of at most ceil(tasks / 32) (so a single large
// potential still spreads over every XCD, while typical batch members stay
// whole); chunks go by decreasing work to the least-loaded bin.  Each bin
// keeps the chunks in task order (a potential's tasks stay together, so its
// table stays in that XCD's L2); with `longest_first` the tasks of each
// chunk go by decreasing work (stable).  Returns the concatenated bins;
// `bins` gets the 9 offsets and `max_bin` the largest bin's size.
std::vector<WaveTask> xcd_binned(const std::vector<WaveTask>& tasks, const std::vector<int32_t>& cum,
                                 const std::vector<double>& work, bool longest_first, std::vector<int32_t>& bins,
                                 int32_t& max_bin)
{
    struct Chunk {
        int32_t t0, t1;
        double work;
    };
    std::vector<Chunk> chunks;
    const int32_t cmax = std::max<int32_t>(1, (static_cast<int32_t>(tasks.size()) + 31) / 32);
    for (size_t k = 0; k + 1 < cum.size(); ++k)
        for (int32_t t = cum[k]; t < cum[k + 1]; t += cmax) {
            const int32_t e = std::min(cum[k + 1], t + cmax);
            double w = 0.0;
            for (int32_t i = t; i < e; ++i)
                w += work[static_cast<size_t>(i)];
            chunks.push_back({t, e, w});
        }
    std::vector<size_t> by_work(chunks.size());
    std::iota(by_work.begin(), by_work.end(), size_t(0));
    std::stable_sort(by_work.begin(), by_work.end(),
                     [&](size_t a, size_t b) { return chunks[a].work > chunks[b].work; });
    std::vector<int32_t> chunk_bin(chunks.size(), 0);
    double load[8] = {};
    for (size_t c : by_work) {
        const int x = static_cast<int>(std::min_element(load, load + 8) - load);
        chunk_bin[c] = x;
        load[x] += chunks[c].work;
    }
    std::vector<WaveTask> out;
    out.reserve(tasks.size());
    bins.assign(9, 0);
    max_bin = 0;
    for (int x = 0; x < 8; ++x) {
        bins[static_cast<size_t>(x)] = static_cast<int32_t>(out.size());
        for (size_t c = 0; c < chunks.size(); ++c) {  // chunks are in task order
            if (chunk_bin[c] != x)
                continue;
            std::vector<int32_t> ids;
            for (int32_t t = chunks[c].t0; t < chunks[c].t1; ++t)
                ids.push_back(t);
            if (longest_first)
                std::stable_sort(ids.begin(), ids.end(), [&](int32_t a, int32_t b) {
                    return work[static_cast<size_t>(a)] > work[static_cast<size_t>(b)];
                });
            for (int32_t t : ids)
                out.push_back(tasks[static_cast<size_t>(t)]);
        }
        max_bin = std::max(max_bin, static_cast<int32_t>(out.size()) - bins[static_cast<size_t>(x)]);
    }
    bins[8] = static_cast<int32_t>(out.size());
    return out;
}

int load_batch(epseon_ctx* ctx, int32_t count, const epseon_potential_desc* desc, const double* v_host,
               const double* v_dev, const int64_t* v_offsets)
{
    if (const int rc = check_ctx(ctx); rc != EPSEON_OK)
        return rc;
    if (count < 0 ||
        (count > 0 && (desc == nullptr || (v_host == nullptr && v_dev == nullptr) || v_offsets == nullptr)))
        return fail(EPSEON_ERR_INVALID, "invalid potential batch");
    EPSEON_HIP_TRY(hipStreamSynchronize(ctx->stream));
    EPSEON_HIP_TRY(hipStreamSynchronize(ctx->stream2));  // tables of the previous load
    ctx->tables_pending = false;
    // The captured solve points into the arena that this load reuses.
    ++ctx->loads;
    if (ctx->solve_exec != nullptr) {
        EPSEON_HIP_TRY(hipGraphExecDestroy(ctx->solve_exec));
        ctx->solve_exec = nullptr;
    }
    ctx->loaded = ctx->scanned = ctx->refined = ctx->normalized = false;
    ctx->scan_timed = ctx->refine_timed = ctx->normalize_timed = ctx->vjp_timed = false;
    ctx->solve_timed = false;
    ctx->d_stage = nullptr;
    ctx->stage_elems = 0;
    ctx->d_ckpt = nullptr;
    ctx->d_psi_mats = nullptr;
    ctx->desc.assign(desc, desc + count);
    ctx->rec.assign(static_cast<size_t>(count), PotentialRecord{});

    ctx->arena.reset();
    Arena& a = ctx->arena;
    char msg[256];
    // Descriptor checks (no V access).
    std::vector<double> betas(static_cast<size_t>(std::max(count, 1)));
    int64_t user_span = 0;
    for (int32_t p = 0; p < count; ++p) {
        const epseon_potential_desc& d = desc[p];
        if (d.n_points < 4 || d.n_points > (int64_t(1) << 30)) {
            std::snprintf(msg, sizeof msg, "potential %d: point count %lld outside [4, 2^30]", p,
                          static_cast<long long>(d.n_points));
            return fail(EPSEON_ERR_INVALID, msg);
        }
        if (d.n_energies < 1 || d.n_energies > (int64_t(1) << 30)) {
            std::snprintf(msg, sizeof msg, "potential %d: trial energy count %lld outside [1, 2^30]",
                          p, static_cast<long long>(d.n_energies));
            return fail(EPSEON_ERR_INVALID, msg);
        }
        if (!(std::isfinite(d.r_min) && std::isfinite(d.r_max) && d.r_max > d.r_min)) {
            std::snprintf(msg, sizeof msg, "potential %d: need finite min_r < max_r", p);
            return fail(EPSEON_ERR_INVALID, msg);
        }
        if (!(std::isfinite(d.reduced_mass) && d.reduced_mass > 0.0)) {
            std::snprintf(msg, sizeof msg, "potential %d: reduced mass must be positive", p);
            return fail(EPSEON_ERR_INVALID, msg);
        }
        if (d.level_min < 0 || d.level_count < 0) {
            std::snprintf(msg, sizeof msg, "potential %d: negative level range", p);
            return fail(EPSEON_ERR_INVALID, msg);
        }
        if (v_offsets[p] < 0) {
            std::snprintf(msg, sizeof msg, "potential %d: negative V offset", p);
            return fail(EPSEON_ERR_INVALID, msg);
        }
        const double h = (d.r_max - d.r_min) / static_cast<double>(d.n_points - 1);
        betas[static_cast<size_t>(p)] = (h * h * d.reduced_mass) / (12.0 * kHbar2Over2Amu);
        user_span = std::max<int64_t>(user_span, v_offsets[p] + d.n_points);
    }

    // V statistics: min V, its first index, the wall-trimmed Dirichlet points
    // (the last point left of the minimum and the first point right of it
    // where V >= V_min + 1/(2 beta); inside, 1 + beta (E - V_i) > 1/2 for
    // every E >= V_min) and the first non-finite value.
    std::vector<PotentialStats> vstats(static_cast<size_t>(std::max(count, 1)));
    if (v_host != nullptr) {
        for (int32_t p = 0; p < count; ++p) {
            const int64_t n = desc[p].n_points;
            const double* vp = v_host + v_offsets[p];
            PotentialStats& st = vstats[static_cast<size_t>(p)];
            st = PotentialStats{0.0, 0, 0, n - 1, n};
            for (int64_t i = 0; i < n; ++i) {
                if (!std::isfinite(vp[i])) {
                    st.bad = i;
                    break;
                }
                if (vp[i] < vp[st.imin])
                    st.imin = i;
            }
            if (st.bad < n)
                continue;
            st.vmin = vp[st.imin];
            const double wall = st.vmin + 0.5 / betas[static_cast<size_t>(p)];
            for (int64_t i = st.imin; i >= 0; --i)
                if (vp[i] >= wall) {
                    st.first = i;
                    break;
                }
            for (int64_t i = st.imin; i < n; ++i)
                if (vp[i] >= wall) {
                    st.last = i;
                    break;
                }
        }
    } else if (count > 0) {
        std::vector<int64_t> counts(static_cast<size_t>(count));
        for (int32_t p = 0; p < count; ++p)
            counts[static_cast<size_t>(p)] = desc[p].n_points;
        int64_t* d_off = nullptr;
        int64_t* d_cnt = nullptr;
        double* d_beta = nullptr;
        PotentialStats* d_st = nullptr;
        EPSEON_HIP_TRY(arena_array(a, static_cast<size_t>(count), &d_off));
        EPSEON_HIP_TRY(arena_array(a, static_cast<size_t>(count), &d_cnt));
        EPSEON_HIP_TRY(arena_array(a, static_cast<size_t>(count), &d_beta));
        EPSEON_HIP_TRY(arena_array(a, static_cast<size_t>(count), &d_st));
        EPSEON_HIP_TRY(hipMemcpyAsync(d_off, v_offsets, count * sizeof(int64_t), hipMemcpyHostToDevice,
                                      ctx->stream));
        EPSEON_HIP_TRY(hipMemcpyAsync(d_cnt, counts.data(), count * sizeof(int64_t),
                                      hipMemcpyHostToDevice, ctx->stream));
        EPSEON_HIP_TRY(hipMemcpyAsync(d_beta, betas.data(), count * sizeof(double),
                                      hipMemcpyHostToDevice, ctx->stream));
        EPSEON_HIP_TRY(launch_potential_stats(ctx->stream, v_dev, d_off, d_cnt, d_beta, count, d_st));
        EPSEON_HIP_TRY(hipMemcpyAsync(vstats.data(), d_st, count * sizeof(PotentialStats),
                                      hipMemcpyDeviceToHost, ctx->stream));
        EPSEON_HIP_TRY(hipStreamSynchronize(ctx->stream));
    }

    int64_t pts = 0, en = 0, lv = 0, psi_total = 0, seg_total = 0, ckpt_total = 0;
    int32_t max_points = 0, max_seg_rows = 0;
    double scan_pu = 0.0;
    for (int32_t p = 0; p < count; ++p) {
        const epseon_potential_desc& d = desc[p];
        const PotentialStats& st = vstats[static_cast<size_t>(p)];
        if (st.bad < d.n_points) {
            std::snprintf(msg, sizeof msg, "potential %d: V[%lld] is not finite", p,
                          static_cast<long long>(st.bad));
            return fail(EPSEON_ERR_INVALID, msg);
        }
        const int64_t imin = st.imin;
        const double vmin = st.vmin;
        if (!(std::isfinite(d.e_top) && d.e_top > vmin)) {
            std::snprintf(msg, sizeof msg,
                          "potential %d: energy window top %.17g is not above min V %.17g", p,
                          d.e_top, vmin);
            return fail(EPSEON_ERR_INVALID, msg);
        }
        PotentialRecord& r = ctx->rec[p];
        r.h = (d.r_max - d.r_min) / static_cast<double>(d.n_points - 1);
        r.beta = betas[static_cast<size_t>(p)];
        const int64_t first = st.first, last = st.last;
        r.first_point = first;
        r.e_lo = vmin;
        r.e_hi = d.e_top;
        r.de = (r.e_hi - r.e_lo) / static_cast<double>(d.n_energies + 1);
        r.abs_tol = 1e-15 * (r.e_hi - r.e_lo);
        r.n_points = static_cast<int32_t>(last - first + 1);
        r.n_energies = static_cast<int32_t>(d.n_energies);
        r.level_min = d.level_min;
        r.level_count = d.level_count;
        r.stride = node_stride(r.beta, r.e_hi, vmin);
        // f_max = 1 + beta (E_hi - V_min) <= 3/2 keeps E_hi below the wall
        // threshold and bounds the per-step growth of the recurrence.
        const double f_max = 1.0 + r.beta * (r.e_hi - vmin);
        r.flags = (f_max <= 1.5 && r.n_points >= 4) ? EPSEON_POT_OK : EPSEON_POT_UNRESOLVED;
        if (r.n_points < 4)
            r.n_points = 4;
        // Two-sided eigenfunction shots meet at min V (inside the classically
        // allowed region of every level); points m and m+1 must be interior.
        r.match = static_cast<int32_t>(
            std::clamp<int64_t>(imin - first, 1, static_cast<int64_t>(r.n_points) - 3));
        r.full_points = static_cast<int32_t>(d.n_points);
        // Scan far-end scale: w_{N-1} = 12^(n-2) v_{N-1}, 12^(n-2) =
        // far_scale * 2^far_exp with far_scale in [1, 2).
        {
            const double l2 = static_cast<double>(r.n_points - 2) * std::log2(12.0);
            const double fl = std::floor(l2);
            r.far_exp = static_cast<int32_t>(fl);
            r.far_scale = std::exp2(l2 - fl);
        }
        r.seg_row = static_cast<int32_t>(seg_total);
        seg_total += seg_rows(r.n_points, r.flags, r.level_count);
        max_seg_rows = std::max(max_seg_rows, seg_rows(r.n_points, r.flags, r.level_count));
        if (seg_total > INT32_MAX)
            return fail(EPSEON_ERR_INVALID, "batch too large for the segment table");
        // Eigenfunction shots run in segments of psi_seg_len steps; the
        // normalisation pass leaves a checkpoint per level at each segment
        // start (slot 0 of each side is unused).
        r.psi_seg_len = psi_seg_len(r.n_points);
        r.psi_segs_left = (r.match + r.psi_seg_len - 1) / r.psi_seg_len;
        r.psi_segs_right = (r.n_points - 2 - r.match + r.psi_seg_len - 1) / r.psi_seg_len;
        r.psi_ckpt = ckpt_total;
        if (r.flags == EPSEON_POT_OK)
            ckpt_total += static_cast<int64_t>(d.level_count) * (r.psi_segs_left + r.psi_segs_right);
        r.user_offset = v_offsets[p];
        r.psi_offset = psi_total;
        psi_total += static_cast<int64_t>(d.level_count) * d.n_points;
        r.v_offset = pts;
        r.ut_offset = pts + first;
        r.energy_offset = en;
        r.level_offset = lv;
        // 64-byte aligned rows with 64 points of slack: the shot prefetches up
        // to 48 points past the last full 32-point super-block.
        pts += ((d.n_points + 7) & ~int64_t(7)) + 64;
        en += d.n_energies;
        lv += d.level_count;
        max_points = std::max(max_points, static_cast<int32_t>(d.n_points));
        if (r.flags == EPSEON_POT_OK)
            scan_pu += static_cast<double>(r.n_points - 2) * static_cast<double>(d.n_energies);
    }
    ctx->total_points = pts;
    ctx->total_energies = en;
    ctx->total_levels = lv;
    ctx->max_points = max_points;
    ctx->scan_pu_full = scan_pu;
    ctx->user_span = user_span;
    ctx->psi_total = psi_total;
    ctx->ckpt_total = ckpt_total;

    // Work lists: one wavefront per 64 trial energies (scan) or 64 levels
    // (refine), longest grids first so the dispatcher drains large items early.
    std::vector<int32_t> order(static_cast<size_t>(count));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return ctx->rec[a].n_points > ctx->rec[b].n_points;
    });
    ctx->order_pos.assign(static_cast<size_t>(count), 0);
    for (int32_t k = 0; k < count; ++k)
        ctx->order_pos[static_cast<size_t>(order[static_cast<size_t>(k)])] = k;
    // Segmented refinement uses one 64-lane block per level.
    std::vector<WaveTask> scan_tasks, refine_tasks, level_tasks, lane1_tasks, lane2_tasks, group_tasks[2];
    std::vector<PsiTask> vjp_lists[4], psi_tasks;
    ctx->cum_psi_tasks.assign(1, 0);
    ctx->cum_lane1_tasks.assign(1, 0);
    ctx->cum_lane2_tasks.assign(1, 0);
    ctx->cum_scan_tasks.assign(1, 0);
    ctx->cum_refine_tasks.assign(1, 0);
    ctx->cum_level_tasks.assign(1, 0);
    ctx->cum_group_tasks[0].assign(1, 0);
    ctx->cum_group_tasks[1].assign(1, 0);
    ctx->cum_scan_work.assign(1, 0.0);
    for (int32_t p : order) {
        const int32_t groups = (ctx->rec[p].level_count + 63) / 64;
        const PotentialRecord& r = ctx->rec[p];
        {
            // Unresolved potentials get one task (side 0, segment 0), which
            // zeroes their rows and gradient.
            std::vector<PsiTask>& vl = vjp_lists[std::clamp(groups, 1, 4) - 1];
            const bool ok = r.flags == EPSEON_POT_OK;
            for (int32_t side = 0; side < (ok ? 2 : 1); ++side) {
                const int32_t segs = !ok ? 1 : (side == 0 ? r.psi_segs_left : r.psi_segs_right);
                for (int32_t sg = 0; sg < segs; ++sg) {
                    vl.push_back({p, 0, side, sg});
                    for (int32_t l = 0; l < r.level_count; l += 64)
                        psi_tasks.push_back({p, l, side, sg});
                }
            }
            ctx->cum_psi_tasks.push_back(static_cast<int32_t>(psi_tasks.size()));
        }
        if (r.flags == EPSEON_POT_OK)
            for (int32_t j = 0; j < r.n_energies; j += 64)
                scan_tasks.push_back({p, j});
        for (int32_t l = 0; l < r.level_count; l += 64)
            refine_tasks.push_back({p, l});
        for (int32_t l = 0; l < r.level_count; ++l)
            level_tasks.push_back({p, l});
        for (int32_t g = 0; g < 2; ++g)
            for (int32_t l = 0; l < r.level_count; l += 2 << g)
                group_tasks[g].push_back({p, l});
        if (r.level_count > 64)
            for (int32_t l = 0; l < r.level_count; l += 128)
                lane2_tasks.push_back({p, l});
        else if (r.level_count > 0)
            lane1_tasks.push_back({p, 0});
        ctx->cum_lane1_tasks.push_back(static_cast<int32_t>(lane1_tasks.size()));
        ctx->cum_lane2_tasks.push_back(static_cast<int32_t>(lane2_tasks.size()));
        ctx->cum_scan_tasks.push_back(static_cast<int32_t>(scan_tasks.size()));
        ctx->cum_refine_tasks.push_back(static_cast<int32_t>(refine_tasks.size()));
        ctx->cum_level_tasks.push_back(static_cast<int32_t>(level_tasks.size()));
        for (int32_t g = 0; g < 2; ++g)
            ctx->cum_group_tasks[g].push_back(static_cast<int32_t>(group_tasks[g].size()));
        ctx->cum_scan_work.push_back(
            ctx->cum_scan_work.back() +
            (r.flags == EPSEON_POT_OK ? static_cast<double>(r.n_points) * r.n_energies : 0.0));
    }
    ctx->n_scan_tasks = static_cast<int32_t>(scan_tasks.size());
    // Work per task: trial energies times recurrence steps (whole grid) for
    // the scan, grid points for a level's refinement.
    std::vector<double> scan_task_work(scan_tasks.size()), level_task_work(level_tasks.size());
    for (size_t t = 0; t < scan_tasks.size(); ++t) {
        const PotentialRecord& r = ctx->rec[static_cast<size_t>(scan_tasks[t].potential)];
        scan_task_work[t] = static_cast<double>(std::min(64, r.n_energies - scan_tasks[t].first)) *
                            static_cast<double>(r.n_points - 2);
    }
    for (size_t t = 0; t < level_tasks.size(); ++t)
        level_task_work[t] = static_cast<double>(ctx->rec[static_cast<size_t>(level_tasks[t].potential)].n_points);
    std::vector<double> group_task_work[2];
    for (int32_t g = 0; g < 2; ++g) {
        group_task_work[g].resize(group_tasks[g].size());
        for (size_t t = 0; t < group_tasks[g].size(); ++t) {
            const PotentialRecord& r = ctx->rec[static_cast<size_t>(group_tasks[g][t].potential)];
            group_task_work[g][t] =
                static_cast<double>(r.n_points) * std::min(2 << g, r.level_count - group_tasks[g][t].first);
        }
    }
    std::vector<int32_t> xcd_bins, xcd_level_bins, xcd_group_bins[2];
    const std::vector<WaveTask> xcd_tasks =
        xcd_binned(scan_tasks, ctx->cum_scan_tasks, scan_task_work, false, xcd_bins, ctx->xcd_max_bin);
    const std::vector<WaveTask> xcd_level_tasks = xcd_binned(level_tasks, ctx->cum_level_tasks, level_task_work,
                                                             false, xcd_level_bins, ctx->xcd_max_level_bin);
    std::vector<WaveTask> xcd_group_tasks[2];
    for (int32_t g = 0; g < 2; ++g)
        xcd_group_tasks[g] = xcd_binned(group_tasks[g], ctx->cum_group_tasks[g], group_task_work[g], false,
                                        xcd_group_bins[g], ctx->xcd_max_group_bin[g]);
    ctx->n_refine_tasks = static_cast<int32_t>(refine_tasks.size());
    ctx->n_level_tasks = static_cast<int32_t>(level_tasks.size());
    for (int32_t g = 0; g < 2; ++g)
        ctx->n_group_tasks[g] = static_cast<int32_t>(group_tasks[g].size());

    const size_t np = static_cast<size_t>(std::max<int64_t>(pts, 1));
    const size_t ne = static_cast<size_t>(std::max<int64_t>(en, 1));
    const size_t nl = static_cast<size_t>(std::max<int64_t>(lv, 1));
    // V in the caller's layout: one upload of the whole span from the host,
    // or the caller's device buffer itself (read only during this call).
    const double* v_src = v_dev;
    if (v_host != nullptr) {
        EPSEON_HIP_TRY(arena_array(a, static_cast<size_t>(std::max<int64_t>(user_span, 1)), &ctx->d_v));
        v_src = ctx->d_v;
    }
    EPSEON_HIP_TRY(arena_array(a, np, &ctx->d_ut));
    EPSEON_HIP_TRY(arena_array(a, np, &ctx->d_uv));
    EPSEON_HIP_TRY(arena_array(a, np, &ctx->d_uvr));
    EPSEON_HIP_TRY(arena_array(a, static_cast<size_t>(std::max<int64_t>(seg_total, 1)) * 64, &ctx->d_segtab_full));
    const size_t nw = std::max<size_t>(static_cast<size_t>(count) * kBands, 1);
    int2* d_raw_windows = nullptr;
    EPSEON_HIP_TRY(arena_array(a, nw, &d_raw_windows));
    EPSEON_HIP_TRY(arena_array(a, nw, &ctx->d_win_full));
    EPSEON_HIP_TRY(arena_array(a, nw, &ctx->d_win_band));
    EPSEON_HIP_TRY(arena_array(a, nl, &ctx->d_brackets));
    EPSEON_HIP_TRY(arena_array(a, nl, &ctx->d_points));
    EPSEON_HIP_TRY(arena_array(a, std::max<size_t>(count, 1), &ctx->d_rec));
    EPSEON_HIP_TRY(arena_array(a, std::max<size_t>(scan_tasks.size(), 1), &ctx->d_scan_tasks));
    EPSEON_HIP_TRY(arena_array(a, std::max<size_t>(xcd_tasks.size(), 1), &ctx->d_xcd_scan_tasks));
    EPSEON_HIP_TRY(arena_array(a, size_t(9), &ctx->d_xcd_bins));
    EPSEON_HIP_TRY(arena_array(a, std::max<size_t>(xcd_level_tasks.size(), 1), &ctx->d_xcd_level_tasks));
    EPSEON_HIP_TRY(arena_array(a, size_t(9), &ctx->d_xcd_level_bins));
    for (int32_t g = 0; g < 2; ++g) {
        EPSEON_HIP_TRY(arena_array(a, std::max<size_t>(group_tasks[g].size(), 1), &ctx->d_group_tasks[g]));
        EPSEON_HIP_TRY(arena_array(a, std::max<size_t>(xcd_group_tasks[g].size(), 1), &ctx->d_xcd_group_tasks[g]));
        EPSEON_HIP_TRY(arena_array(a, size_t(9), &ctx->d_xcd_group_bins[g]));
    }
    EPSEON_HIP_TRY(arena_array(a, std::max<size_t>(refine_tasks.size(), 1), &ctx->d_refine_tasks));
    EPSEON_HIP_TRY(arena_array(a, std::max<size_t>(level_tasks.size(), 1), &ctx->d_level_tasks));
    EPSEON_HIP_TRY(arena_array(a, std::max<size_t>(lane1_tasks.size(), 1), &ctx->d_lane1_tasks));
    EPSEON_HIP_TRY(arena_array(a, std::max<size_t>(lane2_tasks.size(), 1), &ctx->d_lane2_tasks));
    EPSEON_HIP_TRY(arena_array(a, ne, &ctx->d_nodes));
    EPSEON_HIP_TRY(arena_array(a, ne, &ctx->d_far));
    EPSEON_HIP_TRY(arena_array(a, ne, &ctx->d_resc));
    {
        // One block: levels (f64), status, iterations (i32) and the summary
        // (3 i64), each 256-byte aligned, read back whole; then the steps
        // (i32), which stay on the device.
        const size_t a8 = (nl * sizeof(double) + 255) & ~size_t(255);
        const size_t a4 = (nl * sizeof(int32_t) + 255) & ~size_t(255);
        ctx->level_block_bytes = a8 + 2 * a4 + 256;
        EPSEON_HIP_TRY(arena_array(a, ctx->level_block_bytes + a4, &ctx->d_level_block));
        ctx->d_levels = reinterpret_cast<double*>(ctx->d_level_block);
        ctx->d_status = reinterpret_cast<int32_t*>(ctx->d_level_block + a8);
        ctx->d_iters = reinterpret_cast<int32_t*>(ctx->d_level_block + a8 + a4);
        ctx->d_summary = reinterpret_cast<int64_t*>(ctx->d_level_block + a8 + 2 * a4);
        ctx->d_steps = reinterpret_cast<int32_t*>(ctx->d_level_block + ctx->level_block_bytes);
    }
    EPSEON_HIP_TRY(arena_array(a, nl, &ctx->d_norms));
    EPSEON_HIP_TRY(arena_array(a, nl, &ctx->d_grad_levels));
    for (int w = 0; w < 4; ++w) {
        ctx->n_vjp[w] = static_cast<int32_t>(vjp_lists[w].size());
        EPSEON_HIP_TRY(arena_array(a, std::max<size_t>(vjp_lists[w].size(), 1), &ctx->d_vjp_lists[w]));
        if (!vjp_lists[w].empty())
            EPSEON_HIP_TRY(hipMemcpyAsync(ctx->d_vjp_lists[w], vjp_lists[w].data(),
                                          vjp_lists[w].size() * sizeof(PsiTask),
                                          hipMemcpyHostToDevice, ctx->stream));
    }
    EPSEON_HIP_TRY(arena_array(a, std::max<size_t>(psi_tasks.size(), 1), &ctx->d_psi_tasks));
    if (!psi_tasks.empty())
        EPSEON_HIP_TRY(hipMemcpyAsync(ctx->d_psi_tasks, psi_tasks.data(), psi_tasks.size() * sizeof(PsiTask),
                                      hipMemcpyHostToDevice, ctx->stream));

    if (v_host != nullptr && user_span > 0)
        EPSEON_HIP_TRY(hipMemcpyAsync(ctx->d_v, v_host, static_cast<size_t>(user_span) * sizeof(double),
                                      hipMemcpyHostToDevice, ctx->stream));
    if (count > 0)
        EPSEON_HIP_TRY(hipMemcpyAsync(ctx->d_rec, ctx->rec.data(),
                                      static_cast<size_t>(count) * sizeof(PotentialRecord),
                                      hipMemcpyHostToDevice, ctx->stream));
    if (!scan_tasks.empty())
        EPSEON_HIP_TRY(hipMemcpyAsync(ctx->d_scan_tasks, scan_tasks.data(),
                                      scan_tasks.size() * sizeof(WaveTask), hipMemcpyHostToDevice,
                                      ctx->stream));
    if (!xcd_tasks.empty())
        EPSEON_HIP_TRY(hipMemcpyAsync(ctx->d_xcd_scan_tasks, xcd_tasks.data(),
                                      xcd_tasks.size() * sizeof(WaveTask), hipMemcpyHostToDevice,
                                      ctx->stream));
    if (!xcd_level_tasks.empty())
        EPSEON_HIP_TRY(hipMemcpyAsync(ctx->d_xcd_level_tasks, xcd_level_tasks.data(),
                                      xcd_level_tasks.size() * sizeof(WaveTask), hipMemcpyHostToDevice,
                                      ctx->stream));
    EPSEON_HIP_TRY(hipMemcpyAsync(ctx->d_xcd_level_bins, xcd_level_bins.data(), 9 * sizeof(int32_t),
                                  hipMemcpyHostToDevice, ctx->stream));
    for (int32_t g = 0; g < 2; ++g) {
        if (!group_tasks[g].empty())
            EPSEON_HIP_TRY(hipMemcpyAsync(ctx->d_group_tasks[g], group_tasks[g].data(),
                                          group_tasks[g].size() * sizeof(WaveTask), hipMemcpyHostToDevice,
                                          ctx->stream));
        if (!xcd_group_tasks[g].empty())
            EPSEON_HIP_TRY(hipMemcpyAsync(ctx->d_xcd_group_tasks[g], xcd_group_tasks[g].data(),
                                          xcd_group_tasks[g].size() * sizeof(WaveTask), hipMemcpyHostToDevice,
                                          ctx->stream));
        EPSEON_HIP_TRY(hipMemcpyAsync(ctx->d_xcd_group_bins[g], xcd_group_bins[g].data(), 9 * sizeof(int32_t),
                                      hipMemcpyHostToDevice, ctx->stream));
    }
    EPSEON_HIP_TRY(hipMemcpyAsync(ctx->d_xcd_bins, xcd_bins.data(), 9 * sizeof(int32_t),
                                  hipMemcpyHostToDevice, ctx->stream));
    if (!refine_tasks.empty())
        EPSEON_HIP_TRY(hipMemcpyAsync(ctx->d_refine_tasks, refine_tasks.data(),
                                      refine_tasks.size() * sizeof(WaveTask),
                                      hipMemcpyHostToDevice, ctx->stream));
    if (!level_tasks.empty())
        EPSEON_HIP_TRY(hipMemcpyAsync(ctx->d_level_tasks, level_tasks.data(),
                                      level_tasks.size() * sizeof(WaveTask), hipMemcpyHostToDevice,
                                      ctx->stream));
    if (!lane1_tasks.empty())
        EPSEON_HIP_TRY(hipMemcpyAsync(ctx->d_lane1_tasks, lane1_tasks.data(),
                                      lane1_tasks.size() * sizeof(WaveTask), hipMemcpyHostToDevice,
                                      ctx->stream));
    if (!lane2_tasks.empty())
        EPSEON_HIP_TRY(hipMemcpyAsync(ctx->d_lane2_tasks, lane2_tasks.data(),
                                      lane2_tasks.size() * sizeof(WaveTask), hipMemcpyHostToDevice,
                                      ctx->stream));
    EPSEON_HIP_TRY(hipMemsetAsync(ctx->d_nodes, 0, ne * sizeof(int32_t), ctx->stream));
    EPSEON_HIP_TRY(hipMemsetAsync(ctx->d_far, 0, ne * sizeof(double), ctx->stream));
    EPSEON_HIP_TRY(hipMemsetAsync(ctx->d_resc, 0, ne * sizeof(int32_t), ctx->stream));
    EPSEON_HIP_TRY(launch_shift_potential(ctx->stream, v_src, ctx->d_rec, count, max_points,
                                          ctx->d_ut, ctx->d_uv, ctx->d_uvr));
    // Band windows: computed on the device, read back to size the segment
    // tables and to give each window its far-end scale.
    std::vector<int2> raw(static_cast<size_t>(count) * kBands);
    EPSEON_HIP_TRY(launch_band_windows(ctx->stream, ctx->d_uv, ctx->d_rec, count, d_raw_windows));
    if (count > 0)
        EPSEON_HIP_TRY(hipMemcpyAsync(raw.data(), d_raw_windows, raw.size() * sizeof(int2),
                                      hipMemcpyDeviceToHost, ctx->stream));
    EPSEON_HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::vector<BandWindow> win_full(raw.size()), win_band(raw.size());
    std::vector<TableJob> jobs_full, jobs_band;
    int64_t band_total = 0;
    int32_t max_rows = max_seg_rows;
    double scan_pu_band = 0.0;
    for (int32_t p = 0; p < count; ++p) {
        const PotentialRecord& r = ctx->rec[static_cast<size_t>(p)];
        const bool tables = seg_rows(r.n_points, r.flags, r.level_count) > 0;
        if (tables)
            jobs_full.push_back({p, 0, r.n_points, r.seg_row});
        for (int32_t b = 0; b < kBands; ++b) {
            const size_t i = static_cast<size_t>(p) * kBands + static_cast<size_t>(b);
            win_full[i] = BandWindow{0, r.n_points - 1, r.far_exp, r.seg_row, r.far_scale};
            BandWindow w{raw[i].x, raw[i].y, 0, 0, 1.0};
            const int32_t n = w.hi - w.lo + 1;
            const double l2 = static_cast<double>(n - 2) * std::log2(12.0);
            const double fl = std::floor(l2);
            w.far_exp = static_cast<int32_t>(fl);
            w.far_scale = std::exp2(l2 - fl);
            if (b > 0 && win_band[i - 1].lo == w.lo && win_band[i - 1].hi == w.hi) {
                w.seg_row = win_band[i - 1].seg_row;
            } else if (tables) {
                w.seg_row = static_cast<int32_t>(band_total);
                jobs_band.push_back({p, w.lo, n, w.seg_row});
                band_total += seg_rows(n, 0, 1);
                max_rows = std::max(max_rows, seg_rows(n, 0, 1));
                if (band_total > INT32_MAX)
                    return fail(EPSEON_ERR_INVALID, "batch too large for the segment tables");
            }
            win_band[i] = w;
        }
    }
    // Windowed scan: work per task (trial energies times its band window's
    // steps), its total, and its own XCD bins with the longest tasks first.
    std::vector<double> band_task_work(scan_tasks.size());
    for (size_t t = 0; t < scan_tasks.size(); ++t) {
        const WaveTask& task = scan_tasks[t];
        const PotentialRecord& r = ctx->rec[static_cast<size_t>(task.potential)];
        const BandWindow& w = win_band[static_cast<size_t>(task.potential) * kBands +
                                       static_cast<size_t>(host_group_band(r, task.first))];
        band_task_work[t] = static_cast<double>(std::min(64, r.n_energies - task.first)) *
                            static_cast<double>(w.hi - w.lo - 1);
        scan_pu_band += band_task_work[t];
    }
    std::vector<int32_t> band_bins;
    const std::vector<WaveTask> band_tasks =
        xcd_binned(scan_tasks, ctx->cum_scan_tasks, band_task_work, true, band_bins, ctx->xcd_max_bin_band);
    EPSEON_HIP_TRY(arena_array(a, std::max<size_t>(band_tasks.size(), 1), &ctx->d_xcd_scan_tasks_band));
    EPSEON_HIP_TRY(arena_array(a, size_t(9), &ctx->d_xcd_bins_band));
    if (!band_tasks.empty())
        EPSEON_HIP_TRY(hipMemcpyAsync(ctx->d_xcd_scan_tasks_band, band_tasks.data(),
                                      band_tasks.size() * sizeof(WaveTask), hipMemcpyHostToDevice, ctx->stream));
    EPSEON_HIP_TRY(hipMemcpyAsync(ctx->d_xcd_bins_band, band_bins.data(), 9 * sizeof(int32_t),
                                  hipMemcpyHostToDevice, ctx->stream));
    ctx->scan_pu_band = scan_pu_band;
    ctx->h_win_full = win_full;
    ctx->h_win_band = win_band;
    ctx->segtab_ok = max_rows <= INT32_MAX / 512;
    EPSEON_HIP_TRY(arena_array(a, static_cast<size_t>(std::max<int64_t>(band_total, 1)) * 64, &ctx->d_segtab_band));
    TableJob* d_jobs = nullptr;
    EPSEON_HIP_TRY(arena_array(a, std::max<size_t>(jobs_full.size() + jobs_band.size(), 1), &d_jobs));
    ctx->d_table_jobs = d_jobs;
    ctx->n_jobs_full = static_cast<int32_t>(jobs_full.size());
    ctx->n_jobs_band = static_cast<int32_t>(jobs_band.size());
    ctx->table_max_rows = max_rows;
    ctx->tables_full = ctx->tables_band = false;
    if (count > 0) {
        EPSEON_HIP_TRY(hipMemcpyAsync(ctx->d_win_full, win_full.data(), win_full.size() * sizeof(BandWindow),
                                      hipMemcpyHostToDevice, ctx->stream));
        EPSEON_HIP_TRY(hipMemcpyAsync(ctx->d_win_band, win_band.data(), win_band.size() * sizeof(BandWindow),
                                      hipMemcpyHostToDevice, ctx->stream));
    }
    if (!jobs_full.empty())
        EPSEON_HIP_TRY(hipMemcpyAsync(d_jobs, jobs_full.data(), jobs_full.size() * sizeof(TableJob),
                                      hipMemcpyHostToDevice, ctx->stream));
    if (!jobs_band.empty())
        EPSEON_HIP_TRY(hipMemcpyAsync(d_jobs + jobs_full.size(), jobs_band.data(),
                                      jobs_band.size() * sizeof(TableJob), hipMemcpyHostToDevice,
                                      ctx->stream));
    if (ctx->segtab_ok && jobs_full.size() + jobs_band.size() > 0) {
        // The set the current tuning's solve will use, built on stream2 while
        // the next scan runs; the refinement waits for it (tables_pending).
        const int32_t mode = choose_refine_mode(ctx);
        if (mode == EPSEON_REFINE_SEGMENTED && segmented_lanes(ctx) == 64) {
            EPSEON_HIP_TRY(hipEventRecord(ctx->ev_fork, ctx->stream));
            EPSEON_HIP_TRY(hipStreamWaitEvent(ctx->stream2, ctx->ev_fork, 0));
            EPSEON_HIP_TRY(build_tables(ctx, ctx->stream2, windows_for(ctx, mode)));
            EPSEON_HIP_TRY(hipEventRecord(ctx->ev_tables, ctx->stream2));
            ctx->tables_pending = true;
        }
    }
    // The host buffers are pageable and owned by the caller: finish the
    // copies before returning.
    EPSEON_HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->loaded = true;
    return EPSEON_OK;
}

}  // namespace

int epseon_load_potentials(epseon_ctx* ctx, int32_t count, const epseon_potential_desc* desc,
                           const double* v, const int64_t* v_offsets)
{
    return load_batch(ctx, count, desc, v, nullptr, v_offsets);
}

int epseon_load_potentials_device(epseon_ctx* ctx, int32_t count, const epseon_potential_desc* desc,
                                  const double* v_device, const int64_t* v_offsets)
{
    return load_batch(ctx, count, desc, nullptr, v_device, v_offsets);
}

int epseon_potential_info_get(epseon_ctx* ctx, int32_t p, epseon_potential_info* info)
{
    if (ctx == nullptr || info == nullptr)
        return fail(EPSEON_ERR_INVALID, "null argument");
    if (!ctx->loaded)
        return fail(EPSEON_ERR_STATE, "no potentials loaded");
    if (p < 0 || p >= static_cast<int32_t>(ctx->rec.size()))
        return fail(EPSEON_ERR_INVALID, "potential index out of range");
    const PotentialRecord& r = ctx->rec[p];
    info->beta = r.beta;
    info->h = r.h;
    info->e_lo = r.e_lo;
    info->e_hi = r.e_hi;
    info->stride = r.stride;
    info->flags = r.flags;
    info->first_point = r.first_point;
    info->point_count = r.n_points;
    info->far_exponent = r.far_exp;
    info->reserved = 0;
    return EPSEON_OK;
}

namespace {

// Windows pay when the scan has several wavefronts per SIMD: a SIMD then
// takes new wavefronts as short-window ones finish.  With about one per SIMD
// (a single large potential) the full-window wavefronts set the time anyway.
bool windows_enabled(const epseon_ctx* ctx)
{
    if (ctx->windows != 2)
        return ctx->windows == 1;
    return static_cast<int64_t>(ctx->n_scan_tasks) >= kWindowsMinScanWavesPerSimd * 4 * ctx->compute_units;
}

int32_t choose_refine_mode(const epseon_ctx* ctx)
{
    int32_t mode = ctx->refine_mode;
    if (mode == EPSEON_REFINE_AUTO)
        mode = !windows_enabled(ctx) && ctx->total_levels >= kLanePerLevelMinLevels ? EPSEON_REFINE_LANE
                                                                                     : EPSEON_REFINE_SEGMENTED;
    return mode;
}

// Segmented blocks: up to 4 wavefronts per level while the batch has too few
// levels to give every SIMD a wavefront (unless fixed by the tuning).
int32_t segmented_lanes(const epseon_ctx* ctx)
{
    if (ctx->segmented_lanes != 0)
        return ctx->segmented_lanes;
    int32_t lanes = 64;
    while (lanes < 256 && ctx->total_levels * (lanes / 64) * 2 <= 2048)
        lanes *= 2;
    return lanes;
}

// Levels per segmented block: in 64-lane blocks four (each table load then
// feeds the shots of every level of the group in the same window) once the
// batch has at least kQuadMinLevelsPerCu levels per compute unit, two from
// kPairMinLevelsPerCu, so that the groups still give every SIMD several
// blocks; else one (unless fixed by the tuning).
int32_t segmented_levels(const epseon_ctx* ctx, int32_t lanes)
{
    if (lanes != 64)
        return 1;
    if (ctx->segmented_levels != 0)
        return ctx->segmented_levels;
    if (ctx->total_levels >= kQuadMinLevelsPerCu * ctx->compute_units)
        return 4;
    return ctx->total_levels >= kPairMinLevelsPerCu * ctx->compute_units ? 2 : 1;
}

// Window set of a scan followed by a refinement in `mode`: the band windows
// when enabled, except for lane mode (whose wavefronts walk one table for
// all their levels), which shoots over the whole grid.
bool windows_for(const epseon_ctx* ctx, int32_t mode) { return windows_enabled(ctx) && mode != EPSEON_REFINE_LANE; }

// Start of a scan: fix the window set it (and the refinement after it) uses.
void begin_scan(epseon_ctx* ctx, int32_t mode)
{
    ctx->scan_windows = windows_for(ctx, mode);
    ctx->scan_point_updates = ctx->scan_windows ? ctx->scan_pu_band : ctx->scan_pu_full;
}

const BandWindow* scan_window_set(const epseon_ctx* ctx)
{
    return ctx->scan_windows ? ctx->d_win_band : ctx->d_win_full;
}

// Scan of the potentials [k0, k1) in task order on `stream`; the whole batch
// goes through the per-XCD bins.
hipError_t run_scan(epseon_ctx* ctx, hipStream_t stream, size_t k0, size_t k1)
{
    const int32_t t0 = ctx->cum_scan_tasks[k0];
    const int32_t t1 = ctx->cum_scan_tasks[k1];
    const BandWindow* win = scan_window_set(ctx);
    if (ctx->xcd_scan && k0 == 0 && k1 == ctx->cum_scan_tasks.size() - 1) {
        if (ctx->scan_windows)
            return launch_scan(stream, ctx->d_ut, ctx->d_rec, win, ctx->d_xcd_scan_tasks_band, t1 - t0,
                               ctx->d_xcd_bins_band, ctx->xcd_max_bin_band, ctx->waves_per_block,
                               ctx->d_nodes, ctx->d_far, ctx->d_resc);
        return launch_scan(stream, ctx->d_ut, ctx->d_rec, win, ctx->d_xcd_scan_tasks, t1 - t0,
                           ctx->d_xcd_bins, ctx->xcd_max_bin, ctx->waves_per_block, ctx->d_nodes,
                           ctx->d_far, ctx->d_resc);
    }
    return launch_scan(stream, ctx->d_ut, ctx->d_rec, win, ctx->d_scan_tasks + t0, t1 - t0, nullptr, 0,
                       ctx->waves_per_block, ctx->d_nodes, ctx->d_far, ctx->d_resc);
}

// Make `stream` wait for the segment tables of the last load (once), and,
// for a refinement in 64-lane segmented blocks (`mode`, `lanes`), build the
// set the last scan's windows need if the load did not.
hipError_t wait_tables(epseon_ctx* ctx, hipStream_t stream, int32_t mode, int32_t lanes)
{
    if (ctx->tables_pending) {
        ctx->tables_pending = false;
        if (const hipError_t e = hipStreamWaitEvent(stream, ctx->ev_tables, 0); e != hipSuccess)
            return e;
    }
    if (mode != EPSEON_REFINE_SEGMENTED || lanes != 64)
        return hipSuccess;
    return build_tables(ctx, stream, ctx->scan_windows);
}

// Refinement of the levels of potentials [k0, k1) in task order, in the
// windows of the last scan.  Lane mode with one level per lane runs the
// 64-level tasks on `stream`; with two, the paired tasks (potentials with
// more than 64 levels) run on `stream` and the single ones on `stream_b`,
// which may be the same stream.
hipError_t run_refine(epseon_ctx* ctx, hipStream_t stream, hipStream_t stream_b, int32_t mode,
                      int32_t lanes, double rel_tol, int32_t max_iter, size_t k0, size_t k1)
{
    const BandWindow* win = scan_window_set(ctx);
    const int32_t t0 = ctx->cum_level_tasks[k0];
    const int32_t t1 = ctx->cum_level_tasks[k1];
    if (const hipError_t e = launch_brackets(stream, ctx->d_rec, win, ctx->d_level_tasks + t0, t1 - t0,
                                             ctx->d_nodes, ctx->d_far, ctx->d_resc, ctx->d_brackets,
                                             ctx->d_points);
        e != hipSuccess)
        return e;
    if (mode == EPSEON_REFINE_LANE && ctx->levels_per_lane == 1) {
        const int32_t r0 = ctx->cum_refine_tasks[k0];
        const int32_t r1 = ctx->cum_refine_tasks[k1];
        return launch_refine(stream, 1, ctx->d_uv, ctx->d_rec, win, ctx->d_refine_tasks + r0, r1 - r0,
                             ctx->waves_per_block, ctx->d_brackets, ctx->d_points, rel_tol, max_iter,
                             ctx->d_levels, ctx->d_status, ctx->d_iters, ctx->d_steps);
    }
    if (mode == EPSEON_REFINE_LANE) {
        const int32_t a0 = ctx->cum_lane2_tasks[k0];
        const int32_t a1 = ctx->cum_lane2_tasks[k1];
        const int32_t b0 = ctx->cum_lane1_tasks[k0];
        const int32_t b1 = ctx->cum_lane1_tasks[k1];
        hipError_t e = launch_refine(stream, 2, ctx->d_uv, ctx->d_rec, win, ctx->d_lane2_tasks + a0, a1 - a0,
                                     ctx->waves_per_block, ctx->d_brackets, ctx->d_points, rel_tol, max_iter,
                                     ctx->d_levels, ctx->d_status, ctx->d_iters, ctx->d_steps);
        if (e != hipSuccess)
            return e;
        if (stream_b != stream) {
            // The single-level tasks run on stream_b: it waits for the brackets.
            if ((e = hipEventRecord(ctx->ev_join_aux[0], stream)) != hipSuccess)
                return e;
            if ((e = hipStreamWaitEvent(stream_b, ctx->ev_join_aux[0], 0)) != hipSuccess)
                return e;
        }
        return launch_refine(stream_b, 1, ctx->d_uv, ctx->d_rec, win, ctx->d_lane1_tasks + b0, b1 - b0,
                             ctx->waves_per_block, ctx->d_brackets, ctx->d_points, rel_tol, max_iter,
                             ctx->d_levels, ctx->d_status, ctx->d_iters, ctx->d_steps);
    }
    const double* tab = lanes == 64 && ctx->segtab_ok ? (ctx->scan_windows ? ctx->d_segtab_band : ctx->d_segtab_full)
                                                     : nullptr;
    const bool whole = ctx->xcd_scan && k0 == 0 && k1 == ctx->cum_level_tasks.size() - 1;
    const int32_t lpb = segmented_levels(ctx, lanes);
    ctx->last_refine_levels = lpb;
    if (lpb == 2 || lpb == 4) {
        const int32_t g = lpb == 2 ? 0 : 1;
        const int32_t p0 = ctx->cum_group_tasks[g][k0];
        const int32_t p1 = ctx->cum_group_tasks[g][k1];
        if (whole)
            return launch_refine_segmented(stream, lanes, lpb, ctx->d_uv, tab, ctx->d_rec, win,
                                           ctx->d_xcd_group_tasks[g], p1 - p0, ctx->d_xcd_group_bins[g],
                                           ctx->xcd_max_group_bin[g], ctx->d_brackets, ctx->d_points, rel_tol,
                                           max_iter, ctx->d_levels, ctx->d_status, ctx->d_iters, ctx->d_steps);
        return launch_refine_segmented(stream, lanes, lpb, ctx->d_uv, tab, ctx->d_rec, win,
                                       ctx->d_group_tasks[g] + p0, p1 - p0, nullptr, 0, ctx->d_brackets,
                                       ctx->d_points, rel_tol, max_iter, ctx->d_levels, ctx->d_status, ctx->d_iters,
                                       ctx->d_steps);
    }
    if (whole)
        return launch_refine_segmented(stream, lanes, 1, ctx->d_uv, tab, ctx->d_rec, win, ctx->d_xcd_level_tasks,
                                       t1 - t0, ctx->d_xcd_level_bins, ctx->xcd_max_level_bin, ctx->d_brackets,
                                       ctx->d_points, rel_tol, max_iter, ctx->d_levels, ctx->d_status, ctx->d_iters,
                                       ctx->d_steps);
    return launch_refine_segmented(stream, lanes, 1, ctx->d_uv, tab, ctx->d_rec, win, ctx->d_level_tasks + t0,
                                   t1 - t0, nullptr, 0, ctx->d_brackets, ctx->d_points, rel_tol, max_iter,
                                   ctx->d_levels, ctx->d_status, ctx->d_iters, ctx->d_steps);
}

int check_refine_args(double rel_tol, int32_t max_iter)
{
    if (!(rel_tol >= 0.0 && std::isfinite(rel_tol)))
        return fail(EPSEON_ERR_INVALID, "relative tolerance must be finite and >= 0");
    if (max_iter < 1)
        return fail(EPSEON_ERR_INVALID, "max_iter must be >= 1");
    return EPSEON_OK;
}

size_t n_potentials(const epseon_ctx* ctx) { return ctx->cum_scan_tasks.size() - 1; }

// Number of potentials (in task order) of the first pipeline group: the
// longest grids holding up to `fraction` of the scan work.  0 disables the
// pipeline.
size_t split_point(const epseon_ctx* ctx, double fraction)
{
    const size_t np = n_potentials(ctx);
    if (!(fraction > 0.0) || np < 2)
        return 0;
    const double limit = fraction * ctx->cum_scan_work.back();
    size_t k = 0;
    while (k + 1 < np && ctx->cum_scan_work[k + 1] <= limit)
        ++k;
    return k;
}

// The refinement summary of the statistics, on the context's stream after
// every refinement (outside any recorded graph).
hipError_t enqueue_summary(epseon_ctx* ctx)
{
    return launch_level_summary(ctx->stream, ctx->d_rec, static_cast<int32_t>(ctx->rec.size()), ctx->d_status,
                                ctx->d_iters, ctx->d_steps, ctx->d_summary);
}

// Record the single-stream solve (scan, brackets, refinement of the whole
// batch) on the context's stream as a graph and instantiate it.  The
// kernels take only device pointers of this load and by-value arguments;
// none of the recorded work is a memset or a copy.
int capture_solve(epseon_ctx* ctx, int32_t mode, int32_t lanes, double rel_tol, int32_t max_iter)
{
    if (ctx->solve_exec != nullptr) {
        EPSEON_HIP_TRY(hipGraphExecDestroy(ctx->solve_exec));
        ctx->solve_exec = nullptr;
    }
    const size_t np = n_potentials(ctx);
    // No wait on an event recorded outside the capture inside it.
    EPSEON_HIP_TRY(wait_tables(ctx, ctx->stream, mode, lanes));
    EPSEON_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    hipError_t e = run_scan(ctx, ctx->stream, 0, np);
    if (e == hipSuccess)
        e = run_refine(ctx, ctx->stream, ctx->stream, mode, lanes, rel_tol, max_iter, 0, np);
    hipGraph_t graph = nullptr;
    const hipError_t e_end = hipStreamEndCapture(ctx->stream, &graph);
    if (e == hipSuccess)
        e = e_end;
    if (e == hipSuccess)
        e = hipGraphInstantiate(&ctx->solve_exec, graph, nullptr, nullptr, 0);
    if (graph != nullptr)
        (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        ctx->solve_exec = nullptr;
        return hip_fail(e, "capturing the solve graph");
    }
    return EPSEON_OK;
}

}  // namespace

int epseon_scan(epseon_ctx* ctx)
{
    if (const int rc = check_ctx(ctx); rc != EPSEON_OK)
        return rc;
    if (!ctx->loaded)
        return fail(EPSEON_ERR_STATE, "epseon_scan before epseon_load_potentials");
    begin_scan(ctx, choose_refine_mode(ctx));
    EPSEON_HIP_TRY(hipEventRecord(ctx->ev[0], ctx->stream));
    EPSEON_HIP_TRY(run_scan(ctx, ctx->stream, 0, n_potentials(ctx)));
    EPSEON_HIP_TRY(hipEventRecord(ctx->ev[1], ctx->stream));
    ctx->scanned = true;
    ctx->scan_timed = true;
    ctx->refined = false;
    ctx->normalized = false;
    return EPSEON_OK;
}

int epseon_refine(epseon_ctx* ctx, double rel_tol, int32_t max_iter)
{
    if (const int rc = check_ctx(ctx); rc != EPSEON_OK)
        return rc;
    if (!ctx->scanned)
        return fail(EPSEON_ERR_STATE, "epseon_refine before epseon_scan");
    if (const int rc = check_refine_args(rel_tol, max_iter); rc != EPSEON_OK)
        return rc;
    // A windowed scan is refined in segmented mode whatever the mode setting.
    int32_t mode = choose_refine_mode(ctx);
    if (ctx->scan_windows && mode == EPSEON_REFINE_LANE)
        mode = EPSEON_REFINE_SEGMENTED;
    const int32_t lanes = segmented_lanes(ctx);
    ctx->last_refine_mode = mode;
    ctx->last_refine_lanes = mode == EPSEON_REFINE_LANE ? 1 : lanes;
    EPSEON_HIP_TRY(wait_tables(ctx, ctx->stream, mode, lanes));
    EPSEON_HIP_TRY(hipEventRecord(ctx->ev[2], ctx->stream));
    if (mode == EPSEON_REFINE_LANE && ctx->levels_per_lane == 2) {
        // Paired and single lane tasks run concurrently on the two streams.
        EPSEON_HIP_TRY(hipEventRecord(ctx->ev_fork, ctx->stream));
        EPSEON_HIP_TRY(hipStreamWaitEvent(ctx->stream2, ctx->ev_fork, 0));
        EPSEON_HIP_TRY(run_refine(ctx, ctx->stream, ctx->stream2, mode, lanes, rel_tol, max_iter, 0,
                                  n_potentials(ctx)));
        EPSEON_HIP_TRY(hipEventRecord(ctx->ev_join, ctx->stream2));
        EPSEON_HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
    } else {
        // One launch on the context's stream: no cross-stream waits between
        // the scan and the refinement.
        EPSEON_HIP_TRY(run_refine(ctx, ctx->stream, ctx->stream, mode, lanes, rel_tol, max_iter, 0,
                                  n_potentials(ctx)));
    }
    EPSEON_HIP_TRY(hipEventRecord(ctx->ev[3], ctx->stream));
    EPSEON_HIP_TRY(enqueue_summary(ctx));
    ctx->refined = true;
    ctx->refine_timed = true;
    ctx->normalized = false;
    ctx->normalize_timed = ctx->vjp_timed = false;
    ctx->host_levels_valid = false;
    return EPSEON_OK;
}

int epseon_set_refine_mode(epseon_ctx* ctx, int32_t mode)
{
    if (ctx == nullptr)
        return fail(EPSEON_ERR_INVALID, "null context");
    if (mode != EPSEON_REFINE_AUTO && mode != EPSEON_REFINE_LANE && mode != EPSEON_REFINE_SEGMENTED)
        return fail(EPSEON_ERR_INVALID, "unknown refine mode");
    ctx->refine_mode = mode;
    return EPSEON_OK;
}

int epseon_set_solve_split(epseon_ctx* ctx, double fraction)
{
    if (ctx == nullptr)
        return fail(EPSEON_ERR_INVALID, "null context");
    if (!(fraction >= 0.0 && fraction < 1.0))
        return fail(EPSEON_ERR_INVALID, "split fraction must be in [0, 1)");
    ctx->solve_split = fraction;
    return EPSEON_OK;
}

int epseon_solve(epseon_ctx* ctx, double rel_tol, int32_t max_iter)
{
    if (const int rc = check_ctx(ctx); rc != EPSEON_OK)
        return rc;
    if (!ctx->loaded)
        return fail(EPSEON_ERR_STATE, "epseon_solve before epseon_load_potentials");
    if (const int rc = check_refine_args(rel_tol, max_iter); rc != EPSEON_OK)
        return rc;
    const size_t np = n_potentials(ctx);
    const size_t k = split_point(ctx, ctx->solve_split);
    const int32_t mode = choose_refine_mode(ctx);
    const int32_t lanes = segmented_lanes(ctx);
    ctx->last_solve_graph = false;
    if (k == 0 && !(mode == EPSEON_REFINE_LANE && ctx->levels_per_lane == 2)) {
        // Scan and refinement back to back on one stream, timed as a whole
        // (events between the two launches would add a barrier each).
        ctx->last_refine_mode = mode;
        ctx->last_refine_lanes = mode == EPSEON_REFINE_LANE ? 1 : lanes;
        begin_scan(ctx, mode);
        // A graph is recorded when a solve repeats the previous one's key (so
        // a load-then-solve loop, where every solve follows a new load,
        // launches directly) and replayed while the key holds.
        const SolveKey key{ctx->loads,
                           mode,
                           lanes,
                           ctx->waves_per_block,
                           ctx->xcd_scan ? 1 : 0,
                           ctx->scan_windows ? 1 : 0,
                           max_iter,
                           mode == EPSEON_REFINE_SEGMENTED ? segmented_levels(ctx, lanes) : 1,
                           rel_tol};
        const bool have = ctx->solve_exec != nullptr && ctx->graph_key == key;
        const bool repeat = ctx->has_last_key && ctx->last_key == key;
        ctx->last_key = key;
        ctx->has_last_key = true;
        if (ctx->graph && ctx->n_scan_tasks + ctx->n_level_tasks > 0 && (have || repeat)) {
            if (!have) {
                if (const int rc = capture_solve(ctx, mode, lanes, rel_tol, max_iter); rc != EPSEON_OK)
                    return rc;
                ctx->graph_key = key;
            }
            EPSEON_HIP_TRY(hipEventRecord(ctx->ev_solve0, ctx->stream));
            EPSEON_HIP_TRY(hipGraphLaunch(ctx->solve_exec, ctx->stream));
            EPSEON_HIP_TRY(hipEventRecord(ctx->ev_solve1, ctx->stream));
            ctx->last_solve_graph = true;
        } else {
            EPSEON_HIP_TRY(hipEventRecord(ctx->ev_solve0, ctx->stream));
            EPSEON_HIP_TRY(run_scan(ctx, ctx->stream, 0, np));
            EPSEON_HIP_TRY(wait_tables(ctx, ctx->stream, mode, lanes));
            EPSEON_HIP_TRY(run_refine(ctx, ctx->stream, ctx->stream, mode, lanes, rel_tol, max_iter, 0, np));
            EPSEON_HIP_TRY(hipEventRecord(ctx->ev_solve1, ctx->stream));
        }
        EPSEON_HIP_TRY(enqueue_summary(ctx));
        ctx->scanned = ctx->refined = true;
        ctx->scan_timed = ctx->refine_timed = false;
        ctx->normalized = false;
        ctx->normalize_timed = ctx->vjp_timed = false;
        ctx->host_levels_valid = false;
        ctx->last_split_used = 0.0;
        ctx->solve_timed = true;
        return EPSEON_OK;
    }
    if (k == 0) {
        EPSEON_HIP_TRY(hipEventRecord(ctx->ev_solve0, ctx->stream));
        if (const int rc = epseon_scan(ctx); rc != EPSEON_OK)
            return rc;
        if (const int rc = epseon_refine(ctx, rel_tol, max_iter); rc != EPSEON_OK)
            return rc;
        EPSEON_HIP_TRY(hipEventRecord(ctx->ev_solve1, ctx->stream));
        ctx->last_split_used = 0.0;
        ctx->solve_timed = true;
        return EPSEON_OK;
    }
    // Pipelined solve: the longest grids (potentials [0, k) in task order)
    // are scanned and refined on the first stream while the rest are scanned
    // and refined on the second, so the long refinement shots of the first
    // group overlap the scan of the second.  Every potential still runs the
    // same kernels on the same data: results are identical to epseon_scan +
    // epseon_refine.
    ctx->last_refine_mode = mode;
    ctx->last_refine_lanes = mode == EPSEON_REFINE_LANE ? 1 : lanes;
    begin_scan(ctx, mode);
    EPSEON_HIP_TRY(hipEventRecord(ctx->ev_solve0, ctx->stream));
    EPSEON_HIP_TRY(hipEventRecord(ctx->ev_fork, ctx->stream));
    EPSEON_HIP_TRY(hipStreamWaitEvent(ctx->stream2, ctx->ev_fork, 0));
    EPSEON_HIP_TRY(run_scan(ctx, ctx->stream, 0, k));
    EPSEON_HIP_TRY(run_scan(ctx, ctx->stream2, k, np));
    // stream2 built the tables before its scan; the first stream waits.
    EPSEON_HIP_TRY(wait_tables(ctx, ctx->stream, mode, lanes));
    EPSEON_HIP_TRY(run_refine(ctx, ctx->stream, ctx->stream, mode, lanes, rel_tol, max_iter, 0, k));
    EPSEON_HIP_TRY(run_refine(ctx, ctx->stream2, ctx->stream2, mode, lanes, rel_tol, max_iter, k, np));
    EPSEON_HIP_TRY(hipEventRecord(ctx->ev_join, ctx->stream2));
    EPSEON_HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
    EPSEON_HIP_TRY(hipEventRecord(ctx->ev_solve1, ctx->stream));
    EPSEON_HIP_TRY(enqueue_summary(ctx));
    ctx->scanned = ctx->refined = true;
    ctx->scan_timed = ctx->refine_timed = false;
    ctx->normalized = false;
    ctx->normalize_timed = ctx->vjp_timed = false;
    ctx->host_levels_valid = false;
    ctx->solve_timed = true;
    ctx->last_split_used = ctx->cum_scan_work[k] / std::max(ctx->cum_scan_work.back(), 1.0);
    return EPSEON_OK;
}

int epseon_band_windows_get(epseon_ctx* ctx, int32_t p, int32_t* lo, int32_t* hi, int32_t* far_exponent)
{
    if (ctx == nullptr || lo == nullptr || hi == nullptr || far_exponent == nullptr)
        return fail(EPSEON_ERR_INVALID, "null argument");
    if (!ctx->loaded)
        return fail(EPSEON_ERR_STATE, "no batch loaded");
    if (p < 0 || static_cast<size_t>(p) >= ctx->rec.size())
        return fail(EPSEON_ERR_INVALID, "potential index out of range");
    const bool band = ctx->scanned ? ctx->scan_windows : windows_for(ctx, choose_refine_mode(ctx));
    const std::vector<BandWindow>& w = band ? ctx->h_win_band : ctx->h_win_full;
    for (int32_t b = 0; b < kBands; ++b) {
        const BandWindow& x = w[static_cast<size_t>(p) * kBands + static_cast<size_t>(b)];
        lo[b] = x.lo;
        hi[b] = x.hi;
        far_exponent[b] = x.far_exp;
    }
    return EPSEON_OK;
}

int epseon_get_tuning(epseon_ctx* ctx, epseon_tuning* t)
{
    if (ctx == nullptr || t == nullptr)
        return fail(EPSEON_ERR_INVALID, "null argument");
    t->refine_mode = ctx->refine_mode;
    t->levels_per_lane = ctx->levels_per_lane;
    t->scan_waves_per_block = ctx->waves_per_block;
    t->segmented_lanes = ctx->segmented_lanes;
    t->xcd_scan = ctx->xcd_scan ? 1 : 0;
    t->windows = ctx->windows;
    t->graph = ctx->graph;
    t->segmented_levels = ctx->segmented_levels;
    t->solve_split = ctx->solve_split;
    return EPSEON_OK;
}

int epseon_set_tuning(epseon_ctx* ctx, const epseon_tuning* t)
{
    if (ctx == nullptr || t == nullptr)
        return fail(EPSEON_ERR_INVALID, "null argument");
    if (t->refine_mode != EPSEON_REFINE_AUTO && t->refine_mode != EPSEON_REFINE_LANE &&
        t->refine_mode != EPSEON_REFINE_SEGMENTED)
        return fail(EPSEON_ERR_INVALID, "unknown refine mode");
    if (t->levels_per_lane != 1 && t->levels_per_lane != 2)
        return fail(EPSEON_ERR_INVALID, "levels_per_lane must be 1 or 2");
    if (t->scan_waves_per_block < 1 || t->scan_waves_per_block > 4)
        return fail(EPSEON_ERR_INVALID, "scan_waves_per_block must be in [1, 4]");
    if (t->segmented_lanes != 0 && t->segmented_lanes != 64 && t->segmented_lanes != 128 &&
        t->segmented_lanes != 256)
        return fail(EPSEON_ERR_INVALID, "segmented_lanes must be 0, 64, 128 or 256");
    if (!(t->solve_split >= 0.0 && t->solve_split < 1.0))
        return fail(EPSEON_ERR_INVALID, "solve_split must be in [0, 1)");
    if (t->xcd_scan != 0 && t->xcd_scan != 1)
        return fail(EPSEON_ERR_INVALID, "xcd_scan must be 0 or 1");
    if (t->windows < 0 || t->windows > 2)
        return fail(EPSEON_ERR_INVALID, "windows must be 0, 1 or 2");
    if (t->graph != 0 && t->graph != 1)
        return fail(EPSEON_ERR_INVALID, "graph must be 0 or 1");
    if (t->segmented_levels != 0 && t->segmented_levels != 1 && t->segmented_levels != 2 &&
        t->segmented_levels != 4)
        return fail(EPSEON_ERR_INVALID, "segmented_levels must be 0, 1, 2 or 4");
    ctx->refine_mode = t->refine_mode;
    ctx->levels_per_lane = t->levels_per_lane;
    ctx->waves_per_block = t->scan_waves_per_block;
    ctx->windows = t->windows;
    ctx->segmented_lanes = t->segmented_lanes;
    ctx->xcd_scan = t->xcd_scan != 0;
    ctx->graph = t->graph;
    ctx->segmented_levels = t->segmented_levels;
    ctx->solve_split = t->solve_split;
    return EPSEON_OK;
}

namespace {

// Device time of one solve under the context's current tuning (best of
// `reps` runs after one untimed run).
int time_solve(epseon_ctx* ctx, double rel_tol, int32_t max_iter, int reps, double* ms_out)
{
    double best = 1e300;
    for (int r = 0; r <= reps; ++r) {
        if (const int rc = epseon_solve(ctx, rel_tol, max_iter); rc != EPSEON_OK)
            return rc;
        EPSEON_HIP_TRY(hipEventSynchronize(ctx->ev_solve1));
        float ms = 0.0f;
        EPSEON_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev_solve0, ctx->ev_solve1));
        if (r > 0)
            best = std::min(best, static_cast<double>(ms));
    }
    *ms_out = best;
    return EPSEON_OK;
}

}  // namespace

int epseon_autotune(epseon_ctx* ctx, double rel_tol, int32_t max_iter, int32_t flags,
                    epseon_tuning* best_out)
{
    if (const int rc = check_ctx(ctx); rc != EPSEON_OK)
        return rc;
    if (!ctx->loaded)
        return fail(EPSEON_ERR_STATE, "epseon_autotune before epseon_load_potentials");
    if (const int rc = check_refine_args(rel_tol, max_iter); rc != EPSEON_OK)
        return rc;
    epseon_tuning best;
    (void)epseon_get_tuning(ctx, &best);
    if (best.refine_mode == EPSEON_REFINE_AUTO)
        best.refine_mode = choose_refine_mode(ctx);
    if (best.windows == 2)
        best.windows = windows_enabled(ctx) ? 1 : 0;
    if (best.segmented_levels == 0)
        best.segmented_levels = segmented_levels(ctx, segmented_lanes(ctx));
    double best_ms = 0.0;
    if (const int rc = epseon_set_tuning(ctx, &best); rc != EPSEON_OK)
        return rc;
    if (const int rc = time_solve(ctx, rel_tol, max_iter, 2, &best_ms); rc != EPSEON_OK)
        return rc;
    // Coordinate search, one knob at a time, keeping a candidate only when it
    // is at least 2% faster.
    auto trial = [&](epseon_tuning cand) -> int {
        if (const int rc = epseon_set_tuning(ctx, &cand); rc != EPSEON_OK)
            return rc;
        double ms = 0.0;
        if (const int rc = time_solve(ctx, rel_tol, max_iter, 2, &ms); rc != EPSEON_OK)
            return rc;
        if (ms < 0.98 * best_ms) {
            best = cand;
            best_ms = ms;
        }
        return EPSEON_OK;
    };
    for (int32_t w : {2, 4}) {
        epseon_tuning c = best;
        c.scan_waves_per_block = w;
        if (const int rc = trial(c); rc != EPSEON_OK)
            return rc;
    }
    if (!(flags & EPSEON_TUNE_KEEP_RESULTS) && best.refine_mode != EPSEON_REFINE_LANE) {
        epseon_tuning c = best;
        c.windows = 1 - best.windows;
        if (const int rc = trial(c); rc != EPSEON_OK)
            return rc;
    }
    {
        epseon_tuning c = best;
        c.xcd_scan = 1 - best.xcd_scan;
        if (const int rc = trial(c); rc != EPSEON_OK)
            return rc;
    }
    if (!(flags & EPSEON_TUNE_KEEP_RESULTS)) {
        epseon_tuning c = best;
        c.refine_mode = best.refine_mode == EPSEON_REFINE_LANE ? EPSEON_REFINE_SEGMENTED : EPSEON_REFINE_LANE;
        if (const int rc = trial(c); rc != EPSEON_OK)
            return rc;
    }
    if (best.refine_mode == EPSEON_REFINE_LANE) {
        epseon_tuning c = best;
        c.levels_per_lane = 3 - best.levels_per_lane;
        if (const int rc = trial(c); rc != EPSEON_OK)
            return rc;
    } else {
        if (!(flags & EPSEON_TUNE_KEEP_RESULTS))
            for (int32_t lanes : {64, 128, 256}) {
                epseon_tuning c = best;
                c.segmented_lanes = lanes;
                if (const int rc = trial(c); rc != EPSEON_OK)
                    return rc;
            }
        // Level groups leave every level's arithmetic unchanged.
        if ((best.segmented_lanes == 0 ? segmented_lanes(ctx) : best.segmented_lanes) == 64) {
            const int32_t current = best.segmented_levels;
            for (int32_t levels : {1, 2, 4}) {
                if (levels == current)
                    continue;
                epseon_tuning c = best;
                c.segmented_levels = levels;
                if (const int rc = trial(c); rc != EPSEON_OK)
                    return rc;
            }
        }
    }
    for (double split : {0.15, 0.3, 0.5}) {
        epseon_tuning c = best;
        c.solve_split = split;
        if (const int rc = trial(c); rc != EPSEON_OK)
            return rc;
    }
    {
        epseon_tuning c = best;
        c.graph = 1 - best.graph;
        if (const int rc = trial(c); rc != EPSEON_OK)
            return rc;
    }
    if (const int rc = epseon_set_tuning(ctx, &best); rc != EPSEON_OK)
        return rc;
    if (best_out)
        *best_out = best;
    return EPSEON_OK;
}

int epseon_ctx_synchronize(epseon_ctx* ctx)
{
    if (const int rc = check_ctx(ctx); rc != EPSEON_OK)
        return rc;
    EPSEON_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return EPSEON_OK;
}

namespace {

// Read the level block back (one copy into pinned memory, then a stream
// synchronisation); h_levels, h_status and h_iters point into it, and the
// summary is copied out of it.
int fetch_levels(epseon_ctx* ctx)
{
    const size_t n = static_cast<size_t>(ctx->total_levels);
    ctx->h_levels = nullptr;
    ctx->h_status = ctx->h_iters = nullptr;
    if (n > 0) {
        const size_t bytes = ctx->level_block_bytes;
        if (ctx->h_pinned_bytes < bytes) {
            if (ctx->h_pinned != nullptr) {
                EPSEON_HIP_TRY(hipHostFree(ctx->h_pinned));
                ctx->h_pinned = nullptr;
                ctx->h_pinned_bytes = 0;
            }
            EPSEON_HIP_TRY(hipHostMalloc(&ctx->h_pinned, bytes, hipHostMallocDefault));
            ctx->h_pinned_bytes = bytes;
        }
        EPSEON_HIP_TRY(hipMemcpyAsync(ctx->h_pinned, ctx->d_level_block, ctx->level_block_bytes,
                                      hipMemcpyDeviceToHost, ctx->stream));
    }
    EPSEON_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (n > 0) {
        const char* h = static_cast<const char*>(ctx->h_pinned);
        const auto at = [&](const void* d) {
            return h + (static_cast<const char*>(d) - static_cast<const char*>(ctx->d_level_block));
        };
        ctx->h_levels = reinterpret_cast<const double*>(h);
        ctx->h_status = reinterpret_cast<const int32_t*>(at(ctx->d_status));
        ctx->h_iters = reinterpret_cast<const int32_t*>(at(ctx->d_iters));
        std::memcpy(ctx->h_summary, at(ctx->d_summary), sizeof ctx->h_summary);
    }
    ctx->host_levels_valid = true;
    return EPSEON_OK;
}

}  // namespace

namespace {

int ensure_stage(epseon_ctx* ctx, size_t elems)
{
    if (ctx->stage_elems >= elems && ctx->d_stage != nullptr)
        return EPSEON_OK;
    EPSEON_HIP_TRY(arena_array(ctx->arena, std::max<size_t>(elems, 1), &ctx->d_stage));
    ctx->stage_elems = elems;
    return EPSEON_OK;
}

int check_mem(int32_t mem)
{
    if (mem != EPSEON_MEM_HOST && mem != EPSEON_MEM_DEVICE)
        return fail(EPSEON_ERR_INVALID, "memory kind must be EPSEON_MEM_HOST or EPSEON_MEM_DEVICE");
    return EPSEON_OK;
}

}  // namespace

int epseon_normalize(epseon_ctx* ctx)
{
    if (const int rc = check_ctx(ctx); rc != EPSEON_OK)
        return rc;
    if (!ctx->refined)
        return fail(EPSEON_ERR_STATE, "epseon_normalize before epseon_refine");
    if (ctx->d_ckpt == nullptr) {
        const size_t slots = static_cast<size_t>(std::max<int64_t>(ctx->ckpt_total, 1));
        EPSEON_HIP_TRY(arena_array(ctx->arena, slots, &ctx->d_ckpt));
        EPSEON_HIP_TRY(arena_array(ctx->arena, slots, &ctx->d_psi_mats));
    }
    EPSEON_HIP_TRY(hipEventRecord(ctx->ev[4], ctx->stream));
    EPSEON_HIP_TRY(launch_normalize(ctx->stream, ctx->d_uv, ctx->d_uvr, ctx->d_rec, ctx->d_psi_tasks,
                                    ctx->cum_psi_tasks.back(), ctx->d_refine_tasks, ctx->n_refine_tasks,
                                    ctx->d_levels, ctx->d_status, ctx->d_psi_mats, ctx->d_ckpt,
                                    ctx->d_norms));
    EPSEON_HIP_TRY(hipEventRecord(ctx->ev[5], ctx->stream));
    ctx->normalized = true;
    ctx->normalize_timed = true;
    return EPSEON_OK;
}

namespace {

// Eigenfunctions of potentials in task-order positions [k0, k1) into `psi`
// (host or device), whose element 0 is eigenfunction value `base`.
int write_wavefunctions(epseon_ctx* ctx, size_t k0, size_t k1, int64_t base, size_t elems,
                        double* psi, int32_t mem)
{
    if (!ctx->normalized)
        if (const int rc = epseon_normalize(ctx); rc != EPSEON_OK)
            return rc;
    if (elems == 0)
        return EPSEON_OK;
    double* dst = psi;
    if (mem == EPSEON_MEM_HOST) {
        if (const int rc = ensure_stage(ctx, elems); rc != EPSEON_OK)
            return rc;
        dst = ctx->d_stage;
    }
    const int32_t t0 = ctx->cum_psi_tasks[k0];
    const int32_t t1 = ctx->cum_psi_tasks[k1];
    EPSEON_HIP_TRY(launch_eigenfunctions(ctx->stream, ctx->d_uv, ctx->d_uvr, ctx->d_rec,
                                         ctx->d_psi_tasks + t0, t1 - t0, ctx->d_levels, ctx->d_status,
                                         ctx->d_norms, ctx->d_ckpt, base, dst));
    if (mem == EPSEON_MEM_HOST)
        EPSEON_HIP_TRY(hipMemcpyAsync(psi, dst, elems * sizeof(double), hipMemcpyDeviceToHost,
                                      ctx->stream));
    EPSEON_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return EPSEON_OK;
}

}  // namespace

int epseon_get_wavefunctions(epseon_ctx* ctx, int32_t potential, double* psi, int32_t mem)
{
    if (const int rc = check_ctx(ctx); rc != EPSEON_OK)
        return rc;
    if (const int rc = check_mem(mem); rc != EPSEON_OK)
        return rc;
    if (!ctx->refined)
        return fail(EPSEON_ERR_STATE, "epseon_get_wavefunctions before epseon_refine");
    if (potential < 0 || potential >= static_cast<int32_t>(ctx->rec.size()))
        return fail(EPSEON_ERR_INVALID, "potential index out of range");
    if (psi == nullptr)
        return fail(EPSEON_ERR_INVALID, "null output pointer");
    const PotentialRecord& r = ctx->rec[potential];
    const size_t elems = static_cast<size_t>(r.level_count) * static_cast<size_t>(r.full_points);
    const size_t k = static_cast<size_t>(ctx->order_pos[static_cast<size_t>(potential)]);
    return write_wavefunctions(ctx, k, k + 1, r.psi_offset, elems, psi, mem);
}

int epseon_get_all_wavefunctions(epseon_ctx* ctx, double* psi, int32_t mem)
{
    if (const int rc = check_ctx(ctx); rc != EPSEON_OK)
        return rc;
    if (const int rc = check_mem(mem); rc != EPSEON_OK)
        return rc;
    if (!ctx->refined)
        return fail(EPSEON_ERR_STATE, "epseon_get_all_wavefunctions before epseon_refine");
    if (psi == nullptr && ctx->psi_total > 0)
        return fail(EPSEON_ERR_INVALID, "null output pointer");
    return write_wavefunctions(ctx, 0, n_potentials(ctx), 0, static_cast<size_t>(ctx->psi_total), psi,
                               mem);
}

int epseon_level_vjp(epseon_ctx* ctx, const double* grad_levels, double* grad_v, int32_t mem)
{
    if (const int rc = check_ctx(ctx); rc != EPSEON_OK)
        return rc;
    if (const int rc = check_mem(mem); rc != EPSEON_OK)
        return rc;
    if (!ctx->refined)
        return fail(EPSEON_ERR_STATE, "epseon_level_vjp before epseon_refine");
    if (grad_v == nullptr || (grad_levels == nullptr && ctx->total_levels > 0))
        return fail(EPSEON_ERR_INVALID, "null gradient pointer");
    if (!ctx->normalized)
        if (const int rc = epseon_normalize(ctx); rc != EPSEON_OK)
            return rc;
    const double* g = grad_levels;
    double* out = grad_v;
    if (mem == EPSEON_MEM_HOST) {
        if (ctx->total_levels > 0)
            EPSEON_HIP_TRY(hipMemcpyAsync(ctx->d_grad_levels, grad_levels,
                                          static_cast<size_t>(ctx->total_levels) * sizeof(double),
                                          hipMemcpyHostToDevice, ctx->stream));
        g = ctx->d_grad_levels;
        if (const int rc = ensure_stage(ctx, static_cast<size_t>(ctx->user_span)); rc != EPSEON_OK)
            return rc;
        out = ctx->d_stage;
    }
    EPSEON_HIP_TRY(hipEventRecord(ctx->ev[6], ctx->stream));
    // The four block-width classes touch disjoint potentials: one stream each.
    const hipStream_t streams[4] = {ctx->stream, ctx->stream2, ctx->stream3, ctx->stream4};
    EPSEON_HIP_TRY(hipEventRecord(ctx->ev_fork, ctx->stream));
    for (int w = 1; w < 4; ++w)
        EPSEON_HIP_TRY(hipStreamWaitEvent(streams[w], ctx->ev_fork, 0));
    for (int w = 0; w < 4; ++w)
        EPSEON_HIP_TRY(launch_level_vjp(streams[w], w + 1, ctx->d_uv, ctx->d_uvr, ctx->d_rec,
                                        ctx->d_vjp_lists[w], ctx->n_vjp[w], ctx->d_levels,
                                        ctx->d_status, ctx->d_norms, ctx->d_ckpt, g, out));
    for (int w = 1; w < 4; ++w) {
        EPSEON_HIP_TRY(hipEventRecord(ctx->ev_join_aux[w - 1], streams[w]));
        EPSEON_HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->ev_join_aux[w - 1], 0));
    }
    EPSEON_HIP_TRY(hipEventRecord(ctx->ev[7], ctx->stream));
    ctx->vjp_timed = true;
    if (mem == EPSEON_MEM_HOST) {
        std::vector<double> tmp(static_cast<size_t>(ctx->user_span));
        if (!tmp.empty())
            EPSEON_HIP_TRY(hipMemcpyAsync(tmp.data(), out, tmp.size() * sizeof(double),
                                          hipMemcpyDeviceToHost, ctx->stream));
        EPSEON_HIP_TRY(hipStreamSynchronize(ctx->stream));
        for (const PotentialRecord& r : ctx->rec)
            std::memcpy(grad_v + r.user_offset, tmp.data() + r.user_offset,
                        static_cast<size_t>(r.full_points) * sizeof(double));
    } else {
        EPSEON_HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    return EPSEON_OK;
}

int epseon_level_vjp_noalloc(epseon_ctx* ctx, const double* grad_levels_device, double* grad_v_device)
{
    if (const int rc = check_ctx(ctx); rc != EPSEON_OK)
        return rc;
    if (!ctx->refined)
        return fail(EPSEON_ERR_STATE, "epseon_level_vjp_noalloc before epseon_refine");
    if (grad_v_device == nullptr || (grad_levels_device == nullptr && ctx->total_levels > 0))
        return fail(EPSEON_ERR_INVALID, "null gradient pointer");
    // Segment maps and checkpoints of the normalisation pass, both carved
    // from the device's scratch array.
    const size_t slots = static_cast<size_t>(std::max<int64_t>(ctx->ckpt_total, 1));
    const size_t ckpt_bytes = (slots * sizeof(PsiCheckpoint) + 255) & ~size_t(255);
    const size_t need = ckpt_bytes + slots * sizeof(PsiSegMatrix);
    if (need > kNoAllocScratchBytes) {
        char msg[256];
        std::snprintf(msg, sizeof msg,
                      "epseon_level_vjp_noalloc: the batch needs %zu bytes of scratch (%zu eigenfunction-shot "
                      "segments), above the %zu-byte maximum",
                      need, slots, kNoAllocScratchBytes);
        return fail(EPSEON_ERR_INVALID, msg);
    }
    void* base = scratch_of(ctx->device);
    if (base == nullptr)
        return fail(EPSEON_ERR_DEVICE, "epseon_level_vjp_noalloc: no scratch on this device");
    PsiCheckpoint* ckpt = static_cast<PsiCheckpoint*>(base);
    PsiSegMatrix* mats = reinterpret_cast<PsiSegMatrix*>(static_cast<char*>(base) + ckpt_bytes);
    // The scratch holds whatever the last call on this device left, so the
    // normalisation always runs; every slot a kernel reads is written first.
    // The context's own checkpoints (epseon_normalize) are left as they are.
    EPSEON_HIP_TRY(hipEventRecord(ctx->ev[6], ctx->stream));
    EPSEON_HIP_TRY(launch_normalize(ctx->stream, ctx->d_uv, ctx->d_uvr, ctx->d_rec, ctx->d_psi_tasks,
                                    ctx->cum_psi_tasks.back(), ctx->d_refine_tasks, ctx->n_refine_tasks,
                                    ctx->d_levels, ctx->d_status, mats, ckpt, ctx->d_norms));
    const hipStream_t streams[4] = {ctx->stream, ctx->stream2, ctx->stream3, ctx->stream4};
    EPSEON_HIP_TRY(hipEventRecord(ctx->ev_fork, ctx->stream));
    for (int w = 1; w < 4; ++w)
        EPSEON_HIP_TRY(hipStreamWaitEvent(streams[w], ctx->ev_fork, 0));
    for (int w = 0; w < 4; ++w)
        EPSEON_HIP_TRY(launch_level_vjp(streams[w], w + 1, ctx->d_uv, ctx->d_uvr, ctx->d_rec,
                                        ctx->d_vjp_lists[w], ctx->n_vjp[w], ctx->d_levels, ctx->d_status,
                                        ctx->d_norms, ckpt, grad_levels_device, grad_v_device));
    for (int w = 1; w < 4; ++w) {
        EPSEON_HIP_TRY(hipEventRecord(ctx->ev_join_aux[w - 1], streams[w]));
        EPSEON_HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->ev_join_aux[w - 1], 0));
    }
    EPSEON_HIP_TRY(hipEventRecord(ctx->ev[7], ctx->stream));
    ctx->vjp_timed = true;
    EPSEON_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return EPSEON_OK;
}

int epseon_get_levels(epseon_ctx* ctx, double* levels, int32_t* status, int32_t* iterations)
{
    if (const int rc = check_ctx(ctx); rc != EPSEON_OK)
        return rc;
    if (!ctx->refined)
        return fail(EPSEON_ERR_STATE, "epseon_get_levels before epseon_refine");
    const size_t n = static_cast<size_t>(ctx->total_levels);
    if (const int rc = fetch_levels(ctx); rc != EPSEON_OK)
        return rc;
    if (n > 0 && levels)
        std::memcpy(levels, ctx->h_levels, n * sizeof(double));
    if (n > 0 && status)
        std::memcpy(status, ctx->h_status, n * sizeof(int32_t));
    if (n > 0 && iterations)
        std::memcpy(iterations, ctx->h_iters, n * sizeof(int32_t));
    return EPSEON_OK;
}

int epseon_get_scan(epseon_ctx* ctx, int32_t* nodes, double* far_end, int32_t* rescales)
{
    if (const int rc = check_ctx(ctx); rc != EPSEON_OK)
        return rc;
    if (!ctx->scanned)
        return fail(EPSEON_ERR_STATE, "epseon_get_scan before epseon_scan");
    const size_t n = static_cast<size_t>(ctx->total_energies);
    if (n > 0) {
        if (nodes)
            EPSEON_HIP_TRY(hipMemcpyAsync(nodes, ctx->d_nodes, n * sizeof(int32_t),
                                          hipMemcpyDeviceToHost, ctx->stream));
        if (far_end)
            EPSEON_HIP_TRY(hipMemcpyAsync(far_end, ctx->d_far, n * sizeof(double),
                                          hipMemcpyDeviceToHost, ctx->stream));
        if (rescales)
            EPSEON_HIP_TRY(hipMemcpyAsync(rescales, ctx->d_resc, n * sizeof(int32_t),
                                          hipMemcpyDeviceToHost, ctx->stream));
    }
    EPSEON_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return EPSEON_OK;
}

int epseon_get_brackets(epseon_ctx* ctx, epseon_bracket* out)
{
    if (const int rc = check_ctx(ctx); rc != EPSEON_OK)
        return rc;
    if (!ctx->refined)
        return fail(EPSEON_ERR_STATE, "epseon_get_brackets before epseon_refine");
    if (out == nullptr && ctx->total_levels > 0)
        return fail(EPSEON_ERR_INVALID, "null output pointer");
    std::vector<Bracket> br(static_cast<size_t>(ctx->total_levels));
    if (!br.empty())
        EPSEON_HIP_TRY(hipMemcpyAsync(br.data(), ctx->d_brackets, br.size() * sizeof(Bracket),
                                      hipMemcpyDeviceToHost, ctx->stream));
    EPSEON_HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (const PotentialRecord& r : ctx->rec) {
        for (int32_t l = 0; l < r.level_count; ++l) {
            const Bracket& b = br[static_cast<size_t>(r.level_offset + l)];
            epseon_bracket& o = out[r.level_offset + l];
            o = epseon_bracket{};
            if (r.flags != 0 || !b.found) {
                o.e_lo = o.e_hi = std::nan("");
                continue;
            }
            o.e_lo = b.el;
            o.e_hi = b.eh;
            o.nodes_lo = b.cl;
            o.nodes_hi = b.ch;
            o.lo_valid = b.lo_ok;
            o.found = 1;
            o.band = b.band;
        }
    }
    return EPSEON_OK;
}

int epseon_get_sizes(epseon_ctx* ctx, int32_t* potentials, int64_t* points, int64_t* energies,
                     int64_t* levels)
{
    if (ctx == nullptr)
        return fail(EPSEON_ERR_INVALID, "null context");
    if (potentials)
        *potentials = static_cast<int32_t>(ctx->rec.size());
    if (points) {
        int64_t s = 0;
        for (const epseon_potential_desc& d : ctx->desc)
            s += d.n_points;
        *points = s;
    }
    if (energies)
        *energies = ctx->total_energies;
    if (levels)
        *levels = ctx->total_levels;
    return EPSEON_OK;
}

int epseon_get_eigenfunction_size(epseon_ctx* ctx, int64_t* elements)
{
    if (ctx == nullptr || elements == nullptr)
        return fail(EPSEON_ERR_INVALID, "null argument");
    *elements = ctx->psi_total;
    return EPSEON_OK;
}

int epseon_get_stats(epseon_ctx* ctx, epseon_stats* stats)
{
    if (const int rc = check_ctx(ctx); rc != EPSEON_OK)
        return rc;
    if (stats == nullptr)
        return fail(EPSEON_ERR_INVALID, "null stats pointer");
    std::memset(stats, 0, sizeof *stats);
    stats->scan_windows = ctx->scanned && ctx->scan_windows ? 1 : 0;
    EPSEON_HIP_TRY(hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    if (ctx->scan_timed) {
        EPSEON_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        stats->scan_ms = ms;
        stats->scan_point_updates = ctx->scan_point_updates;
        stats->scan_launches = ctx->n_scan_tasks > 0 ? 1 : 0;
    }
    if (ctx->solve_timed) {
        EPSEON_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev_solve0, ctx->ev_solve1));
        stats->solve_ms = ms;
        stats->solve_split = ctx->last_split_used;
        stats->scan_point_updates = ctx->scan_point_updates;
        stats->solve_graph = ctx->last_solve_graph ? 1 : 0;
    }
    if (ctx->refined) {
        if (ctx->refine_timed) {
            EPSEON_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
            stats->refine_ms = ms;
        }
        // The summary (three integer sums over the levels, computed on the
        // device after the refinement) comes back with the levels.
        if (!ctx->host_levels_valid)
            if (const int rc = fetch_levels(ctx); rc != EPSEON_OK)
                return rc;
        stats->refine_point_updates = static_cast<double>(ctx->h_summary[0]);
        stats->levels_converged = ctx->h_summary[1];
        stats->refine_mode = ctx->last_refine_mode;
        stats->refine_lanes = ctx->last_refine_lanes;
        stats->refine_levels_per_block = ctx->last_refine_mode == EPSEON_REFINE_SEGMENTED ? ctx->last_refine_levels : 1;
        // Eigenfunction passes shoot every converged level once over its
        // trimmed grid (n - 2 steps split between the two inward shots).
        const double shot_pu = static_cast<double>(ctx->h_summary[2]);
        if (ctx->normalize_timed) {
            EPSEON_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]));
            stats->normalize_ms = ms;
            stats->normalize_point_updates = shot_pu;
        }
        if (ctx->vjp_timed) {
            EPSEON_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[6], ctx->ev[7]));
            stats->vjp_ms = ms;
            stats->vjp_point_updates = shot_pu;
        }
    }
    return EPSEON_OK;
}
