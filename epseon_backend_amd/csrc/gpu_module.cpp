// gpu_module.cpp -- the Python module _libepseon_gpu on top of libepseon_hip.
//
// Same classes, method names, keyword names and error behaviour as the
// reference module (cpp/gpu/source/epseon/gpu/python/api.cpp:114-494 and the
// wrappers in cpp/gpu/include/epseon/gpu/python/api.hpp:26-309):
//   EpseonComputeContext.create / get_vulkan_version / get_physical_device_info
//     / get_device_interface
//   ComputeDeviceInterface.get_task_configurator(precision) / submit_task(cfg)
//   TaskConfiguratorFloat32|Float64.set_hardware_config / set_morse_potential
//     / set_vibwa_algorithm / is_configured   (each setter returns self)
//   MorsePotentialConfig(dissociation_energy, equilibrium_bond_distance,
//     well_width, min_r, max_r, point_count)
//   TaskHandleFloat32|Float64.get_status_message / is_done / wait
// The Vulkan device-property structs become the same Python attribute names
// filled from hipDeviceProp_t.  What the reference lacks and this module adds
// (results, cancel, per-potential options, more potential sources) is
// additive: no existing name changes meaning.
//
// The work itself runs in the C ABI (include/epseon_hip.h); this file is host
// C++ only and never touches device memory.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <exception>
#include <memory>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <stop_token>
#include <string>
#include <thread>
#include <tuple>
#include <utility>
#include <variant>
#include <vector>

#include "epseon_hip.h"

namespace py = pybind11;

namespace epseon::gpu {

// ------------------------------------------------------------ precision

enum class Precision { Float32, Float64 };

// Case-insensitive "float32"/"float64" (reference: enums.cpp:34-53); any
// other string raises ValueError with the reference's message (enums.cpp:13-14).
Precision parse_precision(const std::string& text)
{
    std::string lower(text);
    std::transform(lower.begin(), lower.end(), lower.begin(),
                   [](unsigned char c) { return static_cast<char>(std::tolower(c)); });
    if (lower == "float32")
        return Precision::Float32;
    if (lower == "float64")
        return Precision::Float64;
    throw py::value_error("Invalid PrecisionType literal in string: \"" + lower + "\"");
}

[[noreturn]] void throw_capi(int rc, const char* what)
{
    std::string msg = std::string(what) + " failed (" + std::to_string(rc) + "): " + epseon_last_error();
    throw std::runtime_error(msg);
}

void check_capi(int rc, const char* what)
{
    if (rc != EPSEON_OK)
        throw_capi(rc, what);
}

// Version integers of the HIP runtime/driver (major*1e7 + minor*1e5 + patch)
// rendered with four components like the reference's Vulkan version strings
// (common.cpp:9-14).
std::string version_string(int v)
{
    const int major = v / 10000000;
    const int minor = (v / 100000) % 100;
    const int patch = v % 100000;
    return std::to_string(major) + "." + std::to_string(minor) + "." + std::to_string(patch) + ".0";
}

// ------------------------------------------------------ device properties

struct PhysicalDeviceSparseProperties {};

struct PhysicalDeviceLimits {
    uint32_t max_compute_shared_memory_size = 0;
    std::tuple<uint32_t, uint32_t, uint32_t> max_compute_work_group_count{0, 0, 0};
    uint32_t max_compute_work_group_invocations = 0;
    uint32_t max_compute_work_group_size = 0;
};

struct PhysicalDeviceProperties {
    std::string api_version;
    std::string driver_version;
    uint32_t vendor_id = 0;
    uint32_t device_id = 0;
    std::string device_type;
    std::string device_name;
    std::vector<uint8_t> pipeline_cache_uuid;
    PhysicalDeviceLimits limits;
    PhysicalDeviceSparseProperties sparse_properties;
    // Additive: the HIP view of the same device.
    std::string arch;
    int32_t compute_units = 0;
    int32_t pci_device_id = 0;
};

struct MemoryHeap {
    uint64_t size = 0;
    std::vector<std::string> flags;
};

struct MemoryType {
    uint32_t heap_index = 0;
    std::vector<std::string> flags;
};

struct PhysicalDeviceMemoryProperties {
    std::vector<MemoryHeap> memory_heaps;
    std::vector<MemoryType> memory_types;
};

struct PhysicalDeviceInfo {
    PhysicalDeviceProperties device_properties;
    PhysicalDeviceMemoryProperties memory_properties;
};

PhysicalDeviceInfo query_device(int32_t ordinal)
{
    epseon_device_info d{};
    check_capi(epseon_device_info_get(ordinal, &d), "epseon_device_info_get");
    PhysicalDeviceInfo info;
    PhysicalDeviceProperties& p = info.device_properties;
    p.api_version = version_string(d.runtime_version);
    p.driver_version = version_string(d.driver_version);
    p.vendor_id = static_cast<uint32_t>(d.vendor_id);
    // device_id is the HIP ordinal: get_device_interface(device_id) selects by
    // it (the reference matched the Vulkan PCI product id, which is the same
    // for every card of a multi-GPU node).
    p.device_id = static_cast<uint32_t>(ordinal);
    p.device_type = "DISCRETE_GPU";
    p.device_name = d.name;
    p.pipeline_cache_uuid.assign(d.uuid, d.uuid + 16);
    p.limits.max_compute_shared_memory_size = static_cast<uint32_t>(d.shared_mem_per_block);
    p.limits.max_compute_work_group_count = {static_cast<uint32_t>(d.max_grid_dim[0]),
                                             static_cast<uint32_t>(d.max_grid_dim[1]),
                                             static_cast<uint32_t>(d.max_grid_dim[2])};
    p.limits.max_compute_work_group_invocations = static_cast<uint32_t>(d.max_threads_per_block);
    p.limits.max_compute_work_group_size = static_cast<uint32_t>(d.max_block_dim[0]);
    p.arch = d.arch;
    p.compute_units = d.compute_units;
    p.pci_device_id = d.pci_device_id;
    info.memory_properties.memory_heaps.push_back(
        MemoryHeap{static_cast<uint64_t>(d.total_global_mem), {"DEVICE_LOCAL"}});
    info.memory_properties.memory_types.push_back(MemoryType{0, {"DEVICE_LOCAL"}});
    return info;
}

int32_t device_count()
{
    int32_t n = 0;
    check_capi(epseon_device_count(&n), "epseon_device_count");
    return n;
}

// --------------------------------------------------------- configuration

// Reference: HardwareConfig (task_configurator/hardware_config.hpp:11-76).
// potential_buffer_size caps the grid points of one potential, group_size
// the potentials solved per launch, allocation_block_size the device arena
// block in bytes.
struct HardwareConfig {
    uint32_t potential_buffer_size = 0;
    uint32_t group_size = 0;
    uint32_t allocation_block_size = 0;
};

// Reference: MorsePotentialConfig<double> (potential_source.hpp:102-183).
struct MorsePotentialConfig {
    double dissociation_energy = 0;
    double equilibrium_bond_distance = 0;
    double well_width = 0;
    double min_r = 0;
    double max_r = 0;
    uint32_t point_count = 0;
};

// Additive: Morse plus a Gaussian barrier B exp(-((r - r_b)/w)^2).
struct DoubleWellPotentialConfig {
    MorsePotentialConfig morse;
    double barrier_height = 0;
    double barrier_position = 0;
    double barrier_width = 1;
};

// Additive: a V table on the uniform grid [min_r, max_r].
struct TabulatedPotentialConfig {
    std::vector<double> values;
    double min_r = 0;
    double max_r = 0;
    double asymptote = 0;
};

// One potential after the configurator's precision has been applied.
struct PotentialSpec {
    enum class Kind { Morse, DoubleWell, Tabulated } kind = Kind::Morse;
    double de = 0, re = 0, a = 0, min_r = 0, max_r = 0;
    double barrier_height = 0, barrier_position = 0, barrier_width = 1;
    double asymptote = 0;
    uint32_t point_count = 0;
    std::shared_ptr<const std::vector<double>> table;
};

// Reference: VibwaAlgorithmConfig (algorithm_config.hpp:75-199).
struct VibwaConfig {
    double mass_atom_0 = 0, mass_atom_1 = 0, integration_step = 0, min_distance_to_asymptote = 0;
    uint32_t min_level = 0, max_level = 0;
};

struct SolverOptions {
    std::optional<int64_t> trial_energies;
    std::vector<int64_t> trial_energy_counts;
    std::vector<double> reduced_masses;
    double rel_tol = 1e-10;
    int32_t max_iterations = 200;
    bool eigenfunctions = false;
};

template <typename FP>
double as_fp(double x)
{
    return static_cast<double>(static_cast<FP>(x));
}

// Full snapshot of a configurator, taken at submit so later Python calls on
// the builder cannot change a running task.
struct TaskSpec {
    Precision precision = Precision::Float64;
    HardwareConfig hardware;
    std::vector<PotentialSpec> potentials;
    VibwaConfig vibwa;
    SolverOptions options;
};

struct ConfiguratorState {
    std::optional<HardwareConfig> hardware;
    std::optional<std::vector<PotentialSpec>> potentials;
    std::optional<VibwaConfig> vibwa;
    SolverOptions options;
};

// ------------------------------------------------------------------ task

double tabulate_point(const PotentialSpec& s, double h, uint32_t i)
{
    const double r = s.min_r + static_cast<double>(i) * h;
    if (s.kind == PotentialSpec::Kind::Tabulated)
        return (*s.table)[i];
    const double x = 1.0 - std::exp(-s.a * (r - s.re));
    double v = s.de * (x * x);
    if (s.kind == PotentialSpec::Kind::DoubleWell) {
        const double z = (r - s.barrier_position) / s.barrier_width;
        v += s.barrier_height * std::exp(-(z * z));
    }
    return v;
}

double window_top(const PotentialSpec& s, double margin)
{
    const double top = s.kind == PotentialSpec::Kind::Tabulated ? s.asymptote : s.de;
    return top - margin;
}

uint32_t point_count(const PotentialSpec& s)
{
    return s.kind == PotentialSpec::Kind::Tabulated ? static_cast<uint32_t>(s.table->size())
                                                    : s.point_count;
}

struct TaskResults {
    int32_t n_potentials = 0;
    int32_t level_count = 0;
    std::vector<double> levels;
    std::vector<int32_t> level_status;
    std::vector<int32_t> iterations;
    std::vector<std::vector<int32_t>> nodes;
    std::vector<std::vector<double>> far_end;
    std::vector<std::vector<int32_t>> rescales;
    std::vector<int32_t> strides;
    std::vector<int32_t> potential_flags;
    std::vector<epseon_bracket> brackets;  // per level, potential-major
    double scan_ms = 0, refine_ms = 0, total_ms = 0;
    double scan_point_updates = 0, refine_point_updates = 0;
    bool has_eigenfunctions = false;
    std::vector<std::vector<double>> eigenfunctions;  // per potential, level_count x n_points
    std::vector<int64_t> eigenfunction_points;
};

class TaskCancelled : public std::exception {
  public:
    const char* what() const noexcept override { return "task cancelled"; }
};

class Task {
  public:
    Task(int32_t device, TaskSpec spec) : device_(device), spec_(std::move(spec)) {}
    Task(const Task&) = delete;
    Task& operator=(const Task&) = delete;
    ~Task()
    {
        if (worker_.joinable()) {
            worker_.request_stop();
            worker_.join();
        }
    }

    // Reference: TaskHandle::startWorker (task_handle.hpp:94-101).
    void start()
    {
        if (is_running())
            throw std::runtime_error("One worker is already running, can't start another one.");
        done_.store(false, std::memory_order_release);
        started_.store(true, std::memory_order_release);
        worker_ = std::jthread([this](std::stop_token st) { run(st); });
    }

    bool is_done() const { return done_.load(std::memory_order_acquire); }
    bool is_started() const { return started_.load(std::memory_order_acquire); }
    bool is_running() const { return is_started() && !is_done(); }

    bool cancel()
    {
        if (is_running())
            return worker_.request_stop();
        return false;
    }

    void wait()
    {
        {
            py::gil_scoped_release release;
            std::lock_guard<std::mutex> lk(join_mutex_);
            if (worker_.joinable())
                worker_.join();
        }
        rethrow_if_failed();
    }

    std::string status() const
    {
        std::lock_guard<std::mutex> lk(status_mutex_);
        return status_;
    }

    const TaskResults& results() const
    {
        if (!is_done())
            throw std::runtime_error("Task has not finished; call wait() first.");
        rethrow_if_failed();
        if (cancelled_)
            throw std::runtime_error("Task was cancelled; no results.");
        return results_;
    }

    Precision precision() const { return spec_.precision; }

  private:
    void set_status(std::string s)
    {
        std::lock_guard<std::mutex> lk(status_mutex_);
        status_ = std::move(s);
    }

    void rethrow_if_failed() const
    {
        if (error_)
            std::rethrow_exception(error_);
    }

    void run(std::stop_token st)
    {
        try {
            execute(st);
            set_status("done");
        } catch (const TaskCancelled&) {
            cancelled_ = true;
            set_status("cancelled");
        } catch (const std::exception& e) {
            error_ = std::make_exception_ptr(std::runtime_error(e.what()));
            set_status(std::string("failed: ") + e.what());
        } catch (...) {
            error_ = std::make_exception_ptr(std::runtime_error("unknown error in task worker"));
            set_status("failed: unknown error");
        }
        done_.store(true, std::memory_order_release);
        started_.store(false, std::memory_order_release);
    }

    static void poll(const std::stop_token& st)
    {
        if (st.stop_requested())
            throw TaskCancelled();
    }

    // Stands in for VibwaAlgorithm<FP>::run (vibwa.hpp:605-637): the same
    // stop-token polls between phases, but the phases do the work.
    void execute(const std::stop_token& st)
    {
        poll(st);
        const TaskSpec& s = spec_;
        const size_t np = s.potentials.size();
        const double mu = s.vibwa.mass_atom_0 * s.vibwa.mass_atom_1 /
                          (s.vibwa.mass_atom_0 + s.vibwa.mass_atom_1);
        if (s.vibwa.max_level < s.vibwa.min_level)
            throw std::runtime_error("max_level must not be below min_level.");
        const int32_t level_count = static_cast<int32_t>(s.vibwa.max_level - s.vibwa.min_level + 1);
        if (!s.options.reduced_masses.empty() && s.options.reduced_masses.size() != np)
            throw std::runtime_error("reduced_masses must give one value per potential.");
        if (!s.options.trial_energy_counts.empty() && s.options.trial_energy_counts.size() != np)
            throw std::runtime_error("trial_energy_counts must give one value per potential.");
        for (const PotentialSpec& p : s.potentials)
            if (s.hardware.potential_buffer_size != 0 &&
                point_count(p) > s.hardware.potential_buffer_size)
                throw std::runtime_error(
                    "Potential point count " + std::to_string(point_count(p)) +
                    " exceeds potential_buffer_size " +
                    std::to_string(s.hardware.potential_buffer_size) + ".");
        const int64_t default_m =
            s.options.trial_energies.value_or(std::max<int64_t>(256, ((16 * level_count + 63) / 64) * 64));
        const size_t group = s.hardware.group_size ? s.hardware.group_size : np;

        TaskResults r;
        r.n_potentials = static_cast<int32_t>(np);
        r.level_count = level_count;
        r.levels.assign(np * level_count, std::nan(""));
        r.level_status.assign(np * level_count, EPSEON_LEVEL_UNBOUND);
        r.iterations.assign(np * level_count, 0);
        r.brackets.assign(np * level_count, epseon_bracket{});
        r.nodes.resize(np);
        r.far_end.resize(np);
        r.rescales.resize(np);
        r.strides.assign(np, 0);
        r.potential_flags.assign(np, 0);
        r.has_eigenfunctions = s.options.eigenfunctions;
        if (r.has_eigenfunctions) {
            r.eigenfunctions.resize(np);
            r.eigenfunction_points.assign(np, 0);
        }

        set_status("opening device");
        epseon_ctx* raw = nullptr;
        check_capi(epseon_ctx_create(device_, s.hardware.allocation_block_size, &raw),
                   "epseon_ctx_create");
        std::unique_ptr<epseon_ctx, int (*)(epseon_ctx*)> ctx(raw, epseon_ctx_destroy);

        const auto t_start = std::chrono::steady_clock::now();
        for (size_t g0 = 0; g0 < np; g0 += group) {
            const size_t g1 = std::min(np, g0 + group);
            poll(st);
            set_status("tabulating potentials " + std::to_string(g0) + ".." + std::to_string(g1 - 1));
            std::vector<epseon_potential_desc> desc(g1 - g0);
            std::vector<int64_t> offsets(g1 - g0);
            int64_t total = 0;
            for (size_t p = g0; p < g1; ++p) {
                offsets[p - g0] = total;
                total += point_count(s.potentials[p]);
            }
            std::vector<double> v(static_cast<size_t>(std::max<int64_t>(total, 1)));
            tabulate(g0, g1, offsets, v);
            for (size_t p = g0; p < g1; ++p) {
                const PotentialSpec& ps = s.potentials[p];
                epseon_potential_desc& d = desc[p - g0];
                d.n_points = point_count(ps);
                d.n_energies = s.options.trial_energy_counts.empty() ? default_m
                                                                     : s.options.trial_energy_counts[p];
                d.r_min = ps.min_r;
                d.r_max = ps.max_r;
                d.reduced_mass = s.options.reduced_masses.empty() ? mu : s.options.reduced_masses[p];
                d.e_top = window_top(ps, s.vibwa.min_distance_to_asymptote);
                d.level_min = static_cast<int32_t>(s.vibwa.min_level);
                d.level_count = level_count;
            }
            poll(st);
            set_status("uploading potentials");
            check_capi(epseon_load_potentials(ctx.get(), static_cast<int32_t>(g1 - g0), desc.data(),
                                              v.data(), offsets.data()),
                       "epseon_load_potentials");
            poll(st);
            set_status("scanning trial energies");
            check_capi(epseon_scan(ctx.get()), "epseon_scan");
            check_capi(epseon_ctx_synchronize(ctx.get()), "epseon_ctx_synchronize");
            poll(st);
            set_status("refining levels");
            check_capi(epseon_refine(ctx.get(), s.options.rel_tol, s.options.max_iterations),
                       "epseon_refine");
            set_status("reading results");
            const size_t nl = (g1 - g0) * static_cast<size_t>(level_count);
            check_capi(epseon_get_levels(ctx.get(), r.levels.data() + g0 * level_count,
                                         r.level_status.data() + g0 * level_count,
                                         r.iterations.data() + g0 * level_count),
                       "epseon_get_levels");
            check_capi(epseon_get_brackets(ctx.get(), r.brackets.data() + g0 * level_count),
                       "epseon_get_brackets");
            (void)nl;
            int64_t ne = 0;
            for (size_t p = g0; p < g1; ++p)
                ne += desc[p - g0].n_energies;
            std::vector<int32_t> nodes(static_cast<size_t>(ne)), resc(static_cast<size_t>(ne));
            std::vector<double> far(static_cast<size_t>(ne));
            check_capi(epseon_get_scan(ctx.get(), nodes.data(), far.data(), resc.data()),
                       "epseon_get_scan");
            int64_t o = 0;
            for (size_t p = g0; p < g1; ++p) {
                const int64_t m = desc[p - g0].n_energies;
                r.nodes[p].assign(nodes.begin() + o, nodes.begin() + o + m);
                r.far_end[p].assign(far.begin() + o, far.begin() + o + m);
                r.rescales[p].assign(resc.begin() + o, resc.begin() + o + m);
                o += m;
                epseon_potential_info info{};
                check_capi(epseon_potential_info_get(ctx.get(), static_cast<int32_t>(p - g0), &info),
                           "epseon_potential_info_get");
                r.strides[p] = info.stride;
                r.potential_flags[p] = info.flags;
            }
            if (s.options.eigenfunctions) {
                poll(st);
                set_status("normalising eigenfunctions");
                int64_t total_psi = 0;
                check_capi(epseon_get_eigenfunction_size(ctx.get(), &total_psi),
                           "epseon_get_eigenfunction_size");
                std::vector<double> psi(static_cast<size_t>(std::max<int64_t>(total_psi, 1)));
                check_capi(epseon_get_all_wavefunctions(ctx.get(), psi.data(), EPSEON_MEM_HOST),
                           "epseon_get_all_wavefunctions");
                int64_t q = 0;
                for (size_t p = g0; p < g1; ++p) {
                    const int64_t cnt = static_cast<int64_t>(level_count) * desc[p - g0].n_points;
                    r.eigenfunctions[p].assign(psi.begin() + q, psi.begin() + q + cnt);
                    r.eigenfunction_points[p] = desc[p - g0].n_points;
                    q += cnt;
                }
            }
            epseon_stats stats{};
            check_capi(epseon_get_stats(ctx.get(), &stats), "epseon_get_stats");
            r.scan_ms += stats.scan_ms;
            r.refine_ms += stats.refine_ms;
            r.scan_point_updates += stats.scan_point_updates;
            r.refine_point_updates += stats.refine_point_updates;
        }
        r.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start)
                         .count();
        results_ = std::move(r);
    }

    // V_i on each grid, host FP64 (std::exp), split over host threads for
    // large batches.
    void tabulate(size_t g0, size_t g1, const std::vector<int64_t>& offsets, std::vector<double>& v)
    {
        const size_t n = g1 - g0;
        const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        const unsigned workers = static_cast<unsigned>(std::min<size_t>(hw, n));
        auto body = [&](unsigned w) {
            for (size_t q = w; q < n; q += workers) {
                const PotentialSpec& ps = spec_.potentials[g0 + q];
                const uint32_t npts = point_count(ps);
                const double h = (ps.max_r - ps.min_r) / static_cast<double>(npts - 1);
                double* out = v.data() + offsets[q];
                for (uint32_t i = 0; i < npts; ++i)
                    out[i] = tabulate_point(ps, h, i);
            }
        };
        if (workers <= 1) {
            body(0);
            return;
        }
        std::vector<std::thread> pool;
        for (unsigned w = 0; w < workers; ++w)
            pool.emplace_back(body, w);
        for (std::thread& t : pool)
            t.join();
    }

    int32_t device_;
    TaskSpec spec_;
    std::atomic<bool> done_{false};
    std::atomic<bool> started_{false};
    bool cancelled_ = false;
    std::exception_ptr error_;
    TaskResults results_;
    mutable std::mutex status_mutex_;
    std::string status_ = "pending";
    std::mutex join_mutex_;
    std::jthread worker_;
};

// ------------------------------------------------------ python wrappers

template <typename FP>
class TaskHandle {
  public:
    explicit TaskHandle(std::shared_ptr<Task> task) : task_(std::move(task))
    {
        // Reference: the Python handle starts the worker on construction
        // (api.hpp:32-36).
        if (!task_->is_running())
            task_->start();
    }

    std::string get_status_message() const { return task_->status(); }
    bool is_done() const { return task_->is_done(); }
    bool is_running() const { return task_->is_running(); }
    bool cancel() { return task_->cancel(); }
    void wait() { task_->wait(); }

    py::array_t<FP> get_levels() const
    {
        const TaskResults& r = task_->results();
        py::array_t<FP> out({static_cast<py::ssize_t>(r.n_potentials),
                             static_cast<py::ssize_t>(r.level_count)});
        FP* p = out.mutable_data();
        for (size_t i = 0; i < r.levels.size(); ++i)
            p[i] = static_cast<FP>(r.levels[i]);
        return out;
    }

    py::array_t<int32_t> get_level_status() const { return shaped(task_->results().level_status); }
    py::array_t<int32_t> get_iterations() const { return shaped(task_->results().iterations); }

    std::vector<py::array_t<int32_t>> get_node_counts() const
    {
        return ragged(task_->results().nodes);
    }
    std::vector<py::array_t<double>> get_far_end_values() const
    {
        return ragged(task_->results().far_end);
    }
    std::vector<py::array_t<int32_t>> get_rescale_counts() const
    {
        return ragged(task_->results().rescales);
    }
    std::vector<int32_t> get_node_strides() const { return task_->results().strides; }

    py::dict get_brackets() const
    {
        const TaskResults& r = task_->results();
        const std::vector<py::ssize_t> shape{static_cast<py::ssize_t>(r.n_potentials),
                                             static_cast<py::ssize_t>(r.level_count)};
        py::array_t<double> e_lo(shape), e_hi(shape);
        py::array_t<int32_t> n_lo(shape), n_hi(shape), found(shape);
        for (size_t i = 0; i < r.brackets.size(); ++i) {
            const epseon_bracket& b = r.brackets[i];
            e_lo.mutable_data()[i] = b.e_lo;
            e_hi.mutable_data()[i] = b.e_hi;
            n_lo.mutable_data()[i] = b.nodes_lo;
            n_hi.mutable_data()[i] = b.nodes_hi;
            found.mutable_data()[i] = b.found;
        }
        py::dict d;
        d["e_lo"] = e_lo;
        d["e_hi"] = e_hi;
        d["nodes_lo"] = n_lo;
        d["nodes_hi"] = n_hi;
        d["found"] = found;
        return d;
    }
    std::vector<int32_t> get_potential_flags() const { return task_->results().potential_flags; }

    std::vector<py::array_t<FP>> get_wavefunctions() const
    {
        const TaskResults& r = task_->results();
        if (!r.has_eigenfunctions)
            throw std::runtime_error(
                "Eigenfunctions were not requested: call set_eigenfunction_output(True) before submit_task.");
        std::vector<py::array_t<FP>> out;
        out.reserve(r.eigenfunctions.size());
        for (size_t p = 0; p < r.eigenfunctions.size(); ++p) {
            py::array_t<FP> a({static_cast<py::ssize_t>(r.level_count),
                               static_cast<py::ssize_t>(r.eigenfunction_points[p])});
            FP* d = a.mutable_data();
            for (size_t i = 0; i < r.eigenfunctions[p].size(); ++i)
                d[i] = static_cast<FP>(r.eigenfunctions[p][i]);
            out.push_back(std::move(a));
        }
        return out;
    }

    py::dict get_timings() const
    {
        const TaskResults& r = task_->results();
        py::dict d;
        d["scan_ms"] = r.scan_ms;
        d["refine_ms"] = r.refine_ms;
        d["total_ms"] = r.total_ms;
        d["scan_point_updates"] = r.scan_point_updates;
        d["refine_point_updates"] = r.refine_point_updates;
        return d;
    }

  private:
    py::array_t<int32_t> shaped(const std::vector<int32_t>& v) const
    {
        const TaskResults& r = task_->results();
        py::array_t<int32_t> out({static_cast<py::ssize_t>(r.n_potentials),
                                  static_cast<py::ssize_t>(r.level_count)});
        std::copy(v.begin(), v.end(), out.mutable_data());
        return out;
    }

    template <typename T>
    static std::vector<py::array_t<T>> ragged(const std::vector<std::vector<T>>& rows)
    {
        std::vector<py::array_t<T>> out;
        out.reserve(rows.size());
        for (const auto& row : rows) {
            py::array_t<T> a(static_cast<py::ssize_t>(row.size()));
            std::copy(row.begin(), row.end(), a.mutable_data());
            out.push_back(std::move(a));
        }
        return out;
    }

    std::shared_ptr<Task> task_;
};

using TaskHandleFloat32 = TaskHandle<float>;
using TaskHandleFloat64 = TaskHandle<double>;

template <typename FP>
class TaskConfigurator {
  public:
    TaskConfigurator() : state_(std::make_shared<ConfiguratorState>()) {}

    // Reference: set_hardware_config (api.hpp:156-165).
    TaskConfigurator& set_hardware_config(uint32_t potential_buffer_size, uint32_t group_size,
                                          uint32_t allocation_block_size)
    {
        state_->hardware = HardwareConfig{potential_buffer_size, group_size, allocation_block_size};
        return *this;
    }

    // Reference: set_morse_potential (api.hpp:169-218), including its
    // same-point-count rule and message.
    TaskConfigurator& set_morse_potential(const std::vector<MorsePotentialConfig>& configurations)
    {
        std::optional<uint32_t> count;
        for (const MorsePotentialConfig& c : configurations) {
            if (count.has_value() && *count != c.point_count) {
                char msg[192];
                std::snprintf(msg, sizeof msg,
                              "All Morse potentials must have same point count, but previous ones "
                              "had %u and current one has.",
                              *count);
                throw std::runtime_error(msg);
            }
            count = c.point_count;
        }
        return set_morse_potential_batch(configurations);
    }

    // Additive: Morse potentials with independent grids.
    TaskConfigurator& set_morse_potential_batch(const std::vector<MorsePotentialConfig>& configurations)
    {
        std::vector<PotentialSpec> v;
        v.reserve(configurations.size());
        for (const MorsePotentialConfig& c : configurations)
            v.push_back(morse_spec(c));
        state_->potentials = std::move(v);
        return *this;
    }

    // Additive: Morse + Gaussian barrier potentials.
    TaskConfigurator& set_double_well_potential(
        const std::vector<DoubleWellPotentialConfig>& configurations)
    {
        std::vector<PotentialSpec> v;
        v.reserve(configurations.size());
        for (const DoubleWellPotentialConfig& c : configurations) {
            PotentialSpec s = morse_spec(c.morse);
            s.kind = PotentialSpec::Kind::DoubleWell;
            s.barrier_height = as_fp<FP>(c.barrier_height);
            s.barrier_position = as_fp<FP>(c.barrier_position);
            s.barrier_width = as_fp<FP>(c.barrier_width);
            v.push_back(s);
        }
        state_->potentials = std::move(v);
        return *this;
    }

    // Additive: tabulated V(r) on uniform grids.
    TaskConfigurator& set_tabulated_potential(const std::vector<TabulatedPotentialConfig>& configurations)
    {
        std::vector<PotentialSpec> v;
        v.reserve(configurations.size());
        for (const TabulatedPotentialConfig& c : configurations) {
            if (c.values.size() < 4)
                throw std::runtime_error("A tabulated potential needs at least 4 points.");
            PotentialSpec s;
            s.kind = PotentialSpec::Kind::Tabulated;
            s.min_r = as_fp<FP>(c.min_r);
            s.max_r = as_fp<FP>(c.max_r);
            s.asymptote = as_fp<FP>(c.asymptote);
            auto table = std::make_shared<std::vector<double>>(c.values.size());
            for (size_t i = 0; i < c.values.size(); ++i)
                (*table)[i] = as_fp<FP>(c.values[i]);
            s.table = std::move(table);
            s.point_count = static_cast<uint32_t>(c.values.size());
            v.push_back(std::move(s));
        }
        state_->potentials = std::move(v);
        return *this;
    }

    // Reference: set_vibwa_algorithm (api.hpp:222-241).
    TaskConfigurator& set_vibwa_algorithm(double mass_atom_0, double mass_atom_1, double integration_step,
                                          double min_distance_to_asymptote, uint32_t min_level,
                                          uint32_t max_level)
    {
        state_->vibwa = VibwaConfig{as_fp<FP>(mass_atom_0), as_fp<FP>(mass_atom_1),
                                    as_fp<FP>(integration_step), as_fp<FP>(min_distance_to_asymptote),
                                    min_level, max_level};
        return *this;
    }

    // Additive solver options.
    TaskConfigurator& set_trial_energies(int64_t count)
    {
        if (count < 1)
            throw std::runtime_error("The trial energy count must be positive.");
        state_->options.trial_energies = count;
        return *this;
    }

    TaskConfigurator& set_trial_energy_counts(const std::vector<int64_t>& counts)
    {
        for (int64_t c : counts)
            if (c < 1)
                throw std::runtime_error("Trial energy counts must be positive.");
        state_->options.trial_energy_counts = counts;
        return *this;
    }

    TaskConfigurator& set_reduced_masses(const std::vector<double>& masses)
    {
        for (double m : masses)
            if (!(m > 0.0))
                throw std::runtime_error("Reduced masses must be positive.");
        state_->options.reduced_masses.clear();
        for (double m : masses)
            state_->options.reduced_masses.push_back(as_fp<FP>(m));
        return *this;
    }

    TaskConfigurator& set_tolerance(double relative, int32_t max_iterations)
    {
        if (!(relative >= 0.0) || !std::isfinite(relative))
            throw std::runtime_error("The relative tolerance must be finite and non-negative.");
        if (max_iterations < 1)
            throw std::runtime_error("max_iterations must be positive.");
        state_->options.rel_tol = relative;
        state_->options.max_iterations = max_iterations;
        return *this;
    }

    TaskConfigurator& set_eigenfunction_output(bool enabled)
    {
        state_->options.eigenfunctions = enabled;
        return *this;
    }

    bool is_configured() const
    {
        return state_->hardware.has_value() && state_->potentials.has_value() &&
               state_->vibwa.has_value();
    }

    TaskSpec snapshot() const
    {
        TaskSpec s;
        s.precision = std::is_same_v<FP, float> ? Precision::Float32 : Precision::Float64;
        s.hardware = *state_->hardware;
        s.potentials = *state_->potentials;
        s.vibwa = *state_->vibwa;
        s.options = state_->options;
        return s;
    }

  private:
    static PotentialSpec morse_spec(const MorsePotentialConfig& c)
    {
        if (c.point_count < 4)
            throw std::runtime_error("A potential needs at least 4 grid points.");
        PotentialSpec s;
        s.kind = PotentialSpec::Kind::Morse;
        s.de = as_fp<FP>(c.dissociation_energy);
        s.re = as_fp<FP>(c.equilibrium_bond_distance);
        s.a = as_fp<FP>(c.well_width);
        s.min_r = as_fp<FP>(c.min_r);
        s.max_r = as_fp<FP>(c.max_r);
        s.point_count = c.point_count;
        return s;
    }

    std::shared_ptr<ConfiguratorState> state_;
};

using TaskConfiguratorFloat32 = TaskConfigurator<float>;
using TaskConfiguratorFloat64 = TaskConfigurator<double>;
using TaskConfiguratorVariant = std::variant<TaskConfiguratorFloat32, TaskConfiguratorFloat64>;
using TaskHandleVariant = std::variant<TaskHandleFloat32, TaskHandleFloat64>;

// Reference: ComputeDeviceInterface (api.hpp:265-290, device_interface.hpp:22-51).
class ComputeDeviceInterface {
  public:
    explicit ComputeDeviceInterface(int32_t ordinal) : ordinal_(ordinal) {}

    TaskConfiguratorVariant get_task_configurator(const std::string& precision) const
    {
        switch (parse_precision(precision)) {
        case Precision::Float32: return TaskConfiguratorFloat32{};
        case Precision::Float64: return TaskConfiguratorFloat64{};
        }
        throw std::runtime_error("Unreachable.");
    }

    template <typename FP>
    TaskHandleVariant submit_task(const TaskConfigurator<FP>& cfg) const
    {
        if (!cfg.is_configured())
            throw std::runtime_error("TaskConfigurator submitted for execution before fully configured.");
        auto task = std::make_shared<Task>(ordinal_, cfg.snapshot());
        return TaskHandleVariant{TaskHandle<FP>{task}};
    }

    int32_t device_id() const { return ordinal_; }

  private:
    int32_t ordinal_;
};

// Reference: EpseonComputeContext (api.hpp:292-309, compute_context.cpp:41-117).
class EpseonComputeContext {
  public:
    static EpseonComputeContext create()
    {
        EpseonComputeContext c;
        c.device_count_ = device_count();
        return c;
    }

    std::string get_vulkan_version() const
    {
        epseon_device_info d{};
        if (device_count_ > 0 && epseon_device_info_get(0, &d) == EPSEON_OK)
            return version_string(d.runtime_version);
        return "0.0.0.0";
    }

    std::vector<PhysicalDeviceInfo> get_physical_device_info() const
    {
        std::vector<PhysicalDeviceInfo> out;
        for (int32_t i = 0; i < device_count_; ++i)
            out.push_back(query_device(i));
        return out;
    }

    ComputeDeviceInterface get_device_interface(uint32_t device_id) const
    {
        if (static_cast<int64_t>(device_id) >= device_count_)
            throw std::runtime_error("Device not available.");
        return ComputeDeviceInterface(static_cast<int32_t>(device_id));
    }

  private:
    int32_t device_count_ = 0;
};

template <typename FP>
void bind_configurator(py::module_& m, const char* name)
{
    using C = TaskConfigurator<FP>;
    py::class_<C>(m, name, "Builder for configuring GPU compute task.")
        .def(py::init<>())
        .def("set_hardware_config", &C::set_hardware_config, py::arg("potential_buffer_size"),
             py::arg("group_size"), py::arg("allocation_block_size"),
             "Set hardware configuration for a GPU compute task.", py::return_value_policy::reference)
        .def("set_morse_potential", &C::set_morse_potential, py::arg("configurations"),
             "Set potential data source configuration for GPU compute task.",
             py::return_value_policy::reference)
        .def("set_vibwa_algorithm", &C::set_vibwa_algorithm, py::arg("mass_atom_0"),
             py::arg("mass_atom_1"), py::arg("integration_step"), py::arg("min_distance_to_asymptote"),
             py::arg("min_level"), py::arg("max_level"),
             "Set algorithm configuration for a GPU compute task.", py::return_value_policy::reference)
        .def("is_configured", &C::is_configured,
             "Check if this instance is fully configured, i.e. it has been assigned a valid hardware "
             "configuration, potential source and algorithm config.")
        .def("set_morse_potential_batch", &C::set_morse_potential_batch, py::arg("configurations"),
             "Set Morse potentials whose grids may differ in point count.",
             py::return_value_policy::reference)
        .def("set_double_well_potential", &C::set_double_well_potential, py::arg("configurations"),
             "Set Morse potentials with a Gaussian barrier.", py::return_value_policy::reference)
        .def("set_tabulated_potential", &C::set_tabulated_potential, py::arg("configurations"),
             "Set tabulated potentials on uniform grids.", py::return_value_policy::reference)
        .def("set_trial_energies", &C::set_trial_energies, py::arg("count"),
             "Number of trial energies scanned per potential.", py::return_value_policy::reference)
        .def("set_trial_energy_counts", &C::set_trial_energy_counts, py::arg("counts"),
             "Per-potential trial energy counts.", py::return_value_policy::reference)
        .def("set_reduced_masses", &C::set_reduced_masses, py::arg("masses"),
             "Per-potential reduced masses (amu), overriding mass_atom_0/1.",
             py::return_value_policy::reference)
        .def("set_tolerance", &C::set_tolerance, py::arg("relative"), py::arg("max_iterations") = 200,
             "Relative level tolerance and refinement iteration cap.",
             py::return_value_policy::reference)
        .def("set_eigenfunction_output", &C::set_eigenfunction_output, py::arg("enabled") = true,
             "Also compute the normalised eigenfunction of every level (sum psi^2 h = 1).",
             py::return_value_policy::reference);
}

template <typename FP>
void bind_handle(py::module_& m, const char* name, const char* doc)
{
    using H = TaskHandle<FP>;
    py::class_<H>(m, name, doc)
        .def("get_status_message", &H::get_status_message,
             "Get status message explaining current execution stage.")
        .def("is_done", &H::is_done, "Check if task already finished execution.")
        .def("wait", &H::wait, "Block and wait for task to finish.")
        .def("is_running", &H::is_running, "Check if the task is still running.")
        .def("cancel", &H::cancel, "Request cancellation; the worker stops between phases.")
        .def("get_levels", &H::get_levels, "Refined levels, shape (potentials, levels); NaN if absent.")
        .def("get_level_status", &H::get_level_status, "Per-level status codes (0 = converged).")
        .def("get_iterations", &H::get_iterations, "Refinement shots per level.")
        .def("get_node_counts", &H::get_node_counts, "Scan node counts, one array per potential.")
        .def("get_far_end_values", &H::get_far_end_values, "Scan far-end values w_{N-1}.")
        .def("get_rescale_counts", &H::get_rescale_counts, "Scan rescale counts.")
        .def("get_node_strides", &H::get_node_strides, "Node-sampling stride per potential.")
        .def("get_brackets", &H::get_brackets,
             "Each level's bracket from the scan: dict of (potentials, levels) arrays e_lo, e_hi, "
             "nodes_lo, nodes_hi, found.")
        .def("get_potential_flags", &H::get_potential_flags, "Per-potential flags (1 = unresolved).")
        .def("get_timings", &H::get_timings, "Device times and point-update counts.")
        .def("get_wavefunctions", &H::get_wavefunctions,
             "Normalised eigenfunctions, one (levels, points) array per potential; zero rows for absent "
             "levels.");
}

}  // namespace epseon::gpu

using namespace epseon::gpu;

PYBIND11_MODULE(_libepseon_gpu, m)
{
    m.doc() = "Sub package for interacting with GPU compute capabilities.";

    py::class_<PhysicalDeviceSparseProperties>(m, "PhysicalDeviceSparseProperties",
                                               "Sparse resource properties (none on this backend).");

    py::class_<PhysicalDeviceLimits>(m, "PhysicalDeviceLimits",
                                     "Physical device limits - mostly max counts of different resources.")
        .def_readonly("max_compute_shared_memory_size", &PhysicalDeviceLimits::max_compute_shared_memory_size)
        .def_readonly("max_compute_work_group_count", &PhysicalDeviceLimits::max_compute_work_group_count)
        .def_readonly("max_compute_work_group_invocations",
                      &PhysicalDeviceLimits::max_compute_work_group_invocations)
        .def_readonly("max_compute_work_group_size", &PhysicalDeviceLimits::max_compute_work_group_size);

    py::class_<PhysicalDeviceProperties>(m, "PhysicalDeviceProperties",
                                         "Properties of a physical device.")
        .def_readonly("api_version", &PhysicalDeviceProperties::api_version)
        .def_readonly("driver_version", &PhysicalDeviceProperties::driver_version)
        .def_readonly("vendor_id", &PhysicalDeviceProperties::vendor_id)
        .def_readonly("device_id", &PhysicalDeviceProperties::device_id)
        .def_readonly("device_type", &PhysicalDeviceProperties::device_type)
        .def_readonly("device_name", &PhysicalDeviceProperties::device_name)
        .def_readonly("pipeline_cache_uuid", &PhysicalDeviceProperties::pipeline_cache_uuid)
        .def_readonly("limits", &PhysicalDeviceProperties::limits)
        .def_readonly("sparse_properties", &PhysicalDeviceProperties::sparse_properties)
        .def_readonly("arch", &PhysicalDeviceProperties::arch)
        .def_readonly("compute_units", &PhysicalDeviceProperties::compute_units)
        .def_readonly("pci_device_id", &PhysicalDeviceProperties::pci_device_id);

    py::class_<MemoryHeap>(m, "MemoryHeap", "A device memory heap.")
        .def_readonly("size", &MemoryHeap::size)
        .def_readonly("flags", &MemoryHeap::flags);

    py::class_<MemoryType>(m, "MemoryType", "A device memory type.")
        .def_readonly("heap_index", &MemoryType::heap_index)
        .def_readonly("flags", &MemoryType::flags);

    py::class_<PhysicalDeviceMemoryProperties>(m, "PhysicalDeviceMemoryProperties",
                                               "Properties of physical device memory.")
        .def_readonly("memory_heaps", &PhysicalDeviceMemoryProperties::memory_heaps)
        .def_readonly("memory_types", &PhysicalDeviceMemoryProperties::memory_types);

    py::class_<PhysicalDeviceInfo>(m, "PhysicalDeviceInfo", "Container for physical device info.")
        .def_readonly("device_properties", &PhysicalDeviceInfo::device_properties)
        .def_readonly("memory_properties", &PhysicalDeviceInfo::memory_properties);

    bind_handle<float>(m, "TaskHandleFloat32", "Handle object for referencing single precision GPU compute task.");
    bind_handle<double>(m, "TaskHandleFloat64", "Handle object for referencing double precision GPU compute task.");

    py::class_<MorsePotentialConfig>(m, "MorsePotentialConfig", "Configuration of single Morse potential curve.")
        .def(py::init([](double de, double re, double a, double min_r, double max_r, uint32_t n) {
                 return MorsePotentialConfig{de, re, a, min_r, max_r, n};
             }),
             py::arg("dissociation_energy"), py::arg("equilibrium_bond_distance"), py::arg("well_width"),
             py::arg("min_r"), py::arg("max_r"), py::arg("point_count"),
             "Create instance of MorsePotentialConfig class.")
        .def_readonly("dissociation_energy", &MorsePotentialConfig::dissociation_energy)
        .def_readonly("equilibrium_bond_distance", &MorsePotentialConfig::equilibrium_bond_distance)
        .def_readonly("well_width", &MorsePotentialConfig::well_width)
        .def_readonly("min_r", &MorsePotentialConfig::min_r)
        .def_readonly("max_r", &MorsePotentialConfig::max_r)
        .def_readonly("point_count", &MorsePotentialConfig::point_count);

    py::class_<DoubleWellPotentialConfig>(m, "DoubleWellPotentialConfig",
                                          "Morse potential plus a Gaussian barrier B exp(-((r - r_b)/w)^2).")
        .def(py::init([](double de, double re, double a, double b, double rb, double w, double min_r,
                         double max_r, uint32_t n) {
                 if (!(w > 0.0))
                     throw std::runtime_error("barrier_width must be positive.");
                 return DoubleWellPotentialConfig{MorsePotentialConfig{de, re, a, min_r, max_r, n}, b, rb, w};
             }),
             py::arg("dissociation_energy"), py::arg("equilibrium_bond_distance"), py::arg("well_width"),
             py::arg("barrier_height"), py::arg("barrier_position"), py::arg("barrier_width"),
             py::arg("min_r"), py::arg("max_r"), py::arg("point_count"));

    py::class_<TabulatedPotentialConfig>(m, "TabulatedPotentialConfig",
                                         "V(r) values on the uniform grid [min_r, max_r]; levels are "
                                         "searched below `asymptote`.")
        .def(py::init([](std::vector<double> values, double min_r, double max_r, double asymptote) {
                 return TabulatedPotentialConfig{std::move(values), min_r, max_r, asymptote};
             }),
             py::arg("values"), py::arg("min_r"), py::arg("max_r"), py::arg("asymptote"));

    bind_configurator<float>(m, "TaskConfiguratorFloat32");
    bind_configurator<double>(m, "TaskConfiguratorFloat64");

    py::class_<ComputeDeviceInterface>(m, "ComputeDeviceInterface", "Interface to particular GPU device.")
        .def("get_task_configurator", &ComputeDeviceInterface::get_task_configurator,
             "Get builder instance for configuring GPU compute task.")
        .def("submit_task", &ComputeDeviceInterface::submit_task<float>,
             "Submit task for execution. Will raise RuntimeError upon receiving not fully configured "
             "TaskConfigurator.")
        .def("submit_task", &ComputeDeviceInterface::submit_task<double>,
             "Submit task for execution. Will raise RuntimeError upon receiving not fully configured "
             "TaskConfigurator.")
        .def_property_readonly("device_id", &ComputeDeviceInterface::device_id);

    py::class_<EpseonComputeContext>(m, "EpseonComputeContext", "GPU compute context handle.")
        .def_static("create", &EpseonComputeContext::create, "Create instance of EpseonComputeContext.")
        .def("get_vulkan_version", &EpseonComputeContext::get_vulkan_version,
             "API version string (HIP runtime version on this backend).")
        .def("get_physical_device_info", &EpseonComputeContext::get_physical_device_info,
             "Get information about available physical devices.")
        .def("get_device_interface", &EpseonComputeContext::get_device_interface,
             "Get interface for running algorithms on a device.");
}
