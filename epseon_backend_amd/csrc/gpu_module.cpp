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
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <stop_token>
#include <string>
#include <thread>
#include <tuple>
#include <type_traits>
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

// Reference: PotentialFileLoader (potential_source.hpp:43-100), whose data
// getter returns {} (:89-91) and whose file format the reference never
// defines.  Format read here, one potential per file:
//   * text: one grid point per line, "r V" (r in Angstrom, V in cm^-1,
//     separated by spaces, tabs or one comma); blank lines and lines starting
//     with '#' are skipped, except "# asymptote <value>", which sets the
//     energy the levels are searched below (default: V at the last point);
//   * .npy (by file name): a little-endian float64 array of shape (N, 2),
//     C order, rows (r, V); asymptote: V at the last point.
// The r column must be strictly increasing and uniform: every r_i within
// 1e-4 h of r_0 + i h, h = (r_{N-1} - r_0)/(N - 1).  At least 4 points, all
// values finite.
struct LoadedPotential {
    std::vector<double> r, v;
    double asymptote = 0;
};

[[noreturn]] void file_error(const std::string& file, const std::string& what)
{
    throw std::runtime_error("Potential file " + file + ": " + what);
}

LoadedPotential parse_potential_text(const std::string& file)
{
    FILE* f = std::fopen(file.c_str(), "rb");
    if (f == nullptr)
        file_error(file, "cannot be opened");
    std::string text;
    char buf[1 << 16];
    size_t got = 0;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0)
        text.append(buf, got);
    std::fclose(f);
    LoadedPotential out;
    std::optional<double> asymptote;
    size_t line_no = 0, pos = 0;
    while (pos <= text.size()) {
        size_t end = text.find('\n', pos);
        if (end == std::string::npos)
            end = text.size();
        std::string line = text.substr(pos, end - pos);
        pos = end + 1;
        ++line_no;
        if (!line.empty() && line.back() == '\r')
            line.pop_back();
        size_t b = line.find_first_not_of(" \t");
        if (b == std::string::npos)
            continue;
        if (line[b] == '#') {
            std::string rest = line.substr(b + 1);
            size_t k = rest.find_first_not_of(" \t");
            if (k != std::string::npos && rest.compare(k, 9, "asymptote") == 0) {
                const char* c = rest.c_str() + k + 9;
                char* stop = nullptr;
                const double x = std::strtod(c, &stop);
                if (stop == c || !std::isfinite(x))
                    file_error(file, "line " + std::to_string(line_no) + ": bad asymptote value");
                asymptote = x;
            }
            continue;
        }
        const char* c = line.c_str() + b;
        char* stop = nullptr;
        const double r = std::strtod(c, &stop);
        if (stop == c)
            file_error(file, "line " + std::to_string(line_no) + ": expected two numbers \"r V\"");
        c = stop;
        while (*c == ' ' || *c == '\t')
            ++c;
        if (*c == ',')
            ++c;
        const double v = std::strtod(c, &stop);
        if (stop == c)
            file_error(file, "line " + std::to_string(line_no) + ": expected two numbers \"r V\"");
        c = stop;
        while (*c == ' ' || *c == '\t')
            ++c;
        if (*c != '\0')
            file_error(file, "line " + std::to_string(line_no) + ": unexpected text after \"r V\"");
        out.r.push_back(r);
        out.v.push_back(v);
    }
    if (out.v.empty())
        file_error(file, "no data lines");
    out.asymptote = asymptote.value_or(out.v.back());
    return out;
}

LoadedPotential parse_potential_npy(const std::string& file)
{
    FILE* f = std::fopen(file.c_str(), "rb");
    if (f == nullptr)
        file_error(file, "cannot be opened");
    std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);
    unsigned char magic[8];
    if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "\x93NUMPY", 6) != 0)
        file_error(file, "not a .npy file");
    size_t header_len = 0;
    if (magic[6] == 1) {
        unsigned char l[2];
        if (std::fread(l, 1, 2, f) != 2)
            file_error(file, "truncated header");
        header_len = l[0] | (size_t(l[1]) << 8);
    } else if (magic[6] == 2 || magic[6] == 3) {
        unsigned char l[4];
        if (std::fread(l, 1, 4, f) != 4)
            file_error(file, "truncated header");
        header_len = l[0] | (size_t(l[1]) << 8) | (size_t(l[2]) << 16) | (size_t(l[3]) << 24);
    } else {
        file_error(file, "unsupported .npy version");
    }
    std::string header(header_len, '\0');
    if (std::fread(header.data(), 1, header_len, f) != header_len)
        file_error(file, "truncated header");
    auto field = [&](const char* key) -> std::string {
        const size_t k = header.find(std::string("'") + key + "'");
        if (k == std::string::npos)
            file_error(file, std::string("header lacks '") + key + "'");
        size_t c = header.find(':', k);
        return c == std::string::npos ? std::string() : header.substr(c + 1);
    };
    const std::string descr = field("descr");
    if (descr.find("'<f8'") == std::string::npos)
        file_error(file, "the array must be little-endian float64 ('<f8')");
    if (field("fortran_order").find("False") == std::string::npos)
        file_error(file, "the array must be in C order");
    const std::string shape = field("shape");
    const size_t open = shape.find('('), close = shape.find(')');
    if (open == std::string::npos || close == std::string::npos)
        file_error(file, "bad shape");
    long long rows = -1, cols = -1;
    if (std::sscanf(shape.substr(open + 1, close - open - 1).c_str(), "%lld , %lld", &rows, &cols) != 2 ||
        cols != 2 || rows < 1)
        file_error(file, "the array must have shape (N, 2)");
    std::vector<double> data(static_cast<size_t>(rows) * 2);
    if (std::fread(data.data(), sizeof(double), data.size(), f) != data.size())
        file_error(file, "truncated data");
    LoadedPotential out;
    out.r.resize(static_cast<size_t>(rows));
    out.v.resize(static_cast<size_t>(rows));
    for (size_t i = 0; i < static_cast<size_t>(rows); ++i) {
        out.r[i] = data[2 * i];
        out.v[i] = data[2 * i + 1];
    }
    out.asymptote = out.v.back();
    return out;
}

bool ends_with(const std::string& s, const std::string& suffix)
{
    return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}

LoadedPotential read_potential_file(const std::string& file)
{
    LoadedPotential p = ends_with(file, ".npy") ? parse_potential_npy(file) : parse_potential_text(file);
    const size_t n = p.v.size();
    if (n < 4)
        file_error(file, "needs at least 4 grid points, has " + std::to_string(n));
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(p.r[i]) || !std::isfinite(p.v[i]))
            file_error(file, "point " + std::to_string(i) + " is not finite");
    const double h = (p.r[n - 1] - p.r[0]) / static_cast<double>(n - 1);
    if (!(h > 0.0))
        file_error(file, "r must increase");
    for (size_t i = 0; i < n; ++i)
        if (std::fabs(p.r[i] - (p.r[0] + static_cast<double>(i) * h)) > 1e-4 * h)
            file_error(file, "r is not a uniform grid (point " + std::to_string(i) + ")");
    return p;
}

class PotentialFileLoader {
  public:
    PotentialFileLoader() = default;
    explicit PotentialFileLoader(std::vector<std::string> file_names) : file_names_(std::move(file_names)) {}

    const std::vector<std::string>& file_names() const { return file_names_; }

    // Reference: equals (potential_source.hpp:80-87): same file names.
    bool equals(const PotentialFileLoader& other) const { return file_names_ == other.file_names_; }

    // Reference: get_potential_data (potential_source.hpp:89-91): the V
    // values of every file, in file order.
    std::vector<std::vector<double>> get_potential_data() const
    {
        std::vector<std::vector<double>> out;
        out.reserve(file_names_.size());
        for (const std::string& f : file_names_)
            out.push_back(read_potential_file(f).v);
        return out;
    }

    std::vector<LoadedPotential> read() const
    {
        std::vector<LoadedPotential> out;
        out.reserve(file_names_.size());
        for (const std::string& f : file_names_)
            out.push_back(read_potential_file(f));
        return out;
    }

  private:
    std::vector<std::string> file_names_;
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

// Scan results of every trial energy of a task, potential-major.
struct ScanArrays {
    std::vector<int32_t> nodes, rescales;
    std::vector<double> far_end;
};

using EigenRows = std::shared_ptr<const std::vector<double>>;

struct TaskResults {
    int32_t n_potentials = 0;
    int32_t level_count = 0;
    std::vector<double> levels;
    std::vector<int32_t> level_status;
    std::vector<int32_t> iterations;
    std::vector<int64_t> energy_offsets;  // first trial energy of each potential, plus the total
    std::vector<int64_t> point_counts;    // grid points of each potential
    std::vector<int32_t> strides;
    std::vector<int32_t> potential_flags;
    std::vector<epseon_bracket> brackets;  // per level, potential-major
    double solve_ms = 0, total_ms = 0;
    double scan_point_updates = 0, refine_point_updates = 0;
    bool has_eigenfunctions = false;
    // Filled group by group when the task ran in several groups (the device
    // context is reloaded per group); for a one-group task read from the
    // context the task keeps, on first request.
    std::shared_ptr<const ScanArrays> scan;
    std::vector<EigenRows> eigenfunctions;  // per potential, level_count x n_points
};

class TaskCancelled : public std::exception {
  public:
    const char* what() const noexcept override { return "task cancelled"; }
};

using CtxPtr = std::unique_ptr<epseon_ctx, int (*)(epseon_ctx*)>;

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

    // Reference: TaskHandle::cancel (task_handle.hpp:136-144): a stop request
    // the worker honours at its next phase boundary.
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

    // Scan arrays of the whole task (read from the device on first request
    // when the task ran as one group).
    std::shared_ptr<const ScanArrays> scan_arrays()
    {
        const TaskResults& r = results();
        std::lock_guard<std::mutex> lk(fetch_mutex_);
        if (!results_.scan) {
            auto sa = std::make_shared<ScanArrays>();
            const size_t ne = static_cast<size_t>(r.energy_offsets.back());
            sa->nodes.resize(ne);
            sa->far_end.resize(ne);
            sa->rescales.resize(ne);
            check_capi(epseon_get_scan(ctx_.get(), sa->nodes.data(), sa->far_end.data(), sa->rescales.data()),
                       "epseon_get_scan");
            results_.scan = std::move(sa);
        }
        return results_.scan;
    }

    // Eigenfunctions of potential p (read from the device on first request
    // when the task ran as one group).
    EigenRows eigenfunctions(size_t p)
    {
        const TaskResults& r = results();
        if (!r.has_eigenfunctions)
            throw std::runtime_error(
                "Eigenfunctions were not requested: call set_eigenfunction_output(True) before submit_task.");
        if (p >= r.eigenfunctions.size())
            throw py::index_error("potential index out of range");
        std::lock_guard<std::mutex> lk(fetch_mutex_);
        if (!results_.eigenfunctions[p]) {
            auto rows = std::make_shared<std::vector<double>>(static_cast<size_t>(r.level_count) *
                                                              static_cast<size_t>(r.point_counts[p]));
            if (!rows->empty())
                check_capi(epseon_get_wavefunctions(ctx_.get(), static_cast<int32_t>(p), rows->data(),
                                                    EPSEON_MEM_HOST),
                           "epseon_get_wavefunctions");
            results_.eigenfunctions[p] = std::move(rows);
        }
        return results_.eigenfunctions[p];
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
            ctx_.reset();
            set_status("cancelled");
        } catch (const std::exception& e) {
            error_ = std::make_exception_ptr(std::runtime_error(e.what()));
            ctx_.reset();
            set_status(std::string("failed: ") + e.what());
        } catch (...) {
            error_ = std::make_exception_ptr(std::runtime_error("unknown error in task worker"));
            ctx_.reset();
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
        const size_t group = s.hardware.group_size ? s.hardware.group_size : std::max<size_t>(np, 1);
        const bool one_group = np <= group;

        TaskResults r;
        r.n_potentials = static_cast<int32_t>(np);
        r.level_count = level_count;
        r.levels.assign(np * level_count, std::nan(""));
        r.level_status.assign(np * level_count, EPSEON_LEVEL_UNBOUND);
        r.iterations.assign(np * level_count, 0);
        r.brackets.assign(np * level_count, epseon_bracket{});
        r.strides.assign(np, 0);
        r.potential_flags.assign(np, 0);
        r.point_counts.assign(np, 0);
        r.energy_offsets.assign(np + 1, 0);
        for (size_t p = 0; p < np; ++p) {
            r.point_counts[p] = point_count(s.potentials[p]);
            r.energy_offsets[p + 1] =
                r.energy_offsets[p] + (s.options.trial_energy_counts.empty() ? default_m
                                                                             : s.options.trial_energy_counts[p]);
        }
        r.has_eigenfunctions = s.options.eigenfunctions;
        if (r.has_eigenfunctions)
            r.eigenfunctions.assign(np, nullptr);
        std::shared_ptr<ScanArrays> scan;
        if (!one_group) {
            scan = std::make_shared<ScanArrays>();
            const size_t ne = static_cast<size_t>(r.energy_offsets.back());
            scan->nodes.resize(ne);
            scan->far_end.resize(ne);
            scan->rescales.resize(ne);
        }

        set_status("opening device");
        epseon_ctx* raw = nullptr;
        check_capi(epseon_ctx_create(device_, s.hardware.allocation_block_size, &raw),
                   "epseon_ctx_create");
        CtxPtr ctx(raw, epseon_ctx_destroy);

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
            tabulate(g0, g1, offsets, v, st);
            poll(st);
            for (size_t p = g0; p < g1; ++p) {
                const PotentialSpec& ps = s.potentials[p];
                epseon_potential_desc& d = desc[p - g0];
                d.n_points = point_count(ps);
                d.n_energies = r.energy_offsets[p + 1] - r.energy_offsets[p];
                d.r_min = ps.min_r;
                d.r_max = ps.max_r;
                d.reduced_mass = s.options.reduced_masses.empty() ? mu : s.options.reduced_masses[p];
                d.e_top = window_top(ps, s.vibwa.min_distance_to_asymptote);
                d.level_min = static_cast<int32_t>(s.vibwa.min_level);
                d.level_count = level_count;
            }
            set_status("uploading potentials");
            check_capi(epseon_load_potentials(ctx.get(), static_cast<int32_t>(g1 - g0), desc.data(),
                                              v.data(), offsets.data()),
                       "epseon_load_potentials");
            poll(st);
            set_status("solving: scan, brackets, refinement");
            check_capi(epseon_solve(ctx.get(), s.options.rel_tol, s.options.max_iterations), "epseon_solve");
            set_status("reading results");
            const size_t l0 = g0 * static_cast<size_t>(level_count);
            check_capi(epseon_get_levels(ctx.get(), r.levels.data() + l0, r.level_status.data() + l0,
                                         r.iterations.data() + l0),
                       "epseon_get_levels");
            check_capi(epseon_get_brackets(ctx.get(), r.brackets.data() + l0), "epseon_get_brackets");
            for (size_t p = g0; p < g1; ++p) {
                epseon_potential_info info{};
                check_capi(epseon_potential_info_get(ctx.get(), static_cast<int32_t>(p - g0), &info),
                           "epseon_potential_info_get");
                r.strides[p] = info.stride;
                r.potential_flags[p] = info.flags;
            }
            epseon_stats stats{};
            check_capi(epseon_get_stats(ctx.get(), &stats), "epseon_get_stats");
            r.solve_ms += stats.solve_ms;
            r.scan_point_updates += stats.scan_point_updates;
            r.refine_point_updates += stats.refine_point_updates;
            if (one_group)
                break;
            // Several groups: this group's arrays are read before the next load.
            const size_t e0 = static_cast<size_t>(r.energy_offsets[g0]);
            check_capi(epseon_get_scan(ctx.get(), scan->nodes.data() + e0, scan->far_end.data() + e0,
                                       scan->rescales.data() + e0),
                       "epseon_get_scan");
            if (s.options.eigenfunctions) {
                poll(st);
                set_status("normalising eigenfunctions");
                for (size_t p = g0; p < g1; ++p) {
                    auto rows = std::make_shared<std::vector<double>>(static_cast<size_t>(level_count) *
                                                                      static_cast<size_t>(r.point_counts[p]));
                    if (!rows->empty())
                        check_capi(epseon_get_wavefunctions(ctx.get(), static_cast<int32_t>(p - g0), rows->data(),
                                                            EPSEON_MEM_HOST),
                                   "epseon_get_wavefunctions");
                    r.eigenfunctions[p] = std::move(rows);
                }
            }
        }
        if (one_group && s.options.eigenfunctions && np > 0) {
            // Normalise now, so that reading the eigenfunctions later is a copy.
            poll(st);
            set_status("normalising eigenfunctions");
            check_capi(epseon_normalize(ctx.get()), "epseon_normalize");
            check_capi(epseon_ctx_synchronize(ctx.get()), "epseon_ctx_synchronize");
        }
        r.scan = scan;
        r.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start)
                         .count();
        results_ = std::move(r);
        // A one-group task keeps its context: scan arrays and eigenfunctions
        // are read from the device when first requested.
        if (one_group)
            ctx_ = std::move(ctx);
    }

    // V_i on each grid, host FP64 (std::exp), split over host threads for
    // large batches; the workers stop early when a stop is requested.
    void tabulate(size_t g0, size_t g1, const std::vector<int64_t>& offsets, std::vector<double>& v,
                  const std::stop_token& st)
    {
        const size_t n = g1 - g0;
        const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        const unsigned workers = static_cast<unsigned>(std::min<size_t>(hw, n));
        auto body = [&](unsigned w) {
            for (size_t q = w; q < n; q += workers) {
                if (st.stop_requested())
                    return;
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
    CtxPtr ctx_{nullptr, epseon_ctx_destroy};
    std::mutex fetch_mutex_;
    mutable std::mutex status_mutex_;
    std::string status_ = "pending";
    std::mutex join_mutex_;
    std::jthread worker_;
};

// ------------------------------------------------------ python wrappers

// Read-only numpy view of `count` elements at `offset` of a shared buffer,
// shaped `shape`; the array keeps the buffer alive.
template <typename T>
py::array_t<T> view_of(const std::shared_ptr<const std::vector<T>>& owner, size_t offset,
                       const std::vector<py::ssize_t>& shape)
{
    auto* keep = new std::shared_ptr<const std::vector<T>>(owner);
    py::capsule base(keep, [](void* q) { delete static_cast<std::shared_ptr<const std::vector<T>>*>(q); });
    py::array_t<T> a(shape, owner->data() + offset, base);
    a.attr("setflags")(py::arg("write") = false);
    return a;
}

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
        const auto sa = task_->scan_arrays();
        return ragged(std::shared_ptr<const std::vector<int32_t>>(sa, &sa->nodes));
    }
    std::vector<py::array_t<double>> get_far_end_values() const
    {
        const auto sa = task_->scan_arrays();
        return ragged(std::shared_ptr<const std::vector<double>>(sa, &sa->far_end));
    }
    std::vector<py::array_t<int32_t>> get_rescale_counts() const
    {
        const auto sa = task_->scan_arrays();
        return ragged(std::shared_ptr<const std::vector<int32_t>>(sa, &sa->rescales));
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

    // One potential's eigenfunctions, (levels, points).  Float64 handles get
    // a read-only view of the task's copy; float32 handles a rounded copy.
    py::array_t<FP> wavefunctions_of(size_t p) const
    {
        const TaskResults& r = task_->results();
        const EigenRows rows = task_->eigenfunctions(p);
        const std::vector<py::ssize_t> shape{static_cast<py::ssize_t>(r.level_count),
                                             static_cast<py::ssize_t>(r.point_counts[p])};
        if constexpr (std::is_same_v<FP, double>) {
            return view_of<double>(rows, 0, shape);
        } else {
            py::array_t<FP> a(shape);
            FP* d = a.mutable_data();
            for (size_t i = 0; i < rows->size(); ++i)
                d[i] = static_cast<FP>((*rows)[i]);
            return a;
        }
    }

    py::object get_wavefunctions(std::optional<int64_t> potential) const
    {
        const TaskResults& r = task_->results();
        if (!r.has_eigenfunctions)
            throw std::runtime_error(
                "Eigenfunctions were not requested: call set_eigenfunction_output(True) before submit_task.");
        if (potential.has_value()) {
            if (*potential < 0 || *potential >= r.n_potentials)
                throw py::index_error("potential index out of range");
            return wavefunctions_of(static_cast<size_t>(*potential));
        }
        py::list out;
        for (size_t p = 0; p < static_cast<size_t>(r.n_potentials); ++p)
            out.append(wavefunctions_of(p));
        return out;
    }

    py::dict get_timings() const
    {
        const TaskResults& r = task_->results();
        py::dict d;
        d["solve_ms"] = r.solve_ms;
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

    // One read-only view per potential into the task's scan array.
    template <typename T>
    std::vector<py::array_t<T>> ragged(const std::shared_ptr<const std::vector<T>>& all) const
    {
        const TaskResults& r = task_->results();
        std::vector<py::array_t<T>> out;
        out.reserve(static_cast<size_t>(r.n_potentials));
        for (size_t p = 0; p < static_cast<size_t>(r.n_potentials); ++p) {
            const int64_t e0 = r.energy_offsets[p];
            out.push_back(view_of<T>(all, static_cast<size_t>(e0), {static_cast<py::ssize_t>(r.energy_offsets[p + 1] - e0)}));
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

    // Additive: potentials read from files (reference: the configurator's
    // setPotentialSource with a PotentialFileLoader, task_configurator.hpp:
    // 98-101).  The files are read now, so a bad file raises here.
    TaskConfigurator& set_potential_file_loader(const PotentialFileLoader& loader)
    {
        std::vector<TabulatedPotentialConfig> v;
        for (LoadedPotential& p : loader.read())
            v.push_back(TabulatedPotentialConfig{std::move(p.v), p.r.front(), p.r.back(), p.asymptote});
        return set_tabulated_potential(v);
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
        .def("set_potential_file_loader", &C::set_potential_file_loader, py::arg("loader"),
             "Set potentials read from the loader's files (text \"r V\" columns or (N, 2) float64 .npy).",
             py::return_value_policy::reference)
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
        .def("get_wavefunctions", &H::get_wavefunctions, py::arg("potential") = py::none(),
             "Normalised eigenfunctions: a (levels, points) array for `potential`, or one per potential "
             "when it is None; zero rows for absent levels.  Read from the device on first request.");
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

    py::class_<PotentialFileLoader>(m, "PotentialFileLoader",
                                    "Potential source reading one tabulated V(r) per file.")
        .def(py::init<>())
        .def(py::init<std::vector<std::string>>(), py::arg("file_names"))
        .def_property_readonly("file_names", &PotentialFileLoader::file_names)
        .def("equals", &PotentialFileLoader::equals, py::arg("other"))
        .def("__eq__", &PotentialFileLoader::equals)
        .def(
            "get_potential_data",
            [](const PotentialFileLoader& l) {
                std::vector<py::array_t<double>> out;
                for (const std::vector<double>& v : l.get_potential_data()) {
                    py::array_t<double> a(static_cast<py::ssize_t>(v.size()));
                    std::copy(v.begin(), v.end(), a.mutable_data());
                    out.push_back(std::move(a));
                }
                return out;
            },
            "V values of every file, one array per file.");

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
