/*
 * epseon_hip.h -- C ABI of the MI355X Numerov-shooting library (libepseon_hip.so).
 *
 * The reference hosts this path behind its pybind11 module _libepseon_gpu
 * (cpp/gpu/source/epseon/gpu/python/api.cpp:114-494) and the plugin call
 * Algorithm<FP>::run(stop_token, TaskHandle*) (cpp/gpu/include/epseon/gpu/
 * algorithms/algorithm.hpp:34), whose only implementation,
 * VibwaAlgorithm<FP>::run (algorithms/vibwa.hpp:605-637), creates a device,
 * allocates buffers and descriptor sets and returns without dispatching.
 * Every entry point below replaces one step of that pipeline; the comment on
 * each names the reference code it stands in for.
 *
 * Conventions: plain C types only; every call returns EPSEON_OK (0) or an
 * error code, and epseon_last_error() gives the message of the last failing
 * call on the calling thread.  Host pointers everywhere unless a name ends in
 * _device.  A context owns one HIP stream and one device arena; calls on one
 * context must not overlap (the pybind layer runs one context per task).
 *
 * Units: cm^-1 for energies, Angstrom for distances, amu for masses.
 */
#ifndef EPSEON_HIP_H
#define EPSEON_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define EPSEON_OK 0
#define EPSEON_ERR_INVALID 1   /* bad argument */
#define EPSEON_ERR_DEVICE 2    /* HIP runtime error or no such device */
#define EPSEON_ERR_STATE 3     /* call out of order (e.g. refine before scan) */
#define EPSEON_ERR_MEMORY 4    /* device allocation failed */

/* ---- per-level status ------------------------------------------------ */
#define EPSEON_LEVEL_OK 0         /* converged to the requested tolerance */
#define EPSEON_LEVEL_UNBOUND 1    /* no level with this node count below e_top */
#define EPSEON_LEVEL_MAXITER 2    /* iteration cap reached; value is the bracket midpoint */
#define EPSEON_LEVEL_UNRESOLVED 3 /* potential rejected: grid too coarse for the window */

/* ---- refinement modes (epseon_set_refine_mode) ------------------------- */
#define EPSEON_REFINE_AUTO 0       /* SEGMENTED with windows in use; else LANE for >= 131072 levels */
#define EPSEON_REFINE_LANE 1       /* one lane per level, sequential shots */
#define EPSEON_REFINE_SEGMENTED 2  /* one wavefront per level, shots split over 64 grid segments */

/* ---- memory kinds of output/input pointers ------------------------------ */
#define EPSEON_MEM_HOST 0          /* pageable or pinned host memory */
#define EPSEON_MEM_DEVICE 1        /* device memory of the context's device */

/* ---- per-potential flags ----------------------------------------------- */
#define EPSEON_POT_OK 0
#define EPSEON_POT_UNRESOLVED 1   /* 1 + beta (E_top - V_min) > 3/2, or fewer than 4 points left */

/* Device description.  Replaces the Vulkan property structs the reference
 * exposes through ComputeContext::getPhysicalDevicesInfo
 * (cpp/gpu/source/epseon/gpu/compute_context.cpp:94-104; bound at
 * python/api.cpp:118-329). */
typedef struct epseon_device_info {
    char name[256];
    char arch[64];                     /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
    int32_t runtime_version;           /* hipRuntimeGetVersion */
    int32_t driver_version;            /* hipDriverGetVersion */
    int32_t vendor_id;                 /* 0x1002 */
    int32_t pci_device_id;             /* PCI device number of the card */
    int32_t compute_units;
    int32_t max_threads_per_block;
    int32_t max_block_dim[3];
    int32_t max_grid_dim[3];
    int64_t shared_mem_per_block;      /* bytes */
    int64_t shared_mem_per_cu;         /* bytes */
    int64_t total_global_mem;          /* bytes */
    int32_t clock_khz;
    int32_t warp_size;
    uint8_t uuid[16];
} epseon_device_info;

/* One potential of a batch.  Replaces the per-slot inputs of the reference:
 * the Morse record's grid (MorsePotentialConfig min_r/max_r/point_count,
 * task_configurator/potential_source.hpp:102-183) and the Vibwa record's
 * masses, asymptote margin and level range (task_configurator/
 * algorithm_config.hpp:75-99). */
typedef struct epseon_potential_desc {
    int64_t n_points;      /* N >= 4; grid r_i = r_min + i*h, h = (r_max - r_min)/(N - 1) */
    int64_t n_energies;    /* M >= 1 trial energies, E_j = E_lo + (j+1)(e_top - E_lo)/(M+1) */
    double r_min;
    double r_max;
    double reduced_mass;   /* mu = m0*m1/(m0+m1), amu */
    double e_top;          /* top of the energy window; E_lo is min_i V_i */
    int32_t level_min;     /* first level (node count) to refine */
    int32_t level_count;   /* number of levels; the reference's max_level - min_level + 1
                              (algorithm_config.hpp:177) */
} epseon_potential_desc;

/* Per-potential values the library derived at load time. */
typedef struct epseon_potential_info {
    double beta;           /* h^2 mu / (12 K) */
    double h;              /* grid step */
    double e_lo;           /* window floor = min_i V_i */
    double e_hi;           /* window top = e_top */
    int32_t stride;        /* node-sampling stride B */
    int32_t flags;         /* EPSEON_POT_* */
    int64_t first_point;   /* left Dirichlet point of the trimmed grid (index on the full grid) */
    int64_t point_count;   /* points of the trimmed grid, both Dirichlet points included */
    int32_t far_exponent;  /* scan far ends: w_{N-1} = far_end * 2^(far_exponent - 480 rescales) */
    int32_t reserved;
} epseon_potential_info;

/* Counters and device times of the last calls on a context. */
typedef struct epseon_stats {
    double scan_ms;               /* device time of the last epseon_scan (HIP events) */
    double refine_ms;             /* device time of the last epseon_refine */
    double normalize_ms;          /* device time of the last epseon_normalize */
    double scan_point_updates;    /* recurrence steps of the last scan: sum over trial energies of the
                                     points of their shot's window - 2 (N - 2 without windows) */
    double refine_point_updates;  /* sum over levels of shots * (window points - 2) */
    double normalize_point_updates; /* sum over converged levels of (N - 2) */
    int64_t scan_launches;        /* kernel launches of the last scan */
    int64_t levels_converged;     /* levels with EPSEON_LEVEL_OK after the last refine */
    int32_t refine_mode;          /* EPSEON_REFINE_LANE or _SEGMENTED used by the last refine */
    int32_t refine_lanes;         /* threads per level of the last segmented refine (1: lane mode) */
    double vjp_ms;                /* device time of the last epseon_level_vjp */
    double vjp_point_updates;     /* sum over converged levels of (N - 2) per adjoint pass */
    double solve_ms;              /* device time of the last epseon_solve (both streams) */
    double solve_split;           /* scan-work fraction of the first pipeline group (0: sequential) */
    int32_t scan_windows;         /* 1 when the last scan (and its refinement) used band windows */
    int32_t solve_graph;          /* 1 when the last epseon_solve replayed a captured HIP graph */
    int32_t refine_levels_per_block; /* levels sharing a block in the last segmented refine (1 or 2) */
    int32_t reserved;
} epseon_stats;

typedef struct epseon_ctx epseon_ctx;

/* Message of the last failing call on this thread ("" if none). */
const char* epseon_last_error(void);

/* sizeof of the public structs, in this order: epseon_device_info,
 * epseon_potential_desc, epseon_potential_info, epseon_stats, epseon_tuning,
 * epseon_bracket (at most
 * `count` values are written).  Lets a foreign-function binding verify its
 * struct layouts at load time.  No counterpart in the reference (its binding
 * is compiled against the headers). */
int epseon_struct_sizes(int64_t* sizes, int32_t count);

/* Number of visible HIP devices.  Replaces the physical-device enumeration of
 * ComputeContext::getPhysicalDevicesInfo (compute_context.cpp:94-104). */
int epseon_device_count(int32_t* count);

/* Properties of device `device` (an ordinal).  Same source as above. */
int epseon_device_info_get(int32_t device, epseon_device_info* info);

/* Open a context on device `device`: one HIP stream plus a device arena that
 * allocates in blocks of `block_bytes` (0 = 64 MiB).  Replaces the logical
 * device and VMA allocator VibwaAlgorithm::run builds per task
 * (vibwa.hpp:611 and :654-674, :621-626), sized by HardwareConfig's
 * allocation_block_size (task_configurator/hardware_config.hpp:11-76). */
int epseon_ctx_create(int32_t device, uint64_t block_bytes, epseon_ctx** out);

/* Release the context, its stream and all device memory. */
int epseon_ctx_destroy(epseon_ctx* ctx);

/* Upload a batch of `count` potentials: `v` holds the concatenated V tables
 * (cm^-1), potential p at v[v_offsets[p] .. v_offsets[p] + n_points).  The
 * library derives beta, the window and the node stride, moves the Dirichlet
 * points in from the walls to the last/first points where V >= V_min +
 * 1/(2 beta) (so 1 + beta (E - V_i) > 1/2 on the grid that is shot), rejects
 * potentials with 1 + beta (e_top - V_min) > 3/2 (flag EPSEON_POT_UNRESOLVED), builds the
 * shifted table u_i = beta V_i - 1 on the device and lays out every result
 * buffer.  Replaces the staging upload plus the five device-only buffers and
 * the output buffer per slot (ShaderResources::create, vibwa.hpp:153-233;
 * layout from getShaderBufferRequirements, algorithm_config.hpp:173-198),
 * and MorsePotentialGenerator::get_potential_data, which returns {} in the
 * reference (potential_source.hpp:223-225). */
int epseon_load_potentials(epseon_ctx* ctx, int32_t count, const epseon_potential_desc* desc,
                           const double* v, const int64_t* v_offsets);

/* Same as epseon_load_potentials with V already in device memory of the
 * context's device (`v_device`, e.g. a framework tensor): min V, the wall
 * trimming and the finiteness check run on the device with the host loader's
 * exact semantics (first index of the minimum, same threshold arithmetic), so
 * results are identical to a host upload of the same values.  `v_offsets` is
 * a host array.  V is read only during the call.  Same counterpart as
 * epseon_load_potentials; the reference stages every potential through a
 * host-visible buffer (vibwa.hpp:176-187). */
int epseon_load_potentials_device(epseon_ctx* ctx, int32_t count, const epseon_potential_desc* desc,
                                  const double* v_device, const int64_t* v_offsets);

/* Derived values of potential p after epseon_load_potentials.  far_exponent
 * is that of whole-grid scans; windowed scans: epseon_band_windows_get. */
int epseon_potential_info_get(epseon_ctx* ctx, int32_t p, epseon_potential_info* info);

/* Energy bands per potential: the window [e_lo, e_hi] splits into bands with
 * tops E_b = fma(b + 1, (e_hi - e_lo) / EPSEON_BANDS, e_lo). */
#define EPSEON_BANDS 16

/* Shot windows of potential p for the last scan (or, before any scan, for
 * the current tuning): for each band b, the Dirichlet points lo[b] and hi[b]
 * on the trimmed grid and the exponent of the scan's far ends in it (a trial
 * energy E_j of the scan group j / 64 whose highest energy lies in band b has
 * w = far_end * 2^(far_exponent[b] - 480 rescales)).  Windowed: hi[b] is the
 * first point past the last one with V <= E_b where the sum of the growth
 * bound g(x) = 2^(floor((k-1)/2) - 1), x = 12 beta (V_i - E_b) = m 2^k with m
 * in [1/2, 1), reaches 20 (else the grid end); lo[b] likewise on the left.
 * Whole grid: lo = 0, hi = points - 1.  Each array holds EPSEON_BANDS
 * entries.  No counterpart in the reference, whose shots span the grid. */
int epseon_band_windows_get(epseon_ctx* ctx, int32_t p, int32_t* lo, int32_t* hi, int32_t* far_exponent);

/* Numerov scan: shoot every trial energy of every potential, recording node
 * count, far-end mantissa and scale count (w_{N-1} = far_end *
 * 2^(far_exponent - 480 rescales), far_exponent from epseon_potential_info_get).  Stream-ordered; returns
 * after the launch (epseon_ctx_synchronize waits).  This is the dispatch that
 * would sit between descriptor update and "Thread finished" in
 * VibwaAlgorithm::run (vibwa.hpp:630-636). */
int epseon_scan(epseon_ctx* ctx);

/* Bracket every requested level between adjacent trial energies by node count
 * and refine it (count-interpolated node-count cuts until isolated, then regula
 * falsi with Anderson-Bjorck scaling on the far-end value) until the bracket is narrower than
 * rel_tol*|E| + abs_tol, abs_tol = 1e-15 (e_top - E_lo).  Refinement shots
 * use the difference form of the recurrence (on beta V, carrying
 * beta (E - V) at full relative precision), in the mode chosen with
 * epseon_set_refine_mode.  Requires a prior epseon_scan.  Same dispatch site
 * as above; its result is the reference's output buffer of level_count
 * values per slot (algorithm_config.hpp:189). */
int epseon_refine(epseon_ctx* ctx, double rel_tol, int32_t max_iter);

/* Select the refinement mode (EPSEON_REFINE_*).  No counterpart in the
 * reference, whose algorithm record has no such knob. */
int epseon_set_refine_mode(epseon_ctx* ctx, int32_t mode);

/* Scan and refine (same results as epseon_scan followed by epseon_refine).
 * Pipelined over two streams when the split fraction is above 0: the
 * potentials with the longest grids, holding up to that fraction of the scan
 * work, are scanned and refined on one stream while the rest are scanned and
 * refined on the other, so the longest refinement shots overlap scan work.
 * epseon_get_stats reports solve_ms for the whole solve; scan_ms and
 * refine_ms stay 0 (time the phases with epseon_scan and epseon_refine).
 * Same dispatch site as epseon_scan. */
int epseon_solve(epseon_ctx* ctx, double rel_tol, int32_t max_iter);

/* Split fraction of the pipelined epseon_solve, in [0, 1) (default 0: scan
 * then refine).  No counterpart in the reference
 * (its HardwareConfig, hardware_config.hpp:11-76, sizes buffers only). */
int epseon_set_solve_split(epseon_ctx* ctx, double fraction);

/* Launch configuration of a context.  Every field only changes speed, except
 * refine_mode, segmented_lanes and windows, which change the levels within
 * the refinement tolerance (lane and segmented shots round differently;
 * windowed shots stop at other Dirichlet points) and, for windows, the scan's
 * far-end values (node counts stay the same). */
typedef struct epseon_tuning {
    int32_t refine_mode;          /* EPSEON_REFINE_* */
    int32_t levels_per_lane;      /* lane mode: 1, or 2 for potentials with more than 64 levels */
    int32_t scan_waves_per_block; /* wavefronts per scan workgroup, 1..4 */
    int32_t segmented_lanes;      /* segmented mode: 0 (from the batch size), 64, 128 or 256 */
    int32_t xcd_scan;             /* 1: whole-batch scans and segmented refinements bin potentials per
                                     XCD (default), 0: in order */
    int32_t windows;              /* 1: shots stop 15 guaranteed e-folds past the outermost turning
                                     points of their energy band's top (EPSEON_BANDS bands per
                                     potential); 0: every shot spans the trimmed grid; 2 (default):
                                     1 when the scan has at least 8 wavefronts (64 trial energies
                                     each) per SIMD of the device, else 0.  Lane-mode refinement
                                     always shoots the whole grid, so with windows in use AUTO
                                     selects SEGMENTED and a windowed scan is refined segmented. */
    int32_t graph;                /* 1 (default): when a solve repeats the previous solve's batch,
                                     tuning, rel_tol and max_iter, epseon_solve records its launches
                                     (scan, brackets, refinement) as a HIP graph, and replays the
                                     graph while they stay the same (the first solve after a load
                                     launches directly); 0: every call launches the kernels one by
                                     one.  Pipelined (solve_split > 0) and paired-lane solves
                                     always launch directly.  Results are bit-identical either
                                     way. */
    int32_t segmented_levels;     /* segmented mode with 64-lane blocks: levels per block.  4 or 2:
                                     that many consecutive levels of a potential share a block, and
                                     those that shoot in the same window in a round walk it
                                     together, every grid value loaded once for all of them; 1: one
                                     level per block; 0 (default): 4 when the batch has at least 96
                                     levels per compute unit, 2 from 24, else 1.  Results are
                                     bit-identical either way. */
    double solve_split;           /* see epseon_set_solve_split */
} epseon_tuning;

#define EPSEON_TUNE_KEEP_RESULTS 1 /* autotune only knobs that leave results bit-identical */

/* Current tuning of the context. */
int epseon_get_tuning(epseon_ctx* ctx, epseon_tuning* tuning);

/* Set all knobs at once (validated). */
int epseon_set_tuning(epseon_ctx* ctx, const epseon_tuning* tuning);

/* Time candidate tunings on the loaded batch (each a full epseon_solve, best
 * of two after a warm-up run, device time by HIP events), keep the fastest
 * and write it to `best` (may be NULL).  Coordinate search over scan
 * workgroup size and XCD binning, shot windows, refinement mode, levels per lane / segmented
 * block size, the pipeline split and graph replay (an automatic choice is first resolved to
 * its current value).  With EPSEON_TUNE_KEEP_RESULTS the windows, the
 * refinement mode and the segmented block size stay fixed.  The levels and scan results afterwards are
 * those of the last candidate run: call epseon_solve again.  No counterpart
 * in the reference, whose HardwareConfig (hardware_config.hpp:11-76) is set
 * by hand. */
int epseon_autotune(epseon_ctx* ctx, double rel_tol, int32_t max_iter, int32_t flags,
                    epseon_tuning* best);

/* Wait for all work queued on the context's stream. */
int epseon_ctx_synchronize(epseon_ctx* ctx);

/* Eigenfunction normalisation of every converged level: two inward shots at
 * the refined E (difference form of the Numerov recurrence on y), matched by
 * least squares at the point of min V and normalised to sum_i psi_i^2 h = 1,
 * psi > 0 next to the left wall.  Stream-ordered.  epseon_get_wavefunctions
 * and epseon_level_vjp run it when the levels changed since the last call.
 * The reference has no eigenfunction step; it is the "normalise psi" stage
 * its algorithm record is configured for (algorithm_config.hpp:75-199) but
 * never runs (vibwa.hpp:605-637). */
int epseon_normalize(epseon_ctx* ctx);

/* Normalised eigenfunctions of potential `potential`: level_count rows of
 * n_points values (the full grid passed at load), row l = psi of level
 * level_min + l, zero outside the trimmed window and for levels whose status
 * is not EPSEON_LEVEL_OK.  `mem` says whether `psi` is host or device memory.
 * Synchronises the stream.  The per-slot output buffer of the reference
 * (algorithm_config.hpp:189) holds one value per level; this is the
 * eigenfunction output that layout has no room for. */
int epseon_get_wavefunctions(epseon_ctx* ctx, int32_t potential, double* psi, int32_t mem);

/* Eigenfunctions of every potential in one launch: potential p's
 * level_count x n_points block starts at element sum_{q<p} level_count_q *
 * n_points_q (layout as epseon_get_wavefunctions, potentials in load order;
 * total from epseon_get_eigenfunction_size).  Same counterpart as above. */
int epseon_get_all_wavefunctions(epseon_ctx* ctx, double* psi, int32_t mem);

/* Number of doubles epseon_get_all_wavefunctions writes. */
int epseon_get_eigenfunction_size(epseon_ctx* ctx, int64_t* elements);

/* Adjoint of the levels with respect to the tabulated potential (Hellmann-
 * Feynman, exact for the discrete Numerov eigenproblem):
 *   grad_v[v_offsets[p] + i] = sum_l grad_levels[level_offset(p) + l] * psi_{p,l}(i)^2 * h_p
 * over the converged levels l of each potential p; every element of
 * [v_offsets[p], v_offsets[p] + n_points) is written (zero outside the
 * trimmed window) and nothing else.  grad_levels has the layout of
 * epseon_get_levels; grad_v the layout of the V batch passed to
 * epseon_load_potentials.  Both are host or both device memory (`mem`).
 * Eigenfunctions are recomputed, never stored.  Synchronises the stream.
 * No counterpart in the reference (it has no gradient path). */
int epseon_level_vjp(epseon_ctx* ctx, const double* grad_levels, double* grad_v, int32_t mem);

/* Maximum scratch of epseon_level_vjp_noalloc, bytes per device. */
#define EPSEON_NOALLOC_SCRATCH_BYTES (64u << 20)

/* epseon_level_vjp for device pointers that allocates nothing: no hipMalloc,
 * hipFree or other allocating call, on the device or the host.  The
 * normalisation pass's segment maps and shot checkpoints (80 + 32 bytes per
 * eigenfunction-shot segment of every level, segments from
 * epseon_potential_info point counts) live in a library-owned device array of
 * EPSEON_NOALLOC_SCRATCH_BYTES per device, held until the process ends;
 * batches that need more are rejected with EPSEON_ERR_INVALID.  The
 * normalisation always runs (the array keeps the last call's contents), the
 * context's own eigenfunction state is left as it was, and the result equals
 * epseon_level_vjp(ctx, grad_levels_device, grad_v_device, EPSEON_MEM_DEVICE)
 * bit for bit.  Every call on a device shares that device's array: calls
 * that overlap on one device (from several threads or contexts) race.
 * Synchronises the stream.  No counterpart in the reference; it is for
 * callers that cannot allocate device memory (serving loops, code that must
 * not allocate between set-up and teardown). */
int epseon_level_vjp_noalloc(epseon_ctx* ctx, const double* grad_levels_device, double* grad_v_device);

/* Copy the refined levels out, in potential order, level_count per potential
 * (the reference's per-slot output buffer, algorithm_config.hpp:189).  Any of
 * the pointers may be NULL.  Synchronises the stream. */
int epseon_get_levels(epseon_ctx* ctx, double* levels, int32_t* status, int32_t* iterations);

/* Copy the scan results out, in potential order, n_energies per potential.
 * Any of the pointers may be NULL.  Synchronises the stream. */
int epseon_get_scan(epseon_ctx* ctx, int32_t* nodes, double* far_end, int32_t* rescales);

/* Bracket of one level as the refinement started from it. */
typedef struct epseon_bracket {
    double e_lo, e_hi;     /* ends (cm^-1): the last trial energy with at most v nodes (or the
                              window floor min V when lo_valid is 0) and the first with more */
    int32_t nodes_lo;      /* node count of the scan shot at e_lo (0 at the floor) */
    int32_t nodes_hi;      /* node count of the scan shot at e_hi */
    int32_t lo_valid;      /* 1 when e_lo is a trial energy */
    int32_t found;         /* 0: no trial energy has more than v nodes (the level lies above the
                              window top) or the potential is unresolved; the rest is then unset */
    int32_t band;          /* energy band whose shot window the refinement used */
    int32_t reserved;
} epseon_bracket;

/* Brackets of the last refinement (its bracket pass), in level order, the
 * potentials' level_count each.  The bracketing stage has no code in the
 * reference; it would sit in VibwaAlgorithm<FP>::run (vibwa.hpp:605-637).
 * Synchronises the stream. */
int epseon_get_brackets(epseon_ctx* ctx, epseon_bracket* out);

/* Totals: number of potentials, grid points, trial energies and levels of the
 * loaded batch (any pointer may be NULL). */
int epseon_get_sizes(epseon_ctx* ctx, int32_t* potentials, int64_t* points, int64_t* energies,
                     int64_t* levels);

/* Counters and device times of the last phases. */
int epseon_get_stats(epseon_ctx* ctx, epseon_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* EPSEON_HIP_H */
