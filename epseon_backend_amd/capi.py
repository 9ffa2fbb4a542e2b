"""ctypes binding of the C ABI in ``include/epseon_hip.h`` (libepseon_hip.so).

This is the thin Python-side binding a maintainer would add next to the
reference's pybind11 module (``cpp/gpu/source/epseon/gpu/python/api.cpp``);
the pybind11 module ``_libepseon_gpu`` in this package wraps the same library
for the reference's builder/handle API.  Arrays cross the boundary as numpy
buffers (host memory); device memory stays inside the library's arena.

The library is loaded from this package's ``lib/`` directory.  If it is
missing the import fails with an error naming the build step -- there is no
fallback implementation.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from pathlib import Path
from typing import Sequence

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libepseon_hip.so"

EPSEON_OK = 0
EPSEON_ERR_INVALID = 1
EPSEON_ERR_DEVICE = 2
EPSEON_ERR_STATE = 3
EPSEON_ERR_MEMORY = 4

LEVEL_OK = 0
LEVEL_UNBOUND = 1
LEVEL_MAXITER = 2
LEVEL_UNRESOLVED = 3

POT_OK = 0
POT_UNRESOLVED = 1

REFINE_AUTO = 0
REFINE_LANE = 1
REFINE_SEGMENTED = 2

MEM_HOST = 0
MEM_DEVICE = 1


class EpseonError(RuntimeError):
    """Error reported by libepseon_hip (carries the C status code)."""

    def __init__(self, code: int, message: str) -> None:
        super().__init__(message)
        self.code = code


class DeviceInfo(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 256),
        ("arch", ctypes.c_char * 64),
        ("runtime_version", ctypes.c_int32),
        ("driver_version", ctypes.c_int32),
        ("vendor_id", ctypes.c_int32),
        ("pci_device_id", ctypes.c_int32),
        ("compute_units", ctypes.c_int32),
        ("max_threads_per_block", ctypes.c_int32),
        ("max_block_dim", ctypes.c_int32 * 3),
        ("max_grid_dim", ctypes.c_int32 * 3),
        ("shared_mem_per_block", ctypes.c_int64),
        ("shared_mem_per_cu", ctypes.c_int64),
        ("total_global_mem", ctypes.c_int64),
        ("clock_khz", ctypes.c_int32),
        ("warp_size", ctypes.c_int32),
        ("uuid", ctypes.c_uint8 * 16),
    ]


class PotentialDesc(ctypes.Structure):
    _fields_ = [
        ("n_points", ctypes.c_int64),
        ("n_energies", ctypes.c_int64),
        ("r_min", ctypes.c_double),
        ("r_max", ctypes.c_double),
        ("reduced_mass", ctypes.c_double),
        ("e_top", ctypes.c_double),
        ("level_min", ctypes.c_int32),
        ("level_count", ctypes.c_int32),
    ]


class PotentialInfo(ctypes.Structure):
    _fields_ = [
        ("beta", ctypes.c_double),
        ("h", ctypes.c_double),
        ("e_lo", ctypes.c_double),
        ("e_hi", ctypes.c_double),
        ("stride", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("first_point", ctypes.c_int64),
        ("point_count", ctypes.c_int64),
        ("far_exponent", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("scan_ms", ctypes.c_double),
        ("refine_ms", ctypes.c_double),
        ("normalize_ms", ctypes.c_double),
        ("scan_point_updates", ctypes.c_double),
        ("refine_point_updates", ctypes.c_double),
        ("normalize_point_updates", ctypes.c_double),
        ("scan_launches", ctypes.c_int64),
        ("levels_converged", ctypes.c_int64),
        ("refine_mode", ctypes.c_int32),
        ("refine_lanes", ctypes.c_int32),
        ("vjp_ms", ctypes.c_double),
        ("vjp_point_updates", ctypes.c_double),
        ("solve_ms", ctypes.c_double),
        ("solve_split", ctypes.c_double),
        ("scan_windows", ctypes.c_int32),
        ("solve_graph", ctypes.c_int32),
        ("refine_levels_per_block", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class Bracket(ctypes.Structure):
    _fields_ = [
        ("e_lo", ctypes.c_double),
        ("e_hi", ctypes.c_double),
        ("nodes_lo", ctypes.c_int32),
        ("nodes_hi", ctypes.c_int32),
        ("lo_valid", ctypes.c_int32),
        ("found", ctypes.c_int32),
        ("band", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


BRACKET_DTYPE = np.dtype([("e_lo", np.float64), ("e_hi", np.float64), ("nodes_lo", np.int32),
                          ("nodes_hi", np.int32), ("lo_valid", np.int32), ("found", np.int32),
                          ("band", np.int32), ("reserved", np.int32)])


class Tuning(ctypes.Structure):
    _fields_ = [
        ("refine_mode", ctypes.c_int32),
        ("levels_per_lane", ctypes.c_int32),
        ("scan_waves_per_block", ctypes.c_int32),
        ("segmented_lanes", ctypes.c_int32),
        ("xcd_scan", ctypes.c_int32),
        ("windows", ctypes.c_int32),
        ("graph", ctypes.c_int32),
        ("segmented_levels", ctypes.c_int32),
        ("solve_split", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


TUNE_KEEP_RESULTS = 1
BANDS = 16  # energy bands of the shot windows (EPSEON_BANDS)

_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_int32_p = ctypes.POINTER(ctypes.c_int32)
_c_int64_p = ctypes.POINTER(ctypes.c_int64)

# Every entry point of include/epseon_hip.h with its argument types.
SIGNATURES: dict[str, tuple[object, list[object]]] = {
    "epseon_last_error": (ctypes.c_char_p, []),
    "epseon_struct_sizes": (ctypes.c_int, [_c_int64_p, ctypes.c_int32]),
    "epseon_device_count": (ctypes.c_int, [_c_int32_p]),
    "epseon_band_windows_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _c_int32_p, _c_int32_p, _c_int32_p]),
    "epseon_device_info_get": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(DeviceInfo)]),
    "epseon_ctx_create": (ctypes.c_int, [ctypes.c_int32, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
    "epseon_ctx_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "epseon_load_potentials": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(PotentialDesc), _c_double_p, _c_int64_p],
    ),
    "epseon_load_potentials_device": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(PotentialDesc), ctypes.c_void_p, _c_int64_p],
    ),
    "epseon_potential_info_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(PotentialInfo)]),
    "epseon_scan": (ctypes.c_int, [ctypes.c_void_p]),
    "epseon_refine": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int32]),
    "epseon_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int32]),
    "epseon_set_refine_mode": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "epseon_set_solve_split": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "epseon_get_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Tuning)]),
    "epseon_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Tuning)]),
    "epseon_autotune": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(Tuning)],
    ),
    "epseon_ctx_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "epseon_get_levels": (ctypes.c_int, [ctypes.c_void_p, _c_double_p, _c_int32_p, _c_int32_p]),
    "epseon_get_scan": (ctypes.c_int, [ctypes.c_void_p, _c_int32_p, _c_double_p, _c_int32_p]),
    "epseon_get_brackets": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "epseon_get_sizes": (ctypes.c_int, [ctypes.c_void_p, _c_int32_p, _c_int64_p, _c_int64_p, _c_int64_p]),
    "epseon_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Stats)]),
    "epseon_normalize": (ctypes.c_int, [ctypes.c_void_p]),
    "epseon_get_wavefunctions": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32]),
    "epseon_level_vjp": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32]),
    "epseon_level_vjp_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "epseon_get_all_wavefunctions": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32]),
    "epseon_get_eigenfunction_size": (ctypes.c_int, [ctypes.c_void_p, _c_int64_p]),
}

_LIB: ctypes.CDLL | None = None


def _share_runtime_with_torch() -> None:
    """Load torch's HIP runtime first when torch is installed.

    torch ships its own libamdhip64.so with the same SONAME as ROCm's.  If
    this library were loaded first, a later ``import torch`` would map a
    second HIP runtime into the process; importing torch first makes both
    resolve to one runtime.
    """
    if os.environ.get("EPSEON_NO_TORCH_PRELOAD"):
        return
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load_library() -> ctypes.CDLL:
    """Load libepseon_hip.so (once) and declare every C entry point."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        msg = (
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python -m epseon_backend_amd._build or __graft_entry__.build())"
        )
        raise ImportError(msg)
    _share_runtime_with_torch()
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (restype, argtypes) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    sizes = (ctypes.c_int64 * len(ABI_STRUCTS))()
    if lib.epseon_struct_sizes(sizes, len(ABI_STRUCTS)) != EPSEON_OK:
        raise ImportError("libepseon_hip.so does not report its struct sizes")
    for cls, size in zip(ABI_STRUCTS, sizes):
        if ctypes.sizeof(cls) != size:
            msg = f"ABI mismatch: {cls.__name__} is {ctypes.sizeof(cls)} bytes here, {size} in the library"
            raise ImportError(msg)
    _LIB = lib
    return lib


# Public structs in the order epseon_struct_sizes reports them.
ABI_STRUCTS = (DeviceInfo, PotentialDesc, PotentialInfo, Stats, Tuning, Bracket)


def _check(rc: int) -> None:
    if rc != EPSEON_OK:
        lib = load_library()
        raise EpseonError(rc, lib.epseon_last_error().decode(errors="replace"))


def device_count() -> int:
    lib = load_library()
    n = ctypes.c_int32(0)
    _check(lib.epseon_device_count(ctypes.byref(n)))
    return int(n.value)


def device_info(device: int) -> DeviceInfo:
    lib = load_library()
    info = DeviceInfo()
    _check(lib.epseon_device_info_get(device, ctypes.byref(info)))
    return info


@dataclass
class Potential:
    """One potential of a batch: its V table on the uniform grid and options."""

    values: np.ndarray
    r_min: float
    r_max: float
    reduced_mass: float
    e_top: float
    n_energies: int
    level_min: int = 0
    level_count: int = 1


def _as_f64(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


class Context:
    """One device context (stream + arena) of libepseon_hip."""

    def __init__(self, device: int = 0, block_bytes: int = 0) -> None:
        self._lib = load_library()
        handle = ctypes.c_void_p()
        _check(self._lib.epseon_ctx_create(device, block_bytes, ctypes.byref(handle)))
        self._h = handle
        self.n_potentials = 0
        self.level_counts: list[int] = []
        self.energy_counts: list[int] = []
        self.point_counts: list[int] = []
        self.v_offsets: list[int] = []

    def close(self) -> None:
        if self._h:
            self._lib.epseon_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self) -> Context:
        return self

    def __exit__(self, *exc: object) -> None:
        self.close()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    def load(self, potentials: Sequence[Potential]) -> None:
        """Upload a batch (epseon_load_potentials)."""
        n = len(potentials)
        descs = (PotentialDesc * max(n, 1))()
        offsets = np.zeros(max(n, 1), dtype=np.int64)
        total = 0
        for i, p in enumerate(potentials):
            offsets[i] = total
            total += len(p.values)
        v = np.zeros(max(total, 1), dtype=np.float64)
        for i, p in enumerate(potentials):
            vals = _as_f64(p.values)
            v[offsets[i] : offsets[i] + len(vals)] = vals
            d = descs[i]
            d.n_points = len(vals)
            d.n_energies = int(p.n_energies)
            d.r_min = float(p.r_min)
            d.r_max = float(p.r_max)
            d.reduced_mass = float(p.reduced_mass)
            d.e_top = float(p.e_top)
            d.level_min = int(p.level_min)
            d.level_count = int(p.level_count)
        _check(
            self._lib.epseon_load_potentials(
                self._h,
                n,
                descs,
                v.ctypes.data_as(_c_double_p),
                offsets.ctypes.data_as(_c_int64_p),
            )
        )
        self.n_potentials = n
        self.level_counts = [int(p.level_count) for p in potentials]
        self.energy_counts = [int(p.n_energies) for p in potentials]
        self.point_counts = [len(p.values) for p in potentials]
        self.v_offsets = [int(o) for o in offsets[:n]]

    def load_device(self, potentials: Sequence[Potential], v_ptr: int, offsets: Sequence[int]) -> None:
        """Upload a batch whose V tables are already in device memory at
        v_ptr (doubles, potential p at offsets[p]; epseon_load_potentials_device).
        ``Potential.values`` may be any sized object: only its length is used."""
        n = len(potentials)
        descs = (PotentialDesc * max(n, 1))()
        offs = np.zeros(max(n, 1), dtype=np.int64)
        for i, p in enumerate(potentials):
            offs[i] = int(offsets[i])
            d = descs[i]
            d.n_points = len(p.values)
            d.n_energies = int(p.n_energies)
            d.r_min = float(p.r_min)
            d.r_max = float(p.r_max)
            d.reduced_mass = float(p.reduced_mass)
            d.e_top = float(p.e_top)
            d.level_min = int(p.level_min)
            d.level_count = int(p.level_count)
        _check(self._lib.epseon_load_potentials_device(self._h, n, descs, ctypes.c_void_p(v_ptr),
                                                       offs.ctypes.data_as(_c_int64_p)))
        self.n_potentials = n
        self.level_counts = [int(p.level_count) for p in potentials]
        self.energy_counts = [int(p.n_energies) for p in potentials]
        self.point_counts = [len(p.values) for p in potentials]
        self.v_offsets = [int(o) for o in offs[:n]]

    def potential_info(self, p: int) -> PotentialInfo:
        info = PotentialInfo()
        _check(self._lib.epseon_potential_info_get(self._h, p, ctypes.byref(info)))
        return info

    def band_windows(self, p: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(lo, hi, far_exponent) of potential p's BANDS shot windows (those
        of the last scan)."""
        lo = np.empty(BANDS, dtype=np.int32)
        hi = np.empty(BANDS, dtype=np.int32)
        fe = np.empty(BANDS, dtype=np.int32)
        _check(self._lib.epseon_band_windows_get(self._h, p, lo.ctypes.data_as(_c_int32_p),
                                                 hi.ctypes.data_as(_c_int32_p), fe.ctypes.data_as(_c_int32_p)))
        return lo, hi, fe

    def scan(self) -> None:
        _check(self._lib.epseon_scan(self._h))

    def refine(self, rel_tol: float = 1e-10, max_iter: int = 200) -> None:
        _check(self._lib.epseon_refine(self._h, rel_tol, max_iter))

    def set_refine_mode(self, mode: int) -> None:
        _check(self._lib.epseon_set_refine_mode(self._h, mode))

    def tuning(self) -> Tuning:
        t = Tuning()
        _check(self._lib.epseon_get_tuning(self._h, ctypes.byref(t)))
        return t

    def set_tuning(self, t: Tuning) -> None:
        _check(self._lib.epseon_set_tuning(self._h, ctypes.byref(t)))

    def autotune(self, rel_tol: float = 1e-10, max_iter: int = 200, keep_results: bool = False) -> Tuning:
        """Time candidate launch configurations on the loaded batch and keep
        the fastest (epseon_autotune)."""
        t = Tuning()
        flags = TUNE_KEEP_RESULTS if keep_results else 0
        _check(self._lib.epseon_autotune(self._h, rel_tol, max_iter, flags, ctypes.byref(t)))
        return t

    def set_solve_split(self, fraction: float) -> None:
        """Scan-work fraction of the first group of the pipelined solve (0: sequential)."""
        _check(self._lib.epseon_set_solve_split(self._h, fraction))

    def solve(self, rel_tol: float = 1e-10, max_iter: int = 200) -> None:
        _check(self._lib.epseon_solve(self._h, rel_tol, max_iter))

    def synchronize(self) -> None:
        _check(self._lib.epseon_ctx_synchronize(self._h))

    def levels(
        self, out: tuple[np.ndarray, np.ndarray, np.ndarray] | None = None
    ) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(levels, status, iterations), concatenated in potential order.
        ``out`` may hold three C-contiguous arrays (float64, int32, int32) of
        at least the total level count to fill instead of new ones, so that a
        loop of solves reuses its result memory."""
        n = sum(self.level_counts)
        if out is None:
            lv = np.empty(max(n, 1), dtype=np.float64)
            st = np.empty(max(n, 1), dtype=np.int32)
            it = np.empty(max(n, 1), dtype=np.int32)
        else:
            lv, st, it = out
            for a, dt in ((lv, np.float64), (st, np.int32), (it, np.int32)):
                if a.dtype != dt or not a.flags.c_contiguous or a.size < max(n, 1):
                    raise ValueError("levels(out=...): arrays must be C-contiguous float64, int32, int32 "
                                     f"with at least {max(n, 1)} elements")
        _check(
            self._lib.epseon_get_levels(
                self._h,
                lv.ctypes.data_as(_c_double_p),
                st.ctypes.data_as(_c_int32_p),
                it.ctypes.data_as(_c_int32_p),
            )
        )
        return lv[:n], st[:n], it[:n]

    def scan_results(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(node counts, far-end values, rescale counts) in potential order."""
        n = sum(self.energy_counts)
        nodes = np.empty(max(n, 1), dtype=np.int32)
        far = np.empty(max(n, 1), dtype=np.float64)
        resc = np.empty(max(n, 1), dtype=np.int32)
        _check(
            self._lib.epseon_get_scan(
                self._h,
                nodes.ctypes.data_as(_c_int32_p),
                far.ctypes.data_as(_c_double_p),
                resc.ctypes.data_as(_c_int32_p),
            )
        )
        return nodes[:n], far[:n], resc[:n]

    def brackets(self) -> np.ndarray:
        """Brackets of the last refinement, one record per level in level order
        (fields e_lo, e_hi, nodes_lo, nodes_hi, lo_valid, found, band)."""
        n = sum(self.level_counts)
        out = np.zeros(max(n, 1), dtype=BRACKET_DTYPE)
        _check(self._lib.epseon_get_brackets(self._h, out.ctypes.data_as(ctypes.c_void_p)))
        return out[:n]

    def stats(self) -> Stats:
        st = Stats()
        _check(self._lib.epseon_get_stats(self._h, ctypes.byref(st)))
        return st

    def normalize(self) -> None:
        """Eigenfunction normalisation of every converged level (epseon_normalize)."""
        _check(self._lib.epseon_normalize(self._h))

    def wavefunctions(self, p: int) -> np.ndarray:
        """Normalised eigenfunctions of potential p, shape (level_count, n_points)."""
        out = np.empty((self.level_counts[p], self.point_counts[p]), dtype=np.float64)
        _check(self._lib.epseon_get_wavefunctions(self._h, p, out.ctypes.data, MEM_HOST))
        return out

    def wavefunctions_device(self, p: int, out_ptr: int) -> None:
        """Write potential p's eigenfunctions to device memory at out_ptr
        (level_count * n_points doubles)."""
        _check(self._lib.epseon_get_wavefunctions(self._h, p, ctypes.c_void_p(out_ptr), MEM_DEVICE))

    def eigenfunction_size(self) -> int:
        n = ctypes.c_int64()
        _check(self._lib.epseon_get_eigenfunction_size(self._h, ctypes.byref(n)))
        return int(n.value)

    def all_wavefunctions(self) -> list[np.ndarray]:
        """Eigenfunctions of every potential (one launch), as (level_count, n_points) arrays."""
        out = np.empty(max(self.eigenfunction_size(), 1), dtype=np.float64)
        _check(self._lib.epseon_get_all_wavefunctions(self._h, out.ctypes.data, MEM_HOST))
        res, o = [], 0
        for lc, n in zip(self.level_counts, self.point_counts):
            res.append(out[o : o + lc * n].reshape(lc, n))
            o += lc * n
        return res

    def all_wavefunctions_device(self, out_ptr: int) -> None:
        """Write every eigenfunction to device memory at out_ptr (eigenfunction_size() doubles)."""
        _check(self._lib.epseon_get_all_wavefunctions(self._h, ctypes.c_void_p(out_ptr), MEM_DEVICE))

    @property
    def v_span(self) -> int:
        """Length of the V layout passed to load (and of an adjoint output)."""
        return max((o + n for o, n in zip(self.v_offsets, self.point_counts)), default=0)

    def level_vjp(self, grad_levels: np.ndarray) -> np.ndarray:
        """sum_l grad_levels[l] dE_l/dV in the layout of the loaded V batch
        (epseon_level_vjp, host memory)."""
        g = _as_f64(grad_levels)
        if g.shape != (sum(self.level_counts),):
            raise ValueError(f"grad_levels must have shape ({sum(self.level_counts)},), got {g.shape}")
        out = np.zeros(max(self.v_span, 1), dtype=np.float64)
        _check(self._lib.epseon_level_vjp(self._h, g.ctypes.data, out.ctypes.data, MEM_HOST))
        return out[: self.v_span]

    def level_vjp_device(self, grad_levels_ptr: int, grad_v_ptr: int) -> None:
        """Device-memory form of level_vjp: both pointers address device
        buffers of this context's device."""
        _check(self._lib.epseon_level_vjp(self._h, ctypes.c_void_p(grad_levels_ptr),
                                          ctypes.c_void_p(grad_v_ptr), MEM_DEVICE))

    def level_vjp_noalloc(self, grad_levels_ptr: int, grad_v_ptr: int) -> None:
        """level_vjp_device without any allocation: the normalisation's
        checkpoints live in the library's per-device scratch array
        (epseon_level_vjp_noalloc; calls overlapping on one device race)."""
        _check(self._lib.epseon_level_vjp_noalloc(self._h, ctypes.c_void_p(grad_levels_ptr),
                                                  ctypes.c_void_p(grad_v_ptr)))
