"""Type stub of the MI355X ``_libepseon_gpu`` module (gpu_module.cpp).

Same names and keyword names as the reference module; members marked
"additive" do not exist there.
"""
from __future__ import annotations

from typing import Literal, Sequence, overload

import numpy as np
import numpy.typing as npt

class PhysicalDeviceSparseProperties:
    """Sparse resource properties (none on this backend)."""

class PhysicalDeviceLimits:
    max_compute_shared_memory_size: int
    max_compute_work_group_count: tuple[int, int, int]
    max_compute_work_group_invocations: int
    max_compute_work_group_size: int

class PhysicalDeviceProperties:
    api_version: str
    driver_version: str
    vendor_id: int
    device_id: int  # HIP device ordinal
    device_type: Literal["INTEGRATED_GPU", "DISCRETE_GPU", "VIRTUAL_GPU", "CPU", "OTHER"]
    device_name: str
    pipeline_cache_uuid: list[int]
    limits: PhysicalDeviceLimits
    sparse_properties: PhysicalDeviceSparseProperties
    arch: str  # additive: gcnArchName
    compute_units: int  # additive
    pci_device_id: int  # additive

class MemoryHeap:
    size: int
    flags: list[str]

class MemoryType:
    heap_index: int
    flags: list[str]

class PhysicalDeviceMemoryProperties:
    memory_heaps: list[MemoryHeap]
    memory_types: list[MemoryType]

class PhysicalDeviceInfo:
    device_properties: PhysicalDeviceProperties
    memory_properties: PhysicalDeviceMemoryProperties

class MorsePotentialConfig:
    dissociation_energy: float
    equilibrium_bond_distance: float
    well_width: float
    min_r: float
    max_r: float
    point_count: int
    def __init__(
        self,
        dissociation_energy: float,
        equilibrium_bond_distance: float,
        well_width: float,
        min_r: float,
        max_r: float,
        point_count: int,
    ) -> None: ...

class DoubleWellPotentialConfig:  # additive
    def __init__(
        self,
        dissociation_energy: float,
        equilibrium_bond_distance: float,
        well_width: float,
        barrier_height: float,
        barrier_position: float,
        barrier_width: float,
        min_r: float,
        max_r: float,
        point_count: int,
    ) -> None: ...

class TabulatedPotentialConfig:  # additive
    def __init__(self, values: Sequence[float], min_r: float, max_r: float, asymptote: float) -> None: ...

class PotentialFileLoader:
    """One tabulated V(r) per file: text "r V" lines (optional "# asymptote
    <x>") or a (N, 2) float64 .npy; uniform r grid, at least 4 points."""
    file_names: list[str]
    @overload
    def __init__(self) -> None: ...
    @overload
    def __init__(self, file_names: Sequence[str]) -> None: ...
    def equals(self, other: PotentialFileLoader) -> bool: ...
    def get_potential_data(self) -> list[npt.NDArray[np.float64]]: ...

class _TaskConfigurator:
    def __init__(self) -> None: ...
    def set_hardware_config(
        self, potential_buffer_size: int, group_size: int, allocation_block_size: int
    ) -> _TaskConfigurator: ...
    def set_morse_potential(self, configurations: Sequence[MorsePotentialConfig]) -> _TaskConfigurator: ...
    def set_vibwa_algorithm(
        self,
        mass_atom_0: float,
        mass_atom_1: float,
        integration_step: float,
        min_distance_to_asymptote: float,
        min_level: int,
        max_level: int,
    ) -> _TaskConfigurator: ...
    def is_configured(self) -> bool: ...
    # additive
    def set_morse_potential_batch(self, configurations: Sequence[MorsePotentialConfig]) -> _TaskConfigurator: ...
    def set_double_well_potential(self, configurations: Sequence[DoubleWellPotentialConfig]) -> _TaskConfigurator: ...
    def set_tabulated_potential(self, configurations: Sequence[TabulatedPotentialConfig]) -> _TaskConfigurator: ...
    def set_potential_file_loader(self, loader: PotentialFileLoader) -> _TaskConfigurator: ...
    def set_trial_energies(self, count: int) -> _TaskConfigurator: ...
    def set_trial_energy_counts(self, counts: Sequence[int]) -> _TaskConfigurator: ...
    def set_reduced_masses(self, masses: Sequence[float]) -> _TaskConfigurator: ...
    def set_tolerance(self, relative: float, max_iterations: int = 200) -> _TaskConfigurator: ...
    def set_eigenfunction_output(self, enabled: bool = True) -> _TaskConfigurator: ...

class TaskConfiguratorFloat32(_TaskConfigurator): ...
class TaskConfiguratorFloat64(_TaskConfigurator): ...

class _TaskHandle:
    def get_status_message(self) -> str: ...
    def is_done(self) -> bool: ...
    def wait(self) -> None: ...
    # additive
    def is_running(self) -> bool: ...
    def cancel(self) -> bool: ...
    def get_level_status(self) -> npt.NDArray[np.int32]: ...
    def get_iterations(self) -> npt.NDArray[np.int32]: ...
    def get_node_counts(self) -> list[npt.NDArray[np.int32]]: ...
    def get_far_end_values(self) -> list[npt.NDArray[np.float64]]: ...
    def get_rescale_counts(self) -> list[npt.NDArray[np.int32]]: ...
    def get_node_strides(self) -> list[int]: ...
    def get_brackets(self) -> dict[str, npt.NDArray]: ...
    def get_potential_flags(self) -> list[int]: ...
    def get_timings(self) -> dict[str, float]: ...

class TaskHandleFloat32(_TaskHandle):
    def get_levels(self) -> npt.NDArray[np.float32]: ...  # additive
    @overload
    def get_wavefunctions(self, potential: None = None) -> list[npt.NDArray[np.float32]]: ...  # additive
    @overload
    def get_wavefunctions(self, potential: int) -> npt.NDArray[np.float32]: ...

class TaskHandleFloat64(_TaskHandle):
    def get_levels(self) -> npt.NDArray[np.float64]: ...  # additive
    @overload
    def get_wavefunctions(self, potential: None = None) -> list[npt.NDArray[np.float64]]: ...  # additive
    @overload
    def get_wavefunctions(self, potential: int) -> npt.NDArray[np.float64]: ...

class ComputeDeviceInterface:
    device_id: int
    def get_task_configurator(
        self, precision: str
    ) -> TaskConfiguratorFloat32 | TaskConfiguratorFloat64: ...
    @overload
    def submit_task(self, configurator: TaskConfiguratorFloat32) -> TaskHandleFloat32: ...
    @overload
    def submit_task(self, configurator: TaskConfiguratorFloat64) -> TaskHandleFloat64: ...

class EpseonComputeContext:
    @staticmethod
    def create() -> EpseonComputeContext: ...
    def get_vulkan_version(self) -> str: ...
    def get_physical_device_info(self) -> list[PhysicalDeviceInfo]: ...
    def get_device_interface(self, device_id: int) -> ComputeDeviceInterface: ...
