"""CPU oracle for the Numerov-shooting path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package; the product path
(``epseon_backend_amd``) never does.

It wraps ``numerov_oracle.c`` (built by ``oracle/Makefile`` into
``libnumerov_oracle.so``), a plain-C FP64 restatement of the numerics the HIP
kernels implement.  The reference ships no implementation of this path
(``VibwaAlgorithm<FP>::run``, ``cpp/gpu/include/epseon/gpu/algorithms/
vibwa.hpp:605-637``, never dispatches), so the C file writes the
specification down; its functions are pinned in ``tests/`` against closed
forms (the discrete particle in a box, analytic Morse levels), dense
symmetric eigensolves of the discrete problem, and high-precision mpmath
trajectories committed under ``tests/golden/``.
"""
from __future__ import annotations

import ctypes
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "libnumerov_oracle.so"
K_CM = 16.857629191640175  # hbar^2/(2 amu A^2) in cm^-1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lp = ctypes.POINTER(ctypes.c_int64)
BANDS = 16  # energy bands of the shot windows (numerov_oracle.c, EO_BANDS)
_LIB: ctypes.CDLL | None = None


class Bracket(ctypes.Structure):
    _fields_ = [
        ("e_lo", ctypes.c_double),
        ("e_hi", ctypes.c_double),
        ("w_lo", ctypes.c_double),
        ("w_hi", ctypes.c_double),
        ("k_lo", ctypes.c_int32),
        ("k_hi", ctypes.c_int32),
        ("c_lo", ctypes.c_int32),
        ("c_hi", ctypes.c_int32),
        ("lo_valid", ctypes.c_int32),
        ("found", ctypes.c_int32),
        ("rate", ctypes.c_double),
        ("band", ctypes.c_int32),
        ("lo_foreign", ctypes.c_int32),
        ("n_pts", ctypes.c_int32),
        ("i_hi", ctypes.c_int32),
        ("e_pt", ctypes.c_double * 12),
        ("w_pt", ctypes.c_double * 12),
        ("k_pt", ctypes.c_int32 * 12),
    ]


def build() -> Path:
    """Compile the oracle with its Makefile (no-op when up to date)."""
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True, capture_output=True)
    return LIB_PATH


def _lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        build()
    lib = ctypes.CDLL(str(LIB_PATH))
    d, i32, i64 = ctypes.c_double, ctypes.c_int32, ctypes.c_int64
    sig = {
        "eo_num_threads": (i32, []),
        "eo_set_num_threads": (None, [i32]),
        "eo_grid_step": (d, [d, d, i64]),
        "eo_beta": (d, [d, d]),
        "eo_shifted_potential": (None, [_dp, i64, d, _dp]),
        "eo_scaled_potential": (None, [_dp, i64, d, _dp]),
        "eo_scan_table": (None, [_dp, i64, d, _dp]),
        "eo_far_scale": (None, [i64, _dp, _ip]),
        "eo_model_exp_value": (d, [d]),
        "eo_model_log_value": (d, [d]),
        "eo_model_sqrt_value": (d, [d]),
        "eo_shoot_diff": (i32, [_dp, i64, d, i32, _dp, _ip]),
        "eo_morse": (None, [d, d, d, d, d, i64, _dp]),
        "eo_trim": (i32, [_dp, i64, d, d, ctypes.POINTER(i64), ctypes.POINTER(i64)]),
        "eo_trial_energy": (d, [i64, d, d]),
        "eo_trial_spacing": (d, [d, d, i64]),
        "eo_node_stride": (i32, [d, d, d]),
        "eo_shoot": (i32, [_dp, i64, d, i32, _dp, _ip]),
        "eo_shoot_full_count": (i32, [_dp, i64, d]),
        "eo_trajectory": (None, [_dp, i64, d, _dp]),
        "eo_scan": (None, [_dp, i64, d, d, d, i64, i32, _lp, _ip, _dp, _ip]),
        "eo_bracket_level": (None, [_ip, _dp, _ip, i64, d, d, d, i32, _lp, _ip, ctypes.POINTER(Bracket)]),
        "eo_window": (None, [_dp, i64, d, ctypes.POINTER(i64), ctypes.POINTER(i64)]),
        "eo_band_windows": (None, [_dp, i64, d, d, d, _lp]),
        "eo_full_windows": (None, [i64, _lp]),
        "eo_band_top": (d, [i32, d, d]),
        "eo_group_band": (i32, [i64, i64, d, d, d]),
        "eo_solve_windows": (None, [_dp, i64, d, d, d, i32, _lp, _ip]),
        "eo_refine_level": (
            i32,
            [_dp, i64, d, i32, i32, ctypes.POINTER(Bracket), d, d, i32, i32, i32, _dp],
        ),
        "eo_solve": (
            None,
            [_dp, _dp, i64, d, d, d, i64, i32, i32, i32, d, d, i32, i32, i32, i32, _ip, _dp, _ip, _dp, _ip],
        ),
        "eo_segment_count": (i32, [i64, i32]),
        "eo_segment_start": (i64, [i64, i32, i32]),
        "eo_shoot_segmented": (None, [_dp, i64, d, i32, i32, _dp, _ip, _ip]),
        "eo_match_point": (i64, [_dp, i64, i64]),
        "eo_psi_norm": (None, [_dp, i64, d, i64, d, _dp, _ip]),
        "eo_psi": (None, [_dp, i64, d, i64, d, _dp]),
        "eo_level_vjp": (None, [_dp, i64, _dp, _dp, i32, i64, d, _dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def _f64(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def num_threads() -> int:
    """OpenMP threads the oracle's parallel loops use."""
    return int(_lib().eo_num_threads())


def set_num_threads(n: int) -> None:
    """Threads of the oracle's parallel loops from now on."""
    _lib().eo_set_num_threads(int(n))


def grid_step(r_min: float, r_max: float, n: int) -> float:
    return float(_lib().eo_grid_step(r_min, r_max, n))


def beta(h: float, mu: float) -> float:
    return float(_lib().eo_beta(h, mu))


def morse(de: float, re: float, a: float, r_min: float, h: float, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    _lib().eo_morse(de, re, a, r_min, h, n, out.ctypes.data_as(_dp))
    return out


def shifted_potential(v: np.ndarray, b: float) -> np.ndarray:
    v = _f64(v)
    out = np.empty_like(v)
    _lib().eo_shifted_potential(v.ctypes.data_as(_dp), len(v), b, out.ctypes.data_as(_dp))
    return out


def trim(v: np.ndarray, b: float, e_hi: float) -> tuple[int, int, bool]:
    """(first, last, usable): the Dirichlet points after wall trimming."""
    v = _f64(v)
    first = ctypes.c_int64()
    last = ctypes.c_int64()
    ok = _lib().eo_trim(v.ctypes.data_as(_dp), len(v), b, e_hi, ctypes.byref(first), ctypes.byref(last))
    return int(first.value), int(last.value), bool(ok)


def scaled_potential(v: np.ndarray, b: float) -> np.ndarray:
    v = _f64(v)
    out = np.empty_like(v)
    _lib().eo_scaled_potential(v.ctypes.data_as(_dp), len(v), b, out.ctypes.data_as(_dp))
    return out


def trial_spacing(e_lo: float, e_hi: float, m: int) -> float:
    return float(_lib().eo_trial_spacing(e_lo, e_hi, m))


def trial_energies(e_lo: float, e_hi: float, m: int) -> np.ndarray:
    de = trial_spacing(e_lo, e_hi, m)
    lib = _lib()
    return np.array([lib.eo_trial_energy(j, e_lo, de) for j in range(m)], dtype=np.float64)


def node_stride(b: float, e_hi: float, v_min: float) -> int:
    return int(_lib().eo_node_stride(b, e_hi, v_min))


def shoot(pt: np.ndarray, e: float, stride: int) -> tuple[int, float, int]:
    """(node count, far-end mantissa, scale-up count r) of one scan shot at
    probe e = beta*E over the scan table pt (w_{N-1} = mantissa 2^(I - 480 r))."""
    pt = _f64(pt)
    w = ctypes.c_double()
    k = ctypes.c_int32()
    c = _lib().eo_shoot(pt.ctypes.data_as(_dp), len(pt), e, stride, ctypes.byref(w), ctypes.byref(k))
    return int(c), float(w.value), int(k.value)


def far_scale(n: int) -> tuple[float, int]:
    """(c, I) with 12^(n-2) = c 2^I: the far-end scale of an n-point scan shot."""
    c = ctypes.c_double()
    i = ctypes.c_int32()
    _lib().eo_far_scale(n, ctypes.byref(c), ctypes.byref(i))
    return float(c.value), int(i.value)


def scan_far_end(pt: np.ndarray, e: float, stride: int) -> float:
    """w_{N-1} of a scan shot as a float (small problems only: may overflow)."""
    _, w, r = shoot(pt, e, stride)
    _, i = far_scale(len(pt))
    return float(np.ldexp(w, i - 480 * r))


def shoot_diff(uv: np.ndarray, e: float, stride: int) -> tuple[int, float, int]:
    """(node count, w_{N-1}, rescales) of one difference-form shot."""
    uv = _f64(uv)
    w = ctypes.c_double()
    k = ctypes.c_int32()
    c = _lib().eo_shoot_diff(uv.ctypes.data_as(_dp), len(uv), e, stride, ctypes.byref(w), ctypes.byref(k))
    return int(c), float(w.value), int(k.value)


def shoot_full_count(ut: np.ndarray, e: float) -> int:
    ut = _f64(ut)
    return int(_lib().eo_shoot_full_count(ut.ctypes.data_as(_dp), len(ut), e))


def trajectory(ut: np.ndarray, e: float) -> np.ndarray:
    ut = _f64(ut)
    out = np.empty_like(ut)
    _lib().eo_trajectory(ut.ctypes.data_as(_dp), len(ut), e, out.ctypes.data_as(_dp))
    return out


def window(uv: np.ndarray, e: float) -> tuple[int, int]:
    """(lo, hi): Dirichlet points of the shot window at probe e = beta*E."""
    uv = _f64(uv)
    lo = ctypes.c_int64()
    hi = ctypes.c_int64()
    _lib().eo_window(uv.ctypes.data_as(_dp), len(uv), e, ctypes.byref(lo), ctypes.byref(hi))
    return int(lo.value), int(hi.value)


def band_windows(uv: np.ndarray, b: float, e_lo: float, e_hi: float, windows: bool = True) -> np.ndarray:
    """(BANDS, 2) int64 array of the band windows (the whole grid in every
    band when ``windows`` is False)."""
    uv = _f64(uv)
    win = np.empty((BANDS, 2), dtype=np.int64)
    far = np.empty(BANDS, dtype=np.int32)
    _lib().eo_solve_windows(uv.ctypes.data_as(_dp), len(uv), b, e_lo, e_hi, int(windows),
                            win.ctypes.data_as(_lp), far.ctypes.data_as(_ip))
    return win


def band_far_exps(win: np.ndarray) -> np.ndarray:
    """Far-end exponents of the band windows."""
    return np.array([far_scale(int(hi - lo + 1))[1] for lo, hi in win], dtype=np.int32)


def band_top(band: int, e_lo: float, e_hi: float) -> float:
    return float(_lib().eo_band_top(band, e_lo, e_hi))


def group_band(j: int, m: int, e_lo: float, e_hi: float) -> int:
    """Band of the scan group (64 trial energies) holding trial energy j."""
    return int(_lib().eo_group_band(j, m, e_lo, trial_spacing(e_lo, e_hi, m), e_hi))


def scan(pt: np.ndarray, b: float, e_lo: float, e_hi: float, m: int, stride: int,
         win: np.ndarray | None = None):
    """(nodes, far-end mantissas, scale-up counts) of every trial energy, each
    shot in its group's band window (``win``; the whole grid when None)."""
    ut = _f64(pt)
    if win is None:
        win = np.array([[0, len(ut) - 1]] * BANDS, dtype=np.int64)
    win = np.ascontiguousarray(win, dtype=np.int64)
    nodes = np.empty(m, dtype=np.int32)
    far = np.empty(m, dtype=np.float64)
    resc = np.empty(m, dtype=np.int32)
    _lib().eo_scan(
        ut.ctypes.data_as(_dp), len(ut), b, e_lo, e_hi, m, stride, win.ctypes.data_as(_lp),
        nodes.ctypes.data_as(_ip), far.ctypes.data_as(_dp), resc.ctypes.data_as(_ip),
    )
    return nodes, far, resc


def bracket_level(nodes, far, resc, e_lo: float, e_hi: float, v: int, far_exp,
                  win: np.ndarray | None = None) -> Bracket:
    """Bracket of level v from scan arrays.  ``win`` holds the band windows
    of the scan (None: one window, the whole grid, for every band) and
    ``far_exp`` their far-end exponents (one for all bands, or one each)."""
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    far = _f64(far)
    resc = np.ascontiguousarray(resc, dtype=np.int32)
    fe = np.ascontiguousarray(np.broadcast_to(np.asarray(far_exp, dtype=np.int32), (BANDS,)))
    if win is None:
        win = np.zeros((BANDS, 2), dtype=np.int64)
    win = np.ascontiguousarray(win, dtype=np.int64)
    br = Bracket()
    m = len(nodes)
    _lib().eo_bracket_level(
        nodes.ctypes.data_as(_ip), far.ctypes.data_as(_dp), resc.ctypes.data_as(_ip), m,
        e_lo, trial_spacing(e_lo, e_hi, m), e_hi, v, win.ctypes.data_as(_lp), fe.ctypes.data_as(_ip),
        ctypes.byref(br),
    )
    return br


def refine_level(uv, b: float, stride: int, v: int, br: Bracket, rel_tol: float, abs_tol: float,
                 max_iter: int = 200, segmented: bool = False, lanes: int = 64) -> tuple[float, int]:
    uv = _f64(uv)
    lv = ctypes.c_double()
    it = _lib().eo_refine_level(
        uv.ctypes.data_as(_dp), len(uv), b, stride, v, ctypes.byref(br), rel_tol, abs_tol,
        max_iter, int(segmented), lanes, ctypes.byref(lv),
    )
    return float(lv.value), int(it)


@dataclass
class Problem:
    """One potential prepared exactly as the library prepares it."""

    v: np.ndarray
    r_min: float
    r_max: float
    mu: float
    e_top: float
    m: int

    @property
    def n(self) -> int:
        return len(self.v)

    @property
    def h(self) -> float:
        return grid_step(self.r_min, self.r_max, self.n)

    @property
    def beta(self) -> float:
        return beta(self.h, self.mu)

    @property
    def e_lo(self) -> float:
        return float(np.min(self.v))

    @property
    def trimmed(self) -> tuple[int, int, bool]:
        """(first, last, usable) of the wall-trimmed grid."""
        return trim(self.v, self.beta, self.e_top)

    @property
    def usable(self) -> bool:
        return self.trimmed[2]

    @property
    def ut(self) -> np.ndarray:
        """beta V - 1 on the trimmed grid (the table of the division-free w-form:
        trajectory and full-count checks)."""
        first, last, _ = self.trimmed
        return shifted_potential(self.v[first : last + 1], self.beta)

    @property
    def pt(self) -> np.ndarray:
        """(beta V - 1)/12 on the trimmed grid (the scan shots' table)."""
        first, last, _ = self.trimmed
        v = _f64(self.v[first : last + 1])
        out = np.empty_like(v)
        _lib().eo_scan_table(v.ctypes.data_as(_dp), len(v), self.beta, out.ctypes.data_as(_dp))
        return out

    @property
    def far_exp(self) -> int:
        """Binary exponent of the scan far-end scale of the trimmed grid."""
        first, last, _ = self.trimmed
        return far_scale(last - first + 1)[1]

    @property
    def uv(self) -> np.ndarray:
        """beta V on the trimmed grid (the refinement shots' table)."""
        first, last, _ = self.trimmed
        return scaled_potential(self.v[first : last + 1], self.beta)

    @property
    def stride(self) -> int:
        return node_stride(self.beta, self.e_top, self.e_lo)

    @property
    def abs_tol(self) -> float:
        return 1e-15 * (self.e_top - self.e_lo)

    def windows(self, on: bool = True) -> np.ndarray:
        """(BANDS, 2) band windows (lo, hi) on the trimmed grid."""
        return band_windows(self.uv, self.beta, self.e_lo, self.e_top, on)

    def far_exps(self, on: bool = True) -> np.ndarray:
        """Far-end exponents of the band windows."""
        return band_far_exps(self.windows(on))

    def scan(self, on: bool = True):
        """(nodes, far, resc) of the windowed (or whole-grid) scan."""
        return scan(self.pt, self.beta, self.e_lo, self.e_top, self.m, self.stride, self.windows(on))

    def scan_shot(self, j: int, on: bool = True) -> tuple[int, float, int]:
        """The scan's shot of trial energy j: (nodes, far-end mantissa, r)."""
        lo, hi = self.windows(on)[group_band(j, self.m, self.e_lo, self.e_top)]
        e = self.beta * float(trial_energies(self.e_lo, self.e_top, self.m)[j])
        return shoot(self.pt[lo : hi + 1], e, self.stride)

    def refine(self, nodes, far, resc, v: int, rel_tol: float = 1e-10, max_iter: int = 200,
               segmented: bool = False, lanes: int = 64, on: bool = True) -> tuple[float, int]:
        """Refine level v from scan arrays as the solve does (bracket, band
        window, refinement): (level or NaN, iterations)."""
        win = self.windows(on)
        br = bracket_level(nodes, far, resc, self.e_lo, self.e_top, v, band_far_exps(win), win)
        if not br.found:
            return float("nan"), 0
        lo, hi = win[br.band]
        return refine_level(self.uv[lo : hi + 1], self.beta, self.stride, v, br, rel_tol, self.abs_tol,
                            max_iter, segmented, lanes)

    @property
    def match(self) -> int:
        """Matching point of the two inward eigenfunction shots (min V,
        relative to the trimmed grid, clamped to [1, n_trim - 3])."""
        first, last, _ = self.trimmed
        return int(_lib().eo_match_point(_f64(self.v).ctypes.data_as(_dp), first, last - first + 1))


def segment_count(n: int, lanes: int = 64) -> int:
    return int(_lib().eo_segment_count(n, lanes))


def shoot_segmented(uv: np.ndarray, e: float, stride: int, count: bool = True, lanes: int = 64):
    """(far-end mantissa, binary exponent, node count or None) of a segmented
    difference-form shot."""
    uv = _f64(uv)
    w = ctypes.c_double()
    k = ctypes.c_int32()
    c = ctypes.c_int32()
    _lib().eo_shoot_segmented(
        uv.ctypes.data_as(_dp), len(uv), e, stride, lanes, ctypes.byref(w), ctypes.byref(k),
        ctypes.byref(c) if count else None,
    )
    return float(w.value), int(k.value), (int(c.value) if count else None)


def solve(prob: Problem, v_min: int, n_levels: int, rel_tol: float = 1e-10, max_iter: int = 200,
          segmented: bool = False, lanes: int = 64, windows: bool = True):
    """Scan + bracket + refine, as epseon_solve does: returns
    (nodes, far, resc, levels, iterations).  ``segmented`` selects the
    refinement with segmented shots over blocks of ``lanes`` threads;
    ``windows`` the band-windowed shots (else every shot spans the grid)."""
    ut = prob.pt
    uv = prob.uv
    m = prob.m
    nodes = np.empty(m, dtype=np.int32)
    far = np.empty(m, dtype=np.float64)
    resc = np.empty(m, dtype=np.int32)
    lv = np.empty(max(n_levels, 1), dtype=np.float64)
    it = np.empty(max(n_levels, 1), dtype=np.int32)
    _lib().eo_solve(
        ut.ctypes.data_as(_dp), uv.ctypes.data_as(_dp), len(ut), prob.beta, prob.e_lo, prob.e_top, m, prob.stride,
        v_min, n_levels, rel_tol, prob.abs_tol, max_iter, int(segmented), lanes, int(windows),
        nodes.ctypes.data_as(_ip), far.ctypes.data_as(_dp), resc.ctypes.data_as(_ip),
        lv.ctypes.data_as(_dp), it.ctypes.data_as(_ip),
    )
    return nodes, far, resc, lv[:n_levels], it[:n_levels]


def morse_levels(mu: float, de: float, a: float, v: np.ndarray | int) -> np.ndarray:
    """Analytic Morse levels E_v = 2a sqrt(K' D_e)(v+1/2) - a^2 K' (v+1/2)^2, K' = K/mu."""
    kp = K_CM / mu
    x = np.asarray(v, dtype=np.float64) + 0.5
    return 2.0 * a * np.sqrt(kp * de) * x - a * a * kp * x * x


def psi_norm(prob: Problem, level: float) -> tuple[float, float, int, int]:
    """(a_left, a_right, k_left, k_right) of the eigenfunction at E = level."""
    uv = prob.uv
    a = np.zeros(2)
    k = np.zeros(2, dtype=np.int32)
    _lib().eo_psi_norm(uv.ctypes.data_as(_dp), len(uv), prob.beta * level, prob.match, prob.h,
                       a.ctypes.data_as(_dp), k.ctypes.data_as(_ip))
    return float(a[0]), float(a[1]), int(k[0]), int(k[1])


def eigenfunction(prob: Problem, level: float) -> np.ndarray:
    """Normalised eigenfunction at E = level on the full grid (zero outside
    the open trimmed window), sum psi^2 h = 1."""
    uv = prob.uv
    first, last, _ = prob.trimmed
    out = np.zeros(prob.n)
    win = np.zeros(len(uv))
    _lib().eo_psi(uv.ctypes.data_as(_dp), len(uv), prob.beta * level, prob.match, prob.h,
                  win.ctypes.data_as(_dp))
    out[first : last + 1] = win
    return out


def level_vjp(prob: Problem, levels: np.ndarray, grad: np.ndarray) -> np.ndarray:
    """sum_l grad[l] psi_l^2 h on the full grid over the finite levels."""
    uv = prob.uv
    first, last, _ = prob.trimmed
    e = _f64(np.where(np.isfinite(levels), prob.beta * np.asarray(levels, dtype=np.float64), np.nan))
    g = _f64(grad)
    win = np.zeros(len(uv))
    _lib().eo_level_vjp(uv.ctypes.data_as(_dp), len(uv), e.ctypes.data_as(_dp), g.ctypes.data_as(_dp),
                        len(e), prob.match, prob.h, win.ctypes.data_as(_dp))
    out = np.zeros(prob.n)
    out[first : last + 1] = win
    return out


def model_functions() -> tuple:
    """(exp, log, sqrt) of the refinement model, for the tests."""
    lib = _lib()
    return lib.eo_model_exp_value, lib.eo_model_log_value, lib.eo_model_sqrt_value
