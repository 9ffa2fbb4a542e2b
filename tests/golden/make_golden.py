"""Generate the golden vectors in tests/golden/ (run: python tests/golden/make_golden.py).

Everything here is computed with mpmath at 50 significant digits from the
textbook form of Numerov's method, which is not the form the oracle or the
kernels use:

    f_i = 1 + (h^2 mu / (12 K)) (E - V_i)
    f_{i+1} y_{i+1} = (12 - 10 f_i) y_i - f_{i-1} y_{i-1},   y_0 = 0, y_1 = 1

(the oracle runs the division-free product form w_i = y_i f_i prod_{j<i} f_j
and a difference form).  V_i is the Morse potential evaluated in mpmath on the
grid r_i = r_min + i h; the FP64 table the tests feed the oracle is the
correctly rounded value of the same expression, so the comparison measures
the oracle's arithmetic, not a different potential.

Outputs:
  numerov_trajectories.json  y_i for small grids at a few energies
  morse_levels.json          analytic Morse levels of the two single-potential
                             workloads (H2-like and I2-like) in 30 digits
  potential_tables.json      V(r_i) and the Morse parameter derivatives at ~20
                             grid points of the two single-potential grids and
                             of one double well, r_i the FP64 grid values
                             r_min + i h (h = (r_max - r_min)/(N - 1) in FP64)
                             taken exactly
"""
from __future__ import annotations

import json
from pathlib import Path

import mpmath as mp

mp.mp.dps = 50
HERE = Path(__file__).resolve().parent
# hbar^2/(2 amu A^2) in cm^-1 from CODATA 2018 h, c, u (exact h and c).
H = mp.mpf("6.62607015e-34")
U = mp.mpf("1.66053906660e-27")
C = mp.mpf("299792458")
K = H / (8 * mp.pi**2 * U * C) * mp.mpf(10) ** 18


def morse(de, re, a, r):
    x = 1 - mp.e ** (-a * (r - re))
    return de * x * x


def trajectory(de, re, a, mu, r_min, r_max, n, e):
    h = (r_max - r_min) / (n - 1)
    beta = h * h * mu / (12 * K)
    v = [morse(de, re, a, r_min + i * h) for i in range(n)]
    f = [1 + beta * (e - vi) for vi in v]
    y = [mp.mpf(0), mp.mpf(1)]
    for i in range(1, n - 1):
        y.append(((12 - 10 * f[i]) * y[i] - f[i - 1] * y[i - 1]) / f[i + 1])
    return v, f, y


def main() -> None:
    cases = []
    # (D_e, r_e, a, mu, r_min, r_max, N, energies): a light and a heavy Morse
    # on short grids, energies below, near and above their lowest levels.
    specs = [
        ("38297", "0.7414", "1.9426", "0.50391", "0.35", "1.6", 96, ["1500", "2167.25", "4000", "9000"]),
        ("12440", "2.666", "1.8676", "63.452", "2.45", "2.95", 160, ["60", "107.135", "300", "700"]),
    ]
    for de, re, a, mu, rmin, rmax, n, energies in specs:
        de, re, a, mu, rmin, rmax = map(mp.mpf, (de, re, a, mu, rmin, rmax))
        for es in energies:
            e = mp.mpf(es)
            v, f, y = trajectory(de, re, a, mu, rmin, rmax, n, e)
            cases.append(
                {
                    "de": str(de), "re": str(re), "a": str(a), "mu": str(mu),
                    "r_min": str(rmin), "r_max": str(rmax), "n": n, "energy": es,
                    "v": [mp.nstr(x, 25) for x in v],
                    "f": [mp.nstr(x, 25) for x in f],
                    "y": [mp.nstr(x, 25) for x in y],
                }
            )
    (HERE / "numerov_trajectories.json").write_text(json.dumps({"K": mp.nstr(K, 30), "cases": cases}))

    levels = {}
    for name, (de, a, mu, nv) in {
        "h2": ("38297", "1.9426", "0.50391", 17),
        "i2": ("12440", "1.8676", "63.452", 116),
    }.items():
        de, a, mu = map(mp.mpf, (de, a, mu))
        kp = K / mu
        out = []
        for v in range(nv):
            x = v + mp.mpf("0.5")
            out.append(mp.nstr(2 * a * mp.sqrt(kp * de) * x - a * a * kp * x * x, 30))
        levels[name] = out
    (HERE / "morse_levels.json").write_text(json.dumps(levels, indent=1))


def potential_tables() -> None:
    """Golden V(r_i) and dV/d(D_e, a, r_e) at FP64 grid points, 30 digits."""
    # (name, D_e, r_e, a, r_min, r_max, N, barrier B, r_b, w) -- the double
    # well is the first member of the stress batch (its float parameters).
    specs = [
        ("h2", 38297.0, 0.7414, 1.9426, 0.2, 5.0, 10000, 0.0, 0.0, 1.0),
        ("i2", 12440.0, 2.666, 1.8676, 2.0, 8.0, 200000, 0.0, 0.0, 1.0),
        ("double_well", 8193.433993653385, 2.0423502771804785, 1.616745917580468, 1.4423502771804784,
         5.753508557993871, 1000000, 2013.4310468218557, 2.4928887652620624, 0.10133013237059119),
    ]
    out = {}
    for name, de, re, a, r_min, r_max, n, bb, rb, w in specs:
        h = (r_max - r_min) / (n - 1)  # FP64, as the tabulating code computes it
        idx = sorted({0, 1, 2, 3, 17, 100, 1000, n // 7, n // 5, n // 3, n // 2, 2 * n // 3,
                      3 * n // 4, 9 * n // 10, n - 100, n - 3, n - 2, n - 1})
        rows = []
        for i in idx:
            r = r_min + i * h  # FP64 grid value
            R, DE, RE, A = mp.mpf(r), mp.mpf(de), mp.mpf(re), mp.mpf(a)
            ex = mp.e ** (-A * (R - RE))
            x = 1 - ex
            v = DE * x * x
            if bb != 0.0:
                z = (R - mp.mpf(rb)) / mp.mpf(w)
                v += mp.mpf(bb) * mp.e ** (-(z * z))
            rows.append({
                "i": i, "r": repr(r), "v": mp.nstr(v, 30),
                "dv_dde": mp.nstr(x * x, 30),
                "dv_da": mp.nstr(2 * DE * x * ex * (R - RE), 30),
                "dv_dre": mp.nstr(-2 * DE * A * x * ex, 30),
            })
        out[name] = {"de": repr(de), "re": repr(re), "a": repr(a), "r_min": repr(r_min),
                     "r_max": repr(r_max), "n": n, "barrier": repr(bb), "r_b": repr(rb),
                     "width": repr(w), "points": rows}
    (HERE / "potential_tables.json").write_text(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
    potential_tables()
