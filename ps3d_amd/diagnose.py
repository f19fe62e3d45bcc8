"""Host-side diagnostics of a model state (numpy, interior points).

energy() is the kinetic + potential energy of diagnose.F90:11-35 (the
N0 > 0 branch uses 0.5 b^2 / N0^2, the N0 = 0 branch b (k - 1/2) dz), the
quantity the precision comparisons of the mixed form are stated for.
"""
from __future__ import annotations

import numpy as np


def _int(a, prm):
    return a[2:2 + prm.nx, 2:2 + prm.ny, 2:2 + prm.nz]


def energy(f: dict, prm) -> tuple[float, float]:
    """(ke, pe) summed over the interior, diagnose.F90:14-33."""
    u, v, w, b = (_int(f[n], prm) for n in ("u", "v", "w", "b"))
    ke = float(np.sum(0.5 * (u ** 2 + v ** 2 + prm.dsqr * w ** 2)))
    if prm.N0 > 0.0:
        pe = float(np.sum(0.5 * b ** 2 / prm.N0 ** 2))
    else:
        zk = (np.arange(1, prm.nz + 1) - 0.5) * prm.dz
        pe = float(np.sum(b * zk[None, None, :]))
    return ke, pe


def total_energy(f: dict, prm) -> float:
    ke, pe = energy(f, prm)
    return ke + pe


def rel_l2(x: np.ndarray, ref: np.ndarray) -> float:
    """||x - ref||_2 / ||ref||_2 (0 when both vanish)."""
    d = float(np.linalg.norm((x - ref).ravel()))
    n = float(np.linalg.norm(ref.ravel()))
    return d / n if n > 0 else d


def compare(f: dict, ref: dict, prm) -> dict:
    """rel L2 per prognostic field (interior) and of the total energy."""
    out = {n: rel_l2(_int(f[n], prm), _int(ref[n], prm)) for n in ("u", "v", "w", "b", "p")}
    e, er = total_energy(f, prm), total_energy(ref, prm)
    out["energy"] = abs(e - er) / abs(er) if er != 0 else abs(e - er)
    return out
