"""Pinning of the optimal-balance restatement (oracle/balance.py OptBalance,
diag_opt_balance.f90:55-378) on the CPU.

The ramp by its closed form and symmetry; the procedure by what the
mathematics fixes: in a linear, frictionless model the geostrophic mode is a
steady solution of the discrete equations, so backward integration, the
projection on it, forward integration and the exchange of the base point
leave exactly the geostrophic part of the initial state -- the zero-order
balance of diag_balance.f90:109-131 (itself pinned in
test_balance_oracle.py).  A dropped projection, a wrong sign in the exchange
or a missing restore fails these checks.  The time-average variant keeps a
geostrophic state fixed (every average of a steady state is the state).
"""
import math

import numpy as np
import pytest

from ps3d_amd import configs


def test_ramp_closed_form():
    from oracle.balance import ramp
    assert ramp(0.0) == 0.0
    assert ramp(1.0) == 1.0
    assert ramp(0.5) == 0.5
    th = np.linspace(0.01, 0.99, 99)
    r = np.array([ramp(t) for t in th])
    assert np.all(np.diff(r) >= 0) and r[0] < 1e-40 and r[-1] == 1.0
    assert np.allclose(r + np.array([ramp(1 - t) for t in th]), 1.0, atol=1e-15)
    t = 0.3
    assert ramp(t) == math.exp(-1 / t) / (math.exp(-1 / t) + math.exp(-1 / (1 - t)))


def _linear(lid, seed=3):
    prm, fields = configs.random_state(16, 16, 8, seed=seed, lid=lid, Ah=0.0, Av=0.0, Kh=0.0, Kv=0.0,
                                       Ahbi=0.0, Avbi=0.0, Khbi=0.0, Kvbi=0.0, enable_nonlinear=False)
    return prm, fields


def _oracle(prm, fields):
    import oracle
    orc = oracle.OracleModel(prm)
    for n, a in fields.items():
        orc.f[n][...] = a
    return orc


def _rel(a, b, prm, scale=None):
    """relative L2 difference on the interior; scale: the norm to divide by
    when b may vanish (the geostrophic w under a rigid lid is zero)"""
    I = (slice(2, 2 + prm.nx), slice(2, 2 + prm.ny), slice(2, 2 + prm.nz))
    den = np.linalg.norm(b[I]) if scale is None else scale
    return np.linalg.norm(a[I] - b[I]) / max(den, 1e-300)


def _scale(fields, prm, n):
    I = (slice(2, 2 + prm.nx), slice(2, 2 + prm.ny), slice(2, 2 + prm.nz))
    return np.linalg.norm(fields[n][I])


@pytest.mark.parametrize("lid", [True, False])
def test_linear_model_gives_the_geostrophic_part(lid):
    from oracle.balance import Balance, OptBalance
    prm, fields = _linear(lid)
    orc = _oracle(prm, fields)
    bal = Balance(orc)
    bal.zero_order()
    g = [bal.r[n] for n in ("u0", "v0", "w0", "b0")]
    before = {n: orc.f[n].copy() for n in ("u", "v", "w", "b", "p")}
    ub, norms = OptBalance(orc, period=12 * prm.dt, max_iterations=2, tol=1e-30).run()
    for n, a, b_ in zip("uvwb", ub, g):
        assert _rel(a, b_, prm, _scale(fields, prm, n)) < 1e-11, n
    # the state is restored
    for n, a in before.items():
        assert np.array_equal(orc.f[n], a), n
    assert len(norms) == 2 and norms[1] < 1e-11
    # not a trivial pass: the state has a large unbalanced part
    assert _rel(fields["u"], g[0], prm) > 1e-3
    orc.close()


@pytest.mark.parametrize("lid", [True, False])
def test_time_average_keeps_a_geostrophic_state(lid):
    from oracle.balance import Balance, OptBalance
    prm, fields = _linear(lid, seed=8)
    orc = _oracle(prm, fields)
    bal = Balance(orc)
    bal.zero_order()
    for n, q in zip(("u", "v", "w", "b"), ("u0", "v0", "w0", "b0")):
        orc.f[n][...] = bal.r[q]
    g = [bal.r[q].copy() for q in ("u0", "v0", "w0", "b0")]
    ub, norms = OptBalance(orc, period=8 * prm.dt, max_iterations=2, tol=1e-30, time_average=True,
                           average=5 * prm.dt, average_times=2).run()
    for n, a, b_ in zip("uvwb", ub, g):
        assert _rel(a, b_, prm, _scale(fields, prm, n)) < 1e-11, n
    assert max(norms) < 1e-11
    orc.close()
