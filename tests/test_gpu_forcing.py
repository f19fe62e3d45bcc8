"""The set_forcing hook on the GPU (ps3d_set_forcing, ps3d_set_wavemaker).

Parity: the device wavemaker (for_src/config.f90:224-237) inside the step
against the oracle with the same hook, bit for bit over many steps (the
oracle's wavemaker is pinned against the closed form in
test_forcing_oracle.py); a Python hook doing the same forcing through host
copies (the host round trip) against the device wavemaker, bit for bit;
ps3d_device_ptr against ps3d_get_field; the FP64-only rule.
"""
import ctypes
import math

import numpy as np
import pytest

from conftest import interior

pytestmark = pytest.mark.gpu

CASES = {
    "lid_fused": dict(nx=32, ny=24, nz=12, lid=True, Avbi=0.0, Kvbi=0.0),
    "periodic": dict(nx=24, ny=20, nz=10, lid=False),
}


def _setup(case, seed=4):
    import oracle
    from ps3d_amd import configs
    from ps3d_amd.model import Model
    c = dict(CASES[case])
    nx, ny, nz, lid = c.pop("nx"), c.pop("ny"), c.pop("nz"), c.pop("lid")
    prm, fields = configs.random_state(nx, ny, nz, seed=seed, lid=lid, **c)
    gpu = Model(prm)
    orc = oracle.OracleModel(prm)
    for n, a in fields.items():
        gpu.f[n][...] = a
        orc.f[n][...] = a
    gpu.push()
    re, im = configs.wavemaker_amplitudes(prm)
    return prm, gpu, orc, re, im


@pytest.mark.parametrize("case", list(CASES))
def test_wavemaker_matches_oracle(case):
    prm, gpu, orc, re, im = _setup(case)
    print(f"wavemaker case={case}", flush=True)
    omega = -2.0 * math.pi / (40 * prm.dt)
    gpu.set_wavemaker(omega, re, im)
    orc.set_wavemaker(omega, re, im)
    steps = 25
    gpu.run(steps)
    orc.run(steps)
    gpu.pull()
    for n in ("u", "v", "w", "b", "p"):
        assert np.array_equal(gpu.f[n], orc.f[n]), n
    # the forcing did something: the unforced run differs
    plain = type(gpu)(prm)
    from ps3d_amd import configs
    _, fields = configs.random_state(prm.nx, prm.ny, prm.nz, seed=4, lid=prm.enable_vertical_boundaries,
                                     **{k: v for k, v in CASES[case].items() if k not in ("nx", "ny", "nz", "lid")})
    for n, a in fields.items():
        plain.f[n][...] = a
    plain.push()
    plain.run(steps)
    plain.pull()
    assert not np.array_equal(plain.f["u"], gpu.f["u"])
    gpu.clear_wavemaker()
    plain.close()
    gpu.close()
    orc.close()


def test_host_round_trip_hook_matches_device_wavemaker():
    prm, dev, orc, re, im = _setup("lid_fused", seed=9)
    orc.close()
    from ps3d_amd.model import Model
    host = Model(prm)
    dev.pull()
    for n in ("u", "v", "w", "b"):
        host.f[n][...] = dev.f[n]
    host.push()
    omega = 3.1e-3
    dev.set_wavemaker(omega, re, im)
    calls = []

    def hook(m, itt, tau, stream):
        m.synchronize()
        m.pull(["u", "v", "w", "b"])
        t = prm.dt * itt
        cs, sn = math.cos(omega * t), math.sin(omega * t)
        for n in ("u", "v", "w", "b"):
            m.f[n][...] = m.f[n] + prm.dt * (cs * re[n] + sn * im[n])
        m.push(["u", "v", "w", "b"])
        calls.append((itt, tau))

    host.set_forcing(hook)
    steps = 8
    print("host round trip hook", flush=True)
    dev.run(steps)
    host.run(steps)
    assert [c[0] for c in calls] == list(range(steps))
    dev.pull()
    host.pull()
    for n in ("u", "v", "w", "b", "p"):
        assert np.array_equal(dev.f[n], host.f[n]), n
    host.set_forcing(None)
    host.close()
    dev.close()


def test_device_ptr_addresses_the_fields():
    prm, gpu, orc, re, im = _setup("periodic")
    orc.close()
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize()
    seen = {}

    def hook(m, itt, tau, stream):
        m.synchronize()
        for n, slot in (("u", None), ("b", None), ("du", tau), ("db", tau)):
            buf = np.empty(prm.shape, order="F")
            rc = hip.hipMemcpy(buf.ctypes.data, m.device_ptr(n, slot), buf.nbytes, 2)   # device to host
            assert rc == 0
            seen[n] = buf
        m.pull(["u", "b", "du", "db"])
        seen["pulled"] = {n: m.f[n].copy() for n in ("u", "b", "du", "db")}
        seen["tau"] = tau

    gpu.set_forcing(hook)
    gpu.run(1)
    tau = seen["tau"]
    for n in ("u", "b"):
        assert np.array_equal(seen[n], seen["pulled"][n]), n
    for n in ("du", "db"):
        assert np.array_equal(seen[n], seen["pulled"][n][..., tau - 1]), n
    gpu.close()


def test_forcing_is_fp64_only():
    from ps3d_amd import configs
    from ps3d_amd._lib import PS3D_ERR_STATE
    from ps3d_amd.model import Model
    from ps3d_amd.params import PS3D_MIXED
    prm, fields = configs.jets(32, 32, 16, seed=1)
    m = Model(prm.replace(precision=PS3D_MIXED))
    with pytest.raises(Exception) as e:
        m.set_forcing(lambda *a: None)
    assert getattr(e.value, "code", PS3D_ERR_STATE) == PS3D_ERR_STATE
    m.close()


def test_failing_hook_stops_the_step():
    prm, gpu, orc, re, im = _setup("periodic")
    orc.close()

    def hook(m, itt, tau, stream):
        raise ValueError("boom")

    gpu.set_forcing(hook)
    with pytest.raises(Exception):
        gpu.integrate()
    assert isinstance(gpu._forcing_error, ValueError)
    gpu.close()
