"""tos_oc_update_noalloc (the OC update that allocates nothing: partials and scalars in a
fixed device array per GPU) against tos_oc_update on a context holding the same densities
and sensitivities, through the C ABI.  Two calls in a row, the second on a smaller
domain, so the second must not see the first call's scratch contents.  The threshold
bisection's decisions depend only on the volumes' comparison with the target, so equal
decisions give the same threshold and bit-identical densities; the two entry points
partition the volume sums differently, so only an exact tie could differ."""
import ctypes

import numpy as np
import pytest

from topoptspline3d_amd.context import Context

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from topoptspline3d_amd.backend import gpu_available, hip_api
    if not gpu_available():
        pytest.skip("no GPU")
    return hip_api()


def domain(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    solid = (rng.random((nz, ny, nx)) > 0.25).astype(np.uint8)
    solid[:, :, 0] = 1
    shell = ((rng.random((nz, ny, nx)) > 0.9) & (solid > 0)).astype(np.uint8)
    support = np.zeros((nz + 1, ny + 1, nx + 1), np.uint8)
    support[:, :, 0] = 1
    rho = np.where(solid > 0, rng.uniform(0.05, 1.0, solid.shape), 0.0).astype(np.float32).ravel()
    sens = (-rng.uniform(0.0, 2.0, solid.size) * (rng.random(solid.size) > 0.05)).astype(np.float32)
    return solid, shell, support, rho, sens


def by_context(hip, solid, shell, support, rho, sens, vgoal, move, damp):
    c = Context(hip, solid, support, shell=shell)
    c.set_material(1e6, 0.4, 3.0, 1e-3)
    c.set_density(rho)
    c.set_sensitivity(sens)
    g, steps = c.oc_update(vgoal, move, damp)
    out = c.density()
    c.close()
    return out, g, steps


def by_noalloc(hip, solid, shell, rho, sens, vgoal, move, damp, out_dev=None):
    import torch
    dev = torch.device("cuda")
    act = torch.from_numpy(solid.ravel().copy()).to(dev)
    shl = torch.from_numpy(shell.ravel().copy()).to(dev)
    r = torch.from_numpy(rho).to(dev)
    s = torch.from_numpy(sens).to(dev)
    o = torch.full_like(r, -7.0)
    torch.cuda.synchronize()
    g, steps = ctypes.c_double(), ctypes.c_int()
    rc = hip.oc_update_noalloc(rho.size, act.data_ptr(), shl.data_ptr(), r.data_ptr(), s.data_ptr(), o.data_ptr(),
                               vgoal, move, damp, 1e-3, None, ctypes.byref(g), ctypes.byref(steps))
    assert rc == 0, rc
    torch.cuda.synchronize()
    return o.cpu().numpy(), g.value, steps.value


@pytest.mark.parametrize("dims", [((40, 20, 24), (13, 7, 9))])
def test_noalloc_oc_matches_context_oc(hip, dims):
    for i, (nx, ny, nz) in enumerate(dims):
        solid, shell, support, rho, sens = domain(nx, ny, nz, 5 + i)
        print(f"[noalloc] call {i}: {nx}x{ny}x{nz}", flush=True)
        ref, g_ref, st_ref = by_context(hip, solid, shell, support, rho, sens, 0.45, 0.05, 0.5)
        out, g, st = by_noalloc(hip, solid, shell, rho, sens, 0.45, 0.05, 0.5)
        act = solid.ravel() > 0
        assert st == st_ref and g == g_ref, (st, st_ref, g, g_ref)
        assert np.array_equal(out[act], ref[act])
        assert np.array_equal(out[~act], rho[~act])  # inactive elements copied through
        assert np.all(out[act & (shell.ravel() > 0)] == 1.0)


def test_noalloc_oc_rejects_sizes_beyond_its_scratch(hip):
    from topoptspline3d_amd import _abi
    n = 67108864 + 1  # TOS_NOALLOC_MAX_ELEMENTS + 1: rejected on the host before any launch
    rc = hip.oc_update_noalloc(n, None, None, 16, 16, 16, 0.4, 0.03, 0.5, 1e-3, None, None, None)
    assert rc == -7 and _abi.ERRORS[rc] == "beyond capacity"
    rc = hip.oc_update_noalloc(-1, None, None, 16, 16, 16, 0.4, 0.03, 0.5, 1e-3, None, None, None)
    assert rc == -1
