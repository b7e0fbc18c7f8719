"""The fine-level Gauss-Seidel schedules (eight colour passes, the colour-slab wavefront,
row-paired colours with a workgroup barrier, colours paired inside a wave) must perform
exactly the same updates: V-cycles and solves are compared bit for bit between sweep
modes 0-3 on irregular domains (slabs of one and of several node planes, rows of one and
of several wave segments, rows starting at x = 0)."""
import numpy as np
import pytest

from topoptspline3d_amd.context import Context
from topoptspline3d_amd.domains import cantilever

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from topoptspline3d_amd.backend import gpu_available, hip_api
    if not gpu_available():
        pytest.skip("no GPU")
    return hip_api()


@pytest.mark.parametrize("dims,seed,zsup", [((40, 20, 24), 0, False), ((30, 16, 69), 1, False),
                                            ((300, 12, 10), 2, False), ((260, 10, 9), 3, True)])
def test_wavefront_schedule_is_bit_identical(hip, dims, seed, zsup):
    """zsup: supports on z = 0 instead of x = 0, so that rows start at x = 0; the wide
    grids have several in-wave pair segments per row."""
    print(f"\n[schedule] dims={dims} seed={seed} zsup={zsup}", flush=True)
    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    solid, support, f = cantilever(nx, ny, nz)
    if zsup:
        support = np.zeros_like(support)
        support.reshape(nz + 1, ny + 1, nx + 1)[0] = 1
    solid = (rng.random(solid.shape) > 0.25).astype(np.uint8)
    solid[:, :, 0] = 1
    solid[0, :, -1] = 1
    c = Context(hip, solid, support)
    c.set_material(1e6, 0.4, 3.0, 1e-3)
    c.set_loads(f)
    c.set_density(rng.uniform(0.05, 1.0, solid.size).astype(np.float32))
    c.update_operator()
    r = rng.normal(size=3 * c.nn)
    out = {}
    for mode in (0, 1, 2, 3):
        c.set_sweep_mode(mode)
        z0 = c.vcycle(r, precision=0)
        z1 = c.vcycle(r, precision=1)
        # (the long cantilevers need more than maxit iterations: compared unconverged)
        st = c.solve(rtol=1e-8, maxit=200, precision=1, warm_start=False, check=False)
        out[mode] = (z0, z1, c.displacement(), st["iterations"][0])
    for m in (1, 2, 3):
        assert np.array_equal(out[0][0], out[m][0]), m
        assert np.array_equal(out[0][1], out[m][1]), m
        assert np.array_equal(out[0][2], out[m][2]), m
        assert out[0][3] == out[m][3], m


@pytest.mark.parametrize("precision", [1, 3])
def test_graph_replay_matches_eager(hip, precision):
    """The FCG iteration body is replayed from a captured hipGraph unless kernels are being
    timed (precision 3: four bodies, one per direction buffer); eager (timed) and replayed
    solves must agree bit for bit, call after call, with operator updates (eager work)
    between the replays."""
    import ctypes
    from topoptspline3d_amd import _abi
    nx, ny, nz = 36, 18, 20
    rng = np.random.default_rng(3)
    solid, support, f = cantilever(nx, ny, nz)
    c = Context(hip, solid, support)
    c.set_material(1e6, 0.4, 3.0, 1e-3)
    c.set_loads(f)
    ks = _abi.KernelStats()
    for rep in range(3):
        print(f"\n[graph replay] rep {rep}", flush=True)
        c.set_density(rng.uniform(0.1, 1.0, solid.size).astype(np.float32))
        c.update_operator()
        out = []
        for timed in (True, False, False):
            if timed:
                hip.profile(c.h, 1, ctypes.byref(ks))
            st = c.solve(rtol=1e-8, maxit=200, precision=precision, warm_start=False)
            if timed:
                hip.profile(c.h, 0, ctypes.byref(ks))
            out.append((c.displacement().copy(), st["iterations"][0]))
        assert out[0][1] == out[1][1] == out[2][1]
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][0], out[2][0])


def test_large_level_schedules_identical(hip):
    """Levels of at least 2^20 free nodes form the residual after the zero-start sweep from
    the later-colour neighbours only, per colour (mode 0) or per colour pair (mode 3):
    V-cycles must agree bit for bit."""
    nx, ny, nz = 128, 96, 96
    rng = np.random.default_rng(11)
    solid, support, f = cantilever(nx, ny, nz)
    solid = (rng.random(solid.shape) > 0.05).astype(np.uint8)
    solid[:, :, 0] = 1
    c = Context(hip, solid, support)
    assert c.info()["n_free_nodes"][0] >= 1 << 20
    c.set_material(1e6, 0.4, 3.0, 1e-3)
    c.set_loads(f)
    c.set_density(rng.uniform(0.05, 1.0, solid.size).astype(np.float32))
    c.update_operator()
    r = rng.normal(size=3 * c.nn)
    out = {}
    for mode in (0, 3):
        print(f"\n[schedule] large level, mode {mode}", flush=True)
        c.set_sweep_mode(mode)
        out[mode] = (c.vcycle(r, precision=1), c.vcycle(r, precision=2))
    assert np.array_equal(out[0][0], out[3][0])
    assert np.array_equal(out[0][1], out[3][1])


def test_large_stencil_level_residual_identical(hip):
    """A stored-stencil level above 2^20 free nodes (level 1 of a 206^3 box) forms its
    residual after the zero-start sweep from the later colours only, per colour (mode 0)
    or per colour pair (mode 3): V-cycles must agree bit for bit, and repeat bit for bit."""
    n = 206
    solid, support, f = cantilever(n, n, n)
    c = Context(hip, solid, support)
    assert c.info()["n_free_nodes"][1] >= 1 << 20
    rng = np.random.default_rng(12)
    c.set_material(1e6, 0.4, 3.0, 1e-3)
    c.set_loads(f)
    c.set_density(rng.uniform(0.05, 1.0, solid.size).astype(np.float32))
    c.update_operator()
    r = rng.normal(size=3 * c.nn)
    out = {}
    for mode in (0, 3):
        print(f"\n[schedule] large stencil level, mode {mode}", flush=True)
        c.set_sweep_mode(mode)
        out[mode] = (c.vcycle(r, precision=1), c.vcycle(r, precision=2))
    assert np.array_equal(out[0][0], out[3][0])
    assert np.array_equal(out[0][1], out[3][1])
    # rows of two segments on the fine level: the paired passes must not depend on the
    # order in which the waves of a row's segments run
    for rep in range(3):
        assert np.array_equal(c.vcycle(r, precision=1), out[3][0]), rep
        assert np.array_equal(c.vcycle(r, precision=2), out[3][1]), rep


@pytest.mark.parametrize("dims,seed", [((300, 12, 10), 2), ((40, 20, 24), 0)])
def test_fcg_dots_from_the_sweep(hip, dims, seed):
    """At precision 3 with the in-wave paired schedule, z.r and z.q come from the waves of
    the last fine post-smoothing sweep; under the eight-colour schedule (the same V-cycle
    updates bit for bit) from a pass of their own. Only the summation order of the two
    dots differs: after 2, 3 and 4 FCG iterations (the first beta uses them at the second)
    the iterates agree to rounding, where a node counted twice or missed (rows of several
    segments on the first domain, of one on the second) would move beta by about 1e-4 and
    the iterate with it; the full solves take the same iterations to the same solution
    within what the 1e-8 tolerance leaves open (later betas can come from nearly
    cancelling z.q sums, whose last bits depend on the summation order: 1.7e-8 after 5
    iterations on the first domain, 3e-10 at convergence)."""
    rng = np.random.default_rng(seed)
    solid, support, f = cantilever(*dims)
    c = Context(hip, solid, support)
    c.set_material(1e6, 0.4, 3.0, 1e-3)
    c.set_loads(f)
    c.set_density(rng.uniform(0.1, 1.0, solid.size).astype(np.float32))
    c.update_operator()
    for maxit, tol in ((2, 1e-12), (3, 1e-12), (4, 1e-12), (300, 1e-7)):
        out = {}
        for mode in (0, 3):
            c.set_sweep_mode(mode)
            st = c.solve(rtol=1e-8, maxit=maxit, precision=3, warm_start=False, check=False)
            out[mode] = (c.displacement().copy(), st["iterations"][0])
        d = np.linalg.norm(out[3][0] - out[0][0]) / np.linalg.norm(out[0][0])
        print(f"\n[fcg dots] {dims} maxit {maxit}: iterations {out[0][1]} / {out[3][1]}, "
              f"relative difference {d:.2e}", flush=True)
        assert out[0][1] == out[3][1]
        assert d <= tol
