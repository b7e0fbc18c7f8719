"""GPU checks of the learned coarse-level correction through the C ABI: the BF16 MFMA
forward and backward kernels against the NumPy oracle (oracle/unet_ref.py, same BF16
rounding points), and what the trained network gains inside the solver.

Tolerances: forward and gradients to 3e-2 relative L2 (BF16 storage, FP32 vs FP64
accumulation order); the loss to 1e-2 relative."""
import os
import sys

import numpy as np
import pytest

from topoptspline3d_amd.context import Context
from topoptspline3d_amd.domains import cantilever

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import unet_ref as U  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from topoptspline3d_amd.backend import gpu_available, hip_api
    if not gpu_available():
        pytest.skip("no GPU")
    return hip_api()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def make_inputs(seed, shape, B=2):
    rng = np.random.default_rng(seed)
    feat = rng.normal(size=(B,) + shape + (4,)).astype(np.float32)
    feat[..., 3] = rng.uniform(-1, 0, size=(B,) + shape)
    tgt = rng.normal(size=(B,) + shape + (3,)).astype(np.float32)
    mask = (rng.random((B,) + shape) > 0.15).astype(np.uint8)
    return feat, tgt, mask


def oracle_params(seed):
    params = U.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    params = [(W, rng.normal(0, 0.05, b.shape)) for W, b in params]
    params[-1] = (params[-1][0] * 10, params[-1][1])
    theta = U.pack(params).astype(np.float32)
    return U.unpack(theta.astype(np.float64)), theta


@pytest.mark.parametrize("shape", [(5, 7, 21), (9, 6, 11)])
def test_unet_forward_matches_oracle(hip, shape):
    from topoptspline3d_amd.learned import UNet
    print(f"\n[unet forward] shape={shape}", flush=True)
    params, theta = oracle_params(11)
    net = UNet(hip, seed=1)
    net.set_params(theta)
    feat, _, _ = make_inputs(3, shape)
    y = net.forward(feat)
    y_ref, _ = U.forward(params, feat.astype(np.float64), bf=True)
    err = rel_l2(y, y_ref)
    print(f"rel L2 {err:.3e}")
    assert err < 3e-2


@pytest.mark.parametrize("shape", [(5, 7, 21), (8, 5, 9)])
def test_unet_gradient_matches_oracle(hip, shape):
    from topoptspline3d_amd.learned import UNet
    print(f"\n[unet gradient] shape={shape}", flush=True)
    params, theta = oracle_params(12)
    net = UNet(hip, seed=1)
    net.set_params(theta)
    feat, tgt, mask = make_inputs(4, shape)
    loss, g = net.loss_grad(feat, tgt, mask)
    loss_ref, grads_ref = U.loss_and_grad(params, feat.astype(np.float64), tgt.astype(np.float64), mask, bf=True)
    assert abs(loss - loss_ref) <= 1e-2 * loss_ref
    off = 0
    for li, (dW, db) in enumerate(grads_ref):
        nw, nb = dW.size, db.size
        ew = rel_l2(g[off:off + nw], dW.ravel())
        eb = rel_l2(g[off + nw:off + nw + nb], db)
        print(f"layer {U.LAYERS[li]}: dW {ew:.2e} db {eb:.2e}")
        assert ew < 3e-2 and eb < 3e-2, (U.LAYERS[li], ew, eb)
        off += nw + nb
    assert off == g.size


def test_learned_correction_trains_and_is_safe(hip):
    """Log level-2 pairs from the solver on a 64x32x32 cantilever over a few OC
    iterations, train on 80 %, and compare held-out loss and FCG iterations.  The trained
    network learns the level-2 error (held-out loss falls by well over 30 %) and its
    corrections are kept, but the FCG iteration count does not fall (13 with and without
    it when measured): after the coarse-grid correction the level-2 error left is small
    next to the fine-level error that sets the iteration count.  The test therefore checks
    what the guarded correction promises -- the same solution to the same tolerance at no
    more than two extra iterations -- and reports the counts; the 511 benchmark measures
    the learned and plain runs over the same iterations."""
    from topoptspline3d_amd import learned as LN
    nx, ny, nz = 64, 32, 32
    solid, support, f = cantilever(nx, ny, nz)
    ctx = Context(hip, solid, support)
    ctx.set_material(1e6, 0.4, 3.0, 1e-3)
    ctx.set_loads(f)
    box = LN.level_box(ctx, 2)
    ds = LN.PairDataset(hip, box, 600)
    LN.log_pairs(ctx, ds, level=2, n_inner=8, stride=1)
    rng = np.random.default_rng(0)
    rho = np.full(nx * ny * nz, 0.4, np.float32)
    its_log = []
    for it in range(8):
        ctx.set_density(np.clip(rho * rng.uniform(0.6, 1.4, rho.size), 1e-3, 1).astype(np.float32))
        ctx.update_operator()
        st = ctx.solve(rtol=1e-6, maxit=200, precision=1, warm_start=False)
        its_log.append(st["iterations"][0])
    LN.log_pairs(ctx, None)
    n = len(ds)
    print(f"\n[learned] {n} pairs from {its_log} FCG iterations, box {box}", flush=True)
    assert n >= 40
    ntr = int(0.8 * n)
    net = LN.UNet(hip, seed=3)
    before = net.evaluate(ds, ntr, n - ntr)
    lf, ll = net.train(ds, 0, ntr, steps=300, batch=8, lr=2e-3, seed=1)
    after = net.evaluate(ds, ntr, n - ntr)
    print(f"held-out loss {before:.4f} -> {after:.4f}; train {lf:.4f} -> {ll:.4f}")
    assert after < 0.7 * before
    # same system from zero: plain vs learned-corrected FCG
    st0 = ctx.solve(rtol=1e-6, maxit=200, precision=1, warm_start=False)
    c0 = ctx.compliance()[0]
    LN.attach(ctx, net, level=2)
    st1 = ctx.solve(rtol=1e-6, maxit=200, precision=1, warm_start=False)
    c1 = ctx.compliance()[0]
    tried, acc, step = LN.stats(ctx)
    LN.attach(ctx, None)
    print(f"FCG iterations plain {st0['iterations'][0]} learned {st1['iterations'][0]}; "
          f"corrections kept {acc}/{tried}, mean step {step:.3f}; "
          f"compliance rel diff {abs(c1 - c0) / c0:.2e}")
    assert st1["rel_res"][0] <= 1e-6
    assert abs(c1 - c0) <= 1e-5 * c0
    assert tried > 0 and acc > 0 and step > 0
    assert st1["iterations"][0] <= st0["iterations"][0] + 2


def test_zero_and_garbage_networks_fall_back_to_plain_mg(hip):
    """A network with all-zero parameters proposes e = 0 (e.r = 0: never kept), so the
    solve is plain MG-PCG: same iteration count and solution.  A network with huge random
    parameters proposes garbage; the guard keeps only energy-reducing steps in the
    accepted window, and flexible CG still converges to the same solution."""
    from topoptspline3d_amd import learned as LN
    solid, support, f = cantilever(64, 32, 32)  # 4 levels: level 2 has one below
    ctx = Context(hip, solid, support)
    ctx.set_material(1e6, 0.4, 3.0, 1e-3)
    ctx.set_loads(f)
    rng = np.random.default_rng(4)
    ctx.set_density(rng.uniform(0.05, 1.0, solid.size).astype(np.float32))
    ctx.update_operator()
    st0 = ctx.solve(rtol=1e-6, maxit=200, precision=1, warm_start=False)
    u0, c0 = ctx.displacement(), ctx.compliance()[0]
    net = LN.UNet(hip, seed=5)
    cases = {"zero": np.zeros(net.n_params, np.float32),
             "garbage": rng.normal(0, 100, net.n_params).astype(np.float32)}
    for tag, theta in cases.items():
        net.set_params(theta)
        LN.attach(ctx, net, level=2)
        st = ctx.solve(rtol=1e-6, maxit=200, precision=1, warm_start=False)
        tried, kept, _ = LN.stats(ctx)
        LN.attach(ctx, None)
        c = ctx.compliance()[0]
        print(f"{tag}: iterations {st['iterations'][0]} (plain {st0['iterations'][0]}), tried {tried} kept {kept}")
        assert st["rel_res"][0] <= 1e-6 and abs(c - c0) <= 1e-5 * c0
        if tag == "zero":
            assert kept == 0 and tried > 0
            assert st["iterations"][0] == st0["iterations"][0]
            assert np.array_equal(ctx.displacement(), u0)


def test_precision_3_refuses_an_attached_network(hip):
    """The FP32-residual solve (precision 3) has no learned-correction path: with a
    network attached it returns TOS_E_ARG; after detaching it solves."""
    from topoptspline3d_amd import learned as LN
    from topoptspline3d_amd.context import TosError
    solid, support, f = cantilever(64, 32, 32)  # 4 levels: level 2 has one below
    ctx = Context(hip, solid, support)
    ctx.set_material(1e6, 0.4, 3.0, 1e-3)
    ctx.set_loads(f)
    ctx.set_density(np.full(solid.size, 0.4, np.float32))
    ctx.update_operator()
    net = LN.UNet(hip, seed=5)
    LN.attach(ctx, net, level=2)
    with pytest.raises(TosError, match="precision 3"):
        ctx.solve(rtol=1e-6, maxit=200, precision=3)
    LN.attach(ctx, None)
    st = ctx.solve(rtol=1e-6, maxit=200, precision=3)
    assert st["rel_res"][0] <= 1e-6
