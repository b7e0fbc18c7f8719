"""Worst-case load search (tos_worst_compliance, the modified power method of
optimization/optimization.cpp:213-351) on the GPU through the C ABI, against the FP64
oracle on the same seeded start force, and by size-independent properties at a larger
size.

Tolerances: start force 1e-15 absolute (unit vector; the norm is summed in another
order); FP64 cycles: same number of steps (within one), compliance 1e-8 relative, force
2e-4 absolute (one step of the stop rule's force change); FP32 / FP16-stencil cycles:
compliance within 1e-4 of the FP64 result (the stop rule's own accuracy); at the larger
size the returned force is checked against an accurate solve: f . K^-1 f within 1e-3 of
the reported compliance and the projected displacement parallel to f within 1e-2."""
import numpy as np
import pytest

from oracle_api import oracle_api
from topoptspline3d_amd.context import Context
from topoptspline3d_amd.domains import cantilever, face_load_region

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from topoptspline3d_amd.backend import gpu_available, hip_api
    if not gpu_available():
        pytest.skip("no GPU")
    return hip_api()


def make(api, nx, ny, nz, seed=0, void=0.0):
    solid, support, f = cantilever(nx, ny, nz)
    rng = np.random.default_rng(seed)
    if void:
        solid = (rng.random(solid.shape) > void).astype(np.uint8)
        solid[:, :, 0] = 1
        solid[:, :, -1] = 1
    c = Context(api, solid, support)
    c.set_material(1e6, 0.4, 3.0, 1e-3)
    c.set_loads(f)
    c.set_density(rng.uniform(0.1, 1.0, solid.size).astype(np.float32))
    c.update_operator()
    return c


def test_start_force_matches_oracle(hip):
    dims = (24, 12, 12)
    print(f"\n[worst start] dims={dims}", flush=True)
    ids, nrm = face_load_region(*dims, "z+", inset=2)
    for normals in (None, nrm + np.array([0.3, -0.2, 0.0])):  # tilted normals: a real projection
        out = []
        for api in (hip, oracle_api()):
            c = make(api, *dims, seed=1)
            st = c.worst_compliance(ids, normals=normals, seed=11, max_iter=0)
            assert st["iterations"] == 0
            out.append((c.get_loads(0), c.displacement(0)))
        assert np.abs(out[0][0] - out[1][0]).max() <= 1e-15
        assert not out[0][1].any()


@pytest.mark.parametrize("face,with_normals,void", [("x+", False, 0.0), ("z+", True, 0.0), ("x+", True, 0.2)])
def test_power_method_matches_oracle(hip, face, with_normals, void):
    dims = (24, 12, 12)
    print(f"\n[worst vs oracle] dims={dims} face={face} normals={with_normals} void={void}", flush=True)
    ids, nrm = face_load_region(*dims, face, inset=1)
    res = []
    for api in (hip, oracle_api()):
        c = make(api, *dims, seed=2, void=void)
        st = c.worst_compliance(ids, normals=nrm if with_normals else None, seed=5, precision=0)
        res.append((st, c.get_loads(0), c.displacement(0), c.compliance()[0]))
    (sg, fg, ug, cg), (so, fo, uo, co) = res
    print(f"gpu: its {sg['iterations']} c {sg['compliance']:.10e} fch {sg['fch']:.2e} res {sg['rel_res']:.2e}; "
          f"oracle: its {so['iterations']} c {so['compliance']:.10e}", flush=True)
    assert sg["fch"] <= 1e-4 and sg["rel_res"] <= 1e-2
    assert abs(sg["iterations"] - so["iterations"]) <= 1
    if sg["iterations"] == so["iterations"]:
        assert abs(sg["compliance"] - so["compliance"]) <= 1e-8 * abs(so["compliance"])
        assert np.abs(fg - fo).max() <= 2e-4
    else:
        assert abs(sg["compliance"] - so["compliance"]) <= 1e-4 * abs(so["compliance"])
    assert abs(cg - sg["compliance"]) <= 1e-12 * abs(cg)  # load case 0 holds the result
    assert np.count_nonzero(fg.reshape(3, -1)[:, np.setdiff1d(np.arange(fg.size // 3), ids)]) == 0


@pytest.mark.parametrize("precision", [1, 2])
def test_mixed_precision_cycles_reach_same_worst_load(hip, precision):
    dims = (24, 12, 12)
    print(f"\n[worst precision {precision}] dims={dims}", flush=True)
    ids, nrm = face_load_region(*dims, "x+", inset=1)
    c = make(hip, *dims, seed=3)
    s0 = c.worst_compliance(ids, seed=9, precision=0)
    s1 = c.worst_compliance(ids, seed=9, precision=precision)
    print(f"fp64 its {s0['iterations']} c {s0['compliance']:.8e}; precision {precision} its {s1['iterations']} "
          f"c {s1['compliance']:.8e}", flush=True)
    assert s1["fch"] <= 1e-4 and s1["rel_res"] <= 1e-2
    assert abs(s1["compliance"] - s0["compliance"]) <= 1e-4 * s0["compliance"]


def test_large_worst_load_is_an_eigenpair(hip):
    """128x64x64 cantilever: the returned force f and an accurate u = K^-1 f satisfy
    f . u ~ reported compliance and P u ~ (f . u) f (P: the normal projection)."""
    dims = (128, 64, 64)
    print(f"\n[worst large] dims={dims}", flush=True)
    ids, nrm = face_load_region(*dims, "x+", inset=0)
    c = make(hip, *dims, seed=4)
    c.set_galerkin_precision(1)
    c.update_operator()
    st = c.worst_compliance(ids, normals=nrm, seed=21, precision=1)
    its = st["iterations"]
    print(f"its {its} c {st['compliance']:.8e} fch {st['fch']:.2e} res {st['rel_res']:.2e} "
          f"{st['seconds'] * 1e3 / max(its, 1):.2f} ms/step", flush=True)
    st2 = c.worst_compliance(ids, normals=nrm, seed=21, precision=1)
    assert st2["iterations"] == its and st2["compliance"] == st["compliance"]  # deterministic
    f = c.get_loads(0)
    sv = c.solve(rtol=1e-10, maxit=300, precision=0, warm_start=True)
    assert sv["rel_res"][0] <= 1e-10
    u = c.displacement(0)
    cu = f @ u
    print(f"accurate f.K^-1 f {cu:.8e}", flush=True)
    assert abs(cu - st["compliance"]) <= 1e-3 * cu
    Pu = np.zeros_like(u)
    Pu.reshape(3, -1)[0, ids] = u.reshape(3, -1)[0, ids]  # normal (1, 0, 0)
    assert np.linalg.norm(Pu - cu * f) <= 1e-2 * np.linalg.norm(Pu)


@pytest.mark.parametrize("case,reso", [("bench_cube24x12x12/config11.json", 48), ("bench_fertility/config1.json", 64)])
def test_bench_config_domain_matches_oracle(hip, case, reso):
    """The reference's bench inputs (mesh voxelised, regions and load field from the JSON):
    the box cube24x12x12/config11 at resolution 48 and the irregular fertility mesh
    (config1) at 64; two iterations of the worst-case loop and of the fixed-load OC loop,
    GPU against the oracle."""
    import os
    from oracle_api import ROOT
    from topoptspline3d_amd import mesh_domain
    from topoptspline3d_amd.drivers import run_oc, run_worst_oc
    d = mesh_domain.load_config(os.path.join(ROOT, "tests", "golden", *case.split("/")), reso)
    print(f"\n[bench config] solid {d['solid'].shape} supports {int(d['support'].sum())} "
          f"load nodes {d['load_nodes'].size}", flush=True)
    rg = run_worst_oc(hip, 0, 0, 0, iters=2, domain=d)
    ro = run_worst_oc(oracle_api(), 0, 0, 0, iters=2, domain=d)
    print("worst gpu", rg["c"], rg["its"], "oracle", ro["c"], ro["its"], flush=True)
    assert np.allclose(rg["c"], ro["c"], rtol=1e-6, atol=0)
    dom = (d["solid"], d["support"], d["f"], d["shell"])
    og = run_oc(hip, 0, 0, 0, iters=2, rtol=1e-8, precision=1, domain=dom)
    oo = run_oc(oracle_api(), 0, 0, 0, iters=2, rtol=1e-8, precision=0, domain=dom)
    print("oc gpu", og["c"], "oracle", oo["c"], flush=True)
    assert np.allclose(og["c"], oo["c"], rtol=1e-6, atol=0)


def test_bad_load_regions_rejected_like_the_oracle(hip):
    """repeated, supported, out-of-range load nodes and zero normals: TOS_E_ARG from both
    libraries; an operator-less context: TOS_E_STATE; the context stays usable after"""
    from topoptspline3d_amd.context import TosError
    dims = (24, 12, 12)
    ids, nrm = face_load_region(*dims, "x+")
    for api in (hip, oracle_api()):
        c = make(api, *dims)
        nn = c.nn
        for nodes, normals in ((np.r_[ids, ids[:1]], None), (np.array([0]), None), (np.array([nn]), None),
                               (ids, np.zeros((ids.size, 3)))):
            with pytest.raises(TosError) as e:
                c.worst_compliance(nodes, normals=normals)
            assert e.value.code == -1
        st = c.worst_compliance(ids, normals=nrm, seed=1)
        assert st["fch"] <= 1e-4 and st["rel_res"] <= 1e-2 and st["compliance"] > 0
        solid, support, _ = cantilever(*dims)
        fresh = Context(api, solid, support)
        fresh.set_material(1e6, 0.4, 3.0, 1e-3)
        with pytest.raises(TosError) as e:
            fresh.worst_compliance(ids)
        assert e.value.code == -6
