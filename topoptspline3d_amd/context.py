"""numpy-facing wrapper of one design-domain context of the C ABI (include/tos_api.h).

`Context(api, ...)` works over any library bound with `_abi.bind`; the product's
`topoptspline3d_amd.backend.hip_api()` is the HIP library.
"""
import ctypes as C

import numpy as np

from . import _abi


class TosError(RuntimeError):
    def __init__(self, code, where, msg):
        super().__init__(f"{where}: {_abi.ERRORS.get(code, code)} ({code}): {msg}")
        self.code = code


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class Context:
    """One voxel design domain: hierarchy, loads, densities, sensitivities, spline."""

    def __init__(self, api, solid, support=None, shell=None, h=1.0, n_levels=0,
                 coarsen_threshold=0, skip_layer=False):
        self.api = api
        solid = np.ascontiguousarray(solid, dtype=np.uint8)
        nz, ny, nx = solid.shape  # arrays are indexed [z, y, x] (x fastest in memory)
        self.shape_e = (nz, ny, nx)
        self.shape_n = (nz + 1, ny + 1, nx + 1)
        self._keep = [solid]
        d = _abi.GridDesc()
        d.nx, d.ny, d.nz, d.h = nx, ny, nz, float(h)
        d.solid = _p(solid, C.c_uint8)
        if support is not None:
            support = np.ascontiguousarray(support, dtype=np.uint8)
            assert support.shape == self.shape_n
            d.support = _p(support, C.c_uint8)
            self._keep.append(support)
        if shell is not None:
            shell = np.ascontiguousarray(shell, dtype=np.uint8)
            assert shell.shape == self.shape_e
            d.shell = _p(shell, C.c_uint8)
            self._keep.append(shell)
        d.n_levels = int(n_levels)
        d.coarsen_threshold = int(coarsen_threshold)
        d.skip_layer = int(bool(skip_layer))
        h_ = C.c_void_p()
        rc = api.ctx_create(C.byref(d), C.byref(h_))
        if rc != 0:
            raise TosError(rc, "ctx_create", "")
        self.h = h_
        self.ne = nx * ny * nz
        self.nn = (nx + 1) * (ny + 1) * (nz + 1)
        self.K = 1

    # -------------------------------------------------------------------------------
    def _chk(self, rc, where):
        if rc != 0:
            msg = self.api.last_error(self.h)
            raise TosError(rc, where, msg.decode() if msg else "")
        return rc

    def close(self):
        if self.h:
            self.api.ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        i = _abi.Info()
        self._chk(self.api.get_info(self.h, C.byref(i)), "get_info")
        n = i.n_levels
        return {
            "n_levels": n,
            "level_dims": [tuple(i.level_dims[l]) for l in range(n)],
            "n_active_elements": [i.n_active_elements[l] for l in range(n)],
            "n_free_nodes": [i.n_free_nodes[l] for l in range(n)],
            "n_elements": i.n_elements, "n_nodes": i.n_nodes,
            "coarse_dofs": i.coarse_dofs, "n_cases": i.n_cases, "n_coeffs": i.n_coeffs,
            "has_support": bool(i.has_support),
        }

    def node_flags(self, level=0):
        dims = self.info()["level_dims"][level]
        out = np.zeros((dims[2] + 1) * (dims[1] + 1) * (dims[0] + 1), np.uint8)
        self._chk(self.api.get_node_flags(self.h, level, _p(out, C.c_uint8)), "get_node_flags")
        return out

    def set_material(self, E=1e6, nu=0.4, penal=3.0, rho_min=1e-3):
        self._chk(self.api.set_material(self.h, E, nu, penal, rho_min), "set_material")

    def template_matrix(self):
        ke = np.zeros(576)
        self._chk(self.api.get_template_matrix(self.h, _p(ke, C.c_double)), "get_template_matrix")
        return ke.reshape(24, 24)

    def set_loads(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        K = f.size // (3 * self.nn)
        assert f.size == K * 3 * self.nn
        self._chk(self.api.set_loads(self.h, K, _p(f, C.c_double)), "set_loads")
        self.K = K

    def get_loads(self, k=0):
        f = np.zeros(3 * self.nn)
        self._chk(self.api.get_loads(self.h, k, _p(f, C.c_double)), "get_loads")
        return f

    def set_density(self, rho):
        rho = np.ascontiguousarray(rho, dtype=np.float32).ravel()
        assert rho.size == self.ne
        self._chk(self.api.set_density(self.h, _p(rho, C.c_float)), "set_density")

    def density(self):
        r = np.zeros(self.ne, np.float32)
        self._chk(self.api.get_density(self.h, _p(r, C.c_float)), "get_density")
        return r

    def update_operator(self):
        self._chk(self.api.update_operator(self.h), "update_operator")

    def stencil(self, level):
        dims = self.info()["level_dims"][level]
        nn = (dims[0] + 1) * (dims[1] + 1) * (dims[2] + 1)
        s = np.zeros(243 * nn)
        self._chk(self.api.get_stencil(self.h, level, _p(s, C.c_double)), "get_stencil")
        return s.reshape(27, 9, nn)

    def apply(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        y = np.zeros(3 * self.nn)
        self._chk(self.api.apply_operator(self.h, _p(u, C.c_double), _p(y, C.c_double)), "apply_operator")
        return y

    def vcycle(self, r, precision=0, pre=1, post=1):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros(3 * self.nn)
        self._chk(self.api.vcycle(self.h, _p(r, C.c_double), _p(z, C.c_double), precision, pre, post), "vcycle")
        return z

    def solve(self, rtol=1e-8, maxit=200, precision=0, pre=1, post=1, warm_start=True,
              use_pcg=True, check=True):
        o = _abi.SolveOpts(rtol, maxit, precision, pre, post, int(warm_start), int(use_pcg))
        st = _abi.SolveStats()
        rc = self.api.solve(self.h, C.byref(o), C.byref(st))
        if check:
            self._chk(rc, "solve")
        return {"rc": rc, "iterations": [st.iterations[k] for k in range(self.K)],
                "rel_res": [st.rel_res[k] for k in range(self.K)], "seconds": st.seconds}

    def solve_batched(self, rtol=1e-6, maxit=200, pre=1, post=1, warm_start=False, check=True):
        """all load cases in one interleaved FCG with an FP32 V-cycle (graph-captured)"""
        o = _abi.SolveOpts(rtol, maxit, 1, pre, post, int(warm_start), 1)
        st = _abi.SolveStats()
        rc = self.api.solve_batched(self.h, C.byref(o), C.byref(st))
        if check:
            self._chk(rc, "solve_batched")
        return {"rc": rc, "iterations": [st.iterations[k] for k in range(self.K)],
                "rel_res": [st.rel_res[k] for k in range(self.K)], "seconds": st.seconds}

    def set_sweep_mode(self, mode):
        """fine GS schedule: 0 colour passes, 1 colour-slab wavefront, 2 row pairs, 3 in-wave pairs
        (FP32 cycles); all perform the same updates"""
        self._chk(self.api.set_sweep_mode(self.h, int(mode)), "set_sweep_mode")

    def set_galerkin_precision(self, prec):
        """coarse operators built into FP64 (0, default) or FP32 storage (1, for precision 1/2
        solves; FP64 stencils then built on demand); takes effect at the next update_operator"""
        self._chk(self.api.set_galerkin_precision(self.h, int(prec)), "set_galerkin_precision")

    def set_graph_replay(self, enable):
        """replay the captured FCG iteration body (default) or launch kernel by kernel"""
        self._chk(self.api.set_graph_replay(self.h, int(bool(enable))), "set_graph_replay")

    def set_start_recycling(self, depth):
        """warm starts projected onto span{u, earlier solutions} (tos_set_start_recycling)"""
        self._chk(self.api.set_start_recycling(self.h, int(depth)), "set_start_recycling")

    def batch_tuning(self):
        buf = C.create_string_buffer(512)
        self._chk(self.api.batch_tuning(self.h, buf, 512), "batch_tuning")
        return buf.value.decode()

    def displacement(self, k=0):
        u = np.zeros(3 * self.nn)
        self._chk(self.api.get_displacement(self.h, k, _p(u, C.c_double)), "get_displacement")
        return u

    def set_displacement(self, u, k=0):
        u = np.ascontiguousarray(u, dtype=np.float64)
        self._chk(self.api.set_displacement(self.h, k, _p(u, C.c_double)), "set_displacement")

    def compliance(self):
        c = np.zeros(max(self.K, 1))
        self._chk(self.api.compliance(self.h, _p(c, C.c_double)), "compliance")
        return c

    def volume(self):
        v = C.c_double()
        self._chk(self.api.volume(self.h, C.byref(v)), "volume")
        return v.value

    def worst_compliance(self, nodes, normals=None, seed=0, precision=0, pre=1, post=1, max_iter=500,
                         fch_tol=1e-4, res_tol=1e-2, check=True):
        """worst-case load over the load nodes by the modified power method (modifiedPM,
        optimization/optimization.cpp:213-351); load case 0 then holds the worst force and
        its displacement.  normals: (n, 3) force directions, or None for a free direction."""
        nodes = np.ascontiguousarray(nodes, dtype=np.int64).ravel()
        o = _abi.WorstOpts()
        o.n_load, o.nodes, o.seed = nodes.size, _p(nodes, C.c_int64), int(seed)
        if normals is not None:
            normals = np.ascontiguousarray(normals, dtype=np.float64).reshape(nodes.size, 3)
            o.normals = _p(normals, C.c_double)
        o.precision, o.pre, o.post, o.max_iter = int(precision), int(pre), int(post), int(max_iter)
        o.fch_tol, o.res_tol = float(fch_tol), float(res_tol)
        st = _abi.WorstStats()
        rc = self.api.worst_compliance(self.h, C.byref(o), C.byref(st))
        if check:
            self._chk(rc, "worst_compliance")
        return {"rc": rc, "iterations": st.iterations, "compliance": st.compliance, "fch": st.fch,
                "rel_res": st.rel_res, "seconds": st.seconds}

    def balance_loads(self, nodes, normals=None, k=0):
        """forceProject (optimization/projection.cpp:289-350): keep load case k's force on the
        given nodes only, along the normals when given, and on a floating body without its
        resultant force and moment"""
        nodes = np.ascontiguousarray(nodes, dtype=np.int64).ravel()
        nrm = None
        if normals is not None:
            nrm = np.ascontiguousarray(normals, dtype=np.float64).reshape(nodes.size, 3)
        self._chk(self.api.balance_loads(self.h, k, nodes.size, _p(nodes, C.c_int64),
                                         _p(nrm, C.c_double) if nrm is not None else None), "balance_loads")

    def probe_element_mfma(self, reps=10):
        """times y = K u (FP32) by the matrix-free VALU kernel and by MFMA element products
        (products only / with a float-atomic scatter); returns ms each and the relative
        difference of the scattered result"""
        out = np.zeros(4)
        self._chk(self.api.probe_element_mfma(self.h, int(reps), _p(out, C.c_double)), "probe_element_mfma")
        return {"valu_ms": out[0], "mfma_products_ms": out[1], "mfma_scatter_ms": out[2], "rel_diff": out[3]}

    def remove_rigid(self, k=0):
        """displacementProject (optimization/projection.cpp:374-395), floating bodies only"""
        self._chk(self.api.remove_rigid(self.h, k), "remove_rigid")

    def sensitivity(self, k=0):
        self._chk(self.api.sensitivity(self.h, k), "sensitivity")

    def get_sensitivity(self, which=0):
        s = np.zeros(self.ne, np.float32)
        self._chk(self.api.get_sensitivity(self.h, which, _p(s, C.c_float)), "get_sensitivity")
        return s

    def set_sensitivity(self, s, which=0):
        s = np.ascontiguousarray(s, dtype=np.float32).ravel()
        self._chk(self.api.set_sensitivity(self.h, which, _p(s, C.c_float)), "set_sensitivity")

    def init_vol_sensitivity(self):
        self._chk(self.api.init_vol_sensitivity(self.h), "init_vol_sensitivity")

    def filter(self, radius, which=0):
        self._chk(self.api.filter(self.h, which, float(radius)), "filter")

    def element_compliance(self, k=0):
        out = np.zeros(self.ne)
        self._chk(self.api.element_compliance(self.h, k, _p(out, C.c_double)), "element_compliance")
        return out

    def oc_update(self, vgoal, move=0.03, damp=0.5):
        g = C.c_double()
        n = C.c_int()
        self._chk(self.api.oc_update(self.h, vgoal, move, damp, C.byref(g), C.byref(n)), "oc_update")
        return g.value, n.value

    def oc_update_projected(self, vgoal, move=0.03, damp=0.5, beta=8.0):
        """OC on the design variable behind the Heaviside projection (updateDensities2,
        optimization/optimization.cu:462-557); returns (threshold, trials, max change,
        mean change of the design variable)"""
        g = C.c_double()
        n = C.c_int()
        ch = np.zeros(2)
        self._chk(self.api.oc_update_projected(self.h, vgoal, move, damp, beta, C.byref(g), C.byref(n),
                                               _p(ch, C.c_double)), "oc_update_projected")
        return g.value, n.value, float(ch[0]), float(ch[1])

    def project_density(self, beta):
        self._chk(self.api.project_density(self.h, beta), "project_density")

    def projection_sensitivity(self, beta, which=0):
        self._chk(self.api.projection_sensitivity(self.h, which, beta), "projection_sensitivity")

    # spline ---------------------------------------------------------------------------
    def spline_init(self, mode, order, param):
        self._chk(self.api.spline_init(self.h, mode, order, param), "spline_init")
        self.n_coeffs = self.info()["n_coeffs"]
        return self.n_coeffs

    def set_coeffs(self, c):
        c = np.ascontiguousarray(c, dtype=np.float32).ravel()
        assert c.size == self.n_coeffs
        self._chk(self.api.spline_set_coeffs(self.h, _p(c, C.c_float)), "spline_set_coeffs")

    def coeffs(self):
        c = np.zeros(self.n_coeffs, np.float32)
        self._chk(self.api.spline_get_coeffs(self.h, _p(c, C.c_float)), "spline_get_coeffs")
        return c

    def spline_eval(self):
        self._chk(self.api.spline_eval(self.h), "spline_eval")

    def spline_adjoint(self, which=0):
        self._chk(self.api.spline_adjoint(self.h, which), "spline_adjoint")

    def oc_update_coeffs(self, vgoal, move=0.03, damp=0.5, cmin=1e-3, cmax=1.0):
        g = C.c_double()
        n = C.c_int()
        self._chk(self.api.oc_update_coeffs(self.h, vgoal, move, damp, cmin, cmax, C.byref(g), C.byref(n)),
                  "oc_update_coeffs")
        return g.value, n.value

    def surface_constraints(self, cells=64, beta=8.0, print_angle=0.75 * np.pi, ss_mode=4, drip_mode=4,
                            p_norm=8.0, hfunction_c=10.0, sigmoid_c=0.09, drip_beta=128.0, drip_mu=32.0,
                            drip_angle=160.0 / 180.0 * np.pi, drip_alpha=4.0, drip_beta_scale=10.0,
                            drip_version=0):
        """self-supporting and drip constraints on the spline's surface points (deal_surface_points,
        optimization/optimization.cpp:1103-1128); gradients via coeff_sens(2) / coeff_sens(3)"""
        o = _abi.SurfaceOpts(int(cells), float(beta), float(print_angle), int(ss_mode), int(drip_mode),
                             float(p_norm), float(hfunction_c), float(sigmoid_c), float(drip_beta),
                             float(drip_mu), float(drip_angle), float(drip_alpha), float(drip_beta_scale),
                             int(drip_version))
        st = _abi.SurfaceStats()
        self._chk(self.api.surface_constraints(self.h, C.byref(o), C.byref(st)), "surface_constraints")
        return {"n_points": st.n_points, "n_down": st.n_down, "direction": st.direction,
                "ss_value": st.ss_value, "drip_value": st.drip_value}

    def surface_points(self, n):
        p = np.zeros((n, 3))
        nv = np.zeros((n, 3))
        self._chk(self.api.surface_points(self.h, n, _p(p, C.c_double), _p(nv, C.c_double)), "surface_points")
        return p, nv

    def coeff_sens(self, which=0):
        c = np.zeros(self.n_coeffs, np.float32)
        self._chk(self.api.spline_get_coeff_sens(self.h, which, _p(c, C.c_float)), "spline_get_coeff_sens")
        return c


class MMA:
    """Method of moving asymptotes (MMA::mma_t, optimization/mma_t.h:159-266)."""

    def __init__(self, api, n, m, xmin, xmax, move=0.5):
        self.api = api
        self.n, self.m = n, m
        lo = np.ascontiguousarray(np.broadcast_to(np.asarray(xmin, np.float64), (n,)))
        hi = np.ascontiguousarray(np.broadcast_to(np.asarray(xmax, np.float64), (n,)))
        h = C.c_void_p()
        rc = api.mma_create(n, m, _p(lo, C.c_double), _p(hi, C.c_double), float(move), C.byref(h))
        if rc != 0:
            raise TosError(rc, "mma_create", "")
        self.h = h

    def close(self):
        if self.h:
            self.api.mma_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_x(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        rc = self.api.mma_set_x(self.h, _p(x, C.c_double))
        if rc:
            raise TosError(rc, "mma_set_x", "")

    def x(self):
        x = np.zeros(self.n)
        rc = self.api.mma_get_x(self.h, _p(x, C.c_double))
        if rc:
            raise TosError(rc, "mma_get_x", "")
        return x

    def update(self, df, dg, g):
        df = np.ascontiguousarray(df, dtype=np.float64)
        dg = np.ascontiguousarray(dg, dtype=np.float64).ravel()
        g = np.ascontiguousarray(np.atleast_1d(g), dtype=np.float64)
        cv = C.c_int()
        rc = self.api.mma_update(self.h, _p(df, C.c_double), _p(dg, C.c_double), _p(g, C.c_double), C.byref(cv))
        if rc:
            raise TosError(rc, "mma_update", "")
        return bool(cv.value)

    def step_on_surface(self, ctx, sens_scale, vol_scale, vfrac, ss_scale, drip_scale):
        """m = 3: volume, self-supporting and drip constraints (tos_ctx_mma_step_surface)"""
        cv = C.c_int()
        rc = self.api.ctx_mma_step_surface(ctx.h, self.h, sens_scale, vol_scale, vfrac, ss_scale, drip_scale,
                                           C.byref(cv))
        if rc:
            raise TosError(rc, "ctx_mma_step_surface", "")
        return bool(cv.value)

    def step_on(self, ctx, sens_scale, vol_scale, vfrac):
        cv = C.c_int()
        rc = self.api.ctx_mma_step(ctx.h, self.h, sens_scale, vol_scale, vfrac, C.byref(cv))
        if rc:
            raise TosError(rc, "ctx_mma_step", "")
        return bool(cv.value)
