"""ctypes description of the C ABI declared in include/tos_api.h.

`bind(lib, prefix)` attaches argument/return types to every entry point of a loaded
shared library and returns a namespace of callables without the prefix.  The product
binds the HIP library with prefix ``tos_``.
"""
import ctypes as C
import types

u8p = C.POINTER(C.c_uint8)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int)


class GridDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("h", C.c_double),
                ("solid", u8p), ("support", u8p), ("shell", u8p),
                ("n_levels", C.c_int), ("coarsen_threshold", C.c_int), ("skip_layer", C.c_int)]


class Info(C.Structure):
    _fields_ = [("n_levels", C.c_int), ("level_dims", (C.c_int * 3) * 16),
                ("n_active_elements", C.c_int64 * 16), ("n_free_nodes", C.c_int64 * 16),
                ("n_elements", C.c_int64), ("n_nodes", C.c_int64),
                ("coarse_dofs", C.c_int), ("n_cases", C.c_int), ("n_coeffs", C.c_int64),
                ("has_support", C.c_int)]


class SolveOpts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("maxit", C.c_int), ("precision", C.c_int),
                ("pre", C.c_int), ("post", C.c_int), ("warm_start", C.c_int),
                ("use_pcg", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int * 64), ("rel_res", C.c_double * 64),
                ("seconds", C.c_double)]


class WorstOpts(C.Structure):
    _fields_ = [("n_load", C.c_int), ("nodes", C.POINTER(C.c_int64)), ("normals", C.POINTER(C.c_double)),
                ("seed", C.c_uint64), ("precision", C.c_int), ("pre", C.c_int), ("post", C.c_int),
                ("max_iter", C.c_int), ("fch_tol", C.c_double), ("res_tol", C.c_double)]


class WorstStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("compliance", C.c_double), ("fch", C.c_double),
                ("rel_res", C.c_double), ("seconds", C.c_double)]


class SurfaceOpts(C.Structure):
    _fields_ = [("cells", C.c_int), ("beta", C.c_double), ("print_angle", C.c_double), ("ss_mode", C.c_int),
                ("drip_mode", C.c_int), ("p_norm", C.c_double), ("hfunction_c", C.c_double),
                ("sigmoid_c", C.c_double), ("drip_beta", C.c_double), ("drip_mu", C.c_double),
                ("drip_angle", C.c_double), ("drip_alpha", C.c_double), ("drip_beta_scale", C.c_double),
                ("drip_version", C.c_int)]


class SurfaceStats(C.Structure):
    _fields_ = [("n_points", C.c_int64), ("n_down", C.c_int64), ("direction", C.c_double),
                ("ss_value", C.c_double), ("drip_value", C.c_double)]


class KernelStats(C.Structure):
    _fields_ = [("n_classes", C.c_int), ("names", (C.c_char * 32) * 20),
                ("total_ms", C.c_double * 20), ("launches", C.c_int64 * 20)]


vp = C.c_void_p
vpp = C.POINTER(C.c_void_p)

# name -> (restype, argtypes)
SIGNATURES = {
    "ctx_create": (C.c_int, [C.POINTER(GridDesc), vpp]),
    "ctx_destroy": (C.c_int, [vp]),
    "last_error": (C.c_char_p, [vp]),
    "get_info": (C.c_int, [vp, C.POINTER(Info)]),
    "get_node_flags": (C.c_int, [vp, C.c_int, u8p]),
    "set_material": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_double]),
    "get_template_matrix": (C.c_int, [vp, f64p]),
    "set_loads": (C.c_int, [vp, C.c_int, f64p]),
    "get_loads": (C.c_int, [vp, C.c_int, f64p]),
    "set_density": (C.c_int, [vp, f32p]),
    "get_density": (C.c_int, [vp, f32p]),
    "update_operator": (C.c_int, [vp]),
    "get_stencil": (C.c_int, [vp, C.c_int, f64p]),
    "apply_operator": (C.c_int, [vp, f64p, f64p]),
    "vcycle": (C.c_int, [vp, f64p, f64p, C.c_int, C.c_int, C.c_int]),
    "solve": (C.c_int, [vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "get_displacement": (C.c_int, [vp, C.c_int, f64p]),
    "set_displacement": (C.c_int, [vp, C.c_int, f64p]),
    "compliance": (C.c_int, [vp, f64p]),
    "volume": (C.c_int, [vp, f64p]),
    "worst_compliance": (C.c_int, [vp, C.POINTER(WorstOpts), C.POINTER(WorstStats)]),
    "balance_loads": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "remove_rigid": (C.c_int, [vp, C.c_int]),
    "probe_element_mfma": (C.c_int, [vp, C.c_int, C.POINTER(C.c_double)]),
    "sensitivity": (C.c_int, [vp, C.c_int]),
    "get_sensitivity": (C.c_int, [vp, C.c_int, f32p]),
    "set_sensitivity": (C.c_int, [vp, C.c_int, f32p]),
    "init_vol_sensitivity": (C.c_int, [vp]),
    "filter": (C.c_int, [vp, C.c_int, C.c_double]),
    "element_compliance": (C.c_int, [vp, C.c_int, f64p]),
    "oc_update": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, f64p, i32p]),
    "oc_update_projected": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_double, f64p, i32p, f64p]),
    "project_density": (C.c_int, [vp, C.c_double]),
    "projection_sensitivity": (C.c_int, [vp, C.c_int, C.c_double]),
    "spline_init": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
    "spline_set_coeffs": (C.c_int, [vp, f32p]),
    "spline_get_coeffs": (C.c_int, [vp, f32p]),
    "spline_eval": (C.c_int, [vp]),
    "spline_adjoint": (C.c_int, [vp, C.c_int]),
    "spline_get_coeff_sens": (C.c_int, [vp, C.c_int, f32p]),
    "oc_update_coeffs": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, f64p, i32p]),
    "mma_create": (C.c_int, [C.c_int, C.c_int, f64p, f64p, C.c_double, vpp]),
    "mma_destroy": (C.c_int, [vp]),
    "mma_set_x": (C.c_int, [vp, f64p]),
    "mma_get_x": (C.c_int, [vp, f64p]),
    "mma_update": (C.c_int, [vp, f64p, f64p, f64p, i32p]),
    "ctx_mma_step": (C.c_int, [vp, vp, C.c_double, C.c_double, C.c_double, i32p]),
    "ctx_mma_step_surface": (C.c_int, [vp, vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, i32p]),
    "surface_constraints": (C.c_int, [vp, C.POINTER(SurfaceOpts), C.POINTER(SurfaceStats)]),
    "surface_points": (C.c_int, [vp, C.c_int64, f64p, f64p]),
    "profile": (C.c_int, [vp, C.c_int, C.POINTER(KernelStats)]),
    "solve_batched": (C.c_int, [vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "set_sweep_mode": (C.c_int, [vp, C.c_int]),
    "set_galerkin_precision": (C.c_int, [vp, C.c_int]),
    "set_graph_replay": (C.c_int, [vp, C.c_int]),
    "set_start_recycling": (C.c_int, [vp, C.c_int]),
    "batch_tuning": (C.c_int, [vp, C.c_char_p, C.c_int]),
}

# learned coarse-level correction (HIP library only; the network oracle is numpy)
i64p = C.POINTER(C.c_int64)
LEARN_SIGNATURES = {
    "unet_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_uint64, vpp]),
    "unet_destroy": (C.c_int, [vp]),
    "unet_last_error": (C.c_char_p, [vp]),
    "unet_num_params": (C.c_int, [vp, i64p]),
    "unet_get_params": (C.c_int, [vp, f32p]),
    "unet_set_params": (C.c_int, [vp, f32p]),
    "unet_forward": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, f32p, f32p]),
    "unet_loss_grad": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, f32p, f32p, u8p, f64p, f32p]),
    "dataset_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, vpp]),
    "dataset_destroy": (C.c_int, [vp]),
    "dataset_size": (C.c_int, [vp, i32p]),
    "dataset_get": (C.c_int, [vp, C.c_int, f32p, f32p, u8p]),
    "unet_train": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64, f64p, f64p]),
    "unet_eval": (C.c_int, [vp, vp, C.c_int, C.c_int, f64p]),
    "ctx_learn_attach": (C.c_int, [vp, vp, C.c_int, C.c_int]),
    "ctx_learn_log": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int]),
    "ctx_learn_stats": (C.c_int, [vp, i32p, i32p, f64p]),
}

# entry points on caller-owned device arrays (HIP library only)
DEVICE_SIGNATURES = {
    "oc_update_noalloc": (C.c_int, [C.c_int64, vp, vp, vp, vp, vp, C.c_double, C.c_double, C.c_double,
                                    C.c_double, vp, f64p, i32p]),
}

ERRORS = {0: "ok", -1: "bad argument", -2: "HIP error", -3: "not converged",
          -4: "NaN/Inf", -5: "singular coarsest system", -6: "call order", -7: "beyond capacity"}


def bind(lib, prefix, learned=False):
    """learned=True: also the HIP library's own tables (network, device-array entry points)"""
    ns = types.SimpleNamespace()
    table = dict(SIGNATURES, **LEARN_SIGNATURES, **DEVICE_SIGNATURES) if learned else SIGNATURES
    for name, (res, args) in table.items():
        fn = getattr(lib, prefix + name)
        fn.restype = res
        fn.argtypes = args
        setattr(ns, name, fn)
    ns._lib = lib
    ns._prefix = prefix
    return ns


def exported_names(prefix, learned=True):
    names = list(SIGNATURES) + (list(LEARN_SIGNATURES) + list(DEVICE_SIGNATURES) if learned else [])
    return [prefix + n for n in names]
