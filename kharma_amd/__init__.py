"""kharma_amd -- MI355X-native GRMHD step (KHARMA's explicit step on gfx950).

Python host binding over the C ABI in include/kharma_amd.h.  Device memory and
streams come from PyTorch; every computation runs in the HIP kernels of
libkharma_amd.so.  There is no CPU fallback: importing the runtime raises if
the library was not built, and calls raise if the GPU is unavailable.

Layout of every cell array: float64 [8, n3, n2, n1] (entire block, ghosts included),
variables rho, u, u1, u2, u3, B1, B2, B3 -- the layout of KHARMA's VariablePack.
"""
import ctypes
import os
import re

from . import problems  # noqa: F401  (host-side initial conditions)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkharma_amd.so")
HEADER = os.path.abspath(os.path.join(HERE, "..", "include", "kharma_amd.h"))

NV = 8
PREC_FP64 = 0
PREC_SPLIT = 1
RECON = {"donor_cell_c": 0, "linear_mc": 1, "weno5": 2, "ppm": 3}
FRAME = {"fluid": 0, "normal": 1, "drift": 2}
COORDS = {"cartesian_minkowski": 0, "ks": 1}
TRANSFORM = {"null": 0, "mks": 1, "fmks": 2}
BC = {"periodic": 0, "outflow": 1, "reflecting": 2}
DRIVER = {"kharma": 0, "simple": 1}
INVERTER = {"onedw": 0, "kastaun": 1}
# b_field/solver and b_field/ct_scheme (b_ct.cpp:66-68)
BSOLVER = {"flux_ct": 0, "face_ct": 1}
CT_SCHEME = {"bs99": 0, "gs05_0": 1, "gs05_c": 2, "sg07": 3}


class KhParams(ctypes.Structure):
    _fields_ = [
        ("nx1", ctypes.c_int), ("nx2", ctypes.c_int), ("nx3", ctypes.c_int), ("ng", ctypes.c_int),
        ("x1min", ctypes.c_double), ("x1max", ctypes.c_double),
        ("x2min", ctypes.c_double), ("x2max", ctypes.c_double),
        ("x3min", ctypes.c_double), ("x3max", ctypes.c_double),
        ("periodic1", ctypes.c_int), ("periodic2", ctypes.c_int), ("periodic3", ctypes.c_int),
        ("gam", ctypes.c_double),
        ("cfl", ctypes.c_double), ("dt_min", ctypes.c_double), ("dt_start", ctypes.c_double),
        ("max_dt_increase", ctypes.c_double), ("tlim", ctypes.c_double),
        ("recon", ctypes.c_int), ("use_hlle", ctypes.c_int),
        ("floors_on", ctypes.c_int), ("recon_floors", ctypes.c_int), ("floor_frame", ctypes.c_int),
        ("rho_min_geom", ctypes.c_double), ("u_min_geom", ctypes.c_double),
        ("rho_min_const", ctypes.c_double), ("u_min_const", ctypes.c_double), ("r_char", ctypes.c_double),
        ("bsq_over_rho_max", ctypes.c_double), ("bsq_over_u_max", ctypes.c_double),
        ("u_over_rho_max", ctypes.c_double), ("ktot_max", ctypes.c_double), ("gamma_max", ctypes.c_double),
        ("use_r_char", ctypes.c_int), ("temp_adjust_u", ctypes.c_int),
        ("iter_max", ctypes.c_int), ("err_tol", ctypes.c_double),
        ("fix_average", ctypes.c_int), ("fix_atmo", ctypes.c_int),
        ("integrator", ctypes.c_int), ("coords", ctypes.c_int), ("precision", ctypes.c_int),
        ("a", ctypes.c_double), ("transform", ctypes.c_int),
        ("hslope", ctypes.c_double), ("mks_smooth", ctypes.c_double), ("poly_xt", ctypes.c_double),
        ("poly_alpha", ctypes.c_double), ("fmks_zero_point", ctypes.c_double),
        ("bc_inner_x1", ctypes.c_int), ("bc_outer_x1", ctypes.c_int),
        ("bc_inner_x2", ctypes.c_int), ("bc_outer_x2", ctypes.c_int),
        ("check_inflow_inner_x1", ctypes.c_int), ("check_inflow_outer_x1", ctypes.c_int),
        ("zero_polar_flux", ctypes.c_int), ("fix_polar_flux", ctypes.c_int), ("fix_corner_inner", ctypes.c_int),
        ("driver_type", ctypes.c_int),
        ("kill_on_large_divb", ctypes.c_int), ("kill_on_divb_over", ctypes.c_double),
        ("inverter", ctypes.c_int),
        ("fofc", ctypes.c_int), ("fofc_polar_cells", ctypes.c_int), ("fofc_eh_buffer", ctypes.c_double),
        ("b_solver", ctypes.c_int), ("ct_scheme", ctypes.c_int),
    ]


class KhTorusParams(ctypes.Structure):
    _fields_ = [
        ("rin", ctypes.c_double), ("rmax", ctypes.c_double), ("kappa", ctypes.c_double),
        ("u_jitter", ctypes.c_double), ("rng_seed", ctypes.c_int),
        ("beta_min", ctypes.c_double), ("min_A", ctypes.c_double),
    ]


class KhVars(ctypes.Structure):
    """kh_vars: per-variable base pointers (rho, u, u1, u2, u3, B1, B2, B3)."""
    _fields_ = [("var", ctypes.c_void_p * 8)]


class KhVarMap(ctypes.Structure):
    """kh_varmap: pack indices of rho, u, the first velocity and the first field component."""
    _fields_ = [("RHO", ctypes.c_int), ("UU", ctypes.c_int), ("U1", ctypes.c_int), ("B1", ctypes.c_int)]


class KharmaError(RuntimeError):
    pass


_lib = None


def header_functions():
    """Names of the functions declared in include/kharma_amd.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(kh_\w+)\s*\(", txt, flags=re.M)))


def lib():
    """Load libkharma_amd.so (build it with kharma_amd.build or __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise KharmaError(f"{LIB_PATH} is missing: build it with `python -m kharma_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, i, d, l = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long
    pvp = ctypes.POINTER(ctypes.c_void_p)
    pv = ctypes.POINTER(KhVars)
    sig = {
        "kh_last_error": (ctypes.c_char_p, []),
        "kh_version": (i, []),
        "kh_block_create": (i, [ctypes.POINTER(KhParams), pvp]),
        "kh_block_destroy": (None, [vp]),
        "kh_block_set_stream": (i, [vp, vp]),
        "kh_block_shape": (i, [vp, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(l)]),
        "kh_get_flux": (i, [vp, i, vp, vp, vp, vp]),
        "kh_flux_ct": (i, [vp, vp, vp]),
        "kh_flux_divergence": (i, [vp, vp, vp]),
        "kh_weighted_sum": (i, [vp, vp, vp, vp, d, d, d, vp]),
        "kh_add_geo_source": (i, [vp, vp, vp]),
        "kh_utop": (i, [vp, vp, vp, vp, vp]),
        "kh_apply_floors": (i, [vp, vp, vp, vp, vp]),
        "kh_fix_utop": (i, [vp, vp, vp, vp]),
        "kh_ptou": (i, [vp, vp, vp]),
        "kh_sync_bounds": (i, [vp, vp, i]),
        "kh_apply_bounds": (i, [vp, vp, vp]),
        "kh_torus_prims": (i, [ctypes.POINTER(KhParams), ctypes.POINTER(KhTorusParams), vp]),
        "kh_geometry_host": (i, [ctypes.POINTER(KhParams), i, i, vp]),
        "kh_estimate_timestep": (i, [vp, vp, vp, d, ctypes.POINTER(d)]),
        "kh_estimate_timestep_noalloc": (i, [vp, vp, vp, d, vp]),
        "kh_max_divb": (i, [vp, vp, ctypes.POINTER(d)]),
        "kh_block_face_count": (i, [vp, ctypes.POINTER(l)]),
        "kh_ct_calculate_emf": (i, [vp, vp, vp, vp]),
        "kh_ct_calculate_emf_v": (i, [vp, pv, pv, vp]),
        "kh_ct_add_source": (i, [vp, vp, vp]),
        "kh_ct_weighted_sum_face": (i, [vp, vp, vp, d, d, vp]),
        "kh_ct_sync_faces": (i, [vp, vp]),
        "kh_ct_utop": (i, [vp, vp, vp, vp]),
        "kh_ct_utop_v": (i, [vp, vp, pv, pv]),
        "kh_ct_max_divb": (i, [vp, vp, ctypes.POINTER(d)]),
        "kh_driver_init_faces": (i, [vp, vp, vp]),
        "kh_driver_get_faces": (i, [vp, vp]),
        "kh_driver_create": (i, [ctypes.POINTER(KhParams), pvp]),
        "kh_driver_destroy": (None, [vp]),
        "kh_driver_set_stream": (i, [vp, vp]),
        "kh_driver_init_prims": (i, [vp, vp]),
        "kh_driver_step": (i, [vp, i]),
        "kh_driver_set_clock": (i, [vp, d, d, i]),
        "kh_driver_get_state": (i, [vp, vp, vp]),
        "kh_driver_get_scalars": (i, [vp, ctypes.POINTER(d), ctypes.POINTER(d), ctypes.POINTER(d),
                                      ctypes.POINTER(i), ctypes.POINTER(i)]),
        "kh_driver_get_dt_history": (i, [vp, ctypes.POINTER(d), i]),
        "kh_driver_get_flags": (i, [vp, vp, vp]),
        "kh_driver_block": (i, [vp, pvp]),
        "kh_driver_arrays": (i, [vp, pvp, pvp]),
        "kh_driver_set_kernel_timing": (i, [vp, i]),
        "kh_driver_kernel_times": (i, [vp, ctypes.POINTER(d), ctypes.POINTER(i), i]),
        "kh_adjoint_create": (i, [ctypes.POINTER(KhParams), pvp]),
        "kh_adjoint_destroy": (None, [vp]),
        "kh_adjoint_set_stream": (i, [vp, vp]),
        "kh_adjoint_step": (i, [vp, vp, vp, d, vp]),
        "kh_adjoint_step_recorded": (i, [vp, vp, vp, vp, d, vp]),
        "kh_adjoint_set_deferred_checks": (i, [vp, i]),
        "kh_adjoint_check": (i, [vp]),
        "kh_adjoint_last_times": (i, [vp, ctypes.POINTER(d), ctypes.POINTER(d)]),
        "kh_adjoint_flux": (i, [vp, vp, vp, vp]),
        "kh_adjoint_flux_div": (i, [vp, vp, vp]),
        "kh_adjoint_utop": (i, [vp, vp, vp, vp]),
        "kh_adjoint_ptou": (i, [vp, vp, vp, vp]),
        "kh_adjoint_bounds": (i, [vp, vp, vp]),
        "kh_adjoint_fix_flux": (i, [vp, vp, vp]),
        "kh_adjoint_geo_source": (i, [vp, vp, vp, vp]),
        "kh_adjoint_floors": (i, [vp, vp, vp, vp]),
        "kh_adjoint_fixup": (i, [vp, vp, vp, vp]),
        "kh_adjoint_set_face_flux_objective": (i, [vp, i, i, i, d]),
        "kh_driver_set_face_flux_objective": (i, [vp, i, i, i, d]),
        "kh_driver_set_stage_record": (i, [vp, vp]),
        "kh_driver_stage_record_bytes": (i, [vp, ctypes.POINTER(ctypes.c_size_t)]),
        "kh_driver_check": (i, [vp]),
        "kh_driver_save_restart": (i, [vp, ctypes.c_char_p]),
        "kh_driver_load_restart": (i, [vp, ctypes.c_char_p]),
        "kh_mesh_save_restart": (i, [vp, ctypes.c_char_p]),
        "kh_mesh_load_restart": (i, [vp, ctypes.c_char_p]),
        "kh_mesh_create": (i, [ctypes.POINTER(KhParams), i, i, i, pvp]),
        "kh_mesh_destroy": (None, [vp]),
        "kh_mesh_set_stream": (i, [vp, vp]),
        "kh_mesh_shape": (i, [vp, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i),
                              ctypes.POINTER(l)]),
        "kh_mesh_init_prims": (i, [vp, vp]),
        "kh_mesh_step": (i, [vp, i]),
        "kh_mesh_set_graph": (i, [vp, i]),
        "kh_mesh_set_packed": (i, [vp, i]),
        "kh_mesh_set_clock": (i, [vp, d, d, i]),
        "kh_mesh_get_state": (i, [vp, vp, vp]),
        "kh_mesh_get_scalars": (i, [vp, ctypes.POINTER(d), ctypes.POINTER(d), ctypes.POINTER(d),
                                    ctypes.POINTER(i), ctypes.POINTER(i)]),
        "kh_mesh_get_dt_history": (i, [vp, ctypes.POINTER(d), i]),
        "kh_mesh_check": (i, [vp]),
        "kh_mesh_block_arrays": (i, [vp, i, pvp, pvp]),
        "kh_mesh_block": (i, [vp, i, pvp]),
        "kh_vars_from_pack": (i, [ctypes.POINTER(ctypes.c_void_p), i, ctypes.POINTER(KhVarMap), ctypes.POINTER(KhVars)]),
        "kh_get_flux_v": (i, [vp, i, pv, pv, vp, vp]),
        "kh_flux_ct_v": (i, [vp, pv, vp]),
        "kh_flux_divergence_v": (i, [vp, pv, pv]),
        "kh_add_geo_source_v": (i, [vp, pv, pv]),
        "kh_weighted_sum_v": (i, [vp, pv, pv, pv, d, d, d, pv]),
        "kh_utop_v": (i, [vp, pv, pv, vp, vp]),
        "kh_apply_floors_v": (i, [vp, pv, pv, vp, vp]),
        "kh_fix_utop_v": (i, [vp, pv, pv, vp]),
        "kh_ptou_v": (i, [vp, pv, pv]),
        "kh_apply_bounds_v": (i, [vp, pv, pv]),
        "kh_sync_bounds_v": (i, [vp, pv, i]),
        "kh_max_divb_v": (i, [vp, pv, ctypes.POINTER(d)]),
        "kh_flags_to_real": (i, [vp, vp, vp, l]),
        "kh_flags_from_real": (i, [vp, vp, vp, l]),
        "kh_determine_floors_v": (i, [vp, pv, vp]),
        "kh_mark_fofc": (i, [vp, vp, vp]),
        "kh_fofc_v": (i, [vp, pv, vp, pv, vp, vp]),
        "kh_driver_objective": (i, [vp, ctypes.POINTER(d), i]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(status):
    if status != 0:
        raise KharmaError(f"kharma_amd error {status}: {lib().kh_last_error().decode()}")


def make_params(cfg):
    """kh_params from a run configuration dict (see kharma_amd.problems)."""
    p = KhParams()
    p.nx1, p.nx2, p.nx3 = cfg["nx"]
    p.ng = cfg.get("ng", 4)
    (p.x1min, p.x1max), (p.x2min, p.x2max), (p.x3min, p.x3max) = cfg["bounds"]
    bcs = cfg.get("boundaries", {})
    p.bc_inner_x1 = BC[bcs.get("inner_x1", "periodic")]
    p.bc_outer_x1 = BC[bcs.get("outer_x1", "periodic")]
    p.bc_inner_x2 = BC[bcs.get("inner_x2", "periodic")]
    p.bc_outer_x2 = BC[bcs.get("outer_x2", "periodic")]
    p.periodic1 = int(p.bc_inner_x1 == 0)
    p.periodic2 = int(p.bc_inner_x2 == 0)
    p.periodic3 = 1
    p.check_inflow_inner_x1 = int(cfg.get("check_inflow_inner_x1", False))
    p.check_inflow_outer_x1 = int(cfg.get("check_inflow_outer_x1", False))
    p.zero_polar_flux = int(cfg.get("zero_polar_flux", False))
    p.fix_polar_flux = int(cfg.get("fix_polar_flux", False))
    p.fix_corner_inner = int(cfg.get("fix_corner_inner", False))
    p.driver_type = DRIVER[cfg.get("driver", "kharma")]
    p.kill_on_large_divb = int(cfg.get("kill_on_large_divb", True))
    p.kill_on_divb_over = cfg.get("kill_on_divb_over", 1e-3)
    p.inverter = INVERTER[cfg.get("inverter", "onedw")]
    p.fofc = int(cfg.get("fofc", False))
    p.fofc_polar_cells = cfg.get("fofc_polar_cells", 0)
    p.fofc_eh_buffer = cfg.get("fofc_eh_buffer", 0.1)
    p.b_solver = BSOLVER[cfg.get("b_solver", "flux_ct")]
    p.ct_scheme = CT_SCHEME[cfg.get("ct_scheme", "gs05_c")]
    p.gam = cfg["gamma"]
    p.cfl = cfg.get("cfl", 0.9)
    p.dt_min = cfg.get("dt_min", 1e-5)
    p.dt_start = cfg.get("dt_start", p.dt_min)
    p.max_dt_increase = cfg.get("max_dt_increase", 2.0)
    p.tlim = cfg.get("tlim", 1e30)
    p.recon = RECON[cfg.get("recon", "weno5")]
    p.use_hlle = int(cfg.get("riemann", "llf") == "hlle")
    p.floors_on = int(cfg.get("floors", True))
    p.recon_floors = int(cfg.get("recon_floors", p.floors_on and cfg.get("recon", "weno5") == "weno5"))
    p.floor_frame = FRAME[cfg.get("floor_frame", "drift")]
    fl = cfg.get("floor_values", {})
    p.rho_min_geom = fl.get("rho_min_geom", 1e-6)
    p.u_min_geom = fl.get("u_min_geom", 1e-8)
    p.rho_min_const = fl.get("rho_min_const", 1e-8)
    p.u_min_const = fl.get("u_min_const", 1e-10)
    p.r_char = fl.get("r_char", 10.0)
    p.bsq_over_rho_max = fl.get("bsq_over_rho_max", 1e20)
    p.bsq_over_u_max = fl.get("bsq_over_u_max", 1e20)
    p.u_over_rho_max = fl.get("u_over_rho_max", 1e20)
    p.ktot_max = fl.get("ktot_max", 1e20)
    p.gamma_max = fl.get("gamma_max", 50.0)
    p.use_r_char = int(fl.get("use_r_char", False))
    p.temp_adjust_u = int(fl.get("temp_adjust_u", False))
    # inverter/iter_max and err_tol default per solver (inverter/inverter.cpp:65-70):
    # 25 and 1e-12 for kastaun, 8 and 1e-8 for onedw.
    kastaun = p.inverter == INVERTER["kastaun"]
    p.iter_max = cfg.get("iter_max", 25 if kastaun else 8)
    p.err_tol = cfg.get("err_tol", 1e-12 if kastaun else 1e-8)
    p.fix_average = int(cfg.get("fix_average", True))
    p.fix_atmo = int(cfg.get("fix_atmo", True))
    p.integrator = {"rk2": 0, "vl2": 1}[cfg.get("integrator", "rk2")]
    p.coords = COORDS[cfg.get("coords", "cartesian_minkowski")]
    p.precision = {"fp64": PREC_FP64, "split": PREC_SPLIT}[cfg.get("precision", "fp64")]
    p.a = cfg.get("a", 0.0)
    p.transform = TRANSFORM[cfg.get("transform", "null")]
    p.hslope = cfg.get("hslope", 0.3)
    p.mks_smooth = cfg.get("mks_smooth", 0.5)
    p.poly_xt = cfg.get("poly_xt", 0.82)
    p.poly_alpha = cfg.get("poly_alpha", 14.0)
    p.fmks_zero_point = cfg.get("fmks_zero_point", 0.0)
    return p


GEOM_LOCI = {"center": 0, "face1": 1, "face2": 2, "corner": 3}


def geometry_host(cfg, loc, comp):
    """One component of the block's geometry cache, computed on the host
    ([n2+1, n1+1] float64; see kh_geometry_host for the component numbering)."""
    import numpy as np
    p = make_params(cfg)
    n1 = p.nx1 + 2 * p.ng
    n2 = p.nx2 + 2 * p.ng if p.nx2 > 1 else 1
    out = np.empty((n2 + 1, n1 + 1))
    _check(lib().kh_geometry_host(ctypes.byref(p), GEOM_LOCI.get(loc, loc), comp,
                                  ctypes.c_void_p(out.ctypes.data)))
    return out


def torus_prims(cfg):
    """Initial primitives of the torus problem (host float64 [8, n3, n2, n1], interior filled)."""
    import numpy as np
    p = make_params(cfg)
    t = KhTorusParams()
    tc = cfg["torus"]
    t.rin, t.rmax, t.kappa = tc["rin"], tc["rmax"], tc["kappa"]
    t.u_jitter, t.rng_seed = tc["u_jitter"], tc["rng_seed"]
    t.beta_min, t.min_A = tc["beta_min"], tc["min_A"]
    ng = p.ng
    shape = (p.nx3 + 2 * ng if p.nx3 > 1 else 1, p.nx2 + 2 * ng if p.nx2 > 1 else 1, p.nx1 + 2 * ng)
    P = np.zeros((NV,) + shape)
    _check(lib().kh_torus_prims(ctypes.byref(p), ctypes.byref(t), ctypes.c_void_p(P.ctypes.data)))
    return P


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise KharmaError("kharma_amd needs a HIP device (torch.cuda.is_available() is False)")
    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def vars_of(arrays):
    """kh_vars from 8 device tensors, one per variable (rho, u, u1, u2, u3, B1, B2, B3),
    each [n3, n2, n1] and contiguous; the tensors may live anywhere."""
    arrays = list(arrays)
    if len(arrays) != NV:
        raise KharmaError(f"a pack view needs {NV} variables, got {len(arrays)}")
    v = KhVars()
    for n, a in enumerate(arrays):
        if not a.is_contiguous():
            raise KharmaError("pack variables must be contiguous [n3, n2, n1] arrays")
        v.var[n] = a.data_ptr()
    return v


def _vars3(F):
    arr = (KhVars * 3)()
    for d, Fd in enumerate(F):
        arr[d] = vars_of(Fd)
    return arr


class Block:
    """One meshblock's geometry + index ranges; task-level entry points."""

    def __init__(self, cfg):
        self.cfg = cfg
        self._p = make_params(cfg)
        h = ctypes.c_void_p()
        _check(lib().kh_block_create(ctypes.byref(self._p), ctypes.byref(h)))
        self.h = h
        n1, n2, n3, nc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_long()
        _check(lib().kh_block_shape(self.h, ctypes.byref(n1), ctypes.byref(n2), ctypes.byref(n3), ctypes.byref(nc)))
        self.shape = (n3.value, n2.value, n1.value)
        self.nc = nc.value
        self.split = self._p.precision == PREC_SPLIT

    def __del__(self):
        try:
            lib().kh_block_destroy(self.h)
        except Exception:
            pass

    @property
    def flux_dtype(self):
        torch = _torch()
        return torch.float32 if self.split else torch.float64

    def _bind(self):
        _check(lib().kh_block_set_stream(self.h, _stream()))

    def new_fluxes(self):
        torch = _torch()
        F = torch.zeros((3, NV) + self.shape, dtype=self.flux_dtype, device="cuda")
        cmax = torch.zeros((3,) + self.shape, dtype=self.flux_dtype, device="cuda")
        cmin = torch.zeros((3,) + self.shape, dtype=self.flux_dtype, device="cuda")
        return F, cmax, cmin

    def get_flux(self, d, P, F, cmax, cmin):
        self._bind()
        _check(lib().kh_get_flux(self.h, d, _ptr(P), _ptr(F[d - 1]), _ptr(cmax[d - 1]), _ptr(cmin[d - 1])))

    def flux_ct(self, F, emf):
        self._bind()
        _check(lib().kh_flux_ct(self.h, _ptr(F), _ptr(emf)))

    def flux_divergence(self, F, dUdt):
        self._bind()
        _check(lib().kh_flux_divergence(self.h, _ptr(F), _ptr(dUdt)))

    def add_geo_source(self, P, dUdt):
        self._bind()
        _check(lib().kh_add_geo_source(self.h, _ptr(P), _ptr(dUdt)))

    def weighted_sum(self, Usub, Ufull, dUdt, gam0, gam1, beta_dt, Uout):
        self._bind()
        _check(lib().kh_weighted_sum(self.h, _ptr(Usub), _ptr(Ufull), _ptr(dUdt), gam0, gam1, beta_dt, _ptr(Uout)))

    def utop(self, U, P, pflag, fflag):
        self._bind()
        _check(lib().kh_utop(self.h, _ptr(U), _ptr(P), _ptr(pflag), _ptr(fflag)))

    def apply_floors(self, U, P, pflag, fflag):
        self._bind()
        _check(lib().kh_apply_floors(self.h, _ptr(U), _ptr(P), _ptr(pflag), _ptr(fflag)))

    def fix_utop(self, U, P, pflag):
        self._bind()
        _check(lib().kh_fix_utop(self.h, _ptr(U), _ptr(P), _ptr(pflag)))

    def ptou(self, P, U):
        self._bind()
        _check(lib().kh_ptou(self.h, _ptr(P), _ptr(U)))

    def sync_bounds(self, A, nvar):
        self._bind()
        _check(lib().kh_sync_bounds(self.h, _ptr(A), nvar))

    def apply_bounds(self, U, P):
        self._bind()
        _check(lib().kh_apply_bounds(self.h, _ptr(U), _ptr(P)))

    def estimate_timestep(self, cmax, cmin, dt_last):
        self._bind()
        out = ctypes.c_double()
        _check(lib().kh_estimate_timestep(self.h, _ptr(cmax), _ptr(cmin), dt_last, ctypes.byref(out)))
        return out.value

    def estimate_timestep_noalloc(self, cmax, cmin, dt_last, out):
        """kh_estimate_timestep_noalloc: the clipped timestep into the device tensor `out`
        (one float64), asynchronously, with no allocation."""
        self._bind()
        _check(lib().kh_estimate_timestep_noalloc(self.h, _ptr(cmax), _ptr(cmin), dt_last, _ptr(out)))

    def max_divb(self, U):
        self._bind()
        out = ctypes.c_double()
        _check(lib().kh_max_divb(self.h, _ptr(U), ctypes.byref(out)))
        return out.value

    # ---- face-centred CT (B_CT): face / edge tensors [3, n3+1, n2+1, n1+1] ----
    @property
    def face_shape(self):
        n3, n2, n1 = self.shape
        return (3, n3 + 1, n2 + 1, n1 + 1)

    def ct_calculate_emf(self, P, F, emf):
        self._bind()
        _check(lib().kh_ct_calculate_emf(self.h, _ptr(P), _ptr(F), _ptr(emf)))

    def ct_calculate_emf_v(self, P, F, emf):
        self._bind()
        Pv, Fv = vars_of(P), _vars3(F)
        _check(lib().kh_ct_calculate_emf_v(self.h, ctypes.byref(Pv), Fv, _ptr(emf)))

    def ct_add_source(self, emf, dfB):
        self._bind()
        _check(lib().kh_ct_add_source(self.h, _ptr(emf), _ptr(dfB)))

    def ct_weighted_sum_face(self, x, y, w1, w2, z):
        self._bind()
        _check(lib().kh_ct_weighted_sum_face(self.h, _ptr(x), _ptr(y), w1, w2, _ptr(z)))

    def ct_sync_faces(self, fB):
        self._bind()
        _check(lib().kh_ct_sync_faces(self.h, _ptr(fB)))

    def ct_utop(self, fB, U, P):
        self._bind()
        _check(lib().kh_ct_utop(self.h, _ptr(fB), _ptr(U), _ptr(P)))

    def ct_utop_v(self, fB, U, P):
        self._bind()
        Uv, Pv = vars_of(U), vars_of(P)
        _check(lib().kh_ct_utop_v(self.h, _ptr(fB), ctypes.byref(Uv), ctypes.byref(Pv)))

    def ct_max_divb(self, fB):
        self._bind()
        out = ctypes.c_double()
        _check(lib().kh_ct_max_divb(self.h, _ptr(fB), ctypes.byref(out)))
        return out.value

    # ---- pack view: each argument is 8 per-variable tensors (fluxes: 3 x 8) ----
    def get_flux_v(self, d, P, F_d, cmax_d, cmin_d):
        self._bind()
        _check(lib().kh_get_flux_v(self.h, d, ctypes.byref(vars_of(P)), ctypes.byref(vars_of(F_d)),
                                   _ptr(cmax_d), _ptr(cmin_d)))

    def flux_ct_v(self, F, emf):
        self._bind()
        _check(lib().kh_flux_ct_v(self.h, _vars3(F), _ptr(emf)))

    def flux_divergence_v(self, F, dUdt):
        self._bind()
        _check(lib().kh_flux_divergence_v(self.h, _vars3(F), ctypes.byref(vars_of(dUdt))))

    def add_geo_source_v(self, P, dUdt):
        self._bind()
        _check(lib().kh_add_geo_source_v(self.h, ctypes.byref(vars_of(P)), ctypes.byref(vars_of(dUdt))))

    def weighted_sum_v(self, Usub, Ufull, dUdt, gam0, gam1, beta_dt, Uout):
        self._bind()
        _check(lib().kh_weighted_sum_v(self.h, ctypes.byref(vars_of(Usub)), ctypes.byref(vars_of(Ufull)),
                                       ctypes.byref(vars_of(dUdt)), gam0, gam1, beta_dt, ctypes.byref(vars_of(Uout))))

    def utop_v(self, U, P, pflag, fflag):
        self._bind()
        _check(lib().kh_utop_v(self.h, ctypes.byref(vars_of(U)), ctypes.byref(vars_of(P)), _ptr(pflag), _ptr(fflag)))

    def apply_floors_v(self, U, P, pflag, fflag):
        self._bind()
        _check(lib().kh_apply_floors_v(self.h, ctypes.byref(vars_of(U)), ctypes.byref(vars_of(P)), _ptr(pflag),
                                       _ptr(fflag)))

    def fix_utop_v(self, U, P, pflag):
        self._bind()
        _check(lib().kh_fix_utop_v(self.h, ctypes.byref(vars_of(U)), ctypes.byref(vars_of(P)), _ptr(pflag)))

    def ptou_v(self, P, U):
        self._bind()
        _check(lib().kh_ptou_v(self.h, ctypes.byref(vars_of(P)), ctypes.byref(vars_of(U))))

    def apply_bounds_v(self, U, P):
        self._bind()
        _check(lib().kh_apply_bounds_v(self.h, ctypes.byref(vars_of(U)), ctypes.byref(vars_of(P))))

    def sync_bounds_v(self, A):
        """A: 1..8 per-variable tensors."""
        self._bind()
        A = list(A)
        v = KhVars()
        for n, a in enumerate(A):
            v.var[n] = a.data_ptr()
        _check(lib().kh_sync_bounds_v(self.h, ctypes.byref(v), len(A)))

    def max_divb_v(self, U):
        self._bind()
        out = ctypes.c_double()
        _check(lib().kh_max_divb_v(self.h, ctypes.byref(vars_of(U)), ctypes.byref(out)))
        return out.value

    def flags_to_real(self, flag, real_flag):
        self._bind()
        _check(lib().kh_flags_to_real(self.h, _ptr(flag), _ptr(real_flag), flag.numel()))

    def flags_from_real(self, real_flag, flag):
        self._bind()
        _check(lib().kh_flags_from_real(self.h, _ptr(real_flag), _ptr(flag), flag.numel()))

    # ---- first-order flux correction at the task level ----
    def determine_floors_v(self, P, fflag):
        self._bind()
        _check(lib().kh_determine_floors_v(self.h, ctypes.byref(vars_of(P)), _ptr(fflag)))

    def mark_fofc(self, fflag, fofcflag):
        self._bind()
        _check(lib().kh_mark_fofc(self.h, _ptr(fflag), _ptr(fofcflag)))

    def fofc_v(self, P, fofcflag, F, cmax, cmin):
        self._bind()
        _check(lib().kh_fofc_v(self.h, ctypes.byref(vars_of(P)), _ptr(fofcflag), _vars3(F), _ptr(cmax), _ptr(cmin)))


def vars_from_pack(pack, varmap):
    """kh_vars_from_pack: `pack` is a list of per-variable device tensors (a pack's
    variables in pack order), `varmap` a dict with RHO, UU, U1, B1 pack indices."""
    arr = (ctypes.c_void_p * len(pack))(*[t.data_ptr() for t in pack])
    m = KhVarMap(varmap["RHO"], varmap["UU"], varmap["U1"], varmap["B1"])
    out = KhVars()
    _check(lib().kh_vars_from_pack(arr, len(pack), ctypes.byref(m), ctypes.byref(out)))
    return out


class Driver:
    """KHARMA's default stage driver for one periodic block, run on the GPU."""

    def __init__(self, cfg):
        self.cfg = cfg
        self._p = make_params(cfg)
        h = ctypes.c_void_p()
        _check(lib().kh_driver_create(ctypes.byref(self._p), ctypes.byref(h)))
        self.h = h
        b = ctypes.c_void_p()
        _check(lib().kh_driver_block(self.h, ctypes.byref(b)))
        n1, n2, n3, nc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_long()
        _check(lib().kh_block_shape(b, ctypes.byref(n1), ctypes.byref(n2), ctypes.byref(n3), ctypes.byref(nc)))
        self.shape = (n3.value, n2.value, n1.value)
        self.nc = nc.value

    def __del__(self):
        try:
            lib().kh_driver_destroy(self.h)
        except Exception:
            pass

    def _bind(self):
        _check(lib().kh_driver_set_stream(self.h, _stream()))

    def init_prims(self, P):
        torch = _torch()
        P = torch.as_tensor(P, dtype=torch.float64).to("cuda").contiguous()
        assert tuple(P.shape) == (NV,) + self.shape, (P.shape, self.shape)
        self._bind()
        _check(lib().kh_driver_init_prims(self.h, _ptr(P)))
        self._P_init = P  # keep alive until the copy is done

    def init_faces(self, P, fB):
        """kh_driver_init_faces: primitives P (B ignored) and face fields fB
        [3, n3+1, n2+1, n1+1] of a face_ct run."""
        torch = _torch()
        P = torch.as_tensor(P, dtype=torch.float64).to("cuda").contiguous()
        fB = torch.as_tensor(fB, dtype=torch.float64).to("cuda").contiguous()
        n3, n2, n1 = self.shape
        assert tuple(P.shape) == (NV,) + self.shape and tuple(fB.shape) == (3, n3 + 1, n2 + 1, n1 + 1)
        self._bind()
        _check(lib().kh_driver_init_faces(self.h, _ptr(P), _ptr(fB)))
        self._P_init, self._fB_init = P, fB

    def faces(self):
        """The base container's face fields [3, n3+1, n2+1, n1+1] (face_ct runs)."""
        torch = _torch()
        n3, n2, n1 = self.shape
        fB = torch.empty((3, n3 + 1, n2 + 1, n1 + 1), dtype=torch.float64, device="cuda")
        self._bind()
        _check(lib().kh_driver_get_faces(self.h, _ptr(fB)))
        return fB

    def step(self, n=1):
        self._bind()
        _check(lib().kh_driver_step(self.h, n))

    def set_clock(self, time, dt, ncycle=0):
        self._bind()
        _check(lib().kh_driver_set_clock(self.h, time, dt, ncycle))

    def state(self):
        torch = _torch()
        P = torch.empty((NV,) + self.shape, dtype=torch.float64, device="cuda")
        U = torch.empty_like(P)
        self._bind()
        _check(lib().kh_driver_get_state(self.h, _ptr(P), _ptr(U)))
        return P, U

    def save_restart(self, path):
        """kh_driver_save_restart: the state (P, U) and clock to a restart file."""
        self._bind()
        _check(lib().kh_driver_save_restart(self.h, os.fsencode(path)))

    def load_restart(self, path):
        """kh_driver_load_restart: continue from a restart file of the same mesh."""
        self._bind()
        _check(lib().kh_driver_load_restart(self.h, os.fsencode(path)))

    def flags(self):
        """(pflag, fflag) of the last stage as int32 device tensors [n3, n2, n1]."""
        torch = _torch()
        pf = torch.empty(self.shape, dtype=torch.int32, device="cuda")
        ff = torch.empty_like(pf)
        self._bind()
        _check(lib().kh_driver_get_flags(self.h, _ptr(pf), _ptr(ff)))
        return pf, ff

    def scalars(self):
        t, dt, divb = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        nc, nf = ctypes.c_int(), ctypes.c_int()
        self._bind()
        _check(lib().kh_driver_get_scalars(self.h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(divb),
                                           ctypes.byref(nc), ctypes.byref(nf)))
        return dict(time=t.value, dt=dt.value, divb=divb.value, ncycle=nc.value, fails=nf.value)

    KERNEL_KINDS = ("flux_x1", "flux_x2", "flux_x3", "bounds", "update_s1", "update_s2", "fixup", "divb", "finalize")

    def set_kernel_timing(self, enable):
        _check(lib().kh_driver_set_kernel_timing(self.h, int(bool(enable))))

    def kernel_times(self):
        """{kind: (total_ms, launches)} since the last call (synchronises)."""
        n = len(self.KERNEL_KINDS)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int * n)()
        _check(lib().kh_driver_kernel_times(self.h, ms, cnt, n))
        return {k: (ms[q], cnt[q]) for q, k in enumerate(self.KERNEL_KINDS)}

    def check(self):
        """Raise KharmaError if a step's corner divB exceeded kill_on_divb_over
        (kh_driver_check; synchronises)."""
        self._bind()
        _check(lib().kh_driver_check(self.h))

    def new_stage_record(self):
        """A device buffer (uint8 CUDA tensor) for one step's record."""
        n = ctypes.c_size_t()
        _check(lib().kh_driver_stage_record_bytes(self.h, ctypes.byref(n)))
        return _torch().empty(n.value, dtype=_torch().uint8, device="cuda")

    def set_stage_record(self, rec=None):
        """Record every following step into rec (kh_driver_set_stage_record);
        None: recording off.  rec must outlive the recording."""
        self._bind()
        _check(lib().kh_driver_set_stage_record(self.h, _ptr(rec)))

    def set_face_flux_objective(self, dir, var, face, weight=1.0):
        """Accumulate J += weight sum_stages w_s dt sum F_dir[var](face) dA during steps
        (kh_driver_set_face_flux_objective; dir = 0 switches it off)."""
        _check(lib().kh_driver_set_face_flux_objective(self.h, int(dir), int(var), int(face), float(weight)))

    def objective(self, reset=False):
        v = ctypes.c_double()
        self._bind()
        _check(lib().kh_driver_objective(self.h, ctypes.byref(v), int(bool(reset))))
        return v.value

    def dt_history(self, n):
        arr = (ctypes.c_double * n)()
        self._bind()
        _check(lib().kh_driver_get_dt_history(self.h, arr, n))
        return list(arr)


class Mesh:
    """A uniform mesh of nb1 x nb2 x nb3 meshblocks on the GPU (kh_mesh_*): the
    default stage driver over all blocks.  cfg describes the whole mesh; states
    read and write in the layout of one block of the whole mesh."""

    def __init__(self, cfg, nblocks):
        self.cfg = cfg
        self._p = make_params(cfg)
        nb = tuple(nblocks) + (1,) * (3 - len(nblocks))
        h = ctypes.c_void_p()
        _check(lib().kh_mesh_create(ctypes.byref(self._p), nb[0], nb[1], nb[2], ctypes.byref(h)))
        self.h = h
        self.nblocks = nb
        nbk, n1, n2, n3, nc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_long()
        _check(lib().kh_mesh_shape(self.h, ctypes.byref(nbk), ctypes.byref(n1), ctypes.byref(n2), ctypes.byref(n3),
                                   ctypes.byref(nc)))
        self.block_shape = (n3.value, n2.value, n1.value)
        ndim, _ = problems.block_shape(cfg)
        ng = self._p.ng
        self.shape = (self._p.nx3 + 2 * ng if ndim > 2 else 1, self._p.nx2 + 2 * ng if ndim > 1 else 1,
                      self._p.nx1 + 2 * ng)

    def __del__(self):
        try:
            lib().kh_mesh_destroy(self.h)
        except Exception:
            pass

    def _bind(self):
        _check(lib().kh_mesh_set_stream(self.h, _stream()))

    def init_prims(self, P):
        torch = _torch()
        P = torch.as_tensor(P, dtype=torch.float64).to("cuda").contiguous()
        assert tuple(P.shape) == (NV,) + self.shape, (P.shape, self.shape)
        self._bind()
        _check(lib().kh_mesh_init_prims(self.h, _ptr(P)))
        self._P_init = P

    def step(self, n=1):
        self._bind()
        _check(lib().kh_mesh_step(self.h, n))

    def set_graph(self, enable=True):
        """kh_mesh_set_graph: replay each step as a captured HIP graph."""
        _check(lib().kh_mesh_set_graph(self.h, int(bool(enable))))

    def set_packed(self, enable=True):
        """kh_mesh_set_packed: each kernel once over all blocks (default) or once per block."""
        _check(lib().kh_mesh_set_packed(self.h, int(bool(enable))))

    def set_clock(self, time, dt, ncycle=0):
        self._bind()
        _check(lib().kh_mesh_set_clock(self.h, time, dt, ncycle))

    def state(self):
        torch = _torch()
        P = torch.empty((NV,) + self.shape, dtype=torch.float64, device="cuda")
        U = torch.empty_like(P)
        self._bind()
        _check(lib().kh_mesh_get_state(self.h, _ptr(P), _ptr(U)))
        return P, U

    def scalars(self):
        t, dt, divb = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        nc, nf = ctypes.c_int(), ctypes.c_int()
        self._bind()
        _check(lib().kh_mesh_get_scalars(self.h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(divb),
                                         ctypes.byref(nc), ctypes.byref(nf)))
        return dict(time=t.value, dt=dt.value, divb=divb.value, ncycle=nc.value, fails=nf.value)

    def dt_history(self, n):
        arr = (ctypes.c_double * n)()
        self._bind()
        _check(lib().kh_mesh_get_dt_history(self.h, arr, n))
        return list(arr)

    def check(self):
        self._bind()
        _check(lib().kh_mesh_check(self.h))

    def save_restart(self, path):
        self._bind()
        _check(lib().kh_mesh_save_restart(self.h, os.fsencode(path)))

    def load_restart(self, path):
        self._bind()
        _check(lib().kh_mesh_load_restart(self.h, os.fsencode(path)))


RESTART_MAGIC = b"KHAMDRST"


def read_restart(path):
    """A restart file (kh_driver_save_restart / kh_mesh_save_restart) on the host:
    dict(time, dt, ncycle, params (KhParams), P, U (float64 [8, n3, n2, n1]), and
    fB (float64 [3, n3+1, n2+1, n1+1]) for a face_ct run)."""
    import numpy as np
    raw = open(path, "rb").read()
    if raw[:8] != RESTART_MAGIC:
        raise KharmaError(f"{path}: not a restart file of this library")
    version, hbytes = np.frombuffer(raw, dtype=np.uint32, count=2, offset=8)
    n1, n2, n3, ncycle = np.frombuffer(raw, dtype=np.int32, count=4, offset=16)
    time, dt = np.frombuffer(raw, dtype=np.float64, count=2, offset=32)
    params = KhParams.from_buffer_copy(raw[48:48 + ctypes.sizeof(KhParams)])
    cells = 8 * int(n1) * int(n2) * int(n3)
    fshape = (3, int(n3) + 1, int(n2) + 1, int(n1) + 1)
    faces = int(np.prod(fshape)) if params.b_solver == BSOLVER["face_ct"] else 0
    data = np.frombuffer(raw, dtype=np.float64, count=2 * cells + faces, offset=int(hbytes))
    shape = (8, int(n3), int(n2), int(n1))
    out = dict(time=float(time), dt=float(dt), ncycle=int(ncycle), version=int(version), params=params,
               P=data[:cells].reshape(shape).copy(), U=data[cells:2 * cells].reshape(shape).copy())
    if faces:
        out["fB"] = data[2 * cells:].reshape(fshape).copy()
    return out


class AdjointStep:
    """Reverse-mode adjoint of one driver step (kh_adjoint_step): maps dL/dP_{n+1}
    to dL/dP_n for the step taken from P_n with timestep dt, recomputing the
    forward step from P_n.  FP64, WENO5 + LLF; Minkowski or Kerr-Schild blocks
    with the forward step's boundaries (floors on the Kerr-Schild path)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self._p = make_params(dict(cfg, precision="fp64"))
        h = ctypes.c_void_p()
        _check(lib().kh_adjoint_create(ctypes.byref(self._p), ctypes.byref(h)))
        self.h = h

    def __del__(self):
        try:
            lib().kh_adjoint_destroy(self.h)
        except Exception:
            pass

    def step(self, P, dt, lamP, P_next=None):
        """In place: lamP (float64 CUDA tensor [8, n3, n2, n1]) <- dL/dP_n for the
        step from P taken with dt.  P_next, the step's result, spares the
        recomputation of its second stage (its flags must have been checked)."""
        _check(lib().kh_adjoint_set_stream(self.h, _stream()))
        _check(lib().kh_adjoint_step(self.h, _ptr(P), _ptr(P_next), float(dt), _ptr(lamP)))

    def step_recorded(self, P, rec, dt, lamP, P_next):
        """step() from the record the driver kept of the step from P to P_next
        (Driver.set_stage_record), without recomputing its stages."""
        _check(lib().kh_adjoint_set_stream(self.h, _stream()))
        _check(lib().kh_adjoint_step_recorded(self.h, _ptr(P), _ptr(rec), _ptr(P_next), float(dt), _ptr(lamP)))

    def set_deferred_checks(self, on):
        """Deferred checks (kh_adjoint_set_deferred_checks): steps stop
        synchronising; check() reports what they found."""
        _check(lib().kh_adjoint_set_stream(self.h, _stream()))
        _check(lib().kh_adjoint_set_deferred_checks(self.h, int(bool(on))))

    def check(self):
        """Raise KharmaError for zones the deferred steps did not cover (kh_adjoint_check)."""
        _check(lib().kh_adjoint_set_stream(self.h, _stream()))
        _check(lib().kh_adjoint_check(self.h))

    def _call(self, fn, *args):
        _check(lib().kh_adjoint_set_stream(self.h, _stream()))
        _check(getattr(lib(), fn)(self.h, *[_ptr(x) for x in args]))

    def flux(self, P, lamF, lamP):
        """lamP = sum_d (dF_d/dP)^T lamF[d] (entire block)."""
        self._call("kh_adjoint_flux", P, lamF, lamP)

    def flux_div(self, lam_dUdt, lamF):
        """lamF = (d dUdt / dF)^T lam_dUdt, flux-CT included."""
        self._call("kh_adjoint_flux_div", lam_dUdt, lamF)

    def utop(self, P, lamP, lamU):
        """lamU = (dU/dP)^-T lamP at each interior zone."""
        self._call("kh_adjoint_utop", P, lamP, lamU)

    def ptou(self, P, lamU, lamP):
        """lamP += (dU/dP)^T lamU at each interior zone."""
        self._call("kh_adjoint_ptou", P, lamU, lamP)

    def bounds(self, P, lamP):
        """In place: ghost-zone adjoints of lamP moved onto the zones the ghost fill reads."""
        self._call("kh_adjoint_bounds", P, lamP)

    def fix_flux(self, F_fixed, lamF):
        """In place: lamF <- (dF_fixed/dF)^T lamF for the physical-boundary flux fixes."""
        self._call("kh_adjoint_fix_flux", F_fixed, lamF)

    def geo_source(self, P, lam_dUdt, lamP):
        """lamP += (dS/dP)^T lam_dUdt at each interior zone (geometric source)."""
        self._call("kh_adjoint_geo_source", P, lam_dUdt, lamP)

    def floors(self, P_in, lamP_out, lamP_in):
        """lamP_in = (dP_out/dP_in)^T lamP_out for the floors of each interior zone."""
        self._call("kh_adjoint_floors", P_in, lamP_out, lamP_in)

    def fixup(self, P_fixed, pflag, lamP):
        """In place: lamP <- (dP_fixed/dP)^T lamP for the inversion fixup of the zones pflag > 0."""
        self._call("kh_adjoint_fixup", P_fixed, pflag, lamP)

    def set_face_flux_objective(self, dir, var, face, weight=1.0):
        """Seed every following step with the derivative of the driver's face-flux objective."""
        _check(lib().kh_adjoint_set_face_flux_objective(self.h, int(dir), int(var), int(face), float(weight)))

    def last_times(self):
        f, a = ctypes.c_double(), ctypes.c_double()
        _check(lib().kh_adjoint_last_times(self.h, ctypes.byref(f), ctypes.byref(a)))
        return f.value, a.value
