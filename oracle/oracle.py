"""ctypes wrapper over oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The oracle is the CPU FP64 restatement of KHARMA's GRMHD step that the HIP
path is checked against.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
# Same source compiled with FMA contraction: measures a workload's roundoff sensitivity.
LIB_FMA_PATH = os.path.join(HERE, "liboracle_fma.so")
# The same source on forward-mode dual numbers: exact directional derivatives (tangent-linear).
LIB_TANGENT_PATH = os.path.join(HERE, "liboracle_tangent.so")

NV = 8


class OrParams(ctypes.Structure):
    _fields_ = [
        ("nx1", ctypes.c_int), ("nx2", ctypes.c_int), ("nx3", ctypes.c_int), ("ng", ctypes.c_int),
        ("x1min", ctypes.c_double), ("x1max", ctypes.c_double),
        ("x2min", ctypes.c_double), ("x2max", ctypes.c_double),
        ("x3min", ctypes.c_double), ("x3max", ctypes.c_double),
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
        ("integrator", ctypes.c_int), ("spherical", ctypes.c_int),
        ("a", ctypes.c_double), ("transform", ctypes.c_int),
        ("hslope", ctypes.c_double), ("mks_smooth", ctypes.c_double), ("poly_xt", ctypes.c_double),
        ("poly_alpha", ctypes.c_double), ("fmks_zero_point", ctypes.c_double),
        ("bc_inner_x1", ctypes.c_int), ("bc_outer_x1", ctypes.c_int),
        ("bc_inner_x2", ctypes.c_int), ("bc_outer_x2", ctypes.c_int),
        ("check_inflow_inner_x1", ctypes.c_int), ("check_inflow_outer_x1", ctypes.c_int),
        ("zero_polar_flux", ctypes.c_int), ("fix_polar_flux", ctypes.c_int), ("fix_corner_inner", ctypes.c_int),
        ("periodic1", ctypes.c_int), ("periodic2", ctypes.c_int), ("periodic3", ctypes.c_int),
        ("driver", ctypes.c_int), ("inverter", ctypes.c_int),
        ("fofc", ctypes.c_int), ("fofc_polar_cells", ctypes.c_int), ("fofc_eh_buffer", ctypes.c_double),
        ("face_ct", ctypes.c_int), ("ct_scheme", ctypes.c_int),
    ]


_libs = {}


def build(force=False):
    """Compile liboracle.so (and liboracle_fma.so) from oracle/kharma_oracle.cpp with the committed Makefile."""
    src = os.path.getmtime(os.path.join(HERE, "kharma_oracle.cpp"))
    stale = any(not os.path.exists(p) or os.path.getmtime(p) < src for p in (LIB_PATH, LIB_FMA_PATH, LIB_TANGENT_PATH))
    if force or stale:
        subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []), check=True)


def lib(variant=""):
    if variant not in _libs:
        path = {"fma": LIB_FMA_PATH, "tangent": LIB_TANGENT_PATH}.get(variant, LIB_PATH)
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        vp, d, i, dp = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
        L.or_create.restype = vp
        L.or_create.argtypes = [ctypes.POINTER(OrParams)]
        L.or_destroy.argtypes = [vp]
        L.or_ncells.restype = ctypes.c_size_t
        L.or_ncells.argtypes = [vp]
        L.or_init_state.argtypes = [vp, dp]
        L.or_set_state.argtypes = [vp, dp, dp]
        L.or_get_state.argtypes = [vp, dp, dp]
        L.or_set_dt.argtypes = [vp, d]
        L.or_get_dt.restype = d
        L.or_get_dt.argtypes = [vp]
        L.or_get_time.restype = d
        L.or_get_time.argtypes = [vp]
        L.or_step.restype = i
        L.or_step.argtypes = [vp, i, dp]
        L.or_get_flux.argtypes = [vp, i]
        for name in ("or_flux_ct", "or_flux_divergence", "or_utop", "or_apply_floors", "or_fix_utop", "or_ptou",
                     "or_add_geo_source", "or_fofc_stage1"):
            getattr(L, name).argtypes = [vp]
        L.or_estimate_timestep.restype = d
        L.or_estimate_timestep.argtypes = [vp]
        L.or_max_divb.restype = d
        L.or_max_divb.argtypes = [vp]
        L.or_get_array.argtypes = [vp, i, dp]
        L.or_set_array.argtypes = [vp, i, dp]
        L.or_weno5.argtypes = [d, d, d, d, d, dp, dp]
        L.or_reconstruct.argtypes = [i, d, d, d, d, d, dp, dp]
        L.or_num_threads.restype = i
        L.or_get_geometry.restype = ctypes.c_long
        L.or_get_geometry.argtypes = [vp, i, dp]
        L.or_apply_bcs.argtypes = [vp]
        L.or_set_geometry.argtypes = [vp, i, dp]
        L.or_fix_flux.argtypes = [vp]
        L.or_plant.argtypes = [i]
        L.or_weighted_sum.argtypes = [vp, dp, dp, d, d, d, dp]
        L.or_sync_vars.argtypes = [vp, dp, i]
        L.or_init_state_faces.argtypes = [vp, dp, dp]
        L.or_nfaces.restype = ctypes.c_size_t
        L.or_nfaces.argtypes = [vp]
        for name in ("or_ct_calculate_emf", "or_ct_add_source", "or_ct_block_utop", "or_ct_sync_faces"):
            getattr(L, name).argtypes = [vp]
        L.or_ct_max_divb.restype = d
        L.or_ct_max_divb.argtypes = [vp]
        L.or_ct_weighted_sum_face.argtypes = [vp, dp, dp, d, d, dp]
        if variant == "tangent":
            L.or_set_tangent.argtypes = [vp, dp, dp]
            L.or_get_tangent.argtypes = [vp, dp, dp]
        _libs[variant] = L
    return _libs[variant]


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


RECON = {"donor_cell_c": 0, "linear_mc": 1, "weno5": 2, "ppm": 3}
FRAME = {"fluid": 0, "normal": 1, "drift": 2}
TRANSFORM = {"null": 0, "mks": 1, "fmks": 2}
BC = {"periodic": 0, "outflow": 1, "reflecting": 2}
DRIVER = {"kharma": 0, "simple": 1}
CT_SCHEME = {"bs99": 0, "gs05_0": 1, "gs05_c": 2, "sg07": 3}


def make_params(cfg):
    """Build OrParams from a run configuration dict (kharma_amd.problems layout)."""
    p = OrParams()
    p.nx1, p.nx2, p.nx3 = cfg["nx"]
    p.ng = cfg.get("ng", 4)
    (p.x1min, p.x1max), (p.x2min, p.x2max), (p.x3min, p.x3max) = cfg["bounds"]
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
    p.iter_max = cfg.get("iter_max", 8)
    p.err_tol = cfg.get("err_tol", 1e-8)
    p.fix_average = int(cfg.get("fix_average", True))
    p.fix_atmo = int(cfg.get("fix_atmo", True))
    p.integrator = {"rk2": 0, "vl2": 1}[cfg.get("integrator", "rk2")]
    p.spherical = int(cfg.get("coords", "cartesian_minkowski") == "ks")
    p.a = cfg.get("a", 0.0)
    p.transform = TRANSFORM[cfg.get("transform", "null")]
    p.hslope = cfg.get("hslope", 0.3)
    p.mks_smooth = cfg.get("mks_smooth", 0.5)
    p.poly_xt = cfg.get("poly_xt", 0.82)
    p.poly_alpha = cfg.get("poly_alpha", 14.0)
    p.fmks_zero_point = cfg.get("fmks_zero_point", 0.0)
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
    p.driver = DRIVER[cfg.get("driver", "kharma")]
    p.inverter = {"onedw": 0, "kastaun": 1}[cfg.get("inverter", "onedw")]
    p.fofc = int(cfg.get("fofc", False))
    p.fofc_polar_cells = cfg.get("fofc_polar_cells", 0)
    p.fofc_eh_buffer = cfg.get("fofc_eh_buffer", 0.1)
    p.face_ct = int(cfg.get("b_solver", "flux_ct") == "face_ct")
    p.ct_scheme = CT_SCHEME[cfg.get("ct_scheme", "gs05_c")]
    return p


class OracleBlock:
    """One meshblock advanced by the CPU restatement."""

    def __init__(self, cfg, variant=""):
        self.cfg = cfg
        self._p = make_params(cfg)
        self.L = lib(variant)
        self.h = self.L.or_create(ctypes.byref(self._p))
        ng = self._p.ng
        nx1, nx2, nx3 = cfg["nx"]
        self.ndim = 3 if nx3 > 1 else (2 if nx2 > 1 else 1)
        n1 = nx1 + 2 * ng
        n2 = nx2 + 2 * ng if self.ndim > 1 else 1
        n3 = nx3 + 2 * ng if self.ndim > 2 else 1
        self.shape = (n3, n2, n1)
        self.nc = n1 * n2 * n3
        # face / edge arrays of B_CT: [3][n3+1][n2+1][n1+1]
        self.fshape = (3, n3 + 1, n2 + 1, n1 + 1)

    def __del__(self):
        try:
            self.L.or_destroy(self.h)
        except Exception:
            pass

    def init_state(self, P):
        P = np.ascontiguousarray(P, dtype=np.float64)
        assert P.shape == (NV,) + self.shape
        self.L.or_init_state(self.h, _ptr(P))

    def set_tangent(self, dP, dU):
        """Tangent-linear variant: seed the directions of the base state P, U."""
        dP = np.ascontiguousarray(dP, dtype=np.float64)
        dU = np.ascontiguousarray(dU, dtype=np.float64)
        self.L.or_set_tangent(self.h, _ptr(dP), _ptr(dU))

    def tangent(self):
        """Tangent-linear variant: the tangents (dP, dU) of the base state."""
        dP = np.empty((NV,) + self.shape)
        dU = np.empty((NV,) + self.shape)
        self.L.or_get_tangent(self.h, _ptr(dP), _ptr(dU))
        return dP, dU

    def set_state(self, P, U):
        P = np.ascontiguousarray(P, dtype=np.float64)
        U = np.ascontiguousarray(U, dtype=np.float64)
        self.L.or_set_state(self.h, _ptr(P), _ptr(U))

    def state(self):
        P = np.empty((NV,) + self.shape)
        U = np.empty((NV,) + self.shape)
        self.L.or_get_state(self.h, _ptr(P), _ptr(U))
        return P, U

    def step(self, n=1):
        dts = np.zeros(n)
        self.L.or_step(self.h, n, _ptr(dts))
        return dts

    @property
    def dt(self):
        return self.L.or_get_dt(self.h)

    @dt.setter
    def dt(self, v):
        self.L.or_set_dt(self.h, v)

    @property
    def time(self):
        return self.L.or_get_time(self.h)

    def array(self, which):
        if which in (12, 13, 14):
            out = np.empty(self.fshape)
            self.L.or_get_array(self.h, which, _ptr(out))
            return out
        sizes = {0: (3, NV), 1: (3,), 2: (3,), 3: (NV,), 4: (), 5: (), 6: (3,), 7: (NV,), 8: (NV,),
                 9: (), 10: (NV,), 11: (), 15: (3,)}
        out = np.empty(sizes[which] + self.shape)
        self.L.or_get_array(self.h, which, _ptr(out))
        return out

    def set_array(self, which, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.L.or_set_array(self.h, which, _ptr(a))

    def get_flux(self, d):
        self.L.or_get_flux(self.h, d)

    def flux_ct(self):
        self.L.or_flux_ct(self.h)

    def fofc_stage1(self):
        """First-order flux correction of stage 1 on the current fluxes; marks,
        trial primitives and floor margins then read as array(9), (10), (11)."""
        self.L.or_fofc_stage1(self.h)

    def flux_divergence(self):
        self.L.or_flux_divergence(self.h)

    def add_geo_source(self):
        self.L.or_add_geo_source(self.h)

    def utop(self):
        self.L.or_utop(self.h)

    def apply_floors(self):
        self.L.or_apply_floors(self.h)

    def fix_utop(self):
        self.L.or_fix_utop(self.h)

    def ptou(self):
        self.L.or_ptou(self.h)

    def estimate_timestep(self):
        return self.L.or_estimate_timestep(self.h)

    def max_divb(self):
        return self.L.or_max_divb(self.h)

    def weighted_sum(self, Usub, Ufull, gam0, gam1, beta_dt):
        """gam0 Usub + gam1 Ufull + beta_dt dUdt (the block's dUdt) -> new array."""
        a = np.ascontiguousarray(Usub, dtype=np.float64)
        b = np.ascontiguousarray(Ufull, dtype=np.float64)
        out = np.empty_like(a)
        self.L.or_weighted_sum(self.h, _ptr(a), _ptr(b), gam0, gam1, beta_dt, _ptr(out))
        return out

    # ---- B_CT (face-centred constrained transport)
    def init_state_faces(self, P, fB):
        """Initial state with face fields fB [3][n3+1][n2+1][n1+1] (the B of P is ignored)."""
        P = np.ascontiguousarray(P, dtype=np.float64)
        fB = np.ascontiguousarray(fB, dtype=np.float64)
        assert P.shape == (NV,) + self.shape and fB.shape == self.fshape
        self.L.or_init_state_faces(self.h, _ptr(P), _ptr(fB))

    def faces(self):
        return self.array(12)

    def ct_calculate_emf(self):
        self.L.or_ct_calculate_emf(self.h)

    def ct_add_source(self):
        self.L.or_ct_add_source(self.h)

    def ct_block_utop(self):
        self.L.or_ct_block_utop(self.h)

    def ct_sync_faces(self):
        self.L.or_ct_sync_faces(self.h)

    def ct_max_divb(self):
        return self.L.or_ct_max_divb(self.h)

    def ct_weighted_sum_face(self, x, y, w1, w2):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        z = np.empty_like(x)
        self.L.or_ct_weighted_sum_face(self.h, _ptr(x), _ptr(y), w1, w2, _ptr(z))
        return z

    def sync_vars(self, A):
        """The periodic exchange of the block with itself on A [nvar, n3, n2, n1] -> new array."""
        out = np.ascontiguousarray(A, dtype=np.float64).copy()
        self.L.or_sync_vars(self.h, _ptr(out), out.shape[0])
        return out

    def apply_bcs(self):
        self.L.or_apply_bcs(self.h)

    def fix_flux(self):
        self.L.or_fix_flux(self.h)

    def set_geometry(self, which, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.L.or_get_geometry(self.h, which, None)
        self.L.or_set_geometry(self.h, which, _ptr(a))

    def geometry(self, which):
        """Geometry cache: 0 gcov, 1 gcon, 2 gdet, 3 gdet*Gamma, 4 centre radius (see or_get_geometry)."""
        n = self.L.or_get_geometry(self.h, which, None)
        out = np.empty(n)
        self.L.or_get_geometry(self.h, which, _ptr(out))
        gn2, gn1 = self.shape[1] + 1, self.shape[2] + 1
        nloc = 5
        shapes = {0: (nloc, gn2, gn1, 4, 4), 1: (nloc, gn2, gn1, 4, 4), 2: (nloc, gn2, gn1),
                  3: (gn2, gn1, 4, 4, 4), 4: (gn2, gn1)}
        return out.reshape(shapes[which])


def weno5(x):
    l, r = ctypes.c_double(), ctypes.c_double()
    lib().or_weno5(*[float(v) for v in x], ctypes.byref(l), ctypes.byref(r))
    return l.value, r.value


def reconstruct(name, x):
    l, r = ctypes.c_double(), ctypes.c_double()
    lib().or_reconstruct(RECON[name], *[float(v) for v in x], ctypes.byref(l), ctypes.byref(r))
    return l.value, r.value


def num_threads():
    return lib().or_num_threads()


class planted:
    """Context manager: plant the numbered known mistake in the restatement
    (or_plant), so a test can show that its pin fails on it."""

    def __init__(self, code, variant=""):
        self.code, self.variant = code, variant

    def __enter__(self):
        lib(self.variant).or_plant(self.code)
        return self

    def __exit__(self, *exc):
        lib(self.variant).or_plant(0)
        return False
