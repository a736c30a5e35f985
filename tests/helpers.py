"""Shared comparison helpers for the parity tests."""
import numpy as np

from kharma_amd import problems as pr


def rel_l1(a, b):
    """Per-variable relative L1 distance sum|a-b| / sum|b| over the leading axis.

    Variables whose reference is identically zero report the absolute mean instead."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    out = []
    for v in range(a.shape[0]):
        d = np.abs(a[v] - b[v]).sum()
        n = np.abs(b[v]).sum()
        out.append(d / n if n > 0 else d / max(b[v].size, 1))
    return np.array(out)


def interior_of(A, cfg):
    return A[(slice(None),) + pr.interior(cfg)]


def face_slices(cfg, d):
    """Index box (k, j, i slices) of the faces normal to d written by the HIP flux kernel."""
    ng = cfg.get("ng", 4)
    ndim, _ = pr.block_shape(cfg)
    nx1, nx2, nx3 = cfg["nx"]
    is_, ie = ng, ng + nx1 - 1
    js, je = (ng, ng + nx2 - 1) if ndim > 1 else (0, 0)
    ks, ke = (ng, ng + nx3 - 1) if ndim > 2 else (0, 0)
    i0, i1 = (is_, ie + 1) if d == 1 else (is_ - 1, ie + 1)
    j0, j1 = (js, je)
    k0, k1 = (ks, ke)
    if ndim > 1:
        j0, j1 = (js, je + 1) if d == 2 else (js - 1, je + 1)
    if ndim > 2:
        k0, k1 = (ks, ke + 1) if d == 3 else (ks - 1, ke + 1)
    return (slice(k0, k1 + 1), slice(j0, j1 + 1), slice(i0, i1 + 1))


SYM = [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]


_GEOMETRY = {}


def _library_geometry(ob, cfg):
    """The library's host geometry in the oracle's layout, built once per
    configuration (each kh_geometry_host call builds the whole cache)."""
    import kharma_amd as kh
    key = repr(sorted((k, repr(v)) for k, v in cfg.items()))
    if key not in _GEOMETRY:
        _GEOMETRY.clear()
        gcov = ob.geometry(0)
        gcon = ob.geometry(1)
        gdet = ob.geometry(2)
        for ol, loc in ((0, "face1"), (1, "face2"), (2, "center"), (3, "center"), (4, "corner")):
            for c, (a, b) in enumerate(SYM):
                v = kh.geometry_host(cfg, loc, c)
                gcov[ol, ..., a, b] = v
                gcov[ol, ..., b, a] = v
                w = kh.geometry_host(cfg, loc, 10 + c)
                gcon[ol, ..., a, b] = w
                gcon[ol, ..., b, a] = w
            gdet[ol] = kh.geometry_host(cfg, loc, 20)
        conn = np.stack([kh.geometry_host(cfg, "center", 22 + n) for n in range(64)], axis=-1)
        _GEOMETRY[key] = (gcov, gcon, gdet, conn.reshape(conn.shape[:2] + (4, 4, 4)),
                          kh.geometry_host(cfg, "center", 88))
    return _GEOMETRY[key]


def install_library_geometry(ob, cfg):
    """Give the oracle block the library's geometry cache, so that a step
    comparison sees identical metric inputs (the caches themselves are compared
    in test_geometry_torus.py).  Oracle loci: face1, face2, face3, centre, corner."""
    for which, a in enumerate(_library_geometry(ob, cfg)):
        ob.set_geometry(which, a)
