"""The pack view of the C ABI: the task entry points reading and writing the 8
variables through per-variable pointers (kh_vars), as KHARMA keeps prims.rho,
prims.u, prims.uvec, prims.B and their cons.* counterparts in separate fields
(grmhd/grmhd.cpp:170-185, b_flux_ct/b_flux_ct.cpp:116-118, types.hpp:126-218).

Every *_v entry point is run on variables scattered through one pool in a
permuted order with gaps between them, and must equal the contiguous entry point
on the same values bit for bit: both run the same kernels, so any difference is
an addressing mistake.  Also here: kh_weighted_sum in place (AddStateUpdate's
form, kharma_driver.cpp:366-370) against numpy, kh_sync_bounds against the
periodic images, the VarMap helper and the Real-valued flag conversion.
"""
import numpy as np
import pytest

from kharma_amd import problems as pr

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NV = 8


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    import kharma_amd
    kharma_amd.lib()


CASES = {
    "ot2d": lambda: pr.orszag_tang_config(n=64),
    "fast3d": lambda: pr.mhdmode_config(n=16, ndim=3, mode="fast", amp=0.05),
    "torus2d": lambda: pr.torus_config(64, 48, 1),
    "torus3d": lambda: pr.torus_config(32, 24, 8),
}


def _state(name, steps=3):
    """A state a few driver steps into the run (ghosts filled, U consistent)."""
    import kharma_amd as kh
    cfg = CASES[name]()
    gen = {"ot2d": pr.orszag_tang_prims, "fast3d": pr.mhdmode_prims}.get(name, pr.torus_prims)
    d = kh.Driver(cfg)
    d.init_prims(torch.from_numpy(gen(cfg)).cuda())
    d.step(steps)
    P, U = d.state()
    torch.cuda.synchronize()
    return cfg, P.clone(), U.clone()


class Scattered:
    """8 variables placed in one pool in a permuted order, `gap` elements apart."""

    def __init__(self, shape, dtype, seed, gap=96):
        nc = int(np.prod(shape))
        rng = np.random.default_rng(seed)
        self.perm = rng.permutation(NV)
        self.pool = torch.full((NV * (nc + gap) + gap,), float("nan"), dtype=dtype, device="cuda")
        self.vars = []
        for v in range(NV):
            off = gap + int(self.perm[v]) * (nc + gap)
            self.vars.append(self.pool[off:off + nc].view(shape))

    def load(self, A):
        for v in range(NV):
            self.vars[v].copy_(A[v])
        return self

    def stack(self):
        return torch.stack([x for x in self.vars])


def _scatter(A, seed):
    return Scattered(tuple(A.shape[1:]), A.dtype, seed).load(A)


def _fluxes(blk, seed):
    F, cmax, cmin = blk.new_fluxes()
    Fs = [Scattered(blk.shape, blk.flux_dtype, seed + d) for d in range(3)]
    for d in range(3):
        Fs[d].load(torch.zeros((NV,) + blk.shape, dtype=blk.flux_dtype, device="cuda"))
    return F, cmax, cmin, Fs


def _same(a, b):
    torch.cuda.synchronize()
    a, b = a.detach().cpu(), b.detach().cpu()
    return torch.equal(torch.nan_to_num(a, nan=7.0), torch.nan_to_num(b, nan=7.0))


@pytest.mark.parametrize("case", ["ot2d", "fast3d", "torus2d", "torus3d"])
def test_flux_ct_divergence_source_pack_equals_contiguous(case):
    import kharma_amd as kh
    cfg, P, U = _state(case)
    print(f"[pack flux/ct/div] {case}", flush=True)
    blk = kh.Block(cfg)
    nd = pr.block_shape(cfg)[0]
    Ps = _scatter(P, 1)
    F, cmax, cmin, Fs = _fluxes(blk, 10)
    cmax_v, cmin_v = torch.zeros_like(cmax), torch.zeros_like(cmin)
    for d in range(1, nd + 1):
        blk.get_flux(d, P, F, cmax, cmin)
        blk.get_flux_v(d, Ps.vars, Fs[d - 1].vars, cmax_v[d - 1], cmin_v[d - 1])
    for d in range(nd):
        assert _same(Fs[d].stack(), F[d]), d
    assert _same(cmax_v, cmax) and _same(cmin_v, cmin)
    emf = torch.zeros((3,) + blk.shape, dtype=blk.flux_dtype, device="cuda")
    emf_v = torch.zeros_like(emf)
    blk.flux_ct(F, emf)
    blk.flux_ct_v([f.vars for f in Fs], emf_v)
    assert _same(emf_v, emf)
    for d in range(nd):
        assert _same(Fs[d].stack(), F[d]), d
    dU = torch.zeros((NV,) + blk.shape, dtype=torch.float64, device="cuda")
    dUs = _scatter(dU, 2)
    blk.flux_divergence(F, dU)
    blk.flux_divergence_v([f.vars for f in Fs], dUs.vars)
    blk.add_geo_source(P, dU)
    blk.add_geo_source_v(Ps.vars, dUs.vars)
    assert _same(dUs.stack(), dU)


@pytest.mark.parametrize("case", ["ot2d", "torus2d", "torus3d"])
def test_utop_floors_fixup_ptou_bounds_divb_pack_equals_contiguous(case):
    import kharma_amd as kh
    cfg, P, U = _state(case)
    print(f"[pack utop/floors/fixup/ptou/bounds] {case}", flush=True)
    blk = kh.Block(cfg)
    gen = torch.Generator(device="cpu").manual_seed(5)
    guess = P.clone()
    guess[:5] *= 1.0 + 1e-3 * torch.randn(guess[:5].shape, generator=gen, dtype=torch.float64).cuda()
    Ubad = U.clone()
    ng = cfg["ng"]
    Ubad[(0,) + tuple(ng + 2 if s > 1 else 0 for s in blk.shape)] = -1.0   # a planted failed inversion
    Pc, Uc = guess.clone(), Ubad.clone()
    Ps, Us = _scatter(guess, 3), _scatter(Ubad, 4)
    pf, ff = (torch.zeros(blk.shape, dtype=torch.int32, device="cuda") for _ in range(2))
    pfv, ffv = torch.zeros_like(pf), torch.zeros_like(ff)
    blk.utop(Uc, Pc, pf, ff)
    blk.utop_v(Us.vars, Ps.vars, pfv, ffv)
    assert _same(Ps.stack(), Pc) and _same(pfv, pf) and _same(ffv, ff)
    assert int((pf > 0).sum()) >= 1
    blk.apply_floors(Uc, Pc, pf, ff)
    blk.apply_floors_v(Us.vars, Ps.vars, pfv, ffv)
    assert _same(Ps.stack(), Pc) and _same(Us.stack(), Uc) and _same(ffv, ff) and _same(pfv, pf)
    blk.fix_utop(Uc, Pc, pf)
    blk.fix_utop_v(Us.vars, Ps.vars, pfv)
    assert _same(Ps.stack(), Pc) and _same(Us.stack(), Uc)
    blk.ptou(Pc, Uc)
    blk.ptou_v(Ps.vars, Us.vars)
    assert _same(Us.stack(), Uc)
    blk.apply_bounds(Uc, Pc)
    blk.apply_bounds_v(Us.vars, Ps.vars)
    assert _same(Ps.stack(), Pc) and _same(Us.stack(), Uc)
    assert blk.max_divb_v(Us.vars) == blk.max_divb(Uc)


def test_weighted_sum_in_place_and_pack():
    """AddStateUpdate's WeightedSumData pair with the output aliasing an input
    (kharma_driver.cpp:366-370), contiguous and per variable, against numpy."""
    import kharma_amd as kh
    cfg, P, U = _state("fast3d")
    blk = kh.Block(cfg)
    gen = torch.Generator(device="cpu").manual_seed(9)
    Usub = U.clone()
    Ufull = U * (1 + 1e-2 * torch.randn(U.shape, generator=gen, dtype=torch.float64).cuda())
    dU = torch.randn(U.shape, generator=gen, dtype=torch.float64).cuda()
    g0, g1, bdt = 0.5, 0.5, 0.0137
    a, f, d = (x.cpu().numpy() for x in (Usub, Ufull, dU))
    expect = (g0 * a + g1 * f) * 1.0 + bdt * d   # the two WeightedSumData calls, in order
    out_sep = torch.empty_like(U)
    blk.weighted_sum(Usub, Ufull, dU, g0, g1, bdt, out_sep)
    inplace = Ufull.clone()
    blk.weighted_sum(Usub, inplace, dU, g0, g1, bdt, inplace)
    Ss, Fs, Ds = _scatter(Usub, 11), _scatter(Ufull, 12), _scatter(dU, 13)
    blk.weighted_sum_v(Ss.vars, Fs.vars, Ds.vars, g0, g1, bdt, Fs.vars)   # in place, per variable
    torch.cuda.synchronize()
    got = out_sep.cpu().numpy()
    # the device contracts the products into fused multiply-adds: within 2 ulps of the terms
    scale = np.abs(g0 * a) + np.abs(g1 * f) + np.abs(bdt * d)
    assert np.all(np.abs(got - expect) <= 4.5e-16 * scale)
    assert np.array_equal(inplace.cpu().numpy(), got)       # in place: the same result
    assert np.array_equal(Fs.stack().cpu().numpy(), got)    # per variable, in place


@pytest.mark.parametrize("case", ["ot2d", "fast3d"])
def test_sync_bounds_fills_periodic_images(case):
    """kh_sync_bounds / kh_sync_bounds_v: every ghost zone of a periodic block equals
    its interior image, the interior untouched (the exchange behind AddBoundarySync,
    kharma_driver.cpp:138-183, for a block that is its own neighbour)."""
    import kharma_amd as kh
    cfg = CASES[case]()
    blk = kh.Block(cfg)
    gen = torch.Generator(device="cpu").manual_seed(3)
    A = torch.randn((11,) + blk.shape, generator=gen, dtype=torch.float64).cuda()
    ref = A.cpu().numpy().copy()
    interior = pr.interior(cfg)
    idx = []
    for ax, sl in enumerate(interior):
        n = blk.shape[ax]
        if sl.stop - sl.start == n:
            idx.append(np.arange(n))
        else:
            idx.append(sl.start + (np.arange(n) - sl.start) % (sl.stop - sl.start))
    expect = ref[:, idx[0][:, None, None], idx[1][None, :, None], idx[2][None, None, :]]
    blk.sync_bounds(A, 11)                       # 11 contiguous variables
    sc = _scatter(torch.from_numpy(ref[:8]).cuda(), 21)
    blk.sync_bounds_v(sc.vars[:5])                # the first 5 variables of a pack
    torch.cuda.synchronize()
    assert np.array_equal(A.cpu().numpy(), expect)
    got = sc.stack().cpu().numpy()
    assert np.array_equal(got[:5], expect[:5])
    assert np.array_equal(got[5:], ref[5:8])      # variables beyond nvar untouched


def test_vars_from_pack_and_real_flags():
    import kharma_amd as kh
    cfg = CASES["fast3d"]()
    blk = kh.Block(cfg)
    # a pack with 11 variables: cons.u at 0, cons.rho at 2, cons.uvec at 5..7, cons.B at 8..10
    pack = [torch.full(blk.shape, float(v), dtype=torch.float64, device="cuda") for v in range(11)]
    v = kh.vars_from_pack(pack, dict(RHO=2, UU=0, U1=5, B1=8))
    want = [2, 0, 5, 6, 7, 8, 9, 10]
    assert [v.var[n] for n in range(8)] == [pack[w].data_ptr() for w in want]
    with pytest.raises(kh.KharmaError):
        kh.vars_from_pack(pack, dict(RHO=2, UU=0, U1=-1, B1=8))
    with pytest.raises(kh.KharmaError):
        kh.vars_from_pack(pack, dict(RHO=2, UU=0, U1=5, B1=9))   # B3 would be index 11
    flags = torch.randint(-3, 1 << 20, blk.shape, dtype=torch.int32, device="cuda")
    real = torch.empty(blk.shape, dtype=torch.float64, device="cuda")
    back = torch.empty_like(flags)
    blk.flags_to_real(flags, real)
    blk.flags_from_real(real, back)
    torch.cuda.synchronize()
    assert torch.equal(real.cpu(), flags.cpu().double())
    assert torch.equal(back.cpu(), flags.cpu())
