"""Several meshblocks on one GPU (kh_mesh_*): a mesh split into nb1 x nb2 x nb3
equal blocks, advanced by the default stage driver with the ghost zones of
shared faces kept current and the timestep / divB reduced over all blocks
(KHARMADriver::AddBoundarySync, kharma_driver.cpp:138-183;
SetGlobalTimeStep, kharma_driver.cpp:439-460).

Every zone does the same arithmetic on the same inputs as in the one-block
driver (coordinates from the mesh's global index, the same fluxes from ghost
values equal to the neighbour's interior), so the split mesh must equal the
one-block mesh bit for bit: periodic modes (Minkowski), the Kerr-Schild torus
with blocks on its physical x1 / x2 faces, the split-precision form, and the
torus at the benchmark's layout (192x128x128 in 16 blocks of 96x64x32).  One
mesh step also meets the per-step bar against the oracle on the torus.
"""
import numpy as np
import pytest

from kharma_amd import problems as pr
from tests.helpers import install_library_geometry, interior_of, rel_l1

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    import kharma_amd
    kharma_amd.lib()


def _one_block(cfg, P0, nsteps):
    import kharma_amd as kh
    d = kh.Driver(cfg)
    d.init_prims(torch.from_numpy(P0).cuda())
    d.step(nsteps)
    P, U = d.state()
    torch.cuda.synchronize()
    return P.cpu(), U.cpu(), d.dt_history(nsteps), d.scalars()


def _mesh(cfg, P0, nsteps, nb):
    import kharma_amd as kh
    m = kh.Mesh(cfg, nb)
    m.init_prims(torch.from_numpy(P0).cuda())
    m.step(nsteps)
    P, U = m.state()
    torch.cuda.synchronize()
    return P.cpu(), U.cpu(), m.dt_history(nsteps), m.scalars()


CASES = {
    "ot2d_2x2": (lambda: pr.orszag_tang_config(n=128), pr.orszag_tang_prims, (2, 2, 1), 20),
    "ot2d_4x1": (lambda: pr.orszag_tang_config(n=128), pr.orszag_tang_prims, (4, 1, 1), 20),
    "fast3d_2x2x2": (lambda: pr.mhdmode_config(n=32, ndim=3, mode="fast", amp=1e-2), pr.mhdmode_prims, (2, 2, 2), 8),
    "fast3d_split_2x1x2": (lambda: dict(pr.mhdmode_config(n=32, ndim=3, mode="fast", amp=1e-2), precision="split"),
                           pr.mhdmode_prims, (2, 1, 2), 8),
    "torus2d_2x2": (lambda: pr.torus_config(64, 48, 1), pr.torus_prims, (2, 2, 1), 10),
    "torus3d_2x2x2": (lambda: pr.torus_config(64, 48, 32), pr.torus_prims, (2, 2, 2), 6),
    "torus3d_3x1x2": (lambda: pr.torus_config(48, 24, 16), pr.torus_prims, (3, 1, 2), 6),
    "torus3d_split_2x2x1": (lambda: dict(pr.torus_config(64, 48, 16), precision="split"), pr.torus_prims,
                            (2, 2, 1), 6),
    # fofc/on (kharma_driver.cpp:291-337): marks synced into the neighbours' ghosts,
    # polar rows on the blocks of the physical X2 faces only
    "torus2d_fofc_2x2": (lambda: dict(pr.torus_config(64, 48, 1), fofc=True, fofc_polar_cells=2), pr.torus_prims,
                         (2, 2, 1), 10),
    "torus3d_fofc_2x2x2": (lambda: dict(pr.torus_config(48, 32, 16), fofc=True, fofc_polar_cells=1),
                           pr.torus_prims, (2, 2, 2), 6),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_mesh_equals_one_block(case):
    mk, gen, nb, nsteps = CASES[case]
    cfg = mk()
    P0 = gen(cfg)
    print(f"[mesh] {case}: {nb} blocks, {nsteps} steps", flush=True)
    P1, U1, dt1, s1 = _one_block(cfg, P0, nsteps)
    Pm, Um, dtm, sm = _mesh(cfg, P0, nsteps, nb)
    assert dtm == dt1
    assert sm["time"] == s1["time"] and sm["ncycle"] == s1["ncycle"] and sm["fails"] == s1["fails"]
    assert sm["divb"] == s1["divb"]
    # every zone of the mesh (interior, periodic images and physical ghosts)
    assert torch.equal(Pm, P1), (Pm - P1).abs().max()
    assert torch.equal(Um, U1), (Um - U1).abs().max()


def test_mesh_torus_benchmark_layout():
    """The torus at its benchmark size in the layout of the benchmark's blocks:
    192x128x128 as 2 x 2 x 4 blocks of 96x64x32."""
    cfg = pr.torus_config(192, 128, 128)
    P0 = pr.torus_prims(cfg)
    P1, U1, dt1, _ = _one_block(cfg, P0, 3)
    Pm, Um, dtm, _ = _mesh(cfg, P0, 3, (2, 2, 4))
    assert dtm == dt1
    assert torch.equal(Pm, P1) and torch.equal(Um, U1)


def test_mesh_torus_step_against_oracle(oracle_lib):
    """One mesh step (2 x 2 blocks of the 2-D torus) from a state 5 steps into
    the run, against the oracle's step: the per-step bar of the torus tests."""
    import kharma_amd as kh
    cfg = pr.torus_config(64, 48, 1)
    P0 = pr.torus_prims(cfg)
    ob = oracle_lib.OracleBlock(cfg)
    install_library_geometry(ob, cfg)
    ob.init_state(P0)
    ob.step(5)
    P, U = ob.state()
    t, dt = ob.time, ob.dt
    m = kh.Mesh(cfg, (2, 2))
    m.init_prims(torch.from_numpy(P).cuda())
    m.set_clock(t, dt, 5)
    m.step(1)
    Pg, _ = m.state()
    ref = oracle_lib.OracleBlock(cfg)
    install_library_geometry(ref, cfg)
    ref.set_state(P, U)
    ref.dt = dt
    ref.step(1)
    Po, _ = ref.state()
    fma = oracle_lib.OracleBlock(cfg, variant="fma")
    install_library_geometry(fma, cfg)
    fma.set_state(P, U)
    fma.dt = dt
    fma.step(1)
    sens = rel_l1(interior_of(fma.state()[0], cfg), interior_of(Po, cfg))
    r = rel_l1(interior_of(Pg.cpu().numpy(), cfg), interior_of(Po, cfg))
    print(f"[mesh vs oracle] rel L1 {r}; oracle FMA sensitivity {sens}", flush=True)
    assert np.all(r[[0, 1, 5, 6, 7]] < 1e-12), r
    assert np.all(r <= np.maximum(1e-12, 10 * sens)), (r, sens)


@pytest.mark.parametrize("case", ["torus3d_2x2x2", "fast3d_split_2x1x2", "torus3d_fofc_2x2x2"])
def test_mesh_graph_replay_equals_eager(case):
    """kh_mesh_set_graph: every replay of the captured step equals the eager
    step bit for bit, with the program's own eager work (state read-outs,
    synchronisation) between the replays."""
    import kharma_amd as kh
    mk, gen, nb, _ = CASES[case]
    cfg = mk()
    P0 = torch.from_numpy(gen(cfg)).cuda()
    a = kh.Mesh(cfg, nb)
    a.init_prims(P0)
    g = kh.Mesh(cfg, nb)
    g.init_prims(P0)
    g.set_graph(True)
    for n in range(4):
        a.step(1)
        g.step(1)
        Pa, Ua = a.state()
        Pg, Ug = g.state()
        torch.cuda.synchronize()
        assert torch.equal(Pa, Pg) and torch.equal(Ua, Ug), (case, n)
    g.step(3)
    a.step(3)
    assert torch.equal(a.state()[0], g.state()[0])
    assert a.dt_history(7) == g.dt_history(7) and a.scalars() == g.scalars()


@pytest.mark.parametrize("case", sorted(CASES))
def test_mesh_packed_equals_per_block(case):
    """kh_mesh_set_packed: the step launched once over all blocks (the default)
    equals the step launched once per block, bit for bit, state, flags' effects,
    timestep history and scalars."""
    import kharma_amd as kh
    mk, gen, nb, _ = CASES[case]
    cfg = mk()
    P0 = torch.from_numpy(gen(cfg)).cuda()
    a = kh.Mesh(cfg, nb)
    a.set_packed(False)
    a.init_prims(P0)
    p = kh.Mesh(cfg, nb)
    p.init_prims(P0)
    for n in range(3):
        a.step(2)
        p.step(2)
        Pa, Ua = a.state()
        Pp, Up = p.state()
        torch.cuda.synchronize()
        assert torch.equal(Pa, Pp) and torch.equal(Ua, Up), (case, n)
    assert a.dt_history(6) == p.dt_history(6) and a.scalars() == p.scalars()


def test_mesh_rejects_bad_layouts():
    import kharma_amd as kh
    cfg = pr.mhdmode_config(n=32, ndim=3, mode="fast")
    kh.Mesh(cfg, (1, 1, 1))
    for nb, what in (((3, 1, 1), "divide evenly"), ((8, 1, 1), "nx >= 2 ng")):
        with pytest.raises(kh.KharmaError, match=what):
            kh.Mesh(cfg, nb)
    with pytest.raises(kh.KharmaError, match="inactive axis"):
        kh.Mesh(pr.orszag_tang_config(n=64), (2, 2, 2))
