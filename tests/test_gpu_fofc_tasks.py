"""The first-order flux correction through its task-level entry points, against
the oracle, with the floors the torus atmosphere was set at.

The chain is KHARMADriver::AddFOFC (driver/kharma_driver.cpp:292-337) as a
KHARMA task list would call it: the raw fluxes' divergence, the trial update
(WeightedSumData), the trial state's U->P (Inverter + B_FluxCT), the floor check
(Floors::DetermineGRMHDFloors, floors/floors.cpp:197-231), Flux::MarkFOFC
(flux/fofc.cpp:45-111) and Flux::FOFC (flux/fofc.cpp:113-226).

Marks: a zone whose trial state lies within rounding of a floor threshold takes
its mark from the last bits of that state, which the GPU's reciprocal-based
arithmetic and the oracle's IEEE divisions round differently.  The marks must
agree everywhere else; the zones where they may differ are those whose trial
state is within 1e-9 (relative) of a threshold, and the test reports how many
there are.  The replacement fluxes are then compared with the marks held equal:
the GPU's kh_fofc_v runs on the oracle's marks.
"""
import numpy as np
import pytest

from kharma_amd import problems as pr
from tests.helpers import face_slices, install_library_geometry

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NV = 8


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    import kharma_amd
    kharma_amd.lib()


def _cuda(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def _np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().double().numpy()


CASES = {
    # the torus at the floors its atmosphere was set at, horizon buffer and two polar rows
    "torus2d": lambda: dict(pr.torus_config(64, 48, 1), fofc=True, fofc_polar_cells=2),
    "torus3d": lambda: dict(pr.torus_config(32, 24, 8), fofc=True, fofc_polar_cells=1),
    # a flat 3-D mode whose trial densities straddle a constant floor
    "fast3d": lambda: dict(pr.mhdmode_config(n=16, ndim=3, mode="fast", amp=0.05), driver="kharma", fofc=True,
                           floor_values=dict(rho_min_const=1.0, u_min_const=1e-10)),
}


def _setup(oracle_lib, name, nsteps):
    cfg = CASES[name]()
    spherical = cfg.get("coords") == "ks"
    P0 = pr.torus_prims(cfg) if spherical else pr.mhdmode_prims(cfg)
    ob = oracle_lib.OracleBlock(cfg)
    if spherical:
        install_library_geometry(ob, cfg)
    ob.init_state(P0)
    ob.step(nsteps)
    P, U = ob.state()
    dt = ob.dt
    ob = oracle_lib.OracleBlock(cfg)
    if spherical:
        install_library_geometry(ob, cfg)
    ob.set_state(P, U)
    ob.dt = dt
    return cfg, P, U, dt, ob


@pytest.mark.parametrize("case", ["torus2d", "torus3d", "fast3d"])
def test_fofc_task_chain_against_oracle(oracle_lib, case):
    import kharma_amd as kh
    cfg, P, U, dt, ob = _setup(oracle_lib, case, 4)
    print(f"[fofc tasks] {case} dt {dt:.3e}", flush=True)
    blk = kh.Block(cfg)
    nd = ob.ndim
    F, cmax, cmin = blk.new_fluxes()
    Pd, Ud = _cuda(P), _cuda(U)
    for d in range(1, nd + 1):
        ob.get_flux(d)
        blk.get_flux(d, Pd, F, cmax, cmin)
    ob.fofc_stage1()
    mark_o = ob.array(9).astype(np.int32)
    trial_o = ob.array(10)
    margin = ob.array(11)
    Fo, mxo, mno = ob.array(0), ob.array(1), ob.array(2)
    # the task chain on the GPU, each variable in its own array
    Fv = [[F[d, v] for v in range(NV)] for d in range(3)]
    Pv = [Pd[v] for v in range(NV)]
    dU = torch.zeros_like(Pd)
    blk.flux_divergence_v(Fv, [dU[v] for v in range(NV)])
    Ug = torch.zeros_like(Ud)
    blk.weighted_sum_v([Ud[v] for v in range(NV)], [Ud[v] for v in range(NV)], [dU[v] for v in range(NV)],
                       0.0, 1.0, 1.0 * dt, [Ug[v] for v in range(NV)])
    Pg = Pd.clone()
    pflag = torch.zeros(blk.shape, dtype=torch.int32, device="cuda")
    fflag = torch.zeros_like(pflag)
    blk.utop_v([Ug[v] for v in range(NV)], [Pg[v] for v in range(NV)], pflag, fflag)
    blk.determine_floors_v([Pg[v] for v in range(NV)], fflag)
    mark_g = torch.zeros_like(pflag)
    blk.mark_fofc(fflag, mark_g)
    # trial states agree to the inversion's rounding
    inter = pr.interior(cfg)
    sl = (slice(None),) + inter
    tg, to = _np(Pg)[sl], trial_o[sl]
    rel = np.abs(tg - to) / (np.abs(to) + 1e-300)
    assert np.median(rel[:5]) < 1e-12, np.median(rel[:5])
    # marks: equal wherever the trial state is not within rounding of a floor threshold
    mg, mo, mar = _np(mark_g)[inter].astype(np.int32), mark_o[inter], margin[inter]
    near = mar < 1e-9
    differ = mg != mo
    n_marked = int(mo.sum())
    print(f"[fofc tasks] {case}: {n_marked} of {mo.size} zones marked; {int(near.sum())} within 1e-9 of a "
          f"floor threshold; {int(differ.sum())} marks differ", flush=True)
    assert n_marked > 0
    assert not np.any(differ & ~near), np.argwhere(differ & ~near)[:10]
    # replacement fluxes with the marks held equal (the oracle's marks on the GPU)
    mark_eq = _cuda(mark_o, torch.int32)
    blk.fofc_v(Pv, mark_eq, Fv, cmax, cmin)
    Fg, mxg, mng = _np(F), _np(cmax), _np(cmin)
    for d in range(1, nd + 1):
        fs = face_slices(cfg, d)
        a, b = Fg[d - 1][(slice(None),) + fs], Fo[d - 1][(slice(None),) + fs]
        scale = np.abs(b).max() + 1e-300
        err = np.abs(a - b).max(axis=(1, 2, 3)) / scale
        assert np.all(err < 1e-12), (d, err)
        sm = max(np.abs(mxo[d - 1][fs]).max(), np.abs(mno[d - 1][fs]).max())
        assert np.max(np.abs(mxg[d - 1][fs] - mxo[d - 1][fs])) <= 1e-12 * sm
        assert np.max(np.abs(mng[d - 1][fs] - mno[d - 1][fs])) <= 1e-12 * sm
