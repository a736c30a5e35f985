"""kh_estimate_timestep_noalloc, the allocation-free variant of
kh_estimate_timestep (GRMHD::EstimateTimestep + SetGlobalTimeStep's clip,
grmhd/grmhd.cpp:310-339, kharma_driver.cpp:439-460) for callers that cannot
allocate: its partials live in a static device array shared by every call on
the GPU.  It must give the allocating entry point's timestep bit for bit, on two
calls in a row with the second block smaller than the first (the second call
reads only partials it wrote), and reject a block too large for the array on
the host, before any launch.
"""
import numpy as np
import pytest

from kharma_amd import problems as pr

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    import kharma_amd
    kharma_amd.lib()


def _speeds(cfg, gen, steps=2):
    import kharma_amd as kh
    d = kh.Driver(cfg)
    d.init_prims(torch.from_numpy(gen(cfg)).cuda())
    d.step(steps)
    P, _ = d.state()
    blk = kh.Block(cfg)
    F, cmax, cmin = blk.new_fluxes()
    for dr in range(1, pr.block_shape(cfg)[0] + 1):
        blk.get_flux(dr, P, F, cmax, cmin)
    return blk, cmax, cmin


def test_noalloc_timestep_equals_allocating_entry_point():
    big = (pr.mhdmode_config(n=32, ndim=3, mode="fast", amp=1e-2), pr.mhdmode_prims)      # 128 work-groups
    small = (dict(pr.orszag_tang_config(n=32), precision="split"), pr.orszag_tang_prims)  # 4 work-groups, FP32 speeds
    for name, (cfg, gen) in (("fast3d 32^3 fp64", big), ("Orszag-Tang 32^2 split", small)):
        blk, cmax, cmin = _speeds(cfg, gen)
        for dt_last in (1e-5, 1e9):   # the max_dt_increase clip, then the CFL value
            print(f"[noalloc dt] {name} dt_last {dt_last:g}", flush=True)
            ref = blk.estimate_timestep(cmax, cmin, dt_last)
            out = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
            blk.estimate_timestep_noalloc(cmax, cmin, dt_last, out)
            torch.cuda.synchronize()
            got = float(out.item())
            assert got == ref, (name, dt_last, got, ref)
        assert ref < 1e9


def test_noalloc_timestep_rejects_oversized_block():
    import kharma_amd as kh
    cfg = pr.mhdmode_config(n=260, ndim=3, mode="fast")     # 260^3 interior zones > 65536 * 256
    blk = kh.Block(cfg)
    dummy = torch.zeros(1, dtype=torch.float64, device="cuda")
    with pytest.raises(kh.KharmaError, match="static scratch"):
        blk.estimate_timestep_noalloc(dummy, dummy, 1e-3, dummy)
    assert np.isfinite(dummy.item())
