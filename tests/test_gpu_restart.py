"""Restart dumps (kh_driver_save_restart / kh_mesh_save_restart).

The reference's restart test (tests/restart/run.sh) runs the torus, restarts
from a dump of the Metadata::Restart fields and requires the two final states
to agree to 1e-9 relative.  Here the dump restores U and P exactly and the
clock, so a run of N steps and a run of k steps + dump + reload + N - k steps
must agree bit for bit: on the one-block torus, on a mesh of blocks, and across
the two (a mesh's dump continued by the one-block driver and back).
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


def _state(run):
    P, U = run.state()
    torch.cuda.synchronize()
    return P.cpu(), U.cpu()


@pytest.mark.parametrize("prec", ["fp64", "split"])
def test_driver_restart_is_exact(tmp_path, prec):
    import kharma_amd as kh
    cfg = dict(pr.torus_config(48, 32, 16), precision=prec)
    P0 = torch.from_numpy(pr.torus_prims(cfg)).cuda()
    a = kh.Driver(cfg)
    a.init_prims(P0)
    a.step(7)
    ref = _state(a)
    ref_s = a.scalars()
    b = kh.Driver(cfg)
    b.init_prims(P0)
    b.step(3)
    path = str(tmp_path / "torus.rst")
    b.save_restart(path)
    f = kh.read_restart(path)
    Pb, Ub = _state(b)
    assert f["ncycle"] == 3 and np.array_equal(f["P"], Pb.numpy()) and np.array_equal(f["U"], Ub.numpy())
    c = kh.Driver(cfg)
    c.load_restart(path)
    c.step(4)
    got = _state(c)
    s = c.scalars()
    print(f"[restart] {prec}: t {s['time']:.6e} vs {ref_s['time']:.6e}, cycle {s['ncycle']}", flush=True)
    assert s["ncycle"] == ref_s["ncycle"] and s["time"] == ref_s["time"] and s["dt"] == ref_s["dt"]
    assert torch.equal(got[0], ref[0]) and torch.equal(got[1], ref[1])


def test_mesh_restart_is_exact_and_portable(tmp_path):
    import kharma_amd as kh
    cfg = pr.torus_config(64, 48, 32)
    P0 = torch.from_numpy(pr.torus_prims(cfg)).cuda()
    m = kh.Mesh(cfg, (2, 2, 2))
    m.init_prims(P0)
    m.step(6)
    ref = _state(m)
    m2 = kh.Mesh(cfg, (2, 2, 2))
    m2.init_prims(P0)
    m2.step(2)
    path = str(tmp_path / "mesh.rst")
    m2.save_restart(path)
    m3 = kh.Mesh(cfg, (2, 2, 2))
    m3.load_restart(path)
    m3.step(4)
    got = _state(m3)
    assert torch.equal(got[0], ref[0]) and torch.equal(got[1], ref[1])
    # the mesh's dump continued by the one-block driver, and that driver's dump by another layout
    d = kh.Driver(cfg)
    d.load_restart(path)
    d.step(2)
    path2 = str(tmp_path / "one_block.rst")
    d.save_restart(path2)
    m4 = kh.Mesh(cfg, (1, 2, 4))
    m4.load_restart(path2)
    m4.step(2)
    got2 = _state(m4)
    assert torch.equal(got2[0], ref[0]) and torch.equal(got2[1], ref[1])


def test_restart_rejects_another_mesh(tmp_path):
    import kharma_amd as kh
    cfg = pr.torus_config(48, 32, 16)
    d = kh.Driver(cfg)
    d.init_prims(torch.from_numpy(pr.torus_prims(cfg)).cuda())
    path = str(tmp_path / "a.rst")
    d.save_restart(path)
    other = kh.Driver(pr.torus_config(64, 32, 16))
    with pytest.raises(kh.KharmaError, match="zones differ"):
        other.load_restart(path)
    with pytest.raises(kh.KharmaError, match="cannot open"):
        d.load_restart(str(tmp_path / "missing.rst"))
