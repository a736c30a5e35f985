"""Benchmark of the GRMHD step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload torus|fastmode] [--no-cpu]

Default workload (torus): 3-D Fishbone-Moncrief torus around a Kerr black hole,
a = 0.9375, FMKS coordinates, 192 x 128 x 128 zones in one meshblock (4 ghost
zones), r_out = 1000, gamma = 13/9, SANE seed field normalised to beta = 100,
internal energy jittered by +-4 %, WENO5 + LLF, flux-CT, RK2, onedw FP64
inverter, drift-frame floors, outflow / reflecting / periodic boundaries --
the precision form at production shape.  `--workload fastmode` runs the 3-D
linear fast mode on the periodic unit cube at --n^3 (default 128).

Both forms run back to back from the same initial state:
  * plain FP64,
  * split precision (FP32 reconstruction / fluxes / EMFs over FP64 U, inverter,
    geometry, dt),
and the line reports the split form's relative L1 deviation from the FP64 run
after the same number of steps.  A "step" is one full RK2 step; a zone-cycle
is one interior zone advanced by one step.  `value` is the split form's
zone-cycles/s with the state resident in HBM before the timed region starts.

Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "zone-cycles/s on 1 MI355X: FP32-split vs FP64 plain; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0


def _bench_mesh(cfg, P0, nb, steps, warmup, packed=True):
    """zone-cycles timing of the same workload split into nb blocks (kh_mesh_*):
    HIP events on the stream around `steps` steps after `warmup`; packed: each
    kernel launched once over all blocks (kh_mesh_set_packed), else once per block."""
    import torch
    import kharma_amd as kh
    m = kh.Mesh(cfg, nb)
    m.set_packed(packed)
    m.init_prims(torch.from_numpy(P0).cuda())
    m.step(warmup)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    m.step(steps)
    stop.record()
    torch.cuda.synchronize()
    m.check()
    return start.elapsed_time(stop)


def _bench_form(cfg, P0, steps, warmup):
    import torch
    import kharma_amd as kh
    drv = kh.Driver(cfg)
    drv.init_prims(torch.from_numpy(P0).cuda())
    drv.step(warmup)
    torch.cuda.synchronize()
    drv.set_kernel_timing(True)
    drv.kernel_times()  # reset accumulators
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    drv.step(steps)
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop)
    ktimes = drv.kernel_times()
    drv.set_kernel_timing(False)
    drv.check()   # the reference's stop on a large divB (b_field/kill_on_large_divb)
    s = drv.scalars()
    P, _ = drv.state()
    torch.cuda.synchronize()
    return dict(ms=ms, kernels=ktimes, scalars=s, P=P)


# Algorithmic HBM bytes per interior zone for each kernel of the step: the FP64
# state contract of one RK2 step (U and P in and out, 8 FP64 each) apportioned
# to the kernel that owns it.  Stage 1 reads P0, U0 and writes U1, P1 (4 x 64 B);
# stage 2 reads P1, U1, U0 and writes U2, P2 (5 x 64 B): 576 B per zone-cycle.
# Face fluxes, signal speeds, flags and ghost images are implementation
# intermediates and count only in the measured traffic.
STATE_BYTES = {"update_s1": 256, "update_s2": 320}
STEP_STATE_BYTES = 576


def _kernel_io_bytes(kind, prec_bytes):
    """Bytes per interior zone that `kind` must move as this implementation splits the
    step (every array it reads or writes touched once; ghost images excluded)."""
    fb = prec_bytes
    if kind.startswith("flux_x"):
        # P (FP32 shadow in the split form) in, 8 fluxes out, + 1 signal speed at stage 2
        return 8 * fb + 8 * fb + (fb if kind.endswith("s2") else 0)
    if kind == "update_s1":
        # 3 x 8 face fluxes, U, 5-var P guess in; U, P, FP32 P shadow, 2 flags out
        return 24 * fb + 64 + 40 + 64 + 64 + (32 if fb == 4 else 0) + 8
    if kind == "update_s2":
        return _kernel_io_bytes("update_s1", fb) + 64 + 3 * fb   # + U0, + 3 signal speeds
    return None


def _valu_fractions(tag):
    """Compute-side roofline of the split form's kernels from the committed SQ
    counter summary (profiles/valu_<workload>.json): VALU instructions issued as
    a fraction of the chip's issue capacity (one wave64 instruction per SIMD per
    4 cycles, which for FP64 FMA is the 78.6 TFLOP/s vector peak) and the share
    of the resident waves' cycles spent waiting on memory.  The update pair is
    weighted by the kernels' times."""
    f = os.path.join(ROOT, "profiles", f"valu_{tag}.json")
    if not os.path.exists(f):
        return None
    k = json.load(open(f))["kernels"]["split"]
    out = {"peak": "78.6 TFLOP/s FP64 vector (VALU issue capacity)", "source": os.path.relpath(f, ROOT)}
    for ro in ("combine_s2", "recover_s2", "flux_x1_s2", "flux_x2_s2", "flux_x3_s2"):
        if ro in k:
            out[ro] = {"valu_issue_frac": k[ro]["valu_issue_frac"], "wait_frac": k[ro]["wait_any_per_wave_cycle"]}
    if "combine_s2" in k and "recover_s2" in k:
        a, b = k["combine_s2"], k["recover_s2"]
        ta, tb = a["avg_us_under_counters"], b["avg_us_under_counters"]
        out["update_s2_valu_issue_frac"] = (a["valu_issue_frac"] * ta + b["valu_issue_frac"] * tb) / (ta + tb)
    return out


def _cpu_model():
    """The host CPU's model name (from /proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _cpu_baseline(cfg, P0, what, budget_s=12.0, max_steps=8, warm=1, fB=None):
    from oracle import oracle
    oracle.build()
    ob = oracle.OracleBlock(cfg)
    if fB is not None:
        ob.init_state_faces(P0, fB)
    else:
        ob.init_state(P0)
    ob.step(warm)  # warm-up (first touch, dt ramp start)
    nz = cfg["nx"][0] * cfg["nx"][1] * cfg["nx"][2]
    t0 = time.perf_counter()
    n = 0
    while n < max_steps and (time.perf_counter() - t0) < budget_s:
        ob.step(1)
        n += 1
    dt = time.perf_counter() - t0
    return dict(value=nz * n / dt, unit="zone-cycles/s", cores=oracle.num_threads(), kind="port",
                cpu_model=_cpu_model(),
                sample=f"{n} RK2 steps of the same {what} after {warm} warm-up step(s), "
                       f"CPU restatement of KHARMA's algorithm (C++ FP64, OpenMP over k,j)")


def _other_configs(with_cpu):
    """The CPU-runnable configurations beside the headline: C1 (2-D Orszag-Tang
    256^2, FP64, 100 steps), C2 at 64^3 (3-D fast mode, FP64 and split) and the
    Orszag-Tang vortex with face-centred CT (gs05_c, HLLE, 256^2, both forms),
    zone-cycles/s on the GPU (HIP events around the steps, after 2 warm-up
    steps) and of the CPU restatement on a bounded sample (2 warm-up steps)."""
    import torch
    import kharma_amd as kh
    from kharma_amd import problems as pr
    out = {}
    ot_ct = dict(pr.orszag_tang_config(n=256, init="reference", riemann="hlle"), b_solver="face_ct", ct_scheme="gs05_c")
    runs = [("c1_orszag_tang_256", pr.orszag_tang_config(n=256), pr.orszag_tang_prims, 100, ("fp64",)),
            ("c2_fast_mode_64", pr.mhdmode_config(n=64, ndim=3, mode="fast", one_period=False), pr.mhdmode_prims, 50,
             ("fp64", "split")),
            # the reference's face-CT Orszag-Tang setting (pars/tests/orszag_tang.par:37-43) at 256^2
            ("orszag_tang_256_face_ct_gs05_c", ot_ct, pr.orszag_tang_prims, 100, ("fp64", "split"))]
    for name, cfg, gen, steps, forms in runs:
        P0 = gen(cfg)
        fB0 = pr.orszag_tang_faces(cfg) if cfg.get("b_solver") == "face_ct" else None
        nz = cfg["nx"][0] * cfg["nx"][1] * cfg["nx"][2]
        rec = {"zones": nz, "steps": steps}
        for f in forms:
            d = kh.Driver(dict(cfg, precision=f))
            if fB0 is not None:
                d.init_faces(torch.from_numpy(P0).cuda(), torch.from_numpy(fB0).cuda())
            else:
                d.init_prims(torch.from_numpy(P0).cuda())
            d.step(2)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record()
            d.step(steps)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1)
            rec[f"gpu_{f}"] = {"zone_cycles_per_s": nz * steps / (ms * 1e-3), "ms_per_step": ms / steps}
        if with_cpu:
            rec["cpu"] = _cpu_baseline(cfg, P0, name.replace("_", " "), budget_s=6.0, max_steps=steps, warm=2, fB=fB0)
            rec["gpu_fp64_over_cpu"] = rec["gpu_fp64"]["zone_cycles_per_s"] / rec["cpu"]["value"]
        out[name] = rec
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="torus", choices=["torus", "fastmode"])
    ap.add_argument("--n", type=int, default=128, help="fastmode: zones per side")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiation-form measurement")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the differentiation-form measurements (Alfven 64^3 and 2-D torus) and print "
                         "their results (profiling runs)")
    ap.add_argument("--forms", default="fp64,split", help="comma list of fp64,split (profiling runs use one)")
    ap.add_argument("--fofc", action="store_true",
                    help="profiling mode (--forms fp64): run the step with the first-order flux correction on")
    ap.add_argument("--mesh", default=None,
                    help="profiling mode: run the torus as this block layout (e.g. 2x2x4) with the mesh "
                         "driver (each kernel once over all blocks) and print its timings")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="after the runs, copy a buffer of this many MiB on the device: a known-size "
                         "reference for scaling HBM byte counters (FETCH_SIZE / WRITE_SIZE)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": a.gpus}))
        sys.exit(2)

    import numpy as np
    import torch
    import kharma_amd as kh
    from kharma_amd import problems as pr

    if a.adjoint_only:
        print(json.dumps({"adjoint": _bench_adjoint(ndirs=8), "adjoint_torus": _bench_adjoint_torus(ndirs=2)}))
        return

    if a.workload == "torus":
        cfg = pr.torus_config(192, 128, 128)
        P0 = pr.torus_prims(cfg)
        tag = "torus192"
        what = "FM torus 192x128x128 block"
        desc = ("3-D Fishbone-Moncrief torus, Kerr a=0.9375, FMKS, 192x128x128 in 1 meshblock, r_out 1000, "
                "gamma 13/9, SANE field beta 100, u jitter +-4%, WENO5+LLF, flux-CT, RK2, onedw FP64 inverter, "
                "drift floors, outflow/reflecting/periodic boundaries")
    else:
        cfg = pr.mhdmode_config(n=a.n, ndim=3, mode="fast", one_period=False)
        P0 = pr.mhdmode_prims(cfg)
        tag = f"fastmode{a.n}"
        what = f"3-D fast mode {a.n}^3 block"
        desc = (f"3-D linear MHD fast mode, periodic unit cube {a.n}^3 in 1 meshblock, gamma 4/3, "
                "WENO5+LLF, flux-CT, RK2, onedw FP64 inverter, drift floors")
    nz = cfg["nx"][0] * cfg["nx"][1] * cfg["nx"][2]

    forms = a.forms.split(",")
    if a.mesh:
        nb = tuple(int(x) for x in a.mesh.split("x"))
        for f in forms:
            ms = _bench_mesh(dict(cfg, precision=f), P0, nb, a.steps, a.warmup)
            print(json.dumps({"form": f, "blocks": a.mesh, "ms_per_step": ms / a.steps,
                              "zone_cycles_per_s": nz * a.steps / (ms * 1e-3)}))
        return
    if forms != ["fp64", "split"]:
        # profiling mode: run the requested forms only, print their timings, no JSON contract line
        for f in forms:
            r = _bench_form(dict(cfg, precision=f, fofc=a.fofc), P0, a.steps, a.warmup)
            print(json.dumps({"form": f, "fofc": a.fofc, "ms_per_step": r["ms"] / a.steps,
                              "zone_cycles_per_s": nz * a.steps / (r["ms"] * 1e-3)}))
        if a.copy_mb:
            _calibration_copy(a.copy_mb)
        return
    r64 = _bench_form(dict(cfg, precision="fp64"), P0, a.steps, a.warmup)
    r32 = _bench_form(dict(cfg, precision="split"), P0, a.steps, a.warmup)
    # cost of the first-order flux correction (fofc/on with its defaults) on the same
    # workload, FP64 form, default driver only
    fofc = None
    if cfg.get("driver", "kharma") == "kharma":
        nf, wf = min(a.steps, 50), min(a.warmup, 10)
        rf = _bench_form(dict(cfg, precision="fp64", fofc=True), P0, nf, wf)
        rb = _bench_form(dict(cfg, precision="fp64"), P0, nf, wf)
        fofc = {"options": "fofc/on, eh_buffer 0.1, polar_cells 0", "steps": nf, "warmup": wf,
                "ms_per_step": rf["ms"] / nf, "plain_ms_per_step": rb["ms"] / nf,
                "over_plain": rf["ms"] / rb["ms"]}
    # the benchmark's own block layout: 192x128x128 as 16 blocks of 96x64x32
    layout = None
    if a.workload == "torus":
        nb = (2, 2, 4)
        layout = {"blocks": "2x2x4 blocks of 96x64x32 (ghost zones per block 4, each block's own)",
                  "launches": "fp64/split: each kernel once over all 16 blocks; *_per_block: once per block"}
        for f in ("fp64", "split"):
            for packed in (True, False):
                ms = _bench_mesh(dict(cfg, precision=f), P0, nb, a.steps, a.warmup, packed)
                layout[f + ("" if packed else "_per_block")] = {"zone_cycles_per_s": nz * a.steps / (ms * 1e-3),
                                                                "ms_per_step": ms / a.steps}
        if fofc is not None:   # the same first-order flux correction over the 16 blocks
            nf, wf = fofc["steps"], fofc["warmup"]
            ms = _bench_mesh(dict(cfg, precision="fp64", fofc=True), P0, nb, nf, wf)
            layout["fofc_fp64"] = {"steps": nf, "ms_per_step": ms / nf,
                                   "over_plain": (ms / nf) / layout["fp64"]["ms_per_step"]}
    adj = None if a.no_adjoint else _bench_adjoint()
    adj_gr = None if a.no_adjoint else _bench_adjoint_torus()
    zc64 = nz * a.steps / (r64["ms"] * 1e-3)
    zc32 = nz * a.steps / (r32["ms"] * 1e-3)

    # deviation of the split form from the FP64 form after the same step count
    sl = (slice(None),) + pr.interior(cfg)
    P64 = r64["P"].double().cpu().numpy()[sl]
    P32 = r32["P"].double().cpu().numpy()[sl]
    rel = (np.abs(P32 - P64).sum(axis=(1, 2, 3)) / np.abs(P64).sum(axis=(1, 2, 3)).clip(1e-300))

    # one-step deviation: one split step and one FP64 step from the same (FP64) state
    dt_last = r64["scalars"]["dt"]
    one = {}
    for f in ("fp64", "split"):
        drv = kh.Driver(dict(cfg, precision=f))
        drv.init_prims(r64["P"])
        drv.set_clock(r64["scalars"]["time"], dt_last, r64["scalars"]["ncycle"])
        drv.step(1)
        one[f] = drv.state()[0].double().cpu().numpy()[sl]
        del drv
    rel1 = (np.abs(one["split"] - one["fp64"]).sum(axis=(1, 2, 3)) /
            np.abs(one["fp64"]).sum(axis=(1, 2, 3)).clip(1e-300))

    # dominant kernel of the reported (split) form
    kt = {k: v for k, v in r32["kernels"].items() if v[1] > 0}
    dom = max(kt, key=lambda k: kt[k][0])
    tot_ms, nl = kt[dom]
    avg_ms = tot_ms / nl
    per_zone = STATE_BYTES.get(dom)
    algo = per_zone * nz if per_zone else None
    achieved = (algo / (avg_ms * 1e-3) / 1e9) if algo else None
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"hbm_traffic_{tag}.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("kernels", {}).get("split", {}).get(dom, {}).get("hbm_bytes_per_launch")
    io = _kernel_io_bytes(dom, 4)
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": traffic,
            "kernel": dom, "avg_launch_ms": avg_ms, "algorithmic_bytes_per_launch": algo,
            "algorithmic_bytes_per_zone": per_zone, "zones_per_launch": nz,
            "kernel_io_bytes_per_launch": io * nz if io else None,
            "valu": _valu_fractions(tag)}
    step_bytes = STEP_STATE_BYTES * nz
    kernel_ms = {k: {"total_ms": v[0], "launches": v[1], "avg_ms": (v[0] / v[1]) if v[1] else None}
                 for k, v in r32["kernels"].items()}
    kernel_ms64 = {k: {"total_ms": v[0], "launches": v[1], "avg_ms": (v[0] / v[1]) if v[1] else None}
                   for k, v in r64["kernels"].items()}

    out = {
        "metric": METRIC,
        "value": zc32,
        "unit": "zone-cycles/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r32["ms"] / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64",
        "data": "synthetic",
        "config": {"workload": desc, "zones": nz, "ghosts": cfg["ng"], "forms": ["fp64", "split"],
                   "layout": "one meshblock (value); the 16-block layout is reported in layout_16_blocks"},
        "fp64": {"value": zc64, "ms_per_step": r64["ms"] / a.steps, "dtype": "f64"},
        "split_over_fp64": zc32 / zc64,
        "split_rel_l1_vs_fp64": {"after_steps": a.steps + a.warmup, "max": float(rel.max()),
                                 "per_var": [float(x) for x in rel],
                                 "one_step_max": float(rel1.max()), "one_step_per_var": [float(x) for x in rel1]},
        "fofc_fp64": fofc,
        "layout_16_blocks": layout,
        "adjoint_over_forward": adj["adjoint_over_forward"] if adj else None,
        "adjoint": adj,
        "adjoint_torus": adj_gr,
        "step_state_roofline": {"bytes_per_zone_cycle": STEP_STATE_BYTES, "achieved_GBs_fp64": step_bytes * a.steps / (r64["ms"] * 1e-3) / 1e9,
                                "achieved_GBs_split": step_bytes * a.steps / (r32["ms"] * 1e-3) / 1e9,
                                "peak": HBM_PEAK_GBS},
        "roofline": roof,
        "kernels_split": kernel_ms,
        "kernels_fp64": kernel_ms64,
        "sim_time": {"fp64": r64["scalars"]["time"], "split": r32["scalars"]["time"],
                     "divb_fp64": r64["scalars"]["divb"], "divb_split": r32["scalars"]["divb"]},
    }
    if not a.no_cpu:
        out["cpu_baseline"] = _cpu_baseline(cfg, P0, what)
    else:
        out["cpu_baseline"] = None
    out["other_configs"] = _other_configs(not a.no_cpu)
    print(json.dumps(out))
    if a.copy_mb:
        _calibration_copy(a.copy_mb)


def _bench_adjoint(n=64, nsteps=20, ndirs=64, seed=20261202, h=1e-7):
    """Differentiation form on the Alfven mode: gradient of L = |P_N - P_target|^2
    after nsteps FP64 RK2 steps (P_target: the mode at phase + pi/4) with respect
    to the initial primitive field, by the adjoint sweep (per-step checkpoints);
    adjoint/forward GPU time per step; check against central differences (step h).

    Directions: the mode is uniform along x3 and rho, u, B1 are uniform
    everywhere, where WENO5's smoothness indicators are exactly zero and any
    finite perturbation moves its nonlinear weights by O(1), so differences of
    the literal state are taken in ndirs x3-independent white-noise directions in
    the wave's fields u3, B3 (the directions along which the map is smooth at
    this state).  The same objective is then checked from a generic smooth state
    (the mode plus a 1e-3 Fourier mode in every variable, uniform nowhere) in
    smooth random and white-noise directions over all 8 variables."""
    import math
    import numpy as np
    import torch
    from kharma_amd import problems as pr
    from kharma_amd.adjoint import Evolution, generic_state, noise_direction, smooth_direction

    cfg = pr.mhdmode_config(n=n, ndim=3, mode="alfven", amp=1e-4, one_period=False)
    # start at the CFL step of the fast speed (~0.66 c for this background) instead of dt_min
    cfg["dt_start"] = 0.9 / (n * 3 * 0.7)
    evo = Evolution(cfg)
    inner = (slice(None),) + pr.interior(cfg)
    Pt = torch.as_tensor(pr.mhdmode_state(cfg, 0.0, phase=math.pi / 4), device="cuda")
    rng = np.random.default_rng(seed)

    def grad_and_check(P0, dirs):
        dts = evo.run(P0, nsteps)[1]          # the CFL timestep sequence, then held fixed
        PN, _, ck, _ = evo.run(P0, nsteps, dts, checkpoints=True)
        lam = evo._embed(2.0 * (PN - Pt))
        rec = adj = 0.0
        recs = ck["records"]   # every step's stage intermediates, recorded by the forward run
        for k in reversed(range(nsteps)):
            evo.adj.step_recorded(ck[k], recs[k], dts[k], lam, ck[k + 1])
            f, g_ = evo.adj.last_times()
            rec += f
            adj += g_
        g = lam[inner].cpu().numpy()
        loss = lambda P: float(((evo.run(P, nsteps, dts)[0] - Pt) ** 2).sum())
        fd, ad = [], []
        for v in dirs:
            vt = torch.as_tensor(v, device="cuda")
            fd.append((loss(P0 + h * vt) - loss(P0 - h * vt)) / (2 * h))
            ad.append(float((g * v).sum()))
        fd, ad = np.array(fd), np.array(ad)
        err = np.abs(fd - ad)
        chk = {"directions": len(fd), "max_abs_err_over_max_fd": float(err.max() / np.abs(fd).max()),
               "median_rel_err": float(np.median(err / np.abs(fd).clip(1e-300)))}
        return dts, rec / nsteps, adj / nsteps, chk

    # literal C4 state
    P0 = torch.as_tensor(pr.mhdmode_prims(cfg)[inner].copy(), device="cuda")
    # warm-up: the first launches of each kernel load its code object
    _w = evo.run(P0, 2, checkpoints=True)
    _l = evo._embed(torch.zeros_like(P0))
    for k in reversed(range(2)):
        evo.adj.step(_w[2][k], _w[1][k], _l, _w[2][k + 1])
        evo.adj.step_recorded(_w[2][k], _w[2]["records"][k], _w[1][k], _l, _w[2][k + 1])
    dirs = []
    for _ in range(ndirs):
        v = noise_direction(cfg, rng, (pr.U3, pr.B3))
        v[:] = v[:, :1]
        dirs.append(v / np.linalg.norm(v))
    dts, rec_ms, adj_ms, chk_c4 = grad_and_check(P0, dirs)
    # forward GPU time per step (events on the stream the driver runs on, around
    # the steps alone; median of three runs), plain and with the stage-1 record
    # the differentiable forward run keeps (a copy of the stage-1 state and flags per step)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    rec_buf = evo.drv.new_stage_record()

    def fwd_time(record):
        runs = []
        for _ in range(3):
            evo.drv.init_prims(evo._embed(P0))
            if record:
                evo.drv.set_stage_record(rec_buf)
            torch.cuda.synchronize()
            e0.record()
            t = 0.0
            for k in range(nsteps):
                evo.drv.set_clock(t, dts[k], k)
                evo.drv.step(1)
                t += dts[k]
            e1.record()
            torch.cuda.synchronize()
            evo.drv.set_stage_record()
            runs.append(e0.elapsed_time(e1) / nsteps)
        return sorted(runs)[1]
    fwd_plain_ms = fwd_time(False)
    fwd_ms = fwd_time(True)
    # generic smooth state
    Pg = torch.as_tensor(generic_state(cfg, rng), device="cuda")
    dirs = [smooth_direction(cfg, rng) for _ in range(16)] + [noise_direction(cfg, rng, range(pr.NV)) for _ in range(8)]
    _, _, _, chk_gen = grad_and_check(Pg, dirs)
    back_ms = rec_ms + adj_ms
    return {"workload": f"Alfven mode {n}^3, amp 1e-4, {nsteps} RK2 steps, FP64, per-step checkpoints "
                        "with each step's stage intermediates recorded",
            "loss": "|P_N - P_target|^2, P_target = mode at phase + pi/4",
            "forward_ms_per_step": fwd_ms, "forward_plain_ms_per_step": fwd_plain_ms,
            "backward_ms_per_step": back_ms,
            "backward_checkpoint_load_ms_per_step": rec_ms, "backward_transpose_ms_per_step": adj_ms,
            "adjoint_over_forward": back_ms / fwd_ms, "fd_step": h, "seed": seed,
            "grad_check": {"c4_state_x3_independent_u3_B3_noise": chk_c4,
                           "generic_state_smooth_and_noise_all_vars": chk_gen}}


def _bench_adjoint_torus(nx1=256, nx2=128, nsteps=500, every=25, ndirs=8, seed=20261203, h=1e-5):
    """2-D FMKS torus (the 3-D torus physics at 256 x 128), FP64, nsteps steps:
    J = time-integrated mass flux through the x1 face nearest r = 2.5 (the
    face-1 density flux, RK-weighted over the stages, times dx2 dx3), and its
    gradient with respect to the internal-energy perturbation field delta
    (u = u_0 (1 + delta), the seeded +-4 % jitter) by the adjoint sweep with
    two-level (revolve-style) checkpointing every `every` steps.  Timed with
    HIP events: the plain forward run, the forward run with checkpoints, the
    backward sweep (segment re-runs that record each step's stage
    intermediates, then adjoint steps from those records), and the adjoint
    steps' own record-load / transpose times.  Checked against central differences of
    the driver along `ndirs` random directions of delta (relative step h)."""
    import math
    import numpy as np
    import torch
    from kharma_amd import problems as pr
    from kharma_amd.adjoint import Evolution, evolve_objective, face_index

    cfg = pr.torus_config(nx1, nx2, 1)
    inner = (slice(None),) + pr.interior(cfg)
    face = face_index(cfg, 1, math.log(2.5))
    r_face = math.exp(cfg["bounds"][0][0] + (face - cfg["ng"]) * (cfg["bounds"][0][1] - cfg["bounds"][0][0]) / nx1)
    evo = Evolution(cfg, objective=dict(dir=1, var=pr.RHO, face=face, weight=1.0))
    P0 = torch.as_tensor(pr.torus_prims(cfg)[inner].copy(), device="cuda")
    u0 = torch.as_tensor(pr.torus_prims(dict(cfg, torus=dict(cfg["torus"], u_jitter=0.0)))[inner][pr.UU].copy(),
                         device="cuda")
    dts = evo.run(P0, nsteps)[1]
    # warm-up: the first launches of each kernel load its code object
    _Pw = P0.clone().requires_grad_(True)
    evolve_objective(_Pw, evo, 2, dts[:2], every=1)[1].backward()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]

    def plain_forward():
        evo.drv.init_prims(evo._embed(P0))
        t = 0.0
        for k in range(nsteps):
            evo.drv.set_clock(t, dts[k], k)
            evo.drv.step(1)
            t += dts[k]

    def timed_forward():
        evo.drv.init_prims(evo._embed(P0))
        torch.cuda.synchronize()
        ev[0].record()
        t = 0.0
        for k in range(nsteps):
            evo.drv.set_clock(t, dts[k], k)
            evo.drv.step(1)
            t += dts[k]
        ev[1].record()
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1])

    plain_forward()
    fwd_ms = sorted(timed_forward() for _ in range(3))[1]
    # gradient: forward with checkpoints, backward sweep (coverage checked once at its end)
    Pg = P0.clone().requires_grad_(True)
    torch.cuda.synchronize()
    ev[0].record()
    _, J = evolve_objective(Pg, evo, nsteps, dts, every=every)
    ev[1].record()
    J.backward()
    ev[2].record()
    torch.cuda.synchronize()
    ckpt_fwd_ms = ev[0].elapsed_time(ev[1])
    back_ms = ev[1].elapsed_time(ev[2])
    # the adjoint steps' own record-load / transpose times: a second sweep whose
    # steps synchronise and time themselves
    rec = adj = 0.0
    orig = {name: getattr(evo.adj, name) for name in ("step", "step_recorded")}

    def timed(name):
        def call(*args):
            nonlocal rec, adj
            orig[name](*args)
            f, g_ = evo.adj.last_times()
            rec += f
            adj += g_
        return call
    for name in orig:
        setattr(evo.adj, name, timed(name))
    evo.deferred_checks = False
    P2 = P0.clone().requires_grad_(True)
    evolve_objective(P2, evo, nsteps, dts, every=every)[1].backward()
    evo.deferred_checks = True
    for name, fn in orig.items():
        setattr(evo.adj, name, fn)
    assert torch.equal(P2.grad, Pg.grad)
    g_delta = (Pg.grad[pr.UU] * u0).cpu().numpy()      # dJ/d delta
    Jv = float(J.detach())
    # central differences along random directions of delta (unit max norm)
    rng = np.random.default_rng(seed)
    fd, ad = [], []
    for _ in range(ndirs):
        v = rng.uniform(-1.0, 1.0, size=g_delta.shape)
        vt = torch.as_tensor(v, device="cuda")
        Pp = P0.clone()
        Pp[pr.UU] += h * vt * u0
        Pm = P0.clone()
        Pm[pr.UU] -= h * vt * u0
        fd.append((evo.run(Pp, nsteps, dts)[3] - evo.run(Pm, nsteps, dts)[3]) / (2 * h))
        ad.append(float((g_delta * v).sum()))
    fd, ad = np.array(fd), np.array(ad)
    err = np.abs(fd - ad) / np.maximum(np.abs(fd), np.abs(ad)).clip(1e-300)
    smooth = _torus_objective_smooth_check(cfg, face, u0, seed)
    fwd_step = fwd_ms / nsteps
    return {"workload": f"2-D FMKS torus {nx1}x{nx2} (a=0.9375, rin 6, rmax 12, gamma 13/9, beta 100, +-4% u jitter), "
                        f"FP64, {nsteps} RK2 steps, drift-frame floors, outflow/reflecting boundaries",
            "objective": f"time-integrated mass flux through the x1 face at r = {r_face:.4f} (index {face})",
            "J": Jv, "gradient": "dJ/d delta, u = u_0 (1 + delta) (the jitter field), all interior zones",
            "grad_abs_max": float(np.abs(g_delta).max()), "grad_l2": float(np.linalg.norm(g_delta)),
            "checkpoints": {"every": every, "held": nsteps // every + 1, "segment_states": every},
            "forward_ms_per_step": fwd_step,
            "gradient_ms": ckpt_fwd_ms + back_ms,
            "gradient_over_forward": (ckpt_fwd_ms + back_ms) / fwd_ms,
            "backward_ms_per_step": back_ms / nsteps,
            "backward_over_forward": back_ms / fwd_ms,
            "adjoint_step_ms": (rec + adj) / nsteps, "adjoint_step_record_load_ms": rec / nsteps,
            "adjoint_step_transpose_ms": adj / nsteps,
            "adjoint_step_over_forward_step": (rec + adj) / nsteps / fwd_step,
            "grad_check": {"directions": int(ndirs), "fd_rel_step": h, "seed": seed,
                           "max_rel_err": float(err.max()), "median_rel_err": float(np.median(err)),
                           "fd": [float(x) for x in fd], "adjoint": [float(x) for x in ad]},
            "grad_check_smooth": smooth}


def _torus_objective_smooth_check(cfg, face, u0, seed, nsteps=6, every=3, ndirs=4, h=1e-7):
    """The same objective and its gradient with respect to the initial state
    (dJ/d delta above is its u component times u_0) where central differences
    are well posed: the torus with its atmosphere at 1.5 times its floors (no
    zone floored in the steps), small smooth modes in every variable (the bare
    atmosphere is uniform in theta, where WENO5's smoothness indicators vanish
    and finite differences measure the weights of the perturbation), and the
    recovery converged to 1e-14; nsteps steps with checkpoints every `every`
    (segment re-runs); smooth directions in all variables, each scaled by the
    variable's local magnitude."""
    import numpy as np
    import torch
    from kharma_amd import problems as pr
    from kharma_amd.adjoint import Evolution, evolve_objective, smooth_torus_state

    cs = dict(cfg, err_tol=1e-14, iter_max=30)
    rng = np.random.default_rng(seed + 1)
    P, direction = smooth_torus_state(cs, rng)
    P0 = torch.as_tensor(P, device="cuda")
    evo = Evolution(cs, objective=dict(dir=1, var=pr.RHO, face=face, weight=1.0))
    dts = evo.run(P0, nsteps)[1]
    Pg = P0.clone().requires_grad_(True)
    _, J = evolve_objective(Pg, evo, nsteps, dts, every=every)
    J.backward()
    g = Pg.grad.cpu().numpy()
    fd, ad = [], []
    for _ in range(ndirs):
        v = direction()
        vt = torch.as_tensor(v, device="cuda")
        fd.append((evo.run(P0 + h * vt, nsteps, dts)[3] - evo.run(P0 - h * vt, nsteps, dts)[3]) / (2 * h))
        ad.append(float((g * v).sum()))
    fd, ad = np.array(fd), np.array(ad)
    err = np.abs(fd - ad) / np.maximum(np.abs(fd), np.abs(ad)).clip(1e-300)
    return {"state": "smooth modes in all variables, atmosphere at 1.5x its floors, recovery to 1e-14",
            "gradient": "dJ/dP_0, all variables", "steps": nsteps, "every": every,
            "directions": int(ndirs), "fd_rel_step": h, "max_rel_err": float(err.max()),
            "median_rel_err": float(np.median(err))}


def _calibration_copy(mb):
    """One device-to-device copy of `mb` MiB (reads and writes mb MiB of HBM)."""
    import torch
    n = mb * (1 << 20) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calibration_copy_mib": mb}))


if __name__ == "__main__":
    main()
