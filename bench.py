"""Benchmark of the virtual-CBCT chain on one MI355X.

Headline: forward projection rays x energy-bins per second on the full-size plain workload — a 25 cm water
cylinder on a 512 x 512 x 128 grid (0.5 mm in-plane, 1 mm slices, six uint16 elemental maps), 360 angles,
400 x 400 detector with a 115 mm half-fan offset, 91 energy bins (10-100 keV).  One step = one forward
projection of all 360 angles x 160 000 rays x 91 bins with the volume resident in HBM.

Also reported on the same line: FDK (filter + backprojection to 270 x 270 x 128) voxel updates per second,
adjoint gradients per second on the differentiable workload (256^2 x 64 FP32 maps, 90 angles, 256^2
detector), the roofline of the dominant kernel (algorithmic bytes from the per-ray segment counts), and the
reference's own CPU projector timed on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "forward rays*bins/s, FDK GUPS, adjoint grads/s vs plain GPU at <=1e-4 rel err"
HBM_PEAK_GBS = 8000.0
# Tile shapes fixed on the command line (--tile, --adjoint-tile, --batch-grid) instead of autotuned: profiled
# runs fix the shapes the autotuner picked in an unprofiled run, so that the profiler sees no candidate launches.
FIXED = {"tile": None, "adjoint_tile": None, "batch_grid": None}


def emit(line):
    """Prints the JSON line, with every tile-shape decision of the run (candidates and timings per kernel and
    geometry) under "autotune_log"."""
    from virtualcbct_amd.batch import autotune_log
    line["autotune_log"] = [l for l in autotune_log().strip().split("\n") if l]
    print(json.dumps(line), flush=True)


def projector(*args):
    """A Projector with the tile shapes given on the command line (else autotuned)."""
    from virtualcbct_amd.projector import Projector
    P = Projector(*args)
    if FIXED["tile"]:
        P.set_tile(*FIXED["tile"])
    if FIXED["adjoint_tile"]:
        P.set_adjoint_tile(*FIXED["adjoint_tile"])
    return P


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_c2(td):
    from virtualcbct_amd import io, phantoms
    w = phantoms.WORKLOADS["c2"]
    atten, pdf = phantoms.synthetic_spectrum(td, 100)
    vol = phantoms.water_cylinder(w["N"], w["Nz"], w["s"], w["radius"])
    ang, xs, ys = phantoms.flexmap_angles(w["n_ang"], w["xshift_proj"])
    cs, sn, xc, yc = io.trig_table(w["W"], w["H"], ang, xs, ys)
    fm = f"{td}/fdk.flexmap"
    phantoms.write_flexmap(fm, w["n_ang"], w["xshift_fdk"])
    a_fdk, xs_fdk, ys_fdk = io.read_fdk_angles(fm, w["n_ang"])
    return dict(w=w, atten=atten, pdf=pdf, vol=vol, ang=ang, cs=cs, sn=sn, xc=xc, yc=yc, a_fdk=a_fdk, xs_fdk=xs_fdk,
                ys_fdk=ys_fdk)


def read_profile_value(name, key):
    """A value recorded as `key: value` in a committed rocprofv3 summary under profiles/ (None if absent)."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith(key + ":"):
            return float(line.split(":")[1])
    return None


def read_traffic():
    """HBM bytes per launch of the forward kernel from the committed rocprofv3 PMC summary, if any (FETCH_SIZE and
    WRITE_SIZE of the kernel, each scaled by the same counter's reading for a plain copy of known size)."""
    return read_profile_value("forward_c2_traffic.txt", "traffic_bytes_per_launch")


def host_cpu():
    """The host's CPU as the CPU baselines report it: the logical CPU count (nproc), the CPUs this process may run
    on, OMP_NUM_THREADS, the lscpu model name, and the thread count used (the CPUs this process may run on,
    capped by OMP_NUM_THREADS where that is set, as on a GPU host shared between GPUs)."""
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    if model is None and os.path.exists("/proc/cpuinfo"):
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = min(affinity, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else affinity
    return {"nproc": os.cpu_count(), "affinity_cpus": affinity, "OMP_NUM_THREADS": omp, "model": model,
            "threads": threads}


def cpu_baseline(inp, seconds_target=15.0):
    """The reference's CPU projector (P/reprojection.cpp, built from its sources into oracle/_ref) on a bounded
    sample of (angle, bin) pairs of the same workload, OpenMP over pairs on every CPU this process may use (the
    function itself is single-threaded per call).  Falls back to the FP64 oracle when the reference build is
    absent."""
    from oracle import oracle as O
    w = inp["w"]
    cpu = host_cpu()
    nthreads = cpu["threads"]
    g = O.Geometry(N=w["N"], s=w["s"], Nz=w["Nz"], t=w["t"], H=w["H"], W=w["W"], cos_t=inp["cs"], sin_t=inp["sn"],
                   xc=inp["xc"], yc=inp["yc"])
    rng = np.random.default_rng(0)
    if O.ref_available():
        # calibrate with one pair per thread, then size the sample to ~seconds_target
        pairs = [(int(a), int(k)) for a, k in zip(rng.integers(0, w["n_ang"], nthreads), rng.integers(0, 91, nthreads))]
        sec, _ = O.ref_reprojection_timed(g, inp["vol"], inp["ang"], 1000.0, inp["atten"], inp["pdf"], pairs, nthreads)
        n = max(nthreads, int(nthreads * seconds_target / max(sec, 1e-3)) // nthreads * nthreads)
        n = min(n, 8 * nthreads)
        pairs = [(int(a), int(k)) for a, k in zip(rng.integers(0, w["n_ang"], n), rng.integers(0, 91, n))]
        sec, _ = O.ref_reprojection_timed(g, inp["vol"], inp["ang"], 1000.0, inp["atten"], inp["pdf"], pairs, nthreads)
        rate = n * w["W"] * w["H"] / sec
        return {"value": rate, "unit": "rays*bins/s", "cores": nthreads, "kind": "reference",
                "sample": f"{n} (angle, bin) pairs of the 360x91 workload, each a full 400x400 detector "
                          f"(P/reprojection.cpp CPU twin, one call per pair), {sec:.1f} s", "host": cpu}
    # oracle: one traversal per ray then 91 bins; time a slab of rays of one angle
    sub = O.Geometry(N=w["N"], s=w["s"], Nz=w["Nz"], t=w["t"], H=w["H"], W=w["W"], cos_t=inp["cs"][:1],
                     sin_t=inp["sn"][:1], xc=inp["xc"][:1], yc=inp["yc"][:1])
    t0 = time.perf_counter()
    O.forward_collapsed(sub, inp["vol"], inp["atten"], inp["pdf"], 1000.0)
    sec = time.perf_counter() - t0
    return {"value": w["W"] * w["H"] * 91 / sec, "unit": "rays*bins/s", "cores": nthreads, "kind": "port",
            "sample": f"1 angle x 160000 rays x 91 bins with the FP64 oracle, {sec:.1f} s", "host": cpu}


def bench_fdk(line, inp, dev, stream, out):
    import torch
    from virtualcbct_amd import fdk
    w = inp["w"]
    # FDK on the projections just computed, through a plan (tables built once, backprojector autotuned)
    proj = out
    filt = torch.empty_like(proj)
    rec = torch.zeros((w["Z"], w["R"], w["R"]), dtype=torch.float32, device=dev)
    plan = fdk.FdkPlan(w["H"], w["W"], w["n_ang"], w["R"], w["Z"], w["zstart"], w["zstep"], w["rscale"], 1000,
                       inp["a_fdk"], inp["xs_fdk"], inp["ys_fdk"])
    plan.filter(proj, filt)
    plan.backproject(filt, rec)
    torch.cuda.synchronize()
    nrep = 5
    e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    tf, tb = [], []
    for _ in range(nrep):
        rec.zero_()
        e[0].record(stream)
        plan.filter(proj, filt)
        e[1].record(stream)
        plan.backproject(filt, rec)
        e[2].record(stream)
        torch.cuda.synchronize()
        tf.append(e[0].elapsed_time(e[1]))
        tb.append(e[1].elapsed_time(e[2]))
    # half-fan weighting + padding of the same projections (host wall clock: table upload, kernel, output
    # allocation, synchronise)
    import time
    tw = []
    for _ in range(nrep):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fdk.offset_weight_device(proj, inp["xs_fdk"], inp["ys_fdk"])
        tw.append((time.perf_counter() - t0) * 1e3)
    # the same scan through a half-fan plan (weighting + padding to 630 columns inside the filter call)
    plan_hf = fdk.FdkPlan(w["H"], w["W"], w["n_ang"], w["R"], w["Z"], w["zstart"], w["zstep"], w["rscale"], 1000,
                          inp["a_fdk"], inp["xs_fdk"], inp["ys_fdk"], half_fan=True)
    filt_hf = torch.empty(plan_hf.shape_filtered, dtype=torch.float32, device=dev)
    plan_hf.filter(proj, filt_hf)
    torch.cuda.synchronize()
    thf, thb = [], []
    for _ in range(nrep):
        rec.zero_()
        e[0].record(stream)
        plan_hf.filter(proj, filt_hf)
        e[1].record(stream)
        plan_hf.backproject(filt_hf, rec)
        e[2].record(stream)
        torch.cuda.synchronize()
        thf.append(e[0].elapsed_time(e[1]))
        thb.append(e[1].elapsed_time(e[2]))
    plan_hf.close()
    del filt_hf
    # the same scan through a plan with the opt-in band-limited ramp (window alpha = 1, Ram-Lak), reported
    # separately: the reference's kernel (alpha = 0.5) is the parity path
    plan_rl = fdk.FdkPlan(w["H"], w["W"], w["n_ang"], w["R"], w["Z"], w["zstart"], w["zstep"], w["rscale"], 1000,
                          inp["a_fdk"], inp["xs_fdk"], inp["ys_fdk"], window_alpha=1.0)
    plan_rl.filter(proj, filt)
    torch.cuda.synchronize()
    trl, trb = [], []
    for _ in range(nrep):
        rec.zero_()
        e[0].record(stream)
        plan_rl.filter(proj, filt)
        e[1].record(stream)
        plan_rl.backproject(filt, rec)
        e[2].record(stream)
        torch.cuda.synchronize()
        trl.append(e[0].elapsed_time(e[1]))
        trb.append(e[1].elapsed_time(e[2]))
    plan_rl.close()
    # the reference's CPU backprojection (F/reconstructImage.cpp, built from its sources; OpenMP over slices)
    # on a bounded sample of the same reconstruction: whole slices from the centre up, all 360 views, a
    # multiple of the thread count (one slice per thread at a time)
    cpu_ref = None
    from oracle import oracle as O
    if O.ref_available():
        fh = filt.cpu().numpy()
        nthreads = host_cpu()["threads"]
        _, sec1 = O.ref_reconstruct_image(w["R"], 1, 0, w["zstep"], w["rscale"], fh, 1000, inp["a_fdk"],
                                          inp["xs_fdk"])
        nz = int(min(w["Z"], nthreads * max(1, round(6.0 / max(sec1, 1e-3)))))
        _, sec = O.ref_reconstruct_image(w["R"], nz, 0, w["zstep"], w["rscale"], fh, 1000, inp["a_fdk"],
                                         inp["xs_fdk"])
        upd = w["R"] * w["R"] * nz * w["n_ang"]
        cpu_ref = {"gups": upd / sec / 1e9, "cores": nthreads, "kind": "reference",
                   "sample": f"{nz} slices of the 270^2 reconstruction from all 360 views, {sec:.1f} s"}
        log(f"cpu reference backprojection: {cpu_ref['gups']:.3f} GUPS ({cpu_ref['sample']})")
        del fh
    # the reference's CPU filter (F/filterSinogram.cpp, OpenMP over projections) on a bounded sample of the
    # same projections: a multiple of the thread count, sized to a few seconds
    cpu_filt = None
    if O.ref_available():
        nthreads = host_cpu()["threads"]
        ph = proj.cpu().numpy()
        _, s1 = O.ref_filter_sinogram(ph[:nthreads], inp["xs_fdk"][:nthreads], inp["ys_fdk"][:nthreads])
        na = int(min(w["n_ang"], nthreads * max(1, round(4.0 / max(s1, 1e-3)))))
        _, secf = O.ref_filter_sinogram(ph[:na], inp["xs_fdk"][:na], inp["ys_fdk"][:na])
        cpu_filt = {"rows_per_s": na * w["H"] / secf, "cores": nthreads, "kind": "reference",
                    "sample": f"{na} of the {w['n_ang']} projections ({w['H']} rows of {w['W']}), {secf:.1f} s"}
        log(f"cpu reference filter: {cpu_filt['rows_per_s']:.3e} rows/s ({cpu_filt['sample']})")
        del ph
    from virtualcbct_amd.batch import autotune_log
    updates = w["R"] * w["R"] * w["Z"] * w["n_ang"]
    line["fdk"] = {"gups": updates / (np.median(tb) * 1e-3) / 1e9, "bp_ms": float(np.median(tb)),
                   "filter_ms": float(np.median(tf)), "offset_weight_ms": float(np.median(tw)),
                   "half_fan_plan": {"filter_ms": float(np.median(thf)), "bp_ms": float(np.median(thb)),
                                     "width": plan_hf.width},
                   "ram_lak_plan": {"window_alpha": 1.0, "filter_ms": float(np.median(trl)),
                                    "bp_ms": float(np.median(trb))},
                   "cpu_reference_backprojection": cpu_ref, "cpu_reference_filter": cpu_filt,
                   "filter_rows_per_s": w["n_ang"] * w["H"] / (np.median(tf) * 1e-3),
                   "updates": updates, "shape": [w["Z"], w["R"], w["R"]],
                   # the backprojector is bound by VALU issue (its taps hit in L2), so its bound is stated in
                   # instructions: VALU wave-instructions per launch from the committed counter profile, as
                   # lane-instructions per voxel update, and the VALU busy fraction they imply at this run's time
                   "bp_bound": fdk_bound(updates, float(np.median(tb))),
                   "autotune": [l for l in autotune_log().strip().split("\n") if l.startswith("bp ")]}
    log(f"fdk: filter {np.median(tf):.3f} ms, backprojection {np.median(tb):.3f} ms, "
        f"{line['fdk']['gups']:.1f} GUPS, offset weighting {np.median(tw):.3f} ms; half-fan plan: filter "
        f"{np.median(thf):.3f} ms, backprojection {np.median(thb):.3f} ms")
    plan.close()
    del filt, rec


def fdk_bound(updates, bp_ms):
    """The backprojector's bound from its counters (profiles/fdk_c2_counters.md, the plain plan's launch):
    VALU-issue time = wave-instructions x 4 cycles / (SIMDs x clock)."""
    valu = read_profile_value("fdk_c2_counters.md", "bp_valu_wave_instructions_per_launch")
    simds, clock_hz = 256 * 4, 2.4e9  # MI355X: 256 CUs x 4 SIMDs, 2.4 GHz peak engine clock
    out = {"bound": "valu", "source": "profiles/fdk_c2_counters.md (rocprofv3 SQ_INSTS_VALU)",
           "hbm_equivalent_frac": 16.0 * updates / (bp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "hbm_equivalent_note": "16 B per update (4 FP32 taps) over the 8 TB/s peak; above 1 because the taps "
                                  "hit in L2, so HBM is not the bound"}
    if valu:
        valu_ms = valu * 4 / simds / clock_hz * 1e3
        out.update({"valu_wave_instructions_per_launch": valu,
                    "lane_instructions_per_update": valu * 64 / updates,
                    "valu_issue_ms": valu_ms, "valu_busy_frac": valu_ms / bp_ms})
    return out


def bench_detector(line, dev, stream, e, out):
    import torch
    from virtualcbct_amd import detector
    # detector model on the headline projections (360 x 400 x 400): Poisson counts of I0 = 1e5 quanta per
    # open-field pixel plus a flat 5 % scatter field, normalised for the filter
    sc = torch.full_like(out, 0.05)
    noisy = torch.empty_like(out)
    detector.detector_counts(out, sc, I0=1.0e5, seed=1, out=noisy)
    torch.cuda.synchronize()
    t = []
    for r in range(5):
        e[0].record(stream)
        detector.detector_counts(out, sc, I0=1.0e5, seed=2 + r, out=noisy)
        e[1].record(stream)
        torch.cuda.synchronize()
        t.append(e[0].elapsed_time(e[1]))
    ms = float(np.median(t))
    line["detector"] = {"ms": ms, "pixels": out.numel(), "pixels_per_s": out.numel() / (ms * 1e-3), "I0": 1.0e5,
                        "scatter": "flat 0.05"}
    log(f"detector model: {ms:.3f} ms for {out.numel()} pixels")
    del sc, noisy


def bench_adjoint(line, inp, dev, stream, e):
    import torch
    # adjoint on the differentiable workload (256^2 x 64 FP32 maps, 90 angles, 256^2 detector)
    from virtualcbct_amd import phantoms, io
    N4, Z4, n4, D4 = 256, 64, 90, 256
    ang4 = np.deg2rad(-180.0 + 4.0 * np.arange(n4) + 1e-4)
    c4, s4, x4, y4 = io.trig_table(D4, D4, ang4, np.zeros(n4), np.zeros(n4))
    P4 = projector(N4, 0.1, Z4, 0.1, D4, D4, n4, c4, s4, x4, y4)
    maps = torch.from_numpy(phantoms.water_cylinder_f32(N4, Z4, 0.1, 10.0)).to(dev).contiguous()
    P4.set_spectrum(inp["atten"], inp["pdf"])
    # target: log projection of the phantom with densities x (1 + 0.05 N(0, 1)) (seed 4000)
    rng = np.random.default_rng(4000)
    P4.set_volume(maps * torch.from_numpy((1.0 + 0.05 * rng.standard_normal(maps.shape)).astype(np.float32)).to(dev))
    tgt = P4.logproj(1.0).clone()
    P4.set_volume(maps)
    P4.adjoint(tgt, 1.0)
    torch.cuda.synchronize()
    tl = []
    for _ in range(5):
        e[0].record(stream)
        P4.logproj(1.0)
        e[1].record(stream)
        torch.cuda.synchronize()
        tl.append(e[0].elapsed_time(e[1]))
    ta = []
    for _ in range(5):
        e[0].record(stream)
        grad, loss = P4.adjoint(tgt, 1.0)
        e[1].record(stream)
        torch.cuda.synchronize()
        ta.append(e[0].elapsed_time(e[1]))
    segs4 = float(P4.segment_counts().double().sum().item())
    rays4 = n4 * D4 * D4
    adj_bytes = 48.0 * segs4 + 8.0 * rays4
    line["adjoint"] = {"grads_per_s": 1e3 / float(np.median(ta)), "ms": float(np.median(ta)),
                       "workload": "256^2 x 64 FP32 maps (1 mm), 90 angles, 256^2 detector, 91 bins; target from "
                                   "densities x (1 + 0.05 N(0,1))",
                       "roofline_frac": adj_bytes / (np.median(ta) * 1e-3) / 1e9 / HBM_PEAK_GBS,
                       "logproj_ms": float(np.median(tl)),
                       "adjoint_channels": P4.adjoint_channels,
                       "adjoint_over_logproj": float(np.median(ta) / np.median(tl)),
                       "segments": segs4}
    log(f"adjoint: {np.median(ta):.3f} ms/gradient")
    P4.close()


def bench_batched(line, inp, dev, stream, e, NE):
    import torch
    # batched: eight phantoms of uneven length (64 .. 256 slices), 360 angles, 91 bins — one batched launch
    # (grid form chosen by the autotuner) and its hipGraph replay against the plain per-phantom launches
    from virtualcbct_amd import phantoms, io
    from virtualcbct_amd.batch import Batch, autotune_log
    zs = (64, 96, 128, 160, 192, 224, 240, 256)
    ang3, xs3, ys3 = phantoms.flexmap_angles(360)
    c3, s3, x3, y3 = io.trig_table(400, 400, ang3, xs3, ys3)
    projs, outs3 = [], []
    for i, Z in enumerate(zs):
        volt = phantoms.ellipsoid_phantom_torch(512, Z, 0.05, 0.1, 3000 + i, device=dev)
        Pi = projector(512, 0.05, Z, 0.1, 400, 400, 360, c3, s3, x3, y3)
        Pi.set_volume(volt)
        Pi.set_spectrum(inp["atten"], inp["pdf"])
        del volt
        projs.append(Pi)
        outs3.append(torch.empty(Pi.shape_proj, dtype=torch.float32, device=dev))
    torch.cuda.synchronize()
    for Pi, o in zip(projs, outs3):
        Pi.forward(1000.0, out=o)
    torch.cuda.synchronize()
    plain_ms = []
    for _ in range(2):
        e[0].record(stream)
        for Pi, o in zip(projs, outs3):
            Pi.forward(1000.0, out=o)
        e[1].record(stream)
        torch.cuda.synchronize()
        plain_ms.append(e[0].elapsed_time(e[1]))
    ref0 = outs3[3].clone()
    B = Batch()
    if FIXED["batch_grid"] is not None:
        B.set_grid(FIXED["batch_grid"])
    for Pi, o in zip(projs, outs3):
        B.add_forward(Pi, o, 1000.0)
    B.run()  # finalises and autotunes the grid
    torch.cuda.synchronize()
    batch_ms = []
    for _ in range(2):
        e[0].record(stream)
        B.run()
        e[1].record(stream)
        torch.cuda.synchronize()
        batch_ms.append(e[0].elapsed_time(e[1]))
    same = bool(torch.equal(outs3[3], ref0))
    B.capture()
    graph_ms = []
    for _ in range(2):
        e[0].record(stream)
        B.replay()
        e[1].record(stream)
        torch.cuda.synchronize()
        graph_ms.append(e[0].elapsed_time(e[1]))
    rays3 = 360 * 400 * 400 * len(zs)
    line["batched"] = {"workload": "8 ellipsoid phantoms 512^2 x {64..256} slices, 360 angles, 400^2, 91 bins",
                       "plain_ms": float(min(plain_ms)), "batched_ms": float(min(batch_ms)),
                       "graph_ms": float(min(graph_ms)),
                       "plain_rays_bins_per_s": rays3 * NE / (min(plain_ms) * 1e-3),
                       "batched_rays_bins_per_s": rays3 * NE / (min(batch_ms) * 1e-3),
                       "speedup": float(min(plain_ms) / min(batch_ms)), "identical_to_plain": same,
                       "batch_grid": ("flat" if B.blocks == 0 else "sequence of tuned per-phantom launches" if B.blocks == -2
                                     else f"persistent {B.blocks} blocks"), "autotune": autotune_log().strip().split("\n")}
    log(f"batched: plain {min(plain_ms):.2f} ms, batched {min(batch_ms):.2f} ms, graph {min(graph_ms):.2f} ms")
    B.close()
    for Pi in projs:
        Pi.close()
    del outs3


def bench_decompose(line, td, dev, stream, e, iters=4):
    """Dual-energy decomposition: four problems (512^2 x {64, 96, 112, 128} FP32 maps, 360 angles, 400^2) under
    80 and 120 kV spectra, optimised together by Adam.  The fused, graph-replayed path (per problem one
    traversal and one scatter for both spectra, the Adam update on the device) runs `iters` iterations after one
    eager warm-up and one capture; the per-problem autograd path (one log projection and one backward pass per
    problem and spectrum, torch Adam) is timed on two iterations beside it."""
    import time
    import torch
    from virtualcbct_amd import io, phantoms
    from virtualcbct_amd.decompose import FusedDecomposition, decompose_autograd
    spectra = {kv: phantoms.synthetic_spectrum(td, kv) for kv in (80, 120)}
    ang, xs, ys = phantoms.flexmap_angles(360)
    c, s, x, y = io.trig_table(400, 400, ang, xs, ys)
    problems = []
    for i, Z in enumerate((64, 96, 112, 128)):
        vol = phantoms.ellipsoid_phantom_torch(512, Z, 0.05, 0.1, 5000 + i, device=dev)
        truth = (vol.permute(1, 0, 2, 3).float() / 1000.0).contiguous()  # [mat][z][y][x] g/cc
        del vol
        projs, targets = [], []
        for kv in (80, 120):
            P = projector(512, 0.05, Z, 0.1, 400, 400, 360, c, s, x, y)
            P.set_spectrum(*spectra[kv])
            P.set_volume(truth)
            targets.append(P.logproj(1.0).clone())
            projs.append(P)
        problems.append({"projectors": projs, "targets": targets, "shape": tuple(truth.shape), "xfactor": 1.0})
        del truth
    init = [torch.full(p["shape"], 0.2, device=dev) for p in problems]
    # per-problem autograd path
    decompose_autograd(problems, iters=1, lr=0.02, init=init)  # warm-up (allocations)
    torch.cuda.synchronize()
    e[0].record(stream)
    decompose_autograd(problems, iters=2, lr=0.02, init=init)
    e[1].record(stream)
    torch.cuda.synchronize()
    autograd_ms = e[0].elapsed_time(e[1]) / 2
    # fused, graph-replayed path: the eager first iteration and the capture outside the timing
    st = FusedDecomposition(problems, lr=0.02, init=init, graph=True)
    st.step()
    st.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e[0].record(stream)
    for _ in range(iters):
        st.step()
    e[1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e[0].elapsed_time(e[1]) / iters
    hist = st.losses()
    line["decompose"] = {"workload": "4 problems 512^2 x {64,96,112,128} FP32 maps, 360 angles, 400^2, 80+120 kV "
                                     "(71+111 bins), Adam lr 0.02",
                         "iteration_ms": ms, "iterations_timed": iters, "iterations_total": st.iteration,
                         "wall_s_timed": wall, "problem_gradients_per_s": 4 * 1e3 / ms,
                         "spectrum_gradients_per_s": 8 * 1e3 / ms,
                         "per_problem_autograd_iteration_ms": autograd_ms,
                         "speedup_vs_per_problem": autograd_ms / ms,
                         "channels": st.channels, "graph_replayed": st.graph is not None,
                         "loss_first": hist[0], "loss_last": hist[-1],
                         "loss_every_10": [h for k, h in enumerate(hist) if k % 10 == 0]}
    log(f"decompose: fused + graph {ms:.1f} ms per Adam iteration over {iters} iterations ({wall:.1f} s wall), "
        f"per-problem autograd path {autograd_ms:.1f} ms; loss {hist[0]:.4e} -> {hist[-1]:.4e}")
    for p in problems:
        for P in p["projectors"]:
            P.close()


def bench_phantom(line, td, dev, stream, e):
    """Organ-label phantom -> elemental maps: a 543 x 271 x 100 label volume (the ICRP adult-male slice size
    and the reference converter's slice count) to (a) the reference's per-slice layout [100][6][271][543] u16 and
    (b) a 512 x 512 x 100 projector volume (centred crop / pad fused in).  CPU beside it: the oracle's
    restatement of the reference's voxel loop (every organ row tested per voxel) on the same labels."""
    import torch
    from oracle import oracle as O
    from virtualcbct_amd import io, material_maps, phantoms
    media, organs = phantoms.write_material_tables(os.path.join(td, "tables"), seed=3)
    lab_h = phantoms.synthetic_labels(543, 271, 100, seed=1)
    lab = torch.from_numpy(lab_h).to(dev)
    lut, _ = material_maps.material_table(media, organs)
    nvox = lab_h.size
    maps = material_maps.labels_to_maps(lab, lut)
    ang, xs, ys = phantoms.flexmap_angles(1)
    c, s, x, y = io.trig_table(16, 16, ang, xs, ys)
    P = projector(512, 0.05, 100, 0.2, 16, 16, 1, c, s, x, y)
    P.set_volume_labels(lab, lut)
    torch.cuda.synchronize()
    nrep = 20
    t = {}
    for name, fn in (("maps", lambda: material_maps.labels_to_maps(lab, lut, out=maps)),
                     ("projector_volume", lambda: P.set_volume_labels(lab, lut))):
        e[0].record(stream)
        for _ in range(nrep):
            fn()
        e[1].record(stream)
        torch.cuda.synchronize()
        t[name] = e[0].elapsed_time(e[1]) / nrep
    c0, r0 = material_maps.centred_placement(543, 271, 512)
    read_b = 100 * min(271, 512) * min(543, 512)  # labels inside the placed window
    pv = 512 * 512 * 100
    res = {"workload": "labels 543x271x100 u8 -> maps [100][6][271][543] u16 | projector volume 512x512x100 "
                       f"(16-byte voxels, placement col0={c0} row0={r0})",
           "maps_ms": t["maps"], "maps_gb_s": 13.0 * nvox / (t["maps"] * 1e-3) / 1e9,
           "projector_volume_ms": t["projector_volume"],
           "projector_volume_gb_s": (read_b + 16.0 * pv) / (t["projector_volume"] * 1e-3) / 1e9,
           "maps_voxels_per_s": nvox / (t["maps"] * 1e-3)}
    res["maps_roofline_frac"] = res["maps_gb_s"] / HBM_PEAK_GBS
    res["projector_volume_roofline_frac"] = res["projector_volume_gb_s"] / HBM_PEAK_GBS
    w = O.material_weights(media, True, 1)
    org = O.organ_table(organs)
    t0 = time.perf_counter()
    O.labels_to_maps(lab_h, w, org)
    sec = time.perf_counter() - t0
    res["cpu_port"] = {"voxels_per_s": nvox / sec, "cores": host_cpu()["threads"],
                       "sample": f"the whole 543x271x100 volume, {len(org[0])} organ rows, {sec:.2f} s"}
    line["phantom"] = res
    log(f"phantom: maps {t['maps']:.3f} ms ({res['maps_gb_s']:.0f} GB/s), projector volume "
        f"{t['projector_volume']:.3f} ms ({res['projector_volume_gb_s']:.0f} GB/s), CPU {sec:.2f} s")
    P.close()


def bench_iterative(line, dev, stream, e):
    """Iterative-reconstruction pieces on a 256 x 256 x 64 image (1 mm voxels), 360 angles, 400 x 400 detector:
    lengthproj (voxel-driven chord backprojection, FP64 geometry per voxel and angle) and one IR_ImageCBCT
    update.  CPU beside it: the reference's own lengthproj (compiled from its sources, OpenMP over slices) on
    16 slices x 36 angles of the same geometry."""
    import torch
    from oracle import oracle as O
    from virtualcbct_amd import io, iterative, phantoms
    N, Nz, H, W, nA = 256, 64, 400, 400, 360
    ang, xs, ys = phantoms.flexmap_angles(nA)
    c, s_, x, y = io.trig_table(W, H, ang, xs, ys)
    P = projector(N, 0.1, Nz, 0.1, H, W, nA, c, s_, x, y)
    rng = np.random.default_rng(0)
    proj_h = rng.integers(20000, 60000, (nA, H, W)).astype(np.uint16)
    proj = torch.from_numpy(proj_h.view(np.int16)).to(dev)
    npf = iterative.lengthproj(P, proj)
    before = torch.from_numpy(rng.integers(0, 3000, (Nz, N, N)).astype(np.int16)).to(dev)
    npf_re = npf * 1.01
    out = iterative.IR_ImageCBCT(before, npf, npf_re, 0.1, 0.3, 0.05)
    torch.cuda.synchronize()
    reps = 3
    e[0].record(stream)
    for _ in range(reps):
        iterative.lengthproj(P, proj, out=npf)
    e[1].record(stream)
    for _ in range(reps):
        iterative.IR_ImageCBCT(before, npf, npf_re, 0.1, 0.3, 0.05, out=out)
    e[2].record(stream)
    torch.cuda.synchronize()
    lp_ms = e[0].elapsed_time(e[1]) / reps
    ir_ms = e[1].elapsed_time(e[2]) / reps
    pairs = N * N * Nz * nA
    res = {"workload": f"image {N}x{N}x{Nz} (1 mm), {nA} angles, {W}x{H} detector",
           "lengthproj_ms": lp_ms, "lengthproj_voxel_angles_per_s": pairs / (lp_ms * 1e-3),
           "ir_update_ms": ir_ms, "ir_update_voxels_per_s": N * N * Nz / (ir_ms * 1e-3),
           "ir_update_gb_s": (N * N * Nz * (2 + 4 + 4 + 4)) / (ir_ms * 1e-3) / 1e9}
    if O.ref_available():
        nz_c, na_c = 16, 36
        sel = np.arange(0, nA, nA // na_c)[:na_c]
        g = O.Geometry(N=N, s=0.1, Nz=nz_c, t=0.1, H=H, W=W, cos_t=c[sel], sin_t=s_[sel], xc=x[sel], yc=y[sel])
        t0 = time.perf_counter()
        O.ref_lengthproj(g, proj_h[sel], ang[sel])
        sec = time.perf_counter() - t0
        res["cpu_reference"] = {"voxel_angles_per_s": N * N * nz_c * na_c / sec,
                                "cores": min(nz_c, host_cpu()["threads"]),
                                "sample": f"P/lengthproj.cpp on {N}x{N}x{nz_c} x {na_c} angles, {sec:.2f} s"}
    # the MLEM + TV loop on the forward and adjoint kernels (A x: log projection of the image's water maps with
    # one energy bin; A^T: the adjoint scatter; the reference's update), 5 iterations after one
    from virtualcbct_amd.iterative import MlemTV
    truth = torch.from_numpy(phantoms.water_cylinder_f32(N, Nz, 0.1, 10.0)[3] / 0.888 * 1000).to(torch.int16).to(dev)
    mu = np.full(6, 0.2)
    solver = MlemTV(P, mu, torch.zeros(P.shape_proj, dtype=torch.float32, device=dev))
    b = solver.A(truth).clone()
    solver = MlemTV(P, mu, b, weight_tv=1e-5)
    x = torch.full((Nz, N, N), 800, dtype=torch.int16, device=dev)
    x = solver.step(x).to(torch.int16)
    torch.cuda.synchronize()
    e[0].record(stream)
    for _ in range(5):
        x = solver.step(x).to(torch.int16)
    e[1].record(stream)
    torch.cuda.synchronize()
    mlem_ms = e[0].elapsed_time(e[1]) / 5
    res["mlem_tv"] = {"workload": f"image {N}x{N}x{Nz} (1 mm), {nA} angles, {W}x{H} detector; A x: one-bin log "
                                  "projection of water maps, A^T: adjoint scatter, update: IR_ImageCBCT",
                      "iteration_ms": mlem_ms, "channels": P.adjoint_channels,
                      "misfit": float((solver.A(x) - b).norm() / b.norm())}
    line["iterative"] = res
    log(f"iterative: lengthproj {lp_ms:.2f} ms ({res['lengthproj_voxel_angles_per_s']:.3e} voxel-angles/s), "
        f"IR update {ir_ms:.3f} ms, MLEM-TV iteration on the kernels {mlem_ms:.2f} ms")
    P.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="headline forward only (used under the profiler)")
    ap.add_argument("--only", choices=["adjoint", "batched", "decompose", "fdk", "phantom", "iterative",
                                                 "detector"], default=None,
                    help="run one secondary section alone and print its JSON (used under the profiler)")
    ap.add_argument("--decompose-iters", type=int, default=200,
                    help="Adam iterations of `--only decompose` (the default bench line times 4)")
    ap.add_argument("--tile", default=None, help="fix the forward tile shape, row_groups,cols,order,split")
    ap.add_argument("--adjoint-tile", default=None, help="fix the adjoint kernels' tile shape (4 waves)")
    ap.add_argument("--batch-grid", type=int, default=None,
                    help="fix the batched forward's grid (0 flat, n persistent blocks, -2 sequence)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device copy (scale for the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    for k in ("tile", "adjoint_tile"):
        v = getattr(args, k)
        if v:
            FIXED[k] = tuple(int(t) for t in v.split(","))
    FIXED["batch_grid"] = args.batch_grid
    if args.copy_calibration:
        import torch
        n = 1 << 28  # 2^28 floats = 1 GiB
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)

    import torch
    from virtualcbct_amd import fdk

    dev = torch.device("cuda:0")
    td = tempfile.mkdtemp(prefix="vcbct_bench_")
    log("building synthetic c2 inputs ...")
    inp = build_c2(td)
    w = inp["w"]
    P = projector(w["N"], w["s"], w["Nz"], w["t"], w["H"], w["W"], w["n_ang"], inp["cs"], inp["sn"], inp["xc"],
                  inp["yc"])
    P.set_volume(inp["vol"])
    P.set_spectrum(inp["atten"], inp["pdf"])
    out = torch.empty(P.shape_proj, dtype=torch.float32, device=dev)
    NE = len(inp["pdf"])
    stream = torch.cuda.current_stream()

    if args.only in ("fdk", "detector"):  # projections of the headline workload, then the stage after them
        out = torch.empty(P.shape_proj, dtype=torch.float32, device=dev)
        P.forward(1000.0, out=out)
        line = {"section": args.only}
        if args.only == "fdk":
            bench_fdk(line, inp, dev, stream, out)
        else:
            bench_detector(line, dev, stream, [torch.cuda.Event(enable_timing=True) for _ in range(2)], out)
        emit(line)
        return
    if args.only:
        line = {"section": args.only}
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        if args.only == "adjoint":
            bench_adjoint(line, inp, dev, stream, e)
        elif args.only == "decompose":
            bench_decompose(line, td, dev, stream, e, iters=args.decompose_iters)
        elif args.only == "phantom":
            bench_phantom(line, td, dev, stream, e)
        elif args.only == "iterative":
            bench_iterative(line, dev, stream, e)
        else:
            bench_batched(line, inp, dev, stream, e, NE)
        emit(line)
        return

    log(f"forward warmup x{args.warmup} ...")
    for _ in range(args.warmup):
        P.forward(1000.0, out=out)
    torch.cuda.synchronize()
    # timed region: K forward projections, HIP events around each launch on the launch stream
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        evs[i][0].record(stream)
        P.forward(1000.0, out=out)
        evs[i][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = [a.elapsed_time(b) for a, b in evs]
    ms_step = wall * 1e3 / args.steps
    rays = w["n_ang"] * w["H"] * w["W"]
    value = rays * NE / (wall / args.steps)
    log(f"forward: {ms_step:.3f} ms/step (kernel avg {np.mean(kms):.3f} ms), {value:.4e} rays*bins/s")

    # roofline of the forward kernel: 12 B of voxel data per charged segment + 4 B output per ray
    segs = float(P.segment_counts().double().sum().item())
    torch.cuda.synchronize()
    alg_bytes = 12.0 * segs + 4.0 * rays
    k_avg_s = float(np.mean(kms)) * 1e-3
    # what the kernel gathers per segment: the u16 record (16 B) or 2 / 4 FP32 spectral channels
    nfc = P.forward_channels
    rec, rec_bytes = {0: ("uint4", 16), 2: ("float2", 8), 1: ("float2", 8)}.get(nfc, ("Ch4", 16))
    gathered = rec_bytes * segs + 4.0 * rays
    achieved = alg_bytes / k_avg_s / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(),
                "traffic_source": "profiles/forward_c2_traffic.txt: rocprofv3 FETCH_SIZE and WRITE_SIZE of this "
                                  "kernel, scaled by their readings for a plain 1 GiB copy (not measured in this run)",
                "kernel": f"k_forward<0, {rec}, {P.tile[0] * P.tile[1]}>", "segments_per_launch": segs,
                "algorithmic_bytes_per_launch": alg_bytes,
                "algorithmic_bytes_note": "12 B per charged segment (six u16 maps) + 4 B per ray, the traversal's "
                                          "definition; the kernel gathers the volume's spectral channels instead",
                "gathered_bytes_per_segment": rec_bytes,
                "gathered_bytes_per_launch": gathered,
                "gathered_gbs": gathered / k_avg_s / 1e9}
    from virtualcbct_amd.batch import autotune_log
    tile = {"row_groups": P.tile[0], "cols": P.tile[1], "order": P.tile[2], "split": P.tile[3],
            "autotune": [l for l in autotune_log().strip().split("\n") if l.startswith("fwd ")]}
    log(f"forward tile (autotuned): {P.tile}")
    line = {"metric": METRIC, "value": value, "unit": "rays*bins/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "precision": "FP32 line integrals (u16 x FP32 segment length), FP32 spectral terms and sum; FP64 ray "
                         "set-up, plane positions, crossings and segment midpoints (they choose voxels); FP64 "
                         "terms for rays whose FP32 exponentials all underflow",
            "config": {"workload": "25 cm water cylinder, 512x512x128 voxels (0.5 mm x 0.5 mm x 1 mm), 6 uint16 "
                                   "elemental maps, 360 angles, 400x400 detector, 115 mm half-fan offset, "
                                   "91 energy bins 10-100 keV, FP32 projections"},
            "roofline": roofline, "tile": tile,
            # context only (vs_baseline stays null: the reference publishes no benchmark)
            "reference_gpu_comment": {
                "rays_bins_per_s": 360 * 400 * 400 * 101 / 289.0,
                "source": "P/reprojectionOnGPU.cu:975-983, a developer comment beside the shipped launch shape: "
                          "4 min 49 s on an unnamed GeForce, read as 360 angles x 400^2 rays x 101 bins (the "
                          "comment states neither); an inference, not a published number"}}

    if not args.no_extras:
        bench_fdk(line, inp, dev, stream, out)
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        bench_detector(line, dev, stream, e, out)
        bench_adjoint(line, inp, dev, stream, e)

    if not args.no_extras:
        bench_batched(line, inp, dev, stream, e, NE)
        bench_decompose(line, td, dev, stream, e)
        bench_phantom(line, td, dev, stream, e)
        bench_iterative(line, dev, stream, e)

    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(inp)
        log(f"cpu baseline: {line['cpu_baseline']['value']:.4e} rays*bins/s on {line['cpu_baseline']['cores']} threads")
    P.close()
    emit(line)


if __name__ == "__main__":
    main()
