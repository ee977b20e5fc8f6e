"""Throughput of the polyenergetic cone-beam forward projection on one MI355X.

One step = one full scan of the 3-material 512^3 configuration: 360 views of
a 1024x768 detector, K = 1082 samples per ray (306.33 G samples), every
material marched once per ray and the 101-bin spectrum (grid + CsI response)
applied per ray; outputs I and IAbsorb stay in HBM.  Inputs are synthetic
(seeded phantom, Kramers spectrum, model attenuation tables) and resident in
HBM before timing starts.

The headline is the mixed-precision projector (volumes stored in FP16,
interpolation, line integrals and exp in FP32, energy sums in compensated
FP32); the same scan through the FP32 projector with the same launch options
and through the plain FP32 projector (corners gathered from HBM/L2, no LDS
staging) is timed beside it.  --dtype f32 / bf16 time the other storage types.

Outside the timed region the line also carries, unless --no-extras:
  - the error against the FP64 oracle of the FP16, BF16 and FP32 projectors
    over 8 views spread over the scan (max and mean);
  - the 128^3 single-material scan: the plain and staged FP32 GPU projectors
    and the oracle's FP32 mirror on this host's cores;
  - the differentiable 3 x 256^3 projector: backward / forward time, peak
    device memory;
  - the batched workload of 8 uneven scans: aggregate Gsamples/s of the
    persistent graph launch against the same scans run one after the other
    through the plain FP32 projector.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gsamples/s & views/s, 3-mat 512³ poly+grid; mixed vs FP32 at ≤1e-3 rel err"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(setup, vols, kw, target_s: float):
    """The oracle's FP32 mirror (march + spectrum), OpenMP on this host, on a
    bounded number of views of the same workload."""
    import numpy as np
    from oracle import oracle as O
    g = O.geom_from_setup(setup)
    p = O.paras(g)
    threads = O.num_threads()
    det = np.stack([O.det_response(1.0, float(m), kw["scint_thickness"]) for m in kw["scint_mu"]])

    def run(v0, nv):
        t = time.perf_counter()
        L = O.march_f32(g, p, vols, v0, nv, threads=threads)
        O.spectrum_f32(L, kw["atten"], kw["spectrum"], kw["i0_val"], kw["grid"], det, threads=threads)
        return time.perf_counter() - t

    t1 = run(0, 1)
    n = int(max(1, min(16, round(target_s / max(t1, 1e-3)) - 1)))
    t2 = run(1, n) if n > 0 else 0.0
    views = 1 + n
    secs = t1 + t2
    spv = setup.det[0] * setup.det[1] * p.n_steps
    return {"value": round(views * spv / secs / 1e9, 4), "unit": "Gsamples/s", "cores": threads,
            "kind": "port",
            "sample": f"views 0-{views - 1} of {setup.views} ({views * spv / 1e9:.2f} G samples), "
                      f"march + spectrum, {secs:.1f} s on {threads} threads"}


def det_table(kw):
    import numpy as np
    from oracle import oracle as O
    return np.stack([O.det_response(kw.get("det_factor", 1.0), float(m), kw["scint_thickness"])
                     for m in kw["scint_mu"]])


ACC_VIEWS = (0, 45, 90, 135, 180, 225, 270, 315)


def accuracy(projs, setup, vols, kw, views=ACC_VIEWS):
    """Relative error against the FP64 evaluation of the method (oracle,
    CPU) over `views` of the scan, for each projector of `projs` (name ->
    Projector): |d| / max(|ref|, f max|ref|) with f = 1e-3 for I and 1e-2 for
    p = ln(I0/I), max and mean over every ray of those views."""
    import numpy as np
    import torch
    from oracle import oracle as O
    g = O.geom_from_setup(setup)
    p = O.paras(g)
    det = det_table(kw)
    ref_I, ref_p = [], []
    for v in views:
        L64 = O.march_f64(g, p, vols, v, 1, threads=O.num_threads())
        I64, _, I064 = O.spectrum_f64(L64, kw["atten"], kw["spectrum"], kw["i0_val"], kw["grid"], det)
        ref_I.append(I64[0])
        ref_p.append(np.log(I064 / I64[0]))
        log(f"accuracy: FP64 oracle view {v} done")
    ref_I, ref_p = np.stack(ref_I), np.stack(ref_p)

    def rel(a, b, f):
        return np.abs(a - b) / np.maximum(np.abs(b), f * np.abs(b).max())

    out = {"views": list(views), "metric": "|d| / max(|ref|, f max|ref|) vs FP64 oracle, f = 1e-3 (I), "
                                           "1e-2 (p = ln(I0/I)), over every ray of the views"}
    for name, pr in projs.items():
        I = np.stack([pr.forward(v, 1)[0][0].cpu().numpy().astype(np.float64) for v in views])
        I0 = pr.flat_field().cpu().numpy().astype(np.float64)
        torch.cuda.synchronize()
        eI, ep = rel(I, ref_I, 1e-3), rel(np.log(I0[None] / I), ref_p, 1e-2)
        out[name] = {"I_max": float(eI.max()), "I_mean": float(eI.mean()),
                     "p_max": float(ep.max()), "p_mean": float(ep.mean())}
        log(f"accuracy {name}: {out[name]}")
    return out


def config1_figures(steps, warmup, cpu_seconds):
    """The 128^3 single-material scan (256x256, 90 views, K = 272, 61 bins,
    energy-integrating CsI, no grid): the plain FP32 GPU projector (every
    corner from HBM/L2) and the LDS-staged one, against the oracle's FP32
    mirror on this host's cores (a bounded sample of views), with the GPU's
    error against that mirror on the views the CPU ran."""
    import numpy as np
    import torch

    import ct_forward_projection_amd as P
    from ct_forward_projection_amd import synthetic as S
    from oracle import oracle as O
    s = S.config1()
    vol = S.cylinder_phantom()
    kw = P.spectral_inputs(s, *s.det)
    out = {"workload": "128^3 water/ellipsoid phantom (1 mm), 256x256 @ 0.8 mm, 90 views, K=272, "
                       "80 kVp 61 bins, CsI E*eta, no grid"}
    with P.Projector(P.scan_params(**s.scan_kwargs())) as pr:
        pr.set_volumes([vol])
        pr.set_spectral(**kw)
        samples = s.views * pr.samples_per_view
        out["samples_per_scan"] = samples
        stream = torch.cuda.current_stream()
        for march in ("global", "lds"):
            pr.set_tuning(march=march)
            for _ in range(warmup):
                pr.forward()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(steps):
                I, IA = pr.forward()
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / steps
            key = "gpu_fp32_plain" if march == "global" else "gpu_fp32_staged"
            out[key] = {"gsamples_per_s": round(samples / ms / 1e6, 2), "views_per_s": round(s.views / ms * 1e3, 1),
                        "ms_per_scan": round(ms, 3)}
            log(f"config 1 {key}: {out[key]}")
        I_gpu = I.cpu().numpy()
    g = O.geom_from_setup(s)
    p = O.paras(g)
    threads = O.num_threads()
    spec = np.asarray(kw["spectrum"], np.float32) * np.arange(1, len(kw["spectrum"]) + 1, dtype=np.float32)
    det = det_table(kw)
    t = time.perf_counter()
    n = 0
    errs = []
    while n < s.views and (n == 0 or time.perf_counter() - t < cpu_seconds):
        L = O.march_f32(g, p, [vol], n, 1, threads=threads)
        Ic, _, _ = O.spectrum_f32(L, kw["atten"], spec, kw["i0_val"], None, det, threads=threads)
        ref = Ic[0].astype(np.float64)
        errs.append(float((np.abs(I_gpu[n] - ref) / np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max())).max()))
        n += 1
    secs = time.perf_counter() - t
    spv = s.det[0] * s.det[1] * p.n_steps
    out["cpu_fp32_mirror"] = {"gsamples_per_s": round(n * spv / secs / 1e9, 4),
                              "views_per_s": round(n / secs, 3), "cores": threads,
                              "sample": f"views 0-{n - 1} of {s.views}, march + spectrum, {secs:.1f} s"}
    out["gpu_vs_cpu_fp32_max_rel_err"] = max(errs)
    out["speedup_gpu_plain_vs_cpu"] = round(out["gpu_fp32_plain"]["gsamples_per_s"] /
                                            out["cpu_fp32_mirror"]["gsamples_per_s"], 1)
    log(f"config 1 cpu: {out['cpu_fp32_mirror']}, gpu vs cpu {max(errs):.2e}")
    return out


def run_adjoint(args):
    print(json.dumps(adjoint_figures(args.steps, args.warmup)), flush=True)


def adjoint_figures(steps, warmup):
    """Differentiable projector (3 x 256^3, 512x512, 180 views, grid + CsI):
    one weighted-least-squares step on log projections against a Poisson-noisy
    target (1e5 photons per pixel).  Forward = layout + projection launch;
    backward = layout + adjoint launch + the loss's elementwise ops.  Returns
    both times, their ratio and peak device memory."""
    import numpy as np
    import torch

    import ct_forward_projection_amd as P
    from ct_forward_projection_amd import autograd as A
    from ct_forward_projection_amd import synthetic as S

    setup = S.config4()
    truth = np.stack(S.three_material_phantom(setup.n_vox[0], seed=41))
    kw = P.spectral_inputs(setup, *setup.det)
    dev = torch.device("cuda", 0)
    proj = P.Projector(P.scan_params(**setup.scan_kwargs()))
    proj.set_spectral(**kw)
    I0 = proj.flat_field().double()
    truth_t = torch.from_numpy(truth).to(dev)
    with torch.no_grad():
        It, _ = A.project(truth_t, proj)
    counts = A.poisson_target(It, seed=3).double()
    p_t = A.log_projection(counts, I0)
    rng = np.random.default_rng(5)
    est = (truth_t * torch.from_numpy(rng.uniform(0.9, 1.1, truth.shape).astype(np.float32)).to(dev))
    est.requires_grad_(True)
    del It, truth_t
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()

    def step(timed):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        est.grad = None
        ev[0].record(stream)
        I, _ = A.project(est, proj)
        ev[1].record(stream)
        loss = A.wls_loss(A.log_projection(I.double(), I0), p_t, counts)
        loss.backward()
        ev[2].record(stream)
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2]), float(loss)

    for i in range(warmup):
        step(False)
        log(f"adjoint: warmup {i + 1}/{warmup} done")
    torch.cuda.reset_peak_memory_stats(dev)
    fw, bw, losses = [], [], []
    for i in range(steps):
        f, b, l = step(True)
        fw.append(f)
        bw.append(b)
        losses.append(l)
        log(f"adjoint: step {i + 1}: forward {f:.2f} ms, backward {b:.2f} ms, loss {l:.6g}")
    peak = torch.cuda.max_memory_allocated(dev) + proj.device_bytes
    f, b = sum(fw) / len(fw), sum(bw) / len(bw)
    samples = setup.views * proj.samples_per_view
    out = {"metric": "differentiable projector: backward / forward time", "value": round(b / f, 3),
           "unit": "ratio", "higher_is_better": False, "steps": steps, "warmup": warmup,
           "forward_ms": round(f, 2), "backward_ms": round(b, 2),
           "forward_gsamples_per_s": round(samples / (f / 1e3) / 1e9, 2),
           "peak_device_bytes": int(peak), "projector_device_bytes": proj.device_bytes,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "3-material 256^3 (1 mm), 512x512 @ 0.8 mm, 180 views, 120 kVp "
                                  "101 bins, Pb grid 10:1, CsI; WLS on log projections vs "
                                  "Poisson target at 1e5 photons/pixel",
                      "tuning": proj.get_tuning()}}
    proj.close()
    return out


def batch_specs(seed: int = 2024):
    from ct_forward_projection_amd import synthetic as S
    return S.config5_specs(seed)


def batch_setup(spec):
    from ct_forward_projection_amd import synthetic as S
    return S.config5_setup(spec)


def run_batch(args):
    print(json.dumps(batch_figures(args.steps, args.warmup)), flush=True)


def batch_figures(steps, warmup):
    """Eight uneven scans in one persistent, graph-captured launch (FP16
    storage, each scan autotuned), against the same eight run one after the
    other through the plain FP32 projector."""
    import torch

    import ct_forward_projection_amd as P

    dev = torch.device("cuda", 0)
    specs = batch_specs()
    projs, plain, items = [], [], []
    tuned = []
    for sp in specs:
        s, vols = batch_setup(sp)
        kw = P.spectral_inputs(s, *s.det)
        pr = P.Projector(P.scan_params(**s.scan_kwargs()))
        pr.set_volumes(vols, storage_dtype=P.CTFP_F16)
        pr.set_spectral(**kw)
        tuned.append(pr.autotune())
        I = torch.empty((s.views, s.det[1], s.det[0]), device=dev)
        items.append((pr, 0, s.views, I, torch.empty_like(I)))
        projs.append(pr)
        pl = P.Projector(P.scan_params(**s.scan_kwargs()))
        pl.set_volumes(vols, storage_dtype=P.CTFP_F32)
        pl.set_spectral(**kw)
        pl.set_tuning(march="global")
        plain.append(pl)
        log(f"scan {len(projs)}: {sp} tuned {tuned[-1]}")
    batch = P.Batch(items)
    samples = batch.samples
    stream = torch.cuda.current_stream()

    def timed(fn, label):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        log(f"{label}: {ms:.2f} ms per pass")
        return ms

    ms_batch = timed(lambda: batch.run(stream), "batch (FP16, persistent graph)")
    outs = [(torch.empty_like(I), torch.empty_like(IA)) for _, _, _, I, IA in items]

    def sequential(ps):
        for pr, (I, IA), it in zip(ps, outs, items):
            pr.forward(0, it[2], I, IA, stream=stream)

    ms_plain = timed(lambda: sequential(plain), "sequential plain FP32")
    ms_seq16 = timed(lambda: sequential(projs), "sequential FP16 (tuned, one launch each)")
    # the batch and the per-scan FP16 launches compute the same projections
    same = all(torch.equal(I, o[0]) and torch.equal(IA, o[1]) for (_, _, _, I, IA), o in zip(items, outs))
    diffs = []
    for (_, _, _, I, IA), o in zip(items, outs):
        d = (I - o[0]).abs()
        diffs.append({"max_abs": float(d.max()), "max_rel": float((d / o[0].abs().clamp_min(1e-30)).max()),
                      "n_diff": int((d != 0).sum()), "nan": int(torch.isnan(I).sum())})
    log(f"batch vs per-scan launches: {diffs}")
    out = {"metric": "batched uneven workload: aggregate Gsamples/s", "value": round(samples / ms_batch / 1e6, 2),
           "unit": "Gsamples/s", "higher_is_better": True, "steps": steps, "warmup": warmup,
           "ms_per_pass": round(ms_batch, 2), "samples_per_pass": samples,
           "plain_fp32_sequential_value": round(samples / ms_plain / 1e6, 2),
           "speedup_vs_plain_fp32_sequential": round(ms_plain / ms_batch, 3),
           "fp16_sequential_value": round(samples / ms_seq16 / 1e6, 2),
           "batch_equals_per_scan_launches": same, "batch_vs_per_scan": diffs,
           "dtype": "f16", "data": "synthetic",
           "config": {"workload": "8 seeded scans: cubic phantoms 128-512 (0.5 mm), 1-3 materials, "
                                  "detectors 256x256-1024x768, 60-360 views, 80-140 kVp, grid on/off",
                      "scans": specs, "tuning": tuned}}
    batch.close()
    for p in projs + plain:
        p.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--views", type=int, default=0, help="views per step (0 = the full 360)")
    ap.add_argument("--dtype", choices=["f32", "f16", "bf16"], default="f16",
                    help="storage precision of the resident volumes (f16: the mixed-precision "
                         "projector; f32 and the global-gather FP32 projector are timed beside it)")
    ap.add_argument("--march", choices=["lds", "global"], default="lds",
                    help="corner gathers from LDS-staged chunks or straight from HBM/L2")
    ap.add_argument("--chunk", type=int, default=0, help="samples per staged chunk (0 = library default)")
    ap.add_argument("--lds-kib", type=int, default=0, help="LDS per workgroup in KiB (0 = default)")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the 8-view error check against the FP64 oracle")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the 128^3, differentiable 256^3 and batched figures")
    ap.add_argument("--compare-steps", type=int, default=5,
                    help="timed steps of the FP32 projectors beside the headline (at most --steps)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true",
                    help="do not time the FP32 projectors beside a 16-bit run")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="also launch the copy kernel over this many MiB (counter calibration)")
    ap.add_argument("--workload", choices=["scan", "adjoint", "batch"], default="scan",
                    help="scan: the 512^3 forward projection (headline); adjoint: one "
                         "gradient step of the differentiable 256^3 projector; batch: eight "
                         "uneven scans in one persistent launch")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "views shard with no exchange, see the sharding design notes", file=sys.stderr)
        sys.exit(2)

    if args.workload == "adjoint":
        run_adjoint(args)
        return
    if args.workload == "batch":
        run_batch(args)
        return

    import numpy as np
    import torch

    import ct_forward_projection_amd as P
    from ct_forward_projection_amd import synthetic as S

    setup = S.config2()
    t = time.perf_counter()
    vols = list(S.three_material_phantom(512))
    kw = P.spectral_inputs(setup, *setup.det)
    log(f"inputs ready in {time.perf_counter() - t:.1f} s")
    dtypes = {"f32": P.CTFP_F32, "f16": P.CTFP_F16, "bf16": P.CTFP_BF16}
    proj = P.Projector(P.scan_params(**setup.scan_kwargs()))
    proj.set_volumes(vols, storage_dtype=dtypes[args.dtype])
    proj.set_spectral(**kw)
    tune = {"march": args.march}
    if args.chunk > 0:
        tune["chunk"] = args.chunk
    if args.lds_kib > 0:
        tune["lds_bytes"] = args.lds_kib * 1024
    proj.set_tuning(**tune)
    views = setup.views if args.views <= 0 else min(args.views, setup.views)
    nu, nv = setup.det
    dev = torch.device("cuda", 0)
    I = torch.empty((views, nv, nu), dtype=torch.float32, device=dev)
    IA = torch.empty_like(I)
    stream = torch.cuda.current_stream()
    samples = views * proj.samples_per_view
    n_mat = len(vols)
    # one voxel-equivalent per material per sample, in the stored precision
    bytes_per_sample = (4 if args.dtype == "f32" else 2) * n_mat

    if args.calibrate_copy_mib > 0:
        nb = args.calibrate_copy_mib << 20
        a = torch.empty(nb // 4, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        from ct_forward_projection_amd import _lib as L
        L.check(L.lib().ctfp_bench_copy(a.data_ptr(), b.data_ptr(), nb, stream.cuda_stream))
        torch.cuda.synchronize()
        log(f"copy kernel: {nb} bytes read, {nb} bytes written")

    def timed(pr, label, steps, warmup):
        for i in range(warmup):
            pr.forward(0, views, I, IA, stream=stream)
            torch.cuda.synchronize()
            log(f"{label}: warmup {i + 1}/{warmup} done")
        starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
        ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            starts[i].record(stream)
            pr.forward(0, views, I, IA, stream=stream)
            ends[i].record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kern = [s.elapsed_time(e) for s, e in zip(starts, ends)]
        log(f"{label}: timed {steps} steps, wall {wall:.3f} s, per-launch ms {kern}")
        return wall, kern

    marched = proj.marched_samples(0, views)
    evaluated = proj.evaluated_samples(0, views)
    wall, kern_ms = timed(proj, args.dtype, args.steps, args.warmup)
    ms_per_step = wall * 1e3 / args.steps

    # stage split: the same launch without the spectrum (line integrals of
    # the M materials written instead of I and IAbsorb), outside the timed run
    Lbuf = torch.empty((views, len(vols), nv, nu), dtype=torch.float32, device=dev)
    proj.line_integrals(0, views, out=Lbuf, stream=stream)
    march_ms = []
    for _ in range(3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        proj.line_integrals(0, views, out=Lbuf, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        march_ms.append(e0.elapsed_time(e1))
    del Lbuf
    log(f"march only (line integrals): per-launch ms {march_ms}")
    gsps = samples / (ms_per_step / 1e3) / 1e9
    avg_launch_s = sum(kern_ms) / len(kern_ms) / 1e3
    achieved = samples * bytes_per_sample / avg_launch_s / 1e9

    # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE /
    # WRITE_SIZE runs of this workload (not measured in this run)
    traffic, traffic_src = None, None
    tfile = os.path.join(ROOT, "profiles", "march_config2_traffic.json")
    if os.path.exists(tfile) and views == setup.views:
        with open(tfile) as f:
            entry = json.load(f).get(f"{args.dtype}_{args.march}")
        if entry:
            traffic = entry.get("bytes_per_launch")
            traffic_src = "profiles/march_config2_traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by " \
                          "the copy-kernel calibration; a committed measurement, not this run)"

    result = {
        "metric": METRIC, "value": round(gsps, 3), "unit": "Gsamples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": args.dtype,
        "data": "synthetic",
        "config": {"workload": "3-material 512^3 (water/bone/iodine, 0.5 mm), 1024x768 "
                               "@ 0.4 mm, 360 views, K=1082, 120 kVp 101 bins, Pb grid 10:1 "
                               "40 l/cm, CsI response",
                   "views_per_step": views, "samples_per_step": samples,
                   "materials": n_mat, "volume_dtype": args.dtype, "tuning": proj.get_tuning()},
        "views_per_s": round(views / (ms_per_step / 1e3), 2),
        "median_launch_ms": round(sorted(kern_ms)[len(kern_ms) // 2], 3),
        # samples inside the rays' support windows, and those the march
        # evaluates after skipping chunks whose cells are all zero (the rest
        # are exact zeros of the reference's lattice and count as done)
        "window_fraction": round(marched / samples, 4),
        "evaluated_fraction": round(evaluated / samples, 4),
        "stages_ms": {"march_only": round(sorted(march_ms)[1], 3),
                      "spectrum": round(sorted(kern_ms)[len(kern_ms) // 2] - sorted(march_ms)[1], 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_source": traffic_src,
                     "limiter": "VALU instruction issue (rocprofv3 SQ counters: "
                                "profiles/march_config2_f16.md), not HBM bandwidth"},
    }
    if args.dtype != "f32" and not args.no_fp32:
        # the LDS-staged FP32 projector with the same launch options, and the
        # plain FP32 projector (every corner gathered from HBM/L2, no staging)
        ref = P.Projector(P.scan_params(**setup.scan_kwargs()))
        ref.set_volumes(vols, storage_dtype=P.CTFP_F32)
        ref.set_spectral(**kw)
        ref.set_tuning(**tune)
        cs = max(1, min(args.steps, args.compare_steps))
        wall32, _ = timed(ref, "f32", cs, 1)
        result["staged_fp32_value"] = round(samples / (wall32 / cs) / 1e9, 3)
        result["speedup_vs_staged_fp32"] = round(result["value"] / result["staged_fp32_value"], 3)
        ref.set_tuning(march="global")
        wallg, _ = timed(ref, "f32 global", cs, 1)
        result["plain_fp32_value"] = round(samples / (wallg / cs) / 1e9, 3)
        result["speedup_vs_plain_fp32"] = round(result["value"] / result["plain_fp32_value"], 3)
        result["compare_steps"] = cs
        ref.set_tuning(**tune)
        if not args.no_accuracy:
            other = "bf16" if args.dtype == "f16" else "f16"
            proj2 = P.Projector(P.scan_params(**setup.scan_kwargs()))
            proj2.set_volumes(vols, storage_dtype=dtypes[other])
            proj2.set_spectral(**kw)
            result["rel_err_vs_fp64"] = accuracy({args.dtype: proj, other: proj2, "f32": ref}, setup, vols, kw)
            proj2.close()
        ref.close()
    elif not args.no_accuracy:
        result["rel_err_vs_fp64"] = accuracy({args.dtype: proj}, setup, vols, kw)
    del I, IA
    proj.close()
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline(setup, vols, kw, args.cpu_seconds)
    else:
        result["cpu_baseline"] = None
    if not args.no_extras:
        del vols
        torch.cuda.empty_cache()
        log("extras: 128^3 single-material scan ...")
        result["config1_single_material"] = config1_figures(max(3, min(args.steps, 10)), 1,
                                                            min(args.cpu_seconds, 8.0))
        log("extras: differentiable 3 x 256^3 projector ...")
        result["config4_differentiable"] = adjoint_figures(3, 1)
        torch.cuda.empty_cache()
        log("extras: batched uneven workload ...")
        result["config5_batched"] = batch_figures(3, 1)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
