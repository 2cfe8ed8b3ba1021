#!/usr/bin/env python3
"""Benchmark of the undulator integration step on one MI355X.

A step is one pass of the hot path over the whole time window: transverse
half step, space charge, longitudinal RK4 push, transverse half step, source
deposition, ADI field advance, slippage and the per-step P(z)/|b(z)|
diagnostics -- Gencore::run's loop body with output_step=1.  The default
workload is config 3: time-dependent SASE over 4096 slices with a Gaussian
current (peak 4 kA), 2220..16384 macroparticles per slice (40.1 M), 255x255
field, long- and short-range space charge, run in the mixed FP32 form and,
on the same loaded beam, in the FP64 plain form.  Beam and field are
resident in HBM when the timed region starts.  Both forms run the whole
lattice (509 steps); the line reports the mixed form's throughput, its
ratio to the FP64 form's and its P(z), |b(z)|, final-field and spectrum
errors against the FP64 form.

Prints one JSON line.  The CPU baseline is the FP64 oracle (a clean-room
C++ restatement of the reference's loops, OpenMP over slices like the
reference's MPI slice blocks) timed on this host on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "genesis-1.3-version4_amd", "python"))

import numpy as np  # noqa: E402

import g4mi  # noqa: E402
import g4mi_workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0


build_workload = W.build_workload
make_context = W.make_context


def gpu_measure(cfg, steps_k, warmup, adi_kernel=-1, precision=g4mi.FP64, breakdown=20, refine=None,
                mixed_adi=None, fft=False, work=None, whole_lattice=False, sched=()):
    """Times `steps_k` steps after `warmup` on a context of `cfg` in the
    given precision.  With whole_lattice the rest of the lattice runs after
    the timed steps (untimed) so the records and the final field cover the
    whole undulator line."""
    t_load = time.perf_counter()
    off, beam, field, current, steps = work if work is not None else build_workload(cfg)
    ctx = make_context(cfg, off, beam, field, current, precision)
    t_load = time.perf_counter() - t_load
    if fft:
        ctx.set_field_solver(fft=True)
    if adi_kernel >= 0 and precision == g4mi.FP64:
        ctx.set_schedule("adi_kernel", adi_kernel)
    if precision == g4mi.FP32:
        if refine is not None:
            ctx.set_schedule("adi_refine", refine)
        if mixed_adi is not None:
            ctx.set_schedule("mixed_adi", mixed_adi)
    for key, val, form in sched:
        if form is None or form == precision:
            ctx.set_schedule(key, val)
    nst = len(steps)
    extra = breakdown if precision == g4mi.FP32 else 0
    if warmup + steps_k + extra > nst:
        raise SystemExit(f"warmup+steps must be <= {nst - extra} (lattice length)")
    run_steps = steps if whole_lattice else steps[:warmup + steps_k + extra]
    out = W.run_lattice(ctx, run_steps, warmup=warmup, timed=steps_k, breakdown=extra)
    sched = ctx.schedule()
    t0 = warmup + extra  # the mixed form's breakdown steps run before its timed ones
    timed = steps[t0:t0 + steps_k]
    dz = sum(s.dz for s in timed)
    ntot = int(off[-1])
    f, first = ctx.download_field() if whole_lattice else (None, None)
    ctx.close()
    return dict(seconds=out["seconds"], ntot=ntot, dz=dz, kt=out["kt"], recs=out["recs"],
                schedule=sched, load_s=t_load, field=f, first=first, timed0=t0)


# device kernel behind each timed kernel group (the traffic table names kernels)
KERNEL_OF = {g4mi.FP64: {"adi": "k_adi_band", "push": "k_push", "transverse": "k_transverse",
                         "deposit": "k_deposit"},
             "fp64_fused": {"adi": "k_adi_band", "push": "k_fp64_main", "transverse": "k_fp64_front"},
             g4mi.FP32: {"adi": "k_mix_adi_stream", "push": "k_mix_main", "transverse": "k_mix_front"}}


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    FETCH_SIZE/WRITE_SIZE summary (profiles/pmc_traffic.md), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cols) == 6 and cols[0] == workload and kernel in cols[1]:
            try:
                return float(cols[5])
            except ValueError:
                return None
    return None


def measured_traffic_sum(workload, kernel):
    """Sum of the HBM bytes per launch of every kernel whose name contains
    `kernel` on `workload` (profiles/pmc_traffic.md), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.md")
    if not os.path.exists(path):
        return None
    tot = 0.0
    for line in open(path):
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cols) == 6 and cols[0] == workload and kernel in cols[1]:
            try:
                tot += float(cols[5])
            except ValueError:
                return None
    return tot or None


FP64_MFMA_PEAK_TFLOPS = 78.6


def roofline(cfg, kt, ntot, precision=g4mi.FP64, schedule=None):
    """Dominant kernel and its algorithmic bytes per launch.
    FP64 plain form:
      push: read gamma, theta, x, y, px, py + write gamma, theta = 64 B/particle;
      ADI (x then y sweep of every cell): field read + write = 32 B/cell;
      deposit: read x, y, theta, gamma = 32 B/particle;
      transverse (two half steps): 2 x (read 5 + write 4 doubles) = 144 B/particle.
    Mixed FP32 form:
      main (space charge + push + R56 + transverse + deposit): read and write the
      6 FP32 coordinates = 48 B/particle;
      ADI: field read + write = 16 B/cell;
      front (transverse + moments): read 5 + write 4 floats = 36 B/particle."""
    cells = cfg["nslice"] * cfg["ngrid"] ** 2
    top = max((k for k in kt if k != "steps"), key=kt.get)
    if schedule and schedule.get("field_solver") == "fft" and top == "adi" and \
            schedule.get("propagation") == "fft-conv":
        # spectral step as FFT convolutions of every row, then every column:
        # the field read + written once = 32 B/cell, as the ADI
        ms = kt["adi"]
        gbs = 32.0 * cells / (ms * 1e-3) / 1e9
        traffic = measured_traffic_sum(cfg["name"] + "-fft", "k_fft_conv")  # both passes of a step
        return "adi", dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                           frac=round(gbs / HBM_PEAK_GBS, 4), traffic=int(traffic) if traffic else None,
                           kernel="field (fft)",
                           device_kernel="k_fft_conv", ms_per_launch=round(ms, 5),
                           algorithmic_bytes=int(32.0 * cells))
    if schedule and schedule.get("field_solver") == "fft" and top == "adi":
        # spectral step: two complex n x n x n products per slice (E M^T, then
        # M T), 8 real flops per complex multiply-add, on FP64 MFMA
        flops = 16.0 * cfg["ngrid"] ** 3 * cfg["nslice"]
        ms = kt["adi"]
        tf = flops / (ms * 1e-3) / 1e12
        return "adi", dict(bound="mfma", achieved=round(tf, 2), peak=FP64_MFMA_PEAK_TFLOPS, unit="TFLOP/s",
                           frac=round(tf / FP64_MFMA_PEAK_TFLOPS, 4), traffic=None, kernel="field (fft)",
                           device_kernel="k_cgemm", ms_per_launch=round(ms, 5),
                           algorithmic_flops=int(flops))
    fused = precision == g4mi.FP64 and bool(schedule and schedule.get("fp64_fused"))
    if fused:
        # front: read gamma, x, y, px, py + write x, y, px, py = 72 B/particle;
        # main: read and write all six coordinates = 96 B/particle
        per = {"push": 96.0 * ntot, "adi": 32.0 * cells, "transverse": 72.0 * ntot}
    elif precision == g4mi.FP64:
        per = {"push": 64.0 * ntot, "adi": 32.0 * cells, "deposit": 32.0 * ntot,
               "transverse": 144.0 * ntot}
    else:
        per = {"push": 48.0 * ntot, "adi": 16.0 * cells, "transverse": 36.0 * ntot}
    name = max(per, key=lambda k: kt[k])
    ms = kt[name]
    gbs = per[name] / (ms * 1e-3) / 1e9
    dev = KERNEL_OF["fp64_fused" if fused else precision].get(name)
    if name == "adi" and precision == g4mi.FP32 and schedule:
        dev = {"band32-wide": "k_mix_adi_wide", "band32-stream": "k_mix_adi_stream",
               "band32-tile": "k_mix_adi<"}.get(schedule.get("adi"), dev)
    if name == "adi" and precision == g4mi.FP64 and schedule:
        # the ADI kernel the tuner chose for the undulator step (the timed steps)
        und = (schedule.get("adi_by_delz") or {}).get(f"{W.DELZ:.6g}", schedule.get("adi"))
        dev = {"band": "k_adi_band", "band-wide": "k_adi_wide64"}.get(und)  # traffic rows: banded kernels
    traffic = measured_traffic(cfg["name"], dev) if dev else None
    return name, dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=round(gbs / HBM_PEAK_GBS, 4),
                      traffic=int(traffic) if traffic else None,
                      kernel=name, device_kernel=dev, ms_per_launch=round(ms, 5),
                      algorithmic_bytes=int(per[name]))


def _cpu_model():
    """CPU model name and physical core count of this host (/proc/cpuinfo,
    as lscpu reports them)."""
    model, cores = "unknown", set()
    phys = core = None
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model == "unknown":
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k and phys is not None:
                cores.add((phys, core))
                phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    return model, len(cores) or None


def cpu_baseline(cfg, work, nsteps=2, nslices=(512, 32)):
    """The FP64 oracle on this host's CPU over the first `nsteps` steps of a
    block of the workload's central slices (the same loaded particles),
    once with the threads of this GPU's CPU share and once on one core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctypes as C
    import oracle_py as O
    off, beam, _, current, _ = work
    ns = cfg["nslice"]
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    steps = O.unroll(W.fodo_elements(), W.DELZ, W.LAMBDA_R, W.GAMMA0)
    model, phys = _cpu_model()
    res = []
    for nsl, threads in ((min(ns, nslices[0]), share), (min(ns, nslices[1]), 1)):
        i0 = (ns - nsl) // 2
        a, b = int(off[i0]), int(off[i0 + nsl])
        counts = np.diff(off[i0:i0 + nsl + 1])
        sim = O.Sim(counts, cfg["ngrid"], cfg["gridmax"], W.LAMBDA_R, timerun=cfg["timerun"],
                    sc=(cfg["sc_nz"], cfg["sc_nphi"], cfg["sc_ngrid"], 0.0),
                    longrange=cfg["longrange"], threads=threads)
        for k in ("gamma", "theta", "x", "y", "px", "py"):
            getattr(sim, k)[:] = beam[k][a:b]
        sim.current[:] = current[i0:i0 + nsl]
        t0 = time.perf_counter()
        for i in range(nsteps):
            O.lib().orc_step_all(C.byref(sim.s), C.byref(steps[i]), W.GAMMA0)
            sim.diag()
        dt = time.perf_counter() - t0
        res.append(dict(value=round((b - a) * nsteps / dt, 1), threads=threads, slices=nsl,
                        macroparticles=b - a, seconds=round(dt, 2)))
    multi, single = res
    return dict(value=multi["value"], unit="macroparticle-steps/s", cores=multi["threads"],
                kind="port", cpu_model=model, host_physical_cores=phys,
                single_core=dict(value=single["value"], cores=1,
                                 sample=f"{single['slices']} central slices, {single['macroparticles']} "
                                        f"macroparticles, {nsteps} steps, {single['seconds']} s"),
                sample=f"the first {nsteps} steps of {cfg['name']} (FP64 oracle, OpenMP over slices) on "
                       f"its {multi['slices']} central slices ({multi['macroparticles']} of the workload's "
                       f"loaded macroparticles, {cfg['ngrid']}^2 field), {multi['seconds']} s on "
                       f"{multi['threads']} threads: the CPU share of one GPU on this host")


def fno_probe(ngrid=151, batch=16, reps=5):
    """Throughput of the learned field step at its full shape (width 32,
    16x16 modes, 4 layers, hidden 64): forward and forward+backward passes
    on random data."""
    f = g4mi.Fno(ngrid, width=32, modes=16, layers=4, hidden=64, max_batch=batch, seed=1)
    P = ngrid * ngrid
    rng = np.random.default_rng(0)
    inp = rng.normal(size=(batch, P, 4)).astype(np.float32)
    tgt = rng.normal(size=(batch, P, 2)).astype(np.float32)
    f.forward(inp)
    t0 = time.perf_counter()
    for _ in range(reps):
        f.forward(inp)
    tf = (time.perf_counter() - t0) / reps
    f.loss_grad(inp, tgt)
    t0 = time.perf_counter()
    for _ in range(reps):
        f.loss_grad(inp, tgt)
    tb = (time.perf_counter() - t0) / reps
    out = dict(fno_probe=dict(ngrid=ngrid, batch=batch, params=f.param_count(),
                              forward_ms=round(tf * 1e3, 3), forward_backward_ms=round(tb * 1e3, 3),
                              samples_per_s_train=round(batch / tb, 1)))
    f.close()
    print(json.dumps(out), flush=True)
    return 0


def deck_example1(reps=2):
    """Wall time of the reference's Example1 deck through the deck runner
    (parsing, lattice matching, beam and field generation, 1104 steps with a
    record after each including the spectral block, the .out.h5 file), the
    whole-program figure the reference publishes for it
    (examples/Example1-SteadyState/README.md: 29.5481 s on one MPI rank)."""
    import tempfile
    tmp = tempfile.mkdtemp(prefix="g4mi_ex1_")
    lat = os.path.join(tmp, "Example1.lat")
    with open(lat, "w") as f:
        f.write(W.EXAMPLE1_LAT.format(extra=""))
    path = os.path.join(tmp, "Example1.in")
    with open(path, "w") as f:
        f.write(W.EXAMPLE1_IN.format(root=os.path.join(tmp, "ex1"), lat=lat))
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        g4mi.run_deck(path, device=0, quiet=True)
        times.append(time.perf_counter() - t0)
    out = os.path.join(tmp, "ex1.out.h5")
    pend = None
    try:
        pend = float(g4mi.h5_read(out, "Field/power")[-1, 0])
    except Exception:
        pass
    print(json.dumps(dict(example1_deck=dict(
        seconds=round(min(times), 3), seconds_each=[round(t, 3) for t in times], p_end_w=pend,
        output_h5=os.path.exists(out), reference_published_s=29.5481,
        reference_source="examples/Example1-SteadyState/README.md:142,180 (whole program, one MPI rank)"))),
        flush=True)
    return 0


def _run_steps(ctx, steps):
    """Advance ctx over `steps` without records."""
    if not steps:
        return
    arr = (g4mi.StepParams * len(steps))(*steps)
    g4mi._chk(g4mi.lib().g4mi_run(ctx.h, arr, len(steps), W.GAMMA0, None), "g4mi_run")


def _run_power(ctx, steps):
    """Advance ctx over `steps`; slice power after each step, [len(steps)][ns]."""
    r = ctx.run(steps, W.GAMMA0, record=("power",))
    return r["power"][1:]


def config4(a):
    """Learned field step (config 4).  Training pairs: mixed-form config 1
    runs (1024 slices x 8192, 151^2) with shot-noise seeds seed0.. ; on every
    `step_stride`-th step the ADI's (E, S, E_next) of every `slice_stride`-th
    slice are captured until `pairs` are held.  Training: Adam with a cosine
    schedule, the last 5 % of the pairs (the last seeds' pairs) held out.
    Deployment: config 3's beam (4096 slices, Gaussian current, uneven
    counts) on a 151^2 grid in the mixed form with the network in place of
    the ADI, ADI on every 8th step and when the energy-balance residual
    exceeds 1e-3; P(z) against the FP64 GPU form (the oracle's arithmetic to
    1e-10) and against the mixed form with the ADI alone."""
    t_start = time.perf_counter()
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    ng = 151
    steps = g4mi.parse_lattice(W.lattice_text(), "FEL", W.DELZ, W.LAMBDA_R, W.GAMMA0)
    nst = len(steps)
    npairs, nseeds = a.fno_pairs, a.fno_seeds
    per_seed = -(-npairs // nseeds)
    cfg1 = dict(W.CONFIG1, precision="fp32")
    slice_stride = 64
    per_step = -(-cfg1["nslice"] // slice_stride)
    step_stride = max(1, (nst * per_step) // per_seed)
    fno = g4mi.Fno(ng, width=32, modes=16, layers=4, hidden=64, max_batch=a.fno_batch, seed=1)
    fno.reserve(npairs)
    t0 = time.perf_counter()
    for k in range(nseeds):
        if fno.pair_count() >= npairs:
            break
        cfg = dict(cfg1, seed=a.fno_seed0 + k)
        off, beam, field, current, _ = build_workload(cfg)
        ctx = make_context(cfg, off, beam, field, current, g4mi.FP32)
        del beam, field
        for i in range(0, nst, step_stride):
            ctx.attach_fno(fno, 1, every=slice_stride)
            _run_steps(ctx, steps[i:i + 1])
            ctx.attach_fno(None, 0)
            _run_steps(ctx, steps[i + 1:i + step_stride])
        ctx.sync()
        ctx.close()
        log(f"capture seed {cfg['seed']}: {fno.pair_count()} pairs, {time.perf_counter() - t0:.1f} s")
    t_capture = time.perf_counter() - t0
    n = fno.pair_count()
    nh = max(1, n // 20)
    ntr = n - nh
    base = fno.eval(ntr, n, a.fno_train_batch)
    p0 = fno.get_params()
    fno.set_params(np.zeros_like(p0))
    zero = fno.eval(ntr, n, a.fno_train_batch)  # field-unchanged prediction
    fno.set_params(p0)
    t0 = time.perf_counter()
    tls, hls = [], []
    for e in range(a.fno_epochs):
        tl, hl = fno.train(epochs=1, batch=a.fno_train_batch, lr=a.fno_lr, nholdout=nh, seed=11,
                           first_epoch=e, total_epochs=a.fno_epochs)
        tls.append(float(tl[0]))
        hls.append(float(hl[0]))
        log(f"epoch {e}: train loss {tl[0]:.4e}, hold-out loss {hl[0]:.4e}, "
            f"{time.perf_counter() - t0:.1f} s")
    t_train = time.perf_counter() - t0

    # deployment on config 3's beam at 151^2
    cfg3 = dict(W.CONFIG3, ngrid=ng)
    nd = min(nst, a.fno_deploy_steps) if a.fno_deploy_steps > 0 else nst
    dsteps = steps[:nd]
    off, beam, field, current, _ = build_workload(cfg3)
    runs = {}
    for name, prec in (("fp64", g4mi.FP64), ("mixed_adi", g4mi.FP32)):
        ctx = make_context(cfg3, off, beam, field, current, prec)
        ctx.sync()
        t0 = time.perf_counter()
        runs[name] = _run_power(ctx, dsteps)
        ctx.sync()
        runs[name + "_s"] = time.perf_counter() - t0
        ctx.close()
        log(f"deploy {name}: {runs[name + '_s']:.2f} s")
    ctx = make_context(cfg3, off, beam, field, current, g4mi.FP32)
    del beam, field
    ctx.attach_fno(fno, 2, every=8, threshold=1e-3)
    ctx.sync()
    t0 = time.perf_counter()
    parts = []
    for i in range(0, nd, 16):
        parts.append(_run_power(ctx, dsteps[i:i + 16]))
        log(f"deploy fno: step {min(nd, i + 16)}/{nd}, {time.perf_counter() - t0:.1f} s")
    ctx.sync()
    t_fno = time.perf_counter() - t0
    stats, judge = ctx.fno_stats(), ctx.fno_judge()
    ctx.close()
    fno.close()
    pf = np.concatenate(parts).sum(axis=1)
    p64, pm = runs["fp64"].sum(axis=1), runs["mixed_adi"].sum(axis=1)
    floor = 1e-6 * p64.max()  # relative errors where P is above 1e-6 of its maximum
    mask = p64 > floor

    def rel(x):
        return float(np.max(np.abs(x[mask] / p64[mask] - 1))) if mask.any() else None

    out = {
        "metric": "learned field step: P(z) error and fallback rate",
        "config": {"workload": "config4: FNO (4 layers, width 32, 16x16 modes, hidden 64) on "
                               "151x151; pairs from mixed-form config 1 runs; deployed on config 3's "
                               "beam (4096 slices, Gaussian current) at 151x151",
                   "pairs": n, "holdout_pairs": nh, "seeds": f"{a.fno_seed0}..{a.fno_seed0 + nseeds - 1}",
                   "capture": f"every {step_stride}th step, every {slice_stride}th slice",
                   "epochs": a.fno_epochs, "train_batch": a.fno_train_batch, "lr": a.fno_lr,
                   "deploy_steps": nd},
        "dtype": "bf16 MFMA, f32 accumulate (network); f32 mixed form (beam, field)",
        "data": "synthetic: Hammersley quiet loading + shot noise, as LoadBeam",
        "training": {"capture_s": round(t_capture, 1), "train_s": round(t_train, 1),
                     "samples_per_s": round(ntr * a.fno_epochs / t_train, 1),
                     "train_loss": tls, "holdout_loss": hls,
                     "holdout_loss_untrained": base, "holdout_loss_field_unchanged": zero},
        "deployment": {"fno_steps": stats["fno_steps"], "adi_steps": stats["adi_steps"],
                       "fallback_rate": round(stats["adi_steps"] / max(1, nd), 4),
                       "forced_every_8": -(-nd // 8),
                       "residual_fallbacks": stats["adi_steps"] - -(-nd // 8),
                       "mean_residual_fno": stats["mean_residual_fno"],
                       "mean_residual_adi": stats["mean_residual_adi"],
                       "judge": judge,
                       "p_rel_err_max_fno_vs_fp64": rel(pf),
                       "p_rel_err_end_fno_vs_fp64": float(abs(pf[-1] / p64[-1] - 1)),
                       "p_rel_err_max_mixed_adi_vs_fp64": rel(pm),
                       "p_rel_err_end_mixed_adi_vs_fp64": float(abs(pm[-1] / p64[-1] - 1)),
                       "p_end_w": {"fp64": float(p64[-1]), "mixed_adi": float(pm[-1]), "fno": float(pf[-1])},
                       "s_per_step": {"fno_run": round(t_fno / nd, 5),
                                      "mixed_adi_run": round(runs["mixed_adi_s"] / nd, 5),
                                      "fp64_run": round(runs["fp64_s"] / nd, 5)}},
        "total_s": round(time.perf_counter() - t_start, 1),
    }
    print(json.dumps(out), flush=True)
    return 0


def copy_probe(mib):
    """A plain device-to-device copy of a known size: its FETCH_SIZE and
    WRITE_SIZE per dispatch, against mib MiB, give the counters' scale."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_mib": mib, "ok": bool(dst[-1].item() == 1.0)}), flush=True)
    return 0


def _line(cfg, m, steps, warmup, prec):
    """Rate, per-kernel times and roofline of one measured form."""
    rate = m["ntot"] * steps / m["seconds"]
    cells = cfg["nslice"] * cfg["ngrid"] ** 2
    _, roof = roofline(cfg, m["kt"], m["ntot"], prec, m["schedule"])
    return dict(value=round(rate, 1), ms_per_step=round(1e3 * m["seconds"] / steps, 4),
                field_cell_updates_per_s=round(cells * steps / m["seconds"], 1),
                s_per_undulator_metre=round(m["seconds"] / m["dz"], 5),
                kernel_ms_per_step={k: round(v, 5) for k, v in m["kt"].items() if k != "steps"},
                roofline=roof, schedule=m["schedule"])


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=80,
                    help="timed steps; the default keeps them inside the first undulator module")
    ap.add_argument("--warmup", type=int, default=20,
                    help="untimed steps first; the mixed form's schedule tuner takes the first 18 steps")
    ap.add_argument("--config", choices=("config1", "config0", "config2", "config3", "config4"),
                    default="config3",
                    help="config3 (default) and config2 run the mixed form and, on the same beam, "
                         "the FP64 form over the whole lattice; config4 runs the learned field step")
    ap.add_argument("--precision", choices=("fp64", "fp32"), default=None,
                    help="override the configuration's precision (one form only)")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--adi-kernel", type=int, default=-1,
                    help="fix the ADI kernel (0 Thomas, 1 fused, 2 fused+prefetch, 3 banded); "
                         "default: autotuned")
    ap.add_argument("--field-solver", choices=("adi", "fft"), default="adi",
                    help="FP64 form: the ADI (default) or the spectral field solver")
    ap.add_argument("--adi-refine", type=int, default=None, choices=(0, 1),
                    help="mixed form: FP64 residual correction in the ADI (default: the library's)")
    ap.add_argument("--mixed-adi", type=int, default=None, choices=(0, 1, 2),
                    help="mixed form: 0 tile ADI kernel, 1 streaming strips, 2 whole-width streaming "
                         "(default: the library's)")
    ap.add_argument("--set", action="append", default=[], metavar="[FORM:]KEY=VAL",
                    help="a schedule entry (g4mi_set_schedule) for the measured contexts, e.g. "
                         "fp32:main_nt=384 or fp64:adi_kernel=3; repeatable")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (the reference "
                         "point that scales the FETCH_SIZE/WRITE_SIZE counters) and exit")
    ap.add_argument("--fno-probe", action="store_true", help="only time the learned field step")
    ap.add_argument("--deck-example1", action="store_true",
                    help="only time the reference's Example1 deck through the deck runner (whole program)")
    ap.add_argument("--fno-pairs", type=int, default=200000)
    ap.add_argument("--fno-seeds", type=int, default=100)
    ap.add_argument("--fno-seed0", type=int, default=100)
    ap.add_argument("--fno-epochs", type=int, default=3)
    ap.add_argument("--fno-lr", type=float, default=1e-3)
    ap.add_argument("--fno-batch", type=int, default=64, help="slices per inference pass")
    ap.add_argument("--fno-train-batch", type=int, default=16)
    ap.add_argument("--fno-deploy-steps", type=int, default=0, help="0: the whole lattice")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the slice-block sharding "
                                   "across GPUs is a design, not an implementation",
                          "n_gpus": a.gpus}))
        return 2
    if a.copy_probe:
        return copy_probe(a.copy_probe)
    if a.fno_probe:
        return fno_probe()
    if a.deck_example1:
        return deck_example1()
    if a.config == "config4":
        return config4(a)
    cfg = dict(W.CONFIGS[a.config])
    compare = a.precision is None and cfg["precision"] == "fp32"
    if a.precision:
        cfg["precision"] = a.precision
    prec = g4mi.FP32 if cfg["precision"] == "fp32" else g4mi.FP64
    log = lambda msg: print(msg, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.perf_counter()
    work = build_workload(cfg)
    log(f"workload {cfg['name']}: {int(work[0][-1])} macroparticles loaded in "
        f"{time.perf_counter() - t0:.1f} s")
    sched = []
    for item in a.set:
        form, _, kv = item.rpartition(":")
        key, _, val = kv.partition("=")
        sched.append((key, int(val), {"fp32": g4mi.FP32, "fp64": g4mi.FP64}.get(form)))
    kw = dict(adi_kernel=a.adi_kernel, refine=a.adi_refine, mixed_adi=a.mixed_adi,
              fft=a.field_solver == "fft", work=work, whole_lattice=compare, sched=sched)
    ref = None
    if compare:
        ref = gpu_measure(cfg, a.steps, a.warmup, precision=g4mi.FP64, **kw)
        log(f"FP64 form: {1e3 * ref['seconds'] / a.steps:.3f} ms per step")
    m = gpu_measure(cfg, a.steps, a.warmup, precision=prec, **kw)
    log(f"{cfg['precision']} form: {1e3 * m['seconds'] / a.steps:.3f} ms per step")
    main_line = _line(cfg, m, a.steps, a.warmup, prec)
    if prec == g4mi.FP64:
        dtype = "f64"
    else:
        refine = m["schedule"].get("adi_refine", 0)
        dtype = "f32 (f64 slice phase base, f64 sums" + (", f64 ADI residual correction)" if refine else ")")
    out = {
        "metric": "macroparticle-steps/s",
        "value": main_line["value"],
        "unit": "macroparticle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": main_line["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic: Hammersley quiet loading + shot noise (seed 7), as LoadBeam",
        "config": {"workload": W.describe(cfg) +
                   ("" if prec == g4mi.FP64 else
                    (", FP64 ADI residual correction" if m["schedule"].get("adi_refine", 0)
                     else ", no FP64 ADI residual correction")),
                   "slices": cfg["nslice"], "macroparticles": m["ntot"], "ngrid": cfg["ngrid"],
                   "precision": cfg["precision"], "parallelism": "single GPU"},
        "field_cell_updates_per_s": main_line["field_cell_updates_per_s"],
        "s_per_undulator_metre": main_line["s_per_undulator_metre"],
        "kernel_ms_per_step": main_line["kernel_ms_per_step"],
        "roofline": main_line["roofline"],
        "schedule": main_line["schedule"],
        "setup_s": round(m["load_s"], 2),
    }
    if ref is not None:
        fl = _line(cfg, ref, a.steps, a.warmup, g4mi.FP64)
        out["vs_fp64"] = round(main_line["value"] / fl["value"], 3)
        out["fp64_form"] = fl
        e = W.form_errors(ref["recs"], m["recs"], ref["field"], ref["first"], m["field"], m["first"])
        out["errors_vs_fp64"] = dict({k: (round(v, 8) if isinstance(v, float) else v) for k, v in e.items()},
                                     span=f"all {e['steps']} steps of the lattice (timed steps "
                                          f"{m['timed0']}..{m['timed0'] + a.steps - 1} inside it)",
                                     bar=0.01)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg, work, a.cpu_steps)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
