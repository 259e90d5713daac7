"""Throughput of the Y/C-separation hot path on one MI355X.

One step = one batched decode through the C ABI (hvd_gpu_decode_fields_u16)
of 64 NTSC 910x263 uint16 fields already resident in HBM: IRE unpack ->
burst lock-in -> active crop + carrier -> FFT holographic init -> IRLS 4 x
CG 25 (cg_tol 1e-12, so every iteration runs) -> Y and chi.  `value` is
fields/s of that batched path.  Also reported: the plain path (one field per
call), the mixed 32 NTSC + 32 PAL batch (stop at relative residual 1e-6 or
40 CG iterations per stage), integer-SAD block matching, 512 mixed fields streamed in 8-field motion
windows (block matching + batched decode per window), 200 Adam steps on the
regularisation weights of 16 fields through the solve's adjoint, a live CG
roofline from HIP events recorded on the kernels' own stream, and the
hvd-core CPU baseline (its own lock-in + VariationalRefine compiled from source) timed on
this host, which doubles as a check that the GPU solve matches hvd-core on
identical inputs.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import re
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "hvd-cvbs-decoding_amd"))

HBM_PEAK_GBS = 8000.0
CG_BYTES_PER_PX_ITER = 64  # algorithmic: chi r+w, g r+w, d r+w, 4 weights read
BATCH = 64
CFG0 = dict(cg_iterations=100, irls_outer=4, fast=0, cg_tol=1e-12)
CFG2 = dict(cg_iterations=160, irls_outer=4, fast=0, cg_tol=1e-6)


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def make_fields(torch, synth, n, pal_every=0, seed0=1000):
    raws, geoms = [], []
    for k in range(n):
        pal = pal_every and (k % pal_every == pal_every - 1)
        raw = (synth.pal_field if pal else synth.ntsc_field)(seed0 + k)[0]
        raws.append(torch.from_numpy(raw.view(np.int16)).to("cuda:0"))
        geoms.append(synth.PAL if pal else synth.NTSC)
    return raws, geoms


def autotune(eng, geom, n):
    """Run the library's autotuner for n fields of this geometry's active shape
    (records the fastest strip height and solve-group size in the context)."""
    t0 = time.perf_counter()
    rows, gpx, ns = eng.autotune_solver(geom.lines, geom.width, geom.uses_vswitch, n)
    dt = time.perf_counter() - t0
    log(f"[bench] autotune {geom.lines}x{geom.width} x{n}: strip rows {rows}, group {gpx} px, "
        f"{ns:.4f} ns/px-iteration ({dt:.1f} s)")
    return {"fields": n, "strip_rows": rows, "group_pixels": gpx, "ns_per_pixel_iteration": round(ns, 5),
            "tuning_s": round(dt, 2)}


def traffic_from_profile():
    """HBM bytes per CG iteration (both launches) measured with rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE, as recorded in profiles/; None if absent."""
    p = ROOT / "profiles" / "cg_traffic_ntsc910x263_batch64.md"
    if not p.exists():
        return None
    m = re.search(r"traffic_bytes_per_cg_iteration:\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores() -> int:
    """Cores this process may use (its affinity mask), at most the 16 of the GPU's host share."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(torch, gpu, synth, eng, n_fields=6):
    """hvd-core's own BurstLockinPhase + VariationalRefine (compiled from the
    reference sources with OpenMP, oracle/_ref) on n config-0 fields.  The
    holographic init needs FFTW, which this image lacks, so the init is not
    part of the timed CPU work (it is ~4% of the field's algorithmic bytes);
    its chi0 comes from the GPU init and feeds both solvers, which makes the
    comparison of their outputs exact-input."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle
    if not oracle.REFLIB_OMP.exists():
        return None
    ref = oracle.Reference(openmp=True)
    g = synth.NTSC
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    fields, chi0s, ss = [], [], []
    for k in range(n_fields):
        raw = synth.ntsc_field(5000 + k)[0]
        f = synth.ire_from_raw(raw)
        fields.append(f)
    # GPU lock-in + init give the shared solver input
    dfs = [torch.from_numpy(f).to("cuda:0") for f in fields]
    th_d, _ = eng.burst_lockin_phase(dfs, [g] * n_fields)
    fal = g.first_active_field_line
    planes = []
    for f, th in zip(dfs, th_d):
        s = f[fal:fal + g.lines, g.active_video_start:g.active_video_end].contiguous()
        chi = torch.zeros((g.lines, g.width, 2), device="cuda:0")
        planes.append(eng.plane(s, chi, None, theta=th[fal:fal + g.lines].contiguous(), a0=g.active_video_start,
                                sample_rate=g.sample_rate))
    eng.holographic_init(planes)
    torch.cuda.synchronize()
    for p in planes:
        s, chi = p._keep[0], p._keep[1]
        ss.append(s.cpu().numpy())
        c = chi.cpu().numpy()
        chi0s.append(c[..., 0].astype(np.float64) + 1j * c[..., 1])
    cfg = oracle.Cfg.make(**CFG0)
    t0 = time.perf_counter()
    outs, cars = [], []
    for f, s, chi0 in zip(fields, ss, chi0s):
        th = ref.burst_lockin_phase(f, g)
        car = ref.make_carrier(th[fal:fal + g.lines], None, g)
        outs.append(ref.variational_refine(s, car, chi0, cfg))
        cars.append(car)
    dt = time.perf_counter() - t0
    # the same solve on the GPU from the same chi0 and the same hvd-core carrier
    errs = []
    for (y_ref, chi_ref), s, chi0, car in zip(outs, ss, chi0s, cars):
        ds = torch.from_numpy(s).to("cuda:0")
        dchi = torch.from_numpy(np.stack([chi0.real, chi0.imag], -1).astype(np.float32)).to("cuda:0")
        dcar = torch.from_numpy(np.stack([car.real, car.imag], -1).astype(np.float32)).to("cuda:0")
        dy = torch.zeros_like(ds)
        eng.variational_refine([eng.plane(ds, dchi, dy, carrier=dcar)], gpu.SolverParams.make(**CFG0))
        torch.cuda.synchronize()
        c = dchi.cpu().numpy()
        errs.append(max(float(np.abs(dy.cpu().numpy() - y_ref).max()),
                        float(np.abs(c[..., 0] + 1j * c[..., 1] - chi_ref).max())))
    # mode (ii): parallel_across_fields (sequence.cpp:1079-1085), every field
    # single-threaded, one field per core
    across = cpu_baseline_across_fields(ref, g, fields, ss, chi0s, cfg)
    one_at_a_time = {"value": round(n_fields / dt, 3), "unit": "fields/s", "cores": cores, "kind": "reference",
                     "sample": f"{n_fields} NTSC 910x263 fields: hvd-core BurstLockinPhase + VariationalRefine "
                               f"(IRLS 4 x CG 25), OpenMP {cores} threads, one field at a time; "
                               f"init excluded (FFTW absent)"}
    # the baseline is the better of hvd-core's two CPU modes; both are kept
    best = across if across and across["value"] > one_at_a_time["value"] else one_at_a_time
    return dict(best, mode="across_fields" if best is across else "one_field_at_a_time",
                gpu_vs_reference_max_abs_err_ire=float(max(errs)), one_field_at_a_time=one_at_a_time,
                across_fields=across, cpu_model=cpu_model(), nproc=os.cpu_count(), affinity_cores=host_cores())


def cpu_baseline_across_fields(ref, g, fields, ss, chi0s, cfg):
    """hvd-core's second CPU mode: fields decoded concurrently, each on one
    thread (OpenMP teams of one), as many at once as the host share has
    cores; the same fields repeated to fill them."""
    import ctypes
    from concurrent.futures import ThreadPoolExecutor
    try:
        gomp = ctypes.CDLL("libgomp.so.1")
    except OSError:
        return None
    workers = host_cores()
    fal = g.first_active_field_line
    jobs = [(fields[k % len(fields)], ss[k % len(ss)], chi0s[k % len(chi0s)]) for k in range(workers)]

    def one(job):
        gomp.omp_set_num_threads(1)
        f, s, chi0 = job
        th = ref.burst_lockin_phase(f, g)
        car = ref.make_carrier(th[fal:fal + g.lines], None, g)
        ref.variational_refine(s, car, chi0, cfg)

    with ThreadPoolExecutor(max_workers=workers) as ex:
        t0 = time.perf_counter()
        list(ex.map(one, jobs))
        dt = time.perf_counter() - t0
    return {"value": round(len(jobs) / dt, 3), "unit": "fields/s", "cores": workers, "kind": "reference",
            "sample": f"{len(jobs)} NTSC 910x263 fields: hvd-core BurstLockinPhase + VariationalRefine "
                      f"(IRLS 4 x CG 25), {workers} at once, each on one thread (parallel_across_fields); "
                      f"init excluded (FFTW absent)"}


def training_bench(torch, gpu, synth, eng, steps=200, n_fields=16):
    """Regularisation weights by gradient on synthetic ground truth: n_fields
    NTSC fields in FP32 (lock-in, carrier and holographic init on the GPU),
    then `steps` Adam steps on (log lambda_Y, log lambda_C), each a forward
    IRLS 4 x CG 25 solve of the whole batch and its adjoint (IRLS weights
    recomputed in the backward), loss = MSE of Y, U, V against the truth."""
    from hvd_amd import autograd as HA
    g = synth.NTSC
    s_l, car_l, chi_l, yt_l, ct_l = [], [], [], [], []
    for k in range(n_fields):
        raw, sc, _ = synth.ntsc_field(7000 + k)
        field = eng.sample_to_ire_u16(torch.from_numpy(raw.view(np.int16)).to("cuda:0"), 20480.0, 61440.0)
        th, _ = eng.burst_lockin_phase([field], [g])
        fal, a0 = g.first_active_field_line, g.active_video_start
        s = field[fal:fal + g.lines, a0:a0 + g.width].contiguous()
        car = eng.make_carrier(th[0][fal:fal + g.lines].contiguous(), None, g)
        chi0 = torch.empty((g.lines, g.width, 2), dtype=torch.float32, device="cuda:0")
        y0 = torch.empty_like(s)
        eng.holographic_init([gpu.Engine.plane(s, chi0, y=y0, carrier=car)])
        s_l.append(s), car_l.append(car), chi_l.append(chi0)
        yt_l.append(torch.from_numpy(sc.y.astype(np.float32)))
        ct_l.append(torch.from_numpy(np.stack([sc.chi.real, sc.chi.imag], -1).astype(np.float32)))
    S, CAR, CHI0 = torch.stack(s_l), torch.stack(car_l), torch.stack(chi_l)
    YT, CT = torch.stack(yt_l).to("cuda:0"), torch.stack(ct_l).to("cuda:0")
    ly = torch.tensor(0.0, device="cuda:0", requires_grad=True)
    lc = torch.tensor(0.0, device="cuda:0", requires_grad=True)
    opt = torch.optim.Adam([ly, lc], lr=0.02)
    npx = g.lines * g.width

    def step():
        opt.zero_grad()
        y, chi = HA.refine(S, CAR, CHI0, ly, lc, eng, irls_outer=4, cg_iterations=100, fast=0, cg_tol=1e-12,
                           adj_iterations=25, adj_tol=1e-6)
        loss = (((y - YT) ** 2).sum() + ((chi - CT) ** 2).sum()) / (npx * n_fields)
        loss.backward()
        opt.step()
        return float(loss.detach())

    first = step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    last = first
    for _ in range(steps - 1):
        last = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(steps - 1, 1)
    return {"fields": n_fields, "adam_steps": steps, "ms_per_step": round(dt * 1e3, 2),
            "fields_per_s_fwd_bwd": round(n_fields / dt, 1), "loss_first": round(first, 5),
            "loss_last": round(last, 5), "lambda_y": round(float(torch.exp(ly.detach())), 4),
            "lambda_c": round(float(torch.exp(lc.detach())), 4)}


def streaming_bench(torch, gpu, synth, eng, n_windows=64, window=8, unique=16):
    """512 mixed fields (3 NTSC : 1 PAL) in 8-field motion windows: per
    window, integer-SAD block matching of its 7 consecutive pairs and the
    batched field decode (lock-in, FFT init, IRLS 4 x CG <= 40 to relative
    residual 1e-6).  Noise 0.2 / 1 / 3 IRE, chroma up to 75 IRE, one window
    in five with fast motion (<= 12 px/field).  `unique` distinct windows are
    generated and streamed round-robin from HBM."""
    rng = np.random.default_rng(11)
    wins = []
    for k in range(unique):
        pal = k % 4 == 3
        gg = synth.PAL if pal else synth.NTSC
        fast = k % 5 == 4
        raws, _ = synth.motion_sequence(9000 + k, n_fields=window, g=gg, noise_ire=float([0.2, 1.0, 3.0][k % 3]),
                                        max_rect_v=12 if fast else 8, saturation=float(rng.uniform(30.0, 75.0)))
        d = [torch.from_numpy(r.view(np.int16)).to("cuda:0") for r in raws]
        ys = [torch.empty((gg.lines, gg.width), device="cuda:0") for _ in d]
        chis = [torch.empty((gg.lines, gg.width, 2), device="cuda:0") for _ in d]
        wins.append((d, [gg] * window, ys, chis))
    p = gpu.SolverParams.make(**CFG2)

    def run(w, plain=False):
        d, geoms, ys, chis = w
        eng.block_match_sad_u16(d[:-1], d[1:])
        if plain:
            for k in range(len(d)):
                eng.decode_fields_u16([d[k]], [geoms[k]], p, outputs=([ys[k]], [chis[k]]))
        else:
            eng.decode_fields_u16(d, geoms, p, outputs=(ys, chis))

    for w in wins[:2]:
        run(w)
    torch.cuda.synchronize()
    eng.set_profiling(True)
    t0 = time.perf_counter()
    for i in range(n_windows):
        run(wins[i % unique])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    cg_ms, _, pxit = eng.cg_timing()
    eng.set_profiling(False)
    t0 = time.perf_counter()
    nplain = 8
    for i in range(nplain):
        run(wins[i % unique], plain=True)
    torch.cuda.synchronize()
    dtp = time.perf_counter() - t0
    return {"fields": n_windows * window, "windows": n_windows, "fields_per_s": round(n_windows * window / dt, 1),
            "plain_fields_per_s": round(nplain * window / dtp, 1),
            "cg_algorithmic_gbs": round(CG_BYTES_PER_PX_ITER * pxit / (cg_ms * 1e-3) / 1e9, 1)}


def temporal_bench(torch, gpu, eng, reps=10):
    """VariationalRefine with two temporal neighbours (the 3-D solve) on an
    NTSC field's active grid, 263 x 758, the headline's solver settings
    (4 IRLS stages x 25 CG iterations, every iteration run).  Fields/s; the
    CG iterations' own time from HIP events around each stage's dir + step
    launches (the rest of the call, per-stage weights and the fixed set-up,
    reported apart); their algorithmic bytes: the 2-D 64 B per pixel plus,
    per neighbour, its composite, carrier and temporal weight read (16 B)."""
    h, w, nn = 263, 758, 2
    r, c = np.mgrid[0:h, 0:w].astype(np.float64)
    car = np.exp(1j * (0.5 * np.pi * c + 0.3 * r))
    y = 30 + 20 * np.sin(0.35 * (r + c))
    chi = 10 * np.sin(0.4 * (c - r)) - 6j * np.cos(0.3 * (c + r))
    s = y + (chi * car).real
    chi0 = chi + 3 * np.sin(0.9 * c) * (1 - 1j)

    def d(a):
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to("cuda:0")

    def cd(z):
        return d(np.stack([z.real, z.imag], -1))

    nbrs = []
    for j in range(nn):
        cw = car * np.exp(0.5j * np.pi * (j + 1))
        nbrs.append((d(y + (chi * cw).real + 0.3 * np.sin(1.7 * r + j)), cd(cw), d(np.full((h, w), 0.8))))
    ds, dcar, dchi0 = d(s), cd(car), cd(chi0)
    dchi, dy = torch.empty_like(dchi0), torch.empty_like(ds)
    p = gpu.SolverParams.make(**CFG0)
    pl = gpu.Engine.plane(ds, dchi, dy, carrier=dcar)

    def one(stats=False):
        dchi.copy_(dchi0)
        return eng.variational_refine_temporal(pl, nbrs, p, 0.5, 2.0, want_stats=stats)

    iters = int(one(True).sum())  # iterations the solve actually runs (it is deterministic)
    torch.cuda.synchronize()
    eng.set_profiling(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        one()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    cg_ms, launched, _ = eng.cg_timing()
    eng.set_profiling(False)
    us_it = 1e3 * cg_ms / max(launched, 1)
    bytes_it = (CG_BYTES_PER_PX_ITER + 16 * nn) * h * w
    # the same solve for a batch of 16 such fields in one call (the window
    # driver solves a colour of its Gauss-Seidel sweep as one batch)
    nb_f = 16
    planes, outs = [], []
    for _ in range(nb_f):
        c = torch.empty_like(dchi0)
        planes.append(gpu.Engine.plane(ds, c, torch.empty_like(ds), carrier=dcar))
        outs.append(c)

    def batch():
        for c in outs:
            c.copy_(dchi0)
        eng.variational_refine_temporal_batch(planes, [nbrs] * nb_f, p, 0.5, 2.0)

    batch()
    torch.cuda.synchronize()
    eng.set_profiling(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        batch()
    torch.cuda.synchronize()
    dtb = (time.perf_counter() - t0) / reps
    cg_ms_b, launched_b, _ = eng.cg_timing()
    eng.set_profiling(False)
    us_it_b = 1e3 * cg_ms_b / max(launched_b, 1)
    return {"fields_per_s": round(1.0 / dt, 1), "ms_per_field": round(1e3 * dt, 3),
            "us_per_cg_iteration": round(us_it, 2),
            "ms_per_field_outside_cg_iterations": round(1e3 * dt - cg_ms / reps, 3),
            "cg_algorithmic_gbs": round(bytes_it / (us_it * 1e-6) / 1e9, 1),
            "neighbours": nn, "cg_iterations": iters,
            "batch16": {"fields_per_s": round(nb_f / dtb, 1), "us_per_cg_iteration": round(us_it_b, 2),
                        "cg_algorithmic_gbs": round(nb_f * bytes_it / (us_it_b * 1e-6) / 1e9, 1),
                        "frac_of_hbm_peak": round(nb_f * bytes_it / (us_it_b * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)}}


WINDOW_FRAMES, WINDOW_CORE = 8, (1, 7)  # HvdConfig chunk_frames 6 + chunk_overlap 1 on each side


def window3d_frames(synth, seed=61):
    """A window of 8 NTSC frames of one scene (field-sequential int16 10-bit,
    0.5 IRE noise).  The block matching searches every candidate whatever the
    content, so a static scene costs what a moving one does."""
    fpd = synth.FRAME_NTSC
    return fpd, synth.frames_s16(fpd, WINDOW_FRAMES, seed=seed)


def window3d_bench(torch, gpu, synth, eng, reps=3):
    """DecodeFrameSequenceWindow with the 3-D window driver on 8 NTSC frames
    (16 fields; HvdConfig's chunk of 6 core frames plus one overlap frame on
    each side), temporal strength 1: noise calibration, two Gauss-Seidel
    passes in fast mode (pair motion, predicted + verified long offsets,
    trajectory snap, motion-compensated neighbours, coherence and odd-offset
    gates, SynthReference anchor + joint solve from pass 1), output-fidelity
    luma, weave and requantise.  HvdConfig's solver defaults (cg_iterations
    2) and cg_iterations 40.  Core frames/s per call, the frames already in
    HBM."""
    fpd, frames = window3d_frames(synth)
    dfr = [torch.from_numpy(f).to("cuda:0") for f in frames]
    fp = gpu.FrameParams(**{k: v for k, v in fpd.items() if k != "extra_sample_lines"})
    out = {}
    for cg in (2, 40):
        sp = gpu.SolverParams.make(cg_iterations=cg)
        bp = gpu.BridgeParams.make(temporal_strength=1.0)
        wp = gpu.WindowParams.make()
        d = {}
        eng.decode_frame_window_s16(dfr, WINDOW_CORE[0], WINDOW_CORE[1], fp, sp, bp, wparams=wp, diag=d)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            eng.decode_frame_window_s16(dfr, WINDOW_CORE[0], WINDOW_CORE[1], fp, sp, bp, wparams=wp)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        ncore = WINDOW_CORE[1] - WINDOW_CORE[0]
        out[f"cg{cg}"] = {"core_frames_per_s": round(ncore / dt, 2), "window_fields_per_s": round(2 * WINDOW_FRAMES / dt, 2),
                          "ms_per_window": round(1e3 * dt, 2), "resolved_strength": round(d["resolved_strength"], 3),
                          "sigma_ire": round(d["sigma_ire"], 3)}
    return out


def window3d_cpu_baseline(torch, gpu, synth, eng):
    """hvd-core's own DecodeFieldWindowWithInits (oracle/_ref, OpenMP) on the
    same 16 fields with HvdConfig's defaults (cg_iterations 2), temporal
    strength 1.  HolographicInit needs FFTW, absent here: the inits come
    from the library (hvd_gpu_holographic_init on hvd-core's float carriers)
    and the CPU time is the driver's alone."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle
    import bridge as OB
    if not oracle.REFLIB_OMP.exists():
        return None
    ref = oracle.Reference(openmp=True)
    fpd, frames = window3d_frames(synth)
    fp_o = OB.FrameParams(**fpd)
    g = OB.field_geometry(fp_o, synth.Geometry)
    fal = g.first_active_field_line
    fields, inits, ys, par = [], [], [], []
    for fr in frames:
        for k, fld in enumerate(OB.split_frame_fields(fr, fp_o)):
            th = ref.burst_lockin_phase(fld, g)
            car = ref.make_carrier(th[fal:fal + g.lines], None, g).astype(np.complex64)
            s = np.ascontiguousarray(fld[fal:fal + g.lines, g.active_video_start:g.active_video_end], np.float32)
            ds = torch.from_numpy(s).to("cuda:0")
            dchi = torch.zeros((g.lines, g.width, 2), device="cuda:0")
            dy = torch.zeros_like(ds)
            dcar = torch.from_numpy(np.stack([car.real, car.imag], -1).astype(np.float32)).to("cuda:0")
            eng.holographic_init([eng.plane(ds, dchi, dy, carrier=dcar, sample_rate=g.sample_rate)])
            c = dchi.cpu().numpy()
            fields.append((s, car))
            inits.append((c[..., 0] + 1j * c[..., 1]).astype(np.complex64))
            ys.append(dy.cpu().numpy())
            par.append(k)
    cfg = oracle.SeqCfg.make(cg_iterations=2, temporal_strength=1.0)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    ref.decode_window_with_inits(fields, par, inits, cfg, inits_y=ys)
    dt = time.perf_counter() - t0
    ncore = WINDOW_CORE[1] - WINDOW_CORE[0]
    return {"value": round(ncore / dt, 3), "unit": "core frames/s", "cores": cores, "kind": "reference",
            "sample": f"one window of {len(fields)} NTSC fields: hvd-core DecodeFieldWindowWithInits "
                      f"(HvdConfig defaults, strength 1), OpenMP {cores} threads; inits and lock-in excluded",
            "s_per_window": round(dt, 3)}


def plugin_path_bench(torch, gpu, synth, eng, reps=5):
    """DecodeFrameSequenceWindow through libhvd_core.so as the decode-orc
    plugin calls it per chunk (hvd_chroma_decoder_stage.cpp:700-760): host
    int16 frames in, YcFrameS16 (luma, chroma, U, V) out, host copies
    included, one HvdEngine across calls, HvdConfig defaults (cg_iterations
    2).  lib/hvd_window_bench runs it in a child process.  Two chunk shapes
    on the 3-D window's NTSC frames: the temporal switch off (6 frames, no
    overlap) and on (6 core frames + 1 overlap frame per side, strength 1).
    Beside each, the C ABI call (hvd_gpu_decode_frame_window_ex_s16) on the
    same frames already in HBM with the same settings."""
    import subprocess
    import tempfile
    exe = ROOT / "hvd-cvbs-decoding_amd" / "lib" / "hvd_window_bench"
    if not exe.exists():
        return None
    fpd, frames = window3d_frames(synth)
    kv = [f"{k}={','.join(str(e) for e in v) if k == 'extra_sample_lines' else v}" for k, v in fpd.items()]
    fp = gpu.FrameParams(**{k: v for k, v in fpd.items() if k != "extra_sample_lines"})
    dfr = [torch.from_numpy(f).to("cuda:0") for f in frames]
    out = {}
    with tempfile.TemporaryDirectory() as td:
        path = pathlib.Path(td) / "frames.s16"
        np.concatenate([np.asarray(f, np.int16).reshape(-1) for f in frames]).tofile(path)
        for name, nfr, (cb, ce), temporal in (("temporal_off_chunk6", 6, (0, 6), "off"),
                                              ("temporal_on_chunk6_overlap1", WINDOW_FRAMES, WINDOW_CORE, "1")):
            r = subprocess.run([str(exe), str(path), str(nfr), str(cb), str(ce), temporal, str(reps)] + kv,
                               capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError(f"hvd_window_bench failed ({r.returncode}): {r.stderr.strip()[-400:]}")
            cpp = json.loads(r.stdout.strip().splitlines()[-1])
            sp = gpu.SolverParams.make()
            bp = gpu.BridgeParams.make(temporal_strength=-1.0 if temporal == "off" else float(temporal))
            wp = gpu.WindowParams.make()
            eng.decode_frame_window_s16(dfr[:nfr], cb, ce, fp, sp, bp, wparams=wp)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                eng.decode_frame_window_s16(dfr[:nfr], cb, ce, fp, sp, bp, wparams=wp)
            torch.cuda.synchronize()
            abi_ms = 1e3 * (time.perf_counter() - t0) / reps
            out[name] = {"frames": nfr, "core_frames": ce - cb, "cpp_api_ms_per_call": cpp["ms_per_call"],
                         "cpp_api_ms_best_call": cpp["ms_best_call"], "cpp_api_ms_first_call": cpp["ms_first_call"],
                         "cpp_api_core_frames_per_s": cpp["core_frames_per_s"],
                         "c_abi_resident_ms_per_call": round(abi_ms, 3),
                         "cpp_over_c_abi": round(cpp["ms_per_call"] / abi_ms, 3)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--only-batched", action="store_true", help="time the headline workload only (profiling)")
    ap.add_argument("--no-autotune", action="store_true", help="use the built-in launch shapes (no autotuner)")
    ap.add_argument("--no-graphs", action="store_true", help="launch the CG iterations one by one (no HIP graphs)")
    ap.add_argument("--solver-tuning", default="", metavar="ROWS,GROUP_PX",
                    help="fix the solver's strip height and solve-group pixels instead of autotuning")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 512 MiB device-to-device copy and exit (reference point for PMC byte counters)")
    ap.add_argument("--only-line", choices=("window3d", "temporal"), default="",
                    help="run one secondary line alone and print it (profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU layout is described in "
              "DESIGN.md and is not measured.", file=sys.stderr)
        sys.exit(2)

    import torch
    from hvd_amd import gpu, synth

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    if args.calibrate_copy:
        n = 512 << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copied {n} bytes device to device")
        return
    eng = gpu.Engine(0)
    if args.only_line:
        line = window3d_bench(torch, gpu, synth, eng) if args.only_line == "window3d" else temporal_bench(torch, gpu, eng)
        eng.close()
        print(json.dumps({args.only_line: line}), flush=True)
        return
    g = synth.NTSC
    px_field = g.lines * g.width

    log(f"[bench] generating {BATCH} NTSC fields")
    raws, geoms = make_fields(torch, synth, BATCH)
    p0 = gpu.SolverParams.make(**CFG0)
    ys = [torch.empty((g.lines, g.width), device="cuda:0") for _ in range(BATCH)]
    chis = [torch.empty((g.lines, g.width, 2), device="cuda:0") for _ in range(BATCH)]

    def step():
        eng.decode_fields_u16(raws, geoms, p0, outputs=(ys, chis))

    eng.set_graphs(not args.no_graphs)
    tuned = []
    if args.solver_tuning:
        rows, gpx = (int(v) for v in args.solver_tuning.split(","))
        eng.set_solver_tuning(rows, gpx)
        tuned.append({"fixed": True, "strip_rows": rows, "group_pixels": gpx})
    elif not args.no_autotune:  # before any timing: the tuner's own runs are not part of a step
        tuned.append(autotune(eng, g, BATCH))
        if not args.only_batched:
            # the other lines' batches: mixed (32 + 32), training (16), streaming windows (8), plain (1)
            for geom, n in ((g, 32), (g, 16), (g, 8), (g, 1), (synth.PAL, 32), (synth.PAL, 8), (synth.PAL, 1)):
                tuned.append(autotune(eng, geom, n))

    log(f"[bench] warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    eng.set_profiling(True)
    log(f"[bench] timing {args.steps} batched steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    cg_ms, cg_iters, pxit = eng.cg_timing()
    eng.set_profiling(False)
    fields_s = BATCH * args.steps / dt
    ms_step = 1000.0 * dt / args.steps
    avg_pair_ms = cg_ms / cg_iters  # per CG iteration of one solve group
    px_per_pair = pxit / cg_iters  # pixels actually iterated per group-iteration
    achieved = CG_BYTES_PER_PX_ITER * px_per_pair / (avg_pair_ms * 1e-3) / 1e9
    log(f"[bench] batched: {fields_s:.1f} fields/s, {ms_step:.2f} ms/step; CG iteration {avg_pair_ms*1e3:.1f} us, "
        f"{achieved:.0f} GB/s algorithmic")
    traffic = traffic_from_profile()
    result = {
        "metric": "NTSC fields/s at <=1e-4 err, batched vs plain on 1 GPU; % of HBM peak",
        "value": round(fields_s, 2), "unit": "fields/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (f64 reductions)", "data": "synthetic",
        "config": {"workload": "ntsc_910x263_batched64_lockin_fftinit_irls4x25", "fields_per_step": BATCH,
                   "field": "NTSC 4fsc 910x263 uint16, active 760x242", "irls_outer": 4, "cg_inner": 25,
                   "cg_tol": 1e-12, "motion_estimation": False, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "CG iteration (one k_cg_iter launch per solve group; k_cg_finish once per IRLS stage)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "avg_us_per_iteration": round(avg_pair_ms * 1e3, 2),
                     "algorithmic_bytes_per_iteration": int(CG_BYTES_PER_PX_ITER * px_per_pair)},
        "solver_tuning": tuned if tuned else "built-in policy",
        "hip_graphs": not args.no_graphs,
    }
    if not args.only_batched:
        # plain path: one field per call
        nplain = 16
        for k in range(2):
            eng.decode_fields_u16([raws[k]], [g], p0, outputs=([ys[k]], [chis[k]]))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(nplain):
            eng.decode_fields_u16([raws[k]], [g], p0, outputs=([ys[k]], [chis[k]]))
        torch.cuda.synchronize()
        plain = nplain / (time.perf_counter() - t0)
        result["plain_fields_per_s"] = round(plain, 2)
        result["batched_over_plain"] = round(fields_s / plain, 3)
        log(f"[bench] plain: {plain:.1f} fields/s")
        # mixed batch: 32 NTSC + 32 PAL, relative residual 1e-6 or 40 iterations per stage
        mraws, mgeoms = make_fields(torch, synth, 64, pal_every=2, seed0=3000)
        mys = [torch.empty((gg.lines, gg.width), device="cuda:0") for gg in mgeoms]
        mchis = [torch.empty((gg.lines, gg.width, 2), device="cuda:0") for gg in mgeoms]
        p2 = gpu.SolverParams.make(**CFG2)
        eng.decode_fields_u16(mraws, mgeoms, p2, outputs=(mys, mchis))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            eng.decode_fields_u16(mraws, mgeoms, p2, outputs=(mys, mchis))
        torch.cuda.synchronize()
        result["mixed_32ntsc_32pal_fields_per_s"] = round(3 * 64 / (time.perf_counter() - t0), 2)
        log(f"[bench] mixed: {result['mixed_32ntsc_32pal_fields_per_s']} fields/s")
        # integer-SAD block matching: 8 consecutive fields -> 7 pairs, 16x8 blocks, +/-16 x +/-4
        seq, _ = synth.motion_sequence(77, n_fields=8)
        dseq = [torch.from_numpy(r.view(np.int16)).to("cuda:0") for r in seq]
        eng.block_match_sad_u16(dseq[:-1], dseq[1:])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(20):
            eng.block_match_sad_u16(dseq[:-1], dseq[1:])
        torch.cuda.synchronize()
        sad_pairs = 20 * 7 / (time.perf_counter() - t0)
        result["sad_bm_field_pairs_per_s"] = round(sad_pairs, 1)
        result["sad_bm_gops"] = round(sad_pairs * 910 * 263 * 33 * 9 * 3 / 1e9, 1)  # sub, abs, add per candidate px
        log(f"[bench] SAD BM: {sad_pairs:.0f} field pairs/s")
    if not args.only_batched:
        log("[bench] streaming: 512 mixed fields in 8-field windows")
        result["streaming_512_mixed_windows8"] = streaming_bench(torch, gpu, synth, eng)
        log(f"[bench] streaming: {result['streaming_512_mixed_windows8']}")
        log("[bench] training: 16 NTSC fields, 200 Adam steps on log lambda")
        result["training_16ntsc_adam200"] = training_bench(torch, gpu, synth, eng)
        log(f"[bench] training: {result['training_16ntsc_adam200']}")
    if not args.only_batched:
        log("[bench] 3-D window: DecodeFrameSequenceWindow, 8 NTSC frames")
        result["frame_window_3d_ntsc8"] = window3d_bench(torch, gpu, synth, eng)
        log(f"[bench] 3-D window: {result['frame_window_3d_ntsc8']}")
        log("[bench] plugin path: DecodeFrameSequenceWindow through libhvd_core.so on host frames")
        result["plugin_path_ntsc"] = plugin_path_bench(torch, gpu, synth, eng)
        log(f"[bench] plugin path: {result['plugin_path_ntsc']}")
        log("[bench] temporal: NTSC-sized field with 2 temporal neighbours")
        result["temporal_2nbr_ntsc758x263"] = temporal_bench(torch, gpu, eng)
        log(f"[bench] temporal: {result['temporal_2nbr_ntsc758x263']}")
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (hvd-core reference build)")
        cb = cpu_baseline(torch, gpu, synth, eng)
        result["cpu_baseline"] = cb
        if not args.only_batched:
            log("[bench] CPU baseline of the 3-D window (hvd-core reference build)")
            result["frame_window_3d_ntsc8"]["cpu_baseline"] = window3d_cpu_baseline(torch, gpu, synth, eng)
        if cb:
            log(f"[bench] CPU: {cb['value']} fields/s on {cb['cores']} threads; GPU vs hvd-core "
                f"max |d| {cb['gpu_vs_reference_max_abs_err_ire']:.2e} IRE")
    else:
        result["cpu_baseline"] = None
    eng.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
