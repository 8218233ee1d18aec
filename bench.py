"""Benchmark of the plain FP32 CLOUDSC kernel on one MI355X.

Workload: plain CLOUDSC with FP32 fields, NLEV = 137, 655,360 columns (a
TCo399-sized grid) in IFS blocked layout, seeded synthetic atmospheres
generated on the device.  One step = one launch of cloudsc_blocked_mixed_c
(FP32 state and arithmetic, FP64 water budget: within 1e-5 of the FP64 build;
--variant sp times cloudsc_blocked_sp_c, all FP32, ~1e-2 from FP64) over all
columns, preceded by the device copy that restores PLUDE (CLOUDSC rescales it
in place, so every step starts from the same inputs).  Inputs are resident in
HBM before the timed region starts.  The other variant is timed the same way
and reported under its name.

Reported besides the throughput:
  roofline      HBM-bound: algorithmic bytes per launch (4*(49*NLEV+20) per
                column) / average kernel duration from HIP events on the
                kernel's stream; traffic from the rocprofv3 FETCH_SIZE /
                WRITE_SIZE summary committed under profiles/ (null if absent)
  err_vs_fp64   per-field relative RMS of the headline result against the FP64
                CPU oracle on the first 2048 columns
  cpu_baseline  the CPU oracle (C++/OpenMP restatement, FP32) on a sample of
                the same workload, OpenMP over NPROMA=32 blocks, SCHEDULE(DYNAMIC,1)
  hybrid        the learned surrogate: training throughput, inference and the
                hybrid (surrogate + rerun of gated columns) against plain FP32
  satadj        CLOUD_SATADJ (the saturation adjustment ahead of CLOUDSC), FP32
                and mixed, on the same grid: kernel time, algorithmic GB/s and
                fraction of HBM peak, error against the FP64 oracle, CPU oracle
  budget        CLOUDSC with the LCLDBUD budget diagnostics into PEXTRA (68
                fields) against the plain launch on the same grid
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NGPTOT, KLEV = 655360, 137
METRIC = "columns/sec at NLEV=137 (plain FP32 vs surrogate-hybrid) + % HBM peak, err vs FP64"


def bytes_per_column(klev, word=4):
    # 26*KLEV+5 words read, 23*KLEV+15 written on the default path
    return word * (49 * klev + 20)


def read_traffic(nproma, name="cloudsc_sp_l137_655360_traffic.md"):
    """HBM bytes per dispatch measured by rocprofv3 (profiles/<name>)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    txt = open(path).read()
    m = re.search(r"nproma\s*=\s*(\d+).*?hbm_bytes_per_dispatch\s*=\s*([0-9.eE+]+)", txt, re.S)
    if not m or int(m.group(1)) != nproma:
        return None
    return float(m.group(2))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _reblock(a, nproma_from, nproma_to):
    """A blocked array (NGPBLKS, ..., NPROMA) re-cut into blocks of `nproma_to` columns."""
    import numpy as np

    nb = a.shape[0]
    cols = np.moveaxis(a, -1, 1).reshape((nb * nproma_from,) + a.shape[1:-1])
    n = cols.shape[0] // nproma_to * nproma_to
    return np.ascontiguousarray(np.moveaxis(cols[:n].reshape((n // nproma_to, nproma_to) + a.shape[1:-1]), 1, -1))


def cpu_child(workdir):
    """The CPU baseline proper, in a process of its own (no torch, no HIP): OpenMP binds its threads to cores
    (OMP_PROC_BIND=close, OMP_PLACES=cores from the parent) over the whole CPU share, which a process whose
    OpenMP runtime already started could not do.  Reads the inputs and settings the parent wrote to `workdir`,
    prints one JSON line."""
    affinity = len(os.sched_getaffinity(0))  # before the OpenMP runtime binds this thread to its first core
    import numpy as np

    import oracle
    from openifs_amd import cloudsc as C

    with open(os.path.join(workdir, "settings.json")) as fh:
        cfg = json.load(fh)
    with open(os.path.join(workdir, "params.bin"), "rb") as fh:
        params = C.CloudscParams.from_buffer_copy(fh.read())
    base = {n: np.load(os.path.join(workdir, n + ".npy")) for n in C.INPUT_NAMES}
    nthreads, nproma0 = cfg["threads"], cfg["nproma"]

    def arrays(nproma, ncols):
        arr = {n: _reblock(v, nproma0, nproma)[: ncols // nproma] for n, v in base.items()}
        for name, kind, role in C.BLOCKED_ARRAYS:
            if role == "o":
                arr[name] = np.empty(C.blocked_shape(kind, nproma, KLEV, ncols // nproma), dtype=np.float32)
        return arr

    def run(arr, nproma, ncols):
        oracle.cloudsc_blocked("sp", params, ncols, nproma, KLEV, 900.0, arr, nthreads=nthreads)

    ncols = cfg["sample_cols"]
    arr = arrays(nproma0, ncols)
    plude0 = arr["plude"].copy()
    run(arr, nproma0, ncols)  # warm-up
    times = []
    while (sum(times) < cfg["min_seconds"] or len(times) < 5) and len(times) < 200:
        arr["plude"][...] = plude0
        t0 = time.perf_counter()
        run(arr, nproma0, ncols)
        times.append(time.perf_counter() - t0)
    per_nproma = {nproma0: ncols / sorted(times)[len(times) // 2]}
    for npr in cfg["sweep_npromas"]:
        n = cfg["sweep_cols"] // npr * npr
        a = arrays(npr, n)
        p0 = a["plude"].copy()
        run(a, npr, n)
        ts = []
        for _ in range(3):
            a["plude"][...] = p0
            t0 = time.perf_counter()
            run(a, npr, n)
            ts.append(time.perf_counter() - t0)
        per_nproma[npr] = n / sorted(ts)[1]
    print(json.dumps({"per_nproma": {str(k): v for k, v in per_nproma.items()}, "seconds": sum(times),
                      "passes": len(times), "affinity": affinity}))


def cpu_baseline(params, sample_cols, nthreads, min_seconds, npromas=(8, 16, 64, 128), sweep_cols=16384):
    """FP32 CPU oracle over `sample_cols` columns of the workload in NPROMA = 32 blocks, in a child process with
    OMP_PROC_BIND=close, OMP_PLACES=cores and `nthreads` threads: one warm-up pass, then passes (PLUDE restored
    before each) until `min_seconds` have elapsed and at least 5 passes ran; then `sweep_cols` columns per other
    NPROMA (median of 3).  Returns (columns/s of the median NPROMA-32 pass, seconds, passes, {nproma: columns/s},
    CPUs the child could run on)."""
    import shutil
    import subprocess
    import tempfile

    import numpy as np
    import torch

    from openifs_amd import cloudsc as C

    nproma = 32
    f = C.BlockedFields(sample_cols, nproma, KLEV, "sp", "cuda", names=set(C.INPUT_NAMES))
    C.synth_device(f, params, 1234)
    workdir = tempfile.mkdtemp(prefix="cloudsc_cpu_", dir="/dev/shm" if os.path.isdir("/dev/shm") else None)
    try:
        for n in C.INPUT_NAMES:
            np.save(os.path.join(workdir, n + ".npy"), f[n].cpu().numpy())
        del f
        torch.cuda.empty_cache()
        with open(os.path.join(workdir, "params.bin"), "wb") as fh:
            fh.write(bytes(params))
        with open(os.path.join(workdir, "settings.json"), "w") as fh:
            json.dump({"threads": nthreads, "nproma": nproma, "sample_cols": sample_cols,
                       "min_seconds": min_seconds, "sweep_npromas": list(npromas),
                       "sweep_cols": min(sweep_cols, sample_cols)}, fh)
        env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores", OMP_NUM_THREADS=str(nthreads))
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", workdir], env=env,
                           stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=1200)
        if r.returncode != 0:
            raise RuntimeError(f"CPU baseline child failed ({r.returncode}): {r.stderr[-2000:]}")
        res = json.loads(r.stdout.strip().splitlines()[-1])
    finally:
        shutil.rmtree(workdir, ignore_errors=True)
    per_nproma = {int(k): v for k, v in res["per_nproma"].items()}
    return per_nproma[nproma], res["seconds"], res["passes"], per_nproma, res["affinity"]


def mlp_flops_per_column(din, dout):
    """Forward and training (forward + backward) FLOPs per column of the padded MLP."""
    fwd = sum(2 * a * b for a, b in zip(din, dout))
    bwd = sum(4 * a * b for a, b in zip(din, dout)) - 2 * din[0] * dout[0]  # no input gradient of layer 0
    return fwd, fwd + bwd


# surrogate training recipe of the benchmark (and of the full-size training test)
TRAIN = {"lr": 1e-3, "epochs": 2, "flux_weight": 4.0, "surface_weight": 50.0}


def surrogate_bench(params, train_cols, hybrid_cols, rerun_fraction, gate_fraction, infer_batch):
    """Trains the surrogate (fresh seeded columns, the plain FP32 kernel as teacher) and times the hybrid
    against the plain kernel on the same `hybrid_cols` columns."""
    import numpy as np
    import torch

    from openifs_amd import cloudsc as C
    from openifs_amd import surrogate as S

    dev = "cuda:0"
    sur = S.Surrogate(KLEV, hidden=1024, hidden_layers=3, max_batch=infer_batch, seed=7)
    fwd_flop, train_flop = mlp_flops_per_column(sur.mlp.din, sur.mlp.dout)

    # training: batches of 4096 columns, Adam, two epochs over `train_cols` columns, cosine learning rate from 1e-3,
    # precipitation-flux targets weighted x4 and the surface flux x200 (x4 as a flux, x50 on top)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = S.train_surrogate(sur, params, train_cols, chunk=262144, lr=TRAIN["lr"], seed0=5000, batch=4096,
                               device=dev, epochs=TRAIN["epochs"], cosine=True, flux_weight=TRAIN["flux_weight"],
                               surface_weight=TRAIN["surface_weight"])
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    trained_cols = len(losses) * 262144

    # hybrid workload: plain reference first (also its time), then the hybrid on the same inputs
    f = C.BlockedFields(hybrid_cols, 128, KLEV, "sp", dev)
    C.synth_device(f, params, 31337)
    plude0 = f["plude"].clone()
    scores = sur.gate_scores(f, params)
    thr = float(torch.quantile(scores[: min(hybrid_cols, 1 << 24)].float().cpu(), 1.0 - gate_fraction)) \
        if gate_fraction > 0 else float("inf")
    del scores
    stream = torch.cuda.current_stream()

    def timed(run, reps=3):
        run()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            f["plude"].copy_(plude0)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            r = run()
            e1.record(stream)
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return float(np.median(ts)), r

    plain_ms, _ = timed(lambda: C.cloudsc_blocked(f, params, 900.0, stream=stream.cuda_stream))
    sp_ref = (f["pfplsl"][:, -1, :].double() + f["pfplsn"][:, -1, :].double()).clone()
    tq_ref = f["ptendency_loc_q"].clone()
    tt_ref = f["ptendency_loc_t"].clone()
    net_ms, _ = timed(lambda: sur.predict(f, params))
    sp_net = f["pfplsl"][:, -1, :].double() + f["pfplsn"][:, -1, :].double()
    err_net = float(torch.linalg.vector_norm(sp_net - sp_ref) / torch.linalg.vector_norm(sp_ref))
    hyb_ms, n_plain = timed(lambda: sur.hybrid(f, params, rerun_fraction=rerun_fraction, rerun_seed=1,
                                                gate_threshold=thr))
    sp_h = f["pfplsl"][:, -1, :].double() + f["pfplsn"][:, -1, :].double()

    def rel(a, b):
        return float(torch.linalg.vector_norm((a.double() - b.double()).flatten()) /
                     torch.linalg.vector_norm(b.double().flatten()))

    res = {
        "workload": f"surrogate-hybrid CLOUDSC, NLEV={KLEV}, {hybrid_cols} columns, FP32 state, BF16 MLP "
                    f"1918-1024-1024-1024-822 (padded {sur.n_in}-...-{sur.n_out})",
        "columns_per_s": hybrid_cols / (hyb_ms / 1e3),
        "hybrid_ms": hyb_ms,
        "plain_ms_same_inputs": plain_ms,
        "plain_columns_per_s_same_inputs": hybrid_cols / (plain_ms / 1e3),
        "speedup_vs_plain": plain_ms / hyb_ms,
        "network_only_ms": net_ms,
        "network_tflops": fwd_flop * hybrid_cols / (net_ms / 1e3) / 1e12,
        "network_mfma_frac": fwd_flop * hybrid_cols / (net_ms / 1e3) / 2.5e15,
        "plain_columns": n_plain,
        "plain_fraction": n_plain / hybrid_cols,
        "rerun_fraction": rerun_fraction,
        "gate_threshold": thr,
        "surface_precip_relrms_vs_plain": rel(sp_h, sp_ref),
        "surface_precip_relrms_network_only": err_net,
        "tendency_t_relrms_vs_plain": rel(f["ptendency_loc_t"], tt_ref),
        "tendency_q_relrms_vs_plain": rel(f["ptendency_loc_q"], tq_ref),
        "training": {"columns": trained_cols, "distinct_columns": trained_cols // TRAIN["epochs"],
                     "seconds": t_train, "columns_per_s": trained_cols / t_train,
                     "tflops": train_flop * trained_cols / t_train / 1e12, "batch": 4096, **TRAIN,
                     "schedule": "cosine to lr/50", "final_loss": losses[-1] if losses else None,
                     "first_loss": losses[0] if losses else None},
    }
    del f, plude0, sp_ref, tq_ref, tt_ref
    sur.close()
    torch.cuda.empty_cache()
    return res


def satadj_bench(params, nproma, steps, warmup, cpu_threads, cpu_seconds):
    """CLOUD_SATADJ (the saturation adjustment ahead of CLOUDSC), FP32 and mixed, on the same 655,360 x L137
    workload: kernel time from HIP events on the launch stream, algorithmic bytes, accuracy of the first 2048
    columns against the FP64 oracle, and the FP32 oracle on 16 CPU threads over a sample."""
    import numpy as np
    import torch

    import oracle
    from openifs_amd import cloudsc as C
    from openifs_amd import satadj as S

    res = {}
    bpc = S.algorithmic_bytes(KLEV, params.ncldtop)
    ins = [n for n, k, r in S.SATADJ_BLOCKED if r == "i"]
    outs = [n for n, k, r in S.SATADJ_BLOCKED if r == "o"]
    for prec in ("sp", "mixed"):
        f = S.SatadjFields(NGPTOT, nproma, KLEV, prec, "cuda:0")
        S.synth_device(f, params, 20261203)
        stream = torch.cuda.current_stream()
        for _ in range(warmup):
            S.satadj_blocked(f, params, 900.0, stream=stream.cuda_stream)
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        for a, b in ev:
            a.record(stream)
            S.satadj_blocked(f, params, 900.0, stream=stream.cuda_stream)
            b.record(stream)
        torch.cuda.synchronize()
        ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
        nb = 2048 // nproma
        inp = {n: f[n][:nb].cpu().numpy().astype(np.float64) for n in ins}
        got = {n: f[n][:nb].cpu().numpy().astype(np.float64) for n in outs}
        cpu_inp = {n: f[n][:nb].cpu().numpy().copy() for n in ins} if prec == "sp" else None
        del f
        torch.cuda.empty_cache()
        ref = dict(inp)
        for n, kind, role in S.SATADJ_BLOCKED:
            if role == "o":
                ref[n] = np.full(C.blocked_shape(kind, nproma, KLEV, nb), np.nan)
        oracle.satadj_blocked("dp", params, nb * nproma, nproma, KLEV, 900.0, ref, nthreads=cpu_threads)
        err = {}
        for n in outs:
            a, r = got[n], ref[n]
            if n == "ptendency_loc_cld":
                a, r = a[:, :2], r[:, :2]
            m = ~np.isnan(r)
            err[n] = float(np.sqrt(np.sum((a[m] - r[m]) ** 2) / max(np.sum(r[m] ** 2), 1e-300)))
        gbs = bpc * NGPTOT / (ms / 1e3) / 1e9
        res[prec] = {"kernel_ms": ms, "columns_per_s": NGPTOT / (ms / 1e3),
                     "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": gbs / HBM_PEAK_GBS,
                                  "traffic": read_traffic(nproma, "satadj_sp_l137_655360_traffic.md")},
                     "err_vs_fp64": err}
        if cpu_inp is not None and cpu_seconds > 0:
            # FP32 oracle over the 2048 sampled columns, repeated for about cpu_seconds
            arr = dict(cpu_inp)
            for n, kind, role in S.SATADJ_BLOCKED:
                if role == "o":
                    arr[n] = np.zeros(C.blocked_shape(kind, nproma, KLEV, nb), dtype=np.float32)
            passes, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < cpu_seconds:
                oracle.satadj_blocked("sp", params, nb * nproma, nproma, KLEV, 900.0, arr, nthreads=cpu_threads)
                passes += 1
            dt = time.perf_counter() - t0
            res["cpu_baseline"] = {"value": passes * nb * nproma / dt, "unit": "columns/s", "cores": cpu_threads,
                                   "kind": "port", "sample": f"{nb * nproma} columns x {passes} passes ({dt:.1f} s), "
                                                             "FP32 oracle, OpenMP over NPROMA blocks"}
    res["algorithmic_bytes_per_column"] = bpc
    res["workload"] = f"CLOUD_SATADJ, NLEV={KLEV}, {NGPTOT} columns, NPROMA={nproma}, dt=900 s"
    res["fused"] = fused_bench(params, nproma, steps)
    return res


def fused_bench(params, nproma, steps):
    """The CALLPAR sequence CLOUD_SATADJ -> TENDENCY_CML increment -> TENDENCY_SATADJ copy -> CLOUDSC in FP32:
    separate launches (the adjustment kernel, the increment and copy as device element-wise operations, the CLOUDSC
    kernel) against the single fused launch.  Inputs (PLUDE, the CML tendencies) are restored outside the timed
    events before every step."""
    import numpy as np
    import torch

    from openifs_amd import cloudsc as C
    from openifs_amd import satadj as S

    f = S.FusedFields(NGPTOT, nproma, KLEV, "sp", "cuda:0")
    S.synth_device(f, params, 20261204)
    keep = {n: f[n].clone() for n in ("plude", "ptendency_cml_t", "ptendency_cml_q", "ptendency_cml_a",
                                      "ptendency_cml_cld")}
    stream = torch.cuda.current_stream()

    def restore():
        for n, x in keep.items():
            f[n].copy_(x, non_blocking=True)

    def unfused():
        S.satadj_blocked(f, params, 900.0, stream=stream.cuda_stream)
        f["ptendency_satadj_t"].copy_(f["ptendency_loc_t"])
        f["ptendency_satadj_q"].copy_(f["ptendency_loc_q"])
        f["ptendency_satadj_a"].copy_(f["ptendency_loc_a"])
        f["ptendency_satadj_cld"][:, :2].copy_(f["ptendency_loc_cld"][:, :2])
        f["ptendency_cml_t"].add_(f["ptendency_satadj_t"])
        f["ptendency_cml_q"].add_(f["ptendency_satadj_q"])
        f["ptendency_cml_a"].add_(f["ptendency_satadj_a"])
        f["ptendency_cml_cld"][:, :2].add_(f["ptendency_satadj_cld"][:, :2])
        C.cloudsc_blocked(f, params, 900.0, stream=stream.cuda_stream)

    def fused():
        S.satadj_cloudsc_blocked(f, params, 900.0, stream=stream.cuda_stream)

    out = {}
    for name, run in (("unfused_ms", unfused), ("fused_ms", fused)):
        ts = []
        for i in range(steps + 2):
            restore()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            run()
            b.record(stream)
            torch.cuda.synchronize()
            if i >= 2:
                ts.append(a.elapsed_time(b))
        out[name] = float(np.mean(ts))
    bpc = S.fused_bytes_per_column(KLEV, params.ncldtop)
    gbs = bpc * NGPTOT / (out["fused_ms"] / 1e3) / 1e9
    out.update({"speedup": out["unfused_ms"] / out["fused_ms"], "columns_per_s": NGPTOT / (out["fused_ms"] / 1e3),
                "algorithmic_bytes_per_column": bpc,
                "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": gbs / HBM_PEAK_GBS, "traffic": None},
                "workload": f"CLOUD_SATADJ + CML increment + CLOUDSC FP32, NLEV={KLEV}, {NGPTOT} columns, "
                            f"NPROMA={nproma}"})
    del f, keep
    torch.cuda.empty_cache()
    return out


def budget_bytes_per_column(klev, ncldtop, word=4):
    """Algorithmic HBM bytes per column of the budget launch (C, L, I, T and vertical-integral budgets): the plain
    CLOUDSC bytes, the seven budget inputs on the levels the scheme solves and KPBLTYPE/KCTOP, and one read and
    one write of every PEXTRA word the budgets touch: 52 fields on the solved levels (C 7, L 17, I 15, T 13) and
    57 slots of the vertical-integral field."""
    nlev = klev - ncldtop + 1
    return bytes_per_column(klev, word) + word * (7 * nlev + 2 + 2 * (52 * nlev + 57))


def budget_bench(params, nproma, steps):
    """The budget-instrumented CLOUDSC launch (LCLDBUDC/L/I/T and LCLDBUD_VERTINT into PEXTRA, 68 fields) against
    the plain launch, FP32, on the same 655,360 x L137 workload; PLUDE restored outside the timed events."""
    import numpy as np
    import torch

    from openifs_amd import cloudsc as C

    f = C.BlockedFields(NGPTOT, nproma, KLEV, "sp", "cuda:0")
    C.synth_device(f, params, 20261205)
    plude0 = f["plude"].clone()
    flags = C.BUD_C | C.BUD_L | C.BUD_I | C.BUD_T | C.BUD_VERTINT
    q = type(params).from_buffer_copy(params)
    q.lcldbud = flags
    g = torch.Generator(device="cuda:0")
    g.manual_seed(5)
    sh2, sh1 = (f.ngpblks, KLEV, nproma), (f.ngpblks, nproma)
    bud = {n: torch.randn(sh2, generator=g, device="cuda:0") * 1e-5 for n, k in C.BUDGET_ARRAYS if k == "2d"}
    bud["kpbltype"] = torch.randint(0, 4, sh1, generator=g, device="cuda:0", dtype=torch.int32)
    bud["kctop"] = torch.randint(1, KLEV + 1, sh1, generator=g, device="cuda:0", dtype=torch.int32)
    bud["pextra"] = torch.zeros(C.extra_shape(nproma, KLEV, C.budget_fields(flags), f.ngpblks), device="cuda:0")
    stream = torch.cuda.current_stream()
    out = {}
    for name, run in (("plain_ms", lambda: C.cloudsc_blocked(f, params, 900.0, stream=stream.cuda_stream)),
                      ("budget_ms", lambda: C.cloudsc_blocked_diag(f, bud, q, 900.0, stream=stream.cuda_stream))):
        ts = []
        for i in range(steps + 2):
            f["plude"].copy_(plude0, non_blocking=True)
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            run()
            b.record(stream)
            torch.cuda.synchronize()
            if i >= 2:
                ts.append(a.elapsed_time(b))
        out[name] = float(np.mean(ts))
    bpc = budget_bytes_per_column(KLEV, params.ncldtop)
    gbs = bpc * NGPTOT / (out["budget_ms"] / 1e3) / 1e9
    out.update({"overhead": out["budget_ms"] / out["plain_ms"], "columns_per_s": NGPTOT / (out["budget_ms"] / 1e3),
                "algorithmic_bytes_per_column": bpc,
                "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": gbs / HBM_PEAK_GBS, "traffic": None},
                "workload": f"CLOUDSC FP32 with the C, L, I, T and vertical-integral budgets (KFLDX = 68), "
                            f"NLEV={KLEV}, {NGPTOT} columns, NPROMA={nproma}"})
    del f, bud, plude0
    torch.cuda.empty_cache()
    return out


def cloudsc2_bytes_per_column(klev, word=4):
    """Algorithmic HBM bytes per column of CLOUDSC2: PAPHP1 and four flux outputs on KLEV+1 half levels, 14 inputs
    and 6 outputs on KLEV levels, each read or written once."""
    return word * (20 * klev + 5 * (klev + 1))


def layer_bytes_per_column(klev, word=4):
    """Algorithmic HBM bytes per column of CLOUD_S_LAYER (LEPCLD2) with every optional output: PRS1 and 14 level
    inputs read, 10 level outputs (7 tendencies, PQSAT, PCLC, PCOVPTOT), 4 half-level fluxes and PPRECTYPE/PFZRA
    written."""
    return word * (14 * klev + (klev + 1) + 10 * klev + 4 * (klev + 1) + 2)


def layer_bench(params, nproma, steps, lepcld2=1):
    """CLOUD_S_LAYER (SATUR + CLOUDSC2 or CLOUDST + wrap-up, one launch) in FP32 on 655,360 x L137, every optional
    output written; kernel time from HIP events on the launch stream."""
    import numpy as np
    import torch

    from openifs_amd import cloudsc2 as C2

    q2 = C2.default_params2(params)
    q2.lepcld2 = lepcld2
    src = C2.synth_device(params, 20261207, NGPTOT, nproma, KLEV, "sp", "cuda:0")
    f = {lay: src[c2] for lay, c2 in C2.LAYER_FROM_CLOUDSC2.items()}
    f["pa"] = torch.rand(src["pqm1"].shape, device="cuda:0")
    ngpblks = (NGPTOT + nproma - 1) // nproma
    for name, kind, role, opt in C2.LAYER_ARRAYS:
        if role == "o":
            f[name] = torch.empty(C2.layer_shape(kind, nproma, KLEV, ngpblks), device="cuda:0")
    ceta = torch.from_numpy(C2.reference_eta(KLEV)).to("cuda:0", torch.float32)
    stream = torch.cuda.current_stream()
    run = lambda: C2.cloud_s_layer_blocked(f, params, q2, ceta, NGPTOT, nproma, KLEV, 900.0, "sp",
                                           stream=stream.cuda_stream)
    for _ in range(2):
        run()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for a, b in ev:
        a.record(stream)
        run()
        b.record(stream)
    torch.cuda.synchronize()
    ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    del f, src
    torch.cuda.empty_cache()
    # CLOUDST reads neither PMFU/PMFD nor the condensate tendencies, and also writes PQLWP/PQIWP
    bpc = layer_bytes_per_column(KLEV) + (0 if lepcld2 else 4 * (2 * KLEV - 4 * KLEV))
    gbs = bpc * NGPTOT / (ms / 1e3) / 1e9
    return {"kernel_ms": ms, "columns_per_s": NGPTOT / (ms / 1e3), "algorithmic_bytes_per_column": bpc,
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS,
                         "traffic": read_traffic(nproma, f"cloud_s_layer_{'cloudsc2' if lepcld2 else 'cloudst'}"
                                                         "_sp_l137_655360_traffic.md")},
            "workload": f"CLOUD_S_LAYER FP32 (SATUR, {'CLOUDSC2' if lepcld2 else 'CLOUDST'}, wrap-up), NLEV={KLEV}, "
                        f"{NGPTOT} columns, NPROMA={nproma}"}


def cloudsc2_bench(params, nproma, steps, cpu_threads, cpu_seconds):
    """CLOUDSC2 (the simplified scheme of the linearised physics) in FP32 on 655,360 x L137: kernel time from HIP
    events on the launch stream, algorithmic bytes, accuracy of the first 2048 columns against the FP64 oracle, and
    the FP32 oracle on the host cores over those columns."""
    import numpy as np
    import torch

    import oracle
    from openifs_amd import cloudsc2 as C2

    q2 = C2.default_params2(params)
    f = C2.synth_device(params, 20261206, NGPTOT, nproma, KLEV, "sp", "cuda:0")
    ceta = torch.from_numpy(C2.reference_eta(KLEV)).to("cuda:0", torch.float32)
    stream = torch.cuda.current_stream()
    run = lambda: C2.cloudsc2_blocked(f, params, q2, ceta, NGPTOT, nproma, KLEV, 900.0, "sp",
                                      stream=stream.cuda_stream)
    for _ in range(2):
        run()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for a, b in ev:
        a.record(stream)
        run()
        b.record(stream)
    torch.cuda.synchronize()
    ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    nb = 2048 // nproma
    inp = {n: f[n][:nb].cpu().numpy().copy() for n in C2.INPUT_NAMES}
    got = {n: f[n][:nb].double().cpu().numpy() for n in C2.OUTPUT_NAMES}
    del f
    torch.cuda.empty_cache()
    ref = {n: v.astype(np.float64) for n, v in inp.items()}
    for n in C2.OUTPUT_NAMES:
        ref[n] = np.full(got[n].shape, np.nan)
    oracle.cloudsc2_blocked("dp", params, q2, C2.reference_eta(KLEV), nb * nproma, nproma, 1, KLEV, False, 900.0, ref,
                            nthreads=cpu_threads)
    err = {}
    for n in C2.OUTPUT_NAMES:
        a, r = got[n], ref[n]
        err[n] = float(np.sqrt(np.sum((a - r) ** 2) / max(np.sum(r ** 2), 1e-300)))
    bpc = cloudsc2_bytes_per_column(KLEV)
    gbs = bpc * NGPTOT / (ms / 1e3) / 1e9
    res = {"kernel_ms": ms, "columns_per_s": NGPTOT / (ms / 1e3), "algorithmic_bytes_per_column": bpc,
           "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": gbs / HBM_PEAK_GBS,
                        "traffic": read_traffic(nproma, "cloudsc2_sp_l137_655360_traffic.md")},
           "err_vs_fp64": err,
           "workload": f"CLOUDSC2 FP32 (CUADJTQ final adjustment), NLEV={KLEV}, {NGPTOT} columns, NPROMA={nproma}, "
                       f"dt=900 s"}
    if cpu_seconds > 0:
        arr = dict(inp)
        for n in C2.OUTPUT_NAMES:
            arr[n] = np.zeros(got[n].shape, dtype=np.float32)
        passes, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_seconds:
            oracle.cloudsc2_blocked("sp", params, q2, C2.reference_eta(KLEV), nb * nproma, nproma, 1, KLEV, False,
                                    900.0, arr, nthreads=cpu_threads)
            passes += 1
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": passes * nb * nproma / dt, "unit": "columns/s", "cores": cpu_threads,
                               "kind": "port", "sample": f"{nb * nproma} columns x {passes} passes ({dt:.1f} s), "
                                                         "FP32 oracle, OpenMP over NPROMA blocks"}
    res["layer"] = layer_bench(params, nproma, steps, lepcld2=1)
    res["layer_cloudst"] = layer_bench(params, nproma, steps, lepcld2=0)
    return res


def plain_leg(C, params, variant, nproma, steps, warmup, cpu_threads, dev):
    """Plain CLOUDSC in `variant` (sp or mixed) on the C2 grid: per-field error of the first 2048 columns against
    the FP64 CPU oracle, then `steps` timed steps (PLUDE restore + one blocked launch), kernel time from HIP events
    on the launch stream."""
    import numpy as np
    import torch

    import oracle

    f = C.BlockedFields(NGPTOT, nproma, KLEV, variant, dev)
    C.synth_device(f, params, 20261202)
    plude0 = f["plude"].clone()
    nb = 2048 // nproma
    inp = {n: f[n][:nb].cpu().numpy() for n in C.INPUT_NAMES}
    C.cloudsc_blocked(f, params, 900.0)
    torch.cuda.synchronize()
    out = {n: f[n][:nb].cpu().numpy() for n in C.OUTPUT_NAMES}
    ref = {k: (v.astype(np.float64) if v.dtype.kind == "f" else v.copy()) for k, v in inp.items()}
    for name, kind, role in C.BLOCKED_ARRAYS:
        if role == "o":
            ref[name] = np.full(C.blocked_shape(kind, nproma, KLEV, nb), np.nan)
    oracle.cloudsc_blocked("dp", params, nb * nproma, nproma, KLEV, 900.0, ref, nthreads=cpu_threads)

    def rr(a, r):
        a = a.astype(np.float64)
        m = ~np.isnan(r)
        return float(np.sqrt(np.sum((a[m] - r[m]) ** 2) / np.sum(r[m] ** 2)))

    err = {n: rr(out[n], ref[n]) for n in ("ptendency_loc_t", "ptendency_loc_q", "ptendency_loc_a", "pfplsl",
                                           "pfplsn", "pcovptot")}
    for s, nm in enumerate(("l", "i", "r", "s")):
        err[f"ptendency_loc_cld_{nm}"] = rr(out["ptendency_loc_cld"][:, s], ref["ptendency_loc_cld"][:, s])
    sp = out["pfplsl"][:, -1].astype(np.float64) + out["pfplsn"][:, -1]
    sr = ref["pfplsl"][:, -1] + ref["pfplsn"][:, -1]
    err["surface_precip"] = float(np.sqrt(np.sum((sp - sr) ** 2) / np.sum(sr ** 2)))

    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        f["plude"].copy_(plude0, non_blocking=True)
        C.cloudsc_blocked(f, params, 900.0, stream=stream.cuda_stream)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        f["plude"].copy_(plude0, non_blocking=True)
        ev[i][0].record(stream)
        C.cloudsc_blocked(f, params, 900.0, stream=stream.cuda_stream)
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    launch = None
    try:
        from openifs_amd import autotune

        launch = dict(zip(("block", "waves_per_simd", "klev_compiled"), autotune.get_launch(variant, KLEV)))
    except Exception:  # noqa: BLE001 - informational
        pass
    del f, plude0
    torch.cuda.empty_cache()
    ms = wall * 1e3 / steps
    gbs = bytes_per_column(KLEV) * NGPTOT / (kern_ms / 1e3) / 1e9
    return {"variant": variant, "ms_per_step": ms, "kernel_ms": kern_ms, "columns_per_s": NGPTOT / (ms / 1e3),
            "achieved_gbs": gbs, "frac": gbs / HBM_PEAK_GBS, "launch": launch, "err_vs_fp64": err}


def host_cpus():
    """CPUs this process may run on."""
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


# CPU threads one GPU's share of a GPU host provides to its process (OMP_NUM_THREADS is set to it there)
CPU_SHARE = 16


HOST_CPUS = host_cpus()  # taken at import, before any OpenMP runtime of this process binds its threads


def cpu_threads(args):
    return args.cpu_threads or min(CPU_SHARE, int(os.environ.get("OMP_NUM_THREADS", CPU_SHARE)), HOST_CPUS)


def main():
    global NGPTOT, KLEV
    if len(sys.argv) == 3 and sys.argv[1] == "--cpu-child":
        cpu_child(sys.argv[2])
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nproma", type=int, default=128)
    ap.add_argument("--variant", choices=("mixed", "sp"), default="mixed",
                    help="headline variant: mixed (FP32, FP64 water budget; within 1e-5 of FP64) or sp (all FP32)")
    ap.add_argument("--no-second", action="store_true", help="do not time the other variant")
    ap.add_argument("--columns", type=int, default=NGPTOT,
                    help="columns of the plain workload (profiling other grid sizes; the benchmark is 655,360)")
    ap.add_argument("--klev", type=int, default=KLEV, help="levels of the plain workload (the benchmark is 137)")
    ap.add_argument("--cpu-sample", type=int, default=65536)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-hybrid", action="store_true")
    ap.add_argument("--no-satadj", action="store_true")
    ap.add_argument("--no-budget", action="store_true")
    ap.add_argument("--no-cloudsc2", action="store_true")
    ap.add_argument("--train-cols", type=int, default=20_000_000)
    ap.add_argument("--hybrid-cols", type=int, default=6_599_680)
    ap.add_argument("--rerun-fraction", type=float, default=0.01)
    ap.add_argument("--gate-fraction", type=float, default=0.02)
    ap.add_argument("--infer-batch", type=int, default=65536)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (CLOUDSC is column-local; --gpus {args.gpus} "
              "is not supported). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from openifs_amd import cloudsc as C

    NGPTOT, KLEV = args.columns, args.klev
    dev = "cuda:0"
    torch.cuda.set_device(0)
    params = C.default_params(900.0, KLEV)
    head = plain_leg(C, params, args.variant, args.nproma, args.steps, args.warmup, args.cpu_threads, dev)
    other = "sp" if args.variant == "mixed" else "mixed"
    second = None if args.no_second else plain_leg(C, params, other, args.nproma, max(5, min(args.steps, 20)), 2,
                                                   args.cpu_threads, dev)
    ms_per_step, kern_ms = head["ms_per_step"], head["kernel_ms"]
    value = NGPTOT / (ms_per_step / 1e3)
    bpc = bytes_per_column(KLEV)
    achieved = bpc * NGPTOT / (kern_ms / 1e3) / 1e9
    traffic = read_traffic(args.nproma, f"cloudsc_{args.variant}_l{KLEV}_{NGPTOT}_traffic.md")

    hybrid = None
    if not args.no_hybrid:
        hybrid = surrogate_bench(params, args.train_cols, args.hybrid_cols, args.rerun_fraction, args.gate_fraction,
                                 args.infer_batch)

    satadj = None
    if not args.no_satadj:
        satadj = satadj_bench(params, args.nproma, max(5, min(args.steps, 20)), 2,
                              cpu_threads(args), 0.0 if args.no_cpu else 3.0)

    budget = None
    if not args.no_budget:
        budget = budget_bench(params, args.nproma, max(5, min(args.steps, 10)))

    cloudsc2 = None
    if not args.no_cloudsc2:
        cloudsc2 = cloudsc2_bench(params, args.nproma, max(5, min(args.steps, 20)),
                                  cpu_threads(args), 0.0 if args.no_cpu else 3.0)

    cpu = None
    if not args.no_cpu:
        nthreads = cpu_threads(args)
        cps, cdt, passes, per_nproma, child_cpus = cpu_baseline(params, args.cpu_sample, nthreads, args.cpu_seconds)
        best = max(per_nproma, key=per_nproma.get)
        cpu = {"value": cps, "unit": "columns/s", "cores": nthreads, "kind": "port",
               "sample": f"{args.cpu_sample} columns of the same synthetic distributions, median of {passes} passes "
                         f"({cdt:.1f} s) after one warm-up, FP32 oracle (C++/OpenMP restatement of CLOUDSC), "
                         f"NPROMA=32 blocks, SCHEDULE(DYNAMIC,1)",
               "cpu_model": cpu_model(), "host_cpus": os.cpu_count(), "host_cpus_available": child_cpus,
               "threads": nthreads, "omp_proc_bind": "close", "omp_places": "cores",
               "columns_per_s_by_nproma": {str(k): v for k, v in sorted(per_nproma.items())},
               "best_nproma": best}

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "columns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "precision": {"mixed": "FP32 storage and arithmetic, FP64 water budget of each level",
                      "sp": "FP32 throughout (the ifs_sp build)"}[args.variant],
        "data": "synthetic (seeded atmospheres generated on the device; distributions in include/cloudsc_synth.h)",
        "config": {"workload": f"plain CLOUDSC FP32 ({args.variant}), NLEV={KLEV}, {NGPTOT} columns"
                               f"{' (TCo399-sized)' if NGPTOT == 655360 else ''}, "
                               f"NPROMA={args.nproma}, dt=900 s", "columns": NGPTOT, "nlev": KLEV,
                   "nproma": args.nproma, "variant": args.variant, "launch": head["launch"],
                   "parallelism": "single GPU"},
        "kernel_ms": kern_ms,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_column": bpc},
        "err_vs_fp64": head["err_vs_fp64"],
        other: second,
        "hybrid": hybrid,
        "satadj": satadj,
        "budget": budget,
        "cloudsc2": cloudsc2,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
