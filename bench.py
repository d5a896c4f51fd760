"""Benchmark: map-level -2lnL + gradient evaluations per second, FP64.

Workload (default): nside=1024, lmax=2048 (4,198,401 packed a_lm, 12,582,912
pixels), Gaussian beam FWHM 10 arcmin, white noise sigma 20 uK, 20 degree
latitude cut, C_l = 1e3/(l(l+1)).  One step = one likelihood + gradient
evaluation (cmbl_value_and_grad_device) with x, the data, the weights and all
tables already resident in HBM.

Prints ONE JSON line.  `roofline` is for the dominant kernel (the slower of the
two FP64 Legendre kernels), with algorithmic work 8 * R * Nalm FLOP per launch
(R = 2*nside ring pairs, Nalm = (lmax+1)(lmax+2)/2; 4 FLOP for the recurrence
step and 4 for the complex accumulation per (l, m, ring pair)), timed with HIP
events on the evaluation stream during the timed region.  `cpu_baseline` times
the CPU oracle (C FP64 restatement, OpenMP) on one evaluation of the same
workload.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "likelihood+gradient evals/sec at Nside=1024,lmax=2048 FP64; FP64 roofline %"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0

from cases import CONFIG1, CONFIG2, CONFIG4, CONFIG5, CONFIG5_CHAIN, batch_configs, device_case  # noqa: E402

CONFIGS = {1: CONFIG1, 2: CONFIG2, 4: CONFIG4, 5: CONFIG5}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_case(cfg, seed=101):
    """Inputs of the workload; the data map is synthesised on the GPU."""
    return device_case(cfg, seed)


def bench_batch(args):
    """64 uneven maps: one batched evaluation (one launch per stage over all
    maps) against the same maps evaluated one after another."""
    import torch

    import cosmopp_amd as cm

    cfgs = batch_configs()
    dev = torch.device("cuda:0")
    maps, xs, gs = [], [], []
    log("bench: building %d maps" % len(cfgs))
    for c in cfgs:
        mp, inp = build_case(c, c["seed"])
        maps.append(mp)
        xs.append(torch.from_numpy(inp["x"]).to(dev))
        gs.append(torch.empty_like(xs[-1]))
    # the per-map comparison runs each map with its single-map tuning (the
    # tuned table) rather than the batch members' one ring pair per lane
    singles = [build_case(dict(c, tuning=None), c["seed"])[0] for c in cfgs]
    f = torch.empty(len(maps), dtype=torch.float64, device=dev)
    batch = cm.Batch(maps)
    sp = torch.cuda.current_stream().cuda_stream
    xp, gp = [x.data_ptr() for x in xs], [g.data_ptr() for g in gs]

    def batched():
        batch.value_and_grad_device(xp, f.data_ptr(), gp, sp)

    def single():
        for i, mp in enumerate(singles):
            mp.value_and_grad_device(xp[i], f.data_ptr() + 8 * i, gp[i], sp)

    rates = {}
    for name, fn in (("batched", batched), ("per_map", single)):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        rates[name] = len(maps) * args.steps / (time.perf_counter() - t0)
        log("bench: %s %.2f map-evals/s" % (name, rates[name]))
    flop = sum(16.0 * 2 * c["nside"] * (c["lmax"] + 1) * (c["lmax"] + 2) / 2 for c in cfgs)
    out = {
        "metric": "map-level -2lnL + gradient evaluations per second over a batch of 64 uneven maps, FP64",
        "value": round(rates["batched"], 3),
        "unit": "map-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(len(maps) / rates["batched"] * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "64 maps, lmax ~ U{256,512,768,1024,1536}, nside = 2^round(log2(lmax/2)), "
                               "FWHM = 5400/lmax arcmin, sigma ~ U[10,50] uK, cut ~ U[10,30] deg",
                   "maps": [[c["nside"], c["lmax"]] for c in cfgs], "parallelism": "single GPU"},
        "per_map_value": round(rates["per_map"], 3),
        "batched_over_per_map": round(rates["batched"] / rates["per_map"], 3),
        "legendre_flop_per_batch_eval": flop,
        "legendre_tflops_batched": round(flop * rates["batched"] / len(maps) / 1e12, 3),
    }
    batch.close()
    for mp in maps + singles:
        mp.close()
    print(json.dumps(out), flush=True)


def bench_c5_chains(args):
    """The 16 chain maps of config 5 (nside 256 / lmax 512, FWHM 40 arcmin,
    sigma 40 uK, 20 degree cut, independent seeds), evaluated as one batch."""
    import torch

    import cosmopp_amd as cm

    cfg = CONFIG5_CHAIN
    dev = torch.device("cuda:0")
    maps, xs, gs = [], [], []
    for i in range(16):
        mp, inp = build_case(cfg, 2000 + 10 * i)
        maps.append(mp)
        xs.append(torch.from_numpy(inp["x"]).to(dev))
        gs.append(torch.empty_like(xs[-1]))
    f = torch.empty(len(maps), dtype=torch.float64, device=dev)
    batch = cm.Batch(maps)
    sp = torch.cuda.current_stream().cuda_stream
    xp, gp = [x.data_ptr() for x in xs], [g.data_ptr() for g in gs]
    for _ in range(args.warmup):
        batch.value_and_grad_device(xp, f.data_ptr(), gp, sp)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.value_and_grad_device(xp, f.data_ptr(), gp, sp)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    batch.close()
    for mp in maps:
        mp.close()
    return {"maps": 16, "nside": 256, "lmax": 512, "ms_per_batch": round(ms, 4),
            "map_evals_per_s": round(16 / ms * 1e3, 3)}


def bench_pol(args, cfg):
    """Spin-2 synthesis (HEALPix alm2map_pol: T, E, B -> T, Q, U) on device
    buffers at the configuration's nside / lmax; one step = one call."""
    import numpy as np
    import torch

    import cosmopp_amd as cm

    nside, lmax = cfg["nside"], cfg["lmax"]
    log("bench: pol nside=%d lmax=%d" % (nside, lmax))
    mp = cm.Map(nside, lmax)
    rng = np.random.default_rng(17)
    n = (lmax + 1) ** 2
    dev = torch.device("cuda:0")
    ins = [torch.from_numpy(rng.standard_normal(n)).to(dev) for _ in range(3)]
    outs = [torch.empty(mp.npix, dtype=torch.float64, device=dev) for _ in range(3)]
    sp = torch.cuda.current_stream().cuda_stream
    ptrs = [v.data_ptr() for v in ins] + [o.data_ptr() for o in outs]
    for _ in range(args.warmup):
        mp.alm2map_pol_device(*ptrs, sp)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mp.alm2map_pol_device(*ptrs, sp)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    R, nalm = 2 * nside, (lmax + 1) * (lmax + 2) // 2
    # algorithmic: T 8 R Nalm; Q/U 4 (recurrence) + 16 (a^E F1, a^B F2, a^B F1, a^E F2) per (l, m, pair)
    flop = 28.0 * R * nalm
    rate = args.steps / elapsed
    out = {
        "metric": "spin-2 synthesis (alm2map_pol T,E,B -> T,Q,U) calls/sec, FP64",
        "value": round(rate, 3),
        "unit": "calls/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "alm2map_pol nside=%d lmax=%d" % (nside, lmax), "nside": nside, "lmax": lmax},
        "algorithmic_flop_per_call": flop,
        "algorithmic_tflops": round(flop * rate / 1e12, 3),
        "frac_fp64_peak": round(flop * rate / 1e12 / FP64_PEAK_TFLOPS, 4),
        "executed_legendre_flop_per_call": 9 * 8.0 * R * nalm,
    }
    mp.close()
    print(json.dumps(out), flush=True)


TRAFFIC_FILE = "profiles/pmc_traffic.json"


def bench_loops(args):
    """Device-resident optimisation and sampling at nside 1024 / lmax 2048:
    200 L-BFGS iterations (m = 10, More-Thuente) from x = 0 and 500 HMC
    leapfrog steps (50 trajectories of a fixed 10 steps of 0.05; masses from
    the prior and noise curvature per l).  Overhead = loop time over (its
    likelihood evaluations x the time of one evaluation) - 1."""
    import numpy as np
    import torch

    import cosmopp_amd as cm

    cfg = CONFIGS[4]
    nside, lmax = cfg["nside"], cfg["lmax"]
    log("bench: loops nside=%d lmax=%d" % (nside, lmax))
    mp, inp = build_case(cfg)
    dev = torch.device("cuda:0")
    n = (lmax + 1) ** 2
    x = torch.from_numpy(inp["x"]).to(dev)
    g = torch.empty_like(x)
    f = torch.empty(1, dtype=torch.float64, device=dev)
    sp = torch.cuda.current_stream().cuda_stream
    for _ in range(3):
        mp.value_and_grad_device(x.data_ptr(), f.data_ptr(), g.data_ptr(), sp)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(20):
        mp.value_and_grad_device(x.data_ptr(), f.data_ptr(), g.data_ptr(), sp)
    torch.cuda.synchronize()
    eval_s = (time.perf_counter() - t0) / 20
    out = {"metric": "device-resident L-BFGS and HMC loops, nside=%d lmax=%d FP64" % (nside, lmax),
           "eval_ms": round(eval_s * 1e3, 4), "n_gpus": 1, "dtype": "f64", "data": "synthetic"}
    # L-BFGS: 200 iterations (epsilon tiny so that no convergence test stops it)
    x0 = torch.zeros(n, dtype=torch.float64, device=dev)
    res = torch.empty_like(x0)
    lb = cm.LBFGS(mp, x0.data_ptr(), m=10)
    lb.minimize(res.data_ptr(), epsilon=1e-300, gnorm_tol=0.0, max_iter=11)  # graph capture of every variant
    lb.set_starting(x0.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fl, it, ev = lb.minimize(res.data_ptr(), epsilon=1e-300, gnorm_tol=0.0, max_iter=199)
    t_lb = time.perf_counter() - t0
    ev_total = ev  # trial evaluations of the line searches
    out["lbfgs"] = {"iterations": it, "evaluations": ev_total, "seconds": round(t_lb, 4), "final_neg2lnL": fl,
                    "ms_per_iteration": round(t_lb / it * 1e3, 4),
                    "overhead_vs_evaluations": round(t_lb / (ev_total * eval_s) - 1, 4)}
    lb.close()
    log("bench: lbfgs %d iterations %d evaluations %.3f s" % (it, ev_total, t_lb))
    # HMC: 50 trajectories x 10 fixed leapfrog steps of 0.05
    l = np.floor(np.sqrt(np.arange(n))).astype(np.int64)
    l -= (l * l > np.arange(n))
    cl, b = inp["cl"], inp["beam"]
    wbar = float(np.mean(inp["w"]))
    curv = np.where(l >= 2, 2.0 / np.maximum(cl[l], 1e-300), 0.0) + 2.0 * b[l] ** 2 * wbar * 12 * nside * nside / (
        4 * math.pi)
    masses = torch.from_numpy(curv).to(dev)
    start = torch.from_numpy(inp["x"]).to(dev)
    hm = cm.HMC(mp, start.data_ptr(), masses.data_ptr(), 0.05, 10, 4242, fixed=True)
    hm.run(1)  # graph capture
    st0 = hm.state()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hm.run(50)
    t_hmc = time.perf_counter() - t0
    st = hm.state()
    evs = st["evaluations"] - st0["evaluations"]
    out["hmc"] = {"trajectories": 50, "leapfrog_steps": 500, "steps": "fixed: tau = 0.05, 10 per trajectory",
                  "evaluations": evs, "accepted": st["accepted"] - st0["accepted"], "seconds": round(t_hmc, 4),
                  "ms_per_leapfrog_step": round(t_hmc / 500 * 1e3, 4),
                  "overhead_vs_evaluations": round(t_hmc / (evs * eval_s) - 1, 4),
                  "momenta": "reference draws (whitenoise seeds) on 12 host threads, up to 12 trajectories ahead"}
    hm.close()
    log("bench: hmc 500 steps %.3f s" % t_hmc)
    mp.close()
    print(json.dumps(out), flush=True)


def bench_gibbs(args):
    """CMBGibbsSampler steps on the GPU at the reference sampler test's sizes
    (source/test_cmb_gibbs.cpp:64-99: nside 16 / Gibbs lmax 30 with a uniform
    mask, nside 32 / lmax 50 with the reference's mask1.fits), and the Gibbs
    CG at nside 1024 / lmax 2048 (ms per iteration).  The oracle's sampler is
    timed beside the first two on the host (a bounded number of steps)."""
    import numpy as np

    import cosmopp_amd as cm
    from cases import gibbs_case

    def synth(nside, L, a, beam):
        mp = cm.Map(nside, L)
        mp.set_beam(beam)
        m = mp.forward(a)
        mp.close()
        return m

    m, ns, nested = cm.read_map(os.path.join(ROOT, "tests", "golden", "mask1.fits"))
    ring = cm.reorder(m, ns, to_ring=True) if nested else m
    out = {"metric": "Gibbs sampler steps/s (CMBGibbsSampler::step) and Gibbs CG ms/iteration, FP64",
           "n_gpus": 1, "dtype": "f64", "data": "synthetic sky, white noise; mask1.fits from the reference's tests"}
    for i, steps, cpu_steps in ((0, 40, 20), (4, 6, 1)):
        c = gibbs_case(i, ring_mask=ring if i == 4 else None, synth=synth)
        log("bench: gibbs subtest %d nside=%d lmax=%d" % (i, c["nside"], c["lmax"]))
        gs = cm.GibbsSampler(c["map"], c["noise"], c["mask"], c["sigma"], c["lmax"], c["fwhm"], c["cl0"],
                             seed=c["seed"])
        gs.step()
        t0 = time.perf_counter()
        its = 0
        for _ in range(steps):
            gs.step()
            its += gs.iterations
        dt = (time.perf_counter() - t0) / steps
        gs.close()
        rec = {"nside": c["nside"], "lmax": c["lmax"], "mask": "mask1.fits" if i == 4 else "uniform",
               "steps_per_s": round(1.0 / dt, 3), "ms_per_step": round(dt * 1e3, 3),
               "cg_iterations_per_step": its / steps, "ms_per_cg_iteration": round(dt * 1e3 / (its / steps), 4)}
        if not args.no_cpu_baseline:
            import oracle

            og = oracle.GibbsSampler(c["nside"], c["lmax"], c["map"], c["noise"], c["mask"], c["sigma"], c["beam"],
                                     c["cl0"], c["seed"])
            t0 = time.perf_counter()
            for _ in range(cpu_steps):
                og.step()
            cdt = (time.perf_counter() - t0) / cpu_steps
            rec["cpu_baseline"] = {"steps_per_s": round(1.0 / cdt, 4), "kind": "port", "cores": oracle.num_threads(),
                                   "sample": "%d step(s) of the oracle's CMBGibbsSampler restatement" % cpu_steps}
        out["sampler_subtest%d" % i] = rec
    # Gibbs CG at the headline geometry: 300 iterations
    nside, lmax, sigma = 1024, 2048, 20.0
    log("bench: gibbs CG nside=%d lmax=%d" % (nside, lmax))
    g = cm.Gibbs(nside, lmax, cm.cl_shape(lmax), cm.beam(lmax, 10.0 / 60.0),
                 cm.mask_latitude(nside, math.radians(20.0)), sigma)
    rng = np.random.default_rng(3)
    npix = 12 * nside * nside
    b = g.rhs(rng.standard_normal(npix) * 50, rng.standard_normal(npix), rng.standard_normal(npix))
    for its in (16, 300):
        t0 = time.perf_counter()
        try:
            g.solve(b, precision=1e-30, max_iter=its)  # runs the full count
        except cm.CmblError as e:
            if e.code != cm.CMBL_ERR_NOCONV:
                raise
        dt = time.perf_counter() - t0
    g.close()
    out["gibbs_cg_n1024_l2048"] = {"iterations": 300, "ms_per_iteration": round(dt * 1e3 / 300, 4)}
    print(json.dumps(out), flush=True)


def read_traffic(nside, lmax, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE passes, calibrated on a plain copy), if
    that file holds this kernel and workload; the counters need a profiler run
    of their own, so the bench line reads them rather than measuring them."""
    path = os.path.join(ROOT, TRAFFIC_FILE)
    if not os.path.exists(path):
        return None
    try:
        t = json.load(open(path))
        key = "%s_n%d_l%d" % (kernel, nside, lmax)
        return t.get(key)
    except Exception:
        return None


def cpu_baseline(cfg, inputs, budget_s, result=None):
    """Time the oracle (CPU FP64 restatement) on the same workload.  When
    `result` is a dict, the oracle's last (-2lnL, gradient) go into it."""
    import oracle

    nside, lmax = cfg["nside"], cfg["lmax"]
    threads = oracle.num_threads()
    times = []
    t_start = time.time()
    while True:
        t0 = time.perf_counter()
        fo, go = oracle.neg2lnL(nside, lmax, inputs["x"], inputs["beam"], inputs["cl"], inputs["d"], inputs["w"])
        times.append(time.perf_counter() - t0)
        if result is not None:
            result["f"], result["g"] = fo, go
        if time.time() - t_start + times[-1] > budget_s or len(times) >= 5:
            break
    times.sort()
    med = times[len(times) // 2]
    return {
        "value": 1.0 / med,
        "unit": "evals/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d full value+gradient evaluation(s) of the same workload (median %.2f s), "
                  "CPU oracle oracle/cmbl_oracle.c, %d OpenMP threads" % (len(times), med, threads),
    }


def cpu_table(args):
    """The CPU oracle at configs 1, 2 and 4, on one thread and on every host
    thread this process may use: bounded samples of full value + gradient
    evaluations (the GPU only synthesises the input maps)."""
    import os
    import platform

    import oracle

    full = oracle.num_threads()
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    rows = []
    for c in (1, 2, 4):
        cfg = CONFIGS[c]
        mp, inputs = build_case(cfg)
        mp.close()
        for threads in (full, 1):
            oracle.set_threads(threads)
            budget = args.cpu_budget if threads > 1 else 3 * args.cpu_budget
            b = cpu_baseline(cfg, inputs, budget)
            rows.append({"config": c, "nside": cfg["nside"], "lmax": cfg["lmax"], "threads": threads,
                         "evals_per_s": b["value"], "sample": b["sample"]})
            log("cpu table: config %d, %d threads: %.4g evals/s" % (c, threads, b["value"]))
        oracle.set_threads(full)
    print(json.dumps({"cpu_table": rows, "cpu_model": model, "nproc": os.cpu_count(),
                      "threads_available": full}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=4, choices=sorted(CONFIGS) + [3],
                    help="workload: 4 (default, the headline), 1, 2, 5 (main map), 3 (64-map batch)")
    ap.add_argument("--pol", action="store_true",
                    help="time the spin-2 synthesis (cmbl_alm2map_pol_device) at the config's nside/lmax")
    ap.add_argument("--graph", action="store_true",
                    help="also time the evaluation captured in a HIP graph and replayed (launch-bound sizes)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-table", action="store_true",
                    help="only time the CPU oracle at configs 1, 2, 4 on one and on all threads")
    ap.add_argument("--loops", action="store_true",
                    help="time the device-resident L-BFGS (200 iterations) and HMC (500 leapfrog steps) at config 4")
    ap.add_argument("--gibbs", action="store_true",
                    help="time Gibbs sampler steps (nside 16 and 32, the reference test's sizes) and the Gibbs CG")
    ap.add_argument("--chains", action="store_true",
                    help="only time config 5's 16 chain maps (nside 256 / lmax 512) as one batch")
    ap.add_argument("--host-buffers", action="store_true",
                    help="time cmbl_value_and_grad with x and g in host memory (PCIe-inclusive rate)")
    ap.add_argument("--cpu-budget", type=float, default=40.0, help="seconds for the CPU baseline leg")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (calibrates HBM byte counters)")
    args = ap.parse_args()

    if args.copy_probe:
        import torch

        n = args.copy_probe * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(max(args.steps, 1)):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8,
                          "copies": max(args.steps, 1)}))
        return

    if args.gpus != 1:
        print("bench.py measures one MI355X only.  Several GPUs "
              "would take whole maps each (independent maps, LPT by Legendre cost) or m-slices of one map "
              "with two all-to-all transposes around the ring stage; that layout is not implemented.  "
              "Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import cosmopp_amd as cm

    if args.cpu_table:
        cpu_table(args)
        return
    if args.config == 3:
        bench_batch(args)
        return
    if args.loops:
        bench_loops(args)
        return
    if args.gibbs:
        bench_gibbs(args)
        return
    if args.chains:
        out = {"metric": "map-level -2lnL + gradient evaluations per second over config 5's 16 chain maps "
                         "as one batch, FP64", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
               "dtype": "f64", "data": "synthetic"}
        out.update(bench_c5_chains(args))
        print(json.dumps(out), flush=True)
        return
    cfg = CONFIGS[args.config]
    if args.pol:
        bench_pol(args, cfg)
        return
    nside, lmax = cfg["nside"], cfg["lmax"]
    log("bench: building inputs nside=%d lmax=%d" % (nside, lmax))
    mp, inputs = build_case(cfg)
    dev = torch.device("cuda:0")
    x = torch.from_numpy(inputs["x"]).to(dev)
    g = torch.empty_like(x)
    f = torch.empty(1, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    if args.host_buffers:
        xh = inputs["x"]
        for _ in range(args.warmup):
            mp.value_and_grad(xh)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fh, gh = mp.value_and_grad(xh)
        rate = args.steps / (time.perf_counter() - t0)
        print(json.dumps({"metric": "host-buffer value_and_grad evals/s (x in, g out over PCIe)",
                          "value": round(rate, 3), "unit": "evals/s", "nside": nside, "lmax": lmax,
                          "steps": args.steps, "neg2lnL": fh}), flush=True)
        mp.close()
        return

    log("bench: warmup %d" % args.warmup)
    for _ in range(args.warmup):
        mp.value_and_grad_device(x.data_ptr(), f.data_ptr(), g.data_ptr(), sp)
    torch.cuda.synchronize()

    graph_rate = None
    if args.graph:
        # the same evaluation captured once in a HIP graph (on a side stream:
        # capture needs a non-default stream) and replayed
        side = torch.cuda.Stream()
        side.wait_stream(stream)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=side):
            mp.value_and_grad_device(x.data_ptr(), f.data_ptr(), g.data_ptr(), side.cuda_stream)
        graph.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            graph.replay()
        torch.cuda.synchronize()
        graph_rate = args.steps / (time.perf_counter() - t0)
        del graph
        log("bench: graph replay %.1f evals/s" % graph_rate)

    mp.set_timing(True)
    log("bench: timing %d steps" % args.steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mp.value_and_grad_device(x.data_ptr(), f.data_ptr(), g.data_ptr(), sp)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stages, nev = mp.stage_times()
    mp.set_timing(False)
    fval = f.item()
    if not math.isfinite(fval):
        raise RuntimeError("non-finite -2lnL")

    elapsed = t1 - t0
    ms_per_step = elapsed / args.steps * 1e3
    value = args.steps / elapsed
    stage_avg = {k: v / max(nev, 1) for k, v in stages.items()}

    R = 2 * nside
    nalm = (lmax + 1) * (lmax + 2) // 2
    flop_per_launch = 8.0 * R * nalm
    dom = "adjoint" if stage_avg["adjoint"] >= stage_avg["synthesis"] else "synthesis"
    t_dom = stage_avg[dom] * 1e-3
    achieved = flop_per_launch / t_dom / 1e12
    kname = "k_adj" if dom == "adjoint" else "k_synth"
    traffic = read_traffic(nside, lmax, kname)
    roofline = {
        "bound": "fp64_valu",
        "compute": "FP64 VALU (no MFMA; MI355X FP64 vector peak equals its FP64 matrix peak, 78.6 TFLOP/s)",
        "kernel": "%s (Legendre %s)" % (kname, dom),
        "achieved": round(achieved, 3),
        "peak": FP64_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
        "traffic": traffic,
        "traffic_source": TRAFFIC_FILE if traffic is not None else None,
        "flop_per_launch": flop_per_launch,
        "avg_launch_ms": round(stage_avg[dom], 4),
        "legendre_pair_frac": round(2 * flop_per_launch / ((stage_avg["adjoint"] + stage_avg["synthesis"]) * 1e-3)
                                    / 1e12 / FP64_PEAK_TFLOPS, 4),
    }
    # ring stage (HBM-bound): algorithmic bytes 32 Nring (L+1) (parity sums
    # in, ring transforms out, 16 B per (ring, m) each way) + 16 Npix (d, w)
    ring_bytes = 32.0 * (4 * nside - 1) * (lmax + 1) + 16.0 * 12 * nside * nside
    roofline["ring_stage_bytes"] = ring_bytes
    roofline["ring_stage_hbm_frac"] = round(ring_bytes / (stage_avg["ring"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    ring_traffic = read_traffic(nside, lmax, "k_ring")
    if ring_traffic is not None:
        roofline["ring_traffic"] = ring_traffic
        roofline["ring_traffic_ratio"] = round(ring_traffic / ring_bytes, 3)

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        log("bench: cpu baseline (oracle)")
        ref = {}
        cpu = cpu_baseline(cfg, inputs, args.cpu_budget, ref)
        # the oracle's -2lnL and gradient at the same x against this run's
        import numpy as np

        gd = g.cpu().numpy()
        parity = {"f_rel": abs(fval - ref["f"]) / abs(ref["f"]),
                  "g_rel": float(np.linalg.norm(gd - ref["g"]) / np.linalg.norm(ref["g"])),
                  "tolerance": 1e-10}
        log("bench: parity vs oracle f_rel %.3e g_rel %.3e" % (parity["f_rel"], parity["g_rel"]))

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "map-level -2lnL + gradient (synthesis, ring DFTs, weighting, adjoint), "
                        "nside=%d lmax=%d FWHM=%.1f arcmin sigma=%g uK, %g deg latitude cut"
                        % (nside, lmax, cfg["fwhm_deg"] * 60, cfg["sigma"], cfg["cut_deg"]),
            "nside": nside,
            "lmax": lmax,
            "nparams": (lmax + 1) ** 2,
            "npix": 12 * nside * nside,
            "parallelism": "single GPU",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
        "stages_ms": {k: round(v, 4) for k, v in stage_avg.items()},
        "neg2lnL": fval,
    }
    if graph_rate is not None:
        out["graph_replay_evals_per_s"] = round(graph_rate, 3)
    mp.close()
    if args.config == 5:
        log("bench: config 5 chain maps")
        ch = bench_c5_chains(args)
        out["chains"] = ch
        # one config-5 step: the main map and the 16 chain maps once each
        out["config5_steps_per_s"] = round(1e3 / (ms_per_step + ch["ms_per_batch"]), 3)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
