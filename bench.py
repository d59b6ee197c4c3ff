"""Benchmark of the MI355X CNRK2 step (one JSON line on stdout).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-steps C] [--workload NAME]

Workload (default): plain plane Poiseuille DNS, FP64, constant bulk flux,
grid 128x97x128, Lx=4pi, Lz=4pi/3, nu=1/2800 (bulk Re 5600), dt=0.005,
base flow 1.5(1-y^2), bulk velocity (1, 0), seeded divergence-free
perturbation of amplitude 0.05 with kappa^-4 energy spectrum.
A step is one CNRK2 step! call (three stages) on device-resident fields.

value = steps/s on one GPU.  roofline = the dominant kernel class (by device
time) with its algorithmic bytes per launch over its event-timed mean launch
duration; roofline.traffic = that kernel's HBM bytes per launch measured by
rocprofv3 FETCH_SIZE + WRITE_SIZE (copy-calibrated), read from
profiles/hbm_traffic_<workload>.json when present.  cpu_baseline = the oracle (C++ restatement of the reference CPU
step, all host cores) on a bounded sample of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "channelflow.jl_amd")]

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "poiseuille128": dict(N=(128, 97, 128), L=(4 * math.pi, 4 * math.pi / 3), nu=1 / 2800, dt=0.005,
                          profile="poiseuille15", bulk=(1.0, 0.0), amp=0.05,
                          desc="plane Poiseuille 128x97x128 FP64 bulk-flux, Lx=4pi Lz=4pi/3 nu=1/2800 dt=0.005"),
    "couette32": dict(N=(32, 33, 32), L=(2 * math.pi / 1.14, 2 * math.pi / 2.5), nu=1 / 400, dt=0.01,
                      profile="couette", bulk=None, amp=0.1,
                      desc="plane Couette 32x33x32 FP64, Lx=2pi/1.14 Lz=2pi/2.5 Re=400 dt=0.01"),
}

CLASSES = ["y_prologue", "z_inverse", "x_pass", "z_forward", "y_stage"]
LAUNCHES_PER_STEP = [1, 3, 3, 3, 3]


def algorithmic_bytes(Nx, Ny, Nz):
    """Per-launch algorithmic HBM bytes (one read of every input, one write of
    every output, complex128) of each kernel class, and the canonical
    per-step figure 16(87 S + 54 Sz) used for the step-level roofline."""
    Nxh, Nzp = Nx // 2 + 1, (3 * Nz + 1) // 2
    S = Nxh * Nz * Ny
    Sz = Nxh * Nzp * Ny
    per_launch = {
        "y_prologue": 16 * 8 * S,             # U (3) -> V, Omega_x, Omega_z (5)
        "z_inverse": 16 * (5 * S + 6 * Sz),   # 5 fields (k,l,y) -> 6 (k,zp,y) (V_x also as dV_x/dz)
        "x_pass": 16 * 9 * Sz,                # 6 in, 3 out
        "z_forward": 16 * (3 * Sz + 3 * S),   # 3 fields (k,zp,y) -> (k,l,y)
        "y_stage": 16 * 55 * S / 3,           # stage mean of (7+12, 10+12, 10+4) S
    }
    step_canonical = 16 * (87 * S + 54 * Sz)
    return per_launch, step_canonical, S, Sz


MIXED = dict(N=(256, 193, 256), L=(4 * math.pi, 4 * math.pi / 3), nu=1 / 2800, dt=0.0025, amp=0.1, bulk=(1.0, 0.0),
             desc="plane Poiseuille 256x193x256, Lx=4pi Lz=4pi/3 nu=1/2800 dt=0.0025, bulk velocity (1, 0), "
                  "amplitude 0.1; FP32 transforms and nonlinear term, FP64 solves")


def mixed_bytes(Nx, Ny, Nz):
    """Algorithmic bytes of the mixed step: FP64 state traffic 16*39 S, FP32
    y-physical and z-padded traffic 8*(48 S + 54 Sz); and of the FP64 step."""
    Nxh, Nzp = Nx // 2 + 1, (3 * Nz + 1) // 2
    S, Sz = Nxh * Nz * Ny, Nxh * Nzp * Ny
    return 16 * 39 * S + 8 * (48 * S + 54 * Sz), 16 * (87 * S + 54 * Sz)


def wall_shear(g, U, nu, base):
    """nu (u'(-1) - u'(+1)) / 2 of the total mean streamwise velocity
    (examples/mkm590/run.jl:236): Chebyshev sums over the (0,0) column."""
    import numpy as np
    a = U[0, :, 0, 0].real.cpu().numpy() + base
    n = np.arange(g.Ny, dtype=float)
    dp = float(np.sum(n ** 2 * a))
    dm = float(np.sum((-1.0) ** (n + 1) * n ** 2 * a))
    return nu * (dm - dp) / 2


def bench_mixed(cf, steps):
    """FP32-mixed step against the plain FP64 step on the same seeded state:
    steps/s of each, and the relative error of energy, dissipation and wall
    shear after `steps` steps."""
    import numpy as np
    import torch

    w = MIXED
    g = cf.Grid(*w["N"], *w["L"])
    prof = lambda y: 1.5 * (1 - y ** 2)  # noqa: E731
    res = {}
    states = {}
    s0 = None
    for prec in ("fp64", "mixed"):
        p = cf.ChannelFlowProblem(g, prof, w["nu"], w["dt"], bulkvelocity=w["bulk"], precision=prec)
        cf.tune(p)  # both steps with their timed tile choice
        if s0 is None:
            s0 = cf.perturbation_state(p, w["amp"], seed=7)
        st = s0.copy()
        cf.step_(p.scheme, p.nlterm, st.velocity, st.stagepressure, 0.0, **p.constraint)  # warm-up step
        st = s0.copy()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        cf.step_(p.scheme, p.nlterm, st.velocity, st.stagepressure, 0.0, nsteps=steps, **p.constraint)
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / steps
        base = base_coefficients(p)
        res[prec] = {"ms_per_step": round(ms, 4), "steps_per_s": round(1000.0 / ms, 2),
                     "E": cf.kinetic_energy(p, st.velocity), "D": cf.dissipation_rate(p, st.velocity),
                     "wall_shear": wall_shear(g, st.velocity, w["nu"], base)}
        states[prec] = st
        del p
    mb, fb = mixed_bytes(*w["N"])
    f64, mx = res["fp64"], res["mixed"]
    rel = lambda k: abs(mx[k] - f64[k]) / abs(f64[k])  # noqa: E731
    dU = float((states["mixed"].velocity - states["fp64"].velocity).abs().max())
    return {
        "workload": w["desc"], "steps": steps,
        "mixed_steps_per_s": mx["steps_per_s"], "fp64_steps_per_s": f64["steps_per_s"],
        "speedup_vs_fp64": round(mx["steps_per_s"] / f64["steps_per_s"], 3),
        "mixed_hbm_frac": round(mb * mx["steps_per_s"] / 1e9 / HBM_PEAK_GBS, 4),
        "fp64_hbm_frac": round(fb * f64["steps_per_s"] / 1e9 / HBM_PEAK_GBS, 4),
        "rel_err_energy": rel("E"), "rel_err_dissipation": rel("D"), "rel_err_wall_shear": rel("wall_shear"),
        "max_abs_dU": dU,
    }


ENSEMBLE = dict(n=64, grids=[(16, 17, 16), (24, 33, 24), (32, 33, 32), (48, 49, 48)], Re=(300.0, 500.0),
                L=(2 * math.pi / 1.14, 2 * math.pi / 2.5), dt=0.01, amp=0.1,
                desc="64 plane Couette instances, grids drawn (seed 2024) from 16x17x16, 24x33x24, 32x33x32, "
                     "48x49x48, Re uniform in [300, 500], Lx=2pi/1.14 Lz=2pi/2.5 dt=0.01, amplitude 0.1 "
                     "(seed 1000+i); one graph-captured ensemble step, autotuned tiles")


# The reference's published complete-step timings (benchmarks/results/final/
# cpu-1.csv and gpu.csv, min over 100 samples; docs/src/benchmarks.md:38-44):
# N -> (Xeon Gold 6336Y 1 thread, A100 80GB PCIe) seconds, grid (N, N+1, N).
REFERENCE_STEP_S = {8: (0.000261355, 0.002532956), 16: (0.002208649, 0.002612679), 32: (0.016417331, 0.002889341),
                    64: (0.169632803, 0.005745282), 128: (1.62903821, 0.019938163),
                    192: (6.581443052, 0.052922028), 256: (15.013734888, 0.115872041)}


def bench_reference_grids(cf, sizes, samples=100, warmup=5):
    """The reference's own step benchmark on this device: plane Couette,
    rotating form, zero pressure gradient, nu = 1/400, dt = 0.002,
    Lx = Lz = 2pi, grid (N, N+1, N), the smooth initial field of
    benchmarks/initial.jl:2-13 (projected, with its pressure). Protocol of
    benchmarks/step.jl:31-53: restore the initial state, then time one step
    (host clock around the call and a device synchronisation); five warm-up
    samples, then the minimum and mean over `samples`."""
    import numpy as np
    import torch
    rows = []
    for N in sizes:
        g = cf.Grid(N, N + 1, N, 2 * math.pi, 2 * math.pi)
        prob = cf.CouetteFlow(g, 1 / 400, 0.002)
        cf.tune(prob)
        y, z, x = g.points()
        one = np.ones((g.Ny, g.Nzp, g.Nxp))
        u = np.stack([0.1 * (1 - y ** 2) * np.cos(x) * np.sin(z) * one,
                      0.05 * (1 - y ** 2) ** 2 * np.cos(z) * one,
                      0.2 * y * (1 - y ** 2) * np.sin(z) * one])
        U0 = cf.project_(cf.FFT(prob, torch.from_numpy(u).cuda()), prob)
        P0 = cf.pressure(U0, prob)
        U, P = U0.clone(), P0.clone()
        times = []
        for k in range(warmup + samples):
            U.copy_(U0)
            P.copy_(P0)
            torch.cuda.synchronize()
            a = time.perf_counter()
            cf.step_(prob.scheme, prob.nlterm, U, P, 0.0, **prob.constraint)
            torch.cuda.synchronize()
            if k >= warmup:
                times.append(time.perf_counter() - a)
        assert bool(torch.isfinite(torch.view_as_real(U)).all()), "non-finite velocity"
        tmin, tmean = min(times), sum(times) / len(times)
        row = {"grid": [N, N + 1, N], "min_ms": round(1e3 * tmin, 4), "mean_ms": round(1e3 * tmean, 4)}
        if N in REFERENCE_STEP_S:
            cpu_s, a100_s = REFERENCE_STEP_S[N]
            row.update({"reference_cpu_1thread_ms": round(1e3 * cpu_s, 3), "reference_a100_ms": round(1e3 * a100_s, 3),
                        "speedup_vs_a100": round(a100_s / tmin, 2), "speedup_vs_cpu_1thread": round(cpu_s / tmin, 1)})
        rows.append(row)
        progress(f"reference grid {N}x{N + 1}x{N}: min {1e3 * tmin:.3f} ms, mean {1e3 * tmean:.3f} ms")
        del prob, U0, P0, U, P
        torch.cuda.empty_cache()
    return {"workload": "the reference's step benchmark: plane Couette, rotating form, zero pressure gradient, "
                        "nu=1/400, dt=0.002, Lx=Lz=2pi, grid (N, N+1, N), smooth initial field; min and mean of "
                        f"{samples} single timed steps after {warmup} warm-ups, each from the restored initial state",
            "reference_hardware": "Xeon Gold 6336Y 1 thread (FFTW) and A100 80GB PCIe (CUDA.jl), "
                                  "docs/src/benchmarks.md:38-44",
            "rows": rows}


def bench_ensemble(cf, steps, plain_steps=20):
    """Batched ensemble (one graph launch per ensemble step) against the plain
    path (each problem advanced by its own step_ calls): instance-steps/s."""
    import numpy as np
    import torch

    w = ENSEMBLE
    rng = np.random.default_rng(2024)
    probs, states, total_bytes, drawn = [], [], 0, []
    for i in range(w["n"]):
        dims = w["grids"][int(rng.integers(len(w["grids"])))]
        Re = float(rng.uniform(*w["Re"]))
        g = cf.Grid(*dims, *w["L"])
        p = cf.CouetteFlow(g, 1.0 / Re, w["dt"])
        probs.append(p)
        states.append(cf.random_state(p, w["amp"], seed=1000 + i))
        total_bytes += algorithmic_bytes(*dims)[1]
        drawn.append(dims)
    plain_states = [s.copy() for s in states]
    t0 = time.perf_counter()
    ens = cf.Ensemble(probs, states, tune=True)
    setup_s = time.perf_counter() - t0
    launches, _ = ens.info()
    ens.advance(10)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    ens.advance(steps)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / steps
    ok = all(bool(torch.isfinite(torch.view_as_real(s.velocity)).all()) for s in states)
    # plain path: every instance stepped by its own step_ calls, one after another
    for p, s in zip(probs, plain_states):
        cf.step_(p.scheme, p.nlterm, s.velocity, s.stagepressure, 0.0, nsteps=2, **p.constraint)
    torch.cuda.synchronize()
    a.record()
    for p, s in zip(probs, plain_states):
        cf.step_(p.scheme, p.nlterm, s.velocity, s.stagepressure, 0.0, nsteps=plain_steps, **p.constraint)
    b.record()
    torch.cuda.synchronize()
    ms_plain = a.elapsed_time(b) / plain_steps
    counts = {f"{d[0]}x{d[1]}x{d[2]}": drawn.count(d) for d in w["grids"]}
    del ens
    return {
        "workload": w["desc"], "instances": w["n"], "grid_counts": counts, "steps": steps,
        "instance_steps_per_s": round(w["n"] * 1000.0 / ms, 1),
        "ms_per_ensemble_step": round(ms, 4),
        "plain_instance_steps_per_s": round(w["n"] * 1000.0 / ms_plain, 1),
        "speedup_vs_plain": round(ms_plain / ms, 2),
        "launches_per_ensemble_step": launches, "plain_launches_per_ensemble_step": 13 * w["n"],
        "ensemble_step_algorithmic_bytes": total_bytes,
        "hbm_frac": round(total_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "setup_s_with_tuning": round(setup_s, 2), "finite": ok,
    }


CLOSURE = dict(fine=(128, 97, 128), coarse=(32, 33, 32), L=(4 * math.pi, 4 * math.pi / 3), nu=1 / 2800, dt=0.005,
               bulk=(1.0, 0.0), amp=0.3,
               desc="learned subgrid closure: coarse Poiseuille 32x33x32 (Lx=4pi Lz=4pi/3 nu=1/2800 dt=0.005, bulk "
                    "velocity (1, 0)); MLP 10-64-64-64-3 GELU, BF16 weights/activations on MFMA, FP32 accumulation; "
                    "trained on box-filtered snapshots of a 128x97x128 FP64 plain run (perturbation amplitude 0.3)")


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bench_closure(cf, snapshots, every, eval_steps, epochs, sample_every=10):
    """Config of the learned closure: data from the fine plain run, training
    on the device, and coarse runs with and without the closure judged
    against the filtered DNS over the next `eval_steps` steps."""
    import numpy as np
    import torch
    from channelflow import closure as CL

    w = CLOSURE
    prof = lambda y: 1.5 * (1 - y ** 2)  # noqa: E731
    gf, gc = cf.Grid(*w["fine"], *w["L"]), cf.Grid(*w["coarse"], *w["L"])
    fine = cf.ChannelFlowProblem(gf, prof, w["nu"], w["dt"], bulkvelocity=w["bulk"])
    coarse = cf.ChannelFlowProblem(gc, prof, w["nu"], w["dt"], bulkvelocity=w["bulk"])
    st = cf.perturbation_state(fine, w["amp"], seed=11)
    t0 = time.perf_counter()
    Xs, Ts = [], []
    for k in range(snapshots):
        cf.step_(fine.scheme, fine.nlterm, st.velocity, st.stagepressure, 0.0, nsteps=every, **fine.constraint)
        X, T = CL.snapshot_samples(fine, coarse, st.velocity)
        Xs.append(X)
        Ts.append(T)
        if (k + 1) % 250 == 0:
            torch.cuda.synchronize()
            progress(f"closure: {k + 1}/{snapshots} snapshots")
    # interleaved split: every tenth snapshot is held out, so the held-out
    # set spans the whole sampled window (a final block would see only the
    # run's last, possibly different, state)
    test_ids = [k for k in range(snapshots) if k % 10 == 9] or [snapshots - 1]
    train_ids = [k for k in range(snapshots) if k not in set(test_ids)] or [0]
    Xtr, Ttr = torch.cat([Xs[k] for k in train_ids]), torch.cat([Ts[k] for k in train_ids])
    Xte, Tte = torch.cat([Xs[k] for k in test_ids]), torch.cat([Ts[k] for k in test_ids])
    del Xs, Ts
    torch.cuda.synchronize()
    data_s = time.perf_counter() - t0
    net = CL.SubgridMLP(seed=1)
    net.fit_normalization(Xtr, Ttr)
    t0 = time.perf_counter()
    progress(f"closure: training on {Xtr.shape[0]} samples")
    hist = net.train(Xtr, Ttr, epochs=epochs, batch=8192, lr=2e-3)
    progress(f"closure: training loss by epoch {hist}")
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    train_samples = Xtr.shape[0]
    test_loss, test_base = net.evaluate(Xte, Tte)
    del Xtr, Ttr, Xte, Tte
    torch.cuda.empty_cache()
    # a posteriori: filtered DNS and coarse runs from the same filtered state
    Ub = cf.box_filter(fine, coarse, st.velocity)
    # the closure as a device-side forcing (graph-replayed step), and the same
    # closure called back from the host at every stage (same kernels, same result)
    runs = {"plain": (Ub.clone(), None), "closure": (Ub.clone(), net.forcing(coarse)),
            "closure_mean": (Ub.clone(), net.forcing(coarse, modes="mean")),
            "closure_host": (Ub.clone(), net.forcing(coarse, device=False))}
    stats = {k: cf.ProfileStatistics(gc.Ny) for k in ("dns", "plain", "closure", "closure_mean", "closure_host")}
    ms = {}
    for name, (U, force) in runs.items():
        P = cf.pressure(U, coarse)
        torch.cuda.synchronize()
        tt = 0.0
        for s in range(0, eval_steps, sample_every):
            a = time.perf_counter()
            cf.step_(coarse.scheme, coarse.nlterm, U, P, s * w["dt"], nsteps=sample_every, forcing=force,
                     **coarse.constraint)
            torch.cuda.synchronize()
            tt += time.perf_counter() - a
            stats[name].sample(coarse, U, (s + sample_every) * w["dt"])
        ms[name] = 1000.0 * tt / eval_steps
        progress(f"closure: coarse {name} run done ({ms[name]:.3f} ms/step)")
    for s in range(0, eval_steps, sample_every):
        cf.step_(fine.scheme, fine.nlterm, st.velocity, st.stagepressure, 0.0, nsteps=sample_every, **fine.constraint)
        stats["dns"].sample(coarse, cf.box_filter(fine, coarse, st.velocity), (s + sample_every) * w["dt"])
    ref_m, ref_c = stats["dns"].profiles()

    def err(name):
        m, c = stats[name].profiles()
        em = float(np.linalg.norm(m[:, 0] - ref_m[:, 0]) / np.linalg.norm(ref_m[:, 0]))
        ec = [float(np.linalg.norm(c[:, k] - ref_c[:, k]) / max(np.linalg.norm(ref_c[:, k]), 1e-300)) for k in range(6)]
        return em, ec

    em_p, ec_p = err("plain")
    em_c, ec_c = err("closure")
    em_m, ec_m = err("closure_mean")
    names = ["uu", "vv", "ww", "uv", "uw", "vw"]
    return {
        "workload": w["desc"], "snapshots": snapshots, "steps_between_snapshots": every,
        "train_samples": train_samples, "epochs": epochs, "train_loss_by_epoch": [round(h, 5) for h in hist],
        "train_samples_per_s": round(epochs * train_samples / train_s, 1), "data_s": round(data_s, 2),
        "train_s": round(train_s, 2),
        "heldout_loss": round(test_loss, 5), "heldout_loss_no_model": round(test_base, 5),
        "heldout_gain": round(1.0 - test_loss / test_base, 4),
        "eval_steps": eval_steps,
        "coarse_plain_steps_per_s": round(1000.0 / ms["plain"], 1),
        "coarse_closure_steps_per_s": round(1000.0 / ms["closure"], 1),
        "coarse_closure_mean_steps_per_s": round(1000.0 / ms["closure_mean"], 1),
        "coarse_closure_host_callback_steps_per_s": round(1000.0 / ms["closure_host"], 1),
        "closure_device_equals_host_callback": bool(torch.equal(runs["closure"][0], runs["closure_host"][0])),
        "mean_profile_rel_err": {"plain": round(em_p, 5), "closure": round(em_c, 5),
                                 "closure_mean": round(em_m, 5)},
        "reynolds_stress_rel_err": {"plain": dict(zip(names, [round(e, 4) for e in ec_p])),
                                    "closure": dict(zip(names, [round(e, 4) for e in ec_c])),
                                    "closure_mean": dict(zip(names, [round(e, 4) for e in ec_m]))},
    }


def base_coefficients(problem):
    """Chebyshev coefficients of the problem's base profile (host DCT-I of its
    Lobatto-point values, chebcoeffs normalisation)."""
    import numpy as np
    v = np.asarray(problem.baseflow, dtype=float)
    P = len(v) - 1
    j = np.arange(P + 1)
    c = np.array([(v * np.cos(np.pi * n * j / P) * np.where((j == 0) | (j == P), 0.5, 1.0)).sum() for n in range(P + 1)])
    c *= 2.0 / P
    c[0] /= 2
    c[P] /= 2
    return c


def make_problem(cf, w):
    Nx, Ny, Nz = w["N"]
    g = cf.Grid(Nx, Ny, Nz, *w["L"])
    prof = (lambda y: 1.5 * (1 - y ** 2)) if w["profile"] == "poiseuille15" else (lambda y: y)
    kw = {"bulkvelocity": w["bulk"]} if w["bulk"] else {"pressuregradient": (0.0, 0.0)}
    return g, prof, kw, cf.ChannelFlowProblem(g, prof, w["nu"], w["dt"], **kw)


def _timed_samples(ref, U0, P0, warmup, samples):
    """benchmarks/step.jl:31-53: setup outside timing, the state restored
    before every sample, `warmup` untimed steps, then `samples` timed single
    steps; returns the per-step seconds."""
    import numpy as np
    secs = []
    for i in range(warmup + samples):
        U, P = np.array(U0, copy=True), np.array(P0, copy=True)
        t0 = time.perf_counter()
        ref.step_inplace(U, P, 1)
        if i >= warmup:
            secs.append(time.perf_counter() - t0)
    return secs


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# the reference's published single-thread step (Julia + FFTW, Xeon Gold 6336Y):
# plane Couette, rotational form, nu = 1/400, dt = 0.002, Lx = Lz = 2 pi,
# grid 96x97x96 (benchmarks/results/xeon-1.csv:8)
REF_1THREAD_96 = 0.6229


def cpu_baseline(w, prof, kw, U0, P0, warmup, samples, samples_1t):
    """The oracle (C++ restatement of the reference CPU step) timed on the
    host: the same workload on all threads (min and mean over `samples`
    steps), and single-threaded on the reference's own 96x97x96 benchmark
    case, whose Julia + FFTW time is published, to state the restatement's
    speed against the reference's."""
    import numpy as np
    from oracle import oracle as O

    Nx, Ny, Nz = w["N"]
    go = O.Grid(Nx, Ny, Nz, *w["L"])
    threads = os.cpu_count() or 1
    O.set_threads(min(threads, 16))
    ref = O.Problem(go, prof, w["nu"], w["dt"], **kw)
    secs = _timed_samples(ref, U0, P0, warmup, samples)
    used = O.num_threads()
    out = {"min_s": round(min(secs), 4), "mean_s": round(float(np.mean(secs)), 4), "threads": used,
           "warmup": warmup, "samples": samples}
    one = None
    if samples_1t > 0:
        O.set_threads(1)
        g96 = O.Grid(96, 97, 96, 2 * math.pi, 2 * math.pi)
        r96 = O.Problem(g96, lambda y: y, 1 / 400, 0.002)
        rng = np.random.default_rng(0)
        U96, P96 = r96.random_state(0.1 * rng.standard_normal((3,) + g96.phys_shape))
        s1 = _timed_samples(r96, U96, P96, 1, samples_1t)
        one = {"case": "plane Couette 96x97x96, Lx=Lz=2pi, nu=1/400, dt=0.002, rotational (the reference's "
                       "benchmark case)",
               "threads": 1, "warmup": 1, "samples": samples_1t, "min_s": round(min(s1), 4),
               "mean_s": round(float(np.mean(s1)), 4), "reference_julia_fftw_1thread_s": REF_1THREAD_96,
               "oracle_over_reference": round(min(s1) / REF_1THREAD_96, 2)}
        O.set_threads(used)
    return out, one


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=20, help="timed oracle samples, all threads (0: skip)")
    ap.add_argument("--cpu-warmup", type=int, default=5)
    ap.add_argument("--cpu-steps-1thread", type=int, default=3,
                    help="timed single-thread oracle samples on the reference's 96x97x96 case")
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--workload", default="poiseuille128", choices=sorted(WORKLOADS))
    ap.add_argument("--mixed-steps", type=int, default=500,
                    help="steps of the FP32-mixed 256x193x256 comparison (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=1000,
                    help="steps of the 64-instance batched ensemble (0: skip)")
    ap.add_argument("--closure-snapshots", type=int, default=2000,
                    help="filtered snapshots of the learned-closure workload (0: skip)")
    ap.add_argument("--closure-every", type=int, default=5, help="fine steps between snapshots")
    ap.add_argument("--closure-eval-steps", type=int, default=5000)
    ap.add_argument("--closure-epochs", type=int, default=2)
    ap.add_argument("--reference-grids", default="8,16,32,64,128,192,256",
                    help="N list of the reference's (N, N+1, N) Couette step benchmark ('' : skip)")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run N device copies of 256 MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")  # 256 MiB
        dst = torch.empty_like(src)
        for _ in range(args.copy_calibration):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each": src.numel() * 8, "copies": args.copy_calibration}))
        return 0
    if args.gpus != 1:
        print(f"bench.py: this library measures one GPU only (requested --gpus {args.gpus}); "
              "the multi-GPU pencil decomposition is a design note, not built.", file=sys.stderr)
        return 2

    import torch

    import channelflow as cf

    cf.require_device()
    w = WORKLOADS[args.workload]
    g, prof, kw, prob = make_problem(cf, w)
    tiles = cf.tune(prob)  # tile shapes by timing (setup; results do not depend on them)
    full_tiles = cf.get_tiles(prob)
    state = cf.perturbation_state(prob, w["amp"], seed=42)
    U, P = state.velocity, state.stagepressure
    U0, P0 = U.cpu().numpy(), P.cpu().numpy()
    t = 0.0
    for _ in range(args.warmup):
        t = cf.step_(prob.scheme, prob.nlterm, U, P, t, **prob.constraint)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(args.steps):
        t = cf.step_(prob.scheme, prob.nlterm, U, P, t, **prob.constraint)
    stop.record()
    torch.cuda.synchronize()
    ms_total = start.elapsed_time(stop)
    ms_per_step = ms_total / args.steps
    steps_per_s = 1000.0 / ms_per_step
    assert bool(torch.isfinite(torch.view_as_real(U)).all()), "non-finite velocity"

    # per-kernel-class device times (HIP events on the launch stream)
    acc = [0.0] * 5
    for _ in range(args.profile_steps):
        ms = cf.step_timed(prob.scheme, U, P, t, **prob.constraint)
        acc = [a + m for a, m in zip(acc, ms)]
    per_step_ms = [a / max(args.profile_steps, 1) for a in acc]
    if args.profile_steps == 0:
        per_step_ms = [1.0] * 5  # roofline not measured in this run
    dom = max(range(5), key=lambda i: per_step_ms[i])
    per_launch_bytes, step_bytes, S, Sz = algorithmic_bytes(*w["N"])
    launch_ms = per_step_ms[dom] / LAUNCHES_PER_STEP[dom]
    achieved = per_launch_bytes[CLASSES[dom]] / (launch_ms * 1e-3) / 1e9
    breakdown = {CLASSES[i]: {"ms_per_step": round(per_step_ms[i], 5),
                              "gbs": round(per_launch_bytes[CLASSES[i]] * LAUNCHES_PER_STEP[i] /
                                           (per_step_ms[i] * 1e-3) / 1e9, 1)} for i in range(5)}
    step_gbs = step_bytes * steps_per_s / 1e9

    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"hbm_traffic_{args.workload}.json")
    if os.path.exists(tpath):  # rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, copy-calibrated
        with open(tpath) as f:
            rec = json.load(f)["kernels"].get(CLASSES[dom])
        if rec:
            traffic = rec["fetch"] + rec["write"]

    cpu = None
    if args.cpu_steps > 0:
        allc, one = cpu_baseline(w, prof, kw, U0, P0, args.cpu_warmup, args.cpu_steps, args.cpu_steps_1thread)
        cpu = {"value": round(1.0 / allc["min_s"], 5), "unit": "steps/s", "cores": allc["threads"], "kind": "port",
               "sample": f"the oracle on the same workload: {allc['warmup']} warm-up steps, then "
                         f"{allc['samples']} single steps each from the restored initial state "
                         f"(benchmarks/step.jl:31-53 protocol); value = 1/min, {allc['threads']} OpenMP threads",
               "min_s": allc["min_s"], "mean_s": allc["mean_s"], "nproc": os.cpu_count(), "cpu_model": cpu_model(),
               "gpu_over_cpu": round(steps_per_s * allc["min_s"], 1), "one_thread": one}

    forms = {}
    del prob, state, U, P
    progress(f"headline done: {steps_per_s:.1f} steps/s")
    if args.mixed_steps > 0:
        torch.cuda.empty_cache()
        try:
            forms["mixed"] = bench_mixed(cf, args.mixed_steps)
        except Exception as e:  # the headline line must still print
            forms["mixed"] = {"error": f"{type(e).__name__}: {e}"}
    if args.ensemble_steps > 0:
        torch.cuda.empty_cache()
        try:
            forms["batched"] = bench_ensemble(cf, args.ensemble_steps)
        except Exception as e:
            forms["batched"] = {"error": f"{type(e).__name__}: {e}"}
    if args.reference_grids:
        torch.cuda.empty_cache()
        try:
            forms["reference_grids"] = bench_reference_grids(cf, [int(n) for n in args.reference_grids.split(",")])
        except Exception as e:
            forms["reference_grids"] = {"error": f"{type(e).__name__}: {e}"}
    if args.closure_snapshots > 0:
        torch.cuda.empty_cache()
        try:
            forms["closure"] = bench_closure(cf, args.closure_snapshots, args.closure_every, args.closure_eval_steps,
                                             args.closure_epochs)
        except Exception as e:
            forms["closure"] = {"error": f"{type(e).__name__}: {e}"}

    out = {
        "metric": "CNRK2 steps/s & %HBM roofline, 128×97×128 FP64; mixed/batched/closure vs plain",
        "value": round(steps_per_s, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded divergence-free perturbation, kappa^-4 spectrum, amplitude 0.05)",
        "config": {"workload": w["desc"], "grid": list(w["N"]), "form": "rotational",
                   "constraint": "bulkvelocity (1.0, 0.0)" if w["bulk"] else "pressuregradient (0, 0)",
                   "tiles": {"z_tile_k": tiles[0], "z_points_per_thread": tiles[1], "x_waves_per_block": tiles[2],
                             "x_variant": ["six", "split", "async"][full_tiles[2]],
                             "y_coefs_per_lane": full_tiles[4]},
                   "step_algorithmic_bytes": step_bytes,
                   "step_hbm_frac": round(step_gbs / HBM_PEAK_GBS, 4),
                   "kernel_breakdown": breakdown},
        "roofline": {"bound": "hbm", "kernel": CLASSES[dom], "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic},
        "cpu_baseline": cpu,
        "forms": forms,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
