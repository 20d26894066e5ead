#!/usr/bin/env python3
"""Throughput of the P3M Yukawa NVE loop on one MI355X: chosen vs plain form.

Workload (default): N = 1,000,000 unit-charge ions, box L = (4 pi N / 3)^(1/3),
Gamma = 100 (T = 0.01), kappa = 2, 128^3 mesh, order-6 assignment, optimal
Green function, uniform positions (RandomState(1)) and Maxwellian velocities
(RandomState(2)), dt = 0.1, NVE Velocity-Verlet with per-step K, U, E, T and
n(q, t) for the 20 smallest q along each axis.  alpha = 0.40 and rc = 7.0
hold the RMS force error to at most 1e-5 of the mean force on the
equilibrated liquid (Gamma 100), where the mean force is ~20x smaller than on
the uniform start and the budget binds.  The error against a tight FP64
reference (autotune.reference_settings) is measured at setup on both states (the
liquid: the same start after 100 steps at dt/10 and --eq-steps steps with a
velocity rescale every 10 steps and a per-ion speed clip) and reported.

Two forms run the same workload from the same seeded state:
  plain  - FP64 everywhere (the reference's arithmetic);
  chosen - FP32 PP pair math, assignment weights and FFT/Green/ik passes,
           FP64 charge-mesh atomics, energy/momentum reductions and positions.
`value` is the chosen form's throughput; the plain form's throughput, the
ratio and both force errors are reported in `config`.

Timed region (per form): `steps` fused NVE steps (kick-drift, binning, PP,
assignment, FFT solve, gather + kick, FP64 reductions, n(q, t)) with every
input already in HBM, bracketed by synchronisations; HIP events on the
kernels' stream give the per-stage durations that feed the roofline entry.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "md-force-routines_amd"))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s at RMS force err<=1e-5, chosen vs plain FP64 P3M, 1 GPU"
PEAK_HBM = 8.0e12          # B/s, MI355X
PEAK_FP64 = 78.6e12        # FLOP/s, MI355X dense FP64
PEAK_FP32 = 157.3e12       # FLOP/s, MI355X dense FP32


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--mesh", type=int, default=128)
    ap.add_argument("--kappa", type=float, default=2.0)
    ap.add_argument("--gamma", type=float, default=100.0)
    ap.add_argument("--alpha", type=float, default=0.40)
    ap.add_argument("--rc", type=float, default=7.0)
    ap.add_argument("--dt", type=float, default=0.1)
    ap.add_argument("--nq", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-err", action="store_true", help="skip the force-error measurement")
    ap.add_argument("--no-liquid", action="store_true",
                    help="measure the force error on the uniform start only (no equilibrated state)")
    ap.add_argument("--cpu-steps", type=int, default=3,
                    help="NVE steps of the all-core CPU baseline (the C restatement on all host threads)")
    ap.add_argument("--cpu-steps1", type=int, default=1,
                    help="NVE steps of the single-thread CPU baseline (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="threads of the all-core baseline (0 = all)")
    ap.add_argument("--skin", type=float, default=0.05,
                    help="Verlet pair-list skin of the PP stage with --no-autotune (0 = off); the cells of "
                         "rc + skin set the skin actually used (the largest they allow).  The autotuner picks "
                         "the skin, PP waves and PM block shapes by timing")
    ap.add_argument("--workload", choices=("n1e6", "n8e6", "replicas", "mlp", "radial", "drift", "sqw", "cpu"),
                    default="n1e6",
                    help="n1e6: one N = 1e6 system (default); replicas: 64 seeded replicas, N 2,000-50,000, "
                         "autotuned, stepped by one graph launch per --graph-steps steps")
    ap.add_argument("--replicas", type=int, default=64)
    ap.add_argument("--replica-meshes", type=lambda t: tuple(int(x) for x in t.split(",")), default=(32, 64, 128),
                    help="mesh sizes the replica autotuner considers (comma separated)")
    ap.add_argument("--mlp-epochs", type=int, default=20)
    ap.add_argument("--drift-steps", type=int, default=10000)
    ap.add_argument("--drift-skin", type=float, default=0.0,
                    help="pair-list skin of both forms in --workload drift (0 = no lists)")
    ap.add_argument("--eq-steps", type=int, default=300)
    ap.add_argument("--uniform-steps", type=int, default=1000)
    ap.add_argument("--start", choices=("both", "uniform", "equilibrated"), default="both",
                    help="drift workload: start(s) of the trajectories")
    ap.add_argument("--autotune", action=argparse.BooleanOptionalAction, default=True,
                    help="alpha and rc from the autotuner (FP32-mixed, budget 1e-5, the given mesh, tuned on the "
                         "equilibrated liquid); --no-autotune uses --alpha/--rc")
    ap.add_argument("--graph-steps", type=int, default=50, help="0 = no graph (eager launches)")
    ap.add_argument("--cpu-configs", default="cfg1,cfg2",
                    help="cpu workload: comma-separated cfg1 (= cfg1-1t and cfg1-all: one thread, all threads "
                         "and the GPU), cfg2")
    ap.add_argument("--streams", type=int, default=16, help="worker streams of the replica graph")
    ap.add_argument("--form", choices=("both", "chosen", "plain"), default="both",
                    help="forms to time; value is the chosen form's unless only plain runs")
    ap.add_argument("--pp-waves", type=int, default=0,
                    help="PP waves a workgroup without --autotune (4, 8, 12, 16; 0 = the context's choice)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 1 GiB device-to-device copy (counter calibration for rocprofv3 --pmc)")
    return ap.parse_args()


PMC_FILE = os.path.join(ROOT, "profiles", "p3m_mixed_n1e6_m128_pmc_traffic.md")


def pmc_traffic(stage, form, N, M, rc, alpha):
    """HBM bytes per launch of the PP kernel from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE scaled by a 1 GiB copy), when this run
    is the workload that summary measured; else None."""
    if not (stage == "pp" and form == "mixed"):
        return None
    try:
        for line in open(PMC_FILE):
            if line.startswith("traffic k_pp<float"):
                t = line.split(">", 1)[1].split()  # key value pairs after the kernel name
                kv = dict(zip(t[0::2], map(float, t[1::2])))
                # the PP kernel's traffic depends on N, the mesh and the cells
                # (rc), not on alpha
                if (kv.get("N"), kv.get("M"), kv.get("rc")) != (N, M, rc):
                    return None
                return (kv["fetch_kib"] * kv["fetch_factor"] + kv["write_kib"] * kv["write_factor"]) * 1024.0
    except OSError:
        return None
    return None


def host_cpu():
    """Model name and thread counts of the host CPU."""
    model = platform.processor() or "cpu"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    return {"model": model, "nproc": avail, "cpu_count": os.cpu_count()}


def stage_work(stage, N, M, rc):
    """Algorithmic work of one launch of a stage: (bytes, flops).

    Plain FP64: bin 60 B/particle; PP 56 B/particle and 61 FLOP per unique
    in-cutoff pair (N rc^3 / 2 pairs at unit Wigner-Seitz density);
    assignment 32 B/particle + 8 B/node; FFT solve 100 B/node (R2C 16,
    Green x ik 36, three C2R 48); gather 56 B/particle + 24 B/node;
    kick-drift 120 B/particle (read pos, vel, acc; write pos, vel); n(q)
    24 B/particle.
    """
    M3 = M ** 3
    table = {
        "bin": (60.0 * N, 0.0),
        "pp": (56.0 * N, 61.0 * N * rc ** 3 / 2.0),
        "assign": (32.0 * N + 8.0 * M3, 0.0),
        "fft": (100.0 * M3, 0.0),
        "gather": (56.0 * N + 24.0 * M3, 0.0),
        "drift": (120.0 * N, 0.0),
        "modes": (24.0 * N, 0.0),
        "finalize": (0.0, 0.0),
    }
    return table[stage]


def run_replicas(a):
    """Replica ensemble: 64 replicas, N_r ~ U{2000..50000}, Gamma_r in {10, 50,
    100, 175}, kappa_r in {0.5, 1, 2, 3} (RandomState(3) draws in that order),
    positions RandomState(1000 + r), velocities RandomState(2000 + r) at
    T_r = 1/Gamma_r.  Each replica is first equilibrated from that start
    (100 steps at dt/10, then --eq-steps steps with a velocity rescale every
    10 steps and a per-ion speed clip, on a provisional setting: 32^3 or 64^3
    mesh, rc = min(5, L/3), alpha = 3.2/rc), and its (mesh, alpha, rc) is
    then taken from the autotuner (meshes --replica-meshes, default 32^3,
    64^3 and 128^3, in order of modelled cost) on the equilibrated liquid, where an RMS
    force error <= 1e-5 of the mean force is the harder bar (FP32-mixed).
    With --autotune (default) the pair-list skins are then chosen by
    measurement (ReplicaBatch.tune_skins: per replica the fastest alone of no
    lists and skins 0.25, 0.5 and 1, then kept if the ensemble with them
    beats the ensemble without lists).
    Equilibration and tuning are outside the timed region.  Timed: `steps`
    NVE steps of all replicas from their equilibrated states with n(q, t)
    (20 q per axis) recorded on device, one graph launch per --graph-steps
    steps."""
    import numpy as np
    import torch

    from mdforce import autotune, mlp_train
    from mdforce.batch import ReplicaBatch
    from mdforce.engine import P3M
    from mdforce.initialize_pos_vel import initial_split

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    rs = np.random.RandomState(3)
    R = a.replicas
    Ns = rs.randint(2000, 50001, R)
    gammas = rs.choice([10.0, 50.0, 100.0, 175.0], R)
    kappas = rs.choice([0.5, 1.0, 2.0, 3.0], R)
    systems, P, V, A = [], [], [], []
    t_tune = time.perf_counter()
    errs = []
    for r in range(R):
        N = int(Ns[r])
        L = (4.0 * math.pi * N / 3.0) ** (1.0 / 3.0)
        pos, vel = initial_split(1000 + r, 2000 + r, N, [L, L, L], 1.0 / gammas[r])
        p = torch.from_numpy(pos).to(dev)
        v = torch.from_numpy(vel).to(dev)
        rc_eq = min(5.0, L / 3.0 * (1.0 - 1e-9))
        eq = P3M(N, 32 if N <= 10000 else 64, L, float(kappas[r]), 3.2 / rc_eq, rc_eq, precision="mixed",
                 device=dev)
        eq.build_green()
        mlp_train.equilibrate(eq, p, v, 1.0 / gammas[r], a.eq_steps, a.dt, vcap=4.0, soft_steps=100)
        eq.close()
        ch = autotune.tune(p, L, float(kappas[r]), budget=1e-5, meshes=a.replica_meshes, precision="mixed",
                           device=dev)
        errs.append(ch.error)
        systems.append(dict(n=N, L=L, kappa=float(kappas[r]), mesh=ch.mesh, alpha=ch.alpha, rc=ch.rc,
                            precision="mixed"))
        P.append(p)
        V.append(v)
        print(f"[tune] replica {r}: N={N} Gamma={gammas[r]:g} kappa={kappas[r]:g} -> mesh {ch.mesh} "
              f"alpha {ch.alpha} rc {ch.rc} err {ch.error:.2e}", file=sys.stderr, flush=True)
    t_tune = time.perf_counter() - t_tune
    b = ReplicaBatch(systems, n_streams=a.streams, device=dev)
    for r, rep in enumerate(b.replicas):
        acc, _ = rep.force(P[r])
        A.append(acc.clone())
    nq = a.nq
    skins = [0.0] * R
    if a.autotune:
        t_sk = time.perf_counter()
        skins = b.tune_skins(P, V, A, a.dt, nq=nq,
                             log=lambda m: print(f"[skin] {m}", file=sys.stderr, flush=True))
        t_tune += time.perf_counter() - t_sk
    K = min(a.graph_steps, a.steps)
    if a.warmup:
        b.run(P, V, A, a.warmup, a.dt, nq=nq, graph_steps=min(K, a.warmup))
    energies = [torch.empty((a.steps, 4), dtype=torch.float64, device=dev) for _ in range(R)]
    nqt = [torch.empty((a.steps, nq, 3), dtype=torch.complex128, device=dev) for _ in range(R)]
    b.run(P, V, A, K, a.dt, nq=nq, energies=energies, nqt=nqt, graph_steps=K)  # capture outside the timing
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    b.run(P, V, A, a.steps, a.dt, nq=nq, energies=energies, nqt=nqt, graph_steps=K)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = start.elapsed_time(stop)
    total_n = int(sum(Ns))
    value = total_n * a.steps / (gpu_ms / 1e3)
    drift = []
    for r in range(R):
        e = energies[r].cpu().numpy()
        drift.append(float(np.max(np.abs(e[:, 0] - e[0, 0])) / (Ns[r] / gammas[r])))
    nq_ok = all(bool(torch.isfinite(torch.view_as_real(t)).all()) for t in nqt)
    lstats = [rep.list_stats() for rep in b.replicas]
    out = {
        "metric": METRIC + " (replica ensemble)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": gpu_ms / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{R} replicas, N {int(Ns.min())}-{int(Ns.max())} (sum {total_n}), "
                        "Gamma {10,50,100,175}, kappa {0.5,1,2,3}, FP32-mixed, equilibrated, autotuned "
                        "mesh/alpha/rc",
            "graph_steps": K, "graph_mode": ("segmented" if b.graph_mode == 0 else "per-replica streams"),
            "streams": a.streams, "nq": nq, "dt": a.dt,
            "sum_N_rc3": float(sum(s["n"] * s["rc"] ** 3 for s in systems)),
            "tuned_error_max": max(errs), "tuned_error_mean": float(np.mean(errs)),
            "meshes": sorted(set(s["mesh"] for s in systems)),
            "rc_range": [min(s["rc"] for s in systems), max(s["rc"] for s in systems)],
            "tuned_on": "equilibrated liquid", "eq_and_tune_s": t_tune, "wall_s": wall,
            "pair_lists": {"replicas_with_lists": sum(1 for x in skins if x > 0),
                           "skin_range": [min(skins), max(skins)],
                           "builds": sum(st["builds"] for st in lstats),
                           "list_steps": sum(st["list_steps"] for st in lstats)},
            "energy_drift_per_NT_max": max(drift), "nqt_finite": nq_ok,
            "parallelism": "single GPU, replicas overlapped on worker streams inside one hipGraph",
        },
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def run_mlp(a):
    """Learned short-range correction.  Training (outside the timed region):
    16 states Gamma in {10, 50, 100, 175} x kappa in {0.5, 1, 2, 3}, 10
    equilibrated snapshots each of N = 125,000 (2e7 force residuals) between
    tight FP64 P3M (rc 8, 128^3) and the reduced setting (rc 3.5, 64^3,
    alpha 0.6; mesh spacing 1.26, that of N = 1e6 on 128^3).  Evaluation at
    N = 1e6, Gamma 100, kappa 2, equilibrated: RMS force error against a tight
    FP64 P3M (autotune.reference_settings) of the reduced setting with and
    without the correction, the estimate of the plain P3M judge (128^3 at the
    bench's alpha and rc, which meet 1e-5 in the liquid) and the resulting
    form, and the NVE throughput of the corrected reduced step and of the
    plain FP64 step."""
    import numpy as np
    import torch

    from mdforce import mlp_train as T
    from mdforce.mlp import PairMLP, default_meta

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    Ntr, rc_red, a_red = 125000, 3.5, 0.6
    red = lambda k: (64, a_red, rc_red)
    tight = lambda k: (128, 0.55, 8.0)
    prop = lambda k: (64, 0.48, 5.5)
    states = [(g, k) for g in (10.0, 50.0, 100.0, 175.0) for k in (0.5, 1.0, 2.0, 3.0)]
    t0 = time.perf_counter()
    snaps = T.make_snapshots(Ntr, states, 10, 20, 300, red=red, tight=tight, prop=prop, seed0=100, log=log)
    t_data = time.perf_counter() - t0
    rms = T.residual_rms(snaps)
    engs = {k: T.Setting(Ntr, *red(k), k).eng for k in (0.5, 1.0, 2.0, 3.0)}
    net = PairMLP(1_000_000, 50_000_000, default_meta(2.0, a_red, rc_red, out_scale=rms / 5.0), device=dev)
    t0 = time.perf_counter()
    hist = T.train(net, engs, snaps, a.mlp_epochs, 2e-3, lr_end=2e-4, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    del snaps
    # evaluation system: N = 1e6, Gamma 100, kappa 2, equilibrated
    N, kappa, gamma = 1_000_000, 2.0, 100.0
    ev = T.make_snapshots(N, [(gamma, kappa)], 1, 20, 300, red=lambda k: (128, a_red, rc_red),
                          tight=lambda k: (256, 0.5, 9.0), prop=lambda k: (128, 0.48, 5.5), seed0=5000, log=log)[0]
    pos = ev["pos"]
    from mdforce.autotune import reference_settings
    Sr = T.Setting(N, 128, a_red, rc_red, kappa)
    St = T.Setting(N, *reference_settings(T.box_length(N)), kappa)
    Fref = T.forces(St.eng, pos)
    St.close()
    meanF = float(torch.sqrt((Fref * Fref).sum(1)).mean())
    Fr = T.forces(Sr.eng, pos)
    net.set_meta(f_kappa=kappa / 3.0, f_alpha=a_red)
    Fc = Fr + net.correction(Sr.eng, pos)
    e_red = float(torch.sqrt(((Fr - Fref) ** 2).sum(1).mean())) / meanF
    e_cor = float(torch.sqrt(((Fc - Fref) ** 2).sum(1).mean())) / meanF
    Sp = T.Setting(N, 128, a.alpha, a.rc, kappa)
    form, est = T.choose_form(net, Sr.eng, Sp.eng, pos, kappa, a_red, 1e-5)

    def timed(eng, steps):
        p = pos.clone()
        v = torch.zeros_like(p)
        acc, _ = eng.force(p)
        acc = acc.clone()
        eng.md_run(p, v, acc, 2, 0.1)
        torch.cuda.synchronize()
        s0 = torch.cuda.Event(enable_timing=True)
        s1 = torch.cuda.Event(enable_timing=True)
        s0.record()
        eng.md_run(p, v, acc, steps, 0.1)
        s1.record()
        torch.cuda.synchronize()
        return s0.elapsed_time(s1) / steps

    net.attach(Sr.eng)
    ms_learned = timed(Sr.eng, a.steps)
    net.detach(Sr.eng)
    ms_red = timed(Sr.eng, a.steps)
    ms_plain = timed(Sp.eng, a.steps)
    used = ms_learned if form == "learned" else ms_plain
    out = {
        "metric": METRIC + " (learned correction)",
        "value": N / (used / 1e3),
        "unit": "particle-steps/s",
        "n_gpus": 1, "steps": a.steps, "warmup": 2, "ms_per_step": used,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16/f64", "data": "synthetic",
        "config": {
            "workload": "learned PP correction, train N=125000 x 160 snapshots, evaluate N=1e6 Gamma 100 kappa 2",
            "reduced": {"mesh": 128, "alpha": a_red, "rc": rc_red},
            "residuals": 160 * Ntr, "epochs": a.mlp_epochs, "loss_first_last": [hist[0], hist[-1]],
            "data_s": t_data, "train_s": t_train,
            "err_reduced_rel_meanF": e_red, "err_corrected_rel_meanF": e_cor, "gain": e_red / e_cor,
            "judge_estimate": est, "budget": 1e-5, "form_used": form,
            "ms_step_learned": ms_learned, "ms_step_reduced_no_mlp": ms_red, "ms_step_plain_fp64": ms_plain,
            "parallelism": "single GPU",
        },
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def run_radial(a):
    """How far a learned radial pair correction carries: N = 1e6, Gamma 100,
    kappa 2, the bench's equilibrated liquid (snapshot A) and 60 NVE steps
    later (snapshot B).  For each setting (128^3; alpha, rc below) the
    residual F_tight - F_setting on A (tight FP64 P3M, autotune.reference_settings)
    is fitted by a radial pair force f(r) = sum_k c_k exp(-(r - mu_k)^2 / 2 w^2),
    16 Gaussians over [0.7, rc] (linear least squares, mlp_train.fit_radial);
    the fit is judged on B.  Such an f(r) can be folded into the PP pair table
    at no cost per step; it is worth doing only if it takes a cheaper setting
    below the 1e-5 budget.  The line reports each setting's error on B before
    and after the correction and its measured ms a step (FP32-mixed, launch
    choices tuned)."""
    import numpy as np
    import torch

    from mdforce import autotune, mlp, mlp_train
    from mdforce.engine import P3M
    from mdforce.initialize_pos_vel import initial_split

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    N, kappa, gamma, dt = 1_000_000, 2.0, 100.0, 0.1
    L = (4.0 * math.pi * N / 3.0) ** (1.0 / 3.0)
    pos0, vel0 = initial_split(1, 2, N, [L, L, L], 1.0 / gamma)
    p = torch.from_numpy(pos0).to(dev)
    v = torch.from_numpy(vel0).to(dev)
    eq = P3M(N, 128, L, kappa, 0.42, 7.0, precision="mixed", device=dev)
    eq.build_green()
    mlp_train.equilibrate(eq, p, v, 1.0 / gamma, a.eq_steps, dt, vcap=4.0, soft_steps=100)
    pA = p.clone()
    acc, _ = eq.force(p)
    acc = acc.clone()
    eq.md_run(p, v, acc, 60, dt)
    pB, vB = p.clone(), v.clone()
    eq.close()
    refA = autotune.reference_force(P3M, pA, L, kappa, device=dev)
    refB = autotune.reference_force(P3M, pB, L, kappa, device=dev)
    rows = []
    for alpha, rc in ((0.4329, 6.75), (0.42, 7.0), (0.45, 6.5), (0.5, 6.0), (0.55, 6.0), (0.6, 5.5)):
        e = P3M(N, 128, L, kappa, alpha, rc, device=dev)
        e.build_green()
        FA = e.force(pA, None, af=1.0)[0].clone()
        FB = e.force(pB, None, af=1.0)[0].clone()
        net = mlp.PairMLP(N, int(1.25 * 4.19 * rc ** 3 * N / (L ** 3 / N) + 1000), mlp.default_meta(kappa, alpha, rc),
                          device=dev)
        centers = torch.linspace(0.7, rc, 16, dtype=torch.float64, device=dev)
        width = float(centers[1] - centers[0])
        GA, npairs = mlp_train.radial_design(net, e, pA, centers, width)
        c = mlp_train.fit_radial(GA, refA - FA)
        del GA
        GB, _ = mlp_train.radial_design(net, e, pB, centers, width)
        corrB = torch.einsum("nka,k->na", GB, c)
        del GB
        eB, eBc = autotune.force_error(FB, refB), autotune.force_error(FB + corrB, refB)
        net.close()
        e.close()
        em = P3M(N, 128, L, kappa, alpha, rc, precision="mixed", device=dev)
        em.build_green()
        tl = autotune.tune_launch(em, pB, vB, dt, steps=20)
        em.close()
        rows.append({"alpha": alpha, "rc": rc, "pairs_per_ion": npairs / N, "err_A": autotune.force_error(FA, refA),
                     "err_B": eB, "err_B_corrected": eBc, "gain": eB / eBc, "ms_per_step_mixed": tl["ms_per_step"],
                     "launch": tl["requested"]})
        print(f"[radial] alpha {alpha} rc {rc}: error {eB:.3e} -> {eBc:.3e} (x{eB / eBc:.2f}); "
              f"{tl['ms_per_step']:.4f} ms/step", file=sys.stderr, flush=True)
        del FA, FB, corrB
        torch.cuda.empty_cache()
    ok = [r for r in rows if r["err_B"] <= 1e-5]
    okc = [r for r in rows if r["err_B_corrected"] <= 1e-5]
    best = min(ok, key=lambda r: r["ms_per_step_mixed"])
    bestc = min(okc, key=lambda r: r["ms_per_step_mixed"])
    out = {
        "metric": METRIC + " (radial pair correction)",
        "value": N / (bestc["ms_per_step_mixed"] / 1e3), "unit": "particle-steps/s", "n_gpus": 1, "steps": 20,
        "warmup": 2, "ms_per_step": bestc["ms_per_step_mixed"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "radial correction fitted on an N=1e6 liquid snapshot, judged 60 steps later",
                   "settings": rows, "fastest_within_budget": {"alpha": best["alpha"], "rc": best["rc"]},
                   "fastest_within_budget_corrected": {"alpha": bestc["alpha"], "rc": bestc["rc"]},
                   "parallelism": "single GPU"},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def run_cpu(a):
    """The CPU baselines of the measurement plan, on this host, beside the GPU.

    cfg1 in full: N = 1e4, Gamma 100, kappa 2, 64^3, alpha 0.85, rc 6.2,
    dt 0.1, the seeded start, 500 steps with the velocity rescale of
    thermostat_p3m_ord6.py:5-25 (first branch), then 1000 NVE steps with n(q,t)
    for the 20 smallest q along x; through the C restatement on one thread
    (serial build) and on all threads (OpenMP build), and the plain FP64 GPU
    path on the same protocol.  cfg2: 5 NVE steps of the N = 1e6 liquid
    (128^3, alpha 0.4329, rc 6.75) on one thread and on all threads.
    Throughput counts the NVE steps only.  One JSON line with every row."""
    import torch

    from oracle import c_oracle

    c_oracle.build(ref=False)
    host = host_cpu()
    nthr = c_oracle.lib_omp(0)[1]
    dev = torch.device("cuda", 0)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    rows = []
    want = set(a.cpu_configs.split(","))
    if want & {"cfg1", "cfg1-1t", "cfg1-all"}:
        rows += cpu_cfg1_rows(want, nthr, dev, log)
    if "cfg2" in want:
        rows += cpu_cfg2_rows(a, nthr, dev, log)
    main = rows[1] if len(rows) > 1 else rows[0]
    out = {
        "metric": "CPU baselines of the P3M NVE step (C restatement of the reference algorithm)",
        "value": main["particle_steps_per_s"], "unit": "particle-steps/s", "n_gpus": 1,
        "steps": 5 if want == {"cfg2"} else 1000, "warmup": 0 if want == {"cfg2"} else 500,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "cfg1 in full and cfg2 for 5 steps on the host CPU, beside the GPU",
                   "host": host, "rows": rows, "parallelism": "host threads"},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def cpu_cfg1_rows(want, nthr, dev, log):
    """cfg1 in full on one thread (cfg1-1t), on all threads and the GPU (cfg1-all)."""
    import numpy as np
    import torch

    from mdforce import autotune
    from mdforce.engine import P3M
    from mdforce.initialize_pos_vel import initial_split
    from oracle import c_oracle
    rows = []
    N, kappa, gamma, M, alpha, rc, dt = 10000, 2.0, 100.0, 64, 0.85, 6.2, 0.1
    T = 1.0 / gamma
    L = (4.0 * math.pi * N / 3.0) ** (1.0 / 3.0)
    pos0, vel0 = initial_split(1, 2, N, [L, L, L], T)
    G_k, _ = c_oracle.gf_opt(kappa, alpha, 6, 3, M, L)
    q = 2.0 * math.pi / L * np.arange(1, 21, dtype=np.float64)
    Z = np.ones(N)

    def cpu_cfg1(threads, label):
        p, v = pos0.copy(), vel0.copy()
        _, acc, _, _, _ = c_oracle.force_pot(p, Z, L, kappa, alpha, rc, G_k, M)
        t0 = time.perf_counter()
        for it in range(500):
            c_oracle.md_steps(p, v, acc, Z, 1, dt, L, kappa, alpha, rc, G_k, M, threads=threads)
            Tn = float(np.sum(v ** 2)) / N  # thermostat_p3m_ord6.py:12-17
            v *= math.sqrt(T / Tn)
            if it % 100 == 99:
                log(f"[cpu] cfg1 {label}: equilibration step {it + 1}")
        t_eq = time.perf_counter() - t0
        en_all, t_nve = [], 0.0
        for chunk in range(10):
            t0 = time.perf_counter()
            en, _ = c_oracle.md_steps(p, v, acc, Z, 100, dt, L, kappa, alpha, rc, G_k, M, q=q, threads=threads)
            t_nve += time.perf_counter() - t0
            en_all.append(en)
            log(f"[cpu] cfg1 {label}: NVE step {100 * (chunk + 1)} ({t_nve:.1f} s)")
        en = np.concatenate(en_all)
        drift = float(np.max(np.abs(en[:, 0] - en[0, 0])) / (N * T))
        return p, {"config": "cfg1 N=1e4 (500 eq + 1000 NVE)", "form": f"CPU C restatement, {label}",
                   "particle_steps_per_s": N * 1000 / t_nve, "nve_s": t_nve, "eq_s": t_eq,
                   "drift_per_NT": drift}

    if want & {"cfg1", "cfg1-1t"}:
        _, r = cpu_cfg1(None, "1 thread (serial build, -O2)")
        r["threads"] = 1
        rows.append(r)
    if not want & {"cfg1", "cfg1-all"}:
        return rows
    _, r = cpu_cfg1(0, f"{nthr} threads (OpenMP build, -O3)")
    r["threads"] = nthr
    rows.append(r)
    # the GPU plain form on the same protocol
    e = P3M(N, M, L, kappa, alpha, rc, device=dev)
    e.set_green(torch.from_numpy(G_k).to(dev))
    p = torch.from_numpy(pos0).to(dev)
    v = torch.from_numpy(vel0).to(dev)
    acc, _ = e.force(p)
    acc = acc.clone()
    for it in range(500):
        e.md_run(p, v, acc, 1, dt)
        ks = (v * v).sum().reshape(1)
        e.thermostat(v, ks, T, it)
    qd = torch.from_numpy(q).to(dev)
    e.md_run(p, v, acc, 5, dt, q=qd)
    torch.cuda.synchronize()
    s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s0.record()
    en, _ = e.md_run(p, v, acc, 1000, dt, q=qd)
    s1.record()
    torch.cuda.synchronize()
    ms = s0.elapsed_time(s1)
    enh = en.cpu().numpy()
    ref = autotune.reference_force(P3M, p, L, kappa, device=dev)
    a1, _ = e.force(p, None, af=1.0)
    rows.append({"config": "cfg1 N=1e4 (500 eq + 1000 NVE)", "form": "plain FP64 GPU (MI355X)",
                 "particle_steps_per_s": N * 1000 / (ms / 1e3), "nve_s": ms / 1e3,
                 "drift_per_NT": float(np.max(np.abs(enh[:, 0] - enh[0, 0])) / (N * T)),
                 "force_error_rel_meanF": autotune.force_error(a1, ref)})
    e.close()
    log(f"[cpu] cfg1 GPU plain: {ms:.1f} ms for 1000 steps")

    return rows


def cpu_cfg2_rows(a, nthr, dev, log):
    """5 NVE steps of the N = 1e6 liquid on one thread and on all threads."""
    import numpy as np
    import torch

    from mdforce import mlp_train
    from mdforce.engine import P3M
    from mdforce.initialize_pos_vel import initial_split
    from oracle import c_oracle
    rows = []
    kappa, T, dt = 2.0, 0.01, 0.1
    N, M, alpha, rc = 1_000_000, 128, 0.43288602577760327, 6.75
    L = (4.0 * math.pi * N / 3.0) ** (1.0 / 3.0)
    pos0, vel0 = initial_split(1, 2, N, [L, L, L], T)
    p = torch.from_numpy(pos0).to(dev)
    v = torch.from_numpy(vel0).to(dev)
    eq = P3M(N, M, L, kappa, alpha, rc, precision="mixed", device=dev)
    eq.build_green()
    mlp_train.equilibrate(eq, p, v, T, a.eq_steps, dt, vcap=4.0, soft_steps=100)
    eq.close()
    e = P3M(N, M, L, kappa, alpha, rc, device=dev)
    G_full, _ = e.build_green(6, 3, want_full=True)
    acc0 = e.force(p)[0].cpu().numpy()
    e.close()
    Gh = G_full.cpu().numpy()
    ph, vh = p.cpu().numpy(), v.cpu().numpy()
    qh = 2.0 * math.pi / L * np.arange(1, 21, dtype=np.float64)
    for threads, label in ((None, "1 thread (serial build, -O2)"), (0, f"{nthr} threads (OpenMP build, -O3)")):
        pp, vv, aa = ph.copy(), vh.copy(), acc0.copy()
        t0 = time.perf_counter()
        for k in range(5):
            c_oracle.md_steps(pp, vv, aa, np.ones(N), 1, dt, L, kappa, alpha, rc, Gh, M, q=qh, threads=threads)
            log(f"[cpu] cfg2 {label}: step {k + 1} ({time.perf_counter() - t0:.1f} s)")
        tc = time.perf_counter() - t0
        rows.append({"config": "cfg2 N=1e6 liquid, 128^3, alpha 0.4329, rc 6.75 (5 NVE steps)",
                     "form": f"CPU C restatement, {label}", "threads": 1 if threads is None else nthr,
                     "particle_steps_per_s": N * 5 / tc, "nve_s": tc})
    return rows


def run_drift(a):
    """Energy conservation and force error along a trajectory, both forms.

    Two starts of N = 1e6 (Gamma 100, kappa 2, 128^3, alpha/rc of the n1e6
    workload):
      uniform       NVE straight from the seeded uniform positions and
                    Maxwellian velocities (the n1e6 workload's own start),
                    --uniform-steps steps (1000: the release of the random
                    start's close pairs heats this system without bound at
                    dt 0.1, so its energy is not conserved by either form);
      equilibrated  the same start after --eq-steps steps with a velocity
                    rescale to T every 10 steps (thermostat_p3m_ord6.py:15-18),
                    preceded by 100 steps at dt/10 and with each ion's speed
                    clipped to 4 sqrt(3T) at every rescale, so that the release
                    of the closest random pairs leaves no fast ions behind.
      From the equilibrated start each form runs --drift-steps NVE steps.
    Every 100 steps of every trajectory the force is compared with tight FP64
    P3M (autotune.reference_settings: 512^3, alpha 0.32, rc 15 at N = 1e6)
    on the current positions.  Reported per start and form: max |E(t) - E(0)| /
    (N T), the slope of a linear fit of (E - E(0)) / (N T) per 1000 steps,
    the force error (RMS over mean |F|) every 100 steps, the largest speed at
    the end, and ms per step."""
    import numpy as np
    import torch

    from mdforce import mlp_train as T_
    from mdforce.engine import P3M
    from mdforce.initialize_pos_vel import initial_split

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    N, M = a.n, a.mesh
    L = (4.0 * math.pi * N / 3.0) ** (1.0 / 3.0)
    T = 1.0 / a.gamma
    pos0, vel0 = initial_split(1, 2, N, [L, L, L], T)
    from mdforce.autotune import reference_settings
    ref_set = reference_settings(L)
    ref = P3M(N, ref_set[0], L, a.kappa, ref_set[1], ref_set[2], device=dev)
    ref.build_green()
    starts = {}
    if a.start in ("uniform", "both"):
        starts["uniform"] = (torch.from_numpy(pos0).to(dev), torch.from_numpy(vel0).to(dev))
    if a.start in ("equilibrated", "both"):
        pos = torch.from_numpy(pos0).to(dev)
        vel = torch.from_numpy(vel0).to(dev)
        eq = P3M(N, M, L, a.kappa, a.alpha, a.rc, device=dev)
        eq.build_green()
        T_.equilibrate(eq, pos, vel, T, a.eq_steps, a.dt, vcap=4.0, soft_steps=100)
        eq.close()
        starts["equilibrated"] = (pos, vel)
    chunk = 100
    out = {}
    for sname, (pos, vel) in starts.items():
        nchunk = max(1, (a.uniform_steps if sname == "uniform" else a.drift_steps) // chunk)
        out[sname] = {}
        for prec in ("fp64", "mixed"):
            e = P3M(N, M, L, a.kappa, a.alpha, a.rc, precision=prec, device=dev)
            e.build_green()
            if a.drift_skin > 0:
                e.set_skin(a.drift_skin)
            p, v = pos.clone(), vel.clone()
            acc, _ = e.force(p)
            acc = acc.clone()
            en_all, errs = [], []
            t_ms = 0.0
            for c in range(nchunk):
                fr, _ = ref.force(p)
                d = acc - fr  # both af (pp + pm) with the default af
                errs.append(float(torch.sqrt((d * d).sum(1).mean()) / torch.sqrt((fr * fr).sum(1)).mean()))
                s0 = torch.cuda.Event(enable_timing=True)
                s1 = torch.cuda.Event(enable_timing=True)
                s0.record()
                en, _ = e.md_run(p, v, acc, chunk, a.dt)
                s1.record()
                torch.cuda.synchronize()
                t_ms += s0.elapsed_time(s1)
                en_all.append(en[:, 0].cpu().numpy().copy())
                if c % 10 == 0:
                    print(f"[drift] {sname} {prec} step {(c + 1) * chunk}: dE/(N T) "
                          f"{(en_all[-1][-1] - en_all[0][0]) / (N * T):.3e} T {float(en[-1, 3]):.5f} "
                          f"vmax {float(v.norm(dim=1).max()):.3f} err {errs[-1]:.2e}", file=sys.stderr, flush=True)
            E = np.concatenate(en_all)
            de = (E - E[0]) / (N * T)
            steps = np.arange(1, len(E) + 1)
            slope = float(np.polyfit(steps, de, 1)[0] * 1000.0)
            out[sname]["chosen" if prec == "mixed" else "plain"] = {
                "ms_per_step": t_ms / (nchunk * chunk), "energy_drift_max_per_NT": float(np.abs(de).max()),
                "energy_drift_end_per_NT": float(de[-1]), "energy_drift_slope_per_NT_per_1000_steps": slope,
                "force_error_max": max(errs), "force_error_mean": float(np.mean(errs)),
                "force_error_every_100_steps": [float(f"{x:.3e}") for x in errs],
                "vmax_end": float(v.norm(dim=1).max()), "steps": nchunk * chunk}
            e.close()
    ref.close()
    first = out.get("equilibrated", next(iter(out.values())))
    res = {
        "metric": METRIC + " (energy drift over a trajectory)",
        "value": N / (first["chosen"]["ms_per_step"] / 1e3), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": first["chosen"]["steps"], "warmup": a.eq_steps, "ms_per_step": first["chosen"]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"N={N}, {M}^3, Gamma {a.gamma:g}, kappa {a.kappa:g}, NVE trajectories of both "
                               f"forms from " + ", ".join(f"the {k} start ({v['chosen']['steps']} steps)"
                                                          for k, v in out.items()),
                   "alpha": a.alpha, "rc": a.rc, "dt": a.dt, "pair_list_skin": a.drift_skin, "starts": out,
                   "parallelism": "single GPU"},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(res), flush=True)


def run_sqw(a):
    """S(q, omega) post-processing of an ensemble's n(q, t) on the device.

    cfg1's shape (s_of_q_w_script.py:6-25): 20 runs, Nq = int(2.5 / dq) = 13 q
    (the series n_q_t[:, 3 i, :] of its 39 columns, 3 axes each) of Npd = 16000
    steps, N = 1e4, dt = 0.05.
    Synthetic complex n(q, t) already resident in HBM; a step is one complete
    S(q, omega) of the ensemble (every FFT stage and the |X|^2 pass).  The
    roofline counts the algorithmic bytes: the 3 Nq used columns read once
    (16 B a sample) and S written once (8 B per run, q and omega).
    The CPU baseline is the NumPy oracle (numpy.fft, one thread) on the same
    ensemble, repeated for at least 5 s.
    """
    import numpy as np
    import torch
    from mdforce.sqw import sqw
    dev = torch.device("cuda", 0)
    n_avg, Nq, npd, N, dt = 20, 13, 16000, 10000, 0.05
    T = npd * dt
    g = torch.Generator(device=dev).manual_seed(1)
    x = torch.randn((n_avg, npd, 3 * Nq, 3), dtype=torch.complex128, device=dev, generator=g)
    out = torch.empty((n_avg, Nq, npd), dtype=torch.float64, device=dev)
    st = torch.cuda.current_stream(dev)
    for _ in range(a.warmup):
        sqw(x, N, dt, T, q_stride=3, nq=Nq, out=out)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(a.steps):
        sqw(x, N, dt, T, q_stride=3, nq=Nq, out=out)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.steps
    samples = n_avg * Nq * 3 * npd
    byts = samples * 16 + n_avg * Nq * npd * 8
    cpu = None
    if not a.no_cpu:
        from oracle import sqw_oracle as O
        xs = x.cpu().numpy()
        reps, t0 = 0, time.perf_counter()
        while reps < 3 or time.perf_counter() - t0 < 5.0:
            O.s_of_q_w_avg_mat(xs, Nq, np.zeros((n_avg, Nq, npd)), n_avg, N, dt, T)
            reps += 1
        tc = (time.perf_counter() - t0) / reps
        cpu = {"value": samples / tc, "unit": "samples/s", "cores": 1, "kind": "port",
               "sample": f"oracle/sqw_oracle.py s_of_q_w_avg_mat (numpy.fft) on the whole ensemble, "
                         f"{tc:.3f} s a call, {reps} calls"}
    res = {
        "metric": "S(q,w) of an n(q,t) ensemble: time samples transformed per s, 1 GPU",
        "value": samples / (ms / 1e3), "unit": "samples/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"S(q,w): {n_avg} runs x {Nq} q x 3 axes x {npd} steps (FP64 mixed-radix FFT "
                               f"along t, mean over axes)", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": byts / (ms / 1e3) / 1e9, "peak": 8000.0, "unit": "GB/s",
                     "frac": byts / (ms / 1e3) / 1e9 / 8000.0, "traffic": None,
                     "kernel": "whole S(q,w) call (7 FFT stages + |X|^2)"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (every box here has one GPU)",
              file=sys.stderr)
        sys.exit(2)
    if a.workload == "sqw":
        run_sqw(a)
        return
    if a.workload == "replicas":
        run_replicas(a)
        return
    if a.workload == "mlp":
        run_mlp(a)
        return
    if a.workload == "radial":
        run_radial(a)
        return
    if a.workload == "cpu":
        run_cpu(a)
        return
    if a.workload == "drift":
        run_drift(a)
        return
    if a.workload == "n8e6":
        # large weakly screened system: N = 8e6, Gamma 50, kappa 0.5, 256^3, dt 0.05, autotuned alpha/rc
        a.n, a.mesh, a.gamma, a.kappa, a.dt, a.autotune = 8_000_000, 256, 50.0, 0.5, 0.05, True
    import numpy as np
    import torch

    from mdforce.engine import P3M
    from mdforce.initialize_pos_vel import initial_split

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    N, M = a.n, a.mesh
    L = (4.0 * math.pi * N / 3.0) ** (1.0 / 3.0)
    T = 1.0 / a.gamma
    pos0, vel0 = initial_split(1, 2, N, [L, L, L], T)

    dq = 2.0 * math.pi / L
    # Equilibrated liquid of the same system: the force budget is relative to
    # the mean force, which is ~20x smaller there than on the uniform start.
    # The timed loop starts from this liquid too, so that the step time, the
    # force error and the energy drift all refer to one state.
    pos_liq = vel_liq = None
    if not a.no_liquid:
        from mdforce import mlp_train
        t_eq = time.perf_counter()
        pos_liq = torch.from_numpy(pos0).to(dev)
        vel_liq = torch.from_numpy(vel0).to(dev)
        eq = P3M(N, M, L, a.kappa, a.alpha, a.rc, precision="mixed", device=dev)
        eq.build_green()
        mlp_train.equilibrate(eq, pos_liq, vel_liq, T, a.eq_steps, a.dt, vcap=4.0, soft_steps=100)
        eq.close()
        torch.cuda.synchronize()
        print(f"[bench] equilibrated {a.eq_steps} + 100 steps in {time.perf_counter() - t_eq:.1f} s",
              file=sys.stderr, flush=True)
    tuned = None
    if a.autotune:
        from mdforce import autotune
        t_tune = time.perf_counter()
        # the configured mesh first; twice as fine only if no (alpha, rc) meets the budget on it
        # cheapest (alpha, rc) of the cost model on the configured mesh (twice as
        # fine only if none meets the budget), then its rc neighbours with
        # refined alpha, timed with their launch choices: the fastest wins
        tp = pos_liq if pos_liq is not None else torch.from_numpy(pos0).to(dev)
        tv = vel_liq if vel_liq is not None else torch.from_numpy(vel0).to(dev)
        ch = autotune.tune_measured(tp, tv, L, a.kappa, a.dt, budget=1e-5, meshes=(M,), fallback_meshes=(2 * M,),
                                    device=dev, log=lambda m: print(m, file=sys.stderr, flush=True))
        a.alpha, a.rc = ch.alpha, ch.rc
        if ch.mesh != M:
            print(f"[bench] no (alpha, rc) on {M}^3 meets the budget on this state; tuned on {ch.mesh}^3",
                  file=sys.stderr, flush=True)
            M = a.mesh = ch.mesh
        tuned = {"mesh": ch.mesh, "alpha": ch.alpha, "rc": ch.rc, "error": ch.error, "timed": ch.tried,
                 "tuned_on": "equilibrated liquid" if pos_liq is not None else "uniform start",
                 "tune_s": time.perf_counter() - t_tune}
    if a.calib_copy:
        src = torch.ones(128 * 1024 * 1024, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)  # reads and writes 1 GiB each
        torch.cuda.synchronize()
        del src, dst

    refs = {}
    ref_set = None
    if not a.no_err:
        from mdforce.autotune import reference_settings
        ref_set = reference_settings(L)
        ref = P3M(N, ref_set[0], L, a.kappa, ref_set[1], ref_set[2], device=dev)
        ref.build_green(6, 3)
        refs["uniform_start"] = (torch.from_numpy(pos0).to(dev), ref.force(torch.from_numpy(pos0).to(dev))[0].clone())
        if pos_liq is not None:
            refs["equilibrated"] = (pos_liq, ref.force(pos_liq)[0].clone())
        ref.close()
        del ref
        torch.cuda.synchronize()

    def force_error(eng):
        """RMS force error of eng against the reference on each state."""
        out = {}
        for k, (p, acc_t) in refs.items():
            acc, _ = eng.force(p)
            d = acc - acc_t
            rms_d = float(torch.sqrt(torch.mean(torch.sum(d * d, 1))))
            rms_f = float(torch.sqrt(torch.mean(torch.sum(acc_t * acc_t, 1))))
            mean_f = float(torch.mean(torch.sqrt(torch.sum(acc_t * acc_t, 1))))
            out[k] = {"rms_rel_rms": rms_d / rms_f, "rms_rel_meanF": rms_d / mean_f, "mean_F": mean_f}
        if out:
            out["max_rel_meanF"] = max(v["rms_rel_meanF"] for v in out.values())
            out["reference"] = f"FP64 P3M {ref_set[0]}^3, alpha {ref_set[1]:.3f}, rc {ref_set[2]:.2f}"
        return out

    start_pos = pos_liq if pos_liq is not None else torch.from_numpy(pos0).to(dev)
    start_vel = vel_liq if vel_liq is not None else torch.from_numpy(vel0).to(dev)

    def run_form(precision):
        eng = P3M(N, M, L, a.kappa, a.alpha, a.rc, precision=precision, device=dev)
        G_full, _ = eng.build_green(6, 3, want_full=True)
        err = force_error(eng)
        tl = None
        if a.autotune:
            # skin (and cell size), PP waves and PM block shapes by timing on this state
            from mdforce import autotune
            tl = autotune.tune_launch(eng, start_pos, start_vel, a.dt, steps=20)
            print(f"[bench] {precision}: launch tuned to {tl['requested']} ({tl['ms_per_step']:.4f} ms/step in "
                  f"{len(tl['timings'])} timings)", file=sys.stderr, flush=True)
        else:
            if a.skin > 0:
                eng.set_skin(a.skin)
            if a.pp_waves:
                eng.set_launch(a.pp_waves)
        skin = tl["skin"] if tl else a.skin
        pp_cfg = eng.config()
        pos = start_pos.clone()
        vel = start_vel.clone()
        acc, _ = eng.force(pos)
        acc = acc.clone()
        torch.cuda.synchronize()
        acc_init = acc.cpu().numpy() if precision == "fp64" else None
        if a.warmup:
            eng.md_run(pos, vel, acc, a.warmup, a.dt, q_step=dq, nq=a.nq)
        energies = torch.empty((a.steps, 4), dtype=torch.float64, device=dev)
        nqt = torch.empty((a.steps, a.nq, 3), dtype=torch.complex128, device=dev)
        torch.cuda.synchronize()
        ls0 = eng.list_stats() if skin > 0 else None
        # the timed steps carry two events a step, around the PP stage (the
        # dominant kernel's duration for the roofline); the other stages are
        # timed afterwards in steps of their own, with events at every stage
        eng.profile_start(events=4 * (a.steps + 4), pp_only=True)
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        start.record()
        eng.md_run(pos, vel, acc, a.steps, a.dt, q_step=dq, nq=a.nq, energies=energies, nqt=nqt)
        stop.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gpu_ms = start.elapsed_time(stop)
        pp_timed = eng.profile_read()["pp"]
        eng.profile_stop()
        ls1 = eng.list_stats() if skin > 0 else None
        ns = min(a.steps, 20)
        eng.profile_start(events=16 * (ns + 4))
        eng.md_run(pos, vel, acc, ns, a.dt, q_step=dq, nq=a.nq)
        stages = eng.profile_read()
        eng.profile_stop()
        stages["pp_timed"] = pp_timed
        lists = None
        launch = eng.launch_config()
        if skin > 0:
            lists = {"skin_requested": skin, "cells_x": ls1["cells_x"],
                     "skin_effective": L / ls1["cells_x"] - a.rc, "list_cap": ls1["list_cap"],
                     "builds_timed": ls1["builds"] - ls0["builds"],
                     "list_steps_timed": ls1["list_steps"] - ls0["list_steps"]}
        en_h = energies.cpu().numpy()
        drift = float(np.max(np.abs(en_h[:, 0] - en_h[0, 0])) / (N * T))
        eng.close()
        return {"gpu_ms": gpu_ms, "wall": wall, "stages": stages, "err": err, "drift": drift, "G_full": G_full,
                "pp_config": pp_cfg, "lists": lists, "launch": launch,
                "launch_tuning": ({"chosen": tl["requested"], "timings_ms": [[w, ms] for w, ms in tl["timings"]]}
                                  if tl else None),
                "acc0": acc_init, "value": N * a.steps / (gpu_ms / 1e3)}

    forms = {"both": ("fp64", "mixed"), "chosen": ("mixed",), "plain": ("fp64",)}[a.form]
    res = {f: run_form(f) for f in forms}
    main_form = "mixed" if "mixed" in res else "fp64"
    R = res[main_form]
    value = R["value"]
    ms_per_step = R["gpu_ms"] / a.steps
    stages = R["stages"]

    # roofline of the dominant stage of the reported form
    dom = max((s for s in stages if s not in ("finalize", "pp_timed")), key=lambda s: stages[s][0])
    # the PP kernel's launches inside the timed region itself
    tot_ms, cnt = stages["pp_timed"] if dom == "pp" else stages[dom]
    avg_s = tot_ms / max(cnt, 1) / 1e3
    by, fl = stage_work(dom, N, M, a.rc)
    if fl > 0:
        # the pair terms run on the vector ALUs (no matrix instructions)
        peak = PEAK_FP32 if main_form == "mixed" else PEAK_FP64
        roof = {"bound": "valu", "achieved": fl / avg_s / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s"}
    else:
        roof = {"bound": "hbm", "achieved": by / avg_s / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = pmc_traffic(dom, main_form, N, M, a.rc, a.alpha)
    pp_name = (f"k_pp<{'float' if main_form == 'mixed' else 'double'}, {R['pp_config']['pp_waves']}, "
               f"{'true' if R['lists'] else 'false'}>")
    roof["kernel"] = {"pp": pp_name, "fft": "fft passes", "gather": "k_gather", "assign": "k_assign",
                      "bin": "bin kernels", "drift": "k_kick_drift", "modes": "k_modes_grid"}.get(dom, dom)
    roof["avg_launch_ms"] = avg_s * 1e3

    def stage_ms(st):
        return {s: round(v[0] / max(v[1], 1), 4) for s, v in st.items() if v[1] and s != "pp_timed"}

    form_cfg = {}
    for f, r in res.items():
        form_cfg["chosen" if f == "mixed" else "plain"] = {
            "particle_steps_per_s": r["value"], "ms_per_step": r["gpu_ms"] / a.steps, "force_error": r["err"],
            "energy_drift_per_NT": r["drift"], "stage_ms": stage_ms(r["stages"]), "wall_s": r["wall"],
            "pp_config": r["pp_config"], "pair_lists": r["lists"], "launch": r["launch"],
            "launch_tuning": r["launch_tuning"]}
    if len(res) == 2:
        form_cfg["chosen_over_plain"] = res["mixed"]["value"] / res["fp64"]["value"]
    G_full = R["G_full"]
    acc0 = res["fp64"]["acc0"] if "fp64" in res else None
    q = torch.arange(1, a.nq + 1, dtype=torch.float64, device=dev) * dq

    cpu = None
    if not a.no_cpu:
        from oracle import c_oracle
        c_oracle.build(ref=False)
        Gh = G_full.cpu().numpy()
        p0h, v0h = start_pos.cpu().numpy(), start_vel.cpu().numpy()
        if acc0 is None:  # initial FP64 force, equal to the C restatement's to 1e-12
            e64 = P3M(N, M, L, a.kappa, a.alpha, a.rc, device=dev)
            e64.build_green(6, 3)
            acc0 = e64.force(start_pos)[0].cpu().numpy()
            e64.close()
        qh = q.cpu().numpy()
        host = host_cpu()
        state = "equilibrated liquid" if pos_liq is not None else "uniform start"

        def cpu_run(steps, threads):
            p_h, v_h, a_h = p0h.copy(), v0h.copy(), acc0.copy()
            t0 = time.perf_counter()
            c_oracle.md_steps(p_h, v_h, a_h, np.ones(N), steps, a.dt, L, a.kappa, a.alpha, a.rc, Gh, M, q=qh,
                              threads=threads)
            return time.perf_counter() - t0

        nthr = c_oracle.lib_omp(a.cpu_threads)[1]
        print(f"[bench] CPU baseline: {a.cpu_steps} steps on {nthr} threads", file=sys.stderr, flush=True)
        tc = cpu_run(a.cpu_steps, a.cpu_threads)
        cpu = {"value": N * a.cpu_steps / tc, "unit": "particle-steps/s", "cores": nthr, "kind": "port",
               "sample": f"{a.cpu_steps} NVE steps of the same N={N}, {M}^3 workload from the {state}, the C "
                         f"restatement of the reference algorithm (oracle/c/p3m_oracle.c, OpenMP, -O3) on "
                         f"{nthr} threads, {tc:.1f} s; host {host['model']}, nproc {host['nproc']}"}
        if a.cpu_steps1 > 0:
            t1 = cpu_run(a.cpu_steps1, None)
            cpu["single_thread"] = {"value": N * a.cpu_steps1 / t1, "cores": 1,
                                    "sample": f"{a.cpu_steps1} NVE step(s), serial build (-O2), {t1:.1f} s"}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if main_form == "mixed" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"P3M Yukawa OCP NVE, N={N}, {M}^3 mesh (Gamma={a.gamma:g}, kappa={a.kappa:g}), "
                        f"{'FP32-mixed' if main_form == 'mixed' else 'plain FP64'}",
            "form": "chosen (FP32-mixed)" if main_form == "mixed" else "plain",
            "N": N, "mesh": M, "L": L, "kappa": a.kappa, "Gamma": a.gamma, "alpha": a.alpha, "rc": a.rc,
            "dt": a.dt, "nq": a.nq, "forms": form_cfg, "autotuned": tuned, "parallelism": "single GPU",
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
