"""Benchmark: RBE + LJ force path on the 200,000-particle LJ electrolyte, one MI355X.

Workload (the configuration the headline metric is quoted on): 200,000
particles at rho* = 0.8 (5 % +1, 5 % -1, 90 % neutral LJ solvent), uniform
positions with overlap rejection (d_min = 0.85 sigma), LJ cutoff 2.5 sigma,
real-space Coulomb cutoff 3.5 sigma, skin 0.3 sigma, RBL short range
(r_core = 2.5, neighbor_sample_num = 30), RBE k-space with alpha = 0.09
(g_ewald = 0.3/sigma) and P = 200 importance-sampled k-vectors per step drawn
from the seeded stream, FP32 pair and e^{ik.r} terms, FP64 reductions.

A step is one MD step: one full force evaluation (skin check, conditional
neighbour rebuild, RBL pair forces, RBE k-space, assembly) with the leapfrog
update of velocities and positions fused into the assembly kernel, so rebuilds
happen at their natural rate.  Inputs are resident in HBM before the timed
region.

Prints ONE JSON line.  --gpus above 1 exits: this build measures one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-steps/s, RBE+LJ force, 200k LJ-salt; %HBM peak; force err vs FP64"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_pair_traffic_electrolyte200k.md")


# the sources that define the pair kernel; the traffic profile records their
# hash, and a profile taken on other sources is not used
PAIR_SOURCES = ("rbmd_amd/csrc/kernels.hip", "rbmd_amd/csrc/device.hpp",
                "rbmd_amd/csrc/engine.hpp")


def pair_source_hash():
    import hashlib

    h = hashlib.sha256()
    for rel in PAIR_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def measured_traffic(atoms):
    """HBM bytes per k_pair launch from the committed rocprofv3 --pmc summary
    (read and write bytes per dispatch from the memory-request counters), when
    it was taken on this workload with the current pair-kernel sources; None
    otherwise."""
    try:
        text = open(TRAFFIC_FILE).read()
    except OSError:
        return None
    vals = {}
    for line in text.splitlines():
        if ":" in line and line.strip().startswith("- "):
            key, _, v = line.strip()[2:].partition(":")
            vals[key.strip()] = v.strip()
    try:
        if int(vals["atoms"]) != atoms or vals["pair_source_sha256"] != pair_source_hash():
            return None
        return int(vals["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--atoms", type=int, default=200000)
    ap.add_argument("--dt", type=float, default=0.002)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--workload",
                    choices=["electrolyte", "adjoint", "replicas", "water", "lj", "md"],
                    default="electrolyte",
                    help="electrolyte: the headline force step (default); adjoint: "
                         "forward and force-matching adjoint on the 2,000,000-particle "
                         "electrolyte with P = 400; replicas: 256 independent "
                         "electrolytes of 500-8000 atoms, batched vs per replica; water: "
                         "72,981-atom SPC/E, CVFF + RBL + RBSOG; lj: 4,000-atom FP64 LJ fcc, "
                         "100 NVE steps (each prints its own JSON line)")
    ap.add_argument("--tune-cache", default=None,
                    help="replicas: JSON cache of the batch autotuner (default: measure)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB three times (a "
                         "plain copy kernel of known size, to scale HBM counters), and "
                         "gather it three times in a random order of 16-byte rows (a "
                         "gather kernel of known size)")
    ap.add_argument("--print-source-hash", action="store_true",
                    help="print the hash of the pair-kernel sources and exit")
    return ap.parse_args()


def bench_adjoint(args):
    """Forward evaluation and its force-matching adjoint (dL/dr and per-type
    dL/dQ, dL/deps, dL/dsigma) on the 2,000,000-particle LJ electrolyte, RBL +
    RBE with P = 400: GPU time of each, and their ratio.  dL/dQ shifts the
    charged atoms of a type (the neutral solvent's charge is fixed at 0); the
    line also carries the time with the neutral type's dL/dQ at Q = 0
    (neutral_q: Coulomb pairs of charged and neutral atoms and the neutral
    atoms' structure sums added)."""
    import torch

    from rbmd_amd import ForceContext, rbe_sample_host, systems

    n = args.atoms if args.atoms != 200000 else 2000000
    dev = torch.device("cuda:0")
    # a seeded box of n / 8 particles replicated 2 x 2 x 2 (RBMD's replicate)
    s = systems.replicate(systems.electrolyte(n // 8, seed=5), (2, 2, 2))
    n = s.n
    cfg = systems.ljsalt_config(P=400)
    ctx = ForceContext(cfg, device=0)
    ctx.set_system(s.box_lo, s.box_len, s.types, s.ids, s.charges, s.eps, s.sigma)
    p = torch.tensor(s.pos.astype(np.float32), device=dev)
    px, py, pz = (p[:, d].contiguous() for d in range(3))
    f = [torch.zeros(n, dtype=torch.float32, device=dev) for _ in range(3)]
    G = torch.randn((3, n), dtype=torch.float32, device=dev, generator=torch.Generator(
        device=dev).manual_seed(3))
    dr = [torch.zeros(n, dtype=torch.float32, device=dev) for _ in range(3)]
    dparam = torch.zeros(9, dtype=torch.float64, device=dev)
    m = torch.tensor(rbe_sample_host(0.09, s.box_len, 400, 7, 0).reshape(-1),
                     dtype=torch.int32, device=dev)
    steps, warm = max(1, min(args.steps, 20)), max(1, min(args.warmup, 5))

    def fwd(k):
        ctx.force(k, px, py, pz, *f, m_samples=m)

    def vjp(k, neutral_q=False):
        ctx.force_vjp(k, px, py, pz, G[0], G[1], G[2], *dr, dparam, neutral_q=neutral_q)

    for k in range(warm):
        fwd(k)
        vjp(k)
        vjp(k, True)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record()
    for k in range(steps):
        fwd(k)
    ev[1].record()
    for k in range(steps):
        fwd(k)
        vjp(k)
    ev[2].record()
    for k in range(steps):
        fwd(k)
        vjp(k, True)
    ev[3].record()
    torch.cuda.synchronize()
    t_f = ev[0].elapsed_time(ev[1]) / steps
    t_b = ev[1].elapsed_time(ev[2]) / steps - t_f
    t_bn = ev[2].elapsed_time(ev[3]) / steps - t_f
    line = {
        "metric": "force-matching adjoint / forward GPU time, 2M LJ-salt RBE P=400",
        "value": round(t_b / t_f, 3), "unit": "ratio", "n_gpus": 1, "steps": steps,
        "warmup": warm, "higher_is_better": False, "dtype": "f32", "data": "synthetic",
        "forward_ms": round(t_f, 4), "adjoint_ms": round(t_b, 4),
        "adjoint_neutral_q_ms": round(t_bn, 4), "ratio_neutral_q": round(t_bn / t_f, 3),
        "config": {"workload": f"LJ electrolyte RBL+RBE, {n} atoms, P=400, adjoint of "
                               "L = sum |F - F_t|^2 w.r.t. positions and per-type Q, "
                               "eps, sigma", "atoms": n, "P": 400},
    }
    print(json.dumps(line))


def bench_replicas(args):
    """256 independent LJ electrolytes, N_r = round(exp(U[ln 500, ln 8000]))
    (seed 4; composition, density and cutoffs of the 200k electrolyte), RBL +
    RBE with P = 100 k-samples per replica and a fresh k-sample every step.
    Times one force evaluation of every replica three ways: the replica batch
    as one graph replay (launch shape from its autotuner), the same batch
    kernels launched one replica at a time, and the single-system engine with
    one context per replica (its own graph per replica)."""
    import torch

    from rbmd_amd import BatchContext, ForceContext, rbe_sample_host, systems

    dev = torch.device("cuda:0")
    reps = systems.replica_set(256, seed=4)
    P = 100
    cfg = systems.ljsalt_config(P=P)
    n = sum(s.n for s in reps)
    b = BatchContext(cfg, device=0)
    b.set_systems(reps)
    pos = torch.tensor(np.concatenate([s.pos for s in reps]).astype(np.float32), device=dev)
    px, py, pz = (pos[:, d].contiguous() for d in range(3))
    f = [torch.zeros(n, dtype=torch.float32, device=dev) for _ in range(3)]
    steps, warm = max(1, min(args.steps, 20)), max(1, min(args.warmup, 5))
    # k-samples of every replica for each evaluation, drawn on the host beforehand
    pool = []
    for k in range(4):
        ms = [rbe_sample_host(0.09, s.box_len, P, 100 + r, k) for r, s in enumerate(reps)]
        pool.append(torch.tensor(np.concatenate(ms).reshape(-1).astype(np.int32), device=dev))
    m = torch.empty_like(pool[0])
    m.copy_(pool[0])
    shape, tune_us = b.autotune(0, px, py, pz, *f, m_samples=m, cache_path=args.tune_cache)

    def timed(fn):
        for k in range(warm):
            fn(k)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for k in range(steps):
            fn(warm + k)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    def batched(k):
        m.copy_(pool[k % 4])
        b.force(k, px, py, pz, *f, m_samples=m)

    def per_replica(k):
        m.copy_(pool[k % 4])
        b.force(k, px, py, pz, *f, m_samples=m, per_replica=True)

    t_graph = timed(batched)
    t_plain = timed(per_replica)
    # the single-system engine, one context per replica (graph replay each)
    ctxs, views = [], []
    off = b.offsets
    for r, s in enumerate(reps):
        c = ForceContext(cfg, device=0)
        c.set_system(s.box_lo, s.box_len, s.types, s.ids, s.charges, s.eps, s.sigma)
        a, e = int(off[r]), int(off[r + 1])
        ctxs.append(c)
        views.append((px[a:e], py[a:e], pz[a:e], f[0][a:e], f[1][a:e], f[2][a:e],
                      m[3 * P * r:3 * P * (r + 1)]))

    def contexts(k):
        m.copy_(pool[k % 4])
        for c, v in zip(ctxs, views):
            c.force(k, *v[:6], m_samples=v[6])

    t_ctx = timed(contexts)
    line = {
        "metric": "batched replicas: speedup of one graph-replayed batch evaluation over "
                  "per-replica launches (256 LJ-salt boxes, 500-8000 atoms, RBE P=100)",
        "value": round(t_plain / t_graph, 2), "unit": "x", "n_gpus": 1, "steps": steps,
        "warmup": warm, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
        "ms_batched_graph": round(t_graph, 4), "ms_per_replica_launches": round(t_plain, 4),
        "ms_per_replica_contexts": round(t_ctx, 4),
        "speedup_vs_contexts": round(t_ctx / t_graph, 2),
        "atom_steps_per_s_batched": round(n / (t_graph * 1e-3), 1),
        "tuned_shape": shape, "tune_us": round(tune_us, 2),
        "config": {"workload": f"256 LJ electrolytes, {n} atoms (500-8000 each, seed 4), "
                               "RBL + RBE, P=100 per replica", "atoms": n, "replicas": 256,
                   "P": P},
    }
    print(json.dumps(line))


def bench_water(args):
    """72,981-atom rigid SPC/E water (24,327 molecules on a seeded lattice, random
    orientations, 33.4 / nm^3), CVFF with special_bonds [0, 0, 0], RBL (r_core
    6.5 A, 100 samples, rc 10 A, skin 2 A) and RBSOG level 1 with P = 500 (512):
    force evaluations at fixed positions with the seeded k-sample stream and a
    neighbour rebuild forced every 10th evaluation; the force error against
    the FP64 oracle (pair + SOG k-space) and the oracle's CPU time for one
    evaluation."""
    import torch

    import oracle
    from rbmd_amd import ForceContext, systems
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from refs import reference_force, rel_rms, cfg_qqr2e

    dev = torch.device("cuda:0")
    s = systems.spce_water(24327, seed=3)
    cfg = systems.cvff_water_config(coulomb="RBSOG", alpha=0.0784, P=500, rbsog_level=1)
    ctx = ForceContext(cfg, device=0)
    ctx.set_system(s.box_lo, s.box_len, s.types, s.ids, s.charges, s.eps, s.sigma)
    ctx.set_special(s.nspecial, s.special)
    n = s.n
    P = ctx.num_samples()
    p = torch.tensor(s.pos.astype(np.float32), device=dev)
    px, py, pz = (p[:, d].contiguous() for d in range(3))
    f = [torch.zeros(n, dtype=torch.float32, device=dev) for _ in range(3)]
    steps, warm = max(1, min(args.steps, 200)), max(1, min(args.warmup, 20))
    for k in range(warm):
        ctx.force(k, px, py, pz, *f, rebuild=(k % 10 == 0))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(warm, warm + steps):
        ctx.force(k, px, py, pz, *f, rebuild=(k % 10 == 0))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # force error at a fixed explicit sample
    sset = oracle.sog_setup(1, 10.0, s.box_len)
    m, idx = oracle.sog_sample(31, 0, P, s.box_len, sset)
    mt = torch.tensor(np.concatenate([m.reshape(-1), idx]).astype(np.int32), device=dev)
    ctx.force(1, px, py, pz, *f, m_samples=mt, num_k=P, sync=True)
    got = np.stack([t.double().cpu().numpy() for t in f], 1)
    pos32 = s.pos.astype(np.float32)
    tc = time.perf_counter()
    ref = reference_force(s, pos32, cfg, 1, None, with_kspace=False)[0]
    fk, _, _ = oracle.sog_kspace_f64(pos32.astype(np.float64), s.charges, s.box_len, m, idx,
                                     sset, qqr2e=cfg_qqr2e(cfg))
    t_cpu = time.perf_counter() - tc
    ferr = rel_rms(got, ref + fk)
    th = ctx.thermo()
    # the dominant kernel, k_pair_b (k_pair is not launched when every atom is
    # charged), timed by HIP events around the pair phase over a repeat of the
    # timed evaluations.  Algorithmic bytes per launch: 5 B per list-B entry
    # (4 B index + 1 B type) + 36 B per atom (16 B position/charge, 4 B type,
    # 16 B force); the kernel is bound by VALU issue, not by these bytes.
    ctx.kernel_times()
    for k in range(warm + steps, warm + 2 * steps):
        ctx.force(k, px, py, pz, *f, rebuild=(k % 10 == 0), timing=True)
    torch.cuda.synchronize()
    pair_us = ctx.kernel_times()["pair"]
    algo = 5 * int(th["pairs_b"]) + 36 * n
    ach = algo / (pair_us * 1e-6) / 1e9
    roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_pair_b", "kernel_us": round(pair_us, 2),
                "algorithmic_bytes": algo,
                "limiter": "VALU issue (profiles/valu_counters_water73k.md)"}
    line = {
        "metric": "atom-steps/s, CVFF SPC/E water, RBL + RBSOG, 72,981 atoms",
        "value": round(n * steps / elapsed, 1), "unit": "atom-steps/s", "n_gpus": 1,
        "steps": steps, "warmup": warm, "ms_per_step": round(1e3 * elapsed / steps, 4),
        "higher_is_better": True, "dtype": "f32", "data": "synthetic",
        "force_rel_rms_vs_fp64": float(ferr),
        "config": {"workload": "SPC/E water 24,327 molecules, CVFF special_bonds [0,0,0], "
                               "RBL r_core 6.5 A / 100 samples, rc 10 A, RBSOG level 1 "
                               "P=512, positions fixed, rebuild every 10th evaluation",
                   "atoms": n, "P": P, "rbl_frequency": th["rbl_frequency"],
                   "pairs_a": th["pairs_a"], "pairs_b": th["pairs_b"]},
        "roofline": roofline,
        "cpu_baseline": {"value": round(n / t_cpu, 1), "unit": "atom-steps/s", "cores": 1,
                         "kind": "port",
                         "sample": "one evaluation (pair forces with the cell search, SOG "
                                   "k-space in FP64) of the same configuration"},
    }
    print(json.dumps(line))


def bench_lj(args):
    """Plain LJ fluid: 4,000 atoms (10 x 10 x 10 fcc, rho* 0.8442, +-0.05 sigma
    jitter, Maxwell velocities at T* = 1), rc 2.5, skin 0.3, full Verlet list,
    FP64, energy and virial every step; 100 NVE steps at dt = 0.005 (the force
    fused with the leapfrog update, unit masses).  Reports the GPU rate, the
    total-energy drift over the run, the force error against the FP64 oracle
    and the oracle's CPU time (the reference has no CPU path)."""
    import torch

    import oracle
    from rbmd_amd import ForceContext, systems

    dev = torch.device("cuda:0")
    s = systems.lj_fcc(cells=10, seed=1)
    cfg = systems.lj_config(precision="fp64")
    ctx = ForceContext(cfg, device=0)
    ctx.set_system(s.box_lo, s.box_len, s.types, s.ids, s.charges, s.eps, s.sigma)
    ctx.set_masses(np.ones(1))
    n, dt, steps = s.n, 0.005, 100
    pos = torch.tensor(s.pos.T.copy(), dtype=torch.float64, device=dev)
    vel = torch.tensor(s.vel.T.copy(), dtype=torch.float64, device=dev)
    frc = torch.zeros((3, n), dtype=torch.float64, device=dev)
    # initial force, then half a kick back so the leapfrog velocities are v(-dt/2)
    ctx.force(0, pos[0], pos[1], pos[2], frc[0], frc[1], frc[2], energy=True, sync=True)
    e0 = ctx.thermo()["e_vdwl"] + 0.5 * float((vel ** 2).sum())
    vel.sub_(frc, alpha=0.5 * dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(1, steps + 1):
        ctx.force_leapfrog(k, pos[0], pos[1], pos[2], vel[0], vel[1], vel[2], frc[0],
                           frc[1], frc[2], dt, energy=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # total energy at the end: the force at the final positions and v(t) from
    # the half-step velocities
    ctx.force(steps + 1, pos[0], pos[1], pos[2], frc[0], frc[1], frc[2], energy=True,
              sync=True)
    v_t = vel + 0.5 * dt * frc
    e1 = ctx.thermo()["e_vdwl"] + 0.5 * float((v_t ** 2).sum())
    P = pos.T.double().cpu().numpy()
    F = frc.T.double().cpu().numpy()
    tc = time.perf_counter()
    ref, _, _ = oracle.verlet_force_f64(P, s.charges, s.types, s.eps, s.sigma, s.box_len, 2.5)
    t_cpu = time.perf_counter() - tc
    ferr = float(np.sqrt(((F - ref) ** 2).sum() / (ref ** 2).sum()))
    line = {
        "metric": "atom-steps/s, LJ fcc 4,000 atoms, FP64 full list with energy and virial",
        "value": round(n * steps / elapsed, 1), "unit": "atom-steps/s", "n_gpus": 1,
        "steps": steps, "warmup": 1, "ms_per_step": round(1e3 * elapsed / steps, 4),
        "higher_is_better": True, "dtype": "f64", "data": "synthetic",
        "energy_drift_rel": abs(e1 - e0) / abs(e0), "force_rel_rms_vs_fp64": ferr,
        "config": {"workload": "LJ fcc 10x10x10 (4,000 atoms), rho* 0.8442, rc 2.5, skin "
                               "0.3, Verlet-List, FP64, NVE dt 0.005, 100 steps",
                   "atoms": n, "rebuilds": ctx.thermo()["rebuilds"]},
        "cpu_baseline": {"value": round(n / t_cpu, 1), "unit": "atom-steps/s", "cores": 1,
                         "kind": "port",
                         "sample": "one FP64 force evaluation of the final configuration "
                                   "(cell search + full pair list)"},
    }
    print(json.dumps(line))


def bench_md(args):
    """The whole NVT step of the 200k electrolyte (RBL + RBE, P = 200): the
    Nose-Hoover chain (temperature [1, 1, 1]), velocity Verlet with image
    flags and the force, through rbmdx_md_run, against the same number of
    force-only evaluations.  Positions start from the warmed-up electrolyte
    (capped-velocity relaxation of the random start, not timed)."""
    import torch

    from rbmd_amd import ForceContext, systems

    dev = torch.device("cuda:0")
    s = systems.electrolyte(args.atoms, seed=2)
    cfg = systems.ljsalt_config()
    cfg["execution"] = {"ensemble": "NVT", "integration_type": "vv",
                        "temp_ctrl_type": "NOSE_HOOVER", "temperature": [1.0, 1.0, 1.0],
                        "timestep": args.dt}
    ctx = ForceContext(cfg, device=0)
    ctx.set_system(s.box_lo, s.box_len, s.types, s.ids, s.charges, s.eps, s.sigma)
    ctx.set_masses(np.ones(len(s.eps)))
    n, dt = s.n, args.dt
    x = [torch.tensor(s.pos[:, d].copy(), dtype=torch.float32, device=dev) for d in range(3)]
    v = [torch.tensor(s.vel[:, d].copy(), dtype=torch.float32, device=dev) * 0.5
         for d in range(3)]
    f = [torch.zeros(n, dtype=torch.float32, device=dev) for _ in range(3)]
    img = [torch.zeros(n, dtype=torch.int32, device=dev) for _ in range(3)]
    vmax = 0.05 / dt
    for k in range(args.warmup):
        ctx.force(k, *x, *f)
        for d in range(3):
            v[d].add_(f[d], alpha=dt).clamp_(-vmax, vmax)
            x[d].add_(v[d], alpha=dt)
    ctx.md_init(x, v, f, image=img)
    ctx.md_run(20, x, v, f, image=img)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.md_run(args.steps, x, v, f, image=img)
    torch.cuda.synchronize()
    t_md = time.perf_counter() - t0
    st = ctx.md_state()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ctx.force(10 ** 6 + k, *x, *f)
    torch.cuda.synchronize()
    t_force = time.perf_counter() - t0
    line = {
        "metric": "atom-steps/s, NVT Nose-Hoover MD step (integrator + RBL/RBE force), 200k",
        "value": round(n * args.steps / t_md, 1), "unit": "atom-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_md / args.steps, 4),
        "force_only_ms_per_step": round(1e3 * t_force / args.steps, 4),
        "higher_is_better": True, "dtype": "f32 (thermostat f64)", "data": "synthetic",
        "temperature_end": st["temperature"], "thermostat_energy": st["thermostat_energy"],
        "config": {"workload": f"LJ electrolyte RBL+RBE, {n} atoms, P=200, NVT "
                               f"Nose-Hoover [1, 1, 1], dt {dt}", "atoms": n},
    }
    print(json.dumps(line))


def main():
    args = parse()
    if args.print_source_hash:
        print(pair_source_hash())
        return
    if args.workload == "md":
        if args.gpus != 1:
            print("bench.py measures one GPU only", file=sys.stderr)
            sys.exit(2)
        bench_md(args)
        return
    if args.workload in ("adjoint", "replicas", "water", "lj"):
        if args.gpus != 1:
            print("bench.py measures one GPU only", file=sys.stderr)
            sys.exit(2)
        {"adjoint": bench_adjoint, "replicas": bench_replicas, "water": bench_water,
         "lj": bench_lj}[args.workload](args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); how the "
              "path would shard over several GPUs is a design note, not a measurement",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from rbmd_amd import ForceContext, systems

    dev = torch.device("cuda:0")
    sys_ = systems.electrolyte(args.atoms, seed=2)
    cfg = systems.ljsalt_config()
    ctx = ForceContext(cfg, device=0)
    ctx.set_system(sys_.box_lo, sys_.box_len, sys_.types, sys_.ids, sys_.charges,
                   sys_.eps, sys_.sigma)
    n = sys_.n
    pos = torch.tensor(sys_.pos.T.copy(), dtype=torch.float32, device=dev)  # (3, N)
    vel = torch.tensor(sys_.vel.T.copy(), dtype=torch.float32, device=dev) * 0.5
    frc = torch.zeros((3, n), dtype=torch.float32, device=dev)
    dt = args.dt
    vmax = 0.05 / dt  # cap during warm-up only: relaxes overlaps of the random start

    def warm_step(k):
        ctx.force(k, pos[0], pos[1], pos[2], frc[0], frc[1], frc[2])
        vel.add_(frc, alpha=dt)
        vel.clamp_(-vmax, vmax)
        pos.add_(vel, alpha=dt)

    # warm-up: W untimed steps, the first W - 20 with the velocity cap (relaxes
    # the random start), the last 20 (or all W when fewer) with the timed call
    n_leap = min(args.warmup, 20)
    for k in range(args.warmup - n_leap):
        warm_step(k)
    torch.cuda.synchronize()

    def md_steps(first, count, timing):
        for k in range(first, first + count):
            # force evaluation + leapfrog update fused in one call (unit masses)
            ctx.force_leapfrog(k, pos[0], pos[1], pos[2], vel[0], vel[1], vel[2], frc[0],
                               frc[1], frc[2], dt, timing=timing)

    # the timed call itself, warmed up: its graph (keyed on the velocity
    # buffers and dt) is captured here, not inside the timed region
    md_steps(args.warmup - n_leap, n_leap, False)
    torch.cuda.synchronize()
    first = args.warmup

    # timed region: no timing events (each event adds a gap on the stream)
    th0 = ctx.thermo()
    t0 = time.perf_counter()
    md_steps(first, args.steps, False)
    host_elapsed = time.perf_counter() - t0  # enqueue time on the host
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    th1 = ctx.thermo()
    value = n * args.steps / elapsed
    rebuilds = th1["rebuilds"] - th0["rebuilds"]
    # the same run continued for the same number of steps with per-phase HIP
    # events on the launch streams: average kernel durations
    ctx.kernel_times()
    t1 = time.perf_counter()
    md_steps(first + args.steps, args.steps, True)
    torch.cuda.synchronize()
    elapsed_ev = time.perf_counter() - t1
    kt = ctx.kernel_times()
    th2 = ctx.thermo()

    # host cost of one call, measured over 10 calls from an idle device (fewer
    # than the 16 in-flight evaluations the context allows, so no call waits)
    def host_cost(eager):
        base = first + 2 * args.steps + (100 if eager else 0)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for k in range(base, base + 10):
            ctx.force_leapfrog(k, pos[0], pos[1], pos[2], vel[0], vel[1], vel[2], frc[0],
                               frc[1], frc[2], dt, eager=eager)
        h = (time.perf_counter() - t) / 10
        torch.cuda.synchronize()
        return h

    host_graph_us = 1e6 * min(host_cost(False) for _ in range(3))
    host_eager_us = 1e6 * min(host_cost(True) for _ in range(3))

    # dominant kernel: the pair kernel k_pair.  Algorithmic bytes per launch,
    # the compulsory-traffic model of the pair pass: 4 B per pair-list entry
    # (list A + list B; this build stores list-A entries in 2 B) + 36 B per
    # atom (16 B position/charge tile, 4 B type, 16 B force); neighbour
    # positions assumed cache- or LDS-resident.
    pairs_a, pairs_b = th2["pairs_a"], th2["pairs_b"]
    bytes_pair = 4.0 * (pairs_a + pairs_b) + 36.0 * n
    t_pair = kt.get("pair", 0.0) * 1e-6
    achieved = bytes_pair / t_pair / 1e9 if t_pair > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(n),
                "kernel": "k_pair", "kernel_us": round(kt.get("pair", 0.0), 2),
                "algorithmic_bytes": int(bytes_pair), "pairs_a": int(pairs_a),
                "pairs_b": int(pairs_b)}

    if args.calib_copy_mb > 0:
        src = torch.empty(args.calib_copy_mb << 18, dtype=torch.float32, device=dev)
        src.fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        # gather: every 16-byte row read once, in a random order (index_select
        # reads the int64 index, 8 B per row, and writes the rows in order)
        rows = src.view(-1, 4)
        perm = torch.randperm(rows.shape[0], device=dev)
        out = torch.empty_like(rows)
        torch.cuda.synchronize()
        for _ in range(3):
            torch.index_select(rows, 0, perm, out=out)
        torch.cuda.synchronize()
        del src, dst, rows, perm, out

    # force error vs FP64 and the CPU baseline (oracle restatement, one core)
    cpu = None
    ferr = None
    if not args.no_cpu:
        cpu, ferr = cpu_leg(ctx, sys_, cfg, pos, args.cpu_steps)

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "LJ electrolyte RBL+RBE, 200k atoms, P=200", "atoms": n,
                   "charged": th1["charged"], "rho": 0.8, "lj_cut": 2.5, "coul_cut": 3.5,
                   "r_core": 2.5, "neighbor_sample_num": 30,
                   "rbl_frequency": th1["rbl_frequency"], "alpha": 0.09, "P": 200,
                   "skin": 0.3, "dt": dt, "rebuilds_in_timed_region": rebuilds,
                   "parallelism": "single GPU"},
        "kernel_us": {k: round(v, 2) for k, v in kt.items()},
        "ms_per_step_with_timing_events": round(1e3 * elapsed_ev / args.steps, 4),
        "host_enqueue_ms_per_step": round(1e3 * host_elapsed / args.steps, 4),
        "host_us_per_call": {"graph": round(host_graph_us, 1), "eager": round(host_eager_us, 1)},
        "force_rel_rms_vs_fp64": ferr,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo, as lscpu reports it)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(ctx, sys_, cfg, pos, cpu_steps):
    """Oracle CPU restatement timed on a bounded sample, and the GPU force error
    against the oracle's FP64 evaluation at the same positions and k-sample."""
    import torch

    import oracle
    from rbmd_amd import rbe_sample_host
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from refs import reference_force, rel_rms

    pos32 = pos.T.contiguous().cpu().numpy().astype(np.float32)
    pos32 = np.mod(pos32, sys_.box_len.astype(np.float32)).astype(np.float32)
    L = sys_.box_len
    n = sys_.n
    P = 200
    step = 10 ** 6
    m = rbe_sample_host(0.09, L, P, 777, 0)
    dev = pos.device
    p = torch.tensor(pos32.T.copy(), device=dev)
    f = torch.zeros_like(p)
    mt = torch.tensor(m.reshape(-1), dtype=torch.int32, device=dev)
    ctx.force(step, p[0], p[1], p[2], f[0], f[1], f[2], m_samples=mt, sync=True)
    got = f.T.double().cpu().numpy()
    ref = reference_force(sys_, pos32, cfg, step, m, prec=0)[0]
    ferr = rel_rms(got, ref)

    nb = cfg["hyper_parameters"]["neighbor"]
    freq = oracle.rbl_frequency(n, L, nb["r_core"], nb["cut_off"], nb["neighbor_sample_num"])
    params = oracle.pair_params(nb["r_core"], nb["cut_off"], nb["lj_cut_off"], 0.09, 1.0, freq,
                                step, 1, 1, 1)
    w = oracle.rbe_S(0.09, L) / P
    q32 = sys_.charges.astype(np.float32)
    # the GPU run rebuilds its lists about once in 8 steps: the CPU sample is
    # 8 steps with one rebuild, on one thread and on the host's cores (at
    # most 16, the CPU share of one GPU)
    cores = min(16, len(os.sched_getaffinity(0)))
    runs = {}
    for threads in (1, cores):
        t0 = time.perf_counter()
        oracle.cpu_force_steps(pos32, q32, sys_.types, sys_.ids, sys_.eps, sys_.sigma, L,
                               params, nb["skin"], m, w, cpu_steps, 8, threads)
        runs[threads] = time.perf_counter() - t0
    el = runs[cores]
    cpu = {"value": round(n * cpu_steps / el, 1), "unit": "atom-steps/s", "cores": cores,
           "kind": "port",
           "sample": f"{cpu_steps} force evaluations of the same 200k-atom configuration, "
                     "a list rebuild every 8 (RBL candidate list rc + skin, FP32 pair, RBE "
                     "k-space with FP32 phases and FP64 sums), OpenMP over atoms",
           "seconds": round(el, 3),
           "value_1_thread": round(n * cpu_steps / runs[1], 1),
           "seconds_1_thread": round(runs[1], 3),
           "cpu_model": cpu_model()}
    return cpu, ferr


if __name__ == "__main__":
    main()
