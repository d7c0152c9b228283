#!/usr/bin/env python3
"""PIC step throughput on one MI355X.

Workload (the scaled plain-PIC configuration): 128^3 nodes, dx = 1e-2 m,
Dirichlet phi = 0 walls, kinetic e- (Te = 2 eV) and O+ (16 amu, Ti = 0.1 eV),
20M macro-particles each, uniform positions, Maxwellian velocities, uniform
weight w = n0 L^3 / N with n0 = 1e12 m^-3, dt = 1e-10 s.  One step is
  push + sort e-, O+  (CIC gather + kick + drift + absorb, counting sort by
                      cell every step: pic_push_sort, a histogram pass and a
                      pass that pushes again into the sorted position)
  deposit e-, O+    (number density), charge density,
  Poisson solve to a 1e-6 relative residual (FP64), warm-started from the
  previous phi: the learned-preconditioned flexible CG when its trained
  weights are present (the same step with red-black SOR is timed beside
  it), E-field.
value = particle pushes per second over the K timed steps, inputs resident
in HBM.  Also reported: per-kernel HIP-event times, the roofline of the
dominant kernel, cold-start Poisson time to 1e-6 for each solver, and the
reference's own multithreaded CPU path timed on this host.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "engineering-degree-in-plasma-simulations_amd")
sys.path.insert(0, PKG)

QE = 1.602176565e-19
ME = 9.10938215e-31
AMU = 1.660538921e-27
EV_K = 11604.52
KB = 1.380648e-23
SIGMA_TABLE = os.path.join(PKG, "data", "sigma_el_O.dat")
HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0  # dense BF16 (MI355X)

# algorithmic bytes per unit of work (per particle, or per node and iteration)
ALG_BYTES = {
    "push": 24.0,      # histogram pass of the push: read x, y, z, vx, vy, vz (FP32)
    "scatter": 64.0,   # push into the sorted position: the 32-byte particle read and written
    "deposit": 16.0,   # x, y, z, w
    "pcg": 44.0,       # FP32 inner PCG: per node and iteration (88 B in FP64)
    "sor": 24.0,       # FP64 phi r/w + rho read per node and sweep
}
# the learned solver's flexible-CG vector updates per node and iteration: p, q
# (FP64) from z (FP32) and the previous p (28 B), then x, r updated from p, q (48 B)
ALG_BYTES_FCG = 76.0
KERNEL_NAMES = {
    "push": "k_push<false,2>", "scatter": "k_push_scatter", "deposit": "k_dep_cells + k_dep_nodes",
    "pcg": "k_pcg1<float> + k_pcg2<float>", "sor": "k_sor_split (one launch per colour)",
}



def measured_traffic(kernel):
    """HBM bytes per launch of a kernel from the committed PMC summary
    (profiles/pmc_traffic.json: 2 x FETCH_SIZE + 1 x WRITE_SIZE, calibrated
    on a device copy), or None where it was not measured."""
    tpath = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(tpath):
        return None
    return json.load(open(tpath)).get(kernel, {}).get("bytes_per_launch")

def unet_flops(channels, coarsen, n):
    """algorithmic FLOPs of one U-Net forward (2 per multiply-add; every
    convolution at its true channel counts, no padding) on the network grid
    of an n^3 node grid: input 2 -> C0 and B_0 on level 0, A_l and B_l on
    level l > 0, per decoder level a 1x1 up-convolution and Dec_l, output
    C0 -> 1"""
    vox = [(n // coarsen // 2 ** l) ** 3 for l in range(len(channels))]
    f = 2 * 27 * 2 * channels[0] * vox[0] + 2 * 27 * channels[0] * 1 * vox[0]  # in, out
    for l, c in enumerate(channels):
        if l:
            f += 2 * 27 * channels[l - 1] * c * vox[l]  # A_l
        f += 2 * 27 * c * c * vox[l]  # B_l
    for l in range(len(channels) - 1):
        f += 2 * channels[l + 1] * channels[l] * vox[l + 1]  # 1x1 up-convolution
        f += 2 * 27 * channels[l] * channels[l] * vox[l]  # Dec_l
    return f


def host_info():
    """nproc and the CPU model of this host (recorded with every CPU baseline)"""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def cpu_baseline(threads, particles, steps):
    """The reference's threaded electron push + serial deposit on this host."""
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    sample = "%d electrons on the 128^3 grid, %d steps, threaded push + deposit" % (particles,
                                                                                   steps)
    if os.path.exists(harness):
        t0 = time.time()
        out = subprocess.run([harness, "bench", "128", str(particles), str(steps), str(threads),
                              "0"], capture_output=True, text=True, timeout=600)
        line = [ln for ln in out.stdout.splitlines() if ln.startswith("{")][-1]
        r = json.loads(line)
        per_particle = r["push_s_per_particle"] + r["deposit_s_per_particle"]
        return {"value": 1.0 / per_particle, "unit": "particle pushes/s", "cores": threads,
                "kind": "reference", "sample": sample, "host": host_info(),
                "note": "the reference's ThreadPool push removes no particle (its deletion "
                        "buffers are std::bind copies, ThreadPool.h:63); the box keeps them inside",
                "push_only_pushes_per_s": r["pushes_per_s"], "wall_s": round(time.time() - t0, 1),
                # serial reference loops on the same electrons (the secondary rows)
                "load_particles_per_s": 1.0 / r["load_s_per_particle"],
                "moments_particles_per_s": 1.0 / r["moments_s_per_particle"],
                "diag_particles_per_s": 1.0 / r["diag_s_per_particle"]}
    # the oracle's v3-structured threaded push (reference binary not built)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    n = 128
    g = O.make_grid(n, n, n, [0, 0, 0], [1.27, 1.27, 1.27])
    rng = np.random.default_rng(3)
    p = np.zeros((particles, 7))
    p[:, :3] = rng.random((particles, 3)) * 1.27 * 0.98 + 0.0127
    p[:, 3:6] = rng.normal(0, 5e5, (particles, 3))
    p[:, 6] = 1.0
    p = p.reshape(-1)
    ef = np.zeros(3 * O.nn(g))
    t0 = time.time()
    for _ in range(steps):
        O.push_threaded(g, ef, p, -QE, ME, 1e-10, threads)
    dt = time.time() - t0
    return {"value": particles * steps / dt, "unit": "particle pushes/s", "cores": threads,
            "kind": "port", "sample": sample.replace("push + deposit", "push"), "host": host_info()}


def cpu_baseline_dsmc(nxy, nz, n_neu, n_ele, steps, dt):
    """The reference's MC_MEX_Ionization::apply (serial: its cell sort and the
    NTC pair loop) on a box of the discharge workload's cell densities."""
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.path.exists(harness):
        return None
    out = subprocess.run([harness, "bench_mcc", str(nxy), str(nz), str(n_neu), str(n_ele), str(steps),
                          repr(dt), SIGMA_TABLE], capture_output=True, text=True, timeout=900)
    r = json.loads([ln for ln in out.stdout.splitlines() if ln.startswith("{")][-1])
    return {"value": r["particles_per_s"], "unit": "DSMC particles/s", "cores": 1, "kind": "reference",
            "sample": "%d cells, %d O2 + %d e- at the workload's cell densities, %d apply calls (median)"
                      % (r["cells"], n_neu, n_ele, steps), "apply_s": r["apply_s"]}


def dsmc_bench(pg, args, cpu):
    """The DSMC step of the discharge workload on one GPU: 256 x 64 x 64 nodes,
    O2 at 10 Pa / 300 K as variable-weight neutrals, electrons at 10 eV and
    ions, 50M particles in all; a step is pic_mcc_apply (the neutrals' and
    electrons' births of the last step indexed by cell, or a join or whole
    sort, then one thread per cell running the NTC groups, collisions,
    ionization and births), as the reference's apply sorts and pairs.  The
    per-stage profile repeats the same steps on a second, identical set-up
    (same seeds and step keys), so its stages cover the timed steps' work.
    Returns the section of the bench line."""
    nxy, nz, dx = 64, 256, 1e-4
    L = np.array([dx * (nxy - 1), dx * (nxy - 1), dx * (nz - 1)])
    box = 0.98 * L
    vol = float(np.prod(box))
    n_gas = 10.0 / (KB * 300.0)
    Nn, Ne, Ni = args.dsmc_neutrals, args.dsmc_electrons, args.dsmc_ions
    w_n = n_gas * vol / Nn
    w_e = w_n / 1000.0
    dt = args.dsmc_dt
    t = np.loadtxt(SIGMA_TABLE)

    def setup():
        ctx = pg.Context(nxy, nxy, nz, [0, 0, 0], L)
        # elastic splits append ~5% of the neutrals per step: room for the
        # warm-up and timed steps
        sn = ctx.species_add(32 * AMU, 0.0, w_n, int(Nn * (1.0 + 0.06 * (args.dsmc_warmup + args.dsmc_steps + 2))))
        si = ctx.species_add(32 * AMU, QE, w_e, Ni + Ne // 2 + 1024)
        se = ctx.species_add(ME, -QE, w_e, Ne + Ne // 2 + 1024)
        c = 0.5 * L
        ctx.load_box_thermal(sn, c, box, n_gas, 300.0, seed=21, mode=0)
        ctx.load_box_thermal(se, c, box, Ne * w_e / vol, 10.0 * EV_K, seed=22, mode=0)
        ctx.load_box_thermal(si, c, box, Ni * w_e / vol, 300.0, seed=23, mode=0)
        mid = ctx.mcc_create(sn, si, se, 12.07 * QE, t[:, 0], t[:, 1])
        for step in range(args.dsmc_warmup):  # also brings W sigma v_max to its working value
            ctx.mcc_apply(mid, dt, 7, step)
        return ctx, sn, se, mid

    ctx, sn, se, mid = setup()
    step = args.dsmc_warmup
    K = args.dsmc_steps
    n0 = ctx.count(sn) + ctx.count(se)
    ctx.sync()
    t0 = time.perf_counter()
    tot = {"collisions": 0, "ionizations": 0, "groups": 0}
    for _ in range(K):
        st = ctx.mcc_apply(mid, dt, 7, step)
        step += 1
        for k in tot:
            tot[k] += st[k]
    ctx.sync()
    T = time.perf_counter() - t0
    n1 = ctx.count(sn) + ctx.count(se)
    ctx.close()
    # the same steps again on an identical set-up, with HIP events around
    # every launch, averaged
    ctx, sn, se, mid = setup()
    step = args.dsmc_warmup
    ctx.profile(True)
    ctx.profile_reset()
    st = {"collisions": 0, "ionizations": 0, "groups": 0}
    for _ in range(K):
        r = ctx.mcc_apply(mid, dt, 7, step)
        step += 1
        for k in st:
            st[k] += r[k] / K
    ctx.sync()
    prof = {k: (v[0] / K, v[1]) for k, v in ctx.profile_read().items()}
    ctx.profile(False)
    ncells = (nxy - 1) * (nxy - 1) * (nz - 1)
    mcc_ms = prof["mcc"][0]
    # k_mcc bytes: both species' cell offsets (8 B per cell), the neutral and
    # electron of every NTC group (w and v: 32 B), per accepted collision the
    # electron velocity and neutral weight written back (16 B) and one 32-B
    # birth (split neutral or ion), per ionization one more (the electron)
    alg = 8.0 * ncells + 32.0 * st["groups"] + 48.0 * st["collisions"] + 32.0 * st["ionizations"]
    ach = alg / (mcc_ms * 1e-3) / 1e9
    roof_mcc = {"bound": "hbm", "kernel": "k_mcc", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS, "traffic": measured_traffic("k_mcc"), "alg_bytes_per_launch": alg,
                "avg_launch_ms": mcc_ms}
    # the sorts (the appended births' index, a join or a whole sort): their
    # time only -- how many particles a sort moved is not visible here
    sc_ms, sc_n = prof["scatter"]
    so_ms = prof["sort"][0]
    out = {
        "metric": "DSMC particles/s (neutrals + electrons through pic_mcc_apply: sort + pair selection "
                  "+ collisions)",
        "value": (n0 + n1) / 2.0 * K / T, "unit": "DSMC particles/s", "steps": K, "warmup": args.dsmc_warmup,
        "ms_per_step": T / K * 1e3,
        "config": {"workload": "DSMC ionization, discharge cell densities: %dx%dx%d nodes, O2 10 Pa / "
                               "300 K, e- 10 eV, %d + %d + %d particles, dt %.1e s, 64-point sigma_el "
                               "table" % (nz, nxy, nxy, Nn, Ne, Ni, dt),
                   "cells": ncells},
        "particles_start": n0, "particles_end": n1,
        "per_step": {k: v / K for k, v in tot.items()},
        "stages_ms_per_step": {k: round(v[0], 4) for k, v in prof.items() if v[1]},
        "stages_profiled_steps": K,
        "roofline": roof_mcc,
        "sorts_ms_per_step": sc_ms + so_ms,
        "dominant": "k_mcc" if mcc_ms >= sc_ms + so_ms else "sorts",
        "cpu_baseline": cpu,
    }
    ctx.close()
    return out


def discharge_bench(pg, args):
    """The DSMC ionization discharge as a loop (the reference's main.cpp:
    177-288 over the build's classes' calls): 64 x 64 x 256 nodes, dx =
    1e-4 m, Rectangle plates of 0.1 L at both z ends at +300 V and 0 V, O2 at
    10 Pa / 300 K between them as variable-weight neutrals, electrons at
    10 eV and O2+ ions (50M particles in all).  The field is solved once
    (v3 never re-solves in its loop, main.cpp:172-173, :260-261).  A step:
    MC_MEX_Ionization::apply, the electrons pushed every step, the ions every
    10th with 10 dt (recycled into neutrals at the plates), the neutrals
    every 100th with 50 dt (diffuse reflection at the plates), the pushed
    species' densities and the charge density; every 50th step a merge of
    every species keeps the particle count bounded (the reference merges
    every 50th step past the 250th, main.cpp:179)."""
    nxy, nz, dx = 64, 256, 1e-4
    L = np.array([dx * (nxy - 1), dx * (nxy - 1), dx * (nz - 1)])
    dt = args.discharge_dt
    ctx = pg.Context(nxy, nxy, nz, [0, 0, 0], L)
    ctx.set_dt(dt)
    ctx.add_rectangle([L[0] / 2, L[1] / 2, 0.0], [L[0], L[1], 0.1 * L[2]], 300.0)
    ctx.add_rectangle([L[0] / 2, L[1] / 2, L[2]], [L[0], L[1], 0.1 * L[2]], 0.0)
    gas = np.array([L[0], L[1], 0.9 * L[2]]) * 0.999
    c = 0.5 * L
    vol = float(np.prod(gas))
    n_gas = 10.0 / (KB * 300.0)
    Nn, Ne, Ni = args.dsmc_neutrals, args.dsmc_electrons, args.dsmc_ions
    w_n = n_gas * vol / Nn
    w_e = w_n / 1000.0
    sn = ctx.species_add(32 * AMU, 0.0, w_n, int(1.6 * Nn))
    si = ctx.species_add(32 * AMU, QE, w_e, 2 * Ni)
    se = ctx.species_add(ME, -QE, w_e, 2 * Ne)
    ctx.load_box_thermal(sn, c, gas, n_gas, 300.0, seed=31, mode=0)
    ctx.load_box_thermal(se, c, gas, Ne * w_e / vol, 10.0 * EV_K, seed=32, mode=0)
    ctx.load_box_thermal(si, c, gas, Ni * w_e / vol, 300.0, seed=33, mode=0)
    ctx.set_surface(sn, pg.SURF_REFLECT, seed=34)
    ctx.set_surface(si, pg.SURF_RECYCLE, sn, seed=35)
    t = np.loadtxt(SIGMA_TABLE)
    # W sigma v_rel max from the start at its bound over the table (the
    # reference's 1e-14 max(mpw0), Interactions.cpp:534, is 1/100 of it
    # here: no cell would reach half an NTC group and the maximum would never
    # be raised)
    m_red = 32 * AMU * ME / (32 * AMU + ME)
    wsv0 = w_n * float(np.max(t[:, 1] * np.sqrt(2.0 * t[:, 0] * QE / m_red)))
    mid = ctx.mcc_create(sn, si, se, 12.07 * QE, t[:, 0], t[:, 1], wsv_max_init=wsv0)
    for s_ in (sn, si, se):
        ctx.deposit(s_)
    ctx.charge_density()
    st = ctx.poisson(method=pg.SOLVER_SOR_RB, tol=1e-6, tol_kind=pg.TOL_REL, max_it=50000,
                     omega=2.0 / (1.0 + np.sin(np.pi / (nz - 1))), check_every=25)
    ctx.efield()
    counts0 = {"neutrals": ctx.count(sn), "ions": ctx.count(si), "electrons": ctx.count(se)}
    tot = {"collisions": 0, "ionizations": 0, "groups": 0, "merged_bins": 0, "merge_removed": 0}

    def step(ts):
        m = ctx.mcc_apply(mid, dt, 9, ts)
        for k in ("collisions", "ionizations", "groups"):
            tot[k] += m[k]
        ctx.push_sort(se, dt)
        ctx.deposit(se)
        if ts % 10 == 0:
            ctx.push_sort(si, 10 * dt)
            ctx.deposit(si)
        if ts % 100 == 0:
            ctx.push_sort(sn, 50 * dt)  # main.cpp:229-234
            ctx.deposit(sn)
        ctx.charge_density()
        if ts % 50 == 0 and ts > 0:
            for s_ in (sn, si, se):
                r = ctx.merge(s_, seed=40 + ts)
                tot["merged_bins"] += r["bins"]
                tot["merge_removed"] += r["removed"]

    K = args.discharge_steps
    ctx.sync()
    t0 = time.perf_counter()
    for ts in range(1, K + 1):
        step(ts)
    ctx.sync()
    T = time.perf_counter() - t0
    counts1 = {"neutrals": ctx.count(sn), "ions": ctx.count(si), "electrons": ctx.count(se)}
    # per-stage device times over the next 100 steps (HIP events around
    # every launch; a separate pass so they stay out of the timed loop)
    PK = 100
    ctx.profile(True)
    ctx.profile_reset()
    for ts in range(K + 1, K + PK + 1):
        step(ts)
    ctx.sync()
    prof = ctx.profile_read()
    ctx.profile(False)
    ctx.close()
    return {"workload": "DSMC ionization discharge: %dx%dx%d nodes, +300 V / 0 V plates, O2 10 Pa / "
                        "300 K, %d + %d + %d particles, dt %.1e s, subcycled ions (10 dt) and neutrals "
                        "(50 dt), solve once, merge every 50 steps" % (nz, nxy, nxy, Nn, Ne, Ni, dt),
            "steps": K, "ms_per_step": T / K * 1e3, "initial_solve": {"iterations": st["iterations"],
                                                                      "rel_residual": st["rel_residual"]},
            "particles_start": counts0, "particles_end": counts1,
            "per_step": {k: v / K for k, v in tot.items()},
            "stages_ms_per_step": {k: round(v[0] / PK, 4) for k, v in prof.items() if v[1]},
            "stages_profiled_steps": PK}


def c5_bench(pg, args):
    """The full loop at scale with Boltzmann electrons: 256^3 nodes, dx =
    1e-2 m, Dirichlet walls, 200M kinetic O+ (0.1 eV), Boltzmann electrons
    n0 = 1e12 m^-3, Te0 = 2 eV in a Newton-Boltzmann field solve to a 1e-6
    relative FP64 residual every step (learned inner solver with the shipped
    network, which is fully convolutional and runs on the 128^3 network grid
    of this 256^3 node grid; FP32-inner and FP64 Newton-PCG beside it).  A
    step: push + sort of the ions, their density, the charge density, the
    Newton solve warm-started from the extrapolated potential, the E-field.
    The final FP64 residual of the nonlinear problem is recomputed here."""
    n = args.c5_nodes
    L = 1e-2 * (n - 1)
    N = args.c5_particles
    n0, Te0, dt = 1e12, 2.0, 1e-10
    ctx = pg.Context(n, n, n, [0, 0, 0], [L, L, L])
    ctx.dirichlet_walls()
    ctx.set_dt(dt)
    w = 1.5 * n0 * L ** 3 / N  # ions at 1.5 n0: a positive space charge against n0 electrons
    si = ctx.species_add(16 * AMU, QE, w, int(N * 1.02) + 1024)
    t0 = time.perf_counter()
    ctx.load_box_thermal(si, [L / 2] * 3, [L] * 3, 1.5 * n0, 0.1 * EV_K, seed=51, mode=1)
    ctx.sort(si)
    t_load = time.perf_counter() - t0
    weights = os.path.join(PKG, "data", "unet_128.f32")
    have_net = os.path.exists(weights) and os.path.exists(weights + ".json")
    if have_net:
        ctx.unet_load(weights)
    methods = [("learned", pg.SOLVER_LEARNED)] if have_net else []
    methods += [("pcg_mixed", pg.SOLVER_PCG_MIXED), ("pcg_fp64", pg.SOLVER_PCG)]
    out = {"workload": "%d^3 nodes, %d O+ kinetic, Boltzmann electrons n0 = 1e12, Te0 = 2 eV, Newton "
                       "to 1e-6 relative FP64 residual every step" % (n, N),
           "load_sort_s": t_load, "solvers": {}}
    # every solver starts from the same converged potential of the loaded
    # state (a cold FP64 solve, untimed), with the extrapolation history reset
    ctx.deposit(si)
    ctx.charge_density()
    ctx.poisson(method=pg.SOLVER_PCG, tol=1e-6, tol_kind=pg.TOL_REL, max_it=20000, n0=n0, Te0=Te0)
    phi_start = ctx.get_field(pg.FIELD_PHI)
    for name, method in methods:
        kw = dict(method=method, tol=1e-6, tol_kind=pg.TOL_REL, max_it=20000, n0=n0, Te0=Te0,
                  flags=pg.SOLVE_EXTRAPOLATE)
        its = []

        def step():
            ctx.push_sort(si, dt)
            ctx.deposit(si)
            ctx.charge_density()
            st = ctx.poisson(**kw)
            if not st["converged"]:
                raise RuntimeError("C5 Newton solve did not converge: %r" % st)
            its.append((st["newton_iterations"], st["iterations"], st["seconds"]))
            ctx.efield()

        ctx.set_field(pg.FIELD_PHI, phi_start)
        step()  # warm-up
        its.clear()
        ctx.profile(True)
        ctx.profile_reset()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.c5_steps):
            step()
        ctx.sync()
        T = time.perf_counter() - t0
        prof = ctx.profile_read()
        ctx.profile(False)
        phi = ctx.get_field(pg.FIELD_PHI)
        rho = ctx.get_field(pg.FIELD_RHO)
        out["solvers"][name] = {
            "ms_per_step": T / args.c5_steps * 1e3, "pushes_per_s": N * args.c5_steps / T,
            "newton_iterations_per_step": float(np.mean([a for a, _, _ in its])),
            "inner_iterations_per_step": float(np.mean([b for _, b, _ in its])),
            "solve_ms_per_step": float(np.mean([c for _, _, c in its])) * 1e3,
            "final_fp64_residual": nonlinear_residual(n, L, phi, rho, 0.0, n0, Te0),
            "stages_ms_per_step": {k: round(v[0] / args.c5_steps, 4) for k, v in prof.items() if v[1]},
            "note": "profiled with HIP events around every launch"}
        print("c5 %s: %.1f ms per step" % (name, out["solvers"][name]["ms_per_step"]), file=sys.stderr)
    ctx.close()
    return out


def nonlinear_residual(n, L, phi, rho, phi0, n0, Te0):
    """||F|| / ||b|| over the interior nodes, F = A phi + rho/eps0 - q n0
    exp((phi - phi0)/Te0)/eps0 (FP64, here on the host)"""
    eps0 = 8.85418782e-12
    f = phi.reshape(n, n, n)
    r = rho.reshape(n, n, n)
    ih2 = 1.0 / (L / (n - 1)) ** 2
    c = f[1:-1, 1:-1, 1:-1]
    A = (f[2:, 1:-1, 1:-1] + f[:-2, 1:-1, 1:-1] + f[1:-1, 2:, 1:-1] + f[1:-1, :-2, 1:-1] +
         f[1:-1, 1:-1, 2:] + f[1:-1, 1:-1, :-2] - 6 * c) * ih2
    b = -r[1:-1, 1:-1, 1:-1] / eps0
    F = A - b - QE * n0 * np.exp((c - phi0) / Te0) / eps0
    return float(np.sqrt(np.sum(F ** 2) / np.sum(b ** 2)))


def cpu_baseline_newton(n, threads):
    """The reference's Newton-Boltzmann PCG solve (solveNRPCG) on an n^3 box."""
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.path.exists(harness):
        return None
    out = subprocess.run([harness, "bench_newton", str(n), str(threads), "1e-6"], capture_output=True,
                         text=True, timeout=600)
    r = json.loads([ln for ln in out.stdout.splitlines() if ln.startswith("{")][-1])
    return {"value": r["solve_s"], "unit": "s per solve", "cores": threads, "kind": "reference",
            "sample": "%d^3 nodes, Dirichlet walls, uniform ion charge 1.5 n0 q, Boltzmann electrons "
                      "n0 = 1e12, Te0 = 2 eV, reference criterion 1e-6" % n,
            "converged": bool(r["converged"])}


def timed(ctx, fn, reps=3):
    """median wall time of fn() between stream synchronisations (s)"""
    ts = []
    for _ in range(reps):
        ctx.sync()
        t0 = time.perf_counter()
        fn()
        ctx.sync()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def newton_solve(pg, ctx, n0=1e12, Te0=2.0):
    """Newton-Boltzmann solve of the uniform-ion problem from phi = 0 (as the
    reference harness's): returns (seconds, stats)"""
    ctx.set_field(pg.FIELD_RHO, np.full(ctx.nn, QE * 1.5 * n0))
    ctx.set_field(pg.FIELD_PHI, np.zeros(ctx.nn))
    ctx.sync()
    t0 = time.perf_counter()
    st = ctx.poisson(method=pg.SOLVER_PCG, tol=1e-6, tol_kind=pg.TOL_REF_RMS, max_it=20000, n0=n0, Te0=Te0)
    ctx.sync()
    return time.perf_counter() - t0, st


def extras_bench(pg, ctx, args, se, N, n, cpu, cpu_newton, x0, xm):
    """The secondary rows on the main workload's context: density moments of
    one species (N particles), the diagnostic sums, the loader, and the
    Newton-Boltzmann solve (here at n^3 and at the CPU sample's 64^3)."""
    out = {}
    nn = ctx.nn

    def mom():
        ctx.clear_samples(se)
        ctx.deposit(se, pg.DEPOSIT_MOMENTS)

    t = timed(ctx, mom)
    # particle: cell word, offsets, velocity, weight (32 B); nodes: 7 FP64 channels written
    alg = 32.0 * N + 56.0 * nn
    out["moments"] = {"row": "a21 sampleMoments", "particles": N, "seconds": t, "particles_per_s": N / t,
                      "roofline": {"bound": "hbm", "achieved": alg / t / 1e9, "peak": HBM_PEAK_GBS,
                                   "unit": "GB/s", "frac": alg / t / 1e9 / HBM_PEAK_GBS,
                                   "alg_bytes_per_call": alg},
                      "cpu_baseline": ({"value": cpu["moments_particles_per_s"], "unit": "particles/s",
                                        "cores": 1, "kind": "reference", "sample": cpu["sample"]}
                                       if cpu and "moments_particles_per_s" in cpu else None)}

    def diag():
        ctx.diag_sums(se)
        ctx.field_energy()

    t = timed(ctx, diag)
    alg = 32.0 * N + 16.0 * nn  # every particle read; E (float4) per node
    out["diagnostics"] = {"row": "a22 micro count, momentum, KE, PE, field energy", "particles": N,
                          "seconds": t, "particles_per_s": N / t,
                          "roofline": {"bound": "hbm", "achieved": alg / t / 1e9, "peak": HBM_PEAK_GBS,
                                       "unit": "GB/s", "frac": alg / t / 1e9 / HBM_PEAK_GBS,
                                       "alg_bytes_per_call": alg},
                          "note": "includes the host reads of the sums",
                          "cpu_baseline": ({"value": cpu["diag_particles_per_s"], "unit": "particles/s",
                                            "cores": 1, "kind": "reference", "sample": cpu["sample"]}
                                           if cpu and "diag_particles_per_s" in cpu else None)}
    # loader: a fresh species of N electrons, the reference's sampler (mode 0)
    sl = ctx.species_add(ME, -QE, 1.0, N + 1024)
    L = np.asarray(xm, float) - np.asarray(x0, float)
    vol = float(np.prod(0.98 * L))
    ctx.sync()
    t0 = time.perf_counter()
    got = ctx.load_box_thermal(sl, np.asarray(x0, float) + 0.5 * L, 0.98 * L, N * 1.0 / vol, 2.0 * EV_K,
                               seed=5, mode=0)
    ctx.sync()
    t = time.perf_counter() - t0
    alg = 32.0 * got
    out["loader"] = {"row": "a23 loadParticleBoxThermal", "particles": got, "seconds": t,
                     "particles_per_s": got / t,
                     "roofline": {"bound": "hbm", "achieved": alg / t / 1e9, "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "frac": alg / t / 1e9 / HBM_PEAK_GBS, "alg_bytes_per_call": alg},
                     "note": "wall time of the call, including its host-side checks",
                     "cpu_baseline": ({"value": cpu["load_particles_per_s"], "unit": "particles/s", "cores": 1,
                                       "kind": "reference", "sample": cpu["sample"]}
                                      if cpu and "load_particles_per_s" in cpu else None)}
    # Newton-Boltzmann (FP64 outer Newton, FP64 Jacobi-PCG inner): at n^3 here
    phi_keep, rho_keep = ctx.get_field(pg.FIELD_PHI), ctx.get_field(pg.FIELD_RHO)
    t, st = newton_solve(pg, ctx)
    its = st["iterations"]
    alg = 88.0 * nn * its  # the two fused PCG kernels per inner iteration (FP64)
    res = {"row": "a11 solveNRPCG", "grid": "%d^3" % n, "seconds": t, "newton_iterations": st["newton_iterations"],
           "pcg_iterations": its, "converged": st["converged"],
           "roofline": {"bound": "hbm", "kernel": "k_pcg1 + k_pcg2", "achieved": alg / t / 1e9,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": alg / t / 1e9 / HBM_PEAK_GBS,
                        "note": "whole solve time, cache-resident vectors (16.8 MB each)"}}
    ctx.set_field(pg.FIELD_RHO, rho_keep)
    ctx.set_field(pg.FIELD_PHI, phi_keep)
    # the same problem at the CPU sample's size
    c64 = pg.Context(64, 64, 64, [0, 0, 0], [0.63, 0.63, 0.63])
    c64.dirichlet_walls()
    newton_solve(pg, c64)  # first call allocates the solver's vectors
    t64, st64 = newton_solve(pg, c64)
    c64.close()
    res["at_64cubed"] = {"seconds": t64, "newton_iterations": st64["newton_iterations"],
                         "pcg_iterations": st64["iterations"], "converged": st64["converged"]}
    res["cpu_baseline"] = cpu_newton
    if cpu_newton:
        res["speedup_vs_cpu_at_64cubed"] = cpu_newton["value"] / t64
    out["newton_boltzmann"] = res
    return out


def poisson_256(pg, weights, n=256):
    """Cold solves to a 1e-6 relative residual at 256^3 nodes on a blob
    density: the shipped 128^3-node network, fully convolutional, runs on the
    128^3 network grid of this grid; red-black SOR needs about twice the
    sweeps of 128^3, the two-level cycle about the same number of
    iterations.  Each solve runs twice, the second is timed."""
    L = 1e-2 * (n - 1)
    c = pg.Context(n, n, n, [0, 0, 0], [L] * 3)
    c.dirichlet_walls()
    c.unet_load(weights)
    c.set_field(pg.FIELD_RHO, blob_density(n, seed=4242))
    zeros = np.zeros(c.nn)
    res = {"grid": "%d^3" % n}
    for name, kw in (("pcg_fp64", dict(method=pg.SOLVER_PCG)), ("pcg_mixed", dict(method=pg.SOLVER_PCG_MIXED)),
                     ("sor_rb", dict(method=pg.SOLVER_SOR_RB, check_every=25,
                                     omega=2.0 / (1.0 + np.sin(np.pi / (n - 1))))),
                     ("learned", dict(method=pg.SOLVER_LEARNED))):
        for _ in range(2):
            c.set_field(pg.FIELD_PHI, zeros)
            st = c.poisson(tol=1e-6, tol_kind=pg.TOL_REL, max_it=50000, **kw)
        res[name] = {"seconds": round(st["seconds"], 5), "iterations": st["iterations"],
                     "rel_residual": st["rel_residual"], "converged": st["converged"]}
        print("poisson 256^3 blobs %s: %.4f s, %d it, rel %.2e" % (name, st["seconds"], st["iterations"],
                                                                  st["rel_residual"]), file=sys.stderr)
    if res["learned"]["converged"]:
        t = res["learned"]["seconds"]
        res["learned_speedup_vs_sor_rb"] = round(res["sor_rb"]["seconds"] / t, 3)
        res["learned_speedup_vs_pcg_fp64"] = round(res["pcg_fp64"]["seconds"] / t, 3)
    c.close()
    return res


def blob_density(n, seed=12345):
    """six Gaussian blobs (widths 3 to 0.15 n nodes) plus white noise at a
    tenth of their amplitude, zero on the walls; C/m^3, node order i-fastest"""
    rng = np.random.default_rng(seed)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    rho = np.zeros((n, n, n))
    for _ in range(6):
        c = rng.uniform(4, n - 5, 3)
        s = rng.uniform(3, 0.15 * n)
        rho += rng.choice([-1, 1]) * np.exp(-((i - c[0]) ** 2 + (j - c[1]) ** 2 + (k - c[2]) ** 2)
                                            / (2 * s * s))
    rho += 0.1 * rng.normal(size=rho.shape)
    rho[0, :, :] = rho[-1, :, :] = rho[:, 0, :] = rho[:, -1, :] = rho[:, :, 0] = rho[:, :, -1] = 0
    return 1e-8 * rho.reshape(-1)


def copy_calibration():
    """A plain copy of known size, for scaling FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    n = 256 * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--particles", type=int, default=20_000_000, help="per species")
    ap.add_argument("--nodes", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-poisson-compare", action="store_true")
    ap.add_argument("--cpu-particles", type=int, default=20_000_000)  # one species of the workload
    ap.add_argument("--cpu-steps", type=int, default=20)  # about 10 s of host work
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--step-solver", default="auto",  # learned when its weights exist, else sor_rb
                    choices=["auto", "sor_rb", "pcg_mixed", "pcg_fp64", "learned"])
    # each step's solve starts from 2 phi_n - phi_(n-1) ("extrapolate") or from phi_n ("previous")
    ap.add_argument("--warm-start", default="extrapolate", choices=["extrapolate", "previous"])
    # the DSMC step of the discharge workload, reported beside the PIC line
    ap.add_argument("--no-dsmc", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the secondary rows (moments, "
                                                             "diagnostics, loader, Newton-Boltzmann)")
    ap.add_argument("--only-dsmc", action="store_true", help="print the DSMC section alone")
    ap.add_argument("--only-discharge", action="store_true", help="print the discharge-loop section alone")
    ap.add_argument("--only-c5", action="store_true", help="print the C5 section alone")
    ap.add_argument("--dsmc-neutrals", type=int, default=40_000_000)
    ap.add_argument("--dsmc-electrons", type=int, default=5_000_000)
    ap.add_argument("--dsmc-ions", type=int, default=5_000_000)
    ap.add_argument("--dsmc-steps", type=int, default=10)
    ap.add_argument("--dsmc-warmup", type=int, default=3)
    ap.add_argument("--dsmc-dt", type=float, default=2e-9)  # a few NTC groups per cell
    # the discharge as a loop (plates, subcycling, surfaces, merge)
    ap.add_argument("--no-discharge", action="store_true")
    ap.add_argument("--discharge-steps", type=int, default=200)
    ap.add_argument("--discharge-dt", type=float, default=5e-10)  # about one NTC group per cell
    # the Boltzmann-electron loop at scale
    ap.add_argument("--no-c5", action="store_true")
    ap.add_argument("--c5-nodes", type=int, default=256)
    ap.add_argument("--c5-particles", type=int, default=200_000_000)
    ap.add_argument("--c5-steps", type=int, default=3)
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU; this build has no multi-GPU path (--gpus %d refused)"
              % args.gpus, file=sys.stderr)
        return 2

    cpu = None
    cpu_dsmc = None
    cpu_newton = None
    if not args.no_cpu_baseline:  # before the GPU is initialised
        # the reference's default is all cores but one (Config.cpp:68-75), of
        # the 16 host cores a one-GPU job may use here
        threads = max(1, min(16, os.cpu_count() or 1) - 1)
        if not (args.only_dsmc or args.only_discharge or args.only_c5):
            cpu = cpu_baseline(threads, args.cpu_particles, args.cpu_steps)
            print("cpu baseline: %.3e pushes/s (%s, %d threads)" % (cpu["value"], cpu["kind"],
                                                                     threads), file=sys.stderr)
        if not args.no_dsmc and not (args.only_discharge or args.only_c5):
            # 32 x 32 x 128 cells (an eighth of the workload's) at its densities
            f = 32 * 32 * 128 / (63 * 63 * 255)
            cpu_dsmc = cpu_baseline_dsmc(33, 129, int(args.dsmc_neutrals * f), int(args.dsmc_electrons * f),
                                         args.dsmc_warmup + 7, args.dsmc_dt)
            if cpu_dsmc:
                print("cpu baseline dsmc: %.3e particles/s (reference apply, 1 thread)" % cpu_dsmc["value"],
                      file=sys.stderr)
        if not (args.only_dsmc or args.only_discharge or args.only_c5) and not args.no_extras:
            cpu_newton = cpu_baseline_newton(64, threads)

    import pic_gpu as pg

    if args.only_dsmc:
        d = dsmc_bench(pg, args, cpu_dsmc)
        d["n_gpus"] = 1
        d["higher_is_better"] = True
        print(json.dumps(d))
        return 0
    if args.only_discharge or args.only_c5:
        d = discharge_bench(pg, args) if args.only_discharge else c5_bench(pg, args)
        d["n_gpus"] = 1
        print(json.dumps(d))
        return 0

    n = args.nodes
    dx = 1e-2
    L = dx * (n - 1)
    x0 = np.zeros(3)
    xm = np.full(3, L)
    N = args.particles
    n0 = 1e12
    w = n0 * L ** 3 / N
    dt = 1e-10
    ctx = pg.Context(n, n, n, x0, xm)
    ctx.dirichlet_walls()
    ctx.set_dt(dt)
    cap = int(N * 1.02) + 1024
    se = ctx.species_add(ME, -QE, w, cap)
    si = ctx.species_add(16 * AMU, QE, w, cap)
    centre = 0.5 * (x0 + xm)
    ctx.load_box_thermal(se, centre, xm - x0, n0, 2.0 * EV_K, seed=11, mode=1)
    ctx.load_box_thermal(si, centre, xm - x0, n0, 0.1 * EV_K, seed=12, mode=1)
    species = (se, si)
    for s in species:
        ctx.sort(s)
        ctx.deposit(s)
    ctx.charge_density()
    weights = os.path.join(PKG, "data", "unet_%d.f32" % n)
    have_net = os.path.exists(weights) and os.path.exists(weights + ".json")
    net_meta = None
    if have_net:
        net_meta = ctx.unet_load(weights)
    elif args.step_solver == "learned":
        raise SystemExit("--step-solver learned needs %s(.json)" % weights)
    step_solver = args.step_solver
    if step_solver == "auto":
        step_solver = "learned" if have_net else "sor_rb"
    omega_opt = 2.0 / (1.0 + np.sin(np.pi / (n - 1)))

    def solver_kw(name):
        kw = {"sor_rb": dict(method=pg.SOLVER_SOR_RB, omega=omega_opt, check_every=10),
              "pcg_mixed": dict(method=pg.SOLVER_PCG_MIXED),
              "pcg_fp64": dict(method=pg.SOLVER_PCG),
              "learned": dict(method=pg.SOLVER_LEARNED, check_every=1)}[name]
        kw.update(tol=1e-6, tol_kind=pg.TOL_REL, max_it=20000,
                  flags=pg.SOLVE_EXTRAPOLATE if args.warm_start == "extrapolate" else 0)
        return kw

    solve_kw = solver_kw(step_solver)
    ctx.poisson(**solve_kw)
    ctx.efield()
    print("loaded %d + %d particles" % (ctx.count(se), ctx.count(si)), file=sys.stderr)

    poisson = None
    if not args.no_poisson_compare:
        phi_keep = ctx.get_field(pg.FIELD_PHI)
        rho_keep = ctx.get_field(pg.FIELD_RHO)
        zeros = np.zeros(ctx.nn)
        # cold start on the step's own charge density and on a smooth
        # blob-plus-noise density; each solve runs twice, the second is timed
        cases = {"pic_density": rho_keep, "blobs": blob_density(n)}
        poisson = {"grid": "%d^3" % n, "start": "phi = 0", "tol_rel": 1e-6,
                   "weights": os.path.relpath(weights, ROOT) if have_net else None}
        runs = [("pcg_fp64", dict(method=pg.SOLVER_PCG)),
                ("pcg_mixed", dict(method=pg.SOLVER_PCG_MIXED)),
                ("sor_rb", dict(method=pg.SOLVER_SOR_RB, check_every=25,
                                omega=2.0 / (1.0 + np.sin(np.pi / (n - 1)))))]
        if have_net:
            runs.append(("learned", dict(method=pg.SOLVER_LEARNED)))
        for cname, rho in cases.items():
            ctx.set_field(pg.FIELD_RHO, rho)
            res = {}
            for name, kw in runs:
                for _ in range(2):
                    ctx.set_field(pg.FIELD_PHI, zeros)
                    st = ctx.poisson(tol=1e-6, tol_kind=pg.TOL_REL, max_it=20000, **kw)
                res[name] = {"seconds": round(st["seconds"], 5), "iterations": st["iterations"],
                             "rel_residual": st["rel_residual"], "converged": st["converged"]}
                print("poisson %s %s: %.4f s, %d it, rel %.2e" % (
                    cname, name, st["seconds"], st["iterations"], st["rel_residual"]), file=sys.stderr)
            if have_net and res["learned"]["converged"]:
                t = res["learned"]["seconds"]
                res["learned_speedup_vs_sor_rb"] = round(res["sor_rb"]["seconds"] / t, 3)
                res["learned_speedup_vs_pcg_fp64"] = round(res["pcg_fp64"]["seconds"] / t, 3)
            poisson[cname] = res
        ctx.set_field(pg.FIELD_RHO, rho_keep)
        ctx.set_field(pg.FIELD_PHI, phi_keep)
        if have_net:
            poisson["blobs_256"] = poisson_256(pg, weights)

    solve_iters = []

    def step(kw=solve_kw):
        for s in species:
            ctx.push_sort(s, dt)
        for s in species:
            ctx.deposit(s)
        ctx.charge_density()
        st = ctx.poisson(**kw)
        if not st["converged"]:
            raise RuntimeError("Poisson solve did not converge: %r" % st)
        solve_iters.append(st["iterations"])
        ctx.efield()

    for _ in range(args.warmup):
        step()
    solve_iters.clear()
    ctx.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step()
    ctx.sync()
    t1 = time.perf_counter()
    T = t1 - t0
    iters_main = float(np.mean(solve_iters)) if solve_iters else None
    # per-stage device times (HIP events around every launch) from a separate
    # pass, so the event records stay out of the timed steps
    PROF_STEPS = 3
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(PROF_STEPS):
        step()
    ctx.sync()
    prof = ctx.profile_read()
    ctx.profile(False)
    K = args.steps
    n_alive = ctx.count(se) + ctx.count(si)
    value = 2 * N * K / T  # pushes issued (both species, every step)

    # the same step with the plain red-black SOR solve, for comparison
    compare = None
    if step_solver != "sor_rb":
        kw = solver_kw("sor_rb")
        step(kw)
        solve_iters.clear()
        ctx.sync()
        c0 = time.perf_counter()
        for k in range(K):
            step(kw)
        ctx.sync()
        c1 = time.perf_counter()
        compare = {"sor_rb": {"value": 2 * N * K / (c1 - c0), "ms_per_step": (c1 - c0) / K * 1e3,
                              "poisson_iterations_per_step": float(np.mean(solve_iters))}}
        print("same step with sor_rb: %.3e pushes/s" % compare["sor_rb"]["value"], file=sys.stderr)

    stages = {k: {"ms_per_step": round(v[0] / PROF_STEPS, 4), "launches": v[1]}
              for k, v in prof.items() if v[1]}
    # dominant kernel by total device time among the bandwidth kernels
    cand = {k: v for k, v in prof.items() if k in ALG_BYTES and v[1]}
    dom = max(cand, key=lambda k: cand[k][0])
    tot_ms, launches = cand[dom]
    avg_s = tot_ms * 1e-3 / launches
    if dom == "pcg":
        units = ctx.nn
    elif dom == "sor":
        units = ctx.nn  # one profiled launch = one full red-black sweep
    else:
        units = N
    per_unit = ALG_BYTES[dom]
    if dom == "pcg" and step_solver == "learned":
        per_unit = ALG_BYTES_FCG
    alg = per_unit * units
    achieved = alg / avg_s / 1e9
    # the network's forward as a whole against the dense BF16 MFMA peak
    roof_net = None
    if net_meta is not None and prof.get("unet", (0, 0))[1]:
        fwd_s = prof["unet"][0] * 1e-3 / prof["unet"][1]
        fl = unet_flops(net_meta["channels"], net_meta["coarsen"], n)
        ach = fl / fwd_s / 1e12
        roof_net = {"bound": "mfma", "kernel": "U-Net forward (k_conv_in, k_conv3_rb, k_conv3_small, "
                                               "k_up1_add, k_conv_out, k_pool)",
                    "achieved": ach, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": ach / MFMA_BF16_PEAK_TFLOPS, "traffic": None,
                    "alg_flops_per_forward": fl, "avg_forward_ms": fwd_s * 1e3}
    traffic = measured_traffic(dom)
    # every bandwidth stage of the step against the HBM peak
    rooflines = {}
    for k in ("push", "scatter", "deposit"):
        if prof.get(k, (0, 0))[1]:
            launches_k = prof[k][1]
            avg_k = prof[k][0] * 1e-3 / launches_k
            if k == "push":  # the histogram pass (one launch per species)
                b = ALG_BYTES["push"] * N
            elif k == "scatter":
                b = ALG_BYTES["scatter"] * N
            else:  # one deposit call (k_dep_cells + k_dep_nodes): 16 B per particle + 8 B per node
                b = ALG_BYTES["deposit"] * N + 8.0 * ctx.nn
            rooflines[k] = {"kernel": KERNEL_NAMES[k], "achieved": b / avg_k / 1e9, "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": b / avg_k / 1e9 / HBM_PEAK_GBS, "alg_bytes_per_launch": b,
                            "avg_launch_ms": avg_k * 1e3, "traffic": measured_traffic(k)}
    deposits_per_s = None
    if "deposit" in rooflines:
        deposits_per_s = N / (rooflines["deposit"]["avg_launch_ms"] * 1e-3)

    line = {
        "metric": "particle pushes/s + time-to-1e-6 Poisson residual, learned vs plain, 1 GPU",
        "value": value,
        "unit": "particle pushes/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": T / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 particles, f64 Poisson residual",
        "data": "synthetic (uniform positions, Maxwellian velocities, seeded Philox)",
        "config": {"workload": "scaled plain PIC, %d^3 nodes, 2 x %d particles, counting sort "
                               "every step, Poisson to 1e-6 rel. residual every step" % (n, N),
                   "grid_nodes": n ** 3, "particles_per_species": N, "dt": dt,
                   "poisson_step_solver": step_solver, "poisson_warm_start": args.warm_start},
        "particles_alive": n_alive,
        "poisson_iterations_per_step": iters_main,
        "step_solver_compare": compare,
        "stages": stages,
        "roofline": {"bound": "hbm",
                     "kernel": ("k_fcg_pq + k_fcg_update" if dom == "pcg" and step_solver == "learned"
                                else KERNEL_NAMES[dom]),
                     "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "alg_bytes_per_launch": alg,
                     "avg_launch_ms": avg_s * 1e3},
        "roofline_network": roof_net,
        "rooflines": rooflines,
        "deposits_per_s": deposits_per_s,
        "poisson": poisson,
        "cpu_baseline": cpu,
    }
    if not args.no_extras:
        line["extras"] = extras_bench(pg, ctx, args, se, N, n, cpu, cpu_newton, x0, xm)
        print("extras: %s" % json.dumps({k: round(v.get("seconds", 0), 5) for k, v in line["extras"].items()}),
              file=sys.stderr)
    ctx.close()
    if not args.no_dsmc:
        line["dsmc"] = dsmc_bench(pg, args, cpu_dsmc)
        print("dsmc: %.3e particles/s, %.3f ms per step" % (line["dsmc"]["value"], line["dsmc"]["ms_per_step"]),
              file=sys.stderr)
    if not args.no_discharge:
        line["discharge_loop"] = discharge_bench(pg, args)
        print("discharge loop: %.3f ms per step" % line["discharge_loop"]["ms_per_step"], file=sys.stderr)
    if not args.no_c5:
        line["c5"] = c5_bench(pg, args)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
