#!/usr/bin/env python3
"""Benchmark: particle-updates/s of the CUBE PM+PP step on one MI355X.

Workload (the 512^3 configuration): 512^3 particles in a 400 Mpc/h periodic
box, nc=256 coarse cells (1024^3 fine mesh), 2^3 tiles (ncb=6 buffers; the
fine mesh is one periodic 1024^3 mesh on the GPU), int8 positions, int16
velocities, PP on (pp_range=1, rsoft=0.3), Zel'dovich ICs at z=49 from
seeded white noise (seed 42) and an Eisenstein-Hu no-wiggle LCDM spectrum
(Om=0.31, Ob=0.049, h=0.68, ns=0.96, sigma8=0.81).

One step is the reference's step loop body (CUBEnu/work/main/main.f90:102-118):
timestep (host) -> update_x -> buffer_grid -> buffer_x -> ext_pp_force ->
particle_mesh -> buffer_v.  Inputs are resident on the GPU before timing.

Prints ONE JSON line.  `roofline` is for the dominant kernel, from HIP-event
timings on the library's stream; `step_roofline` is the whole step against
the algorithmic bytes B_step = N_p*beta_p + N_f*54 + N_c*86; `cpu_baseline`
times the oracle (the C restatement of the reference path, OpenMP, with
scipy FFTs) on every host thread of this process's CPU share, one full step
from the same 512^3 initial state.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-updates/s at 512^3 (PM+PP step); batched adjoint grads/s vs plain"
HBM_PEAK = 8.0e12
FP64_PEAK_TF = 78.6  # MI355X FP64, TFLOP/s
FLOP_PER_PAIR = 33


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_state_gpu(g, box, z_i, seed, particles=True):
    """Zel'dovich ICs generated on the GPU straight into the context (the
    host copy of the state in `host`; particles=False: its cell arrays)."""
    t = time.time()
    a, sigma = g.make_ics(box, z_i, seed)
    log(f"[bench] GPU Zel'dovich ICs: {g.nplocal} particles in {time.time() - t:.1f}s, sigma_vi={sigma:.5g}")
    st = g.get_state(particles=particles)
    nc = g.nc
    # IC velocity constraint and vz_max (WU/initial_conditions.f90:631-641),
    # from the decoded cell velocities plus residuals bounded by the codec
    vf = st["vfield"]
    vz = float(np.abs(vf[:, 2]).max()) + 3 * sigma * 2.5
    vmax = float(np.abs(vf).max()) + 3 * sigma * 2.5
    return dict(a=a, sigma_vi=sigma, np=st["nplocal"], vz_max=vz,
                dt_vmax=float(np.float32(0.9) * np.float32(20) / np.float32(vmax)), host=st)


class Step:
    """The step loop body of W/main.f90:102-118 on the GPU context."""

    def __init__(self, cube, st, mass_p, omega_m=0.31):
        from cafproject_amd.timestep import Clock
        self.g = cube
        self.mass = mass_p
        self.clock = Clock(st["a"], [0.0], omega_m, 1 - omega_m)
        self.sigma = st["sigma_vi"]
        self.vz_max = st["vz_max"]
        self.dtc = dict(dt_fine=1000.0, dt_coarse=1000.0, dt_pp=1000.0, dt_vmax=st["dt_vmax"])
        self.first = None  # (dt_old, dt, a_mid) of the first step, for the CPU baseline

    def __call__(self):
        g, c = self.g, self.clock
        c.timestep(**self.dtc)
        if self.first is None:
            self.first = dict(dt_old=float(c.dt_old), dt=float(c.dt), a_mid=float(c.a_mid))
        vs = g.update_x(float(c.dt_old), float(c.dt), self.vz_max, self.sigma)
        g.buffer_grid()
        g.buffer_x()
        a_mid, dt = float(c.a_mid), float(c.dt)
        GG = np.float32(1.0) / np.float32(6.0) / np.float32(np.pi)
        if g.pp_on:
            f2pp = g.ext_pp_force(a_mid, dt, self.sigma, self.mass)
            self.dtc["dt_pp"] = float(np.float32(1.0) / (np.sqrt(np.float32(f2pp)) * np.float32(a_mid) * GG))
        pm = g.particle_mesh(a_mid, dt, self.sigma, vs["sigma_vi_new"], self.mass)
        self.sigma = vs["sigma_vi_new"]
        g.buffer_v()
        self.dtc.update(dt_fine=pm["dt_fine"], dt_coarse=pm["dt_coarse"], dt_vmax=pm["dt_vmax"])
        self.vz_max = pm["vz_max"]
        return vs, pm


def stage_bytes(np_, nf3, nc3, bx, bv):
    """Algorithmic bytes per launch of each stage (the B_step split: every
    stage reads its inputs once and writes its outputs once)."""
    return {
        "drift": np_ * (2 * bx + 2 * bv) + nc3 * 32,
        "pp": np_ * (bx + 2 * bv),
        "fine_deposit": np_ * bx + nf3 * 4,
        "fine_fft": nf3 * 8,
        "fine_inverse": nf3 * 30,
        "fine_kick": np_ * (bx + 2 * bv) + nf3 * 12,
        "coarse_deposit": np_ * bx + nc3 * 4,
        "coarse_force": nc3 * 38,
        "coarse_kick": np_ * (bx + 2 * bv) + nc3 * 12,
    }


MAIN_KERNEL = {
    "drift": "k_drift_dest+k_place_ranked+k_sort_small+k_drift_final2",
    "pp": "k_pp_blk", "fine_deposit": "k_fine_deposit_cc", "fine_kick": "k_kick<fine>",
    "coarse_deposit": "k_coarse_deposit_b2", "coarse_kick": "k_kick<coarse>",
    "fine_fft": "k_r2c_x+k_c2c_r", "fine_inverse": "k_green_r+k_c2c_r+k_c2r_x3",
    "coarse_force": "coarse fft kernels",
}
ONE_PASS_KICKS = "k_kick<fine+coarse> (both kicks in one particle pass)"
# stages whose kernels share the GPU with other streams' (ext_pp_force forms
# the fine and coarse PM forces on side streams beside the PP kernel)
CONCURRENT = {"pp": "fine_deposit+fine_fft+fine_inverse+coarse_deposit+coarse_force",
              "fine_deposit": "pp+coarse_deposit+coarse_force", "coarse_deposit": "pp+fine_deposit+fine_fft",
              "coarse_force": "pp+fine_fft"}
# kernel-name prefixes in the rocprofv3 PMC traffic table, per single-kernel stage
PMC_NAMES = {"pp": ["k_pp_blk<"], "fine_deposit": ["k_fine_deposit_cc<"], "fine_kick": ["k_kick<"],
             "coarse_deposit": ["k_coarse_deposit_b2<"], "coarse_kick": ["k_kick<"]}
PMC_TABLE = os.path.join(ROOT, "profiles", "pmpp512_step_pmc_traffic.md")


def pmc_traffic(stage, nc):
    """HBM bytes per launch of a stage's kernel from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE table (measured on the 512^3 workload), or None."""
    if nc != 256 or stage not in PMC_NAMES or not os.path.exists(PMC_TABLE):
        return None
    rows = {}
    for line in open(PMC_TABLE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[0].startswith("`"):
            rows[cells[0].strip("`")] = float(cells[5]) * 1e9
    fine = stage == "fine_kick"
    for name, b in rows.items():
        if any(name.startswith(p) for p in PMC_NAMES[stage]):
            if name.startswith("k_kick<") and ("true" in name) != fine:
                continue
            return b
    return None


def host_cores():
    """Threads this process may use on the host (its CPU share) and the CPU model."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return n, model


def cpu_baseline(state, nc, nnt, box, dts, reps=1, izipx=1):
    """The oracle (the C restatement of the reference path, OpenMP over the
    reference's own parallel loops, scipy FFTs with the same thread count
    standing in for threaded FFTW) on every host thread this process has,
    on the same 512^3 initial state the GPU stepped: one full step
    (update_x, buffers, PP, particle_mesh, buffer_v) per repetition, the
    median reported, with per-stage times."""
    from oracle import oracle as O
    threads, model = host_cores()
    O.FFT_WORKERS = threads
    O.lib().oc_set_threads(threads)
    t = time.time()
    oc = O.OracleCube(nc, nnt, izipx=izipx, izipv=2, pp_on=True, has_pid=False)
    oc.init_kernels()
    n = int(state["nplocal"])
    log(f"[bench] oracle context + Green's functions in {time.time() - t:.1f}s ({threads} threads, {model})")
    m = float(np.float32((4 * nc) ** 3 / n))
    sig = state["sigma_vi"]
    times, stages = [], []
    for r in range(reps):
        oc.set_state(state["xp"], state["vp"], None, state["rhoc"], state["vfield"])
        oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
        st = {}
        t0 = time.perf_counter()
        s2, _ = oc.update_x(dts["dt_old"], dts["dt"], state["vz_max"], sig)
        t1 = time.perf_counter()
        oc.buffer_grid(); oc.buffer_x()
        t2 = time.perf_counter()
        oc.ext_pp_force(dts["a_mid"], dts["dt"], sig, m)
        t3 = time.perf_counter()
        oc.particle_mesh(dts["a_mid"], dts["dt"], sig, s2, m)
        t4 = time.perf_counter()
        oc.buffer_v()
        t5 = time.perf_counter()
        st = dict(update_x=t1 - t0, buffers=(t2 - t1) + (t5 - t4), ext_pp_force=t3 - t2, particle_mesh=t4 - t3)
        times.append(t5 - t0)
        stages.append(st)
        log(f"[bench] oracle step {r}: {t5 - t0:.1f} s " + " ".join(f"{k}={v:.2f}" for k, v in st.items()))
    k = int(np.argsort(times)[len(times) // 2])
    tmed = times[k]
    return dict(value=n / tmed, unit="particle-updates/s", cores=threads, kind="port",
                sample=f"one full oracle step (update_x, buffers, ext_pp_force, particle_mesh, buffer_v) of the "
                       f"same {n}-particle initial state (nc={nc}, {nnt}^3 tiles, x{izipx}v2, PP on), median of "
                       f"{reps}: {tmed:.1f} s on {threads} host threads ({model}); scipy FFTs on the same threads",
                seconds=tmed, stage_seconds={a: round(b, 3) for a, b in stages[k].items()}, cpu_model=model)


def large_box(steps=3, warmup=1):
    """The large single-GPU box: 1024^3 particles, 800 Mpc/h, nc=512 (2048^3
    fine mesh, compact layout), 4^3 tiles, int16 x and v, PP on, Zel'dovich
    ICs at z=49 (seed 7): step time, particle-updates/s and each stage's
    algorithmic GB/s (per step: in the compact layout the inverse FFT
    stages launch once per force component)."""
    from cafproject_amd.cube import Cube
    nc, nnt, box = 512, 4, 800.0
    t = time.time()
    g = Cube(nc, nnt, izipx=2, izipv=2, has_pid=False, pp_on=True, pp_range=1, rsoft=0.3)
    g.init_kernels()
    st = make_state_gpu(g, box, 49.0, 7, particles=False)
    del st["host"]
    np_ = st["np"]
    log(f"[bench] large box ready in {time.time() - t:.1f}s")
    step = Step(g, st, float(np.float32((4 * nc) ** 3 / np_)))
    for i in range(warmup):
        step()
    g.sync()
    g.profile(True)
    g.timers(reset=True)
    t0 = time.perf_counter()
    for i in range(steps):
        step()
    g.sync()
    el = time.perf_counter() - t0
    tm = g.timers()
    g.profile(False)
    ms = {k: v[0] / steps for k, v in tm.items()}
    ms["fine_inverse"] = sum(ms[k] for k in ("fine_green", "fine_invy", "fine_c2r"))
    sb = stage_bytes(np_, (4 * nc) ** 3, nc ** 3, 6, 6)
    if tm["coarse_kick"][1] == 0 and tm["fine_kick"][1] > 0:
        sb["fine_kick"] += sb.pop("coarse_kick")
    stages = {k: dict(ms_per_step=round(ms[k], 3), bytes=int(b),
                      gbps=round(b / (ms[k] * 1e-3) / 1e9, 1) if ms[k] > 0 else None) for k, b in sb.items()}
    out = dict(workload="CUBE PM+PP step, 1024^3 particles, 800 Mpc/h, nc=512 (2048^3 fine mesh, compact layout), "
                        "4^3 tiles, int16 x / int16 v, PP on, z=49 Zel'dovich EH no-wiggle, seed 7",
               particles=np_, steps=steps, warmup=warmup, ms_per_step=1e3 * el / steps,
               value=np_ * steps / el, unit="particle-updates/s", stages=stages, a_final=float(step.clock.a))
    del step, g
    return out


def batched_gradients(iters=3):
    """The second half of the metric: forward (10 KDK steps) + MSE loss +
    adjoint gradients of an uneven ensemble of 48 differentiable-PM boxes
    (16x32^3, 16x64^3, 8x96^3, 8x128^3 particles, fine mesh 2x the particle
    grid, 50-200 Mpc/h, sigma8 in [0.7, 0.9], seeds 1000-1047), as one
    replayed HIP graph against the per-box loop of the same kernels on one
    stream.  grads/s = boxes / wall time of one ensemble pass."""
    import torch
    from cafproject_amd import dpm
    t = time.time()
    b, tens = dpm.make_ensemble()
    nbox = len(tens)
    shapes = b.tune()
    log(f"[bench] ensemble of {nbox} boxes built and tuned in {time.time() - t:.1f}s")
    res = {}
    for name, graph in (("per_box_loop", False), ("graph", True)):
        b.run(graph=graph)  # warm-up (and capture)
        b.sync()
        t0 = time.perf_counter()
        for _ in range(iters):
            b.run(graph=graph)
        b.sync()
        dt = (time.perf_counter() - t0) / iters
        res[name] = dict(ms_per_pass=1e3 * dt, grads_per_s=nbox / dt)
        log(f"[bench] ensemble {name}: {1e3 * dt:.1f} ms per pass, {nbox / dt:.1f} grads/s")
    torch.cuda.synchronize()
    loss = [float(x["loss"].item()) for x in tens]
    b.close()
    return {"metric": "batched adjoint grads/s (forward + adjoint gradient per box)", "boxes": nbox,
            "workload": "48 boxes: 16x32^3, 16x64^3, 8x96^3, 8x128^3 particles, fine mesh 2x particle grid, "
                        "10 KDK steps a=0.1->1, MSE loss, per-step checkpoints",
            "graph_grads_per_s": res["graph"]["grads_per_s"], "loop_grads_per_s": res["per_box_loop"]["grads_per_s"],
            "speedup": res["graph"]["grads_per_s"] / res["per_box_loop"]["grads_per_s"], "iters": iters,
            "detail": res, "tuned_shapes": sorted({(int(a), int(c)) for a, c in shapes}),
            "loss_finite": bool(np.all(np.isfinite(loss)))}


Z_BINS = (49.0, 20.0, 10.0, 5.0, 3.0, 2.0, 1.0, 0.5, 0.25, 0.0)
Z_PROBES = (1.0, 0.0)


def pair_count(st, nc, nnt):
    """PP pair visits of a state: the sum over particles of the particles in
    their 27 neighbouring fine cells (the particle itself included), the
    (2*pp_range+1)^3 walk of W/ext_pp_force.f90:94-115 at pp_range=1; counted
    on the GPU with torch from the cell-sorted codes."""
    import torch
    dev = torch.device("cuda")
    nt, nf = nc // nnt, 4 * nc
    rhoc = torch.from_numpy(st["rhoc"]).to(dev)
    cell = torch.repeat_interleave(torch.arange(rhoc.numel(), device=dev, dtype=torch.int64), rhoc.long())
    loc = [cell % nt, (cell // nt) % nt, (cell // nt ** 2) % nt]
    t = cell // nt ** 3
    tile = [t % nnt, (t // nnt) % nnt, t // nnt ** 2]
    del cell, t
    codes = torch.from_numpy(st["xp"].view(np.uint8) if st["xp"].dtype == np.int8 else
                             (st["xp"].view(np.uint16) >> 8).astype(np.uint8)).to(dev).long()
    f = [4 * (tile[a] * nt + loc[a]) + (codes[:, a] >> 6) for a in range(3)]
    del loc, tile, codes
    n = torch.bincount((f[2] * nf + f[1]) * nf + f[0], minlength=nf ** 3).view(nf, nf, nf).to(torch.int32)
    del f
    s = n.clone()
    for ax in (2, 1, 0):
        s = s + torch.roll(s, 1, ax) + torch.roll(s, -1, ax)
    pairs = int((n.long() * s.long()).sum().item())
    del n, s
    torch.cuda.empty_cache()
    return pairs


def alone(st, nc, nnt, izipx, a_mid, dt, sigma, mass, drift_args, reps=3):
    """PP, both deposits and the drift on a copy of the state in a second
    context without Green's functions (so nothing runs beside PP): fastest
    of reps."""
    from cafproject_amd.cube import Cube
    g2 = Cube(nc, nnt, izipx=izipx, izipv=2, has_pid=False, pp_on=True, pp_range=1, rsoft=0.3)
    best = dict(pp=1e30, fine_deposit=1e30, coarse_deposit=1e30, drift=1e30)
    for _ in range(reps):
        g2.set_state(st["xp"], st["vp"], None, st["rhoc"], st["vfield"])
        g2.sync()
        for k, fn in (("pp", lambda: g2.ext_pp_force(a_mid, dt, sigma, mass)),
                      ("fine_deposit", lambda: g2.fine_deposit(mass)), ("coarse_deposit", lambda: g2.coarse_deposit(mass)),
                      ("drift", lambda: g2.update_x(*drift_args))):
            t = time.perf_counter()
            fn()
            g2.sync()
            best[k] = min(best[k], 1e3 * (time.perf_counter() - t))
    del g2
    return {k: round(v, 2) for k, v in best.items()}


def to_z0(nc, nnt, box, izipx, pp_on, max_steps=3000):
    """The whole run of the headline workload, z=49 -> 0 with the
    reference's step criterion (W/main.f90:88-145): every step's wall time
    and stage times (PP, deposits, drift, kicks) binned by redshift, the
    particle count after every step, P(k) of the final state, and at z = 1
    and z = 0 the state's PP pair count with PP, the deposits and the drift
    timed alone (outside the run's clock)."""
    from cafproject_amd.cube import Cube
    g = Cube(nc, nnt, izipx=izipx, izipv=2, has_pid=False, pp_on=pp_on, pp_range=1, rsoft=0.3)
    g.init_kernels()
    st = make_state_gpu(g, box, 49.0, 42, particles=False)
    n = st["np"]
    del st["host"]
    mass = float(np.float32((4 * nc) ** 3 / n))
    step = Step(g, st, mass)
    g.profile(True)
    g.timers(reset=True)
    rows = []
    probes = []
    t0 = time.perf_counter()
    zp = list(Z_PROBES)
    while not step.clock.final_step and len(rows) < max_steps:
        a0 = float(step.clock.a)
        t = time.perf_counter()
        step()
        g.sync()
        ms = 1e3 * (time.perf_counter() - t)
        tm = {k: v[0] / max(v[1], 1) for k, v in g.timers(reset=True).items()}
        if g.nplocal != n:
            raise RuntimeError(f"particle count {g.nplocal} != {n} after step {len(rows) + 1}")
        a1 = float(step.clock.a)
        rows.append(dict(z=1 / a1 - 1, ms=ms, pp=tm.get("pp", 0.0), drift=tm["drift"],
                         fine_deposit=tm["fine_deposit"], coarse_deposit=tm["coarse_deposit"],
                         kick=tm["fine_kick"] + tm["coarse_kick"], da=a1 - a0))
        if zp and (1 / a1 - 1 <= zp[0] + 1e-6):
            # PP and the deposits alone on this state, and its pair count
            # (outside the run's clock)
            tp = time.perf_counter()
            st_h = g.get_state()
            pairs = pair_count(st_h, nc, nnt)
            al = alone(st_h, nc, nnt, izipx, float(step.clock.a), float(step.clock.dt), step.sigma, mass,
                       (float(step.clock.dt_old), float(step.clock.dt), step.vz_max, step.sigma))
            del st_h
            probes.append(dict(z=round(1 / a1 - 1, 4), pairs=pairs, pairs_per_particle=round(pairs / n, 2),
                               alone_ms=al, pp_pairs_per_s=pairs / (al["pp"] * 1e-3)))
            log(f"[bench] z0 run probe at z={1 / a1 - 1:.3f}: {pairs / n:.1f} pairs per particle, alone {al}")
            zp.pop(0)
            t0 += time.perf_counter() - tp
        if len(rows) % 10 == 0 or step.clock.final_step:
            log(f"[bench] z0 run step {len(rows)} a={a1:.4f} z={1 / a1 - 1:.3f} {ms:.1f} ms "
                f"(pp {rows[-1]['pp']:.1f}, deposits {tm['fine_deposit']:.1f}/{tm['coarse_deposit']:.1f})")
    total = time.perf_counter() - t0
    g.profile(False)
    bins = []
    for zhi, zlo in zip(Z_BINS[:-1], Z_BINS[1:]):
        sel = [r for r in rows if zlo <= r["z"] < zhi or (zhi == Z_BINS[0] and r["z"] >= zhi)]
        if not sel:
            continue
        mean = {k: round(sum(r[k] for r in sel) / len(sel), 2)
                for k in ("ms", "pp", "drift", "fine_deposit", "coarse_deposit", "kick")}
        bins.append(dict(z_range=[zhi, zlo], steps=len(sel), ms_per_step=mean["ms"],
                         max_ms=round(max(r["ms"] for r in sel), 2),
                         stage_ms={k: v for k, v in mean.items() if k != "ms"},
                         pp_max_ms=round(max(r["pp"] for r in sel), 2)))
    k, d2, cnt = g.power(2 * nc, box)
    pk = {f"{kk:.3f}": float(d2[int(np.argmin(np.abs(k - kk)))]) for kk in (0.05, 0.1, 0.3, 1.0)}
    del step, g
    return {
        "metric": "wall seconds of the run z=49 -> 0 (512^3 PM+PP)", "value": total, "unit": "s",
        "n_gpus": 1, "steps": len(rows), "higher_is_better": False, "dtype": "f32", "data": "synthetic",
        "ms_per_step": 1e3 * total / max(len(rows), 1), "vs_baseline": None,
        "config": {"workload": f"CUBE PM+PP run z=49 -> 0, {round(n ** (1 / 3))}^3 particles, {box:g} Mpc/h, "
                               f"nc={nc}, {nnt}^3 tiles, int{8 * izipx} x / int16 v, PP {'on' if pp_on else 'off'}, "
                               f"Zel'dovich EH no-wiggle, seed 42", "particles": n},
        "a_final": 1 / (1 + rows[-1]["z"]) if rows else None,
        "particles_conserved": True, "by_redshift": bins, "alone_at": probes,
        "delta2_z0": {"k_h_per_mpc": list(pk), "delta2": list(pk.values())},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nc", type=int, default=256)
    ap.add_argument("--nnt", type=int, default=2)
    ap.add_argument("--box", type=float, default=400.0)
    ap.add_argument("--no-pp", action="store_true")
    ap.add_argument("--izipx", type=int, default=1, choices=(1, 2), help="bytes per position component")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-gradient ensemble measurement")
    ap.add_argument("--batched-iters", type=int, default=3)
    ap.add_argument("--cpu-reps", type=int, default=3, help="oracle steps timed for the CPU baseline (median)")
    ap.add_argument("--no-large", action="store_true", help="skip the 1024^3-particle large-box measurement")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 1 GiB device-to-device copy kernel (torch), the reference "
                         "point for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE counters")
    ap.add_argument("--to-z0", action="store_true",
                    help="instead of the timed steps: the whole run z=49 -> 0, step times by redshift")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU (the whole box on one device); "
              f"--gpus {args.gpus} is not supported. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.to_z0:
        print(json.dumps(to_z0(args.nc, args.nnt, args.box, args.izipx, not args.no_pp)), flush=True)
        return
    from cafproject_amd.cube import Cube
    if args.calibrate_copy:
        import torch
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] calibration copy: 2 x 1 GiB read + 1 GiB written")
        del src, dst
        torch.cuda.empty_cache()
    nc, nnt = args.nc, args.nnt
    izipx, izipv = args.izipx, 2
    pp_on = not args.no_pp
    t = time.time()
    g = Cube(nc, nnt, izipx=izipx, izipv=izipv, has_pid=False, pp_on=pp_on, pp_range=1, rsoft=0.3)
    g.init_kernels()
    log(f"[bench] context + Green's functions in {time.time() - t:.1f}s")
    st = make_state_gpu(g, args.box, 49.0, 42)
    np_ = st["np"]
    mass = float(np.float32((4 * nc) ** 3 / np_))
    step = Step(g, st, mass)
    for i in range(args.warmup):
        t = time.perf_counter()
        step()
        log(f"[bench] warmup step {i} {1e3 * (time.perf_counter() - t):.1f} ms a={float(step.clock.a):.5f}")
    g.sync()

    def pairs_now():
        """PP pair visits of the current state (outside the timed region)."""
        if not pp_on:
            return None
        try:
            return pair_count(g.get_state(), nc, nnt)
        except Exception as e:  # the count must not hide the measurement
            log(f"[bench] pair count skipped: {e}")
            return None

    pairs0 = pairs_now()
    g.profile(True)
    g.timers(reset=True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
    g.sync()
    elapsed = time.perf_counter() - t0
    tm = g.timers()
    g.profile(False)
    pairs1 = pairs_now()
    ms_step = 1e3 * elapsed / args.steps
    value = np_ * args.steps / elapsed
    bx, bv = 3 * izipx, 3 * izipv
    nf3, nc3 = (4 * nc) ** 3, nc ** 3
    beta = 7 * bx + 8 * bv if pp_on else 6 * bx + 6 * bv
    B_step = np_ * beta + nf3 * 54 + nc3 * 86
    sb = stage_bytes(np_, nf3, nc3, bx, bv)
    if not pp_on:
        sb.pop("pp")
    stage_ms = {k: v[0] / max(v[1], 1) for k, v in tm.items()}
    stage_ms["fine_inverse"] = sum(stage_ms[k] for k in ("fine_green", "fine_invy", "fine_c2r"))
    # particle_mesh runs both kicks as one pass, timed as fine_kick: its
    # algorithmic bytes are the two kick stages' of the B_step model (the
    # reference's fine and coarse kicks each read x, v and write v)
    one_pass = tm["coarse_kick"][1] == 0 and tm["fine_kick"][1] > 0
    if one_pass:
        sb["fine_kick"] += sb.pop("coarse_kick")
    per_launch = {k: dict(ms=round(stage_ms[k], 4), bytes=int(b),
                          gbps=round(b / (stage_ms[k] * 1e-3) / 1e9, 1) if stage_ms[k] > 0 else None)
                  for k, b in sb.items()}
    # the dominant kernel by time among the single-kernel stages; a stage
    # that runs beside other streams' work (PP, the deposits) is timed by
    # its event span, which includes the time it shares the GPU
    single = [k for k in ("fine_deposit", "pp", "fine_kick", "coarse_deposit", "coarse_kick") if k in per_launch]

    def kernel_entry(k):
        d = per_launch[k]
        ach = d["bytes"] / (d["ms"] * 1e-3)
        tr = pmc_traffic(k, nc)
        return {"bound": "hbm", "kernel": (ONE_PASS_KICKS if one_pass and k == "fine_kick" else MAIN_KERNEL[k]),
                "stage": k, "concurrent_with": CONCURRENT.get(k),
                "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": ach / HBM_PEAK,
                "traffic": tr, "traffic_source": os.path.relpath(PMC_TABLE, ROOT) if tr else None,
                "algorithmic_bytes_per_launch": d["bytes"], "avg_launch_ms": d["ms"]}

    dom = max(single, key=lambda k: per_launch[k]["ms"])
    roofline = kernel_entry(dom)
    if dom == "pp" and pairs0 and pairs1:
        # PP is FP64 pair arithmetic: the reference's pair force counts 33
        # FP64 operations (W/ext_pp_force.f90:103-110: xvec21 3, |r|^2 5 and
        # its sqrt 1, rcut = r/16 and the cutoff polynomial 10, r^3 2, the 3
        # quotients, mass and pcut products 6, the 3 sums; the position
        # decode and the *0 of soft pairs not counted); pairs per step from
        # the states before and after the window
        pairs = 0.5 * (pairs0 + pairs1)
        tf = pairs * FLOP_PER_PAIR / (per_launch["pp"]["ms"] * 1e-3) / 1e12
        roofline["fp64"] = {"achieved": tf, "peak": FP64_PEAK_TF, "unit": "TFLOP/s", "frac": tf / FP64_PEAK_TF,
                            "flop_per_pair": FLOP_PER_PAIR, "pairs_per_step": [pairs0, pairs1],
                            "pairs_per_s": pairs / (per_launch["pp"]["ms"] * 1e-3)}
    alone = [k for k in single if k not in CONCURRENT and k != dom]
    if alone:
        roofline["secondary"] = kernel_entry(max(alone, key=lambda k: per_launch[k]["ms"]))
    out = {
        "metric": METRIC, "value": value, "unit": "particle-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"CUBE PM+PP step, {round(np_ ** (1 / 3))}^3 particles, {args.box:g} Mpc/h, "
                               f"nc={nc} ({4 * nc}^3 fine mesh), {nnt}^3 tiles, int{8 * izipx} x / int16 v, "
                               f"PP {'on' if pp_on else 'off'}, z=49 Zel'dovich EH no-wiggle, seed 42",
                   "particles": np_, "nc": nc, "nnt": nnt, "nf": 4 * nc, "pp": pp_on,
                   "izipx": izipx, "izipv": izipv},
        "roofline": roofline,
        "step_roofline": {"B_step": B_step, "achieved_GBps": B_step / (ms_step * 1e-3) / 1e9,
                          "frac": B_step / (ms_step * 1e-3) / HBM_PEAK},
        "stages": per_launch,
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "a_final": float(step.clock.a),
    }
    first = step.first
    del step, g
    if not args.no_large:
        out["large_box"] = large_box()
    if not args.no_batched:
        out["batched"] = batched_gradients(args.batched_iters)
    if not args.no_cpu_baseline:
        try:
            host = dict(st["host"], sigma_vi=st["sigma_vi"], vz_max=st["vz_max"])
            out["cpu_baseline"] = cpu_baseline(host, nc, nnt, args.box, first, reps=args.cpu_reps, izipx=izipx)
            out["cpu_baseline"]["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
        except Exception as e:  # the baseline must not hide the GPU number
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
