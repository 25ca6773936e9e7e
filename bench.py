#!/usr/bin/env python3
"""Benchmark of the FastPM particle-mesh step on one MI355X.

Workload: a single box of side 1000 Mpc/h with 256^3 particles on a 512^3 force
mesh (B = 2), flat LCDM (Om0 = 0.3, h = 0.7, ns = 0.96, sigma8 = 0.8) with
the no-wiggle Eisenstein-Hu linear power, white-noise seed 7, 2LPT at a = 0.1
and 40 kick-drift-kick steps linear in a to a = 1, CIC paint/readout with
window compensation.  Inputs are synthetic (generated on the device).

A step is one K, D, F, K cycle; the timed steps run as Solver.nbody runs
them without a monitor (a step boundary's closing kick, the next opening kick
and the drift in one pass, bit for bit the separate updates; the Force's
delta_k not stored since nobody receives it).  The float32 run (integer
lattice index plus float32 displacement) is timed over --steps steps after
--warmup untimed steps, then run on to a = 1; the float64 plain run (one transform and one
readout per force component) is timed the same way.  The final states are
compared: P(k) relative error for k < k_Nyq/2, r(k), and the largest periodic
displacement difference.  Further legs (each can be skipped): the float64
oracle's KDK step on the host cores (cpu_baseline), 64 boxes of uneven sizes
replayed from a captured graph (ensemble), the differentiable model's forward
and adjoint pass at 128^3 / 256^3 (gradient) and 512^3 particles on a
1024^3 TSC mesh (large_box).  One JSON line is printed.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fastpm-python_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=38)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nc", type=int, default=256)
    ap.add_argument("--B", type=int, default=2)
    ap.add_argument("--box", type=float, default=1000.0)
    ap.add_argument("--nsteps", type=int, default=40, help="steps from a=0.1 to a=1")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--no-fp64", action="store_true", help="skip the float64 plain run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the 64-box batched ensemble (graph replay) measurement")
    ap.add_argument("--no-grad", action="store_true",
                    help="skip the differentiable-model (forward + adjoint) measurement")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the large box: 512^3 particles on 1024^3, TSC, FP32 vs FP64")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 1 GiB device copy (reference kernel for HBM counters)")
    return ap.parse_args()


def run_sim(args, dtype, fused, timed, resampler="cic", a0=0.1):
    """Runs the whole simulation; times `timed` steps after `args.warmup`."""
    import numpy as np
    import torch
    from fastpm.core import Solver, leapfrog, KernelTimer
    from fastpm.pm import ParticleMesh
    from fastpm.cosmology import Cosmology, LinearPower

    cosmo = Cosmology(Om0=0.3, h=0.7, n_s=0.96, sigma8=0.8, Ob0=0.05, Tcmb0=2.7255)
    Pk = LinearPower(cosmo)
    pm = ParticleMesh(Nmesh=[args.nc] * 3, BoxSize=args.box, dtype=dtype, resampler=resampler)
    solver = Solver(pm, cosmo, B=args.B, compensate=True, fused=fused, autotune=fused)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dlin = solver.linear(solver.whitenoise(args.seed), Pk)
    Q = pm.generate_uniform_particle_grid(shift=0.0)
    state = solver.lpt(dlin, Q, a=a0, order=2)
    del dlin
    torch.cuda.synchronize()
    t_ic = time.perf_counter() - t0
    stages = np.linspace(a0, 1.0, args.nsteps + 1)
    actions = list(leapfrog(stages))
    step = solver.nbodystep
    step.run(*actions[0], state, None)  # initial force
    body = actions[1:]
    W, K = args.warmup, timed
    if W + K > args.nsteps:
        raise SystemExit("--warmup + --steps exceeds the %d steps of the run" % args.nsteps)
    # the actions of whole steps, as Solver.nbody applies them (no monitor:
    # kicks and drifts at step boundaries fused)
    step.run_actions(body[: 4 * W], state)
    timer = KernelTimer()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step.timer = timer
    step.run_actions(body[4 * W: 4 * (W + K)], state)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    step.timer = None
    kern = timer.summary()
    step.run_actions(body[4 * (W + K):], state)
    torch.cuda.synchronize()
    return dict(state=state, solver=solver, elapsed=elapsed, kern=kern, t_ic=t_ic,
                npart=len(Q), nmesh=int(solver.fpm.Nmesh[0]))


def gates(r32, r64, args):
    """P(k) error, r(k) and displacement error of float32 against float64."""
    import numpy as np
    import torch
    from fastpm.pm import RealField
    from fastpm.utils import power

    s64, s32 = r64["state"], r32["state"]
    fpm64 = r64["solver"].fpm
    m64 = fpm64.paint(s64.particles, mass=1.0, resampler="cic")
    m32 = r32["solver"].fpm.paint(s32.particles, mass=1.0, resampler="cic")
    m32d = RealField(fpm64, m32.value.double().contiguous())
    del m32
    c64, c32 = m64.r2c(), m32d.r2c()
    k, p64, _ = power(c64, nbins=32)
    _, p32, _ = power(c32, nbins=32)
    _, _, _, r = power(c32, c64, nbins=32)
    knyq = np.pi * r64["nmesh"] / args.box
    sel = (k < 0.5 * knyq) & np.isfinite(p64)
    dx = (s32.X.double() - s64.X + args.box / 2) % args.box - args.box / 2
    return dict(pk_rel_err_max=float(np.max(np.abs(p32[sel] / p64[sel] - 1))),
                rk_min=float(np.min(r[sel])), max_disp_err_mpch=float(dx.abs().max()),
                max_disp_err_cells=float(dx.abs().max()) * r64["nmesh"] / args.box)


def readout_kernel_name(fpm, order):
    """The compiled stencil-readout kernel that fpm's variant fields select
    (as rocprofv3 names it)."""
    v = fpm.variant
    nt = {0: 1024 if order == 2 else 512, 1: 256, 2: 512, 3: 1024}[v["readout"]]
    span = 2 if v["span"] == 1 and (int(fpm.Nmesh[2]) // 16) % 2 == 0 else 1
    if span == 2:
        nt = max(nt, 512)
    return "k_readout_grad_tiles<float, %d, %d, %s, %d>" % (
        order, nt, "true" if v["rho"] == 0 else "false", span)


def _profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` at this workload from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (profiles/traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t["kernels"][kernel]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def run_ensemble():
    """64 boxes of {32,48,64,96,128}^3 particles on twice-finer meshes, sides
    U[100, 800] Mpc/h, sigma8 U[0.7, 0.9], seeds 0..63, 2LPT at a = 0.1 and
    20 steps in FP32; each step of all boxes is one captured graph replay
    (the previous step's closing kick, opening kick and drift in one pass,
    then the force)."""
    import numpy as np
    import torch
    from fastpm.ensemble import Ensemble, config_specs

    stages = np.linspace(0.1, 1.0, 21)
    out = None
    for batched in (True, False):
        ens = Ensemble(config_specs(64), stages, batched=batched)
        t0 = time.perf_counter()
        ens.initialize()
        torch.cuda.synchronize()
        t_init = time.perf_counter() - t0
        ens.step_eager()
        torch.cuda.synchronize()
        ens.capture()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(ens.nsteps - 1):
            ens.graph.replay()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        ens.finish()  # the last closing kick (each replay began with the previous one)
        n = ens.nsteps - 1
        if batched:
            out = dict(workload="64 boxes, particle grids {32,48,64,96,128}^3, meshes 2x, 20 steps, "
                                "FP32, graph replay per step, batched launches (%d plans)"
                                % len(ens.plans),
                       particles=int(ens.npart),
                       cells=int(sum(int(b.solver.fpm.Nmesh.prod()) for b in ens.boxes)),
                       value=ens.npart * n / el, unit="particle-steps/s",
                       ms_per_step=el / n * 1e3, timed_steps=n, init_seconds=t_init)
        else:
            # the same step with each box's own launches on 16 streams
            out["per_box_streams_ms_per_step"] = el / n * 1e3
        del ens
        torch.cuda.empty_cache()
    return out


def region_bandwidth(kern, Np, Nc, Nk, bpe=4):
    """Algorithmic bytes of each timed region of one KDK step of the float32
    stencil path (per call) and the rate they imply: kick/drift 36 N_p, the
    fused kicks + drift of a step boundary 60 N_p (P, F, S in; P, S out), bin
    12 N_p (S read), the step boundary inside the next Force's binning pass
    60 N_p (its S output is binned from registers), paint 4 N_c, transforms
    8 N_c (density in, potential out; without a monitor the spectrum is not
    stored), readout 8 N_c + 28 N_p."""
    s = bpe / 4.0
    model = {"kick": 36 * Np, "drift": 36 * Np, "kick_drift": 60 * Np, "bin": 12 * Np,
             "kick_drift_bin": 60 * Np,
             "paint": 4 * Nc,
             "fft": 8 * Nc, "readout": 8 * Nc + 28 * Np}
    out = {}
    for k, (c, m) in kern.items():
        if k in model:
            gbs = model[k] * s / (m * 1e-3) / 1e9
            out[k] = {"bytes": int(model[k] * s), "GB_s": gbs, "frac": gbs / HBM_PEAK_GBS}
    return out


def run_gradient():
    """The differentiable forward model: L = 500 Mpc/h, 128^3 particles on a
    256^3 mesh (B = 2), 10 KDK steps from a = 0.1, float32, loss against the
    density of seed 99 divided by P(k); times one forward pass with per-step
    checkpoints and one adjoint pass to the white-noise gradient."""
    import numpy as np
    import torch
    from fastpm.core import Solver
    from fastpm.pm import ParticleMesh
    from fastpm.cosmology import Cosmology, LinearPower
    from fastpm.adjoint import ForwardModel, DensityLoss

    cosmo = Cosmology(Om0=0.3, h=0.7, n_s=0.96, sigma8=0.8)
    Pk = LinearPower(cosmo)
    pm = ParticleMesh(Nmesh=[128] * 3, BoxSize=500.0, dtype="f4")
    solver = Solver(pm, cosmo, B=2)
    model = ForwardModel(solver, Pk, np.linspace(0.1, 1.0, 11))
    loss = DensityLoss(solver, Pk, None)
    loss.target = loss.density_k(model.forward(solver.whitenoise(99)))
    w = solver.whitenoise(7)
    for _ in range(2):  # warm-up
        L, gS = loss(model.forward(w))
        model.backward(gS)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = model.forward(w)
    L, gS = loss(st)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    g = model.backward(gS)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    n = 128 ** 3 * 10
    return dict(workload="L=500 Mpc/h, 128^3 particles, 256^3 mesh, 10 steps, FP32, "
                         "density loss / P(k), gradient w.r.t. the white-noise modes",
                forward_ms=(t1 - t0) * 1e3, adjoint_ms=(t2 - t1) * 1e3,
                particle_steps_per_s=n / (t2 - t0), loss=L,
                grad_norm=float(torch.linalg.vector_norm(g.value).item()))


def run_large(args):
    """512^3 particles on a 1024^3 mesh, L = 2000 Mpc/h, TSC with window
    compensation, 2LPT at a = 0.05, 40 steps to a = 1: FP32 against FP64
    (plain) P(k) and displacements."""
    import copy

    la = copy.copy(args)
    la.nc, la.box, la.B, la.nsteps, la.warmup, la.steps = 512, 2000.0, 2, 40, 1, 4
    res = {}
    for dt, fused in (("f4", True), ("f8", False)):
        r = run_sim(la, dt, fused, la.steps, resampler="tsc", a0=0.05)
        res[dt] = r
    g = gates(res["f4"], res["f8"], la)
    out = dict(workload="L=2000 Mpc/h, 512^3 particles, 1024^3 mesh, TSC with compensation, "
                        "2LPT at a=0.05, 40 steps",
               fp32_ms_per_step=res["f4"]["elapsed"] / la.steps * 1e3,
               fp64_plain_ms_per_step=res["f8"]["elapsed"] / la.steps * 1e3,
               fp32_kernels_ms={k: {"calls": c, "mean_ms": m}
                                for k, (c, m) in res["f4"]["kern"].items()},
               accuracy=g)
    nm = res["f4"]["nmesh"]
    out["fp32_kernels_bandwidth"] = region_bandwidth(res["f4"]["kern"], res["f4"]["npart"],
                                                     nm ** 3, nm * nm * (nm // 2 + 1))
    return out


def cpu_baseline(args, r64):
    """One KDK step of the FP64 oracle at the same workload, from the GPU
    float64 state, on this host's cores."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import fastpm_oracle as O

    st = r64["state"]
    cores = os.cpu_count() or 1
    O.set_fft_workers(min(cores, 16))
    osolver = O.Solver(args.nc, args.box, 0.3, B=args.B, resampler="cic", deconv=2)
    ost = O.State(O.lattice((args.nc,) * 3, (args.box,) * 3))
    ost.S[...] = st.S.cpu().numpy()
    ost.P[...] = st.P.cpu().numpy()
    ost.F[...] = st.F.cpu().numpy()
    ost.a = dict(S=1.0, P=1.0, F=1.0)
    a0, a1 = 1.0, 1.0 + 1.0 / args.nsteps
    acts = list(O.leapfrog([a0, a1]))[1:]
    t0 = time.perf_counter()
    osolver.nbody(ost, acts)
    el = time.perf_counter() - t0
    return dict(value=len(ost.Q) / el, unit="particle-steps/s", cores=min(cores, 16), kind="port",
                sample="one KDK step (K, D, F, K) of the float64 numpy/scipy oracle at 256^3 "
                       "particles / 512^3 mesh, %.1f s wall" % el)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (requested --gpus %d)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        sys.exit(1)
    if args.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    print("bench: float32 fused run", file=sys.stderr, flush=True)
    r32 = run_sim(args, "f4", True, args.steps)
    ms32 = r32["elapsed"] / args.steps * 1e3
    Np = r32["npart"]
    Nc = r32["nmesh"] ** 3
    value = Np * args.steps / r32["elapsed"]
    out = {
        "metric": "particle-steps/s at 256³ ptcl/512³ mesh, FP32 vs FP64 GPU, P(k) rel err ≤1e-3",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms32, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "single box L=1000 Mpc/h, %d^3 particles, %d^3 mesh (B=%d), CIC "
                               "with window compensation, 2LPT at a=0.1, %d KDK steps to a=1"
                               % (args.nc, r32["nmesh"], args.B, args.nsteps),
                   "particles": Np, "mesh": Nc, "seed": args.seed},
    }
    # roofline of the dominant kernel (readout: one kernel per call)
    kern = r32["kern"]
    out["kernels_ms"] = {k: {"calls": c, "mean_ms": m} for k, (c, m) in kern.items()}
    out["kernels_bandwidth"] = region_bandwidth(
        kern, Np, Nc, r32["nmesh"] ** 2 * (r32["nmesh"] // 2 + 1))
    ro_calls, ro_ms = kern["readout"]
    # density and potential meshes read once, S in, RHO and F out
    ro_bytes = 8 * Nc + 28 * Np
    ach = ro_bytes / (ro_ms * 1e-3) / 1e9
    kname = readout_kernel_name(r32["solver"].fpm, 2)
    out["roofline"] = {"kernel": kname + " (RHO + four-point-gradient force, CIC)",
                       "bound": "hbm",
                       "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": ach / HBM_PEAK_GBS, "traffic": _profiled_traffic(kname),
                       "algorithmic_bytes": ro_bytes}
    Nk = r32["nmesh"] ** 2 * (r32["nmesh"] // 2 + 1)
    # compulsory bytes of this dataflow (kicks and drift of a step boundary
    # in one pass, one inverse transform of the potential, stencil in the
    # readout, the spectrum not stored without a monitor) and of the
    # reference's (separate kicks and drift, delta_k stored, three gradient
    # transforms, four-mesh readout)
    step_bytes = 20 * Nc + 100 * Np
    ref_bytes = 36 * Nc + 32 * Nk + 148 * Np
    out["step_roofline"] = {"algorithmic_bytes": step_bytes,
                            "achieved_GBs": step_bytes / (ms32 * 1e-3) / 1e9,
                            "frac": step_bytes / (ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                            "reference_dataflow_bytes": ref_bytes,
                            "reference_dataflow_GBs": ref_bytes / (ms32 * 1e-3) / 1e9}
    out["ic_seconds"] = r32["t_ic"]
    if not args.no_fp64:
        print("bench: float64 plain run", file=sys.stderr, flush=True)
        r64 = run_sim(args, "f8", False, args.steps)
        ms64 = r64["elapsed"] / args.steps * 1e3
        out["fp64_plain"] = {"ms_per_step": ms64, "value": Np * args.steps / r64["elapsed"],
                             "kernels_ms": {k: {"calls": c, "mean_ms": m}
                                            for k, (c, m) in r64["kern"].items()}}
        out["speedup_vs_fp64_plain"] = ms64 / ms32
        print("bench: accuracy gates", file=sys.stderr, flush=True)
        out["accuracy"] = gates(r32, r64, args)
        del r32
        torch.cuda.empty_cache()
        if not args.no_cpu:
            print("bench: CPU oracle baseline", file=sys.stderr, flush=True)
            out["cpu_baseline"] = cpu_baseline(args, r64)
        del r64
        torch.cuda.empty_cache()
    if not args.no_ensemble:
        print("bench: batched ensemble", file=sys.stderr, flush=True)
        out["ensemble"] = run_ensemble()
    if not args.no_grad:
        print("bench: differentiable model", file=sys.stderr, flush=True)
        out["gradient"] = run_gradient()
    if not args.no_large:
        print("bench: large box", file=sys.stderr, flush=True)
        out["large_box"] = run_large(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
