"""Benchmark of the MI355X tracer-transport step and SSH solve.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c4]

The default workload is the 1.3 M-node, 70-level mesh with 10 tracers. One
"step" is one solve_tracers_ale pass over every tracer of the workload (AB2,
MFCT/QR4C/FCT advection, implicit vertical diffusion), replayed from a
captured hipGraph with all fields resident in HBM. `value` is tracer-updates/s
(wet 3-D nodes x tracers x steps / wall time of the K timed steps). The SSH
solve (diagonally preconditioned CG to 1e-10 relative residual, x0 = 0) is
timed separately and reported as ssh_solve_ms; beside it, the learned
preconditioner in flexible CG, trained on this workload's operator from the
model's own right-hand sides, against plain CG on a held-out right-hand side
(ssh_learned).

Also reported: the per-kernel profile of the step (hipEvents on the library
stream), the roofline of the dominant kernel, and the oracle's OpenMP build
timed on a bounded sample of the same workload on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def kernel_bytes(name, mesh, ntr):
    """Algorithmic bytes of one launch of a step kernel (all tracers of the batch).

    Every array a kernel must read or write is counted once over the entries it
    covers: wet layers N3, wet interfaces NI = N3 + N2, edge levels E3 (the
    nu12..nl12 ranges), element levels F3; FP64 = 8 B, int32 = 4 B. Values of
    neighbouring columns that a gather reads again are not counted twice.
    Per-tracer terms scale with ntr; mesh and dynamics terms do not.
    """
    N2, E2, F2, L = mesh.nod2D, mesh.edge2D, mesh.elem2D, mesh.nl - 1
    N3, E3, F3 = mesh.n3, mesh.e3, mesh.f3
    NI = N3 + N2
    lists = 4 * (N2 + 1) + 8 * E2           # node -> edge gather lists
    if name == "kA_edge":                    # T^n, T^n-1 ; LO_h, A_h | uv, helem, edge/elem geometry
        return ntr * (16 * N3 + 16 * E3) + 24 * F3 + 88 * E2 + 76 * F2 + 24 * N2
    if name == "kB_node":                    # LO_h, T^n, T^n-1 ; LO, A_v | w, w_e, area, zbar, Z, h, h_new, areasvol
        return ntr * (8 * E3 + 16 * N3 + 8 * N3 + 8 * NI) + 32 * NI + 32 * N3 + lists + 8 * E2 + 8 * N2
    if name == "kC_node":                    # LO, T^n, A_v, A_h ; R+, R- | h_new, areasvol, edge info
        return ntr * (16 * N3 + 8 * NI + 8 * E3 + 16 * N3) + 16 * N3 + lists + 32 * E2 + 8 * N2
    if name == "kD_node":                    # R+-, A_v, A_h, LO, T^n ; T^n+1, T^n -> valuesold | h, h_new, areasvol, area, Kv
        return ntr * (16 * N3 + 8 * NI + 8 * E3 + 16 * N3 + 16 * N3) + 32 * N3 + 8 * NI + lists + 16 * E2 + 16 * N2
    # staged per-stage kernels (FGPU_STEP=staged)
    if name == "k_init_ab":
        return ntr * 32 * N2 * L
    if name == "k_tr_xy":
        return ntr * (8 * N3 + 16 * F3) + 68 * F2
    if name == "k_edge_flux":
        return ntr * (16 * N3 + 16 * F3 + 16 * E3) + 24 * F3 + 88 * E2 + 8 * F2
    if name == "k_node_lo":
        return ntr * (8 * E3 + 16 * N3 + 16 * N3) + 16 * NI + 32 * N3 + 16 * NI + lists + 8 * E2
    if name == "k_elem_aux":
        return ntr * (16 * N3 + 16 * F2 * L) + 12 * F2
    if name == "k_node_limit":
        return ntr * (16 * F3 + 8 * N3 + 8 * NI + 8 * E3 + 16 * N3) + 16 * N3 + lists + 8 * E2
    if name == "k_edge_limit":
        return ntr * (16 * E3 + 16 * N3) + 16 * E2
    if name == "k_node_update":
        return ntr * (8 * NI + 16 * N3 + 8 * E3 + 16 * N3 + 8 * N3) + 24 * N3 + lists + 8 * E2
    if name == "k_vdiff":
        return ntr * 16 * N3 + 8 * (N3 + 2 * NI) + 8 * NI + 8 * N2
    # the steps either side of the transport (tracer independent)
    if name == "k_vert_vel":                 # uv, helem, area, hnode_new ; W, W_e, W_i, CFL_z | edge geometry
        return 24 * F3 + 8 * NI + 8 * N3 + 32 * NI + lists + 32 * 2 * E2 + 40 * E2
    if name == "k_ssh_rhs_edge":             # uv, uv_rhs, helem ; edge sums | edge geometry, element levels
        return 40 * F3 + 40 * E2 + 8 * F2 + 8 * E2
    if name == "k_ssh_rhs_node":             # edge sums, ssh_rhs_old ; ssh_rhs | lists
        return 8 * E2 + lists + 8 * N2 + 8 * N2
    return None


def kernel_source_digest():
    """Digest of the fused step's kernel source: a committed traffic
    measurement holds only for the kernels it was measured on."""
    import hashlib
    h = hashlib.sha256()
    for f in ("fgpu_step.hip", "fgpu_vert.h", "fgpu_internal.h"):
        with open(os.path.join(ROOT, "fesom2_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def measured_traffic(config, kernel):
    """HBM bytes per launch of `kernel` from profiles/traffic_<config>.txt
    (rocprofv3 FETCH_SIZE/WRITE_SIZE, scaled by the copy-kernel calibration),
    or None when the file was measured on other kernel sources than these."""
    tfile = os.path.join(ROOT, "profiles", f"traffic_{config}.txt")
    if not os.path.exists(tfile):
        return None, "no measurement for this workload"
    digest, traffic = None, None
    for line in open(tfile):
        parts = line.split()
        if line.startswith("# source digest") and parts:
            digest = parts[-1]
        elif len(parts) >= 2 and parts[0] == kernel:
            traffic = float(parts[1])
    if digest != kernel_source_digest():
        return None, f"{os.path.relpath(tfile, ROOT)} was measured on other kernel sources"
    return traffic, os.path.relpath(tfile, ROOT)


def step_model_bytes(mesh, ntr):
    """Whole-step algorithmic bytes of the transport model: per tracer 88 N3 + 16 E3
    (state in/out, LO, A_h, A_v, R+-), shared 24 F3 + 72 N3 + mesh terms."""
    N2, E2, F2 = mesh.nod2D, mesh.edge2D, mesh.elem2D
    N3, E3, F3 = mesh.n3, mesh.e3, mesh.f3
    return ntr * (88 * N3 + 16 * E3) + 24 * F3 + 72 * N3 + 72 * E2 + 80 * F2 + 48 * N2


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(mesh, state, budget_s=15.0):
    """The oracle's OpenMP build (C port of the reference path, columns, edges
    and elements in parallel, bit-identical to the serial oracle) on whole
    steps of the same mesh and state, on OMP_NUM_THREADS host threads (16, the
    GPU box's CPU share, when unset). The sample keeps as many of the tracers
    as about budget_s allows at an assumed 2e6 tracer-updates/s per thread (at
    least T and S) and runs whole steps until budget_s has passed."""
    threads = int(os.environ.get("OMP_NUM_THREADS", "16"))
    os.environ["OMP_NUM_THREADS"] = str(threads)   # read when the OpenMP runtime starts
    from oracle.oracle import Oracle
    orc = Oracle(mesh, omp=True)
    k = max(min(2, state.ntr), min(state.ntr, int(budget_s * 2.0e6 * threads / max(mesh.n3, 1))))
    v = state.values[:k].copy()
    vo = state.valuesold[:k].copy()
    t0 = time.perf_counter()
    n = 0
    while True:
        orc.step(state, v, vo)
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s or n >= 50:
            break
    upd = k * mesh.n3 * n / el
    which = f"all {state.ntr} tracers" if k == state.ntr else f"{k} of the {state.ntr} tracers"
    return {"value": upd, "unit": "tracer-updates/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"{n} full steps of the same mesh and state ({which}), {el:.1f} s, the oracle's "
                      f"OpenMP build (gcc -O3, no FMA contraction, bit-identical to the serial oracle), "
                      f"{threads} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c4", choices=["c0", "c1", "c2", "c4"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ssh-reps", type=int, default=5)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-preconditioner SSH solve")
    ap.add_argument("--learn-steps", type=int, default=200,
                    help="plain-solver steps whose right-hand sides (every 4th) train the network")
    ap.add_argument("--learn-epochs", type=int, default=20)
    ap.add_argument("--shape", default=None,
                    help="TPB,TCH,CPB[,KD[,KC[,KA[,KB]]]]: fixed launch shape of the fused step (variant of its "
                         "last kernel: 0 lock-step, 1 sweep wave, 2 one wave per column; tracers per block of the "
                         "(tracer, level) bound kernel, 0 for the per-column one; levels per lane of the "
                         "edge-flux and of the LO kernel, 1 or 2) instead of the autotuner's (a profiled run then "
                         "holds only that shape's launches)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the multi-GPU path is design-only "
                                   "(halo exchange over RCCL is not built)", "n_gpus": args.gpus}))
        sys.exit(2)

    if args.shape:
        sv = [int(x) for x in args.shape.split(",")]
        os.environ["FGPU_STEP_TPB"], os.environ["FGPU_STEP_TCH"], os.environ["FGPU_STEP_CPB"] = (
            str(sv[0]), str(sv[1]), str(sv[2]))
        os.environ["FGPU_STEP_KD"] = ["lockstep", "sweepwave", "wavelocal"][sv[3] if len(sv) > 3 else 0]
        os.environ["FGPU_STEP_KC"] = str(sv[4] if len(sv) > 4 else 0)
        os.environ["FGPU_STEP_KA"] = str(sv[5] if len(sv) > 5 else 1)
        os.environ["FGPU_STEP_KB"] = str(sv[6] if len(sv) > 6 else 1)
    from fesom2_amd import gpu, meshgen as mg
    t0 = time.perf_counter()
    mesh, state, _ = mg.make_config(args.config)
    t_gen = time.perf_counter() - t0
    print(f"[bench] {args.config}: N2={mesh.nod2D} E2={mesh.edge2D} F2={mesh.elem2D} nl={mesh.nl} "
          f"N3={mesh.n3} ntr={state.ntr} (generated in {t_gen:.1f}s)", file=sys.stderr, flush=True)
    ctx = gpu.context_for(mesh, state)
    dt = state.dt
    if args.shape:
        shape = ctx.step_shape()
        print(f"[bench] fixed shape: {shape[0]} tracers per block, {shape[1]} per Thomas chunk, {shape[2]} "
              f"columns per block", file=sys.stderr, flush=True)
    else:
        tpb, tune_ms = ctx.autotune_step(dt, 3)
        shape = ctx.step_shape()
        print(f"[bench] autotuner: {shape[0]} tracers per block, {shape[1]} per Thomas chunk, {shape[2]} columns "
              f"per block, last kernel variant {ctx.step_variants()} ({tune_ms:.3f} ms/step)", file=sys.stderr,
              flush=True)
    ctx_variant, ctx_kc, ctx_ka, ctx_kb = ctx.step_variants()
    ctx.solve_tracers_ale(dt, max(1, args.warmup))
    ctx.synchronize()
    t0 = time.perf_counter()
    ms_dev = ctx.solve_tracers_ale_timed(dt, args.steps)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    updates = state.ntr * mesh.n3 * args.steps
    value = updates / wall
    print(f"[bench] {args.steps} steps: wall {wall*1e3:.2f} ms, device {ms_dev:.2f} ms", file=sys.stderr, flush=True)

    # per-kernel profile of a few more steps
    ctx.profile_enable(True)
    nprof = min(args.steps, 5)
    ctx.solve_tracers_ale(dt, nprof)
    prof = ctx.profile()
    ctx.profile_enable(False)
    kernels = {}
    for name, (ms, cnt) in prof.items():
        b = kernel_bytes(name, mesh, state.ntr)
        avg = ms / max(cnt, 1)
        kernels[name] = {"avg_ms": round(avg, 4), "launches": cnt,
                         "alg_GBs": round(b / (avg * 1e-3) / 1e9, 1) if b else None}
    dom = max(prof.items(), key=lambda kv: kv[1][0])[0]
    dom_avg_ms = prof[dom][0] / prof[dom][1]
    dom_bytes = kernel_bytes(dom, mesh, state.ntr)
    achieved = dom_bytes / (dom_avg_ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic(args.config, dom)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": traffic_src, "alg_bytes_per_launch": int(dom_bytes)}
    step_bytes = sum(kernel_bytes(n, mesh, state.ntr) or 0 for n in prof)
    step_ms = sum(ms for ms, c in prof.values()) / nprof

    # the steps either side of the transport: vert_vel_ale and compute_ssh_rhs_ale
    # (device-resident uv; uv_rhs and ssh_rhs_old are uploaded once before)
    rng = np.random.default_rng(7)
    uv_rhs = rng.normal(0.0, 0.01, state.uv.shape)
    rhs_old = np.zeros(mesh.nod2D)
    ctx.set_uv(state.uv)
    ctx.vert_vel_ale(dt)
    ctx.compute_ssh_rhs_ale(uv_rhs, rhs_old)
    ctx.profile_enable(True)
    for _ in range(nprof):
        ctx.vert_vel_ale(dt)
        ctx.compute_ssh_rhs_ale(uv_rhs, rhs_old)
    prof_dyn = ctx.profile()
    ctx.profile_enable(False)
    dyn_kernels = {}
    for name, (ms, cnt) in prof_dyn.items():
        b = kernel_bytes(name, mesh, state.ntr)
        avg = ms / max(cnt, 1)
        dyn_kernels[name] = {"avg_ms": round(avg, 4), "launches": cnt,
                             "alg_GBs": round(b / (avg * 1e-3) / 1e9, 1) if b else None}
    print(f"[bench] vert_vel_ale / compute_ssh_rhs_ale: {dyn_kernels}", file=sys.stderr, flush=True)

    # SSH solve
    ctx.init_stiff_mat_ale(dt)
    ctx.ssh_solve_preconditioner(2)
    rhs = mg.ssh_rhs(mesh)
    ctx.ssh_upload(np.zeros(mesh.nod2D), rhs)
    ctx.ssh_solve_resident(1e-10, 5000, 1)
    sst = ctx.ssh_solve_resident(1e-10, 5000, args.ssh_reps)
    print(f"[bench] ssh: {sst}", file=sys.stderr, flush=True)

    # learned preconditioner in flexible CG beside plain diagonal CG, on a
    # held-out right-hand side of the model's own steps (another seed's run)
    learned = None
    if not args.no_learned:
        from fesom2_amd import ssh_learned as sl
        t0 = time.perf_counter()
        rb, re = sl.model_training_set(ctx, mesh, state, nsteps=args.learn_steps, every=4, seed=1)
        print(f"[bench] learned SSH: {rb.shape[0]} training pairs in {time.perf_counter() - t0:.1f}s",
              file=sys.stderr, flush=True)
        ctx.nn_init(0)
        losses = []
        for e0 in range(0, args.learn_epochs, 5):   # the device's Adam state carries over between calls
            ne = min(5, args.learn_epochs - e0)
            losses.extend(ctx.nn_train(rb, re, epochs=ne, batch=8, lr=2e-3))
            print(f"[bench] learned SSH: epoch {e0 + ne}/{args.learn_epochs}, loss {losses[-1]:.4f}",
                  file=sys.stderr, flush=True)
        t_train = time.perf_counter() - t0
        hb, _ = sl.model_rhs_sequence(ctx, mesh, state, nsteps=20, every=20, seed=7)
        ctx.ssh_solve_preconditioner(2)
        ctx.ssh_upload(np.zeros(mesh.nod2D), hb[0])
        ctx.ssh_solve_resident(1e-10, 5000, 1)
        pst = ctx.ssh_solve_resident(1e-10, 5000, args.ssh_reps)
        ctx.ssh_solve_learned_resident(1e-10, 500, 5000, 1)
        lst = ctx.ssh_solve_learned_resident(1e-10, 500, 5000, args.ssh_reps)
        learned = {"rhs": "held-out step of a model run (compute_ssh_rhs_ale), x0 = 0",
                   "plain_cg_ms": round(pst["ms"], 3), "plain_cg_iters": pst["iters"],
                   "learned_fcg_ms": round(lst["ms"], 3), "learned_fcg_iters": lst["iters"],
                   "learned_fallback_iters": lst["fallback_iters"], "learned_converged": lst["converged"],
                   "speedup_vs_plain": round(pst["ms"] / lst["ms"], 3),
                   "training": f"{rb.shape[0]} residual/error pairs from {args.learn_steps} plain-solver steps, "
                               f"{args.learn_epochs} epochs, loss {losses[0]:.3f} -> {losses[-1]:.3f}, "
                               f"{t_train:.1f} s with data"}
        print(f"[bench] learned SSH: {learned}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(mesh, state)
        print(f"[bench] cpu baseline: {cpu}", file=sys.stderr, flush=True)
    ctx.close()

    out = {
        "metric": "tracer-updates/s (3D node x tracer) + SSH solve ms at 1e-10 rel residual",
        "value": value, "unit": "tracer-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded mesh, T/S, velocity)",
        "config": {"workload": f"{args.config}: {mesh.nod2D}-node jittered band mesh, {mesh.nl-1} levels, "
                               f"{state.ntr} tracers, MFCT/QR4C/FCT + implicit vertical diffusion, dt={dt:g}s",
                   "nod2D": mesh.nod2D, "elem2D": mesh.elem2D, "edge2D": mesh.edge2D, "levels": mesh.nl - 1,
                   "n3": mesh.n3, "tracers": state.ntr, "parallelism": "single GPU"},
        "ssh_solve_ms": round(sst["ms"], 3), "ssh_iters": sst["iters"], "ssh_converged": sst["converged"],
        "ssh_precond": "diagonal (Jacobi)",
        "ssh_learned": learned,
        "device_ms_per_step": ms_dev / args.steps,
        "autotuned_shape": {"tracers_per_block": shape[0], "thomas_chunk": shape[1], "columns_per_block": shape[2],
                            "last_kernel": ["lock-step", "sweep wave", "one wave per column"][ctx_variant],
                            "bound_kernel_tracers_per_block": ctx_kc,
                            "edge_kernel_levels_per_lane": ctx_ka, "lo_kernel_levels_per_lane": ctx_kb},
        "step_alg_GBs": round(step_bytes / (step_ms * 1e-3) / 1e9, 1),
        "step_model_bytes": int(step_model_bytes(mesh, state.ntr)),
        "step_model_frac_of_hbm": round(step_model_bytes(mesh, state.ntr) / (wall / args.steps) / 1e9 / HBM_PEAK_GBS, 4),
        "kernels": kernels,
        "dynamics_kernels": dyn_kernels,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
