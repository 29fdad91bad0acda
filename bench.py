"""Benchmark of the MI355X time loop: GLL-point x mode updates per second.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload N] [--no-cpu]

One step = one pass of Newmark::solve's loop body (Newmark.cpp:47-93) over the
whole synthetic mesh resident in HBM: Newmark update, source, element stiffness
(with on-device assembly), solid-fluid coupling and station recording.  The
timed region is bracketed by a device synchronisation and timed with HIP events
on the domain's stream.  Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GLL-pt·mode updates/s + wall-time to 1% misfit, learned-Nu vs full-Nu, 1 GPU"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    0: "shell-4000q-solid-iso1d-nu4",
    1: "shell-20000q-solid+fluid-ti3d-nu8..128",
}


def cpu_baseline(mesh, seconds=20.0):
    """Time the FP32 oracle (a restatement of the reference's CPU loop) on a
    bounded number of steps of the same workload, with all host threads."""
    from oracle.refcpu import RefDomain
    threads = min(16, os.cpu_count() or 1)
    ref = RefDomain(mesh, double=False, threads=threads)
    # probe one step, then size the sample to ~`seconds`
    t0 = time.perf_counter()
    ref.run(1, 1, 0)
    one = time.perf_counter() - t0
    nsample = int(max(1, min(mesh.d.nstep - 1, seconds / max(one, 1e-6))))
    t0 = time.perf_counter()
    ref.run(2, nsample, 0)
    el = time.perf_counter() - t0
    upd = mesh.updates_per_step()
    return {"value": upd * nsample / el, "unit": "GLL-pt·mode updates/s", "cores": threads, "kind": "port",
            "sample": "%d steps of the same mesh (steps 2..%d), FP32 oracle, %d threads, %.1f s"
                      % (nsample, nsample + 1, threads, el)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the whole mesh lives in one GPU's HBM); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import axisem3d_amd as ax

    which = args.workload
    t0 = time.perf_counter()
    mesh = ax.ShellMesh(which=which)
    t_mesh = time.perf_counter() - t0
    nstep_needed = args.warmup + args.steps
    if nstep_needed > mesh.d.nstep:
        raise SystemExit("bench.py: warmup + steps = %d exceeds the %d-step source-time function"
                         % (nstep_needed, mesh.d.nstep))
    dom = ax.Domain(mesh)
    bytes_step, upd_step = dom.step_cost()
    dom.resetZero()
    if args.warmup:
        dom.run(1, args.warmup, 0, use_graph=True)
    # timed region: HIP events on the domain's stream around graph replays of
    # the step; the library synchronises the stream before and after
    dom.run(args.warmup + 1, args.steps, 0, use_graph=True)
    ms = dom.last_run_ms()
    # kernel-class timing (same steps again, profiled) for the roofline figure
    dom.resetZero()
    if args.warmup:
        dom.run(1, args.warmup, 0)
    prof = dom.run_profiled(args.warmup + 1, args.steps)
    el_ms = prof["element_ms"] / prof["steps"]
    nm_ms = prof["newmark_ms"] / prof["steps"]
    by_el, by_nm = dom.kernel_bytes()
    cand = {"element": (el_ms, by_el), "newmark": (nm_ms, by_nm)}
    dom_name = max(cand, key=lambda k: cand[k][0])
    kms, kby = cand[dom_name]
    achieved = kby / (kms * 1e-3) / 1e9
    value = upd_step * args.steps / (ms * 1e-3)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "GLL-pt·mode updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": WORKLOADS.get(which, str(which)),
            "quads": mesh.d.nelem,
            "gll_points": mesh.d.npoint,
            "updates_per_step": upd_step,
            "algorithmic_bytes_per_step": bytes_step,
            "dt_s": mesh.d.dt,
            "parallelism": "1 GPU, whole mesh in HBM",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel_ms_per_step": kms,
            "algorithmic_bytes_per_launch": kby,
            "step_bytes_frac": bytes_step / (ms * 1e-3 / args.steps) / 1e9 / HBM_PEAK_GBS,
        },
        "mesh_build_s": t_mesh,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mesh, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
