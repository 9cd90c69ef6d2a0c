#!/usr/bin/env python3
"""Beam-splitter site-update throughput on one MI355X: grouped FP32 vs plain FP64.

Workload: lossy boson-sampling MPO, M=64 modes, N=12 single photons in modes
0,5,...,55, d=13, chi=1024, eta=0.6; Haar unitaries (seeds 100, 101, ...)
as Clements meshes of 2016 beam splitters each.

A "step" is one brick-wall layer.  The first --warmup layers run untimed
(they bring the bond dimension up to chi); the next --steps layers are
timed, for two paths on the same inputs:
  plain    one instance, complex128 contraction, every update launched on its
           own (the reference's one-update-at-a-time structure)
  grouped  --instances independent instances; every update of a layer across
           all of them in one grouped run (one launch per stage), contraction
           in --precision (default complex64 on FP32 MFMA), SVD / truncation /
           BS elements / norm in FP64
`value` is the grouped path's BS updates/s over its timed layers; the plain
path's rate and the ratio are in `config`.

Also on the JSON line:
  roofline      the grouped run's dominant kernel: algorithmic work per launch
                / its average HIP-event time over the timed region
  cpu_baseline  the FP64 CPU oracle timed on a bounded sample of the same
                updates (first updates of the first timed layer)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "BS updates/s + % MFMA/HBM peak, M=64 χ=1024: FP32-grouped vs plain FP64"
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0}   # MI355X dense
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2, help="timed layers")
    ap.add_argument("--warmup", type=int, default=16, help="untimed layers before the timed ones")
    ap.add_argument("--modes", type=int, default=64)
    ap.add_argument("--photons", type=int, default=12)
    ap.add_argument("--stride", type=int, default=5)
    ap.add_argument("--d", type=int, default=13)
    ap.add_argument("--chi", type=int, default=1024)
    ap.add_argument("--eta", type=float, default=0.6)
    ap.add_argument("--seed", type=int, default=100)
    ap.add_argument("--instances", type=int, default=32)
    ap.add_argument("--precision", default="c64", choices=["c128", "c64", "bf16x3"])
    ap.add_argument("--paths", default="plain,grouped")
    ap.add_argument("--group-gib", type=float, default=24.0, help="workspace bound of one group")
    ap.add_argument("--streams", type=int, default=2, choices=[1, 2, 3, 4],
                    help="launching streams the groups of a layer alternate between")
    ap.add_argument("--first-frac", type=float, default=0.25,
                    help="workspace share of a layer's first group (planned while the device idles)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="budget of the CPU-oracle baseline sample (0 = skip)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="MiB read and written by one elementwise kernel run first (PMC byte calibration)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="time the eigensolver route split on the last warmup layer (grouped path)")
    ap.add_argument("--eig-threshold", type=int, default=-1,
                    help="fixed eigensolver route split for both paths instead of the autotuned one")
    ap.add_argument("--verbose", action="store_true")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stage_table(stages):
    for s, (ms, fl, by, cnt) in stages.items():
        if cnt:
            log(f"[bench]   {s:13s} {ms:10.3f} ms {cnt:6d} launches "
                f"{fl / max(ms, 1e-9) / 1e9:8.3f} TFLOP/s {by / max(ms, 1e-9) / 1e6:9.2f} GB/s")


HBM_BOUND = ("topk", "scatter", "observables", "bs_table", "tridiag")


def measured_traffic(kernel, workload, instances, precision, eig_threshold, launches):
    """HBM bytes per launch of `kernel` measured on this workload by two
    rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE, scaled by a probe kernel
    of known size) and recorded in profiles/<kernel>_*_traffic.json; None if
    no record matches (the same grouping: a per-launch figure holds only for
    the same number of launches)."""
    import glob
    for path in sorted(glob.glob(os.path.join(HERE, "profiles", f"{kernel}_*_traffic.json"))):
        with open(path) as f:
            rec = json.load(f)
        if (rec.get("kernel"), rec.get("workload"), rec.get("instances"), rec.get("precision"),
                rec.get("eig_threshold"), rec.get("launches")) == (kernel, workload, instances,
                                                                   precision, eig_threshold, launches):
            return rec["traffic_bytes_per_launch"], os.path.relpath(path, HERE)
    return None, None


def roofline(stages, precision):
    dom = max((s for s in stages if stages[s][3] > 0), key=lambda s: stages[s][0])
    ms, fl, by, cnt = stages[dom]
    t = ms / cnt / 1e3
    if dom in HBM_BOUND or fl == 0.0:
        ach = by / cnt / t / 1e9 if by else None
        roof = {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": (ach / PEAK_HBM_GBS) if ach else None, "traffic": None}
    else:
        peak = PEAK_TFLOPS["f64"]
        if dom == "contract" and precision != "c128":
            peak = PEAK_TFLOPS["f32"] if precision == "c64" else PEAK_TFLOPS["bf16"]
        ach = fl / cnt / t / 1e12
        roof = {"bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s",
                "frac": ach / peak, "traffic": None}
    roof.update({"kernel": dom, "launches": cnt, "avg_launch_ms": ms / cnt,
                 "stage_ms": {s: round(v[0], 3) for s, v in stages.items() if v[3]}})
    return roof


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the multi-GPU sharding is a "
                                   "design note only -- run with --gpus 1"}))
        sys.exit(2)
    import torch
    from bosonsampling_amd import _capi, engine, mesh

    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a HIP device")
    prec = {"c128": engine.C128, "c64": engine.C64, "bf16x3": engine.BF16X3}[args.precision]
    if args.copy_probe:
        # one elementwise kernel (torch's scalar multiply) that streams n bytes in and n out
        n = args.copy_probe << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: {n} bytes read + {n} bytes written")
        del src, dst
    M, N, d, chi = args.modes, args.photons, args.d, args.chi
    inputs = list(range(0, N * args.stride, args.stride))
    W, K = args.warmup, args.steps
    if W + K > M:
        raise SystemExit("warmup + steps exceeds the number of layers")
    paths = args.paths.split(",")
    B = args.instances
    meshes = [mesh.haar_clements_mesh(M, args.seed + b)[0] for b in range(B)]
    t0 = time.time()
    init = engine.initialize(M, d, chi, inputs, args.eta)
    torch.cuda.synchronize()
    log(f"[bench] init {time.time() - t0:.2f}s")
    max_input = max(inputs)
    result = {}

    # ---------------------------------------------------------------- plain
    snap = None
    if "plain" in paths:
        st = init.clone()

        def run_layer(layer, precision):
            todo = engine.layer_modes(M, layer, max_input)
            if not todo:
                return 0
            U3 = engine.bs_tables(d, np.stack([meshes[0][layer, i] for i, _ in todo]), st.device)
            for k, (i, mode) in enumerate(todo):
                st.update(mode, U3[k], precision)
            return len(todo)

        t1 = time.time()
        plain_tuned = fixed_threshold(args)
        for layer in range(W):
            if args.autotune and args.eig_threshold < 0 and layer == W - 1:
                best, times = engine.autotune_eig_threshold([st], [meshes[0]], inputs, layer,
                                                            precision=engine.C128, grouped=False)
                plain_tuned = {"layer": layer, "eig_threshold": best,
                               "seconds": {str(k): round(v, 3) for k, v in times.items()}}
                log(f"[bench] plain path: autotuned eigensolver split on layer {layer}: "
                    f"threshold {best} {times}")
            run_layer(layer, engine.C128)
        torch.cuda.synchronize()
        log(f"[bench] plain warmup {W} layers {time.time() - t1:.1f}s")
        if args.cpu_seconds > 0:
            snap = st.to_host()
        _capi.timing_enable(True)
        _capi.timing_read(reset=True)
        torch.cuda.synchronize()
        ts = time.time()
        n = sum(run_layer(layer, engine.C128) for layer in range(W, W + K))
        torch.cuda.synchronize()
        el = time.time() - ts
        st_plain = _capi.timing_read(reset=True)
        _capi.timing_enable(False)
        result["plain"] = {"updates": n, "seconds": el, "rate": n / el,
                           "observables": st.observables(), "tuned": plain_tuned}
        log(f"[bench] plain c128: {n} updates in {el:.3f}s = {n / el:.2f} updates/s")
        stage_table(st_plain)
        del st
        torch.cuda.empty_cache()

    # -------------------------------------------------------------- grouped
    roof = None
    if "grouped" in paths:
        states = [init.clone() for _ in range(B)]
        runner = engine.GroupRunner(init.device, max_bytes=int(args.group_gib * (1 << 30)),
                                    streams=args.streams, first_frac=args.first_frac)
        t1 = time.time()
        tuned = fixed_threshold(args)
        for layer in range(W):
            if args.autotune and args.eig_threshold < 0 and layer == W - 1:
                # tune the eigensolver route split on the last warmup layer (clones)
                best, times = engine.autotune_eig_threshold(states, meshes, inputs, layer,
                                                            precision=prec, runner=runner)
                tuned = {"layer": layer, "eig_threshold": best,
                         "seconds": {str(k): round(v, 3) for k, v in times.items()}}
                log(f"[bench] autotuned eigensolver split on layer {layer}: threshold {best} "
                    f"{times}")
            engine.simulate_batch(states, meshes, inputs, layers=[layer], runner=runner,
                                  precision=prec)
            torch.cuda.synchronize()
            if args.verbose:
                log(f"[bench] grouped warmup layer {layer}: {time.time() - t1:.1f}s "
                    f"groups {runner.last_info.get('groups')}")
        log(f"[bench] grouped warmup {W} layers x {B} instances {time.time() - t1:.1f}s")
        _capi.timing_enable(True)
        _capi.timing_read(reset=True)
        torch.cuda.synchronize()
        ts = time.time()
        n = 0
        for layer in range(W, W + K):
            k, _ = engine.simulate_batch(states, meshes, inputs, layers=[layer], runner=runner,
                                         precision=prec)
            n += k
            log(f"[bench] grouped timed layer {layer}: {n} updates {time.time() - ts:.3f}s")
        torch.cuda.synchronize()
        el = time.time() - ts
        st_grp = _capi.timing_read(reset=True)
        _capi.timing_enable(False)
        result["grouped"] = {"updates": n, "seconds": el, "rate": n / el,
                             "observables0": states[0].observables(), "tuned": tuned}
        log(f"[bench] grouped {args.precision}: {n} updates in {el:.3f}s = {n / el:.2f} updates/s")
        stage_table(st_grp)
        roof = roofline(st_grp, args.precision)
        roof["launch_streams"] = args.streams
        if args.streams > 1:
            roof["note"] = (f"groups go round {args.streams} launching streams: a launch's duration "
                            "includes sharing the GPU with the other streams' stages")

    cpu = cpu_baseline(snap, d, chi, meshes[0], inputs, W, args.cpu_seconds) if snap else None
    workload = (f"lossy boson-sampling MPO site updates, M={M} N={N} "
                f"(modes 0,{args.stride},..) d={d} chi={chi} eta={args.eta}; "
                f"layers {W}..{W + K - 1} timed")
    g = result.get("grouped")
    p = result.get("plain")
    if roof is not None:
        thr = (g or {}).get("tuned") or {}
        roof["traffic"], src = measured_traffic(roof["kernel"], workload, B, args.precision,
                                                thr.get("eig_threshold", 96), roof["launches"])
        if src:
            roof["traffic_source"] = src
    main_path = g or p
    out = {
        "metric": METRIC,
        "value": main_path["rate"],
        "unit": "BS updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * main_path["seconds"] / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": (f"{args.precision} contraction, f64 SVD/truncation" if g else "complex128 (f64)"),
        "data": "synthetic: Haar unitaries (seeds %d..%d) as Clements meshes"
                % (args.seed, args.seed + (B - 1 if g else 0)),
        "config": {
            "workload": workload,
            "grouped": ({"instances": B, "precision": args.precision, "updates": g["updates"],
                         "updates_per_s": g["rate"], "autotune": g.get("tuned"),
                         "launch_streams": args.streams} if g else None),
            "plain_fp64": ({"instances": 1, "updates": p["updates"],
                            "updates_per_s": p["rate"], "autotune": p.get("tuned")} if p else None),
            "grouped_over_plain": (g["rate"] / p["rate"]) if (g and p) else None,
            "precision_check": precision_check(g, p, args.precision, W + K),
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def fixed_threshold(args):
    """Set the eigensolver route split given by --eig-threshold (recorded like
    an autotuned one, without timings); None when it is to be autotuned."""
    if args.eig_threshold < 0:
        return None
    from bosonsampling_amd import _capi
    _capi.check(_capi.lib().bsmpo_set_eig_threshold(int(args.eig_threshold)), "set_eig_threshold")
    return {"eig_threshold": int(args.eig_threshold), "fixed": True}


def precision_check(g, p, precision, layers):
    """Instance 0 of the grouped path (every layer in `precision`) against the
    plain complex128 path on the same mesh: output photon-number marginals
    P(n_j = n) and trace after the same layers."""
    if not (g and p):
        return None
    tg, mg = g["observables0"]
    tp, mp = p["observables"]
    return {"precision": precision, "layers": layers, "reference": "plain complex128, same mesh",
            "marginal_max_abs_err": float(np.abs(mg - mp).max()),
            "trace_abs_err": float(abs(tg - tp)), "trace": float(tp)}


def cpu_baseline(snap, d, chi, mm, inputs, layer, budget_s):
    """Time oracle.mpo.apply_update (numpy FP64) on the first updates of
    `layer` from the plain path's GPU snapshot, until ~budget_s of CPU work."""
    from oracle import fock, mpo
    from bosonsampling_amd.engine import layer_modes
    G, L, ch = snap
    st = mpo.MPOState(d, chi, G.copy(), L.copy(), ch.copy(), np.ones(chi))
    todo = layer_modes(G.shape[0], layer, max(inputs))
    n, used = 0, 0.0
    for i, mode in todo:
        U3 = fock.bs_table(d, mm[layer, i])
        t = time.time()
        mpo.apply_update(st, mode, U3)
        used += time.time() - t
        n += 1
        log(f"[bench] cpu oracle update {n} (mode {mode}): {used:.2f}s")
        if used >= budget_s:
            break
    threads = os.environ.get("OPENBLAS_NUM_THREADS") or os.environ.get("OMP_NUM_THREADS")
    cores = int(threads) if threads else len(os.sched_getaffinity(0))
    return {"value": n / used if used > 0 else None, "unit": "BS updates/s", "cores": cores,
            "kind": "port",
            "sample": f"{n} FP64 oracle updates (numpy + LAPACK) of layer {layer}, from the "
                      f"plain path's GPU snapshot, {used:.1f}s"}


if __name__ == "__main__":
    main()
