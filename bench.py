#!/usr/bin/env python3
"""Beam-splitter site-update throughput on one MI355X: grouped FP32 vs plain FP64.

Workload: lossy boson-sampling MPO, M=64 modes, N=12 single photons in modes
0,5,...,55, d=13, chi=1024, eta=0.6; Haar unitaries (seeds 100, 101, ...)
as Clements meshes of 2016 beam splitters each.

A "step" is one brick-wall layer.  The timed layers start at layer
max(--start, --warmup) (default 16: by then the bond dimension has reached
chi on the inner bonds); the --warmup layers before them run untimed, and so
do the layers before those (state preparation).  --steps layers are timed,
for two paths on the same inputs:
  plain    one instance, complex128 contraction, every update launched on its
           own (the reference's one-update-at-a-time structure)
  grouped  --instances independent instances; every update of a layer across
           all of them in one grouped run (one launch per stage), contraction
           in --precision (default complex64 on the FP32 MFMA), SVD /
           truncation / BS elements / norm in FP64
`value` is the grouped path's BS updates/s over its timed layers; the plain
path's rate and the ratio are in `config`, with the mean kept bond dimension
at the start and end of the timed layers.

Also on the JSON line:
  roofline      the grouped run's dominant kernel: algorithmic work per launch
                / its average HIP-event time over the timed region; `stages`
                holds the same figures for every stage
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
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0}   # MI355X dense (MFMA; FP64 vector too)
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4, help="timed layers")
    ap.add_argument("--warmup", type=int, default=2, help="untimed layers just before the timed ones")
    ap.add_argument("--start", type=int, default=16,
                    help="first timed layer (at least --warmup); earlier layers prepare the state")
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
    ap.add_argument("--streams", type=int, default=1, choices=[1, 2, 3, 4],
                    help="launching streams the host-planned run's groups alternate between "
                         "(1: each stage's HIP events time the stage alone, the roofline's "
                         "source; 2 is that path's fastest setting)")
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
    ap.add_argument("--two-stage", type=int, default=-1,
                    help="two-stage reduction threshold (bsmpo_set_two_stage; 0 = off, "
                         "-1 = the library default)")
    ap.add_argument("--graph", type=int, default=1,
                    help="1: also time the grouped path as device-planned layer programs (one "
                         "captured HIP graph per layer); `value` is then that run's rate")
    ap.add_argument("--graph-groups", type=int, default=8, help="groups per layer of the graph run")
    ap.add_argument("--graph-chains", type=int, default=3, choices=[1, 2, 3, 4],
                    help="concurrent chains of groups (scratch arenas) of the graph run")
    ap.add_argument("--graph-headroom", type=float, default=1.3,
                    help="capacities of the graph run's groups over what layer --start needs")
    ap.add_argument("--verbose", action="store_true")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stage_table(stages):
    for s, (ms, fl, by, cnt) in stages.items():
        if cnt:
            log(f"[bench]   {s:13s} {ms:10.3f} ms {cnt:6d} launches "
                f"{fl / max(ms, 1e-9) / 1e9:8.3f} TFLOP/s {by / max(ms, 1e-9) / 1e6:9.2f} GB/s")


# Roof of each stage: ("mfma" | "valu" | "hbm", peak key); None = latency-bound
# chains with no roof (Jacobi sweeps, bisection, inverse iteration).
STAGE_ROOF = {"contract": ("mfma", None), "gram": ("mfma", "f64"), "wmat": ("mfma", "f64"),
              "vectors": ("mfma", "f64"), "backtransform": ("mfma", "f64"),
              "tridiag": ("hbm", None), "topk": ("hbm", None), "scatter": ("hbm", None),
              "observables": ("hbm", None), "bs_table": ("hbm", None),
              "jacobi": None, "bisect": None, "trivec": None}


def stage_entry(name, ms, fl, by, cnt, precision):
    """Algorithmic work of one stage over the timed region against its roof."""
    e = {"ms": round(ms, 3), "launches": cnt, "avg_launch_ms": round(ms / cnt, 4)}
    roof = STAGE_ROOF.get(name)
    t = ms / 1e3
    if roof is None or t <= 0:
        e["bound"] = "latency"
        return e
    kind, key = roof
    if kind == "hbm":
        ach = by / t / 1e9
        e.update({"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                  "frac": ach / PEAK_HBM_GBS, "work_per_launch": by / cnt})
    else:
        if key is None:  # the contraction: the precision's MFMA peak
            key = {"c128": "f64", "c64": "f32", "bf16x3": "bf16"}[precision]
        ach = fl / t / 1e12
        peak = PEAK_TFLOPS[key]
        e.update({"bound": kind, "achieved": ach, "peak": peak, "unit": "TFLOP/s",
                  "frac": ach / peak, "work_per_launch": fl / cnt})
    return e


def roofline(stages, precision):
    table = {s: stage_entry(s, *v, precision) for s, v in stages.items() if v[3] > 0}
    dom = max(table, key=lambda s: table[s]["ms"])
    d = table[dom]
    roof = {k: d.get(k) for k in ("bound", "achieved", "peak", "unit", "frac")}
    roof.update({"traffic": None, "kernel": dom, "launches": d["launches"],
                 "avg_launch_ms": d["avg_launch_ms"], "stages": table})
    return roof


def measured_traffic(kernel, shape, precision, eig_threshold, alg_bytes_per_launch):
    """HBM bytes per launch of `kernel`: the ratio of measured to algorithmic
    bytes recorded in profiles/<kernel>_*_traffic.json for this workload shape,
    precision and eigensolver split (two rocprofv3 --pmc passes, FETCH_SIZE and
    WRITE_SIZE, scaled by a probe kernel of known size), times this run's
    algorithmic bytes per launch; None if no record matches."""
    import glob
    for path in sorted(glob.glob(os.path.join(HERE, "profiles", f"{kernel}_*_traffic.json"))):
        with open(path) as f:
            rec = json.load(f)
        if (rec.get("kernel"), rec.get("shape"), rec.get("precision"),
                rec.get("eig_threshold")) == (kernel, shape, precision, eig_threshold):
            return rec["traffic_over_algorithmic"] * alg_bytes_per_launch, os.path.relpath(path, HERE)
    return None, None


def kept_bond_dimension(st):
    """Mean number of kept singular values (Lambda > 0) over the inner bonds
    that hold any."""
    cnt = (st.Lambda > 0).sum(dim=1).cpu().numpy()
    cnt = cnt[cnt > 0]
    return float(cnt.mean()) if cnt.size else 0.0


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
    if args.two_stage >= 0:
        _capi.check(_capi.lib().bsmpo_set_two_stage(args.two_stage), "set_two_stage")
    two_stage = int(_capi.lib().bsmpo_two_stage_threshold())
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
    T0 = max(args.start, W)          # first timed layer
    if T0 + K > M:
        raise SystemExit("the timed layers run past the last layer of the mesh")
    paths = args.paths.split(",")
    B = args.instances
    meshes = [mesh.haar_clements_mesh(M, args.seed + b)[0] for b in range(B)]
    t0 = time.time()
    init = engine.initialize(M, d, chi, inputs, args.eta)
    torch.cuda.synchronize()
    log(f"[bench] init {time.time() - t0:.2f}s; untimed layers 0..{T0 - 1} "
        f"(warmup {T0 - W}..{T0 - 1}), timed {T0}..{T0 + K - 1}")
    max_input = max(inputs)
    result = {}
    shape = f"M={M} N={N} stride={args.stride} d={d} chi={chi} eta={args.eta} instances={B}"

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
        for layer in range(T0):
            if args.autotune and args.eig_threshold < 0 and layer == T0 - 1:
                best, times = engine.autotune_eig_threshold([st], [meshes[0]], inputs, layer,
                                                            precision=engine.C128, grouped=False)
                plain_tuned = {"layer": layer, "eig_threshold": best,
                               "seconds": {str(k): round(v, 3) for k, v in times.items()}}
                log(f"[bench] plain path: autotuned eigensolver split on layer {layer}: "
                    f"threshold {best} {times}")
            run_layer(layer, engine.C128)
        torch.cuda.synchronize()
        log(f"[bench] plain untimed {T0} layers {time.time() - t1:.1f}s")
        if args.cpu_seconds > 0:
            snap = st.to_host()
        bond0 = kept_bond_dimension(st)
        obs_start = st.observables()
        _capi.timing_enable(True)
        _capi.timing_read(reset=True)
        torch.cuda.synchronize()
        ts = time.time()
        n = sum(run_layer(layer, engine.C128) for layer in range(T0, T0 + K))
        torch.cuda.synchronize()
        el = time.time() - ts
        st_plain = _capi.timing_read(reset=True)
        _capi.timing_enable(False)
        result["plain"] = {"updates": n, "seconds": el, "rate": n / el,
                           "observables": st.observables(), "tuned": plain_tuned,
                           "observables_start": obs_start,
                           "bond": [bond0, kept_bond_dimension(st)]}
        log(f"[bench] plain c128: {n} updates in {el:.3f}s = {n / el:.2f} updates/s")
        stage_table(st_plain)
        del st
        torch.cuda.empty_cache()

    # -------------------------------------------------------------- grouped
    roof = None
    if "grouped" in paths:
        rules = engine.set_precision_rules(prec)  # tie band and drop rule of the precision
        states = [init.clone() for _ in range(B)]
        runner = engine.GroupRunner(init.device, max_bytes=int(args.group_gib * (1 << 30)),
                                    streams=args.streams, first_frac=args.first_frac)
        t1 = time.time()
        tuned = fixed_threshold(args)
        for layer in range(T0):
            if args.autotune and args.eig_threshold < 0 and layer == T0 - 1 and args.graph:
                # tune launch and tile geometry on the last untimed layer (copies of
                # the states, as layer-program replays); both timed runs use it
                runner.works = [None] * len(runner.works)
                torch.cuda.empty_cache()
                best, times = engine.autotune_layer(states, meshes, layer, precision=prec,
                                                    n_groups=args.graph_groups,
                                                    n_streams=args.graph_chains,
                                                    headroom=args.graph_headroom, log=log)
                tuned = {"layer": layer, **best,
                         "seconds": {k: {str(a): round(b, 3) for a, b in v.items()}
                                     for k, v in times.items()}}
            elif args.autotune and args.eig_threshold < 0 and layer == T0 - 1:
                # tune the eigensolver route split on the last untimed layer (clones)
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
                log(f"[bench] grouped untimed layer {layer}: {time.time() - t1:.1f}s "
                    f"groups {runner.last_info.get('groups')}")
        log(f"[bench] grouped untimed {T0} layers x {B} instances {time.time() - t1:.1f}s")
        graph_states = [s.clone() for s in states] if args.graph else None
        bond0 = kept_bond_dimension(states[0])
        g_obs_start = states[0].observables()
        _capi.timing_enable(True)
        _capi.timing_read(reset=True)
        torch.cuda.synchronize()
        ts = time.time()
        n = 0
        for layer in range(T0, T0 + K):
            k, _ = engine.simulate_batch(states, meshes, inputs, layers=[layer], runner=runner,
                                         precision=prec)
            n += k
            log(f"[bench] grouped timed layer {layer}: {n} updates {time.time() - ts:.3f}s")
        torch.cuda.synchronize()
        el = time.time() - ts
        st_grp = _capi.timing_read(reset=True)
        _capi.timing_enable(False)
        result["grouped"] = {"updates": n, "seconds": el, "rate": n / el,
                             "observables0": states[0].observables(), "tuned": tuned,
                             "observables0_start": g_obs_start,
                             "bond": [bond0, kept_bond_dimension(states[0])], "rules": rules}
        engine.set_precision_rules(engine.C128)
        log(f"[bench] grouped {args.precision}: {n} updates in {el:.3f}s = {n / el:.2f} updates/s")
        stage_table(st_grp)
        roof = roofline(st_grp, args.precision)
        roof["launch_streams"] = args.streams
        if args.streams > 1:
            roof["note"] = (f"groups go round {args.streams} launching streams: a launch's duration "
                            "includes sharing the GPU with the other streams' stages")
        if args.graph:
            # the same layers from the same state as layer programs: planned on
            # the device, one captured HIP graph replayed per layer
            final_eager = [s.observables() for s in states]
            del states, runner
            torch.cuda.empty_cache()
            rules = engine.set_precision_rules(prec)
            t1 = time.time()
            lg = engine.LayerGraphRunner(graph_states, meshes, prec, n_groups=args.graph_groups,
                                         n_streams=args.graph_chains, headroom=args.graph_headroom)
            lg._set_caps(lg.needs(T0 % 2))
            lg._create()
            torch.cuda.synchronize()
            log(f"[bench] layer program built in {time.time() - t1:.2f}s "
                f"(groups {args.graph_groups}, work arena {lg.caps.work_bytes / 2**30:.1f} GiB)")
            ts = time.time()
            ng = lg.run(range(T0, T0 + K))
            torch.cuda.synchronize()
            elg = time.time() - ts
            dm = max(float(np.abs(a[1] - b.observables()[1]).max())
                     for a, b in zip(final_eager, graph_states))
            result["graph"] = {"updates": ng, "seconds": elg, "rate": ng / elg,
                               "fallbacks": lg.fallbacks, "groups": args.graph_groups,
                               "chains": args.graph_chains,
                               "marginals_vs_host_planned": dm,
                               "bond": [bond0, kept_bond_dimension(graph_states[0])]}
            engine.set_precision_rules(engine.C128)
            log(f"[bench] graph {args.precision}: {ng} updates in {elg:.3f}s = {ng / elg:.2f} "
                f"updates/s (fallbacks {lg.fallbacks}; marginals vs host-planned {dm:.2e})")
            lg.destroy()
            del lg, graph_states
            torch.cuda.empty_cache()

    cpu = cpu_baseline(snap, d, chi, meshes[0], inputs, T0, args.cpu_seconds) if snap else None
    workload = (f"lossy boson-sampling MPO site updates, M={M} N={N} "
                f"(modes 0,{args.stride},..) d={d} chi={chi} eta={args.eta}; "
                f"layers {T0}..{T0 + K - 1} timed")
    g = result.get("grouped")
    p = result.get("plain")
    if roof is not None and roof["bound"] == "hbm":
        thr = (g or {}).get("tuned") or {}
        alg = roof["stages"][roof["kernel"]]["work_per_launch"]
        roof["traffic"], src = measured_traffic(roof["kernel"], shape, args.precision,
                                                thr.get("eig_threshold", 96), alg)
        if src:
            roof["traffic_source"] = src
    gr = result.get("graph")
    main_path = gr or g or p
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
            "timed_layers": [T0, T0 + K - 1],
            "kept_bond_dimension": {"chi": chi,
                                    "grouped_start_end": g["bond"] if g else None,
                                    "plain_start_end": p["bond"] if p else None},
            "grouped": ({"instances": B, "precision": args.precision, "updates": g["updates"],
                         "updates_per_s": g["rate"], "autotune": g.get("tuned"),
                         "launch_streams": args.streams, "two_stage_min_n": two_stage,
                         "tie_rel": g["rules"][0], "drop_rel": g["rules"][1]}
                        if g else None),
            "plain_fp64": ({"instances": 1, "updates": p["updates"],
                            "updates_per_s": p["rate"], "autotune": p.get("tuned")} if p else None),
            "graph": ({"updates": gr["updates"], "updates_per_s": gr["rate"],
                       "groups_per_layer": gr["groups"], "chains": gr["chains"],
                       "fallbacks": gr["fallbacks"],
                       "marginals_vs_host_planned": gr["marginals_vs_host_planned"],
                       "note": "value: the grouped path as layer programs (planned on the "
                               "device, captured HIP graphs replayed per layer, no host "
                               "round trip); `grouped` is the same work host-planned group "
                               "by group, whose stage events give the roofline"}
                      if gr else None),
            "grouped_over_plain": (main_path["rate"] / p["rate"]) if (g and p) else None,
            "precision_check": precision_check(g, p, args.precision, T0 + K),
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
    out = {"precision": precision, "layers": layers, "reference": "plain complex128, same mesh",
           "marginal_max_abs_err": float(np.abs(mg - mp).max()),
           "trace_abs_err": float(abs(tg - tp)), "trace": float(tp)}
    if "observables0_start" in g and "observables_start" in p:
        (ts, ms), (tq, mq) = g["observables0_start"], p["observables_start"]
        out["at_first_timed_layer"] = {"marginal_max_abs_err": float(np.abs(ms - mq).max()),
                                       "trace_abs_err": float(abs(ts - tq))}
    return out


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
