"""Forward+adjoint evaluations per second at 2M cells, mixed precision vs FP64.

One step = one gradient evaluation of the inverse-problem objective on one
GPU, every array resident in HBM before the timed region starts:

  logK (2M cells) -> K_e = hm(exp(logK_a), exp(logK_b))          (6M edges)
  -> steady solve h (assembly, numeric AMG refresh, AMG-PCG to 1e-10)
  -> J = sum_obs (h - h_obs)^2 over 200 cells, delta = dJ/dh
  -> adjoint solve through the same hierarchy -> dJ/dK_e -> dJ/dlogK.

Workload: 200x200x50 cells over 100x100x25 m, layered Matern(nu=1) logK with
correlation length 50 m, sigma 1, 4 injectors of 1e-4 m^3/s, Dirichlet h = 0
on the top face; observed heads from a second field plus N(0, 1e-3^2) noise.
Steps cycle through 4 pre-generated logK fields so no step reuses an earlier
result.  The FP64 path is timed on the same steps for the ratio.

Prints one JSON line (the driver's contract).  ``--gpus N`` with N > 1 exits:
this build runs on one GPU.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nx", type=int, nargs=3, default=[200, 200, 50])
    ap.add_argument("--cpu-baseline", choices=["auto", "skip"], default="auto")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--workload", choices=["steady2m", "ensemble", "transient", "twophase",
                                           "richards", "transport", "barrier"],
                    default="steady2m",
                    help="steady2m (default, the headline line); ensemble: 96 packed "
                         "realizations on 32x64^3 + 32x96x96x24 + 32x128x128x16 grids; "
                         "transient: 200 backward-Euler steps on the 2M grid with the "
                         "checkpointed adjoint; twophase: IMPES steps on the 2M grid; "
                         "richards: steady Richards solve + gradient on the 2M grid; "
                         "transport: advection-dispersion steps + gradient on the 2M grid")
    ap.add_argument("--sigma", type=float, default=1.0,
                    help="steady2m: standard deviation of the logK fields (the conditioning "
                         "sweep runs 1, 2 and 3; the headline line is sigma 1)")
    ap.add_argument("--inner-reltol", type=float, default=1e-5,
                    help="steady2m: relative tolerance of each FP32 inner solve of the mixed "
                         "precision's refinement")
    ap.add_argument("--graph", action="store_true",
                    help="steady2m and ensemble: replay each PCG iteration as a captured "
                         "hipGraph (DPF_FLAG_GRAPH)")
    ap.add_argument("--tuning", default=None, metavar="FINE,SECOND,DEEPER,RATIO",
                    help="steady2m: fix the Chebyshev degrees and interval ratio of every "
                         "precision instead of autotuning (e.g. 1,1,2,15, to repeat a tuned "
                         "run's configuration under a profiler, whose overheads change the "
                         "autotuner's timings)")
    ap.add_argument("--autotune", action="store_true",
                    help="steady2m and ensemble: let the first solve of each precision time "
                         "the coarse levels' lanes per row and the smoother degrees "
                         "(DPF_FLAG_AUTOTUNE) instead of the library's fixed defaults, which "
                         "measure as fast or faster on these workloads")
    ap.add_argument("--lanes", default=None, metavar="LG1,LG2,...",
                    help="steady2m: pin the lanes-per-row exponent of each split coarse "
                         "level (in level order) instead of autotuning it")
    ap.add_argument("--fp32-storage", action="store_true",
                    help="steady2m: the FP32 V-cycle reads FP32 operator values instead of "
                         "bf16 (DPF_FLAG_FP32_STORAGE)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 256 MiB device copy (the reference kernel that "
                         "scales the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU (asked for --gpus {args.gpus}); nothing in this build "
              "uses more than one GPU.", file=sys.stderr)
        sys.exit(2)

    import torch
    from __graft_entry__ import load_package
    P = load_package()
    from dpfehm_jl_amd import _lib, api

    if args.workload == "ensemble":
        return bench_ensemble(args, P, api, _lib, torch)
    if args.workload == "transient":
        return bench_transient(args, P, api, _lib, torch)
    if args.workload == "twophase":
        return bench_twophase(args, P, api, _lib, torch)
    if args.workload == "richards":
        return bench_richards(args, P, api, _lib, torch)
    if args.workload == "transport":
        return bench_transport(args, P, api, _lib, torch)
    if args.workload == "barrier":
        return bench_barrier(args, _lib)

    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    ns = tuple(args.nx)
    maxs = (100.0 * (ns[0] - 1) / 199.0, 100.0 * (ns[1] - 1) / 199.0, 25.0 * (ns[2] - 1) / 49.0)
    t0 = time.time()
    w = P.workloads.steady_2m(sigma=args.sigma, ns=ns, maxs=maxs)
    log(f"[bench] problem n={w.n} E={w.nedges} built in {time.time() - t0:.1f}s")
    t0 = time.time()
    topo = api.Topology(w.n, w.neighbors, w.areasoverlengths, w.dirichletnodes)
    if args.tuning:
        f, c, d, r = args.tuning.split(",")
        for prec in ("fp64", "mixed", "mixed_pc"):
            topo.set_smoother(prec, int(f), int(c), int(d), float(r))
    if args.lanes:
        split = [l for l, g in enumerate(topo.lanes()[0]) if g > 0]
        for l, g in zip(split, args.lanes.split(",")):
            topo.set_lanes(l, int(g))
    topo_s = time.time() - t0
    log(f"[bench] topology: nfree={topo.nfree} nnz={topo.nnz} levels={topo.levels} "
        f"coarse={topo.coarse_rows} opcx={topo.operator_complexity:.2f} "
        f"host symbolic setup {topo.symbolic_setup_ms:.0f} ms (create {topo_s:.1f}s)")

    stream = torch.cuda.current_stream(dev)
    sptr = ctypes.c_void_p(stream.cuda_stream)
    L = _lib.lib()
    sp = grid_sp = P.grid.grid_spacing(w.mins, w.maxs, w.ns)
    fields = [w.logK] + [P.fields.layered_logK(w.ns, sp, -9.0, 50.0, args.sigma, s)
                         for s in (10, 11, 12)]
    logK_d = [torch.from_numpy(f).to(dev) for f in fields]
    dh_d = torch.from_numpy(w.dirichleths).to(dev)
    Q_d = torch.from_numpy(w.Qs).to(dev)
    obs_d = torch.from_numpy(w.obs - 1).to(dev)
    Ke_d = torch.empty(w.nedges, dtype=torch.float64, device=dev)
    h_d = torch.empty(w.n, dtype=torch.float64, device=dev)
    delta_d = torch.zeros(w.n, dtype=torch.float64, device=dev)
    dK_d = torch.empty(w.nedges, dtype=torch.float64, device=dev)
    dlogK_d = torch.empty(w.n, dtype=torch.float64, device=dev)

    def opts(precision, timing=False):
        o = api._opts({"precision": precision, "reltol": 1e-10, "graph": args.graph,
                       "inner_reltol": args.inner_reltol, "fp32_storage": args.fp32_storage,
                       "autotune": args.autotune},
                      memory=_lib.DPF_DEVICE, stream=sptr)
        if timing:
            o.flags = (o.flags | _lib.DPF_FLAG_KERNEL_TIMING) & ~_lib.DPF_FLAG_GRAPH
        return o

    def p(t):
        return ctypes.c_void_p(t.data_ptr())

    def forward(logK, o, keep=True):
        _lib.check(L.dpf_hm_forward(topo.handle, p(logK), p(Ke_d), ctypes.byref(o)))
        st = _lib.Stats()
        sh = ctypes.c_void_p()
        _lib.check(L.dpf_steady_forward(topo.handle, p(Ke_d), p(dh_d), p(Q_d), p(h_d),
                                        ctypes.byref(o), ctypes.byref(st),
                                        ctypes.byref(sh) if keep else None))
        return st, sh

    # observed heads: FP64 solve on the observation field + noise (set-up, untimed)
    lk_obs = torch.from_numpy(P.workloads.observation_logK(w, seed=4)).to(dev)
    forward(lk_obs, opts("fp64"), keep=False)
    h_obs_d = h_d[obs_d] + torch.from_numpy(P.workloads.observation_noise(w.obs.size)).to(dev)

    def evaluate(k, o, stats):
        logK = logK_d[k % len(logK_d)]
        st, sh = forward(logK, o)
        r = h_d[obs_d] - h_obs_d
        J = torch.sum(r * r)
        delta_d.zero_()
        delta_d[obs_d] = 2.0 * r
        st2 = _lib.Stats()
        code = L.dpf_steady_pullback(sh, p(delta_d), p(dK_d), None, None, ctypes.byref(o),
                                     ctypes.byref(st2))
        L.dpf_state_destroy(sh)
        _lib.check(code)
        _lib.check(L.dpf_hm_pullback(topo.handle, p(logK), p(Ke_d), p(dK_d), p(dlogK_d),
                                     ctypes.byref(o)))
        stats.append((st.as_dict(), st2.as_dict()))
        return J

    if args.copy_calibration:
        src = torch.ones(1 << 26, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] calibration copy: 3 x 256 MiB read + 256 MiB written")

    results = {}
    for precision in ("fp64", "mixed", "mixed_pc"):
        o = opts(precision)
        for k in range(args.warmup):
            evaluate(k, o, [])
        torch.cuda.synchronize()
        stats = []
        t0 = time.perf_counter()
        for k in range(args.steps):
            evaluate(k, o, stats)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        results[precision] = dict(seconds=dt, stats=stats)
        its = [(a["iterations"], b["iterations"]) for a, b in stats]
        log(f"[bench] {precision}: {args.steps / dt:.2f} evals/s ({1e3 * dt / args.steps:.1f} ms/eval) "
            f"iterations fwd/adj {its[:4]} setup {stats[-1][0]['setup_ms']:.2f} ms "
            f"relres {stats[-1][0]['relres']:.2e}/{stats[-1][1]['relres']:.2e}")

    # accuracy of the mixed path against FP64 on the same field (untimed)
    acc = {}
    for precision in ("fp64", "mixed", "mixed_pc"):
        evaluate(0, opts(precision), [])
        torch.cuda.synchronize()
        acc[precision] = (h_d.cpu().numpy().copy(), dlogK_d.cpu().numpy().copy())

    def rel_err(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))
    head_err = rel_err(acc["mixed"][0], acc["fp64"][0])
    grad_err = rel_err(acc["mixed"][1], acc["fp64"][1])
    pc_err = (rel_err(acc["mixed_pc"][0], acc["fp64"][0]), rel_err(acc["mixed_pc"][1], acc["fp64"][1]))
    log(f"[bench] mixed vs fp64: head rel err {head_err:.2e}, gradient rel err {grad_err:.2e}; "
        f"mixed_pc vs fp64: {pc_err[0]:.2e}, {pc_err[1]:.2e}")

    # the same evaluation through the host-buffer interface (numpy arrays in
    # and out: logK, K_e, heads, seed and gradient cross PCIe), mixed
    # precision, for the PCIe-inclusive rate; never the headline value
    def host_eval(k):
        lk = fields[k % len(fields)]
        K = api.logKs2Ks_harmonic(lk, w.neighbors, w.dirichletnodes, w.areasoverlengths)
        h, pb = api.rrule_groundwater_steadystate(K, w.neighbors, w.areasoverlengths,
                                                  w.dirichletnodes, w.dirichleths, w.Qs,
                                                  reltol=1e-10, precision="mixed",
                                                  autotune=args.autotune)
        delta = np.zeros(w.n)
        delta[w.obs - 1] = 2.0 * (h[w.obs - 1] - h_obs_host)
        dK = pb(delta)[1]
        pb.state.close()
        return api.logKs2Ks_harmonic_pullback(lk, K, dK, w.neighbors, w.dirichletnodes,
                                              w.areasoverlengths)
    h_obs_host = h_obs_d.cpu().numpy()
    host_eval(0)
    t0 = time.perf_counter()
    nh = 3
    for k in range(nh):
        host_eval(k)
    host_rate = nh / (time.perf_counter() - t0)
    log(f"[bench] mixed through host buffers (PCIe included): {host_rate:.2f} evals/s")

    # kernel table: the same evaluations again with every launch timed by a
    # dispatch event pair on the stream it runs on, per kernel class, AMG
    # level and value size, with each launch's algorithmic bytes
    ktab = {}
    for precision in ("mixed", "fp64"):
        o = opts(precision, timing=True)
        kernel_table(L, topo, reset=True)
        for k in range(args.steps):
            evaluate(k, o, [])
        torch.cuda.synchronize()
        ktab[precision] = kernel_table(L, topo, reset=True)

    # back-to-back launches of single kernels, for reference
    prof = {}
    for precision, pc in (("mixed", _lib.DPF_MIXED), ("fp64", _lib.DPF_FP64)):
        evaluate(1, opts(precision), [])
        for kern in (0, 1):
            ms = ctypes.c_double()
            by = ctypes.c_double()
            o = opts(precision)
            _lib.check(L.dpf_profile_kernel(topo.handle, pc, kern, 50, ctypes.byref(ms),
                                            ctypes.byref(by), ctypes.byref(o)))
            prof[(precision, kern)] = (ms.value, by.value)
            log(f"[bench] kernel {kern} {precision}: {ms.value * 1e3:.1f} us/launch, "
                f"{by.value / 1e6:.1f} MB -> {by.value / ms.value / 1e6:.0f} GB/s")

    tuned = topo.tuned_degrees()
    lanes, lane_trials = topo.lanes()
    log(f"[bench] Chebyshev degrees (finest, second, deeper, ratio): {tuned}; layout {topo.layout()}; "
        f"lanes per row by level {[1 << g if g else 0 for g in lanes]}; lane trials (us) "
        f"{ {f'{l}/{1 << k}': round(v, 2) for (l, k), v in lane_trials.items()} }")
    value = args.steps / results["mixed"]["seconds"]
    fp64_rate = args.steps / results["fp64"]["seconds"]

    cpu = None
    if args.cpu_baseline == "auto":
        cpu = cpu_baseline(w, fields, h_obs_d.cpu().numpy(), args.cpu_seconds)
    roof, evalroof, tables = roofline_from_tables(ktab, results, args.steps)

    out = {
        "metric": "forward+adjoint evals/s @2M cells, mixed vs FP64 GPU, <=1e-8 rel head err",
        "value": round(value, 3),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * results["mixed"]["seconds"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 refinement",
        "data": "synthetic",
        "config": {
            "workload": f"steady groundwater forward+adjoint, {ns[0]}x{ns[1]}x{ns[2]} cells "
                        f"({w.n} cells, {w.nedges} edges), harmonic-mean K from layered "
                        f"Matern(nu=1, l=50 m, sigma={args.sigma:g}) logK, top-face Dirichlet, "
                        "4 injectors, "
                        "200 observations, AMG-PCG to 1e-10, mixed precision",
            "cells": w.n, "edges": w.nedges, "nfree": topo.nfree, "nnz": topo.nnz,
            "amg_levels": topo.levels, "reltol": 1e-10, "inner_reltol": args.inner_reltol,
        },
        "fp64_evals_per_s": round(fp64_rate, 3),
        "mixed_over_fp64": round(value / fp64_rate, 3),
        "mixed_pc_evals_per_s": round(args.steps / results["mixed_pc"]["seconds"], 3),
        "host_buffers_mixed_evals_per_s": round(host_rate, 3),
        "mixed_pc_rel_err_head_grad": pc_err,
        "head_rel_err_mixed_vs_fp64": head_err,
        "grad_rel_err_mixed_vs_fp64": grad_err,
        "iterations_mixed_fwd_adj": [results["mixed"]["stats"][0][0]["iterations"],
                                     results["mixed"]["stats"][0][1]["iterations"]],
        "iterations_fp64_fwd_adj": [results["fp64"]["stats"][0][0]["iterations"],
                                    results["fp64"]["stats"][0][1]["iterations"]],
        "topology_setup_s": round(topo_s, 3),
        "numeric_setup_ms_per_solve": {
            k: round(results[k]["stats"][-1][0]["setup_ms"], 3) for k in ("fp64", "mixed")},
        "cheb_degrees_finest_second_deeper_ratio": tuned,
        "cheb_pinned": bool(args.tuning),
        "autotuned": args.autotune,
        "lanes_per_row_by_level": [1 << g if g else 0 for g in lanes],
        "lane_trials_us": {f"{l}/{1 << k}": round(v, 2) for (l, k), v in lane_trials.items()},
        "hipgraph_iterations": bool(args.graph),
        "roofline": roof,
        "roofline_eval": evalroof,
        "kernel_table_mixed": tables["mixed"],
        "kernel_table_fp64": tables["fp64"],
        "kernels_back_to_back": {f"{pr}/{k}": {"us": round(v[0] * 1e3, 2),
                                           "GBps": round(v[1] / v[0] / 1e6, 1)}
                             for (pr, k), v in prof.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_ensemble(args, P, api, _lib, torch):
    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    t0 = time.time()
    ms = P.workloads.ensemble_members()
    log(f"[bench] ensemble: {len(ms)} members, {sum(m['n'] for m in ms)} cells, "
        f"generated in {time.time() - t0:.1f}s")
    t0 = time.time()
    topo = api.BatchTopology([(m["n"], m["nb"], m["aol"], m["dn"]) for m in ms])
    topo_s = time.time() - t0
    log(f"[bench] batch topology: nfree={topo.nfree} nnz={topo.nnz} levels={topo.levels} "
        f"coarse={topo.coarse_rows} host symbolic {topo.symbolic_setup_ms:.0f} ms, create {topo_s:.1f}s")
    L = _lib.lib()
    sptr = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    logK = torch.from_numpy(np.concatenate([m["logK"] for m in ms])).to(dev)
    dh = torch.zeros(topo.n, dtype=torch.float64, device=dev)
    Q = torch.from_numpy(np.concatenate([m["Q"] for m in ms])).to(dev)
    obs = torch.from_numpy(np.concatenate([m["obs"] - 1 + c for m, c in
                                           zip(ms, topo.cell0[:-1])])).to(dev)
    Ke = torch.empty(topo.ne, dtype=torch.float64, device=dev)
    h = torch.empty(topo.n, dtype=torch.float64, device=dev)
    delta = torch.zeros(topo.n, dtype=torch.float64, device=dev)
    dK = torch.empty(topo.ne, dtype=torch.float64, device=dev)
    dlogK = torch.empty(topo.n, dtype=torch.float64, device=dev)

    def p(t):
        return ctypes.c_void_p(t.data_ptr())

    def evaluate(o):
        _lib.check(L.dpf_hm_forward(topo.handle, p(logK), p(Ke), ctypes.byref(o)))
        st, sh = _lib.Stats(), ctypes.c_void_p()
        _lib.check(L.dpf_steady_forward(topo.handle, p(Ke), p(dh), p(Q), p(h), ctypes.byref(o),
                                        ctypes.byref(st), ctypes.byref(sh)))
        delta.zero_()
        delta[obs] = 2.0 * h[obs]
        st2 = _lib.Stats()
        code = L.dpf_steady_pullback(sh, p(delta), p(dK), None, None, ctypes.byref(o),
                                     ctypes.byref(st2))
        L.dpf_state_destroy(sh)
        _lib.check(code)
        _lib.check(L.dpf_hm_pullback(topo.handle, p(logK), p(Ke), p(dK), p(dlogK),
                                     ctypes.byref(o)))
        return st.as_dict(), st2.as_dict()

    res = {}
    for precision in ("fp64", "mixed", "mixed_pc"):
        o = api._opts({"precision": precision, "reltol": 1e-10, "graph": args.graph,
                       "autotune": args.autotune},
                      memory=_lib.DPF_DEVICE, stream=sptr)
        for _ in range(args.warmup):
            evaluate(o)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stats = [evaluate(o) for _ in range(args.steps)]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        rr = topo.member_relres()
        res[precision] = (dt, stats, float(rr.max()))
        log(f"[bench] ensemble {precision}: {args.steps / dt:.2f} batch evals/s = "
            f"{len(ms) * args.steps / dt:.1f} member forward+adjoint evals/s; iterations "
            f"{stats[-1][0]['iterations']}/{stats[-1][1]['iterations']} in "
            f"{stats[-1][0]['outer']}/{stats[-1][1]['outer']} refinement steps, worst member "
            f"relres {rr.max():.2e}, tuned {topo.tuned_degrees().get(precision)}")
    value = len(ms) * args.steps / res["mixed"][0]
    out = {
        "metric": "member forward+adjoint evals/s, 96-realization packed ensemble, mixed",
        "value": round(value, 2), "unit": "member evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * res["mixed"][0] / args.steps, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 inner / f64 refinement", "data": "synthetic",
        "config": {"workload": "96 packed realizations: 32x64^3 + 32x96x96x24 + 32x128x128x16 "
                               "cells, sigma 1.5, correlation length U(20,80) m, "
                               "forward+adjoint to 1e-10 per member",
                   "cells": topo.n, "nfree": topo.nfree, "amg_levels": topo.levels},
        "fp64_member_evals_per_s": round(len(ms) * args.steps / res["fp64"][0], 2),
        "mixed_pc_member_evals_per_s": round(len(ms) * args.steps / res["mixed_pc"][0], 2),
        "worst_member_relres": {k: v[2] for k, v in res.items()},
        "topology_setup_s": round(topo_s, 2),
    }
    print(json.dumps(out), flush=True)


def bench_transient(args, P, api, _lib, torch):
    """200 backward-Euler steps of 3600 s on the 2M-cell grid (Ss = 1e-5 1/m,
    injection ramp over the first 50 steps), objective on the final heads at
    the 200 observation cells, checkpointed adjoint (every 20 steps)."""
    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    w = P.workloads.steady_2m()
    topo = api.Topology(w.n, w.neighbors, w.areasoverlengths, w.dirichletnodes)
    L = _lib.lib()
    sptr = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    nsteps, dt = 200, 3600.0
    qs = torch.from_numpy(P.workloads.transient_schedule(nsteps, dt, 50)).to(dev)
    K = torch.from_numpy(w.Ks()).to(dev)
    dh = torch.from_numpy(w.dirichleths).to(dev)
    Q = torch.from_numpy(w.Qs).to(dev)
    Ss = torch.full((w.n,), 1e-5, dtype=torch.float64, device=dev)
    V = torch.from_numpy(w.volumes).to(dev)
    u0 = torch.zeros(topo.nfree, dtype=torch.float64, device=dev)
    uN = torch.empty(topo.nfree, dtype=torch.float64, device=dev)
    free_idx = np.cumsum(topo.isfree) - 1
    obs_f = torch.from_numpy(free_idx[w.obs - 1]).to(dev)
    g = torch.zeros(topo.nfree, dtype=torch.float64, device=dev)
    dK = torch.empty(topo.ne, dtype=torch.float64, device=dev)
    dQ = torch.empty(topo.n, dtype=torch.float64, device=dev)

    def p(t):
        return ctypes.c_void_p(t.data_ptr())

    out = {}
    # the configured checkpointing (every 20 steps) in each precision, then
    # every state kept (201 x 1.96 M x 8 B = 3.2 GB: no recomputation)
    for precision, ck in (("fp64", 20), ("mixed", 20), ("mixed_pc", 20), ("mixed_pc", 1)):
        o = api._opts({"precision": precision, "reltol": 1e-10}, memory=_lib.DPF_DEVICE,
                      stream=sptr)
        st, sh = _lib.Stats(), ctypes.c_void_p()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _lib.check(L.dpf_transient_forward(topo.handle, p(K), p(dh), p(Q), p(Ss), p(V), p(u0),
                                           nsteps, dt, p(qs), ck, p(uN), None, ctypes.byref(o),
                                           ctypes.byref(st), ctypes.byref(sh)))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        g.zero_()
        g[obs_f] = 2.0 * uN[obs_f]
        st2 = _lib.Stats()
        code = L.dpf_transient_pullback(sh, p(g), p(dK), p(dQ), ctypes.byref(o), ctypes.byref(st2))
        L.dpf_state_destroy(sh)
        _lib.check(code)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        key = precision if ck == 20 else f"{precision}_all_states"
        out[key] = dict(forward_s=round(t1 - t0, 3), adjoint_s=round(t2 - t1, 3),
                        forward_iterations=st.iterations, adjoint_iterations=st2.iterations)
        log(f"[bench] transient {key}: forward {t1 - t0:.2f}s ({st.iterations} PCG "
            f"iterations over {nsteps} steps), adjoint {t2 - t1:.2f}s "
            f"({st2.iterations} iterations incl. recomputation; checkpoint every {ck})")
    res = {"metric": "transient forward+adjoint wall time, 2M cells x 200 backward-Euler steps",
           "value": out["mixed"]["forward_s"] + out["mixed"]["adjoint_s"], "unit": "s",
           "n_gpus": 1, "higher_is_better": False, "data": "synthetic",
           "dtype": "f32 inner / f64 refinement", "config": {"workload": "transient 2M, 200 steps, "
           "checkpoint every 20"}, "detail": out}
    print(json.dumps(res), flush=True)


def bench_twophase(args, P, api, _lib, torch):
    """IMPES (src/twophase.jl:238-264) on the 2M-cell grid: K = exp(logK - mean)
    of the 2M field per cell, the 4 injectors at 1 m^3/s each, head 0 on the
    top face, fluid (vw, vo, swc, sor) = (1, 5, 0.1, 0.15), S0 = swc, and dt =
    100 CFL limits of the first pressure field, so a step is one pressure
    solve plus ~100 upstream sub-steps.  Steps are timed with inputs resident
    on the GPU; the upstream kernel is timed alone on the first step's fluxes.
    cpu_baseline: the numpy oracle's sub-step on the same grid."""
    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    w = P.workloads.steady_2m()
    n, ne = w.n, w.nedges
    topo = api.Topology(w.n, w.neighbors, w.areasoverlengths, w.dirichletnodes)
    L = _lib.lib()
    sptr = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    Kc = np.exp(w.logK - w.logK.mean())
    Qh = np.where(w.Qs > 0, 1.0, 0.0)
    fluid_h = np.array([1.0, 5.0, 0.1, 0.15])
    K = torch.from_numpy(Kc).to(dev)
    Q = torch.from_numpy(Qh).to(dev)
    dh = torch.from_numpy(w.dirichleths).to(dev)
    vol = torch.from_numpy(w.volumes).to(dev)
    S0 = torch.full((n,), 0.1, dtype=torch.float64, device=dev)
    Pd = torch.zeros(n, dtype=torch.float64, device=dev)
    Sd = torch.empty(n, dtype=torch.float64, device=dev)
    Vn = torch.empty(ne, dtype=torch.float64, device=dev)
    fl = ctypes.c_void_p(fluid_h.ctypes.data)  # the fluid parameters stay in host memory

    def p(t):
        return ctypes.c_void_p(t.data_ptr())

    # dt from the first pressure field (set-up, untimed)
    o = api._opts({"precision": "fp64", "reltol": 1e-10}, memory=_lib.DPF_DEVICE, stream=sptr)
    mw = ((0.1 - 0.1) / 0.75) ** 2 / 1.0
    mo = (1 - (0.1 - 0.1) / 0.75) ** 2 / 5.0
    Km = torch.from_numpy((mw + mo) * Kc).to(dev)
    _lib.check(L.dpf_twophase_tpfa(topo.handle, p(Km), p(dh), p(Q), p(Pd), p(Vn), 0,
                                   ctypes.byref(o), None))
    torch.cuda.synchronize()
    vn = Vn.cpu().numpy()
    nb0 = w.neighbors - 1
    Vi = np.zeros(n)
    np.add.at(Vi, np.where(vn < 0, nb0[:, 0], nb0[:, 1]), np.abs(vn))
    cfl = (0.75 / 3.0) * np.min(w.volumes / (Vi + Qh))
    dt = 100.0 * cfl
    log(f"[bench] twophase: n={n}, cfl {cfl:.4g}, dt {dt:.4g} (100 CFL limits)")
    out = {}
    for precision in ("fp64", "mixed"):
        o = api._opts({"precision": precision, "reltol": 1e-10}, memory=_lib.DPF_DEVICE,
                      stream=sptr)
        nsub = ctypes.c_int64()
        st = _lib.Stats()

        def run(nt, S_in):
            _lib.check(L.dpf_solvetwophase(topo.handle, p(Pd), p(S_in), p(K), p(dh), p(Q),
                                           p(vol), fl, dt, nt, 0, p(Pd), p(Sd), None, None,
                                           ctypes.byref(nsub), None, ctypes.byref(o),
                                           ctypes.byref(st)))
        run(max(args.warmup, 1), S0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run(args.steps, S0)
        torch.cuda.synchronize()
        dt_s = time.perf_counter() - t0
        out[precision] = dict(steps_per_s=round(args.steps / dt_s, 3),
                              ms_per_step=round(1e3 * dt_s / args.steps, 3),
                              substeps=int(nsub.value), pcg_iterations=st.iterations)
        log(f"[bench] twophase {precision}: {args.steps / dt_s:.2f} IMPES steps/s "
            f"({1e3 * dt_s / args.steps:.1f} ms/step, {nsub.value} sub-steps, "
            f"{st.iterations} PCG iterations over {args.steps} steps)")
        # gradient of the observed pressures after the same steps with respect
        # to the cell permeabilities: forward with the tape, then the pullback
        sh = ctypes.c_void_p()
        seed = torch.zeros(n, dtype=torch.float64, device=dev)
        seed[torch.from_numpy(w.obs - 1).to(dev)] = 1.0
        dKd = torch.empty(n, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _lib.check(L.dpf_solvetwophase(topo.handle, p(Pd), p(S0), p(K), p(dh), p(Q), p(vol), fl,
                                       dt, args.steps, 0, p(Pd), p(Sd), None, None,
                                       ctypes.byref(nsub), ctypes.byref(sh), ctypes.byref(o),
                                       ctypes.byref(st)))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        st2 = _lib.Stats()
        code = L.dpf_solvetwophase_pullback(sh, p(seed), None, p(dKd), None, ctypes.byref(o),
                                            ctypes.byref(st2))
        L.dpf_state_destroy(sh)
        _lib.check(code)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        out[precision].update(taped_forward_s=round(t1 - t0, 3), adjoint_s=round(t2 - t1, 3),
                              adjoint_pcg_iterations=st2.iterations)
        log(f"[bench] twophase {precision} gradient: taped forward {t1 - t0:.3f} s, adjoint "
            f"{t2 - t1:.3f} s ({st2.iterations} adjoint PCG iterations) for {args.steps} steps")
    # the upstream kernel alone, on the first pressure field's fluxes
    _lib.check(L.dpf_twophase_tpfa(topo.handle, p(Km), p(dh), p(Q), p(Pd), p(Vn), 0,
                                   ctypes.byref(o), None))
    Sd.copy_(S0)
    nsub = ctypes.c_int64()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _lib.check(L.dpf_twophase_upstream(topo.handle, p(Sd), fl, p(Q), dt, p(Pd), p(Vn), p(vol),
                                       ctypes.byref(nsub), ctypes.byref(o)))
    torch.cuda.synchronize()
    up_s = time.perf_counter() - t0
    us = 1e6 * up_s / max(nsub.value, 1)
    # algorithmic bytes per sub-step of k_tp_upstream: per cell S and f read and
    # written, Q, pv and inc_ptr read (7 x 8 B); per incidence entry (2 per edge)
    # the neighbour index (4 B) and the upwind-signed flux weight (8 B)
    ninc = 2 * ne
    alg = 56 * n + 12 * ninc
    gbs = alg / (us * 1e-6) / 1e9
    log(f"[bench] upstream alone: {nsub.value} sub-steps, {us:.1f} us/sub-step (incl. the CFL "
        f"reduction and launch gaps), {alg / 1e6:.0f} MB algorithmic -> {gbs:.0f} GB/s")
    # cpu baseline: one numpy oracle sub-step on the same arrays
    cpu = None
    if args.cpu_baseline == "auto":
        from oracle import twophase_oracle as TPO
        Sh = np.full(n, 0.1)
        Ph = Pd.cpu().numpy()
        vnh = Vn.cpu().numpy()
        t0 = time.perf_counter()
        f = TPO.fracflow(Sh, tuple(fluid_h))
        R = TPO.saturation_residual(f, Qh, w.neighbors, Ph, vnh)
        dtx = (dt / max(nsub.value, 1)) / w.volumes
        Sh = np.clip(Sh + (R * dtx + Qh * dtx), 0.1, 0.85)
        cs = time.perf_counter() - t0
        cpu = {"value": round(1.0 / cs, 3), "unit": "sub-steps/s", "cores": 1, "kind": "port",
               "sample": "one upstream sub-step of the numpy oracle on the same 2M-cell arrays"}
        log(f"[bench] cpu baseline: numpy oracle sub-step {cs:.2f} s")
    res = {"metric": "IMPES steps/s, 2M-cell two-phase five-spot (pressure solve + ~100 CFL "
                     "sub-steps per step)",
           "value": out["mixed"]["steps_per_s"], "unit": "steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
           "data": "synthetic", "dtype": "f64 transport; f32 inner / f64 refinement pressure",
           "config": {"workload": "twophase 2M (200x200x50), fluid (1, 5, 0.1, 0.15), 4 "
                                  "injectors, top-face Dirichlet, dt = 100 CFL limits"},
           "detail": out,
           "upstream": {"us_per_substep": round(us, 2), "algorithmic_bytes": alg,
                        "GBps": round(gbs, 1), "peak_GBps": 8000.0, "frac": round(gbs / 8000, 4)},
           "cpu_baseline": cpu}
    print(json.dumps(res), flush=True)


def bench_richards(args, P, api, _lib, torch):
    """Steady Richards (src/richards.jl:156-205) on the 2M-cell grid: the
    example's van Genuchten parameters (alpha 0.375, N 1.25 on every edge,
    examples/richards_steadystate_3d/inputdeck.jl:25-26), geometric-mean K_e
    from the 2M logK field, pressure head -1 m on the top face (unsaturated
    near the top, saturated below) and 4 extraction wells of 1e-5 m^3/s at the
    2M injector cells.  One step = forward Newton solve from psi0 = 0 plus the
    rrule's pullback for the heads at the 200 observation cells; inputs
    resident on the GPU."""
    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    w = P.workloads.steady_2m()
    n, ne = w.n, w.nedges
    topo = api.Topology(w.n, w.neighbors, w.areasoverlengths, w.dirichletnodes)
    L = _lib.lib()
    sptr = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    coords, _, _, _ = P.grid.regulargrid3d((0.0, 0.0, 0.0), w.maxs, w.ns)
    nb0 = w.neighbors - 1
    Ks = torch.from_numpy(np.exp(0.5 * (w.logK[nb0[:, 0]] + w.logK[nb0[:, 1]]))).to(dev)
    dpsi = torch.from_numpy(np.where(np.isin(np.arange(n), w.dirichletnodes - 1), -1.0, 0.0)).to(dev)
    z = torch.from_numpy(np.ascontiguousarray(coords[2])).to(dev)
    al = torch.full((ne,), 0.375, dtype=torch.float64, device=dev)
    Nv = torch.full((ne,), 1.25, dtype=torch.float64, device=dev)
    Q = torch.from_numpy(np.where(w.Qs > 0, -1e-5, 0.0)).to(dev)
    psi0 = torch.zeros(n, dtype=torch.float64, device=dev)
    psi = torch.empty(n, dtype=torch.float64, device=dev)
    delta = torch.zeros(n, dtype=torch.float64, device=dev)
    delta[torch.from_numpy(w.obs - 1).to(dev)] = 1.0
    dK = torch.empty(ne, dtype=torch.float64, device=dev)
    def p(t):
        return ctypes.c_void_p(t.data_ptr())

    def run(precision):
        o = api._opts({"reltol": 1e-10, "precision": precision}, memory=_lib.DPF_DEVICE,
                      stream=sptr)

        def step():
            st, sh = _lib.Stats(), ctypes.c_void_p()
            _lib.check(L.dpf_richards_steady(topo.handle, p(psi0), p(Ks), p(dpsi), p(z), p(al),
                                             p(Nv), p(Q), p(psi), ctypes.byref(o),
                                             ctypes.byref(st), ctypes.byref(sh)))
            st2 = _lib.Stats()
            code = L.dpf_richards_pullback(sh, p(delta), p(dK), None, None, None, None,
                                           ctypes.byref(o), ctypes.byref(st2))
            L.dpf_state_destroy(sh)
            _lib.check(code)
            return st, st2
        for _ in range(max(args.warmup, 1)):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stats = [step() for _ in range(args.steps)]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        a, b = stats[-1]
        log(f"[bench] richards {precision}: {args.steps / dt:.2f} forward+gradient evals/s "
            f"({1e3 * dt / args.steps:.0f} ms), Newton steps {a.outer}, BiCGStab iterations "
            f"{a.iterations} forward / {b.iterations} adjoint, relres {a.relres:.1e}")
        return args.steps / dt, a, b, psi.clone(), dK.clone()

    r64, a, b, psi64, dK64 = run("fp64")
    r32, a32, b32, psi32, dK32 = run("mixed_pc")
    frac = float((psi64 < 0).double().mean())
    rel = lambda x, y: float(torch.linalg.norm(x - y) / torch.linalg.norm(y))  # noqa: E731
    log(f"[bench] richards: {100 * frac:.1f} % of cells unsaturated; FP32 V-cycle vs FP64: "
        f"heads {rel(psi32, psi64):.1e}, dK {rel(dK32, dK64):.1e}")
    res = {"metric": "steady Richards forward+gradient evals/s, 2M cells",
           "value": round(r64, 3), "unit": "evals/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
           "data": "synthetic", "dtype": "f64",
           "config": {"workload": "richards 2M (200x200x50), alpha 0.375, N 1.25, psi = -1 on "
                                  "the top face, 4 extraction wells of 1e-5 m^3/s"},
           "newton_steps": a.outer, "bicgstab_forward": a.iterations,
           "bicgstab_adjoint": b.iterations, "unsaturated_fraction": round(frac, 4),
           "fp32_vcycle_evals_per_s": round(r32, 3),
           "fp32_vcycle_bicgstab_fwd_adj": [a32.iterations, b32.iterations],
           "fp32_vcycle_rel_err_head_grad": [rel(psi32, psi64), rel(dK32, dK64)]}
    print(json.dumps(res), flush=True)


def bench_barrier(args, _lib):
    """Cost of one grid-wide barrier inside a cooperative launch
    (dpf_profile_grid_barrier), the alternative to a kernel boundary between
    dependent phases of a persistent kernel: workgroups taking part on every
    XCD or on one, 1000 rounds of two barriers each."""
    L = _lib.lib()
    rows = []
    for one_xcd in (0, 1):
        for nparts in (8, 32, 64, 128, 256):
            us = ctypes.c_double()
            errs = ctypes.c_int32()
            code = L.dpf_profile_grid_barrier(nparts, one_xcd, 1000, ctypes.byref(us),
                                              ctypes.byref(errs))
            if code != 0:
                log(f"[bench] barrier nparts {nparts} one_xcd {one_xcd}: "
                    f"{L.dpf_last_error().decode()}")
                continue
            rows.append({"workgroups": nparts, "one_xcd": bool(one_xcd),
                         "us_per_barrier": round(us.value, 3), "errors": errs.value})
            log(f"[bench] barrier: {nparts} workgroups {'on one XCD' if one_xcd else 'on all XCDs'}:"
                f" {us.value:.3f} us per barrier, errors {errs.value}")
    print(json.dumps({"metric": "grid barrier cost", "unit": "us", "higher_is_better": False,
                      "rows": rows}), flush=True)


def bench_transport(args, P, api, _lib, torch):
    """Advection-dispersion (src/transport.jl) on the 2M-cell grid: velocity
    (1, 0.5, 0.1) m/s and dispersion 1e-3 m^2/s on every edge, a Gaussian
    plume of 10 m radius, u = 0 on the top face; 50 backward-Euler steps over
    10 s, then the gradient of the concentrations at the 200 observation cells
    in the velocities, dispersion and sources.  Inputs resident on the GPU."""
    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    w = P.workloads.steady_2m()
    n, ne = w.n, w.nedges
    topo = api.Topology(w.n, w.neighbors, w.areasoverlengths, w.dirichletnodes)
    L = _lib.lib()
    sptr = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    coords, _, _, vol = P.grid.regulargrid3d((0.0, 0.0, 0.0), w.maxs, w.ns)

    def d(a):
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    c = [d(coords[k]) for k in range(3)]
    vx, vy, vz = d(np.full(ne, 1.0)), d(np.full(ne, 0.5)), d(np.full(ne, 0.1))
    Ds = d(np.full(ne, 1e-3))
    ud, Q, V = d(np.zeros(n)), d(np.zeros(n)), d(vol)
    r2 = (coords[0] - 30.0) ** 2 + (coords[1] - 30.0) ** 2 + (coords[2] - 12.0) ** 2
    u0 = d(np.where(np.isin(np.arange(n), w.dirichletnodes - 1), 0.0, np.exp(-r2 / 100.0)))
    u1 = torch.empty(n, dtype=torch.float64, device=dev)
    seed = torch.zeros(n, dtype=torch.float64, device=dev)
    seed[torch.from_numpy(w.obs - 1).to(dev)] = 1.0
    g = [torch.empty(ne, dtype=torch.float64, device=dev) for _ in range(4)]
    gq = torch.empty(n, dtype=torch.float64, device=dev)
    o = api._opts({"reltol": 1e-10}, memory=_lib.DPF_DEVICE, stream=sptr)

    def p(t):
        return ctypes.c_void_p(t.data_ptr())
    nsteps = 50

    def step():
        st, sh = _lib.Stats(), ctypes.c_void_p()
        _lib.check(L.dpf_transport_steps(topo.handle, p(u0), p(vx), p(vy), p(vz), p(Ds), p(ud),
                                         p(Q), p(V), p(c[0]), p(c[1]), p(c[2]), 1, 0.0, 10.0,
                                         nsteps, p(u1), None, ctypes.byref(o), ctypes.byref(st),
                                         ctypes.byref(sh)))
        torch.cuda.synchronize()
        tf = time.perf_counter()
        st2 = _lib.Stats()
        code = L.dpf_transport_pullback(sh, p(seed), p(g[0]), p(g[1]), p(g[2]), p(g[3]), p(gq),
                                        ctypes.byref(o), ctypes.byref(st2))
        L.dpf_state_destroy(sh)
        _lib.check(code)
        return st, st2, tf
    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st, st2, tf = step()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    mass0 = float((u0 * V).sum())
    mass1 = float((u1 * V).sum())
    log(f"[bench] transport: {nsteps} backward-Euler steps {tf - t0:.3f} s ({st.iterations} "
        f"BiCGStab iterations), gradient {t2 - tf:.3f} s ({st2.iterations}); mass "
        f"{mass0:.4g} -> {mass1:.4g}")
    res = {"metric": "advection-dispersion forward+gradient wall time, 2M cells x 50 backward-"
                     "Euler steps", "value": round(t2 - t0, 3), "unit": "s", "n_gpus": 1,
           "higher_is_better": False, "data": "synthetic", "dtype": "f64",
           "config": {"workload": "transport 2M (200x200x50), v = (1, 0.5, 0.1), D = 1e-3, "
                                  "50 steps over 10 s"},
           "forward_s": round(tf - t0, 3), "gradient_s": round(t2 - tf, 3),
           "bicgstab_forward": st.iterations, "bicgstab_adjoint": st2.iterations}
    print(json.dumps(res), flush=True)


def kernel_table(L, topo, reset=False):
    """The topology's kernel table (dpf_kernel_table) as a list of dicts."""
    from dpfehm_jl_amd import _lib
    cap = 256
    recs = (_lib.KernelRecord * cap)()
    n = ctypes.c_int32()
    _lib.check(L.dpf_kernel_table(topo.handle, recs, cap, ctypes.byref(n), 1 if reset else 0))
    return [dict(kernel=r.kernel.decode(), level=r.level, w=r.value_bytes, launches=r.launches,
                 ms=r.ms, bytes=r.bytes) for r in recs[:min(n.value, cap)]]


KERNEL_SYMBOL = {"cheb_sweep": "k_cheb", "residual": "k_residual", "restrict": "k_restrict",
                 "prolong": "k_prolong", "prolong_coarse_solve": "k_prolong_coarse",
                 "cg_pq": "k_cg_pq", "cg_xr": "k_cg_xr", "finish": "k_finish",
                 "outer_residual": "k_outer_residual", "galerkin_ap": "k_prod_gather",
                 "galerkin_rap": "k_prod_gather", "galerkin_ap_rows": "k_ap_rows",
                 "galerkin_rap_rows": "k_rap_rows", "assemble": "k_assemble"}


def kernel_label(r):
    typ = {4: "float", 8: "double"}.get(r["w"], "?")
    sym = KERNEL_SYMBOL.get(r["kernel"], r["kernel"])
    lev = f" level {r['level']}" if r["level"] >= 0 else ""
    return f"{sym}<{typ}> ({r['kernel']}{lev})"


def roofline_from_tables(ktab, results, steps):
    """roofline: the kernel class with the most time per evaluation in the
    mixed path, achieved = its algorithmic bytes / its summed launch time
    (dispatch events, over every launch of `steps` evaluations).
    roofline_eval: B_eval / t_eval, the algorithmic bytes of every launch of
    one evaluation over the wall time of one evaluation in the timed region
    (no event timing there), for both precisions."""
    peak = 8000.0
    mix = sorted(ktab["mixed"], key=lambda r: -r["ms"])
    dom = mix[0]
    ach = dom["bytes"] / (dom["ms"] * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": kernel_label(dom), "achieved": round(ach, 1),
            "peak": peak, "unit": "GB/s", "frac": round(ach / peak, 4),
            "traffic": traffic_per_launch(dom),
            "us_per_launch": round(1e3 * dom["ms"] / dom["launches"], 2),
            "bytes_per_launch": round(dom["bytes"] / dom["launches"]),
            "launches_per_eval": round(dom["launches"] / steps, 1),
            "share_of_kernel_time": round(dom["ms"] / sum(r["ms"] for r in mix), 4)}
    evalroof = {}
    for prec in ("mixed", "fp64"):
        tab = ktab[prec]
        b_eval = sum(r["bytes"] for r in tab) / steps
        k_eval = sum(r["ms"] for r in tab) / steps * 1e-3
        t_eval = results[prec]["seconds"] / steps
        evalroof[prec] = {"B_eval_bytes": round(b_eval), "t_eval_s": round(t_eval, 6),
                          "achieved": round(b_eval / t_eval / 1e9, 1),
                          "frac": round(b_eval / t_eval / 1e9 / peak, 4),
                          "kernel_busy_s": round(k_eval, 6),
                          "achieved_over_kernel_busy": round(b_eval / k_eval / 1e9, 1)}
    tables = {}
    for prec in ("mixed", "fp64"):
        rows = []
        for r in sorted(ktab[prec], key=lambda r: -r["ms"])[:20]:
            rows.append({"kernel": kernel_label(r),
                         "launches_per_eval": round(r["launches"] / steps, 1),
                         "us_per_eval": round(1e3 * r["ms"] / steps, 1),
                         "us_per_launch": round(1e3 * r["ms"] / r["launches"], 2),
                         "bytes_per_launch": round(r["bytes"] / r["launches"]),
                         "traffic_per_launch": traffic_per_launch(r),
                         "GBps": round(r["bytes"] / (r["ms"] * 1e-3) / 1e9, 1),
                         "frac": round(r["bytes"] / (r["ms"] * 1e-3) / 1e9 / peak, 3)})
        tables[prec] = rows
    table = tables["mixed"]
    for r in table[:12]:
        log(f"[bench] {r['kernel']:55s} {r['launches_per_eval']:6.1f}/eval "
            f"{r['us_per_eval']:8.1f} us/eval {r['us_per_launch']:7.2f} us {r['GBps']:7.1f} GB/s")
    for prec, e in evalroof.items():
        log(f"[bench] {prec}: B_eval {e['B_eval_bytes'] / 1e9:.2f} GB, t_eval {1e3 * e['t_eval_s']:.2f} ms "
            f"-> {e['achieved']:.0f} GB/s ({e['frac']:.3f} of peak); kernel busy "
            f"{1e3 * e['kernel_busy_s']:.2f} ms")
    return roof, evalroof, tables


def traffic_per_launch(rec):
    """Measured HBM bytes per launch of a kernel class (FETCH_SIZE + WRITE_SIZE
    of rocprofv3 --pmc, scaled by the same counters on a plain copy of known
    size) from the committed profiles/hbm_traffic_2m.json, or None if that
    kernel has no measurement there."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_2m.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    key = f"{rec['kernel']}/{rec['level']}/{rec['w']}"
    ent = t.get("kernels", {}).get(key)
    return ent.get("bytes_per_launch") if ent else None


def cpu_baseline(w, fields, h_obs, budget_s):
    """The multi-core FP64 CPU restatement (oracle/amg_cpu.cpp: the same
    discretisation, SA-AMG-PCG to 1e-10, OpenMP on every host core), not the
    reference (pure Julia, not installed).  Mesh set-up once, then whole
    forward + adjoint evaluations (map, assembly, hierarchy refresh, both
    solves, gradient) on the same fields as the GPU steps until the time
    budget is spent."""
    from oracle import amg_cpu
    t0 = time.perf_counter()
    cpu = amg_cpu.CpuAmg(w.n, w.neighbors, w.areasoverlengths, w.dirichletnodes)
    t_setup = time.perf_counter() - t0
    n = 0
    t0 = time.perf_counter()
    its = None
    while True:
        J, g, info = cpu.eval(fields[n % len(fields)], w.dirichleths, w.Qs, w.obs, h_obs,
                              reltol=1e-10)
        its = (int(info[0]), int(info[1]))
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 50:
            break
    rate = n / el
    model = amg_cpu.cpu_model()
    log(f"[bench] cpu baseline: {cpu.threads} threads ({model}), mesh set-up {t_setup:.1f}s, "
        f"{n} evals in {el:.1f}s -> {rate:.3f} evals/s, iterations {its}")
    cpu.close()
    return {"value": round(rate, 4), "unit": "evals/s", "cores": cpu.threads, "kind": "port",
            "cpu_model": model,
            "sample": f"{n} whole forward+adjoint evaluations of the same 2M-cell workload and "
                      "fields (logK map, assembly, SA-AMG refresh, forward and adjoint PCG to "
                      "1e-10, gradient) by the FP64 C++ restatement with OpenMP on every host "
                      "core (oracle/amg_cpu.cpp); not the reference, which is Julia and not "
                      "installed; mesh set-up excluded, as on the GPU"}


if __name__ == "__main__":
    main()
