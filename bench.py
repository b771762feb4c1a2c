#!/usr/bin/env python3
"""Benchmark: unique characteristic sets per second on one MI355X.

Workload (one step = one full ltl_learn call, i.e. one pass of the hot path
over the instance until the minimal formula is found): 500 positive + 500
negative synthetic traces, 5 atoms, lengths U{10..32}, per-atom bit rates
U[0.3, 0.7], labels from the planted G(a -> F b) & ~(c U d); unit costs, max
cost 14, dedup table of 2^28 slots.  The instance is generated (seed 1) and
uploaded before the timed region; each call re-initialises its device state.

Prints one JSON line.  `value` = unique sets stored per second of whole-call
wall time (every set deduplicated and stored with its parents; the last
generated level, which no later pass reads in full, without its words and
summaries: `words_deferred_per_step`; the level below it, when large, without
its words: `words_unwritten_per_step`); `ms_per_step` = time to the minimal
formula.  `roofline` is the dominant kernel's HBM traffic per launch
(rocprofv3 FETCH + WRITE bytes, calibrated on a device copy, from
profiles/traffic.json) over its HIP-event duration against the 8 TB/s HBM
peak (`achieved`, `frac`), with its algorithmic rate beside it (`alg_gbs`:
operand sets counted once per candidate and operand, so re-reads served by
L2 count too); `cpu_baseline` times
the CPU oracle (the reference enumeration restated in C, OpenMP, every set
stored in full) on a bounded sample of the same instance (levels up to cost
9, about 5 s on 16 threads).  The extras (skipped with --no-extra) time
the learned order against the plain one (configs 2-5), the batched launch
against one call per instance (config 4), and an instance of traces longer
than 64 positions (4 words per trace).

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""
import argparse
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "ltl-learning-on-gpus_amd")
HBM_PEAK = 8.0e12
METRIC = "unique char-sets/sec; time-to-min-formula (learned vs plain order)"


def load_package():
    name = "ltl_learning_on_gpus_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def traffic_for(kernel):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    k = t.get("kernels", {}).get(kernel)
    if not k:
        return None
    return int(k["bytes_per_dispatch"])


def cpu_baseline(inst, max_cost, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    oi = O.Instance(inst.num_pos, inst.num_neg, inst.lens, inst.atoms, inst.names)
    t = time.perf_counter()
    r = O.enumerate_(oi, max_cost=max_cost, threads=threads)
    dt = time.perf_counter() - t
    uniques = r["n_stored"] - inst.atoms.shape[0]
    return uniques / dt, dt, uniques


def host_cores():
    """Cores this process may use (affinity, capped by OMP_NUM_THREADS when
    the environment sets it: the GPU box grants its jobs a share of the
    host), and what the host reports."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    use = min(aff, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else aff
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return use, {"nproc": os.cpu_count(), "affinity": aff, "omp_num_threads": omp, "cpu_model": model}


def time_to_min_cpu_gpu(ltl, threads, budget_gb):
    """Configurations 1 and 3 to their minimal formula: the oracle on the host
    (one thread and `threads`) beside ltl_learn (best of 3 calls)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    out = {}
    L = ltl.Learner(budget=int(budget_gb * 1e9))
    try:
        for cfg in (1, 3):
            inst = ltl.generate(cfg, seed=1)
            oi = O.Instance(inst.num_pos, inst.num_neg, inst.lens, inst.atoms, inst.names)
            row = {}
            for th in sorted({1, threads}):
                t = time.perf_counter()
                r = O.enumerate_(oi, max_cost=inst.max_cost, hamming_k=inst.hamming_k, threads=th)
                row[f"cpu_s_{th}_threads"] = round(time.perf_counter() - t, 4)
            g = min((L.learn(inst, max_cost=inst.max_cost, hamming_k=inst.hamming_k, hash_slots_log2=0)
                     for _ in range(3)), key=lambda x: x["seconds"])
            assert (g["status"], g["cost"]) == (r["status"], r["cost"])
            row.update(gpu_s=round(g["seconds"], 4), cost=r["cost"], formula=g["formula"])
            out[f"config{cfg}"] = row
    finally:
        L.close()
    return out


def calibrate_copy(nbytes=1 << 30):
    """One device-to-device copy of nbytes (the runtime's copy kernel): the
    reference reading for rocprofv3's FETCH_SIZE / WRITE_SIZE scale."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"calibrate_copy_bytes": nbytes}), flush=True)


def kernel_kind(name):
    """The library's kernel kind (ltl_get_kernel_stats names) of a kernel
    symbol in a rocprofv3 trace; other kernels by their own name."""
    import re
    if "rocprim" in name:
        return "scan_finish"
    if "copyBuffer" in name:
        return "copy"
    m = re.search(r"k_(\w+?)(<[^>]*>)?\(", name)
    if not m:
        return name[:40]
    k = m.group(1)
    op = re.match(r"<(\d+)", m.group(2) or "")
    op = int(op.group(1)) if op else None
    unary, binary = {1, 4, 5, 6}, {2, 3}
    if k in ("gen1", "gen_xfg", "gen1m", "gen_xfgm"):
        return "gen_unary"
    if k == "gen2":
        return "gen_until" if op == 7 else "gen_and_or"
    if k == "gen2m":
        return "gen_until"
    if k in ("gen2f", "gen2fs", "gen2fm"):  # the fused & | U pass (one word, with successor tests, K words)
        return "gen_binary"
    if k in ("emit", "emitm"):
        return "emit_unary" if op in unary else "emit_and_or" if op in binary else "emit_until"
    if k == "emit_meta":
        return "emit_meta"
    if k in ("decide", "dedup_small", "dedup_lb", "rx_tail"):
        return "decide"
    if k in ("finish", "compact"):
        return "scan_finish"
    if k == "chk1":
        return "check_unary"
    if k == "chk2":
        return "check_and_or"
    if k in ("chkU", "chkUm"):
        return "check_until"
    return k


def pmc_summary(base):
    """Per kernel kind, per dispatch, from rocprofv3 --pmc runs under `base`
    (subfolders fetch, write: FETCH_SIZE / WRITE_SIZE of one bench.py call;
    calf, calw: the same counters for --calibrate-copy; any other subfolder:
    more counters of the bench call, cal<name>: of the copy): HBM bytes
    scaled by the copy's calibration, and the raw counters."""
    import csv
    import glob as _glob

    def load(sub):
        rows = {}
        path = os.path.join(base, sub, "run_counter_collection.csv")
        if not os.path.exists(path):
            return rows
        for r in csv.DictReader(open(path)):
            k = kernel_kind(r["Kernel_Name"])
            d = rows.setdefault(k, {"dispatches": set(), "counters": {}})
            d["dispatches"].add(r["Dispatch_Id"])
            d["counters"][r["Counter_Name"]] = d["counters"].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        return rows

    subs = sorted(os.path.basename(p) for p in _glob.glob(os.path.join(base, "*")) if os.path.isdir(p))
    data = {s: load(s) for s in subs}
    copy_bytes = 1 << 30
    calf = data.get("calf", {}).get("copy", {}).get("counters", {}).get("FETCH_SIZE")
    calw = data.get("calw", {}).get("copy", {}).get("counters", {}).get("WRITE_SIZE")
    fs = copy_bytes / (calf * 1024) if calf else None
    ws = copy_bytes / (calw * 1024) if calw else None
    out = {"calibration": {"copy_bytes": copy_bytes, "fetch_kib": calf, "write_kib": calw,
                           "fetch_scale": round(fs, 4) if fs else None, "write_scale": round(ws, 4) if ws else None},
           "kernels": {}, "copy_counters": {}}
    for sub in subs:
        if sub.startswith("cal") and sub not in ("calf", "calw"):
            c = data[sub].get("copy")
            if c:
                out["copy_counters"].update({k: v / len(c["dispatches"]) for k, v in c["counters"].items()})
    kinds = set()
    for sub in subs:
        if not sub.startswith("cal"):
            kinds.update(data[sub])
    for k in sorted(kinds):
        e = {}
        for sub in subs:
            if sub.startswith("cal") or k not in data[sub]:
                continue
            d = data[sub][k]
            n = len(d["dispatches"])
            e["dispatches"] = n
            for c, v in d["counters"].items():
                e.setdefault("counters", {})[c] = round(v / n, 1)
        cnt = e.get("counters", {})
        if "FETCH_SIZE" in cnt and fs:
            e["fetch_bytes_per_dispatch"] = int(cnt["FETCH_SIZE"] * 1024 * fs)
        if "WRITE_SIZE" in cnt and ws:
            e["write_bytes_per_dispatch"] = int(cnt["WRITE_SIZE"] * 1024 * ws)
        if "fetch_bytes_per_dispatch" in e and "write_bytes_per_dispatch" in e:
            e["bytes_per_dispatch"] = e["fetch_bytes_per_dispatch"] + e["write_bytes_per_dispatch"]
        out["kernels"][k] = e
    return out


def learned_vs_plain(ltl, L, args):
    """Time to the minimal formula, plain vs learned order, at equal cost:
    configuration 2 (steps as above), configuration 3, every configuration-4
    instance (one call each) and the configuration-5 instances of seed 1, up
    to their planted size.  The scorer is trained on the device first
    (outside the timed calls) from logs of configurations 1-3 and of the
    configuration-5 instances of seeds 2 and 3 (held out from the ones
    timed)."""
    import numpy as np
    t = time.perf_counter()
    rows = ltl.collect_logs(learner=L, per_block=256, config2_max_cost=8, config3_max_cost=6)
    rows = np.concatenate([rows, ltl.collect_logs(learner=L, per_block=1024, config1_seeds=(), config2_max_cost=1,
                                                  config3_max_cost=1, config5_seeds=(2, 3))])
    t_log = time.perf_counter() - t
    t = time.perf_counter()
    trained, loss, n_rows = ltl.train_scorer(rows, seed=2, epochs=20)
    t_train = time.perf_counter() - t
    untrained = ltl.Scorer(seed=5)

    def best_of(inst, k, scorer, **opts):
        L.set_scorer(scorer)
        order = ltl.ORDER_PLAIN if scorer is None else ltl.ORDER_LEARNED
        best = None
        for _ in range(k):
            r = L.learn(inst, order=order, **opts)
            if best is None or r["seconds"] < best["seconds"]:
                best = r
        L.set_scorer(None)
        return best

    out = {"scorer": {"log_rows": int(n_rows), "log_s": round(t_log, 3), "train_s": round(t_train, 3),
                      "final_loss": round(float(loss), 6),
                      "trained_on": "plain-order logs of configs 1-3 (seeded) and of the config-5 "
                                    "instances of seeds 2 and 3"}}
    inst2 = ltl.generate(2, seed=1)
    p = best_of(inst2, max(args.steps, 1), None, max_cost=args.max_cost)
    q = best_of(inst2, max(args.steps, 1), trained, max_cost=args.max_cost)
    assert q["status"] == p["status"] and q["cost"] == p["cost"]
    out["config2"] = {"plain_s": round(p["seconds"], 4), "learned_s": round(q["seconds"], 4),
                      "speedup": round(p["seconds"] / q["seconds"], 3), "cost": p["cost"],
                      "plain_formula": p["formula"], "learned_formula": q["formula"]}
    # (configurations 3-5 name no table size: the table starts small and grows
    # with the levels, instead of clearing 2^28 slots per call)
    inst3 = ltl.generate(3, seed=1)
    p = best_of(inst3, 2, None, max_cost=inst3.max_cost, hamming_k=inst3.hamming_k, hash_slots_log2=0)
    q = best_of(inst3, 2, trained, max_cost=inst3.max_cost, hamming_k=inst3.hamming_k, hash_slots_log2=0)
    assert q["status"] == p["status"] and q["cost"] == p["cost"]
    out["config3"] = {"plain_s": round(p["seconds"], 4), "learned_s": round(q["seconds"], 4),
                      "speedup": round(p["seconds"] / q["seconds"], 3), "cost": p["cost"],
                      "hamming_k": inst3.hamming_k, "plain_formula": p["formula"], "learned_formula": q["formula"]}
    rows5, tot_p, tot_l, tot_u, faster, ranks, shares = [], 0.0, 0.0, 0.0, 0, [], []
    for idx in range(args.config5):
        inst = ltl.generate(5, seed=1, index=idx)
        mc5 = args.config5_cost or inst.max_cost
        p = best_of(inst, 2, None, max_cost=mc5, hash_slots_log2=0)
        q = best_of(inst, 2, trained, max_cost=mc5, hash_slots_log2=0)
        u = best_of(inst, 2, untrained, max_cost=mc5, hash_slots_log2=0)
        assert (q["status"], q["cost"]) == (p["status"], p["cost"]) == (u["status"], u["cost"])
        tot_p += p["seconds"]
        tot_l += q["seconds"]
        tot_u += u["seconds"]
        faster += q["seconds"] < p["seconds"]
        rows5.append([idx, mc5, p["status"], p["cost"], round(p["seconds"] * 1e3, 2), round(q["seconds"] * 1e3, 2),
                      round(u["seconds"] * 1e3, 2)])
        if p["status"] == 1 and p["cost"] >= 9:
            # where the ranking can act: the ranked check rounds' share of the
            # device time of a learned-order call (kernel timing on)
            for name, sc in (("trained", trained), ("untrained", untrained)):
                L.set_scorer(sc)
                r = L.learn(inst, order=ltl.ORDER_LEARNED, max_cost=mc5, hash_slots_log2=0, profile=1)
                ks = L.kernel_stats()
                shares.append([idx, name, round(ks["check_ranked"]["seconds"] * 1e3, 2),
                               round(r["kernel_seconds"] * 1e3, 2), r["checked"]])
            L.set_scorer(None)
        if p["status"] == 1 and p["cost"] >= 4:
            rk = ltl.first_hit_ranks(L, inst, {"trained": trained, "untrained": untrained}, min(mc5, 11),
                                     max_rows=1 << 20)
            if rk:
                ranks.append(dict(index=idx, **rk))
    # Configuration 4: every instance, one call each, plain vs learned.
    insts4 = [ltl.generate(4, seed=1, index=i) for i in range(256)]
    mc4 = max(i.max_cost for i in insts4)
    t4p, t4l, faster4, same4, slower4 = 0.0, 0.0, 0, 0, []
    for idx, inst in enumerate(insts4):
        p = best_of(inst, 2, None, max_cost=mc4, hash_slots_log2=0)
        q = best_of(inst, 2, trained, max_cost=mc4, hash_slots_log2=0)
        same4 += (q["status"], q["cost"]) == (p["status"], p["cost"])
        t4p += p["seconds"]
        t4l += q["seconds"]
        faster4 += q["seconds"] < p["seconds"]
        if q["seconds"] >= p["seconds"]:
            slower4.append([idx, p["cost"], round(p["seconds"] * 1e3, 3), round(q["seconds"] * 1e3, 3)])
    out["config4"] = {"instances": len(insts4), "max_cost": mc4, "plain_s": round(t4p, 4),
                      "learned_s": round(t4l, 4), "same_status_and_cost": same4, "learned_faster": faster4,
                      "not_faster": "[index, cost, plain ms, learned ms]", "not_faster_rows": slower4[:32]}
    out["config5"] = {"instances": args.config5, "max_cost": args.config5_cost or "planted size",
                      "plain_s": round(tot_p, 4), "learned_s": round(tot_l, 4), "learned_untrained_s": round(tot_u, 4),
                      "learned_faster": faster,
                      "per_instance_ms": "[index, max cost, status, cost, plain, learned, learned untrained]",
                      "rows": rows5, "first_hit_rank": ranks,
                      "ranked_share": "[index, scorer, check_ranked ms, all kernels ms, candidates checked]",
                      "ranked_share_rows": shares}
    trained.close()
    untrained.close()
    return out


def scorer_throughput(n):
    """Forward and gradient of the scorer over n random feature rows (wall
    time of the C ABI calls, host copies included; kernel times come from
    a kernel trace of this run).  FLOPs per row: 2 * (32*128 + 128*128 + 128)
    forward, twice that again for the backward products."""
    import numpy as np
    ltl = load_package()
    sc = ltl.Scorer(seed=1)
    rng = np.random.default_rng(0)
    x = rng.uniform(0, 1, (n, ltl.SCORER_FEATURES)).astype(np.float32)
    t = rng.uniform(0, 1, n).astype(np.float32)
    sc.forward(x[:1024])
    t0 = time.perf_counter()
    sc.forward(x)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    sc.grad(x, t)
    t_grad = time.perf_counter() - t0
    sc.close()
    fl = 2 * (32 * 128 + 128 * 128 + 128)
    print(json.dumps({"scorer_rows": n, "forward_s": round(t_fwd, 4), "grad_s": round(t_grad, 4),
                      "forward_flop": fl * n, "grad_flop": 3 * fl * n}), flush=True)


def batch_vs_loop(ltl, args):
    """Configuration 4: 256 instances, one call per instance vs the batched
    persistent launch with autotuned parameters."""
    insts = [ltl.generate(4, seed=1, index=i) for i in range(256)]
    mc = max(i.max_cost for i in insts)
    L = ltl.Learner(budget=int(args.budget_gb * 1e9))
    try:
        for i in insts[:8]:
            L.learn(i, max_cost=mc)
        t = time.perf_counter()
        loop = [L.learn(i, max_cost=mc) for i in insts]
        t_loop = time.perf_counter() - t
        trials = L.autotune_batch(insts, max_cost=mc)
        best = L.batch_params()
        t = time.perf_counter()
        res = L.learn_batch(insts, max_cost=mc)
        t_batch = time.perf_counter() - t
        st = L.batch_stats()
    finally:
        L.close()
    same = sum(a["cost"] == b["cost"] and a["status"] == b["status"] for a, b in zip(loop, res))
    return {"instances": len(insts), "max_cost": mc, "loop_s": round(t_loop, 4), "batched_s": round(t_batch, 4),
            "loop_inst_per_s": round(len(insts) / t_loop, 1), "batched_inst_per_s": round(len(insts) / t_batch, 1),
            "speedup": round(t_loop / t_batch, 2), "same_status_and_cost": same, "tuned": best,
            "trials": len(trials), "on_device": st["on_device"], "deferred": st["deferred"], "host": st["host"]}


def dc_runs(ltl, args):
    """Divide and conquer (the reference's RandSplit and DetSplit drivers,
    windows of 64 traces) on 16 problems of 300 + 300 traces (4 atoms,
    lengths U{8..24}, labelled by planted formulas of size 9, max cost 5 per
    window): one ltl_learn_dc call per problem vs one ltl_learn_dc_batch call
    for all 16 (windows of all problems learned together, matches on the
    device).  Wall time of the calls."""
    insts = [ltl.generate_custom(900 + k, 300, 300, 4, 8, 24, 0.5, planted_size=9) for k in range(16)]
    L = ltl.Learner(budget=int(args.budget_gb * 1e9) or (16 << 30), max_cost=5, hash_slots_log2=0)
    out = {"problems": len(insts), "traces": insts[0].W, "window": 64, "max_cost": 5}
    try:
        for algo, name in ((ltl.DC_RAND, "rand"), (ltl.DC_DET, "det")):
            L.learn_dc(insts[0], window=64, algo=algo, seed=1)
            t = time.perf_counter()
            one = [L.learn_dc(i, window=64, algo=algo, seed=1) for i in insts]
            t_one = time.perf_counter() - t
            t = time.perf_counter()
            bat = L.learn_dc_batch(insts, window=64, algo=algo, seed=1)
            t_bat = time.perf_counter() - t
            same = sum(a["formula"] == b["formula"] and a["calls"] == b["calls"] for a, b in zip(one, bat))
            out[name] = {"one_by_one_s": round(t_one, 4), "batched_s": round(t_bat, 4),
                         "speedup": round(t_one / t_bat, 2), "same_formula_and_calls": same,
                         "window_calls": sum(a["calls"] for a in one),
                         "mean_formula_cost": round(sum(a["cost"] for a in one) / len(one), 2)}
    finally:
        L.close()
    return out


def long_instance(ltl, seed, P, N, A, lo, hi, planted, pool=4000):
    """P + N traces of U{lo..hi} positions (K = ceil(hi / 64) words each),
    per-atom bit rates U[0.3, 0.7], labelled by `planted` (product host
    evaluator) from a seeded pool."""
    import numpy as np
    rng = np.random.default_rng(seed)
    K = (hi + 63) // 64
    lens = rng.integers(lo, hi + 1, size=pool).astype(np.uint8)
    rates = rng.uniform(0.3, 0.7, size=A)
    bits = rng.random((A, pool, 64 * K)) < rates[:, None, None]
    bits &= np.arange(64 * K)[None, None, :] < lens[None, :, None].astype(np.int64)
    shifts = np.arange(64, dtype=np.uint64)
    words = (bits.reshape(A, pool, K, 64).astype(np.uint64) << shifts).sum(axis=3, dtype=np.uint64)
    every = ltl.Instance(pool, 0, lens, words.reshape(A, pool * K), K=K)
    holds = (ltl.eval_formula(every, planted)[::K] & np.uint64(1)).astype(bool)
    pi, ni = np.nonzero(holds)[0][:P], np.nonzero(~holds)[0][:N]
    idx = np.concatenate([pi, ni])
    return ltl.Instance(len(pi), len(ni), lens[idx], words[:, idx, :].reshape(A, len(idx) * K), K=K)


def long_traces(ltl, args):
    """Traces longer than 64 positions (K = 4 words per trace): 100 + 100
    traces of U{65..250} positions, 5 atoms, config-2's planted formula."""
    inst = long_instance(ltl, 1, 100, 100, 5, 65, 250, "(G((~(a))|(F(b))))&(~((c)U(d)))")
    L = ltl.Learner(budget=int(args.budget_gb * 1e9), hash_slots_log2=28, max_cost=14, neg_type=1)
    try:
        L.learn(inst)
        L.learn(inst, profile=1)  # every kernel kind timed (a barrier per launch), untimed
        stats = L.kernel_stats()
        best = None
        for _ in range(3):
            t = time.perf_counter()
            r = L.learn(inst)
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
        dt = best
    finally:
        L.close()
    dom = max(stats, key=lambda k: stats[k]["seconds"])
    d = stats[dom]
    return {"workload": "100+100 traces, 5 atoms, lengths U{65..250} (4 words per trace), per-atom rate "
                        "U[0.3,0.7], planted G(a->F b)&~(c U d), unit costs, max cost 14, seed 1, plain order",
            "words_per_trace": inst.K, "value": round((r["n_stored"] - 5) / dt, 1), "unit": "unique CS/s",
            "seconds": round(dt, 4), "seconds_is": "best of 3 calls, no kernel timing",
            "min_cost": r["cost"], "formula": r["formula"],
            "uniques": r["n_stored"], "words_deferred": r["words_deferred"], "words_unwritten": r["words_unwritten"],
            "store_overflow": bool(r["on_the_fly"]),
            "kernels_ms": {k: round(v["seconds"] * 1e3, 3) for k, v in stats.items() if v["seconds"] > 0},
            # (the algorithmic byte model: operand sets counted once per
            # candidate and operand, re-reads served by L2 included, so the
            # fraction of the 8 TB/s peak can exceed 1; no traffic counters)
            "dominant": {"kernel": dom, "alg_gbs": round(d["alg_bytes"] / d["seconds"] / 1e9, 1),
                         "alg_frac": round(d["alg_bytes"] / d["seconds"] / HBM_PEAK, 4)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--budget-gb", type=float, default=0.0, help="device pool; 0 = 75%% of free memory")
    ap.add_argument("--max-cost", type=int, default=14)
    ap.add_argument("--cpu-cost", type=int, default=9, help="levels timed for the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0 = the cores this process may use (affinity, capped by OMP_NUM_THREADS)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the learned-order and batch measurements")
    ap.add_argument("--config5", type=int, default=16, help="configuration-5 instances compared")
    ap.add_argument("--config5-cost", type=int, default=0, help="0 = each instance's planted size")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run one 1 GiB device copy (PMC calibration)")
    ap.add_argument("--scorer-rows", type=int, default=0,
                    help="only run the scorer's forward and gradient on this many random rows")
    ap.add_argument("--dc", action="store_true", help="only run the divide-and-conquer comparison")
    ap.add_argument("--long-traces", action="store_true",
                    help="only run the instance of traces longer than 64 positions")
    ap.add_argument("--pmc-summary", metavar="DIR",
                    help="only summarize rocprofv3 --pmc runs under DIR (see pmc_summary) as JSON")
    args = ap.parse_args()
    if args.pmc_summary:
        print(json.dumps(pmc_summary(args.pmc_summary), indent=1))
        return
    if args.long_traces:
        print(json.dumps(long_traces(load_package(), args)), flush=True)
        return
    if args.dc:
        print(json.dumps(dc_runs(load_package(), args)), flush=True)
        return
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.scorer_rows:
        scorer_throughput(args.scorer_rows)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    ltl = load_package()
    inst = ltl.generate(2, seed=1)
    L = ltl.Learner(budget=int(args.budget_gb * 1e9), hash_slots_log2=28, max_cost=args.max_cost, neg_type=1)
    # Every kernel kind timed once, before the timed steps (profile 1: a
    # dispatch barrier per launch, about 5 ms per call), to find the dominant
    # kind; inside the timed steps HIP events bracket that kind's launches
    # only (profile 3), about 16 launches per call.
    L.learn(inst)  # (the first call reserves the device pool)
    t = time.perf_counter()
    L.learn(inst, profile=1)
    wall_profiled = time.perf_counter() - t
    kinds = L.kernel_stats()
    all_kinds = {k: round(v["seconds"] * 1e3, 3) for k, v in kinds.items() if v["seconds"] > 0}
    dom = max(all_kinds, key=lambda k: all_kinds[k])
    prof = dict(profile=3, profile_kinds=1 << list(kinds).index(dom))
    for _ in range(args.warmup):
        L.learn(inst, **prof)
    stats = {}
    wall, uniques, results = 0.0, 0, []
    for _ in range(args.steps):
        t = time.perf_counter()
        r = L.learn(inst, **prof)
        wall += time.perf_counter() - t
        uniques += r["n_stored"] - inst.atoms.shape[0]
        results.append(r)
        for k, v in L.kernel_stats().items():
            s = stats.setdefault(k, dict(launches=0, seconds=0.0, alg_bytes=0, min_bytes=0))
            for f in s:
                s[f] += v[f]
    r0 = results[0]
    assert all(r["formula"] == r0["formula"] and r["cost"] == r0["cost"] for r in results)
    d = stats[dom]
    achieved = d["alg_bytes"] / d["seconds"] / 1e9
    traffic = traffic_for(dom)
    avg_s = d["seconds"] / max(d["launches"], 1)
    # achieved / frac: the dominant kernel's HBM traffic per launch (calibrated
    # rocprofv3 FETCH_SIZE + WRITE_SIZE, profiles/traffic.json) over this run's
    # average launch, a physical fraction of the 8 TB/s peak.  The algorithmic
    # byte model (operand sets counted once per candidate and operand, so
    # re-reads served by L2 count too) is reported beside it.
    alg_per = d["alg_bytes"] / max(d["launches"], 1)
    phys = traffic / avg_s if traffic else None
    roof = {"bound": "hbm", "kernel": dom,
            "achieved": round(phys / 1e9, 1) if phys else None, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": round(phys / HBM_PEAK, 4) if phys else None,
            "traffic": traffic,
            "avg_launch_ms": round(avg_s * 1e3, 4),
            "alg_bytes_per_launch": int(alg_per),
            "alg_gbs": round(achieved, 1),
            "compulsory_gbs": round(d["min_bytes"] / d["seconds"] / 1e9, 1)}
    out = {
        "metric": METRIC, "value": round(uniques / wall, 1), "unit": "unique CS/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": "config2: 500+500 traces, 5 atoms, lengths U{10..32}, per-atom rate "
                               "U[0.3,0.7], planted G(a->F b)&~(c U d), unit costs, max cost 14, "
                               "2^28 dedup slots, seed 1, plain order",
                   "traces": inst.W, "atoms": int(inst.atoms.shape[0]), "max_cost": args.max_cost},
        "time_to_min_formula_s": round(wall / args.steps, 4),
        "min_cost": r0["cost"], "formula": r0["formula"], "uniques_per_step": r0["n_stored"], "words_deferred_per_step": r0["words_deferred"],
        "words_unwritten_per_step": r0["words_unwritten"],
        # sets stored with their words and summaries (the deferred last level,
        # stored with parents only, and the level below it when stored without
        # words, left out), per second of call time
        "stored_full_per_s": round((uniques - (r0["words_deferred"] + r0["words_unwritten"]) * args.steps) / wall, 1),
        "candidates_per_step": r0["candidates"], "checked_per_step": r0["checked"],
        # (timed steps: the dominant kind only; every kind: one profiled call)
        "kernels_ms_per_step": {dom: round(d["seconds"] / args.steps * 1e3, 3)},
        "kernels_ms_all_kinds": all_kinds, "ms_all_kinds_profiled": round(wall_profiled * 1e3, 3),
        "roofline": roof,
    }
    L.close()
    if not args.no_extra:
        L = ltl.Learner(budget=int(args.budget_gb * 1e9), hash_slots_log2=28, max_cost=args.max_cost, neg_type=1)
        out["learned_vs_plain"] = learned_vs_plain(ltl, L, args)
        L.close()
        out["config4_batch"] = batch_vs_loop(ltl, args)
        out["divide_and_conquer"] = dc_runs(ltl, args)
        out["long_traces"] = long_traces(ltl, args)
    if not args.no_cpu:
        avail, host = host_cores()
        threads = args.cpu_threads or avail
        v, dt, u = cpu_baseline(inst, args.cpu_cost, threads)
        out["cpu_baseline"] = {"value": round(v, 1), "unit": "unique CS/s", "cores": threads,
                               "kind": "port", "seconds": round(dt, 3),
                               "sample": f"same instance, levels up to cost {args.cpu_cost} "
                                         f"({u} unique sets), exact dedup, OpenMP",
                               "host": host}
        # one thread, on a smaller sample (levels up to cost 8, about 10 s)
        v1, dt1, u1 = cpu_baseline(inst, min(args.cpu_cost, 8), 1)
        out["cpu_baseline"]["single_thread"] = {"value": round(v1, 1), "cores": 1, "seconds": round(dt1, 3),
                                                "sample": f"levels up to cost {min(args.cpu_cost, 8)} "
                                                          f"({u1} unique sets)"}
        out["cpu_baseline"]["time_to_min"] = time_to_min_cpu_gpu(ltl, threads, args.budget_gb or 8)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
