"""Throughput benchmark: items/s of the greedy spectral packer on one MI355X.

Workload (default): 240x123x100 int32 tray, 350 seeded unions of 3-6
ellipsoids/capsules (bounding boxes 20-50 voxels, fill 0.30-0.40), 24
orientations, proximity from the rounded Euclidean distance to all occupied
voxels, FP64 score combine.  A step is one item: its orientation search, the
placement when a feasible offset exists, and the refresh of the tray's
distance field and spectra.

Two loops run over the same K items (default all 350, volume order) from an
empty tray, after W warmup items on a throwaway tray:
  plain    every fitting orientation searched in one batched launch,
  learned  the proposer's top-k orientations searched (full search when its
           confidence is low or no proposal is feasible).
The metric's value is the learned loop's items PLACED per second (items that
fit nowhere are decided too, and count in items_processed_per_s only) when
the learned loop's final density over a whole run of all items is within 0.5
percentage points of the plain loop's, else the plain loop's.  When fewer
items are timed, both loops also run untimed over all items for that gate.

Prints one JSON line with the roofline of the dominant kernel of the reported
loop (HIP events on the library's stream), the whole loop's fraction of the
HBM roofline, and a CPU baseline of the reference CPU path restated in
oracle/ref_cpu_path.py (a prefix packed in full and compared with the GPU's
placements, a 1-thread figure, and the 64^3 boxes workload in full).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "items placed/s @240×123×100 tray, 24 rot; learned vs plain at ≤0.5pp density"
HBM_PEAK_GBS = 8000.0


KERNELS = ["dist_zy", "dist_x", "inc_zy", "inc_x", "inc_stats", "learn_gate", "learn_pick", "tray_slab_fwd", "tray_rows_fwd", "tray_x_fwd", "item_slab_fwd",
           "item_rows_fwd", "pencil_y_fwd", "corr_x", "score_yz", "pencil_y_inv", "rows_score",
           "select", "verify", "decide", "place", "direct_search", "gather_od", "proposer",
           "proposer_items", "record", "spec_add", "feas_begin", "feas_x", "feas_yz", "feas_decide",
           "tray_columns", "tray_fine", "tile_heights", "tray_masks_inc", "core_columns", "item_columns",
           "run_lists", "item_masks", "column_test", "pre_masks", "pre_column_test", "coarse_decide",
           "post"]


def device_sync():
    hip = ctypes.CDLL("libamdhip64.so")
    rc = hip.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed ({rc})")


def kernel_bytes(name: str, Ph: int, V: int, m: float) -> float:
    """Algorithmic HBM bytes of one launch, on the tray size itself.

    Scored search, per orientation: read the two cached complex64 tray
    spectra (16*Ph), write the two products after the x pass (16*Ph) and read
    them back in the yz pass (16*Ph), Ph = Nx*Ny*(Nz//2+1).  The x-pass kernel
    does the first two for a batch of m orientations (the spectra once), the
    yz-pass kernel the third.
    Feasibility pass, per orientation pair (m = pairs): the x pass reads the
    tray spectrum once (16*Ph) and writes one complex64 product per pair
    (16*Ph); the yz pass reads it back.
    Feasibility pairs are counted on the device: pairs the column test
    excluded in both orientations are not transformed.
    """
    if name in ("score_yz", "feas_yz"):
        return 16.0 * Ph * m
    if name in ("corr_x", "feas_x"):
        return 16.0 * Ph + 16.0 * Ph * m
    raise KeyError(name)


def bench_uneven(args) -> int:
    """Independent trays (sides log-uniform in [48, 256], z <= 128), 30-200
    blob items each, 24 orientations, interlocking-free check after every
    placement, all instances packed concurrently (spk_pack_many: a pool of
    host workers, one context and stream per instance).  value = items
    placed per second over all instances; items processed per second is
    reported beside it.  The loop runs twice from empty trays: kernel events
    off (value), then on, one instance after another (kernel table, roofline
    of the scored search)."""
    import psacking_amd as P
    from psacking_amd import workloads as W
    from psacking_amd.rotations import orientation_maps

    ws = [W.uneven_instance(i) for i in range(args.instances)]
    perms, flips = orientation_maps(24)
    lists, orders = [], []
    for w in ws:
        vols = [int((it > 0).sum()) for it in w.items]
        lists.append([np.ascontiguousarray(it.astype(np.int32)) for it in w.items])
        orders.append(np.array([int(i) for i in np.argsort(vols)[::-1]], np.int32))

    def sessions_for(gen):
        ss = []
        for w in ws:
            s = P.Session(0)
            s.set_option("plan", args.plan)
            s.set_option("certificate", args.certificate)
            for kv in args.opt:
                k, v = kv.split("=")
                s.set_option(k, float(v))
            s.set_tray(np.zeros(w.tray_size, np.int32), args.metric, args.sources, gen)
            ss.append(s)
        return ss

    t0 = time.perf_counter()
    sessions = sessions_for(0)
    device_sync()
    t_setup = time.perf_counter() - t0
    if args.warmup > 0:
        nw = [min(args.warmup, len(l)) for l in lists]
        P.pack_many(sessions, [l[:k] for l, k in zip(lists, nw)],
                    [np.arange(k, dtype=np.int32) for k in nw], perms, flips, "cube", True)
        device_sync()

    def run(timing):
        ss = sessions_for(1)
        for s in ss:
            s.timing_reset()
            s.timing(timing)
        # with kernel events the instances run one after another: events on
        # concurrent streams would also time the other streams' kernels
        ss[0].set_option("many_workers", 1 if timing else args.many_workers)
        device_sync()
        t1 = time.perf_counter()
        res = P.pack_many(ss, lists, orders, perms, flips, "cube", True)
        device_sync()
        dt = time.perf_counter() - t1
        for s in ss:
            s.timing(False)
        return ss, res, dt

    sessions, res, dt = run(False)
    n_items = sum(len(l) for l in lists)
    placed = sum(1 for rl in res for r in rl if r["found"])
    fallback = sum(1 for rl in res for r in rl if r["path"] != 0)
    dens = [float((s.tray() > 0).sum()) / s.tray().size for s in sessions]
    print(f"[bench] {len(ws)} instances, {n_items} items ({placed} placed), {dt:.3f} s, "
          f"setup {t_setup:.3f} s, fallbacks {fallback}, mean tray density "
          f"{np.mean(dens):.4f}", file=sys.stderr, flush=True)
    ev_sessions, ev_res, dt_ev = run(True)
    key = lambda r: (r["found"], tuple(r["position"]), r["orientation"], r["score"])
    assert all([key(x) for x in a] == [key(x) for x in b] for a, b in zip(res, ev_res)), \
        "uneven runs with and without kernel events differ"
    kt = {}
    for s in ev_sessions:
        for n in KERNELS + ["ff_seed", "ff_tile", "ff_batch_init", "ff_unpack", "masked_argmin"]:
            ms, cnt = s.timing_get(n)
            a = kt.setdefault(n, [0.0, 0])
            a[0] += ms
            a[1] += cnt
    busy = sum(v[0] for v in kt.values())
    for n, (ms, cnt) in sorted(kt.items(), key=lambda kv: -kv[1][0]):
        if cnt:
            print(f"[kernel:uneven] {n:14s} {ms:9.2f} ms  {cnt:6d} launches  "
                  f"{ms / cnt * 1e3:8.1f} us/launch  {100 * ms / max(busy, 1e-9):5.1f} %",
                  file=sys.stderr)
    # roofline of the scored search's yz pass: algorithmic bytes 16 Ph per
    # searched orientation on each instance's own tray, over its event time
    b_alg = t_alg = 0.0
    n_launch = 0
    for s, w in zip(ev_sessions, ws):
        nx, ny, nz = w.tray_size
        ms, cnt = s.timing_get("score_yz")
        b_alg += 16.0 * nx * ny * (nz // 2 + 1) * s.timing_units("score_yz")
        t_alg += ms / 1e3
        n_launch += cnt
    achieved = b_alg / t_alg / 1e9 if t_alg > 0 else 0.0
    cpu = None
    if args.cpu_items > 0:
        # the oracle's interlocking-free search (cached FP64 spectra, exact
        # labelling flood fill) on the first item of the largest instance,
        # all 24 orientations, on an empty tray
        from oracle import spk_oracle as O
        big = max(range(len(ws)), key=lambda i: int(np.prod(ws[i].tray_size)))
        w = ws[big]
        item = lists[big][int(orders[big][0])]
        t1 = time.perf_counter()
        tray = np.zeros(w.tray_size, np.int32)
        st = O.TrayState(tray, O.distance_field(tray, args.metric, args.sources))
        for g in O.get_orientations(item, 24):
            g = O.as_grid(g)
            if any(g.shape[d] > w.tray_size[d] for d in range(3)):
                continue
            O.search_state(st, g, "cube", O.interlocking_mask_state(st, g))
        sec = time.perf_counter() - t1
        cores = len(os.sched_getaffinity(0))
        if os.environ.get("OMP_NUM_THREADS", "").isdigit():
            cores = max(1, min(cores, int(os.environ["OMP_NUM_THREADS"])))
        cpu = {"value": 1.0 / sec, "unit": "items/s", "cores": cores, "kind": "port",
               "sample": (f"the oracle's interlocking-free search (FP64 scipy FFT on cached "
                          f"spectra, labelling flood fill) of the first item of instance {big} "
                          f"({'x'.join(map(str, w.tray_size))}), 24 orientations, on an empty "
                          f"tray: {sec:.1f} s for one item")}
        print(f"[cpu] {cpu}", file=sys.stderr, flush=True)
    out = {
        "metric": f"items placed/s over {len(ws)} concurrent uneven trays, 24 rot, interlocking-free",
        "value": placed / dt,
        "unit": "items/s",
        "n_gpus": 1,
        "steps": n_items,
        "warmup": args.warmup,
        "ms_per_step": dt / n_items * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": (f"uneven: {len(ws)} trays (seeds 4..{3 + len(ws)}), sides log-uniform "
                         f"48-256 (z<=128), {n_items} blob items, 24 orientations, "
                         f"interlocking-free, proximity {args.metric}/{args.sources}"),
            "items_placed": placed,
            "items_processed_per_s": n_items / dt,
            "exact_fallbacks": fallback,
            "mean_density": float(np.mean(dens)),
            "kernel_ms": {n: round(v[0], 3) for n, v in kt.items() if v[1]},
        },
        "roofline": {"bound": "hbm", "kernel": "score_yz", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "algorithmic_bytes": b_alg,
                     "avg_launch_us": t_alg / max(1, n_launch) * 1e6,
                     "unit_of_work": "orientation (16 Ph bytes each, Ph of its own tray)"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=350)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="readme",
                    choices=["readme", "blobs", "boxes", "uneven"])
    ap.add_argument("--instances", type=int, default=48,
                    help="uneven workload: number of independent trays")
    ap.add_argument("--many-workers", type=int, default=8,
                    help="uneven workload: host workers of spk_pack_many")
    ap.add_argument("--metric", default="edt", choices=["edt", "l1"])
    ap.add_argument("--sources", default="nonzero", choices=["nonzero", "eq1"])
    ap.add_argument("--cpu-items", type=int, default=2,
                    help="CPU baseline: items of the timed sequence packed in full on the host "
                         "(0 = skip the baseline)")
    ap.add_argument("--cpu-budget", type=float, default=12.0,
                    help="CPU baseline: seconds of 1-thread searches for the 1-thread figure")
    ap.add_argument("--plan", type=int, default=0, choices=[0, 1, 2],
                    help="FFT plan: 0 auto, 1 yz slabs, 2 z rows + y pencils")
    ap.add_argument("--proposer", default="auto", choices=["auto", "plain", "learned"],
                    help="auto: time the plain and the learned loop and report the learned one "
                         "when its density is within 0.5 pp of the plain one")
    ap.add_argument("--certificate", type=int, default=0, choices=[0, 1],
                    help="free-space certificate before the feasibility pass (exact)")
    ap.add_argument("--opt", action="append", default=[], metavar="NAME=VALUE",
                    help="extra context option (spk_set_option), repeatable")
    ap.add_argument("--topk", type=int, default=4)
    ap.add_argument("--min-conf", type=float, default=0.0)
    ap.add_argument("--weights", default=os.path.join(ROOT, "psacking_amd", "weights",
                                                      "proposer.npy"))
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device copies of MIB MiB (HBM counter calibration)")
    args = ap.parse_args(argv)

    if args.copy_calibration:
        import torch
        n = args.copy_calibration << 20
        a = torch.ones(n, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch_read": n, "copy_bytes_per_launch_write": n,
                          "launches": 5}), flush=True)
        return 0

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        return 2
    if args.workload == "uneven":
        return bench_uneven(args)

    import psacking_amd as P
    from psacking_amd import proposer as PR
    from psacking_amd import workloads as W
    from psacking_amd.rotations import orientation_maps

    w = W.by_name(args.workload)
    n_orient = w.num_orientations
    items = [np.ascontiguousarray(it.astype(np.int32)) for it in w.items]
    vols = [int((it > 0).sum()) for it in items]
    order = [int(i) for i in np.argsort(vols)[::-1]]
    K = max(1, min(args.steps, len(items)))
    W_ = max(0, min(args.warmup, len(items)))
    perms, flips = orientation_maps(n_orient)
    penalty = "cube" if n_orient > 1 else "pow"
    Nx, Ny, Nz = w.tray_size
    V = Nx * Ny * Nz
    Ph = Nx * Ny * (Nz // 2 + 1)
    want_plain = args.proposer in ("auto", "plain")
    want_learned = args.proposer in ("auto", "learned") and n_orient > 1
    nn = None
    if want_learned:
        if os.path.exists(args.weights):
            nn = PR.Proposer(0, 64, np.load(args.weights, allow_pickle=False))
        elif args.proposer == "learned":
            raise FileNotFoundError(args.weights)
        else:
            want_learned = False

    def fresh_session():
        s = P.Session(0)
        s.set_option("plan", args.plan)
        s.set_option("certificate", args.certificate)
        for kv in args.opt:
            k, v = kv.split("=")
            s.set_option(k, float(v))
        t0 = time.perf_counter()
        s.set_tray(np.zeros(w.tray_size, np.int32), args.metric, args.sources, 0)
        device_sync()
        return s, time.perf_counter() - t0

    # warmup on a throwaway tray: the first W items of the sequence
    if W_:
        warm = [items[i] for i in order[:W_]]
        if want_plain:
            s, _ = fresh_session()
            s.pack(warm, np.arange(W_, dtype=np.int32), perms, flips, penalty, False, 0)
        if want_learned:
            s, _ = fresh_session()
            PR.pack_learned(s, nn, warm, np.arange(W_), perms, flips, penalty, args.topk,
                            args.min_conf)
        device_sync()

    timed = [items[i] for i in order[:K]]
    runs = {}

    def finish(name, s, dt, res, t_setup):
        t = s.tray()
        kt = {n: s.timing_get(n) for n in KERNELS}
        units = {n: s.timing_units(n) for n in ("corr_x", "score_yz", "feas_x", "feas_yz")}
        # orientations the scored search ran, counted on the device (the
        # column test's pruned orientations and items decided with no FFT
        # are not counted); feasibility pairs run at about a third of one
        runs[name] = {"dt": dt, "res": res, "density": float((t > 0).sum()) / V, "kt": kt,
                      "units": units,
                      "setup": t_setup, "searched": units["score_yz"],
                      "feas_pairs": units["feas_yz"],
                      "placed": sum(1 for r in res if r["found"]),
                      "plan": {"padded": list(s.plan()), "tile": s.tile()[0],
                               "split": bool(s.tile()[1])},
                      "fallbacks": sum(1 for r in res if r["path"] != 0),
                      "max_resid": max(float(r["residual"]) for r in res)}

    # Each loop runs twice on a fresh tray: the timed run with per-kernel HIP
    # events off (the reported value), then an identical run with them on for
    # the kernel table and the roofline.  Both runs must place identically.
    def run_plain(timing):
        s, t_setup = fresh_session()
        s.timing_reset()
        s.timing(timing)
        device_sync()
        t1 = time.perf_counter()
        res = s.pack(timed, np.arange(K, dtype=np.int32), perms, flips, penalty, False, 0)
        device_sync()
        dt = time.perf_counter() - t1
        s.timing(False)
        return s, t_setup, dt, res

    def run_learned(timing):
        s, t_setup = fresh_session()
        s.timing_reset()
        s.timing(timing)
        device_sync()
        t1 = time.perf_counter()
        out = PR.pack_learned(s, nn, timed, np.arange(K), perms, flips, penalty, args.topk,
                              args.min_conf)
        device_sync()
        dt = time.perf_counter() - t1
        s.timing(False)
        return (s, t_setup, dt) + tuple(out)

    def identical(a, b):
        key = lambda r: (r["found"], tuple(r["position"]), r["orientation"], r["score"])
        return [key(x) for x in a] == [key(x) for x in b]

    if want_plain:
        _, _, dt, res0 = run_plain(False)
        s, t_setup, dt_ev, res = run_plain(True)
        assert identical(res0, res), "plain runs with and without kernel events differ"
        finish("plain", s, dt, res, t_setup)
        runs["plain"]["dt_events"] = dt_ev
    if want_learned:
        _, _, dt, res0, _, _ = run_learned(False)
        s, t_setup, dt_ev, res, props, counts = run_learned(True)
        assert identical(res0, res), "learned runs with and without kernel events differ"
        finish("learned", s, dt, res, t_setup)
        runs["learned"]["counts"] = [int(c) for c in counts]
        runs["learned"]["dt_events"] = dt_ev

    for name, r in runs.items():
        busy = sum(v[0] for v in r["kt"].values())
        for n in KERNELS:
            if r["kt"][n][1]:
                print(f"[kernel:{name}] {n:14s} {r['kt'][n][0]:9.2f} ms  {r['kt'][n][1]:6d} launches  "
                      f"{r['kt'][n][0] / max(1, r['kt'][n][1]) * 1e3:8.1f} us/launch  "
                      f"{100 * r['kt'][n][0] / max(busy, 1e-9):5.1f} %", file=sys.stderr)
        print(f"[bench:{name}] {K} items ({r['placed']} placed), {r['dt']:.3f} s = "
              f"{r['placed'] / r['dt']:.1f} items placed/s, {K / r['dt']:.1f} items processed/s "
              f"({r['dt_events']:.3f} s with kernel events), "
              f"kernels busy {busy / 1e3:.3f} s, density "
              f"{100 * r['density']:.2f} %, fallbacks {r['fallbacks']}, max residual "
              f"{r['max_resid']:.4f}" + (f", counts {r['counts']}" if "counts" in r else ""),
              file=sys.stderr, flush=True)

    # Density gate and agreement.  The gate compares the final densities of a
    # whole run over every item: when fewer items are timed, both loops also
    # run untimed over all of them.  Agreement counts only items that at
    # least one loop placed.
    def agreement(a, b):
        rows = [(x, y) for x, y in zip(a, b) if x["found"] or y["found"]]
        same = [x["found"] == y["found"] and tuple(x["position"]) == tuple(y["position"]) and
                x["orientation"] == y["orientation"] for x, y in rows]
        first = next((i for i, v in enumerate(same) if not v), len(same))
        return (float(np.mean(same)) if same else 1.0), first, len(rows)

    agree = None
    if "plain" in runs and "learned" in runs:
        a, b = runs["plain"]["res"], runs["learned"]["res"]
        ident, first, n_rows = agreement(a, b)
        agree = {"placements_identical": ident, "first_divergence": first,
                 "items_placed_by_either": n_rows,
                 "density_plain": runs["plain"]["density"],
                 "density_learned": runs["learned"]["density"],
                 "items_placed_per_s_plain": runs["plain"]["placed"] / runs["plain"]["dt"],
                 "items_placed_per_s_learned": runs["learned"]["placed"] / runs["learned"]["dt"],
                 "speedup": runs["plain"]["dt"] / runs["learned"]["dt"]}
        if K < len(items):
            full = [items[i] for i in order]
            s, _ = fresh_session()
            rp = s.pack(full, np.arange(len(full), dtype=np.int32), perms, flips, penalty,
                        False, 0)
            dp = float((s.tray() > 0).sum()) / V
            s, _ = fresh_session()
            rl = PR.pack_learned(s, nn, full, np.arange(len(full)), perms, flips, penalty,
                                 args.topk, args.min_conf)[0]
            dl = float((s.tray() > 0).sum()) / V
            fi, ff, fn = agreement(rp, rl)
            agree["full_run"] = {"items": len(full),
                                 "placed_plain": sum(1 for r in rp if r["found"]),
                                 "placed_learned": sum(1 for r in rl if r["found"]),
                                 "density_plain": dp, "density_learned": dl,
                                 "placements_identical": fi, "first_divergence": ff}
        else:
            dp, dl = runs["plain"]["density"], runs["learned"]["density"]
        agree["gate_items"] = len(items)
        agree["density_loss_pp"] = 100 * (dp - dl)
        print(f"[bench] learned vs plain: {agree}", file=sys.stderr, flush=True)

    # the metric's value: the learned loop's items placed per second when its
    # density (whole run) is within 0.5 pp of the plain method's, else the
    # plain method's
    if "learned" in runs and (agree is None or agree["density_loss_pp"] <= 0.5):
        used = "learned"
    else:
        used = "plain" if "plain" in runs else "learned"
    r = runs[used]
    cands = ("score_yz", "corr_x", "feas_x", "feas_yz")
    dom = max(cands, key=lambda n: r["kt"][n][0])
    ms_dom, n_dom = r["kt"][dom]
    avg_s = (ms_dom / max(1, n_dom)) / 1e3
    m_avg = r["units"][dom] / max(1, n_dom)
    ph_dom = Ph
    achieved = kernel_bytes(dom, ph_dom, V, m_avg) / avg_s / 1e9 if avg_s > 0 else 0.0
    # measured HBM bytes per launch of this kernel from the committed counter
    # runs (FETCH_SIZE/WRITE_SIZE scaled by a copy kernel), scaled to this
    # run's units per launch
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{w.name}_{Nx}x{Ny}x{Nz}_{used}.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tk = json.load(fh)["kernels"].get(dom)
        if tk and tk["units_per_launch"] > 0:
            traffic = tk["hbm_bytes_per_launch"] * m_avg / tk["units_per_launch"]
    # whole-loop fraction: algorithmic bytes of the work the loop did (every
    # orientation its scored search ran: B_OS = 48 Ph each; every feasibility
    # pair: 16 Ph; every placement: B_place = 48 Ph + 32 V) over the wall time
    # at HBM peak
    b_loop = (48.0 * Ph * r["searched"] + 16.0 * Ph * r["feas_pairs"] +
              (48.0 * Ph + 32.0 * V) * r["placed"])
    loop_frac = b_loop / r["dt"] / (HBM_PEAK_GBS * 1e9)

    cpu = None
    if args.cpu_items > 0:
        cpu = cpu_baseline(args, w, items, order, perms, flips, P)

    out = {
        "metric": METRIC,
        "value": r["placed"] / r["dt"],
        "unit": "items/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W_,
        "ms_per_step": r["dt"] / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": (f"{w.name}: {Nx}x{Ny}x{Nz} int32 tray, {len(items)} seeded items "
                         f"(seed {w.seed}), {n_orient} orientations, proximity "
                         f"{args.metric}/{args.sources}, FP32 FFT + FP64 score"),
            "path": used,
            "value_is": "items placed / wall seconds over the timed items",
            "proposer": (None if used != "learned" else
                         {"topk": args.topk, "min_conf": args.min_conf,
                          "counts": runs["learned"].get("counts"),
                          "weights": os.path.relpath(args.weights, ROOT)}),
            "learned_vs_plain": agree,
            "items_timed": K,
            "items_placed_timed": r["placed"],
            "items_processed_per_s": K / r["dt"],
            "tray_density": r["density"],
            "exact_fallbacks": r["fallbacks"],
            "max_fft_residual": r["max_resid"],
            "setup_s": r["setup"],
            "fft_plan": r["plan"],
            "loop_roofline": {"algorithmic_bytes": b_loop, "orientations_searched": r["searched"],
                              "feasibility_pairs": r["feas_pairs"],
                              "achieved_GBps": b_loop / r["dt"] / 1e9, "frac": loop_frac},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes": kernel_bytes(dom, ph_dom, V, m_avg),
            "avg_launch_us": avg_s * 1e6,
            "units_per_launch": m_avg,
            "unit_of_work": "orientation pair" if dom.startswith("feas") else "orientation",
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def fft_library() -> str:
    """What the CPU baseline's FFT is: FFTW3 when the host has it (the
    reference's CPU branch links it), else pocketfft through scipy.fft."""
    import subprocess
    try:
        ld = subprocess.run(["ldconfig", "-p"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL,
                            text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        ld = ""
    libs = sorted({l.split()[0] for l in ld.splitlines() if "libfftw3" in l})
    if libs:
        return f"FFTW3 present ({', '.join(libs)}) but not bound from Python; pocketfft used"
    return "pocketfft (scipy.fft, complex128); ldconfig lists no libfftw3"


def cpu_baseline(args, w, items, order, perms, flips, P):
    """The reference CPU path restated as written (oracle/ref_cpu_path.py),
    timed on this host:
      * value: a prefix of the timed items (24 orientations, L1 field of
        value-1 voxels, cube penalty, (2N+1)-padded C2C double FFTs with a
        plan per call) packed in full with every host core this process may
        use (OMP_NUM_THREADS caps it, as on the GPU host); its placements are
        compared with the GPU loop on the same items and settings;
      * value_1thread: the same search at one thread, from the orientation
        searches of the prefix's first item that fit a bounded budget;
      * config1: the 64^3 tray with 20 boxes, 1 orientation, packed in full
        with every core, placements compared with the GPU's."""
    from oracle import ref_cpu_path as R
    from psacking_amd import workloads as W
    nproc = len(os.sched_getaffinity(0))
    cores = nproc
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        cores = max(1, min(nproc, int(os.environ["OMP_NUM_THREADS"])))
    n = min(args.cpu_items, len(order))
    pre = [items[i] for i in order[:n]]
    t0 = time.perf_counter()
    cpu_out, cpu_tray = R.pack_as_written(pre, w.tray_size, w.num_orientations, workers=cores)
    dt = time.perf_counter() - t0
    s = P.Session(0)
    s.set_tray(np.zeros(w.tray_size, np.int32), "l1", "eq1", 0)
    g = s.pack(pre, np.arange(n, dtype=np.int32), perms, flips, "cube", False, 0)
    match = all(c[0] == bool(r["found"]) and (not c[0] or (
        tuple(c[1]) == tuple(r["position"]) and c[2] == r["orientation"] and c[3] == r["score"]))
        for c, r in zip(cpu_out, g)) and np.array_equal(s.tray(), cpu_tray)
    # one thread: a bounded sample of the same searches, extrapolated to an item
    sec_item, nsearch, spent = R.time_item_sample(pre[0], np.zeros(w.tray_size, np.int32),
                                                  w.num_orientations, workers=1,
                                                  budget_s=args.cpu_budget)
    # config 1 in full
    b = W.boxes()
    bv = [int((it > 0).sum()) for it in b.items]
    border = [int(i) for i in np.argsort(bv)[::-1]]
    bi = [b.items[i] for i in border]
    t1 = time.perf_counter()
    c1_out, c1_tray = R.pack_as_written(bi, b.tray_size, 1, workers=cores)
    dt1 = time.perf_counter() - t1
    gb = P.BinPacker(b.tray_size, num_orientations=1).pack_voxels(bi, sort_by_volume=False)
    c1_match = np.array_equal(gb.tray, c1_tray) and all(
        c[0] == p.success and (not c[0] or (tuple(c[1]) == tuple(p.position) and c[3] == p.score))
        for c, p in zip(c1_out, gb.placements))
    placed = sum(1 for c in cpu_out if c[0])
    cpu = {"value": placed / dt, "unit": "items/s", "cores": cores, "kind": "port",
           "sample": (f"reference CPU path restated as written ({fft_library()}, C2C double at "
                      f"2N+1, plan per call): the first {n} items of the timed sequence "
                      f"({placed} placed), {w.num_orientations} orientations, L1/eq1 field, "
                      f"packed in full in {dt:.1f} s on {cores} threads (nproc {nproc})"),
           "items_processed_per_s": n / dt,
           "placements_match_gpu": bool(match),
           "value_1thread": 1.0 / sec_item,
           "sample_1thread": (f"{nsearch} of {w.num_orientations} orientation searches of the "
                              f"first item at 1 thread ({spent:.1f} s), extrapolated to one item"),
           "config1": {"workload": "boxes: 64^3 tray, 20 boxes (seed 1), 1 orientation",
                       "items_per_s": len(bi) / dt1,
                       "items_placed": sum(1 for c in c1_out if c[0]),
                       "seconds": dt1, "cores": cores,
                       "placements_match_gpu": bool(c1_match)}}
    print(f"[cpu] {cpu}", file=sys.stderr, flush=True)
    return cpu


if __name__ == "__main__":
    sys.exit(main())
