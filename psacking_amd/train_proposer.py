"""Train the orientation proposer on records of the plain packer.

    python -m psacking_amd.train_proposer --records 100000 --epochs 6

1. Item pool: unions of 2-5 ellipsoids (semi-axes U[3,10] voxels, centres in
   a 30^3 box) with the whole shape scaled by s ~ U[1, 1.7], cropped to the
   bounding box, sides <= 64.  Seeded (default 5).
2. Records: runs of the plain greedy loop on an empty 240x123x100 tray, each
   over 600 pool items drawn at random and sorted by volume (largest first),
   24 orientations, proximity from the rounded Euclidean distance field.  For
   every item the device records the tray features it saw and each
   orientation's best score (spk_pack_records); a run stops when a chunk of
   50 items places fewer than 5.  Items with no feasible orientation are not
   records.
   Each record keeps the tray features in a dense device store; the item
   enters as its 24 oriented feature grids.
3. Training: soft-target cross entropy (proposer.soft_targets), Adam, batch
   256, learning rate 1e-3 decayed linearly to 1e-4 over the last 40 % of the
   steps; every step is the library's own forward, backward and Adam kernels.
   The last 5 % of runs are held out: the report gives their loss and how
   often the best orientation is the proposer's top-1 / within its top-4.

Weights go to psacking_amd/weights/proposer.npy (float32, spk_nn_set_params
layout); a JSON report to --report.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

TRAY = (240, 123, 100)


def _pool_item(seed: int) -> np.ndarray:
    from psacking_amd import workloads as W
    rng = np.random.default_rng(seed)
    while True:
        s = rng.uniform(1.0, 1.7)
        g = W.blob(rng, 2, 5, 3.0 * s, 10.0 * s, box=int(round(30 * s)))
        if max(g.shape) <= 64 and g.sum() > 0:
            return g


def item_pool(n: int, seed: int, workers: int) -> list:
    seeds = [seed * 1_000_003 + i for i in range(n)]
    with Pool(workers) as p:
        return p.map(_pool_item, seeds, chunksize=16)


def log(*a):
    print("[train]", *a, flush=True)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--records", type=int, default=100000)
    ap.add_argument("--pool", type=int, default=4000)
    ap.add_argument("--per-run", type=int, default=600)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--workers", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--out", default=os.path.join(HERE, "weights", "proposer.npy"))
    ap.add_argument("--report", default="")
    ap.add_argument("--max-record-seconds", type=float, default=480.0,
                    help="stop generating records after this long")
    ap.add_argument("--max-train-seconds", type=float, default=300.0,
                    help="stop training after the epoch that passes this long")
    args = ap.parse_args(argv)

    t0 = time.time()
    pool = item_pool(args.pool, args.seed, args.workers)
    vols = np.array([int(g.sum()) for g in pool])
    log(f"item pool: {len(pool)} items, volume {vols.mean():.0f} mean, "
        f"sides {np.mean([g.shape for g in pool], axis=0).round(1)} mean, {time.time() - t0:.1f} s")

    # GPU work starts here (after the worker pool is gone)
    import psacking_amd as P
    from psacking_amd import proposer as PR
    from psacking_amd.rotations import orientation_maps

    perms, flips = orientation_maps(24)
    nn = PR.Proposer(0, args.batch, PR.init_params(args.seed))
    # every pool item in its 24 orientations: rows 24*i .. 24*i+23
    ifeat = PR.DeviceArray((len(pool) * PR.N_LOGITS, PR.ITEM_FEATURES), np.uint16)
    for i, g in enumerate(pool):
        nn.item_features_all(g, perms, flips, ifeat, i * PR.N_LOGITS)
    log(f"oriented item features: {time.time() - t0:.1f} s")
    cap = args.records + 50
    tfeat = PR.DeviceArray((cap, PR.TRAY_FEATURES), np.uint16)
    scratch = PR.DeviceArray((50, PR.TRAY_FEATURES), np.uint16)
    rng = np.random.default_rng(args.seed)
    rec_item, rec_score, rec_run = [], [], []
    row = 0
    run = 0
    t1 = time.time()
    last = t1
    while (row < args.records and time.time() - t1 < args.max_record_seconds):
        pick = rng.integers(0, len(pool), args.per_run)
        pick = pick[np.argsort(-vols[pick], kind="stable")]
        s = P.Session(0)
        s.set_tray(np.zeros(TRAY, np.int32), "edt", "nonzero", 0)
        placed = 0
        items_run, scores_run, rows_run = [], [], []
        for c0 in range(0, len(pick), 50):
            chunk = pick[c0:c0 + 50]
            res, _, sc = PR.pack_records(s, [pool[i] for i in chunk], np.arange(len(chunk)),
                                         perms, flips, "cube", id_offset=placed, feats=scratch)
            n_ok = sum(1 for r in res if r["found"])
            placed += n_ok
            keep = [j for j in range(len(chunk)) if np.isfinite(sc[j]).any()]
            keep = keep[:cap - row]
            tfeat.gather_from(scratch, keep, row0=row)
            for j in keep:
                items_run.append(int(chunk[j]))
                scores_run.append(sc[j])
                rows_run.append(row)
                row += 1
            if n_ok < 5 or row >= cap:
                break
        rec_item.append(np.array(rows_run, dtype=np.int64))
        rec_score.append((np.array(items_run, dtype=np.int64),
                          np.array(scores_run, dtype=np.float64).reshape(-1, len(perms))))
        rec_run.append(run)
        run += 1
        if time.time() - last > 20:
            last = time.time()
            log(f"records {row} from {run} runs, {time.time() - t1:.0f} s")
    n_runs = len(rec_item)
    n_val = max(1, n_runs // 20)
    def cat(runs):
        rows = np.concatenate([rec_item[r] for r in runs])
        items = np.concatenate([rec_score[r][0] for r in runs])
        scores = np.concatenate([rec_score[r][1] for r in runs])
        return rows, items, scores
    tr_rows, tr_items, tr_scores = cat(range(n_runs - n_val))
    va_rows, va_items, va_scores = cat(range(n_runs - n_val, n_runs))
    t_rec = time.time() - t1
    log(f"records: {len(tr_rows)} train + {len(va_rows)} held out from {n_runs} runs, "
        f"{t_rec:.0f} s")

    q_tr, m_tr = PR.soft_targets(tr_scores)
    q_va, m_va = PR.soft_targets(va_scores)
    dq_tr = PR.DeviceArray.from_numpy(q_tr)
    dm_tr = PR.DeviceArray.from_numpy(m_tr)
    B = args.batch
    bi = PR.DeviceArray((B * PR.N_LOGITS, PR.ITEM_FEATURES), np.uint16)
    orow = np.arange(PR.N_LOGITS)

    def item_rows(items):
        return (np.asarray(items)[:, None] * PR.N_LOGITS + orow[None, :]).reshape(-1)
    bt = PR.DeviceArray((B, PR.TRAY_FEATURES), np.uint16)
    bq = PR.DeviceArray((B, PR.N_LOGITS), np.float32)
    bm = PR.DeviceArray((B,), np.int32)

    def evaluate():
        losses, top1, top4, n = [], 0, 0, 0
        for b0 in range(0, len(va_rows), B):
            sl = slice(b0, min(b0 + B, len(va_rows)))
            k = sl.stop - sl.start
            bi.gather_from(ifeat, item_rows(np.pad(va_items[sl], (0, B - k))))
            bt.gather_from(tfeat, np.pad(va_rows[sl], (0, B - k)))
            z = nn.forward(bi, bt, batch=k)
            s = va_scores[sl]
            best = np.argmin(s, axis=1)
            order = np.argsort(-z, axis=1, kind="stable")
            top1 += int((order[:, 0] == best).sum())
            top4 += int((order[:, :4] == best[:, None]).any(axis=1).sum())
            zz = np.where(np.isfinite(s), z, -np.inf)
            mx = zz.max(axis=1, keepdims=True)
            lse = mx[:, 0] + np.log(np.exp(zz - mx).sum(axis=1))
            qq = q_va[sl]
            losses.append(-(qq * np.where(np.isfinite(s), z - lse[:, None], 0)).sum(axis=1))
            n += k
        return float(np.concatenate(losses).mean()), top1 / n, top4 / n

    steps_per_epoch = len(tr_rows) // B
    total = steps_per_epoch * args.epochs
    hist = []
    l0, a1, a4 = evaluate()
    log(f"untrained: held-out loss {l0:.4f}, top-1 {a1:.3f}, top-4 {a4:.3f}")
    hist.append({"epoch": 0, "val_loss": l0, "top1": a1, "top4": a4})
    t2 = time.time()
    step = 0
    for ep in range(args.epochs):
        perm = rng.permutation(len(tr_rows))
        run_loss = []
        for k in range(steps_per_epoch):
            idx = perm[k * B:(k + 1) * B]
            frac = step / max(1, total)
            lr = args.lr if frac < 0.6 else args.lr * (1.0 - (frac - 0.6) / 0.4 * 0.9)
            bi.gather_from(ifeat, item_rows(tr_items[idx]))
            bt.gather_from(tfeat, tr_rows[idx])
            bq.gather_from(dq_tr, idx)
            bm.gather_from(dm_tr, idx)
            run_loss.append(nn.train_step(bi, bt, bq, bm, lr))
            step += 1
            if time.time() - last > 20:
                last = time.time()
                log(f"epoch {ep + 1} step {k}/{steps_per_epoch} loss {np.mean(run_loss[-50:]):.4f}")
        vl, a1, a4 = evaluate()
        hist.append({"epoch": ep + 1, "train_loss": float(np.mean(run_loss)), "val_loss": vl,
                     "top1": a1, "top4": a4})
        log(f"epoch {ep + 1}: train loss {np.mean(run_loss):.4f}, held-out loss {vl:.4f}, "
            f"top-1 {a1:.3f}, top-4 {a4:.3f}, {time.time() - t2:.0f} s")
        if time.time() - t2 > args.max_train_seconds:
            log("training time limit reached")
            break
    t_train = time.time() - t2
    nn.save(args.out)
    log(f"weights -> {args.out}")
    rep = {"records_train": int(len(tr_rows)), "records_heldout": int(len(va_rows)),
           "runs": n_runs, "pool": len(pool), "seed": args.seed, "epochs": args.epochs,
           "batch": B, "steps": step, "record_seconds": t_rec, "train_seconds": t_train,
           "steps_per_second": step / max(t_train, 1e-9), "history": hist}
    if args.report:
        os.makedirs(os.path.dirname(os.path.abspath(args.report)), exist_ok=True)
        with open(args.report, "w") as f:
            json.dump(rep, f, indent=1)
    print(json.dumps({k: v for k, v in rep.items() if k != "history"}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
