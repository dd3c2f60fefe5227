"""Write the oracle's placements at the large workloads as test fixtures.

    python tests/golden/make_large_fixtures.py      (about 15 minutes, 8 cores)

Runs the CPU oracle (oracle/spk_oracle.py, greedy loop on cached FP64
spectra) once, here, so the GPU tests only compare against stored data:

  readme   240x123x100 tray, the first 8 items of the seeded 350-item
           workload (seed 3) in descending-volume order, 24 orientations,
           proximity edt/nonzero and l1/eq1;
  uneven   three of the 48 uneven instances (seeds 4 + i): the 200x211x128
           tray (i = 24) and two small ones (i = 21, 40), the first 6 items of
           each in descending-volume order, 24 orientations,
           interlocking-free, proximity edt/nonzero.

Each run stores the order, the items' SHA-256 (to catch a change of the
generator), every outcome (found, position, orientation, score) and the
SHA-256 and occupied count of the final tray.  Output:
tests/golden/large_fixtures.json.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

README_ITEMS = 8
UNEVEN = (24, 21, 40)
UNEVEN_ITEMS = 6


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.int32).tobytes()).hexdigest()


def volume_order(items):
    vols = [int((it > 0).sum()) for it in items]
    return [int(i) for i in np.argsort(vols)[::-1]]


def run(name, tray_size, items, order, n_orient, interlocking_free, metric, sources):
    from oracle import spk_oracle as O
    t0 = time.time()
    sel = [items[i] for i in order]
    r = O.pack_voxels(sel, tray_size, num_orientations=n_orient,
                      interlocking_free=interlocking_free, order=list(range(len(sel))),
                      metric=metric, sources=sources)
    out = []
    for p in r.placements:      # sorted by index into sel = processing order
        out.append([bool(p.success), list(p.position) if p.success else None,
                    int(p.orientation_index) if p.success else -1,
                    float(p.score) if p.success else None])
    print(f"[fixture] {name} {metric}/{sources}: {r.num_placed}/{len(sel)} placed, "
          f"{time.time() - t0:.0f} s", flush=True)
    return {"tray_size": list(tray_size), "n_orient": n_orient,
            "interlocking_free": interlocking_free, "metric": metric, "sources": sources,
            "order": [int(i) for i in order], "item_sha256": [sha(it) for it in sel],
            "outcomes": out, "num_placed": int(r.num_placed),
            "tray_sha256": sha(r.tray), "tray_occupied": int((r.tray > 0).sum())}


def main() -> int:
    from psacking_amd import workloads as W
    fx = {"readme": {}, "uneven": {}}
    w = W.readme()
    order = volume_order(w.items)[:README_ITEMS]
    for metric, sources in (("edt", "nonzero"), ("l1", "eq1")):
        fx["readme"][f"{metric}/{sources}"] = run("readme", w.tray_size, w.items, order, 24,
                                                  False, metric, sources)
    for i in UNEVEN:
        u = W.uneven_instance(i)
        order = volume_order(u.items)[:UNEVEN_ITEMS]
        fx["uneven"][str(i)] = run(u.name, u.tray_size, u.items, order, 24, True, "edt",
                                   "nonzero")
    path = os.path.join(HERE, "large_fixtures.json")
    with open(path, "w") as fh:
        json.dump(fx, fh, indent=1)
    print(f"[fixture] wrote {path}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
