"""Writes tests/golden/*.json from the CPU oracle.

The reference (uSAP) cannot be built here (its Taskflow submodule is absent)
and has no tests or golden vectors of its own, so these fixtures are the
oracle's outputs on seeded synthetic graphs.  They pin the oracle against
regressions and let GPU tests check the engine without re-running the oracle.

Run: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import usap_amd  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402


def dump(name, obj):
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(obj, f, separators=(",", ":"))
        f.write("\n")


def c1_fixture(ds_mode):
    g = usap_amd.config_graph("c1")
    o = Oracle(g, threads=8, seed=1, ds_mode=ds_mode)
    lab0, B0 = o.scc()
    o.set_labels(lab0, B0)
    fx0, S0 = o.entropy()
    best, dS = o.block_merge(0)
    assert o.carry_out(B0 // 2) == 0
    lab1 = o.labels()
    batch = o.nodal_batch(0, 250, 0)
    lab2 = o.labels()
    o2 = Oracle(g, threads=8, seed=1, ds_mode=ds_mode)
    o2.set_labels(lab0, B0)
    B, lab, st = o2.partition()
    return {
        "graph": {"config": "c1", "N": g.N, "E": g.E,
                  "out_idx_sum": int(g.out_idx.astype(np.int64).sum())},
        "params": {"seed": 1, "ds_mode": ds_mode, "init": 0},
        "scc": {"B": B0, "labels": lab0.tolist()},
        "entropy": {"data_fx": str(fx0), "S": S0},
        "merge0": {"best_s": best.tolist(), "dS_fx": [str(d) for d in dS],
                   "B_to_merge": B0 // 2, "labels_after": lab1.tolist()},
        "nodal0": {"beg": 0, "end": 250, "stream": 0, "s": batch["s"].tolist(),
                   "dS_fx": [str(d) for d in batch["dS"]], "inf": batch["inf"].tolist(),
                   "accept": batch["accept"].tolist(), "sum_fx": str(batch["sum"]),
                   "moves": batch["moves"], "labels_after": lab2.tolist()},
        "partition": {"B": B, "labels": lab.tolist(), "S": st["S"],
                      "sweeps": st["sweeps"], "batches": st["batches"], "moves": st["moves"]},
    }


def main():
    dump("c1_ref.json", c1_fixture(0))
    dump("c1_exact.json", c1_fixture(1))
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".json")))


if __name__ == "__main__":
    main()
