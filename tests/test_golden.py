"""Golden fixtures (tests/golden/*.json, written by make_golden.py from the
oracle): the oracle must keep reproducing them (CPU) and the GPU engine must
reproduce them through the C-ABI (gpu).  Plus a quality pin of the oracle on
the reference's own GraphChallenge graphs when the reference checkout is
present (it is not on GPU machines)."""
import json
import os

import numpy as np
import pytest

from conftest import ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden")
FIXTURES = ["c1_ref.json", "c1_exact.json"]


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def check_graph(usap, fx):
    g = usap.config_graph(fx["graph"]["config"])
    assert (g.N, g.E) == (fx["graph"]["N"], fx["graph"]["E"])
    assert int(g.out_idx.astype(np.int64).sum()) == fx["graph"]["out_idx_sum"]
    return g


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_reproduces_golden(usap, orc, name):
    fx = load(name)
    g = check_graph(usap, fx)
    p = fx["params"]
    o = orc.Oracle(g, threads=4, seed=p["seed"], ds_mode=p["ds_mode"])
    lab0, B0 = o.scc()
    assert B0 == fx["scc"]["B"] and lab0.tolist() == fx["scc"]["labels"]
    o.set_labels(lab0, B0)
    assert str(o.entropy()[0]) == fx["entropy"]["data_fx"]
    best, dS = o.block_merge(0)
    assert best.tolist() == fx["merge0"]["best_s"]
    assert [str(d) for d in dS] == fx["merge0"]["dS_fx"]
    o.carry_out(fx["merge0"]["B_to_merge"])
    assert o.labels().tolist() == fx["merge0"]["labels_after"]
    nb = fx["nodal0"]
    r = o.nodal_batch(nb["beg"], nb["end"], nb["stream"])
    assert r["s"].tolist() == nb["s"] and r["accept"].tolist() == nb["accept"]
    assert [str(d) for d in r["dS"]] == nb["dS_fx"]
    assert o.labels().tolist() == nb["labels_after"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", FIXTURES)
def test_engine_reproduces_golden(usap, name):
    fx = load(name)
    g = check_graph(usap, fx)
    p = fx["params"]
    eng = usap.Engine(g, seed=p["seed"], ds_mode=p["ds_mode"], init=p["init"])
    eng.set_labels(np.array(fx["scc"]["labels"], np.int32), fx["scc"]["B"])
    assert str(eng.entropy()[0]) == fx["entropy"]["data_fx"]
    best, dS = eng.merge_propose(0)
    assert best.tolist() == fx["merge0"]["best_s"]
    assert [str(d) for d in dS] == fx["merge0"]["dS_fx"]
    eng.merge_apply(fx["merge0"]["B_to_merge"])
    assert eng.labels().tolist() == fx["merge0"]["labels_after"]
    nb = fx["nodal0"]
    r = eng.nodal_batch(nb["beg"], nb["end"], nb["stream"])
    assert r["s"].tolist() == nb["s"] and r["accept"].tolist() == nb["accept"]
    fin = [i == 0 for i in nb["inf"]]
    assert [str(d) for d, f in zip(r["dS"], fin) if f] == [d for d, f in zip(nb["dS_fx"], fin) if f]
    assert eng.labels().tolist() == nb["labels_after"]
    lab, st = eng.partition()
    pt = fx["partition"]
    assert st["B"] == pt["B"] and lab.tolist() == pt["labels"] and st["S"] == pt["S"]
    assert st["sweeps"] == pt["sweeps"] and st["batches"] == pt["batches"]


REF_DATA = "/root/reference/Dataset/static/lowOverlap_lowBlockSizeVar"


@pytest.mark.parametrize("n,pp,pr", [(1000, 0.998, 0.998), (5000, 1.0, 1.0)])
def test_oracle_quality_on_reference_datasets(usap, orc, n, pp, pr):
    """docs/HPEC2023.pdf p.9 reports pairwise precision/recall for these two
    graphs (Low-Low 1K: .998/.998, 5K: 1.000/1.000).  The partitioner is
    stochastic, so the pin is the median over seeds: within 0.02 in exact dS
    mode; the reference's own dS formula (ref mode, which leaves out the
    in-edge updates of the (r,s) corner) lands lower, within 0.1."""
    base = os.path.join(REF_DATA, f"static_lowOverlap_lowBlockSizeVar_{n}_nodes")
    if not os.path.exists(base + ".tsv"):
        pytest.skip("reference datasets not present")
    g = usap.load_tsv(base + ".tsv", base + "_truePartition.tsv")
    for mode, slack in ((1, 0.02), (0, 0.1)):
        precs, recs = [], []
        for seed in range(5):
            o = orc.Oracle(g, threads=8, seed=seed, ds_mode=mode)
            o.set_labels(*o.scc())
            B, lab, st = o.partition()
            assert st["error"] == 0
            prec, rec, _, _ = usap.pairwise_precision_recall(g.truth, lab)
            precs.append(prec)
            recs.append(rec)
        assert np.median(precs) >= pp - slack and np.median(recs) >= pr - slack, (mode, precs, recs)
