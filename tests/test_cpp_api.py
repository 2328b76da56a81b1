"""The C++ front end (include/usap/graph_p.hpp) compiles against the C-ABI
library and behaves like the reference's Graph_P: construct from a TSV
prefix, partition(), get_partitions()."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT


@pytest.fixture(scope="module")
def cpp_bin(tmp_path_factory):
    from usap_amd import _lib
    out = tmp_path_factory.mktemp("cpp") / "graph_p_main"
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "graph_p_main.cpp"), "-o", str(out),
                    "-L", libdir, "-lusap", f"-Wl,-rpath,{libdir}"], check=True)
    return str(out)


def write_tsv(path, g):
    src, dst = g.edges()
    with open(str(path) + ".tsv", "w") as f:
        for s, d in zip(src, dst):
            f.write(f"{s + 1}\t{d + 1}\t1\n")
    with open(str(path) + "_truePartition.tsv", "w") as f:
        for i, b in enumerate(g.truth):
            f.write(f"{i + 1}\t{b + 1}\n")


def test_cpp_front_end_loads(usap, cpp_bin, tmp_path):
    g = usap.config_graph("c1")
    write_tsv(tmp_path / "c1", g)
    r = subprocess.run([cpp_bin, str(tmp_path / "c1"), "load"], capture_output=True, text=True,
                       check=True)
    assert r.stdout.split()[:6] == ["nodes", str(g.N), "edges", str(g.E), "truth", str(g.N)]


@pytest.mark.gpu
def test_cpp_front_end_partitions_like_the_engine(usap, cpp_bin, tmp_path):
    g = usap.config_graph("c1")
    write_tsv(tmp_path / "c1", g)
    r = subprocess.run([cpp_bin, str(tmp_path / "c1"), "partition", "5"], capture_output=True,
                       text=True, check=True, timeout=120)
    lines = r.stdout.splitlines()
    B = int(lines[1].split()[1])
    lab = np.array([int(x) for x in lines[2:]], np.int32)
    eng = usap.Engine(g, seed=5)
    lab2, st = eng.partition()
    assert B == st["B"]
    np.testing.assert_array_equal(lab, lab2)
