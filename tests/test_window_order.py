"""The semi-external window order (csrc/window_order.hpp), compiled with g++ and checked on
patterns of cached and streamed windows: a permutation, a cached window first whenever any is,
each kind in index order, and the streamed windows spread evenly over the pass."""
import os
import random
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "graphee_amd", "csrc")


@pytest.fixture(scope="module")
def order_bin(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("wo") / "window_order")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-Wall", f"-I{CSRC}",
                           os.path.join(HERE, "native", "window_order_main.cpp"), "-o", out])
    return out


def orders(binary, patterns):
    r = subprocess.run([binary, *patterns], capture_output=True, text=True, check=True)
    return [[int(x) for x in line.split()] for line in r.stdout.strip().split("\n")]


def test_window_order_properties(order_bin):
    rng = random.Random(7)
    pats = ["c", "s", "cs", "sc", "ccss", "ssss", "cccc", "sccccccs", "cs" * 20, "c" + "s" * 44]
    pats += ["".join(rng.choice("cs") for _ in range(rng.randint(1, 90))) for _ in range(60)]
    for pat, o in zip(pats, orders(order_bin, pats)):
        n = len(pat)
        assert sorted(o) == list(range(n)), pat                       # a permutation
        if "c" in pat:
            assert pat[o[0]] == "c", pat                              # a cached window first
        for kind in "cs":
            seq = [w for w in o if pat[w] == kind]
            assert seq == sorted(seq), pat                            # each kind in index order
        ns = pat.count("s")
        streamed = 0
        for k, w in enumerate(o, 1):
            streamed += pat[w] == "s"
            assert streamed == (k * ns) // n, (pat, k)                # spread evenly


def test_window_order_budget_one_is_identity(order_bin):
    # every window streamed (a device budget of one ring): the index order
    (o,) = orders(order_bin, ["s" * 17])
    assert o == list(range(17))
