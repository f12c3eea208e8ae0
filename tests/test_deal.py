"""Host-only checks of the batch workgroup dealer (sparseimaging_amd/csrc/deal.h).

Compiled with g++ into a small harness; no GPU.  The dealer must give every
problem at least one workgroup, never index outside the launch, number each
problem's workgroups 0 .. nb-1 exactly once, and keep a problem on one XCD
(workgroup b runs on XCD b mod 8) when there are at least 8 problems -- also
when the launch size is not a multiple of 8.
"""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("deal") / "deal_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-o", exe, os.path.join(HERE, "native", "deal_check.cpp")],
                   check=True)
    return exe


def _deal(exe, nblk, work):
    inp = f"{nblk} {len(work)} " + " ".join(repr(float(w)) for w in work)
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    if out[0] == "fail":
        return None, None
    nb = [int(t) for t in out[0].split()[1:]]
    blk = [tuple(int(t) for t in ln.split()) for ln in out[1:] if ln.strip()]
    return nb, blk


@pytest.mark.parametrize("nblk", [512, 509, 300, 257, 64, 17, 9, 8, 5])
@pytest.mark.parametrize("npb", [1, 3, 7, 8, 9, 36, 61, 64])
def test_deal_invariants(harness, nblk, npb):
    rng = np.random.default_rng(nblk * 100 + npb)
    work = rng.uniform(1.0, 10.0, npb)
    nb, blk = _deal(harness, nblk, work)
    if npb > nblk:
        assert nb is None
        return
    assert len(blk) == nblk and len(nb) == npb
    assert min(nb) >= 1 and sum(nb) <= nblk
    seen = {}
    for b, (p, j) in enumerate(blk):
        if p < 0:
            continue
        assert 0 <= p < npb and 0 <= j < nb[p]
        seen.setdefault(p, set()).add(j)
        if npb >= 8:
            assert b % 8 == min(bb for bb, (pp, _) in enumerate(blk) if pp == p) % 8, "problem spans XCDs"
    for p in range(npb):
        assert seen[p] == set(range(nb[p]))


def test_deal_proportional(harness):
    """Within an XCD the split follows the work per iteration."""
    work = [1.0] * 8 + [7.0] * 8
    nb, blk = _deal(harness, 512, work)
    assert sum(nb) == 512
    assert min(nb[8:]) > max(nb[:8])
