"""Pins of the oracle's two graph builders by hand-traced builds.

The expected graphs below were worked out by hand from the reference's rules;
the trace is written beside each case so it can be checked line by line.  Each
check is chosen so that a plausible mistake in the restatement fails it, and
test_planted_mistakes_are_caught builds the oracle with each such mistake
planted (oracle/pann_oracle.c, ORACLE_MUTANT) and shows that some pin here
fails for every one of them.

Vamana (vamana/index.h:63-319):
  * the batch schedule 1, 2, 4, ... up to 0.02 n, then batches of 0.02 n
    (:202-234), checked as numbers;
  * a 7-point build with R = 2, L = 2, alpha = 1.2 and a fixed insertion
    order, whose reverse-edge step hits newsize == R twice (:292-303);
  * robustPrune's `alpha * d(p*, p') <= d(p, p')` at equality (:111).
HNSW (HNSW.hpp):
  * the insert batches ceil(b * base) + 1, capped at 0.02 n (:776-792);
  * a 4-point build with hand-chosen levels, where prune_heuristic meets
    d(c, r) == alpha * d(c, u) (strict `<` keeps c, :455-483), a reverse-edge
    list that reaches exactly m (append, not cut, :966) and one that overflows
    (cut to the m nearest, :967-978).
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

ORACLE_DIR = os.path.join(ROOT, "oracle")


def _lists(adj):
    return [[int(x) for x in adj[i, 1:1 + adj[i, 0]]] for i in range(adj.shape[0])]


# ------------------------------------------------------------------ schedules
def test_vamana_batch_schedule(oracle):
    # n = 7: max batch = (size_t)(0.02 * 7) = 0, which falls back to n (:208-209);
    # 2^0 = 1 <= 7: [0, 1); 2^1 = 2 <= 7: [1, 3); 2^2 = 4 <= 7: [3, min(7, 7)) -> done
    assert oracle.vamana_batches(7) == [(0, 1), (1, 3), (3, 7)]
    # n = 100: max batch 2; [0, 1), [1, 3) while 2^inc <= 2, then [3, 5), [5, 7), ...
    s = oracle.vamana_batches(100)
    assert s[:5] == [(0, 1), (1, 3), (3, 5), (5, 7), (7, 9)]
    assert s[-2:] == [(97, 99), (99, 100)] and len(s) == 2 + 49
    # n = 1000: max batch 20; doubling through 2^4 = 16 <= 20 ([15, 31)), then
    # steps of 20 from 31: [31, 51) ... [971, 991), [991, 1000)
    s = oracle.vamana_batches(1000)
    assert s[:7] == [(0, 1), (1, 3), (3, 7), (7, 15), (15, 31), (31, 51), (51, 71)]
    assert s[-1] == (991, 1000) and len(s) == 5 + 49


def test_vamana_schedule_of_the_gpu_driver_is_the_oracles(oracle):
    from parlayann_amd.vamana import batch_schedule
    for n in [1, 2, 7, 49, 50, 100, 1000, 3001, 123457]:
        assert batch_schedule(n, n) == oracle.vamana_batches(n), n


def test_hnsw_batch_schedule(oracle):
    from parlayann_amd import hnsw
    # n = 4: 0.02 n = 0, taken as 1 (the reference would loop forever): one point a batch
    assert oracle.hnsw_batches(4) == [(1, 2), (2, 3), (3, 4)]
    # n = 1000: limit 20; ends ceil(2 b) + 1: 3, 7, 15, 31, then b + 20
    s = oracle.hnsw_batches(1000)
    assert s[:7] == [(1, 3), (3, 7), (7, 15), (15, 31), (31, 51), (51, 71), (71, 91)]
    assert s[-1] == (991, 1000)
    for n in [2, 4, 49, 50, 1000, 12345]:
        assert hnsw.batch_bounds(n) == oracle.hnsw_batches(n), n


# ---------------------------------------------------------- Vamana hand trace
# 1-D points (squared distances are exact integers), ids 0..6 at
#   x = 0, 9, 3, 6, 1, 12, 4;  insertion order 3, 5, 2, 6, 1, 4, 0;
# R = 2, L = 2, alpha = 1.2, start point 0, limit = n, degree_limit = R.
# With L = 2 the deferral threshold ceil(0.125 * 2) = 1 defers only empty
# candidate lists.  Batches: [3], [5, 2], [6, 1, 4, 0].
#
# batch 1, insert 3 (x=6): frontier (0,36); 0 has no edges -> visited {0}.
#   prune -> [0].  G[3] = [0].  Reverse: 0 gets [3], newsize 1 <= 2: G[0] = [3].
# batch 2 (both search the graph above):
#   insert 5 (x=12): visit 0 -> candidate (3,36); frontier (3,36),(0,144);
#     visit 3: 0 seen.  visited (3,36),(0,144).  prune: take 3; 1.2*d(3,0)=43.2
#     <= 144 drops 0.  -> [3]
#   insert 2 (x=3): visit 0 -> (3,9); frontier (0,9),(3,9); visit 3.
#     visited (0,9),(3,9).  prune: take 0; 1.2*d(0,3)=43.2 > 9 keeps 3 -> [0, 3]
#   G[5] = [3], G[2] = [0, 3].  Reverse groups: 0 <- [2], 3 <- [5, 2].
#   0: newsize 1 + 1 = 2 <= R: append -> G[0] = [2, 3]   (newsize == R)
#   3: newsize 2 + 1 = 3 > R: robustPrune(3, {5, 2} + out {0}):
#     sorted (2,9),(0,36),(5,36); take 2; 1.2*d(2,0)=10.8 <= 36 drops 0;
#     1.2*d(2,5)=97.2 > 36 keeps 5; take 5 -> G[3] = [2, 5]
# batch 3 (all search G[0]=[2,3], G[2]=[0,3], G[3]=[2,5], G[5]=[3]):
#   insert 6 (x=4): visit 0 -> (2,1),(3,4); frontier (2,1),(3,4); visit 2:
#     0, 3 seen (defer, empty); visit 3: 2 seen, 5 at 64 >= cutoff 4.
#     visited (2,1),(3,4),(0,16).  prune: take 2; 3 kept (10.8 > 4); 0 dropped
#     (10.8 <= 16); take 3 -> [2, 3]
#   insert 1 (x=9): visit 0 -> (3,9),(2,36); visit 3 -> (5,9) < cutoff 36;
#     frontier (3,9),(5,9); visit 5: 3 seen.  visited (3,9),(5,9),(0,81).
#     prune: take 3; 5 kept (43.2 > 9); 0 dropped (43.2 <= 81); take 5 -> [3, 5]
#   insert 4 (x=1): visit 0 -> (2,4),(3,25); frontier (0,1),(2,4); visit 2.
#     visited (0,1),(2,4).  prune: take 0; 2 kept (10.8 > 4) -> [0, 2]
#   insert 0 (x=0, the start point itself): visit 0 -> (2,9),(3,36); frontier
#     (0,0),(2,9); visit 2.  visited (0,0),(2,9).  prune with 0's out-edges
#     {2, 3}: sorted (0,0),(2,9),(3,36) after dedup; 0 is p: skipped; take 2;
#     1.2*d(2,3)=10.8 <= 36 drops 3 -> [2]
#   G[6] = [2, 3], G[1] = [3, 5], G[4] = [0, 2], G[0] = [2].
#   Reverse groups (flattened in batch order): 0 <- [4]; 2 <- [6, 4, 0];
#     3 <- [6, 1]; 5 <- [1].
#   0: 1 + 1 = 2 <= R: append -> G[0] = [4, 2]            (newsize == R)
#   2: 3 + 2 > R: robustPrune(2, {6,4,0} + out {0,3}): sorted (6,1),(4,4),(0,9),
#     (3,9); take 6; 4 kept (1.2*9 > 4); 0 kept (1.2*16 > 9); 3 dropped
#     (1.2*4 <= 9); take 4 -> G[2] = [6, 4]
#   3: 2 + 2 > R: robustPrune(3, {6,1} + out {2,5}): sorted (6,4),(1,9),(2,9),
#     (5,36); take 6; 1 kept (30 > 9); 2 dropped (1.2 <= 9); 5 kept; take 1
#     -> G[3] = [6, 1]
#   5: 1 + 1 = 2 <= R: append -> G[5] = [1, 3]            (newsize == R)
# Final sort by distance (:180-185) changes no list: G[1] = [3, 5] is a tie at
# 9, kept in id order.
VAMANA_X = np.array([[0], [9], [3], [6], [1], [12], [4]], np.float32)
VAMANA_ORDER = [3, 5, 2, 6, 1, 4, 0]
VAMANA_EXPECT = [[4, 2], [3, 5], [6, 4], [6, 1], [0, 2], [1, 3], [2, 3]]


@pytest.mark.parametrize("sort_neighbors", [False, True])
def test_vamana_hand_traced_build(oracle, sort_neighbors):
    perm = np.array(VAMANA_ORDER, np.uint32)[None, :]
    adj = oracle.build_vamana(VAMANA_X, "l2", 2, 2, 1.2, perm, sort_neighbors=sort_neighbors,
                              nthreads=1)
    assert _lists(adj) == VAMANA_EXPECT


def test_robust_prune_prunes_at_equality(oracle):
    """alpha * d(p*, p') <= d(p, p') (index.h:111): with alpha = 2, d(p*, p') = 1
    and a candidate distance d(p, p') = 2 exactly, p' goes; one ulp below 2 it stays."""
    X = np.array([[0.0], [1.0], [2.0]], np.float32)  # p = 0, p* = 1, p' = 2: d(1, 2) = 1
    adj = np.zeros((3, 3), np.uint32)
    cand = np.array([1, 2], np.uint32)
    out, _ = oracle.robust_prune(adj, X, "l2", 0, cand, np.array([1.0, 2.0], np.float32), 2.0, 2)
    assert out.tolist() == [1]
    below = np.nextafter(np.float32(2.0), np.float32(0.0))
    out, _ = oracle.robust_prune(adj, X, "l2", 0, cand, np.array([1.0, below], np.float32), 2.0,
                                 2)
    assert out.tolist() == [1, 2]


# ------------------------------------------------------------ HNSW hand trace
# 2-D points, ids in insertion order: 0 (2,0), 1 (1,2), 2 (0,0), 3 (-3,0);
# levels 1, 0, 1, 0; m = 1 (level 0 keeps 2m = 2, level 1 keeps m = 1);
# efc = 8 (every reachable node ends in the frontier); alpha = 1.
# n = 4, so 0.02 n rounds to 0 and each batch is one point.
#   node 0 is the first node and the entrance (level 1).
#   insert 1 (level 0): descent on level 1 from 0 (no edges) keeps 0.  Level 0:
#     frontier (0,5) -> select [0].  G0[1] = [0]; 0 gets 1: 0 + 1 <= 2 -> G0[0] = [1].
#   insert 2 (level 1): level 1 from 0: frontier (0,4) -> select (m = 1) [0];
#     G1[2] = [0]; 0 gets 2 on level 1: 0 + 1 = 1 <= 1 -> G1[0] = [2].
#     Level 0 from 0: visit 0 -> (1,5): frontier (0,4),(1,5).  prune_heuristic:
#     take 0; c = 1: d(1,0) = 1 + 4 = 5, alpha * d(1,2) = 5: 5 < 5 is false, so 1
#     is kept -> G0[2] = [0, 1].  Reverse: 0 has 1 + 1 = 2 = 2m -> append: G0[0] =
#     [1, 2] (a cut would sort them by distance to 0: [2, 1]); 1 -> G0[1] = [0, 2].
#   insert 3 (level 0): level-1 descent from 0 (25): 0's level-1 edge 2 at 9
#     replaces it; 2's edge 0 seen -> start from 2.  Level 0 from 2: frontier
#     (2,9),(1,20),(0,25).  prune: take 2; 1 goes (d(1,2) = 5 < 20); 0 goes
#     (d(0,2) = 4 < 25) -> G0[3] = [2].  Reverse: 2 has 2 + 1 > 2: cut to the 2
#     nearest of 0 (4), 1 (5), 3 (9) -> G0[2] = [0, 1].
HNSW_X = np.array([[2, 0], [1, 2], [0, 0], [-3, 0]], np.float32)
HNSW_LEVELS = np.array([1, 0, 1, 0], np.uint32)
HNSW_EXPECT = ([[1, 2], [0, 2], [0, 1], [2]], [[2], [], [0], []])


def test_hnsw_hand_traced_build(oracle):
    adj, ent = oracle.build_hnsw(HNSW_X, "l2", 1, 8, 1.0, HNSW_LEVELS, nthreads=1)
    assert ent == 0
    assert len(adj) == 2
    assert (_lists(adj[0]), _lists(adj[1])) == HNSW_EXPECT


# ------------------------------------------------------------ planted mistakes
# mutant -> the pins meant to catch it (ORACLE_MUTANT in oracle/pann_oracle.c)
PINS = {
    1: ["vamana_build"],        # reverse edges append only while newsize < R
    2: ["prune_equality"],      # robustPrune prunes only when alpha * d < d
    3: ["vamana_schedule", "vamana_build"],  # one batch
    4: ["vamana_schedule", "vamana_build"],  # batches of one
    5: ["prune_add"],           # robustPrune without p's out-neighbours
    6: ["vamana_build_unsorted"],  # reverse edges: old neighbours first
    7: ["hnsw_build"],          # prune_heuristic prunes at d(c, r) == alpha d(c, u)
    8: ["hnsw_build"],          # reverse edges cut when the list reaches m
    9: ["hnsw_build"],          # reverse-edge cut keeps the farthest
    10: ["hnsw_schedule"],      # batch end ceil(b * base) without + 1
}


def _run_pin(oracle, name):
    if name == "vamana_build":
        test_vamana_hand_traced_build(oracle, True)
    elif name == "vamana_build_unsorted":
        test_vamana_hand_traced_build(oracle, False)
    elif name == "vamana_schedule":
        test_vamana_batch_schedule(oracle)
    elif name == "prune_equality":
        test_robust_prune_prunes_at_equality(oracle)
    elif name == "prune_add":
        from test_oracle_pins import test_prune_adds_out_neighbours_and_dedups
        test_prune_adds_out_neighbours_and_dedups(oracle)
    elif name == "hnsw_build":
        test_hnsw_hand_traced_build(oracle)
    elif name == "hnsw_schedule":
        test_hnsw_batch_schedule(oracle)
    else:
        raise KeyError(name)


@pytest.mark.parametrize("mutant", sorted(PINS))
def test_planted_mistakes_are_caught(oracle, tmp_path, mutant):
    so = str(tmp_path / f"liboracle_mutant{mutant}.so")
    subprocess.run(["make", "-s", "-B", "-C", ORACLE_DIR, "mutant", f"ORACLE_MUTANT={mutant}",
                    f"OUT={so}"], check=True)
    failed = []
    with oracle.use_library(so):
        for name in PINS[mutant]:
            try:
                _run_pin(oracle, name)
            except AssertionError:
                failed.append(name)
    assert failed, f"no pin caught planted mistake {mutant}"
    # and the unplanted oracle passes the same pins
    for name in PINS[mutant]:
        _run_pin(oracle, name)
