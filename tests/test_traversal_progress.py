"""Termination of the speculative traversal loop, checked on the CPU.

A model of one wave (64 lanes) running the persistent traversal of
hikari_amd/csrc/hk_kernels.hip: the refill of idle lanes (k_extend /
k_intersect_persistent), the node phase of trav_round with its wave ballots
(a lane parks one leaf, closest-hit lanes a second; a lane whose next
reference is on its stack pops one entry per iteration, and a culled entry
leaves it popping again in the next iteration; the phase ends when no lane
searches for a first leaf, or when fewer than kLeafEarly do and some lane
holds a leaf), then the leaf phase.  Each lane's ray is a random finite
traversal: every node step or stack pop yields the next reference (an inner
node, a leaf, another pop, or the end of the ray), and any-hit leaves may end
the ray.

The property the kernel relies on: every round of the wave makes progress
(a node step, a stack pop, a leaf test, or a finished ray), so the wave
drains in a bounded number of rounds.  The same model without the "some lane
holds a leaf" condition of the early leaf phase stalls, which is the hang
that condition fixed.
"""
import numpy as np
import pytest

LANES = 64
K_LEAF_EARLY = 8  # kLeafEarly
NODE, LEAF, DONE, POP = 0, 1, 2, 3


class Ray:
    """A random traversal: the references a lane meets, in order."""

    def __init__(self, rng):
        n = int(rng.integers(1, 40))
        kinds = rng.choice([NODE, LEAF, POP], size=n, p=[0.55, 0.25, 0.2])
        self.refs = list(kinds) + [DONE]
        self.pos = 0
        self.hit_p = float(rng.uniform(0.0, 0.5))

    def next(self):
        r = self.refs[min(self.pos, len(self.refs) - 1)]
        self.pos += 1
        return r


def trav_round(lanes, any_hit, early_needs_leaf, rng, stats):
    """One call of trav_round for every lane of the wave (the loop of
    hk_kernels.hip trav_round, lines marked there)."""
    for ln in lanes:
        ln["pend2"] = DONE
    while True:
        for ln in lanes:                               # park a leaf, then one pop
            if ln["ray"] is None:
                continue
            if ln["ref"] == LEAF:
                if ln["pend"] == DONE:
                    ln["pend"], ln["ref"] = LEAF, POP
                elif not any_hit and ln["pend2"] == DONE:
                    ln["pend2"], ln["ref"] = LEAF, POP
            if ln["ref"] == POP:                       # a culled entry yields POP again
                ln["ref"] = ln["ray"].next()
                stats["pops"] += 1
        searching = sum(1 for ln in lanes
                        if ln["ray"] is not None and ln["ref"] in (NODE, POP) and ln["pend"] == DONE)
        holds = any(ln["ray"] is not None and ln["pend"] != DONE for ln in lanes)
        early = searching < K_LEAF_EARLY and (holds or not early_needs_leaf)
        if searching == 0 or early:
            break
        for ln in lanes:                               # node_step: a child, POP or DONE
            if ln["ray"] is not None and ln["ref"] == NODE:
                ln["ref"] = ln["ray"].next()
                stats["nodes"] += 1
    for slot in ("pend", "pend2"):                     # leaf phase
        for ln in lanes:
            if ln["ray"] is not None and ln[slot] != DONE:
                ln[slot] = DONE
                stats["leaves"] += 1
                if any_hit and rng.uniform() < ln["ray"].hit_p:
                    ln["ref"] = DONE
                    ln["pend"] = ln["pend2"] = DONE


def run_wave(seed, any_hit, early_needs_leaf=True, n_rays=600, refill=24, max_rounds=200000):
    """The persistent wave: returns (rounds, stalled)."""
    rng = np.random.default_rng(seed)
    lanes = [{"ray": None, "ref": DONE, "pend": DONE, "pend2": DONE} for _ in range(LANES)]
    next_ray, finished, rounds = 0, 0, 0
    while True:
        idle = [ln for ln in lanes if ln["ray"] is None]
        started = 0
        if next_ray < n_rays and len(idle) >= refill:   # wave_fetch refill
            for ln in idle:
                if next_ray == n_rays:
                    break
                ln.update(ray=Ray(rng), ref=NODE, pend=DONE)
                next_ray += 1
                started += 1
        if all(ln["ray"] is None for ln in lanes):
            if next_ray == n_rays:
                return rounds, False
            continue
        stats = {"nodes": 0, "leaves": 0, "pops": 0}
        trav_round(lanes, any_hit, early_needs_leaf, rng, stats)
        done_now = 0
        for ln in lanes:                               # a finished ray leaves its lane
            if ln["ray"] is not None and ln["ref"] == DONE and ln["pend"] == DONE:
                ln["ray"] = None
                done_now += 1
        finished += done_now
        rounds += 1
        if stats["nodes"] == 0 and stats["leaves"] == 0 and stats["pops"] == 0 and done_now == 0 and started == 0:
            return rounds, True
        if rounds > max_rounds:
            return rounds, True


@pytest.mark.parametrize("any_hit", [False, True])
@pytest.mark.parametrize("seed", range(40))
def test_every_round_makes_progress(seed, any_hit):
    rounds, stalled = run_wave(seed, any_hit)
    assert not stalled, f"wave stalled after {rounds} rounds"


def test_early_leaf_phase_without_held_leaf_stalls():
    """Planted slip: the early leaf phase without "some lane holds a leaf"."""
    stalls = sum(run_wave(seed, False, early_needs_leaf=False, max_rounds=5000)[1] for seed in range(40))
    assert stalls > 0
