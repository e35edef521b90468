"""Index bounds of the batched multi-view frame, emulated on the CPU for the
full-size 50-view set (48 probe faces at 128^2/256^2/512^2 + 2 XR eyes at
2064x2208, log-uniform [32, 4096] subsets of an 8192-light pool).

For every configuration the autotuner can choose (wave tile 8x8 / 16x4,
cluster cell 32 / 64 / 128 px, 16 / 24 / 32 / 48 slices), applied to all 50
views at once and to each eye alone (the autotuner's one-view plans), the
addresses of k_batch_geom, k_batch_mark, k_batch_mark_tail, k_batch_count,
k_batch_scan, k_batch_fill and the persistent shading item walk are
recomputed from the kernels' formulas with the real cameras and lights and
checked against the plan's allocations (tests/batch_emu.py).  Both cell
layouts are covered: views aligned to 256 cells, and packed back to back as
the batched cull laid them out before the alignment.

These checks follow the kernels' source.  They could not see the mark
queue's room check as the compiler had built it: written as a select over a
wrapping difference, it was folded away at -O3, and a light whose sub-queue
was full wrote its chunk entries past its segment (past the queue buffer for
sub-queue 63).  The source now states the room as a saturating subtraction;
tests/test_gpu_parity.py::test_cluster_cull_mark_queue_overflow runs the
overflow on the device.
"""
import numpy as np
import pytest

import batch_emu as E
from helio_amd import abi


@pytest.fixture(scope="module")
def lib():
    return abi.Lib()


@pytest.fixture(scope="module")
def c5(lib):
    return E.c5_views(lib, small=False)


def run_plan(lib, views, cfg, aligned):
    plan = E.build_plan(views, [cfg] * len(views), lib, aligned=aligned)
    E.check_geom(views, plan, aligned=aligned)
    stats = E.check_mark(views, plan)
    E.check_count_scan_fill(views, plan, aligned=aligned)
    E.check_shade_items(views, plan)
    return plan, stats


@pytest.mark.parametrize("aligned", [True, False], ids=["aligned", "packed"])
def test_full_set_every_candidate(lib, c5, aligned):
    views, pool_size = c5
    assert len(views) == 50
    assert sum(v.width * v.height for v in views) == 14_619_648
    seen = set()
    for cfg in E.candidates():
        plan, stats = run_plan(lib, views, cfg, aligned)
        seen.add((cfg[1], cfg[2]))
        assert stats["lights"] > 0 and stats["chunks"] >= stats["lights"]
        for v in views:
            assert v.subset.max() < pool_size
    assert len(seen) == 12


def test_one_view_plans_of_the_eyes(lib, c5):
    views, _ = c5
    eyes = [v for v in views if v.adjoint]
    assert len(eyes) == 2
    for cfg in E.candidates():
        for v in eyes:
            run_plan(lib, [v], cfg, aligned=True)


def test_queue_overflow_stays_in_bounds(lib, c5):
    """Every light of both eyes at the finest grid (cell 32, 48 slices): the
    chunks past a full sub-queue stay with their light's wave; the entries
    read by the tail are exactly those written."""
    views, _ = c5
    eyes = [v for v in views if v.adjoint]
    plan, stats = run_plan(lib, eyes, (3, 32, 48), aligned=True)
    assert stats["queued"] <= E.K_MARK_QUEUES * plan["seg"]


def test_scan_truncates_to_the_pool():
    rng = np.random.default_rng(7)
    for cap in (0, 24):
        raw = rng.integers(0, 5000, 4096)
        o, cnt = E.scan_emulate(raw, cap, 10_000)
        assert (o + cnt <= 10_000).all()
        assert cnt.sum() == min(np.minimum(raw, cap or raw.max()).sum(), 10_000)


def test_floor_div_exact_for_plan_divisors():
    """hl_floor_div is exact below 2^21 for the slice counts and column counts
    of every plan above (and for any divisor up to 130, for any reciprocal
    within one ulp)."""
    x = np.arange(1 << 21, dtype=np.int64)
    for d in (1, 2, 3, 5, 7, 16, 24, 32, 33, 35, 47, 48, 64, 65, 69, 127, 129):
        for q in E.floor_div_f32(x, d):
            assert np.array_equal(q, x // d), d


def test_emulation_constants_match_the_sources():
    """Every K_* constant of the emulation has the value of its constexpr in
    helio_amd/csrc (kMarkInline for K_MARK_INLINE, and so on)."""
    import re
    from pathlib import Path
    src = "".join(p.read_text() for p in sorted((Path(__file__).resolve().parents[1] / "helio_amd" / "csrc")
                                                 .glob("*.[hc]*")))
    defs = dict(re.findall(r"constexpr uint32_t (k[A-Za-z0-9]+) = (\d+)u?;", src))
    names = [n for n in dir(E) if re.fullmatch(r"K_[A-Z0-9_]+", n)]
    assert len(names) >= 9
    for n in names:
        k = "k" + "".join(w.capitalize() for w in n[2:].split("_"))
        assert k in defs, f"{n}: no constexpr {k}"
        assert int(defs[k]) == getattr(E, n), n
