"""Batched execution forms (graph-captured rounds, hierarchical roots, block
shapes) give results identical to the plain path and to the CPU oracle."""
import numpy as np
import pytest

from circlepacking_amd import abi

pytestmark = pytest.mark.gpu

G, H, PER = abi.OPT_GRAPH, abi.OPT_HIER_ROOTS, abi.OPT_PERSISTENT


def same(a, b):
    sa, fa = a
    sb, fb = b
    assert sa.as_dict() == sb.as_dict()
    assert np.array_equal(fa, fb)


@pytest.mark.parametrize("opts", [G, G | H, PER | H])
def test_graph_replays_match_plain_and_oracle(cpl, cplo, opts):
    base = abi.cfg_4param(max_depth=6, grid=(16, 16, 16, 16))
    plain = cpl.refine_run(base)
    ref = cplo.refine_run(base)
    same(plain, ref)
    cfg = abi.with_options(base, opts)
    for rep in range(4):  # capture, re-capture with measured sizes, then replays
        same(cpl.refine_run(cfg), plain)
        # eager work and a synchronisation between replays
        cpl.prim("add_rd", np.arange(1000.0), 1.0)


@pytest.mark.parametrize("block,cpt", [(64, 1), (128, 8), (512, 2), (1024, 4)])
def test_block_shapes(cpl, block, cpt):
    base = abi.cfg_small_2param(raw=True, n=128)
    plain = cpl.refine_run(base)
    for opts in (G, G | H, PER | H):
        same(cpl.refine_run(abi.with_options(base, opts, block, cpt)), plain)


@pytest.mark.parametrize("opts", [G | H, PER | H])
def test_full_grid_batched_forms_match_plain(cpl, opts):
    base = abi.cfg_4param()
    plain = cpl.refine_run(base)
    cfg = abi.with_options(base, opts)
    for _ in range(2):
        same(cpl.refine_run(cfg), plain)


def test_full_grid_depth8_forms_match_plain(cpl):
    """64^4 roots refined to depth 8 (about 9.2e7 evaluations, 8.8e6 items in
    the widest round): graph-captured and persistent forms equal the plain path.
    The unresolved leaves roughly double per extra level (2.8e6, 4.4e6, 8.8e6 at
    depths 6, 7, 8) while the certified count barely moves: those cells fail the
    criteria at every width."""
    base = abi.cfg_4param(max_depth=8)
    plain = cpl.refine_run(base)
    s = plain[0]
    assert s.rounds == 9 and s.unresolved > 2 * 4418579 - 100000
    for opts in (G | H, PER | H):
        cfg = abi.with_options(base, opts)
        for _ in range(2):
            same(cpl.refine_run(cfg), plain)


def test_persistent_small_capacity_spills_and_matches(cpl, monkeypatch):
    """A deep 2x2 refinement forces the per-wave stacks to spill chunks to the global queue."""
    base = abi.cfg_small_2param(raw=True, n=64, max_depth=6)
    plain = cpl.refine_run(base)
    same(cpl.refine_run(abi.with_options(base, PER | H)), plain)


def test_no_digest_keeps_counts(cpl):
    base = abi.cfg_4param(grid=(16, 16, 16, 16))
    a, _ = cpl.refine_run(base)
    b, _ = cpl.refine_run(abi.with_options(base, G | H | abi.OPT_NO_DIGEST))
    da, db = a.as_dict(), b.as_dict()
    for k in ("evaluations", "verdicts", "family_verdicts", "certified", "unresolved"):
        assert da[k] == db[k]
    assert all(x == 0 for x in db["digest"])


def test_autotune_picks_a_listed_shape(cpl):
    cfg = abi.with_options(abi.cfg_4param(grid=(16, 16, 16, 16)), G | H | abi.OPT_NO_DIGEST)
    table = cpl.autotune(cfg, reps=1)
    assert len(table) == 20
    best = min(table, key=lambda t: t[2])
    assert (cfg.block, cfg.cells_per_thread) == best[:2]


def test_graph_runs_report_only_their_families(cpl, cplo):
    # each family chain of the graph form clears its own statistics records, so
    # records that a run with other families left behind must not reach a run
    # without them (with and without per-root flags, i.e. with and without the
    # shared prologue)
    four = abi.with_options(abi.cfg_4param(grid=(16, 16, 16, 16)), G | H)
    cases = [(abi.cfg_4param(grid=(16, 16, 16, 16), families=(abi.MIDDLE, abi.END_3D)), G | H),
             (abi.cfg_small_2param(), G)]
    for cfg, opts in cases:
        cpl.refine_run(four)
        plain = cpl.refine_run(cfg)
        same(plain, cplo.refine_run(cfg))
        for flags in (False, True):
            s, f = cpl.refine_run(abi.with_options(cfg, opts), want_flags=flags)
            assert s.as_dict() == plain[0].as_dict()
            if flags:
                assert np.array_equal(f, plain[1])


def test_graph_recaptures_after_buffers_move(cpl):
    """A captured graph addresses the library's device buffers: a plain run that
    grows one of them (a deeper lambda table), and a shutdown that releases
    them all, must make the next graph run capture again (its results equal the
    first run's) instead of replaying kernels on freed memory."""
    cfg = abi.with_options(abi.cfg_4param(grid=(16, 16, 16, 16), max_depth=4), G | H)
    first = cpl.refine_run(cfg)
    cpl.refine_run(cfg)  # a replay
    # plain form with a lambda table larger than any earlier run's (64 cells x 2^11 levels):
    # the table's buffer is reallocated
    cpl.refine_run(abi.cfg_4param(grid=(64, 2, 2, 2), max_depth=10))
    same(cpl.refine_run(cfg), first)
    assert cpl.lib().cpl_shutdown() == 0
    assert cpl.lib().cpl_init() == 0
    same(cpl.refine_run(cfg), first)
    same(cpl.refine_run(cfg), first)
