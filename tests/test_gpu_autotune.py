"""The MTE tile autotuner and tile settings on the GPU.

roadmc_autotune_ex must try every valid setting of each patch-count bucket
and keep the fastest it measured; the rows of paths per workgroup and the
workgroups per CU must not change any path's end total (each path's products
and draws are the same in any tile), and the k-split of remainder tiles may
move them only within FP32 rounding."""
import numpy as np
import pytest

from transportoptimiser_amd import landscape
from transportoptimiser_amd import native as nat

pytestmark = pytest.mark.gpu

BUCKETS = (128, 256, 512, 1024)
PRECISIONS = [nat.PREC_FP32, nat.PREC_BF16X3]


def _valid_rows(b):
    return [r for r in (32, 64, 128, 256) if r * b <= 32768]


@pytest.fixture
def keep_tiles(ctx):
    saved = {p: ctx.tile_table(p) for p in PRECISIONS}
    yield
    for p, tab in saved.items():
        for b, t in tab.items():
            ctx.set_tile(p, b, t["rows"], t["split"], t["per_cu"])


@pytest.mark.parametrize("prec", PRECISIONS)
def test_autotune_keeps_the_fastest_valid_setting(ctx, keep_tiles, prec):
    print(f"autotune precision {prec}", flush=True)
    trials = ctx.autotune(n_paths=2048, n_years=2, precision=prec)
    for b in BUCKETS:
        tb = [t for t in trials if t["bucket_max_p"] == b]
        assert {t["rows"] for t in tb} == set(_valid_rows(b)), b
        assert all(t["precision"] == prec and t["ms"] > 0 for t in tb)
        assert all(t["per_cu"] in (1, 2) and t["split"] in (0, 1) for t in tb)
        # one trial per distinct setting
        assert len({(t["rows"], t["split"], t["per_cu"]) for t in tb}) == len(tb)
        chosen = [t for t in tb if t["chosen"]]
        assert len(chosen) == 1
        assert chosen[0]["ms"] == min(t["ms"] for t in tb)
        assert ctx.tile(prec, b) == {k: chosen[0][k] for k in ("rows", "split", "per_cu")}


def test_set_tile_rejects_invalid_settings(ctx, keep_tiles):
    for args in [(nat.PREC_FP64, 128, 32, 0, 0), (nat.PREC_FP32, 100, 32, 0, 0),
                 (nat.PREC_FP32, 1024, 64, 0, 0), (nat.PREC_FP32, 256, 48, 0, 0),
                 (nat.PREC_FP32, 256, 32, 2, 0), (nat.PREC_FP32, 256, 32, 0, 3)]:
        with pytest.raises(nat.RoadmcError) as e:
            ctx.set_tile(*args)
        assert e.value.code == nat.ERR_INVALID


def _roads():
    # One road per bucket (P = 100, 200, 300, 700 patches); P = 200 and 300
    # leave remainder tiles at 32 rows, so the split setting matters.
    out = []
    for i, P in enumerate((100, 200, 300, 700)):
        sp = landscape.make_species(np.random.default_rng(50 + i), P, mort_max=0.4,
                                    threshold_ratio=0.7)
        out.append(landscape.RoadProblem([sp], n_years=6, n_paths=1000, seed=9, road_id=i))
    return out


def _run(ctx, roads, prec, rows_of, split, per_cu):
    for b in BUCKETS:
        ctx.set_tile(prec, b, rows_of(b), split, per_cu)
    for r in roads:
        r.precision = prec
    print(f"batch: precision {prec} rows {[rows_of(b) for b in BUCKETS]} split {split} "
          f"per_cu {per_cu}", flush=True)
    res = ctx.simulate_mte_batch(roads, end_pops=True)
    return [o.end_pops for o in res], [o.mte_prob for o in res]


@pytest.mark.parametrize("prec", PRECISIONS)
def test_rows_and_per_cu_do_not_change_results(ctx, keep_tiles, prec):
    roads = _roads()
    base, base_mte = _run(ctx, roads, prec, lambda b: 32, 0, 1)
    for rows_of, per_cu in [(lambda b: max(_valid_rows(b)), 0), (lambda b: 64 if b <= 512 else 32, 2),
                            (lambda b: 32, 2)]:
        ends, mte = _run(ctx, roads, prec, rows_of, 0, per_cu)
        for a, e in zip(base, ends):
            np.testing.assert_array_equal(a, e)
        for a, e in zip(base_mte, mte):
            np.testing.assert_array_equal(a, e)
    # the k-split of remainder tiles: FP32 rounding only
    ends, _ = _run(ctx, roads, prec, lambda b: 32, 1, 1)
    for a, e in zip(base, ends):
        np.testing.assert_allclose(e, a, rtol=2e-6)
