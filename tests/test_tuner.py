"""Host-side checks of the autotuner: search space, features (no GPU), timing database I/O."""
import gzip
import json

import numpy as np

from symboliclib_amd import tuner as TU
from symboliclib_amd import workloads as W


def test_search_spaces():
    cfgs = list(TU.configs())
    assert len(cfgs) == 4 * 3 * 2 * 4 * 3 == 288
    assert len({json.dumps(c, sort_keys=True) for c in cfgs}) == 288
    assert len(list(TU.configs(TU.PRODUCT_SPACE))) == 96


def test_features_of_cloth_and_product_tiles():
    sides, cloths = TU.cloth_program(3, max_vertices=20000)
    t = TU.cloth_tuner(cloths, space=dict(TU.SPACE, lds=[32768]))
    u = t.build(32768)
    cfgs = t._cfgs(32768)
    stats = TU.code_stats(u, cfgs)
    tiles = u.tile_stats()
    F = np.stack([TU.features(stats[(c["groups"], c["wpe"])], tiles, c) for c in cfgs])
    assert F.shape == (72, len(TU.FEATURES)) and np.isfinite(F).all()
    # more output groups: fewer values live per work item, more statements per element
    assert stats[(5, 0)]["max_live"] < stats[(2, 0)]["max_live"]
    assert stats[(5, 0)]["statements"] > stats[(2, 0)]["statements"]
    # tiles cover every output exactly once
    assert int(tiles[:, 0].sum()) == u.fused_stats()["num_outputs"]
    u.close()
    n, cp, ri, a = W.c1_matrix(subdiv=3)
    t = TU.matrix_tuner(cp, ri, a, space=dict(TU.PRODUCT_SPACE, lds=[32768]))
    u = t.build(32768)
    c = t._cfgs(32768)[0]
    f = TU.features(TU.code_stats(u, [c])[(1, 0)], u.tile_stats(), c)
    assert np.isfinite(f).all() and f[TU.FEATURES.index("log2_elements_per_tile")] == 0
    u.close()


def test_timing_database_roundtrip(tmp_path):
    recs = [{"config": {"lds": 32768}, "ms": 0.5 + k, "features": [float(k)] * len(TU.FEATURES)} for k in range(3)]
    path = tmp_path / "db.jsonl"
    TU.write_db(path, {"kind": "cloth_batch", "seed": 1}, recs)
    got = TU.read_db(path)
    assert [r["ms"] for r in got] == [0.5, 1.5, 2.5] and got[0]["program"]["seed"] == 1
    gz = tmp_path / "db.jsonl.gz"
    with gzip.open(gz, "wt") as f:
        f.write(path.read_text())
    X, y = TU.db_arrays(TU.read_db(gz))
    assert X.shape == (3, len(TU.FEATURES)) and np.allclose(np.exp(y), [0.5, 1.5, 2.5])
    assert TU.regret(0.55, recs) == 0.55 / 0.5 - 1
