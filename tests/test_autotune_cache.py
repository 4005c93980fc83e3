"""Host logic of the launch autotuner's cache file (no GPU needed)."""
import json

from gempic_amd import core


def test_key_names_every_input():
    base = dict(V=2, n=4194304, nx=32, deg0=3, deg1=2, dt=0.05, n_steps=80,
                blocks=[0, 512], sort_intervals=[20, 40])
    k0 = core.autotune_key(**base)
    for name, other in [("V", 1), ("n", 1048576), ("nx", 64), ("deg0", 2), ("deg1", 1),
                        ("dt", 0.1), ("n_steps", 40), ("blocks", [0, 1024]),
                        ("sort_intervals", [20]), ("kernels_per_step", 2),
                        ("arithmetic", "fused"), ("device", "gfx950/256cu")]:
        assert core.autotune_key(**dict(base, **{name: other})) != k0, name
    assert core.autotune_key(**base) == k0


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "tune.json")
    assert core.autotune_cache_lookup(path, "k") is None
    res = {"blocks": 512, "sort_interval": 40, "ms_per_step": [[0.1]], "cached": False}
    core.autotune_cache_store(path, "k", res)
    core.autotune_cache_store(path, "k2", dict(res, blocks=1024))
    hit = core.autotune_cache_lookup(path, "k")
    assert hit["blocks"] == 512 and hit["sort_interval"] == 40 and "cached" not in hit
    assert core.autotune_cache_lookup(path, "k2")["blocks"] == 1024
    assert sorted(json.load(open(path))) == ["k", "k2"]


def test_cache_ignores_malformed_entries(tmp_path):
    path = tmp_path / "tune.json"
    path.write_text(json.dumps({"k": {"blocks": "512", "sort_interval": 40}}))
    assert core.autotune_cache_lookup(str(path), "k") is None
