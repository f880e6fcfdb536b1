"""The batch tile-shape table of the autotuner (gismo_amd/autotune.py) on
the CPU: size buckets, class keys, candidate sets and lookups (a class
missing from the table takes the default shape); the committed table, if
any, names only valid shapes."""
import glob
import json
import os

from gismo_amd import autotune


def test_buckets_and_keys():
    assert autotune.bucket(100) == "s" and autotune.bucket(4095) == "s"
    assert autotune.bucket(4096) == "m" and autotune.bucket(32767) == "m"
    assert autotune.bucket(32768) == "l" and autotune.bucket(10 ** 7) == "l"
    assert autotune.class_key("mixed", 3, 4, 50000) == "mixed-d3-p4-l"


def test_candidates():
    mixed = autotune.candidates(True)
    fp64 = autotune.candidates(False)
    assert len(mixed) == 3 * 2 + 4 and len(fp64) == 3 + 4
    assert all(c["half"] == 1 for c in fp64)
    assert {c["lanes"] for c in mixed if c["kind"] == "csr"} == {8, 16, 32, 64}


def test_lookup():
    t = autotune.Table({"mixed-d3-p3-m": {"best": {"kind": "box", "lanes": 4, "half": 0}},
                        "fp64-d2-p2-s": {"best": {"kind": "csr", "lanes": 16, "half": 1}}}, "gfx950")
    assert t.lookup(True, 3, 3, 10000) == ("box", 4, 0)
    assert t.lookup(False, 2, 2, 100) == ("csr", 16, 1)
    assert t.lookup(True, 2, 2, 100) == ("box", 1, 1)      # not in the table: default
    assert autotune.Table.load("gfx000", directory="/nonexistent").lookup(True, 3, 3, 10) == ("box", 1, 1)


def test_committed_tables_are_valid():
    for path in glob.glob(os.path.join(autotune.DEFAULT_DIR, "batch_tiles_*.json")):
        data = json.load(open(path))
        assert data["classes"], path
        for key, e in data["classes"].items():
            b = e["best"]
            assert b["kind"] in ("box", "csr")
            assert b["lanes"] in (autotune.BOX_LANES if b["kind"] == "box" else autotune.CSR_LANES)
            assert b["half"] in (0, 1)
            assert key.startswith(("mixed-", "fp64-"))
            # the chosen shape is the fastest measured one
            best = min(e["us_per_iteration"].values())
            assert e["us_per_iteration"][autotune._name(b)] == best
