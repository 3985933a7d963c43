"""The autotuner's launch-shape table: the library's reader (oks_tune_lookup,
host code, no device needed) against known tables, and the committed gfx950
table's shape and contents.  The table is read once per process, so every
case runs in a fresh interpreter with $OKS_TUNE_FILE pointing at it."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE = os.path.join(ROOT, "openkspace_amd", "tuning", "gfx950.json")


def _lookup(path, queries):
    code = ("import json, sys\n"
            "from openkspace_amd._lib import tune_lookup, OksError\n"
            "out = []\n"
            "for arch, op, key, name in json.loads(sys.argv[1]):\n"
            "    try:\n"
            "        out.append(tune_lookup(arch, op, key, name))\n"
            "    except OksError as e:\n"
            "        out.append('error %d' % e.code)\n"
            "print(json.dumps(out))\n")
    env = dict(os.environ)
    env["OKS_TUNE_FILE"] = path
    r = subprocess.run([sys.executable, "-c", code, json.dumps(queries)], cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_lookup_known_table(tmp_path):
    t = {"arch": "gfx950", "tool": "test", "entries": [
        {"op": "cs", "key": [320, 320], "opts": {"cs_sx": 8, "cs_threads": 256}, "tried": {"a": [1, {"b": None}]}},
        {"op": "ifft_rss", "key": [15, 368, 640, 320, 1], "opts": {"rows_waves": 12}, "ms": 0.5}]}
    p = tmp_path / "t.json"
    p.write_text(json.dumps(t))
    got = _lookup(str(p), [["gfx950", "cs", [320, 320], "cs_sx"], ["gfx950", "cs", [320, 320], "cs_threads"],
                           ["gfx950", "ifft_rss", [15, 368, 640, 320, 1], "rows_waves"],
                           ["gfx950", "cs", [320, 320], "rows_waves"], ["gfx950", "cs", [256, 256], "cs_sx"],
                           ["gfx942", "cs", [320, 320], "cs_sx"]])
    assert got == [8, 256, 12, None, None, None]


def test_malformed_and_missing_tables(tmp_path):
    bad = tmp_path / "bad.json"
    bad.write_text('{"arch": "gfx950", "entries": [{"op": "cs", "key": [1.5], "opts": {}}]}')
    assert _lookup(str(bad), [["gfx950", "cs", [1], "cs_sx"]]) == ["error 1"]
    trunc = tmp_path / "trunc.json"
    trunc.write_text('{"arch": "gfx950", "entries": [')
    assert _lookup(str(trunc), [["gfx950", "cs", [1], "cs_sx"]]) == ["error 1"]
    assert _lookup(str(tmp_path / "absent.json"), [["gfx950", "cs", [320, 320], "cs_sx"]]) == [None]
    assert _lookup("", [["gfx950", "cs", [320, 320], "cs_sx"]]) == [None]


@pytest.mark.skipif(not os.path.exists(TABLE), reason="no committed gfx950 table")
def test_committed_table():
    t = json.load(open(TABLE))
    assert t["arch"] == "gfx950" and t["entries"]
    allowed = {"cs_sx", "cs_threads", "cols_sx", "cols_threads", "rows_waves", "cs_cols_map", "cs_upd_map", "cs_pipe",
               "cs_coop"}
    for e in t["entries"]:
        assert e["op"] in ("ifft_rss", "cs") and set(e["opts"]) <= allowed
        assert e["ms"] <= e["default_ms"]
    ops = {(e["op"], tuple(e["key"])) for e in t["entries"]}
    # the served stream's shapes are covered
    for ny, nx, nc, os_x in [(256, 256, 8, 1), (320, 320, 16, 1), (368, 640, 15, 2), (384, 384, 32, 1)]:
        assert ("cs", (ny, nx)) in ops and ("ifft_rss", (nc, ny, nx, nx // os_x, 1)) in ops
    q = [["gfx950", e["op"], e["key"], k] for e in t["entries"] for k in e["opts"]]
    want = [v for e in t["entries"] for v in e["opts"].values()]
    assert _lookup(TABLE, q) == want
