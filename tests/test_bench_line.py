"""The bench line's contract, checked without a GPU: the metric string is
the headline metric's, no dict literal in bench.py repeats a key (a repeated
key silently overwrites the first), and --gpus above 1 is refused."""

import ast
import io
import json
import os
from contextlib import redirect_stdout

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_no_repeated_keys_in_bench_dicts():
    tree = ast.parse(open(os.path.join(ROOT, "bench.py")).read())
    for node in ast.walk(tree):
        if isinstance(node, ast.Dict):
            keys = [k.value for k in node.keys if isinstance(k, ast.Constant)]
            repeated = {k for k in keys if keys.count(k) > 1}
            assert not repeated, f"bench.py:{node.lineno}: repeated keys {sorted(repeated)}"


def test_line_carries_the_headline_metric():
    with open(os.path.join(ROOT, "BASELINE.json")) as f:
        metric = json.load(f)["metric"]
    buf = io.StringIO()
    with redirect_stdout(buf):
        assert bench.main(["--legs", "cpu", "--no-cpu"]) == 0
    line = json.loads(buf.getvalue().strip().splitlines()[-1])
    assert line["metric"] == metric
    assert line["n_gpus"] == 1 and line["higher_is_better"] is True
    for key in ("roofline", "cpu_baseline", "config", "dtype", "unit", "steps", "warmup"):
        assert key in line


def test_more_than_one_gpu_is_refused(capsys):
    assert bench.main(["--gpus", "2"]) == 2
    assert "one GPU" in capsys.readouterr().err


def test_roofline_fractions_follow_the_counted_bytes_and_instructions():
    # 40 B per FP32 cell-update of the 256^3 tube over a 0.5 ms launch
    r = bench.roofline(256, 40, 0.5, "f32", (32, 24, 24))
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == bench.HBM_PEAK_GBS
    assert r["algorithmic_bytes"] == 40 * 256 ** 3
    assert abs(r["achieved"] - 40 * 256 ** 3 / 0.5e-3 / 1e9) < 0.1
    assert abs(r["frac"] - r["achieved"] / bench.HBM_PEAK_GBS) < 1e-4
    assert r["traffic"] == bench.TRAFFIC[("f32", 256, (32, 24, 24))]
    insts = bench.VALU_INSTS[("f32", 256, (32, 24, 24))]
    v = r["valu_issue"]
    assert v["insts_per_launch"] == insts and v["peak"] == 256 * 4 * 16 * 2.4e9
    assert abs(v["frac"] - insts * 64 / 0.5e-3 / v["peak"]) < 1e-4
    # a tile without counter runs carries no traffic or VALU figures
    r = bench.roofline(256, 40, 0.5, "f32", (64, 4, 8))
    assert r["traffic"] is None and "valu_issue" not in r
