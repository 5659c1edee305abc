"""bench.py's host side: the stored HBM-traffic readings are well formed and
name source files that exist (their hash is what lets the benchmark drop a
stale reading), and a multi-GPU request exits with status 2 before any GPU
work."""
import glob
import json
import os
import re
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_traffic_files_well_formed():
    files = glob.glob(os.path.join(ROOT, "profiles", "*_traffic.json"))
    assert files
    for f in files:
        j = json.load(open(f))
        assert j["kernel"] and j["hbm_bytes_per_launch"] > 0, f
        assert j["fetch_scale"] == 2.0 and j["write_scale"] == 1.0, f
        expect = j["fetch_size_kib"] * 1024 * j["fetch_scale"] + j["write_size_kib"] * 1024 * j["write_scale"]
        assert abs(j["hbm_bytes_per_launch"] - expect) <= 2, f
        assert j["sources"], f
        for rel in j["sources"]:
            assert os.path.isfile(os.path.join(ROOT, rel)), (f, rel)
        assert re.fullmatch(r"[0-9a-f]{64}", j["sources_sha256"]), f


def test_sources_hash_changes_with_content(tmp_path, monkeypatch):
    a = tmp_path / "a.hip"
    a.write_text("one")
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    h1 = bench.sources_sha256(["a.hip"])
    a.write_text("two")
    assert bench.sources_sha256(["a.hip"]) != h1
    assert bench.sources_sha256([]) is None


def test_multi_gpu_request_exits(monkeypatch, capsys):
    monkeypatch.setattr(sys, "argv", ["bench.py", "--gpus", "2"])
    with pytest.raises(SystemExit) as e:
        bench.main()
    assert e.value.code == 2
    assert "one GPU" in capsys.readouterr().err
