"""Host logic of bench.py's default run: the batched, stress, learned and
filters workloads run after the frame and are nested under 'secondary' (their
per-kernel tables dropped); a workload that raises is recorded as an error and
the others still run.  The workload functions are replaced by stand-ins, so no
GPU is touched."""

import argparse
import importlib.util
import os

from conftest import ROOT


def _load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_secondary_lines(monkeypatch, capsys):
    bench = _load_bench()
    seen = []

    def ok(name):
        def run(args):
            seen.append((name, args.secondary, args.no_cpu_baseline))
            return bench._emit({"metric": name, "value": 1.0, "kernels": {"k": {}}}, args)

        return run

    def boom(args):
        seen.append(("learned", args.secondary, args.no_cpu_baseline))
        raise RuntimeError("no model")

    monkeypatch.setattr(bench, "run_batched", ok("batched"))
    monkeypatch.setattr(bench, "run_stress", ok("stress"))
    monkeypatch.setattr(bench, "run_learned", boom)
    monkeypatch.setattr(bench, "run_filters", ok("filters"))
    args = argparse.Namespace(steps=3, warmup=1, no_cpu_baseline=False, no_secondary=False, copy_mib=0)
    res = bench._secondary_lines(args)
    assert list(res) == ["batched", "stress", "learned", "filters"]
    assert [s[0] for s in seen] == ["batched", "stress", "learned", "filters"]
    assert all(sec and nocpu for _, sec, nocpu in seen)
    assert args.no_cpu_baseline is False and not hasattr(args, "secondary")
    for name in ("batched", "stress", "filters"):
        assert res[name]["metric"] == name and "kernels" not in res[name] and "wall_s" in res[name]
    assert res["learned"]["error"] == "RuntimeError: no model"
    # nothing printed to stdout: the default run prints one JSON line itself
    assert capsys.readouterr().out == ""


def test_emit_prints_standalone_line(capsys):
    bench = _load_bench()
    out = bench._emit({"metric": "m", "value": 2.0}, argparse.Namespace())
    assert out["value"] == 2.0
    assert capsys.readouterr().out.strip() == '{"metric": "m", "value": 2.0}'
