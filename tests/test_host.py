"""Host-side logic that needs no GPU: struct layouts of the C ABI against the
ctypes binding, bench.py's command-line contract, the no-fallback rule of the
binding, the rocprof summariser and the physical set-up helpers."""
import ctypes
import json
import os
import sqlite3
import subprocess
import sys

import numpy as np
import pytest

from girih_amd import gpu, physics

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

LAYOUT_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "girih_gpu.h"
#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m))
int main(void) {
  printf("gr_desc %zu\n", sizeof(gr_desc)); F(gr_desc, flags); F(gr_desc, roc2); F(gr_desc, coef);
  printf("gr_stats %zu\n", sizeof(gr_stats)); F(gr_stats, device_bytes); F(gr_stats, x_offset);
  F(gr_stats, tile_variant); F(gr_stats, xcd_order);
  printf("gr_vmodel %zu\n", sizeof(gr_vmodel)); F(gr_vmodel, vmin); F(gr_vmodel, anomaly_center);
  F(gr_vmodel, round_bf16);
  printf("gr_batch_stats %zu\n", sizeof(gr_batch_stats)); F(gr_batch_stats, items); F(gr_batch_stats, nvariants);
  return 0;
}
"""


def test_struct_layouts_match_ctypes(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(LAYOUT_C)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I" + os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                                 check=True).stdout.split("\n") if line)
    py = {"gr_desc": gpu.Desc, "gr_stats": gpu.Stats, "gr_vmodel": gpu.VModel, "gr_batch_stats": gpu.BatchStats}
    for key, val in out.items():
        if "." in key:
            t, m = key.split(".")
            assert getattr(py[t], m).offset == int(val), key
        else:
            assert ctypes.sizeof(py[key]) == int(val), key


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--config", "1,2"], capture_output=True,
                       text=True)
    assert r.returncode == 2 and "--config" in r.stderr


def test_binding_has_no_fallback(monkeypatch):
    monkeypatch.setattr(gpu, "_lib", None)
    monkeypatch.setattr(gpu, "LIB_PATH", os.path.join(REPO, "girih_amd", "lib", "does_not_exist.so"))
    with pytest.raises(RuntimeError, match="not built"):
        gpu.lib()


def _fake_rocpd(path, kernel, durations_ns, counter=None, values_kib=None):
    c = sqlite3.connect(path)
    c.execute("create table top_kernels (name, total_calls, total_duration, average, percentage)")
    c.execute("create table kernels (name, start, duration)")
    c.execute("create table pmc_events (dispatch_id, name, counter_name, counter_value)")
    if durations_ns:
        c.execute("insert into top_kernels values (?,?,?,?,?)",
                  (kernel, len(durations_ns), sum(durations_ns) / 1e3, np.mean(durations_ns) / 1e3, 100.0))
        for n, d in enumerate(durations_ns):
            c.execute("insert into kernels values (?,?,?)", (kernel, n, d))
    for n, v in enumerate(values_kib or []):
        for half in (0.5, 0.5):  # two counter instances per dispatch are summed
            c.execute("insert into pmc_events values (?,?,?,?)", (n, kernel, counter, v * half))
    c.commit()
    c.close()


def test_profile_summary_numbers(tmp_path):
    k = "void gr::sweep_kernel<float>(x)"
    _fake_rocpd(tmp_path / "t.db", k, [2_000_000, 2_000_000])          # 2 ms per launch
    _fake_rocpd(tmp_path / "f.db", k, [], "FETCH_SIZE", [4096.0, 4096.0])  # 4 MiB counted
    _fake_rocpd(tmp_path / "w.db", k, [], "WRITE_SIZE", [2048.0])
    _fake_rocpd(tmp_path / "cf.db", "elementwise_kernel", [], "FETCH_SIZE", [512.0 * 1024])  # 1 GiB reads as 512 MiB
    _fake_rocpd(tmp_path / "cw.db", "elementwise_kernel", [], "WRITE_SIZE", [1024.0 * 1024])
    out = tmp_path / "p.md"
    tj = tmp_path / "traffic.json"
    subprocess.run([sys.executable, "-m", "girih_amd.profsum", "--trace", str(tmp_path / "t.db"),
                    "--fetch", str(tmp_path / "f.db"), "--write", str(tmp_path / "w.db"),
                    "--cfetch", str(tmp_path / "cf.db"), "--cwrite", str(tmp_path / "cw.db"),
                    "--kernel", "sweep_kernel", "--lups", "1000000", "--algo-bytes", "12000000",
                    "--cmd", "x", "--title", "t", "--out", str(out), "--traffic-json", str(tj),
                    "--key", "k@1,2,3,4", "--launch-config", "1,2,3,4"], check=True, cwd=REPO,
                   capture_output=True)
    d = json.loads(tj.read_text())["k@1,2,3,4"]
    assert d["fetch_bytes"] == pytest.approx(2 * 4096 * 1024)  # FETCH scale 2
    assert d["write_bytes"] == pytest.approx(2048 * 1024)
    assert d["avg_launch_us"] == pytest.approx(2000.0)
    assert d["launch_config"] == "1,2,3,4"
    assert "6.00 GB/s" not in out.read_text() and "12.00 B/LUP" in out.read_text()


def test_committed_traffic_index_matches_its_summaries():
    """profiles/traffic.json (what bench.py reads for roofline.traffic) points at committed
    summaries: each entry's key ends in its launch configuration, its summary exists, names
    that configuration in its command, and reports the same average launch time."""
    prof = os.path.join(REPO, "profiles")
    idx = json.load(open(os.path.join(prof, "traffic.json")))
    assert idx
    for key, e in idx.items():
        assert key.endswith("@" + e["launch_config"]), key
        text = open(os.path.join(prof, e["summary"])).read()
        assert e["launch_config"] in text, key
        assert "average duration: %.2f us" % e["avg_launch_us"] in text, key
        assert e["bytes_per_launch"] == pytest.approx(e["fetch_bytes"] + e["write_bytes"]), key


def test_physics_helpers():
    c = physics.fd8_coef()
    assert abs(c[0] + 6 * sum(c[1:])) < 1e-12  # a constant field has zero Laplacian
    x = np.arange(-4, 5, dtype=float)
    w = np.array([c[4], c[3], c[2], c[1], c[0] / 3, c[1], c[2], c[3], c[4]])
    assert abs(np.dot(w, x ** 2) - 2.0) < 1e-12  # exact second derivative of x^2
    assert abs(np.dot(w, x ** 8)) < 1e-9 and abs(np.dot(w, x ** 10)) > 1.0  # exact to degree 9, not 10
    dt = physics.cfl_dt(3000.0, 10.0, 0.4)
    assert physics.roc2_const(3000.0, dt, 10.0) == pytest.approx(0.16)
    r = physics.ricker(15.0, 1e-3, 200)
    assert np.argmax(r) == int(round(1.0 / 15.0 / 1e-3)) and r.max() == pytest.approx(1.0, abs=1e-3)
