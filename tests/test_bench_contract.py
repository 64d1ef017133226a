"""CPU checks of bench.py's contract pieces: the multi-GPU refusal, the
roofline object built from per-launch timings, and the PMC traffic summary."""
import csv
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_roofline_reports_the_dominant_launch():
    import bench
    csr = {"format": "csr", "method": "adaptive"}
    per = {"a": {"bytes": 1000_000_000, "us": 500.0, "column_segments": 1, "layout": csr},
           "b": {"bytes": 300_000_000, "us": 100.0, "column_segments": 4, "layout": csr}}
    r = bench.roofline(per, "adaptive")
    assert r["launch"].startswith("a SpMV") and r["bound"] == "hbm" and r["unit"] == "GB/s"
    assert abs(r["achieved"] - 2000.0) < 1e-6 and abs(r["frac"] - 0.25) < 1e-9
    assert r["peak"] == 8000.0 and "b" in r["other_launches"]
    assert r["other_launches"]["b"]["kernel"].startswith("csr_adaptive_xcd_f64")
    # the autotuner's picks: the diagonal layout and a 2048-unit CSR plan
    per["a"]["layout"] = {"format": "dia", "kernel": 0, "chunk": 0, "seg_cols": 0}
    per["b"]["layout"] = {"format": "csr", "kernel": 0, "chunk": 2048, "seg_cols": 0}
    r = bench.roofline(per, "tuned")
    assert r["kernel"] == "dia_spmv2_f64" and abs(r["achieved"] - 2000.0) < 1e-6
    assert r["other_launches"]["b"]["kernel"].startswith("csr_adaptive_xcd_f64")
    # the sorted-block segments (CSR kernel 1)
    per["b"]["layout"] = dict(per["b"]["layout"], kernel=1)
    r = bench.roofline(per, "tuned")
    assert r["other_launches"]["b"]["kernel"].startswith("csr_sorted_xcd_f64")


def _csv(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Grid_Size", "Counter_Name",
                                          "Counter_Value"])
        w.writeheader()
        for r in rows:
            w.writerow(r)


def test_traffic_summary_scales_by_the_copy(tmp_path):
    import bench
    k = "void bsp::csr_adaptive_f64<2048, false>(int const*)"
    fetch = [{"Kernel_Name": k, "Grid_Size": 256, "Counter_Name": "FETCH_SIZE",
              "Counter_Value": v} for v in (100.0, 110.0, 120.0)]
    write = [{"Kernel_Name": k, "Grid_Size": 256, "Counter_Name": "WRITE_SIZE",
              "Counter_Value": 10.0}]
    copy = "__amd_rocclr_copyBuffer"
    cf = [{"Kernel_Name": copy, "Grid_Size": 1, "Counter_Name": "FETCH_SIZE",
           "Counter_Value": 131072.0}]
    cw = [{"Kernel_Name": copy, "Grid_Size": 1, "Counter_Name": "WRITE_SIZE",
           "Counter_Value": 262144.0}]
    paths = [tmp_path / n for n in ("f.csv", "w.csv", "cf.csv", "cw.csv")]
    for p, rows in zip(paths, (fetch, write, cf, cw)):
        _csv(p, rows)
    out = tmp_path / "t.json"
    bench.traffic_summary(*map(str, paths), out=str(out))
    d = json.load(open(out))
    assert d["read_scale"] == 2.0 and d["write_scale"] == 1.0
    (launch,) = d["launches"].values()
    assert launch["hbm_bytes"] == int((2.0 * 110.0 + 10.0) * 1024)  # median fetch
