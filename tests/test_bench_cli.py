"""bench.py's contract on the CPU: more than one GPU is refused with a
message (nothing runs on a second GPU), before any device work; the CPU
baseline of configs 0, 2 and 3 (the oracle on samples) reports positive rates
and the GPU / CPU ratios of the GPU lines it is given."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_cpu_config_rates():
    sys.path.insert(0, ROOT)
    import bench
    from oracle import oracle as O
    import libftcrypto_amd as ft
    H = bytes(O.Array.from_scalars(ft.random_scalars(1, 3, 5)).to_bytes())
    gpu = {"cfg2_aggregation": {k: {"point_adds_per_s": 1e10} for k in ("reduce", "segmented_sum", "scan")},
           "cfg3_sparse_step_graph": {"graph_ms_per_step": 25.0}}
    r = bench.cpu_config_rates(H, 3, gpu)
    assert r["kind"] == "port" and r["cores"] >= 1
    assert r["cfg0"]["encrypt_ciphertexts_per_s"] > 0 and "gpu_over_cpu" not in r["cfg0"]
    for k in ("reduce", "segmented_sum", "scan"):
        rate = r["cfg2"][k + "_point_adds_per_s"]
        assert rate > 0 and abs(r["cfg2"]["gpu_over_cpu"][k] - 1e10 / rate) < 1e-6 * (1e10 / rate)
    assert r["cfg3"]["nodes_per_s"] > 0
    assert abs(r["cfg3"]["gpu_over_cpu"] - (2**22 / 25e-3) / r["cfg3"]["nodes_per_s"]) < 1e-6 * r["cfg3"]["gpu_over_cpu"]
