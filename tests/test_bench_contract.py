"""bench.py checks that need no GPU: it measures one GPU only, and the committed
line-sum profile it reads (roofline.traffic, VALU issue fraction) describes the
kernel that the default schedule launches."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_committed_profile_matches_default_schedule():
    from irsounderlbl.api import Context
    import inspect
    sig = inspect.signature(Context.set_schedule)
    P, persistent = sig.parameters["tile_points"].default, sig.parameters["persistent"].default
    kernel = ("linesum_persistent_kernel" if persistent else "linesum_kernel") + f"<{P}>"
    with open(os.path.join(ROOT, "profiles", "linesum_full_iasi_pmc.json")) as f:
        J = json.load(f)
    assert J["kernel"] == kernel
    for k in ("hbm_bytes_per_launch", "valu_insts_per_launch", "launch_ms_avg_trace", "valu_flops_fp32_per_launch"):
        assert J[k] > 0
    # the counter-based FP32 FLOP/s of the committed profile is a fraction of the peak
    tflops = 64 * J["valu_flops_fp32_per_launch"] / (J["launch_ms_avg_trace"] * 1e-3) / 1e12
    assert 0 < tflops < 157.3
    with open(os.path.join(ROOT, "profiles", "rte_ils_full_iasi_pmc.json")) as f:
        H = json.load(f)
    assert H["rte"]["kernel"].startswith("rte_kernel") and H["ils"]["kernel"].startswith("ils_block_kernel")
    for k in ("rte", "ils"):
        assert H[k]["hbm_bytes_per_launch"] > 0 and H[k]["launch_ms_avg_trace"] > 0
    with open(os.path.join(ROOT, "profiles", "linesum_full_iasi_kernel_trace.md")) as f:
        trace = f.read()
    assert f"void {kernel}(LinesumArgs)" in trace
    assert f"{J['launch_ms_avg_trace']:.3f}" in trace
