"""bench.py contract checks that need no GPU: the multi-GPU refusal, and that the committed
counter summary the bench line's roofline.traffic comes from parses to the FP64 hopping
term's bytes per launch."""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def test_multi_gpu_request_exits_with_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_traffic_summary_parses():
    import bench

    t = bench.pmc_traffic(ROOT / "profiles" / "dslash_32x32x32x32_pmc.txt", "k_dslash<kq::LinkD, kq::LinkD14, double, 1")
    assert t is not None
    # the FP64 term's compulsory bytes (14-real long links) are 2144 B/site x 524288 sites = 1.124 GB;
    # the counters may only add to that (neighbour re-reads), and not by more than a few percent
    assert 1.124e9 <= t <= 1.25e9
    # the 4-right-hand-side FP32 term (rows staged in LDS): 1184 B/site = 0.621 GB compulsory, ~30 % more fetched
    t4 = bench.pmc_traffic(ROOT / "profiles" / "dslash_32x32x32x32_pmc.txt",
                           "k_dslash_staged<kq::LinkF, kq::LinkF12, float, 4")
    assert t4 is not None and 0.621e9 <= t4 <= 1.0e9


@pytest.mark.parametrize("name", ["dslash_32x32x32x32_pmc.txt", "dslash_32x32x32x32_kernels_only_stats.txt",
                                  "cgm_32x32x32x32_12shift_kernel_stats.txt"])
def test_profiles_state_their_command(name):
    text = (ROOT / "profiles" / name).read_text()
    assert "made by" in text and "rocprofv3" in text and "bench.py" in text


def test_host_cpu_description():
    import bench

    hc = bench.host_cpu()
    assert hc["cpus_allowed"] >= 1
    assert hc["physical_cores"] is None or hc["physical_cores"] >= 1
    assert isinstance(hc["model"], str) and hc["model"]
