"""bench.py's contract, checked without a GPU: it refuses more than one GPU
before touching one, its named workloads are the configurations the metric
is quoted on, its algorithmic bytes per node follow the step's accounting
(four population transfers per direction plus the macroscopic fields), and
its traffic figures come from the committed rocprofv3 summaries."""
import os
import subprocess
import sys

import pytest

import bench
from clip_amd import Config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_before_any_gpu_work():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "single MI355X" in r.stderr
    assert r.stdout == ""


@pytest.mark.parametrize("name,dim,interior", [
    ("bubble3d256", 3, (256, 256, 256)),
    ("jet3d192x192x768", 3, (192, 192, 768)),
    ("drop2d128", 2, (128, 128)),
    ("rti2d", 2, (256, 1024)),
    ("slab192x192x96", 3, (192, 192, 96)),
])
def test_named_workloads_have_the_quoted_sizes(name, dim, interior):
    cfg = Config.from_string(bench.workload_text(name))
    assert cfg.dim == dim
    assert tuple(cfg.interior_shape()) == interior


@pytest.mark.parametrize("precision,per_node", [("f64", 664), ("f32", 332), ("mixed_f16", 180), ("mixed_bf16", 180)])
def test_fused_kernel_bytes_per_node(precision, per_node):
    """k_fused reads and writes f and g (4 Q transfers of the stored type),
    reads phi and reads and writes u (2 D + 1 values of the compute type)."""
    Q, D = 19, 3
    assert 4 * Q * bench.pop_bytes(precision) + (2 * D + 1) * bench.mac_bytes(precision) == per_node


def test_traffic_figures_come_from_committed_summaries():
    for precision in ("f64", "f32", "mixed_f16"):
        t = bench.load_traffic("k_fused", "bubble3d256", precision)
        assert t is not None and t > 0, precision
    assert bench.load_traffic("k_fused", "no_such_workload", "f64") is None


def test_cpu_identification():
    s = bench.cpu_model()
    assert s and "cores available" in s
    assert bench.HOST_CORES >= 1
