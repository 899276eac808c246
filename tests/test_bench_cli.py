"""bench.py's command line and its byte accounting (CPU only: nothing here
loads the HIP library)."""
import importlib.util
import pathlib
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parents[1]


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_help_lists_the_hardware_queue_setting():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--help"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    assert "--hw-queues" in r.stdout


def test_algorithmic_bytes_of_the_headline_matrix():
    b = _bench()
    # R-MAT s22 ef16 seed 2: n = 2^22, nnz = 65,243,842; int32 CSR with FP32
    # values, x and y: nnz (4 + 4) + (n + 1) 4 + 2 n 4, counted by hand
    n, nnz = 1 << 22, 65_243_842
    assert b.csr_bytes(n, n, nnz, 4) == 521_950_736 + 16_777_220 + 33_554_432
    # k right-hand sides add k (x + y) vectors; BF16 values take 2 bytes
    assert b.csr_bytes(n, n, nnz, 2, k=8) == nnz * 6 + (n + 1) * 4 + 8 * 2 * n * 4


def test_traffic_figure_needs_the_profiled_sources():
    b = _bench()
    # a policy other than the one a summary names gets no figure: the
    # workgroup-tile policy 1 and the panel policy without the streaming hint
    assert b.traffic_from_profile(1) is None
    assert b.traffic_from_profile(8) is None
    h = b.kernel_source_hash(9)
    assert len(h) == 16 and h == b.kernel_source_hash(9)
    t = b.traffic_from_profile(9)
    text = b.TRAFFIC_PROFILES["panel"].read_text()
    if ("source_sha16: " + h) in text:
        assert t is not None and t > 0
    else:
        assert t is None
