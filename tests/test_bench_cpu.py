"""bench.py host logic on the CPU: the one-GPU guard and the counter
scaling that turns rocprofv3 FETCH_SIZE / WRITE_SIZE readings into bytes."""
import importlib.util
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_mod", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_traffic_scaling():
    b = load_bench()
    copy_kib = b.COPY_MIB * 1024
    # a copy kernel whose FETCH reads half of what it moves, WRITE all of it
    passes = {"FETCH_SIZE": {"void hymor::k_copy(d2 const*)": copy_kib / 2,
                             "void hymor::k_term_split<8, double, 1, 4, true>(long)": 1000.0},
              "WRITE_SIZE": {"void hymor::k_copy(d2 const*)": float(copy_kib),
                             "void hymor::k_term_split<8, double, 1, 4, true>(long)": 10.0}}
    t, f = b.traffic_from_passes(passes, "k_term_split<8, double, 1, 4, true>")
    assert f["FETCH_SIZE"] == 2.0 and f["WRITE_SIZE"] == 1.0
    assert t == (1000.0 * 2 + 10.0) * 1024
    assert b.traffic_from_passes(passes, "k_term<8, double, 1, false>") == (None, None)
