"""bench.py's targets summary, from a made-up bench line (no GPU)."""
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_targets_report_each_goal_and_skip_missing_legs():
    b = _bench()
    line = {
        "roofline": {"frac": 0.75},
        "step_hbm_frac": 0.69,
        "step_hbm_frac_fp64_state_formula": 0.695,
        "variants": {"fp16": {"speedup_vs_main": 1.8}},
        "precision_error": {"fp16": {"rel_l2_max": 4.5e-3}},
        "adjoint": {"fp32": {"cost_vs_forward": 2.9}, "fp32_whole_trace": {"cost_vs_forward": 2.2},
                    "rel_err_fp32_vs_fp64": 1.6e-6},
    }
    t = b.targets(line, "fp32")
    assert t["fp32_stencil_hbm_frac"]["met"] is True
    assert t["fp32_step_hbm_frac"] == {"value": 0.695, "bar": ">= 0.70", "met": False, "own_bytes": 0.69}
    assert t["fp16_speedup_vs_fp32"]["met"] is True
    assert t["fp16_ir_rel_l2_vs_fp64"]["met"] is False
    assert t["adjoint_cost_vs_forward"] == {"value": 2.9, "bar": "<= 3", "met": True, "whole_trace": 2.2}
    assert t["adjoint_fp32_vs_fp64_rel_err"]["met"] is True
    # legs that did not run, or a main precision other than FP32, leave goals out
    t2 = b.targets({"roofline": {"frac": 0.6}}, "fp16")
    assert t2 == {}
    t3 = b.targets({"roofline": {"frac": 0.6}}, "fp32")
    assert set(t3) == {"fp32_stencil_hbm_frac"} and t3["fp32_stencil_hbm_frac"]["met"] is False


def test_launch_label_says_what_ran():
    b = _bench()
    base = {"graph_launches": 0, "graph_steps": 0, "tail_graph_steps": 0, "single_steps": 0}
    assert b.launch_label(dict(base, graph_launches=1, tail_graph_steps=20)) == "one 20-step HIP graph"
    assert (b.launch_label(dict(base, graph_launches=17, graph_steps=30, tail_graph_steps=20))
            == "16 replays of a 30-step HIP graph + one 20-step HIP graph")
    assert b.launch_label(dict(base, graph_launches=3, graph_steps=30)) == "3 replays of a 30-step HIP graph"
    assert b.launch_label(dict(base, single_steps=20)) == "20 single-launch steps"


def test_more_than_one_gpu_is_refused():
    """The path runs on one GPU: --gpus N > 1 (or a WORLD_SIZE above 1)
    exits with status 2 and says so, before any GPU work."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    env = dict(os.environ, WORLD_SIZE="2")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")], capture_output=True, text=True,
                       timeout=120, env=env)
    assert r.returncode == 2
