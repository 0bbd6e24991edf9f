"""bench.py's command-line contract and its operation count, on the CPU."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_tape_stats_counts_reverse_operations():
    """Per arithmetic instruction: 1 (recompute) + per operand 1 (partial) + 2
    (multiply and accumulate).  x*y + sin(x): MUL 1+6, SIN 1+3, ADD 1+6 = 18;
    the inputs and the output count nothing."""
    sys.path.insert(0, ROOT)
    import bench
    import casadi_amd as ca
    x, y = ca.SX.sym("x"), ca.SX.sym("y")
    f = ca.Function("f", [x, y], [x * y + ca.sin(x)])
    st = bench.tape_stats(f, 2)
    ops = f.instructions()["op"]
    assert st["n_instr"] == ops.size and st["n_arith"] == 3
    assert st["rev_flops"] == 18
    assert st["n_seg"] == int(np.ceil(ops.size / 2))
