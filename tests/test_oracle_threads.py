"""Thread-count determinism of the oracle (CPU only).

The oracle, like the example, runs the M2L, P2P and L2P loops with OpenMP
and adds into shared local expansions and fields through atomic updates of
the generated operators (printers.py:271-275 with atomic=True,
example/example.py:6,16), so with several threads the order of those
additions, and with it the last bits of the result, varies from run to run.
Checked here: one thread gives bit-identical results on repeated runs, and
eight threads agree with one thread to rounding (FMM and Barnes-Hut).
"""
import os
import subprocess
import sys

import numpy as np

_SCRIPT = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[2])
from oracle import oracle as O
from tests import util
pos, q = util.uniform(20000, 41)
t = O.OracleTree(pos, q, 32, 8, 0.5)
np.save(sys.argv[1] + "_fmm.npy", t.compute_field_fmm())
np.save(sys.argv[1] + "_bh.npy", t.compute_field_bh())
"""


def _run(tmp, name, threads):
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_SCHEDULE="dynamic,16")
    out = os.path.join(tmp, name)
    subprocess.run([sys.executable, "-c", _SCRIPT, out, root], env=env, check=True, cwd=root, timeout=300)
    return np.load(out + "_fmm.npy"), np.load(out + "_bh.npy")


def test_oracle_one_thread_is_deterministic_and_matches_many(tmp_path):
    a = _run(str(tmp_path), "t1a", 1)
    b = _run(str(tmp_path), "t1b", 1)
    c = _run(str(tmp_path), "t8", 8)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    for x, z in zip(a, c):
        assert np.linalg.norm(x - z) / np.linalg.norm(x) < 1e-14
