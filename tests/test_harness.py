"""The point-directory runner (reference wire format) end to end."""
import subprocess

import numpy as np
import pytest

from conftest import PKG_DIR
from harness_util import read_answer, read_results, write_point

EXE = PKG_DIR / "lib" / "bytecrt_solution"


def test_runner_usage_and_bad_manifest(tmp_path):
    r = subprocess.run([str(EXE)], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    (tmp_path / "manifest.json").write_text('{"B": 1}')
    r = subprocess.run([str(EXE), str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 2 and "manifest.json has no N" in r.stderr


@pytest.mark.gpu
def test_runner_end_to_end(tmp_path, pkg, orc):
    B, N, W = 2, 8192, 220
    L = (W + 31) // 32
    q = pkg.datagen.modulus(W, 77)
    cases = []
    for c in range(2):
        a = np.stack([pkg.datagen.operands(N, q, 100 * c + i, edges_first=True) for i in range(B)])
        b = np.stack([pkg.datagen.operands(N, q, 100 * c + 50 + i) for i in range(B)])
        cases.append((a, b))
    write_point(tmp_path, B, N, W, q, cases)
    # a third, truncated case must be recorded as crashed, not abort the run
    (tmp_path / "manifest.json").write_text(f'{{"B": {B}, "N": {N}, "W": {W}, "L": {L}, "cases": 3}}')
    bad = tmp_path / "cases" / "000002"
    bad.mkdir(parents=True)
    (bad / "a.bin").write_bytes(b"\0" * 12)
    (bad / "b.bin").write_bytes(b"\0" * 12)
    r = subprocess.run([str(EXE), str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    res = read_results(tmp_path)
    assert res["init_s"] > 0
    st = [c["status"] for c in res["cases"]]
    assert st == ["ok", "ok", "crashed"]
    assert "bytes for" in res["cases"][2]["note"]
    pl = orc.Plan(N, W, q)
    for i, (a, b) in enumerate(cases):
        got = read_answer(tmp_path, i, B, N, L)
        for k in range(B):
            np.testing.assert_array_equal(got[k], pl.product(a[k], b[k]))


def test_call_bench_usage():
    r = subprocess.run([str(PKG_DIR / "lib" / "fherma_call_bench")], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
def test_fherma_run_multi_chunk():
    """fherma_run at N=65,536 with B=5 moves three chunks through the staging
    ring with several output pieces each; products 0 and 4 must equal their
    B=1 runs, and the owned output must have exactly B*N*L words."""
    import json
    r = subprocess.run([str(PKG_DIR / "lib" / "fherma_call_bench"), "5", "65536", "1747", "3", "2"],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["exact"] is True and len(res["ms"]) == 2
