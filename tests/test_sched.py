"""The server's wave scheduler (csrc/bytecrt_sched.hpp) against a simulated
backend on the CPU (csrc/sched_sim.cpp): ring regions never overlap and stay
in bounds, waves respect their size, plan-count and in-flight limits, admission
is strictly first-come-first-served, and every request is served."""
import json
import subprocess

import pytest

from conftest import PKG_DIR


@pytest.fixture(scope="module")
def sim(tmp_path_factory):
    exe = tmp_path_factory.mktemp("sched") / "sched_sim"
    csrc = PKG_DIR / "csrc"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", f"-I{csrc}", str(csrc / "sched_sim.cpp"),
                    "-o", str(exe)], check=True)
    return exe


# seed, requests, plans, arena, pinned in, pinned out, max_wave, max_groups, max_inflight, arrivals/tick
CASES = [
    (1, 3000, 8, 4096, 2048, 1024, 128, 16, 4, 3.0),    # mixed sizes, wraparound with several waves live
    (2, 3000, 3, 100, 50, 25, 8, 2, 2, 1.0),            # tight rings, two plans per wave at most
    (3, 500, 7, 64, 32, 16, 1, 1, 1, 5.0),              # one request per wave, one wave at a time
    (4, 5000, 16, 1 << 20, 1 << 19, 1 << 18, 128, 16, 4, 40.0),  # overload: queue builds, waves fill
    (5, 2000, 12, 1000, 900, 100, 32, 4, 3, 0.3),       # light load, output ring the binding one
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"seed{c[0]}")
def test_scheduler_invariants(sim, case):
    r = subprocess.run([str(sim), *map(str, case)], capture_output=True, text=True, timeout=120)
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert r.returncode == 0 and out["ok"], out
    _, requests, _, *_, max_wave, _, max_inflight, _ = case
    assert out["requests"] == requests
    assert 1 <= out["max_wave"] <= max_wave
    assert 1 <= out["peak_inflight"] <= max_inflight


def run(sim, case):
    r = subprocess.run([str(sim), *map(str, case)], capture_output=True, text=True, timeout=120)
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_scheduler_batches_under_load(sim):
    """Under overload a wave takes everything that fits (many requests per
    wave); under mixed load several waves are in flight at once, wrapping
    round the rings."""
    over = run(sim, CASES[3])
    assert over["max_wave"] > 8 and over["waves"] < over["requests"] / 4
    mixed = run(sim, CASES[0])
    assert mixed["peak_inflight"] == 4 and mixed["max_wave"] > 1


def test_sim_usage(sim):
    r = subprocess.run([str(sim)], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
