"""The C++ host layer (mpt_amd/include/mpt_amd/*.hpp over include/mptg.h):
headers compile standalone, and the examples (planar-arm PRRT, SE(3)
rigid-body PRRT*, the holonomic 2-D point demo with its Scenario's own C++
valid()/link()) plan on the GPU; the last one's trees equal the oracle's."""
import dataclasses
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
EXAMPLE = ROOT / "mpt_amd" / "examples" / "planar_arm.bin"
EXAMPLE_STAR = ROOT / "mpt_amd" / "examples" / "rigid_body_star.bin"
EXAMPLE_HOST = ROOT / "mpt_amd" / "examples" / "holonomic_host.bin"


def test_headers_compile_standalone(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('#include "mpt_amd/planner.hpp"\n#include "mpt_amd/scenarios.hpp"\n'
                   "using P = mpt_amd::Planner<mpt_amd::LinkManipulator<20>,\n"
                   "    mpt_amd::PRRT<mpt_amd::report_stats<true>, mpt_amd::gpu_batch<4096>>>;\n"
                   "static_assert(P::kDim == 20);\n"
                   "using Q = mpt_amd::PRRT<mpt_amd::gpu_batch<64>, mpt_amd::check_order<MPTG_ORDER_SURROGATE>>;\n"
                   "static_assert(Q::batch == 64 && Q::order == MPTG_ORDER_SURROGATE && !Q::stats);\n"
                   "using G = mpt_amd::PRRT<mpt_amd::graph_iterations<8>, mpt_amd::gpu_batch<64>>;\n"
                   "static_assert(G::graph == 8 && G::batch == 64 && mpt_amd::PRRT<>::graph == 0);\n"
                   "using PG = mpt_amd::Planner<mpt_amd::LinkManipulator<5>, G>;\n"
                   "static_assert(!PG::kStar);\n"
                   "using S = mpt_amd::PRRTStar<mpt_amd::rewire_k<20>, mpt_amd::gpu_batch<512>,\n"
                   "    mpt_amd::report_stats<true>>;\n"
                   "static_assert(S::k == 20 && S::batch == 512 && S::stats);\n"
                   "static_assert(mpt_amd::PRRTStar<>::k == 0 && mpt_amd::PRRTStar<>::batch == 8192);\n"
                   "using R = mpt_amd::Planner<mpt_amd::RigidBodyMeshes, S>;\n"
                   "static_assert(R::kStar && R::kDim == 7 && !R::kHostScene);\n"
                   "static_assert(mpt_amd::PRRT<mpt_amd::max_threads<3>>::threads == 3 && mpt_amd::PRRT<>::threads == 0);\n"
                   "static_assert(mpt_amd::PRRTStar<mpt_amd::max_threads<5>, mpt_amd::rewire_k<8>>::threads == 5);\n"
                   "int main() { return 0; }\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-I", str(ROOT / "include"), "-I",
                        str(ROOT / "mpt_amd" / "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_example_is_built():
    assert EXAMPLE.exists() and EXAMPLE_STAR.exists() and EXAMPLE_HOST.exists(), "run __graft_entry__.build()"


@pytest.mark.gpu
def test_example_plans_on_gpu():
    print("planar arm example", flush=True)
    r = subprocess.run([str(EXAMPLE), "3"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], flush=True)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "path_valid=1" in r.stdout and "solved=1" in r.stdout
    assert "graph_equal=1" in r.stdout


@pytest.mark.gpu
def test_rigid_body_star_example_plans_on_gpu():
    print("rigid body PRRT* example", flush=True)
    r = subprocess.run([str(EXAMPLE_STAR), "5"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], flush=True)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "path_valid=1" in r.stdout and "solved=1" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("mode,batch,target,radius", [("prrt1", 1, 600, 1e-6), ("prrt1", 1, 1500, 40.0),
                                                      ("prrt256", 256, 8000, 40.0), ("star64", 64, 3000, 1e-6),
                                                      ("star64", 64, 5000, 40.0)])
def test_host_scenario_example_equals_oracle(tmp_path, mode, batch, target, radius):
    """A Scenario with only C++ valid()/link() (no gpu_scene()) plans through
    mpt_amd::Planner; its tree (and PRRT* path costs) equal the oracle's on
    the same demo scenario, and so does the solution it extracts (waypoints
    and both callback shapes).  A goal radius of 40 finds tens of goals, so
    the choice among them is tested; in PRRT* 12 of 33 goal nodes are rewired
    after they were found."""
    import oracle
    from mpt_amd import scenes
    print("holonomic host-scenario example", mode, flush=True)
    out = tmp_path / "tree.bin"
    r = subprocess.run([str(EXAMPLE_HOST), mode, "7", str(target), "60", "0.05", str(out), repr(radius)],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], flush=True)
    assert r.returncode == 0, r.stderr[-2000:]
    raw = out.read_bytes()
    n = int(np.frombuffer(raw[:8], dtype=np.int64)[0])
    states = np.frombuffer(raw[8:8 + 16 * n], dtype=np.float64).reshape(n, 2)
    parents = np.frombuffer(raw[8 + 16 * n:8 + 20 * n], dtype=np.int32)
    sc = dataclasses.replace(scenes.shapes_scene(), range=60.0, goal_bias=0.05, goal_radius=radius)
    o = oracle.Oracle(sc)
    pos = 8 + 20 * n
    if mode == "star64":
        costs = np.frombuffer(raw[pos:pos + 8 * n], dtype=np.float64)
        pos += 8 * n
        on, op, oc, og, ost, oss = o.prrt_star(7, batch, target, k=20)
        assert np.array_equal(costs, oc)
        path_ids = best_path_star(op, oss["goal_cost0"], oss["goal_parent0"], og)
    else:
        on, op, og, ost = o.prrt(7, batch, target)
        path_ids = best_path_prrt(on, op, og)
    assert np.array_equal(states, on) and np.array_equal(parents, op)
    assert f"link_calls={ost['link_calls']} " in r.stdout
    # the extracted solution equals the reference's choice on the oracle's tree
    npath = int(np.frombuffer(raw[pos:pos + 8], dtype=np.int64)[0])
    path = np.frombuffer(raw[pos + 8:pos + 8 + 16 * npath], dtype=np.float64).reshape(npath, 2)
    edges = int(np.frombuffer(raw[pos + 8 + 16 * npath:pos + 16 + 16 * npath], dtype=np.int64)[0])
    assert len(og) > 0 and npath == len(path_ids) > 1
    assert np.array_equal(path, on[path_ids])
    assert edges == npath - 1 and "callbacks_agree=1" in r.stdout


def best_path_prrt(nodes, parents, goals):
    """PRRT::bestSolution / pathCost / solution (src/mpt/impl/prrt/prrt.hpp:208-234,
    261-273): the goal of least path cost (summed from the goal up, the first
    found on ties), its path from the start."""
    best, best_cost = None, None
    for g in goals:
        cost, n = 0.0, int(g)
        while parents[n] >= 0:
            d = nodes[n] - nodes[parents[n]]
            cost += float(np.sqrt(d[0] * d[0] + d[1] * d[1]))
            n = int(parents[n])
        if best is None or cost < best_cost:
            best, best_cost = int(g), cost
    path, n = [], best
    while n >= 0:
        path.append(n)
        n = int(parents[n])
    return path[::-1]


def best_path_star(parents, cost0, parent0, goals):
    """PRRTStar::foundGoal / solution (src/mpt/impl/prrt_star/prrt_star.hpp:
    203-216, 324-336): the goal edge found with the least cost at the time
    (the first on ties); the path leaves the goal by the parent it had then
    and follows the current tree above it."""
    best = None
    for i in range(len(goals)):
        if best is None or cost0[i] < cost0[best]:
            best = i
    path, n = [int(goals[best])], int(parent0[best])
    while n >= 0:
        path.append(n)
        n = int(parents[n])
    return path[::-1]


@pytest.mark.gpu
def test_scenario_sampler_example(tmp_path):
    """A Scenario with its own sample(rng) (the left half of the canvas, no
    sample for one draw in ten) drives the planner's samples: every node lies
    in the left half, and the slots without a sample are counted nowhere
    (samples = iterations x batch > zero + invalid + link calls)."""
    print("scenario sampler example", flush=True)
    out = tmp_path / "tree.bin"
    r = subprocess.run([str(EXAMPLE_HOST), "sampled256", "9", "3000", "60", "0", str(out)], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout[-2000:], flush=True)
    assert r.returncode == 0, r.stderr[-2000:]
    raw = out.read_bytes()
    n = int(np.frombuffer(raw[:8], dtype=np.int64)[0])
    states = np.frombuffer(raw[8:8 + 16 * n], dtype=np.float64).reshape(n, 2)
    assert n >= 3000 and (states[:, 0] <= 512.0).all() and (states[:, 1] <= 512.0).all()
    st = dict(kv.split("=") for kv in r.stdout.split("\n")[0].split())
    samples, iters = int(st["samples"]), int(st["iterations"])
    used = int(st["zero"]) + int(st["invalid"]) + int(st["link_calls"])
    assert samples == 256 * iters and 0.05 * samples < samples - used < 0.15 * samples
