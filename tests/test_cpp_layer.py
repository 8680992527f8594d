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
@pytest.mark.parametrize("mode,batch,target", [("prrt1", 1, 600), ("prrt256", 256, 5000), ("star64", 64, 3000)])
def test_host_scenario_example_equals_oracle(tmp_path, mode, batch, target):
    """A Scenario with only C++ valid()/link() (no gpu_scene()) plans through
    mpt_amd::Planner; its tree (and PRRT* path costs) equal the oracle's on
    the same demo scenario."""
    import oracle
    from mpt_amd import scenes
    print("holonomic host-scenario example", mode, flush=True)
    out = tmp_path / "tree.bin"
    r = subprocess.run([str(EXAMPLE_HOST), mode, "7", str(target), "60", "0.05", str(out)], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout[-2000:], flush=True)
    assert r.returncode == 0, r.stderr[-2000:]
    raw = out.read_bytes()
    n = int(np.frombuffer(raw[:8], dtype=np.int64)[0])
    states = np.frombuffer(raw[8:8 + 16 * n], dtype=np.float64).reshape(n, 2)
    parents = np.frombuffer(raw[8 + 16 * n:8 + 20 * n], dtype=np.int32)
    sc = dataclasses.replace(scenes.shapes_scene(), range=60.0, goal_bias=0.05)
    o = oracle.Oracle(sc)
    if mode == "star64":
        costs = np.frombuffer(raw[8 + 20 * n:8 + 28 * n], dtype=np.float64)
        on, op, oc, og, ost, oss = o.prrt_star(7, batch, target, k=20)
        assert np.array_equal(costs, oc)
    else:
        on, op, og, ost = o.prrt(7, batch, target)
    assert np.array_equal(states, on) and np.array_equal(parents, op)
    assert f"link_calls={ost['link_calls']} " in r.stdout
