"""The C++ Numerics adapter (include/peanoclaw_amd_numerics.hpp).

CPU: the header compiles against a patch type with a float heap (the
reference's SWE builds store float) and a custom access policy; the recorded
call trace of an adaptive-grid run is well formed; the replay program fails
loudly without a device.  GPU: the replay program runs the recorded calls of
a two-level grid (global steps, then local time stepping) and of a grid that
is refined and coarsened (demanded mesh widths, new patches, switches to
leaves, released patches) through the adapter on heap-backed patches, and
its heap contents and clocks equal the oracle's bitwise."""
import os
import subprocess

import numpy as np
import pytest

from adapter_trace import REPLAY, RecordingBackend, replay

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

FLOAT_HEAP = r"""
#include <array>
#include <vector>
#include "peanoclaw_amd_numerics.hpp"

struct Clock {
  double t = 0, dt = 0, est = 0;
  double getCurrentTime() const { return t; }
  void setCurrentTime(double v) { t = v; }
  double getTimestepSize() const { return dt; }
  void setTimestepSize(double v) { dt = v; }
  double getEstimatedNextTimestepSize() const { return est; }
  void setEstimatedNextTimestepSize(double v) { est = v; }
  double getMinimalNeighborTimeConstraint() const { return 1e300; }
};

struct FloatPatch {
  int id = 0, n = 8;
  Clock clock;
  std::vector<float> u = std::vector<float>(3 * 64 + 4 * 100);
};

// a policy for a patch type without the Peano member names
struct FloatAccess {
  static int key(FloatPatch& p) { return p.id; }
  static int dataIndex(FloatPatch& p) { return p.id; }
  static bool isVirtual(FloatPatch&) { return false; }
  static Clock& timeIntervals(FloatPatch& p) { return p.clock; }
  static float* uNew(FloatPatch& p) { return p.u.data(); }
  static float* uOld(FloatPatch& p, int q) { return p.u.data() + 192 + 100 * q; }
  static float* bathymetry(FloatPatch& p) { return p.u.data() + 492; }
  static void describe(FloatPatch& p, pc_cell_description& d) {
    d.position[0] = d.position[1] = 0.0;
    d.size[0] = d.size[1] = 1.0;
    d.subdivision_factor[0] = d.subdivision_factor[1] = p.n;
    d.ghostlayer_width = 1;
    d.level = 1;
    d.is_virtual = 0;
    d.time = p.clock.t;
    d.timestep_size = p.clock.dt;
  }
};

template class peanoclaw_amd::SWENumerics<FloatPatch, FloatAccess>;

void use(peanoclaw_amd::SWENumerics<FloatPatch, FloatAccess>& num, FloatPatch& a, FloatPatch& b) {
  std::array<int, 2> sz{1, 8}, so{7, 0}, d{-1, 0};
  std::array<bool, 4> f{};
  num.initializePatch(a);
  num.transferGhostlayer(sz, so, d, b, a);
  num.interpolateSolution(sz, d, b, a, true, false, true);
  num.restrictSolution(a, b, false);
  num.postProcessRestriction(b, false);
  num.fillBoundaryLayer(a, 0, false);
  num.copyUNewToUOld(a);
  num.solveTimestep(a, 1.0, false, f);
  (void)num.cfl(a);
  (void)num.getDemandedMeshWidth(a, false);
  num.addPatchToSolution(a);
  num.destroyPatch(a);
}
"""


def test_adapter_header_compiles_for_a_float_heap(tmp_path):
    src = tmp_path / "float_heap.cpp"
    src.write_text(FLOAT_HEAP)
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def _two_level_run(rec, steps=12, lts_iterations=40):
    from peanoclaw_amd.amr import TwoLevelGrid, dam_break_fields
    fn = dam_break_fields(centre=(0.47, 0.53), radius=0.25,
                          hills=[(0.3, 0.6, 0.1, 0.25), (0.7, 0.35, 0.08, 0.2)])
    refined = {(1, 1), (1, 2), (2, 1), (4, 4), (3, 2)}
    g = TwoLevelGrid(rec, 6, 6, 8, refined, fn)
    for _ in range(steps):
        g.step()
    its, leaf_steps = g.run_lts(g.time + g.dt + 0.004, max_iterations=lts_iterations)
    return g, its, leaf_steps


def test_recorded_trace_is_well_formed(tmp_path):
    rec = RecordingBackend()
    g, its, leaf_steps = _two_level_run(rec, steps=2, lts_iterations=3)
    trace, blob = rec.write(str(tmp_path))
    ops = [l.split() for l in open(trace).read().splitlines()]
    arity = {"N": 9, "S": 2, "V": 2, "Z": 1, "R": 2, "P": 1, "C": 1, "T": 8, "I": 9, "W": 3,
             "F": 2, "K": 3, "X": 2, "D": 6, "U": 4, "B": 2, "Q": 1}
    assert ops and all(o[0] in arity and len(o) == 1 + arity[o[0]] for o in ops)
    kinds = {o[0] for o in ops}
    # both level transfers, walls, restriction and the LTS clock checks occur
    assert {"N", "S", "Z", "R", "P", "C", "T", "I", "W", "F", "K", "X"} <= kinds
    # hex floats round-trip exactly
    for o in ops:
        if o[0] == "X":
            assert float.fromhex(o[2]) == float.fromhex(float.fromhex(o[2]).hex())
    n_blob = os.path.getsize(blob) // 8
    assert n_blob == sum(n * n + (n + 2) ** 2 for n in rec.sizes)
    assert leaf_steps > 0


def test_replay_fails_loudly_without_a_device(tmp_path):
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is visible")
    if not os.path.exists(REPLAY):
        pytest.skip("replay program not built (python -m peanoclaw_amd.build)")
    rec = RecordingBackend()
    _two_level_run(rec, steps=1, lts_iterations=1)
    r, _ = replay(rec, str(tmp_path))
    assert r.returncode == 2
    assert "pc_pool_create" in r.stderr


@pytest.mark.gpu
def test_adapter_replay_two_level_bitwise(tmp_path):
    """12 global steps and up to 40 local-time-stepping iterations of a
    6x6x8^2 grid with 5 refined patches: every call through the C++ adapter
    on heap-backed patches; heap contents, clocks and every intermediate CFL
    value equal to the oracle's."""
    assert os.path.exists(REPLAY), "replay program not built (python -m peanoclaw_amd.build)"
    rec = RecordingBackend()
    g, its, leaf_steps = _two_level_run(rec)
    r, out = replay(rec, str(tmp_path))
    print(r.stdout.strip(), r.stderr.strip()[-2000:])
    assert r.returncode == 0, r.stderr
    res = rec.read_result(out)
    same = total = 0
    for h, P in enumerate(rec.p):
        clk, u = res[h]
        ref = np.array(P.unew)
        same += int((u.view(np.uint64) == ref.view(np.uint64)).sum())
        total += u.size
        if not P.s.is_virtual:
            assert clk[0] == P.s.time and clk[1] == P.s.dt, h
    print(f"adapter replay: {len(rec.ops)} calls, {its} LTS iterations ({leaf_steps} leaf steps); "
          f"bitwise {same}/{total}")
    assert same == total


def _regrid_run(rec):
    """A 6 x 6 x 9^2 grid without refinement: a regrid that refines the
    dam-break front, 12 steps, a regrid with a threshold no gradient reaches
    (coarsens every patch whose children are old enough), 6 more steps."""
    from peanoclaw_amd.amr import AmrGrid, dam_break_fields
    fn = dam_break_fields(centre=(0.47, 0.53), radius=0.25, hills=[(0.3, 0.6, 0.1, 0.2)])
    g = AmrGrid(rec, 6, 6, 9, lambda lev, ij: False, fn, max_level=2)
    refined = g.regrid()
    g.run(12)
    coarsened = g.regrid(threshold=1e9)
    g.run(6)
    return g, refined, coarsened


def test_recorded_regrid_trace_has_remesh_calls(tmp_path):
    rec = RecordingBackend()
    g, refined, coarsened = _regrid_run(rec)
    assert refined[0] > 0 and coarsened[1] > 0
    kinds = {l.split()[0] for l in rec.ops}
    assert {"D", "U", "B", "Q", "V"} <= kinds


@pytest.mark.gpu
def test_adapter_replay_regrid_bitwise(tmp_path):
    """Remesh through the adapter: demanded mesh widths on the device equal
    the oracle's, new patches registered with b from the initial data and
    interpolated from their parent, coarsened patches restricted and
    switched to leaves, released children destroyed; the heap contents of
    every live patch equal the oracle's bitwise."""
    assert os.path.exists(REPLAY), "replay program not built (python -m peanoclaw_amd.build)"
    rec = RecordingBackend()
    g, refined, coarsened = _regrid_run(rec)
    r, out = replay(rec, str(tmp_path))
    print(r.stdout.strip(), r.stderr.strip()[-2000:])
    assert r.returncode == 0, r.stderr
    res = rec.read_result(out)
    same = total = 0
    for h, (clk, u) in res.items():
        ref = np.array(rec.p[h].unew)
        same += int((u.view(np.uint64) == ref.view(np.uint64)).sum())
        total += u.size
    print(f"adapter regrid replay: refined {refined}, coarsened {coarsened}, "
          f"{len(rec.released)} patches released; bitwise {same}/{total}")
    assert same == total
