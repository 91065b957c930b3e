"""Recording oracle backend (test infrastructure): runs the adaptive-grid
drivers of peanoclaw_amd.amr on the oracle and writes every Numerics call
they make as a trace that tests/cpp/numerics_replay replays through the C++
adapter (include/peanoclaw_amd_numerics.hpp) on the GPU."""
import os
import subprocess

import numpy as np

from amr_oracle import OracleBackend

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPLAY = os.path.join(ROOT, "tests", "cpp", "numerics_replay")


def hx(v):
    return float(v).hex()


class RecordingBackend(OracleBackend):
    def __init__(self):
        super().__init__()
        self.ops = []
        self.blob = []
        self.nblob = 0
        self.sizes = []
        self.released = set()

    def new_patch(self, pos, size, n, level, virtual, time=0.0, dt=0.0):
        h = super().new_patch(pos, size, n, level, virtual, time, dt)
        self.sizes.append(n)
        self.ops.append(f"N {h} {level} {int(bool(virtual))} {n} {hx(pos[0])} {hx(pos[1])} "
                        f"{hx(size)} {hx(time)} {hx(dt)}")
        return h

    def set_state(self, s, h, b):
        super().set_state(s, h, b)
        self.ops.append(f"S {s} {self._blob(h, b)}")

    def set_virtual(self, s, virtual):
        super().set_virtual(s, virtual)
        self.ops.append(f"V {s} {int(bool(virtual))}")

    def clear(self, s):
        super().clear(s)
        self.ops.append(f"Z {s}")

    def restrict(self, src, dst):
        super().restrict(src, dst)
        self.ops.append(f"R {src} {dst}")

    def post_restrict(self, dst):
        super().post_restrict(dst)
        self.ops.append(f"P {dst}")

    def copy_unew_to_uold(self, s):
        super().copy_unew_to_uold(s)
        self.ops.append(f"C {s}")

    def transfer(self, size, soff, doff, src, dst):
        super().transfer(size, soff, doff, src, dst)
        self.ops.append(f"T {size[0]} {size[1]} {soff[0]} {soff[1]} {doff[0]} {doff[1]} "
                        f"{src} {dst}")

    def interpolate(self, size, doff, src, dst, to_uold, to_current, use_time):
        super().interpolate(size, doff, src, dst, to_uold, to_current, use_time)
        self.ops.append(f"I {size[0]} {size[1]} {doff[0]} {doff[1]} {src} {dst} "
                        f"{int(to_uold)} {int(to_current)} {int(use_time)}")

    def fill_boundary(self, s, dim, upper):
        super().fill_boundary(s, dim, upper)
        self.ops.append(f"W {s} {dim} {int(upper)}")

    def cfl(self, s):
        v = super().cfl(s)
        self.ops.append(f"F {s} {hx(v)}")
        return v

    def clock(self, s):
        t, dt = super().clock(s)
        self.ops.append(f"K {s} {hx(t)} {hx(dt)}")
        return t, dt

    def solve(self, s, max_dt):
        self.ops.append(f"X {s} {hx(max_dt)}")
        return super().solve(s, max_dt)

    def flags(self, slots, threshold, minimal_width, maximal_width):
        # the demanded mesh width of every patch, which the replay checks
        import ctypes

        from oracle_lib import lib
        for s in slots:
            dw = (ctypes.c_double * 2)()
            lib().or_demanded_mesh_width(self.p[s].ref, threshold, minimal_width, maximal_width,
                                         dw)
            self.ops.append(f"D {s} {hx(threshold)} {hx(minimal_width)} {hx(maximal_width)} "
                            f"{hx(dw[0])} {hx(dw[1])}")
        return super().flags(slots, threshold, minimal_width, maximal_width)

    def _blob(self, *arrays):
        off = self.nblob
        for a in arrays:
            a = np.ascontiguousarray(a, dtype=np.float64).ravel()
            self.blob.append(a)
            self.nblob += a.size
        return off

    def switch_to_leaf(self, s, time, dt):
        super().switch_to_leaf(s, time, dt)
        P = self.p[s]
        off = self._blob(np.array(P.uold), np.array(P.unew))
        self.ops.append(f"U {s} {off} {hx(time)} {hx(dt)}")

    def set_b(self, s, b):
        super().set_b(s, b)
        self.ops.append(f"B {s} {self._blob(b)}")

    def release(self, s):
        super().release(s)
        self.ops.append(f"Q {s}")
        self.released.add(s)

    def write(self, directory):
        trace = os.path.join(directory, "trace.txt")
        blob = os.path.join(directory, "blob.bin")
        with open(trace, "w") as f:
            f.write("\n".join(self.ops) + "\n")
        (np.concatenate(self.blob) if self.blob else np.zeros(0)).tofile(blob)
        return trace, blob

    def read_result(self, path):
        """{handle: (clock (3,), uNew (3, n, n))} as written by the replay."""
        raw = np.fromfile(path, dtype=np.float64)
        out, o = {}, 0
        for h, n in enumerate(self.sizes):
            if h in self.released:
                continue
            clk = raw[o:o + 3]
            o += 3
            u = raw[o:o + 3 * n * n].reshape(3, n, n)
            o += 3 * n * n
            out[h] = (clk, u)
        if o != raw.size:
            raise ValueError("replay output has an unexpected size")
        return out


def replay(rec, directory, timeout=300):
    trace, blob = rec.write(directory)
    out = os.path.join(directory, "result.bin")
    r = subprocess.run([REPLAY, trace, blob, out], capture_output=True, text=True,
                       timeout=timeout)
    return r, out


if __name__ == "__main__":
    # Writes the recorded calls of the two-level run (12 global steps, up to
    # 40 local-time-stepping iterations) to DIR/trace.txt and DIR/blob.bin,
    # the inputs of: tests/cpp/numerics_replay DIR/trace.txt DIR/blob.bin OUT
    import sys

    sys.path.insert(0, ROOT)
    from test_numerics_adapter import _two_level_run
    d = sys.argv[1]
    os.makedirs(d, exist_ok=True)
    rec = RecordingBackend()
    _two_level_run(rec)
    rec.write(d)
    print(f"{len(rec.ops)} calls on {len(rec.sizes)} patches written to {d}")
