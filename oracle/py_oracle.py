"""ctypes wrapper of the FP64 CPU oracle (liboracle.so, prefix orc_) and of the
reference compiled from source (_ref/libmmcref.so, prefix ref_).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent / "falling-dice_amd"))
import mmc  # noqa: E402  (struct definitions of the C-ABI only)

ORACLE_LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libmmcref.so"

_libs: dict = {}
FRAME_CB = C.CFUNCTYPE(C.c_int, C.POINTER(mmc.Frame), C.c_int32, C.c_void_p)


def _load(prefix: str):
    if prefix in _libs:
        return _libs[prefix]
    path = ORACLE_LIB if prefix == "orc" else REF_LIB
    if not path.exists():
        raise FileNotFoundError(f"{path} not built (make -C oracle)")
    L = C.CDLL(str(path))
    P, dp = C.c_void_p, C.POINTER(C.c_double)
    i32p = C.POINTER(C.c_int32)

    def f(name, args, res=C.c_int):
        fn = getattr(L, f"{prefix}_{name}")
        fn.argtypes, fn.restype = args, res
        return fn
    f("create", [C.POINTER(mmc.Config)], P)
    f("create_lattice", [C.POINTER(mmc.Config), C.POINTER(mmc.Lattice)], P)
    f("destroy", [P], None)
    f("last_error", [P], C.c_char_p)
    f("info", [P, i32p, i32p, C.POINTER(C.c_int64), dp], None)
    f("lattice", [P, dp, dp, dp, C.POINTER(C.c_uint8), i32p, i32p, dp, dp, dp, dp, dp, dp,
                  C.POINTER(C.c_uint64), i32p, i32p, i32p], None)
    f("objective", [P, dp, C.c_double, C.c_int32, C.POINTER(mmc.ObjectiveOut), dp, dp], C.c_double)
    f("minimize", [P, C.c_double, dp, C.POINTER(mmc.OptStats), C.POINTER(mmc.ObjectiveOut)])
    f("commit", [P])
    f("run", [P, C.POINTER(mmc.Frame), C.c_int32, i32p, C.c_int32])
    f("beam_potential", [P, dp, dp], C.c_double)
    if prefix == "ref":
        f("export_stl", [P, C.c_char_p, C.c_int32])
    if prefix == "orc":
        f("energy_change", [P, dp, dp, C.c_double, dp], None)
        f("set_frame_callback", [P, FRAME_CB, P], None)
    f("capture", [P, C.c_double, dp, dp, P], None)
    f("rigid_rotation_test", [P, dp])
    f("gradient_check", [P, dp])
    f("read_state", [P, dp, dp, dp], None)
    f("set_plastic", [P, dp], None)
    f("set_pose", [P, dp, dp], None)
    f("scalar_return", [C.c_double] * 6 + [dp], None)
    f("vector_return", [dp, dp] + [C.c_double] * 4 + [dp], None)
    f("closest_segments", [dp, dp, dp, dp, dp], None)
    f("quat_exp", [dp, dp], None)
    f("quat_log", [dp, dp], None)
    f("jl_inv_t", [dp, dp, dp], None)
    _libs[prefix] = L
    return L


def available(prefix: str) -> bool:
    return (ORACLE_LIB if prefix == "orc" else REF_LIB).exists()


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


class Solver:
    """A CPU solver instance: prefix 'orc' (the oracle) or 'ref' (the reference)."""

    def __init__(self, cfg: mmc.Config, prefix: str = "orc", lattice: mmc.LatticeArrays | None = None):
        self.L = _load(prefix)
        self.prefix = prefix
        self.cfg = cfg
        self._fn = lambda n: getattr(self.L, f"{prefix}_{n}")
        if lattice is None:
            self.h = self._fn("create")(C.byref(cfg))
        else:
            ls = lattice.struct()
            self.h = self._fn("create_lattice")(C.byref(cfg), C.byref(ls))
        err = self._fn("last_error")(self.h)
        if err:
            raise RuntimeError(err.decode())
        nn, nb, nc = C.c_int32(), C.c_int32(), C.c_int64()
        env = np.zeros(6)
        self._fn("info")(self.h, C.byref(nn), C.byref(nb), C.byref(nc), _dp(env))
        self.n_nodes, self.n_beams, self.n_connected = nn.value, nb.value, nc.value
        self.width, self.height, self.depth, self.area, self.mass, self.top0 = env

    def close(self):
        if getattr(self, "h", None):
            self._fn("destroy")(self.h)
            self.h = None

    __del__ = close

    def lattice(self) -> mmc.LatticeArrays:
        N, B = self.n_nodes, self.n_beams
        a = dict(x=np.zeros(3 * N), q=np.zeros(4 * N), mass=np.zeros(N),
                 locked=np.zeros(6 * N, np.uint8), a=np.zeros(B, np.int32), b=np.zeros(B, np.int32),
                 L0=np.zeros(B), radius=np.zeros(B), area=np.zeros(B), inertia=np.zeros(B),
                 e0=np.zeros(3 * B), plastic=np.zeros(7 * B),
                 connected=np.zeros(self.n_connected, np.uint64),
                 node_layer=np.zeros(N, np.int32), beam_layer=np.zeros(B, np.int32),
                 beam_family=np.zeros(B, np.int32))
        i32 = lambda v: v.ctypes.data_as(C.POINTER(C.c_int32))  # noqa: E731
        self._fn("lattice")(self.h, _dp(a["x"]), _dp(a["q"]), _dp(a["mass"]),
                            a["locked"].ctypes.data_as(C.POINTER(C.c_uint8)), i32(a["a"]), i32(a["b"]),
                            _dp(a["L0"]), _dp(a["radius"]), _dp(a["area"]), _dp(a["inertia"]),
                            _dp(a["e0"]), _dp(a["plastic"]),
                            a["connected"].ctypes.data_as(C.POINTER(C.c_uint64)),
                            i32(a["node_layer"]), i32(a["beam_layer"]), i32(a["beam_family"]))
        return mmc.LatticeArrays(width=self.width, height=self.height, depth=self.depth, **a)

    def objective(self, z, top, trials=False):
        z = np.ascontiguousarray(z, dtype=np.float64)
        out = mmc.ObjectiveOut()
        g = np.zeros(6 * self.n_nodes)
        t = np.zeros(mmc.TRIAL_STRIDE * self.n_beams) if trials else None
        self._fn("objective")(self.h, _dp(z), top, int(trials), C.byref(out), _dp(g), _dp(t))
        if trials:
            return out, g, t.reshape(self.n_beams, mmc.TRIAL_STRIDE)
        return out, g

    def energy_change(self, z0, z1, top):
        """(per-term sum, total difference) of the energy change z0 -> z1."""
        out = np.zeros(2)
        self._fn("energy_change")(self.h, _dp(np.ascontiguousarray(z0, dtype=np.float64)),
                                  _dp(np.ascontiguousarray(z1, dtype=np.float64)), top, _dp(out))
        return out[0], out[1]

    def minimize(self, top, z=None):
        z = np.zeros(6 * self.n_nodes) if z is None else np.array(z, dtype=np.float64)
        st, fin = mmc.OptStats(), mmc.ObjectiveOut()
        self._fn("minimize")(self.h, top, _dp(z), C.byref(st), C.byref(fin))
        return st, fin, z

    def commit(self):
        rc = self._fn("commit")(self.h)
        if rc:
            raise RuntimeError(self._fn("last_error")(self.h).decode())

    def run(self, verbose=False, cap=None, on_frame=None, limit=None):
        """The reference schedule (tare, loading, unloading).  on_frame(frames)
        is called with the frames so far after every accepted frame (oracle
        only); the run ends after `limit` frames when that is given."""
        cap = cap or 4 * (self.cfg.load_steps + self.cfg.unload_steps + 1) + 4096
        frames = (mmc.Frame * cap)()
        n = C.c_int32()
        cb = None
        if on_frame is not None or limit is not None:
            if self.prefix != "orc":
                raise ValueError("frame observers are an oracle feature")
            acc = []

            def observe(frame, count, _user):
                acc.append(np.frombuffer(C.string_at(frame, C.sizeof(mmc.Frame)), dtype=mmc.FRAME_DTYPE).copy())
                if on_frame is not None:
                    on_frame(np.concatenate(acc))
                return 1 if (limit is not None and count >= limit) else 0
            cb = FRAME_CB(observe)
            self._fn("set_frame_callback")(self.h, cb, None)
        rc = self._fn("run")(self.h, frames, cap, C.byref(n), int(verbose))
        if cb is not None:
            self._fn("set_frame_callback")(self.h, FRAME_CB(), None)
        arr = np.frombuffer(bytes(frames)[: min(n.value, cap) * C.sizeof(mmc.Frame)],
                            dtype=mmc.FRAME_DTYPE).copy()
        if rc:
            raise RuntimeError(self._fn("last_error")(self.h).decode())
        return arr

    def beam_potential(self, x, q):
        return self._fn("beam_potential")(self.h, _dp(np.ascontiguousarray(x, np.float64)),
                                          _dp(np.ascontiguousarray(q, np.float64)))

    def export_stl(self, path, segments=12):
        """The reference's Simulator::exportSTL (reference build only)."""
        if self._fn("export_stl")(self.h, str(path).encode(), int(segments)) != 0:
            raise RuntimeError(self._fn("last_error")(self.h).decode())

    def capture(self, top, prev_x=None, prev_q=None):
        """Simulator::capture of the committed state (mmc.SNAPSHOT_DTYPE):
        velocities from the previous poses (node-major x 3N, q 4N) when given."""
        out = np.zeros(self.n_beams + self.n_nodes + mmc.MACHINE_BODIES, mmc.SNAPSHOT_DTYPE)
        px = None if prev_x is None else np.ascontiguousarray(prev_x, np.float64)
        pq = None if prev_q is None else np.ascontiguousarray(prev_q, np.float64)
        self._fn("capture")(self.h, top, _dp(px), _dp(pq), out.ctypes.data_as(C.c_void_p))
        return out

    def rigid_rotation_test(self):
        e = C.c_double()
        p = self._fn("rigid_rotation_test")(self.h, C.byref(e))
        return bool(p), e.value

    def gradient_check(self):
        e = C.c_double()
        p = self._fn("gradient_check")(self.h, C.byref(e))
        return bool(p), e.value

    def read_state(self):
        x, q, p = np.zeros(3 * self.n_nodes), np.zeros(4 * self.n_nodes), np.zeros(7 * self.n_beams)
        self._fn("read_state")(self.h, _dp(x), _dp(q), _dp(p))
        return x, q, p

    def set_plastic(self, p):
        self._fn("set_plastic")(self.h, _dp(np.ascontiguousarray(p, np.float64)))

    def set_pose(self, x, q):
        self._fn("set_pose")(self.h, _dp(np.ascontiguousarray(x, np.float64)),
                             _dp(np.ascontiguousarray(q, np.float64)))


def helper(prefix: str, name: str, *args):
    """Calls one of the closed-form helpers (scalar_return, vector_return,
    closest_segments, quat_exp, quat_log, jl_inv_t) and returns its output array."""
    L = _load(prefix)
    fn = getattr(L, f"{prefix}_{name}")
    sizes = {"scalar_return": 7, "vector_return": 12, "closest_segments": 8, "quat_exp": 4,
             "quat_log": 3, "jl_inv_t": 3}
    out = np.zeros(sizes[name])
    conv = [(_dp(np.ascontiguousarray(a, np.float64)) if isinstance(a, (list, tuple, np.ndarray))
             else C.c_double(a)) for a in args]
    fn(*conv, _dp(out))
    return out
