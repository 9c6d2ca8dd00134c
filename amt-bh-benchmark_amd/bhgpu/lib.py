"""ctypes binding of libbhgpu.so (include/bh_gpu.h).

Loading fails loudly when the library has not been built: there is no
fallback path.
"""
import ctypes
import os

import numpy as np

from .layout import FORCE_PARAMS, PARTICLE, SYSTEM, WALK_NODE

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libbhgpu.so")

BH_OK = 0
STATUS = {
    0: "BH_OK",
    1: "BH_ERR_INVALID",
    2: "BH_ERR_NONFINITE",
    3: "BH_ERR_INCOMPLETE",
    4: "BH_ERR_OOM",
    5: "BH_ERR_HIP",
    6: "BH_ERR_TOO_LARGE",
    7: "BH_ERR_NO_TREE",
}
BH_FP32, BH_FP64 = 0, 1
BH_WALK_GROUPED, BH_WALK_PER_TARGET = 0, 1
BH_LAYOUT_REFERENCE, BH_LAYOUT_BFS = 0, 1
BH_GRAPH_OFF, BH_GRAPH_ON = 0, 1

# Every entry point of include/bh_gpu.h with its ctypes signature.
_P = ctypes.c_void_p
_U64 = ctypes.c_uint64
_U32 = ctypes.c_uint32
_D = ctypes.c_double
_I = ctypes.c_int
SIGNATURES = {
    "bh_ctx_create": ([ctypes.POINTER(_P), _I], _I),
    "bh_ctx_destroy": ([_P], None),
    "bh_last_error": ([_P], ctypes.c_char_p),
    "bh_set_precision": ([_P, _I], _I),
    "bh_set_walk_mode": ([_P, _I], _I),
    "bh_set_graph_mode": ([_P, _I], _I),
    "bh_set_walk_tile": ([_P, _U32], _I),
    "bh_debug_walk_stack_limit": ([_P, _U32], _I),
    "bh_graph_count": ([_P, ctypes.POINTER(_U32)], _I),
    "bh_max_particles": ([], _U64),
    "bh_step": ([_P, _P, _U64, _P, _U32, _U32, _U32, _D, _D, _P], _I),
    "bh_step_device": ([_P, _P, _P, _U64, _P, _U32, _U32, _U32, _D, _D, _P], _I),
    "bh_step_device_noalloc": ([_P, _P, _P, _U64, _P, _U32, _U32, _U32, _D, _D, _P], _I),
    "bh_step_batch": ([_P, _P, _P, _P, _U32, _P], _I),
    "bh_step_batch_host": ([_P, _P, _P, _U32, _P], _I),
    "bh_state_step_batch": ([_P, _P, _U32, _P], _I),
    "bh_state_upload": ([_P, _P, _U64], _I),
    "bh_state_step": ([_P, _P, _U32, _U32, _U32, _D, _D, _P], _I),
    "bh_state_download": ([_P, _P, _U64], _I),
    "bh_state_device_ptr": ([_P], _P),
    "bh_build_tree": ([_P, _P, _U64, _U32, _U32], _I),
    "bh_tree_sizes": ([_P, _P, _P, _P], _I),
    "bh_tree_export": ([_P, _I, _P, _P, _P, _P, _P, _P, _P], _I),
    "bh_accel": ([_P, _P, _U32, _P, _P, _P, _P], _I),
    "bh_direct_sum": ([_P, _P, _U64, _P, _U64, _D, _D, _P], _I),
    "bh_last_step_stats": ([_P, _P], _I),
    "bh_energies": ([_P, _P, _U64, _D, _P, _P], _I),
    "bh_read_csv": ([ctypes.c_char_p, ctypes.POINTER(_P), ctypes.POINTER(_U64), ctypes.c_char_p, ctypes.c_size_t], _I),
    "bh_free": ([_P], None),
    "bh_write_vtk": ([ctypes.c_char_p, _P, _U64, ctypes.c_char_p, ctypes.c_size_t], _I),
    "bh_server_create": ([ctypes.POINTER(_P), _P, _U64, _U32], _I),
    "bh_server_destroy": ([_P], None),
    "bh_server_last_error": ([_P], ctypes.c_char_p),
    "bh_server_footprint": ([_P], _U64),
    "bh_server_submit": ([_P, _P, _P, _P, ctypes.POINTER(_U64)], _I),
    "bh_server_step": ([_P, _P], _I),
    "bh_server_poll": ([_P, _P, _U32, ctypes.POINTER(_U32)], _I),
    "bh_server_pending": ([_P, ctypes.POINTER(_U32), ctypes.POINTER(_U32)], _I),
    "bh_server_clock": ([_P], _D),
}

_lib = None


def load(path=LIB_PATH):
    """The library handle; raises if libbhgpu.so is missing or incomplete."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built: run __graft_entry__.build() or make -C amt-bh-benchmark_amd")
        lib = ctypes.CDLL(path)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


class BHError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class StepTiming(ctypes.Structure):
    _fields_ = [("tree_build_ms", ctypes.c_double), ("traversal_ms", ctypes.c_double), ("exchange_ms", ctypes.c_double)]


class StepStats(ctypes.Structure):
    _fields_ = [("sort_ms", ctypes.c_double), ("tree_ms", ctypes.c_double), ("targets_ms", ctypes.c_double),
                ("walk_ms", ctypes.c_double), ("n_pp", ctypes.c_uint64), ("n_pc", ctypes.c_uint64),
                ("n_nodes", ctypes.c_uint64), ("n_leaves", ctypes.c_uint64), ("n_targets", ctypes.c_uint64),
                ("step_ms", ctypes.c_double), ("n_units", ctypes.c_uint64), ("graph", ctypes.c_uint32),
                ("reruns", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ForceParamsC(ctypes.Structure):
    _fields_ = [("theta", ctypes.c_double), ("eps2", ctypes.c_double), ("G", ctypes.c_double),
                ("use_quadrupole", ctypes.c_int32), ("pad", ctypes.c_int32)]


class Request(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("fp", ForceParamsC), ("max_leaf", ctypes.c_uint32),
                ("group_size", ctypes.c_uint32), ("target_size", ctypes.c_uint32), ("steps", ctypes.c_uint32),
                ("dt", ctypes.c_double)]


class Completion(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_uint64), ("submit_s", ctypes.c_double), ("admit_s", ctypes.c_double),
                ("done_s", ctypes.c_double), ("evaluations", ctypes.c_uint32), ("status", ctypes.c_int32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("active_systems", ctypes.c_uint32), ("pending", ctypes.c_uint32), ("admitted", ctypes.c_uint32),
                ("completed", ctypes.c_uint32), ("stepped_systems", ctypes.c_uint32), ("pad", ctypes.c_uint32),
                ("active_particles", ctypes.c_uint64),
                ("stepped_particles", ctypes.c_uint64), ("used_bytes", ctypes.c_uint64), ("step_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def request(n, theta=0.5, eps2=0.0, G=1.0, use_quadrupole=True, max_leaf=16, group_size=256, target_size=32,
            steps=1, dt=1e-3):
    return Request(n, ForceParamsC(theta, eps2, G, 1 if use_quadrupole else 0, 0), max_leaf, group_size, target_size,
                   steps, dt)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _fp(fp):
    if isinstance(fp, np.ndarray) and fp.dtype == FORCE_PARAMS:
        return fp
    raise TypeError("force params must come from bhgpu.force_params()")


def read_csv(path):
    """read_csv (io.cc:122-178) through the library: Particle records in file order."""
    lib = load()
    buf, n = _P(), _U64()
    err = ctypes.create_string_buffer(512)
    rc = lib.bh_read_csv(os.fsencode(path), ctypes.byref(buf), ctypes.byref(n), err, len(err))
    if rc != BH_OK:
        raise BHError(rc, err.value.decode())
    if n.value == 0:
        return np.zeros(0, PARTICLE)
    try:
        raw = ctypes.string_at(buf.value, n.value * PARTICLE.itemsize)
        return np.frombuffer(raw, dtype=PARTICLE).copy()
    finally:
        lib.bh_free(buf)


def write_vtk(path, particles):
    """write_vtk (io.cc:180-195) through the library."""
    lib = load()
    p = np.ascontiguousarray(particles, dtype=PARTICLE)
    err = ctypes.create_string_buffer(512)
    rc = lib.bh_write_vtk(os.fsencode(path), _ptr(p), len(p), err, len(err))
    if rc != BH_OK:
        raise BHError(rc, err.value.decode())


class Context:
    """A bh_ctx on one GPU."""

    def __init__(self, device=0):
        self._lib = load()
        h = _P()
        rc = self._lib.bh_ctx_create(ctypes.byref(h), device)
        if rc != BH_OK:
            raise BHError(rc, f"bh_ctx_create(device={device}) failed")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.bh_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != BH_OK:
            raise BHError(rc, self._lib.bh_last_error(self._h).decode())

    def set_precision(self, precision):
        self._check(self._lib.bh_set_precision(self._h, precision))

    def set_walk_mode(self, mode):
        self._check(self._lib.bh_set_walk_mode(self._h, mode))

    def set_graph_mode(self, mode):
        self._check(self._lib.bh_set_graph_mode(self._h, mode))

    def set_walk_tile(self, tile):
        self._check(self._lib.bh_set_walk_tile(self._h, tile))

    def debug_walk_stack_limit(self, levels):
        self._check(self._lib.bh_debug_walk_stack_limit(self._h, levels))

    def graph_count(self):
        k = _U32()
        self._check(self._lib.bh_graph_count(self._h, ctypes.byref(k)))
        return k.value

    def step(self, p, fp, max_leaf, group_size, target_size, kick_dt, drift_dt):
        """step_particles: returns (particles in Morton order, StepTiming)."""
        out = np.array(p, dtype=PARTICLE, copy=True)
        t = StepTiming()
        self._check(self._lib.bh_step(self._h, _ptr(out), len(out), _ptr(_fp(fp)), max_leaf, group_size, target_size,
                                      kick_dt, drift_dt, ctypes.byref(t)))
        return out, t

    def step_device(self, d_in, d_out, n, fp, max_leaf, group_size, target_size, kick_dt, drift_dt):
        t = StepTiming()
        self._check(self._lib.bh_step_device(self._h, d_in, d_out, n, _ptr(_fp(fp)), max_leaf, group_size, target_size,
                                             kick_dt, drift_dt, ctypes.byref(t)))
        return t

    def step_device_noalloc(self, d_in, d_out, n, fp, max_leaf, group_size, target_size, kick_dt, drift_dt):
        """bh_step_device_noalloc: the step with its scratch in the library's fixed per-GPU array."""
        t = StepTiming()
        self._check(self._lib.bh_step_device_noalloc(self._h, d_in, d_out, n, _ptr(_fp(fp)), max_leaf, group_size,
                                                     target_size, kick_dt, drift_dt, ctypes.byref(t)))
        return t

    def step_batch(self, p, systems):
        """bh_step_batch_host: systems (SYSTEM records) own consecutive ranges of
        p; returns (particles, each system in its own Morton order, StepTiming)."""
        systems = np.ascontiguousarray(systems, dtype=SYSTEM)
        out = np.array(p, dtype=PARTICLE, copy=True)
        if int(systems["n"].sum()) != len(out):
            raise ValueError("system sizes do not add up to the particle count")
        t = StepTiming()
        self._check(self._lib.bh_step_batch_host(self._h, _ptr(out), _ptr(systems), len(systems), ctypes.byref(t)))
        return out, t

    def step_batch_device(self, d_in, d_out, systems):
        systems = np.ascontiguousarray(systems, dtype=SYSTEM)
        t = StepTiming()
        self._check(self._lib.bh_step_batch(self._h, d_in, d_out, _ptr(systems), len(systems), ctypes.byref(t)))
        return t

    def state_upload(self, p):
        p = np.ascontiguousarray(p, dtype=PARTICLE)
        self._check(self._lib.bh_state_upload(self._h, _ptr(p), len(p)))
        self._state_n = len(p)

    def state_step(self, fp, max_leaf, group_size, target_size, kick_dt, drift_dt):
        t = StepTiming()
        self._check(self._lib.bh_state_step(self._h, _ptr(_fp(fp)), max_leaf, group_size, target_size, kick_dt, drift_dt,
                                            ctypes.byref(t)))
        return t

    def state_step_batch(self, systems):
        systems = np.ascontiguousarray(systems, dtype=SYSTEM)
        t = StepTiming()
        self._check(self._lib.bh_state_step_batch(self._h, _ptr(systems), len(systems), ctypes.byref(t)))
        return t

    def state_download(self):
        out = np.zeros(self._state_n, PARTICLE)
        self._check(self._lib.bh_state_download(self._h, _ptr(out), len(out)))
        return out

    def state_device_ptr(self):
        return self._lib.bh_state_device_ptr(self._h)

    def build_tree(self, p, max_leaf, group_size):
        p = np.ascontiguousarray(p, dtype=PARTICLE)
        self._check(self._lib.bh_build_tree(self._h, _ptr(p), len(p), max_leaf, group_size))
        self._tree_n = len(p)

    def tree_sizes(self):
        nn, ng = _U64(), _U64()
        dom = np.zeros(4)
        self._check(self._lib.bh_tree_sizes(self._h, ctypes.byref(nn), ctypes.byref(ng), _ptr(dom)))
        return nn.value, ng.value, dom

    def export(self, layout=BH_LAYOUT_REFERENCE):
        """The last built tree, as the oracle's export (reference node layout by default)."""
        m, g, dom = self.tree_sizes()
        n = self._tree_n
        out = dict(
            keys=np.zeros(n, np.uint64),
            perm=np.zeros(n, np.uint32),
            particles=np.zeros(n, PARTICLE),
            nodes=np.zeros(m, WALK_NODE),
            center=np.zeros((m, 3)),
            iab=np.zeros((m, 6)),
            groups=np.zeros(g, np.uint32),
        )
        self._check(self._lib.bh_tree_export(self._h, layout, _ptr(out["keys"]), _ptr(out["perm"]), _ptr(out["particles"]),
                                             _ptr(out["nodes"]), _ptr(out["center"]), _ptr(out["iab"]),
                                             _ptr(out["groups"])))
        out["domain"] = dom
        return out

    def export_perm(self):
        """Sorted position -> input index of the last built tree."""
        perm = np.zeros(self._tree_n, np.uint32)
        self._check(self._lib.bh_tree_export(self._h, BH_LAYOUT_BFS, None, _ptr(perm), None, None, None, None, None))
        return perm

    def accel(self, fp, target_size):
        """run_grouped_pass on the last tree: (acc, n_pp, n_pc, totals), Morton order."""
        n = self._tree_n
        acc = np.zeros((n, 3))
        npp = np.zeros(n, np.uint32)
        npc = np.zeros(n, np.uint32)
        tot = np.zeros(2, np.uint64)
        self._check(self._lib.bh_accel(self._h, _ptr(_fp(fp)), target_size, _ptr(acc), _ptr(npp), _ptr(npc), _ptr(tot)))
        return acc, npp, npc, tot

    def last_step_stats(self):
        st = StepStats()
        self._check(self._lib.bh_last_step_stats(self._h, ctypes.byref(st)))
        return st

    def energies(self, p, G=1.0, potential=True):
        """(E_K, E_P) of host particles; E_P is the O(n^2) FP64 pair sum (None if not asked)."""
        p = np.ascontiguousarray(p, dtype=PARTICLE)
        ek, ep = ctypes.c_double(), ctypes.c_double()
        self._check(self._lib.bh_energies(self._h, _ptr(p), len(p), G, ctypes.byref(ek),
                                          ctypes.byref(ep) if potential else None))
        return ek.value, (ep.value if potential else None)

    def direct_sum(self, p, idx, eps2, G=1.0):
        p = np.ascontiguousarray(p, dtype=PARTICLE)
        if idx is None:
            m = len(p)
            acc = np.zeros((m, 3))
            self._check(self._lib.bh_direct_sum(self._h, _ptr(p), len(p), None, 0, eps2, G, _ptr(acc)))
        else:
            idx = np.ascontiguousarray(idx, np.uint32)
            acc = np.zeros((len(idx), 3))
            self._check(self._lib.bh_direct_sum(self._h, _ptr(p), len(p), _ptr(idx), len(idx), eps2, G, _ptr(acc)))
        return acc


class Server:
    """bh_server: continuous batching of independent leapfrog runs on one context."""

    def __init__(self, ctx, mem_cap_bytes=200 * 10 ** 9, max_systems=65535):
        self._lib = load()
        self.ctx = ctx
        h = ctypes.c_void_p()
        rc = self._lib.bh_server_create(ctypes.byref(h), ctx._h, mem_cap_bytes, max_systems)
        if rc != BH_OK:
            raise BHError(rc, "bh_server_create failed")
        self._h = h
        self._results = {}  # ticket -> (result array, keep-alive)

    def close(self):
        if self._h:
            self._lib.bh_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != BH_OK:
            raise BHError(rc, self._lib.bh_server_last_error(self._h).decode())

    def footprint(self, req):
        return self._lib.bh_server_footprint(ctypes.byref(req))

    def submit(self, particles, req):
        """Queues a request; returns its ticket. The result array is handed
        out by poll()."""
        p = np.ascontiguousarray(particles, dtype=PARTICLE)
        out = np.zeros(len(p), PARTICLE)
        t = ctypes.c_uint64()
        self._check(self._lib.bh_server_submit(self._h, _ptr(p), _ptr(out), ctypes.byref(req), ctypes.byref(t)))
        self._results[t.value] = out
        return t.value

    def step(self):
        st = ServerStats()
        self._check(self._lib.bh_server_step(self._h, ctypes.byref(st)))
        return st

    def poll(self, max_items=4096):
        """[(ticket, result particles, Completion)] completed since the last poll."""
        buf = (Completion * max_items)()
        cnt = ctypes.c_uint32()
        self._check(self._lib.bh_server_poll(self._h, buf, max_items, ctypes.byref(cnt)))
        return [(buf[i].ticket, self._results.pop(buf[i].ticket), buf[i]) for i in range(cnt.value)]

    def pending(self):
        p, a = ctypes.c_uint32(), ctypes.c_uint32()
        self._check(self._lib.bh_server_pending(self._h, ctypes.byref(p), ctypes.byref(a)))
        return p.value, a.value

    def clock(self):
        return self._lib.bh_server_clock(self._h)
