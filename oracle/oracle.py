"""ctypes binding of the CPU oracle (oracle/libbfm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product package
(pscbfm_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import subprocess
from ctypes import POINTER, c_char_p, c_int, c_int32, c_uint8, c_uint32, c_uint64, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libbfm_oracle.so"


class OrcSystem(ctypes.Structure):
    _fields_ = [
        ("box", c_int32 * 3),
        ("periodic", c_int32 * 3),
        ("n", c_uint32),
        ("nbr_off", POINTER(c_uint32)),
        ("nbr_idx", POINTER(c_uint32)),
        ("color", POINTER(c_uint32)),
        ("n_colors", c_uint32),
        ("forbidden", POINTER(c_uint8)),
        ("seed", c_uint64),
        ("rng", c_uint32),
    ]


RNG_INT_HASH = 1   # the reference's default Rngs::Hash (Rng::IntHash)
RNG_PHILOX = 3     # Rng::Philox


_lib = None


def build() -> None:
    """Compiles the oracle with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        lib = ctypes.CDLL(str(LIB_PATH))
        lib.orc_philox4x32_10.argtypes = [c_void_p, c_void_p, c_void_p]
        lib.orc_philox4x32_10.restype = None
        lib.orc_move_word.argtypes = [c_uint64, c_uint64, c_uint32]
        lib.orc_move_word.restype = c_uint32
        lib.orc_int_hash.argtypes = [c_uint32]
        lib.orc_int_hash.restype = c_uint32
        lib.orc_substep_seed.argtypes = [c_uint64, c_uint64, c_uint32]
        lib.orc_substep_seed.restype = c_uint32
        lib.orc_direction_word.argtypes = [POINTER(OrcSystem), c_uint64, c_uint32, c_uint32]
        lib.orc_direction_word.restype = c_uint32
        lib.orc_color_order.argtypes = [c_uint64, c_uint64, c_uint32, c_void_p]
        lib.orc_color_order.restype = None
        lib.orc_standard_bondset.argtypes = [c_void_p]
        lib.orc_standard_bondset.restype = None
        lib.orc_check_system.argtypes = [POINTER(OrcSystem), c_void_p, c_char_p, c_int]
        lib.orc_check_system.restype = c_int
        lib.orc_graph_coloring.argtypes = [c_uint32, c_void_p, c_void_p, c_int, c_void_p, POINTER(c_uint32)]
        lib.orc_graph_coloring.restype = c_int
        lib.orc_sweep.argtypes = [POINTER(OrcSystem), c_void_p, c_uint64, c_uint32, c_int,
                                  POINTER(c_uint64), POINTER(c_uint64)]
        lib.orc_sweep.restype = c_int
        lib.orc_sweep_parallel.argtypes = [POINTER(OrcSystem), c_void_p, c_uint64, c_uint32, c_int,
                                           POINTER(c_uint64), POINTER(c_uint64)]
        lib.orc_sweep_parallel.restype = c_int
        lib.orc_random_sequential.argtypes = [POINTER(OrcSystem), c_void_p, c_uint64, c_uint32,
                                              POINTER(c_uint64)]
        lib.orc_random_sequential.restype = c_int
        _lib = lib
    return _lib


def philox(ctr, key) -> list[int]:
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    load().orc_philox4x32_10(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return [int(v) for v in out]


def move_word(seed: int, age: int, mono_id: int) -> int:
    return int(load().orc_move_word(seed, age, mono_id))


def int_hash(a: int) -> int:
    return int(load().orc_int_hash(a & 0xFFFFFFFF))


def substep_seed(seed: int, age: int, substep: int) -> int:
    return int(load().orc_substep_seed(seed & 0xFFFFFFFFFFFFFFFF, age, substep))


def direction_word(sys: "System", age: int, substep: int, mono_id: int) -> int:
    return int(load().orc_direction_word(ctypes.byref(sys.struct), age, substep, mono_id))


def color_order(seed: int, age: int, n_colors: int) -> list[int]:
    out = np.zeros(n_colors, dtype=np.uint32)
    load().orc_color_order(seed, age, n_colors, out.ctypes.data)
    return [int(v) for v in out]


def standard_bondset() -> np.ndarray:
    f = np.zeros(512, dtype=np.uint8)
    load().orc_standard_bondset(f.ctypes.data)
    return f


def csr(n: int, bonds: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Both directions of every undirected bond, grouped by monomer."""
    b = np.asarray(bonds, dtype=np.int64).reshape(-1, 2)
    src = np.concatenate([b[:, 0], b[:, 1]])
    dst = np.concatenate([b[:, 1], b[:, 0]])
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    off = np.zeros(n + 1, dtype=np.uint32)
    np.add.at(off, src + 1, 1)
    off = np.cumsum(off).astype(np.uint32)
    return off, dst.astype(np.uint32)


def directed_csr(n: int, links: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Links as given (source -> target), grouped by source in the given order."""
    b = np.asarray(links, dtype=np.int64).reshape(-1, 2)
    order = np.argsort(b[:, 0], kind="stable")
    src, dst = b[order, 0], b[order, 1]
    off = np.zeros(n + 1, dtype=np.uint32)
    np.add.at(off, src + 1, 1)
    off = np.cumsum(off).astype(np.uint32)
    return off, dst.astype(np.uint32)


def call_order_links(system: dict) -> np.ndarray:
    """Directed links in the order the LeMonADE-style set-up issues them
    (GPUScBFM_AB_Type.h:159-161 sets both directions of every bond): the
    system's "links" as given, else (a, b), (b, a) for every bond (a, b) in
    turn -- the order pscbfm_amd.updater_from_system uses."""
    if system.get("links") is not None:
        return np.asarray(system["links"], dtype=np.int64).reshape(-1, 2)
    b = np.asarray(system["bonds"], dtype=np.int64).reshape(-1, 2)
    both = np.empty((2 * b.shape[0], 2), dtype=np.int64)
    both[0::2] = b
    both[1::2] = b[:, ::-1]
    return both


class System:
    """Holds the arrays behind an OrcSystem so they outlive the struct.

    Without `colors`, the monomers are coloured by the oracle's restatement of
    the reference's colouring over the links in call order."""

    def __init__(self, system: dict, seed: int, colors: np.ndarray | None = None, rng: int = RNG_INT_HASH):
        self.xyz0 = np.ascontiguousarray(system["xyz"], dtype=np.int32).reshape(-1, 3)
        n = self.xyz0.shape[0]
        if system.get("links") is not None:
            # directed links: a bond is tested only by the monomer listing it
            self.off, self.idx = directed_csr(n, system["links"])
        else:
            self.off, self.idx = csr(n, system["bonds"])
        if colors is None:
            c_off, c_idx = directed_csr(n, call_order_links(system))
            colors, nc = graph_coloring(n, c_off, c_idx)
        else:
            colors = np.ascontiguousarray(colors, dtype=np.uint32)
            nc = int(colors.max()) + 1 if n else 1
        self.colors = colors
        self.forbidden = np.ascontiguousarray(system["forbidden"], dtype=np.uint8)
        s = OrcSystem()
        for a in range(3):
            s.box[a] = int(system["box"][a])
            s.periodic[a] = int(system.get("periodic", (1, 1, 1))[a])
        s.n = n
        s.nbr_off = self.off.ctypes.data_as(POINTER(c_uint32))
        s.nbr_idx = self.idx.ctypes.data_as(POINTER(c_uint32))
        s.color = self.colors.ctypes.data_as(POINTER(c_uint32))
        s.n_colors = nc
        s.forbidden = self.forbidden.ctypes.data_as(POINTER(c_uint8))
        s.seed = seed & 0xFFFFFFFFFFFFFFFF
        s.rng = rng
        self.struct = s
        self.n_colors = nc


def graph_coloring(n: int, off: np.ndarray, idx: np.ndarray, uniform: bool = True) -> tuple[np.ndarray, int]:
    """The reference's colouring (graphColoring.tpp) of the directed link
    lists (off, idx) in call order; uniform = the rebalance the updater uses."""
    off = np.ascontiguousarray(off, dtype=np.uint32)
    idx = np.ascontiguousarray(idx, dtype=np.uint32) if len(idx) else np.zeros(1, dtype=np.uint32)
    col = np.zeros(max(n, 1), dtype=np.uint32)
    nc = c_uint32()
    rc = load().orc_graph_coloring(n, off.ctypes.data, idx.ctypes.data, int(bool(uniform)), col.ctypes.data,
                                   ctypes.byref(nc))
    if rc != 0:
        raise RuntimeError("oracle colouring needs more than 255 colours")
    return col[:n], nc.value


def check_system(sys: System, xyz: np.ndarray) -> tuple[int, str]:
    a = np.ascontiguousarray(xyz, dtype=np.int32)
    buf = ctypes.create_string_buffer(512)
    rc = load().orc_check_system(ctypes.byref(sys.struct), a.ctypes.data, buf, 512)
    return rc, buf.value.decode()


def sweep(sys: System, xyz: np.ndarray, age0: int, n_mcs: int, form: int = 0):
    """Returns (new xyz, accepted, proposals)."""
    a = np.ascontiguousarray(xyz, dtype=np.int32).copy()
    acc, prop = c_uint64(), c_uint64()
    rc = load().orc_sweep(ctypes.byref(sys.struct), a.ctypes.data, age0, n_mcs, form,
                          ctypes.byref(acc), ctypes.byref(prop))
    if rc != 0:
        raise RuntimeError(f"orc_sweep failed with {rc}")
    return a, acc.value, prop.value


def sweep_parallel(sys: System, xyz: np.ndarray, age0: int, n_mcs: int, threads: int = 0):
    """The corner-lattice sweep with OpenMP loops; returns (new xyz, accepted, proposals)."""
    a = np.ascontiguousarray(xyz, dtype=np.int32).copy()
    acc, prop = c_uint64(), c_uint64()
    rc = load().orc_sweep_parallel(ctypes.byref(sys.struct), a.ctypes.data, age0, n_mcs, threads,
                                   ctypes.byref(acc), ctypes.byref(prop))
    if rc != 0:
        raise RuntimeError(f"orc_sweep_parallel failed with {rc}")
    return a, acc.value, prop.value


def random_sequential(sys: System, xyz: np.ndarray, age0: int, n_mcs: int):
    a = np.ascontiguousarray(xyz, dtype=np.int32).copy()
    acc = c_uint64()
    rc = load().orc_random_sequential(ctypes.byref(sys.struct), a.ctypes.data, age0, n_mcs, ctypes.byref(acc))
    if rc != 0:
        raise RuntimeError(f"orc_random_sequential failed with {rc}")
    return a, acc.value
