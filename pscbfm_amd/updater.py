"""Python face of the updater: the reference's UpdaterGPUScBFM_AB_Type API
(src/pscbfm/UpdaterGPUScBFM_AB_Type.h:453-514) and its LeMonADE adapter
GPUScBFM_AB_Type (src/pscbfm/GPUScBFM_AB_Type.h:27-286), over the C ABI.
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_int32, c_int64, c_uint32, c_void_p

import numpy as np

from . import _capi

# Random number generators, numbered as the reference's Rng enum
# (UpdaterGPUScBFM_AB_Type.h:413-420).
Rng_IntHash = 1   # the reference's default (.cu:1407)
Rng_Philox = 3



def _mcs(n_mcs: int) -> int:
    """An MCS count for the C ABI's uint32_t (a negative count would wrap)."""
    n = int(n_mcs)
    if n < 0 or n > 0xFFFFFFFF:
        raise ValueError(f"MCS count {n_mcs} out of range [0, 2^32)")
    return n

class UpdaterGPUScBFM_AB_Type:
    """Same setters, call order and errors as the reference updater.

    Errors: InvalidArgument (a ValueError) where the reference throws
    std::invalid_argument, BfmError where it throws std::runtime_error.
    """

    def __init__(self):
        self._lib = _capi.load()
        h = c_void_p()
        _capi.check(self._lib.bfm_create(byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.bfm_destroy(h)
            self._h = None

    # ---- reference setters --------------------------------------------------
    def setGpu(self, gpu: int) -> None:
        _capi.check(self._lib.bfm_set_gpu(self._h, gpu))

    def copyBondSet(self, dx: int, dy: int, dz: int, forbidden: bool) -> None:
        _capi.check(self._lib.bfm_copy_bond_set(self._h, dx, dy, dz, int(bool(forbidden))))

    def setLatticeSize(self, x: int, y: int, z: int) -> None:
        _capi.check(self._lib.bfm_set_lattice_size(self._h, x, y, z))

    def setNrOfAllMonomers(self, n: int) -> None:
        _capi.check(self._lib.bfm_set_nr_of_all_monomers(self._h, n))

    def setAttribute(self, i: int, attribute: int) -> None:
        _capi.check(self._lib.bfm_set_attribute(self._h, i, attribute))

    def setMonomerCoordinates(self, i: int, x: int, y: int, z: int) -> None:
        _capi.check(self._lib.bfm_set_monomer_coordinates(self._h, i, x, y, z))

    def setConnectivity(self, i: int, j: int) -> None:
        _capi.check(self._lib.bfm_set_connectivity(self._h, i, j))

    def setPeriodicity(self, px: bool, py: bool, pz: bool) -> None:
        _capi.check(self._lib.bfm_set_periodicity(self._h, int(px), int(py), int(pz)))

    def setAge(self, age: int) -> None:
        _capi.check(self._lib.bfm_set_age(self._h, age))

    def getAge(self) -> int:
        v = c_int64()
        _capi.check(self._lib.bfm_get_age(self._h, byref(v)))
        return v.value

    def setRng(self, rng: int) -> None:
        _capi.check(self._lib.bfm_set_rng(self._h, rng))

    def getRng(self) -> int:
        v = ctypes.c_int()
        _capi.check(self._lib.bfm_get_rng(self._h, byref(v)))
        return v.value

    def setSeed(self, seed: int) -> None:
        _capi.check(self._lib.bfm_set_seed(self._h, seed & 0xFFFFFFFFFFFFFFFF))

    def setStepsBetweenSortings(self, steps: int) -> None:
        _capi.check(self._lib.bfm_set_steps_between_sortings(self._h, steps))

    def getStepsBetweenSortings(self) -> int:
        v = c_int64()
        _capi.check(self._lib.bfm_get_steps_between_sortings(self._h, byref(v)))
        return v.value

    def initialize(self) -> None:
        _capi.check(self._lib.bfm_initialize(self._h))

    def runSimulationOnGPU(self, n_mcs: int) -> None:
        _capi.check(self._lib.bfm_run_simulation_on_gpu(self._h, _mcs(n_mcs)))

    def runSimulationOnGPUNoAlloc(self, n_mcs: int) -> None:
        """runSimulationOnGPU without any allocation: the copy-back is staged in
        the library's static device scratch (at most bfm_noalloc_max_monomers
        monomers; calls that overlap on one GPU race)."""
        _capi.check(self._lib.bfm_run_simulation_on_gpu_noalloc(self._h, _mcs(n_mcs)))

    def getMonomerPositionInX(self, i: int) -> int:
        v = c_int32()
        _capi.check(self._lib.bfm_get_monomer_position_in_x(self._h, i, byref(v)))
        return v.value

    def getMonomerPositionInY(self, i: int) -> int:
        v = c_int32()
        _capi.check(self._lib.bfm_get_monomer_position_in_y(self._h, i, byref(v)))
        return v.value

    def getMonomerPositionInZ(self, i: int) -> int:
        v = c_int32()
        _capi.check(self._lib.bfm_get_monomer_position_in_z(self._h, i, byref(v)))
        return v.value

    def cleanup(self) -> None:
        _capi.check(self._lib.bfm_cleanup(self._h))

    # ---- extensions ---------------------------------------------------------
    def validate(self) -> None:
        _capi.check(self._lib.bfm_validate(self._h))

    def setMonomerCoordinatesBulk(self, xyz: np.ndarray) -> None:
        a = np.ascontiguousarray(xyz, dtype=np.int32).reshape(-1, 3)
        _capi.check(self._lib.bfm_set_monomer_coordinates_bulk(self._h, a.ctypes.data, a.shape[0]))

    def setConnectivityBulk(self, pairs: np.ndarray) -> None:
        a = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        _capi.check(self._lib.bfm_set_connectivity_bulk(self._h, a.ctypes.data, a.shape[0]))

    def getMonomerPositions(self, n: int) -> np.ndarray:
        out = np.empty((n, 3), dtype=np.int32)
        _capi.check(self._lib.bfm_get_monomer_positions(self._h, out.ctypes.data, n))
        return out

    def getColors(self, n: int) -> tuple[np.ndarray, int]:
        out = np.empty(n, dtype=np.uint32)
        nc = c_uint32()
        _capi.check(self._lib.bfm_get_colors(self._h, out.ctypes.data, n, byref(nc)))
        return out, nc.value

    def getStats(self) -> dict:
        st = _capi.BfmStats()
        _capi.check(self._lib.bfm_get_stats(self._h, byref(st)))
        return {name: getattr(st, name) for name, _ in _capi.BfmStats._fields_}

    def setKernelTiming(self, every: int) -> None:
        """Time the kernels of every `every`-th substep with HIP events (0: off)."""
        _capi.check(self._lib.bfm_set_kernel_timing(self._h, int(every)))

    def readBfmFile(self, path: str) -> None:
        """Configures the (not yet initialized) updater from a LeMonADE BFM
        file: header, bond-vector set, attributes, bonds and the last
        conformation, whose !mcs value becomes the age."""
        _capi.check(self._lib.bfm_read_bfm_file(self._h, str(path).encode()))

    def writeBfmFile(self, path: str, append: bool = False) -> None:
        """Writes a new BFM file with the current state, or appends the
        current conformation as !mcs=<age>."""
        _capi.check(self._lib.bfm_write_bfm_file(self._h, str(path).encode(), int(bool(append))))

    def setSpatialSorting(self, on: bool) -> None:
        """Morton order of positions within each colour (default) or id order;
        the trajectory is the same either way.  Before initialize."""
        _capi.check(self._lib.bfm_set_spatial_sorting(self._h, int(bool(on))))

    def setPerformGroup(self, group: int) -> None:
        """Check waves per perform wave: 2, 4 (default) or 8."""
        _capi.check(self._lib.bfm_set_perform_group(self._h, int(group)))

    def autotune(self, probe_mcs: int, n_colors: int | None = None) -> dict:
        """Chooses the order within colours (Morton or id), then per colour
        segment the check workgroup size and perform group, and the XCD
        placement, from short timed probes, keeping the pick only if it also
        beats the defaults by wall time (the system advances by 23 * probe_mcs
        MCS when the pick is the default, else 27 * probe_mcs; four fewer
        without spatial sorting).  Returns the chosen configuration (per
        colour with n_colors)."""
        g, x = c_uint32(), ctypes.c_int()
        _capi.check(self._lib.bfm_autotune(self._h, int(probe_mcs), byref(g), byref(x)))
        out = {"perform_group": g.value, "xcd_ranges": bool(x.value)}
        if n_colors:
            cfg = [self.launchConfig(c) for c in range(n_colors)]
            out.update({"check_block": [c["check_block"] for c in cfg],
                        "check_items": [c["check_items"] for c in cfg],
                        "perform_groups": [c["perform_group"] for c in cfg],
                        "morton_order": cfg[0]["morton_order"]})
        return out

    def setCheckBlock(self, block: int) -> None:
        """Check workgroup size (64, 128, 256, 512 or 1024 monomers) for every colour segment."""
        _capi.check(self._lib.bfm_set_check_block(self._h, int(block)))

    def setCheckItems(self, items: int) -> None:
        """Monomers per lane of the check kernel (1 or 2) for every colour
        segment (default: 2 for segments of about 2 M monomers or more, else 1)."""
        _capi.check(self._lib.bfm_set_check_items(self._h, int(items)))

    def setSegmentConfig(self, colour: int, block: int, group: int, items: int | None = None) -> None:
        _capi.check(self._lib.bfm_set_segment_config(self._h, int(colour), int(block), int(group)))
        if items is not None:
            _capi.check(self._lib.bfm_set_segment_items(self._h, int(colour), int(items)))

    def launchConfig(self, colour: int) -> dict:
        t, g, x, m, it = c_uint32(), c_uint32(), ctypes.c_int(), ctypes.c_int(), c_uint32()
        _capi.check(self._lib.bfm_get_launch_config(self._h, int(colour), byref(t), byref(g), byref(x), byref(m)))
        _capi.check(self._lib.bfm_get_check_items(self._h, int(colour), byref(it)))
        return {"check_block": t.value, "check_items": it.value, "perform_group": g.value,
                "xcd_ranges": bool(x.value), "morton_order": bool(m.value)}

    def setXcdRanges(self, on: bool) -> None:
        """Contiguous block range per XCD (default) or round-robin placement;
        the trajectory is the same either way."""
        _capi.check(self._lib.bfm_set_xcd_ranges(self._h, int(bool(on))))

    def runResident(self, n_mcs: int) -> None:
        _capi.check(self._lib.bfm_run_resident(self._h, _mcs(n_mcs)))

    def copyBack(self) -> None:
        _capi.check(self._lib.bfm_copy_back(self._h))

    # ---- analyzers on the device (src/AnalyzerMsd.tpp:99-468) ----------------
    def observe(self) -> dict:
        """Bond-vector histogram (by bond-table index), Re^2, Rg^2 and the box
        centre of mass of the resident state, reduced on the device."""
        o = _capi.BfmObservables()
        _capi.check(self._lib.bfm_observe(self._h, byref(o)))
        return observables_dict(o)

    def chainTable(self) -> dict:
        """First, middle and last monomer of every linear chain, as the
        reference's AnalyzerMsd finds them (host only)."""
        n = c_uint32()
        _capi.check(self._lib.bfm_chain_table(self._h, byref(n), None, None, None, 0))
        out = {k: np.empty(n.value, dtype=np.uint32) for k in ("first", "middle", "last")}
        _capi.check(self._lib.bfm_chain_table(self._h, byref(n), out["first"].ctypes.data,
                                              out["middle"].ctypes.data, out["last"].ctypes.data, n.value))
        return out

    def chainSums(self, n_chains: int) -> np.ndarray:
        """Exact per-chain sums of the resident state (structured array)."""
        dt = np.dtype([("s", "<i8", 3), ("q", "<i8"), ("first", "<i4", 3), ("middle", "<i4", 3),
                       ("last", "<i4", 3), ("n", "<u4")])
        assert dt.itemsize == ctypes.sizeof(_capi.BfmChainSums)
        out = np.zeros(n_chains, dtype=dt)
        _capi.check(self._lib.bfm_get_chain_sums(self._h, out.ctypes.data, n_chains))
        return out

    def attachMsd(self, mcs_to_skip: int, delta_t_realization: int, delta_t_evaluate: int,
                  max_subsequences: int) -> None:
        """AnalyzerMsd(...) + initialize(): evaluates once at the current age."""
        _capi.check(self._lib.bfm_msd_attach(self._h, mcs_to_skip, delta_t_realization, delta_t_evaluate,
                                             max_subsequences))

    def executeMsd(self) -> None:
        """AnalyzerMsd::execute at the current age."""
        _capi.check(self._lib.bfm_msd_execute(self._h))

    def msdPoints(self) -> list[dict]:
        n = c_uint32()
        _capi.check(self._lib.bfm_msd_get_points(self._h, None, 0, byref(n)))
        buf = (_capi.BfmMsdPoint * max(n.value, 1))()
        _capi.check(self._lib.bfm_msd_get_points(self._h, buf, n.value, byref(n)))
        return [{"dt": p.dt, "n_values": p.n_values, "mean": list(p.mean), "sg": list(p.sg)}
                for p in buf[: n.value]]

    def writeMsd(self, path: str) -> None:
        """AnalyzerMsd::cleanup's output file."""
        _capi.check(self._lib.bfm_msd_write(self._h, str(path).encode()))


def observables_dict(o: "_capi.BfmObservables") -> dict:
    return {
        "bonds": o.bonds,
        "bond_hist": np.ctypeslib.as_array(o.bond_hist).astype(np.uint64).copy(),
        "n_chains": o.n_chains,
        "re2_sum": o.re2_sum,
        "re2": o.re2,
        "rg2": o.rg2,
        "box_com": np.array(o.box_com[:]),
    }


def updater_from_system(system: dict, seed: int, gpu: int | None = 0, age: int = 0,
                        validate_only: bool = False, spatial_sorting: bool = True,
                        rng: int = Rng_IntHash,
                        steps_between_sortings: int | None = None) -> UpdaterGPUScBFM_AB_Type:
    """Builds an updater the way the LeMonADE adapter does
    (GPUScBFM_AB_Type.h:120-184): box, periodicity, age, positions, both
    directions of every bond (or, if the system has "links", exactly those
    directed links), the 512-entry bond table, then initialize."""
    u = UpdaterGPUScBFM_AB_Type()
    if gpu is not None and not validate_only:
        u.setGpu(gpu)
    u.setAge(age)
    u.setSeed(seed)
    px, py, pz = system.get("periodic", (1, 1, 1))
    u.setPeriodicity(bool(px), bool(py), bool(pz))
    u.setLatticeSize(*[int(v) for v in system["box"]])
    xyz = np.asarray(system["xyz"], dtype=np.int32).reshape(-1, 3)
    u.setNrOfAllMonomers(xyz.shape[0])
    u.setMonomerCoordinatesBulk(xyz)
    if system.get("links") is not None:
        # directed links as given (setConnectivity is directed: a bond is
        # tested only by moves of the monomer that lists it)
        links = np.asarray(system["links"], dtype=np.uint32).reshape(-1, 2)
        if links.size:
            u.setConnectivityBulk(links)
    else:
        bonds = np.asarray(system["bonds"], dtype=np.uint32).reshape(-1, 2)
        if bonds.size:
            both = np.empty((2 * bonds.shape[0], 2), dtype=np.uint32)
            both[0::2] = bonds
            both[1::2] = bonds[:, ::-1]
            u.setConnectivityBulk(both)
    forbidden = np.asarray(system["forbidden"], dtype=np.uint8)
    for dx in range(-4, 4):
        for dy in range(-4, 4):
            for dz in range(-4, 4):
                idx = (dx & 7) | ((dy & 7) << 3) | ((dz & 7) << 6)
                u.copyBondSet(dx, dy, dz, bool(forbidden[idx]))
    u.setSpatialSorting(spatial_sorting)
    u.setRng(rng)
    if steps_between_sortings is not None:
        u.setStepsBetweenSortings(steps_between_sortings)
    if validate_only:
        u.validate()
    else:
        u.initialize()
    return u


class GPUScBFM_AB_Type:
    """LeMonADE-style updater adapter (GPUScBFM_AB_Type.h:27-286).

    `ingredients` needs: box (3 ints), periodic (3 bools), age (int, updated
    by execute), xyz (n x 3 int array, updated by execute), bonds (m x 2
    undirected pairs), forbidden (512 uint8), seed (int).
    """

    def __init__(self, ingredients, n_steps: int, gpu: int = 0):
        self.ingredients = ingredients
        self.n_steps = int(n_steps)
        self.gpu = gpu
        self._u = None

    def initialize(self) -> None:
        ing = self.ingredients
        system = {"box": ing.box, "periodic": ing.periodic, "xyz": ing.xyz, "bonds": ing.bonds,
                  "forbidden": ing.forbidden}
        self._u = updater_from_system(system, seed=ing.seed, gpu=self.gpu, age=ing.age)

    def execute(self) -> bool:
        ing = self.ingredients
        self._u.setAge(ing.age)
        self._u.runSimulationOnGPU(self.n_steps)
        ing.xyz = self._u.getMonomerPositions(np.asarray(ing.xyz).reshape(-1, 3).shape[0])
        ing.age = ing.age + self.n_steps
        return True

    def cleanup(self) -> None:
        if self._u is not None:
            self._u.cleanup()
            self._u = None
