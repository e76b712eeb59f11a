"""Many independent settling cases stepped together on one GPU.

Each case is a lattice of its own size, lattice type and relaxation time with
its own rigid particles; it follows exactly the single-lattice step of LBMD3
(the same kernels' bodies), so a case gives bit-identical results alone and
in a mix.  Cases share pools allocated once; the host admits a case by
allocating ranges, packing its descriptor and initialising its state, and
retires it by freeing the ranges.  One funlbm_batch_step launch sequence
advances every admitted case that still has steps left; a captured HIP graph
of several steps stays valid across admissions because all the kernels read
their work lists and counts from device memory.
"""

import ctypes
import dataclasses
import math
from typing import List, Optional

import numpy as np
import torch

from .. import _native as nat
from ..particle import boxes
from ..particle.markers import markers as marker_points
from ..particle.markers import surface_markers
from .alloc import RangeAllocator

TILE = 256
# smallest grid of a lattice's fused kernel (8 blocks per compute unit)
MIN_LATTICE_GRID = 2048
H_TILES, H_MARKERS, H_PARTICLES, H_CASES, H_STAMP = 0, 1, 2, 3, 4
H_CHUNKS, H_WORDS = 8, 9


@dataclasses.dataclass
class ParticleSpec:
    """A rigid sphere or ellipsoid (the reference's particle config:
    src/funlbm/particle/base/base.py:11-19, sphere.py, ellipsoid.py)."""

    kind: str = "sphere"  # "sphere" or "ellipsoid"
    r: float = 5.0
    abc: tuple = (5.0, 5.0, 5.0)
    rou: float = 1.5
    center: tuple = (0.0, 0.0, 0.0)
    angle: tuple = (math.pi / 2, 0.0, 0.0)  # rotation vector, alpha = pi/2 by default
    markers: str = "cubes"  # "cubes" (the reference's generator) or "fibonacci"


@dataclasses.dataclass
class CaseSpec:
    """One independent solve: a box of nodes with particles, stepped `steps` times."""

    nodes: tuple
    q: int = 19
    tau: float = 0.6
    steps: int = 1000
    particles: List[ParticleSpec] = dataclasses.field(default_factory=list)
    # g of the buoyancy force (1 - rho_p/rho_f) V g: heavy particles move
    # against g, so +z settles them towards z = 0
    gravity: tuple = (0.0, 0.0, 2e-5)
    bcs: tuple = (nat.BC_WALL,) * 6
    uzero_mask: int = 0x3F
    # base seed of the stochastic rounding's dither (16-bit states)
    seed: int = 0
    # immersed-boundary kernel: "cosine" (the reference's) or "peskin4"
    delta: str = "cosine"

    @property
    def n(self):
        return int(np.prod(self.nodes))


_const_cache = {}


def _semi(p: ParticleSpec):
    if p.kind == "sphere":
        return (float(p.r),) * 3
    return tuple(float(v) for v in p.abc)


def prepare_markers(particles) -> None:
    """Generates the markers of every shape among `particles` not seen yet in
    one GPU launch (admission then finds them cached)."""
    surface_markers([_semi(p) for p in particles if p.markers in (None, "cubes")])


def markers_for(p: ParticleSpec):
    """Body-frame markers (float32) and the constants of a particle,
    following Sphere/Ellipsoid._init (src/funlbm/particle/base/sphere.py:15-53,
    ellipsoid.py:107-143)."""
    a, b, c = _semi(p)
    key = (p.kind, a, b, c, p.markers)
    if key not in _const_cache:
        lag = marker_points(p.markers, a, b, c)
        mass = float(np.float32(4.0 / 3.0 * math.pi * a * b * c))
        if p.kind == "sphere":
            area = float(np.float32(4.0 * math.pi * a * a))
            inertia = np.array([a * a] * 3) * np.float32(mass) / 5.0
        else:
            area = float(np.float32(4.0 / 3.0 * math.pi * math.pow(a * b * c, 2.0 / 3.0)))
            inertia = np.array([b * c, a * c, a * b]) * np.float32(mass) / 5.0
        lm = np.float32(np.float32(area) / np.float32(len(lag)))
        _const_cache[key] = (lag, mass, inertia.astype(np.float32).astype(np.float64), lm)
    return _const_cache[key]


def lattice_for(spec: CaseSpec, storage: int, dev_scale: float, rounding: int = nat.ROUND_NEAREST) -> nat.Lattice:
    lat = nat.Lattice()
    lat.q = spec.q
    lat.nx, lat.ny, lat.nz = (int(v) for v in spec.nodes)
    lat.storage = storage
    for i in range(6):
        lat.bc[i] = int(spec.bcs[i])
    lat.uzero_mask = spec.uzero_mask
    lat.tau = float(np.float32(spec.tau))
    lat.rho0 = 1.0
    lat.dev_scale = dev_scale
    lat.rounding = rounding
    lat.seed = int(spec.seed) & 0xFFFFFFFF
    lat.delta = nat.DELTA_CODES[spec.delta]
    return lat


@dataclasses.dataclass
class _Slot:
    spec: CaseSpec
    lat: nat.Lattice
    f_off: int
    node_off: int
    tile_off: int
    n_tiles: int
    m_off: int
    m_count: int
    p_off: int
    p_count: int
    plane: int
    boxes: list  # (cache offset, R) per particle of the case


class BatchEngine:
    """Pools for up to `max_cases` cases holding at most `max_nodes` nodes,
    `max_markers` markers and `max_particles` particles in total.  The
    particles' moment boxes share a cache of `max_box_nodes` nodes (default
    32 per marker slot); a particle whose box does not fit goes without one
    (its markers pull per node, with the same results)."""

    def __init__(self, storage="f16", max_nodes=1 << 24, max_markers=1 << 16, max_particles=256,
                 max_cases=256, max_q=27, traj_cap=64, dev_scale=1024.0, device=None, rounding=None,
                 max_box_nodes=None):
        from ..base.work import device_detect

        self.device = device_detect(device)
        dev = self.device
        self.storage = nat.STORAGE_CODES[storage]
        self.storage_name = storage
        # 16-bit states round stochastically unless told otherwise
        rounding = rounding or ("stochastic" if storage in ("f16", "bf16") else "nearest")
        if rounding not in nat.ROUNDING_CODES:
            raise ValueError(f"unknown rounding {rounding!r}")
        if rounding == "stochastic" and storage not in ("f16", "bf16"):
            raise ValueError("stochastic rounding applies to f16/bf16 storage only")
        # cases of a stochastic engine draw their dither from step_seed(spec.seed, steps done)
        self.rounding = nat.ROUNDING_CODES[rounding]
        self.dev_scale = float(dev_scale)
        self.max_cases = int(max_cases)
        self.traj_cap = int(traj_cap)
        # a case's state planes hold whole tiles plus two slack tiles (up to
        # 3 * 4096 nodes beyond its own)
        elem_cap = (int(max_nodes) + 3 * 4096 * int(max_cases)) * int(max_q)
        tile_cap = int(max_nodes) // TILE + int(max_cases)
        self.elems = RangeAllocator(elem_cap, align=64)
        self.nodes = RangeAllocator(int(max_nodes), align=TILE)
        self.tiles = RangeAllocator(tile_cap)
        self.markers = RangeAllocator(int(max_markers))
        self.parts = RangeAllocator(int(max_particles))
        self.max_particles = int(max_particles)
        box_cap = int(max_box_nodes) if max_box_nodes is not None else 32 * int(max_markers)
        self.box_cap = min(box_cap, (1 << 31) - 1)
        self.box_alloc = RangeAllocator(self.box_cap)
        self.free_cases = list(range(self.max_cases - 1, -1, -1))
        dt = nat.STORAGE_DTYPES[self.storage]
        i32, f32 = torch.int32, torch.float32
        z = lambda n, d: torch.zeros(n, device=dev, dtype=d)  # noqa: E731
        self.pool = [z(elem_cap, dt), z(elem_cap, dt)]
        self.fol = [z(3 * int(max_nodes), f32), z(3 * int(max_nodes), f32)]
        self.owner = z(int(max_nodes), torch.int64)
        self.fol_tiles = z(tile_cap, i32)
        self.origin = [z(3 * int(max_markers), i32), z(3 * int(max_markers), i32)]
        self.lag = z(3 * int(max_markers), f32)
        self.lx = z(3 * int(max_markers), f32)
        self.lu = z(3 * int(max_markers), f32)
        self.lF = z(3 * int(max_markers), f32)
        self.lm = z(int(max_markers), f32)
        self.marker_case = torch.full((int(max_markers),), -1, device=dev, dtype=i32)
        self.marker_particle = z(int(max_markers), i32)
        self.particle_case = torch.full((int(max_particles),), -1, device=dev, dtype=i32)
        self.particle_markers = z(2 * int(max_particles), i32)
        self.pstate = z(int(max_particles) * nat.PSTRIDE, torch.float64)
        self.case_bytes = nat.lib().funlbm_case_bytes()
        self.cases = z(self.max_cases * self.case_bytes, torch.uint8)
        self.case_par = z(self.max_cases, i32)
        self.case_left = z(self.max_cases, i32)
        self.case_done = z(self.max_cases, i32)
        self.case_flag = z(self.max_cases, i32)
        self.origin_ok = z(self.max_cases, i32)
        self.case_marker_begin = z(self.max_cases, i32)
        self.tile_case = torch.full((tile_cap,), -1, device=dev, dtype=i32)
        self.tile_local = z(tile_cap, i32)
        self.hdr = z(H_WORDS, i32)
        self.box_cache = torch.empty(4 * max(self.box_cap, 1), device=dev, dtype=f32)
        self.particle_box = z(3 * int(max_particles), i32)
        self.chunk_particle = torch.full((self.box_cap // boxes.CHUNK + int(max_particles),), -1,
                                         device=dev, dtype=i32)
        self.box_miss = z(int(max_markers) + 1, i32)
        self.traj = z(int(max_particles) * self.traj_cap * 4, f32)
        self.traj_cap_t = torch.tensor([self.traj_cap], device=dev, dtype=i32)
        self.slots = {}
        self.grid = {"tiles": 2048, "markers": 1024, "particles": 256}
        # exact_grids: every eager launch takes its grids from the live counts
        # (one block per active tile of each lattice; the marker and particle
        # grids as autotune sets them).  A captured graph keeps the grids of
        # its capture, so this suits eager stepping of a changing population.
        self.exact_grids = False
        self._cus = torch.cuda.get_device_properties(dev).multi_processor_count
        self._graph = None
        self._graph_steps = 0
        self._graph_mask = 0
        self._mask = 0  # lattices with cases (bit i: D3Q13, 15, 19, 27), bit 4: per-marker claims
        self._lattice_tiles = [0, 0, 0, 0]
        self._dirty = True

    def _h2d(self, arr):
        """A host array on the device, copied asynchronously on the current
        stream through pinned memory: admission and publishing then do not
        wait for steps already in flight (the pinned block is not reused
        before its copy has run)."""
        t = torch.from_numpy(np.ascontiguousarray(arr)).pin_memory()
        return t.to(self.device, non_blocking=True)

    # ------------------------------------------------------------------ admission
    def fits(self, spec: CaseSpec) -> bool:
        m = sum(len(markers_for(p)[0]) for p in spec.particles)
        return (
            bool(self.free_cases)
            and self.elems.fits(spec.q * self._plane(spec))
            and self.nodes.fits(spec.n)
            and self.tiles.fits(-(-spec.n // TILE))
            and (m == 0 or self.markers.fits(m))
            and (not spec.particles or self.parts.fits(len(spec.particles)))
        )

    def _plane(self, spec: CaseSpec) -> int:
        return nat.state_plane(lattice_for(spec, self.storage, self.dev_scale))

    @staticmethod
    def case_bytes_needed(spec: CaseSpec, storage_bytes: int = 2) -> int:
        """HBM held by a case: two state blocks, two FOL blocks, owner words."""
        storage = {2: nat.STORE_F16, 4: nat.STORE_F32, 8: nat.STORE_F64}[storage_bytes]
        plane = nat.state_plane(lattice_for(spec, storage, 1024.0))
        return 2 * spec.q * plane * storage_bytes + spec.n * (2 * 12 + 8)

    def add(self, spec: CaseSpec) -> int:
        """Admits a case; returns its slot id.  Raises MemoryError if it does not fit."""
        if not self.fits(spec):
            raise MemoryError("case does not fit the pools")
        dev = self.device
        n = spec.n
        plane = self._plane(spec)
        f_off = self.elems.alloc(spec.q * plane)
        node_off = self.nodes.alloc(n)
        n_tiles = -(-n // TILE)
        tile_off = self.tiles.alloc(n_tiles)
        infos = [markers_for(p) for p in spec.particles]
        m_count = sum(len(i[0]) for i in infos)
        m_off = self.markers.alloc(m_count) if m_count else 0
        p_off = self.parts.alloc(len(spec.particles)) if spec.particles else 0
        cid = self.free_cases.pop()
        # moment boxes: a particle whose box does not fit the cache goes without
        pboxes = []
        for lag, *_ in infos:
            R = boxes.box_radius(lag)
            off = self.box_alloc.alloc(boxes.box_nodes(R)) if R >= 0 else None
            pboxes.append((off, R) if off is not None else (0, -1))
        lat = lattice_for(spec, self.storage, self.dev_scale, self.rounding)
        desc = (ctypes.c_uint8 * self.case_bytes)()
        nat.call("funlbm_case_pack", lat, f_off, node_off, tile_off, desc)
        self.cases[cid * self.case_bytes : (cid + 1) * self.case_bytes] = self._h2d(
            np.frombuffer(bytes(desc), dtype=np.uint8).copy())
        # state: the reference's init (f = w, its moments, the two collisions
        # before the first stream), as LBMD3 enters its fused steps
        esz = self.pool[0].element_size()
        nat.call("funlbm_init_state", lat, ctypes.c_void_p(self.pool[0].data_ptr() + f_off * esz),
                 nat.stream_handle(dev))
        for j in range(2):
            self.fol[j][3 * node_off : 3 * (node_off + n)].zero_()
        # the case's spreading stamps restart at 1: clear the tile entries and
        # owner words an earlier case left in these ranges
        self.fol_tiles[tile_off : tile_off + n_tiles].zero_()
        self.owner[node_off : node_off + n].zero_()
        # markers and particles
        ps = np.zeros((len(spec.particles), nat.PSTRIDE))
        pm = np.zeros(2 * len(spec.particles), dtype=np.int32)
        lags, lms, mp = [], [], []
        cur = m_off
        for k, (p, (lag, mass, inertia, lm)) in enumerate(zip(spec.particles, infos)):
            lags.append(lag)
            lms.append(np.full(len(lag), lm, np.float32))
            mp.append(np.full(len(lag), p_off + k, np.int32))
            pm[2 * k], pm[2 * k + 1] = cur, cur + len(lag)
            cur += len(lag)
            ps[k, nat.P_CX : nat.P_CX + 3] = np.array(p.center, np.float32)
            ps[k, nat.P_ANGLE : nat.P_ANGLE + 3] = np.array(p.angle, np.float32)
            ps[k, nat.P_MASS] = mass
            ps[k, nat.P_INERTIA : nat.P_INERTIA + 3] = inertia
            ps[k, nat.P_RHO] = p.rou
            ps[k, nat.P_GRAV : nat.P_GRAV + 3] = np.array(spec.gravity, np.float32)
        if m_count:
            sl = slice(3 * m_off, 3 * (m_off + m_count))
            self.lag[sl] = self._h2d(np.concatenate(lags).reshape(-1))
            self.lx[sl].zero_()
            self.lu[sl].zero_()
            self.lF[sl].zero_()
            self.lm[m_off : m_off + m_count] = self._h2d(np.concatenate(lms))
            self.marker_particle[m_off : m_off + m_count] = self._h2d(np.concatenate(mp))
            self.marker_case[m_off : m_off + m_count] = cid
        if spec.particles:
            P = len(spec.particles)
            self.pstate[p_off * nat.PSTRIDE : (p_off + P) * nat.PSTRIDE] = self._h2d(ps.reshape(-1))
            self.particle_markers[2 * p_off : 2 * (p_off + P)] = self._h2d(pm)
            self.particle_case[p_off : p_off + P] = cid
            self.traj[p_off * self.traj_cap * 4 : (p_off + P) * self.traj_cap * 4].zero_()
        dyn = self._h2d(np.array([0, int(spec.steps), 0, 0, 0, m_off], dtype=np.int32))
        self.case_par[cid] = dyn[0]
        self.case_left[cid] = dyn[1]
        self.case_done[cid] = dyn[2]
        self.case_flag[cid] = dyn[3]
        self.origin_ok[cid] = dyn[4]
        self.case_marker_begin[cid] = dyn[5]
        self.slots[cid] = _Slot(spec, lat, f_off, node_off, tile_off, n_tiles, m_off, m_count, p_off,
                                len(spec.particles), plane, pboxes)
        self._dirty = True
        return cid

    def remove(self, cid: int):
        s = self.slots.pop(cid)
        self.case_left[cid] = 0
        if s.m_count:
            self.marker_case[s.m_off : s.m_off + s.m_count] = -1
            self.markers.free(s.m_off, s.m_count)
        if s.p_count:
            self.particle_case[s.p_off : s.p_off + s.p_count] = -1
            self.parts.free(s.p_off, s.p_count)
        for off, R in s.boxes:
            if R >= 0:
                self.box_alloc.free(off, boxes.box_nodes(R))
        self.elems.free(s.f_off, s.spec.q * s.plane)
        self.nodes.free(s.node_off, s.spec.n)
        self.tiles.free(s.tile_off, s.n_tiles)
        self.free_cases.append(cid)
        self._dirty = True

    def _publish(self):
        """Rebuilds the active tile list and the header counts (host -> device)."""
        if not self._dirty:
            return
        # tiles ordered by lattice (D3Q13, 15, 19, 27): one fused kernel per
        # lattice takes its own range of the list
        tc, tl = [], []
        ends = []
        for q in (13, 15, 19, 27):
            for cid, s in sorted(self.slots.items()):
                if s.spec.q == q:
                    tc.append(np.full(s.n_tiles, cid, np.int32))
                    tl.append(np.arange(s.n_tiles, dtype=np.int32))
            ends.append(int(sum(len(a) for a in tc)))
        nt = ends[-1]
        self._lattice_tiles = [b - a for a, b in zip([0] + ends[:-1], ends)]
        self._mask = sum(1 << i for i, q in enumerate((13, 15, 19, 27))
                         if any(s.spec.q == q for s in self.slots.values()))
        # per-marker spreading claims while a particle has no shared-memory box
        if any(not boxes.claims_in_box(R) for s in self.slots.values() for _, R in s.boxes):
            self._mask |= 16
        if nt:
            self.tile_case[:nt] = self._h2d(np.concatenate(tc))
            self.tile_local[:nt] = self._h2d(np.concatenate(tl))
        # moment boxes over particle slots, and the chunk list of the box kernel
        radii = np.full(self.max_particles, -1, dtype=np.int64)
        offs = np.zeros(self.max_particles, dtype=np.int64)
        for cid, s in sorted(self.slots.items()):
            for k, (off, R) in enumerate(s.boxes):
                radii[s.p_off + k], offs[s.p_off + k] = R, off
        rows, chunks, _ = boxes.layout(radii, offs)
        self.particle_box.copy_(self._h2d(rows.reshape(-1).astype(np.int32)))
        nch = len(chunks)
        if nch:
            self.chunk_particle[:nch] = self._h2d(chunks.astype(np.int32))
        # hdr[H_STAMP] (steps launched) stays as the device has it
        head = [nt, self.markers.high_water, self.parts.high_water, self.max_cases]
        self.hdr[:H_STAMP] = self._h2d(np.array(head, dtype=np.int32))
        self.hdr[H_STAMP + 1 :] = self._h2d(np.array(ends[:3] + [nch], dtype=np.int32))
        self._with_ibm = 1 if self.markers.high_water > 0 else 0
        self._dirty = False

    # ------------------------------------------------------------------ stepping
    def _lattice_grid(self):
        """Grid of each lattice's fused kernel: its own tile count, at least
        MIN_LATTICE_GRID blocks (cases admitted after a graph capture are
        then still spread over the chip), at most grid["tiles"]; with
        exact_grids, its own tile count."""
        if self.exact_grids:
            return (ctypes.c_int32 * 4)(*[max(1, int(c)) for c in self._lattice_tiles])
        cap = int(self.grid["tiles"])
        return (ctypes.c_int32 * 4)(*[max(1, min(cap, max(int(c), MIN_LATTICE_GRID)))
                                      for c in self._lattice_tiles])

    def _ibm_grids(self):
        """Marker and particle grids for the current counts: one pass of 16
        markers per block (with 168 k markers this stepped the 256-case batch
        2 % faster than 2048 persistent blocks) and one block per particle, at
        most four per compute unit (the kernels are grid-stride loops)."""
        n_mark = max(1, self.markers.high_water)
        n_part = max(1, self.parts.high_water)
        return -(-n_mark // 16), max(1, min(n_part, 4 * self._cus))

    def _launch(self):
        P = nat.ptr
        g = dict(self.grid)
        if self.exact_grids:
            g["markers"], g["particles"] = self._ibm_grids()
        nat.call(
            "funlbm_batch_step", self.storage, P(self.cases), P(self.hdr), P(self.case_par),
            P(self.case_left), P(self.case_done), P(self.case_flag), P(self.tile_case),
            P(self.tile_local), P(self.marker_case), P(self.marker_particle), P(self.case_marker_begin),
            P(self.particle_case), P(self.particle_markers), P(self.pool[0]), P(self.pool[1]),
            P(self.fol[0]), P(self.fol[1]), P(self.owner), P(self.fol_tiles), P(self.origin[0]),
            P(self.origin[1]), P(self.origin_ok), P(self.lag), P(self.lx), P(self.lu), P(self.lF),
            P(self.lm), P(self.pstate), P(self.traj), P(self.traj_cap_t), int(g["tiles"]),
            int(g["markers"]), int(g["particles"]), 1, P(self.particle_box), P(self.chunk_particle),
            P(self.box_cache), self.box_cap, P(self.box_miss), self._mask, self._lattice_grid(),
            nat.stream_handle(self.device),
        )

    def step(self, n: int = 1):
        """n timesteps of every case with steps left (eager launches)."""
        self._publish()
        for _ in range(n):
            self._launch()

    def autotune(self, steps: int = 10, candidates=None):
        """Chooses the grids of the batched step by timing the first steps of
        the real run with each candidate (results do not depend on the grids:
        every kernel is a grid-stride loop whose per-node and per-marker work
        is independent of them): first the fused kernels' grid (persistent
        grids of 2-16 blocks per compute unit, or one block per tile), then,
        with that one, the marker kernels' grid (4-16 blocks per compute
        unit, or one pass of 16 markers per block).  The winners stay in
        self.grid, for launches with fixed grids (exact_grids off, which this
        call sets).  Returns the timings."""
        self._publish()
        self.exact_grids = False
        dev = self.device
        cus = self._cus
        n_tiles = int(self.hdr[H_TILES].item())
        self.grid["markers"], self.grid["particles"] = self._ibm_grids()

        def time_steps():
            self._launch()  # settle
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(steps):
                self._launch()
            b.record()
            torch.cuda.synchronize(dev)
            return a.elapsed_time(b) / steps

        if candidates is None:
            candidates = sorted({c for c in (2 * cus, 4 * cus, 8 * cus, 16 * cus, n_tiles) if 0 < c <= max(n_tiles, 1)})
        timings = {}
        for c in candidates:
            self.grid["tiles"] = int(c)
            timings[int(c)] = time_steps()
        best = min(timings, key=timings.get)
        self.grid["tiles"] = best
        marker_timings = {}
        if self._with_ibm:
            n_mark = max(1, self.markers.high_water)
            for c in sorted({4 * cus, 8 * cus, 16 * cus, -(-n_mark // 16)}):
                self.grid["markers"] = int(c)
                marker_timings[int(c)] = time_steps()
            self.grid["markers"] = min(marker_timings, key=marker_timings.get)
        return {"grid_tiles": best, "ms_per_step": {str(k): round(v, 4) for k, v in timings.items()},
                "grid_markers": self.grid["markers"],
                "ms_per_step_markers": {str(k): round(v, 4) for k, v in marker_timings.items()},
                "steps_used": (steps + 1) * (len(candidates) + len(marker_timings))}

    def capture(self, n: int):
        """Captures n steps in a HIP graph (replayed by run_graph)."""
        self._publish()
        # warm the launch path outside the capture
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                for _ in range(n):
                    self._launch()
        torch.cuda.current_stream(self.device).wait_stream(s)
        self._graph, self._graph_steps, self._graph_mask = g, n, self._mask
        return g

    def run_graph(self):
        """Replays the captured steps; captures again first when a case of a
        lattice the graph does not launch has been admitted since."""
        self._publish()
        if self._mask & ~self._graph_mask:
            self.capture(self._graph_steps)
        self._graph.replay()

    # ------------------------------------------------------------------ results
    def left(self):
        return self.case_left.cpu().numpy()

    def finished(self):
        left = self.left()
        return [cid for cid in self.slots if left[cid] <= 0]

    def particle_rows(self, cid):
        s = self.slots[cid]
        return self.pstate[s.p_off * nat.PSTRIDE : (s.p_off + s.p_count) * nat.PSTRIDE].view(
            s.p_count, nat.PSTRIDE)

    def trajectory(self, cid):
        """Recorded (cx, cy, cz, step) per particle, ring of traj_cap entries."""
        s = self.slots[cid]
        t = self.traj.view(-1, self.traj_cap, 4)[s.p_off : s.p_off + s.p_count]
        return t

    def fields(self, cid, with_f=False):
        """The reference's u [X,Y,Z,3] and rho [X,Y,Z,1] of a case after its
        last step (and f [X,Y,Z,Q] with `with_f`)."""
        s = self.slots[cid]
        dev = self.device
        n = s.spec.n
        par = int(self.case_par[cid].item())
        a = 1 - par  # state streamed by the last step
        esz = self.pool[0].element_size()
        u = torch.empty((3, n), device=dev, dtype=torch.float32)
        rho = torch.empty((n,), device=dev, dtype=torch.float32)
        f = torch.empty((s.spec.q, n), device=dev, dtype=torch.float32) if with_f else None
        fol_new = self.fol[par][3 * s.node_off :]
        fol_old = self.fol[a][3 * s.node_off :]
        # the stamp of the case's last spreading is its step count
        stamp = self.case_done[cid : cid + 1]
        nat.call(
            "funlbm_materialize", s.lat, ctypes.c_void_p(self.pool[a].data_ptr() + s.f_off * esz),
            nat.ptr(fol_old), nat.ptr(fol_new), nat.ptr(self.fol_tiles[s.tile_off :]), nat.ptr(stamp),
            None, nat.ptr(f), nat.ptr(u), nat.ptr(rho), None, nat.stream_handle(dev),
        )
        X, Y, Z = s.spec.nodes
        out = (u.view(3, X, Y, Z).permute(1, 2, 3, 0), rho.view(X, Y, Z, 1))
        if with_f:
            out += (f.view(s.spec.q, X, Y, Z).permute(1, 2, 3, 0),)
        return out
