"""Driver: Python mirror of drivers::gpu_driver (drivers/gpu/gpu_driver.h:58-182)
over libnbl_hip.so, plus the persistent batched transport entry point."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import call, ptr


def pencil_beam(energy, n, pos=(0.0, 0.0, 5.0), direction=(0.0, 0.0, -1.0), seed=0):
    """One pixel, all n primaries at one point (profile 0 with sigma 0)."""
    px = np.zeros(1, L.PIXEL_DTYPE)
    px["x"], px["y"], px["z"] = pos
    px["energy"] = energy
    px["sigma"] = 0.0
    px["profile"] = 0
    px["n_primaries"] = n
    return dict(pixels=px, dir=direction, seed=seed)


def uniform_beam(energy, n, half, z, direction=(0.0, 0.0, -1.0), seed=0):
    """One pixel, primaries uniform over a (2 half[0]) x (2 half[1]) rectangle."""
    px = np.zeros(1, L.PIXEL_DTYPE)
    px["z"] = z
    px["energy"] = energy
    px["half_x"], px["half_y"] = half
    px["profile"] = 1
    px["n_primaries"] = n
    return dict(pixels=px, dir=direction, seed=seed)


def raster_scan(xs, ys, z, energy, per_pixel, sigma, direction=(0.0, 0.0, -1.0), seed=0):
    """Gaussian beam of std sigma at every (x, y) of the grid; energy and
    per_pixel may be scalars or arrays over pixels (row-major over ys, xs)."""
    X, Y = np.meshgrid(np.asarray(xs, np.float32), np.asarray(ys, np.float32))
    n = X.size
    px = np.zeros(n, L.PIXEL_DTYPE)
    px["x"], px["y"], px["z"] = X.ravel(), Y.ravel(), z
    px["energy"] = np.broadcast_to(np.asarray(energy, np.float32), (n,))
    px["sigma"] = sigma
    px["profile"] = 0
    px["n_primaries"] = np.broadcast_to(np.asarray(per_pixel, np.uint32), (n,))
    return dict(pixels=px, dir=direction, seed=seed)


class Driver:
    """gpu_driver(capacity, intersect, materials, geometry, min_energy, max_energy, seed)."""

    def __init__(self, geometry, materials, capacity=1_000_000, min_energy=0.0, max_energy=None,
                 seed=0x78C7E39B, threads_per_block=0):
        self.geometry, self.materials = geometry, materials
        if max_energy is None:
            max_energy = materials.max_energy()      # gpu_main.cu:349
        self.capacity = capacity
        opts = L.DriverOpts(threads_per_block=threads_per_block)
        h = C.c_void_p()
        call("nbl_driver_create", capacity, geometry.h, materials.h, float(min_energy), float(max_energy),
             int(seed), C.byref(opts), C.byref(h))
        self.h = h
        self._keep = []

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            L.lib.nbl_driver_destroy(self.h)
            self.h = C.c_void_p()

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- gpu_driver methods -------------------------------------------------
    def push(self, particles, tags) -> int:
        particles = np.ascontiguousarray(particles, L.PARTICLE_DTYPE)
        tags = np.ascontiguousarray(tags, np.uint32)
        n = C.c_uint32()
        call("nbl_driver_push", self.h, ptr(particles), ptr(tags), len(particles), C.byref(n))
        return n.value

    def do_iteration(self):
        call("nbl_driver_do_iteration", self.h)

    def get_running_count(self) -> int:
        n = C.c_uint32()
        call("nbl_driver_get_running_count", self.h, C.byref(n))
        return n.value

    def get_detected_count(self) -> int:
        n = C.c_uint32()
        call("nbl_driver_get_detected_count", self.h, C.byref(n))
        return n.value

    @staticmethod
    def _collector(out):
        def cb(p, tag, user):
            arr = np.frombuffer(C.string_at(p, 28), dtype=L.PARTICLE_DTYPE)
            out.append((arr[0].copy(), tag))
        return L.DETECT_FN(cb)

    def flush_detected(self):
        out = []
        fn = self._collector(out)
        call("nbl_driver_flush_detected", self.h, fn, None)
        return out

    def allocate_input_buffers(self, n):
        call("nbl_driver_allocate_input_buffers", self.h, n)

    def push_to_buffer(self, pool):
        """pool: object with get_work(data, particles, tags) filling free entries
        (work_pool::get_work, common/work_pool.h:69-88)."""
        def fill(data_p, parts_p, tags_p, n, user):
            data = np.ctypeslib.as_array(C.cast(data_p, C.POINTER(C.c_uint8)), (n,))
            parts = np.frombuffer((C.c_char * (28 * n)).from_address(parts_p), dtype=L.PARTICLE_DTYPE)
            tags = np.ctypeslib.as_array(C.cast(tags_p, C.POINTER(C.c_uint32)), (n,))
            pool.get_work(data, parts, tags)
        fn = L.POOL_FILL_FN(fill)
        call("nbl_driver_push_to_buffer", self.h, fn, None)

    def push_to_simulation(self):
        call("nbl_driver_push_to_simulation", self.h)

    def buffer_detected(self):
        call("nbl_driver_buffer_detected", self.h)

    def flush_buffered(self):
        out = []
        fn = self._collector(out)
        run = C.c_uint32()
        call("nbl_driver_flush_buffered", self.h, fn, None, C.byref(run))
        return run.value, out

    # --- slot access (tests) ------------------------------------------------
    def read_slots(self):
        n = self.capacity
        s = dict(status=np.zeros(n, np.uint8), particle_idx=np.zeros(n, np.uint32),
                 particles=np.zeros(n, L.PARTICLE_DTYPE), tags=np.zeros(n, np.uint32),
                 material=np.zeros(n, np.int32), last_tri=np.zeros(n, np.uint32), rng=np.zeros(6 * n, np.uint32))
        v = L.SlotView(**{k: ptr(a) for k, a in s.items()})
        call("nbl_driver_read_slots", self.h, C.byref(v))
        return s

    def write_slots(self, **arrays):
        keep = {k: np.ascontiguousarray(a) for k, a in arrays.items()}
        v = L.SlotView(**{k: ptr(a) for k, a in keep.items()})
        call("nbl_driver_write_slots", self.h, C.byref(v))

    # --- persistent batched transport -----------------------------------------
    def set_sampler(self, sampler):
        """Learned ICDF sampler for the persistent transport (None = the tables)."""
        call("nbl_driver_set_sampler", self.h, sampler.h if sampler is not None else None)
        self.sampler = sampler                 # keep it alive while the driver uses it

    def run_persistent(self, beam, block_size=0, blocks_per_cu=0, stack_depth=0, refill_threshold=0,
                       per_pixel=False, hist=False, deposit=None, detected=0, inelastic_batch=0, inelastic_wait=0,
                       overflow_depth=0, allow_drops=False, deposit_records=None, count_work=False):
        """One persistent launch over the pixel batch `beam`.  deposit =
        dict(shape=(nx, ny, nz), lo=(x, y, z), hi=(x, y, z)) turns on the
        energy-deposition tally: out["deposit"] is the [nz, ny, nx] grid (eV),
        out["deposit_total"] and out["deposit_count"] cover every deposit.
        detected = capacity of the detected-electron record buffer (0: off):
        out["detected"] = (particles, tags), out["detected_count"].
        deposit_records = capacity of the energy-deposition record buffer
        (None: off; 0: count only): out["deposit_records"] = (records of
        DEPOSIT_RECORD_DTYPE, tags), out["deposit_records_count"].
        inelastic_batch, inelastic_wait, overflow_depth: nbl_persistent_opts (0 = defaults).
        count_work: fill the work counters (events, lookups, node steps, ...;
        nbl_persistent_opts.count_work), else they are 0.
        A run that stops any electron (a secondary with no stack room left, or
        the step cap) raises NblError unless allow_drops; then out["stopped"]
        counts them."""
        px = np.ascontiguousarray(beam["pixels"], L.PIXEL_DTYPE)
        d = beam.get("dir", (0.0, 0.0, -1.0))
        b = L.PixelBatch(pixels=ptr(px), n_pixels=len(px), dir=L.Vec3(*d), seed=int(beam.get("seed", 0)))
        return self._persistent("nbl_driver_run_persistent", (C.byref(b),), len(px), block_size, blocks_per_cu,
                                stack_depth, refill_threshold, per_pixel, hist, deposit, detected,
                                inelastic_batch, inelastic_wait, overflow_depth, allow_drops, deposit_records,
                                count_work)

    def run_persistent_list(self, primaries, tags=None, block_size=0, blocks_per_cu=0, stack_depth=0,
                            refill_threshold=0, hist=False, deposit=None, detected=0, inelastic_batch=0,
                            inelastic_wait=0, overflow_depth=0, allow_drops=False, deposit_records=None,
                            count_work=False):
        """The persistent transport over explicit primaries (PARTICLE_DTYPE),
        tags default to their indices; per-pixel tallies are not available."""
        prim = np.ascontiguousarray(primaries, L.PARTICLE_DTYPE)
        tags = np.arange(len(prim), dtype=np.uint32) if tags is None else np.ascontiguousarray(tags, np.uint32)
        self._keep_list = (prim, tags)
        return self._persistent("nbl_driver_run_persistent_list", (ptr(prim), ptr(tags), len(prim)), 0, block_size,
                                blocks_per_cu, stack_depth, refill_threshold, False, hist, deposit, detected,
                                inelastic_batch, inelastic_wait, overflow_depth, allow_drops, deposit_records,
                                count_work)

    def run_persistent_list_noalloc(self, dev_primaries, dev_tags, n, dev_detected=None, dev_detected_tags=None,
                                    detected_capacity=0, hist=False, block_size=0, blocks_per_cu=0, stack_depth=0,
                                    refill_threshold=0, inelastic_batch=0, inelastic_wait=0, overflow_depth=0,
                                    allow_drops=False, count_work=False):
        """The persistent transport over n primaries already in device memory
        (raw device addresses of PARTICLE_DTYPE records and uint32 tags),
        through the entry point that allocates nothing: records of detected
        electrons go to the caller's device arrays (first detected_capacity of
        them); the count is out["detected_se"] + out["detected_bse"]."""
        o = L.PersistentOpts(block_size=block_size, blocks_per_cu=blocks_per_cu, stack_depth=stack_depth,
                             refill_threshold=refill_threshold, inelastic_batch=inelastic_batch,
                             inelastic_wait=inelastic_wait, overflow_depth=overflow_depth,
                             allow_drops=int(bool(allow_drops)), count_work=int(bool(count_work)))
        t = L.Tallies()
        h = np.zeros(L.HIST_BINS, np.uint32) if hist else None
        t.hist = ptr(h)
        call("nbl_driver_run_persistent_list_noalloc", self.h, dev_primaries, dev_tags, int(n), dev_detected,
             dev_detected_tags, int(detected_capacity), C.byref(o), C.byref(t))
        out = {f: getattr(t, f) for f, _ in L.Tallies._fields_
               if f not in ("pixel_se", "pixel_bse", "hist", "deposit", "detected", "deposits")}
        if hist:
            out["hist"] = h
        return out

    def _persistent(self, fn, head, n_px, block_size, blocks_per_cu, stack_depth, refill_threshold, per_pixel,
                    hist, deposit, detected, inelastic_batch=0, inelastic_wait=0, overflow_depth=0,
                    allow_drops=False, deposit_records=None, count_work=False):
        o = L.PersistentOpts(block_size=block_size, blocks_per_cu=blocks_per_cu, stack_depth=stack_depth,
                             refill_threshold=refill_threshold, inelastic_batch=inelastic_batch,
                             inelastic_wait=inelastic_wait, overflow_depth=overflow_depth,
                             allow_drops=int(bool(allow_drops)), count_work=int(bool(count_work)))
        t = L.Tallies()
        se = np.zeros(n_px, np.uint32) if per_pixel else None
        bse = np.zeros(n_px, np.uint32) if per_pixel else None
        h = np.zeros(L.HIST_BINS, np.uint32) if hist else None
        t.pixel_se, t.pixel_bse, t.hist = ptr(se), ptr(bse), ptr(h)
        dep = grid = None
        if deposit is not None:
            nx, ny, nz = (int(v) for v in deposit["shape"])
            grid = np.zeros((nz, ny, nx), np.float64)
            dep = L.DepositGrid(nx=nx, ny=ny, nz=nz, lo=L.Vec3(*map(float, deposit["lo"])),
                                hi=L.Vec3(*map(float, deposit["hi"])), energy=ptr(grid))
            t.deposit = C.cast(C.pointer(dep), C.c_void_p)
        det = dparts = dtags = None
        if detected:
            dparts = np.zeros(int(detected), L.PARTICLE_DTYPE)
            dtags = np.zeros(int(detected), np.uint32)
            det = L.DetectionBuffer(particles=ptr(dparts), tags=ptr(dtags), capacity=int(detected))
            t.detected = C.cast(C.pointer(det), C.c_void_p)
        drb = rparts = rtags = None
        if deposit_records is not None:
            rparts = np.zeros(int(deposit_records), L.DEPOSIT_RECORD_DTYPE)
            rtags = np.zeros(int(deposit_records), np.uint32)
            drb = L.DepositBuffer(records=ptr(rparts), tags=ptr(rtags), capacity=int(deposit_records))
            t.deposits = C.cast(C.pointer(drb), C.c_void_p)
        call(fn, self.h, *head, C.byref(o), C.byref(t))
        out = {f: getattr(t, f) for f, _ in L.Tallies._fields_
               if f not in ("pixel_se", "pixel_bse", "hist", "deposit", "detected", "deposits")}
        if drb is not None:
            k = min(drb.count, int(deposit_records))
            out["deposit_records"], out["deposit_records_count"] = (rparts[:k], rtags[:k]), drb.count
        if det is not None:
            k = min(det.count, int(detected))
            out["detected"], out["detected_count"] = (dparts[:k], dtags[:k]), det.count
        if dep is not None:
            out["deposit"], out["deposit_total"], out["deposit_count"] = grid, dep.total, dep.count
        if per_pixel:
            out["pixel_se"], out["pixel_bse"] = se, bse
        if hist:
            out["hist"] = h
        return out


class WorkPool:
    """work_pool (common/work_pool.h:14-123) over numpy arrays of primaries and tags."""

    def __init__(self, particles, tags):
        self.particles = np.ascontiguousarray(particles, L.PARTICLE_DTYPE)
        self.tags = np.ascontiguousarray(tags, np.uint32)
        self.next = 0

    def primaries_to_go(self):
        return len(self.particles) - self.next

    def done(self):
        return self.primaries_to_go() == 0

    def get_work_batch(self, n):
        k = min(n, self.primaries_to_go())
        s = slice(self.next, self.next + k)
        self.next += k
        return self.particles[s], self.tags[s]

    def get_work(self, data, particles, tags):
        free = np.flatnonzero(data == 0)
        k = min(len(free), self.primaries_to_go())
        if k == 0:
            return
        sel = free[:k]
        data[sel] = 1
        particles[sel] = self.particles[self.next:self.next + k]
        tags[sel] = self.tags[self.next:self.next + k]
        self.next += k
