"""bounce_amd: MI355X (gfx950) contact narrowphase for the Bounce physics engine.

Python access to the C ABI in include/bounce_amd.h (libbounce_amd.so, built in-tree by
bounce_amd/build.py).  The HIP library is the only implementation: loading fails loudly
when it is missing, and there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

from .types import (CACHE_DTYPE, MANIFOLD_DTYPE, SHAPE_DTYPE, HullPool, empty_caches, ptr,
                    SPHERE, CAPSULE, HULL, MESH)

__all__ = ["Context", "Graph", "Server", "HullPool", "load_library", "empty_caches", "BounceError",
           "MANIFOLD_DTYPE", "CACHE_DTYPE", "SHAPE_DTYPE", "SPHERE", "CAPSULE", "HULL", "MESH"]

# The in-tree build; B3AMD_LIBRARY names another build of the same library (a variant built
# for comparison) to load instead.
LIB_PATH = os.environ.get("B3AMD_LIBRARY") or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                           "libbounce_amd.so")
_lib = None

OK, ERR_ARGUMENT, ERR_DEVICE, ERR_BUDGET, ERR_UNSUPPORTED, ERR_NO_DEVICE, ERR_OVERFLOW = 0, -1, -2, -3, -4, -5, -6


class BounceError(RuntimeError):
    def __init__(self, code, message):
        super().__init__("bounce_amd error %d: %s" % (code, message))
        self.code = code


def load_library():
    """The HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libbounce_amd.so is not built: run `python bounce_amd/build.py` "
                           "(or __graft_entry__.build()) first")
    lib = ctypes.CDLL(LIB_PATH)
    vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    sig = {
        "b3amd_create": (i32, [i32, u64, ctypes.POINTER(vp)]),
        "b3amd_destroy": (i32, [vp]),
        "b3amd_last_error": (ctypes.c_char_p, []),
        "b3amd_memory_in_use": (u64, [vp]),
        "b3amd_upload_hulls": (i32, [vp, vp]),
        "b3amd_upload_shapes": (i32, [vp, u32, vp]),
        "b3amd_hull_add": (i32, [vp, vp, vp]),
        "b3amd_hull_free": (i32, [vp, u32]),
        "b3amd_hull_registry_info": (i32, [vp, ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(u64),
                                           ctypes.POINTER(u64)]),
        "b3amd_collide_shape_and_shape": (i32, [vp, vp, vp, u32, vp, u32, vp, vp]),
        "b3amd_cache_alloc": (i32, [vp, u32, ctypes.POINTER(vp)]),
        "b3amd_cache_reset": (i32, [vp, vp, u32, u32]),
        "b3amd_cache_free": (i32, [vp, vp]),
        "b3amd_collide": (i32, [vp, u32, vp, vp, vp, vp, vp]),
        "b3amd_collide_device": (i32, [vp, u32, vp, vp, vp, vp, vp]),
        "b3amd_update_contacts": (i32, [vp, u32, vp, vp, vp, u32, vp, vp, vp]),
        "b3amd_synchronize": (i32, [vp]),
        "b3amd_device_alloc": (i32, [vp, u64, ctypes.POINTER(vp)]),
        "b3amd_device_free": (i32, [vp, vp]),
        "b3amd_copy_to_device": (i32, [vp, vp, vp, u64]),
        "b3amd_copy_to_host": (i32, [vp, vp, vp, u64]),
        "b3amd_copy_on_device": (i32, [vp, vp, vp, u64]),
        "b3amd_set_profiling": (i32, [vp, i32]),
        "b3amd_kernel_stats": (i32, [vp, ctypes.c_char_p, ctypes.POINTER(u32), ctypes.POINTER(ctypes.c_double)]),
        "b3amd_reset_kernel_stats": (i32, [vp]),
        "b3amd_upload_mesh": (i32, [vp, u32, vp, u32, vp, ctypes.POINTER(u32)]),
        "b3amd_mesh_tree": (i32, [vp, u32, ctypes.POINTER(u32), vp]),
        "b3amd_collide_mesh": (i32, [vp, u32, u32, vp, vp, vp, vp, vp, vp]),
        "b3amd_collide_mesh_device": (i32, [vp, u32, u32, vp, vp, vp, vp, vp, vp]),
        "b3amd_mesh_contacts_create": (i32, [vp, u32, u32, vp, vp, vp]),
        "b3amd_mesh_contacts_destroy": (i32, [vp]),
        "b3amd_mesh_contacts_collide": (i32, [vp, vp, vp, vp, vp]),
        "b3amd_mesh_contacts_synchronize": (i32, [vp, vp, vp]),
        "b3amd_mesh_contacts_read": (i32, [vp, vp, vp, vp, vp]),
        "b3amd_mesh_contacts_collide_device": (i32, [vp, vp, vp, vp, vp]),
        "b3amd_mesh_contacts_test_overlap_device": (i32, [vp, vp, vp]),
        "b3amd_mesh_contacts_synchronize_device": (i32, [vp, vp, vp]),
        "b3amd_set_sat_tuning": (i32, [vp, u32, u32]),
        "b3amd_get_sat_tuning": (i32, [vp, ctypes.POINTER(u32), ctypes.POINTER(u32)]),
        "b3amd_autotune": (i32, [vp, u32, vp, vp, vp, u32, ctypes.POINTER(ctypes.c_float)]),
        "b3amd_graph_create": (i32, [vp, u32, vp, vp, vp, vp, vp, ctypes.POINTER(vp)]),
        "b3amd_graph_launch": (i32, [vp]),
        "b3amd_graph_destroy": (i32, [vp]),
        "b3amd_server_create": (i32, [vp, u32, u32, u32, ctypes.POINTER(vp)]),
        "b3amd_server_destroy": (i32, [vp]),
        "b3amd_world_create": (i32, [vp, u32, vp, u32, vp, ctypes.POINTER(u32)]),
        "b3amd_world_destroy": (i32, [vp, u32]),
        "b3amd_world_submit": (i32, [vp, u32, vp, vp, ctypes.POINTER(u64)]),
        "b3amd_server_wait": (i32, [vp, u64, ctypes.POINTER(ctypes.c_double)]),
        "b3amd_world_read_manifolds": (i32, [vp, u32, vp]),
        "b3amd_server_replay": (i32, [vp, u32, vp, vp, vp, vp]),
        "b3amd_server_get_stats": (i32, [vp, vp]),
        "b3amd_server_configure": (i32, [vp, u32, u32]),
        "b3amd_broadphase_create": (i32, [vp, ctypes.POINTER(vp)]),
        "b3amd_broadphase_destroy": (i32, [vp]),
        "b3amd_broadphase_create_proxies": (i32, [vp, u32, vp, vp, vp]),
        "b3amd_broadphase_destroy_proxies": (i32, [vp, u32, vp]),
        "b3amd_broadphase_touch_proxies": (i32, [vp, u32, vp]),
        "b3amd_broadphase_move_proxies": (i32, [vp, u32, vp, vp, vp, vp]),
        "b3amd_broadphase_find_pairs": (i32, [vp, ctypes.POINTER(u32), ctypes.POINTER(vp)]),
        "b3amd_broadphase_test_overlap": (i32, [vp, u32, vp, vp]),
        "b3amd_broadphase_info": (i32, [vp, ctypes.POINTER(u32), ctypes.POINTER(u32)]),
        "b3amd_broadphase_read": (i32, [vp, u32, u32, vp, vp, vp]),
        "b3amd_shape_aabbs": (i32, [vp, u32, vp, vp, u32, vp, vp, vp, vp]),
        "b3amd_contact_manager_create": (i32, [vp, vp, ctypes.POINTER(vp)]),
        "b3amd_contact_manager_destroy": (i32, [vp]),
        "b3amd_contact_manager_set_bodies": (i32, [vp, u32, vp]),
        "b3amd_contact_manager_get_bodies": (i32, [vp, vp]),
        "b3amd_contact_manager_add_shapes": (i32, [vp, u32, vp, vp, vp, vp]),
        "b3amd_contact_manager_synchronize": (i32, [vp, u32, vp, vp, vp]),
        "b3amd_contact_manager_find_new_contacts": (i32, [vp, ctypes.POINTER(u32)]),
        "b3amd_contact_manager_update_contacts": (i32, [vp, vp, vp, vp]),
        "b3amd_contact_manager_read": (i32, [vp, u32, ctypes.POINTER(u32), vp, vp, vp, vp]),
        "b3amd_contact_manager_view": (i32, [vp, ctypes.POINTER(u32), ctypes.POINTER(vp), ctypes.POINTER(vp),
                                             ctypes.POINTER(vp)]),
        "b3amd_host_world_create": (i32, [vp, ctypes.POINTER(vp)]),
        "b3amd_host_world_destroy": (i32, [vp]),
        "b3amd_host_world_settings": (i32, [vp, vp, i32, i32]),
        "b3amd_host_world_create_body": (i32, [vp, vp, ctypes.POINTER(u32)]),
        "b3amd_host_world_create_shape": (i32, [vp, u32, vp, ctypes.POINTER(u32)]),
        "b3amd_host_world_step": (i32, [vp, ctypes.c_float, u32, u32]),
        "b3amd_host_world_bodies": (i32, [vp, u32, u32, vp]),
        "b3amd_host_world_contact_count": (i32, [vp, ctypes.POINTER(u32)]),
        "b3amd_set_counting": (i32, [vp, i32]),
        "b3amd_get_counters": (i32, [vp, vp]),
        "b3amd_reset_counters": (i32, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc != OK:
        raise BounceError(rc, load_library().b3amd_last_error().decode())


class DeviceBuffer:
    """Device memory owned by a Context (counts against its HBM budget)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(ctx.lib.b3amd_device_alloc(ctx.handle, self.nbytes, ctypes.byref(p)))
        self.ptr = p

    def upload(self, array):
        a = np.ascontiguousarray(array)
        assert a.nbytes <= self.nbytes
        _check(self.ctx.lib.b3amd_copy_to_device(self.ctx.handle, self.ptr, ptr(a), a.nbytes))

    def download(self, array):
        assert array.flags["C_CONTIGUOUS"] and array.nbytes <= self.nbytes
        _check(self.ctx.lib.b3amd_copy_to_host(self.ctx.handle, ptr(array), self.ptr, array.nbytes))
        return array

    def free(self):
        if self.ptr is not None and self.ctx.handle is not None:
            _check(self.ctx.lib.b3amd_device_free(self.ctx.handle, self.ptr))
        self.ptr = None

    def view(self, offset, nbytes):
        """A byte range of this buffer, usable wherever a DeviceBuffer is (not freeable)."""
        if offset < 0 or nbytes < 0 or offset + nbytes > self.nbytes:
            raise ValueError("view [%d, %d) outside a buffer of %d bytes" % (offset, offset + nbytes, self.nbytes))
        return DeviceView(self, offset, nbytes)


class DeviceView(DeviceBuffer):
    """A byte range inside a DeviceBuffer; the parent owns the memory."""

    def __init__(self, parent, offset, nbytes):
        self.ctx = parent.ctx
        self.parent = parent
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p(parent.ptr.value + int(offset))

    def free(self):
        raise TypeError("a DeviceView does not own its memory")


class CacheArray(DeviceBuffer):
    """Device b3ConvexCache array from Context.cache_alloc."""

    def __init__(self, ctx, p, count):
        self.ctx = ctx
        self.ptr = p
        self.count = count
        self.nbytes = 36 * count

    def reset(self, first=0, count=None):
        count = self.count - first if count is None else count
        if first < 0 or count < 0 or first + count > self.count:
            raise ValueError("reset range outside the array")
        _check(self.ctx.lib.b3amd_cache_reset(self.ctx.handle, self.ptr, first, count))

    def free(self):
        if self.ptr is not None and self.ctx.handle is not None:
            _check(self.ctx.lib.b3amd_cache_free(self.ctx.handle, self.ptr))
        self.ptr = None


class Graph:
    """A captured b3amd_collide_device batch (b3amd_graph_create)."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.handle = handle

    def launch(self):
        _check(self.ctx.lib.b3amd_graph_launch(self.handle))

    def close(self):
        if self.handle is not None:
            _check(self.ctx.lib.b3amd_graph_destroy(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class ServerStats(ctypes.Structure):
    """b3amd_server_stats."""
    _fields_ = [("batches", ctypes.c_uint64), ("requests", ctypes.c_uint64), ("pairs", ctypes.c_uint64),
                ("max_batch_pairs", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("gpu_ms", ctypes.c_double), ("pack_ms", ctypes.c_double)]


class Server:
    """Continuous batching of many worlds' narrowphase steps (b3amd_server_create)."""

    def __init__(self, ctx, max_batch_pairs, max_batch_bodies, slab_slots):
        self.ctx = ctx
        self.lib = ctx.lib
        h = ctypes.c_void_p()
        _check(self.lib.b3amd_server_create(ctx.handle, max_batch_pairs, max_batch_bodies, slab_slots,
                                            ctypes.byref(h)))
        self.handle = h
        self.pair_counts = {}

    def close(self):
        if self.handle is not None:
            _check(self.lib.b3amd_server_destroy(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def create_world(self, body_shapes, body_pairs):
        bs = np.ascontiguousarray(body_shapes, np.uint32)
        bp = np.ascontiguousarray(body_pairs, np.uint32).reshape(-1, 2)
        w = ctypes.c_uint32()
        _check(self.lib.b3amd_world_create(self.handle, len(bs), ptr(bs), len(bp), ptr(bp), ctypes.byref(w)))
        self.pair_counts[w.value] = len(bp)
        return w.value

    def admit_world(self, body_shapes, body_pairs, timeout_ms):
        """create_world that waits up to timeout_ms for slots (FIFO among admissions)."""
        bs = np.ascontiguousarray(body_shapes, np.uint32)
        bp = np.ascontiguousarray(body_pairs, np.uint32).reshape(-1, 2)
        w = ctypes.c_uint32()
        f = self.lib.b3amd_world_admit
        f.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                      ctypes.c_double, ctypes.c_void_p]
        _check(f(self.handle, len(bs), ptr(bs), len(bp), ptr(bp), float(timeout_ms), ctypes.byref(w)))
        self.pair_counts[w.value] = len(bp)
        return w.value

    def destroy_world(self, world):
        _check(self.lib.b3amd_world_destroy(self.handle, world))
        self.pair_counts.pop(world, None)

    def submit(self, world, body_xfs, body_sweeps=None):
        x = np.ascontiguousarray(body_xfs, np.float32)
        sw = None if body_sweeps is None else np.ascontiguousarray(body_sweeps, np.float32)
        t = ctypes.c_uint64()
        _check(self.lib.b3amd_world_submit(self.handle, world, ptr(x), ptr(sw), ctypes.byref(t)))
        return t.value

    def wait(self, ticket):
        """Blocks until the step is done; returns its latency in ms."""
        ms = ctypes.c_double()
        _check(self.lib.b3amd_server_wait(self.handle, ticket, ctypes.byref(ms)))
        return ms.value

    def read_manifolds(self, world):
        out = np.zeros(self.pair_counts[world], MANIFOLD_DTYPE)
        _check(self.lib.b3amd_world_read_manifolds(self.handle, world, ptr(out)))
        return out

    def read_caches(self, world):
        out = empty_caches(self.pair_counts[world])
        _check(self.lib.b3amd_world_read_caches(self.handle, world, ptr(out)))
        return out

    def replay(self, worlds, times_s, body_xfs):
        """Load generator: request i = (worlds[i], body_xfs[i]) at times_s[i]; returns the
        latencies (ms) from each scheduled arrival.  body_xfs holds float32 arrays."""
        worlds = np.ascontiguousarray(worlds, np.uint32)
        times = np.ascontiguousarray(times_s, np.float64)
        ptrs = np.array([a.ctypes.data for a in body_xfs], np.uint64)
        lat = np.zeros(len(worlds), np.float64)
        _check(self.lib.b3amd_server_replay(self.handle, len(worlds), ptr(worlds), ptr(times), ptr(ptrs), ptr(lat)))
        return lat

    def configure(self, batch_pairs=0, max_inflight=2):
        """Batch cap (0: the server maximum) and device pipeline depth (1..3)."""
        _check(self.lib.b3amd_server_configure(self.handle, batch_pairs, max_inflight))

    def stats(self):
        st = ServerStats()
        _check(self.lib.b3amd_server_get_stats(self.handle, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_ if k != "reserved"}


class Broadphase:
    """b3BroadPhase on the device (b3amd_broadphase_*): the reference's proxy ids and
    FindPairs order, fat AABBs in HBM.  Close it before its Context."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.lib = ctx.lib
        h = ctypes.c_void_p()
        _check(self.lib.b3amd_broadphase_create(ctx.handle, ctypes.byref(h)))
        self.handle = h
        self._bufs = {}

    def close(self):
        if self.handle is not None:
            _check(self.lib.b3amd_broadphase_destroy(self.handle))
            self.handle = None
        for b in self._bufs.values():
            b.free()
        self._bufs = {}

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _buf(self, name, nbytes):
        b = self._bufs.get(name)
        if b is None or b.nbytes < nbytes:
            if b is not None:
                b.free()
            b = DeviceBuffer(self.ctx, max(256, int(nbytes)))
            self._bufs[name] = b
        return b

    def create(self, aabbs, users=None):
        """CreateProxy for each AABB [n, 6] in order; returns the proxy ids."""
        a = np.ascontiguousarray(aabbs, np.float32).reshape(-1, 6)
        u = None if users is None else np.ascontiguousarray(users, np.uint32)
        ids = np.zeros(max(1, len(a)), np.uint32)
        _check(self.lib.b3amd_broadphase_create_proxies(self.handle, len(a), ptr(a), ptr(u), ptr(ids)))
        return ids[:len(a)]

    def destroy(self, ids):
        ids = np.ascontiguousarray(ids, np.uint32)
        _check(self.lib.b3amd_broadphase_destroy_proxies(self.handle, len(ids), ptr(ids)))

    def touch(self, ids):
        ids = np.ascontiguousarray(ids, np.uint32)
        _check(self.lib.b3amd_broadphase_touch_proxies(self.handle, len(ids), ptr(ids)))

    def move(self, ids, aabbs, displacements):
        """MoveProxy for host arrays (uploaded, run, moved flags downloaded)."""
        ids = np.ascontiguousarray(ids, np.uint32)
        n = len(ids)
        if n == 0:
            return np.zeros(0, np.uint32)
        a = np.ascontiguousarray(aabbs, np.float32).reshape(n, 6)
        d = np.ascontiguousarray(displacements, np.float32).reshape(n, 3)
        bi, ba, bd, bm = (self._buf(k, s) for k, s in (("ids", 4 * n), ("aabbs", 24 * n), ("disp", 12 * n),
                                                       ("moved", 4 * n)))
        bi.upload(ids), ba.upload(a), bd.upload(d)
        self.move_device(n, bi, ba, bd, bm)
        return bm.download(np.zeros(n, np.uint32))

    def move_device(self, n, d_ids, d_aabbs, d_disps, d_moved=None):
        _check(self.lib.b3amd_broadphase_move_proxies(self.handle, n, d_ids.ptr, d_aabbs.ptr, d_disps.ptr,
                                                      None if d_moved is None else d_moved.ptr))

    def find_pairs_device(self):
        """FindPairs: (count, device address of the [count, 4] pair words)."""
        n = ctypes.c_uint32()
        p = ctypes.c_void_p()
        _check(self.lib.b3amd_broadphase_find_pairs(self.handle, ctypes.byref(n), ctypes.byref(p)))
        return n.value, p

    def find_pairs(self):
        """FindPairs: [k, 4] uint32 (proxy1, proxy2, user1, user2) in AddPair order."""
        n, p = self.find_pairs_device()
        out = np.zeros((n, 4), np.uint32)
        if n:
            _check(self.lib.b3amd_copy_to_host(self.ctx.handle, ptr(out), p, out.nbytes))
        return out

    def test_overlap(self, proxy_pairs):
        pp = np.ascontiguousarray(proxy_pairs, np.uint32).reshape(-1, 2)
        n = len(pp)
        if n == 0:
            return np.zeros(0, np.uint32)
        bp, bo = self._buf("tp", 8 * n), self._buf("to", 4 * n)
        bp.upload(pp)
        _check(self.lib.b3amd_broadphase_test_overlap(self.handle, n, bp.ptr, bo.ptr))
        return bo.download(np.zeros(n, np.uint32))

    def info(self):
        c, cap = ctypes.c_uint32(), ctypes.c_uint32()
        _check(self.lib.b3amd_broadphase_info(self.handle, ctypes.byref(c), ctypes.byref(cap)))
        return c.value, cap.value

    def read(self, first=0, count=None):
        """(fat AABBs [count, 6], user data, flags) of ids first .. first + count - 1."""
        if count is None:
            count = self.info()[1] - first
        a = np.zeros((count, 6), np.float32)
        u = np.zeros(count, np.uint32)
        f = np.zeros(count, np.uint32)
        _check(self.lib.b3amd_broadphase_read(self.handle, first, count, ptr(a), ptr(u), ptr(f)))
        return a, u, f


class ContactManager:
    """b3ContactManager for convex contacts on the device (b3amd_contact_manager_*), over a
    Broadphase; bodies are flag words (b3BodyType | BODY_AWAKE)."""

    def __init__(self, ctx, bp):
        self.ctx, self.bp, self.lib = ctx, bp, ctx.lib
        h = ctypes.c_void_p()
        _check(self.lib.b3amd_contact_manager_create(ctx.handle, bp.handle, ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle is not None:
            _check(self.lib.b3amd_contact_manager_destroy(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_bodies(self, flags):
        f = np.ascontiguousarray(flags, np.uint32)
        _check(self.lib.b3amd_contact_manager_set_bodies(self.handle, len(f), ptr(f)))

    def bodies(self, count):
        out = np.zeros(count, np.uint32)
        _check(self.lib.b3amd_contact_manager_get_bodies(self.handle, ptr(out)))
        return out

    def add_shapes(self, table_shapes, bodies, d_body_xfs):
        ts = np.ascontiguousarray(table_shapes, np.uint32)
        bo = np.ascontiguousarray(bodies, np.uint32)
        ids = np.zeros(max(1, len(ts)), np.uint32)
        _check(self.lib.b3amd_contact_manager_add_shapes(self.handle, len(ts), ptr(ts), ptr(bo), d_body_xfs.ptr, ptr(ids)))
        return ids[:len(ts)]

    def synchronize(self, n, d_shape_ids, d_body_xfs, d_body_sweeps):
        _check(self.lib.b3amd_contact_manager_synchronize(self.handle, n, d_shape_ids.ptr, d_body_xfs.ptr,
                                                          d_body_sweeps.ptr))

    def find_new_contacts(self):
        c = ctypes.c_uint32()
        _check(self.lib.b3amd_contact_manager_find_new_contacts(self.handle, ctypes.byref(c)))
        return c.value

    def update_contacts(self, d_body_xfs, d_body_sweeps=None):
        st = np.zeros(3, np.uint32)
        _check(self.lib.b3amd_contact_manager_update_contacts(self.handle, d_body_xfs.ptr,
                                                              None if d_body_sweeps is None else d_body_sweeps.ptr,
                                                              ptr(st)))
        return st

    def count(self):
        c = ctypes.c_uint32()
        _check(self.lib.b3amd_contact_manager_read(self.handle, 0, ctypes.byref(c), None, None, None, None))
        return c.value

    def read(self):
        """(shape pairs [n, 2], born [n], overlap [n], manifolds, caches) in list order."""
        n = self.count()
        pairs = np.zeros((max(n, 1), 2), np.uint32)
        born = np.zeros(max(n, 1), np.uint32)
        out = np.zeros(max(n, 1), MANIFOLD_DTYPE)
        caches = empty_caches(max(n, 1))
        c = ctypes.c_uint32()
        _check(self.lib.b3amd_contact_manager_read(self.handle, n, ctypes.byref(c), ptr(pairs), ptr(born), ptr(out),
                                                   ptr(caches)))
        return pairs[:n], born[:n] & 0x7FFFFFFF, born[:n] >> 31, out[:n], caches[:n]


class HostWorld:
    """b3World stepped on the host around the device pipeline (b3amd_host_world_*)."""

    def __init__(self, ctx):
        self.ctx, self.lib = ctx, ctx.lib
        h = ctypes.c_void_p()
        _check(self.lib.b3amd_host_world_create(ctx.handle, ctypes.byref(h)))
        self.handle = h
        self.count = 0

    def close(self):
        if self.handle is not None:
            _check(self.lib.b3amd_host_world_destroy(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def settings(self, gravity=None, warm_starting=True, sleeping=False):
        g = None if gravity is None else np.ascontiguousarray(gravity, np.float32)
        _check(self.lib.b3amd_host_world_settings(self.handle, ptr(g), int(warm_starting), int(sleeping)))

    def create_body(self, type=0, position=(0, 0, 0), orientation=(0, 0, 0, 1), awake=True, linear_velocity=(0, 0, 0),
                    angular_velocity=(0, 0, 0), gravity_scale=1.0, linear_damping=0.0, angular_damping=0.0,
                    fixed_rotation=0):
        from .types import BODY_DEF_DTYPE
        d = np.zeros(1, BODY_DEF_DTYPE)
        d["type"], d["awake"], d["fixed_rotation"] = type, int(awake), fixed_rotation
        d["position"], d["orientation"] = position, orientation
        d["linear_velocity"], d["angular_velocity"] = linear_velocity, angular_velocity
        d["gravity_scale"], d["linear_damping"], d["angular_damping"] = gravity_scale, linear_damping, angular_damping
        b = ctypes.c_uint32()
        _check(self.lib.b3amd_host_world_create_body(self.handle, ptr(d), ctypes.byref(b)))
        self.count += 1
        return b.value

    def create_shape(self, body, table_shape, density=0.0, friction=0.3, restitution=0.0, sensor=False):
        from .types import SHAPE_DEF_DTYPE
        d = np.zeros(1, SHAPE_DEF_DTYPE)
        d["table_shape"], d["density"], d["friction"], d["restitution"], d["sensor"] = (
            table_shape, density, friction, restitution, int(sensor))
        s = ctypes.c_uint32()
        _check(self.lib.b3amd_host_world_create_shape(self.handle, body, ptr(d), ctypes.byref(s)))
        return s.value

    def step(self, dt, velocity_iterations=8, position_iterations=2):
        _check(self.lib.b3amd_host_world_step(self.handle, ctypes.c_float(dt), velocity_iterations,
                                              position_iterations))

    def bodies(self):
        from .types import BODY_STATE_DTYPE
        out = np.zeros(max(1, self.count), BODY_STATE_DTYPE)
        _check(self.lib.b3amd_host_world_bodies(self.handle, 0, self.count, ptr(out)))
        return out[:self.count]

    def contact_count(self):
        n = ctypes.c_uint32()
        _check(self.lib.b3amd_host_world_contact_count(self.handle, ctypes.byref(n)))
        return n.value


class MeshContacts:
    """Persistent mesh contacts (b3amd_mesh_contacts_*): b3MeshContact across steps."""

    def __init__(self, ctx, mesh_id, shape_a, xfs):
        shape_a = np.ascontiguousarray(shape_a, np.uint32)
        self.n = len(shape_a)
        self.lib = ctx.lib
        xfs = np.ascontiguousarray(xfs, np.float32).reshape(self.n, 2, 12)
        h = ctypes.c_void_p()
        _check(self.lib.b3amd_mesh_contacts_create(ctx.handle, mesh_id, self.n, ptr(shape_a), ptr(xfs), ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            self.lib.b3amd_mesh_contacts_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def collide(self, xfs, sweeps=None):
        """b3MeshContact::Collide: (manifolds [n][3], manifold counts)."""
        xfs = np.ascontiguousarray(xfs, np.float32).reshape(self.n, 2, 12)
        if sweeps is not None:
            sweeps = np.ascontiguousarray(sweeps, np.float32).reshape(self.n, 4, 4)
        out = np.zeros((self.n, 3), MANIFOLD_DTYPE)
        counts = np.zeros(self.n, np.uint32)
        _check(self.lib.b3amd_mesh_contacts_collide(self.handle, ptr(xfs), ptr(sweeps), ptr(out), ptr(counts)))
        return out, counts

    def test_overlap(self, xfs):
        """b3MeshContact::TestOverlap per contact (sensor contacts); uint32 flags."""
        xfs = np.ascontiguousarray(xfs, np.float32).reshape(self.n, 2, 12)
        out = np.zeros(self.n, np.uint32)
        f = self.lib.b3amd_mesh_contacts_test_overlap
        f.argtypes = [ctypes.c_void_p] * 3
        _check(f(self.handle, ptr(xfs), ptr(out)))
        return out

    def synchronize(self, xfs, centers0):
        """SynchronizeShapes + FindNewPairs for new transforms (centers0: body A's worldCenter0)."""
        xfs = np.ascontiguousarray(xfs, np.float32).reshape(self.n, 2, 12)
        c0 = np.ascontiguousarray(centers0, np.float32).reshape(self.n, 3)
        _check(self.lib.b3amd_mesh_contacts_synchronize(self.handle, ptr(xfs), ptr(c0)))

    def collide_device(self, d_xfs, d_out, d_counts, d_sweeps=None):
        """b3amd_mesh_contacts_collide_device on resident buffers (enqueued, not synchronised)."""
        _check(self.lib.b3amd_mesh_contacts_collide_device(self.handle, d_xfs.ptr, None if d_sweeps is None else
                                                          d_sweeps.ptr, d_out.ptr, d_counts.ptr))

    def test_overlap_device(self, d_xfs, d_overlap):
        _check(self.lib.b3amd_mesh_contacts_test_overlap_device(self.handle, d_xfs.ptr, d_overlap.ptr))

    def synchronize_device(self, d_xfs, d_centers0):
        _check(self.lib.b3amd_mesh_contacts_synchronize_device(self.handle, d_xfs.ptr, d_centers0.ptr))

    def read(self):
        """(triangle counts, lists [n][64], caches [n][64], fat AABBs [n][6])."""
        from .types import CACHE_DTYPE
        cap = Context.MESH_TRIANGLE_CAP
        counts = np.zeros(self.n, np.uint32)
        lists = np.zeros((self.n, cap), np.uint32)
        caches = np.zeros((self.n, cap), CACHE_DTYPE)
        fat = np.zeros((self.n, 6), np.float32)
        _check(self.lib.b3amd_mesh_contacts_read(self.handle, ptr(counts), ptr(lists), ptr(caches), ptr(fat)))
        return counts, lists, caches, fat


class Context:
    """One narrowphase context on one GPU (b3amd_create)."""

    def __init__(self, device=0, hbm_budget_bytes=0):
        self.lib = load_library()
        h = ctypes.c_void_p()
        _check(self.lib.b3amd_create(device, hbm_budget_bytes, ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle is not None:
            _check(self.lib.b3amd_destroy(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def memory_in_use(self):
        return int(self.lib.b3amd_memory_in_use(self.handle))

    def upload_hulls(self, pool):
        s = pool.struct()
        _check(self.lib.b3amd_upload_hulls(self.handle, ctypes.byref(s)))

    def add_hulls(self, pool):
        """Adds a HullPool's hulls to the registry; returns their handles (uint32)."""
        s = pool.struct()
        handles = np.zeros(max(1, pool.count), np.uint32)
        _check(self.lib.b3amd_hull_add(self.handle, ctypes.byref(s), ptr(handles)))
        return handles[:pool.count]

    def free_hull(self, handle):
        _check(self.lib.b3amd_hull_free(self.handle, int(handle)))

    def hull_registry(self):
        """dict: slots, live, used_bytes, capacity_bytes."""
        slots, live = ctypes.c_uint32(), ctypes.c_uint32()
        used, cap = ctypes.c_uint64(), ctypes.c_uint64()
        _check(self.lib.b3amd_hull_registry_info(self.handle, ctypes.byref(slots), ctypes.byref(live),
                                                 ctypes.byref(used), ctypes.byref(cap)))
        return dict(slots=slots.value, live=live.value, used_bytes=used.value, capacity_bytes=cap.value)

    def collide_one(self, xf_a, shape_a, xf_b, shape_b, sweeps=None, cache=None):
        """One pair through b3amd_collide_shape_and_shape; returns the manifold record.
        `cache` (a one-element CACHE_DTYPE array) is updated in place when given."""
        m = np.zeros(1, MANIFOLD_DTYPE)
        a = np.ascontiguousarray(xf_a, np.float32).reshape(12)
        b = np.ascontiguousarray(xf_b, np.float32).reshape(12)
        if sweeps is not None:
            sweeps = np.ascontiguousarray(sweeps, np.float32).reshape(4, 4)
        if cache is not None:
            assert cache.dtype == CACHE_DTYPE and cache.flags["C_CONTIGUOUS"] and len(cache) == 1
        _check(self.lib.b3amd_collide_shape_and_shape(self.handle, ptr(m), ptr(a), int(shape_a), ptr(b), int(shape_b),
                                                      ptr(sweeps), ptr(cache)))
        return m[0]

    def broadphase(self):
        """A device b3BroadPhase on this context."""
        return Broadphase(self)

    def contact_manager(self, bp):
        """A device b3ContactManager over the Broadphase `bp`."""
        return ContactManager(self, bp)

    def host_world(self):
        """A b3World stepped on the host around this context's device pipeline."""
        return HostWorld(self)

    def shape_aabbs(self, shapes, bodies, body_xfs, body_sweeps=None):
        """b3Body::SynchronizeShapes' swept AABB and displacement of shape shapes[i] on body
        bodies[i] (host arrays in and out); with body_sweeps None, the AABB at m_xf."""
        from .types import SWEEP_DTYPE
        sh = np.ascontiguousarray(shapes, np.uint32)
        bo = np.ascontiguousarray(bodies, np.uint32)
        n = len(sh)
        x = np.ascontiguousarray(body_xfs, np.float32).reshape(-1, 12)
        bufs = [DeviceBuffer(self, max(256, a.nbytes)) for a in (sh, bo, x)]
        for b, a in zip(bufs, (sh, bo, x)):
            b.upload(a)
        sw = None
        if body_sweeps is not None:
            s_arr = np.ascontiguousarray(body_sweeps, SWEEP_DTYPE)
            sw = DeviceBuffer(self, max(256, s_arr.nbytes))
            sw.upload(s_arr)
        oa, od = DeviceBuffer(self, max(256, 24 * n)), DeviceBuffer(self, max(256, 12 * n))
        try:
            _check(self.lib.b3amd_shape_aabbs(self.handle, n, bufs[0].ptr, bufs[1].ptr, len(x), bufs[2].ptr,
                                              None if sw is None else sw.ptr, oa.ptr, od.ptr))
            return oa.download(np.zeros((n, 6), np.float32)), od.download(np.zeros((n, 3), np.float32))
        finally:
            for b in bufs + [oa, od] + ([sw] if sw is not None else []):
                b.free()

    def cache_alloc(self, count):
        """Device caches in the new-contact state; a DeviceBuffer-like handle (free with
        cache_free)."""
        p = ctypes.c_void_p()
        _check(self.lib.b3amd_cache_alloc(self.handle, int(count), ctypes.byref(p)))
        return CacheArray(self, p, int(count))

    def upload_shapes(self, shapes):
        shapes = np.ascontiguousarray(shapes, dtype=SHAPE_DTYPE)
        _check(self.lib.b3amd_upload_shapes(self.handle, len(shapes), ptr(shapes)))

    def collide(self, pairs, xfs, sweeps=None, caches=None):
        """b3CollideShapeAndShape over a batch of host arrays; returns manifolds.
        `caches` (CACHE_DTYPE) is updated in place when given."""
        pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        n = pairs.shape[0]
        xfs = np.ascontiguousarray(xfs, np.float32).reshape(n, 2, 12)
        if sweeps is not None:
            sweeps = np.ascontiguousarray(sweeps, np.float32).reshape(n, 4, 4)
        if caches is not None:
            assert caches.dtype == CACHE_DTYPE and caches.flags["C_CONTIGUOUS"] and len(caches) == n
        out = np.zeros(n, MANIFOLD_DTYPE)
        _check(self.lib.b3amd_collide(self.handle, n, ptr(pairs), ptr(xfs), ptr(sweeps), ptr(caches), ptr(out)))
        return out

    def upload_mesh(self, vertices, triangles):
        """b3Mesh upload with its static tree; returns the mesh id."""
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
        mid = ctypes.c_uint32()
        _check(self.lib.b3amd_upload_mesh(self.handle, len(v), ptr(v), len(t), ptr(t), ctypes.byref(mid)))
        return mid.value

    def mesh_tree(self, mesh_id):
        """Static-tree nodes [2T-1][8] (float32 view; words 6-7 are u32 children)."""
        n = ctypes.c_uint32()
        _check(self.lib.b3amd_mesh_tree(self.handle, mesh_id, ctypes.byref(n), None))
        nodes = np.zeros((n.value, 8), np.float32)
        _check(self.lib.b3amd_mesh_tree(self.handle, mesh_id, ctypes.byref(n), ptr(nodes)))
        return nodes

    MESH_TRIANGLE_CAP = 64
    MESH_WIDE_TRIANGLES = 2048  # B3AMD_MESH_WIDE_TRIANGLES: fresh contacts collided whole up to this

    def collide_mesh(self, mesh_id, shape_a, xfs):
        """Fresh mesh contacts: (triangle counts, triangle lists [n][64], manifolds [n][3],
        manifold counts)."""
        shape_a = np.ascontiguousarray(shape_a, np.uint32)
        n = len(shape_a)
        xfs = np.ascontiguousarray(xfs, np.float32).reshape(n, 2, 12)
        counts = np.zeros(n, np.uint32)
        lists = np.zeros((n, self.MESH_TRIANGLE_CAP), np.uint32)
        out = np.zeros((n, 3), MANIFOLD_DTYPE)
        mcount = np.zeros(n, np.uint32)
        _check(self.lib.b3amd_collide_mesh(self.handle, mesh_id, n, ptr(shape_a), ptr(xfs), ptr(counts), ptr(lists),
                                           ptr(out), ptr(mcount)))
        return counts, lists, out, mcount

    def collide_mesh_device(self, mesh_id, n, d_shape_a, d_xfs, d_out, d_manifold_counts, d_triangle_counts=None,
                            d_triangle_lists=None):
        """b3amd_collide_mesh_device on DeviceBuffers (no synchronisation)."""
        _check(self.lib.b3amd_collide_mesh_device(
            self.handle, mesh_id, n, d_shape_a.ptr, d_xfs.ptr, d_triangle_counts.ptr if d_triangle_counts else None,
            d_triangle_lists.ptr if d_triangle_lists else None, d_out.ptr, d_manifold_counts.ptr))

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def copy_on_device(self, dst, src, nbytes):
        _check(self.lib.b3amd_copy_on_device(self.handle, dst.ptr, src.ptr, int(nbytes)))

    def contact_init(self, M, bodies, state, material, counts, manifolds, old_counts, old_manifolds, old_impulses,
                     warm=True):
        """b3amd_contact_init on host arrays (uploaded, run, read back): warm-start matching
        and b3ContactSolver::InitializeConstraints for n contacts of M manifold slots.
        Returns (impulses [n, M], velocity constraints [n, M])."""
        from .types import IMPULSES_DTYPE, VELOCITY_MANIFOLD_DTYPE
        n = len(counts)
        arrays = [np.ascontiguousarray(bodies, np.uint32), np.ascontiguousarray(state), np.ascontiguousarray(material),
                  np.ascontiguousarray(counts, np.uint32), np.ascontiguousarray(manifolds, MANIFOLD_DTYPE),
                  np.ascontiguousarray(old_counts, np.uint32), np.ascontiguousarray(old_manifolds, MANIFOLD_DTYPE),
                  np.ascontiguousarray(old_impulses, IMPULSES_DTYPE)]
        imp = np.zeros((n, M), IMPULSES_DTYPE)
        out = np.zeros((n, M), VELOCITY_MANIFOLD_DTYPE)
        bufs = [self.alloc(max(16, a.nbytes)) for a in arrays] + [self.alloc(imp.nbytes), self.alloc(out.nbytes)]
        try:
            for b, a in zip(bufs, arrays):
                b.upload(a)
            bufs[-2].upload(imp)
            bufs[-1].upload(out)
            f = self.lib.b3amd_contact_init
            f.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32] + [ctypes.c_void_p] * 8 + \
                [ctypes.c_int] + [ctypes.c_void_p] * 2
            _check(f(self.handle, n, M, *[b.ptr for b in bufs[:8]], int(bool(warm)), bufs[8].ptr, bufs[9].ptr))
            self.synchronize()
            return bufs[8].download(imp), bufs[9].download(out)
        finally:
            for b in bufs:
                b.free()

    def test_overlap(self, pairs, xfs, caches=None):
        """b3TestOverlap per pair (sensors); caches (CACHE_DTYPE, optional) are read and
        updated in place.  Returns the overlap flags (uint32: 1 overlapping, 0 apart)."""
        pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        n = len(pairs)
        xfs = np.ascontiguousarray(xfs, np.float32).reshape(n, 2, 12)
        out = np.zeros(n, np.uint32)
        if caches is not None:
            assert caches.dtype == CACHE_DTYPE and len(caches) == n and caches.flags["C_CONTIGUOUS"]
        f = self.lib.b3amd_test_overlap
        f.argtypes = [ctypes.c_void_p, ctypes.c_uint32] + [ctypes.c_void_p] * 4
        _check(f(self.handle, n, ptr(pairs), ptr(xfs), None if caches is None else ptr(caches), ptr(out)))
        return out

    def collide_device(self, n, d_pairs, d_xfs, d_out, d_sweeps=None, d_caches=None):
        """Enqueue a batch on device-resident buffers (no synchronisation)."""
        p = lambda b: None if b is None else b.ptr
        _check(self.lib.b3amd_collide_device(self.handle, n, p(d_pairs), p(d_xfs), p(d_sweeps), p(d_caches),
                                             p(d_out)))

    def update_contacts(self, body_count, d_body_shapes, d_body_xfs, contact_count, d_contacts, d_out,
                        d_body_sweeps=None, d_caches=None):
        """Enqueue one world step from device-resident body state (b3amd_update_contacts):
        contact i joins bodies d_contacts[2i], d_contacts[2i+1]."""
        p = lambda b: None if b is None else b.ptr
        _check(self.lib.b3amd_update_contacts(self.handle, body_count, p(d_body_shapes), p(d_body_xfs),
                                              p(d_body_sweeps), contact_count, p(d_contacts), p(d_caches),
                                              p(d_out)))

    def synchronize(self):
        _check(self.lib.b3amd_synchronize(self.handle))

    def capture(self, n, d_pairs, d_xfs, d_out, d_sweeps=None, d_caches=None):
        """Captures one device batch as a replayable HIP graph."""
        p = lambda b: None if b is None else b.ptr
        h = ctypes.c_void_p()
        _check(self.lib.b3amd_graph_create(self.handle, n, p(d_pairs), p(d_xfs), p(d_sweeps), p(d_caches), p(d_out),
                                           ctypes.byref(h)))
        return Graph(self, h)

    def set_sat_tuning(self, waves_per_simd, tail_chunk):
        """Launch shape of the persistent SAT kernel (b3amd_set_sat_tuning)."""
        _check(self.lib.b3amd_set_sat_tuning(self.handle, int(waves_per_simd), int(tail_chunk)))

    def sat_tuning(self):
        w, k = ctypes.c_uint32(), ctypes.c_uint32()
        _check(self.lib.b3amd_get_sat_tuning(self.handle, ctypes.byref(w), ctypes.byref(k)))
        return w.value, k.value

    def autotune(self, n, d_pairs, d_xfs, d_out, repeats=2):
        """Times every SAT launch shape on a device batch, keeps the fastest; returns
        ((waves_per_simd, tail_chunk), mean ms per batch)."""
        ms = ctypes.c_float()
        _check(self.lib.b3amd_autotune(self.handle, n, d_pairs.ptr, d_xfs.ptr, d_out.ptr, repeats, ctypes.byref(ms)))
        return self.sat_tuning(), ms.value

    COUNTERS = ("convex_calls", "convex_cache_hits", "gjk_calls", "gjk_iters", "gjk_max_iters", "gjk_cache_hits")

    def set_counting(self, on):
        """Statistics counters (b3_convexCalls, b3_convexCacheHits, b3_gjk*) on or off."""
        _check(self.lib.b3amd_set_counting(self.handle, 1 if on else 0))

    def counters(self):
        raw = np.zeros(5, np.uint64)
        _check(self.lib.b3amd_get_counters(self.handle, ptr(raw)))
        vals = [int(raw[0]), int(raw[1]), int(raw[2]), int(raw[3]), int(raw[4]) & 0xFFFFFFFF, int(raw[4]) >> 32]
        return dict(zip(self.COUNTERS, vals))

    def reset_counters(self):
        _check(self.lib.b3amd_reset_counters(self.handle))

    def set_profiling(self, on):
        _check(self.lib.b3amd_set_profiling(self.handle, 1 if on else 0))

    def kernel_stats(self, name):
        """(launches, total_ms) recorded with HIP events on the context stream."""
        n = ctypes.c_uint32()
        ms = ctypes.c_double()
        _check(self.lib.b3amd_kernel_stats(self.handle, name.encode(), ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value

    def reset_kernel_stats(self):
        _check(self.lib.b3amd_reset_kernel_stats(self.handle))
