"""ctypes bindings of the CPU oracle (oracle/liboracle.so) and of the
reference-header harness (oracle/_ref/libref.so).

Test infrastructure: imported by tests, __graft_entry__.smoke() and bench.py's
cpu_baseline leg only.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref.so")
REF_FAST_SO = os.path.join(ROOT, "oracle", "_ref", "libref_fast.so")

PHOTON_DTYPE = np.dtype([("pos", "<f8", 3), ("dir", "<f8", 3), ("tau", "<f8"),
                         ("sigma", "<f8"), ("weight", "<f8")])

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_u64p = C.POINTER(C.c_uint64)


def ptr(a, t):
    return a.ctypes.data_as(t)


class OrcSubgrid(C.Structure):
    _fields_ = [("anchor", C.c_double * 3), ("cell_size", C.c_double * 3),
                ("inv_cell_size", C.c_double * 3), ("ncell", C.c_int * 4),
                ("ngbs", C.c_uint32 * 27), ("n", _dp), ("xn", _dp),
                ("J", _dp)]


class OrcGrid(C.Structure):
    _fields_ = [("box", C.c_double * 6), ("ncell", C.c_int * 3),
                ("nsub", C.c_int * 3), ("nsubgrid", C.c_int),
                ("cells_per_subgrid", C.c_size_t),
                ("ncell_total", C.c_size_t),
                ("sub", C.POINTER(OrcSubgrid)), ("n", _dp), ("xn", _dp),
                ("J", _dp)]


def load_oracle():
    lib = C.CDLL(ORACLE_SO)
    lib.orc_output_to_input_direction.argtypes = [C.c_int]
    lib.orc_direction_offset.argtypes = [C.c_int, _ip]
    lib.orc_direction_from_offset.argtypes = [_ip]
    lib.orc_direction_from_mask.argtypes = [C.c_int]
    lib.orc_get_output_direction.argtypes = [C.POINTER(OrcSubgrid), _ip]
    lib.orc_get_start_index.argtypes = [C.POINTER(OrcSubgrid), _dp, C.c_int,
                                        _ip]
    lib.orc_subgrid_init.argtypes = [C.POINTER(OrcSubgrid), _dp, _ip]
    lib.orc_interact.argtypes = [C.POINTER(OrcSubgrid), C.c_void_p, C.c_int,
                                 _u32p]
    lib.orc_walk.argtypes = [C.POINTER(OrcSubgrid), C.c_void_p, C.c_int,
                             C.c_int, _u32p]
    lib.orc_is_compatible_output_direction.argtypes = [_dp, C.c_int]
    lib.orc_is_compatible_input_direction.argtypes = [_dp, C.c_int]
    lib.orc_grid_create.argtypes = [C.POINTER(OrcGrid), _dp, _ip, _ip]
    lib.orc_grid_destroy.argtypes = [C.POINTER(OrcGrid)]
    lib.orc_grid_source_subgrid.argtypes = [C.POINTER(OrcGrid), _dp]
    lib.orc_grid_source_subgrid.restype = C.c_uint32
    lib.orc_transport_list.argtypes = [C.POINTER(OrcGrid), C.c_void_p, _u32p,
                                       C.c_size_t, _i32p, _u32p, _u32p,
                                       C.c_int]
    lib.orc_transport_list_reemit.argtypes = [
        C.POINTER(OrcGrid), C.c_void_p, _u32p, C.c_size_t, C.c_double,
        C.c_uint64, C.c_uint32, C.c_uint64, _i32p, _u32p, _u32p, _u32p,
        C.c_int]
    lib.orc_reemit_draw.argtypes = [C.c_uint64, C.c_uint32, C.c_double,
                                    C.c_uint64, C.c_uint32, _dp, _dp]
    lib.orc_reemit_draw.restype = C.c_int
    lib.orc_compute_neutral_fraction.argtypes = [C.POINTER(OrcGrid), C.c_double,
                                                 C.c_double, C.c_double,
                                                 C.c_int]
    lib.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
    lib.orc_uniform_open.argtypes = [C.c_uint32, C.c_uint32]
    lib.orc_uniform_open.restype = C.c_double
    lib.orc_source_counts.argtypes = [_dp, C.c_int, C.c_uint64, C.c_uint64,
                                      C.c_uint32, _u64p]
    lib.orc_emit.argtypes = [C.POINTER(OrcGrid), _dp, _dp, C.c_int, C.c_uint64,
                             C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32,
                             C.c_double, C.c_void_p, _u32p]
    lib.orc_emit_problem.argtypes = [C.POINTER(OrcGrid), _dp, _dp, C.c_int,
                                     C.c_uint64, C.c_uint64, C.c_uint64,
                                     C.c_uint64, C.c_uint32, C.c_double,
                                     C.c_uint32, C.c_void_p, _u32p]
    lib.orc_iteration.argtypes = [C.POINTER(OrcGrid), _dp, _dp, C.c_int,
                                  C.c_uint64, C.c_uint64, C.c_uint32,
                                  C.c_double, C.c_double, C.c_int, C.c_int,
                                  _dp]
    lib.orc_iteration.restype = C.c_uint64
    return lib


def load_ref(fast=False):
    path = REF_FAST_SO if fast else REF_SO
    if not os.path.exists(path):
        return None
    lib = C.CDLL(path)
    lib.ref_get_output_direction.argtypes = [_ip, _ip]
    lib.ref_direction_from_mask.argtypes = [C.c_int]
    lib.ref_output_to_input_direction.argtypes = [C.c_int]
    lib.ref_is_compatible_output_direction.argtypes = [_dp, C.c_int]
    lib.ref_is_compatible_input_direction.argtypes = [_dp, C.c_int]
    lib.ref_get_start_index.argtypes = [_dp, _ip, _dp, C.c_int, _ip]
    lib.ref_interact_list.argtypes = [_dp, _ip, _dp, _dp, _dp, C.c_void_p,
                                      C.c_size_t, C.c_int, _i32p, _u32p]
    lib.ref_walk_list.argtypes = [_dp, _ip, _dp, _dp, C.c_void_p, C.c_size_t,
                                  C.c_int, C.c_int, _i32p, _u32p]
    lib.ref_transport_list.argtypes = [_dp, _ip, _ip, _dp, _dp, _dp,
                                       C.c_void_p, _u32p, C.c_size_t, _i32p,
                                       _u32p, _u32p]
    lib.ref_compute_neutral_fraction.argtypes = [_dp, _ip, _dp, _dp, C.c_double,
                                                 C.c_uint, _dp]
    lib.ref_transport_tasks.argtypes = [_dp, _ip, _ip, _dp, _dp, _dp,
                                        C.c_void_p, _u32p, C.c_size_t, C.c_int,
                                        C.c_size_t]
    lib.ref_transport_tasks.restype = C.c_double
    return lib


class OracleGrid:
    """Owner of an orc_grid (subgrid-major field arrays exposed as numpy)."""

    def __init__(self, lib, box, ncell, nsub):
        self.lib = lib
        self.g = OrcGrid()
        b = np.ascontiguousarray(box, dtype=np.float64)
        nc = np.ascontiguousarray(ncell, dtype=np.int32)
        ns = np.ascontiguousarray(nsub, dtype=np.int32)
        rc = lib.orc_grid_create(C.byref(self.g), ptr(b, _dp), ptr(nc, _ip),
                                 ptr(ns, _ip))
        if rc != 0:
            raise ValueError("orc_grid_create failed: %d" % rc)
        n = self.g.ncell_total
        self.n = np.ctypeslib.as_array(self.g.n, shape=(n,))
        self.xn = np.ctypeslib.as_array(self.g.xn, shape=(n,))
        self.J = np.ctypeslib.as_array(self.g.J, shape=(n,))
        self.box = b
        self.ncell = nc
        self.nsub = ns

    def __del__(self):
        try:
            self.lib.orc_grid_destroy(C.byref(self.g))
        except Exception:
            pass

    def source_subgrid(self, pos):
        p = np.ascontiguousarray(pos, dtype=np.float64)
        return self.lib.orc_grid_source_subgrid(C.byref(self.g), ptr(p, _dp))

    def transport(self, photons, start_sub, nthreads=1):
        n = len(photons)
        st = np.zeros(n, np.int32)
        cr = np.zeros(n, np.uint32)
        hp = np.zeros(n, np.uint32)
        ss = np.ascontiguousarray(start_sub, dtype=np.uint32)
        self.lib.orc_transport_list(C.byref(self.g), photons.ctypes.data,
                                    ptr(ss, _u32p), n, ptr(st, _i32p),
                                    ptr(cr, _u32p), ptr(hp, _u32p), nthreads)
        return st, cr, hp

    def transport_reemit(self, photons, start_sub, probability, seed,
                         iteration, first=0, nthreads=1):
        """Transport with reemission (orc_transport_list_reemit); returns
        status, crossings, hops and reemissions per packet."""
        n = len(photons)
        st = np.zeros(n, np.int32)
        cr = np.zeros(n, np.uint32)
        hp = np.zeros(n, np.uint32)
        re = np.zeros(n, np.uint32)
        ss = np.ascontiguousarray(start_sub, dtype=np.uint32)
        self.lib.orc_transport_list_reemit(
            C.byref(self.g), photons.ctypes.data, ptr(ss, _u32p), n,
            float(probability), int(seed), int(iteration), int(first),
            ptr(st, _i32p), ptr(cr, _u32p), ptr(hp, _u32p), ptr(re, _u32p),
            nthreads)
        return st, cr, hp, re

    def compute_neutral_fraction(self, L, N, alpha, form=0):
        self.lib.orc_compute_neutral_fraction(C.byref(self.g), L, float(N),
                                              alpha, form)

    def emit(self, src_pos, src_L, npackets, seed, iteration, sigma=6.3e-22,
             first=0, count=None, problem=None):
        """orc_emit, or orc_emit_problem for problem q of a problem set."""
        if count is None:
            count = npackets - first
        sp = np.ascontiguousarray(src_pos, dtype=np.float64).reshape(-1)
        sl = np.ascontiguousarray(src_L, dtype=np.float64)
        out = np.zeros(count, PHOTON_DTYPE)
        ss = np.zeros(count, np.uint32)
        if problem is None:
            self.lib.orc_emit(C.byref(self.g), ptr(sp, _dp), ptr(sl, _dp), len(sl),
                              npackets, first, count, seed, iteration, sigma,
                              out.ctypes.data, ptr(ss, _u32p))
        else:
            self.lib.orc_emit_problem(C.byref(self.g), ptr(sp, _dp), ptr(sl, _dp),
                                      len(sl), npackets, first, count, seed,
                                      iteration, sigma, int(problem),
                                      out.ctypes.data, ptr(ss, _u32p))
        return out, ss

    def iteration(self, src_pos, src_L, npackets, seed, iteration,
                  sigma=6.3e-22, alpha=2.7e-19, form=0, nthreads=1):
        sp = np.ascontiguousarray(src_pos, dtype=np.float64).reshape(-1)
        sl = np.ascontiguousarray(src_L, dtype=np.float64)
        l1 = C.c_double(0.0)
        nc = self.lib.orc_iteration(C.byref(self.g), ptr(sp, _dp),
                                    ptr(sl, _dp), len(sl), npackets, seed,
                                    iteration, sigma, alpha, form, nthreads,
                                    C.byref(l1))
        return nc, l1.value


def reemit_draws(lib, ks, r, probability, seed, iteration):
    """orc_reemit_draw for packet numbers ks at absorption r: reemitted flag,
    directions and target optical depths."""
    ks = np.asarray(ks, dtype=np.uint64)
    ok = np.zeros(len(ks), bool)
    dirs = np.zeros((len(ks), 3))
    taus = np.zeros(len(ks))
    d = np.zeros(3)
    t = C.c_double()
    for j, k in enumerate(ks):
        ok[j] = lib.orc_reemit_draw(int(k), int(r), float(probability), int(seed),
                                    int(iteration), ptr(d, _dp), C.byref(t)) == 1
        if ok[j]:
            dirs[j] = d
            taus[j] = t.value
    return ok, dirs, taus


def make_subgrid(lib, box, ncell):
    s = OrcSubgrid()
    b = np.ascontiguousarray(box, dtype=np.float64)
    nc = np.ascontiguousarray(ncell, dtype=np.int32)
    lib.orc_subgrid_init(C.byref(s), ptr(b, _dp), ptr(nc, _ip))
    return s


def ref_transport_list(ref, box, ncell, nsub, n, xn, photons, start_sub):
    b = np.ascontiguousarray(box, dtype=np.float64)
    nc = np.ascontiguousarray(ncell, dtype=np.int32)
    ns = np.ascontiguousarray(nsub, dtype=np.int32)
    nn = np.ascontiguousarray(n, dtype=np.float64)
    xx = np.ascontiguousarray(xn, dtype=np.float64)
    J = np.zeros_like(nn)
    m = len(photons)
    st = np.zeros(m, np.int32)
    cr = np.zeros(m, np.uint32)
    hp = np.zeros(m, np.uint32)
    ss = np.ascontiguousarray(start_sub, dtype=np.uint32)
    ref.ref_transport_list(ptr(b, _dp), ptr(nc, _ip), ptr(ns, _ip),
                           ptr(nn, _dp), ptr(xx, _dp), ptr(J, _dp),
                           photons.ctypes.data, ptr(ss, _u32p), m,
                           ptr(st, _i32p), ptr(cr, _u32p), ptr(hp, _u32p))
    return J, st, cr, hp


def ref_interact_list(ref, box, ncell, n, xn, photons, input_direction):
    b = np.ascontiguousarray(box, dtype=np.float64)
    nc = np.ascontiguousarray(ncell, dtype=np.int32)
    nn = np.ascontiguousarray(n, dtype=np.float64)
    xx = np.ascontiguousarray(xn, dtype=np.float64)
    J = np.zeros_like(nn)
    m = len(photons)
    ex = np.zeros(m, np.int32)
    cr = np.zeros(m, np.uint32)
    ref.ref_interact_list(ptr(b, _dp), ptr(nc, _ip), ptr(nn, _dp),
                          ptr(xx, _dp), ptr(J, _dp), photons.ctypes.data, m,
                          input_direction, ptr(ex, _i32p), ptr(cr, _u32p))
    return J, ex, cr


def ref_walk_list(ref, box, ncell, n, xn, photons, input_direction, kind):
    """The reference's propagate (kind 1) / compute_optical_depth (kind 2)."""
    b = np.ascontiguousarray(box, dtype=np.float64)
    nc = np.ascontiguousarray(ncell, dtype=np.int32)
    nn = np.ascontiguousarray(n, dtype=np.float64)
    xx = np.ascontiguousarray(xn, dtype=np.float64)
    p = np.ascontiguousarray(photons, dtype=PHOTON_DTYPE).copy()
    m = len(p)
    ex = np.zeros(m, np.int32)
    cr = np.zeros(m, np.uint32)
    ref.ref_walk_list(ptr(b, _dp), ptr(nc, _ip), ptr(nn, _dp), ptr(xx, _dp),
                      p.ctypes.data, m, input_direction, kind, ptr(ex, _i32p),
                      ptr(cr, _u32p))
    return p, ex, cr


def degenerate_packets(box, ncell, offset, rng, m_random=4):
    """Packets whose walks tie: the 26 directions along the axes, the face
    diagonals and the body diagonals, from cell centres, grid points (cell
    corners) and face centres, plus a few random interior points.  With a box
    of power-of-two size and ncell a power of two every cell face lies on an
    exactly representable coordinate, so the ties are exact.  offset (the
    entry direction's offset, -1/0/+1 per axis): the packets start on that
    side of the box and point inward along it."""
    box = np.asarray(box, dtype=np.float64)
    ncell = np.asarray(ncell)
    dx = box[3:] / ncell
    s2, s3 = 1.0 / np.sqrt(2.0), 1.0 / np.sqrt(3.0)
    dirs = []
    for a in range(3):
        for sg in (1.0, -1.0):
            d = np.zeros(3)
            d[a] = sg
            dirs.append(d)
    for a, b in ((0, 1), (0, 2), (1, 2)):
        for sa in (1.0, -1.0):
            for sb in (1.0, -1.0):
                d = np.zeros(3)
                d[a], d[b] = sa * s2, sb * s2
                dirs.append(d)
    for sx in (1.0, -1.0):
        for sy in (1.0, -1.0):
            for sz in (1.0, -1.0):
                dirs.append(np.array([sx, sy, sz]) * s3)
    starts = []
    for _ in range(3):
        c = rng.integers(0, ncell)
        starts.append(box[:3] + (c + 0.5) * dx)          # cell centre
        g = rng.integers(1, ncell)
        starts.append(box[:3] + g * dx)                  # grid point
        f = (c + 0.5) * dx
        a = rng.integers(0, 3)
        f[a] = g[a] * dx[a]
        starts.append(box[:3] + f)                       # on a cell face
    for _ in range(m_random):
        starts.append(box[:3] + rng.uniform(0, 1, 3) * box[3:])
    p = np.zeros(len(dirs) * len(starts), PHOTON_DTYPE)
    k = 0
    for s in starts:
        for d in dirs:
            pos, dd = s.copy(), d.copy()
            for a in range(3):
                if offset[a]:
                    pos[a] = box[a] + (box[3 + a] if offset[a] > 0 else 0.0)
                    dd[a] = -offset[a] * abs(dd[a])
            p[k]["pos"] = pos
            p[k]["dir"] = dd
            k += 1
    p["tau"] = rng.exponential(3.0, len(p))
    p["sigma"] = 6.3e-22
    p["weight"] = rng.uniform(0.5, 2.0, len(p))
    return p
