"""MI355X photon-packet transport and ionization engine (ctypes front end).

The product is the C-ABI library ``libcmi.so`` (include/cmi.h) built from the
hand-written gfx950 HIP kernels in ``csrc/``.  This module loads it and wraps
each entry point; it has no fallback: if the library is missing or no GPU is
present, calls raise.
"""
import ctypes as C
import os

import numpy as np

__all__ = [
    "CMIError", "Context", "PHOTON_DTYPE", "load_library", "PREC_FP64",
    "PREC_FP32", "ENGINE_WALK", "ENGINE_TILES", "INSIDE", "output_to_input", "Batch",
    "WarmStartNet", "write_neutral_fractions_dat", "write_neutral_fractions_bin",
    "direction_from_mask", "STATS_FIELDS",
]

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libcmi.so")

PREC_FP64 = 0
PREC_FP32 = 1
ENGINE_WALK = 0
ENGINE_TILES = 1
INSIDE = 0
NEIGHBOUR_OUTSIDE = 0xFFFFFFFF
NOALLOC_MAX_PACKETS = 65536  # CMI_NOALLOC_MAX_PACKETS of include/cmi.h

PHOTON_DTYPE = np.dtype([("pos", "<f8", 3), ("dir", "<f8", 3), ("tau", "<f8"),
                         ("sigma", "<f8"), ("weight", "<f8")])

_ERRORS = {
    -1: "CMI_ERR_ARG", -2: "CMI_ERR_HIP", -3: "CMI_ERR_ALLOC",
    -4: "CMI_ERR_OVERFLOW", -5: "CMI_ERR_DIRECTION", -6: "CMI_ERR_STATE",
    -7: "CMI_ERR_NODEVICE",
}


class CMIError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (_ERRORS.get(code, "?"), code, msg))
        self.code = code


class Stats(C.Structure):
    _fields_ = [("packets", C.c_uint64), ("crossings", C.c_uint64),
                ("hops", C.c_uint64), ("absorbed", C.c_uint64),
                ("escaped", C.c_uint64), ("errors", C.c_uint64),
                ("rounds", C.c_uint32), ("reserved", C.c_uint32),
                ("seconds", C.c_double), ("reemitted", C.c_uint64),
                ("walkout_crossings", C.c_uint64), ("walkout_hops", C.c_uint64)]


STATS_FIELDS = [f for f, _ in Stats._fields_ if f != "reserved"]


class Tuning(C.Structure):
    _fields_ = [("tile", C.c_int * 3), ("block_threads", C.c_int),
                ("refill_batch", C.c_int), ("chunk_min", C.c_int),
                ("walk_blocks", C.c_int), ("lds_fp32", C.c_int),
                ("finish_packets", C.c_int)]


_lib = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_u64p = C.POINTER(C.c_uint64)
_vp = C.c_void_p
_fp = C.POINTER(C.c_float)

# entry point -> (restype, argtypes); every symbol include/cmi.h declares
SIGNATURES = {
    "cmi_create": (C.c_int, [_dp, _ip, _ip, C.c_int, C.POINTER(_vp)]),
    "cmi_destroy": (None, [_vp]),
    "cmi_last_error": (C.c_char_p, [_vp]),
    "cmi_set_engine": (C.c_int, [_vp, C.c_int, C.POINTER(Tuning)]),
    "cmi_set_reemission": (C.c_int, [_vp, C.c_double]),
    "cmi_set_reemission_draws": (C.c_int, [_vp, C.c_uint64, C.c_uint32]),
    "cmi_set_fields": (C.c_int, [_vp, _dp, _dp]),
    "cmi_set_intensity": (C.c_int, [_vp, _dp]),
    "cmi_get_fields": (C.c_int, [_vp, _dp, _dp]),
    "cmi_reset_intensity": (C.c_int, [_vp]),
    "cmi_interact_batch": (C.c_int, [_vp, C.c_uint32, C.c_int, _vp, C.c_size_t,
                                      _i32p, _u32p]),
    "cmi_propagate_batch": (C.c_int, [_vp, C.c_uint32, C.c_int, _vp, C.c_size_t,
                                      _i32p, _u32p]),
    "cmi_optical_depth_batch": (C.c_int, [_vp, C.c_uint32, C.c_int, _vp, C.c_size_t,
                                      _i32p, _u32p]),
    "cmi_transport_packets": (C.c_int, [_vp, _vp, _u32p, C.c_size_t, _i32p,
                                         _u32p, _u32p, C.POINTER(Stats)]),
    "cmi_transport_packets_noalloc": (C.c_int, [_vp, _vp, _u32p, C.c_size_t,
                                                 _i32p, _u32p, _u32p,
                                                 C.POINTER(Stats)]),
    "cmi_set_sources": (C.c_int, [_vp, _dp, _dp, C.c_int]),
    "cmi_emit": (C.c_int, [_vp, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                           C.c_uint32, C.c_double, _vp, _u32p]),
    "cmi_transport": (C.c_int, [_vp, C.c_uint64, C.c_uint64, C.c_uint32,
                                C.c_double, C.POINTER(Stats)]),
    "cmi_update_ionization": (C.c_int, [_vp, C.c_double, C.c_uint64, C.c_double,
                                        C.c_int, _dp]),
    "cmi_total_luminosity": (C.c_double, [_vp]),
    "cmi_output_to_input_direction": (C.c_int, [C.c_int]),
    "cmi_direction_from_mask": (C.c_int, [C.c_int]),
    "cmi_get_neighbours": (C.c_int, [_vp, C.c_uint32, _u32p]),
    "cmi_get_sizes": (C.c_int, [_vp, _u64p, _u64p]),
    "cmi_kernel_timing": (C.c_int, [_vp, _dp, _u64p, C.c_int]),
    "cmi_set_kernel_timing": (C.c_int, [_vp, C.c_int]),
    "cmi_synchronize": (C.c_int, [_vp]),
    "cmi_batch_create": (C.c_int, [C.c_int, _dp, _ip, C.POINTER(Tuning),
                                   C.POINTER(_vp)]),
    "cmi_batch_destroy": (None, [_vp]),
    "cmi_batch_last_error": (C.c_char_p, [_vp]),
    "cmi_batch_set_fields": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "cmi_batch_get_fields": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "cmi_batch_set_sources": (C.c_int, [_vp, C.c_int, _dp, _dp, C.c_int]),
    "cmi_batch_set_tuning": (C.c_int, [_vp, C.POINTER(Tuning)]),
    "cmi_batch_iteration": (C.c_int, [_vp, _u64p, C.c_uint64, C.c_uint32,
                                      C.c_double, C.c_int, _dp, _dp, _u64p,
                                      C.POINTER(Stats)]),
    "cmi_write_neutral_fractions_dat": (C.c_int, [C.c_char_p, _dp, _ip, _ip,
                                                  _dp, _dp]),
    "cmi_write_neutral_fractions_bin": (C.c_int, [C.c_char_p, C.c_uint64,
                                                  _dp]),
    "cmi_write_results": (C.c_int, [_vp, C.c_char_p, C.c_char_p]),
    "cmi_net_create": (C.c_int, [C.c_uint64, C.POINTER(_vp)]),
    "cmi_net_destroy": (None, [_vp]),
    "cmi_net_last_error": (C.c_char_p, [_vp]),
    "cmi_net_param_count": (C.c_size_t, [_vp]),
    "cmi_net_get_params": (C.c_int, [_vp, _fp]),
    "cmi_net_set_params": (C.c_int, [_vp, _fp]),
    "cmi_net_set_normalisation": (C.c_int, [_vp, _fp, _fp, C.c_float]),
    "cmi_net_add_samples": (C.c_int, [_vp, _fp, _fp, C.c_int, _ip]),
    "cmi_net_sample_count": (C.c_int, [_vp]),
    "cmi_net_set_dilation": (C.c_int, [_vp, _ip]),
    "cmi_net_set_loss_weight": (C.c_int, [_vp, C.c_float]),
    "cmi_net_set_input_mode": (C.c_int, [_vp, C.c_int]),
    "cmi_net_set_tiled": (C.c_int, [_vp, C.c_int]),
    "cmi_net_get_input_mode": (C.c_int, [_vp]),
    "cmi_net_set_start_rule": (C.c_int, [_vp, C.c_float, C.c_double]),
    "cmi_net_get_start_rule": (C.c_int, [_vp, C.POINTER(C.c_float),
                                         C.POINTER(C.c_double)]),
    "cmi_net_get_dilation": (C.c_int, [_vp, _ip]),
    "cmi_net_train": (C.c_int, [_vp, C.c_int, C.c_int, _ip, C.c_float,
                                C.c_uint64, _fp]),
    "cmi_net_loss_and_grad": (C.c_int, [_vp, _fp, _fp, C.c_int, _ip, _fp,
                                        _fp]),
    "cmi_net_predict": (C.c_int, [_vp, _fp, _ip, _fp]),
    "cmi_net_warm_start": (C.c_int, [_vp, _vp, C.c_double, C.c_uint64]),
    "cmi_net_context_features": (C.c_int, [_vp, _vp, C.c_double, C.c_uint64,
                                           _fp]),
    "cmi_restrict_density": (C.c_int, [_vp, _vp]),
    "cmi_prolong_neutral_fraction": (C.c_int, [_vp, _vp]),
}


def load_library(path=LIB_PATH):
    """Load libcmi.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise CMIError(-6, "libcmi.so not built: run "
                           "python -m parallelcmacionize_amd.build")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib


def _p(a, t):
    return a.ctypes.data_as(t)


def output_to_input(d):
    """TravelDirections::output_to_input_direction (TravelDirections.hpp:89)."""
    return load_library().cmi_output_to_input_direction(int(d))


def direction_from_mask(mask):
    """TravelDirections::get_output_direction(mask) (TravelDirections.hpp:315)."""
    return load_library().cmi_direction_from_mask(int(mask))


class Context:
    """One grid of subgrids on one GPU (owner of a cmi_ctx)."""

    def __init__(self, box, ncell, nsub, precision=PREC_FP64):
        self.lib = load_library()
        b = np.ascontiguousarray(box, dtype=np.float64)
        nc = np.ascontiguousarray(ncell, dtype=np.int32)
        ns = np.ascontiguousarray(nsub, dtype=np.int32)
        h = _vp()
        rc = self.lib.cmi_create(_p(b, _dp), _p(nc, _ip), _p(ns, _ip),
                                 int(precision), C.byref(h))
        if rc != 0:
            raise CMIError(rc, "cmi_create failed")
        self.h = h
        self.box = b
        self.ncell = nc
        self.nsub = ns
        self.precision = precision
        nsg = C.c_uint64()
        ncl = C.c_uint64()
        self.lib.cmi_get_sizes(h, C.byref(nsg), C.byref(ncl))
        self.nsubgrid = nsg.value
        self.ncell_total = ncl.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.cmi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise CMIError(rc, self.lib.cmi_last_error(self.h).decode())

    def set_engine(self, engine, tile=None, block_threads=0, refill_batch=0,
                   chunk_min=0, walk_blocks=0, lds_fp32=0, finish_packets=0):
        t = Tuning()
        if tile is not None:
            for a in range(3):
                t.tile[a] = int(tile[a])
        t.block_threads = int(block_threads)
        t.refill_batch = int(refill_batch)
        t.chunk_min = int(chunk_min)
        t.walk_blocks = int(walk_blocks)
        t.lds_fp32 = int(lds_fp32)
        t.finish_packets = int(finish_packets)
        self._check(self.lib.cmi_set_engine(self.h, int(engine), C.byref(t)))

    def set_reemission(self, probability, seed=None, iteration=0):
        """Reemission probability; seed/iteration (if given) key the draws of
        explicit-list transports on the walk engine (packet i of a list draws
        as packet number i of transport(seed, iteration))."""
        self._check(self.lib.cmi_set_reemission(self.h, float(probability)))
        if seed is not None:
            self._check(self.lib.cmi_set_reemission_draws(self.h, int(seed),
                                                          int(iteration)))

    def set_fields(self, nH=None, xn=None):
        a = None if nH is None else np.ascontiguousarray(nH, dtype=np.float64)
        b = None if xn is None else np.ascontiguousarray(xn, dtype=np.float64)
        for v in (a, b):
            if v is not None and v.size != self.ncell_total:
                raise ValueError("field size %d != %d" % (v.size, self.ncell_total))
        self._check(self.lib.cmi_set_fields(
            self.h, None if a is None else _p(a, _dp),
            None if b is None else _p(b, _dp)))

    def set_intensity(self, J):
        a = np.ascontiguousarray(J, dtype=np.float64)
        if a.size != self.ncell_total:
            raise ValueError("field size")
        self._check(self.lib.cmi_set_intensity(self.h, _p(a, _dp)))

    def reset_intensity(self):
        self._check(self.lib.cmi_reset_intensity(self.h))

    def get_fields(self):
        J = np.zeros(self.ncell_total)
        x = np.zeros(self.ncell_total)
        self._check(self.lib.cmi_get_fields(self.h, _p(J, _dp), _p(x, _dp)))
        return J, x

    def interact_batch(self, subgrid, entry_dir, photons, kind="interact"):
        """One subgrid, explicit packets: kind 'interact' (deposits),
        'propagate' (no deposits) or 'optical_depth' (no stop, tau += path
        optical depth).  Returns the updated packets, exit directions and
        crossings."""
        fn = {"interact": self.lib.cmi_interact_batch,
              "propagate": self.lib.cmi_propagate_batch,
              "optical_depth": self.lib.cmi_optical_depth_batch}[kind]
        p = np.ascontiguousarray(photons, dtype=PHOTON_DTYPE).copy()
        n = len(p)
        ex = np.zeros(n, np.int32)
        cr = np.zeros(n, np.uint32)
        self._check(fn(self.h, int(subgrid), int(entry_dir), p.ctypes.data, n,
                       _p(ex, _i32p), _p(cr, _u32p)))
        return p, ex, cr

    def transport_packets(self, photons, start_sub, noalloc=False):
        """Whole-grid transport of an explicit packet list.  noalloc=True
        calls cmi_transport_packets_noalloc (walk engine, staging through
        the library's fixed device array, at most NOALLOC_MAX_PACKETS)."""
        p = np.ascontiguousarray(photons, dtype=PHOTON_DTYPE).copy()
        ss = np.ascontiguousarray(start_sub, dtype=np.uint32)
        n = len(p)
        st = np.zeros(n, np.int32)
        cr = np.zeros(n, np.uint32)
        hp = np.zeros(n, np.uint32)
        s = Stats()
        fn = (self.lib.cmi_transport_packets_noalloc if noalloc
              else self.lib.cmi_transport_packets)
        self._check(fn(self.h, p.ctypes.data, _p(ss, _u32p), n, _p(st, _i32p),
                       _p(cr, _u32p), _p(hp, _u32p), C.byref(s)))
        return p, st, cr, hp, _stats_dict(s)

    def set_sources(self, pos, L):
        p = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        l = np.ascontiguousarray(L, dtype=np.float64).reshape(-1)
        self._check(self.lib.cmi_set_sources(self.h, _p(p, _dp), _p(l, _dp),
                                             len(l)))

    def total_luminosity(self):
        return self.lib.cmi_total_luminosity(self.h)

    def emit(self, npackets, seed, iteration, sigma=6.3e-22, first=0,
             count=None):
        if count is None:
            count = npackets - first
        out = np.zeros(count, PHOTON_DTYPE)
        ss = np.zeros(count, np.uint32)
        self._check(self.lib.cmi_emit(self.h, npackets, first, count, seed,
                                      iteration, sigma, out.ctypes.data,
                                      _p(ss, _u32p)))
        return out, ss

    def transport(self, npackets, seed, iteration, sigma=6.3e-22):
        s = Stats()
        self._check(self.lib.cmi_transport(self.h, int(npackets), int(seed),
                                           int(iteration), float(sigma),
                                           C.byref(s)))
        return _stats_dict(s)

    def write_results(self, dat_path=None, bin_path=None):
        """neutral_fractions.dat / .bin of the current fields."""
        self._check(self.lib.cmi_write_results(
            self.h, None if dat_path is None else os.fsencode(dat_path),
            None if bin_path is None else os.fsencode(bin_path)))

    def update_ionization(self, luminosity, npackets, alpha, form=0):
        l1 = C.c_double()
        self._check(self.lib.cmi_update_ionization(self.h, float(luminosity),
                                                   int(npackets), float(alpha),
                                                   int(form), C.byref(l1)))
        return l1.value

    def neighbours(self, subgrid):
        out = np.zeros(27, np.uint32)
        self._check(self.lib.cmi_get_neighbours(self.h, int(subgrid),
                                                _p(out, _u32p)))
        return out

    def kernel_timing(self, reset=False):
        avg = C.c_double()
        calls = C.c_uint64()
        self._check(self.lib.cmi_kernel_timing(self.h, C.byref(avg),
                                               C.byref(calls), int(reset)))
        return avg.value, calls.value

    def set_kernel_timing(self, on):
        """Per-launch event timing of the tiles engine's process kernel
        (default on; off saves two stream events per round)."""
        self._check(self.lib.cmi_set_kernel_timing(self.h, int(bool(on))))

    def synchronize(self):
        self._check(self.lib.cmi_synchronize(self.h))

    def restrict_density_to(self, coarse):
        """coarse (same box, ncell an integer factor smaller) gets this
        context's n_H averaged over each block, x_n = 1e-6 and J = 0."""
        rc = self.lib.cmi_restrict_density(self.h, coarse.h)
        if rc != 0:
            raise CMIError(rc, self.lib.cmi_last_error(coarse.h).decode())

    def prolong_neutral_fraction_from(self, coarse):
        """This context's x_n from a coarse context's (trilinear in the
        clamped logit), J = 0."""
        self._check(self.lib.cmi_prolong_neutral_fraction(coarse.h, self.h))


def write_neutral_fractions_dat(path, box, ncell, nsub, nH, xn):
    """neutral_fractions.dat of the reference (host arrays, reference
    layout)."""
    lib = load_library()
    b = np.ascontiguousarray(box, dtype=np.float64)
    nc = np.ascontiguousarray(ncell, dtype=np.int32)
    ns = np.ascontiguousarray(nsub, dtype=np.int32)
    n = np.ascontiguousarray(nH, dtype=np.float64).reshape(-1)
    x = np.ascontiguousarray(xn, dtype=np.float64).reshape(-1)
    if n.size != int(np.prod(nc)) or x.size != n.size:
        raise ValueError("field size")
    rc = lib.cmi_write_neutral_fractions_dat(os.fsencode(path), _p(b, _dp),
                                             _p(nc, _ip), _p(ns, _ip),
                                             _p(n, _dp), _p(x, _dp))
    if rc != 0:
        raise CMIError(rc, "cannot write %s" % path)


def write_neutral_fractions_bin(path, xn):
    """neutral_fractions.bin of the reference (x_n, reference layout)."""
    lib = load_library()
    x = np.ascontiguousarray(xn, dtype=np.float64).reshape(-1)
    rc = lib.cmi_write_neutral_fractions_bin(os.fsencode(path), x.size,
                                             _p(x, _dp))
    if rc != 0:
        raise CMIError(rc, "cannot write %s" % path)


class Batch:
    """A problem set: many independent grids transported together
    (cmi_batch_*).  Cell arrays of a problem are (nx, ny, nz) lattices."""

    def __init__(self, boxes, ncells, tile=16, block_threads=0, refill_batch=0,
                 chunk=0):
        self.lib = load_library()
        b = np.ascontiguousarray(boxes, dtype=np.float64).reshape(-1, 6)
        nc = np.ascontiguousarray(ncells, dtype=np.int32).reshape(-1, 3)
        if len(b) != len(nc):
            raise ValueError("one box per problem")
        t = Tuning()
        t.tile[0] = t.tile[1] = t.tile[2] = int(tile)
        t.block_threads = int(block_threads)
        t.refill_batch = int(refill_batch)
        t.chunk_min = int(chunk)
        h = _vp()
        rc = self.lib.cmi_batch_create(len(b), _p(b, _dp), _p(nc, _ip),
                                       C.byref(t), C.byref(h))
        if rc != 0:
            raise CMIError(rc, "cmi_batch_create failed")
        self.h = h
        self.boxes = b
        self.ncells = nc
        self.nprob = len(b)

    def close(self):
        if getattr(self, "h", None):
            self.lib.cmi_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise CMIError(rc, self.lib.cmi_batch_last_error(self.h).decode())

    def set_fields(self, p, nH=None, xn=None):
        a = None if nH is None else np.ascontiguousarray(nH, dtype=np.float64)
        b = None if xn is None else np.ascontiguousarray(xn, dtype=np.float64)
        self._check(self.lib.cmi_batch_set_fields(
            self.h, int(p), None if a is None else _p(a, _dp),
            None if b is None else _p(b, _dp)))

    def get_fields(self, p):
        n = int(np.prod(self.ncells[p]))
        J = np.zeros(n)
        x = np.zeros(n)
        self._check(self.lib.cmi_batch_get_fields(self.h, int(p), _p(J, _dp),
                                                  _p(x, _dp)))
        shape = tuple(int(v) for v in self.ncells[p])
        return J.reshape(shape), x.reshape(shape)

    def set_sources(self, p, pos, L):
        a = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        l = np.ascontiguousarray(L, dtype=np.float64).reshape(-1)
        self._check(self.lib.cmi_batch_set_sources(self.h, int(p), _p(a, _dp),
                                                   _p(l, _dp), len(l)))

    def set_tuning(self, block_threads=0, refill_batch=0, chunk=0,
                   lds_fp32=0):
        t = Tuning()
        t.block_threads = int(block_threads)
        t.refill_batch = int(refill_batch)
        t.chunk_min = int(chunk)
        t.lds_fp32 = int(lds_fp32)
        self._check(self.lib.cmi_batch_set_tuning(self.h, C.byref(t)))

    def iteration(self, npackets, seed, iteration, alpha, form=1):
        n = np.ascontiguousarray(npackets, dtype=np.uint64)
        l1 = np.zeros(self.nprob)
        secs = np.zeros(self.nprob)
        cross = np.zeros(self.nprob, np.uint64)
        s = Stats()
        self._check(self.lib.cmi_batch_iteration(
            self.h, _p(n, _u64p), int(seed), int(iteration), float(alpha),
            int(form), _p(l1, _dp), _p(secs, _dp), _p(cross, _u64p),
            C.byref(s)))
        return {"l1": l1, "problem_seconds": secs, "problem_crossings": cross,
                **_stats_dict(s)}


class WarmStartNet:
    """The learned warm start (cmi_net_*): 4 conv layers, 32 channels, BF16
    MFMA with FP32 accumulation, trained on the GPU."""

    def __init__(self, seed=1):
        self.lib = load_library()
        h = _vp()
        rc = self.lib.cmi_net_create(int(seed), C.byref(h))
        if rc != 0:
            raise CMIError(rc, "cmi_net_create failed")
        self.h = h
        self.nparam = self.lib.cmi_net_param_count(h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.cmi_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise CMIError(rc, self.lib.cmi_net_last_error(self.h).decode())

    def params(self):
        out = np.zeros(self.nparam, np.float32)
        self._check(self.lib.cmi_net_get_params(self.h, _p(out, _fp)))
        return out

    def set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        assert a.size == self.nparam
        self._check(self.lib.cmi_net_set_params(self.h, _p(a, _fp)))

    def set_normalisation(self, mean, scale, logit_scale=8.0):
        m = np.ascontiguousarray(mean, dtype=np.float32)
        s = np.ascontiguousarray(scale, dtype=np.float32)
        self._check(self.lib.cmi_net_set_normalisation(
            self.h, _p(m, _fp), _p(s, _fp), float(logit_scale)))

    def add_samples(self, features, targets):
        f = np.ascontiguousarray(features, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        dim = np.array(t.shape[1:], np.int32)
        self._check(self.lib.cmi_net_add_samples(self.h, _p(f, _fp), _p(t, _fp),
                                                 t.shape[0], _p(dim, _ip)))

    def sample_count(self):
        return self.lib.cmi_net_sample_count(self.h)

    def set_dilation(self, dilation):
        d = np.ascontiguousarray(dilation, dtype=np.int32)
        assert d.size == 4
        self._check(self.lib.cmi_net_set_dilation(self.h, _p(d, _ip)))

    def set_input_mode(self, mode):
        self._check(self.lib.cmi_net_set_input_mode(self.h, int(mode)))

    def set_tiled(self, on):
        """Forward convolutions through the LDS-tiled kernel where the
        volume divides into its blocks (default), or always through the
        per-run kernel; the outputs are identical."""
        self._check(self.lib.cmi_net_set_tiled(self.h, int(bool(on))))

    def input_mode(self):
        return self.lib.cmi_net_get_input_mode(self.h)

    def set_loss_weight(self, w_too_neutral):
        self._check(self.lib.cmi_net_set_loss_weight(self.h, float(w_too_neutral)))

    def set_start_rule(self, logit_cut, fallback_xn=1e-6):
        self._check(self.lib.cmi_net_set_start_rule(self.h, float(logit_cut),
                                                    float(fallback_xn)))

    def start_rule(self):
        c = C.c_float()
        f = C.c_double()
        self._check(self.lib.cmi_net_get_start_rule(self.h, C.byref(c),
                                                    C.byref(f)))
        return c.value, f.value

    def dilation(self):
        d = np.zeros(4, np.int32)
        self._check(self.lib.cmi_net_get_dilation(self.h, _p(d, _ip)))
        return d

    def train(self, steps, batch, crop, lr, seed=1):
        losses = np.zeros(int(steps), np.float32)
        c = np.ascontiguousarray(crop, dtype=np.int32)
        self._check(self.lib.cmi_net_train(self.h, int(steps), int(batch),
                                           _p(c, _ip), float(lr), int(seed),
                                           _p(losses, _fp)))
        return losses

    def loss_and_grad(self, features, targets):
        f = np.ascontiguousarray(features, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        dim = np.array(t.shape[1:], np.int32)
        loss = np.zeros(1, np.float32)
        g = np.zeros(self.nparam, np.float32)
        self._check(self.lib.cmi_net_loss_and_grad(
            self.h, _p(f, _fp), _p(t, _fp), t.shape[0], _p(dim, _ip),
            _p(loss, _fp), _p(g, _fp)))
        return float(loss[0]), g

    def predict(self, features):
        f = np.ascontiguousarray(features, dtype=np.float32)
        dim = np.array(f.shape[:3], np.int32)
        out = np.zeros(f.shape[:3], np.float32)
        self._check(self.lib.cmi_net_predict(self.h, _p(f, _fp), _p(dim, _ip),
                                             _p(out, _fp)))
        return out

    def warm_start(self, ctx, luminosity, npackets):
        """Set ctx's x_n from the network (cmi_net_warm_start): input mode 0
        reads ctx's J from a first transport of npackets packets, mode 1
        ctx's x_n after a few cheap iterations of npackets packets."""
        self._check(self.lib.cmi_net_warm_start(self.h, ctx.h, float(luminosity),
                                                int(npackets)))

    def context_features(self, ctx, luminosity, npackets):
        """The network's input features of ctx as a (nx, ny, nz, 2) float32
        lattice (cmi_net_context_features): (log10 n_H, log10(jH + 1e-12))
        of ctx's J in input mode 0, (log10 n_H, logit x_n) in mode 1."""
        shape = tuple(int(v) for v in ctx.ncell) + (2,)
        out = np.zeros(shape, np.float32)
        self._check(self.lib.cmi_net_context_features(
            self.h, ctx.h, float(luminosity), int(npackets), _p(out, _fp)))
        return out


def _stats_dict(s):
    return {f: getattr(s, f) for f in STATS_FIELDS}
