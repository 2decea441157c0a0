"""Write the golden fixtures of tests/golden/ with the reference build.

Runs only where the reference checkout exists (it is compiled into
oracle/_ref/libref.so by oracle/Makefile): every expected output below comes
from the reference's own DensitySubGrid / TravelDirections code; inputs are
seeded.  The fixtures are data (inputs and the reference's outputs).

    make -C oracle && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402

PC = 3.086e16


def isotropic(rng, n):
    cost = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    sint = np.sqrt(np.maximum(1 - cost * cost, 0))
    return np.stack([sint * np.cos(phi), sint * np.sin(phi), cost], axis=1)


def entry_packets(rng, n, box, ncell, in_dir, tau_scale):
    """Packets placed on the entry side named by in_dir, moving inward."""
    off = np.zeros(3, int)
    lib = O.load_oracle()
    import ctypes as C
    o = (C.c_int * 3)()
    lib.orc_direction_offset(in_dir, o)
    off[:] = list(o)
    d = isotropic(rng, n)
    for a in range(3):
        if off[a] != 0:
            d[:, a] = -off[a] * np.abs(d[:, a])
    pos = np.empty((n, 3))
    for a in range(3):
        if off[a] < 0:
            pos[:, a] = box[a]
        elif off[a] > 0:
            pos[:, a] = box[a] + box[3 + a]
        else:
            pos[:, a] = box[a] + rng.uniform(0, 1, n) * box[3 + a]
    p = np.zeros(n, O.PHOTON_DTYPE)
    p["pos"] = pos
    p["dir"] = d
    p["tau"] = -np.log(rng.uniform(0, 1, n)) * tau_scale
    p["sigma"] = 6.3e-22
    p["weight"] = 1.0
    return p


def main():
    ref = O.load_ref()
    if ref is None:
        sys.exit("oracle/_ref/libref.so missing: build it with make -C oracle")
    rng = np.random.default_rng(20240601)

    # 1. DensitySubGrid::interact on one 16^3 subgrid, three entry sides,
    #    a clumpy field.
    box = np.array([-5 * PC, -5 * PC, -5 * PC, 2.5 * PC, 2.5 * PC, 2.5 * PC])
    ncell = np.array([16, 16, 16], np.int32)
    nc = 16 ** 3
    n = 1e8 * np.exp(rng.normal(0, 1, nc))
    xn = 10 ** rng.uniform(-5, -1, nc)
    out = {"box": box, "ncell": ncell, "n": n, "xn": xn}
    for k, in_dir in enumerate([0, 22, 1]):  # INSIDE, FACE_X_N, CORNER_PPP
        p = entry_packets(rng, 400, box, ncell, in_dir, 1.0)
        if in_dir == 0:
            p["pos"] = box[:3] + rng.uniform(0, 1, (400, 3)) * box[3:]
        pin = p.copy()
        J, ex, cr = O.ref_interact_list(ref, box, ncell, n, xn, p, in_dir)
        out["in_dir_%d" % k] = np.int32(in_dir)
        out["photons_in_%d" % k] = pin
        out["photons_out_%d" % k] = p
        out["exit_%d" % k] = ex
        out["crossings_%d" % k] = cr
        out["J_%d" % k] = J
    np.savez_compressed(os.path.join(HERE, "interact_16.npz"), **out)

    # 2. Whole-grid transport: 32^3 cells in 2^3 subgrids, two sources,
    #    uniform and clumpy fields; packets with tau spanning thin to thick.
    box = np.array([-5 * PC, -5 * PC, -5 * PC, 10 * PC, 10 * PC, 10 * PC])
    ncell = np.array([32, 32, 32], np.int32)
    nsub = np.array([2, 2, 2], np.int32)
    ntot = 32 ** 3
    out = {"box": box, "ncell": ncell, "nsub": nsub}
    srcs = np.array([[0.0, 0.0, 0.0], [1.3 * PC, -2.1 * PC, 0.7 * PC]])
    lib = O.load_oracle()
    g = O.OracleGrid(lib, box, ncell, nsub)
    for k, (name, nfield, xfield) in enumerate([
            ("uniform", np.full(ntot, 1e8), np.full(ntot, 3e-4)),
            ("clumpy", 1e8 * np.exp(rng.normal(0, 1, ntot)),
             10 ** rng.uniform(-6, -2, ntot))]):
        m = 600
        p = np.zeros(m, O.PHOTON_DTYPE)
        which = rng.integers(0, 2, m)
        p["pos"] = srcs[which]
        p["dir"] = isotropic(rng, m)
        p["tau"] = -np.log(rng.uniform(0, 1, m)) * rng.choice([0.1, 1, 5], m)
        p["sigma"] = 6.3e-22
        p["weight"] = 1.0
        ss = np.array([g.source_subgrid(q) for q in p["pos"]], np.uint32)
        pin = p.copy()
        J, st, cr, hp = O.ref_transport_list(ref, box, ncell, nsub, nfield,
                                             xfield, p, ss)
        out["n_%d" % k] = nfield
        out["xn_%d" % k] = xfield
        out["photons_in_%d" % k] = pin
        out["start_sub_%d" % k] = ss
        out["photons_out_%d" % k] = p
        out["status_%d" % k] = st
        out["crossings_%d" % k] = cr
        out["hops_%d" % k] = hp
        out["J_%d" % k] = J
    np.savez_compressed(os.path.join(HERE, "transport_32.npz"), **out)

    # 3. compute_neutral_fraction (alpha = 4e-19 hard-coded in the
    #    reference), 16^3 cells with J spanning zero to very large.
    box = np.array([0.0, 0.0, 0.0, 0.3 * PC, 0.3 * PC, 0.3 * PC])
    ncell = np.array([16, 16, 16], np.int32)
    n = 1e8 * np.exp(rng.normal(0, 1, nc))
    J = 10 ** rng.uniform(20, 36, nc)
    J[rng.uniform(0, 1, nc) < 0.05] = 0.0
    L, N = 4.26e49, 1000000
    xo = np.zeros(nc)
    import ctypes as C
    ref.ref_compute_neutral_fraction(O.ptr(box, O._dp), O.ptr(ncell, O._ip),
                                     O.ptr(n, O._dp), O.ptr(J, O._dp), L, N,
                                     O.ptr(xo, O._dp))
    np.savez_compressed(os.path.join(HERE, "neutral_fraction_16.npz"),
                        box=box, ncell=ncell, n=n, J=J, L=L, N=N, xn=xo)
    print("fixtures written to", HERE)
    result_files()
    walk_variants()


def walk_variants():
    """5. DensitySubGrid::propagate and compute_optical_depth on one
    anisotropic 12x16x8 subgrid, clumpy field, three entry sides."""
    ref = O.load_ref()
    rng = np.random.default_rng(55)
    box = np.array([1.0 * PC, -2.0 * PC, 0.5 * PC, 1.5 * PC, 2.0 * PC, 1.0 * PC])
    ncell = np.array([12, 16, 8], np.int32)
    nc = int(np.prod(ncell))
    n = 1e8 * np.exp(rng.normal(0, 1, nc))
    xn = 10 ** rng.uniform(-6, -2, nc)
    out = {"box": box, "ncell": ncell, "n": n, "xn": xn}
    for k, in_dir in enumerate([0, 24, 7]):  # INSIDE, FACE_Y_N, CORNER_NNP
        p = entry_packets(rng, 300, box, ncell, in_dir, 0.5)
        if in_dir == 0:
            p["pos"] = box[:3] + rng.uniform(0, 1, (300, 3)) * box[3:]
        out["in_dir_%d" % k] = np.int32(in_dir)
        out["photons_in_%d" % k] = p.copy()
        for kind, name in ((1, "propagate"), (2, "optical_depth")):
            q, ex, cr = O.ref_walk_list(ref, box, ncell, n, xn, p, in_dir, kind)
            out["%s_out_%d" % (name, k)] = q
            out["%s_exit_%d" % (name, k)] = ex
            out["%s_crossings_%d" % (name, k)] = cr
    np.savez_compressed(os.path.join(HERE, "walk_variants_12.npz"), **out)
    print("walk-variant fixture written")


def result_files():
    """4. neutral_fractions.dat as the reference writes it
    (output_intensities through its MemoryMap), 2^3 subgrids of 4^3 cells,
    an anisotropic off-centre box."""
    ref = O.load_ref()
    rng = np.random.default_rng(44)
    import ctypes as C
    import tempfile
    box = np.array([-1.3 * PC, 0.2 * PC, -0.7 * PC, 2.1 * PC, 1.7 * PC, 2.6 * PC])
    ncell = np.array([8, 8, 8], np.int32)
    nsub = np.array([2, 2, 2], np.int32)
    nc = int(np.prod(ncell))
    n = 1e8 * np.exp(rng.normal(0, 1, nc))
    xn = rng.uniform(0, 1, nc)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "neutral_fractions.dat")
        ref.ref_output_dat.restype = C.c_size_t
        size = ref.ref_output_dat(O.ptr(box, O._dp), O.ptr(ncell, O._ip),
                                  O.ptr(nsub, O._ip), O.ptr(n, O._dp),
                                  O.ptr(xn, O._dp), os.fsencode(path))
        data = np.fromfile(path, np.uint8)
    assert data.size == size
    np.savez_compressed(os.path.join(HERE, "neutral_fractions_dat_8.npz"),
                        box=box, ncell=ncell, nsub=nsub, n=n, xn=xn, dat=data)
    print("result-file fixture written")


if __name__ == "__main__":
    if sys.argv[1:] == ["result-files"]:
        result_files()
    elif sys.argv[1:] == ["walk-variants"]:
        walk_variants()
    else:
        main()
