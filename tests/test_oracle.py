"""Pins of the CPU oracle (oracle/cmi_oracle.c) before it is trusted as the
checker of the HIP path: the reference's own known-answer test, golden
fixtures written by the reference build, bit-for-bit agreement with the
reference build where it is present, closed forms and invariants."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

PC = 3.086e16

D = dict(INSIDE=0, CORNER_PPP=1, CORNER_PPN=2, CORNER_PNP=3, CORNER_PNN=4,
         CORNER_NPP=5, CORNER_NPN=6, CORNER_NNP=7, CORNER_NNN=8,
         EDGE_X_PP=9, EDGE_X_PN=10, EDGE_X_NP=11, EDGE_X_NN=12,
         EDGE_Y_PP=13, EDGE_Y_PN=14, EDGE_Y_NP=15, EDGE_Y_NN=16,
         EDGE_Z_PP=17, EDGE_Z_PN=18, EDGE_Z_NP=19, EDGE_Z_NN=20,
         FACE_X_P=21, FACE_X_N=22, FACE_Y_P=23, FACE_Y_N=24, FACE_Z_P=25,
         FACE_Z_N=26)

# Known-answer vectors of testDensitySubGridDirections.cpp:105-134
# (get_output_direction on a 10^3 grid): (ix, iy, iz) -> direction.
OUTPUT_CASES = [
    ((4, 5, 3), "INSIDE"),
    ((-1, 5, 3), "FACE_X_N"), ((10, 5, 3), "FACE_X_P"),
    ((4, -1, 3), "FACE_Y_N"), ((4, 10, 3), "FACE_Y_P"),
    ((4, 5, -1), "FACE_Z_N"), ((4, 5, 10), "FACE_Z_P"),
    ((-1, -1, 3), "EDGE_Z_NN"), ((-1, 10, 3), "EDGE_Z_NP"),
    ((10, -1, 3), "EDGE_Z_PN"), ((10, 10, 3), "EDGE_Z_PP"),
    ((-1, 5, -1), "EDGE_Y_NN"), ((-1, 5, 10), "EDGE_Y_NP"),
    ((10, 5, -1), "EDGE_Y_PN"), ((10, 5, 10), "EDGE_Y_PP"),
    ((4, -1, -1), "EDGE_X_NN"), ((4, -1, 10), "EDGE_X_NP"),
    ((4, 10, -1), "EDGE_X_PN"), ((4, 10, 10), "EDGE_X_PP"),
    ((-1, -1, -1), "CORNER_NNN"), ((-1, -1, 10), "CORNER_NNP"),
    ((-1, 10, -1), "CORNER_NPN"), ((-1, 10, 10), "CORNER_NPP"),
    ((10, -1, -1), "CORNER_PNN"), ((10, -1, 10), "CORNER_PNP"),
    ((10, 10, -1), "CORNER_PPN"), ((10, 10, 10), "CORNER_PPP"),
]

# Known-answer vectors of testDensitySubGridDirections.cpp:138-167
# (get_start_index on the unit box with 10^3 cells):
# direction, position -> three index.
INPUT_CASES = [
    ("INSIDE", (0.5, 0.5, 0.5), (5, 5, 5)),
    ("FACE_X_N", (0., 0.5, 0.5), (0, 5, 5)),
    ("FACE_X_P", (1., 0.5, 0.5), (9, 5, 5)),
    ("FACE_Y_N", (0.5, 0., 0.5), (5, 0, 5)),
    ("FACE_Y_P", (0.5, 1., 0.5), (5, 9, 5)),
    ("FACE_Z_N", (0.5, 0.5, 0.), (5, 5, 0)),
    ("FACE_Z_P", (0.5, 0.5, 1.), (5, 5, 9)),
    ("EDGE_X_NN", (0.5, 0., 0.), (5, 0, 0)),
    ("EDGE_X_NP", (0.5, 0., 1.), (5, 0, 9)),
    ("EDGE_X_PN", (0.5, 1., 0.), (5, 9, 0)),
    ("EDGE_X_PP", (0.5, 1., 1.), (5, 9, 9)),
    ("EDGE_Y_NN", (0., 0.5, 0.), (0, 5, 0)),
    ("EDGE_Y_NP", (0., 0.5, 1.), (0, 5, 9)),
    ("EDGE_Y_PN", (1., 0.5, 0.), (9, 5, 0)),
    ("EDGE_Y_PP", (1., 0.5, 1.), (9, 5, 9)),
    ("EDGE_Z_NN", (0., 0., 0.5), (0, 0, 5)),
    ("EDGE_Z_NP", (0., 1., 0.5), (0, 9, 5)),
    ("EDGE_Z_PN", (1., 0., 0.5), (9, 0, 5)),
    ("EDGE_Z_PP", (1., 1., 0.5), (9, 9, 5)),
    ("CORNER_NNN", (0., 0., 0.), (0, 0, 0)),
    ("CORNER_NNP", (0., 0., 1.), (0, 0, 9)),
    ("CORNER_NPN", (0., 1., 0.), (0, 9, 0)),
    ("CORNER_NPP", (0., 1., 1.), (0, 9, 9)),
    ("CORNER_PNN", (1., 0., 0.), (9, 0, 0)),
    ("CORNER_PNP", (1., 0., 1.), (9, 0, 9)),
    ("CORNER_PPN", (1., 1., 0.), (9, 9, 0)),
    ("CORNER_PPP", (1., 1., 1.), (9, 9, 9)),
]


def _unit_subgrid(lib):
    return O.make_subgrid(lib, [0., 0., 0., 1., 1., 1.], [10, 10, 10])


@pytest.mark.parametrize("idx,name", OUTPUT_CASES)
def test_output_direction_known_answers(oracle, idx, name):
    s = _unit_subgrid(oracle)
    i = np.array(idx, np.int32)
    assert oracle.orc_get_output_direction(C.byref(s), O.ptr(i, O._ip)) == D[name]


@pytest.mark.parametrize("name,pos,expect", INPUT_CASES)
def test_start_index_known_answers(oracle, name, pos, expect):
    s = _unit_subgrid(oracle)
    p = np.array(pos, np.float64)
    out = np.zeros(3, np.int32)
    oracle.orc_get_start_index(C.byref(s), O.ptr(p, O._dp), D[name],
                               O.ptr(out, O._ip))
    assert tuple(out) == expect


def test_output_to_input_is_involution_and_swaps_sides(oracle):
    off = np.zeros(3, np.int32)
    off2 = np.zeros(3, np.int32)
    for d in range(27):
        e = oracle.orc_output_to_input_direction(d)
        assert oracle.orc_output_to_input_direction(e) == d
        oracle.orc_direction_offset(d, O.ptr(off, O._ip))
        oracle.orc_direction_offset(e, O.ptr(off2, O._ip))
        assert (off2 == -off).all()
    # the reference names: FACE_X_P <-> FACE_X_N, CORNER_PNP <-> CORNER_NPN,
    # EDGE_Y_PN <-> EDGE_Y_NP (TravelDirections.hpp:89-152)
    assert oracle.orc_output_to_input_direction(D["FACE_X_P"]) == D["FACE_X_N"]
    assert oracle.orc_output_to_input_direction(D["CORNER_PNP"]) == D["CORNER_NPN"]
    assert oracle.orc_output_to_input_direction(D["EDGE_Y_PN"]) == D["EDGE_Y_NP"]
    assert oracle.orc_output_to_input_direction(0) == 0


def test_mask_table_against_reference(oracle, ref):
    for mask in range(64):
        assert oracle.orc_direction_from_mask(mask) == \
            ref.ref_direction_from_mask(mask), mask
    for d in range(27):
        assert oracle.orc_output_to_input_direction(d) == \
            ref.ref_output_to_input_direction(d)


def test_compatibility_against_reference(oracle, ref):
    rng = np.random.default_rng(5)
    dirs = rng.choice([-0.5, 0.0, 0.5], size=(200, 3))
    for v in dirs:
        v = np.ascontiguousarray(v)
        for d in range(27):
            assert oracle.orc_is_compatible_output_direction(O.ptr(v, O._dp), d) == \
                ref.ref_is_compatible_output_direction(O.ptr(v, O._dp), d)
            assert oracle.orc_is_compatible_input_direction(O.ptr(v, O._dp), d) == \
                ref.ref_is_compatible_input_direction(O.ptr(v, O._dp), d)


def test_start_index_against_reference(oracle, ref):
    rng = np.random.default_rng(6)
    box = np.array([-1.7, 0.3, 2.0, 1.3, 2.9, 0.7])
    nc = np.array([7, 16, 5], np.int32)
    s = O.make_subgrid(oracle, box, nc)
    a = np.zeros(3, np.int32)
    b = np.zeros(3, np.int32)
    for _ in range(3000):
        d = int(rng.integers(0, 27))
        pos = rng.uniform(-0.05, 1.05, 3) * box[3:]
        ia = oracle.orc_get_start_index(C.byref(s), O.ptr(pos, O._dp), d,
                                        O.ptr(a, O._ip))
        ib = ref.ref_get_start_index(O.ptr(box, O._dp), O.ptr(nc, O._ip),
                                     O.ptr(pos, O._dp), d, O.ptr(b, O._ip))
        assert ia == ib and (a == b).all()


def _interact_oracle(oracle, box, ncell, n, xn, photons, in_dir):
    s = O.make_subgrid(oracle, box, ncell)
    nn = np.ascontiguousarray(n, dtype=np.float64).copy()
    xx = np.ascontiguousarray(xn, dtype=np.float64).copy()
    J = np.zeros_like(nn)
    s.n = O.ptr(nn, O._dp)
    s.xn = O.ptr(xx, O._dp)
    s.J = O.ptr(J, O._dp)
    p = photons.copy()
    ex = np.zeros(len(p), np.int32)
    cr = np.zeros(len(p), np.uint32)
    one = np.zeros(1, np.uint32)
    for i in range(len(p)):
        one[0] = 0
        ex[i] = oracle.orc_interact(C.byref(s), p[i:i + 1].ctypes.data, in_dir,
                                    O.ptr(one, O._u32p))
        cr[i] = one[0]
    return p, ex, cr, J


def test_interact_matches_golden(oracle, golden):
    g = golden("interact_16.npz")
    for k in range(3):
        p, ex, cr, J = _interact_oracle(oracle, g["box"], g["ncell"], g["n"],
                                        g["xn"], g["photons_in_%d" % k],
                                        int(g["in_dir_%d" % k]))
        assert (ex == g["exit_%d" % k]).all()
        assert (cr == g["crossings_%d" % k]).all()
        assert (p.view(np.uint8) == g["photons_out_%d" % k].view(np.uint8)).all()
        assert (J == g["J_%d" % k]).all()


def test_transport_matches_golden(oracle, golden):
    g = golden("transport_32.npz")
    for k in range(2):
        grid = O.OracleGrid(oracle, g["box"], g["ncell"], g["nsub"])
        grid.n[:] = g["n_%d" % k]
        grid.xn[:] = g["xn_%d" % k]
        grid.J[:] = 0
        p = g["photons_in_%d" % k].copy()
        st, cr, hp = grid.transport(p, g["start_sub_%d" % k])
        assert (st == g["status_%d" % k]).all()
        assert (cr == g["crossings_%d" % k]).all()
        assert (hp == g["hops_%d" % k]).all()
        assert (p.view(np.uint8) == g["photons_out_%d" % k].view(np.uint8)).all()
        assert (grid.J == g["J_%d" % k]).all()


def test_neutral_fraction_matches_golden(oracle, golden):
    g = golden("neutral_fraction_16.npz")
    grid = O.OracleGrid(oracle, g["box"], g["ncell"], [1, 1, 1])
    grid.n[:] = g["n"]
    grid.J[:] = g["J"]
    grid.compute_neutral_fraction(float(g["L"]), int(g["N"]), 4e-19, form=0)
    assert (grid.xn == g["xn"]).all()
    assert (grid.J == 0).all()


def test_interact_against_live_reference(oracle, ref):
    rng = np.random.default_rng(11)
    box = np.array([1.0 * PC, -2.0 * PC, 0.5 * PC, 1.1 * PC, 0.9 * PC, 1.3 * PC])
    ncell = np.array([8, 12, 16], np.int32)
    nc = int(np.prod(ncell))
    n = 1e8 * np.exp(rng.normal(0, 1.5, nc))
    xn = 10 ** rng.uniform(-7, 0, nc)
    for in_dir in (0, 3, 14, 25):
        m = 300
        p = np.zeros(m, O.PHOTON_DTYPE)
        p["pos"] = box[:3] + rng.uniform(0, 1, (m, 3)) * box[3:]
        d = rng.normal(size=(m, 3))
        p["dir"] = d / np.linalg.norm(d, axis=1)[:, None]
        off = np.zeros(3, np.int32)
        oracle.orc_direction_offset(in_dir, O.ptr(off, O._ip))
        for a in range(3):
            if off[a]:
                p["pos"][:, a] = box[a] + (box[3 + a] if off[a] > 0 else 0.)
                p["dir"][:, a] = -off[a] * np.abs(p["dir"][:, a])
        p["tau"] = -np.log(rng.uniform(0, 1, m)) * 3
        p["sigma"] = 6.3e-22
        p["weight"] = rng.uniform(0.5, 2, m)
        q = p.copy()
        po, ex, cr, J = _interact_oracle(oracle, box, ncell, n, xn, p, in_dir)
        Jr, exr, crr = O.ref_interact_list(ref, box, ncell, n, xn, q, in_dir)
        assert (ex == exr).all() and (cr == crr).all()
        assert (po.view(np.uint8) == q.view(np.uint8)).all()
        assert (J == Jr).all()


DEGENERATE_BOXES = [
    # power-of-two sizes and cell counts: every cell face is exactly
    # representable, so diagonal walks tie exactly at cell edges and corners
    (np.array([-2.0 ** 54, 0.0, 2.0 ** 52, 2.0 ** 55, 2.0 ** 55, 2.0 ** 55]), [16, 16, 16]),
    (np.array([2.0 ** 53, -2.0 ** 55, 0.0, 2.0 ** 54, 2.0 ** 55, 2.0 ** 53]), [8, 32, 4]),
]
DEGENERATE_IN_DIRS = (0, 1, 8, 11, 14, 19, 21, 26)


def degenerate_case(oracle, k, in_dir, thick=False):
    box, ncell = DEGENERATE_BOXES[k]
    rng = np.random.default_rng(1000 * k + 100 * thick + in_dir)
    nc = int(np.prod(ncell))
    n = 1e8 * np.exp(rng.normal(0, 1.0, nc))
    # thin: the diagonal walks reach the box's edges and corners; thick: most
    # packets are absorbed on the way (partial last steps on tied walks)
    xn = 10 ** (rng.uniform(-7, -1, nc) if thick else rng.uniform(-9, -5, nc))
    off = np.zeros(3, np.int32)
    oracle.orc_direction_offset(in_dir, O.ptr(off, O._ip))
    p = O.degenerate_packets(box, ncell, off, rng)
    return box, np.array(ncell, np.int32), n, xn, p


@pytest.mark.parametrize("thick", [False, True])
@pytest.mark.parametrize("k", [0, 1])
def test_degenerate_packets_against_live_reference(oracle, ref, k, thick):
    """Walks along axes, face and body diagonals from cell centres, grid
    points and faces, entering through faces, edges and corners: the
    reference's tie-breaks and face snapping, bit for bit."""
    for in_dir in DEGENERATE_IN_DIRS:
        box, ncell, n, xn, p = degenerate_case(oracle, k, in_dir, thick)
        q = p.copy()
        po, ex, cr, J = _interact_oracle(oracle, box, ncell, n, xn, p, in_dir)
        Jr, exr, crr = O.ref_interact_list(ref, box, ncell, n, xn, q, in_dir)
        assert (ex == exr).all(), in_dir
        assert (cr == crr).all(), in_dir
        assert (po.view(np.uint8) == q.view(np.uint8)).all(), in_dir
        assert (J == Jr).all(), in_dir


def test_neutral_fraction_against_live_reference(oracle, ref):
    rng = np.random.default_rng(12)
    box = np.array([0., 0., 0., 2 * PC, 1 * PC, 3 * PC])
    ncell = np.array([6, 10, 4], np.int32)
    nc = int(np.prod(ncell))
    n = 10 ** rng.uniform(5, 11, nc)
    J = 10 ** rng.uniform(15, 40, nc)
    J[::7] = 0
    xo = np.zeros(nc)
    ref.ref_compute_neutral_fraction(O.ptr(box, O._dp), O.ptr(ncell, O._ip),
                                     O.ptr(n, O._dp), O.ptr(J, O._dp), 3.3e49,
                                     777777, O.ptr(xo, O._dp))
    grid = O.OracleGrid(oracle, box, ncell, [1, 1, 1])
    grid.n[:] = n
    grid.J[:] = J
    grid.compute_neutral_fraction(3.3e49, 777777, 4e-19, form=0)
    assert (grid.xn == xo).all()


# ---- closed forms and invariants (no reference needed) ----

def _grid(oracle, ncell=(16, 16, 16), nsub=(2, 2, 2), xn=0.0, n=1e8):
    box = np.array([-5 * PC, -5 * PC, -5 * PC, 10 * PC, 10 * PC, 10 * PC])
    g = O.OracleGrid(oracle, box, ncell, nsub)
    g.n[:] = n
    g.xn[:] = xn
    g.J[:] = 0
    return g, box


def _random_packets(rng, m, box, tau=1.0):
    p = np.zeros(m, O.PHOTON_DTYPE)
    p["pos"] = box[:3] + rng.uniform(0.01, 0.99, (m, 3)) * box[3:]
    d = rng.normal(size=(m, 3))
    p["dir"] = d / np.linalg.norm(d, axis=1)[:, None]
    p["tau"] = tau
    p["sigma"] = 6.3e-22
    p["weight"] = 1.0
    return p


def test_empty_box_chord_length_and_crossings(oracle):
    """x_n = 0: nothing is absorbed; every packet leaves the box after the
    ray-box chord; the deposits sum to sigma * w * chord; the crossing count
    is 1 + the number of cell faces the straight ray crosses."""
    rng = np.random.default_rng(1)
    g, box = _grid(oracle, ncell=(24, 24, 24), nsub=(3, 2, 1), xn=0.0)
    p = _random_packets(rng, 400, box)
    start = p["pos"].copy()
    ss = [g.source_subgrid(q) for q in start]
    st, cr, hp = g.transport(p, ss)
    assert (st == 1).all()
    d = p["dir"]
    hi = box[:3] + box[3:]
    t = np.where(d > 0, (hi - start) / np.where(d == 0, 1, d),
                 np.where(d < 0, (box[:3] - start) / np.where(d == 0, 1, d),
                          np.inf)).min(axis=1)
    end = start + t[:, None] * d
    assert np.allclose(p["pos"], end, rtol=0, atol=1e-6 * box[3])
    assert np.isclose(g.J.sum(), 6.3e-22 * t.sum(), rtol=1e-12)
    dx = box[3] / 24
    c0 = np.floor((start - box[:3]) / dx)
    c1 = np.floor((end - 1e-9 * dx * d - box[:3]) / dx)
    expect = 1 + np.abs(c1 - c0).sum(axis=1)
    assert (cr == expect).all()


def test_uniform_absorption_length(oracle):
    """x_n = 1, uniform n: a packet with target tau is absorbed after
    tau / (sigma n) metres and deposits sigma * w * that length."""
    rng = np.random.default_rng(2)
    g, box = _grid(oracle, xn=1.0, n=1e5)
    m = 300
    p = _random_packets(rng, m, box)
    p["pos"] = 0.0
    L = rng.uniform(0.05, 4.5, m) * PC
    p["tau"] = L * 6.3e-22 * 1e5
    start = p["pos"].copy()
    ss = [g.source_subgrid(q) for q in start]
    st, cr, hp = g.transport(p, ss)
    assert (st == 0).all()
    assert np.allclose(p["pos"], start + L[:, None] * p["dir"], rtol=0,
                       atol=1e-9 * PC)
    assert np.isclose(g.J.sum(), 6.3e-22 * L.sum(), rtol=1e-10)
    # the reference stores tau_target - tau_done, where tau_done includes the
    # whole optical depth of the last cell: absorbed packets keep tau <= 0
    # (DensitySubGrid.hpp:1268-1278, 1307)
    assert (p["tau"] <= 0).all()


def test_transport_conserves_packets_and_tau(oracle):
    rng = np.random.default_rng(3)
    g, box = _grid(oracle, xn=1e-3)
    g.n[:] = 1e8 * np.exp(rng.normal(0, 1, g.n.size))
    p = _random_packets(rng, 500, box)
    p["tau"] = rng.exponential(1.0, 500)
    tau0 = p["tau"].copy()
    start = p["pos"].copy()
    ss = [g.source_subgrid(q) for q in start]
    st, cr, hp = g.transport(p, ss)
    assert set(np.unique(st)) <= {0, 1}
    # absorbed packets spent their whole tau, escaped ones kept some
    assert (p["tau"][st == 0] <= 0).all()
    assert (p["tau"][st == 1] > 0).all() and (p["tau"] <= tau0 + 1e-15).all()
    path = np.linalg.norm(p["pos"] - start, axis=1)
    assert np.isclose(g.J.sum(), 6.3e-22 * path.sum(), rtol=1e-9)


def test_parallel_transport_matches_serial(oracle):
    rng = np.random.default_rng(4)
    g, box = _grid(oracle, xn=1e-3)
    g.n[:] = 1e8 * np.exp(rng.normal(0, 1, g.n.size))
    p = _random_packets(rng, 2000, box)
    p["tau"] = rng.exponential(1.0, 2000)
    ss = [g.source_subgrid(q) for q in p["pos"]]
    q = p.copy()
    st, cr, hp = g.transport(p, ss, nthreads=1)
    J1 = g.J.copy()
    g.J[:] = 0
    st2, cr2, hp2 = g.transport(q, ss, nthreads=4)
    assert (st == st2).all() and (cr == cr2).all() and (hp == hp2).all()
    assert (p.view(np.uint8) == q.view(np.uint8)).all()
    assert np.allclose(g.J, J1, rtol=1e-13, atol=0)


@pytest.mark.parametrize("form", [0, 1])
def test_ionization_balance_closed_form(oracle, form):
    """x Gamma = n alpha (1 - x)^2 with Gamma = jH; jH = 0 -> x = 1; large
    jH -> x ~ n alpha / jH."""
    rng = np.random.default_rng(7)
    box = np.array([0., 0., 0., 1 * PC, 1 * PC, 1 * PC])
    g = O.OracleGrid(oracle, box, [8, 8, 8], [1, 1, 1])
    g.n[:] = 10 ** rng.uniform(6, 10, g.n.size)
    # jH / (n alpha) from ~1e-4 (nearly neutral) to ~1e9 (x ~ 1e-9)
    g.J[:] = 10 ** rng.uniform(-11, 3, g.n.size)
    g.J[:10] = 0
    L, N, alpha = 4.26e49, 10 ** 6, 2.7e-19
    J = g.J.copy()
    n = g.n.copy()
    g.compute_neutral_fraction(L, N, alpha, form=form)
    x = g.xn.copy()
    vol = (PC / 8) ** 3
    jH = L / N * J / vol
    assert (x[:10] == 1.0).all()
    ratio = jH / (n * alpha)
    # the reference form 1 + aa (1 - sqrt(1 + 2/aa)) has an absolute error of
    # a few eps * aa (cancellation): compare it where x >> eps * aa only
    m = (jH > 0) & ((form == 1) | (ratio < 1e6))
    lhs = x[m] * jH[m]
    rhs = n[m] * alpha * (1 - x[m]) ** 2
    tol = 1e-12 if form == 1 else 1e-5
    assert np.allclose(lhs, rhs, rtol=tol)
    big = (ratio > 1e6) & ((form == 1) | (ratio < 1e8))
    assert big.sum() > 10
    assert np.allclose(x[big], n[big] * alpha / jH[big], rtol=1e-5)
    if form == 0:
        aa = 0.5 * ratio[jH > 0]
        exact = 2.0 / (2.0 + 2 * aa + np.sqrt(2 * aa * (2 * aa + 4.0)))
        assert (np.abs(x[jH > 0] - exact) <= 8 * 2.2e-16 * (1 + aa)).all()


def test_stable_form_agrees_with_reference_form(oracle):
    box = np.array([0., 0., 0., 1 * PC, 1 * PC, 1 * PC])
    g = O.OracleGrid(oracle, box, [8, 8, 8], [1, 1, 1])
    rng = np.random.default_rng(8)
    n = 10 ** rng.uniform(6, 10, g.n.size)
    J = 10 ** rng.uniform(-11, -3, g.n.size)
    out = []
    for form in (0, 1):
        g.n[:] = n
        g.J[:] = J
        g.compute_neutral_fraction(4.26e49, 10 ** 6, 2.7e-19, form=form)
        out.append(g.xn.copy())
    assert np.allclose(out[0], out[1], rtol=1e-6)


KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2,
     [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
     [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(oracle, ctr, key, expect):
    c = np.array(ctr, np.uint32)
    k = np.array(key, np.uint32)
    o = np.zeros(4, np.uint32)
    oracle.orc_philox4x32_10(O.ptr(c, O._u32p), O.ptr(k, O._u32p),
                             O.ptr(o, O._u32p))
    assert list(o) == expect


def test_emission_statistics(oracle):
    g, box = _grid(oracle)
    src = np.array([[0., 0., 0.], [PC, 2 * PC, -PC]])
    L = np.array([3e49, 1e49])
    p, ss = g.emit(src, L, 200001, 9, 3)
    d = p["dir"]
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-12)
    assert np.abs(d.mean(axis=0)).max() < 0.01
    assert np.allclose((d ** 2).mean(axis=0), 1 / 3, atol=0.01)
    assert abs(p["tau"].mean() - 1.0) < 0.01
    assert abs((p["tau"] ** 2).mean() - 2.0) < 0.04
    n0 = int(np.sum(np.all(p["pos"] == src[0], axis=1)))
    assert n0 + int(np.sum(np.all(p["pos"] == src[1], axis=1))) == 200001
    assert abs(n0 - 0.75 * 200001) <= 1


def test_source_counts(oracle):
    L = np.array([1.0, 2.0, 3.0, 0.5])
    out = np.zeros(4, np.uint64)
    oracle.orc_source_counts(O.ptr(L, O._dp), 4, 1000003, 1, 0,
                             O.ptr(out, O._u64p))
    assert out.sum() == 1000003
    base = np.floor(L / L.sum() * 1000003)
    assert ((out - base) >= 0).all() and ((out - base) <= 1).all()


def test_source_counts_sum_exactly(oracle):
    # the emission files packets by this split: it must sum to N for any
    # luminosities, including ones whose rounded shares sit on integers
    rng = np.random.default_rng(5)
    cases = [(np.full(3, 1.0 / 3.0), 3_000_000), (np.full(7, 0.1), 70),
             (np.array([1e-300, 1.0]), 10)]
    for _ in range(200):
        ns = int(rng.integers(1, 40))
        cases.append((10.0 ** rng.uniform(40, 50, ns), int(rng.integers(1, 10**9))))
    for L, n in cases:
        out = np.zeros(len(L), np.uint64)
        oracle.orc_source_counts(O.ptr(np.ascontiguousarray(L), O._dp), len(L), n, 3, 1,
                                 O.ptr(out, O._u64p))
        assert int(out.sum()) == n
        share = L / L.sum() * n
        assert (np.abs(out - share) <= 1.0 + 1e-6 * share).all()


# ---- the reference's other two walk variants -------------------------------

def _walk_oracle(oracle, box, ncell, n, xn, photons, in_dir, kind):
    s = O.make_subgrid(oracle, box, ncell)
    nn = np.ascontiguousarray(n, dtype=np.float64).copy()
    xx = np.ascontiguousarray(xn, dtype=np.float64).copy()
    J = np.zeros_like(nn)
    s.n = O.ptr(nn, O._dp)
    s.xn = O.ptr(xx, O._dp)
    s.J = O.ptr(J, O._dp)
    p = photons.copy()
    ex = np.zeros(len(p), np.int32)
    cr = np.zeros(len(p), np.uint32)
    one = np.zeros(1, np.uint32)
    for i in range(len(p)):
        one[0] = 0
        ex[i] = oracle.orc_walk(C.byref(s), p[i:i + 1].ctypes.data, in_dir, kind,
                                O.ptr(one, O._u32p))
        cr[i] = one[0]
    return p, ex, cr, J


@pytest.mark.parametrize("kind,name", [(1, "propagate"), (2, "optical_depth")])
def test_walk_variants_match_golden(oracle, golden, kind, name):
    g = golden("walk_variants_12.npz")
    for k in range(3):
        p, ex, cr, J = _walk_oracle(oracle, g["box"], g["ncell"], g["n"], g["xn"],
                                    g["photons_in_%d" % k], int(g["in_dir_%d" % k]),
                                    kind)
        assert (ex == g["%s_exit_%d" % (name, k)]).all()
        assert (cr == g["%s_crossings_%d" % (name, k)]).all()
        assert (p.view(np.uint8) == g["%s_out_%d" % (name, k)].view(np.uint8)).all()
        assert (J == 0).all()   # neither variant deposits


def test_walk_variants_against_live_reference(oracle, ref):
    rng = np.random.default_rng(12)
    box = np.array([0.3 * PC, 0.1 * PC, -0.5 * PC, 0.9 * PC, 1.4 * PC, 0.7 * PC])
    ncell = np.array([10, 14, 6], np.int32)
    nc = int(np.prod(ncell))
    n = 1e8 * np.exp(rng.normal(0, 1.5, nc))
    xn = 10 ** rng.uniform(-7, -1, nc)
    for in_dir in (0, 5, 17, 21):
        m = 250
        p = np.zeros(m, O.PHOTON_DTYPE)
        p["pos"] = box[:3] + rng.uniform(0, 1, (m, 3)) * box[3:]
        d = rng.normal(size=(m, 3))
        p["dir"] = d / np.linalg.norm(d, axis=1)[:, None]
        off = np.zeros(3, np.int32)
        oracle.orc_direction_offset(in_dir, O.ptr(off, O._ip))
        for a in range(3):
            if off[a]:
                p["pos"][:, a] = box[a] + (box[3 + a] if off[a] > 0 else 0.)
                p["dir"][:, a] = -off[a] * np.abs(p["dir"][:, a])
        p["tau"] = -np.log(rng.uniform(0, 1, m))
        p["sigma"] = 6.3e-22
        p["weight"] = 1.0
        for kind in (1, 2):
            po, ex, cr, _ = _walk_oracle(oracle, box, ncell, n, xn, p, in_dir, kind)
            q, exr, crr = O.ref_walk_list(ref, box, ncell, n, xn, p, in_dir, kind)
            assert (ex == exr).all() and (cr == crr).all()
            assert (po.view(np.uint8) == q.view(np.uint8)).all()


def test_propagate_is_interact_without_deposits(oracle):
    """Same packet states and exits as interact, J untouched."""
    rng = np.random.default_rng(13)
    box = np.array([0.0, 0.0, 0.0, PC, PC, PC])
    ncell = np.array([8, 8, 8], np.int32)
    n = 1e8 * np.exp(rng.normal(0, 1, 512))
    xn = 10 ** rng.uniform(-5, -1, 512)
    p = np.zeros(400, O.PHOTON_DTYPE)
    p["pos"] = rng.uniform(0, PC, (400, 3))
    d = rng.normal(size=(400, 3))
    p["dir"] = d / np.linalg.norm(d, axis=1)[:, None]
    p["tau"] = -np.log(rng.uniform(0, 1, 400))
    p["sigma"] = 6.3e-22
    p["weight"] = 1.0
    a, exa, cra, Ja = _interact_oracle(oracle, box, ncell, n, xn, p, 0)
    b, exb, crb, Jb = _walk_oracle(oracle, box, ncell, n, xn, p, 0, 1)
    assert (a.view(np.uint8) == b.view(np.uint8)).all()
    assert (exa == exb).all() and (cra == crb).all()
    assert Ja.sum() > 0 and (Jb == 0).all()


def test_optical_depth_of_uniform_box_is_chord(oracle):
    """Uniform n x: the accumulated optical depth is sigma n x times the
    chord from the start point to the box edge, whatever the target."""
    rng = np.random.default_rng(14)
    L = 2.0 * PC
    box = np.array([0.0, 0.0, 0.0, L, L, L])
    ncell = np.array([16, 16, 16], np.int32)
    nval, xval, sig = 1e8, 3e-3, 6.3e-22
    m = 300
    p = np.zeros(m, O.PHOTON_DTYPE)
    p["pos"] = rng.uniform(0.05 * L, 0.95 * L, (m, 3))
    d = rng.normal(size=(m, 3))
    p["dir"] = d / np.linalg.norm(d, axis=1)[:, None]
    p["tau"] = rng.uniform(0, 2, m)
    p["sigma"] = sig
    p["weight"] = 1.0
    q, ex, cr, J = _walk_oracle(oracle, box, ncell, np.full(4096, nval),
                                np.full(4096, xval), p, 0, 2)
    t = np.where(p["dir"] > 0, (L - p["pos"]) / p["dir"], -p["pos"] / p["dir"])
    chord = t.min(axis=1)
    assert np.allclose(q["tau"] - p["tau"], sig * nval * xval * chord, rtol=1e-9)
    assert (ex != 0).all()
    end = p["pos"] + chord[:, None] * p["dir"]
    assert np.allclose(q["pos"], end, rtol=0, atol=1e-9 * L)


def test_problem_emission_is_the_set_emission(oracle):
    """orc_emit_problem: problem 0 is orc_emit; problem q's packet k draws
    the counters of packet number k + q * 2^52 (the high counter word XOR
    q << 20), and its source split is keyed by seed ^ q."""
    from parallelcmacionize_amd import workloads as W
    box = W.box_10pc()
    og = O.OracleGrid(oracle, box, [16] * 3, [2, 2, 2])
    pos3 = [[0.1e16, 0.2e16, -0.3e16], [1e16, -2e16, 3e16], [-4e16, 0.0, 5e15]]
    L3 = [1e49, 3e48, 7e48]
    a, sa = og.emit(pos3, L3, 5001, 9, 2)
    b, sb = og.emit(pos3, L3, 5001, 9, 2, problem=0)
    assert np.array_equal(a, b) and np.array_equal(sa, sb)
    q, N = 5, 3000
    one = og.emit([[0.0, 0.0, 0.0]], [1e49], N, 9, 2, problem=q)[0]
    base = q << 52
    ref = og.emit([[0.0, 0.0, 0.0]], [1e49], base + N, 9, 2, first=base, count=N)[0]
    assert np.array_equal(one, ref)
    p3, _ = og.emit(pos3, L3, N, 9, 2, problem=q)
    counts = np.zeros(3, np.uint64)
    Ls = np.array(L3)
    oracle.orc_source_counts(O.ptr(Ls, O._dp), 3, N, 9 ^ q, 2, O.ptr(counts, O._u64p))
    for s in range(3):
        assert int((p3["pos"][:, 0] == pos3[s][0]).sum()) == int(counts[s])
