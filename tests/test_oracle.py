"""Pins the CPU oracle before anything is checked against it.

Parity status of each oracle function (what pins it):
  velocity-set tables     reference KATs (velocity_set.rs:282-376) + invariants
                          (qbar maps c -> -c, sum w = 1, sum w c c = I/3)
  MRT M / M^-1            reference matrix rows/entries quoted below
                          (d2q9.rs:44-156, d3q15.rs, d3q19.rs) + M M^-1 = I
  equilibrium             reference KAT (node.rs:1087-1100) + moment identities
  density / velocity      reference KATs (node.rs:953-1069, d2q9.rs:270-292,
                          d3q27.rs:124-149)
  BGK                     reference KAT (node.rs:1071-1085) + conservation
  TRT                     TRT(w, w) == BGK(1/w) + conservation + fixed point
  MRT                     conservation; uniform rates == BGK for D3Q15; the
                          D2Q9/D3Q19 closed-form m_eq quirks detected
  D3Q27 MRT (own basis;   no reference (d3q27.rs:86 placeholder,
   not in the reference)  velocity_set.rs:168 todo!()): M rebuilt here from the
                          tensor polynomials, orthogonality, M M^-1 = I,
                          uniform rates == BGK (m_eq = M f_eq), conservation
                          for any rates, one relaxed mode moves only that
                          moment; the library's tables equal the oracle's
  Guo source              sum S = 0, sum S c = (1 - 1/2tau) F
  scalar source           sum S = (1 - 1/2tau_g) S
  Zou-He                  the node reproduces the prescribed rho / u
  streaming               reference KATs (node.rs:1102-1150)
  neighbours, faces, BB   reference KATs (node.rs:853-941, lattice.rs:448-632)
                          + brute force on random masks
  no-slip                 reference KATs (bc.rs:156-246)
  residual                reference KAT (node.rs:1152-1166) + closed form
  whole step incl. moving-wall BB, anti-BB, scalar BCs, adsorption (double
  q_ads update), Zou-He corners: bit-exact agreement with the independent
  pure-Python restatement tests/pyref.py on tiny lattices (its tables and MRT
  bases are its own, checked against the oracle's here); each of those
  functions is also pinned by mathematics in tests/test_oracle_pins.py
  (fixed points, conservation, a hand-stepped adsorption step).
  Not pinned beyond the reference text: the D2Q9 / D3Q19 closed-form m_eq
  values (d2q9.rs:247-264, d3q19.rs:550-587, including the uz + uz term) —
  no mathematical identity fixes a quirk, so they rest on the transcription.
"""
import json
import os

import numpy as np
import pytest

from lbflow_amd import abi
from lbflow_amd.case import Collision, FaceBC, ScalarFaceBC, ScalarSpec, Spec
from lbflow_amd import configs, geometry
from oracle import oracle

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))
VS = {"D2Q9": abi.D2Q9, "D3Q15": abi.D3Q15, "D3Q19": abi.D3Q19, "D3Q27": abi.D3Q27}


@pytest.fixture(scope="module", autouse=True)
def _built():
    oracle.build()
    oracle.load()


# ------------------------------------------------------------------ tables
@pytest.mark.parametrize("name", ["D2Q9", "D3Q19"])
def test_table_kats(name):
    t = oracle.tables(VS[name])
    k = KATS["tables"][name]
    assert t["d"] == k["d"] and t["q"] == k["q"]
    for i, c in k["c"].items():
        assert list(t["c"][int(i)][: t["d"]]) == c
    for i, w in k["w"].items():
        assert t["w"][int(i)] == w
    for i, qb in k["qbar"].items():
        assert t["qbar"][int(i)] == qb


@pytest.mark.parametrize("vs", [abi.D2Q9, abi.D3Q15, abi.D3Q19, abi.D3Q27])
def test_table_invariants(vs):
    t = oracle.tables(vs)
    c, w, qb, d = t["c"], t["w"], t["qbar"], t["d"]
    assert np.all(c[qb] == -c)
    assert abs(w.sum() - 1.0) < 1e-15
    assert np.allclose((w[:, None, None] * c[:, :, None] * c[:, None, :]).sum(0)[:d, :d],
                       np.eye(d) / 3.0, atol=1e-15)
    assert np.allclose((w[:, None] * c).sum(0), 0, atol=1e-16)
    for f in range(2 * d):
        axis, sign = f // 2, (1 if f % 2 else -1)
        want = [i for i in range(t["q"]) if c[i][axis] == sign]
        assert list(t["qfaces"][f]) == want
        nrm = np.zeros(3, int)
        nrm[axis] = -sign
        assert list(c[t["normal"][f]]) == list(nrm)


def test_face_normals_reference_values():
    # face_normal_directions, src/velocity_set.rs:66-71 (D2Q9) and :95-102
    assert list(oracle.tables(abi.D2Q9)["normal"][:4]) == [1, 3, 2, 4]
    assert list(oracle.tables(abi.D3Q19)["normal"]) == [1, 2, 3, 4, 5, 6]


# ----------------------------------------------------------------- MRT basis
def test_mrt_rows_match_reference_text():
    M, Mi = oracle.mrt_matrices(abi.D2Q9)
    assert list(M[1]) == [-4, -1, -1, -1, -1, 2, 2, 2, 2]          # d2q9.rs:46
    assert list(M[4]) == [0, -2, 0, 2, 0, 1, -1, -1, 1]           # d2q9.rs:49
    assert list(M[8]) == [0, 0, 0, 0, 0, 1, -1, 1, -1]           # d2q9.rs:53
    assert Mi[5][4] == 1.0 / 12.0 and Mi[1][1] == -1.0 / 36.0      # d2q9.rs:112-121, 68-77
    M, Mi = oracle.mrt_matrices(abi.D3Q19)
    assert list(M[1]) == [-30] + [-11] * 6 + [8] * 12               # d3q19.rs:75-78
    assert list(M[2]) == [12] + [-4] * 6 + [1] * 12
    assert list(M[10]) == [0, -4, -4, 2, 2, 2, 2, 1, 1, 1, 1, -2, -2, 1, 1, 1, 1, -2, -2]
    assert list(M[17]) == [0, 0, 0, 0, 0, 0, 0, -1, 1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1]
    assert Mi[0][1] == -5.0 / 399.0 and Mi[1][1] == -11.0 / 2394.0
    assert Mi[7][1] == 4.0 / 1197.0 and Mi[7][2] == 1.0 / 252.0 and Mi[7][4] == 0.025
    M, Mi = oracle.mrt_matrices(abi.D3Q15)
    assert list(M[2]) == [16] + [-4] * 6 + [1] * 8
    assert list(M[14]) == [0] * 7 + [1, -1, -1, 1, -1, 1, -1, 1]
    assert Mi[0][0] == 2.0 / 30.0 and Mi[7][2] == 1.0 / 360.0 and Mi[0][2] == 4.0 / 90.0


@pytest.mark.parametrize("vs", [abi.D2Q9, abi.D3Q15, abi.D3Q19, abi.D3Q27])
def test_mrt_inverse(vs):
    M, Mi = oracle.mrt_matrices(vs)
    assert np.abs(M @ Mi - np.eye(len(M))).max() < 1e-15


# D3Q27 MRT basis of this engine: rows p_a(cx) p_b(cy) p_c(cz) with p0 = 1,
# p1 = a, p2 = 3a^2 - 2, taken in this (total order, then lexicographic) order
D3Q27_ORDERS = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (0, 2, 0), (0, 0, 2),
                (1, 1, 0), (1, 0, 1), (0, 1, 1), (2, 1, 0), (2, 0, 1), (1, 2, 0), (0, 2, 1),
                (1, 0, 2), (0, 1, 2), (1, 1, 1), (2, 2, 0), (2, 0, 2), (0, 2, 2), (2, 1, 1),
                (1, 2, 1), (1, 1, 2), (2, 2, 1), (2, 1, 2), (1, 2, 2), (2, 2, 2)]


def test_d3q27_mrt_basis():
    polys = [np.poly1d([1]), np.poly1d([1, 0]), np.poly1d([3, 0, -2])]
    c = oracle.tables(abi.D3Q27)["c"]
    M_ref = np.array([[polys[a](ci[0]) * polys[b](ci[1]) * polys[g](ci[2]) for ci in c]
                      for a, b, g in D3Q27_ORDERS])
    assert sorted(D3Q27_ORDERS) == sorted({(a, b, g) for a in range(3) for b in range(3)
                                           for g in range(3)})
    M, Mi = oracle.mrt_matrices(abi.D3Q27)
    assert np.array_equal(M, M_ref)
    # density and momentum rows; orthogonal rows; M^-1 = M^T diag(1/|row|^2)
    assert np.array_equal(M[0], np.ones(27))
    assert np.array_equal(M[1:4], c.T)
    G = M @ M.T
    assert np.array_equal(G, np.diag(np.diag(G)))
    assert np.array_equal(Mi, M.T / np.diag(G))


def test_d3q27_mrt_one_relaxed_mode():
    # with only s_k non-zero, M (f* - f) = -s_k (m_k - meq_k) e_k: a transposed
    # or reordered M / M^-1 moves other moments too
    M, _ = oracle.mrt_matrices(abi.D3Q27)
    rng = np.random.default_rng(27)
    rho, u = 1.02, np.array([0.03, -0.02, 0.04])
    feq = oracle.k_equilibrium(abi.D3Q27, rho, u)
    f = feq * (1 + rng.uniform(-0.05, 0.05, 27))
    for k in (0, 1, 4, 7, 16, 20, 26):
        s = np.zeros(27)
        s[k] = 1.3
        d = M @ (oracle.k_mrt(abi.D3Q27, rho, u, f, s) - f)
        want = np.zeros(27)
        want[k] = -1.3 * (M[k] @ f - M[k] @ feq)
        assert np.abs(d - want).max() < 1e-14, k


@pytest.mark.parametrize("vs", [abi.D2Q9, abi.D3Q15, abi.D3Q19, abi.D3Q27])
def test_library_mrt_tables_equal_oracle(vs):
    # the GPU library's constexpr tables (lattice.hpp), read on the host
    import ctypes as C
    from lbflow_amd.engine import load_library
    q = abi.Q_OF[vs]
    M, Mi = np.zeros(q * q), np.zeros(q * q)
    dp = C.POINTER(C.c_double)
    assert load_library().lbfx_mrt_matrices(vs, M.ctypes.data_as(dp), Mi.ctypes.data_as(dp)) == 0
    Mo, Mio = oracle.mrt_matrices(vs)
    assert np.array_equal(M.reshape(q, q), Mo) and np.array_equal(Mi.reshape(q, q), Mio)


# --------------------------------------------------------- per-node kernels
def test_equilibrium_kat():
    k = KATS["equilibrium"]
    assert np.abs(oracle.k_equilibrium(abi.D2Q9, k["rho"], k["u"]) - k["feq"]).max() < 1e-12


@pytest.mark.parametrize("vs", [abi.D2Q9, abi.D3Q15, abi.D3Q19, abi.D3Q27])
def test_equilibrium_moments(vs):
    t = oracle.tables(vs)
    d, c = t["d"], t["c"][:, : t["d"]]
    rng = np.random.default_rng(vs)
    for _ in range(5):
        rho, u = rng.uniform(0.8, 1.2), rng.uniform(-0.1, 0.1, d)
        feq = oracle.k_equilibrium(vs, rho, u)
        assert abs(feq.sum() - rho) < 1e-14
        assert np.abs(feq @ c - rho * u).max() < 1e-15
        P = np.einsum("i,ia,ib->ab", feq, c, c)
        assert np.abs(P - (rho / 3 * np.eye(d) + rho * np.outer(u, u))).max() < 1e-14


def test_bgk_kat():
    k = KATS["bgk"]
    assert np.abs(oracle.k_bgk(abi.D2Q9, k["f"], k["feq"], k["tau"]) - k["fstar"]).max() < 1e-12


def _moments(vs, f):
    t = oracle.tables(vs)
    c = t["c"][:, : t["d"]]
    return f.sum(), f @ c


@pytest.mark.parametrize("vs", [abi.D2Q9, abi.D3Q15, abi.D3Q19, abi.D3Q27])
def test_collisions_conserve_and_relate(vs):
    rng = np.random.default_rng(1)
    q = abi.Q_OF[vs]
    t = oracle.tables(vs)
    for _ in range(3):
        rho, u = rng.uniform(0.9, 1.1), rng.uniform(-0.05, 0.05, t["d"])
        feq = oracle.k_equilibrium(vs, rho, u)
        f = feq * (1 + rng.uniform(-0.05, 0.05, q))
        r0, j0 = _moments(vs, f)
        feq = oracle.k_equilibrium(vs, r0, j0 / r0)
        tau = 0.77
        b = oracle.k_bgk(vs, f, feq, tau)
        tr = oracle.k_trt(vs, f, feq, 1 / tau, 1 / tau)
        assert np.abs(b - tr).max() < 1e-15
        for out in (b, oracle.k_trt(vs, f, feq, 1.3, 0.9)):
            r, j = _moments(vs, out)
            assert abs(r - r0) < 1e-14 and np.abs(j - j0).max() < 1e-15
        s = np.full(q, 1 / tau)
        m = oracle.k_mrt(vs, r0, j0 / r0, f, s)
        r, j = _moments(vs, m)
        assert abs(r - r0) < 1e-13 and np.abs(j - j0).max() < 1e-14
        if vs in (abi.D3Q15, abi.D3Q27):
            # m_eq equals M f_eq (D3Q15's closed form; D3Q27 computes it so,
            # kernel.rs:92-105): uniform rates give BGK
            assert np.abs(m - b).max() < 1e-14
        else:
            # D2Q9 / D3Q19 closed forms deviate from M f_eq (d2q9.rs:247-264,
            # d3q19.rs:575-577), so uniform-rate MRT differs from BGK
            assert np.abs(m - b).max() > 1e-6
        assert np.abs(oracle.k_bgk(vs, feq, feq, tau) - feq).max() < 1e-15


def test_meq_quirks():
    M, _ = oracle.mrt_matrices(abi.D3Q19)
    rho, u = 1.03, np.array([0.02, -0.03, 0.04])
    meq = oracle.k_meq(abi.D3Q19, rho, u)
    Mf = M @ oracle.k_equilibrium(abi.D3Q19, rho, u)
    # the `uz + uz` of d3q19.rs:575,577 where the moment of f_eq has uz*uz:
    # m_9 and m_10 = -m_9 / 2 deviate from M f_eq by exactly rho (uz^2 - 2 uz)
    dev = rho * (u[2] * u[2] - 2.0 * u[2])
    assert abs((meq[9] - Mf[9]) - dev) < 1e-15
    assert abs((meq[10] - Mf[10]) + 0.5 * dev) < 1e-15
    for k in (0, 3, 5, 7, 11, 13, 14, 15):
        assert abs(meq[k] - Mf[k]) < 1e-14
    M, _ = oracle.mrt_matrices(abi.D3Q15)
    Mf = M @ oracle.k_equilibrium(abi.D3Q15, rho, u)
    assert np.abs(oracle.k_meq(abi.D3Q15, rho, u) - Mf).max() < 1e-14


@pytest.mark.parametrize("vs", [abi.D2Q9, abi.D3Q19])
def test_guo_and_scalar_source_moments(vs):
    t = oracle.tables(vs)
    d, c = t["d"], t["c"][:, : t["d"]]
    u, F, tau = np.array([0.03, -0.02, 0.01][:d]), np.array([1e-3, 2e-3, -3e-3][:d]), 0.8
    S = oracle.k_momentum_source(vs, u, F, tau)
    assert abs(S.sum()) < 1e-18
    assert np.abs(S @ c - (1 - 0.5 / tau) * F).max() < 1e-18
    Ss = oracle.k_scalar_source(vs, 0.37, 0.6)
    assert abs(Ss.sum() - (1 - 0.5 / 0.6) * 0.37) < 1e-16


@pytest.mark.parametrize("face", range(4))
def test_zou_he_reproduces_prescribed_state(face):
    rng = np.random.default_rng(face)
    f = oracle.k_equilibrium(abi.D2Q9, 1.01, [0.02, -0.01]) * (1 + rng.uniform(-0.01, 0.01, 9))
    c = oracle.tables(abi.D2Q9)["c"][:, :2]
    ux, uy = 0.05, -0.02
    rc, g = oracle.k_zou_he(face, f, abi.KNOWN_UX | abi.KNOWN_UY, 0.0, (ux, uy))
    assert rc == 0
    assert np.abs(g @ c / g.sum() - [ux, uy]).max() < 1e-14
    if face in (abi.WEST, abi.EAST):
        known, vel = abi.KNOWN_DENSITY | abi.KNOWN_UY, (0.0, uy)
    else:
        known, vel = abi.KNOWN_DENSITY | abi.KNOWN_UX, (ux, 0.0)
    rc, g = oracle.k_zou_he(face, f, known, 1.02, vel)
    assert rc == 0 and abs(g.sum() - 1.02) < 1e-14
    tang = 1 if face in (abi.WEST, abi.EAST) else 0
    assert abs((g @ c)[tang] / g.sum() - vel[tang]) < 1e-14
    # the combinations the reference panics on are rejected
    assert oracle.k_zou_he(face, f, abi.KNOWN_DENSITY | abi.KNOWN_UX | abi.KNOWN_UY, 1.0, (0, 0))[0] != 0


def test_density_velocity_kats():
    for k in KATS["velocity_explicit"]["cases"]:
        u = oracle.k_velocity_explicit(VS[k["vs"]], k["rho"], k["f"])
        assert np.abs(u - k["u"]).max() < 1e-12


# ------------------------------------------------------ node-level (lattice)
def _lattice(dim):
    # Lattice::_test_default (src/momentum.rs:72-99): 10x10 / 10^3 cavity, BGK(0.5)
    if dim == 2:
        return configs.cavity(n=10, tau=0.5)
    return Spec(velocity_set=abi.D3Q19, n=(10, 10, 10), collision=Collision.bgk(0.5),
                faces=[FaceBC.no_slip()] * 3 + [FaceBC.bounce_back(1.0, (0.1, 0, 0))] + [FaceBC.no_slip()] * 2)


def idx(n, p):
    return p[0] + n[0] * (p[1] + (n[1] * p[2] if len(p) > 2 else 0))


def test_neighbour_kats():
    L2 = oracle.OracleCase(_lattice(2))
    k = KATS["neighbours"]
    n = (10, 10)
    for key in ("D2Q9_inner", "D2Q9_edge"):
        node = idx(n, k[key]["node"])
        for i, p in enumerate(k[key]["nbr"]):
            assert L2.neighbor(node, i) == idx(n, p)
    node = idx(n, [9, 9])
    assert L2.neighbor(L2.neighbor(node, 3), 2) == idx(n, k["D2Q9_edge"]["nbr_3_then_2"])
    L3 = oracle.OracleCase(_lattice(3))
    n3 = (10, 10, 10)
    for key in ("D3Q19_inner", "D3Q19_edge"):
        node = idx(n3, k[key]["node"])
        for i, p in k[key]["nbr"].items():
            assert L3.neighbor(node, int(i)) == idx(n3, p)


def test_density_velocity_node_kats():
    L2 = oracle.OracleCase(_lattice(2))
    node = idx((10, 10), [3, 7])
    for case in KATS["density"]["cases"][:2]:
        L2.node_set(node, case["f"])
        L2.node_op(0, node)
        assert abs(L2.node_moments(node)[0] - case["rho"]) < 1e-12
    L3 = oracle.OracleCase(_lattice(3))
    n3 = idx((10, 10, 10), [3, 5, 7])
    L3.node_set(n3, [0.1] * 19)
    L3.node_op(0, n3)
    assert abs(L3.node_moments(n3)[0] - 1.9) < 1e-12
    # velocity with density fixed at 1 (the reference tests do not recompute it)
    for case in KATS["velocity_implicit"]["cases"]:
        f = np.array(case["f"])
        t = oracle.tables(VS[case["vs"]])
        u = f @ t["c"][:, : t["d"]] / case["rho"]
        assert np.abs(u - case["u"]).max() < 1e-12


def test_streaming_kats():
    k = KATS["streaming"]
    L2 = oracle.OracleCase(_lattice(2))
    node = idx((10, 10), k["node"])
    for i, v in enumerate(k["case_1"]["neighbour_fstar_fill"]):
        L2.node_set(L2.neighbor(node, i), [v] * 9, which=1)
    L2.node_op(1, node)
    assert list(L2.node_get(node)) == k["case_1"]["f"]
    for i, fs in enumerate(k["case_2"]["neighbour_fstar"]):
        L2.node_set(L2.neighbor(node, i), fs, which=1)
    L2.node_op(1, node)
    assert list(L2.node_get(node)) == k["case_2"]["f"]


def test_no_slip_kats():
    k = KATS["no_slip"]
    L2 = oracle.OracleCase(_lattice(2))
    for face, name in enumerate(["West", "East", "South", "North"]):
        node = idx((10, 10), [0, 5] if face == 0 else [9, 5])
        L2.node_set(node, k["f"])
        L2.node_set(node, k["fstar"], which=1)
        L2.node_op(2, node, face)
        assert list(L2.node_get(node)) == k[name]


def test_lattice_index_and_faces_kats():
    k = KATS["lattice_index"]
    L2 = oracle.OracleCase(_lattice(2))
    n = k["n"]
    for i, p in k["node_index"].items():
        assert int(i) == idx(n, p)
    _, _, faces = L2.link_tables()
    for fi, name in enumerate(["West", "East", "South", "North"]):
        for p in k["faces"][name]:
            assert faces[idx(n, p)] >> fi & 1
    assert faces[idx(n, k["not_on_any_face"])] == 0


def test_residual_kat():
    k = KATS["residuals"]
    assert abs((k["rho"] - 1.0) - k["d_rho"]) < 1e-12
    L = oracle.OracleCase(configs.cavity(n=6, tau=0.8))
    L.initialize_nodes()
    L.main_steps(1, residuals=True)
    r, ru, _ = L.residuals()
    rho, u = L.fields()
    assert abs(r - np.sqrt(((rho - 1.0) ** 2).sum())) < 1e-15
    assert np.abs(ru - np.sqrt((u ** 2).sum(0))).max() < 1e-15


@pytest.mark.parametrize("vs,shape", [(abi.D2Q9, (7, 5)), (abi.D3Q19, (5, 4, 6)),
                                      (abi.D3Q15, (4, 5, 3)), (abi.D3Q27, (3, 4, 5))])
def test_link_tables_brute_force(vs, shape):
    mask = geometry.bernoulli_mask(shape, 0.3, 11)
    d = len(shape)
    sp = Spec(velocity_set=vs, n=shape, collision=Collision.bgk(0.7),
              faces=[FaceBC.periodic()] * (2 * d), node_type=mask)
    L = oracle.OracleCase(sp)
    t, m, f = L.link_tables()
    c = oracle.tables(vs)["c"]
    n = list(shape) + [1] * (3 - d)
    for i in range(len(mask)):
        x, y, z = i % n[0], (i // n[0]) % n[1], i // (n[0] * n[1])
        want = 0
        if not mask[i]:
            for k in range(abi.Q_OF[vs]):
                xx, yy, zz = (x + c[k][0]) % n[0], (y + c[k][1]) % n[1], (z + c[k][2]) % n[2]
                if mask[xx + n[0] * (yy + n[1] * zz)]:
                    want |= 1 << k
        assert m[i] == want
    assert np.array_equal(t, mask)


def test_map_xyz_roundtrip(tmp_path):
    shape = (5, 4, 3)
    mask = geometry.bernoulli_mask(shape, 0.4, 3)
    p = tmp_path / "map.xyz"
    geometry.write_map_xyz(str(p), mask, shape)
    assert np.array_equal(geometry.read_map_xyz(str(p)), mask)
    # first line of the file is the top row (y = ny-1) of slice z = 0
    first = open(p).read().splitlines()[0].split()
    assert first == [str(v) for v in mask.reshape(3, 4, 5)[0, 3]]


def test_rest_state_stays_at_rest():
    L = oracle.OracleCase(configs.periodic_box(n=(6, 5, 4)))
    L.initialize_nodes()
    f0 = L.pops()
    L.main_steps(5)
    assert np.abs(L.pops() - f0).max() < 1e-15


def test_step_count_semantics():
    # the loop exits after main_steps at t = max_iterations + 1 (momentum.rs:273-292)
    L = oracle.OracleCase(configs.cavity(n=8, tau=0.6))
    L.initialize_nodes()
    assert L.run(3) == 5


# -------------------------------------------- whole steps vs pyref (tiny)
def _pyref_cases():
    mask2 = geometry.bernoulli_mask((7, 6), 0.2, 5)
    cases = {
        "cavity_bgk": configs.cavity(n=6, tau=0.6),
        "cavity_trt": configs.cavity(n=6, op=Collision.trt(1.2, 0.9)),
        "cavity_mrt": configs.cavity(n=6, op=Collision.mrt([1.0, 1.4, 1.2, 1.0, 1.1, 1.0, 1.1, 1.3, 1.3])),
        "porous_d3q19_force_bgk": configs.porous_channel(n=(5, 4, 4), op=Collision.bgk(0.7), force=1e-3,
                                                         p_solid=0.2, seed=1),
        "porous_d3q19_trt": configs.porous_channel(n=(5, 4, 4), force=1e-3, p_solid=0.2, seed=3),
        "d3q15_mrt": configs.periodic_box(n=(4, 3, 5), velocity_set=abi.D3Q15,
                                          op=Collision.mrt([1.0, 1.2, 1.3, 1.0, 1.1, 1.0, 1.1, 1.0, 1.1,
                                                            1.25, 1.25, 1.25, 1.25, 1.25, 1.1])),
        "coupled_zou_he": configs.coupled_discs(nx=9, ny=6, count=2, rmin=1, rmax=1.6, seed=1,
                                                precision=abi.F64, k_a=0.05),
        "coupled_zou_he_unknown_only": configs.coupled_discs(nx=9, ny=6, count=2, rmin=1, rmax=1.6,
                                                             seed=2, precision=abi.F64, k_a=0.05,
                                                             zeroth_unknown_only=True),
    }
    sp = configs.cavity(n=7, tau=0.7).replace(n=(7, 6), node_type=mask2)
    sp.faces = [FaceBC.bounce_back(1.0, (0.0, 0.02)), FaceBC.no_slip(), FaceBC.periodic(),
                FaceBC.periodic()]
    sp.scalars = [ScalarSpec(collision=Collision.bgk(0.8),
                             faces=[ScalarFaceBC(abi.SBC_ANTI_BB_NO_FLUX),
                                    ScalarFaceBC(abi.SBC_SECOND_ORDER_NO_FLUX),
                                    ScalarFaceBC(abi.SBC_PERIODIC), ScalarFaceBC(abi.SBC_PERIODIC)],
                             inner_bc=abi.INNER_ANTI_BOUNCE_BACK,
                             initial_value=np.linspace(0, 1, 42), source=1e-3,
                             adsorption=dict(adsorption_rate=0.1, desorption_rate=0.01, q_max=0.5,
                                             initial_q_ads=0.05)),
                  ScalarSpec(collision=Collision.trt(1.1, 0.8),
                             faces=[ScalarFaceBC(abi.SBC_ANTI_BOUNCE_BACK, 0.3),
                                    ScalarFaceBC(abi.SBC_ZEROTH_ORDER_NO_FLUX),
                                    ScalarFaceBC(abi.SBC_PERIODIC), ScalarFaceBC(abi.SBC_PERIODIC)],
                             inner_bc=abi.INNER_BOUNCE_BACK, initial_value=0.2,
                             adsorption=dict(adsorption_rate=0.2, q_max=1.0))]
    cases["two_scalars_mixed_bcs"] = sp
    return cases


@pytest.mark.parametrize("vs", [abi.D2Q9, abi.D3Q15, abi.D3Q19])
def test_python_restatement_tables_agree(vs):
    # pyref writes its tables and MRT bases out itself (literature polynomials):
    # they must be the oracle's numbers, bit for bit
    import pyref
    d, q, c, w, qbar, qfaces, normal = pyref.vs_tables(vs)
    t = oracle.tables(vs)
    assert d == t["d"] and q == t["q"]
    assert np.array_equal(np.array(c), t["c"][:, :d]) and np.array_equal(w, t["w"])
    assert np.array_equal(qbar, t["qbar"]) and np.array_equal(normal, t["normal"][: 2 * d])
    for f in range(2 * d):
        assert sorted(qfaces[f]) == sorted(t["qfaces"][f])
    M, Mi = pyref.mrt_matrices(vs)
    Mo, Mio = oracle.mrt_matrices(vs)
    assert np.array_equal(np.array(M), Mo) and np.array_equal(np.array(Mi), Mio)


@pytest.mark.parametrize("name", list(_pyref_cases()))
def test_oracle_matches_python_restatement(name):
    import pyref
    sp = _pyref_cases()[name]
    L = oracle.OracleCase(sp)
    L.initialize_nodes()
    P = pyref.PyLattice(sp)
    P.initialize_nodes()
    for _ in range(4):
        L.main_steps(1)
        P.main_steps(1)
        assert np.array_equal(L.pops(), np.array(P.pops())), name
        rho, u = L.fields()
        prho, pu = P.fields()
        assert np.array_equal(rho, prho) and np.array_equal(u, np.array(pu))
        for s, S in enumerate(P.sc):
            c, qa = L.scalar(s)
            assert np.array_equal(c, [S["c"][i] for i in range(L.N)])
            assert np.array_equal(L.pops(1 + s), np.array([S["g"][i] for i in range(L.N)]))
            if sp.scalars[s].adsorption:
                assert np.array_equal(qa, [S["q"][i] for i in range(L.N)])
