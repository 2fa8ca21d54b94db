"""GPU parity of the HIP evaluator against the CPU oracle.

Every case calls the library through its C ABI (ctypes) and compares with
the oracle (a restatement of example/tree.cpp + example/calculate.cpp):
  * operators at every compiled order: <= 1e-12 relative (norm-wise);
  * tree: bit-exact centres, half-widths, rmax, counts and leaf membership
    in canonical (octant-path) form;
  * lists: M2L and P2P lists equal as multisets;
  * full FP64 FMM: <= 1e-12 relative to the oracle's compute_field_fmm;
  * mixed FP32 / BF16x3 FMM: <= 1e-6 relative to the oracle's FP64 FMM at
    the same order (field components);
  * a tree reused with new strengths: the same bars against an oracle
    built with those strengths.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests import util

pytestmark = pytest.mark.gpu

TOL = 1e-12
FULL_BODIES_MAXORDER = 8


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _log(msg):
    print(msg, flush=True)


# ------------------------------------------------------------------ operators

@pytest.mark.parametrize("so,outsz", [(0, 4), (1, 3), (0, 1), (0, 3), (2, 4)])
def test_operator_batches_match_oracle(gpu, so, outsz):
    """Every operator at every compiled order (whole generated bodies up to
    the full-body order; chunked bodies and the derivative-block sums
    above it)."""
    import ctypes
    from fmmgen_amd import tree as T
    from fmmgen_amd import _lib
    from fmmgen_amd.codegen.layout import Msize, Lsize
    lib = O.lib(so, outsz)
    mn, mx = ctypes.c_int(), ctypes.c_int()
    _lib.check(_lib.load().fmm_variant_orders(so, outsz, ctypes.byref(mn), ctypes.byref(mx)))
    rng = np.random.default_rng(7)
    nb = 64
    for p in range(mn.value, mx.value):
        ms, ls = Msize(p, so), Lsize(p, so)
        xyz = rng.uniform(-1, 1, (nb, 3))
        xyz_far = xyz * 3 + np.sign(xyz) * 2.0      # keep 1/R well conditioned for M2L / M2P
        cases = {
            "M2M": (xyz, rng.uniform(-1, 1, (nb, ms)), ms, lib.oracle_op_M2M),
            "M2L": (xyz_far, rng.uniform(-1, 1, (nb, ms)), ls, lib.oracle_op_M2L),
            "L2L": (xyz, rng.uniform(-1, 1, (nb, ls)), ls, lib.oracle_op_L2L),
            "L2P": (xyz, rng.uniform(-1, 1, (nb, ls)), outsz, lib.oracle_op_L2P),
            "M2P": (xyz_far, rng.uniform(-1, 1, (nb, ms)), outsz, lib.oracle_op_M2P),
        }
        if so == 0:
            cases["P2M"] = (xyz, rng.uniform(-1, 1, (nb, 1)), ms, lib.oracle_op_P2M)
        for op, (x, inp, osz, fn) in cases.items():
            _log(f"op {op} so={so} p={p}")
            out0 = rng.uniform(-1, 1, (nb, osz))   # the operators accumulate into their output
            out = out0.copy()
            for i in range(nb):
                a = np.ascontiguousarray(inp[i])
                o = np.ascontiguousarray(out[i])
                if op == "P2M":
                    fn(x[i, 0], x[i, 1], x[i, 2], a[0], o.ctypes.data_as(O.c_double_p), p)
                else:
                    fn(x[i, 0], x[i, 1], x[i, 2], a.ctypes.data_as(O.c_double_p), o.ctypes.data_as(O.c_double_p), p)
                out[i] = o
            g = T.op_batch(op, p, _dev(x), _dev(inp), _dev(out0), so, outsz).cpu().numpy()
            assert util.rel_norm(g - out0, out - out0) < TOL, (op, p)


# ------------------------------------------------------------------ tree + lists

TREE_CASES = [
    ("uniform", 5000, 16, 0.5),
    ("uniform", 20000, 64, 0.5),
    ("plummer", 20000, 32, 0.4),
    ("blobs", 20000, 64, 0.5),
    ("uniform", 3000, 8, 0.3),
    ("uniform", 50, 64, 0.5),      # root is a leaf
    ("uniform", 1, 64, 0.5),       # single particle
]


def _positions(kind, n, seed):
    if kind == "uniform":
        return util.uniform(n, seed)[0]
    if kind == "plummer":
        return util.plummer(n, seed)
    if kind == "mixed":
        from fmmgen_amd import workloads
        return workloads.mixed(n, seed)
    return util.blobs(n, seed)


@pytest.mark.parametrize("kind,n,ncrit,theta", TREE_CASES)
def test_tree_and_lists_bit_exact(gpu, kind, n, ncrit, theta):
    from fmmgen_amd import build_tree
    _log(f"tree {kind} n={n} ncrit={ncrit} theta={theta}")
    pos = _positions(kind, n, 3)
    q = np.random.default_rng(4).uniform(-1, 1, (n, 1))
    ot = O.OracleTree(pos, q, ncrit, 4, theta)
    opaths, ocells = util.oracle_canonical(ot)
    gt = build_tree(_dev(pos), _dev(q), n, ncrit, 4, theta)
    geo, meta = gt.export_cells()
    gpaths, gcells = util.gpu_canonical(geo, meta)
    assert set(gcells) == set(ocells)
    for k, v in ocells.items():
        assert gcells[k] == v, k          # bit-exact geometry and counts
    perm, leaf = gt.export_particles()
    oleaf = ot.leaf_of()
    assert [gpaths[c] for c in leaf] == [opaths[c] for c in oleaf]
    # leaf ranges hold caller indices in increasing order
    for c in range(geo.shape[0]):
        if meta[c, 4] == 0:
            b, k = meta[c, 2], meta[c, 3]
            seg = perm[b:b + k]
            assert np.all(np.diff(seg) > 0)
    om2l, op2p = ot.lists()
    gm2l, gp2p = gt.export_lists()
    assert sorted((opaths[s], opaths[t]) for s, t in om2l) == sorted((gpaths[s], gpaths[t]) for s, t in gm2l)
    assert sorted((opaths[t], opaths[s]) for t, s in op2p) == sorted((gpaths[t], gpaths[s]) for t, s in gp2p)


# ------------------------------------------------------------------ full FMM

FMM_CASES = [
    # kind, n, ncrit, order, theta, so
    ("uniform", 20000, 64, 4, 0.5, 0),
    ("uniform", 20000, 64, 8, 0.5, 0),
    ("plummer", 20000, 32, 6, 0.4, 0),
    ("blobs", 30000, 64, 5, 0.5, 0),
    ("plummer", 20000, 32, 6, 0.4, 1),
    ("uniform", 10000, 16, 3, 0.5, 1),
    ("uniform", 20000, 64, 10, 0.5, 0),
    ("blobs", 20000, 64, 12, 0.5, 0),
    ("mixed", 30000, 128, 10, 0.5, 0),    # half uniform, half Plummer a=0.05, ncrit 128
]


@pytest.mark.parametrize("kind,n,ncrit,order,theta,so", FMM_CASES)
def test_fmm_plain_f64_matches_oracle(gpu, kind, n, ncrit, order, theta, so):
    from fmmgen_amd import build_tree
    _log(f"fmm {kind} n={n} ncrit={ncrit} p={order} theta={theta} so={so}")
    pos = _positions(kind, n, 11)
    S = np.random.default_rng(12).uniform(-1, 1, (n, 1)) if so == 0 else util.dipoles(n, 12)
    outsz = 4 if so == 0 else 3
    ot = O.OracleTree(pos, S, ncrit, order, theta, source_order=so, outputsize=outsz)
    Fo = ot.compute_field_fmm()
    Sd = _dev(S)
    gt = build_tree(_dev(pos), Sd, n, ncrit, order, theta, source_order=so, potential=(so == 0))
    F = torch.zeros(n, outsz, dtype=torch.float64, device="cuda")
    gt.compute_field_fmm(F, mode="plain")
    Fg = F.cpu().numpy()
    assert util.rel_norm(Fg, Fo) < TOL
    # per phase: far field only and near field only
    for mask in (8, 63 - 8):
        Fo2 = ot.compute_field_fmm(mask)
        gt.compute_field_fmm(F, mode="plain", phases=mask)
        assert util.rel_norm(F.cpu().numpy(), Fo2) < TOL, mask


def test_fmm_split_p2p_leaves_match_oracle(gpu):
    """A dense cluster beside sparse background: background leaves next to
    the cluster have far more than one segment of sources, so their P2P
    work is spread over several items that add atomically."""
    from fmmgen_amd import build_tree
    rng = np.random.default_rng(31)
    n_c, n_b = 24000, 400
    pos = np.vstack([0.5 + 0.004 * rng.standard_normal((n_c, 3)), rng.random((n_b, 3))])
    n = len(pos)
    q = rng.uniform(-1, 1, (n, 1))
    _log(f"fmm split leaves n={n}")
    ot = O.OracleTree(pos, q, 64, 5, 0.5)
    Fo = ot.compute_field_fmm()
    gt = build_tree(_dev(pos), _dev(q), n, 64, 5, 0.5)
    info = gt.info()
    assert info["p2p_items"] > info["nleaves"]
    for mode, tol in (("plain", TOL), ("f32", 1e-6)):
        F = torch.zeros(n, 4, dtype=torch.float64, device="cuda")
        gt.compute_field_fmm(F, mode=mode)
        err = util.rel_norm(F.cpu().numpy()[:, 1:], Fo[:, 1:])
        _log(f"  {mode} field difference {err:.3e}")
        assert err < tol
        gt.compute_field_fmm(F, mode=mode, phases=8)
        assert util.rel_norm(F.cpu().numpy(), ot.compute_field_fmm(8)) < tol


def test_direct_sampled_matches_oracle(gpu):
    from fmmgen_amd import direct_sum
    pos, q = util.uniform(20000, 21)
    tg = np.random.default_rng(5).choice(20000, 300, replace=False)
    Fo = O.direct_targets(pos, q, tg)
    Fg = direct_sum(_dev(pos), _dev(q), torch.from_numpy(tg).cuda()).cpu().numpy()
    assert util.rel_norm(Fg, Fo) < TOL


# ------------------------------------------------------------------ mixed precision

MIXED_CASES = [
    ("uniform", 20000, 64, 8, 0.5, 0, "f32"),
    ("plummer", 20000, 32, 6, 0.4, 0, "f32"),
    ("plummer", 20000, 32, 6, 0.4, 1, "f32"),
    ("uniform", 20000, 64, 8, 0.5, 0, "bf16x3"),
    ("plummer", 20000, 32, 6, 0.4, 1, "bf16x3"),
    ("blobs", 20000, 64, 12, 0.5, 0, "f32"),
    ("blobs", 20000, 64, 12, 0.5, 0, "bf16x3"),
    ("uniform", 20000, 64, 10, 0.5, 0, "bf16x3"),
    ("mixed", 30000, 128, 10, 0.5, 0, "bf16x3"),
    ("mixed", 30000, 128, 10, 0.5, 0, "f32"),
]


_ORACLE_FMM = {}   # oracle results shared by the two precisions of one case


@pytest.mark.parametrize("kind,n,ncrit,order,theta,so,mode", MIXED_CASES)
def test_fmm_mixed_matches_oracle(gpu, kind, n, ncrit, order, theta, so, mode):
    """Mixed modes (FP32 or BF16x3 M2L on the matrix cores, FP32 compensated
    P2P) stay within 1e-6 (relative, norm-wise, field components) of the
    oracle's FP64 FMM at the same order, and of the HIP FP64 path: far
    below the truncation error of the expansions at these orders."""
    from fmmgen_amd import build_tree
    _log(f"mixed {mode} {kind} n={n} ncrit={ncrit} p={order} so={so}")
    pos = _positions(kind, n, 13)
    S = np.random.default_rng(14).uniform(-1, 1, (n, 1)) if so == 0 else util.dipoles(n, 14)
    outsz = 4 if so == 0 else 3
    fo = 1 if so == 0 else 0
    key = (kind, n, ncrit, order, theta, so)
    if key not in _ORACLE_FMM:
        _ORACLE_FMM[key] = O.OracleTree(pos, S, ncrit, order, theta, source_order=so,
                                        outputsize=outsz).compute_field_fmm()
    Fo = _ORACLE_FMM[key]
    gt = build_tree(_dev(pos), _dev(S), n, ncrit, order, theta, source_order=so, potential=(so == 0))
    Fm = torch.zeros(n, outsz, dtype=torch.float64, device="cuda")
    gt.compute_field_fmm(Fm, mode=mode)
    err = util.rel_norm(Fm.cpu().numpy()[:, fo:], Fo[:, fo:])
    _log(f"  {mode} vs oracle field difference {err:.3e}")
    assert err < 1e-6
    F64 = torch.zeros_like(Fm)
    gt.compute_field_fmm(F64, mode="plain")
    assert util.rel_norm(F64.cpu().numpy(), Fo) < TOL
    assert util.rel_norm(Fm.cpu().numpy()[:, fo:], F64.cpu().numpy()[:, fo:]) < 1e-6


HARMONIC_CASES = [
    # kind, n, ncrit, orders, theta
    ("blobs", 20000, 64, (2, 3, 4, 5, 6, 7), 0.5),
    ("uniform", 20000, 64, (8, 9, 10, 11, 12), 0.5),
    ("mixed", 30000, 128, (10,), 0.5),
]


@pytest.mark.parametrize("kind,n,ncrit,orders,theta", HARMONIC_CASES)
def test_fmm_harmonic_m2l_f64_matches_oracle(gpu, kind, n, ncrit, orders, theta):
    """The harmonic-layout M2L (folded multipoles, local rows with n_z <= 1,
    unfolded afterwards) in FP64 gives the oracle's expansions to rounding
    at every compiled order: the same sums as the generated operator."""
    from fmmgen_amd import build_tree
    pos = _positions(kind, n, 21)
    S = np.random.default_rng(22).uniform(-1, 1, (n, 1))
    gt = build_tree(_dev(pos), _dev(S), n, ncrit, max(orders), theta)
    gt.set_m2l_form("harmonic")
    F = torch.zeros(n, 4, dtype=torch.float64, device="cuda")
    for order in orders:
        _log(f"harmonic f64 {kind} n={n} ncrit={ncrit} p={order}")
        Fo = O.OracleTree(pos, S, ncrit, order, theta).compute_field_fmm()
        gt.set_order(order)
        gt.compute_field_fmm(F, mode="plain")
        err = util.rel_norm(F.cpu().numpy(), Fo)
        _log(f"  vs oracle {err:.3e}")
        assert err < TOL, order


@pytest.mark.parametrize("order,mode", [(8, "f32"), (12, "f32"), (8, "bf16x3"), (12, "bf16x3")])
def test_fmm_mixed_full_form_matches_oracle(gpu, order, mode):
    """The FP32 and BF16x3 GEMMs on the full graded expansions (the form the
    automatic choice replaces by the harmonic one) stay within 1e-6 of the
    oracle, and the two forms agree to FP32 rounding."""
    from fmmgen_amd import build_tree
    n = 20000
    pos = _positions("blobs", n, 23)
    S = np.random.default_rng(24).uniform(-1, 1, (n, 1))
    Fo = O.OracleTree(pos, S, 64, order, 0.5).compute_field_fmm()
    gt = build_tree(_dev(pos), _dev(S), n, 64, order, 0.5)
    out = {}
    for form in ("full", "harmonic"):
        _log(f"{mode} {form} blobs n={n} p={order}")
        gt.set_m2l_form(form)
        F = torch.zeros(n, 4, dtype=torch.float64, device="cuda")
        gt.compute_field_fmm(F, mode=mode)
        out[form] = F.cpu().numpy()
        err = util.rel_norm(out[form][:, 1:], Fo[:, 1:])
        _log(f"  vs oracle {err:.3e}")
        assert err < 1e-6, form
    assert util.rel_norm(out["full"][:, 1:], out["harmonic"][:, 1:]) < 1e-6


# ------------------------------------------------------------------ other variants

VARIANT_CASES = [
    # kind, n, ncrit, order, theta, source order, output size
    ("uniform", 20000, 64, 0, 0.5, 0, 1),     # potential only, monopoles only
    ("plummer", 20000, 32, 7, 0.4, 0, 1),
    ("blobs", 20000, 64, 10, 0.5, 0, 3),      # field only
    ("uniform", 20000, 64, 4, 0.5, 2, 4),     # quadrupole sources
    ("plummer", 20000, 32, 8, 0.4, 2, 4),
]


def _sources(so, n, seed):
    from fmmgen_amd.codegen.layout import sourcesize
    return np.random.default_rng(seed).uniform(-1, 1, (n, sourcesize(so)))


@pytest.mark.parametrize("kind,n,ncrit,order,theta,so,outsz", VARIANT_CASES)
def test_other_variants_match_oracle(gpu, kind, n, ncrit, order, theta, so, outsz):
    """Potential-only and field-only charge variants and quadrupole sources:
    plain FMM and Barnes-Hut <= 1e-12 against the oracle's variant; the
    mixed modes <= 1e-6 on the field components, and on the potential
    1e-6 (FP32 M2L) or 1e-5 (BF16x3 M2L: about 16 significand bits per
    product, and the potential, unlike the field, is mostly far field)."""
    from fmmgen_amd import build_tree
    _log(f"variant so={so} outsz={outsz} {kind} n={n} p={order}")
    pos = _positions(kind, n, 23)
    S = _sources(so, n, 24)
    ot = O.OracleTree(pos, S, ncrit, order, theta, source_order=so, outputsize=outsz)
    gt = build_tree(_dev(pos), _dev(S), n, ncrit, order, theta, source_order=so,
                    potential=outsz in (1, 4), field=outsz in (3, 4))
    F = torch.zeros(n, outsz, dtype=torch.float64, device="cuda")
    Fo = ot.compute_field_fmm()
    gt.compute_field_fmm(F, mode="plain")
    assert util.rel_norm(F.cpu().numpy(), Fo) < TOL
    fo = 1 if outsz == 4 else 0      # field components
    for mode in ("f32", "bf16x3"):
        gt.compute_field_fmm(F, mode=mode)
        Fm = F.cpu().numpy()
        if outsz == 1:
            err = util.rel_norm(Fm, Fo)
            _log(f"  {mode} potential: {err:.3e}")
            assert err < (1e-6 if mode == "f32" else 1e-5), mode
        else:
            err = util.rel_norm(Fm[:, fo:], Fo[:, fo:])
            _log(f"  {mode} field: {err:.3e}")
            assert err < 1e-6, mode
            if outsz == 4:
                ep = util.rel_norm(Fm[:, 0], Fo[:, 0])
                _log(f"  {mode} potential: {ep:.3e}")
                assert ep < (1e-6 if mode == "f32" else 1e-5), mode
    gt.compute_field_bh(F)
    assert util.rel_norm(F.cpu().numpy(), ot.compute_field_bh()) < TOL


# ------------------------------------------------------------------ Barnes-Hut

BH_CASES = [
    # kind, n, ncrit, order, theta, so
    ("uniform", 20000, 64, 4, 0.5, 0),
    ("plummer", 20000, 32, 8, 0.4, 0),
    ("blobs", 20000, 64, 10, 0.3, 0),
    ("blobs", 10000, 64, 12, 0.5, 0),
    ("mixed", 20000, 128, 9, 0.5, 0),
    ("plummer", 20000, 32, 6, 0.4, 1),
    ("uniform", 3000, 16, 5, 0.0, 0),     # every cell opened: the direct sum
    ("uniform", 50, 64, 5, 0.5, 0),       # root is a leaf
    ("uniform", 1, 64, 3, 0.5, 0),
]


@pytest.mark.parametrize("kind,n,ncrit,order,theta,so", BH_CASES)
def test_bh_matches_oracle(gpu, kind, n, ncrit, order, theta, so):
    """Tree::compute_field_bh against the oracle's restatement: every
    opening decision is taken as the reference takes it, so the sums agree
    to FP64 rounding."""
    from fmmgen_amd import build_tree
    _log(f"bh {kind} n={n} ncrit={ncrit} p={order} theta={theta} so={so}")
    pos = _positions(kind, n, 17)
    S = np.random.default_rng(18).uniform(-1, 1, (n, 1)) if so == 0 else util.dipoles(n, 18)
    outsz = 4 if so == 0 else 3
    ot = O.OracleTree(pos, S, ncrit, order, theta, source_order=so, outputsize=outsz)
    Fo = ot.compute_field_bh()
    gt = build_tree(_dev(pos), _dev(S), n, ncrit, order, theta, source_order=so, potential=(so == 0))
    F = torch.full((n, outsz), float("nan"), dtype=torch.float64, device="cuda")
    gt.compute_field_bh(F)
    Fg = F.cpu().numpy()
    if n == 1:
        assert np.all(Fg == 0.0)
        return
    assert util.rel_norm(Fg, Fo) < TOL
    # an FMM evaluation after it still matches (the BH pass leaves no state behind)
    gt.compute_field_fmm(F, mode="plain")
    assert util.rel_norm(F.cpu().numpy(), ot.compute_field_fmm()) < TOL


# ------------------------------------------------------------------ tree reuse

def test_tree_reuse_with_new_strengths_matches_oracle(gpu):
    """One tree, new source strengths bound with fmm_tree_set_sources (the
    positions, tree and lists stay): plain and mixed evaluations equal the
    oracle built with the new strengths, twice in a row."""
    from fmmgen_amd import build_tree
    n, ncrit, order, theta = 20000, 64, 8, 0.5
    pos = util.blobs(n, 41)
    rng = np.random.default_rng(42)
    q1 = rng.uniform(-1, 1, (n, 1))
    gt = build_tree(_dev(pos), _dev(q1), n, ncrit, order, theta)
    F = torch.zeros(n, 4, dtype=torch.float64, device="cuda")
    gt.compute_field_fmm(F, mode="plain")
    for k in range(2):
        q = rng.uniform(-1, 1, (n, 1)) * (k + 2)
        _log(f"reuse: strengths set {k}")
        gt.set_sources(_dev(q))
        Fo = O.OracleTree(pos, q, ncrit, order, theta).compute_field_fmm()
        for mode, tol in (("plain", TOL), ("f32", 1e-6), ("bf16x3", 1e-6)):
            gt.compute_field_fmm(F, mode=mode)
            fo = 0 if mode == "plain" else 1
            err = util.rel_norm(F.cpu().numpy()[:, fo:], Fo[:, fo:])
            _log(f"  {mode}: {err:.3e}")
            assert err < tol, mode
