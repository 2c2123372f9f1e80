"""The plan path's host AMG setup (amg.rs:127-455, restated in host_amg.cpp) against the reference's own setup
tests (amg.rs:799-942): aggregation coverage, the smoothed prolongator against its dense formula, Galerkin
operators against dense triple products, the boolean Galerkin's preserved entry sum, and an SPD hierarchy that
ends on an exactly factored level.  CPU only (host code)."""
import numpy as np
import scipy.sparse as sp


def _path(n):
    A = sp.diags([-np.ones(n - 1), 2.0 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1], format="csr")
    A.sort_indices()
    return A


def _csr(A):
    from researchuv_amd.solver import CsrMatrix

    return CsrMatrix(A.data, A.indices, A.indptr)


def test_pairwise_aggregation_shrinks_and_covers():
    from researchuv_amd.solver import amg_aggregate

    n = 101
    agg, nc = amg_aggregate(_csr(_path(n)), 2)
    assert (agg >= 0).all() and (agg < nc).all()
    assert nc <= (n + 1) // 2 + 1 and nc < n
    assert (np.bincount(agg, minlength=nc) >= 1).all()


def test_smoothed_prolongator_matches_dense_reference():
    """P = (I - omega D^-1 A) B (amg.rs:308-362, omega = 2/3) on a 12-vertex path with SA_AGG = 4 aggregates;
    interior rows reach into both neighbouring aggregates."""
    from researchuv_amd.solver import amg_aggregate, amg_prolongator

    A = _path(12)
    agg, nc = amg_aggregate(_csr(A), 4)
    P = amg_prolongator(_csr(A), agg, nc, smoothed=True).toarray()
    B = sp.csr_matrix((np.ones(12), (np.arange(12), agg)), shape=(12, nc)).toarray()
    D = A.diagonal()
    want = B - (2.0 / 3.0) * (A.toarray() / D[:, None]) @ B
    np.testing.assert_allclose(P, want, atol=1e-12)
    assert np.count_nonzero(P[6]) >= 2
    T = amg_prolongator(_csr(A), agg, nc, smoothed=False).toarray()
    np.testing.assert_array_equal(T, B)


def test_galerkin_matches_dense_triple_product_and_is_symmetric():
    from researchuv_amd.solver import amg_aggregate, amg_galerkin, amg_prolongator

    A = _path(40)
    agg, nc = amg_aggregate(_csr(A), 4)
    P = amg_prolongator(_csr(A), agg, nc, smoothed=True)
    Ac = amg_galerkin(_csr(A), P, nc).toarray()
    want = P.toarray().T @ A.toarray() @ P.toarray()
    np.testing.assert_allclose(Ac, want, atol=1e-9 * (1 + np.abs(want).max()))
    np.testing.assert_allclose(Ac, Ac.T, atol=1e-10)


def test_boolean_galerkin_keeps_the_entry_sum():
    from researchuv_amd.solver import amg_aggregate, amg_galerkin, amg_prolongator

    A = _path(40)
    agg, nc = amg_aggregate(_csr(A), 2)
    P = amg_prolongator(_csr(A), agg, nc, smoothed=False)
    assert abs(amg_galerkin(_csr(A), P, nc).sum() - A.sum()) < 1e-9


def test_hierarchy_levels_are_galerkin_products_of_their_transfers():
    """The plan path's hierarchy (64 x 64 Poisson grid, both forms): each level's operator is the Galerkin product
    of the level above with its transfer, and every aggregate is used."""
    from researchuv_amd.solver import AmgHost

    A = _poisson(64)
    for smoothed in (False, True):
        h = AmgHost(_csr(A), smoothed=smoothed)
        for k in range(len(h) - 1):
            a, b = h.level(k), h.level(k + 1)
            P = a["p"]
            want = (P.T @ a["op"] @ P).toarray()
            np.testing.assert_allclose(b["op"].toarray(), want, atol=1e-9 * (1 + np.abs(want).max()))
            assert len(np.unique(a["agg"])) == a["n_coarse"] == b["op"].shape[0]


def _poisson(side):
    T = sp.diags([-np.ones(side - 1), 2.0 * np.ones(side), -np.ones(side - 1)], [-1, 0, 1])
    A = (sp.kron(T, sp.identity(side)) + sp.kron(sp.identity(side), T)).tocsr()
    A.sort_indices()
    return A


def test_smoothed_hierarchy_stays_spd_and_bottoms_out():
    """64 x 64 Poisson grid (5-point) under smoothed aggregation: every level symmetric with a positive diagonal,
    at least 3 levels, at most 40, the last level without a transfer, exactly factored, at most 512 rows."""
    from researchuv_amd.solver import AmgHost

    A = _poisson(64)
    h = AmgHost(_csr(A), smoothed=True)
    assert 3 <= len(h) <= 40
    last = h.level(len(h) - 1)
    assert last["p"] is None and last["factored"] and last["op"].shape[0] <= 512
    for k in range(len(h)):
        op = h.level(k)["op"]
        assert (op.diagonal() > 0).all(), k
        d = abs(op - op.T)
        assert d.max() <= 1e-8 * (1 + abs(op).max()), k
