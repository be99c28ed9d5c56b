"""Clustering of the all-vs-all distances (odba_cluster, host code) checked
against SciPy's complete linkage, a published implementation of the same
dendrogram: merge heights, the partition at every cut, the label order the
cut assigns, and the reference's cut strategies and medoid rules
(dba.hpp:219-351). The reference's own hclust-cpp is not in its tree, so
the merge order of exactly tied heights stays unpinned."""
import numpy as np
import pytest
from scipy.cluster.hierarchy import fcluster, linkage

from opendba_amd import native as N


def _condensed(rng, n):
    pts = rng.normal(size=(n, 3))
    iu = np.triu_indices(n, 1)
    return np.sqrt(((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1))[iu]


def _same_partition(a, b):
    pa = {}
    for i, x in enumerate(a):
        pa.setdefault(x, set()).add(i)
    pb = {}
    for i, x in enumerate(b):
        pb.setdefault(x, set()).add(i)
    return sorted(map(sorted, pa.values())) == sorted(map(sorted, pb.values()))


@pytest.mark.parametrize("n", [2, 3, 7, 40, 200])
def test_complete_linkage_matches_scipy(n):
    rng = np.random.default_rng(n)
    d = _condensed(rng, n)
    merge, height = N.hclust_complete(d, n)
    Z = linkage(d, method="complete")
    assert np.allclose(np.sort(height), np.sort(Z[:, 2]), rtol=0, atol=0)
    assert np.all(np.diff(height) >= 0)
    for k in range(1, n + 1):
        ours = N.cutree_k(merge, k)
        ref = fcluster(Z, k, criterion="maxclust")
        assert _same_partition(ours, ref), k


def test_merge_matrix_follows_r_convention():
    rng = np.random.default_rng(5)
    n = 30
    merge, height = N.hclust_complete(_condensed(rng, n), n)
    seen = set()
    for s, (a, b) in enumerate(merge, start=1):
        for x in (a, b):
            assert x != 0
            if x < 0:
                assert -x not in seen  # each observation enters once
                seen.add(-x)
            else:
                assert 1 <= x < s  # an earlier step
        # singletons first, otherwise the smaller label first
        if a < 0 < b:
            pass
        elif a < 0 and b < 0:
            assert -a < -b
        else:
            assert a > 0 and b > 0 and a < b
    assert len(seen) == n


def test_cutree_labels_in_first_seen_order():
    rng = np.random.default_rng(9)
    n = 50
    merge, _ = N.hclust_complete(_condensed(rng, n), n)
    for k in (2, 5, 17):
        lab = N.cutree_k(merge, k)
        assert set(lab) == set(range(k))
        first = [lab[i] for i in range(n) if lab[i] not in lab[:i]]
        assert first == list(range(k))


def _walk_dists(rng, n):
    # distances as float32 upper triangle, like odba_dtw_all_pairs output
    return _condensed(rng, n).astype(np.float32)


def test_cluster_cdist_one_is_single_cluster_with_whole_set_medoid():
    rng = np.random.default_rng(1)
    n = 12
    d = _walk_dists(rng, n)
    lens = np.arange(100, 100 + n)
    memb, med = N.cluster(d, lens, 1.0)
    assert (memb == 0).all() and len(med) == 1
    # first strict minimum of the sum of squares in T, host loop order
    D = np.zeros((n, n), np.float32)
    D[np.triu_indices(n, 1)] = d
    D = D + D.T
    sos = [np.float32(0)] * n
    for r in range(n - 1):
        for c in range(r + 1, n):
            v = np.float32(D[r, c] * D[r, c])
            sos[r] = np.float32(sos[r] + v)
            sos[c] = np.float32(sos[c] + v)
    assert med[0] == int(np.argmin(sos))


def test_cluster_height_cut_and_k_cut():
    rng = np.random.default_rng(2)
    n = 40
    d = _walk_dists(rng, n)
    lens = np.full(n, 500)
    Z = linkage(d.astype(np.float64) / float(d.max()), method="complete")
    for cd in (0.2, 0.5, 0.8):
        memb, med = N.cluster(d, lens, cd)
        k = n - int(np.sum(Z[:, 2] < cd))
        assert memb.max() + 1 == k
        assert _same_partition(memb, fcluster(Z, k, criterion="maxclust"))
        for c, m in enumerate(med):
            assert memb[m] == c
    # K = 4 with singletons not counted: at least 4 clusters of 2+ members
    memb, med = N.cluster(d, lens, 4.0)
    counts = np.bincount(memb)
    assert (counts > 1).sum() >= 4 or memb.max() + 1 == n


def test_multi_cluster_medoid_uses_member_ordinals_as_rows():
    """dba.hpp:314-324 reads the distance row of the member's ordinal within
    its cluster and never resets the sums between clusters."""
    n = 6
    d = np.zeros(n * (n - 1) // 2, np.float32)
    idx = {}
    k = 0
    for r in range(n):
        for c in range(r + 1, n):
            idx[(r, c)] = k
            k += 1
    # two obvious groups {0,1,2} and {3,4,5}; values chosen so the ordinal
    # rows (0,1),(0,2),(1,2) decide both clusters' medoids
    for (r, c), k in idx.items():
        same = (r < 3) == (c < 3)
        d[k] = 0.1 if same else 1.0
    d[idx[(0, 1)]] = 0.05
    d[idx[(0, 2)]] = 0.09
    d[idx[(1, 2)]] = 0.08
    memb, med = N.cluster(d, np.full(n, 100), 0.5)
    assert list(memb) == [0, 0, 0, 1, 1, 1]
    # ordinals 0,1,2 read rows (0,1)=.05 (0,2)=.09 (1,2)=.08:
    # sos = [.0106, .0089, .0145] -> ordinal 1; cluster 1 adds the same
    # amounts again -> ordinal 1 again
    assert list(med) == [1, 4]
