"""Expected outputs of the drop-in C++ layer (include/odba/opendba.hpp) and
the openDBA command line, built from the CPU oracle: the stable length
sort, z-norm, all-vs-all, consensus loop and text formats of performDBA
(dba.hpp:863-1189) and setupAndRun (openDBA.cuh:13-200).

Test infrastructure: clusters come from the product's host clustering
(odba_cluster), which tests/test_cluster_cpu.py pins against SciPy.
"""
import numpy as np

import oracle_lib as O


def g6(v):
    """A value as iostream prints it by default (6 significant digits)."""
    return "%.6g" % v


def stem(path):
    """Sequence name from a file path: everything before the first '.'
    (openDBA.cuh:58-59)."""
    return path[: path.index(".")] if "." in path else path


def stable_order(lens):
    return sorted(range(len(lens)), key=lambda k: lens[k])


def consensus(members, cen, os_=0, oe=0, max_rounds=250):
    """The convergence loop of performDBA (dba.hpp:1078-1110): rounds until
    delta == 0, or in open modes until the new centroid equals the one of
    two rounds back."""
    cen = np.array(cen, copy=True)
    prev = np.zeros_like(cen)
    prev2 = np.zeros_like(cen)
    rounds = 0
    for i in range(max_rounds):
        rc, new, _, delta = O.dba_update(cen, members, os_=os_, oe=oe)
        assert rc == 0
        rounds += 1
        if delta == 0:
            return new, rounds
        if os_ or oe:
            if i >= 1 and np.array_equal(new.view(np.uint8), prev2.view(np.uint8)):
                return new, rounds
            if i > 1:
                prev2 = prev.copy()
            if i > 0:
                prev = new.copy()
        cen = new
    return cen, rounds


def cluster_only(seqs, names, os_, oe, norm, cdist):
    """performDBA(CLUSTER_ONLY): sorted (seqs, names), z-normed seqs,
    memberships and medoids (sorted order)."""
    from opendba_amd import native as N
    order = stable_order([len(s) for s in seqs])
    s = [np.array(seqs[k], copy=True) for k in order]
    nm = [names[k] for k in order]
    if norm:
        s = O.znorm(s)[0]
    d = O.all_pairs(s, os_, oe)
    memb, med = N.cluster(d, [len(x) for x in s], cdist)
    return s, nm, list(memb), list(med), d


def membership_lines(names, memb, med, cdist):
    out = ["## cluster distance threshold was " + g6(cdist)]
    for i, nm in enumerate(names):
        out.append(f"{nm}\t{memb[i]}\t{names[med[memb[i]]]}")
    return out


def consensus_lines(seqs, names, memb, med, os_, oe, norm, mu=None, sg=None, first=0, checkpoints=None):
    """The avg.txt lines of the clusters from `first` on (seqs already
    sorted and z-normed; mu / sg their means and sigmas when norm);
    checkpoints maps a cluster to a starting centroid."""
    lines = []
    for c in range(first, max(memb) + 1):
        m = med[c]
        members = [seqs[i] for i in range(len(seqs)) if memb[i] == c]
        cen = seqs[m].copy()
        if len(members) > 1:
            start = checkpoints.get(c, cen) if checkpoints else cen
            cen, _ = consensus(members, start, os_, oe)
        if norm:
            cen = (mu[m] + cen.astype(np.float64) * sg[m]).astype(seqs[m].dtype)
        lines.append("\t".join([names[m]] + [g6(v) for v in cen]))
    return lines
