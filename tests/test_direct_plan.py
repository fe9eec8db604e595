"""CPU tests of the coarsest-level direct solver's symbolic analysis
(dgx_direct_plan_host, host code only): nested dissection of the interior-node
graph, frontal index sets, extend-add maps and the places of the matrix
entries in the fronts.  The plan is pinned by the mathematics: a multifrontal
solve that follows it (children's Schur complements added into the parent
through the plan's maps, each front's pivot block eliminated) must reproduce a
dense direct solve of the assembled interior matrix, for SPD element matrices
drawn at random; and each front's separator must split its subtree (no edge
between the subtrees of two children)."""
import numpy as np
import pytest

from test_gpu_solve import boundary_nodes


def element_data(oracle, pts, max_num_pts):
    import dendro4_amd as dgx  # noqa: F401
    ru = oracle.points2octree(pts, 30, max_num_pts)
    rb = oracle.balance(ru)
    rda = oracle.DA(rb)
    conn, cn, lev, mk = rda.element_view()
    meta = (mk.astype(np.uint32) | (cn.astype(np.uint32) << 8) | (lev.astype(np.uint32) << 11)).astype(np.uint16)
    l2n = rda.local_to_node
    idx = np.nonzero(l2n != 0xFFFFFFFF)[0]
    xyz = np.zeros((rda.n_nodes, 3), np.uint32)
    xyz[l2n[idx]] = rda.octs[idx, :3]
    bd = boundary_nodes(rda)
    return conn.astype(np.uint32), cn.astype(np.uint32), meta, bd, xyz


def random_element_matrices(E, seed):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((E, 8, 8))
    return np.einsum("eij,ekj->eik", B, B) + 0.5 * np.eye(8)[None]  # SPD, [e][row j][col k]


def assemble_interior(conn, cn, Me, bd):
    n = len(bd)
    E = len(conn)
    A = np.zeros((n, n))
    j = np.arange(8)
    for e in range(E):
        ids = conn[e][j ^ cn[e]]  # child-frame vertex j -> node id
        A[np.ix_(ids, ids)] += Me[e]
    inner = np.nonzero(~bd)[0]
    return A[np.ix_(inner, inner)], inner


def multifrontal_solve(P, vals, b):
    """Follows the plan: assemble fronts, extend-add, eliminate each front's
    pivot block (inverse, then Schur complement), forward and backward sweeps.
    b and the result are indexed by node id."""
    F = np.zeros(int(P["entries"]))
    nnz = len(P["a_dst"])
    off = P["a_off"].astype(np.int64)
    sums = np.add.reduceat(vals[P["a_src"]], off[:-1]) if nnz else np.zeros(0)
    sums[off[:-1] == off[1:]] = 0.0
    F[P["a_dst"].astype(np.int64)] = sums
    fr = P["front9"].astype(np.int64)
    lk = P["link3"].astype(np.int64)

    def front(f):
        o, m = fr[f, 0], fr[f, 1]
        return F[o:o + m * m].reshape(m, m)

    for f in range(len(fr)):
        m, s, ch0, nch = fr[f, 1], fr[f, 2], fr[f, 6], fr[f, 7]
        Fm = front(f)
        for l in range(ch0, ch0 + nch):
            c, fmo = lk[l, 0], lk[l, 2]
            cs, cm = fr[c, 2], fr[c, 1]
            fm = P["fmap"][fmo:fmo + cm - cs].astype(np.int64)
            Fm[np.ix_(fm, fm)] += front(c)[cs:, cs:]
        Fi = np.linalg.inv(Fm[:s, :s])
        P12 = Fi @ Fm[:s, s:]
        P21 = -Fm[s:, :s] @ Fi
        Sc = Fm[s:, s:] + P21 @ Fm[:s, s:]
        Fm[:s, :s], Fm[:s, s:], Fm[s:, :s], Fm[s:, s:] = Fi, P12, P21, Sc
    idx = P["idx"].astype(np.int64)
    upd = np.zeros(int(P["upd_len"]))
    ws = np.zeros(int(P["interior"]))
    for f in range(len(fr)):
        m, s, io, uo, wo, ch0, nch = fr[f, 1], fr[f, 2], fr[f, 3], fr[f, 4], fr[f, 5], fr[f, 6], fr[f, 7]
        w = np.zeros(m)
        w[:s] = b[idx[io:io + s]]
        for l in range(ch0, ch0 + nch):
            c, invo = lk[l, 0], lk[l, 1]
            inv = P["inv"][invo:invo + m]
            hit = inv >= 0
            w[hit] += upd[fr[c, 4] + inv[hit]]
        ws[wo:wo + s] = w[:s]
        upd[uo:uo + m - s] = w[s:] + front(f)[s:, :s] @ w[:s]
    x = np.zeros(len(b))
    for f in reversed(range(len(fr))):
        m, s, io, wo = fr[f, 1], fr[f, 2], fr[f, 3], fr[f, 5]
        v = np.concatenate([ws[wo:wo + s], -x[idx[io + s:io + m]]])
        x[idx[io:io + s]] = front(f)[:s] @ v
    return x


@pytest.mark.parametrize("n,mnp,leaf,dense", [(3000, 1, 32, 0), (6000, 4, 64, 0), (250, 1, 128, 2048)])
def test_plan_reproduces_dense_solve(oracle, n, mnp, leaf, dense):
    import dendro4_amd as dgx
    pts = dgx.gen_points(1, 40 + n, n)
    conn, cn, meta, bd, xyz = element_data(oracle, pts, mnp)
    P = dgx.direct_plan(conn, meta, bd, xyz, leaf=leaf, dense=dense)
    ni = int((~bd).sum())
    assert P["interior"] == ni
    # the separators partition the interior nodes, in elimination order
    fr = P["front9"].astype(np.int64)
    S = np.concatenate([P["idx"][fr[f, 3]:fr[f, 3] + fr[f, 2]] for f in range(len(fr))])
    assert np.array_equal(np.sort(S), np.nonzero(~bd)[0])
    assert len(set(S.tolist())) == ni
    assert all(fr[f, 5] == fr[:f, 2].sum() for f in range(len(fr)))
    if dense:
        assert ni <= dense
        assert len(fr) == 1 and fr[0, 1] == fr[0, 2] == ni
    else:
        assert len(fr) > 4 and fr[:, 8].max() >= 2
    # numeric check with random SPD element matrices
    E = len(conn)
    Me = random_element_matrices(E, n)
    vals = np.transpose(Me, (0, 2, 1)).reshape(-1)  # vals[e*64 + k*8 + j] = Me[e][j][k]
    A, inner = assemble_interior(conn, cn, Me, bd)
    b = np.random.default_rng(1).standard_normal(len(bd))
    x = multifrontal_solve(P, vals, b)
    xr = np.linalg.solve(A, b[inner])
    assert np.abs(x[inner] - xr).max() <= 1e-9 * np.abs(xr).max()


def test_separators_split_their_subtrees(oracle):
    """No matrix entry joins two different children's subtrees of a front."""
    import dendro4_amd as dgx
    pts = dgx.gen_points(2, 7, 5000, n_gauss=4, lo=0.2, hi=0.8, sig_lo=0.05)
    conn, cn, meta, bd, xyz = element_data(oracle, pts, 1)
    P = dgx.direct_plan(conn, meta, bd, xyz, leaf=48, dense=0)
    fr = P["front9"].astype(np.int64)
    lk = P["link3"].astype(np.int64)
    nf = len(fr)
    owner = np.full(len(bd), -1)
    for f in range(nf):
        owner[P["idx"][fr[f, 3]:fr[f, 3] + fr[f, 2]]] = f
    parent = np.full(nf, -1)
    for f in range(nf):
        for l in range(fr[f, 6], fr[f, 6] + fr[f, 7]):
            parent[lk[l, 0]] = f
    # subtree root just below the common ancestor, for each pair of adjacent nodes
    def chain(f):
        out = [f]
        while parent[out[-1]] >= 0:
            out.append(parent[out[-1]])
        return out
    chains = [chain(f) for f in range(nf)]
    j = np.arange(8)
    for e in range(len(conn)):
        ids = [i for i in conn[e][j ^ cn[e]] if not bd[i]]
        fs = {owner[i] for i in ids}
        for a in fs:
            for b in fs:
                # one of the two fronts is an ancestor of (or equal to) the other
                assert a in chains[b] or b in chains[a], (a, b)
