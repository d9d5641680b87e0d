"""ORACLE -- test infrastructure only.

Pure-Python restatement of the reference's index-building path, for small
meshes (loops over cells, no vectorisation):
  convGraph + sortGraphCsr      reference src/Graph.c:24-124
  genrcm / root_find / level_set / rcm / degree / ivec_reverse
                                reference src/Rcm.c:56-887 (1-based there,
                                0-based here, same visiting and tie-break
                                rules)
  reord                         reference src/Reord.c:18-112
  numEqV1 / numEqV2             reference src/Numeq.c:70-135
  csrIa / csrJa + sort          reference src/Csr.c:30-96, 202-281
  ellPackJa                     reference src/EllPack.c:26-79
  csr / ellPack value scatter   reference src/Csr.c:549-975,
                                src/EllPack.c:113-149
Pinned against the reference's MatrixMarket fixtures
(test/output/coo/retangulo*_bin.mtx) and, where it is built, the
reference's own compiled code.
"""

CSR, CSRD, CSRC, ELLPACK, CSRDCOO, CSRCCOO = 1, 2, 3, 4, 5, 6


def cell_graph(nelcon, nviz):
    """convGraph + sortGraphCsr: 0-based neighbour lists, ascending."""
    adj = []
    for c in range(len(nviz)):
        row = [int(nelcon[c][f]) - 1 for f in range(int(nviz[c]))
               if int(nelcon[c][f]) != -1]
        adj.append(sorted(row))
    return adj


def _level_set(root, adj, mask):
    ls = [root]
    mask[root] = 0
    starts = []
    begin = 0
    while True:
        end = len(ls)
        starts.append(begin)
        for i in range(begin, end):
            for w in adj[ls[i]]:
                if mask[w]:
                    mask[w] = 0
                    ls.append(w)
        begin = end
        if len(ls) == end:
            break
    for v in ls:
        mask[v] = 1
    return ls, starts


def _root_find(root, adj, mask):
    ls, starts = _level_set(root, adj, mask)
    nlev, size = len(starts), len(ls)
    if nlev == 1 or nlev == size:
        return root
    while True:
        last = ls[starts[nlev - 1]:]
        best, bestdeg = last[0], size
        for v in last:
            d = sum(1 for w in adj[v] if mask[w] > 0)
            if d < bestdeg:
                best, bestdeg = v, d
        root = best
        ls2, starts2 = _level_set(root, adj, mask)
        if len(starts2) <= nlev:
            break
        nlev = len(starts2)
        if size <= nlev:
            break
    return root


def _rcm(root, adj, mask):
    # degrees over unnumbered nodes of the component (BFS from root)
    order, seen = [root], {root}
    deg = {}
    i = 0
    while i < len(order):
        v = order[i]
        d = 0
        for w in adj[v]:
            if mask[w]:
                d += 1
                if w not in seen:
                    seen.add(w)
                    order.append(w)
        deg[v] = d
        i += 1
    mask[root] = 0
    perm = [root]
    head = 0
    while head < len(perm):
        v = perm[head]
        kids = []
        for w in adj[v]:
            if mask[w]:
                mask[w] = 0
                kids.append(w)
        kids.sort(key=lambda w: deg[w])          # Python's sort is stable
        perm.extend(kids)
        head += 1
    perm.reverse()
    return perm


def genrcm(adj):
    """0-based reverse Cuthill-McKee permutation."""
    n = len(adj)
    mask = [1] * n
    perm = []
    for v in range(n):
        if mask[v]:
            root = _root_find(v, adj, mask)
            perm.extend(_rcm(root, adj, mask))
    return perm


def reord(nelcon, nviz, flag):
    n = len(nviz)
    if not flag:
        return [i + 1 for i in range(n)]
    return [v + 1 for v in genrcm(cell_graph(nelcon, nviz))]


def numeq_v1(num):
    idv = [0] * len(num)
    for i, c in enumerate(num):
        idv[c - 1] = i + 1
    return idv, len(num)


def numeq_v2(num, pinned):
    """pinned: set of 0-based cells whose restriction flag is set."""
    idv, neq = [0] * len(num), 0
    for c in num:
        if c - 1 in pinned:
            idv[c - 1] = -1
        else:
            neq += 1
            idv[c - 1] = neq
    return idv, neq


def _flags(storage, unsym):
    if storage == CSR:
        return (bool(unsym), True, True)            # lower, upper, diag
    if storage in (CSRD, CSRDCOO):
        return (True, bool(unsym), False)
    return (True, False, False)                     # CSRC / CSRCCOO


def datastruct(storage, unsym, idv, num, nelcon, nviz, max_viz, neq):
    """(ia, ja, nad) in the reference's layout."""
    if storage == ELLPACK:
        ja = [0] * ((neq + 1) * max_viz)
        nad = 0
        for c1 in num:
            c = c1 - 1
            row = idv[c] - 1
            if row < 0:
                continue
            for f in range(int(nviz[c])):
                nb = int(nelcon[c][f])
                if nb == -1:
                    ja[row * max_viz + f] = row
                    continue
                col = idv[nb - 1] - 1
                if col >= 0:
                    ja[row * max_viz + f] = col
                    nad += 1
        return [max_viz, (neq + 1) * max_viz], ja, nad
    lower, upper, diag = _flags(storage, unsym)
    rows = [None] * neq
    for c1 in num:
        c = c1 - 1
        row = idv[c] - 1
        if row < 0:
            continue
        cols = [row] if diag else []
        for f in range(int(nviz[c])):
            nb = int(nelcon[c][f])
            if nb == -1:
                continue
            col = idv[nb - 1] - 1
            if col < 0:
                continue
            if (lower and col < row) or (upper and col > row):
                cols.append(col)
        rows[row] = sorted(cols)
    ia, ja = [0], []
    for r in rows:
        ja.extend(r)
        ia.append(len(ja))
    return ia, ja, len(ja)


def assemble(storage, unsym, ia, ja, nad, idv, nelcon, nviz, max_viz, lA, lB,
             neq):
    """csr/ellPack scatter; lA[c][f] face coefficients, lA[c][max_viz] the
    diagonal, lB[c] the right-hand side.  Returns (al, ad, b)."""
    nrow = neq + 1 if storage == ELLPACK else neq
    if storage == ELLPACK:
        al = [0.0] * ((neq + 1) * max_viz)
    elif storage in (CSRC, CSRCCOO) and unsym:
        al = [0.0] * (2 * nad)
    else:
        al = [0.0] * nad
    ad = [0.0] * nrow
    b = [0.0] * nrow
    for c in range(len(nviz)):
        row = idv[c] - 1
        if row < 0:
            continue
        b[row] = lB[c]
        if storage == ELLPACK:
            ad[row] = lA[c][max_viz]
            for f in range(int(nviz[c])):
                al[row * max_viz + f] = lA[c][f]
            continue
        if storage == CSR:
            for k in range(ia[row], ia[row + 1]):
                if ja[k] == row:
                    al[k] = lA[c][max_viz]
        else:
            ad[row] = lA[c][max_viz]
        for f in range(int(nviz[c])):
            nb = int(nelcon[c][f])
            if nb == -1:
                continue
            col = idv[nb - 1] - 1
            if col < 0:
                continue
            if storage in (CSRC, CSRCCOO) and unsym and col > row:
                for k in range(ia[col], ia[col + 1]):
                    if ja[k] == row:
                        al[nad + k] = lA[c][f]
            else:
                for k in range(ia[row], ia[row + 1]):
                    if ja[k] == col:
                        al[k] = lA[c][f]
    return al, ad, b


def csrd_to_coo_pattern(ia, ja, neq):
    """csrToCoo (reference src/WriteMtx.c:113-147) pattern: 1-based (row,
    col) pairs, diagonal first in each row, then the row's stored entries."""
    out = []
    for i in range(neq):
        out.append((i + 1, i + 1))
        for k in range(ia[i], ia[i + 1]):
            out.append((i + 1, ja[k] + 1))
    return out
