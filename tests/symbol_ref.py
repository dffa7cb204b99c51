"""Pure-Python restatement of the intra-supernode row reordering, for tests.

Follows the reference's symbol/symbol_reordering.c: the depth of every cblk
in the elimination tree (compute_cblklevel, :47-60), the split level moved
until the upper part holds a sixth to a quarter of the contributions
(:383-418), the upper / lower lists of contributing cblks of every row
(:420-470), and the cheapest insertion into a cycle through a ghost row with
the cut tie-break (symbol_reorder_tsp, :214-345).  Distances are set
differences of Python sets (min(|A ^ B|, stop)), not a merge of sorted
lists, so an unsorted list or a miscounted merge in the C code would not be
repeated here.  Small supernodes only: the insertion is O(rows^2) Python.
"""


def _levels(tree):
    """depth of each cblk, the root(s) at 1 (reference :47-60)."""
    nc = len(tree)
    lev = [0] * nc
    for c in range(nc):
        d, f = 1, tree[c]
        while f != -1:
            d += 1
            f = tree[f]
        lev[c] = d
    return lev


def _dist(vecs, i, j, stop):
    if i == -1:
        return len(vecs[j])
    if j == -1:
        return len(vecs[i])
    return min(len(vecs[i] ^ vecs[j]), stop)


def _tsp(size, lw, up, stop):
    """Order of the rows 0..size-1 (reference symbol_reorder_tsp)."""
    if size < 3:
        return list(range(size))
    inv = [-1, 0] + [None] * (size - 1)
    ln = [0] * (size + 1)
    d = _dist(lw, 0, -1, stop)
    ln[0] = ln[1] = d
    for i in range(1, size):
        lwb = _dist(lw, i, inv[0], stop)
        lwa = _dist(lw, i, inv[1], stop)
        upa = _dist(up, i, inv[1], 1)
        minl = lwb + lwa - ln[0]
        mpos = 1
        mincut = -1
        j = 1
        while j < i:
            upb = upa
            upa = _dist(up, i, inv[j + 1], 1)
            if upb < 1 or upa < 1:
                lwb = _dist(lw, i, inv[j], stop) if lwa == -1 else lwa
                lwa = _dist(lw, i, inv[j + 1], stop)
                l = lwb + lwa - ln[j]
                if l == minl:
                    if lwb < mincut:
                        mincut, minl, mpos = lwb, l, j + 1
                    if lwa < mincut:
                        mincut, minl, mpos = lwa, l, j + 1
                if l < minl:
                    mincut, minl, mpos = lwb, l, j + 1
                    if lwa < mincut:
                        mincut = lwa
                if lwa == 0:
                    mincut, minl, mpos = 0, l, j + 1
                    j = i
            else:
                lwa = -1
            j += 1
        first = _dist(lw, i, inv[0], stop)
        last = _dist(lw, i, inv[i], stop)
        lwb = _dist(lw, i, inv[mpos - 1], stop)
        lwa = _dist(lw, i, inv[mpos], stop)
        l = first + last - ln[i]
        if l < minl:
            minl, mpos = l, i + 1
        if mpos > 0:
            ln[mpos - 1] = lwb
        if mpos < i + 1:
            # insert row i at mpos, shifting the tail by one
            inv[mpos:i + 2] = [i] + inv[mpos:i + 1]
            ln[mpos:i + 2] = [lwa] + ln[mpos:i + 1]
        else:
            inv[i + 1] = i
            ln[i + 1] = first
    elected = 0
    for k in range(size):
        if inv[k] == -1:
            elected = k
    return [inv[(k + 1 + elected) % (size + 1)] for k in range(size)]


def reorder(cblk, blok, rang, tree, peri, split_level=0, stop=2 ** 62):
    """New peritab after the reordering of every supernode.

    cblk: (cblknbr + 1, 3) fcolnum, lcolnum, bloknum of the symbol matrix
    before the reordering; blok: (bloknbr, 4) frownum, lrownum, lcblknm,
    fcblknm; rang, tree, peri: the ordering."""
    nc = len(tree)
    lev = _levels(list(tree))
    maxdepth = max(lev) if nc else 0
    # bloks facing each cblk, off-diagonal ones, in order of their cblk
    facing = [[] for _ in range(nc)]
    for c in range(nc):
        for b in range(cblk[c][2] + 1, cblk[c + 1][2]):
            facing[blok[b][3]].append(b)
    peri = list(peri)
    for c in range(nc):
        f, l = cblk[c][0], cblk[c][1]
        size = l - f + 1
        dw = [0] * maxdepth
        for b in facing[c]:
            dw[lev[blok[b][2]] - 1] += blok[b][1] - blok[b][0] + 1
        split, sign = split_level, 0
        while True:
            up_t = sum(dw[:max(split, 0)])
            lw_t = sum(dw[max(split, 0):])
            if lw_t < 5 * up_t and lw_t > 10 and up_t > 10 and sign <= 0:
                split -= 1
                sign -= 1
                continue
            if lw_t > 3 * up_t and lw_t > 10 and up_t > 10 and sign >= 0:
                split += 1
                sign += 1
                continue
            break
        lw = [set() for _ in range(size)]
        up = [set() for _ in range(size)]
        for b in facing[c]:
            k = blok[b][2]
            tgt = up if lev[k] <= split else lw
            for r in range(blok[b][0], blok[b][1] + 1):
                tgt[r - f].add(k)
        order = _tsp(size, lw, up, stop)
        seg = peri[rang[c]:rang[c] + size]
        peri[rang[c]:rang[c] + size] = [seg[o] for o in order]
    return peri
