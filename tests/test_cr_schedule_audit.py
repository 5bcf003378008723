"""Address audit of the cyclic-reduction kernels on the CPU.

The schedule the runtime builds once per mesh (build_cr_schedule in
colloc_runtime.cpp) is restated here, and every buffer access of
colloc_cr_pair / colloc_cr_subst / colloc_cr_final / colloc_cr_back (the
element ranges their index expressions produce: input rows 2k and 2k+1 of
the segment, output row k, record k, the carried odd row, and the nodes
left / mid / right of the back substitution) is checked against the
allocations cs_create makes (rows0: rows_cap = 2 Nint + 32 B rows of
NY x CRW; crrec: Nint records of NY x RECW; dy: the instance's nodes), for
batches of uneven meshes (instances of one interval included) and every
pairing rule.

The float32 cyclic reduction indexes the same elements in the same buffers
(typed as float): an element range [e0, e1) that fits n doubles is at
bytes [4 e0, 4 e1) <= 8 n, so the check of the element ranges covers both
stored types; the test also checks it in bytes for both.
"""
import numpy as np
import pytest

NY, NP = 40, 40
CRW = 2 * NY + NP + 1
RECW = 3 * NY + NP + 1


def schedule(nodes, subst_spans):
    """The runtime's schedule: per level, the instance of each output row and
    per instance (in_off, out_off, rec_off, nrows, subst); the row base of
    each level's input region."""
    B = len(nodes)
    n = [k - 1 for k in nodes]
    in_off, acc = [], 0
    for k in nodes:
        in_off.append(acc)
        acc += k - 1
    nsub = [0] * B
    if subst_spans > 0:
        for b in range(B):
            while (subst_spans << (nsub[b] + 1)) <= n[b]:
                nsub[b] += 1
    levels = []
    rbase = [0]
    nin = sum(n)
    rec_base = 0
    lev = 0
    while any(v > 1 for v in n):
        oinst, segs, out = [], [], 0
        for b in range(B):
            nout = n[b] // 2 + (n[b] & 1)
            sub = (2 if lev + 1 < nsub[b] else 1) if lev < nsub[b] else 0
            segs.append((in_off[b], out, rec_base, n[b], sub))
            rec_base += n[b] // 2
            oinst += [b] * nout
            in_off[b] = out
            out += nout
            n[b] = nout
        levels.append((oinst, segs, out))
        rbase.append(rbase[-1] + nin)
        nin = out
        lev += 1
    return levels, rbase, nin, rec_base


def audit(nodes, subst_spans):
    B = len(nodes)
    nint = sum(k - 1 for k in nodes)
    noff = np.concatenate([[0], np.cumsum(nodes)])
    rows_cap = 2 * nint + 32 * B
    levels, rbase, nlast, nrec = schedule(nodes, subst_spans)
    assert rbase[-1] + nlast <= rows_cap            # the runtime rejects the mesh otherwise
    assert nrec <= nint
    rows_elems = rows_cap * NY * CRW
    rec_elems = nint * NY * RECW
    in_rows = [nint] + [lv[2] for lv in levels]     # rows of each level's input region
    spans = []                                      # (buffer, first element, end element)
    for l, (oinst, segs, nout) in enumerate(levels):
        base_in = rbase[l] * NY * CRW
        base_out = rbase[l + 1] * NY * CRW
        assert rbase[l] + in_rows[l] <= rbase[l + 1]  # regions do not overlap
        for g in range(nout):
            b = oinst[g]
            io, oo, ro, nrows, sub = segs[b]
            k = g - oo
            npairs = nrows >> 1
            assert 0 <= k <= npairs and (k < npairs or nrows & 1)
            assert oo + k < nout
            if k == npairs:                          # carried odd row
                r = io + nrows - 1
                assert r < in_rows[l]
                spans.append(("rows", base_in + r * NY * CRW, base_in + (r + 1) * NY * CRW))
                spans.append(("rows", base_out + (oo + k) * NY * CRW, base_out + (oo + k + 1) * NY * CRW))
                continue
            ra = io + 2 * k
            assert ra + 1 < io + nrows and ra + 1 < in_rows[l]
            spans.append(("rows", base_in + ra * NY * CRW, base_in + (ra + 2) * NY * CRW))
            spans.append(("rows", base_out + (oo + k) * NY * CRW, base_out + (oo + k + 1) * NY * CRW))
            if not sub:
                spans.append(("rec", (ro + k) * NY * RECW, (ro + k + 1) * NY * RECW))
            # cr_back on this pair: nodes of the instance
            base = noff[b]
            ni = noff[b + 1] - base - 1
            mid = base + ((2 * k + 1) << l)
            left = base + ((2 * k) << l)
            right = base + min((2 * k + 2) << l, ni)
            assert noff[b] <= left < mid < right <= noff[b + 1] - 1, (l, b, k, left, mid, right)
    # cr_final: row b of the last output region (or of the condensation's rows)
    base_last = rbase[-1] * NY * CRW
    for b in range(B):
        spans.append(("rows", base_last + b * NY * CRW, base_last + (b + 1) * NY * CRW))
        assert b < nlast
    cap = {"rows": rows_elems, "rec": rec_elems}
    for buf, e0, e1 in spans:
        assert 0 <= e0 < e1 <= cap[buf], (buf, e0, e1, cap[buf])
        for size in (8, 4):                         # float64 and float32 storage
            assert e1 * size <= cap[buf] * 8
    return len(levels)


@pytest.mark.parametrize("spans", [0, 1, 16, 64])
def test_cr_schedule_accesses_in_bounds(spans):
    rng = np.random.default_rng(spans + 11)
    shapes = [[2], [3], [2, 2, 2], [4097], [257, 2, 1025, 3, 2049]]
    for _ in range(20):
        B = int(rng.integers(1, 12))
        shapes.append(list(rng.integers(2, 2100, B)))
    for nodes in shapes:
        nodes = [int(v) for v in nodes]
        nlev = audit(nodes, spans)
        assert nlev == int(np.ceil(np.log2(max(v - 1 for v in nodes)))) if max(nodes) > 2 else nlev == 0
