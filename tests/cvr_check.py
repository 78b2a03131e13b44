"""Independent checks of a CVR-64 layout (test infrastructure).

reconstruct() walks a layout the way the SpMV kernel consumes it -- values and
columns step by step, segment ends from the column sign bit, destinations from
the segment list, split slots through the split table -- and rebuilds every
row's multiset of (column, value). A correct layout rebuilds the CSR exactly,
so this pins the conversion without restating it.
"""
from __future__ import annotations

import bisect

import numpy as np

SPLIT_CAP = 72
SEG_EXTRA = 66


def row_of(rp, k):
    """Row holding CSR element k."""
    return bisect.bisect_right(rp, k) - 1


def reconstruct(A, lay, S):
    rp = [int(v) for v in A.rowOffsets]
    m, nnz = A.numRows, A.nnz
    nc = int(lay["nchunks"]) if "nchunks" in lay else int(lay["info"]["nchunks"])
    L = int(lay["layout_elems"]) if "layout_elems" in lay else int(lay["info"]["layout_elems"])
    vals = np.asarray(lay["vals"])
    if vals.dtype == np.uint16:
        vals = (vals.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    vals = vals.astype(np.float64)
    cols = np.asarray(lay["cols"]).astype(np.int64)
    seg, chunk_rows = lay["seg"], lay["chunk_rows"]
    split_n, split_tab = lay["split_n"], lay["split_tab"]
    C = 64 * S
    rows = [[] for _ in range(m)]
    errors = []
    for c in range(nc):
        base = c * C
        steps = S if c < nc - 1 else (L - base) // 64
        b, e = c * C, min((c + 1) * C, nnz)
        head_row, tail_row = row_of(rp, b), row_of(rp, e - 1)
        segbase = int(chunk_rows[c]) + SEG_EXTRA * c
        tab = split_tab[c * 2 * SPLIT_CAP:(c + 1) * 2 * SPLIT_CAP]
        slot_row = {}
        for g in range(int(split_n[c])):
            code, sb, se = int(tab[2 * g]), int(tab[2 * g + 1]), int(tab[2 * g + 3])
            r = code if code >= 0 else head_row if code == -1 else tail_row if code == -2 else None
            for s in range(sb, se):
                slot_row[s] = r
        cur = [[] for _ in range(64)]
        segp = 0
        for s in range(steps):
            o = base + s * 64
            cc = cols[o:o + 64]
            vv = vals[o:o + 64]
            for l in range(64):
                cur[l].append((int(cc[l] & 0x7FFFFFFF), float(vv[l])))
            ends = np.nonzero(cc < 0)[0]
            for k, l in enumerate(ends):
                d = int(seg[segbase + segp + k])
                if d >= 0:
                    r = d
                elif ~d in slot_row:
                    r = slot_row[~d]
                else:
                    errors.append(f"chunk {c}: slot {~d} without a split group")
                    r = None
                if r is None:
                    if any(v != 0.0 for _, v in cur[l]):
                        errors.append(f"chunk {c}: discarded segment holds nonzero values")
                else:
                    rows[r].extend(cur[l])
                cur[l] = []
            segp += len(ends)
            if s == steps - 1 and len(ends) != 64:
                errors.append(f"chunk {c}: {len(ends)} lanes end on the last step")
        if "seg_count" in lay and int(lay["seg_count"][c]) != segp:
            errors.append(f"chunk {c}: seg_count {int(lay['seg_count'][c])} != {segp}")
    for r in range(m):
        want = sorted(zip((int(v) for v in A.cols[rp[r]:rp[r + 1]]), (float(v) for v in A.vals[rp[r]:rp[r + 1]])))
        got = sorted(rows[r])
        if lay.get("prec", 0) == 0 and vals.dtype == np.float64:
            if got != want:
                errors.append(f"row {r}: rebuilt {len(got)} entries, CSR has {len(want)}")
        else:
            if [c for c, _ in got] != [c for c, _ in want]:
                errors.append(f"row {r}: column multiset differs")
        if len(errors) > 20:
            break
    return errors
