"""Legacy Libra metadata (the format of TR-source/SpMM/mGCNkernel.cu's host
functions) written from a CSR, for tests.  The reference holds no producer of
this format; this encoder follows the format as the reference's kernels read
it (oracle.decode_libra states that reading):
  * 8-row windows; in each, the columns with at least `density` rows become
    TC blocks of 8 columns (ascending, the last one padded with -1), one TC
    part per window that has any (t_window_row = window index);
  * the other nonzeros go to residual parts per row, in CSR order: rows of at
    most `short` nonzeros to the short list, rows of at most `partsize` to one
    long part, longer rows to parts of `partsize` and a remainder (short when
    it has at most `short` nonzeros);
  * values fp16, masks bit 8 n + k for row n and column k of a block."""
import numpy as np
import torch


def encode(csr, density=4, partsize=32, short=3, tf32=False):
    """tf32: the tf32 members' format (blocks of 4 columns, int32 masks whose bit
    4 n + k is row n, column k; fp32 values)."""
    rowptr, col, val = csr.rowptr.astype(np.int64), csr.col.astype(np.int64), csr.val.astype(np.float32)
    m = rowptr.size - 1
    W = (m + 7) // 8
    t_row_offset, t_block, t_value, t_column, t_window_row, t_binary = [0], [0], [], [], [], []
    tc_set = set()                          # (row, col) taken by TC blocks
    for w in range(W):
        r0, r1 = 8 * w, min(8 * w + 8, m)
        cnt = {}
        for r in range(r0, r1):
            for e in range(rowptr[r], rowptr[r + 1]):
                cnt[col[e]] = cnt.get(col[e], 0) + 1
        tcc = sorted(c for c, n in cnt.items() if n >= density)
        if not tcc:
            continue
        width = 4 if tf32 else 8
        for i in range(0, len(tcc), width):
            chunk = tcc[i: i + width]
            pos = {c: k for k, c in enumerate(chunk)}
            bits = 0
            entries = []
            for r in range(r0, r1):
                for e in range(rowptr[r], rowptr[r + 1]):
                    if col[e] in pos:
                        b = width * (r - r0) + pos[col[e]]
                        bits |= 1 << b
                        entries.append((b, val[e]))
                        tc_set.add((r, int(col[e])))
            entries.sort()
            t_value.extend(v for _, v in entries)
            t_block.append(len(t_value))
            t_column.extend(chunk + [-1] * (width - len(chunk)))
            top = 8 * width
            t_binary.append(bits - (1 << top) if bits >= (1 << (top - 1)) else bits)
        t_row_offset.append(len(t_binary))
        t_window_row.append(w)
    longp = ([0], [], [], [])               # offsets, rows, columns, values
    shortp = ([0], [], [], [])

    def put(lst, r, es):
        lst[1].append(r)
        lst[2].extend(int(col[e]) for e in es)
        lst[3].extend(float(val[e]) for e in es)
        lst[0].append(len(lst[2]))

    for r in range(m):
        es = [e for e in range(rowptr[r], rowptr[r + 1]) if (r, int(col[e])) not in tc_set]
        if not es:
            continue
        if len(es) <= short:
            put(shortp, r, es)
        elif len(es) <= partsize:
            put(longp, r, es)
        else:
            full = len(es) // partsize * partsize
            for i in range(0, full, partsize):
                put(longp, r, es[i: i + partsize])
            rest = es[full:]
            if rest:
                put(shortp if len(rest) <= short else longp, r, rest)
    i32 = lambda x: torch.tensor(x, dtype=torch.int32)     # noqa: E731
    f16 = ((lambda x: torch.tensor(np.asarray(x, np.float32))) if tf32       # noqa: E731
           else (lambda x: torch.tensor(np.asarray(x, np.float32)).half()))
    meta = {
        "t_row_offset": i32(t_row_offset), "t_blockNew_offset": i32(t_block), "t_value": f16(t_value),
        "t_column": i32(t_column), "t_window_row": i32(t_window_row), "t_atomic": i32([0] * len(t_window_row)),
        "t_binary": torch.tensor(t_binary, dtype=torch.int32 if tf32 else torch.int64),
        "c_row_offset": i32(longp[0]), "c_row": i32(longp[1]), "c_atomic": i32([0] * len(longp[1])),
        "c_column": i32(longp[2]), "c_value": f16(longp[3]),
        "c_row_offset_short": i32(shortp[0]), "c_row_short": i32(shortp[1]),
        "c_atomic_short": i32([0] * len(shortp[1])), "c_column_short": i32(shortp[2]), "c_value_short": f16(shortp[3]),
    }
    counts = {"parts_t": len(t_window_row), "parts_c": len(longp[1]), "parts_c_short": len(shortp[1])}
    return meta, counts


def decode(meta):
    import oracle
    g = lambda k: meta[k].numpy()          # noqa: E731
    return oracle.decode_libra(g("t_row_offset"), g("t_blockNew_offset"), meta["t_value"].float().numpy(),
                               g("t_column"), g("t_window_row"), g("t_binary"),
                               g("c_row_offset"), g("c_row"), g("c_column"), meta["c_value"].float().numpy(),
                               g("c_row_offset_short"), g("c_row_short"), g("c_column_short"),
                               meta["c_value_short"].float().numpy(), tf32=meta["t_binary"].dtype == torch.int32)
