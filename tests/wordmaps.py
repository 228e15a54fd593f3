"""Decode instruction words into operand byte maps (the instruction layout
of the reference decoder, src/kernel.cu:1384-1522) -- test helper."""

ELEM = 8


def slot_base(ptr):
    return (ptr + 1) << 44


def _map(ptr, off, rows, cols, ld):
    return (slot_base(ptr) + off * ELEM, ld * ELEM, cols * ELEM, rows)


def operand_maps(w):
    """[(read_only, (ptr, pitch, width, height)), ...] of one execute entry."""
    op = int(w[2])
    w = [int(x) for x in w]
    if op == 1:  # getrf
        return [(False, _map(w[3], w[4], w[6], w[5], w[7]))]
    if op in (2, 3):  # trsml / trsmr: logical B ny_b x nx_b, stored transposed if b_T
        nx_b, ny_b = w[7], w[8]
        r, c = (nx_b, ny_b) if w[12] else (ny_b, nx_b)
        k = w[9]
        return [(False, _map(w[3], w[5], r, c, w[10])), (True, _map(w[4], w[6], k, k, w[11]))]
    if op == 9:  # trsmu
        return [(False, _map(w[3], w[5], w[8], w[7], w[10])), (True, _map(w[4], w[6], w[9], w[9], w[11]))]
    if op == 10:  # trmm
        return [(False, _map(w[3], w[5], w[8], w[7], w[9])), (True, _map(w[4], w[6], w[8], w[8], w[10]))]
    if op in (4, 5):  # gemm
        m, n, k = w[9], w[10], w[11]
        a = (k, m) if w[15] else (m, k)
        b = (n, k) if w[16] else (k, n)
        return [(False, _map(w[3], w[6], m, n, w[12])), (True, _map(w[4], w[7], *a, w[13])),
                (True, _map(w[5], w[8], *b, w[14]))]
    if op == 6:  # gemm_3x
        m, n, k, l = w[11], w[12], w[13], w[14]
        dims = [(m, k, w[19]), (k, l, w[20]), (l, n, w[21])]
        out = [(False, _map(w[3], w[7], m, n, w[15]))]
        for i, (r, c, t) in enumerate(dims):
            out.append((True, _map(w[4 + i], w[8 + i], *((c, r) if t else (r, c)), w[16 + i])))
        return out
    if op == 7:  # gemm_4x
        m, n, k, l, o = w[13], w[14], w[15], w[16], w[17]
        dims = [(m, k, w[23]), (k, l, w[24]), (l, o, w[25]), (o, n, w[26])]
        out = [(False, _map(w[3], w[8], m, n, w[18]))]
        for i, (r, c, t) in enumerate(dims):
            out.append((True, _map(w[4 + i], w[9 + i], *((c, r) if t else (r, c)), w[19 + i])))
        return out
    if op == 8:  # accum
        nx, ny, k1, k2 = w[11], w[12], w[13], w[14]
        return [(False, _map(w[3], w[7], ny, k1, w[15])), (False, _map(w[4], w[8], nx, k1, w[16])),
                (True, _map(w[5], w[9], ny, k2, w[17])), (True, _map(w[6], w[10], nx, k2, w[18]))]
    if op == 0:
        return []
    raise ValueError(f"unknown operation {op}")


def entry_length(op):
    return {0: 3, 1: 8, 2: 13, 3: 13, 4: 17, 5: 17, 6: 23, 7: 29, 8: 21, 9: 13, 10: 11}[op]


def split_stream(stream):
    """[(kind, words)] entries of one worker stream; kind in exec / wait / finish."""
    out, i = [], 0
    while i < len(stream):
        op = int(stream[i])
        if op == 0:
            out.append(("finish", stream[i:i + 1]))
            break
        if op == 2:
            out.append(("wait", stream[i:i + 3]))
            i += 3
        elif op == 1:
            n = entry_length(int(stream[i + 2]))
            out.append(("exec", stream[i:i + n]))
            i += n
        else:
            raise ValueError(f"bad opcode {op} at {i}")
    return out
