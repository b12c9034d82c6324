"""Brute-force compacted-dBG oracle for tiny inputs (pure Python).

TEST INFRASTRUCTURE ONLY. Written from the semantics of the reference's DFA
(classify.rs:147-518) and unitig rules (output.rs:79-145), not from its branch
code: every valid k-mer occurrence contributes an observation (in-label,
out-label) in canonical orientation, the per-side label sets collapse to
single/multi, and two consecutive k-mers of a record belong to one unitig iff
the exit side of the first and the entry side of the second are both single.

State codes follow state.rs:92-113: unvisited 0, MIMO 3, MISO 4+back,
SIMO 8+front, SISO 16+4*front+back.
"""
ACGT = {"A": 0, "C": 1, "G": 2, "T": 3, "a": 0, "c": 1, "g": 2, "t": 3}
COMP = "TGCA"
MULTI = "*"


def rc(s):
    return "".join(COMP[ACGT[c]] for c in reversed(s))


def canon(s):
    u = s.upper()
    r = rc(u)
    return u if u <= r else r


def valid_positions(seq, k):
    out = []
    for i in range(0, len(seq) - k + 1):
        if all(c in ACGT for c in seq[i:i + k]):
            out.append(i)
    return out


def observations(seq, k):
    """Yield (canonical k-mer, in-label, out-label) per valid occurrence."""
    n = len(seq)
    for i in valid_positions(seq, k):
        x = seq[i:i + k].upper()
        c = canon(x)
        fwd = x == c
        p = seq[i - 1].upper() if i > 0 and seq[i - 1] in ACGT else None
        nx = seq[i + k].upper() if i + k < n and seq[i + k] in ACGT else None
        if p is None and nx is None:
            yield c, MULTI, MULTI  # isolated -> MIMO
            continue
        if nx is not None:
            y = seq[i + 1:i + 1 + k].upper()
            if canon(y) == c:  # self loop
                if p is None or x == y:
                    yield c, MULTI, MULTI
                    continue
                nx = None  # inverted repeat: handled as rightmost (prev only)
        # canonical orientation: FWD in=p, out=n; BWD in=comp(n), out=comp(p)
        if fwd:
            a = p if p is not None else MULTI
            b = nx if nx is not None else MULTI
        else:
            a = COMP[ACGT[nx]] if nx is not None else MULTI
            b = COMP[ACGT[p]] if p is not None else MULTI
        yield c, a, b


def states(seqs, k):
    """canonical k-mer -> 5-bit state code."""
    ins, outs = {}, {}
    for s in seqs:
        for c, a, b in observations(s, k):
            ins.setdefault(c, set()).add(a)
            outs.setdefault(c, set()).add(b)
    res = {}
    for c in ins:
        I, O = ins[c], outs[c]
        si = len(I) == 1 and MULTI not in I
        so = len(O) == 1 and MULTI not in O
        if si and so:
            res[c] = 16 + 4 * ACGT[next(iter(I))] + ACGT[next(iter(O))]
        elif so:
            res[c] = 4 + ACGT[next(iter(O))]
        elif si:
            res[c] = 8 + ACGT[next(iter(I))]
        else:
            res[c] = 3
    return res


def state_class(code):
    if code == 3:
        return "MIMO"
    if 4 <= code < 8:
        return "MISO"
    if 8 <= code < 12:
        return "SIMO"
    return "SISO"


def _entry_single(code, fwd):
    """Is the side through which a traversal enters this k-mer single?"""
    cls = state_class(code)
    if cls == "MIMO":
        return False
    if fwd:
        return cls in ("SISO", "SIMO")  # in-side single
    return cls in ("SISO", "MISO")      # traversing rc: enter through out-side


def _exit_single(code, fwd):
    cls = state_class(code)
    if cls == "MIMO":
        return False
    if fwd:
        return cls in ("SISO", "MISO")
    return cls in ("SISO", "SIMO")


def unitigs(seqs, k):
    """Maximal unitigs as (set of canonical unitig strings, per-record tilings).

    A tiling is a list of (canonical unitig string, '+'/'-', start, end) with
    k-mer start/end indices inside the record; gaps are runs of invalid k-mers.
    """
    st = states(seqs, k)
    segs = set()
    tilings = []
    for s in seqs:
        til = []
        pos = valid_positions(s, k)
        run = []
        prev = None
        runs = []
        for i in pos:
            if prev is not None and i == prev + 1:
                run.append(i)
            else:
                if run:
                    runs.append(run)
                run = [i]
            prev = i
        if run:
            runs.append(run)
        for run in runs:
            cur = [run[0]]
            for i in run[1:]:
                x = s[i - 1:i - 1 + k].upper()
                y = s[i:i + k].upper()
                cx, cy = canon(x), canon(y)
                if _exit_single(st[cx], x == cx) and _entry_single(st[cy], y == cy):
                    cur.append(i)
                else:
                    til.append(cur)
                    cur = [i]
            til.append(cur)
        out = []
        for u in til:
            a, b = u[0], u[-1]
            spelled = s[a:b + k].upper()
            c = spelled if spelled <= rc(spelled) else rc(spelled)
            segs.add(c)
            out.append((c, a, b))
        tilings.append(out)
    return segs, tilings
