"""Shared cases of the unary contraction/attraction tests (imported by
test_oracle.py and test_unary_gpu.py)."""


def unary_groups(labels):
    """Groups of an index string with repeats, numbered as the reference's
    expression code pulls them (MIA_Expr.h:695-731, ExprUtil.h:120-240):
    contractions (plain repeated letters) then attractions (repeated '!'
    letters, written upper case here), each by the position of its letter's
    last occurrence."""
    last = {c: p for p, c in enumerate(labels)}
    rep = [c for c in sorted(last, key=last.get) if labels.count(c) > 1]
    con = [c for c in rep if c.islower()]
    att = [c for c in rep if c.isupper()]
    num = {c: g + 1 for g, c in enumerate(con + att)}
    return [num.get(c, 0) for c in labels], [0] * len(con) + [1] * len(att)


UNARY_CASES = [
    # (labels of A, dims, einsum of the dense result in the output mode order:
    #  staying modes in mode order, then attracted ones in group order)
    ("ijiklm", (5, 6, 5, 6, 4, 3), "ijiklm->jklm"),
    ("jij", (5, 6, 5), "jij->i"),
    ("ijiilm", (5, 6, 5, 5, 7, 3), "ijiilm->jlm"),
    ("ijijlm", (5, 6, 5, 6, 4, 3), "ijijlm->lm"),
    ("ijijim", (5, 6, 5, 6, 5, 3), "ijijim->m"),
    ("II", (5, 5), "ii->i"),
    ("IjI", (5, 6, 5), "iji->ji"),
    ("IjIIlm", (5, 6, 5, 5, 7, 3), "ijiilm->jlmi"),
    ("IjIjkm", (5, 6, 5, 6, 5, 3), "ijijkm->kmi"),
    ("IjIjKK", (5, 6, 5, 6, 5, 5), "ijijkk->ik"),
    ("iJiiJm", (5, 6, 5, 5, 6, 3), "ijiijm->mj"),
    ("ii", (7, 7), "ii->"),
]
