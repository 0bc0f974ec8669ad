"""CPU checks of the Python front end's index labels: element-wise marks
(``!``, Index.h:38-57), their upper-case shorthand and ``~`` lowering the
marks a product left at the back (MIA_Expr.h:257-301)."""
import pytest

from libnt_amd.sparse import Atom, parse_labels


def test_parse_labels():
    assert parse_labels("i!jk!!l") == [("i", 0), ("j", 1), ("k", 0), ("l", 2)]
    assert parse_labels("iJ") == [("i", 0), ("j", 1)]
    assert parse_labels("!!I") == [("i", 3)]
    assert parse_labels([("x", 2)]) == [("x", 2)]
    for bad in ("ijkj", "ij!", "i1"):
        with pytest.raises(ValueError):
            parse_labels(bad)


def test_invert_lowers_only_the_product_inter_indices():
    t = ~Atom(None, "ik!j!!l", inter_count=2)
    assert t.tokens == [("i", 0), ("k", 0), ("j", 0), ("l", 1)] and t.inter_count == 0
    assert t.labels == "ikjL"
    assert (~Atom(None, "i!j")).tokens == [("i", 0), ("j", 1)]  # not a product: unchanged
