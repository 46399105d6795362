"""Shared helpers for the parity tests."""
import numpy as np

FIELDS = ("id", "cell", "life", "gender", "birth_time", "age", "last_birth", "mate_id", "num_babies")


def assert_same_state(a, b, where=""):
    """Bit-exact comparison of two canonical (id-ordered, live-only) states."""
    assert len(a["id"]) == len(b["id"]), "%s: %d vs %d live agents" % (where, len(a["id"]), len(b["id"]))
    for k in FIELDS:
        x, y = a[k], b[k]
        if x.dtype.kind == "f":
            same = np.array_equal(x.view(np.uint64), y.view(np.uint64))
        else:
            same = np.array_equal(x, y)
        if not same:
            bad = np.flatnonzero(x != y)[:5]
            raise AssertionError("%s: field %s differs at %s: %s vs %s"
                                 % (where, k, bad, x[bad], y[bad]))
    if "genome" in a or "genome" in b:
        assert np.array_equal(a["genome"], b["genome"]), "%s: genomes differ" % where


def tut_verhulst(sc):
    """Switch a scenario to the tutorial Verhulst values (tut_Sexual.xml)."""
    p = sc["params"]
    p.verh_b0, p.verh_d0, p.verh_theta = 0.8, 0.001, 0.1
    return sc
