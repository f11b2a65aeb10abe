"""Comparison helpers shared by the parity tests."""
import math

import numpy as np

REL_TOL_F64 = 1e-12  # double SUM/AVG vs the Kahan-Neumaier oracle


def oracle_rows(cols, valid, types, dictionaries=None):
    """Oracle table -> list of tuples (None for NULL, VARCHAR ids -> str)."""
    n = len(valid[0]) if valid else 0
    out = []
    for r in range(n):
        row = []
        for c in range(len(cols)):
            if not valid[c][r]:
                row.append(None)
            elif dictionaries and dictionaries.get(c) is not None:
                row.append(dictionaries[c].get(int(cols[c][r])))
            else:
                row.append(cols[c][r].item())
        out.append(tuple(row))
    return out


def _key(row, nkeys):
    return tuple((0, 0) if v is None else (1, v) for v in row[:nkeys])


def assert_same_groups(got_rows, exp_rows, nkeys, types, rel=REL_TOL_F64):
    """Same set of groups; integer cells bit-exact, DOUBLE cells within `rel`."""
    assert len(got_rows) == len(exp_rows), f"{len(got_rows)} groups vs {len(exp_rows)} expected"
    g = sorted(got_rows, key=lambda r: _key(r, nkeys))
    e = sorted(exp_rows, key=lambda r: _key(r, nkeys))
    worst = 0.0
    for a, b in zip(g, e):
        assert _key(a, nkeys) == _key(b, nkeys), f"group keys differ: {a[:nkeys]} vs {b[:nkeys]}"
        for c in range(nkeys, len(a)):
            x, y = a[c], b[c]
            if x is None or y is None:
                assert x is None and y is None, f"NULL mismatch in column {c}: {a} vs {b}"
                continue
            if types[c] == 12:  # DOUBLE
                if x == y or (math.isnan(x) and math.isnan(y)):
                    continue
                err = abs(x - y) / max(abs(y), 1e-300) if y != 0 else abs(x)
                worst = max(worst, err)
                assert err <= rel or (math.isnan(x) and math.isnan(y)), \
                    f"column {c}: {x!r} vs {y!r} (rel {err:.3e}) in group {a[:nkeys]}"
            else:
                assert x == y, f"column {c}: {x} vs {y} in group {a[:nkeys]}"
    return worst


def assert_topn_tie_aware(got_rows, all_rows, order_col, k, types, descending=True, rel=REL_TOL_F64):
    """TopN parity with ties: the multiset of order values equals the exact
    top-k values, every returned group carries its exact aggregates, and the
    groups returned at the boundary value come from its tie set."""
    nk = order_col if order_col < len(types) and order_col > 0 else 1
    vals = sorted((r[order_col] for r in all_rows), reverse=descending)
    exp_vals = vals[:k]
    got_vals = sorted((r[order_col] for r in got_rows), reverse=descending)
    assert got_vals == exp_vals, f"top-{k} values {got_vals} vs {exp_vals}"
    assert len({r[:nk] for r in got_rows}) == len(got_rows), "duplicate groups in TopN"
    by_key = {r[:nk]: r for r in all_rows}
    for r in got_rows:
        assert r[:nk] in by_key, f"returned group {r[:nk]} does not exist"
    assert_same_groups(got_rows, [by_key[r[:nk]] for r in got_rows], nk, types, rel)
    return True
