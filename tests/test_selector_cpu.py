"""The learned selector shipped with the package (no GPU needed)."""
def test_committed_selector_weights():
    # the learned selector shipped with the package: shape, finite values, and
    # the held-out figures it was accepted on
    import numpy as np
    from slothdb_amd import selector as S
    p, meta = S.load()
    assert p.size == S.N_PARAMS and np.all(np.isfinite(p))
    assert np.all(p[S.N_W + S.N_IN:] > 0)  # feature scales
    ho = meta["held_out"]
    assert ho["queries"] >= 100 and ho["path_accuracy"] > 0.8
    assert ho["time_selector_over_best"] < ho["time_default_over_best"]
