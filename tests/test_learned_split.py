"""The learned-proposal workload's train / test split is by base image: no held-out pair's
query image (base k % n_base) appears in a training pair, and every pair is in one side."""
import numpy as np
import pytest

import learned_ransac as lr


@pytest.mark.parametrize("n_pairs,n_base", [(4096, 256), (256, 32), (1024, 64), (100, 7)])
def test_split_by_base_has_no_shared_query_image(n_pairs, n_base):
    tr, te = lr.split_by_base(n_pairs, n_base, 0.75)
    assert len(tr) + len(te) == n_pairs
    assert len(np.intersect1d(tr, te)) == 0
    assert len(np.intersect1d(tr % n_base, te % n_base)) == 0
    assert len(te) > 0 and len(tr) > len(te)


def test_split_needs_both_sides():
    with pytest.raises(ValueError):
        lr.split_by_base(16, 1, 0.75)
