"""Pins tests/label_check.py (the vectorised label rules) against the oracle."""
import numpy as np
import pytest

import oracle_bind as ob
from label_check import labels_from_flags

import efficientstructuralgraphclusteringongpus_amd as sgc


@pytest.mark.parametrize("eps,mu", [(0.5, 5), (0.3, 2), (0.7, 3), (0.8, 0)])
def test_label_rules_match_oracle_on_rmat(eps, mu):
    g = sgc.HostGraph.rmat(13, 16, seed=4)
    rp, adj = np.array(g.row_ptr), np.array(g.adj)
    labels, core, sim, st = ob.orc_scan(rp, adj, eps, mu)
    l2, c2 = labels_from_flags(rp, adj, sim, mu)
    assert np.array_equal(core, c2)
    assert np.array_equal(labels, l2)
