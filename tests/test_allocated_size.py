"""gal_allocated_size (host code): the bytes malloc grants a request, pinned
against the oracle's routing (size -> slice / big slice / block / segment
run) for every compiled instantiation, at every power of two, its
neighbours, and random sizes. The poison allocator places its trailer with
this figure, so an over-estimate would write past the allocation."""
import numpy as np
import pytest

import gallatin_amd as ga
from oracle import oracle
from tests import invariants


@pytest.mark.parametrize("triple", ga.supported_triples())
def test_matches_oracle_routing(triple):
    seg, mn, mx = triple
    sizes = {1, 2, 3}
    for k in range(1, 34):
        sizes |= {(1 << k) - 1, 1 << k, (1 << k) + 1}
    rng = np.random.default_rng(seg ^ mn ^ mx)
    sizes |= set(int(x) for x in rng.integers(1, 1 << 33, 400))
    for s in sorted(sizes):
        expect, _ = invariants.granted_bytes(seg, mn, mx, s, oracle.route)
        assert ga.allocated_size(s, seg, mn, mx) == expect, (triple, s)
