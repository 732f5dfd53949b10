"""Initial place: the device B2B assembly builds the same CSR and right-hand
side as the sequential host assembly (Eigen setFromTriplets semantics:
duplicates summed in triplet order), so every BiCGSTAB iterate, the IP log
and the final positions agree bit for bit."""
import os

import numpy as np
import pytest

from helpers import design
from replace_amd import netgen
from replace_amd.capi import Placer

pytestmark = pytest.mark.gpu


def _ip(nl, p, host):
    if host:
        os.environ["RP_IP_HOST_ASSEMBLY"] = "1"
    try:
        g = Placer(nl, p)
        st = g.initial_place()
        err, hp = g.ip_records()
        return st.iterations, err.tolist(), hp.tolist(), g.positions()
    finally:
        os.environ.pop("RP_IP_HOST_ASSEMBLY", None)


def _cases():
    out = [design(n) for n in ("simple03", "simple04", "ar01", "incremental01")]
    nl = netgen.generate(3000, 3150, seed=77, num_macros=4, name="ipmix")
    out.append((nl, netgen.params_default(target_density=0.7, initial_place_max_iter=6,
                                          initial_place_max_fanout=30)))
    nl, p = netgen.workload("c0", scale=0.25)
    p.initial_place_max_iter = 5
    out.append((nl, p))
    return out


@pytest.mark.parametrize("case", range(6))
def test_device_assembly_matches_host_assembly(case):
    nl, p = _cases()[case]
    d = _ip(nl, p, host=False)
    h = _ip(nl, p, host=True)
    assert d[0] == h[0] and d[1] == h[1] and d[2] == h[2]
    assert np.array_equal(d[3][0], h[3][0]) and np.array_equal(d[3][1], h[3][1])
