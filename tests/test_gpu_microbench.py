"""CDNA4 micro-benchmarks (ssam_microbench): finite, positive, and ordered the way
the kernels' design relies on -- the DPP hop is cheaper than a ds_bpermute or an
LDS round trip, in latency and in throughput."""
import math

import pytest

from ssam_amd import api

pytestmark = pytest.mark.gpu


def test_microbench_values():
    m = api.microbench()
    assert all(math.isfinite(v) and v > 0 for v in m.values()), m
    assert 1 <= m["fma_latency_cycles"] <= 64, m
    assert m["dpp_add_latency_cycles"] < m["bpermute_latency_cycles"], m
    assert m["dpp_add_latency_cycles"] < m["lds_roundtrip_latency_cycles"], m
    assert m["dpp_add_per_cycle_per_cu"] > m["bpermute_per_cycle_per_cu"], m
    assert m["fma_per_cycle_per_cu"] <= 4.5, m  # at most one wave-instruction per SIMD per cycle
