"""The autotuner's candidate set (host only): every tile size, graph length
and occupancy variant at every mesh size, one occupancy variant for FP64."""

from specfempp_amd.autotune import (GRAPH_CHOICES, TILE_CHOICES, WPE_CHOICES, candidates,
                                    label, signature)
from specfempp_amd.config import batched_shots, homogeneous_box, layered_box


def test_candidates_cover_every_layout():
    for cfgs in ([homogeneous_box(nx=8, nz=6, nstep=10)], batched_shots(nshots=2, nstep=10)):
        c = candidates(cfgs)
        assert len(c) == len(TILE_CHOICES) * len(GRAPH_CHOICES) * len(WPE_CHOICES) == 18
        assert {x["tile_elems"] for x in c} == {16, 32, 64}
        assert {x["graph_steps"] for x in c} == {8, 32}
        assert {x["waves_per_eu"] for x in c} == {6, 7, 8}
        assert len({label(x) for x in c}) == 18


def test_fp64_has_the_default_variant_only():
    c = candidates([homogeneous_box(nx=8, nz=6, nstep=10)], precision=8)
    assert len(c) == 6 and {x["waves_per_eu"] for x in c} == {0}
    assert label(c[0]) == "tile64_graph8_default"
    assert label({"tile_elems": 32, "graph_steps": 32, "waves_per_eu": 6}) == "tile32_graph32_wpe6"


def test_signature_separates_workloads_and_precision():
    a = [homogeneous_box(nx=8, nz=6, nstep=10)]
    b = [layered_box(nx=10, nz=8, nstep=10, nsrc=1, nrec=4)]
    assert signature(a, 4) != signature(b, 4)
    assert signature(a, 4) != signature(a, 8)
    assert signature(a, 4) == signature([homogeneous_box(nx=8, nz=6, nstep=10)], 4)
