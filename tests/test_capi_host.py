"""CPU-side checks of the C ABI library: it loads, exports every symbol that
include/ch_swe.h declares, rejects bad configurations without a device, and
its plan-time edge maps (an independent combinatorial derivation) agree with
the oracle's geometric unfolding entry for entry."""
import os
import re

import numpy as np
import pytest

import ch_swe_amd as cs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "ch_swe.h")).read()
    declared = set(re.findall(r"\b(swe_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(cs.EXPORTS)
    L = cs.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert b"gfx950" in L.swe_version()


@pytest.mark.parametrize("N,order", [(3, 21), (7, 42)])
def test_plan_rejects_bad_sizes(N, order):
    with pytest.raises(cs.SweError, match="N >="):
        cs.create_linear_swm_operator(N, order)


def test_plan_rejects_bad_order():
    with pytest.raises(cs.SweError, match="order must be 21, 42 or 63"):
        cs.create_linear_swm_operator(16, 84)
    with pytest.raises(cs.SweError, match="sbp63 needs N >= 12"):
        cs.create_linear_swm_operator(11, 63)
    with pytest.raises(cs.SweError, match="compensated"):
        cs.create_linear_swm_operator(16, 63, compensated=True)


def _entries(m):
    return {tuple(int(m[k][n]) for k in ("tp", "ti", "tj", "tc", "sp", "si", "sj", "sc", "sign"))
            for n in range(len(m["tp"]))}


@pytest.mark.parametrize("N", [4, 9, 16])
@pytest.mark.parametrize("order", [21, 42, 63])
@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_edge_maps_match_oracle(oracle_mod, N, order, kind):
    if N < {21: 4, 42: 8, 63: 12}[order]:
        pytest.skip(f"sbp{order} needs a larger N")
    o = oracle_mod.Oracle(N, order)
    assert _entries(cs.halo_map(N, order, kind)) == _entries(o.halo_map(kind))


@pytest.mark.parametrize("precision", [cs.FP32, cs.FP64])
def test_fused_tables_selfcheck(precision):
    """Packed side table, segment decode and tile order, for uniform and
    edge-aligned tilings (no GPU needed)."""
    for N in list(range(4, 130)) + [191, 192, 255, 256, 383, 384, 511, 512, 1023]:
        cs.check_tables(N, precision)


def test_tile_split_requests_are_validated():
    """swe_config.tile_x0 / tile_y0 outside what the axis leaves are refused
    before any device work, with the admissible range."""
    with pytest.raises(cs.SweError, match=r"tile_x0 = 30 out of range 8 \.\. 24"):
        cs.create_linear_swm_operator(384, 42, cs.FP32, tile_x0=30, tuned=False)
    with pytest.raises(cs.SweError, match="tile_y0 = 9 out of range 8 .. 8"):
        cs.create_linear_swm_operator(384, 42, cs.FP32, tile_y0=9, tuned=False)
    with pytest.raises(cs.SweError, match="tiled uniformly"):
        cs.create_linear_swm_operator(16, 42, cs.FP32, tile_x0=8, tuned=False)
    with pytest.raises(cs.SweError, match="fused path"):
        cs.create_linear_swm_operator(384, 42, cs.FP32, tile_x0=12, unfused=True, tuned=False)


def test_autotune_choices_and_table(tmp_path, monkeypatch):
    """The tuner's candidate widths are the ones the library accepts, and the
    factory looks the tuned entry up by (N, order, precision)."""
    from ch_swe_amd import autotune as at
    assert at.axis_choices(384, 32) == [0] + list(range(8, 25))
    assert at.axis_choices(384, 16) == [0, 8]
    assert at.axis_choices(16, 32) == [0]
    p = tmp_path / "tuning.json"
    p.write_text('{"96/42/fp32": {"tile_x0": 12, "tile_y0": 0, "linear_blocks": false}}')
    monkeypatch.setenv("CH_SWE_TUNING", str(p))
    assert at.lookup(96, 42, cs.FP32)["tile_x0"] == 12
    assert at.lookup(96, 42, cs.FP64) is None
    assert at.key(384, 21, cs.FP64) == "384/21/fp64"
