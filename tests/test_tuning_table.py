"""Tuning table parsing (cc_load_tuning; host only, no GPU): keys the table sets, defaults for the
rest, a missing file gives the defaults, malformed or out-of-range entries are refused."""
import pytest

import cudacompress_amd as cc


def test_defaults_and_partial_table(tmp_path):
    d = cc.defaultTuning()
    assert d["codingBlockSize"] == 128 and d["fusedXY"] == 1 and len(d["designGrid"]) == 4
    p = tmp_path / "t.txt"
    p.write_text("# comment\ndesign_grid 1 2 3 4   # trailing comment\ncoding_block_size 64\nmeasured 123.4\n"
                 "lane_design_k 128\n")
    t = cc.loadTuning(p)
    assert t["designGrid"] == [1, 2, 3, 4] and t["codingBlockSize"] == 64 and t["laneDesignK"] == 128
    assert t["fusedXY"] == d["fusedXY"] and t["xySlicesMax"] == d["xySlicesMax"]
    assert cc.loadTuning(tmp_path / "missing.txt") == d


@pytest.mark.parametrize("text", ["design_grid 1 2\n", "fused_xy 2\n", "coding_block_size 100\n",
                                  "design_grid 0 1 1 1\n", "xy_slices_max x\n", "lane_design_k 100\n"])
def test_bad_tables_are_refused(tmp_path, text):
    p = tmp_path / "bad.txt"
    p.write_text(text)
    with pytest.raises(cc.CCError):
        cc.loadTuning(p)


def test_committed_table_parses():
    import os
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    path = os.path.join(here, "cudacompress_amd", "tuning", "gfx950.txt")
    assert os.path.exists(path), "the gfx950 tuning table ships with the package"
    t = cc.loadTuning(path)
    assert t == cc.loadTuning()  # the library reads the same table
