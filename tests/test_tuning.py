"""The tuning cache's file logic (lapy_amd.tuning): entries merged by matrix
shape, other architectures kept, format checked.  The GPU side (the tuner
itself, save and load through a context) is in test_gpu_batched.py."""

import json

import pytest

from lapy_amd import tuning


def test_merge_replaces_by_shape_and_keeps_other_archs(tmp_path):
    doc = {"format": 1, "archs": {"gfx942": {"workers": 8, "spmm": [{"n": 5, "nnz": 9, "k": 2, "ncol": 8,
                                                                        "block": 64}]}}}
    e1 = {"n": 2562, "nnz": 17922, "k": 50, "ncol": 64, "block": 128}
    e2 = {"n": 642, "nnz": 4482, "k": 50, "ncol": 64, "block": 64}
    tuning.merge(doc, "gfx950", 16, [e1, e2])
    w, ents = tuning.entries_for(doc, "gfx950")
    assert w == 16 and ents == [e2, e1]  # sorted by shape
    # a retune of one shape replaces it; workers 0 keeps the stored count
    tuning.merge(doc, "gfx950", 0, [dict(e1, block=256)])
    w, ents = tuning.entries_for(doc, "gfx950")
    assert w == 16 and [e["block"] for e in ents] == [64, 256]
    assert tuning.entries_for(doc, "gfx942")[0] == 8
    assert tuning.entries_for(doc, "gfx90a") == (0, [])
    p = tmp_path / "t.json"
    p.write_text(json.dumps(doc))
    assert tuning._read(str(p)) == doc


def test_missing_file_is_empty_and_wrong_format_raises(tmp_path):
    assert tuning._read(str(tmp_path / "none.json")) == {"format": 1, "archs": {}}
    p = tmp_path / "bad.json"
    p.write_text(json.dumps({"format": 7, "archs": {}}))
    with pytest.raises(ValueError, match="format"):
        tuning._read(str(p))
