"""The result folder format against a folder the reference wrote.

``tests/golden/evidence_ref/`` was written by the reference's
``OptimizationResult.save`` (src/topogpu/solver.py:46-61,
src/topogpu/evidence.py:23-90) from the inputs of ``evidence_inputs`` (by
tests/golden/make_golden.py, in the build container). The same inputs saved
here must give the same files, byte for byte, and the same manifest.
"""

from __future__ import annotations

import csv
from pathlib import Path

import numpy as np
import pytest

import topogpu_amd as tg
from topogpu_amd.evidence import EvidenceBundle

from golden.evidence_inputs import evidence_inputs

REF = Path(__file__).resolve().parent / "golden" / "evidence_ref"
FILES = ["history.csv", "rho_final.npy", "render_metadata.json", "summary.json", "ARTIFACT_MANIFEST.csv"]


def test_save_matches_reference_folder(tmp_path):
    rho, history, summary = evidence_inputs()
    res = tg.OptimizationResult(tg.cantilever_3d((4, 3, 2)), rho, history, summary)
    b = res.save(tmp_path / "run")
    assert sorted(p.name for p in REF.iterdir()) == sorted(FILES)
    for name in FILES:
        assert (tmp_path / "run" / name).read_bytes() == (REF / name).read_bytes(), name
    with open(REF / "ARTIFACT_MANIFEST.csv", newline="") as fh:
        ref_rows = list(csv.DictReader(fh))
    assert [r["path"] for r in ref_rows] == ["history.csv", "rho_final.npy", "render_metadata.json", "summary.json"]
    assert [{k: str(v) for k, v in r.items()} for r in b.rows] == ref_rows


def test_write_array_and_add(tmp_path):
    b = EvidenceBundle(tmp_path / "x")
    with pytest.raises(ValueError, match="Evidence arrays must use a .npy filename."):
        b.write_array("u.bin", np.zeros(3), "field")
    p = b.write_array("u.npy", np.arange(5.0), "field")
    assert np.array_equal(np.load(p), np.arange(5.0))
    (tmp_path / "x" / "sub").mkdir()
    (tmp_path / "x" / "sub" / "note.txt").write_text("hi")
    b.add(tmp_path / "x" / "sub" / "note.txt", "note")
    b.write_manifest()
    with open(tmp_path / "x" / "ARTIFACT_MANIFEST.csv", newline="") as fh:
        rows = list(csv.DictReader(fh))
    assert [(r["path"], r["role"], r["bytes"]) for r in rows] == [("u.npy", "field", str(p.stat().st_size)),
                                                                 ("sub/note.txt", "note", "2")]
    assert rows[1]["sha256"] == "8f434346648f6b96df89dda901c5176b10a6d83961dd3c1ac88b59b2dc327aa4"
