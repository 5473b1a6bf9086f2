"""Run-artifact writer with a SHA-256 manifest.

File names, roles and the manifest columns match the reference's
``EvidenceBundle`` (src/topogpu/evidence.py:23-90) so that result folders of
the hip backend are interchangeable with the reference's.
"""

from __future__ import annotations

import csv
import hashlib
import json
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any

import numpy as np


def file_sha256(path: Path) -> str:
    digest = hashlib.sha256()
    with open(path, "rb") as fh:
        while block := fh.read(1 << 20):
            digest.update(block)
    return digest.hexdigest()


@dataclass(slots=True)
class EvidenceBundle:
    root: Path
    rows: list[dict[str, Any]] = field(default_factory=list)

    def _target(self, name: str) -> Path:
        self.root = Path(self.root)
        self.root.mkdir(parents=True, exist_ok=True)
        return self.root / name

    def add(self, path: Path, role: str) -> None:
        p = Path(path).resolve()
        self.rows.append({"path": str(p.relative_to(Path(self.root).resolve())), "role": role,
                          "bytes": p.stat().st_size, "sha256": file_sha256(p)})

    def write_summary(self, summary: dict[str, Any]) -> Path:
        path = self._target("summary.json")
        path.write_text(json.dumps(summary, indent=2), encoding="utf-8")
        self.add(path, "summary")
        return path

    def write_history(self, rows: list[dict[str, Any]]) -> Path:
        path = self._target("history.csv")
        columns = sorted({k for r in rows for k in r})
        with open(path, "w", newline="", encoding="utf-8") as fh:
            w = csv.DictWriter(fh, fieldnames=columns)
            w.writeheader()
            w.writerows(rows)
        self.add(path, "history")
        return path

    def write_density(self, rho: np.ndarray) -> Path:
        path = self._target("rho_final.npy")
        np.save(path, rho)
        self.add(path, "density")
        return path

    def write_array(self, name: str, values: np.ndarray, role: str) -> Path:
        if not name.endswith(".npy"):
            raise ValueError("Evidence arrays must use a .npy filename.")
        path = self._target(name)
        np.save(path, values)
        self.add(path, role)
        return path

    def write_render_metadata(self, metadata: dict[str, Any]) -> Path:
        path = self._target("render_metadata.json")
        path.write_text(json.dumps(metadata, indent=2), encoding="utf-8")
        self.add(path, "render_metadata")
        return path

    def write_manifest(self) -> Path:
        path = self._target("ARTIFACT_MANIFEST.csv")
        with open(path, "w", newline="", encoding="utf-8") as fh:
            w = csv.DictWriter(fh, fieldnames=["path", "role", "bytes", "sha256"])
            w.writeheader()
            w.writerows(self.rows)
        return path
