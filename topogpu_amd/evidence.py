"""Result folder of a SIMP run: artifacts plus a SHA-256 manifest.

The on-disk format is the one a reference result folder has
(src/topogpu/evidence.py:23-90, written by OptimizationResult.save,
src/topogpu/solver.py:46-61), so folders from either backend are read by the
same tools:

* every artifact sits directly in the folder, which is created on the first
  write;
* ``summary.json`` (role ``summary``) and ``render_metadata.json`` (role
  ``render_metadata``): ``json.dumps(obj, indent=2)`` -- two-space indent,
  keys in insertion order, ASCII escapes, no trailing newline -- as UTF-8;
* ``history.csv`` (role ``history``): one column per key that appears in any
  row, sorted; one line per row, empty cells for absent keys; CSV with
  minimal quoting and CRLF line ends;
* ``rho_final.npy`` (role ``density``) and ``<name>.npy`` (any role, the name
  must end in ``.npy``): NumPy ``.npy`` files of the arrays as given;
* ``ARTIFACT_MANIFEST.csv``: columns ``path, role, bytes, sha256`` (path
  relative to the folder, size in bytes, lowercase hex digest), one row per
  artifact in the order written, CRLF line ends; it does not list itself.

Here each artifact is encoded to bytes in memory, its manifest row is
computed from those bytes, and the file is written once.
"""

from __future__ import annotations

import csv
import hashlib
import io
import json
from pathlib import Path
from typing import Any

import numpy as np

MANIFEST = "ARTIFACT_MANIFEST.csv"
MANIFEST_COLUMNS = ("path", "role", "bytes", "sha256")


def _csv_bytes(columns, rows) -> bytes:
    """Minimal-quoting CSV with CRLF line ends; missing cells are empty."""
    text = io.StringIO(newline="")
    out = csv.writer(text)
    out.writerow(columns)
    for row in rows:
        out.writerow([row.get(c, "") for c in columns])
    return text.getvalue().encode("utf-8")


def _json_bytes(obj: Any) -> bytes:
    return json.dumps(obj, indent=2).encode("utf-8")


def _npy_bytes(values) -> bytes:
    buf = io.BytesIO()
    np.save(buf, values)
    return buf.getvalue()


def _digest(data: bytes) -> str:
    return hashlib.sha256(data).hexdigest()


class EvidenceBundle:
    """Writes artifacts into ``root`` and keeps their manifest rows in
    ``rows`` until ``write_manifest``."""

    def __init__(self, root, rows: list[dict[str, Any]] | None = None) -> None:
        self.root = Path(root)
        self.rows: list[dict[str, Any]] = [] if rows is None else rows

    def _emit(self, name: str, role: str, data: bytes) -> Path:
        self.root.mkdir(parents=True, exist_ok=True)
        target = self.root / name
        target.write_bytes(data)
        self.rows.append({"path": name, "role": role, "bytes": len(data), "sha256": _digest(data)})
        return target

    def add(self, path, role: str) -> None:
        """Record a file that is already in the folder (or below it)."""
        p = Path(path).resolve()
        data = p.read_bytes()
        self.rows.append({"path": str(p.relative_to(self.root.resolve())), "role": role, "bytes": len(data),
                          "sha256": _digest(data)})

    def write_summary(self, summary: dict[str, Any]) -> Path:
        return self._emit("summary.json", "summary", _json_bytes(summary))

    def write_render_metadata(self, metadata: dict[str, Any]) -> Path:
        return self._emit("render_metadata.json", "render_metadata", _json_bytes(metadata))

    def write_history(self, rows: list[dict[str, Any]]) -> Path:
        columns = sorted(set().union(*(r.keys() for r in rows))) if rows else []
        return self._emit("history.csv", "history", _csv_bytes(columns, rows))

    def write_density(self, rho: np.ndarray) -> Path:
        return self._emit("rho_final.npy", "density", _npy_bytes(rho))

    def write_array(self, name: str, values: np.ndarray, role: str) -> Path:
        if not name.endswith(".npy"):
            raise ValueError("Evidence arrays must use a .npy filename.")
        return self._emit(name, role, _npy_bytes(values))

    def write_manifest(self) -> Path:
        self.root.mkdir(parents=True, exist_ok=True)
        target = self.root / MANIFEST
        target.write_bytes(_csv_bytes(MANIFEST_COLUMNS, self.rows))
        return target
