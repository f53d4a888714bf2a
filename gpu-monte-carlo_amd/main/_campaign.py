"""Shared campaign helpers: preflight of output paths and atomic per-size saves
(main/he3.py:90-110,192-201 and main/kawamura.py:81-93,156-165 of the reference)."""

import os
import tempfile
from pathlib import Path

from utils.data_io import save_runs


def preflight(paths, duplicate_message="Duplicate simulation output paths"):
    """Refuse duplicate or existing outputs before any simulation starts."""
    paths = [Path(p) for p in paths]
    if len(paths) != len(set(paths)):
        raise ValueError(duplicate_message)
    existing = [p for p in paths if p.exists()]
    if existing:
        raise FileExistsError("Output files already exist:\n" + "\n".join(map(str, existing)))
    return paths


def save_atomically(path, L, run):
    """Stage in the destination directory, then link the final name (never overwrites)."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with tempfile.TemporaryDirectory(dir=path.parent) as tmp:
        staging = Path(tmp) / "run.npz"
        save_runs(staging, {L: run})
        os.link(staging, path)
    return path
