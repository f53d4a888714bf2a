"""Run files: one compressed .npz per campaign with a JSON manifest
(utils/data_io.py of the reference, format version 1).

Arrays are stored as ``L{L}_{name}`` (samples with a leading sample axis and a
replica axis in temperature order); the ``manifest`` entry holds
``{"version": 1, "runs": {str(L): {"metadata": ..., "sample_names": [...]}}}``
with numpy metadata converted to JSON values.  Only numpy's non-pickling
loader is used.
"""

import json
from pathlib import Path

import numpy as np

FORMAT_VERSION = 1


def _json_ready(value):
    if isinstance(value, np.ndarray):
        return value.tolist()
    if isinstance(value, np.generic):
        return value.item()
    if isinstance(value, dict):
        return {str(k): _json_ready(v) for k, v in value.items()}
    if isinstance(value, (list, tuple)):
        return [_json_ready(v) for v in value]
    return value


def save_runs(path, runs):
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    arrays = {}
    manifest = {"version": FORMAT_VERSION, "runs": {}}
    for L, run in runs.items():
        samples = run["samples"]
        manifest["runs"][str(L)] = {
            "metadata": _json_ready(run["metadata"]),
            "sample_names": list(samples),
        }
        for name, values in samples.items():
            arrays[f"L{L}_{name}"] = np.asarray(values)
    arrays["manifest"] = np.array(json.dumps(manifest))
    np.savez_compressed(path, **arrays)


def load_runs(path):
    runs = {}
    with np.load(path, allow_pickle=False) as data:
        manifest = json.loads(data["manifest"].item())
        if manifest["version"] != FORMAT_VERSION:
            raise ValueError("Unsupported data format version")
        for key, info in manifest["runs"].items():
            L = int(key)
            runs[L] = {
                "samples": {name: data[f"L{L}_{name}"].copy() for name in info["sample_names"]},
                "metadata": info["metadata"],
            }
    return runs
