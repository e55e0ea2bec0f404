"""SourcePack index build over xc_fast output through the native indexer.

Host-side mirror of ``build_sourcepack`` (fastxc/operators/sourcepack/
builder.py:163-239) with the same arguments, result record and errors: the
per-timestamp ``<sourcepack_dir>/<timestamp>/sourcepack_index.tsv`` and
``_SUCCESS`` files are written by ``bin/sourcepack_index`` (C++,
fastxc_amd/host/sourcepack.cpp), the same code xc_fast runs on a side thread
with ``--sourcepack-dir`` as each timestamp completes (what the reference's
``AsyncSourcePackMaterializer`` does by polling the ``.done`` markers).
There is no Python fallback: without the built executable the call raises
FileNotFoundError.
"""
from __future__ import annotations

import os
import subprocess
from dataclasses import dataclass
from pathlib import Path

_ROOT = Path(__file__).resolve().parent.parent
SOURCEPACK_INDEX_EXE = _ROOT / "bin" / "sourcepack_index"


@dataclass(frozen=True)
class SourcePackBuildResult:
    output_dir: Path
    index_paths: list
    record_count: int
    source_count: int
    bytes_indexed: int


def _xcpack_dir(root: Path) -> Path:
    return root if root.name == "xcpack" else root / "xcpack"


def build_sourcepack(xc_output_root, sourcepack_dir=None, *, sort_within_source: bool = True,
                     progress_file=None) -> SourcePackBuildResult:
    """Indexes every timestamp of ``<xc_output_root>/xcpack`` (or the xcpack
    directory itself); the default output is the xcpack folder's sibling
    ``sourcepack``."""
    xcpack = _xcpack_dir(Path(xc_output_root).expanduser().resolve())
    if not xcpack.is_dir():
        raise FileNotFoundError(f"XC pack directory not found: {xc_output_root} or {xcpack}")
    out = (Path(sourcepack_dir).expanduser().resolve() if sourcepack_dir is not None
           else xcpack.parent / "sourcepack")
    if not SOURCEPACK_INDEX_EXE.is_file():
        raise FileNotFoundError(f"native indexer not built: {SOURCEPACK_INDEX_EXE}")
    cmd = [str(SOURCEPACK_INDEX_EXE), "-I", str(xcpack), "-O", str(out)]
    if not sort_within_source:
        cmd.append("--keep-order")
    if progress_file is not None:
        cmd += ["--progress", str(Path(progress_file).expanduser().resolve())]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"sourcepack_index failed ({proc.returncode}): {proc.stderr.strip()}")
    index_paths, records, sources, nbytes = [], 0, 0, 0
    for d in sorted(p for p in out.iterdir() if p.is_dir()):
        marker, index = d / "_SUCCESS", d / "sourcepack_index.tsv"
        if not (marker.is_file() and index.is_file()):
            continue
        fields = dict(line.split("\t", 1) for line in marker.read_text().splitlines() if "\t" in line)
        index_paths.append(index)
        records += int(fields.get("records", 0))
        sources += int(fields.get("sources", 0))
        nbytes += int(fields.get("bytes_indexed", 0))
    return SourcePackBuildResult(output_dir=out, index_paths=index_paths, record_count=records,
                                 source_count=sources, bytes_indexed=nbytes)
