"""Tight-Inclusion rational-CSV queries: the reference package's reader
(src/python/p2cccd/datasets/tight_inclusion_queries.py:14-290), its names,
rules and messages.

A query is 8 rows (4 vertices at t0, then 4 at t1) of
`x_num,x_den,y_num,y_den,z_num,z_den,truth`: integer numerators and non-zero
denominators kept exactly (Python ints), coordinates as float(num) /
float(den), and a 0/1 ground-truth label equal over the block.  The vertex
order is the query family's: vertex-face (point, triangle v0, v1, v2),
edge-edge (edge a0, a1, edge b0, b1), the (t,u,v) certificate's layout.

`load_queries` reads a whole file into the (n, 8, 3) array the certificate
takes, under either rule: "python" (this module's, the reference package's)
or "cpp" (the reference's C++ query benchmark: long-double quotient rounded
to double, p2cccd_cpp.read_tight_inclusion_csv)."""
from __future__ import annotations

import itertools
import math
from dataclasses import dataclass
from pathlib import Path
from typing import Iterable, Iterator, Literal, Sequence

import numpy as np

TightInclusionQueryKind = Literal["vertex-face", "edge-edge"]
TIGHT_INCLUSION_QUERY_ROWS = 8
TIGHT_INCLUSION_CSV_COLUMNS = 7
_SKIPPED_CASE_DIRS = {"nyu-full-dataset-archives", "Sample-Queries", "visualization"}


@dataclass(frozen=True, slots=True)
class TightInclusionCSVFile:
    case_name: str
    kind: TightInclusionQueryKind
    csv_path: Path
    line_count: int
    query_count: int
    byte_size: int


@dataclass(frozen=True, slots=True)
class TightInclusionPrimitiveQuery:
    case_name: str
    kind: TightInclusionQueryKind
    csv_path: Path
    query_index: int
    vertices_t0_t1: np.ndarray
    ground_truth: bool
    numerators: np.ndarray
    denominators: np.ndarray

    @property
    def vertices_t0(self) -> np.ndarray:
        return self.vertices_t0_t1[:4]

    @property
    def vertices_t1(self) -> np.ndarray:
        return self.vertices_t0_t1[4:]

    @property
    def rational_magnitude_features(self) -> np.ndarray:
        """[max, max, mean, mean] of log1p|numerator| and log1p|denominator|."""
        nums = np.asarray([_log1p_abs(v) for v in self.numerators.flat], dtype=np.float64)
        dens = np.asarray([_log1p_abs(v) for v in self.denominators.flat], dtype=np.float64)
        return np.asarray([float(nums.max(initial=0.0)), float(dens.max(initial=0.0)),
                           float(nums.mean() if nums.size else 0.0), float(dens.mean() if dens.size else 0.0)],
                          dtype=np.float64)


def _log1p_abs(value) -> float:
    x = abs(int(value))
    try:
        return math.log1p(float(x))
    except OverflowError:  # beyond the double range: decimal digits on a log scale
        return len(str(x)) * math.log(10.0)


def default_tight_inclusion_dataset_root(project_root: Path | None = None) -> Path:
    if project_root is not None:
        return project_root / "baseline" / "datasets" / "continuous-collision-detection"
    return Path("src/baseline/datasets/continuous-collision-detection")


def normalize_query_kind(kind: str) -> TightInclusionQueryKind:
    if kind in ("vertex-face", "vf"):
        return "vertex-face"
    if kind in ("edge-edge", "ee"):
        return "edge-edge"
    raise ValueError(f"unsupported Tight-Inclusion query kind: {kind}")


def infer_query_kind(csv_path: Path) -> TightInclusionQueryKind:
    """From a `vertex-face` / `edge-edge` directory in the path, else the
    file name's prefix."""
    parts = set(Path(csv_path).parts)
    for kind in ("vertex-face", "edge-edge"):
        if kind in parts:
            return kind
    name = Path(csv_path).name
    for kind in ("vertex-face", "edge-edge"):
        if name.startswith(kind):
            return kind
    raise ValueError(f"cannot infer Tight-Inclusion query kind from path: {csv_path}")


def infer_case_name(csv_path: Path, dataset_root: Path | None = None) -> str:
    """The first directory under dataset_root, else the directory above the
    nearest `vertex-face` / `edge-edge` directory."""
    csv_path = Path(csv_path)
    if dataset_root is not None:
        try:
            return csv_path.resolve().relative_to(Path(dataset_root).resolve()).parts[0]
        except ValueError:
            pass
    for parent in csv_path.parents:
        if parent.name in ("vertex-face", "edge-edge"):
            return parent.parent.name
    raise ValueError(f"cannot infer Tight-Inclusion case name from path: {csv_path}")


def _at(line_number: int | None) -> str:
    return f" at line {line_number}" if line_number is not None else ""


def parse_rational_vertex_row(line: str, *, line_number: int | None = None):
    """One row -> (vertex float64[3], numerators object[3], denominators
    object[3], truth)."""
    text = line.strip()
    if not text:
        raise ValueError("empty Tight-Inclusion CSV row")
    cols = text.split(",")
    if len(cols) != TIGHT_INCLUSION_CSV_COLUMNS:
        raise ValueError(f"expected 7 columns{_at(line_number)}, got {len(cols)}")
    try:
        v = [int(c) for c in cols]
    except ValueError as exc:
        raise ValueError(f"non-integer rational CSV value{_at(line_number)}: {text}") from exc
    den = np.asarray(v[1:6:2], dtype=object)
    if any(d == 0 for d in den):
        raise ValueError(f"zero denominator in Tight-Inclusion CSV row{_at(line_number)}")
    if v[6] not in (0, 1):
        raise ValueError(f"ground-truth value must be 0 or 1{_at(line_number)}, got {v[6]}")
    num = np.asarray(v[0:6:2], dtype=object)
    vertex = num.astype(np.float64) / den.astype(np.float64)
    return vertex, num, den, bool(v[6])


def parse_query_lines(lines: Sequence[str], *, case_name: str, kind: TightInclusionQueryKind, csv_path: Path,
                      query_index: int, first_line_number: int = 1) -> TightInclusionPrimitiveQuery:
    if len(lines) != TIGHT_INCLUSION_QUERY_ROWS:
        raise ValueError(f"Tight-Inclusion query must contain 8 rows, got {len(lines)}")
    rows = [parse_rational_vertex_row(l, line_number=first_line_number + k) for k, l in enumerate(lines)]
    if len({r[3] for r in rows}) != 1:
        raise ValueError(f"inconsistent truth labels in {csv_path}:{query_index}")
    return TightInclusionPrimitiveQuery(
        case_name=case_name, kind=kind, csv_path=csv_path, query_index=query_index,
        vertices_t0_t1=np.vstack([r[0] for r in rows]).astype(np.float64, copy=False), ground_truth=rows[0][3],
        numerators=np.vstack([r[1] for r in rows]).astype(object, copy=False),
        denominators=np.vstack([r[2] for r in rows]).astype(object, copy=False))


def count_csv_lines(csv_path: Path) -> int:
    """Newline bytes in the file, read in 1 MiB chunks."""
    n = 0
    with Path(csv_path).open("rb") as fh:
        while chunk := fh.read(1 << 20):
            n += chunk.count(b"\n")
    return n


def inspect_tight_inclusion_csv(csv_path: Path, *, dataset_root: Path | None = None) -> TightInclusionCSVFile:
    path = Path(csv_path)
    if not path.exists():
        raise FileNotFoundError(path)
    lines = count_csv_lines(path)
    if lines % TIGHT_INCLUSION_QUERY_ROWS:
        raise ValueError(f"{path} line_count={lines} is not divisible by 8")
    return TightInclusionCSVFile(case_name=infer_case_name(path, dataset_root), kind=infer_query_kind(path),
                                 csv_path=path, line_count=lines, query_count=lines // TIGHT_INCLUSION_QUERY_ROWS,
                                 byte_size=path.stat().st_size)


def discover_tight_inclusion_csv_files(dataset_root: Path, *, cases: Iterable[str] | None = None,
                                       kinds: Iterable[str] = ("vertex-face", "edge-edge"),
                                       inspect: bool = False) -> tuple[TightInclusionCSVFile, ...]:
    """`<root>/<case>/<kind>/*.csv` in sorted order, skipping hidden and
    archive directories."""
    root = Path(dataset_root)
    if not root.exists():
        raise FileNotFoundError(root)
    wanted = set(cases) if cases is not None else None
    kinds = tuple(normalize_query_kind(k) for k in kinds)
    found = []
    for case_dir in sorted(p for p in root.iterdir() if p.is_dir() and not p.name.startswith(".")):
        if (wanted is not None and case_dir.name not in wanted) or case_dir.name in _SKIPPED_CASE_DIRS:
            continue
        for kind in kinds:
            kind_dir = case_dir / kind
            if not kind_dir.exists():
                continue
            for p in sorted(kind_dir.glob("*.csv")):
                found.append(inspect_tight_inclusion_csv(p, dataset_root=root) if inspect else
                             TightInclusionCSVFile(case_name=case_dir.name, kind=kind, csv_path=p, line_count=0,
                                                   query_count=0, byte_size=p.stat().st_size))
    return tuple(found)


def read_tight_inclusion_query(csv_path: Path, query_index: int, *,
                               dataset_root: Path | None = None) -> TightInclusionPrimitiveQuery:
    if query_index < 0:
        raise ValueError("query_index must be non-negative")
    path = Path(csv_path)
    first = query_index * TIGHT_INCLUSION_QUERY_ROWS
    with path.open("r", encoding="ascii") as fh:
        lines = list(itertools.islice(fh, first, first + TIGHT_INCLUSION_QUERY_ROWS))
    if len(lines) != TIGHT_INCLUSION_QUERY_ROWS:
        raise ValueError(f"query_index={query_index} is out of range for {path}")
    return parse_query_lines(lines, case_name=infer_case_name(path, dataset_root), kind=infer_query_kind(path),
                             csv_path=path, query_index=query_index, first_line_number=first + 1)


def iter_tight_inclusion_queries(csv_path: Path, *, dataset_root: Path | None = None, start_query_index: int = 0,
                                 limit: int | None = None) -> Iterator[TightInclusionPrimitiveQuery]:
    if start_query_index < 0:
        raise ValueError("start_query_index must be non-negative")
    if limit is not None and limit < 0:
        raise ValueError("limit must be non-negative")
    path = Path(csv_path)
    case, kind = infer_case_name(path, dataset_root), infer_query_kind(path)
    emitted, index, block = 0, 0, []
    with path.open("r", encoding="ascii") as fh:
        for line_number, line in enumerate(fh, start=1):
            block.append(line)
            if len(block) < TIGHT_INCLUSION_QUERY_ROWS:
                continue
            if index >= start_query_index:
                if limit is not None and emitted >= limit:
                    return
                yield parse_query_lines(block, case_name=case, kind=kind, csv_path=path, query_index=index,
                                        first_line_number=line_number - TIGHT_INCLUSION_QUERY_ROWS + 1)
                emitted += 1
            block = []
            index += 1
    if block:
        raise ValueError(f"{path} has a truncated query block with {len(block)} leftover rows")


def iter_dataset_queries(dataset_root: Path, *, cases: Iterable[str] | None = None,
                         kinds: Iterable[str] = ("vertex-face", "edge-edge"),
                         limit_per_file: int | None = None) -> Iterator[TightInclusionPrimitiveQuery]:
    for f in discover_tight_inclusion_csv_files(dataset_root, cases=cases, kinds=kinds):
        yield from iter_tight_inclusion_queries(f.csv_path, dataset_root=dataset_root, limit=limit_per_file)


def load_queries(csv_path: Path, *, rule: str = "python", limit: int | None = None):
    """A file's queries as (vertices float64[n, 8, 3], truth bool[n]) for the
    certificate.  rule "python": this module's parse; rule "cpp": the C++
    benchmark's (native reader)."""
    if rule == "cpp":
        import p2cccd_cpp

        return p2cccd_cpp.read_tight_inclusion_csv(str(csv_path), int(limit or 0))
    if rule != "python":
        raise ValueError("rule must be 'python' or 'cpp'")
    qs = list(iter_tight_inclusion_queries(csv_path, limit=limit))
    x = np.stack([q.vertices_t0_t1 for q in qs]) if qs else np.zeros((0, 8, 3))
    return x, np.array([q.ground_truth for q in qs], dtype=bool)
