"""External CCD query batches: the reference's adapter contracts
(src/python/p2cccd/datasets/ccd/contracts.py:9-151) -- one query is four
vertices at t0 and t1 of a vertex-face or edge-edge pair -- with the same
validation rules and messages."""
from __future__ import annotations

from dataclasses import dataclass, field
from enum import Enum
from pathlib import Path
from typing import Any

CCD_ADAPTER_SCHEMA_VERSION = 1
Vec3 = tuple[float, float, float]


class CCDQueryFamily(str, Enum):
    """"vf" or "ee"; str() is the value, as for the reference's StrEnum."""

    VERTEX_FACE = "vf"
    EDGE_EDGE = "ee"

    def __str__(self) -> str:
        return self.value

    @property
    def p2cccd_witness_family(self) -> str:
        return "point_triangle" if self is CCDQueryFamily.VERTEX_FACE else "edge_edge"


@dataclass(frozen=True, slots=True)
class ExternalCCDQuery:
    schema_version: int
    source_name: str
    scene_name: str
    batch_id: str
    query_id: int
    source_query_index: int
    family: CCDQueryFamily
    vertices_t0: tuple[Vec3, Vec3, Vec3, Vec3]
    vertices_t1: tuple[Vec3, Vec3, Vec3, Vec3]
    ground_truth_collides: bool | None = None
    box_pair: tuple[int, int] | None = None
    metadata: dict[str, Any] = field(default_factory=dict)

    def __post_init__(self) -> None:
        if self.schema_version != CCD_ADAPTER_SCHEMA_VERSION:
            raise ValueError("unsupported ExternalCCDQuery schema_version")
        if self.query_id < 0:
            raise ValueError("ExternalCCDQuery.query_id must be non-negative")
        if self.source_query_index < 0:
            raise ValueError("ExternalCCDQuery.source_query_index must be non-negative")
        if len(self.vertices_t0) != 4 or len(self.vertices_t1) != 4:
            raise ValueError("CCD queries must store four vertices at t0 and t1")
        for v in (*self.vertices_t0, *self.vertices_t1):
            if len(v) != 3:
                raise ValueError("CCD query vertices must be 3D")


@dataclass(frozen=True, slots=True)
class DatasetQueryBatch:
    schema_version: int
    source_name: str
    scene_name: str
    batch_id: str
    family: CCDQueryFamily
    queries: tuple[ExternalCCDQuery, ...]
    frame_t0: Path | None = None
    frame_t1: Path | None = None
    metadata: dict[str, Any] = field(default_factory=dict)

    def __post_init__(self) -> None:
        validate_query_batch(self)

    @property
    def query_count(self) -> int:
        return len(self.queries)

    @property
    def collision_count(self) -> int:
        return sum(1 for q in self.queries if q.ground_truth_collides is True)

    @property
    def known_label_count(self) -> int:
        return sum(1 for q in self.queries if q.ground_truth_collides is not None)


def validate_query_batch(batch: DatasetQueryBatch) -> DatasetQueryBatch:
    if batch.schema_version != CCD_ADAPTER_SCHEMA_VERSION:
        raise ValueError("unsupported DatasetQueryBatch schema_version")
    for name in ("source_name", "scene_name", "batch_id"):
        if not getattr(batch, name):
            raise ValueError(f"DatasetQueryBatch.{name} is required")
    seen: set[int] = set()
    for q in batch.queries:
        if q.schema_version != batch.schema_version:
            raise ValueError("query schema_version must match batch schema_version")
        if q.source_name != batch.source_name:
            raise ValueError("query source_name must match batch source_name")
        if q.scene_name != batch.scene_name:
            raise ValueError("query scene_name must match batch scene_name")
        if q.batch_id != batch.batch_id:
            raise ValueError("query batch_id must match batch batch_id")
        if q.family != batch.family:
            raise ValueError("query family must match batch family")
        if q.query_id in seen:
            raise ValueError("query_id values must be unique within a batch")
        seen.add(q.query_id)
    return batch
