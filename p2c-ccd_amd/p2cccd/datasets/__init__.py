"""External CCD query data: the adapter contracts (ccd), the
Tight-Inclusion rational-CSV reader (tight_inclusion_queries) and binary-PLY
frames (ply)."""
from .ccd import (  # noqa: F401
    CCD_ADAPTER_SCHEMA_VERSION,
    CCDQueryFamily,
    DatasetQueryBatch,
    ExternalCCDQuery,
    validate_query_batch,
)
from .tight_inclusion_queries import (  # noqa: F401
    TIGHT_INCLUSION_CSV_COLUMNS,
    TIGHT_INCLUSION_QUERY_ROWS,
    TightInclusionCSVFile,
    TightInclusionPrimitiveQuery,
    count_csv_lines,
    default_tight_inclusion_dataset_root,
    discover_tight_inclusion_csv_files,
    infer_case_name,
    infer_query_kind,
    inspect_tight_inclusion_csv,
    iter_dataset_queries,
    iter_tight_inclusion_queries,
    load_queries,
    normalize_query_kind,
    parse_query_lines,
    parse_rational_vertex_row,
    read_tight_inclusion_query,
)
