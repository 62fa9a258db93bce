"""External CCD query data: the adapter contracts (ccd) and the
Tight-Inclusion rational-CSV and binary-PLY readers."""
from .ccd import (  # noqa: F401
    CCD_ADAPTER_SCHEMA_VERSION,
    CCDQueryFamily,
    DatasetQueryBatch,
    ExternalCCDQuery,
    validate_query_batch,
)
