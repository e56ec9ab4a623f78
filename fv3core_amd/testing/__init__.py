"""Savepoint (translate) test harness of the reference, on this build's
components: comparison metric, threshold overrides, translate classes and
the savepoint readers."""
from .savepoint import NpzSerializer, Savepoint, write_savepoint
from .translate import (
    TranslateFortranData2Py,
    TranslateGrid,
    compare_arr,
    load_threshold_overrides,
    process_override,
    sample_wherefail,
    success,
    success_array,
)
from .translate_classes import TRANSLATE_CLASSES

__all__ = [
    "NpzSerializer", "Savepoint", "write_savepoint", "TranslateFortranData2Py", "TranslateGrid",
    "compare_arr", "load_threshold_overrides", "process_override", "sample_wherefail", "success",
    "success_array", "TRANSLATE_CLASSES",
]
