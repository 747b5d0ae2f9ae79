"""bulk_amd -- MI355X (gfx950) algorithm layer: merge, merge_sort_by_key,
inclusive_scan, reduce, reduce_by_key and reduce_intervals behind a C-ABI
(include/bulkx.h), with a batch executor and a serving loop.

The Python modules are the host-side mirror of the reference drivers and
the harness around the native library; the work itself runs in
libbulkx.so's HIP kernels.
"""
from . import _capi  # noqa: F401
from ._capi import BulkError, BulkLibraryMissing  # noqa: F401

__all__ = ["BulkError", "BulkLibraryMissing", "api"]


def __getattr__(name):
    # api imports torch; load it lazily so the C-ABI checks run without it.
    if name == "api":
        import importlib
        return importlib.import_module(__name__ + ".api")
    raise AttributeError(name)
