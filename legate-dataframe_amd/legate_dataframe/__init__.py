"""MI355X (gfx950) hash join / groupby / hash partition with legate-dataframe's Python API.

Same entry points as the reference package (python/legate_dataframe/__init__.py):
``LogicalTable``, ``LogicalColumn``, ``lib.join.join``, ``lib.groupby_aggregation.groupby_aggregation``;
``lib.stream_compaction.distinct``, ``lib.parquet.parquet_read`` / ``parquet_write``; plus
``lib.hash_partition.hash_partition`` for the local step of the hash repartition.
The columns live in HBM; the operators run hand-written HIP kernels through the C ABI of
libldf_hip.so (include/ldf_hip.h) and raise if that library is missing.
"""

__version__ = "0.1.0"

# torch (device memory and streams) is imported before the native libraries: its HIP runtime
# then serves libldf_hip.so's libamdhip64 dependency, so one HIP runtime owns the GPU in this
# process (a second, separately loaded runtime would find no device).
import torch  # noqa: F401,E402

from legate_dataframe.lib.core.column import LogicalColumn  # noqa: F401,E402
from legate_dataframe.lib.core.table import LogicalTable  # noqa: F401,E402
