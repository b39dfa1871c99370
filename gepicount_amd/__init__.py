"""gepicount_amd -- MI355X-native all-pairs Boolean epistasis scan (gEpiCount's hot path).

The compute lives in libgepicount.so (include/gepicount.h, gepicount_amd/csrc/);
this package is its ctypes binding plus a synthetic PLINK panel generator.
"""
from .api import (GEPI_FFUNC_ENTROPY, GEPI_FFUNC_FISHERPV, FUNC_NAMES, RESULT_DTYPE, Engine,
                  GepiError, Panel, PlinkData, Server, device_count, fisher_pval, lib,
                  process_gpu_1, read_plink, score_table)

__all__ = ["GEPI_FFUNC_ENTROPY", "GEPI_FFUNC_FISHERPV", "FUNC_NAMES", "RESULT_DTYPE", "Engine",
           "GepiError", "Server",
           "Panel", "PlinkData", "device_count", "fisher_pval", "lib", "process_gpu_1",
           "read_plink", "score_table"]
