"""gamma_smc_cu_amd: MI355X-native (gfx950) Gamma-SMC pairwise TMRCA decoder.

Drop-in for the plain method of ``gamma_smc_cu``: :func:`infer` and
:func:`infer_blockwise` keep the reference's signatures and result layout,
and ``_core`` exposes the same native surface (``FlowContext``,
``cuda_mem_info``, ``bitpack``, ``unpack``, ``get_device_count``,
``get_device``, ``set_device``, ``gamma_smc_flow_cached_fb``) backed by hand-written HIP
kernels through a C ABI (``include/gamma_smc_amd.h``).
"""

from . import _core
from ._core import (FlowContext, bitpack, cuda_mem_info, gamma_smc_flow_cached_fb, gamma_smc_flow_cached_fwd,
                    unpack)
from .genes import gene_stats
from .infer import infer, infer_blockwise

__all__ = ["FlowContext", "bitpack", "cuda_mem_info", "gamma_smc_flow_cached_fb", "gamma_smc_flow_cached_fwd",
           "gene_stats", "infer", "infer_blockwise", "unpack"]
