"""warpstate_amd — MI355X (gfx950) GPU scan engine with warpstate's scan API.

The compute path is libwarpstate_amd.so (hand-written HIP kernels behind the
C ABI in include/warpstate.h); this package is the Python mirror of the
reference crate's API over that ABI.
"""
from ._native import (  # noqa: F401
    MATCH_DTYPE,
    MODE_OVERLAPPING,
    MODE_STANDARD,
    TAGGED_DTYPE,
    VARIANT_AUTO,
    VARIANT_DFA_GLOBAL,
    VARIANT_DFA_LDS,
    VARIANT_HASH,
    WarpstateError,
)
from .api import (  # noqa: F401
    AutoMatcherConfig,
    BatchPlan,
    CompiledPatternIndex,
    Context,
    GpuMatcher,
    PendingScan,
    PreparedScan,
    HotSwapPatternSet,
    Match,
    PatternSet,
    PatternSetBuilder,
    ScanItem,
    StreamPipeline,
    StreamScanner,
    TaggedMatch,
    Tuning,
    WarpStateer,
    coalesce,
    default_context,
    matches_to_list,
    scan_batch,
    scan_batch_device,
    scan_batch_gpu,
    scan_batch_planned,
)

Error = WarpstateError
