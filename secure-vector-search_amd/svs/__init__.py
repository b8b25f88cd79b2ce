"""gfx950 backend for the encrypted-score matrix-vector path of
sacs-epfl/secure-vector-search (BN flat / BN-IVF over F_p, p = 2^61 - 1,
with Protocol-2 Freivalds verification, EMVP's score GEMV and Tiptoe's
SimplePIR answer). Native code: libsvs_gpu.so (include/svs_gpu.h)."""
from ._lib import GpuError, BudgetTooSmall, lib, LIB_PATH, EXPORTS  # noqa: F401
from .bntm import (  # noqa: F401
    BnTmConfig, BnTmError, BnTmHandle, BnTmIvfConfig, BnTmIvfHandle, BnTmIvfScorer, BnTmScorer,
    BuildOutcome, CommunicationCost, Context, DeviceArray, DimensionMismatch, DimensionTooLarge, EmptyCorpus,
    GpuIndex, Hit, InvalidNprobe, IvfIndex, TooManyCentroids, VerificationFailed, default_context,
    derive_h_seed, domain_separate, fp_to_signed, generate_subspace, probe_route, quantise,
    regenerate_h_s, cache_fingerprint64, load_bntm_ivf_cache,
)
from .serve import Server, ServedResult  # noqa: F401,E402
