"""MI355X-native search backend for the Mars fractal image codec.

The package holds the HIP/CDNA4 range x domain x isometry search behind the
C ABI of ``include/mars_hip.h`` (``libmars_hip.so``) and the host-side mirror
of the reference's ``SearchProvider`` interface
(crates/mars-codec/src/encode.rs:530-602).
"""

from .provider import (Candidate, EncodeParams, HipSearchProvider, RawMoments, SearchOutcome,
                       SearchRequest)

__all__ = ["Candidate", "EncodeParams", "HipSearchProvider", "RawMoments", "SearchOutcome",
           "SearchRequest"]
