"""MI355X (gfx950) warped TSDF integration: a drop-in for VolumeFusion's
per-frame integration path (knn node lookup, warp blending, projective TSDF
update) behind the reference's af:: entry points."""
from ._lib import VFError, load  # noqa: F401

__all__ = ["VFError", "load"]
