"""gsevt: MI355X-native (gfx950 HIP) rasterizer and event-tracking loop of GS-EVT.

Importing the package does not touch the GPU; the HIP library is loaded on
first use (gsevt._lib.load) and there is no CPU fallback.
"""
__all__ = ["camera", "lie", "model", "raster", "renderer", "frame", "tracker", "synthetic", "workloads"]
