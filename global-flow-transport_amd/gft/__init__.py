"""gft: MI355X-native rendering/transport kernels with the op interface of
tum-pbs/Global-Flow-Transport (phitest/render/cuda/ops_loader.py).

The HIP library is loaded lazily on the first op call; ``gft.ops`` holds the
op namespaces (``raymarching_ops``, ``sampling_ops``, ``blending_ops``).
"""
__all__ = ["ops", "camera", "transform", "synthetic"]
