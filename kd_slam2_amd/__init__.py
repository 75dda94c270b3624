"""kd_slam2_amd: MI355X-native scan-alignment hot path of KD-SLAM.

The GPU path lives in libkdhip.so (kd_slam2_amd/csrc, C ABI in include/kdh.h).
This package holds the host-side mirror of the reference aligner interface
(api.py), the seeded synthetic workloads (synth.py) and the build helper.
"""
from . import kdtree_msg  # noqa: F401
from .api import (  # noqa: F401
    CTICPHIP3D,
    Context,
    DeviceTree,
    FixedEntry,
    HostTree,
    ICPHIP3D,
    Server,
    TreeGenerator3D,
    BatchResult,
    PairBatch,
    adapter_align,
    batch_align,
    batch_build,
    batch_build_dual,
    default_params,
    update_cache_f32,
)
