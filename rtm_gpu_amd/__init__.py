"""rtm_gpu_amd -- 2-D acoustic reverse-time migration on one AMD MI355X (gfx950).

  rtm_gpu_amd.hip    ctypes binding of librtm_hip.so, the C ABI of the HIP kernels
  rtm_gpu_amd.host   ctypes binding of librtm_host.so, the reference-style host driver
  rtm_gpu_amd.synth  seeded synthetic models for the benchmark workloads
  rtm_gpu_amd.build  in-tree build of the native libraries
"""
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")

__all__ = ["LIB_DIR"]
