"""
zkgpu: a BN254 prover core for AMD Instinct MI355X (gfx950).

A drop-in for erihhh6/zk-gpu-accelerator's GPU path: BN254_FR.ntt in
zkgpu.field256_gpu, the G1 point kernels in zkgpu.curve_gpu and
msm_pippenger_gpu in zkgpu.msm_gpu keep their names, arguments, return types
and errors, and run hand-written HIP kernels through libzkgpu_hip.so's C ABI.

Modules
-------
gpu_backend  : loads the library, detects the GPU, DeviceArray, streams, events
field        : FIELD_MODULUS (BN254's scalar field r)
curve        : Point, G, P and the Jacobian <-> affine conversions (host types)
ntt          : root-of-unity derivation and validation (host)
msm          : Pippenger window model and validation (host)
field256_gpu : Field256, BN254_FR, BN254_FQ: Montgomery arithmetic and the NTT
curve_gpu    : element-wise G1 add / add_affine / double
msm_gpu      : msm_pippenger_gpu and the device-resident MSM
domain       : Radix2Domain with its transforms on the device
batch        : batched MSMs over a shared SRS and batched NTTs
serve        : continuous-batching proof-job server
workloads    : seeded synthetic inputs for the benchmark configurations

GPU modules are imported lazily: importing zkgpu itself loads no library.
"""

__version__ = "0.1.0"

from zkgpu.curve import G, P, Point  # noqa: E402
from zkgpu.field import FIELD_MODULUS  # noqa: E402

__all__ = ["FIELD_MODULUS", "G", "P", "Point", "__version__"]
