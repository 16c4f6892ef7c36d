"""
zkref: the CPU oracle for the BN254 prover path (TEST INFRASTRUCTURE ONLY).

This package restates, on Python integers, the reference's CPU algorithms
for the accelerated path: the radix-2 NTT and its root derivation
(src/zkgpu/ntt.py), the evaluation domain (src/zkgpu/domain.py), the BN254
G1 group law in affine and Jacobian coordinates (src/zkgpu/curve.py) and
the MSM algorithms including Pippenger (src/zkgpu/msm.py). It is written
from the mathematics and the reference's documented behaviour, not copied.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import it, and only as the checker or the timed CPU baseline. The product
package (zkgpu) never imports it and has no CPU fallback.

A faster C restatement of the same algorithms lives in oracle/c and is
reached through zkref.fast; it is itself checked against this package.
"""
