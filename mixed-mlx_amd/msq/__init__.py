"""msq: MI355X backend of the mixed-precision SPD square root.

``msq.capi`` binds libmsq_hip.so (include/msq_hip.h); ``msq.generators``
builds the seeded synthetic inputs; ``msq.solver`` holds the host-side solver
configurations (plain FP32 path, low-bit tiers, tier selection).
"""
