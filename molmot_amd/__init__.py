"""molmot_amd: MI355X-native (gfx950 HIP) solver path of molmot.

Subpackages mirror the reference layout: ``molecules`` (tables),
``obe`` (beams, rate equations), ``mot`` (simulators, force scans),
``propagation`` (trajectories) and ``hip`` (the batched device API that
molmot_cuda provides in the reference).
"""
__version__ = "0.1.0"
