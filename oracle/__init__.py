"""FP64 CPU oracle of fd2d_noise's solvers (TEST INFRASTRUCTURE ONLY).

Loaded only by tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py, as the
checker. The product path (fd2d_noise_amd) never imports it.
"""
from .oracle import load, solvers, set_threads, primitives  # noqa: F401
