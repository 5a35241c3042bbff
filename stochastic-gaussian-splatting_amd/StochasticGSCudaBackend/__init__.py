"""MI355X-native stochastic Gaussian splatting backend.

Drop-in for the reference package ``StochasticGSCudaBackend`` (its __init__.py:14-15 and
StochasticGSCudaBackend/__init__.py:1-2 export the same five names). Importing fails with
ImportError when libsgs.so has not been built.
"""
from .render import History, Settings, TemporalMode, Tuning, blue_noise, create_k_buffer, render

__all__ = ["render", "Settings", "TemporalMode", "History", "create_k_buffer", "Tuning", "blue_noise"]
