"""MI355X-native overset connectivity engine with TIOGA's C ABI.

The product is the shared library ``libtioga_amd.so`` built from csrc/ by
:mod:`tioga_amd.build`; :mod:`tioga_amd.tioga` is a Python handle on its C
ABI and :mod:`tioga_amd.workloads` generates the synthetic overset meshes.
"""
__all__ = ["build", "tioga", "workloads"]
