"""q_amd -- MI355X-native engine for Q's qdyn MD step (FEP/EVB/LIE).

The compute path lives in the HIP library ``libqmi.so`` (built from
``q_amd/csrc``); ``q_amd.engine`` binds its C-ABI with ctypes.  Importing the
package does not load the library; creating an engine does, and fails loudly
when the library is missing.
"""
__all__ = ["synth"]
