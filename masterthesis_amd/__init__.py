"""MI355X-native engine for ScaFES's explicit time-stepping hot path.

The compute path is the HIP library masterthesis_amd/lib/libscafes_hip.so
(C ABI declared in include/scafes_hip.h); the C++11 host API mirroring
ScaFES's Problem/DataField/GridGlobal/GridSub lives in
masterthesis_amd/host/ScaFES.  `capi` is the ctypes binding used by the
Python harness.
"""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libscafes_hip.so")
HOST_INCLUDE = os.path.join(PKG_DIR, "host")
