"""ctypes binding of libsuperintervals_hip.so (the C-ABI declared in include/*.h).

The library is loaded from this package's folder.  There is no fallback: if the
shared library is missing or fails to load, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsuperintervals_hip.so")

SI_OK = 0
SI_ERR_ARG = 1
SI_ERR_HIP = 2
SI_ERR_NOT_BUILT = 3
SI_ERR_CAPACITY = 4
SI_ERR_NOMEM = 5
SI_ERR_UNSUPPORTED = 6
SI_ERR_MISMATCH = 7
SI_ORDER_DESC = 0
SI_ORDER_VECTOR = 1
SI_INDEX_NONE = 0xFFFFFFFF

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it first (make -C superintervals_amd/csrc, "
        "or __graft_entry__.build())")

# PyTorch for ROCm ships its own HIP runtime (same soname, libamdhip64.so.7, but loaded by
# torch under the name libamdhip64.so).  Loading torch first lets this library bind to
# that runtime; loading this library first would put a second HIP/HSA runtime into the
# process, and whichever initialises later can find no device.
try:
    import torch  # noqa: F401
except ImportError:
    pass

lib = C.CDLL(LIB_PATH)

_vp, _sz, _i, _u64p = C.c_void_p, C.c_size_t, C.c_int, C.POINTER(C.c_uint64)

_SIGS = {
    # si_device.h
    "siGetLastError": (C.c_char_p, []),
    "siVersion": (C.c_char_p, []),
    "siIndexCreate": (_vp, [_i]),
    "siIndexDestroy": (None, [_vp]),
    "siIndexBuild": (_i, [_vp, _vp, _vp, _vp, _sz, _i, _i, _i, _vp]),
    "siIndexSize": (_sz, [_vp]),
    "siIndexCoordBytes": (_i, [_vp]),
    "siIndexDownload": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "siIndexDeviceStarts": (_vp, [_vp]),
    "siIndexDeviceEnds": (_vp, [_vp]),
    "siIndexDeviceBranch": (_vp, [_vp]),
    "siIndexDevicePerm": (_vp, [_vp]),
    "siIndexLastBuildMs": (C.c_double, [_vp]),
    "siUpperBound": (_i, [_vp, _vp, _sz, _vp, _vp]),
    "siCount": (_i, [_vp, _vp, _vp, _sz, _vp, _vp]),
    "siHasOverlaps": (_i, [_vp, _vp, _vp, _sz, _vp, _vp]),
    "siSearchWorkspaceBytes": (_sz, [_sz]),
    "siSearchOffsets": (_i, [_vp, _vp, _vp, _sz, _vp, _vp, _vp, _sz, _vp]),
    "siSearchFill": (_i, [_vp, _vp, _vp, _sz, _i, _vp, _vp, _vp]),
    "siSearchValuesFill": (_i, [_vp, _vp, _vp, _sz, _vp, _vp, _vp]),
    "siCoverage": (_i, [_vp, _vp, _vp, _sz, _vp, _vp, _vp]),
    "siWalkStats": (_i, [_vp, _vp, _vp, _sz, _vp, _vp]),
    "siGatherKeys": (_i, [_vp, _vp, _sz, _vp, _vp, _vp, _vp]),
    "siIntersectPieces": (_i, [_vp, _vp, _vp, _vp, _sz, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "siDifferenceOffsets": (_i, [_vp, _vp, _vp, _sz, _vp, _vp, _vp, _vp, _sz, _vp]),
    "siDifferenceFill": (_i, [_vp, _vp, _vp, _vp, _sz, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "siLearnedCreate": (_vp, [_vp, C.c_uint32]),
    "siLearnedDestroy": (None, [_vp]),
    "siLearnedTrain": (_i, [_vp, _i, C.c_uint32, C.c_float, _vp]),
    "siLearnedParamCount": (_i, []),
    "siLearnedGetParams": (_i, [_vp, _vp]),
    "siLearnedNorm": (_i, [_vp, _vp, _vp]),
    "siLearnedInfo": (_i, [_vp, _vp, _vp, _vp]),
    "siLearnedGradient": (_i, [_vp, _vp, _vp, C.c_uint32, _vp]),
    "siUpperBoundLearned": (_i, [_vp, _vp, _sz, _vp, _vp, _vp]),
    "siTuneSet": (_i, [C.c_char_p, C.c_int64]),
    "siTuneGet": (C.c_int64, [C.c_char_p]),
    "siPlanBytes": (_sz, [_vp, _sz]),
    "siPlanQueries": (_i, [_vp, _vp, _vp, _sz, _vp, _sz, _vp]),
    "siCountPlanned": (_i, [_vp, _vp, _sz, _vp, _vp]),
    "siSearchOffsetsPlanned": (_i, [_vp, _vp, _sz, _vp, _vp, _vp]),
    "siSearchFillPlanned": (_i, [_vp, _vp, _sz, _i, _vp, _vp, _vp]),
    "siPlanCheck": (_i, [_vp, _vp, _vp]),
    "siSearchNoAlloc": (_i, [_vp, _vp, _vp, _sz, _i, _vp, _vp, _vp, C.c_uint64, _vp]),
    "siSetTableCreate": (_vp, [_i, C.c_uint32]),
    "siSetTableDestroy": (None, [_vp]),
    "siSetTablePut": (_i, [_vp, C.c_uint32, _vp]),
    "siSetTablePlanBytes": (_sz, [_vp, _sz]),
    "siSetTableSearchOffsets": (_i, [_vp, _vp, _vp, _vp, _sz, _vp, _vp, _vp, _sz, _vp]),
    "siSetTableSearchFill": (_i, [_vp, _vp, _sz, _i, _vp, _vp, C.c_uint64, _vp]),
    "siSearchValuesFillPlanned": (_i, [_vp, _vp, _sz, _vp, _vp, _vp]),
    "siSearchFillPlannedBounded": (_i, [_vp, _vp, _sz, _i, _vp, _vp, C.c_uint64, _vp]),
    "siSearchValuesFillPlannedBounded": (_i, [_vp, _vp, _sz, _vp, _vp, C.c_uint64, _vp]),
    "siCountHost": (_i, [_vp, _vp, _vp, _sz, _vp]),
    "siSearchHost": (_i, [_vp, _vp, _vp, _sz, _i, _vp, _vp, _vp, _sz, _vp]),
    "siUpperBoundHost": (_i, [_vp, _vp, _sz, _vp]),
    "siHasOverlapsHost": (_i, [_vp, _vp, _vp, _sz, _vp]),
    "siCoverageHost": (_i, [_vp, _vp, _vp, _sz, _vp, _vp]),
    "siWorkloadIntervals": (_i, [_i, C.c_uint64, _sz, _i, _i, _vp, _vp]),
    "siWorkloadQueries": (_i, [_i, C.c_uint64, _sz, _i, _vp, _vp]),
    # c_superintervals.h
    "createSuperIntervals": (_vp, []),
    "destroySuperIntervals": (None, [_vp]),
    "clearSuperIntervals": (None, [_vp]),
    "reserveSuperIntervals": (None, [_vp, _sz]),
    "addInterval": (None, [_vp, C.c_int32, C.c_int32, C.c_int32]),
    "sizeSuperIntervals": (_sz, [_vp]),
    "sortIntervals": (None, [_vp]),
    "indexSuperIntervals": (None, [_vp]),
    "intervalAt": (C.c_bool, [_vp, _sz, _vp]),
    "startAt": (C.c_int32, [_vp, _sz]),
    "endAt": (C.c_int32, [_vp, _sz]),
    "dataAt": (C.c_int32, [_vp, _sz]),
    "upperBound": (_sz, [_vp, C.c_int32]),
    "anyOverlaps": (C.c_bool, [_vp, C.c_int32, C.c_int32]),
    "countOverlaps": (_sz, [_vp, C.c_int32, C.c_int32]),
    "searchValues": (None, [_vp, C.c_int32, C.c_int32, _vp]),
    "searchIdxs": (None, [_vp, C.c_int32, C.c_int32, _vp]),
    "searchKeys": (None, [_vp, C.c_int32, C.c_int32, _vp]),
    "searchItems": (None, [_vp, C.c_int32, C.c_int32, _vp]),
    "searchPoint": (None, [_vp, C.c_int32, _vp]),
    "coverage": (None, [_vp, C.c_int32, C.c_int32, _vp, _vp]),
    "findOverlaps": (None, [_vp, C.c_int32, C.c_int32, _vp, _vp]),
    "clearIndexResult": (None, [_vp]),
    "destroyIndexResult": (None, [_vp]),
    "clearKeyResult": (None, [_vp]),
    "destroyKeyResult": (None, [_vp]),
    "clearItemResult": (None, [_vp]),
    "destroyItemResult": (None, [_vp]),
    "siCountBatch": (_i, [_vp, _vp, _vp, _sz, _vp]),
    "siSearchIdxsBatch": (_i, [_vp, _vp, _vp, _sz, _vp, _vp, _sz, _vp]),
    "siSearchValuesBatch": (_i, [_vp, _vp, _vp, _sz, _vp, _vp, _sz, _vp]),
    "intersection": (_vp, [_vp, _vp, _vp]),
    "difference": (_vp, [_vp, _vp]),
}

for _name, (_res, _args) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


class cIndexResult(C.Structure):
    """c_superintervals.h:98-102"""
    _fields_ = [("data", C.POINTER(C.c_int32)), ("size", C.c_size_t), ("capacity", C.c_size_t)]


class KeyPair(C.Structure):
    _fields_ = [("start", C.c_int32), ("end", C.c_int32)]


class Interval(C.Structure):
    _fields_ = [("start", C.c_int32), ("end", C.c_int32), ("data", C.c_int32)]


class cKeyResult(C.Structure):
    _fields_ = [("data", C.POINTER(KeyPair)), ("size", C.c_size_t), ("capacity", C.c_size_t)]


class cItemResult(C.Structure):
    _fields_ = [("data", C.POINTER(Interval)), ("size", C.c_size_t), ("capacity", C.c_size_t)]


class cSuperIntervals(C.Structure):
    """c_superintervals.h:81-91 plus the trailing device handle."""
    _fields_ = [("starts", C.POINTER(C.c_int32)), ("ends", C.POINTER(C.c_int32)),
                ("data", C.POINTER(C.c_int32)), ("branch", C.POINTER(C.c_size_t)),
                ("size", C.c_size_t), ("capacity", C.c_size_t), ("idx", C.c_size_t),
                ("startSorted", C.c_bool), ("endSorted", C.c_bool), ("dev", C.c_void_p)]


class SIError(RuntimeError):
    def __init__(self, code: int, where: str):
        msg = lib.siGetLastError().decode(errors="replace")
        super().__init__(f"{where} failed with code {code}: {msg}")
        self.code = code


def check(rc: int, where: str) -> None:
    if rc != SI_OK:
        raise SIError(rc, where)


def exported_symbols() -> list:
    return sorted(_SIGS)
