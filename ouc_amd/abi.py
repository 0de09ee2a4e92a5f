"""ctypes mirror of include/ouc.h (the C ABI of the engine).

The structs and constants here must stay byte-compatible with the header;
tests/test_abi.py checks the field layout against the compiled library.
"""
import ctypes as C

OUC_OK = 0
OUC_INVALID_ARG = 1
OUC_OOM = 2
OUC_DEVICE = 3
OUC_NUMERIC_FALLBACK = 4
OUC_IO = 5
OUC_UNSUPPORTED = 6

STATUS_NAMES = {
    0: "OK", 1: "INVALID_ARG", 2: "OOM", 3: "DEVICE", 4: "NUMERIC_FALLBACK", 5: "IO", 6: "UNSUPPORTED",
}

SEM_FAITHFUL = 0
SEM_EXACT = 1
STRIP_FP32 = 0
STRIP_FP16 = 1
BOUNDARY_CPML = 0
BOUNDARY_LEARNED = 1
BOUNDARY_LEARNED_FALLBACK = 2
LAP_STENCIL = 0
LAP_FFT5 = 1
LAP_FFT_CONT = 2
ENGINE_PLAIN = 0
ENGINE_FUSED = 1
ENGINE_SPLIT_PML = 2


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("nz", C.c_int32),
        ("npml", C.c_int32), ("half_order", C.c_int32), ("nt", C.c_int32),
        ("dx", C.c_float), ("dz", C.c_float), ("dt", C.c_float),
        ("R", C.c_double),
        ("semantics", C.c_int32), ("strip_precision", C.c_int32),
        ("boundary", C.c_int32), ("laplacian", C.c_int32), ("engine", C.c_int32),
        ("band", C.c_int32),
    ]


class Model(C.Structure):
    _fields_ = [("vp", C.POINTER(C.c_float)), ("vs", C.POINTER(C.c_float)), ("rho", C.POINTER(C.c_float))]


class Shot(C.Structure):
    _fields_ = [
        ("sx", C.c_int32), ("sz", C.c_int32), ("rec_z", C.c_int32),
        ("fm", C.c_float), ("wavelet", C.POINTER(C.c_float)), ("mute", C.c_int32), ("nt", C.c_int32),
    ]


class Timings(C.Structure):
    _fields_ = [
        ("total_ms", C.c_double), ("forward_ms", C.c_double), ("backward_ms", C.c_double),
        ("image_ms", C.c_double), ("step_launches", C.c_int64), ("step_kernel_ms", C.c_double),
        ("step_kernel_launches", C.c_int64), ("cell_updates", C.c_double),
        ("canonical_bytes", C.c_double), ("own_bytes", C.c_double),
        ("fallback_shots", C.c_int32), ("reserved", C.c_int32), ("boundary_ratio", C.c_double),
    ]


class SegyGather(C.Structure):
    _fields_ = [
        ("shot", C.c_int32), ("ntraces", C.c_int32), ("nt", C.c_int32), ("dt", C.c_float),
        ("sx_m", C.c_float), ("sz_m", C.c_float), ("rec_x0_m", C.c_float), ("rec_dx_m", C.c_float),
        ("rec_z_m", C.c_float), ("component", C.c_char * 8),
    ]


def make_params(nx, nz, npml, half_order, nt, dx, dz, dt, R=1e-3, semantics=SEM_FAITHFUL,
                strip_precision=STRIP_FP32, boundary=BOUNDARY_CPML, laplacian=LAP_STENCIL,
                engine=ENGINE_PLAIN, band=0):
    return Params(int(nx), int(nz), int(npml), int(half_order), int(nt), float(dx), float(dz), float(dt),
                  float(R), int(semantics), int(strip_precision), int(boundary), int(laplacian), int(engine),
                  int(band))


def fptr(a):
    """float32 C-contiguous numpy array -> float*; None -> NULL."""
    if a is None:
        return C.POINTER(C.c_float)()
    assert a.dtype.name == "float32" and a.flags["C_CONTIGUOUS"], "need C-contiguous float32"
    return a.ctypes.data_as(C.POINTER(C.c_float))


def dptr(a):
    if a is None:
        return C.POINTER(C.c_double)()
    assert a.dtype.name == "float64" and a.flags["C_CONTIGUOUS"], "need C-contiguous float64"
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Tile(C.Structure):
    _fields_ = [("core_cols", C.c_int32), ("prefetch", C.c_int32)]
