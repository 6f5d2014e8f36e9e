"""HDF5 + XDMF output and restart of a state (ctypes view of libg5h5.so, include/godunov_five_eq_io.h).

write() stores U[block][var][cell] as one FP64 dataset per conservative variable in global cell
order, with the run time and iteration, and an .xmf next to it; read() restores the block layout
bit for bit.  The variable names are the reference's (models/FiveEq.h:93-135).
"""
import ctypes
import os

import numpy as np

from . import G5_OK, G5Config, G5Error, nvar_of, state_size

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libg5h5.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        d = ctypes.c_double
        L.g5_h5_write.restype = ctypes.c_int
        L.g5_h5_write.argtypes = [ctypes.c_char_p, ctypes.c_char_p, P(G5Config), P(d), ctypes.c_size_t, d,
                                  ctypes.c_int64, ctypes.c_uint32]
        L.g5_h5_read.restype = ctypes.c_int
        L.g5_h5_read.argtypes = [ctypes.c_char_p, P(G5Config), P(d), ctypes.c_size_t, P(d), P(ctypes.c_int64)]
        L.g5_h5_var_name.restype = ctypes.c_char_p
        L.g5_h5_var_name.argtypes = [ctypes.c_int, ctypes.c_int]
        L.g5_h5_last_error.restype = ctypes.c_char_p
        L.g5_h5_last_error.argtypes = []
        _lib = L
    return _lib


def exported_symbols():
    return ["g5_h5_write", "g5_h5_read", "g5_h5_var_name", "g5_h5_last_error"]


def var_names(dim):
    return [lib().g5_h5_var_name(dim, v).decode() for v in range(nvar_of(dim))]


def _check(rc):
    if rc != G5_OK:
        raise G5Error(rc, lib().g5_h5_last_error().decode())


def write(directory, basename, cfg, U, t, iteration, variables=None):
    """Write directory/basename.h5 and .xmf.  variables: names to write (None: all, a checkpoint)."""
    U = np.ascontiguousarray(U, dtype=np.float64).ravel()
    mask = 0
    if variables is not None:
        names = var_names(cfg.dim)
        for name in variables:
            if name not in names:
                raise ValueError(f"unknown variable {name!r}; known: {names}")
            mask |= 1 << names.index(name)
        if mask == 0:
            raise ValueError("no variable selected")
    _check(lib().g5_h5_write(os.fsencode(directory), os.fsencode(basename), ctypes.byref(cfg),
                             U.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), U.size, t, iteration, mask))
    return os.path.join(directory, basename + ".h5")


def read(path, cfg):
    """(U, t, iteration) from a checkpoint written by write() for the same mesh."""
    U = np.empty(state_size(cfg), dtype=np.float64)
    t = ctypes.c_double()
    it = ctypes.c_int64()
    _check(lib().g5_h5_read(os.fsencode(path), ctypes.byref(cfg), U.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                            U.size, ctypes.byref(t), ctypes.byref(it)))
    return U, t.value, it.value
