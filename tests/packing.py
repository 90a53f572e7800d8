"""Packed input rows for the sub-byte tests, built with numpy alone (not with
the oracle's bit loop), in the layout of src/DedispersionTest.cpp:221-243:
sample x of a row sits in byte x // (8 // bits), starting at bit
(x % (8 // bits)) * bits, least significant bit first."""
from __future__ import annotations

import numpy as np


def pack_rows(values: np.ndarray, bits: int, pitch: int) -> np.ndarray:
    """values [rows, n] (each < 2**bits) -> packed bytes [rows, pitch]."""
    assert bits in (1, 2, 4, 8)
    values = np.asarray(values, dtype=np.uint8)
    rows, n = values.shape
    assert int(values.max(initial=0)) < (1 << bits)
    if bits == 8:
        out = np.zeros((rows, pitch), np.uint8)
        out[:, :n] = values
        return out
    per = 8 // bits
    nb = -(-n // per)
    assert nb <= pitch
    v = np.zeros((rows, nb * per), np.uint8)
    v[:, :n] = values
    if bits == 1:
        packed = np.packbits(v, axis=1, bitorder="little")
    else:
        fields = v.reshape(rows, nb, per).astype(np.uint16)
        packed = (fields << (bits * np.arange(per, dtype=np.uint16))).sum(axis=2).astype(np.uint8)
    out = np.zeros((rows, pitch), np.uint8)
    out[:, :nb] = packed
    return out


def unpacked_pitch(n: int, padding: int) -> int:
    return n if padding == 0 else -(-n // padding) * padding


def random_values(rng: np.random.Generator, rows: int, n: int, bits: int) -> np.ndarray:
    """Uniform over the whole field, top bit included."""
    return rng.integers(0, 1 << bits, size=(rows, n), dtype=np.uint8)
