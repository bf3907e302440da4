"""Helpers shared by the tests: conversion of ParticleArrays to the
oracle's FP64 dicts and error measures."""
import numpy as np

FIELDS = ("x", "y", "z", "u", "v", "w", "h", "m", "rho", "p", "cs")


def to_oracle(pa, ptype=None):
    """FP64 copy (user order) of a ParticleArray for the oracle."""
    from oracle import oracle as O
    d = {f: pa.get(f, only_real_particles=False).astype(np.float64) for f in FIELDS}
    if pa.has_array("tag"):
        d["tag"] = pa.get("tag", only_real_particles=False).astype(np.int64)
    return O.make_array(ptype=pa.particle_type if ptype is None else ptype, **d)


def rel_norm(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def max_rel(a, b, floor=0.0):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))) if len(a) else 0.0


def csr_sets(off, nbr):
    return [frozenset(nbr[off[i]:off[i + 1]].tolist()) for i in range(len(off) - 1)]


def csr_sorted_rows(off, nbr):
    return [np.sort(nbr[off[i]:off[i + 1]]) for i in range(len(off) - 1)]
