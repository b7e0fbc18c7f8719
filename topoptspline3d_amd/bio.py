"""Raw binary records in the reference's `bio` format (utils/binaryIO.h:70-111).

A record file is the bare array, native little-endian, no header:
  write_vector(path, v)        v[0], v[1], ...             (binaryIO.h:70-78)
  write_vectors(path, vs)      v0[0] v1[0] v2[0] v0[1] ... (N equal-length arrays,
                               component-interleaved; binaryIO.h:80-91)
  read_vector(path, dtype)     floor(file size / item size) items (binaryIO.h:93-109)

Records the reference's optimisation loops write (optimization/optimization.cpp:739-765,
grid/Grid.cpp:1238-1262, 2136-2170) and the writers here:
  crec_iter / cworst, vrec_iter / vrec, trec_iter / trec   float64 per iteration
  flast   float64 x3 per load node (force on the load region, interleaved)
  ulast   float64 x3 per node (displacement, interleaved)
  coeff   float32 per spline coefficient (x fastest)
Node order: the reference writes its internal colour-sorted node numbering; these writers
use the lexicographic order of the C ABI (load nodes in the order given).
"""
import os

import numpy as np


def write_vector(path, v, dtype=None):
    a = np.ascontiguousarray(v, dtype=dtype)
    with open(path, "wb") as fh:
        fh.write(a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes())


def write_vectors(path, vs, dtype=np.float64):
    vs = [np.asarray(v, dtype=dtype).ravel() for v in vs]
    n = vs[0].size
    if any(v.size != n for v in vs):
        raise ValueError("write_vectors: arrays of different lengths")
    write_vector(path, np.stack(vs, axis=1))


def read_vector(path, dtype=np.float64):
    dt = np.dtype(dtype).newbyteorder("<")
    raw = open(path, "rb").read()
    n = len(raw) // dt.itemsize
    return np.frombuffer(raw[:n * dt.itemsize], dtype=dt).astype(np.dtype(dtype))


def write_support_force(path, ctx, nodes, k=0):
    """flast: load case k's force on the given load nodes (HierarchyGrid::writeSupportForce)"""
    f = ctx.get_loads(k).reshape(3, -1)
    nodes = np.asarray(nodes, dtype=np.int64)
    write_vectors(path, [f[c, nodes] for c in range(3)])


def write_displacement(path, ctx, k=0):
    """ulast: load case k's displacement at every node (HierarchyGrid::writeDisplacement)"""
    u = ctx.displacement(k).reshape(3, -1)
    write_vectors(path, [u[0], u[1], u[2]])


def write_coeff(path, ctx):
    """coeff: the spline coefficients (HierarchyGrid::writeCoeff)"""
    write_vector(path, ctx.coeffs(), np.float32)


def write_history(out_dir, rec, final=False):
    """compliance / volume / time records of a loop (optimization/optimization.cpp:739-762):
    per iteration crec_iter, vrec_iter, trec_iter; at the end cworst, vrec, trec"""
    names = ("cworst", "vrec", "trec") if final else ("crec_iter", "vrec_iter", "trec_iter")
    for name, key in zip(names, ("c", "v", "ms")):
        write_vector(os.path.join(out_dir, name), rec[key], np.float64)
