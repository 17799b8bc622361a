"""Reads datasets of an HDF5 (JLD2) file through the HDF5 project's own
h5dump and h5ls tools, the independent reader the JLD2 writer is checked
against.  The tools are looked up on PATH and in /opt/conda/bin; tests skip
when they are absent."""
import os
import re
import shutil
import subprocess

import numpy as np


def tool(name):
    return shutil.which(name) or (f"/opt/conda/bin/{name}" if os.path.exists(f"/opt/conda/bin/{name}") else None)


def available():
    return tool("h5dump") is not None and tool("h5ls") is not None


def listing(path):
    """{object path: 'Group' or the dataset's dimensions (tuple, () for a scalar)}."""
    out = subprocess.run([tool("h5ls"), "-r", path], capture_output=True, text=True, check=True).stdout
    objs = {}
    for line in out.splitlines():
        name, kind = line.split(None, 1)
        kind = kind.strip()
        if kind == "Group":
            objs[name] = "Group"
        elif kind.startswith("Dataset"):
            inner = kind[kind.index("{") + 1:kind.index("}")]
            objs[name] = () if inner == "SCALAR" else tuple(int(v) for v in inner.split(","))
    return objs


def dataset(path, name):
    """(values in file order as float64, HDF5 dims, datatype name) of one dataset."""
    out = subprocess.run([tool("h5dump"), "-y", "-w", "0", "-m", "%.17g", "-d", name, path], capture_output=True,
                         text=True, check=True).stdout
    dtype = re.search(r"DATATYPE\s+(\S+)", out).group(1)
    space = re.search(r"DATASPACE\s+(SCALAR|SIMPLE \{ \( ([^)]*) \))", out)
    dims = () if space.group(1) == "SCALAR" else tuple(int(v) for v in space.group(2).split(","))
    body = out[out.index("DATA {") + len("DATA {"):]
    body = body[:body.index("}")]
    vals = np.array([float(v) for v in body.replace(",", " ").split()], dtype=np.float64)
    return vals, dims, dtype


def julia_array(path, name):
    """The dataset as the Julia array JLD2 reads: HDF5 dims reversed, the
    file order column-major."""
    vals, dims, _ = dataset(path, name)
    if dims == ():
        return vals[0]
    return vals.reshape(tuple(reversed(dims)), order="F")
