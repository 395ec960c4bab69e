"""Read VTKHDF files back through the HDF5 project's own h5dump tool, so that the files the
library writes are checked by a reader that shares no code with the writer.

Test infrastructure only.
"""
import os
import re
import shutil
import subprocess
import tempfile

import numpy as np

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"

_NUMPY = {"H5T_IEEE_F64LE": "<f8", "H5T_IEEE_F32LE": "<f4", "H5T_STD_I64LE": "<i8", "H5T_STD_U8LE": "u1",
          "H5T_STD_I32LE": "<i4"}


def available():
    return os.path.exists(H5DUMP)


def _run(*args):
    p = subprocess.run([H5DUMP, *args], capture_output=True, text=True, cwd=tempfile.gettempdir())
    if p.returncode != 0:
        raise RuntimeError(f"h5dump {' '.join(args)} failed: {p.stderr.strip()}")
    return p.stdout


def contents(path):
    """{object path: kind} for every group, dataset and attribute of the file."""
    out = {}
    for line in _run("-n", "1", path).splitlines():
        m = re.match(r"\s*(group|dataset|attribute)\s+(\S+)\s*$", line)
        if m:
            out[m.group(2)] = m.group(1)
    return out


def _header(text):
    """(type, dims) from the DATATYPE / DATASPACE lines of an h5dump header; dims is None
    for a scalar dataspace; a fixed-length string type is ("string", size, pad, cset)."""
    if "H5T_STRING" in text:
        size = int(re.search(r"STRSIZE (\d+);", text).group(1))
        pad = re.search(r"STRPAD (\S+);", text).group(1)
        cset = re.search(r"CSET (\S+);", text).group(1)
        ty = ("string", size, pad, cset)
    else:
        ty = re.search(r"DATATYPE\s+(\S+)", text).group(1)
    if re.search(r"DATASPACE\s+SCALAR", text):
        dims = None
    else:
        dims = tuple(int(x) for x in re.search(r"DATASPACE\s+SIMPLE \{ \( ([^)]*) \)", text).group(1).split(","))
    return ty, dims


def dataset_header(path, name):
    return _header(_run("-H", "-d", name, path))


def attribute(path, name):
    """(type, dims, text of the values) of an attribute."""
    text = _run("-a", name, path)
    ty, dims = _header(text)
    data = re.search(r"DATA \{\s*\(0\): (.*?)\s*\}", text, re.S).group(1)
    return ty, dims, data


def dataset(path, name):
    """The values of a dataset, shaped by its dataspace, read as raw little-endian bytes."""
    ty, dims = dataset_header(path, name)
    fd, tmp = tempfile.mkstemp(suffix=".bin")
    os.close(fd)
    try:
        _run("-d", name, "-b", "LE", "-o", tmp, path)
        a = np.fromfile(tmp, dtype=_NUMPY[ty])
    finally:
        os.remove(tmp)
    return a.reshape(() if dims is None else dims)
