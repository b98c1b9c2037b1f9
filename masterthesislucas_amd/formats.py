"""On-disk and wire formats used by the thesis tooling around the t-SNE / Barnes-Hut path.

Readers and writers for the files the reference's harness consumes and produces, so data and
results can move between the two:

- MNIST IDX images (``Loader::loadMNIST``, dataLoaders/loader.h:14-41): big-endian header
  (magic, count, rows, cols), one byte per pixel, read as float32 rows of rows*cols.
- label files (``Loader::loadLabels``, dataLoaders/loader.h:44-53): the raw bytes, one uint8
  label per point.
- the joint-probability matrix P in Matrix Market coordinate form (``Loader::loadPmatrix`` via
  Eigen ``loadMarket``, dataLoaders/loader.h:55-75), returned as the CSR sorted by (row, col)
  that the t-SNE step uses (``setupEfficientPmatrix``, nBodyInstances/tsne.h:462-520).
- ``graphCSV`` result rows (``writeToFile``/``writeToFile3``/``writeToFileN``,
  nBodyInstances/tsneTests.h:1458-1520): one line per sample, values joined by ``,`` in the
  default ``std::ostream`` formatting (6 significant digits), and the file names the harness
  builds with ``fltToStr`` (tsneTests.h:1522-1536).
- embedding state dumps. The reference writes raw ``TsnePoint2D`` records
  (``storePrecomputed``, tsneTests.h:1550-1574), whose layout depends on the compiler; here the
  state is a versioned structure-of-arrays file instead (see ``save_state``).
"""
from __future__ import annotations

import io
import os
import struct

import numpy as np

__all__ = [
    "load_mnist_idx", "save_mnist_idx", "load_labels", "save_labels",
    "load_pmatrix_mtx", "save_pmatrix_mtx", "ostream_str", "write_graph_csv",
    "read_graph_csv", "flt_to_str", "graph_csv_name", "save_state", "load_state",
]


# ---------------------------------------------------------------------------------------------
# MNIST IDX and labels
def load_mnist_idx(path, max_amount=None):
    """Images of an IDX3 file as float32 [count, rows*cols] (pixel bytes converted to float,
    like Loader::loadMNIST); at most ``max_amount`` images."""
    with open(path, "rb") as f:
        head = f.read(16)
        if len(head) < 16:
            raise ValueError(f"{path}: truncated IDX header")
        magic, count, rows, cols = struct.unpack(">IIII", head)
        if magic != 0x00000803:
            raise ValueError(f"{path}: not an IDX3 ubyte file (magic {magic:#010x})")
        if max_amount is not None:
            count = min(count, int(max_amount))
        raw = np.frombuffer(f.read(count * rows * cols), dtype=np.uint8)
    if raw.size != count * rows * cols:
        raise ValueError(f"{path}: file shorter than its header says")
    return raw.astype(np.float32).reshape(count, rows * cols)


def save_mnist_idx(path, images, rows, cols):
    """Write uint8-valued images [count, rows*cols] as an IDX3 file."""
    a = np.asarray(images)
    if a.ndim != 2 or a.shape[1] != rows * cols:
        raise ValueError("images must be [count, rows*cols]")
    if np.any(a < 0) or np.any(a > 255):
        raise ValueError("IDX ubyte images hold values 0..255")
    with open(path, "wb") as f:
        f.write(struct.pack(">IIII", 0x00000803, a.shape[0], rows, cols))
        f.write(a.astype(np.uint8).tobytes())


def load_labels(path):
    """All bytes of a label file as uint8 (Loader::loadLabels keeps the whole file)."""
    with open(path, "rb") as f:
        return np.frombuffer(f.read(), dtype=np.uint8).copy()


def save_labels(path, labels):
    with open(path, "wb") as f:
        f.write(np.asarray(labels, dtype=np.uint8).tobytes())


# ---------------------------------------------------------------------------------------------
# Matrix Market P
def load_pmatrix_mtx(path):
    """Read a Matrix Market coordinate matrix (real, integer or pattern; general or symmetric;
    1-based indices, as Eigen's loadMarket reads it) and return
    (n, row_ptr uint32 [n+1], col uint32 [nnz], val float32 [nnz]) with entries sorted by
    (row, col). Duplicate (row, col) entries are summed. A symmetric file lists one triangle;
    the mirrored entries are added."""
    with open(path, "r") as f:
        banner = f.readline().split()
        if len(banner) < 5 or banner[0].lower() != "%%matrixmarket" or banner[1].lower() != "matrix":
            raise ValueError(f"{path}: not a Matrix Market matrix file")
        if banner[2].lower() != "coordinate":
            raise ValueError(f"{path}: only the coordinate format is supported")
        field, symm = banner[3].lower(), banner[4].lower()
        if field not in ("real", "double", "integer", "pattern"):
            raise ValueError(f"{path}: unsupported field '{field}'")
        if symm not in ("general", "symmetric"):
            raise ValueError(f"{path}: unsupported symmetry '{symm}'")
        line = f.readline()
        while line and line.lstrip().startswith("%"):
            line = f.readline()
        nrows, ncols, nnz = (int(v) for v in line.split())
        if nrows != ncols:
            raise ValueError(f"{path}: P must be square ({nrows} x {ncols})")
        body = f.read()
    cols_expected = 2 if field == "pattern" else 3
    tok = np.array(body.split(), dtype=np.float64) if body.strip() else np.zeros(0)
    if tok.size != nnz * cols_expected:
        raise ValueError(f"{path}: expected {nnz} entries")
    tok = tok.reshape(nnz, cols_expected)
    r = tok[:, 0].astype(np.int64) - 1
    c = tok[:, 1].astype(np.int64) - 1
    v = np.ones(nnz) if field == "pattern" else tok[:, 2]
    if nnz and (r.min() < 0 or c.min() < 0 or r.max() >= nrows or c.max() >= ncols):
        raise ValueError(f"{path}: index out of range")
    if symm == "symmetric":
        off = r != c
        r, c, v = np.concatenate([r, c[off]]), np.concatenate([c, r[off]]), np.concatenate([v, v[off]])
    return _to_csr(nrows, r, c, v)


def _to_csr(n, r, c, v):
    key = r * n + c
    order = np.argsort(key, kind="stable")
    key, v = key[order], v[order]
    uk, start = np.unique(key, return_index=True)
    vals = np.add.reduceat(v, start) if v.size else v
    rows, cols = uk // n, uk % n
    row_ptr = np.zeros(n + 1, np.int64)
    np.add.at(row_ptr, rows + 1, 1)
    return n, np.cumsum(row_ptr).astype(np.uint32), cols.astype(np.uint32), vals.astype(np.float32)


def save_pmatrix_mtx(path, row_ptr, col, val, n=None, precision=9):
    """Write CSR P as a Matrix Market 'coordinate real general' file (1-based)."""
    row_ptr = np.asarray(row_ptr, np.int64)
    n = row_ptr.size - 1 if n is None else int(n)
    col = np.asarray(col, np.int64)
    val = np.asarray(val, np.float64)
    rows = np.repeat(np.arange(n), np.diff(row_ptr))
    buf = io.StringIO()
    buf.write("%%MatrixMarket matrix coordinate real general\n")
    buf.write(f"{n} {n} {col.size}\n")
    fmt = f"%d %d %.{precision}g\n"
    for r, c, v in zip(rows + 1, col + 1, val):
        buf.write(fmt % (r, c, v))
    with open(path, "w") as f:
        f.write(buf.getvalue())


# ---------------------------------------------------------------------------------------------
# graphCSV rows
def ostream_str(x):
    """A number as ``std::ostream << x`` prints it with the default flags and precision 6."""
    if isinstance(x, (int, np.integer)) and not isinstance(x, bool):
        return str(int(x))
    x = float(x)
    if np.isnan(x):
        return "nan" if not np.signbit(x) else "-nan"
    if np.isinf(x):
        return "inf" if x > 0 else "-inf"
    return "%g" % x


def write_graph_csv(path, *columns):
    """One line per index i: columns[0][i],columns[1][i],... (writeToFileN: the shortest
    column sets the number of lines)."""
    if not columns:
        raise ValueError("at least one column")
    m = min(len(c) for c in columns)
    with open(path, "w") as f:
        for i in range(m):
            f.write(",".join(ostream_str(c[i]) for c in columns) + "\n")


def read_graph_csv(path):
    """Columns of a graphCSV file as float64 arrays."""
    rows = [line.strip().split(",") for line in open(path) if line.strip()]
    if not rows:
        return []
    return [np.array([float(r[k]) for r in rows]) for k in range(len(rows[0]))]


def flt_to_str(f):
    """fltToStr: fixed notation (6 decimals, of the float32 value), trailing zeros and a
    trailing point removed."""
    s = "%f" % float(np.float32(f))
    s = s.rstrip("0")
    return s[:-1] if s.endswith(".") else s


def graph_csv_name(kind, method, n, perplexity, dataset, theta=None):
    """File name of a harness result, e.g. kind 'tsneErrorTimestep' with theta, or
    'tsneCalculationtimeTheta' without (tsneTests.h:265-285, 504-524)."""
    name = f"{kind}{method}_point{int(n)}_perp{int(perplexity)}"
    if theta is not None:
        name += f"_theta{flt_to_str(theta)}"
    return name + f"_dataset{dataset}.csv"


# ---------------------------------------------------------------------------------------------
# embedding state dumps
_STATE_MAGIC = b"MTLSTATE"
_STATE_VERSION = 1


def save_state(path, iteration, y, y_prev=None, labels=None):
    """Versioned SoA dump of a t-SNE state: header (magic, version, n, iteration, flags), then
    x[n], y[n] of the current embedding (float32), optionally the previous embedding (for the
    momentum term) and uint8 labels."""
    y = np.asarray(y, np.float32).reshape(-1, 2)
    n = y.shape[0]
    flags = (1 if y_prev is not None else 0) | (2 if labels is not None else 0)
    with open(path, "wb") as f:
        f.write(_STATE_MAGIC)
        f.write(struct.pack("<IQqI", _STATE_VERSION, n, int(iteration), flags))
        f.write(np.ascontiguousarray(y[:, 0]).tobytes())
        f.write(np.ascontiguousarray(y[:, 1]).tobytes())
        if y_prev is not None:
            yp = np.asarray(y_prev, np.float32).reshape(n, 2)
            f.write(np.ascontiguousarray(yp[:, 0]).tobytes())
            f.write(np.ascontiguousarray(yp[:, 1]).tobytes())
        if labels is not None:
            lab = np.asarray(labels, np.uint8)
            if lab.size != n:
                raise ValueError("one label per point")
            f.write(lab.tobytes())


def load_state(path):
    """dict(iteration, y [n,2], y_prev [n,2] or None, labels or None) from save_state."""
    with open(path, "rb") as f:
        if f.read(len(_STATE_MAGIC)) != _STATE_MAGIC:
            raise ValueError(f"{path}: not a state file")
        version, n, iteration, flags = struct.unpack("<IQqI", f.read(24))
        if version != _STATE_VERSION:
            raise ValueError(f"{path}: state version {version} (this reader: {_STATE_VERSION})")

        def arr(dtype, count):
            b = f.read(count * np.dtype(dtype).itemsize)
            if len(b) != count * np.dtype(dtype).itemsize:
                raise ValueError(f"{path}: truncated")
            return np.frombuffer(b, dtype=dtype).copy()

        y = np.stack([arr(np.float32, n), arr(np.float32, n)], 1)
        y_prev = np.stack([arr(np.float32, n), arr(np.float32, n)], 1) if flags & 1 else None
        labels = arr(np.uint8, n) if flags & 2 else None
    return dict(iteration=int(iteration), y=y, y_prev=y_prev, labels=labels)
