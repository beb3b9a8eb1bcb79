"""Reading scan data: the example data set's HDF5 file (examples/example_data.h5: "data_array" (N_Omega, My, N_eta,
N_rings) float32 intensities, "lattice" (3, 3) direct lattice vectors, "hkl_list" (rings, 3), "wavelength"; the
reference notebook reads it with h5py, example_diffractom_pipeline.ipynb cell 3) without an HDF5 library.

read_hdf5 walks the root group of an HDF5 file written with the original file format (superblock version 0, groups
as symbol tables: a version-1 B-tree of symbol-table nodes whose names live in a local heap; version-1 object
headers) and returns its datasets of contiguous or compact layout, with fixed-point or IEEE floating-point
elements of either byte order, as NumPy arrays (large ones as read-only memory maps of the file). Other layouts
(chunked, compressed), later superblock versions and nested groups raise NotImplementedError.
"""
from __future__ import annotations

import struct
from pathlib import Path

import numpy as np

_SIG = b"\x89HDF\r\n\x1a\n"


class _File:
    def __init__(self, path):
        self.path = Path(path)
        self.raw = np.memmap(self.path, dtype=np.uint8, mode="r")
        if bytes(self.raw[:8]) != _SIG:
            raise ValueError(f"{self.path}: not an HDF5 file")
        if int(self.raw[8]) != 0:
            raise NotImplementedError(f"{self.path}: HDF5 superblock version {int(self.raw[8])} (only version 0)")
        self.so, self.sl = int(self.raw[13]), int(self.raw[14])          # sizes of offsets and lengths

    def bytes(self, at, n):
        return bytes(self.raw[at:at + n])

    def u(self, at, n):
        return int.from_bytes(self.bytes(at, n), "little")

    def root_entry(self):
        """The root group's symbol-table entry (after the superblock's 4 addresses)."""
        return 24 + 4 * self.so

    # ------------------------------------------------------------------ groups
    def group_members(self, entry):
        """{name: object header address} of the group whose symbol-table entry is at `entry`."""
        so = self.so
        header = self.u(entry + so, so)
        cache = self.u(entry + 2 * so, 4)
        if cache == 1:                                # B-tree and heap addresses cached in the entry
            btree, heap = self.u(entry + 2 * so + 8, so), self.u(entry + 3 * so + 8, so)
        else:                                         # the symbol-table message (0x11) of the object header
            msg = self.messages(header).get(0x11)
            if msg is None:
                raise NotImplementedError("the root object is not a symbol-table group")
            btree, heap = self.u(msg, so), self.u(msg + so, so)
        if self.bytes(heap, 4) != b"HEAP":
            raise ValueError("bad local heap")
        heap_data = self.u(heap + 8 + 2 * self.sl, so)
        out = {}
        for snod in self._btree_leaves(btree):
            if self.bytes(snod, 4) != b"SNOD":
                raise ValueError("bad symbol-table node")
            n = self.u(snod + 6, 2)
            for i in range(n):
                e = snod + 8 + i * (2 * so + 24)
                name_off = self.u(e, so)
                end = name_off
                while self.raw[heap_data + end] != 0:
                    end += 1
                name = self.bytes(heap_data + name_off, end - name_off).decode()
                out[name] = self.u(e + so, so)
        return out

    def _btree_leaves(self, node):
        """The symbol-table nodes under a version-1 group B-tree node."""
        if self.bytes(node, 4) != b"TREE":
            raise ValueError("bad B-tree node")
        level, used = int(self.raw[node + 5]), self.u(node + 6, 2)
        so, sl = self.so, self.sl
        p = node + 8 + 2 * so                          # first key
        kids = []
        for i in range(used):
            kids.append(self.u(p + sl + i * (sl + so), so))
        if level == 0:
            return kids
        out = []
        for k in kids:
            out += self._btree_leaves(k)
        return out

    # ------------------------------------------------------------------ object headers
    def messages(self, header):
        """{message type: address of its data} of a version-1 object header (continuation blocks followed)."""
        if int(self.raw[header]) != 1:
            raise NotImplementedError(f"object header version {int(self.raw[header])} (only version 1)")
        n = self.u(header + 2, 2)
        blocks = [(header + 16, self.u(header + 8, 4))]
        out = {}
        seen = 0
        while blocks and seen < n:
            at, size = blocks.pop(0)
            end = at + size
            while at + 8 <= end and seen < n:
                mtype, msize = self.u(at, 2), self.u(at + 2, 2)
                data = at + 8
                if mtype == 0x10:                     # continuation: (offset, length)
                    blocks.append((self.u(data, self.so), self.u(data + self.so, self.sl)))
                else:
                    out.setdefault(mtype, data)
                seen += 1
                at = data + msize
        return out

    def dataset(self, header):
        m = self.messages(header)
        if 0x01 not in m or 0x03 not in m or 0x08 not in m:
            raise NotImplementedError("not a dataset (no dataspace, datatype or layout message)")
        shape = self._dataspace(m[0x01])
        dtype = self._datatype(m[0x03])
        count = int(np.prod(shape)) if shape else 1
        lay = m[0x08]
        version = int(self.raw[lay])
        if version == 3:
            cls = int(self.raw[lay + 1])
            if cls == 0:                              # compact: size (2 bytes) then the data
                size = self.u(lay + 2, 2)
                arr = np.frombuffer(self.bytes(lay + 4, size), dtype=dtype, count=count)
            elif cls == 1:                            # contiguous: address, size
                addr = self.u(lay + 2, self.so)
                arr = np.frombuffer(self.raw, dtype=dtype, count=count, offset=addr) if addr != \
                    (1 << (8 * self.so)) - 1 else np.zeros(count, dtype)
            else:
                raise NotImplementedError("chunked HDF5 datasets are not supported")
        elif version in (1, 2):
            rank, cls = int(self.raw[lay + 1]), int(self.raw[lay + 2])
            if cls != 1:
                raise NotImplementedError("only contiguous version-1/2 layouts are supported")
            addr = self.u(lay + 8, self.so)
            arr = np.frombuffer(self.raw, dtype=dtype, count=count, offset=addr)
        else:
            raise NotImplementedError(f"data layout message version {version}")
        return arr.reshape(shape) if shape else arr.reshape(())

    def _dataspace(self, at):
        version, rank, flags = int(self.raw[at]), int(self.raw[at + 1]), int(self.raw[at + 2])
        first = at + (8 if version == 1 else 4)
        return tuple(self.u(first + i * self.sl, self.sl) for i in range(rank))

    def _datatype(self, at):
        cls = int(self.raw[at]) & 0x0F
        bits0 = int(self.raw[at + 1])
        size = self.u(at + 4, 4)
        order = ">" if bits0 & 1 else "<"
        if cls == 0:                                   # fixed-point
            return np.dtype(f"{order}{'i' if bits0 & 0x08 else 'u'}{size}")
        if cls == 1:                                   # IEEE floating point
            return np.dtype(f"{order}f{size}")
        raise NotImplementedError(f"HDF5 datatype class {cls}")


def read_hdf5(path):
    """{name: array} of the datasets in the root group of an HDF5 file (see the module docstring)."""
    f = _File(path)
    out = {}
    for name, header in f.group_members(f.root_entry()).items():
        out[name] = f.dataset(header)
    return out


def load_example_data(path):
    """The example data set (example_diffractom_pipeline.ipynb cell 3): (data (N_Omega, My, N_eta, N_rings)
    float32 with singleton axes squeezed, direct lattice matrix (3, 3), hkl list (rings, 3), wavelength in A)."""
    d = read_hdf5(path)
    for key in ("data_array", "lattice", "hkl_list", "wavelength"):
        if key not in d:
            raise KeyError(f"{path}: no dataset {key!r}")
    data = np.ascontiguousarray(np.squeeze(d["data_array"]), dtype=np.float32)
    return data, np.asarray(d["lattice"], np.float64), np.asarray(d["hkl_list"]), float(d["wavelength"])
