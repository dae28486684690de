"""save_img (src/pmmoto/io/output.py:148-262) for one GPU holding the whole
domain: a VTK XML ImageData file ``<file_name>.vti`` with the image (and any
additional images) as cell data.

The file layout follows what the reference's writer produces
(io/evtk/vtk.py): ``header_type="UInt64"``, little endian, every array
``format="appended"`` in one raw AppendedData block, each preceded by its
byte count as a UInt64, values in Fortran order (x fastest); WholeExtent =
``start .. start + voxels`` per axis (the subdomain's padded start, negative
by the pad widths), Origin = the domain box's lower corner, Spacing = the
resolution.
"""

from __future__ import annotations

import os
import struct

import numpy as np

__all__ = ["save_img"]

_VTK_TYPES = {"int8": "Int8", "uint8": "UInt8", "int16": "Int16", "uint16": "UInt16", "int32": "Int32",
              "uint32": "UInt32", "int64": "Int64", "uint64": "UInt64", "float32": "Float32",
              "float64": "Float64"}


def _host(a, name):
    try:
        import torch
        if isinstance(a, torch.Tensor):
            return a.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    if not isinstance(a, np.ndarray):
        raise TypeError(f"Expected numpy array for {name}, but got {type(a)}")
    return a


def _floats(v):
    return " ".join(repr(float(x)) for x in v)


def save_img(file_name: str, subdomain, img, additional_img=None, **kwargs) -> None:
    """Write ``file_name + ".vti"``.  ``additional_img`` maps names to
    arrays of the image's shape (TypeError for a non-array, ValueError for a
    wrong shape, output.py:238-253).  Several subdomains (dicts of
    subdomains and images, or a decomposed domain) are not supported: the
    whole domain is one array here."""
    if isinstance(subdomain, dict) or isinstance(img, dict) or subdomain.domain.num_subdomains > 1:
        raise NotImplementedError("one GPU holds the whole domain: save the single image")
    arr = _host(img, "img")
    data = {"img": arr}
    for key, value in (additional_img or {}).items():
        v = _host(value, key)
        if v.shape != arr.shape:
            raise ValueError(f"Invalid img size! {key} has shape {v.shape}. Required shape is {arr.shape} ")
        data[key] = v
    folder = os.path.dirname(file_name)
    if folder:
        os.makedirs(folder, exist_ok=True)  # io_utils.check_file_path
    start = [int(s) for s in subdomain.start]
    end = [s + int(n) for s, n in zip(start, arr.shape)]
    ext = " ".join(f"{s} {e}" for s, e in zip(start, end))
    origin = [float(b[0]) for b in subdomain.domain.box]
    spacing = subdomain.domain.resolution
    arrays, offset, head = [], 0, []
    for name, a in data.items():
        if a.dtype == np.bool_:
            a = a.astype(np.uint8)
        t = _VTK_TYPES.get(a.dtype.name)
        if t is None:
            raise TypeError(f"{name}: dtype {a.dtype} has no VTK type")
        head.append(f'<DataArray Name="{name}" NumberOfComponents="1" type="{t}" format="appended" '
                    f'offset="{offset}"/>')
        raw = np.ravel(a, order="F").astype(a.dtype.newbyteorder("<"), copy=False).tobytes()
        arrays.append(raw)
        offset += 8 + len(raw)
    xml = ('<?xml version="1.0"?>\n'
           '<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n'
           f'<ImageData WholeExtent="{ext}" Origin="{_floats(origin)}" Spacing="{_floats(spacing)}">\n'
           f'<Piece Extent="{ext}">\n<CellData>\n' + "\n".join(head) + '\n</CellData>\n</Piece>\n</ImageData>\n'
           '<AppendedData encoding="raw">\n_')
    with open(file_name + ".vti", "wb") as f:
        f.write(xml.encode("utf-8"))
        for raw in arrays:
            f.write(struct.pack("<Q", len(raw)))
            f.write(raw)
        f.write(b"\n</AppendedData>\n</VTKFile>\n")
