"""Minimal GIFTI surface reader/writer (XML + base64, optionally zlib).

The reference reads and writes GIFTI through nibabel
(src/cat_surface_gpu/gpu_pipeline.py:52-110); nibabel is not a dependency
here, so this module implements the subset a surface needs: point-set and
triangle data arrays in ASCII, Base64Binary or GZipBase64Binary encoding,
either endianness and either index order.
"""

from __future__ import annotations

import base64
from dataclasses import dataclass
from pathlib import Path
import xml.etree.ElementTree as ET
import zlib

import numpy as np

INTENT_POINTSET = 1008
INTENT_TRIANGLE = 1009
_INTENT_NAMES = {"NIFTI_INTENT_POINTSET": INTENT_POINTSET,
                 "NIFTI_INTENT_TRIANGLE": INTENT_TRIANGLE,
                 "NIFTI_INTENT_NONE": 0, "NIFTI_INTENT_SHAPE": 2005}
_INTENT_CODES = {v: k for k, v in _INTENT_NAMES.items()}
_DTYPES = {"NIFTI_TYPE_FLOAT32": "f4", "NIFTI_TYPE_FLOAT64": "f8", "NIFTI_TYPE_INT32": "i4",
           "NIFTI_TYPE_UINT8": "u1", "NIFTI_TYPE_INT16": "i2", "NIFTI_TYPE_INT8": "i1",
           "NIFTI_TYPE_UINT16": "u2", "NIFTI_TYPE_UINT32": "u4", "NIFTI_TYPE_INT64": "i8"}


@dataclass(frozen=True)
class GiftiArray:
    data: np.ndarray
    intent: int


def _decode(elem: ET.Element) -> GiftiArray:
    attrs = elem.attrib
    intent_raw = attrs.get("Intent", "NIFTI_INTENT_NONE")
    intent = _INTENT_NAMES.get(intent_raw, None)
    if intent is None:
        intent = int(intent_raw) if intent_raw.isdigit() else 0
    code = _DTYPES.get(attrs.get("DataType", ""), None)
    if code is None:
        raise ValueError(f"unsupported GIFTI DataType {attrs.get('DataType')}")
    dims = [int(attrs[f"Dim{i}"]) for i in range(int(attrs.get("Dimensionality", "1")))]
    encoding = attrs.get("Encoding", "ASCII")
    endian = "<" if attrs.get("Endian", "LittleEndian") == "LittleEndian" else ">"
    text = (elem.findtext("Data") or "").strip()
    count = int(np.prod(dims)) if dims else 0
    if encoding == "ASCII":
        flat = np.asarray(text.split(), dtype=np.dtype(code)) if text else np.zeros(0, code)
    elif encoding in ("Base64Binary", "GZipBase64Binary"):
        raw = base64.b64decode(text)
        if encoding == "GZipBase64Binary":
            raw = zlib.decompress(raw)
        flat = np.frombuffer(raw, dtype=np.dtype(endian + code), count=count)
    else:
        raise ValueError(f"unsupported GIFTI Encoding {encoding}")
    if flat.size != count:
        raise ValueError("GIFTI data array length does not match its dimensions")
    order = "F" if attrs.get("ArrayIndexingOrder") == "ColumnMajorOrder" else "C"
    data = np.asarray(flat.reshape(dims, order=order)).astype(np.dtype(code).newbyteorder("="))
    return GiftiArray(np.ascontiguousarray(data), intent)


def read_gifti(path: str | Path) -> list[GiftiArray]:
    root = ET.parse(str(path)).getroot()
    if root.tag != "GIFTI":
        raise ValueError(f"not a GIFTI file: {path}")
    return [_decode(e) for e in root.findall("DataArray")]


def _encode(arr: np.ndarray, intent: int, code: str) -> str:
    data = np.ascontiguousarray(arr.astype("<" + code))
    payload = base64.b64encode(zlib.compress(data.tobytes(), 6)).decode("ascii")
    dtype_name = {v: k for k, v in _DTYPES.items()}[code]
    dims = "".join(f' Dim{i}="{d}"' for i, d in enumerate(arr.shape))
    intent_name = _INTENT_CODES.get(intent, "NIFTI_INTENT_NONE")
    return (f'  <DataArray Intent="{intent_name}" DataType="{dtype_name}" '
            f'ArrayIndexingOrder="RowMajorOrder" Dimensionality="{arr.ndim}"{dims} '
            f'Encoding="GZipBase64Binary" Endian="LittleEndian" ExternalFileName="" '
            f'ExternalFileOffset="">\n    <MetaData/>\n    <Data>{payload}</Data>\n'
            f'  </DataArray>\n')


def write_gifti_surface(path: str | Path, vertices: np.ndarray, faces: np.ndarray, *,
                        point_intent: int = INTENT_POINTSET,
                        triangle_intent: int = INTENT_TRIANGLE) -> None:
    out = Path(path)
    out.parent.mkdir(parents=True, exist_ok=True)
    text = ('<?xml version="1.0" encoding="UTF-8"?>\n'
            '<!DOCTYPE GIFTI SYSTEM "http://www.nitrc.org/frs/download.php/115/gifti.dtd">\n'
            '<GIFTI Version="1.0" NumberOfDataArrays="2">\n  <MetaData/>\n  <LabelTable/>\n'
            + _encode(np.asarray(vertices, np.float32), point_intent, "f4")
            + _encode(np.asarray(faces, np.int32), triangle_intent, "i4")
            + "</GIFTI>\n")
    out.write_text(text)
