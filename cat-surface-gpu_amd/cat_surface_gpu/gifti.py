"""Minimal GIFTI surface reader/writer (XML + base64, optionally zlib).

The reference reads and writes GIFTI through nibabel
(src/cat_surface_gpu/gpu_pipeline.py:52-110); nibabel is not a dependency
here, so this module implements the subset a surface needs: point-set and
triangle data arrays in ASCII, Base64Binary or GZipBase64Binary encoding,
either endianness and either index order.

Speed (the registration itself takes ~30 ms per hemisphere on MI355X, so file
I/O matters): the payloads of a file's arrays are decoded in parallel threads
(zlib and base64 release the GIL), identical payloads are decoded once, an
array written again unchanged (the output sphere's triangles are the input
sphere's) reuses the payload it was read from, float payloads are stored
uncompressed inside the zlib stream (vertex coordinates compress by ~10% at
level 1, at ~25 MB/s) and integer payloads at level 1.  Every file written is
ordinary GZipBase64Binary GIFTI.
"""

from __future__ import annotations

import base64
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass
import hashlib
from pathlib import Path
import threading
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


# (dtype code, shape, sha1 of the little-endian row-major bytes) -> GZipBase64Binary
# payload text, and sha1 of a payload text -> decoded flat array; both bounded
_PAYLOADS: dict[tuple, str] = {}
_DECODED: dict[str, np.ndarray] = {}
_CACHE_LOCK = threading.Lock()
_CACHE_ENTRIES = 64
_POOL: ThreadPoolExecutor | None = None


def clear_caches() -> None:
    """Forget every remembered payload and decoded array."""

    with _CACHE_LOCK:
        _PAYLOADS.clear()
        _DECODED.clear()


def _pool() -> ThreadPoolExecutor:
    global _POOL
    if _POOL is None:
        _POOL = ThreadPoolExecutor(max_workers=8, thread_name_prefix="gifti")
    return _POOL


def _remember(table: dict, key, value) -> None:
    with _CACHE_LOCK:
        if len(table) >= _CACHE_ENTRIES:
            table.pop(next(iter(table)))
        table[key] = value


def _payload_key(data: np.ndarray, code: str) -> tuple:
    raw = np.ascontiguousarray(data.astype("<" + code, copy=False))
    return (code, tuple(data.shape), hashlib.sha1(memoryview(raw).cast("B")).hexdigest())


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
        text_key = hashlib.sha1(text.encode("ascii", "ignore")).hexdigest() + encoding + code
        flat = _DECODED.get(text_key)
        if flat is None:
            raw = base64.b64decode(text)
            if encoding == "GZipBase64Binary":
                raw = zlib.decompress(raw)
            flat = np.frombuffer(raw, dtype=np.dtype(endian + code), count=count)
            _remember(_DECODED, text_key, flat)
    else:
        raise ValueError(f"unsupported GIFTI Encoding {encoding}")
    if flat.size != count:
        raise ValueError("GIFTI data array length does not match its dimensions")
    order = "F" if attrs.get("ArrayIndexingOrder") == "ColumnMajorOrder" else "C"
    data = np.ascontiguousarray(
        np.asarray(flat.reshape(dims, order=order)).astype(np.dtype(code).newbyteorder("=")))
    if encoding == "GZipBase64Binary" and endian == "<" and order == "C":
        # an unchanged array written back reuses this payload
        _remember(_PAYLOADS, _payload_key(data, code), text)
    return GiftiArray(data, intent)


def _elements(path) -> list[ET.Element]:
    root = ET.parse(str(path)).getroot()
    if root.tag != "GIFTI":
        raise ValueError(f"not a GIFTI file: {path}")
    return root.findall("DataArray")


def read_gifti(path: str | Path) -> list[GiftiArray]:
    """The data arrays of a GIFTI file, decoded in parallel threads."""

    elems = _elements(path)
    if len(elems) < 2:
        return [_decode(e) for e in elems]
    return list(_pool().map(_decode, elems))


def read_gifti_many(paths) -> list[list[GiftiArray]]:
    """read_gifti of several files: every array of every file in parallel threads."""

    paths = [str(p) for p in paths]
    parsed = list(_pool().map(_elements, paths))
    flat = [e for elems in parsed for e in elems]
    decoded = list(_pool().map(_decode, flat))
    out, k = [], 0
    for elems in parsed:
        out.append(decoded[k:k + len(elems)])
        k += len(elems)
    return out


def read_gifti_intents(path: str | Path) -> list[int]:
    """The Intent of each data array, without decoding any data."""

    out = []
    for _, elem in ET.iterparse(str(path), events=("start",)):
        if elem.tag == "DataArray":
            raw = elem.attrib.get("Intent", "NIFTI_INTENT_NONE")
            code = _INTENT_NAMES.get(raw)
            out.append(code if code is not None else (int(raw) if raw.isdigit() else 0))
    return out


def _payload(arr: np.ndarray, code: str, level: int | None = None) -> str:
    key = _payload_key(arr, code)
    if level is None:
        hit = _PAYLOADS.get(key)
        if hit is not None:
            return hit
        level = 0 if code.startswith("f") else 1
    data = np.ascontiguousarray(arr.astype("<" + code))
    text = base64.b64encode(zlib.compress(data.tobytes(), level)).decode("ascii")
    _remember(_PAYLOADS, key, text)
    return text


def _encode(arr: np.ndarray, intent: int, code: str, payload: str | None = None) -> str:
    if payload is None:
        payload = _payload(arr, code)
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
                        triangle_intent: int = INTENT_TRIANGLE,
                        compress_level: int | None = None) -> None:
    """Write a surface as GZipBase64Binary GIFTI; ``compress_level`` (0-9) forces one
    zlib level for both arrays (default: stored floats, level-1 integers, and the
    payload an unchanged array was read from)."""

    out = Path(path)
    out.parent.mkdir(parents=True, exist_ok=True)
    v = np.asarray(vertices, np.float32)
    f = np.asarray(faces, np.int32)
    vp, fp = _pool().map(lambda a: _payload(*a),
                         ((v, "f4", compress_level), (f, "i4", compress_level)))
    text = ('<?xml version="1.0" encoding="UTF-8"?>\n'
            '<!DOCTYPE GIFTI SYSTEM "http://www.nitrc.org/frs/download.php/115/gifti.dtd">\n'
            '<GIFTI Version="1.0" NumberOfDataArrays="2">\n  <MetaData/>\n  <LabelTable/>\n'
            + _encode(v, point_intent, "f4", vp)
            + _encode(f, triangle_intent, "i4", fp)
            + "</GIFTI>\n")
    out.write_text(text)
