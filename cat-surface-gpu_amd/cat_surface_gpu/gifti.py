"""Minimal GIFTI surface reader/writer (XML + base64, optionally zlib).

The reference reads and writes GIFTI through nibabel
(src/cat_surface_gpu/gpu_pipeline.py:52-110); nibabel is not a dependency
here, so this module implements the subset a surface needs: point-set and
triangle data arrays in ASCII, Base64Binary or GZipBase64Binary encoding,
either endianness and either index order.

Speed (the registration itself takes ~30 ms per hemisphere on MI355X, so file
I/O matters): the payloads of a file's arrays are decoded in parallel threads
(zlib and base64 release the GIL), identical payloads of one read (a surface
and its sphere share the triangles) are decoded once, an
array written again unchanged (the output sphere's triangles are the input
sphere's) reuses the payload it was read from, float payloads are stored
uncompressed inside the zlib stream (vertex coordinates compress by ~10% at
level 1, at ~25 MB/s) and integer payloads at level 1.  Every file written is
ordinary GZipBase64Binary GIFTI.  A file of the plain shape GIFTI writers emit
is not run through an XML parser (expat holds the interpreter lock, ~7 ms per
61,442-vertex surface, as long as the inflation of its payloads): its
DataArray start tags and Data texts are located by pattern matching on the
bytes, which gives what the parser gives for such a file; anything else
(comments, entity or character references, a DOCTYPE internal subset, CDATA
holding markup, a Data element with children, or a file not closed by
</GIFTI>) goes through xml.etree.
"""

from __future__ import annotations

import base64
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass
import hashlib
import os
from pathlib import Path
import re
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
# payload text, bounded
_PAYLOADS: dict[tuple, str] = {}
_CACHE_LOCK = threading.Lock()
_CACHE_ENTRIES = 64
_POOL: ThreadPoolExecutor | None = None


def clear_caches() -> None:
    """Forget every remembered payload."""

    with _CACHE_LOCK:
        _PAYLOADS.clear()


def _pool() -> ThreadPoolExecutor:
    global _POOL
    if _POOL is None:
        # one worker per available core, at most 16 (a file's arrays and a
        # batch's files are the parallel work)
        workers = max(2, min(16, len(os.sched_getaffinity(0))))
        _POOL = ThreadPoolExecutor(max_workers=workers, thread_name_prefix="gifti")
    return _POOL


def _remember(table: dict, key, value) -> None:
    with _CACHE_LOCK:
        if len(table) >= _CACHE_ENTRIES:
            table.pop(next(iter(table)))
        table[key] = value


def _payload_key(data: np.ndarray, code: str) -> tuple:
    raw = np.ascontiguousarray(data.astype("<" + code, copy=False))
    return (code, tuple(data.shape), hashlib.sha1(memoryview(raw).cast("B")).hexdigest())


@dataclass
class _Item:
    """One DataArray's attributes, resolved, and its Data text."""

    intent: int
    code: str
    dims: list
    encoding: str
    endian: str
    order: str
    text: str | bytes
    key: tuple | None  # identifies the decoded payload (binary encodings)


def _prepare(item: tuple[dict, str | bytes]) -> _Item:
    attrs, text = item
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
    order = "F" if attrs.get("ArrayIndexingOrder") == "ColumnMajorOrder" else "C"
    text = text.strip()
    key = None
    if encoding in ("Base64Binary", "GZipBase64Binary"):
        btext = text if isinstance(text, bytes) else text.encode("ascii", "ignore")
        key = (encoding, endian + code, hashlib.sha1(btext).hexdigest())
    elif encoding != "ASCII":
        raise ValueError(f"unsupported GIFTI Encoding {encoding}")
    return _Item(intent, code, dims, encoding, endian, order, text, key)


# base64 alphabet -> 6-bit value; 255 elsewhere
_B64_VALUES = np.full(256, 255, np.uint8)
_B64_VALUES[np.frombuffer(b"ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/",
                          np.uint8)] = np.arange(64, dtype=np.uint8)


def _b64decode(text: str | bytes) -> bytes:
    """base64.b64decode for a payload of alphabet characters and end padding, in
    numpy operations, which release the interpreter lock (binascii holds it: the
    decodes of the arrays read in parallel threads ran one at a time); any other
    text goes to base64.b64decode."""

    b = text if isinstance(text, bytes) else text.encode("ascii", "ignore")
    n = len(b)
    if n < 4096 or n % 4:
        return base64.b64decode(b)
    # every group but the last (which may hold padding)
    v = _B64_VALUES[np.frombuffer(b, np.uint8, count=n - 4)]
    if int(v.max()) == 255:
        return base64.b64decode(b)
    q = v.reshape(-1, 4).astype(np.uint32)
    w = (q[:, 0] << 18) | (q[:, 1] << 12) | (q[:, 2] << 6) | q[:, 3]
    out = np.empty((w.size, 3), np.uint8)
    out[:, 0] = w >> 16
    out[:, 1] = (w >> 8) & 255
    out[:, 2] = w & 255
    return out.tobytes() + base64.b64decode(b[n - 4:])


_B64_ALPHABET = np.frombuffer(
    b"ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/", np.uint8)


def _b64encode(data: bytes) -> str:
    """base64.b64encode(data).decode(), in numpy operations for the whole 3-byte
    groups (the interpreter lock is released, so the outputs of several jobs
    encode at once) and base64 for the padded tail."""

    m = len(data) // 3
    if m < 1024:
        return base64.b64encode(data).decode("ascii")
    a = np.frombuffer(data, np.uint8, count=3 * m).reshape(-1, 3).astype(np.uint32)
    w = (a[:, 0] << 16) | (a[:, 1] << 8) | a[:, 2]
    out = np.empty((m, 4), np.uint8)
    out[:, 0] = _B64_ALPHABET[w >> 18]
    out[:, 1] = _B64_ALPHABET[(w >> 12) & 63]
    out[:, 2] = _B64_ALPHABET[(w >> 6) & 63]
    out[:, 3] = _B64_ALPHABET[w & 63]
    return (out.tobytes() + base64.b64encode(data[3 * m:])).decode("ascii")


def _payload_values(it: _Item) -> np.ndarray:
    """The flat values of a DataArray's Data text."""

    if it.encoding == "ASCII":
        text = it.text.decode("ascii") if isinstance(it.text, bytes) else it.text
        return np.asarray(text.split(), dtype=np.dtype(it.code)) if text else np.zeros(0, it.code)
    raw = _b64decode(it.text)
    if it.encoding == "GZipBase64Binary":
        raw = zlib.decompress(raw)
    return np.frombuffer(raw, dtype=np.dtype(it.endian + it.code))


def _finish(it: _Item, flat: np.ndarray) -> GiftiArray:
    count = int(np.prod(it.dims)) if it.dims else 0
    if flat.size != count:
        raise ValueError("GIFTI data array length does not match its dimensions")
    data = np.ascontiguousarray(np.asarray(flat.reshape(it.dims, order=it.order)).astype(
        np.dtype(it.code).newbyteorder("=")))
    if it.encoding == "GZipBase64Binary" and it.endian == "<" and it.order == "C":
        # an unchanged array written back reuses this payload
        _remember(_PAYLOADS, _payload_key(data, it.code),
                  it.text.decode("ascii") if isinstance(it.text, bytes) else it.text)
    return GiftiArray(data, it.intent)


def _decode(item: tuple[dict, str | bytes]) -> GiftiArray:
    """One DataArray from its attributes and the text of its Data element."""

    it = _prepare(item)
    return _finish(it, _payload_values(it))


# The plain file shape: an optional XML declaration and DOCTYPE without an
# internal subset, the GIFTI root, DataArray elements whose attribute values
# hold no references and whose Data element holds only character data.
_ROOT = re.compile(rb"(?:\xef\xbb\xbf)?\s*(?:<\?xml[^>]*\?>)?\s*(?:<!DOCTYPE[^>\[]*>)?\s*<GIFTI[\s>/]")
_ATTRIBUTE = re.compile(rb"""([A-Za-z_:][-\w:.]*)\s*=\s*(?:"([^"<]*)"|'([^'<]*)')""")
_CDATA = re.compile(rb"<!\[CDATA\[(.*?)\]\]>", re.S)


def _plain_markup(part: bytes) -> bool:
    """Whether markup outside the payloads has no comment, entity declaration or
    CDATA section holding markup."""

    if b"<!--" in part or b"<!ENTITY" in part:
        return False
    if b"<![CDATA[" in part:
        for m in _CDATA.finditer(part):
            if b"<" in m.group(1) or b">" in m.group(1):
                return False
    return True


def _scan(raw: bytes) -> list[tuple[dict, bytes]] | None:
    """(attributes, Data text) of each DataArray of a plain-shaped file, or None.

    A payload is passed over by two byte searches (its first "<", which must
    open "</Data>", and "&"); the markup between payloads is checked in full."""

    if not _ROOT.match(raw) or not raw[-64:].rstrip().endswith(b"</GIFTI>"):
        return None
    out, pos, gap = [], 0, 0
    while True:
        i = raw.find(b"<DataArray", pos)
        if i < 0:
            return out if _plain_markup(raw[gap:]) else None
        j = raw.find(b">", i)
        if j < 0 or raw[i + 10:i + 11] not in (b" ", b"\t", b"\r", b"\n", b">"):
            return None
        head = raw[i + 10:j]
        attrs = {}
        for a in _ATTRIBUTE.finditer(head):
            value = a.group(2) if a.group(2) is not None else a.group(3)
            if b"&" in value or b"\t" in value or b"\n" in value or b"\r" in value:
                return None
            attrs[a.group(1).decode("ascii")] = value.decode("utf-8")
        if _ATTRIBUTE.sub(b"", head).strip():
            return None  # self-closing, or something other than attributes
        # the Data child (not DataSpace and the like) before the DataArray's end
        text, q, after = b"", j, j
        while True:
            d = raw.find(b"<Data", q)
            close = raw.find(b"</DataArray", q, d if d >= 0 else len(raw))
            if d < 0 or close >= 0:
                break
            c = raw[d + 5:d + 6]
            if c not in (b" ", b"\t", b"\r", b"\n", b">", b"/"):
                q = d + 5
                continue
            e = raw.find(b">", d)
            if e < 0:
                return None
            tag = raw[d + 5:e].strip()
            if tag == b"/":
                after = e
                break
            if tag:
                return None
            end = raw.find(b"<", e + 1)
            if (end < 0 or raw[end:end + 6] != b"</Data"
                    or raw[end + 6:end + 7] not in (b" ", b"\t", b"\r", b"\n", b">")
                    or raw.find(b"&", e + 1, end) >= 0):
                return None
            if not _plain_markup(raw[gap:e + 1]):
                return None
            text, gap, after = raw[e + 1:end], end, end
            break
        k = raw.find(b"</DataArray", after)
        if k < 0:
            return None
        out.append((attrs, text))
        pos = k + 11


def _elements(path) -> list[tuple[dict, str | bytes]]:
    """(attributes, Data text) of each DataArray element of a GIFTI file."""

    raw = Path(path).read_bytes()
    items = _scan(raw)
    return items if items is not None else _elements_xml(raw, path)


def _elements_xml(raw: bytes, path) -> list[tuple[dict, str]]:
    """_elements through xml.etree (any well-formed file)."""

    root = ET.fromstring(raw)
    if root.tag != "GIFTI":
        raise ValueError(f"not a GIFTI file: {path}")
    return [(e.attrib, e.findtext("Data") or "") for e in root.findall("DataArray")]


def read_gifti(path: str | Path) -> list[GiftiArray]:
    """The data arrays of a GIFTI file, decoded in parallel threads."""

    return read_gifti_many([path])[0]


def read_gifti_many(paths) -> list[list[GiftiArray]]:
    """read_gifti of several files: every array of every file in parallel threads,
    each distinct payload decoded once."""

    paths = [str(p) for p in paths]
    pool = _pool()
    parsed = list(pool.map(_elements, paths))
    items = list(pool.map(_prepare, [e for elems in parsed for e in elems]))
    first: dict[tuple, int] = {}
    unique = []
    for k, it in enumerate(items):
        if it.key is None or it.key not in first:
            if it.key is not None:
                first[it.key] = len(unique)
            unique.append(k)
    values = list(pool.map(lambda k: _payload_values(items[k]), unique))
    slot = {k: n for n, k in enumerate(unique)}
    decoded = list(pool.map(
        lambda k: _finish(items[k], values[slot[k] if k in slot else first[items[k].key]]),
        range(len(items))))
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
    text = _b64encode(zlib.compress(data.tobytes(), level))
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
