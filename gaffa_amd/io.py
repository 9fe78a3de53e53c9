"""Minimal SIGPROC filterbank container (the input type of the dedispersion API).

Mirrors the header fields of gaffa's ``FilterbankHeader`` (include/gaffa/
filterbank.h:28-64) and its default channel order: channels are stored in
ascending frequency, so files with ``foff < 0`` are reversed on read
(src/gaffa/io/filterbank.cpp:74-118).  ``Filterbank.from_array`` wraps
samples already in memory (the synthetic inputs of the benchmark and tests).
"""

from __future__ import annotations

import os
import struct
from dataclasses import dataclass, field

import numpy as np

_INT_KEYS = {"telescope_id", "machine_id", "data_type", "barycentric", "pulsarcentric",
             "ibeam", "nbeams", "npuls", "nbins", "nbits", "nifs", "nchans", "nsamples"}
_DOUBLE_KEYS = {"az_start", "za_start", "src_raj", "src_dej", "tstart", "tsamp", "fch1",
                "foff", "refdm", "period"}
_STRING_KEYS = {"rawdatafile", "source_name"}
_DTYPES = {8: np.uint8, 16: np.uint16, 32: np.float32}


@dataclass
class FilterbankHeader:
    header_size: int = 0
    nsamples: int = 0
    telescope_id: int = 0
    machine_id: int = 0
    data_type: int = 1
    barycentric: int = 0
    pulsarcentric: int = 0
    ibeam: int = 0
    nbeams: int = 0
    npuls: int = 0
    nbins: int = 0
    nbits: int = 0
    nifs: int = 1
    nchans: int = 0
    az_start: float = 0.0
    za_start: float = 0.0
    src_raj: float = 0.0
    src_dej: float = 0.0
    tstart: float = 0.0
    tsamp: float = 0.0
    fch1: float = 0.0
    foff: float = 0.0
    refdm: float = 0.0
    period: float = 0.0
    rawdatafile: str = ""
    source_name: str = ""
    frequency_table: list = field(default_factory=list)
    uses_frequency_table: bool = False


def _read_string(buf: bytes, pos: int) -> tuple[str, int]:
    (n,) = struct.unpack_from("<i", buf, pos)
    if n < 0 or n > 4096:
        raise RuntimeError("Non-standard filterbank file: bad header string length")
    return buf[pos + 4: pos + 4 + n].decode("ascii"), pos + 4 + n


def read_header(buf: bytes) -> FilterbankHeader:
    h = FilterbankHeader()
    key, pos = _read_string(buf, 0)
    if key != "HEADER_START":
        raise RuntimeError("Non-standard filterbank file: missing HEADER_START")
    freqs: list[float] = []
    while True:
        key, pos = _read_string(buf, pos)
        if key == "HEADER_END":
            break
        if key in _STRING_KEYS:
            val, pos = _read_string(buf, pos)
            setattr(h, key, val)
        elif key in _INT_KEYS:
            (val,) = struct.unpack_from("<i", buf, pos)
            pos += 4
            if key != "nsamples":  # recomputed from the file size, as gaffa does
                setattr(h, key, val)
        elif key in _DOUBLE_KEYS:
            (val,) = struct.unpack_from("<d", buf, pos)
            pos += 8
            setattr(h, key, val)
        elif key in ("FREQUENCY_START", "FREQUENCY_END"):
            h.uses_frequency_table = True
        elif key == "fchannel":
            (val,) = struct.unpack_from("<d", buf, pos)
            pos += 8
            freqs.append(val)
        else:
            raise RuntimeError(f"Unknown filterbank header field: {key}")
    h.header_size = pos
    if h.uses_frequency_table:
        h.frequency_table = freqs
        h.nchans = len(freqs)
    if h.nchans <= 0:
        raise RuntimeError("Invalid filterbank nchans")
    if h.nbits not in _DTYPES:
        raise RuntimeError("Unsupported filterbank nbits")
    return h


class Filterbank:
    """Filterbank samples (nsamples, nifs, nchans), channels ascending in frequency."""

    def __init__(self, path: str | os.PathLike | None = None, *, header: FilterbankHeader | None = None,
                 data: np.ndarray | None = None):
        if path is not None:
            with open(path, "rb") as f:
                raw = f.read()
            header = read_header(raw)
            dtype = _DTYPES[header.nbits]
            row = header.nifs * header.nchans * np.dtype(dtype).itemsize
            payload = len(raw) - header.header_size
            header.nsamples = payload // row
            data = np.frombuffer(raw, dtype=dtype, count=header.nsamples * header.nifs * header.nchans,
                                 offset=header.header_size)
            data = data.reshape(header.nsamples, header.nifs, header.nchans)
            if not header.uses_frequency_table:
                header.frequency_table = [header.fch1 + c * header.foff for c in range(header.nchans)]
            ft = header.frequency_table
            descending = len(ft) >= 2 and ft[0] > ft[-1]
            reverse = descending if header.uses_frequency_table else header.foff < 0.0
            if reverse:
                data = data[:, :, ::-1]
                header.frequency_table = list(reversed(ft))
                header.fch1 = header.frequency_table[0]
                if not header.uses_frequency_table:
                    header.foff = abs(header.foff)
            data = np.ascontiguousarray(data)
        if header is None or data is None:
            raise ValueError("Filterbank needs a path or both header and data")
        self.header = header
        self.data = data

    @classmethod
    def from_array(cls, samples: np.ndarray, *, tsamp: float, frequency_mhz) -> "Filterbank":
        """Wrap (nsamples, nchans) or (nsamples, 1, nchans) samples with ascending frequencies."""
        samples = np.asarray(samples)
        if samples.ndim == 2:
            samples = samples.reshape(samples.shape[0], 1, samples.shape[1])
        freq = [float(f) for f in frequency_mhz]
        nbits = {np.dtype(np.uint8): 8, np.dtype(np.uint16): 16, np.dtype(np.float32): 32}[samples.dtype]
        h = FilterbankHeader(nsamples=samples.shape[0], nbits=nbits, nifs=samples.shape[1],
                             nchans=samples.shape[2], tsamp=float(tsamp), fch1=freq[0],
                             foff=(freq[1] - freq[0]) if len(freq) > 1 else 0.0,
                             frequency_table=freq)
        return cls(header=h, data=np.ascontiguousarray(samples))

    def __repr__(self) -> str:
        h = self.header
        return f"<Filterbank nsamples={h.nsamples} nchans={h.nchans} nbits={h.nbits} tsamp={h.tsamp}>"


__all__ = ["Filterbank", "FilterbankHeader", "read_header"]


def write_filterbank(path: str | os.PathLike, samples: np.ndarray, *, tsamp: float, fch1: float,
                     foff: float, source_name: str = "", tstart: float = 0.0) -> None:
    """Write an 8-bit SIGPROC filterbank: ``samples`` is (nsamples, nchans) in
    the file's channel order (channel c at fch1 + c * foff)."""
    from ._core import _write_filterbank_u8

    _write_filterbank_u8(os.fspath(path), np.ascontiguousarray(samples, dtype=np.uint8),
                         tsamp=tsamp, fch1=fch1, foff=foff, source_name=source_name,
                         tstart=tstart)
