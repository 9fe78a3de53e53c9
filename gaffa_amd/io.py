"""SIGPROC filterbank files (gaffa's ``gaffa.io``).

``Filterbank(path, *, channel_order=..., reverse_backend=..., io_buffer_bytes=...,
openmp_min_rows=...)`` keeps gaffa's signature, defaults and messages
(src/gaffa/python/filterbank_bindings.cpp:23-43, 164-172).  The file is read
by the native C++ reader (header checks of io/filterbank.cpp:120-234,
channel reversal of :236-373): channels are delivered in ascending
frequency unless ``channel_order="preserve_file_order"``.  The header
mirrors gaffa's ``FilterbankHeader`` (include/gaffa/filterbank.h:33-64).
``Filterbank.from_array`` wraps samples already in memory (the synthetic
inputs of the benchmark and tests).
"""

from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

DEFAULT_IO_BUFFER_BYTES = 64 * 1024 * 1024  # gaffa's default_filterbank_io_buffer_bytes


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

    def __repr__(self) -> str:
        return (f"<FilterbankHeader nsamples={self.nsamples} nifs={self.nifs} "
                f"nchans={self.nchans} nbits={self.nbits} foff={self.foff:f}>")


def _header(d: dict) -> FilterbankHeader:
    keys = FilterbankHeader.__dataclass_fields__
    return FilterbankHeader(**{k: v for k, v in d.items() if k in keys})


def read_header(path: str | os.PathLike, *,
                channel_order: str = "frequency_ascending") -> FilterbankHeader:
    """The header of a SIGPROC file, with the frequency metadata of the
    delivered channel order."""
    from ._core import _read_filterbank

    hd, _ = _read_filterbank(os.fspath(path), channel_order=channel_order, header_only=True)
    return _header(hd)


class Filterbank:
    """Filterbank samples (nsamples, nifs, nchans) and their header."""

    def __init__(self, path: str | os.PathLike | None = None, *,
                 channel_order: str = "frequency_ascending", reverse_backend: str = "auto",
                 io_buffer_bytes: int = DEFAULT_IO_BUFFER_BYTES, openmp_min_rows: int = 4096,
                 header: FilterbankHeader | None = None, data: np.ndarray | None = None):
        if path is not None:
            from ._core import _read_filterbank

            hd, data = _read_filterbank(os.fspath(path), channel_order=channel_order,
                                        reverse_backend=reverse_backend,
                                        io_buffer_bytes=io_buffer_bytes,
                                        openmp_min_rows=openmp_min_rows)
            header = _header(hd)
        if header is None or data is None:
            raise ValueError("Filterbank needs a path or both header and data")
        self.header = header
        self.data = data

    @property
    def shape(self):
        return self.data.shape

    @property
    def dtype(self):
        return self.data.dtype

    @property
    def nbytes(self):
        return self.data.nbytes

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
        return f"<Filterbank shape=({h.nsamples}, {h.nifs}, {h.nchans}) nbits={h.nbits}>"


__all__ = ["Filterbank", "FilterbankHeader", "read_header"]


def write_filterbank(path: str | os.PathLike, samples: np.ndarray, *, tsamp: float, fch1: float,
                     foff: float, source_name: str = "", tstart: float = 0.0) -> None:
    """Write an 8-bit SIGPROC filterbank: ``samples`` is (nsamples, nchans) in
    the file's channel order (channel c at fch1 + c * foff)."""
    from ._core import _write_filterbank_u8

    _write_filterbank_u8(os.fspath(path), np.ascontiguousarray(samples, dtype=np.uint8),
                         tsamp=tsamp, fch1=fch1, foff=foff, source_name=source_name,
                         tstart=tstart)
