"""Read az/za grid beams stored as UVBeam "beamfits" files into a GridBeam.

matvis loads beams through pyuvdata (``UVBeam.from_file``; the reference's
test fixture is ``src/matvis/data/NF_HERA_Dipole_small.fits``, loaded in
``tests/conftest.py:9-24``).  pyuvdata is not a dependency here, so this
module reads the file format itself:

* FITS (the published standard): 2880-byte blocks of 80-character header
  cards, big-endian data, axis NAXIS1 varying fastest;
* the beamfits layout in the primary HDU, identified by each axis's CTYPE:
  ``AZIMUTH`` and ``ZENANGLE`` (degrees), ``FREQ`` (Hz), ``FEEDIND`` (E-field
  beams, feeds named by ``FEEDLIST``) or ``STOKES`` (power beams), ``IF``
  (one spectral window), ``VECIND`` (E-field basis vectors) and ``COMPLEX``
  (real, imaginary); the header keys ``BTYPE`` (efield / power) and
  ``COORDSYS`` (az_za).  Axis values are CRVAL + (index + 1 - CRPIX) CDELT.

The result has the array layout of a pyuvdata UVBeam:
``data_array`` (Naxes_vec, Nfeeds or Npols, Nfreqs, Naxes2 = za,
Naxes1 = az).  Only az/za beams are read; HEALPix beams raise.
"""

from __future__ import annotations

import os

import numpy as np

from .beams import GridBeam

_BLOCK = 2880
_DTYPES = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}


def _parse_value(raw: str):
    raw = raw.strip()
    if raw.startswith("'"):
        end = raw.find("'", 1)
        while end != -1 and end + 1 < len(raw) and raw[end + 1] == "'":
            end = raw.find("'", end + 2)
        return raw[1:end].replace("''", "'").rstrip()
    val = raw.split("/", 1)[0].strip()
    if val == "T":
        return True
    if val == "F":
        return False
    try:
        return int(val)
    except ValueError:
        pass
    try:
        return float(val.replace("D", "E"))
    except ValueError:
        return val


def read_fits_hdus(path: str | os.PathLike):
    """All HDUs of a FITS file as (header dict, data array or None).

    Image HDUs (the primary array and IMAGE extensions) are decoded with
    BSCALE/BZERO applied; other extensions keep data None.
    """
    with open(path, "rb") as fh:
        blob = fh.read()
    hdus = []
    pos = 0
    while pos + _BLOCK <= len(blob):
        header = {}
        ended = False
        while not ended:
            if pos + _BLOCK > len(blob):
                raise ValueError(f"{path}: header runs past the end of file")
            block = blob[pos:pos + _BLOCK]
            pos += _BLOCK
            for i in range(36):
                card = block[80 * i:80 * (i + 1)].decode("ascii", "replace")
                key = card[:8].strip()
                if key == "END":
                    ended = True
                    break
                if card[8:10] == "= ":
                    header[key] = _parse_value(card[10:])
        if not header:
            break
        bitpix = int(header["BITPIX"])
        naxis = int(header.get("NAXIS", 0))
        shape = [int(header[f"NAXIS{a}"]) for a in range(1, naxis + 1)]
        n = int(np.prod(shape)) if naxis else 0
        pcount = int(header.get("PCOUNT", 0))
        gcount = int(header.get("GCOUNT", 1))
        nbytes = abs(bitpix) // 8 * gcount * (pcount + n)
        data = None
        is_image = "SIMPLE" in header or header.get("XTENSION") == "IMAGE"
        if is_image and n:
            if bitpix not in _DTYPES:
                raise ValueError(f"{path}: unsupported BITPIX {bitpix}")
            raw = np.frombuffer(blob, dtype=_DTYPES[bitpix], count=n,
                                offset=pos)
            native = np.float64 if bitpix < 0 else raw.dtype.newbyteorder("=")
            data = raw.reshape(shape[::-1]).astype(native)
            if "BSCALE" in header or "BZERO" in header:
                data = data * float(header.get("BSCALE", 1.0)) + float(
                    header.get("BZERO", 0.0))
        hdus.append((header, data))
        pos += -(-nbytes // _BLOCK) * _BLOCK
    if not hdus:
        raise ValueError(f"{path}: not a FITS file")
    return hdus


def _axis_values(h: dict, a: int) -> np.ndarray:
    n = int(h[f"NAXIS{a}"])
    crval = float(h.get(f"CRVAL{a}", 0.0))
    crpix = float(h.get(f"CRPIX{a}", 1.0))
    cdelt = float(h.get(f"CDELT{a}", 1.0))
    return crval + (np.arange(n) + 1 - crpix) * cdelt


def _feed_list(raw) -> list[str]:
    if not isinstance(raw, str):
        return ["x", "y"]
    return [s.strip().strip("'\"") for s in raw.strip("[]").split(",") if s.strip()]


def read_beamfits(path: str | os.PathLike) -> GridBeam:
    """Load an az/za beamfits file as a :class:`GridBeam`."""
    (h, data), *_ = read_fits_hdus(path)
    if data is None:
        raise ValueError(f"{path}: the primary HDU holds no beam array")
    coordsys = str(h.get("COORDSYS", "az_za")).strip()
    if coordsys != "az_za":
        raise ValueError(f'{path}: pixel coordinate system must be "az_za", '
                         f'got "{coordsys}"')
    naxis = int(h["NAXIS"])
    ctypes = {str(h.get(f"CTYPE{a}", "")).strip(): a
              for a in range(1, naxis + 1)}

    def np_axis(name):  # numpy axis of a FITS axis (C order reverses them)
        return naxis - ctypes[name]

    for need in ("AZIMUTH", "ZENANGLE", "FREQ"):
        if need not in ctypes:
            raise ValueError(f"{path}: no {need} axis (CTYPE)")
    btype = str(h.get("BTYPE", "efield")).strip().lower()
    arr = np.asarray(data)
    if "COMPLEX" in ctypes:
        c = np_axis("COMPLEX")
        arr = np.take(arr, 0, axis=c) + 1j * np.take(arr, 1, axis=c)
        removed = [ctypes["COMPLEX"]]
    else:
        removed = []

    def cur_axis(name):  # numpy axis after removing earlier axes
        a = np_axis(name)
        return a - sum(1 for r in removed if naxis - r < a)

    if "IF" in ctypes:
        ax = cur_axis("IF")
        if arr.shape[ax] != 1:
            raise ValueError(f"{path}: {arr.shape[ax]} spectral windows; one "
                             "is supported")
        arr = np.take(arr, 0, axis=ax)
        removed.append(ctypes["IF"])
    pol_name = "FEEDIND" if "FEEDIND" in ctypes else "STOKES"
    order = []
    if "VECIND" in ctypes:
        order.append(cur_axis("VECIND"))
    if pol_name in ctypes:
        order.append(cur_axis(pol_name))
    order += [cur_axis("FREQ"), cur_axis("ZENANGLE"), cur_axis("AZIMUTH")]
    arr = np.transpose(arr, order)
    if "VECIND" not in ctypes:
        arr = arr[None]
    if pol_name not in ctypes:
        arr = arr[:, None]
    az = np.deg2rad(_axis_values(h, ctypes["AZIMUTH"]))
    za = np.deg2rad(_axis_values(h, ctypes["ZENANGLE"]))
    freqs = _axis_values(h, ctypes["FREQ"])
    if btype == "power" and not np.iscomplexobj(arr):
        arr = arr.astype(np.float64)
    return GridBeam(arr, az, za, freq_array=freqs, beam_type=btype,
                    feed_array=_feed_list(h.get("FEEDLIST")))
