"""mdz -- MI355X-native (gfx950) decompression for the ZaidQureshi/compression codecs.

The product is the HIP library ``libmdz.so`` behind the C ABI in
``include/mdz.h``.  This package is its Python binding: one module per
reference namespace (``rle_v1``, ...), each mirroring that namespace's
``decompress_gpu`` entry plus a device-resident, stream-ordered form.
"""
from ._lib import DeviceBuffer, Event, MdzError, Stream, device_count, lib  # noqa: F401

__all__ = ["DeviceBuffer", "Event", "MdzError", "Stream", "device_count", "lib"]
