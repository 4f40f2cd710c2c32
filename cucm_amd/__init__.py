"""cucm_amd -- MI355X-native CuCM chunked context-mixing codec.

HIP kernels for gfx950 behind a C ABI (include/cucm.h, libcucm.so); this
package is the host-side mirror of the reference's compress/decompress
interface.
"""
from ._native import CucmError, Params, lib  # noqa: F401
from .codec import (Context, DeviceBuffer, compress, compress_bound, compress_file,  # noqa: F401
                    decompress, decompress_file, default_context, make_params, peek_header)
