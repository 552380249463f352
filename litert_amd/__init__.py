"""MI355X SBS ray-casting library.

``litert_amd.api`` is the ctypes binding of liblrt_hip.so (importing it fails
when the library is not built); ``litert_amd.build`` builds the library.
"""
