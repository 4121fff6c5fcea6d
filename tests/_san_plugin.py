"""pytest plugin for the sanitizer run (tests/test_host_sanitized.py): points
the product's host library and the oracle at their AddressSanitizer /
UndefinedBehaviorSanitizer builds before any test loads them."""

import ctypes
import os


def pytest_configure(config):
    from mpk import _native

    import oracle as O
    host = ctypes.CDLL(os.environ["MPK_SAN_HOST_LIB"])
    for name, (argt, rest) in _native.SIGNATURES.items():
        fn = getattr(host, name, None)
        if fn is not None:  # the host library holds the set-up routines only
            fn.argtypes = argt
            fn.restype = rest
    _native._lib = host
    orc = ctypes.CDLL(os.environ["MPK_SAN_ORACLE_LIB"])
    for name, (argt, rest) in O._SIG.items():
        fn = getattr(orc, name)
        fn.argtypes = argt
        fn.restype = rest
    O._lib = orc
    O.build = lambda: os.environ["MPK_SAN_ORACLE_LIB"]
