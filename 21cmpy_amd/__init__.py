"""MI355X-native semi-numerical reionization hot path (yunfanz/21cmpy drop-in).

The package name starts with a digit, so import it with
``importlib.import_module("21cmpy_amd")`` (or its submodules the same way).
The stages live in ``Programs`` with the reference's module and function
names; the GPU work goes through the C ABI of ``lib21cm_hip.so``
(include/cm21.h) via ``_lib``.
"""
__version__ = "0.1.0"
