"""Registers this directory as the package ``kcfd_amd`` in a fresh
interpreter.  Builder processes started with ``spawn`` (batch.build_cases
after the GPU is in use) run this file first, so that the tasks and results
they exchange, which name ``kcfd_amd`` modules, unpickle there."""

import importlib.util
import os
import sys

if "kcfd_amd" not in sys.modules:
    _d = os.path.dirname(os.path.abspath(__file__))
    _spec = importlib.util.spec_from_file_location("kcfd_amd", os.path.join(_d, "__init__.py"),
                                                   submodule_search_locations=[_d])
    _m = importlib.util.module_from_spec(_spec)
    sys.modules["kcfd_amd"] = _m
    _spec.loader.exec_module(_m)
