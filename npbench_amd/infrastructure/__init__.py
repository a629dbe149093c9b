"""Harness counterpart of npbench/infrastructure (__init__.py:1-15): the
benchmark specs, framework adapters (the plugin boundary), the test runner
and its utilities.  Adapter classes named in framework_info/*.json are
resolved from this module."""

from .benchmark import Benchmark  # noqa: F401
from .framework import Framework, generate_framework  # noqa: F401
from .mi355x_framework import MI355XFramework  # noqa: F401
from .test import Test  # noqa: F401
from . import timeout, utilities  # noqa: F401
