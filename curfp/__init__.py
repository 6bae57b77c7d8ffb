"""``import curfp`` for code written against the reference package
(curfp/__init__.py): this name is the ``curfp_amd`` module itself, and
``curfp._curfp_C`` its native extension, so the reference's calls, keyword
arguments and constants resolve to the MI355X implementation."""
import sys

import curfp_amd as _impl
from curfp_amd import _curfp_C

sys.modules[__name__ + "._curfp_C"] = _curfp_C
sys.modules[__name__] = _impl
