"""Import shim: exposes the package directory ``iblbm-mechanism-resolved_amd/`` as ``iblbm_amd``.

The directory name carries hyphens, which Python cannot import directly; this module loads
the package from it under the importable name and replaces itself in ``sys.modules``.
"""
import importlib.util
import os
import sys

_HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "iblbm-mechanism-resolved_amd")
_spec = importlib.util.spec_from_file_location(
    __name__, os.path.join(_HERE, "__init__.py"), submodule_search_locations=[_HERE])
_mod = importlib.util.module_from_spec(_spec)
sys.modules[__name__] = _mod
_spec.loader.exec_module(_mod)
