"""BINDER's registration-and-sampling path on AMD Instinct MI355X (gfx950).

``binder_amd._core`` mirrors the reference extension module ``BINDER._core``
(Registration, NonLinearTransformation, topology_correction_3D,
compute_jacobian_determinant_3D/2D) on hand-written HIP kernels; as in the
reference package, ``binder_amd.Registration`` is an alias of that module.
``FastRegistration`` is the reference's Python orchestration on top of it.

The HIP library is loaded on first use of any of these names; if
``libbinder_hip.so`` has not been built the access raises ImportError.
"""
import importlib

from . import utils  # noqa: F401


def __getattr__(name):
    if name in ("_core", "Registration", "_backend"):
        core = importlib.import_module("._core", __name__)
        return {"_core": core, "Registration": core, "_backend": core.backend}[name]
    if name == "FastRegistration":
        # importing the submodule binds the package attribute to the module;
        # rebind it to the class, as `from .FastRegistration import
        # FastRegistration` does in the reference package
        cls = importlib.import_module(".FastRegistration", __name__).FastRegistration
        globals()["FastRegistration"] = cls
        return cls
    raise AttributeError(name)
