"""Backend registry, mirroring src/miepy/backends/__init__.py:1-45.

The reference keeps a module-global backend name ('cpu' or 'gpu'),
`set_backend` raises ValueError on any other name, and `backend(name)` is a
context manager restoring the previous name.  This package registers its
MI355X backend under 'mi355x' (the default here) and keeps the name 'cpu'
for the reference's own C++ path: 'cpu' is always a valid name, and its
dispatch module is whatever the host application registers for it (the
reference's pybind module when this backend is plugged into MiePy); this
package ships no CPU implementation, so asking for the 'cpu' module without
one registered raises.  Other implementations can be added with
`register_backend(name, module)`, where `module` provides the dispatch
functions of `miepy_amd.backend` (sphere_aggregate_tmatrix,
solve_linear_system, particle_cross_sections, force_all, torque_all).

As in the reference, the selection is process-global state and not
thread-safe; the serving loop passes its backend explicitly instead.
"""

import contextlib

_registry = {"cpu": None}
_backend = "mi355x"


def register_backend(name, module):
    _registry[name] = module


def available_backends():
    return tuple(_registry)


def get_backend():
    """Return the name of the active backend."""
    return _backend


def get_backend_module(name=None):
    name = _backend if name is None else name
    if name not in _registry:
        raise ValueError(f"Unknown backend '{name}'. Choose one of {sorted(_registry)}.")
    if _registry[name] is None:
        raise NotImplementedError(f"backend '{name}' is the reference's own C++ path; register its module with "
                                  f"register_backend('{name}', module) to dispatch to it")
    return _registry[name]


def set_backend(name):
    global _backend
    if name not in _registry:
        raise ValueError(f"Unknown backend '{name}'. Choose one of {sorted(_registry)}.")
    _backend = name


@contextlib.contextmanager
def backend(name):
    """Temporarily switch the active backend."""
    global _backend
    old = _backend
    set_backend(name)
    try:
        yield
    finally:
        _backend = old
