"""The tenalg plug-in boundary, checked on the CPU against a minimal stand-in
for tensorly's registry (tensorly is not installed here).  The stand-in
implements only the contract the real one documents:

* ``TenalgBackend.__init_subclass__(backend_name=...)`` records the class and
  ``register_method(name, f)`` attaches ``staticmethod(f)``
  (tensorly/tenalg/base_tenalg.py:7-33);
* the manager is the ``tensorly.tenalg`` module's class, knows
  ``available_backend_names`` / ``_loaded_backends`` and resolves every
  function of ``_functions`` through the current backend
  (tensorly/tenalg/__init__.py:16-30, tensorly/backend/__init__.py:198-204,
  319-340).

The registered functions must be the GPU implementations: called with CPU
tensors they raise the no-CPU-path TypeError."""
import sys
import types

import pytest
import torch

FUNCTIONS = ["mode_dot", "multi_mode_dot", "kronecker", "khatri_rao", "inner", "outer", "batched_outer",
             "higher_order_moment", "_tt_matrix_to_tensor", "unfolding_dot_khatri_rao", "tensordot"]


@pytest.fixture
def fake_tensorly(monkeypatch):
    class TenalgBackend:
        _available = {}

        def __init_subclass__(cls, backend_name, **kw):
            super().__init_subclass__(**kw)
            cls._available[backend_name.lower()] = cls
            cls.backend_name = backend_name

        @classmethod
        def register_method(cls, name, func):
            setattr(cls, name, staticmethod(func))

    class CoreTenalgBackend(TenalgBackend, backend_name="core"):
        pass

    for name in FUNCTIONS:
        CoreTenalgBackend.register_method(name, lambda *a, _n=name, **k: ("core", _n))

    class Manager(types.ModuleType):
        available_backend_names = ["core", "einsum"]
        _loaded_backends = {"core": CoreTenalgBackend()}
        _backend = _loaded_backends["core"]

        @classmethod
        def set_backend(cls, name):
            if name not in cls._loaded_backends:
                if name not in cls.available_backend_names:
                    raise ValueError(f"Unknown backend name {name!r}")
                cls._loaded_backends[name] = TenalgBackend._available[name]()
            cls._backend = cls._loaded_backends[name]

        def __getattr__(self, name):
            if name in FUNCTIONS:
                return getattr(type(self)._backend, name)
            raise AttributeError(name)

    pkg = types.ModuleType("tensorly")
    tenalg = Manager("tensorly.tenalg")
    core = types.ModuleType("tensorly.tenalg.core_tenalg")
    core.CoreTenalgBackend = CoreTenalgBackend
    pkg.tenalg = tenalg
    tenalg.core_tenalg = core
    monkeypatch.setitem(sys.modules, "tensorly", pkg)
    monkeypatch.setitem(sys.modules, "tensorly.tenalg", tenalg)
    monkeypatch.setitem(sys.modules, "tensorly.tenalg.core_tenalg", core)
    return tenalg


def test_register_exposes_all_functions(fake_tensorly):
    from tensorly_amd import tenalg as amd
    cls = amd.register("amd")
    assert cls is not None and cls.backend_name == "amd"
    assert "amd" in type(fake_tensorly).available_backend_names
    fake_tensorly.set_backend("amd")
    for name in FUNCTIONS:
        assert callable(getattr(fake_tensorly, name)), name
    # the seven functions off the hot path are inherited from the core backend
    assert fake_tensorly.kronecker() == ("core", "kronecker")
    assert fake_tensorly.tensordot() == ("core", "tensordot")
    # the four on the path are the GPU implementations (no CPU fallback)
    X = torch.zeros((3, 4, 5))
    fs = [torch.zeros((s, 2)) for s in X.shape]
    with pytest.raises(TypeError):
        fake_tensorly.unfolding_dot_khatri_rao(X, (None, fs), 0)
    with pytest.raises(TypeError):
        fake_tensorly.mode_dot(X, torch.zeros((2, 4)), 1)
    with pytest.raises(TypeError):
        fake_tensorly.multi_mode_dot(X, [torch.zeros((2, 3))], modes=[0])
    assert fake_tensorly.khatri_rao.__name__ == "khatri_rao"
    fake_tensorly.set_backend("core")
    assert fake_tensorly.mode_dot() == ("core", "mode_dot")


def test_register_without_tensorly_returns_none(monkeypatch):
    monkeypatch.setitem(sys.modules, "tensorly", None)
    from tensorly_amd import tenalg as amd
    assert amd.register("amd") is None
