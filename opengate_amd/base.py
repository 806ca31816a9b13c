"""Attribute contract of the drop-in objects.

Mirrors the part of opengate.base.GateObject (base.py:148-310, 362) the photon path
uses: each class declares `user_info_defaults` {name: (default, {doc, allowed_values})},
instances get those attributes, unknown attribute names raise, nested parameter groups
(position / direction / energy) are attribute dictionaries.
"""
import copy


class Box(dict):
    """dict with attribute access (the role python-box plays in the reference)."""

    def __getattr__(self, k):
        try:
            return self[k]
        except KeyError:
            raise AttributeError(k) from None

    def __setattr__(self, k, v):
        if k not in self:
            raise AttributeError(f"unknown parameter '{k}' (allowed: {sorted(self.keys())})")
        self[k] = v

    def __deepcopy__(self, memo):
        return Box({k: copy.deepcopy(v, memo) for k, v in self.items()})


class GateObject:
    user_info_defaults = {}
    # attributes set by the framework itself (not user parameters)
    runtime_attributes = ()

    def __init__(self, name=None, **kwargs):
        object.__setattr__(self, "_user_info", {})
        for klass in reversed(type(self).__mro__):
            for k, (default, opts) in getattr(klass, "user_info_defaults", {}).items():
                self._user_info[k] = (copy.deepcopy(default), opts)
        object.__setattr__(self, "name", name)
        for k, v in kwargs.items():
            setattr(self, k, v)

    @property
    def user_info(self):
        return {k: v for k, (v, _) in self._user_info.items()}

    def __getattr__(self, k):
        ui = self.__dict__.get("_user_info")
        if ui is not None and k in ui:
            return ui[k][0]
        raise AttributeError(f"{type(self).__name__} has no attribute '{k}'")

    def __setattr__(self, k, v):
        ui = self.__dict__.get("_user_info")
        if ui is not None and k in ui:
            opts = ui[k][1]
            allowed = opts.get("allowed_values")
            if allowed is not None and v not in allowed:
                raise ValueError(f"{type(self).__name__}.{k} must be one of {allowed}, not {v!r}")
            if opts.get("deprecated"):
                raise AttributeError(f"{type(self).__name__}.{k} is deprecated: {opts['deprecated']}")
            hook = opts.get("setter_hook")
            if hook is not None:
                v = hook(self, v)
            ui[k] = (v, opts)
            return
        if (k.startswith("_") or any(k in c.__dict__ for c in type(self).__mro__)
                or any(k in getattr(c, "runtime_attributes", ()) for c in type(self).__mro__)):
            object.__setattr__(self, k, v)
            return
        raise AttributeError(f"{type(self).__name__} '{self.name}' has no parameter '{k}'")

    def __repr__(self):
        return f"{type(self).__name__}({self.name!r})"
