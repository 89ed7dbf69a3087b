"""Ordered callback points of a PISO step (src/phipict/core/hooks.py:26-205)."""
from __future__ import annotations

from enum import Enum
from typing import Any, Callable, Mapping, Optional


class Hook(str, Enum):
    """A point of a PISO step at which registered callbacks run, in execution order.
    Callbacks receive keyword arguments ``domain``, ``local_step``, ``time_step``,
    ``total_step`` and ``total_time``."""

    PRE = "PRE"
    POST_SCALAR_SETUP = "POST_SCALAR_SETUP"
    PRE_VELOCITY_SETUP = "PRE_VELOCITY_SETUP"
    POST_VELOCITY_SETUP = "POST_VELOCITY_SETUP"
    POST_PREDICTION = "POST_PREDICTION"
    POST_PRESSURE_SETUP = "POST_PRESSURE_SETUP"
    POST_PRESSURE_RESULT = "POST_PRESSURE_RESULT"
    POST_PRESSURE_NON_ORTHO = "POST_PRESSURE_NON_ORTHO"
    POST_VELOCITY_CORRECTION = "POST_VELOCITY_CORRECTION"
    POST = "POST"

    @classmethod
    def parse(cls, name) -> "Hook":
        try:
            return cls(name)
        except ValueError:
            valid = ", ".join(h.value for h in cls)
            raise ValueError(f"Unknown hook {name!r}; valid hooks: {valid}.") from None


HookFn = Callable[..., Any]


class Hooks:
    """Callbacks per :class:`Hook`, run in registration order."""

    def __init__(self, callbacks: "Optional[Hooks | Mapping]" = None) -> None:
        self._callbacks: dict = {}
        if callbacks is None:
            return
        items = callbacks._callbacks.items() if isinstance(callbacks, Hooks) else callbacks.items()
        for hook, fns in items:
            if fns is None:
                continue
            if callable(fns):
                fns = [fns]
            for fn in fns:
                if fn is not None:
                    self.append(hook, fn)

    @classmethod
    def coerce(cls, hooks) -> "Hooks":
        return hooks if isinstance(hooks, Hooks) else cls(hooks)

    def append(self, hook, fn: HookFn) -> "Hooks":
        self._callbacks.setdefault(Hook.parse(hook), []).append(fn)
        return self

    def prepend(self, hook, fn: HookFn) -> "Hooks":
        self._callbacks.setdefault(Hook.parse(hook), []).insert(0, fn)
        return self

    def get(self, hook) -> list:
        return list(self._callbacks.get(Hook.parse(hook), []))

    def run(self, hook, **kwargs) -> None:
        for fn in self._callbacks.get(Hook.parse(hook), []):
            fn(**kwargs)

    def __contains__(self, hook) -> bool:
        return bool(self._callbacks.get(Hook.parse(hook)))
