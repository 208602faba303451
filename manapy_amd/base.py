"""Configuration objects with manapy's semantics.

``Struct`` and ``make_get_conf`` behave like ``manapy/base/base.py:8-13`` and
``:65-261``: a ``Struct`` is an attribute bag whose ``get`` returns a default
for missing keys and raises ``ValueError`` when a required key resolves to
``None``; ``make_get_conf`` lets keyword arguments override a ``Struct``.
Every solver class here builds its options from a ``_parameters`` table the
same way the reference's ``process_conf`` class methods do
(``solvers/advec/system.py:30-59``).
"""
from copy import copy, deepcopy


def make_get_conf(conf, kwargs):
    """Return ``get(name, default, msg)`` reading ``kwargs`` first, then ``conf``
    (``manapy/base/base.py:8-13``)."""

    def _get_conf_item(name, default=None, msg_if_none=None):
        return kwargs.get(name, conf.get(name, default=default,
                                         msg_if_none=msg_if_none))

    return _get_conf_item


def get_default(arg, default, msg_if_none=None):
    """``manapy/base/base.py:15-24``."""
    out = default if arg is None else arg
    if out is None and msg_if_none is not None:
        raise ValueError(msg_if_none)
    return out


class Struct(object):
    """Attribute container (``manapy/base/base.py:65-261``)."""

    def __init__(self, **kwargs):
        if kwargs:
            self.__dict__.update(kwargs)

    def __repr__(self):
        keys = ", ".join("%s=%r" % (k, v) for k, v in sorted(self.__dict__.items()))
        return "Struct(%s)" % keys

    __str__ = __repr__

    def __add__(self, other):
        """Merge: attributes of ``self`` win unless both sides are Structs."""
        new = copy(self)
        for key, val in other.__dict__.items():
            if hasattr(new, key):
                sval = getattr(self, key)
                if isinstance(sval, Struct) and isinstance(val, Struct):
                    setattr(new, key, sval + val)
                else:
                    setattr(new, key, sval)
            else:
                setattr(new, key, val)
        return new

    def __iadd__(self, other):
        for key, val in other.__dict__.items():
            if hasattr(self, key):
                sval = getattr(self, key)
                if isinstance(sval, Struct) and isinstance(val, Struct):
                    setattr(self, key, sval + val)
            else:
                setattr(self, key, val)
        return self

    def to_dict(self):
        return copy(self.__dict__)

    def get(self, key, default=None, msg_if_none=None):
        out = getattr(self, key, default)
        if out is None and msg_if_none is not None:
            raise ValueError(msg_if_none)
        return out

    def update(self, other, **kwargs):
        if other is None:
            return
        if not isinstance(other, dict):
            other = other.to_dict()
        self.__dict__.update(other, **kwargs)

    def set_default(self, key, default=None):
        return self.__dict__.setdefault(key, default)

    def copy(self, deep=False, name=None):
        other = deepcopy(self) if deep else copy(self)
        if hasattr(self, 'name'):
            other.name = get_default(name, self.name + '_copy')
        return other


def process_conf(parameters, conf, kwargs):
    """Build an options ``Struct`` from a ``(name, type, default, required, doc)``
    table, as the reference's per-class ``process_conf`` does
    (``ddm/domain.py:38-68``): a required option that resolves to ``None``
    raises ``ValueError('missing "<name>" in options!')``; a ``'*'`` row admits
    any extra key of ``conf``."""
    get = make_get_conf(conf, kwargs)
    opts = Struct()
    allow_extra = False
    for name, _, default, required, _ in parameters:
        if name == '*':
            allow_extra = True
            continue
        msg = ('missing "%s" in options!' % name) if required else None
        setattr(opts, name, get(name, default, msg))
    if allow_extra:
        other = set(conf.to_dict().keys()).difference(opts.to_dict().keys())
        for name in other:
            setattr(opts, name, get(name, None, None))
    return opts
