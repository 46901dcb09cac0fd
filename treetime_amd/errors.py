"""Exception hierarchy with the reference's names (treetime/__init__.py:8-40)."""


class TreeTimeError(Exception):
    """TreeTimeError class"""


class MissingDataError(TreeTimeError):
    """MissingDataError class raised when tree or alignment are missing"""


class UnknownMethodError(TreeTimeError):
    """MissingDataError class raised when an unknown method is called"""


class NotReadyError(TreeTimeError):
    """NotReadyError class raised when results are requested before inference"""
