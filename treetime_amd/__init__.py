"""MI355X (gfx950) Felsenstein-pruning engine with TreeTime's TreeAnc / GTR surface."""
version = "0.1.0"
__version__ = version

from .errors import MissingDataError, NotReadyError, TreeTimeError, UnknownMethodError  # noqa: F401
from .gtr import GTR  # noqa: F401
from . import seq_utils, config  # noqa: F401
