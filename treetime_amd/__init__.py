"""MI355X (gfx950) Felsenstein-pruning engine with TreeTime's TreeAnc / GTR surface."""
version = "0.1.0"
__version__ = version

from .errors import MissingDataError, NotReadyError, TreeTimeError, UnknownMethodError  # noqa: F401
from .gtr import GTR  # noqa: F401
from .gtr_site_specific import GTR_site_specific  # noqa: F401
from . import seq_utils, config  # noqa: F401
from .treeanc import TreeAnc  # noqa: F401,E402
from .tree import Tree, Clade  # noqa: F401,E402
