"""Numerical constants of the pruning path (values of treetime/config.py:1-49)."""
VERBOSE = 3
BIG_NUMBER = 1e10
TINY_NUMBER = 1e-12
SUPERTINY_NUMBER = 1e-24
MIN_LOG = -1e8
MIN_BRANCH_LENGTH = 1e-3  # fraction of one_mutation used as the lower cut-off of branch lengths
MAX_BRANCH_LENGTH = 4.0
SUCCESS = "success"
ERROR = "error"
